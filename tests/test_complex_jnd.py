"""Complex-JND aggregation (host routine of the library): hand-worked cases of
the minimum and gradient methods of src/ComplexJNDProfile.cpp:19-95."""
import numpy as np


def _run(levels_by_profile, n_freq, n_noises, methods=3, eps=1.0, failed_by_profile=None):
    from cochlearmodel_amd import stages
    n_levels = len(levels_by_profile[0])
    raw = np.zeros(n_freq * n_noises * n_levels)
    failed = np.zeros(raw.size, np.int32)
    # raw order: noise -> level -> frequency (calculated intervals of updateInputProfile)
    for nl in range(n_noises):
        for fl in range(n_freq):
            for pl in range(n_levels):
                idx = (nl * n_levels + pl) * n_freq + fl
                raw[idx] = levels_by_profile[nl * n_freq + fl][pl]
                if failed_by_profile is not None:
                    failed[idx] = failed_by_profile[nl * n_freq + fl][pl]
    return stages.complex_jnd(raw, n_freq, n_levels, n_noises, methods, eps, failed)


def test_interval_map_matches_profile_order():
    from cochlearmodel_amd import _lib
    iv = np.zeros(2 * 3 * 2, np.int32)
    _lib.check(_lib.lib().cm_complex_jnd_intervals(2, 3, 2, _lib.ptr(iv)))
    # profile (noise 1, freq 0), levels 0..2 -> calculated indexes 6, 8, 10
    np.testing.assert_array_equal(iv.reshape(4, 3)[2], [6, 8, 10])
    np.testing.assert_array_equal(iv.reshape(4, 3)[1], [1, 3, 5])


def test_minimum_method_cases():
    profiles = [
        [5.0, 3.0, 4.0, 2.0, 6.0],    # two interior minima: the last one (2.0)
        [5.0, 4.0, 3.0, 2.5, 2.0],    # decreasing: boundary minimum at the top level, warning
        [1.0, 2.0, 3.0, 4.0, 5.0],    # increasing: boundary minimum at the first level, warning
        [2.0, 2.0, 2.0, 2.0, 2.0],    # flat: nothing qualifies -> 1111, warning
    ]
    failed = [[0, 0, 0, 1, 0], [0, 0, 0, 0, 1], [1, 0, 0, 0, 0], [0] * 5]
    val, warn, conv = _run(profiles, 2, 2, methods=1, failed_by_profile=failed)
    np.testing.assert_array_equal(val[0], np.float32([2.0, 2.0, 1.0, 1111.0]))
    np.testing.assert_array_equal(warn[0], [0, 1, 1, 1])
    np.testing.assert_array_equal(conv[0], [1, 1, 1, 0])


def test_gradient_method_cases():
    profiles = [
        [5.0, 5.5, 7.0, 8.0],         # first rise above eps=1 is 5.5 -> 7.0: value 5.5
        [9.0, 8.5, 8.4, 8.0],         # never rises: first level, warning
        [1.0, 2.5, 2.6, 9.0],         # rise at the first step: 1.0
    ]
    val, warn, conv = _run(profiles, 3, 1, methods=2, eps=1.0)
    np.testing.assert_array_equal(val[0], np.float32([5.5, 9.0, 1.0]))
    np.testing.assert_array_equal(warn[0], [0, 1, 0])
    # both methods: outputs stacked [minimum, gradient]
    val2, warn2, _ = _run(profiles, 3, 1, methods=3, eps=1.0)
    assert val2.shape == (2, 3)
    np.testing.assert_array_equal(val2[1], val[0])


def test_failed_signals_per_calculated_interval():
    from cochlearmodel_amd import stages
    # 4 intervals of 3 blocks; interval 3 is the reference
    fb = np.array([0, 1, 0, 0, 0, 0, 1, 1, 1, 1, 0, 0], np.int32)
    refs = np.array([3], np.int32)
    serial = np.array([0, 1, 2, 0], np.int32)
    out = stages.jnd_failed_signals(fb, 4, refs, serial, 3, 3)
    np.testing.assert_array_equal(out, [1, 0, 3])
    # second chunk of the same run (handled = 4 intervals) accumulates into the same counts
    out2 = stages.jnd_failed_signals(np.zeros(3, np.int32), 1, np.zeros(0, np.int32), np.zeros(5, np.int32), 3, 3,
                                     handled=4)
    np.testing.assert_array_equal(out2, [0, 0, 0])
