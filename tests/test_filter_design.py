"""AC filter design of the library (host code): the Parks-McClellan low-pass
against scipy.signal.remez on the same grid, its equiripple alternation, and
the reference's equiripple procedure around it (minimum order, unit DC gain,
error cases of SubModel::analyzeACFilter)."""
import math

import numpy as np
import pytest

CASES = [(57, 600 / 44100, 1600 / 44100, 0.014125, 10.0), (30, 0.1, 0.15, 1.0, 10.0), (101, 0.2, 0.25, 1.0, 10.0),
         (76, 600 / 44100, 1600 / 44100, 0.01, 10.0), (25, 0.1, 0.2, 1.0, 1.0), (200, 0.05, 0.06, 1.0, 5.0),
         (64, 0.1, 0.14, 3.0, 1.0), (151, 0.02, 0.05, 0.2, 10.0)]


def _remez(n, fp, fs, wp, ws):
    from cochlearmodel_amd import _lib
    h = np.zeros(n)
    d = np.zeros(1)
    it = np.zeros(1, np.int32)
    _lib.check(_lib.lib().cm_remez_lowpass(n, fp, fs, wp, ws, 16, _lib.ptr(h), _lib.ptr(d), _lib.ptr(it)))
    return h, float(d[0]), int(it[0])


@pytest.mark.parametrize("n,fp,fs,wp,ws", CASES)
def test_remez_matches_scipy(n, fp, fs, wp, ws):
    signal = pytest.importorskip("scipy.signal")
    h, _, _ = _remez(n, fp, fs, wp, ws)
    ref = signal.remez(n, [0, fp, fs, 0.5], [1, 0], weight=[wp, ws], fs=1.0, grid_density=16, maxiter=200)
    assert np.abs(h - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())
    np.testing.assert_array_equal(h, h[::-1])


@pytest.mark.parametrize("n,fp,fs,wp,ws", CASES[:4])
def test_remez_equiripple_alternation(n, fp, fs, wp, ws):
    """Chebyshev alternation on the design grid: the weighted error of the real
    amplitude reaches +-delta, alternating in sign, at (number of cosine basis
    functions + 1) frequencies, and between grid points exceeds delta by at
    most a grid-discretisation margin."""
    h, delta, _ = _remez(n, fp, fs, wp, ws)
    n_basis = n // 2 if n % 2 == 0 else n // 2 + 1
    step = 0.5 / (16 * n_basis)
    signs = []
    for lo, hi, d, w in ((0.0, fp, 1.0, wp), (fs, 0.5, 0.0, ws)):
        f = np.concatenate([np.arange(lo, hi, step / 8), [hi]])
        A = np.real(np.exp(-2j * np.pi * np.outer(f, np.arange(n) - (n - 1) / 2)) @ h)
        err = w * (d - A)
        assert np.abs(err).max() <= delta * 1.02
        # local extrema of this band that reach delta
        ext = [i for i in range(len(f)) if abs(err[i]) >= delta * (1 - 1e-6) and
               (i == 0 or abs(err[i]) >= abs(err[i - 1])) and (i == len(f) - 1 or abs(err[i]) >= abs(err[i + 1]))]
        signs.extend(np.sign(err[ext]))
    alternations = 1 + np.count_nonzero(np.diff(signs) != 0)
    assert alternations >= n_basis + 1


def _design(fs, spec):
    from cochlearmodel_amd import _lib
    h = np.zeros(4096)
    a = np.zeros(4096)
    n = np.zeros(1, np.int32)
    iir = np.zeros(1, np.int32)
    rc = _lib.lib().cm_design_ac_filter(fs, spec.encode(), _lib.ptr(h), _lib.ptr(a), len(h), _lib.ptr(n),
                                        _lib.ptr(iir))
    if iir[0]:
        return rc, (h[:int(n[0])], a[:int(n[0])])
    return rc, h[:int(n[0])]


def test_ac_filter_design_procedure():
    """Default band at 44.1 kHz: the minimum order of Oppenheim & Schafer 7.104
    with delta_pass = 10^(-3/20), delta_stop = 10^(-60/20) and transition
    1000/22050 is ceil(18.4996 / 0.331108) = 56, i.e. 57 coefficients; the design
    has unit DC gain and is a low-pass."""
    rc, h = _design(44100.0, "type=equiripple,fpass=600,fstop=1600,apass=3,astop=60")
    assert rc == 0 and len(h) == 57
    assert abs(h.sum() - 1.0) < 1e-12
    np.testing.assert_array_equal(h, h[::-1])
    f = np.array([0.0, 300.0, 3000.0, 10000.0]) / 44100
    H = np.abs(np.exp(-2j * np.pi * np.outer(f, np.arange(len(h)))) @ h)
    assert abs(H[0] - 1) < 1e-12 and H[1] > 0.7 and H[2] < 0.02 and H[3] < 0.02
    # keys are case-insensitive, and the design is deterministic
    rc2, h2 = _design(44100.0, "Type=EquiRipple,Fpass=600,Fstop=1600,Apass=3,Astop=60")
    assert rc2 == 0
    np.testing.assert_array_equal(h, h2)


@pytest.mark.parametrize("spec", ["type=elliptic,fpass=600,fstop=1600", "type=equiripple,fc=800,fpass=600,astop=60",
                                  "type=equiripple,fpass=600,fstop=9000,astop=60", "type=equiripple,fpass=600",
                                  "type=equiripple,fc=800", "type=equiripple,fc=800,wstop=100"])
def test_ac_filter_design_refuses(spec):
    from cochlearmodel_amd import _lib
    rc, _ = _design(44100.0, spec)
    assert rc == -3 and _lib.lib().cm_last_error()


@pytest.mark.parametrize("astop", [40, 50, 60, 70, 80])
def test_equiripple_order_search(astop):
    """The tap search of src/SubModel.cpp:325-346 (Fpass/Fstop with Astop): the
    order moves by ceil(fix 10 log10(ripple / delta_stop)) until the stop-band
    ripple is within a factor 3 of delta_stop = 10^(-Astop/20).  Measured deep
    in the stop band (above 5 kHz), where the equiripple error is the design's
    ripple.  At 40 dB the minimum order of O&S 7.104 misses the target, and the
    search adds taps."""
    rc, h = _design(44100.0, f"type=equiripple,fpass=600,fstop=1600,astop={astop}")
    assert rc == 0
    ratio = _gain(h, np.linspace(5000, 22050, 4000), 44100.0).max() / 10 ** (-astop / 20)
    assert 1 / 3 <= ratio <= 3, ratio
    np.testing.assert_array_equal(h, h[::-1])
    assert abs(h.sum() - 1.0) < 1e-12
    if astop == 40:
        dp, ds = 10 ** (-1 / 20), 1e-2       # Apass defaults to 1 dB when only Astop is given
        min_taps = math.ceil((-10 * math.log10(dp * ds) - 13) / (2.324 * math.pi * 1000 / 22050)) + 1
        assert min_taps == 24 and len(h) > min_taps


@pytest.mark.parametrize("astop", [40, 60, 80])
def test_equiripple_fc_with_astop(astop):
    """Fc with Astop (src/SubModel.cpp:181-183, 219-220, 280-324): the transition
    width is searched from 0.15 of Nyquist by bisection towards 0.3 or 0.01
    until the stop-band ripple is within a factor 3 of the target; the
    minimum order follows from the starting width.  The corner correction aims
    the -Apass (1 dB) point at Fc."""
    rc, h = _design(44100.0, f"type=equiripple,fc=800,astop={astop}")
    assert rc == 0
    # the stop band starts at most 0.3 x 22050 above Fc
    ratio = _gain(h, np.linspace(800 + 0.3 * 22050, 22050, 4000), 44100.0).max() / 10 ** (-astop / 20)
    assert 1 / 3 <= ratio <= 3, ratio
    np.testing.assert_array_equal(h, h[::-1])
    assert abs(h.sum() - 1.0) < 1e-12
    g = _gain(h, np.array([0.0, 200.0]), 44100.0)
    assert abs(g[0] - 1) < 1e-12 and g[1] > 10 ** (-1.5 / 20)


def test_equiripple_fc_with_order():
    """Fc with an order and no Astop: zero transition width (the reference's
    bands [0, wc] and [wc, 1] touch), taken here as a stop band starting one
    grid step above wc; the result is a linear-phase design of order + 1 taps
    with unit DC gain, whatever its (poor) selectivity."""
    rc, h = _design(44100.0, "type=equiripple,fc=800,order=60")
    assert rc == 0 and len(h) == 61
    np.testing.assert_array_equal(h, h[::-1])
    assert abs(h.sum() - 1.0) < 1e-12


def _gain(h, f_hz, fs):
    return np.abs(np.exp(-2j * np.pi * np.outer(np.atleast_1d(f_hz) / fs, np.arange(len(h)))) @ h)


def _f3db(h, fs):
    f = np.arange(0.0, 5000.0, 0.5)
    g = _gain(h, f, fs)
    return f[np.nonzero(g > 10 ** (-3 / 20))[0].max()]


def test_window_design_kaiser():
    """Kaiser window branch: minimum length ceil((A - 7.95) / (2.285 dw)) with
    A = 60 dB and dw = pi 1000/22050 is 160; after the corner correction the
    -3 dB point sits at the mid-band cut-off (600 + 1600) / 2 Hz."""
    rc, h = _design(44100.0, "type=window,windowtype=kaiser,fpass=600,fstop=1600,astop=60")
    assert rc == 0 and len(h) == 160
    assert abs(h.sum() - 1.0) < 1e-12
    np.testing.assert_allclose(h, h[::-1], rtol=0, atol=1e-15)
    assert abs(_f3db(h, 44100.0) - 1100.0) < 0.02 * 1100.0
    assert 20 * np.log10(_gain(h, np.arange(1800.0, 22050.0, 25.0), 44100.0).max()) < -45.0


def test_window_design_unwindowed_names_and_fc_form():
    """Window names the reference's FIRFilterWindow does not implement (Hamming
    here) leave the sinc unwindowed, with the Hamming length int(8 pi / dw);
    the Fc form takes the given order."""
    rc, h = _design(44100.0, "type=window,windowtype=hamming,fpass=600,fstop=1600")
    assert rc == 0 and len(h) == int(8 * np.pi / (1000.0 / 22050.0))
    rc2, h2 = _design(44100.0, "type=window,windowtype=none,fpass=600,fstop=1600")
    assert rc2 == 0 and len(h2) == int(4 * np.pi / (1000.0 / 22050.0))
    rc3, h3 = _design(44100.0, "type=window,windowtype=kaiser,fc=800,order=101,astop=60")
    assert rc3 == 0 and len(h3) == 101 and abs(h3.sum() - 1) < 1e-12
    assert abs(_f3db(h3, 44100.0) - 800.0) < 0.05 * 800.0


@pytest.mark.parametrize("spec", ["type=window,fpass=600,fstop=1600", "type=window,windowtype=kaiser,fc=800",
                                  "type=window,windowtype=kaiser,fpass=600,fstop=1600",
                                  "type=window,windowtype=sinc,fpass=600,fstop=1600"])
def test_window_design_refuses(spec):
    rc, _ = _design(44100.0, spec)
    assert rc == -3


def test_butterworth_design():
    """Butterworth branch: for Apass 3 dB, Astop 60 dB, 600/1600 Hz the order
    ceil(ln(sqrt(10^6 - 1) / eps) / ln(1600/600)) with eps^2 = 10^0.3 - 1 is
    ceil(7.045) = 8, i.e. 4 sections and 17 coefficients (5-term section
    products); the response is the bilinear image of the analog Butterworth,
    |H|^2 = 1 / (1 + (tan(pi f / Fs) / tan(pi wc / 2))^16), wc = wp eps^(-1/8),
    and equals scipy.signal.butter of the same order and cut-off."""
    signal = pytest.importorskip("scipy.signal")
    rc, (b, a) = _design(44100.0, "type=butterworth,fpass=600,fstop=1600,apass=3,astop=60")
    assert rc == 0 and len(b) == 17 and len(a) == 17
    assert np.all(b[9:] == 0) and np.all(a[9:] == 0) and a[0] == 1.0
    eps = np.sqrt(10 ** 0.3 - 1)
    wc = 600 / 22050 * eps ** (-1 / 8)
    bs, as_ = signal.butter(8, wc)
    np.testing.assert_allclose(b[:9], bs, rtol=1e-9, atol=1e-20)
    np.testing.assert_allclose(a[:9], as_, rtol=1e-9, atol=1e-12)
    f = np.linspace(10.0, 5000.0, 400)
    sos = signal.tf2sos(b[:9], a[:9])
    _, H = signal.sosfreqz(sos, worN=f, fs=44100.0)
    ref = 1 / np.sqrt(1 + (np.tan(np.pi * f / 44100) / np.tan(np.pi * wc / 2)) ** 16)
    np.testing.assert_allclose(np.abs(H), ref, atol=1e-5)   # direct-form coefficients of order 8 carry ~5e-6
    # stop-band mode: the cut-off comes from Fstop and Astop instead
    rc2, (b2, a2) = _design(44100.0, "type=butterworth,fpass=600,fstop=1600,apass=3,astop=60,mode=stopband")
    wc2 = 1600 / 22050 * (eps / np.sqrt(1e6 - 1)) ** (1 / 8)
    bs2, as2 = signal.butter(8, wc2)
    assert rc2 == 0
    np.testing.assert_allclose(a2[:9], as2, rtol=1e-9, atol=1e-12)
