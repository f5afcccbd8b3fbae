"""GPU parity of the kernel-17 BM solver against the CPU oracle (through the C-ABI)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _flush(msg):
    print(msg, flush=True)


def test_c1_voice_parity(oracle):
    from cochlearmodel_amd import solver, workloads
    _flush("case: C1 voice tone, 5 blocks x 882 + 220 overlap")
    sig = workloads.c1_signal()
    lay = solver.voice_layout(44100.0, 5, 0.02, 0.005)
    assert (lay.block_nodes, lay.overlap_nodes, lay.bufP1) == (882, 220, 5512)
    g = solver.run_bm_voice(sig, lay)
    from cochlearmodel_amd import _lib
    assert _lib.lib().BMOHCKernel_Wait_Threads() == 0
    # oracle on the same inputs
    sc = oracle.default_scalars(44100.0)
    st, mo, me, ws = oracle.block_decomposition(5, lay.bufP1, lay.overlap_nodes, 0)
    np.testing.assert_array_equal(g["blocks"]["start"], st)
    np.testing.assert_array_equal(g["blocks"]["mosi"], mo)
    np.testing.assert_array_equal(g["blocks"]["mesi"], me)
    np.testing.assert_array_equal(g["blocks"]["write_start"], ws)
    items = oracle.make_items(st, mo, me, ws, 1e-10, 1e-15)
    o = oracle.bm_run(sc, oracle.section_params(), np.full(256, 0.5, np.float32), items, g["input"], lay.bufP1,
                      "f32", threads=8, diagnostics=True)
    rows = slice(lay.overlap_nodes, lay.total_nodes)
    vg, vo = g["speeds"][rows], o["speeds"][rows]
    rel = np.linalg.norm(vg - vo) / np.linalg.norm(vo)
    steps_g, steps_o = g["stats"]["steps"].sum(), o["steps"].sum()
    _flush(f"rel L2 {rel:.3e}, exact rows {np.mean(np.all(vg == vo, axis=1)):.4f}, steps {steps_g} vs {steps_o}")
    assert rel <= 1e-4
    assert abs(int(steps_g) - int(steps_o)) <= 0.01 * steps_o
    np.testing.assert_array_equal(g["failed_nodes"][rows], o["failed_nodes"][rows])


def test_relative_error_tolerances():
    """Relative_Error_Parameters != 0 (src/cochlea.cu:857-963): each block's
    tolerances scale with the power of the largest |input| among the first
    2^floor(log2(min(block length, 1024))) samples of the block,
    power_dB = 20 log10(0.1 max / SPLRef), m1tol = MAX_M1_SP_ERROR 10^(M1_SP_Fix_Factor dB),
    tol = MAX_TOLERANCE 10^(Tolerance_Fix_Factor dB), 0 dB for a silent window.
    Blocks of 441 + 220 overlap samples scan 512 samples from b * 441."""
    from cochlearmodel_amd import solver
    _flush("case: relative tolerances, 5 blocks of 441 + 220 overlap, spikes at known places")
    lay = solver.voice_layout(44100.0, 5, 0.01, 0.005)
    assert (lay.block_nodes, lay.overlap_nodes) == (441, 220)
    sig = np.zeros(lay.bufP1, np.float32)
    sig[100] = -0.02              # block 0: |x| max 0.02 -> 40 dB
    sig[441 + 300] = 0.2          # block 1: 60 dB (also in no other window)
    sig[882 + 20] = 0.002         # block 2: 20 dB; also inside block 1's window, below its maximum
    sig[882 + 512] = 2.0          # first sample past block 2's window: block 3 sees it (80 dB)
    # block 4 (samples 1764..2275) is silent: 0 dB
    m1, tol, fm, ft = 1e-15, 1e-10, 0.05, 0.025
    g = solver.run_bm_voice(sig, lay, relative_error=1, max_m1_sp_error=m1, max_tolerance=tol, m1_fix=fm, tol_fix=ft)
    db = np.array([40.0, 60.0, 20.0, 80.0, 0.0])
    _flush(f"m1tol {g['blocks']['m1tol']}, tol {g['blocks']['tol']}")
    np.testing.assert_allclose(g["blocks"]["m1tol"], m1 * 10 ** (fm * db), rtol=2e-5)
    np.testing.assert_allclose(g["blocks"]["tol"], tol * 10 ** (ft * db), rtol=2e-5)
    assert np.all(np.isfinite(g["speeds"]))
    # the absolute branch on the same input: every block at the given values
    g0 = solver.run_bm_voice(sig, lay, relative_error=0, max_m1_sp_error=m1, max_tolerance=tol, m1_fix=fm, tol_fix=ft)
    assert np.all(g0["blocks"]["m1tol"] == np.float32(m1)) and np.all(g0["blocks"]["tol"] == np.float32(tol))


def test_tolerance_profile_relative():
    """setupToleranceProfile (src/cochlea.cu:4256-4287) with relative errors:
    every block of profile i gets MAX 10^(fix * dBSPLSignal_i); the BM launch
    then takes them (calculate_boundary_conditions = 0)."""
    import ctypes as C
    from cochlearmodel_amd import _lib, solver
    from cochlearmodel_amd._lib import check, ptr
    _flush("case: setupToleranceProfile, 2 profiles x 2 blocks")
    L = _lib.lib()
    prof = np.zeros(2, _lib.PROFILE_DTYPE)
    prof["dBSPLSignal"] = [30.0, 70.0]
    m1, tol, fm, ft = 1e-15, 1e-10, 0.05, 0.025
    check(L.setupToleranceProfile(ptr(prof), 1, C.c_float(m1), C.c_float(tol), 1, C.c_float(fm), C.c_float(ft), 2,
                                  0, 2), "setupToleranceProfile")
    lay = solver.voice_layout(44100.0, 4, 0.01, 0.005)
    g = solver.run_bm_voice(np.zeros(lay.bufP1, np.float32), lay, boundary_conditions=0)
    db = np.array([30.0, 30.0, 70.0, 70.0])
    np.testing.assert_allclose(g["blocks"]["m1tol"], m1 * 10 ** (fm * db), rtol=2e-5)
    np.testing.assert_allclose(g["blocks"]["tol"], tol * 10 ** (ft * db), rtol=2e-5)
    assert np.all(g["speeds"] == 0)
    # more blocks launched than the profile covers is refused
    lay5 = solver.voice_layout(44100.0, 5, 0.01, 0.005)
    with pytest.raises(Exception, match="setupToleranceProfile covered"):
        solver.run_bm_voice(np.zeros(lay5.bufP1, np.float32), lay5, boundary_conditions=0)
