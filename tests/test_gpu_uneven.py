"""Uneven batch (signals of different lengths, per-signal OHC and IHC) in one
launch against separate per-signal runs of the CPU oracle, and invariance of
the results under every launch configuration the autotuner can pick."""
import numpy as np
import pytest

FS = 44100.0
pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _cuda():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _signals(n=6, max_s=0.12):
    from cochlearmodel_amd import workloads
    sigs, ohc, ihc, meta = workloads.c3_signals(n, seed=workloads.MASTER_SEED + 33)
    sigs = [s[:int(max_s * FS)] for s in sigs]
    return sigs, ohc, ihc


def test_uneven_batch_matches_separate_runs(oracle):
    torch = _cuda()
    from cochlearmodel_amd import solver, stages
    from cochlearmodel_amd.uneven import UnevenBatch
    sigs, ohc, ihc = _signals()
    _say(f"case: {len(sigs)} signals, lengths {[len(s) for s in sigs]}, per-signal OHC/IHC")
    ub = UnevenBatch(sigs, ohc, ihc, FS)
    sc = solver.default_scalars(FS)
    tune = ub.autotune(sc, trial_stride=2)
    _say(f"autotune: {tune}")
    ub.run_bm(sc, timed=False)
    lam = ub.run_lambda()
    torch.cuda.synchronize()
    v = ub.d_speeds.cpu().numpy()
    lam = lam.cpu().numpy()
    ub.clear_order()
    so = oracle.default_scalars(FS)
    b, _ = stages.load_ac_filter(FS)
    for k in range(len(sigs)):
        lay, base = ub.layouts[k], ub.base[k]
        it = solver.block_table(lay.time_blocks, lay.bufP1, lay.overlap_nodes, 0)
        it["tol"], it["m1tol"] = 1e-10, 1e-15
        o = oracle.bm_run(so, oracle.section_params(), ohc[k], it, ub.input[base:base + lay.bufP1], lay.bufP1,
                          "f32", threads=8)
        np.testing.assert_array_equal(v[base:base + lay.bufP1], o["speeds"], err_msg=f"signal {k}")
        lo = oracle.ihc_lambda(o["speeds"][:lay.total_nodes], lay.total_nodes, lay.block_nodes * lay.time_blocks, b,
                               stages.dc_filter(FS), stages.ihc_damage(float(ihc[k])), stages.aihc_table(),
                               np.array(stages.SPONT_DEFAULT, np.float32), threads=8)
        np.testing.assert_array_equal(lam[:, base:base + lay.total_nodes], lo, err_msg=f"lambda of signal {k}")


def test_launch_configurations_give_identical_results():
    torch = _cuda()
    from cochlearmodel_amd import _lib, solver
    from cochlearmodel_amd.uneven import UnevenBatch
    sigs, ohc, ihc = _signals(n=9, max_s=0.06)
    ub = UnevenBatch(sigs, ohc, ihc, FS)
    _say(f"case: {ub.TB} blocks of {len(sigs)} uneven signals, sections per lane 4 (waves 1/2/4) / 2 / 1 x "
         "Jacobi unroll 1/2/4 x natural/longest-first order, fixed sweeps and both tolerance modes")
    ref = {}
    for spl in (4, 2, 1):
        _lib.check(_lib.lib().cm_set_bm_sections_per_lane(spl))
        for unroll in (2, 1, 4):
            _lib.check(_lib.lib().cm_set_bm_jacobi_unroll(unroll))
            for order in (False, True):
                if order:
                    lens = ub.items["mesi"] - ub.items["start"]
                    ordr = np.argsort(-lens, kind="stable").astype(np.int32)
                    _lib.check(_lib.lib().cm_set_bm_order(_lib.ptr(ordr), len(ordr)))
                else:
                    _lib.check(_lib.lib().cm_set_bm_order(None, 0))
                for wpb in ((1, 2, 4) if spl == 4 else (1,)):
                    for ws, tol in ((0, False), (0, True), (1, True)):
                        sc = solver.default_scalars(FS)
                        if tol:
                            sc.jacobi_mode, sc.tol_rel, sc.check_every, sc.max_sweeps, sc.warm_start = \
                                1, 1e-6, 5, 1000, ws
                        ub.d_speeds.zero_()
                        ub.run_bm(sc, timed=False, waves_per_block=wpb)
                        torch.cuda.synchronize()
                        out = (ub.d_speeds.cpu().numpy().copy(), ub.d_stats.cpu().numpy().copy())
                        key = (ws, tol)
                        what = f"spl {spl} unroll {unroll} wpb {wpb} order {order}"
                        if key not in ref:
                            ref[key] = out
                        else:
                            np.testing.assert_array_equal(out[0], ref[key][0], err_msg=what)
                            np.testing.assert_array_equal(out[1], ref[key][1], err_msg=what)
    ub.clear_order()
    with pytest.raises(_lib.CochleaError):
        _lib.check(_lib.lib().cm_set_bm_sections_per_lane(3))
    # unroll 4 runs the reference's 20 sweeps as five trips and 2 as one pair;
    # tolerance checks every 5 sweeps take one trip and a single sweep
    with pytest.raises(_lib.CochleaError):
        _lib.check(_lib.lib().cm_set_bm_jacobi_unroll(3))
    # a wrong-sized order table is refused, not run
    bad = np.arange(ub.TB - 1, dtype=np.int32)
    _lib.check(_lib.lib().cm_set_bm_order(_lib.ptr(bad), len(bad)))
    with pytest.raises(_lib.CochleaError):
        ub.run_bm(solver.default_scalars(FS), timed=False)
    ub.clear_order()
    dup = np.zeros(ub.TB, np.int32)
    with pytest.raises(_lib.CochleaError):
        _lib.check(_lib.lib().cm_set_bm_order(_lib.ptr(dup), len(dup)))
