"""JND sweep in the Sim_Type 4 layout (profiles -> input -> BM -> IHC/lambda ->
JND -> final dB -> complex-JND aggregation) on the GPU against the CPU oracle
chain on the same generated input, and the per-interval oracle the benchmark
uses for its sampled check against the whole-sweep oracle."""
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


@pytest.fixture(scope="module")
def small_sweep():
    _cuda()
    from cochlearmodel_amd import solver
    from cochlearmodel_amd.jnd_sweep import JndSweep
    sw = JndSweep((1000.0, 3000.0), (20.0, 40.0, 60.0), (1111.0, 40.0), duration=0.04, methods=3)
    _say(f"case: JND sweep 2 freq x 3 levels x 2 noises, {sw.n_prof} intervals, {sw.TB} blocks")
    sw.step(solver.default_scalars(FS), timed_bm=False)
    return sw


def test_sweep_matches_oracle_chain(oracle, small_sweep):
    import torch
    from cochlearmodel_amd import stages
    sw = small_sweep
    res = sw.finish()
    inp = sw.d_input.cpu().numpy()
    st, mo, me, ws = oracle.block_decomposition(sw.TB, sw.bufP1, sw.ov, sw.DF)
    items = oracle.make_items(st, mo, me, ws, 1e-10, 1e-15)
    o = oracle.bm_run(oracle.default_scalars(FS), oracle.section_params(), np.full(256, 0.5, np.float32), items, inp,
                      sw.bufP1, "f32", threads=8, diagnostics=True)
    g = sw.d_speeds.cpu().numpy()
    np.testing.assert_array_equal(sw.d_failed_blocks.cpu().numpy(), o["failed_blocks"])
    np.testing.assert_array_equal(g[:sw.rows], o["speeds"][:sw.rows])
    b, _ = stages.load_ac_filter(FS)
    lo = oracle.ihc_lambda(o["speeds"][:sw.rows], sw.interval, sw.block * sw.TB, b, stages.dc_filter(FS),
                           stages.ihc_damage(8.0), stages.aihc_table(), np.array(stages.SPONT_DEFAULT, np.float32),
                           threads=8)
    np.testing.assert_array_equal(sw.d_lambda.cpu().numpy(), lo)
    _, fra, fai = oracle.jnd(lo, sw.n_prof, sw.interval, sw.head, sw.length, sw.profiles["dA"].astype(np.float32),
                             sw.ref_index, np.full(256, 30000 / 256), np.array(stages.W_DEFAULT, np.float32), FS,
                             0.2 * FS)
    calc = sw.calc
    gra, gai = sw.d_F_RA.cpu().numpy(), sw.d_FAI.cpu().numpy()
    rel_ra = np.max(np.abs(gra[:, calc] / fra[:, calc] - 1))
    rel_ai = np.max(np.abs(gai[:, calc] / fai[:, calc] - 1))
    _say(f"Fisher sums max rel diff: rate {rel_ra:.2e}, AI {rel_ai:.2e}")
    assert rel_ra <= 1e-4 and rel_ai <= 1e-4
    ai_o, rate_o = stages.jnd_final(fra.reshape(-1), fai.reshape(-1), sw.n_prof, sw.n_prof, sw.refs, sw.serial)
    n_calc = len(calc)
    d_rate = np.max(np.abs(res["rate_db"] - rate_o[:n_calc]))
    d_ai = np.max(np.abs(res["ai_db"] - ai_o[:n_calc]))
    _say(f"JND dB max diff: rate {d_rate:.2e}, AI {d_ai:.2e}; rate dB {np.round(res['rate_db'], 2)}")
    assert d_rate <= 0.01 and d_ai <= 0.01
    assert np.all(np.isfinite(res["rate_db"]))
    failed_o = stages.jnd_failed_signals(o["failed_blocks"], sw.n_prof, sw.refs, sw.serial, sw.DF, n_calc)
    np.testing.assert_array_equal(res["failed_signals"], failed_o)
    val_o, warn_o, _ = stages.complex_jnd(rate_o[:n_calc], 2, 3, 2, 3, 1.0, failed_o)
    val, warn, _ = res["complex"]
    _say(f"complex JND gpu {val.tolist()} oracle {val_o.tolist()}")
    np.testing.assert_allclose(val, val_o, atol=0.01)
    np.testing.assert_array_equal(warn, warn_o)
    # the aggregated value is one of the profile's own levels (or 1111)
    raw = res["rate_db"].astype(np.float32)
    for p in range(val.shape[1]):
        assert val[0, p] == np.float32(1111.0) or np.any(raw == val[0, p])


def test_bench_interval_oracle_matches_whole_sweep(oracle, small_sweep):
    """bench.py checks sampled intervals of the full sweep one at a time on the
    CPU; that re-based run must equal the whole-sweep oracle bit for bit."""
    import bench
    from cochlearmodel_amd import stages
    sw = small_sweep
    inp = sw.d_input.cpu().numpy()
    st, mo, me, ws = oracle.block_decomposition(sw.TB, sw.bufP1, sw.ov, sw.DF)
    items = oracle.make_items(st, mo, me, ws, 1e-10, 1e-15)
    o = oracle.bm_run(oracle.default_scalars(FS), oracle.section_params(), np.full(256, 0.5, np.float32), items, inp,
                      sw.bufP1, "f32", threads=8)
    b, _ = stages.load_ac_filter(FS)
    lo = oracle.ihc_lambda(o["speeds"][:sw.rows], sw.interval, sw.block * sw.TB, b, stages.dc_filter(FS),
                           stages.ihc_damage(8.0), stages.aihc_table(), np.array(stages.SPONT_DEFAULT, np.float32),
                           threads=8)
    for j in (0, 5, sw.n_prof - 1):
        speeds, lam, steps = bench.c4_oracle_interval(sw, j, 8)
        rs = sw.interval_rows(j)
        np.testing.assert_array_equal(speeds, o["speeds"][rs], err_msg=f"interval {j}")
        np.testing.assert_array_equal(lam, lo[:, rs], err_msg=f"lambda of interval {j}")
        assert steps > 0


def test_post_bm_graph_replays_match_eager(small_sweep):
    """The IHC/lambda + JND stages captured as one HIP graph: three replays,
    with eager work (fills of the outputs, a BM rerun) between them, each give
    the eager stages' lambda, mean rate and Fisher sums bit for bit."""
    torch = _cuda()
    from cochlearmodel_amd import solver
    sw = small_sweep
    sw.run_lambda()
    sw.run_jnd()
    torch.cuda.synchronize()
    ref = [x.cpu().numpy().copy() for x in (sw.d_lambda, sw.d_mean_rate, sw.d_F_RA, sw.d_FAI)]
    sw.build_graph()
    try:
        for rep in range(3):
            _say(f"graph replay {rep}")
            for x in (sw.d_lambda, sw.d_mean_rate, sw.d_F_RA, sw.d_FAI):
                x.zero_()
            if rep == 1:
                sw.run_bm(solver.default_scalars(FS), timed=False)
            torch.cuda.synchronize()
            sw.run_graph()
            torch.cuda.synchronize()
            got = [x.cpu().numpy() for x in (sw.d_lambda, sw.d_mean_rate, sw.d_F_RA, sw.d_FAI)]
            for g, r in zip(got, ref):
                np.testing.assert_array_equal(g, r)
        t = sw.step(solver.default_scalars(FS), timed_bm=False, use_graph=True)
        assert t["post_bm_graph_ms"] > 0
        np.testing.assert_array_equal(sw.d_F_RA.cpu().numpy(), ref[2])
    finally:
        sw.release_graph()
