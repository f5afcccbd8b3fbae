"""Every benchmark configuration at its full size on the GPU.

Each test runs the HIP path over the whole workload, checks properties that
hold for every block or interval (finite output, at least 22 committed steps
per emitted sample, mean dt within [dt_min, dt_max], failure flags), and
compares a strided sample of blocks, signals or intervals with the CPU oracle
bit for bit.  The sample keeps the oracle's share of the run to seconds.

  256-tone grid   256 tones x 200 ms, 2560 time blocks, BM + IHC/lambda
  uneven batch    1024 seeded signals 20-500 ms, 13,690 blocks, autotuned
  JND sweep       805 intervals x 300 ms, 12,075 blocks, BM + lambda + JND
  learned         4096 held-out tones: learned vs plain and extrapolated guesses
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
FS = 44100.0
SECTIONS = 256


def _say(msg):
    print(msg, flush=True)


def _cuda():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _threads():
    import os
    return min(16, os.cpu_count() or 1)


def _block_properties(stats, items, conv_time=None, rows=None, sc=None):
    """Committed steps per block cover >= 22 per integrated sample (dt <= dt_max
    against the 22 us sample clock); the mean dt of every emitted row lies in
    [dt_min, dt_max]."""
    samples = items["mesi"] - items["start"]
    assert np.all(stats[:, 0] >= 22 * samples), "fewer than 22 steps per sample"
    assert np.all(stats[:, 1] >= 2 * stats[:, 0]), "a committed step ran fewer than two passes"
    if conv_time is not None:
        ct = conv_time[rows]
        assert np.all(ct <= sc.dt_max * (1 + 1e-6)) and np.all(ct >= sc.dt_min)


def _oracle_items(oracle, items, sel, shift=0):
    it = np.zeros(len(sel), oracle.ITEM_DTYPE)
    for f in ("start", "mosi", "mesi", "write_start"):
        it[f] = items[f][sel] - shift
    for f in ("tol", "m1tol", "set"):
        it[f] = items[f][sel]
    return it


def test_tone_grid_full(oracle):
    """256 tones (32 frequencies 250-8000 Hz x 8 levels 20-90 dB), 200 ms,
    Decouple_Filter 10: 2560 blocks in one reference-method launch, then
    IHC/lambda over all 2.26 M rows."""
    torch = _cuda()
    from cochlearmodel_amd import batch, solver, stages, workloads
    f, lv = workloads.c2_grid()
    tb = batch.ToneBatch(f, lv, 0.2)
    _say(f"case: 256-tone grid, {tb.TB} blocks, {tb.rows} rows, reference method")
    assert tb.TB == 2560
    sc = solver.default_scalars(FS)
    tb.generate_input()
    out = tb.bm.run(sc, tb.d_input, tb.items, tb.bufP1, diagnostics=True, d_speeds=tb.d_speeds)
    tb.ihc.run(tb.d_speeds, tb.rows, tb.fir_ivs, d_lambda=tb.d_lambda, d_scratch=tb.d_scratch)
    torch.cuda.synchronize()
    stats = out["stats"].cpu().numpy()
    assert bool(torch.isfinite(tb.d_speeds).all()) and bool(torch.isfinite(tb.d_lambda).all())
    written = tb.block_rows(np.arange(tb.TB))
    _block_properties(stats, tb.items, out["conv_time"].cpu().numpy(), written, sc)
    failed = out["failed_blocks"].cpu().numpy()
    _say(f"committed steps {stats[:, 0].sum()}, failed blocks {int((failed > 0).sum())}")
    # sampled blocks against the oracle, bit for bit
    sel = tb.sample_blocks(16)
    o = oracle.bm_run(oracle.default_scalars(FS), oracle.section_params(), np.full(SECTIONS, 0.5, np.float32),
                      _oracle_items(oracle, tb.items, sel), tb.d_input.cpu().numpy(), tb.bufP1, "f32",
                      threads=_threads(), diagnostics=True)
    rows = tb.block_rows(sel)
    d_rows = torch.from_numpy(rows).cuda()
    np.testing.assert_array_equal(tb.d_speeds.index_select(0, d_rows).cpu().numpy(), o["speeds"][rows])
    np.testing.assert_array_equal(stats[sel, 0], o["steps"])
    np.testing.assert_array_equal(stats[sel, 2], o["sweeps"])
    np.testing.assert_array_equal(failed[sel], o["failed_blocks"])
    # lambda of two whole intervals (first and last tone) from the GPU's BM rows
    b, _ = stages.load_ac_filter(FS)
    for j in (0, tb.n - 1):
        r = slice(j * tb.interval, (j + 1) * tb.interval)
        lo = oracle.ihc_lambda(tb.d_speeds[r].cpu().numpy(), tb.interval, tb.block * tb.TB, b, stages.dc_filter(FS),
                               stages.ihc_damage(8.0), stages.aihc_table(),
                               np.array(stages.SPONT_DEFAULT, np.float32), threads=_threads())
        np.testing.assert_array_equal(tb.d_lambda[:, r].cpu().numpy(), lo, err_msg=f"lambda of tone {j}")
    del tb, out
    torch.cuda.empty_cache()


def test_uneven_batch_full(oracle):
    """1024 seeded uneven signals in one autotuned launch; eight signals spread
    over the batch against their own separate oracle runs (BM and lambda)."""
    torch = _cuda()
    from cochlearmodel_amd import solver, stages, workloads
    from cochlearmodel_amd.uneven import UnevenBatch
    sigs, ohc, ihc, meta = workloads.c3_signals(1024)
    ub = UnevenBatch(sigs, ohc, ihc, FS)
    _say(f"case: uneven batch, {ub.n} signals, {ub.TB} blocks, {ub.rows} rows")
    assert ub.n == 1024 and ub.TB > 13000
    sc = solver.default_scalars(FS)
    tune = ub.autotune(sc, trial_stride=8)
    _say(f"autotune: {tune}")
    ub.run_bm(sc, timed=False)
    lam = ub.run_lambda()
    torch.cuda.synchronize()
    stats = ub.d_stats.view(ub.TB, 3).cpu().numpy()
    assert bool(torch.isfinite(ub.d_speeds).all()) and bool(torch.isfinite(lam).all())
    _block_properties(stats, ub.items)
    b, _ = stages.load_ac_filter(FS)
    so = oracle.default_scalars(FS)
    for k in np.linspace(0, ub.n - 1, 8).round().astype(int):
        lay, base = ub.layouts[k], ub.base[k]
        it = solver.block_table(lay.time_blocks, lay.bufP1, lay.overlap_nodes, 0)
        it["tol"], it["m1tol"] = 1e-10, 1e-15
        o = oracle.bm_run(so, oracle.section_params(), ohc[k], it, ub.input[base:base + lay.bufP1], lay.bufP1, "f32",
                          threads=_threads())
        np.testing.assert_array_equal(ub.d_speeds[base:base + lay.bufP1].cpu().numpy(), o["speeds"],
                                      err_msg=f"signal {k} ({meta[k]})")
        lo = oracle.ihc_lambda(o["speeds"][:lay.total_nodes], lay.total_nodes, lay.block_nodes * lay.time_blocks, b,
                               stages.dc_filter(FS), stages.ihc_damage(float(ihc[k])), stages.aihc_table(),
                               np.array(stages.SPONT_DEFAULT, np.float32), threads=_threads())
        np.testing.assert_array_equal(lam[:, base:base + lay.total_nodes].cpu().numpy(), lo,
                                      err_msg=f"lambda of signal {k}")
    ub.clear_order()
    del ub, lam
    torch.cuda.empty_cache()


def test_jnd_sweep_full(oracle):
    """16 tones x 10 levels x 5 noise conditions, 300 ms intervals: 805
    intervals, 12,075 blocks.  Every calculated JND is finite; four calculated
    intervals and their references against the oracle chain (BM and lambda
    bit for bit, JND within 0.01 dB)."""
    torch = _cuda()
    import bench
    from cochlearmodel_amd import solver, stages
    sw = bench.c4_sweep()
    _say(f"case: JND sweep, {sw.n_prof} intervals, {sw.TB} blocks, {sw.rows} rows")
    assert sw.n_prof == 805 and sw.TB == 12075
    sw.step(solver.default_scalars(FS), timed_bm=False)
    res = sw.finish()
    torch.cuda.synchronize()
    stats = sw.d_stats.view(sw.TB, 3).cpu().numpy()
    assert bool(torch.isfinite(sw.d_lambda).all())
    assert np.all(np.isfinite(res["ai_db"])) and np.all(np.isfinite(res["rate_db"]))
    _block_properties(stats, sw.items)
    n_calc = len(sw.calc)
    pick = sw.calc[np.linspace(0, n_calc - 1, 4).round().astype(int)]
    need = sorted(set(int(j) for j in pick) | set(int(sw.ref_index[j]) for j in pick))
    cpu = {j: bench.c4_oracle_interval(sw, j, _threads()) for j in need}
    for j in need:
        rs = sw.interval_rows(j)
        np.testing.assert_array_equal(sw.d_speeds[rs].cpu().numpy(), cpu[j][0], err_msg=f"BM of interval {j}")
        np.testing.assert_array_equal(sw.d_lambda[:, rs].cpu().numpy(), cpu[j][1], err_msg=f"lambda of interval {j}")
    nIHC = np.full(SECTIONS, 30000.0 / SECTIONS)
    W = np.array(stages.W_DEFAULT, np.float32)
    for j in pick:
        r = int(sw.ref_index[j])
        lam2 = np.concatenate([cpu[int(j)][1], cpu[r][1]], axis=1)
        dA = np.array([sw.profiles["dA"][j], sw.profiles["dA"][r]], np.float32)
        _, fra, fai = oracle.jnd(lam2, 2, sw.interval, sw.head, sw.length, dA, np.array([1, 1], np.int32), nIHC, W,
                                 FS, 0.2 * FS)
        ai, rate = stages.jnd_final(fra.reshape(-1), fai.reshape(-1), 2, 2, np.array([1], np.int32),
                                    np.array([0, 0], np.int32))
        s = int(sw.serial[j])
        _say(f"interval {int(j)}: AI {res['ai_db'][s]:.4f} vs {ai[0]:.4f} dB, rate {res['rate_db'][s]:.4f} vs "
             f"{rate[0]:.4f} dB")
        assert abs(res["ai_db"][s] - ai[0]) <= 0.01 and abs(res["rate_db"][s] - rate[0]) <= 0.01
    del sw
    torch.cuda.empty_cache()


def test_learned_warm_start_held_out_4096(oracle):
    """The committed network on the 4096 held-out tones (16 batches of 256,
    drawn like the 256-tone grid, seeds disjoint from training): fewer Jacobi
    sweeps per step than the plain and the extrapolated guesses, the same
    committed steps within 1 %, and on sampled blocks of the first batch the
    same error against the exact per-pass solve as plain's (within 25 %)."""
    torch = _cuda()
    from cochlearmodel_amd import batch, warmstart
    params, c_out, _ = warmstart.load_model()
    warmstart.upload(params, c_out)
    _say("case: learned warm start on 4096 held-out tones (16 x 256), tau 1e-6")
    tot = {k: np.zeros(3, np.int64) for k in ("plain", "extrapolated", "learned")}
    for bi in range(16):
        f, db = batch.random_tones(256, warmstart.EVAL_SEED + bi)
        tb = batch.ToneBatch(f, db, 0.2, with_lambda=False)
        tb.generate_input()
        keep = {}
        for name, ws in (("plain", 0), ("extrapolated", 1), ("learned", 2)):
            tb.run_bm(warmstart.tol_scalars(1e-6, ws), timed=False)
            torch.cuda.synchronize()
            tot[name] += np.array(tb.stats(), np.int64)
            if bi == 0:
                keep[name] = tb.d_speeds.cpu().numpy()
                assert np.all(np.isfinite(keep[name]))
        if bi == 0:
            sel = tb.sample_blocks(8)
            so = oracle.default_scalars(FS)
            so.jacobi_mode = 2
            o = oracle.bm_run(so, oracle.section_params(), np.full(SECTIONS, 0.5, np.float32),
                              _oracle_items(oracle, tb.items, sel), tb.d_input.cpu().numpy(), tb.bufP1, "f32",
                              threads=_threads())
            rows = tb.block_rows(sel)
            ex = o["speeds"][rows]
            err = {k: float(np.linalg.norm(v[rows] - ex) / np.linalg.norm(ex)) for k, v in keep.items()}
            _say(f"rel L2 against the exact per-pass solve (8 blocks): {err}")
            # all guesses iterate to the same residual tolerance, so their errors
            # sit at the same level (2.4-3.6e-4 measured on these blocks); the
            # guess changes the sweep count, not the accuracy
            assert err["learned"] <= 1.25 * err["plain"] and max(err.values()) <= 5e-4
        del tb
        torch.cuda.empty_cache()
    sps = {k: v[2] / v[0] for k, v in tot.items()}
    _say(f"sweeps/step over 4096 tones: {sps}; committed steps {[int(v[0]) for v in tot.values()]}")
    assert sps["learned"] < sps["extrapolated"] < sps["plain"]
    assert sps["learned"] < 0.5 * sps["plain"]
    assert abs(int(tot["learned"][0]) - int(tot["plain"][0])) <= 0.01 * tot["plain"][0]
