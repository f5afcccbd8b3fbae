"""GPU parity of the IHC/lambda, JND and input-generation stages and of the
batched BM launches against the CPU oracle, all through the C-ABI."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
FS = 44100.0


def _say(msg):
    print(msg, flush=True)


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    n = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (n if n > 0 else 1.0)


@pytest.fixture(scope="module")
def c1_bm():
    from cochlearmodel_amd import solver, workloads
    lay = solver.voice_layout(FS, 5, 0.02, 0.005)
    g = solver.run_bm_voice(workloads.c1_signal(), lay)
    return lay, g


def test_ihc_lambda_reference_sequence_fir(oracle, c1_bm):
    from cochlearmodel_amd import stages
    _say("case: IHC/lambda FIR on C1 BM, reference call sequence")
    lay, g = c1_bm
    rows = lay.total_nodes
    lam, ihc = stages.run_ihc_reference(rows, lay.block_nodes, lay.time_blocks, 0, backup_stage=13)
    b, _ = stages.load_ac_filter(FS)
    lo, st = oracle.ihc_lambda(g["speeds"][:rows], rows, lay.block_nodes * lay.time_blocks, b, stages.dc_filter(FS),
                               stages.ihc_damage(8.0), stages.aihc_table(),
                               np.array(stages.SPONT_DEFAULT, np.float32), stages=True, threads=8)
    _say(f"lambda rel L2 {_rel(lam, lo):.3e} exact {np.mean(lam == lo):.6f}; IHC exact {np.mean(ihc == st[4]):.6f}")
    assert _rel(lam, lo) <= 1e-4
    assert _rel(ihc, st[4]) <= 1e-4
    assert np.all(lam[:, :, 0] == np.array(stages.SPONT_DEFAULT, np.float32)[:, None])


def test_ihc_lambda_iir_decoupled(oracle, c1_bm):
    from scipy.signal import butter
    from cochlearmodel_amd import stages
    _say("case: IHC/lambda IIR (Butterworth 4) with Decouple_Filter=1")
    lay, g = c1_bm
    rows = lay.total_nodes
    b, a = butter(4, 1000 / 22050)
    lam, _ = stages.run_ihc_reference(rows, lay.block_nodes, lay.time_blocks, 1, ac=(b, a))
    lo = oracle.ihc_lambda(g["speeds"][:rows], lay.block_nodes, lay.block_nodes * lay.time_blocks, b,
                           stages.dc_filter(FS), stages.ihc_damage(8.0), stages.aihc_table(),
                           np.array(stages.SPONT_DEFAULT, np.float32), ac_a=a, threads=8)
    _say(f"lambda rel L2 {_rel(lam, lo):.3e} exact {np.mean(lam == lo):.6f}")
    assert _rel(lam, lo) <= 1e-4


def test_ihc_batched_signal_intervals(oracle):
    import torch
    from cochlearmodel_amd import stages
    _say("case: batched IHC over 3 uneven signals with per-signal IHC vectors")
    rng = np.random.default_rng(11)
    lens = [700, 1234, 401]
    starts = np.cumsum([0] + lens[:-1])
    rows = sum(lens)
    bm = (1e-3 * rng.standard_normal((rows, 256))).astype(np.float32)
    ihcv = [8.0, 6.5, 5.0]
    bi = stages.BatchedIHC(FS, ihc_vectors=ihcv)
    ivs = stages.signal_intervals(starts, lens, sets=[0, 1, 2])
    d_lam, _ = bi.run(torch.from_numpy(bm).cuda(), rows, ivs)
    lam = d_lam.cpu().numpy()
    b, _ = stages.load_ac_filter(FS)
    for k, (s0, n) in enumerate(zip(starts, lens)):
        lo = oracle.ihc_lambda(bm[s0:s0 + n], n, n + 1, b, stages.dc_filter(FS), stages.ihc_damage(ihcv[k]),
                               stages.aihc_table(), np.array(stages.SPONT_DEFAULT, np.float32), threads=8)
        np.testing.assert_array_equal(lam[:, s0:s0 + n], lo, err_msg=f"signal {k}")


def test_ihc_one_pass_and_two_pass_paths_agree(oracle):
    """cm_ihc_lambda_run runs AC -> dS -> DC -> lambda in two passes through
    the caller's AC/dS workspace, or in one pass with no workspace.  Both give
    the same lambda and stages, equal to the oracle, bit for bit; a table whose
    second entry continues the first one's windows (kbase > 0) gives the same
    result through the workspace and is refused without one."""
    import torch
    from cochlearmodel_amd import _lib, stages
    _say("case: IHC/lambda one pass vs two passes over 3000 rows (one and two interval entries), stages kept")
    rng = np.random.default_rng(21)
    rows = 3000
    bm = (1e-3 * rng.standard_normal((rows, 256))).astype(np.float32)
    d_bm = torch.from_numpy(bm).cuda()
    bi = stages.BatchedIHC(FS)
    one = stages.signal_intervals([0], [rows])
    two = np.concatenate([stages.signal_intervals([0], [1100]), stages.signal_intervals([1100], [rows - 1100])])
    two["kbase"] = [0, 1100]
    two["L"], two["tla"], two["frdp"] = rows, rows, rows + 1
    runs = [bi.run(d_bm, rows, one, stages=True, workspace=ws) for ws in (True, False)]
    runs.append(bi.run(d_bm, rows, two, stages=True))
    res = [(lam.cpu().numpy(), st.cpu().numpy()) for lam, st in runs]
    b, _ = stages.load_ac_filter(FS)
    lo, so = oracle.ihc_lambda(bm, rows, rows + 1, b, stages.dc_filter(FS), stages.ihc_damage(8.0),
                               stages.aihc_table(), np.array(stages.SPONT_DEFAULT, np.float32), stages=True, threads=8)
    for k, (lam, st) in enumerate(res):
        np.testing.assert_array_equal(lam, lo, err_msg=f"run {k}")
        np.testing.assert_array_equal(st, so, err_msg=f"run {k}")
    with pytest.raises(_lib.CochleaError, match="kbase"):
        bi.run(d_bm, rows, two, workspace=False)


def test_bm_batched_sets_and_launch_shapes(oracle):
    import torch
    from cochlearmodel_amd import solver
    _say("case: batched BM, 6 blocks over 3 signals, 2 gamma sets, waves/block 1,2,4")
    rng = np.random.default_rng(12)
    lens = [900, 1500, 600]
    starts = np.cumsum([0] + lens[:-1])
    rows = sum(lens)
    inp = np.zeros(rows + 2, np.float32)
    for s0, n in zip(starts, lens):
        t = np.arange(n) / FS
        inp[s0:s0 + n] = 10 * 2e-5 * 10 ** (rng.uniform(30, 80) / 20) * np.sin(2 * np.pi * rng.uniform(500, 4000) * t)
    items = np.zeros(3, solver.ITEM_DTYPE)
    for k, (s0, n) in enumerate(zip(starts, lens)):
        items[k] = (s0, s0, s0 + n, s0, 1e-10, 1e-15, k % 2, 0)
    gam = np.stack([np.full(256, 0.5, np.float32), np.linspace(0, 0.5, 256).astype(np.float32)])
    sc = solver.default_scalars(FS)
    bb = solver.BatchedBM(gam)
    d_in = torch.from_numpy(inp).cuda()
    outs = [bb.run(sc, d_in, items, rows, waves_per_block=w, diagnostics=True) for w in (1, 2, 4)]
    v = [o["speeds"].cpu().numpy() for o in outs]
    assert np.array_equal(v[0], v[1]) and np.array_equal(v[0], v[2])
    it_o = np.zeros(3, oracle.ITEM_DTYPE)
    for f in ("start", "mosi", "mesi", "write_start", "tol", "m1tol", "set"):
        it_o[f] = items[f]
    o = oracle.bm_run(oracle.default_scalars(FS), oracle.section_params(), gam, it_o, inp, rows, "f32", threads=8,
                      diagnostics=True)
    _say(f"rel L2 {_rel(v[0], o['speeds']):.3e} exact {np.mean(v[0] == o['speeds']):.6f}")
    assert _rel(v[0], o["speeds"]) <= 1e-4
    st = outs[0]["stats"].cpu().numpy()
    np.testing.assert_array_equal(st[:, 0], o["steps"])
    np.testing.assert_array_equal(outs[0]["failed_nodes"].cpu().numpy(), o["failed_nodes"])
    np.testing.assert_array_equal(outs[0]["conv_loops"].cpu().numpy(), o["conv_loops"])


def test_bm_persistent_launch_is_refused():
    import torch
    from cochlearmodel_amd import solver
    from cochlearmodel_amd._lib import CochleaError
    items = np.zeros(1, solver.ITEM_DTYPE)
    items[0] = (0, 0, 10, 0, 1e-10, 1e-15, 0, 0)
    bb = solver.BatchedBM(np.full(256, 0.5, np.float32))
    with pytest.raises(CochleaError):
        bb.run(solver.default_scalars(FS), torch.zeros(16, device="cuda"), items, 12, persistent=1)


def _c4_small(noises=(1111.0, 40.0), freqs=(1000.0, 3000.0), levels=(30.0, 60.0), D=0.04, Tb=0.02, seed=7):
    """A small JND sweep laid out as the reference's Sim_Type=4 run."""
    from cochlearmodel_amd import stages
    prof, refs, serial, calc, to_ref = stages.jnd_profiles(list(freqs), list(levels), noises=noises)
    L = int(round(FS * D))
    block = int(math.ceil(np.float32(Tb) * np.float32(FS)))
    DF = L // block
    assert DF * block == L
    n_prof = len(prof)
    TB = n_prof * DF
    ov = 220
    bufP1 = ov + block * (TB + 1)
    head = int(round(FS * 0.014))
    length = int(round(FS * (D - 0.014 - 0.002)))
    WN = np.random.default_rng(seed).standard_normal(L).astype(np.float32)
    e1 = int((D - 0.002) * FS)
    return dict(prof=prof, refs=refs, serial=serial, calc=calc, to_ref=to_ref, L=L, block=block, DF=DF, TB=TB, ov=ov,
                bufP1=bufP1, head=head, length=length, WN=WN, e1=e1, n_prof=n_prof)


def _numpy_norms(c):
    WN = c["WN"].astype(np.float64)
    e1 = c["e1"]
    norm = WN[0] * WN[0] + np.sum(WN[1:e1] ** 2)      # sigma accumulation with dc = 0
    return 0.0, math.sqrt(norm), math.sqrt(0.5 * e1)


def test_input_generation_parity(oracle):
    from cochlearmodel_amd import _lib, stages
    from cochlearmodel_amd._lib import check, ptr
    _say("case: generateInputFromProfile vs oracle")
    c = _c4_small()
    L = _lib.lib()
    check(L.InputProfilesArrayInitialization(c["n_prof"], c["L"], 0, 0, 0), "init")
    from cochlearmodel_amd import solver
    sec = solver.section_params()
    g = np.full(256, 0.5, np.float32)
    keep = [np.ascontiguousarray(x, np.float32) for x in (g, sec[0], sec[1], g, g, sec[2], sec[3], sec[4], sec[5],
                                                          sec[6], sec[7])]
    check(L.BMOHCKernel_Init(*[ptr(k) for k in keep], 1.0, 0.0, c["bufP1"] + 10, (c["bufP1"] + 10) * 256,
                             0, 1, 0, 0), "init")
    out = np.zeros(c["n_prof"] * c["L"], np.float32)
    check(L.generateInputFromProfile(ptr(c["prof"]), ptr(c["WN"]), c["L"], None, 0, 0, 0, 0, 0.16, 0.0, 0, c["e1"],
                                     0, c["e1"], c["n_prof"], 0, 0, c["n_prof"], c["ov"], c["L"], c["head"],
                                     c["length"], FS, 1, ptr(out), 0, 1, 1e-15, 1e-10, 0, 0.0, 0.0, c["DF"], 0,
                                     ptr(np.ones(1, np.float32)), 1, ptr(np.ones(1, np.float32)), 1), "generate")
    wn_dc, wn_norm, sig_norm = _numpy_norms(c)
    ref = oracle.generate_input(c["prof"]["dA"], c["prof"]["Wn"], c["prof"]["frequency"], c["WN"], wn_dc, wn_norm,
                                sig_norm, c["L"], c["head"], FS)
    _say(f"input rel L2 {_rel(out, ref):.3e} exact {np.mean(out == ref):.6f}")
    assert _rel(out, ref) <= 1e-6


def test_jnd_chain_end_to_end(oracle):
    """Profiles -> input -> BM (decoupled blocks) -> IHC/lambda -> JND -> dB, all on
    the GPU through the reference-named C-ABI, against the oracle chain."""
    from cochlearmodel_amd import _lib, solver, stages
    from cochlearmodel_amd._lib import check, ptr
    _say("case: small JND sweep end to end (2 freq x 2 levels x 2 noises + refs)")
    c = _c4_small()
    L = _lib.lib()
    sec = solver.section_params()
    g = np.full(256, 0.5, np.float32)
    keep = [np.ascontiguousarray(x, np.float32) for x in (g, sec[0], sec[1], g, g, sec[2], sec[3], sec[4], sec[5],
                                                          sec[6], sec[7])]
    n_in = c["bufP1"] + 2 * c["ov"]
    check(L.BMOHCKernel_Init(*[ptr(k) for k in keep], 1.0, 0.0, n_in, n_in * 256, 0, 1, 0, 0), "init")
    check(L.InputProfilesArrayInitialization(c["n_prof"], c["L"], 0, 0, 0), "init")
    check(L.generateInputFromProfile(ptr(c["prof"]), ptr(c["WN"]), c["L"], None, 0, 0, 0, 0, 0.16, 0.0, 0, c["e1"],
                                     0, c["e1"], c["n_prof"], 0, 0, c["n_prof"], c["ov"], c["L"], c["head"],
                                     c["length"], FS, 0, None, 0, 1, 1e-15, 1e-10, 0, 0.0, 0.0, c["DF"], 0,
                                     ptr(np.ones(1, np.float32)), 1, ptr(np.ones(1, np.float32)), 1), "generate")
    sc = solver.default_scalars(FS)
    rc = L.BMOHCNewKernel(None, 1, sc.w_ohc, sc.time_step, sc.time_step_out, sc.delta_x, sc.alpha_r, 1, 1, 0, sc.Ts,
                          sc.alpha_l, sc.alpha_s, sc.Gme, sc.a0, sc.a1, sc.a2, sc.sigma_ow, sc.eta1, sc.eta2, c["TB"],
                          c["bufP1"], c["ov"], 5000, 1, sc.dt_max, sc.dt_min, c["DF"], 1e-15, 1e-10, 0, 0, 0.0, 0.0,
                          2e-5, 0, 0, 0, 2, 20, 10, 17, 2)
    check(rc, "BMOHCNewKernel")
    rows = c["n_prof"] * c["L"]
    lam, _ = stages.run_ihc_reference(rows, c["block"], c["TB"], c["DF"])
    nIHC = np.full(256, 30000 / 256, np.float64)
    n_full = c["n_prof"]
    F_RA = np.zeros(3 * n_full, np.float32)
    FAI = np.zeros(3 * n_full, np.float32)
    to_ref = c["to_ref"].astype(np.int32)
    serial = c["serial"].astype(np.int32)
    refs = c["refs"].astype(np.int32)
    calc = c["calc"].astype(np.int32)
    check(L.CudaCalculateJND(1, 1, 0, 3 * n_full, 2e-5, FS, 1.0, ptr(nIHC), ptr(calc), len(calc), ptr(refs),
                             len(refs), 0, n_full, n_full, c["head"], c["ov"], c["L"], c["length"], ptr(serial),
                             ptr(to_ref), ptr(F_RA), ptr(FAI), rows * 256, rows * 256, 0.2 * FS, 1, 0, 0, 0, 0),
          "CudaCalculateJND")
    ai, rate = stages.jnd_final(F_RA, FAI, n_full, n_full, refs, serial)
    # oracle chain
    wn_dc, wn_norm, sig_norm = _numpy_norms(c)
    inp = np.zeros(c["bufP1"] + 2, np.float32)
    inp[:rows] = oracle.generate_input(c["prof"]["dA"], c["prof"]["Wn"], c["prof"]["frequency"], c["WN"], wn_dc,
                                       wn_norm, sig_norm, c["L"], c["head"], FS)
    st, mo, me, ws = oracle.block_decomposition(c["TB"], c["bufP1"], c["ov"], c["DF"])
    items = oracle.make_items(st, mo, me, ws, 1e-10, 1e-15)
    o = oracle.bm_run(oracle.default_scalars(FS), oracle.section_params(), g, items, inp, c["bufP1"], "f32", threads=8)
    b, _ = stages.load_ac_filter(FS)
    lo = oracle.ihc_lambda(o["speeds"][:rows], c["L"], c["block"] * c["TB"], b, stages.dc_filter(FS),
                           stages.ihc_damage(8.0), stages.aihc_table(), np.array(stages.SPONT_DEFAULT, np.float32),
                           threads=8)
    _say(f"lambda rel L2 {_rel(lam, lo):.3e}")
    assert _rel(lam, lo) <= 1e-4
    ref_idx = stages.reference_index(n_full, refs, serial, to_ref)
    dA = c["prof"]["dA"].astype(np.float32)
    mr, fra, fai = oracle.jnd(lo, n_full, c["L"], c["head"], c["length"], dA, ref_idx, nIHC,
                              np.array(stages.W_DEFAULT, np.float32), FS, 0.2 * FS)
    ai_o, rate_o = stages.jnd_final(fra.reshape(-1), fai.reshape(-1), n_full, n_full, refs, serial)
    # per-interval mean rates (getCudaMeanRate): FP64 time sums in order on both sides
    mr_g = np.zeros(3 * n_full * 256, np.float32)
    check(L.GeneralKernel_Copy_Results_Bytes(ptr(mr_g), L.getCudaMeanRate(), mr_g.nbytes, 0), "mean rate copy")
    mr_g = mr_g.reshape(3, n_full, 256)
    _say(f"mean rate rel L2 {_rel(mr_g, mr):.3e} exact {np.mean(mr_g == mr):.6f}")
    np.testing.assert_array_equal(mr_g, mr)
    sel = calc
    _say(f"JND AI dB gpu {np.round(ai[serial[sel]], 3)} oracle {np.round(ai_o[serial[sel]], 3)}")
    assert np.max(np.abs(ai[serial[sel]] - ai_o[serial[sel]])) <= 0.1
    assert np.max(np.abs(rate[serial[sel]] - rate_o[serial[sel]])) <= 0.1
    assert np.all(np.isfinite(ai[serial[sel]]))
    # backup stages 1..6 of CudaCalculateJND (TEST_File_Target) land in getCudaBuffer()
    W3 = np.array(stages.W_DEFAULT, np.float32)
    body = np.concatenate([np.arange(i * c["L"] + c["head"], i * c["L"] + c["head"] + c["length"])
                           for i in range(n_full)])
    for k in range(1, 7):
        check(L.CudaCalculateJND(1, 1, 0, 3 * n_full, 2e-5, FS, 1.0, ptr(nIHC), ptr(calc), len(calc), ptr(refs),
                                 len(refs), 0, n_full, n_full, c["head"], c["ov"], c["L"], c["length"], ptr(serial),
                                 ptr(to_ref), ptr(F_RA), ptr(FAI), rows * 256, rows * 256, 0.2 * FS, 1, 0, 0, 0, k),
              f"CudaCalculateJND stage {k}")
        exp = oracle.jnd(lo, n_full, c["L"], c["head"], c["length"], dA, ref_idx, nIHC, W3, FS, 0.2 * FS,
                         backup_stage=k)[3]
        got = np.zeros(exp.size, np.float32)
        check(L.GeneralKernel_Copy_Results_Bytes(ptr(got), L.getCudaBuffer(), got.nbytes, 0), "stage copy")
        got = got.reshape(exp.shape)
        if k in (3, 4):
            got, exp = got[:, body], exp[:, body]
        fin = np.isfinite(exp)
        _say(f"JND stage {k}: {exp.size} values, exact {np.mean(got[fin] == exp[fin]):.6f}")
        np.testing.assert_array_equal(got, exp)


def test_cuda_calculate_jnd_refuses_bad_layouts():
    """CudaCalculateJND checks its layout against the buffers before any launch:
    lambda planes beyond what RunIHCKernel sized, interval bodies past their
    interval, columns past JNDIntervalsFull, serial positions outside the
    calculated-interval table and dA reads past the BM input buffer are refused
    with -3 and a message."""
    from cochlearmodel_amd import _lib, solver, stages
    from cochlearmodel_amd._lib import check, ptr
    _say("case: CudaCalculateJND argument refusals (100-row lambda buffer)")
    L = _lib.lib()
    # fresh holders: the grow-only buffers of earlier tests must not hide a refusal
    check(L.IHCKernel_Free(), "IHCKernel_Free")
    check(L.BMOHCKernel_Free(), "BMOHCKernel_Free")
    sec = solver.section_params()
    g = np.full(256, 0.5, np.float32)
    keep = [np.ascontiguousarray(x, np.float32) for x in (g, sec[0], sec[1], g, g, sec[2], sec[3], sec[4], sec[5],
                                                          sec[6], sec[7])]
    check(L.BMOHCKernel_Init(*[ptr(k) for k in keep], 1.0, 0.0, 400, 400 * 256, 0, 1, 0, 0), "init")
    b, _ = stages.load_ac_filter(FS)
    dc = stages.dc_filter(FS)
    dmg = stages.ihc_damage(8.0)
    nc = np.array([60, 3, 0.1, 0, 0, 0, 0.61, 0.23, 0.16], np.float32)
    check(L.IHCNewKernel(ptr(dmg), ptr(nc), ptr(dc), len(dc), ptr(b), None, 0, len(b), 256, 2, 2e-5, None, 0, 100, 100,
                         100, 50, 0, 0, 500.0, 1.0, 100.0, 1, 1, 0, 0, 0, 0, 1.0, 1, 1, 0, 1, 0), "IHCNewKernel")
    nIHC = np.full(256, 30000 / 256, np.float64)
    F = np.zeros(3 * 4, np.float32)
    refs = np.array([1], np.int32)
    calc = np.array([0], np.int32)

    def call(matrix_rows=100, n_iv=2, full=2, handled=0, head=10, length=30, serial=(0, 0), calcdA=0, overlap=0):
        ser = np.array(serial, np.int32)
        to_ref = np.array([1], np.int32)
        return L.CudaCalculateJND(1, 1, 0, 3 * full, 2e-5, FS, 1.0, ptr(nIHC), ptr(calc), 1, ptr(refs), 1, handled,
                                  full, n_iv, head, overlap, 50, length, ptr(ser), ptr(to_ref), ptr(F), ptr(F),
                                  matrix_rows * 256, matrix_rows * 256, 0.2 * FS, 1, 0, calcdA, 0, 0)

    def msg():
        return L.cm_last_error().decode()

    assert call(matrix_rows=200) == -3 and "lambda buffer" in msg()
    assert call(head=10, length=45) == -3 and "outside an interval" in msg()
    assert call(head=-1) == -3 and "outside an interval" in msg()
    assert call(handled=1) == -3 and "JNDIntervalsFull" in msg()
    assert call(serial=(5, 0)) == -3 and "serial position" in msg()
    assert call(calcdA=1, overlap=1000) == -3 and "dA reads past" in msg()
    check(L.IHCKernel_Free(), "IHCKernel_Free")
    check(L.BMOHCKernel_Free(), "BMOHCKernel_Free")


@pytest.mark.parametrize("tau,ws", [(1e-4, 0), (1e-6, 0), (1e-6, 1)])
def test_bm_tolerance_mode_parity(oracle, tau, ws):
    """Sweep-to-tolerance Jacobi: same residual test and reduction tree on both sides."""
    import torch
    from cochlearmodel_amd import solver
    _say(f"case: tolerance-mode Jacobi tau={tau} warm_start={ws}, 2 signals of 600 samples")
    n = 600
    t = np.arange(n) / FS
    inp = np.zeros(2 * n + 2, np.float32)
    inp[:n] = 10 * 2e-5 * 10 ** (3.0) * np.sin(2 * np.pi * 1000 * t)
    inp[n:2 * n] = 10 * 2e-5 * 10 ** (1.5) * np.sin(2 * np.pi * 3000 * t)
    items = np.zeros(2, solver.ITEM_DTYPE)
    items[0] = (0, 0, n, 0, 1e-10, 1e-15, 0, 0)
    items[1] = (n, n, 2 * n, n, 1e-10, 1e-15, 0, 0)
    sc = solver.default_scalars(FS)
    sc.jacobi_mode, sc.tol_rel, sc.check_every, sc.max_sweeps, sc.warm_start = 1, tau, 4, 1000, ws
    bb = solver.BatchedBM(np.full(256, 0.5, np.float32))
    out = bb.run(sc, torch.from_numpy(inp).cuda(), items, 2 * n)
    v = out["speeds"].cpu().numpy()
    st = out["stats"].cpu().numpy()
    so = oracle.default_scalars(FS)
    so.jacobi_mode, so.tol_rel, so.check_every, so.max_sweeps, so.warm_start = 1, tau, 4, 1000, ws
    it = np.zeros(2, oracle.ITEM_DTYPE)
    for f in ("start", "mosi", "mesi", "write_start", "tol", "m1tol", "set"):
        it[f] = items[f]
    o = oracle.bm_run(so, oracle.section_params(), np.full(256, 0.5, np.float32), it, inp, 2 * n, "f32", threads=2)
    _say(f"rel L2 {_rel(v, o['speeds']):.3e}, sweeps gpu {st[:, 2]} oracle {o['sweeps']}")
    np.testing.assert_array_equal(v, o["speeds"])
    np.testing.assert_array_equal(st[:, 2], o["sweeps"])


def test_lambda_fix_up_kernel():
    """cm_fix_lambda_run (CUDAFIXJND_Lambda): lambda -> max(0, lambda - spont_k),
    buffer -> max(spont_k, lambda), per fibre group, in float32."""
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from cochlearmodel_amd import _lib, stages
    _say("case: lambda fix-up over 3 x 1001 x 256")
    rng = np.random.default_rng(3)
    lam = rng.uniform(-5.0, 200.0, (3, 1001, 256)).astype(np.float32)
    d_lam = torch.from_numpy(lam.copy()).cuda()
    d_buf = torch.empty_like(d_lam)
    sp = np.array(stages.SPONT_DEFAULT, np.float32)
    _lib.check(_lib.lib().cm_fix_lambda_run(_lib.ptr(d_lam), _lib.ptr(d_buf), 1001, _lib.ptr(sp)))
    torch.cuda.synchronize()
    s = sp[:, None, None]
    np.testing.assert_array_equal(d_lam.cpu().numpy(), np.maximum(np.float32(0), lam - s))
    np.testing.assert_array_equal(d_buf.cpu().numpy(), np.maximum(s, lam))


def test_diagnostic_entry_points():
    """viewGPUStatus and the cudaEvents* timers with plain-C handles."""
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import ctypes as C
    from cochlearmodel_amd import _lib
    L = _lib.lib()
    _say("case: diagnostics (memory status, event timer)")
    assert L.viewGPUStatus(16, b"diagnostics test") == 0
    a, b = C.c_void_p(), C.c_void_p()
    _lib.check(L.cudaEventsCreate(C.byref(a), C.byref(b), 1))
    _lib.check(L.cudaEventsStartTimer(a, b, 1))
    x = torch.ones(1 << 20, device="cuda")
    x.mul_(2.0)
    ms = L.cudaEventsStopQueryTimer(a, b, 1, b"timer")
    assert ms >= 0.0
    _lib.check(L.cudaEventsDestroy(a, b))
    assert L.cudaEventsStopQueryTimer(None, None, 0, b"off") == 0.0


def test_hearing_aid_transfer_functions(oracle):
    """generateInputFromProfile's hearing-aid stage (src/cochlea.cu:4104-4150):
    per profile interval an FIR over min(pos, taps) past samples (position 0
    gets 0), then the 'IIR' stage, which adds taps 1.. of its vector over the
    FIR output.  A pure one-sample delay pins the window rule; general taps are
    checked against a float64 restatement."""
    from cochlearmodel_amd import _lib, solver
    from cochlearmodel_amd._lib import check, ptr
    _say("case: hearing-aid FIR/IIR on generated profiles")
    c = _c4_small()
    L = _lib.lib()
    sec = solver.section_params()
    g = np.full(256, 0.5, np.float32)
    keep = [np.ascontiguousarray(x, np.float32) for x in (g, sec[0], sec[1], g, g, sec[2], sec[3], sec[4], sec[5],
                                                          sec[6], sec[7])]
    check(L.BMOHCKernel_Init(*[ptr(k) for k in keep], 1.0, 0.0, c["bufP1"] + 10, (c["bufP1"] + 10) * 256,
                             0, 1, 0, 0), "init")
    check(L.InputProfilesArrayInitialization(c["n_prof"], c["L"], 0, 0, 0), "init")

    def gen(fir, iir):
        fir = np.ascontiguousarray(fir, np.float32)
        iir = np.ascontiguousarray(iir, np.float32)
        out = np.zeros(c["n_prof"] * c["L"], np.float32)
        check(L.generateInputFromProfile(ptr(c["prof"]), ptr(c["WN"]), c["L"], None, 0, 0, 0, 0, 0.16, 0.0, 0,
                                         c["e1"], 0, c["e1"], c["n_prof"], 0, 0, c["n_prof"], c["ov"], c["L"],
                                         c["head"], c["length"], FS, 1, ptr(out), 0, 1, 1e-15, 1e-10, 0, 0.0, 0.0,
                                         c["DF"], 0, ptr(fir), len(fir), ptr(iir), len(iir)), "generate")
        return out.reshape(c["n_prof"], c["L"])

    x0 = gen([1.0], [1.0])
    delay = gen([0.0, 1.0], [1.0])
    np.testing.assert_array_equal(delay[:, 2:], x0[:, 1:-1])
    np.testing.assert_array_equal(delay[:, :2], 0.0)
    h = np.array([0.5, 0.3, -0.2, 0.1], np.float32)
    gi = np.array([1.0, -0.25, 0.125], np.float32)
    got = gen(h, gi)
    x = x0.astype(np.float64)
    tmp = np.zeros_like(x)
    for pos in range(x.shape[1]):
        for i in range(min(pos, len(h))):
            tmp[:, pos] += float(h[i]) * x[:, pos - i]
    ref = tmp.copy()
    for pos in range(x.shape[1]):
        for j in range(1, min(pos + 1, len(gi))):
            ref[:, pos] += float(gi[j]) * tmp[:, pos - j]
    rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    _say(f"hearing aid vs float64 restatement: rel L2 {rel:.2e}")
    assert rel <= 1e-6


@pytest.mark.parametrize("mode", ["reference", "plain", "extrapolated", "learned"])
def test_zero_input_gives_exact_zero_in_every_mode(oracle, mode):
    """Silence in, exact zeros out: BM velocity and pressure never leave 0 in
    the reference method and in tolerance mode with each guess (the residual of
    a zero right-hand side is zero, so no sweep runs).  Step, pass and sweep
    counts equal the oracle's on the same silent blocks; the learned guess is
    refused by its residual gate and matches the extrapolated run."""
    import torch
    from cochlearmodel_amd import solver, warmstart
    _say(f"case: silent input, 3 blocks of 441 samples, {mode}")
    n = 441
    inp = np.zeros(3 * n + 2, np.float32)
    items = np.zeros(3, solver.ITEM_DTYPE)
    for k in range(3):
        items[k] = (k * n, k * n, (k + 1) * n, k * n, 1e-10, 1e-15, 0, 0)
    sc = solver.default_scalars(FS)
    so = oracle.default_scalars(FS)
    if mode != "reference":
        ws = {"plain": 0, "extrapolated": 1, "learned": 2}[mode]
        sc.jacobi_mode, sc.tol_rel, sc.check_every, sc.max_sweeps, sc.warm_start = 1, 1e-6, 4, 1000, ws
        so.jacobi_mode, so.tol_rel, so.check_every, so.max_sweeps, so.warm_start = 1, 1e-6, 4, 1000, min(ws, 1)
    if mode == "learned":
        params, c_out, _ = warmstart.load_model()
        warmstart.upload(params, c_out)
    bb = solver.BatchedBM(np.full(256, 0.5, np.float32))
    d_speeds = torch.full((3 * n, 256), 7.0, dtype=torch.float32, device="cuda")
    out = bb.run(sc, torch.from_numpy(inp).cuda(), items, 3 * n, diagnostics=True, d_speeds=d_speeds)
    v = out["speeds"].cpu().numpy()
    st = out["stats"].cpu().numpy()
    assert np.all(v == 0.0), "silent input must give exact zeros"
    assert np.all(out["failed_nodes"].cpu().numpy() == 0)
    it = np.zeros(3, oracle.ITEM_DTYPE)
    for f in ("start", "mosi", "mesi", "write_start", "tol", "m1tol", "set"):
        it[f] = items[f]
    o = oracle.bm_run(so, oracle.section_params(), np.full(256, 0.5, np.float32), it, inp, 3 * n, "f32", threads=3)
    _say(f"steps gpu {st[:, 0]} oracle {o['steps']}, sweeps gpu {st[:, 2]} oracle {o['sweeps']}")
    np.testing.assert_array_equal(st[:, 0], o["steps"])
    np.testing.assert_array_equal(st[:, 1], o["passes"])
    np.testing.assert_array_equal(st[:, 2], o["sweeps"])
    if mode != "reference":
        assert np.all(st[:, 2] == 0)
