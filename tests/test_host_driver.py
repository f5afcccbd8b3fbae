"""cochlea_run, the C++ command-line host: the parameter-file reader on the
CPU (ConfigFileHandler semantics) and, on the GPU, the .bin stream of a VOICE
run against the Python mirror of Run_Cuda and the CPU oracle."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "cochlearmodel_amd", "cochlea_run")
FS = 44100.0


def _parse(tmp_path, text):
    p = tmp_path / "t.par"
    p.write_bytes(text.encode())
    r = subprocess.run([EXE, "--parse", str(p)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    out = {}
    for line in r.stdout.splitlines():
        k, v = line.split(" = ", 1)
        out[k] = v
    return out


def test_parameter_file_semantics(tmp_path):
    """src/ConfigFileHandler.cpp:32-56: key = value or key = "value" per line,
    first occurrence wins, CRLF accepted, empty or stray-quote values ignored."""
    assert os.path.exists(EXE), "cochlea_run is not built (python -m cochlearmodel_amd.build)"
    kv = _parse(tmp_path, 'A = 1\r\nB=  2 3  4\r\nA = 5\r\nName = "quoted value"\r\n'
                          'no equals here\r\nC_1=\r\nD =x"y\r\nVec = 1.5 abc 7\n  Indented_Key =\t9\n')
    assert kv == {"A": "1", "B": "2 3  4", "Name": "quoted value", "Vec": "1.5 abc 7", "Indented_Key": "9"}


def test_run_without_input_fails_cleanly(tmp_path):
    p = tmp_path / "bad.par"
    p.write_text("Output_Result = x.bin\n")
    r = subprocess.run([EXE, str(p)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "no input" in r.stderr


@pytest.mark.gpu
def test_voice_run_bin_stream(tmp_path, oracle):
    """Stream order and offsets of the .bin (failed nodes/blocks, convergence
    data, BM velocity, three lambdas, timers) against the same run through the
    Python mirror, and the velocity segment against the oracle, bit for bit."""
    from cochlearmodel_amd import _lib, solver, stages, workloads
    print("case: C1 tone through cochlea_run, Convergence_Tests=15, Allowed_Outputs=15", flush=True)
    sig = workloads.c1_signal()
    inp = tmp_path / "c1.f64"
    sig.astype(np.float64).tofile(inp)
    out = tmp_path / "c1.bin"
    (tmp_path / "c1.par").write_text(
        f"Input_File = {inp}\nFs = 44100\nTime_Blocks = 5\nTime_Block_Length = 0.02\nOverlap_Time = 0.005\n"
        f"OHC_Vector = 0.5\nIHC_Vector = 8\nAllowed_Outputs = 15\nConvergence_Tests = 15\n"
        f'Output_Result = "{out}"\n')
    r = subprocess.run([EXE, str(tmp_path / "c1.par")], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    got = np.fromfile(out, np.float32)
    # the same run through the Python mirror, then the library's buffers
    lay = solver.voice_layout(FS, 5, 0.02, 0.005)
    g = solver.run_bm_voice(sig, lay)
    L = _lib.lib()
    bufP1, TB, ov, tot = lay.bufP1, lay.time_blocks, lay.overlap_nodes, lay.total_nodes

    def dev(getter, n, dtype):
        a = np.zeros(n, dtype)
        _lib.check(L.GeneralKernel_Copy_Results_Bytes(_lib.ptr(a), getter(), a.nbytes, 0))
        return a

    ct = np.zeros(bufP1, np.float32)
    _lib.check(L.extractConvergenceTimes(_lib.ptr(ct), bufP1))
    parts = [ct[ov:tot], dev(L.getCudaConvergeSpeedBlocks, TB, np.float32),
             dev(L.getCudaFailedTimeNodes, bufP1, np.int32)[ov:tot].astype(np.float32),
             dev(L.getCudaFailedBlocks, TB, np.int32).astype(np.float32),
             dev(L.getCudaConvergedTimeNodes, bufP1, np.float32)[ov:tot],
             dev(L.getCudaConvergedBlocks, TB, np.float32),
             dev(L.getCudaConvergedJacobyLoopsPerIteration, bufP1, np.float32)[ov:tot],
             dev(L.getCudaConvergedJacobyLoopsPerIterationBlocks, TB, np.float32),
             g["speeds"][ov:tot].ravel()]
    lam, _ = stages.run_ihc_reference(tot, lay.block_nodes, TB, 0)
    parts += [lam[j, ov:tot].ravel() for j in range(3)]
    expect = np.concatenate(parts)
    assert got.size == expect.size + 2          # + timers (BM, IHC/lambda ms)
    np.testing.assert_array_equal(got[:-2], expect)
    # velocity segment against the oracle
    st, mo, me, ws = oracle.block_decomposition(TB, bufP1, ov, 0)
    o = oracle.bm_run(oracle.default_scalars(FS), oracle.section_params(), np.full(256, 0.5, np.float32),
                      oracle.make_items(st, mo, me, ws, 1e-10, 1e-15), g["input"], bufP1, "f32", threads=8)
    off = sum(p.size for p in parts[:8])
    np.testing.assert_array_equal(got[off:off + (tot - ov) * 256], o["speeds"][ov:tot].ravel())


@pytest.mark.gpu
@pytest.mark.parametrize("warm", [1, 2])
def test_voice_run_tolerance_mode(tmp_path, oracle, warm):
    """Jacobi_Tolerance / Warm_Start keys: the extrapolated guess is bit-exact
    against the oracle's same solver; the learned guess (committed network)
    stays within the tolerance-level distance of the exact per-pass solve."""
    from cochlearmodel_amd import solver, workloads
    print(f"case: C1 tone through cochlea_run, tolerance 1e-6, Warm_Start={warm}", flush=True)
    sig = workloads.c1_signal()
    inp = tmp_path / "c1.f64"
    sig.astype(np.float64).tofile(inp)
    out = tmp_path / "c1.bin"
    (tmp_path / "c1.par").write_text(
        f"Input_File = {inp}\nTime_Blocks = 5\nTime_Block_Length = 0.02\nOverlap_Time = 0.005\n"
        f"Allowed_Outputs = 1\nJacobi_Tolerance = 1e-6\nWarm_Start = {warm}\nOutput_Result = {out}\n")
    r = subprocess.run([EXE, str(tmp_path / "c1.par")], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    got = np.fromfile(out, np.float32)
    lay = solver.voice_layout(FS, 5, 0.02, 0.005)
    TB, ov, tot, bufP1 = lay.time_blocks, lay.overlap_nodes, lay.total_nodes, lay.bufP1
    v = got[(tot - ov) + TB:(tot - ov) + TB + (tot - ov) * 256].reshape(tot - ov, 256)
    x = np.zeros(bufP1, np.float32)
    x[:len(sig)] = sig
    st, mo, me, ws = oracle.block_decomposition(TB, bufP1, ov, 0)
    items = oracle.make_items(st, mo, me, ws, 1e-10, 1e-15)
    so = oracle.default_scalars(FS)
    so.jacobi_mode, so.tol_rel, so.check_every, so.max_sweeps, so.warm_start = (1, 1e-6, 4, 1000, 1) if warm == 1 \
        else (2, 1e-6, 4, 1000, 0)
    o = oracle.bm_run(so, oracle.section_params(), np.full(256, 0.5, np.float32), items, x, bufP1, "f32", threads=8)
    ref = o["speeds"][ov:tot]
    o0 = oracle.bm_run(oracle.default_scalars(FS), oracle.section_params(), np.full(256, 0.5, np.float32), items, x,
                       bufP1, "f32", threads=8)["speeds"][ov:tot]
    print(f"driver vs tolerance-mode oracle {np.abs(v - ref).max():.3e}, vs reference-mode oracle "
          f"{np.abs(v - o0).max():.3e}; {r.stdout.strip()}", flush=True)
    if warm == 1:
        np.testing.assert_array_equal(v, ref)
    else:
        rel = np.linalg.norm(v - ref) / np.linalg.norm(ref)
        print(f"learned warm start vs exact solve: rel L2 {rel:.2e}", flush=True)
        assert rel <= 2e-3


def _voice_par(tmp_path, extra):
    out = tmp_path / "stage.bin"
    (tmp_path / "stage.par").write_text(
        f"Fs = 44100\nTime_Blocks = 5\nTime_Block_Length = 0.02\nOverlap_Time = 0.005\nIHC_Vector = 8\n"
        f"Allowed_Outputs = 15\nOutput_Result = {out}\n" + extra)
    return tmp_path / "stage.par", out


def test_resume_needs_stage_data(tmp_path):
    par, _ = _voice_par(tmp_path, "Run_Stage_Calculation = 1\n")
    r = subprocess.run([EXE, str(par)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "need data or file" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("function_filter", [None, "type=equiripple,fpass=600,fstop=1600,apass=3,astop=60",
                                             "type=window,windowtype=kaiser,fpass=600,fstop=1600,astop=60",
                                             "type=butterworth,fpass=600,fstop=2400,apass=3,astop=30"])
def test_resume_from_bm_velocity(tmp_path, oracle, function_filter):
    """Run_Stage_Calculation = 1: the BM stage is skipped, the given velocity is
    uploaded, and the lambda planes equal the oracle's IHC/lambda of it (with
    the committed AC filter, or one designed from a Function_Filter string)."""
    from cochlearmodel_amd import solver, stages
    lay = solver.voice_layout(FS, 5, 0.02, 0.005)
    TB, ov, tot = lay.time_blocks, lay.overlap_nodes, lay.total_nodes
    print(f"case: cochlea_run resumed from BM velocity, {tot} rows, Function_Filter {function_filter}", flush=True)
    rng = np.random.default_rng(11)
    t = np.arange(tot)[:, None] / FS
    vel = (1e-4 * np.sin(2 * np.pi * 900.0 * t + np.linspace(0, 3, 256)[None, :]) *
           np.exp(-np.linspace(0, 4, 256))[None, :] + 1e-6 * rng.standard_normal((tot, 256))).astype(np.float32)
    f = tmp_path / "velocity.f32"
    vel.tofile(f)
    extra = f"Function_Filter = {function_filter}\n" if function_filter else ""
    par, out = _voice_par(tmp_path, f"Run_Stage_Calculation = 1\nRun_Stage_File_name = {f}\n" + extra)
    r = subprocess.run([EXE, str(par)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    got = np.fromfile(out, np.float32)
    b, a = stages.design_ac_filter(FS, function_filter) if function_filter else stages.load_ac_filter(FS)
    lo = oracle.ihc_lambda(vel, tot, lay.block_nodes * TB, b, stages.dc_filter(FS), stages.ihc_damage(8.0),
                           stages.aihc_table(), np.array(stages.SPONT_DEFAULT, np.float32), threads=8, ac_a=a)
    expect = np.concatenate([vel[ov:tot].ravel()] + [lo[j, ov:tot].ravel() for j in range(3)])
    assert got.size == expect.size + 2
    nv = (tot - ov) * 256
    np.testing.assert_array_equal(got[:nv], expect[:nv])
    if a is None:
        np.testing.assert_array_equal(got[nv:-2], expect[nv:])
    else:   # IIR: serial float accumulation, compared as the IIR stage test does
        rel = np.linalg.norm(got[nv:-2] - expect[nv:]) / np.linalg.norm(expect[nv:])
        print(f"IIR lambda rel L2 {rel:.2e}", flush=True)
        assert rel <= 1e-4
    assert got[-2] == 0.0                      # no BM kernel ran


@pytest.mark.gpu
def test_resume_from_lambda(tmp_path):
    """Run_Stage_Calculation = 2: lambda planes are uploaded and fixed up
    (CUDAFIXJND_Lambda); the written planes are the kept buffer
    max(spont, lambda), the velocity stage is empty."""
    from cochlearmodel_amd import solver, stages
    lay = solver.voice_layout(FS, 5, 0.02, 0.005)
    ov, tot = lay.overlap_nodes, lay.total_nodes
    print(f"case: cochlea_run resumed from lambda, {tot} rows", flush=True)
    rng = np.random.default_rng(12)
    lam = rng.uniform(0.0, 120.0, (3, tot, 256)).astype(np.float32)
    f = tmp_path / "lambda.f32"
    lam.tofile(f)
    par, out = _voice_par(tmp_path, f"Run_Stage_Calculation = 2\nRun_Stage_File_name = {f}\n")
    r = subprocess.run([EXE, str(par)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    got = np.fromfile(out, np.float32)
    sp = np.array(stages.SPONT_DEFAULT, np.float32)[:, None, None]
    kept = np.maximum(sp, lam)
    expect = np.concatenate([np.zeros((tot - ov) * 256, np.float32)] + [kept[j, ov:tot].ravel() for j in range(3)])
    assert got.size == expect.size + 2
    np.testing.assert_array_equal(got[:-2], expect)


def test_jnd_run_needs_output_file(tmp_path):
    p = tmp_path / "jnd.par"
    p.write_text("Sim_Type = 4\nCalculate_JND = 1\ntestedFrequencies = 1000\ntestedPowerLevels = 20 40\n")
    r = subprocess.run([EXE, str(p)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "JND File is missing" in r.stderr


@pytest.mark.gpu
def test_complex_jnd_run_files(tmp_path):
    """Sim_Type 4 through cochlea_run: the raw JND file (failed signals, AI and
    rate JND, each also transposed frequency-major) and the final file
    (complex JND per method and profile, warnings) against the same sweep run
    through the library's batched entry points."""
    from cochlearmodel_amd import solver
    from cochlearmodel_amd.jnd_sweep import JndSweep
    print("case: cochlea_run Sim_Type 4, 2 freq x 3 levels x 2 noises, 40 ms intervals", flush=True)
    sw = JndSweep((1000.0, 3000.0), (20.0, 40.0, 60.0), (1111.0, 40.0), duration=0.04, methods=3)
    sw.step(solver.default_scalars(FS), timed_bm=False)
    res = sw.finish()
    wn = tmp_path / "noise.f64"
    sw.WN.astype(np.float64).tofile(wn)
    fin, raw, res_bin = tmp_path / "jnd_final.bin", tmp_path / "jnd_raw.bin", tmp_path / "jnd_out.bin"
    common = ("Fs = 44100\ntestedFrequencies = 1000 3000\ntestedPowerLevels = 20 40 60\ntestedNoises = 1111 40\n"
              f"JND_Interval_Duration = 0.04\nInput_Noise_File = {wn}\nAllowed_Outputs = 15\n")
    (tmp_path / "jnd.par").write_text(
        "Sim_Type = 4\nCalculate_JND = 1\nComplex_JND_Calculation_Types = 3\n" + common +
        f"JND_File_Name = {fin}\nJND_Raw_File_Name = {raw}\nOutput_Result = {res_bin}\n")
    r = subprocess.run([EXE, str(tmp_path / "jnd.par")], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    n = len(sw.calc)
    got_raw = np.fromfile(raw, np.float32)
    assert got_raw.size == 5 * n
    fs_, ai, ai_t, rate, rate_t = (got_raw[k * n:(k + 1) * n] for k in range(5))
    np.testing.assert_array_equal(fs_, res["failed_signals"].astype(np.float32))
    d_ai = np.abs(ai - res["ai_db"].astype(np.float32)).max()
    d_rate = np.abs(rate - res["rate_db"].astype(np.float32)).max()
    print(f"driver vs batched sweep: AI {d_ai:.2e} dB, rate {d_rate:.2e} dB", flush=True)
    assert d_ai <= 1e-3 and d_rate <= 1e-3
    # calc index (noise, level, freq) -> frequency-major columns of noise x level
    nf, nl, nn = 2, 3, 2
    for i in range(n):
        f, rest = i % nf, i // nf
        assert ai_t[f * (nn * nl) + rest] == ai[i] and rate_t[f * (nn * nl) + rest] == rate[i]
    got_fin = np.fromfile(fin, np.float32)
    val, warn, conv = res["complex"]
    assert got_fin.size == 3 * val.size
    np.testing.assert_allclose(got_fin[:val.size], val.reshape(-1), atol=1e-3)
    np.testing.assert_array_equal(got_fin[val.size:2 * val.size], warn.reshape(-1).astype(np.float32))
    np.testing.assert_array_equal(got_fin[2 * val.size:], conv.reshape(-1).astype(np.float32))
    # Output_Result of the profile run: failed nodes and blocks, velocity and the
    # three lambda planes over rows [0, intervals x interval), then two timers
    rows = sw.rows
    stream = np.fromfile(res_bin, np.float32)
    assert stream.size == rows + sw.TB + 4 * rows * 256 + 2
    np.testing.assert_array_equal(stream[rows:rows + sw.TB], sw.d_failed_blocks.cpu().numpy().astype(np.float32))
    vel = stream[rows + sw.TB:rows + sw.TB + rows * 256].reshape(rows, 256)
    np.testing.assert_array_equal(vel, sw.d_speeds[:rows].cpu().numpy())
    lam = stream[rows + sw.TB + rows * 256:-2].reshape(3, rows, 256)
    np.testing.assert_array_equal(lam, sw.d_lambda.cpu().numpy())
    # Sim_Type 3 without JND: the same tones without reference intervals; each
    # interval's velocity equals that of the same profile in the run above,
    # except the interval's last row, whose drive interpolates towards the
    # first sample of whatever interval follows (a reference interval there)
    res3 = tmp_path / "prof_out.bin"
    (tmp_path / "prof.par").write_text("Sim_Type = 3\nCalculate_JND = 0\n" + common + f"Output_Result = {res3}\n")
    r3 = subprocess.run([EXE, str(tmp_path / "prof.par")], capture_output=True, text=True, timeout=300)
    assert r3.returncode == 0, r3.stdout + r3.stderr
    n3 = len(sw.calc)
    rows3 = n3 * sw.interval
    tb3 = n3 * sw.DF
    s3 = np.fromfile(res3, np.float32)
    assert s3.size == rows3 + tb3 + 4 * rows3 * 256 + 2
    vel3 = s3[rows3 + tb3:rows3 + tb3 + rows3 * 256].reshape(n3, sw.interval, 256)
    vel_all = vel.reshape(sw.n_prof, sw.interval, 256)
    for p_i, j in enumerate(sw.calc):
        np.testing.assert_array_equal(vel3[p_i][:-1], vel_all[j][:-1], err_msg=f"profile {p_i}")


def test_stage_names_are_checked_before_any_device_call(tmp_path):
    """TEST_File_Target / Type_TEST_Output: an unknown stage name, a lambda plane
    past the third and a JND stage in a run without Calculate_JND are refused
    while the parameters are read."""
    jnd = ("Sim_Type = 4\nCalculate_JND = 1\ntestedFrequencies = 1000\ntestedPowerLevels = 20 40\n"
           "JND_File_Name = j.bin\nTEST_File_Target = t.bin\n")
    voice = "Input_Signal = 0 0 0 0\nOutput_Result = o.bin\nTEST_File_Target = t.bin\n"
    for text, err in ((jnd + "Type_TEST_Output = dLambdaX\n", "unknown Type_TEST_Output"),
                      (jnd + "Type_TEST_Output = JND_Lambda3\n", "must be smaller than 3"),
                      (voice + "Type_TEST_Output = CRLB_RA\n", "needs Calculate_JND"),
                      (voice, "needs Calculate_JND")):       # the default MeanRate
        p = tmp_path / "s.par"
        p.write_text(text)
        r = subprocess.run([EXE, str(p)], capture_output=True, text=True, timeout=60)
        assert r.returncode == 1 and err in r.stderr, (text, r.stderr)


@pytest.mark.gpu
def test_jnd_stage_dumps(tmp_path, oracle):
    """TEST_File_Target dumps of a Sim_Type 4 run through cochlea_run: the JND
    stages (AvgMeanRate, dMeanRate, dLambda of one plane, preFisherAI,
    preFisherAITimeReduced, CRLB_RA, MeanRate) against the oracle's stages on
    the same lambda, a lambda plane and the IHC stage against the oracle's IHC
    chain on the same BM velocity, bit for bit."""
    from cochlearmodel_amd import solver, stages
    from cochlearmodel_amd.jnd_sweep import JndSweep, N_IHC
    print("case: cochlea_run Sim_Type 4 stage dumps, 2 freq x 3 levels x 2 noises, 40 ms intervals", flush=True)
    sw = JndSweep((1000.0, 3000.0), (20.0, 40.0, 60.0), (1111.0, 40.0), duration=0.04)
    sw.step(solver.default_scalars(FS), timed_bm=False)
    lam = sw.d_lambda.cpu().numpy()
    speeds = sw.d_speeds[:sw.rows].cpu().numpy()
    wn = tmp_path / "noise.f64"
    sw.WN.astype(np.float64).tofile(wn)
    common = ("Sim_Type = 4\nCalculate_JND = 1\nFs = 44100\ntestedFrequencies = 1000 3000\n"
              "testedPowerLevels = 20 40 60\ntestedNoises = 1111 40\nJND_Interval_Duration = 0.04\n"
              f"Input_Noise_File = {wn}\nJND_File_Name = {tmp_path / 'jnd.bin'}\n")
    nIHC = np.full(256, N_IHC, np.float64)
    W3 = np.array(stages.W_DEFAULT, np.float32)
    dA = sw.profiles["dA"].astype(np.float32)

    def oracle_stage(k):
        return oracle.jnd(lam, sw.n_prof, sw.interval, sw.head, sw.length, dA, sw.ref_index, nIHC, W3, FS,
                          0.2 * FS, backup_stage=k)[3]

    body = np.concatenate([np.arange(i * sw.interval + sw.head, i * sw.interval + sw.head + sw.length)
                           for i in range(sw.n_prof)])
    b, _ = stages.load_ac_filter(FS)
    _, ihc_st = oracle.ihc_lambda(speeds, sw.interval, sw.block * sw.TB, b, stages.dc_filter(FS),
                                  stages.ihc_damage(8.0), stages.aihc_table(),
                                  np.array(stages.SPONT_DEFAULT, np.float32), stages=True, threads=8)
    mr = oracle.jnd(lam, sw.n_prof, sw.interval, sw.head, sw.length, dA, sw.ref_index, nIHC, W3, FS, 0.2 * FS)[0]
    cases = {
        "AvgMeanRate": lambda: oracle_stage(1),
        "dMeanRate": lambda: oracle_stage(2),
        "dLambda1": lambda: oracle_stage(3)[1][body],
        "preFisherAI": lambda: oracle_stage(4)[:, body],
        "preFisherAITimeReduced": lambda: oracle_stage(5),
        "CRLB_RA": lambda: oracle_stage(6),
        "MeanRate": lambda: mr,
        "JND_Lambda2": lambda: lam[2],
        "IHC": lambda: ihc_st[4],
    }
    for name, expect in cases.items():
        tgt = tmp_path / f"{name}.bin"
        (tmp_path / "s.par").write_text(common + f"TEST_File_Target = {tgt}\nType_TEST_Output = {name}\n")
        r = subprocess.run([EXE, str(tmp_path / "s.par")], capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, (name, r.stdout + r.stderr)
        exp = expect()
        got = np.fromfile(tgt, np.float32)
        if name == "dLambda1":
            got = got.reshape(sw.rows, 256)[body]
        elif name == "preFisherAI":
            got = got.reshape(3, sw.rows, 256)[:, body]
        else:
            got = got.reshape(exp.shape)
        print(f"{name}: {exp.size} values, exact {np.mean((got == exp) | (np.isnan(got) & np.isnan(exp))):.6f}",
              flush=True)
        np.testing.assert_array_equal(got, exp, err_msg=name)


def test_profile_layout_without_jnd(tmp_path):
    """Calculate_JND = 0: no reference intervals, reference indexes -1."""
    p = tmp_path / "p.par"
    p.write_text("testedFrequencies = 1000 3000\ntestedPowerLevels = 20 40\ntestedNoises = 1111 40\n"
                 "Calculate_JND = 0\n")
    r = subprocess.run([EXE, "--profiles", str(p)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    rows = [line.split() for line in r.stdout.splitlines()]
    assert len(rows) == 8
    assert all(int(x[4]) == 0 and int(x[6]) == -1 for x in rows)
    assert [int(x[5]) for x in rows] == list(range(8))


@pytest.mark.parametrize("one_ref", [1, 0])
def test_jnd_profile_layout_matches_python_mirror(tmp_path, one_ref):
    """The host's CParams::updateInputProfile restatement (cochlea_run
    --profiles) against stages.jnd_profiles: dA, noise power, frequency,
    reference flags and indexes, serial positions."""
    from cochlearmodel_amd import stages
    p = tmp_path / "p.par"
    p.write_text(f"testedFrequencies = 1000.3 3000\ntestedPowerLevels = 20 40.7 60\ntestedNoises = 1111 40\n"
                 f"Generate_One_Ref_Per_Noise_Level = {one_ref}\n")
    r = subprocess.run([EXE, "--profiles", str(p)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    rows = [line.split() for line in r.stdout.splitlines()]
    prof, refs, serial, calc, to_ref = stages.jnd_profiles([1000.3, 3000.0], [20.0, 40.7, 60.0], noises=(1111.0, 40.0),
                                                           one_ref_per_noise=bool(one_ref))
    assert len(rows) == len(prof)
    for g, row in enumerate(rows):
        assert int(row[0]) == g
        assert float(row[1]) == prof[g]["dA"] and float(row[2]) == prof[g]["Wn"]
        assert float(row[3]) == prof[g]["frequency"]
        assert int(row[4]) == prof[g]["isReference"] and int(row[5]) == prof[g]["calculateIndexPosition"]
        assert int(row[6]) == prof[g]["referenceIndex"] and int(row[7]) == serial[g]


@pytest.mark.gpu
def test_sine_run(tmp_path, oracle):
    """Sim_Type 1 (SIM_TYPE_SIN): the tone of CState::gen_input_data (time
    accumulated in steps of 1/Fs, amplitude CONV_dB_TO_AMP in float) through
    the BM stage, bit for bit against the oracle on the same input."""
    import math
    from cochlearmodel_amd import solver
    lay = solver.voice_layout(FS, 5, 0.02, 0.005)
    TB, ov, tot, bufP1 = lay.time_blocks, lay.overlap_nodes, lay.total_nodes, lay.bufP1
    print("case: cochlea_run Sim_Type 1, 1 kHz at 60 dB, 5 time blocks", flush=True)
    out = tmp_path / "sin.bin"
    (tmp_path / "sin.par").write_text(f"Sim_Type = 1\nSin_Freq = 1000\nSin_dB = 60\nFs = 44100\nTime_Blocks = 5\n"
                                      f"Allowed_Outputs = 1\nOutput_Result = {out}\n")
    r = subprocess.run([EXE, str(tmp_path / "sin.par")], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    amp = float(np.float32(10.0) * np.float32(20e-6) * np.power(np.float32(10.0), np.float32(60.0) / np.float32(20.0)))
    x = np.zeros(bufP1, np.float32)
    t, Ts = 0.0, 1.0 / FS
    for i in range(bufP1):
        x[i] = np.float32(amp * math.sin(2.0 * math.pi * 1000.0 * t))
        t += Ts
    st, mo, me, ws = oracle.block_decomposition(TB, bufP1, ov, 0)
    o = oracle.bm_run(oracle.default_scalars(FS), oracle.section_params(), np.full(256, 0.5, np.float32),
                      oracle.make_items(st, mo, me, ws, 1e-10, 1e-15), x, bufP1, "f32", threads=8)
    got = np.fromfile(out, np.float32)
    n = tot - ov
    v = got[n + TB:n + TB + n * 256]                # after failed nodes and blocks
    np.testing.assert_array_equal(v, o["speeds"][ov:tot].ravel())
