"""Benchmark of the cochlear hot path on one MI355X.

Workload (default, --workload c5): the learned warm start against the plain
batched solver on 4096 held-out pure tones (frequency log-uniform 250-8000 Hz,
level uniform 20-90 dB SPL, seeds disjoint from the network's 2048 training
tones), 200 ms each at 44.1 kHz, 256 sections, healthy profiles (OHC 0.5,
IHC 8), FP32, reference layout (Decouple_Filter=10, 20 ms time blocks), all
four Allowed_Outputs kept in HBM (BM velocity + lambda high/medium/low).  The
set is two launches of 2048 tones (20480 time blocks each), all resident.

One step = profile input synthesis + BM/OHC solve + IHC/lambda of one launch,
alternating between the two.  Inputs and tables are resident in HBM before timing
starts.  After the timed steps every solver variant runs on the same resident
inputs and is scored against the CPU oracle's exact per-pass solve on
sampled blocks.  Prints one JSON line.

Other workloads: c1 (one tone through the VOICE call sequence), c2 (the
256-tone grid, 2560 blocks in one launch), c3 (1024 uneven signals), c4 (the
JND sweep).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FS = 44100.0
SECTIONS = 256
PEAK_FP32_TFLOPS = 157.3     # MI355X FP32 (dense) peak
PEAK_HBM_GBS = 8000.0        # MI355X HBM3E


def flops_per_launch(steps: int, passes: int, sweeps: int) -> float:
    """Algorithmic FP32 work of the kernel-17 method: per section-step
    sum_k (C_k + 4 J_k) with C_0 = 63, C_1 = 79, C_{k>=2} = 49."""
    return SECTIONS * (142.0 * steps + 49.0 * (passes - 2 * steps) + 4.0 * sweeps)


def c2_workload(n_signals: int = 256):
    from cochlearmodel_amd import batch, workloads
    f, lv = workloads.c2_grid()
    return batch.ToneBatch(f[:n_signals], lv[:n_signals], 0.2)


TOL = 1e-6   # relative residual of the tolerance-mode Jacobi


def variant_scalars(name: str):
    """Solver settings of each measured variant (cm_bm_scalars)."""
    from cochlearmodel_amd import solver
    sc = solver.default_scalars(FS)
    if name == "reference_20_2":
        return sc
    sc.jacobi_mode, sc.tol_rel, sc.check_every, sc.max_sweeps = 1, TOL, 4, 1000
    sc.warm_start = 2 if name.startswith("learned_tol") else {"plain_tol": 0, "extrapolated_tol": 1}[name]
    return sc


def set_network_every(name: str):
    """The learned variants consult the network on every step (learned_tol) or
    on every k-th step (learned_tol_every<k>), extrapolating in between."""
    from cochlearmodel_amd import _lib
    k = int(name.split("every")[1]) if "every" in name else 1
    _lib.check(_lib.lib().cm_ws_set_every(k), "cm_ws_set_every")


VARIANTS = ("reference_20_2", "plain_tol", "extrapolated_tol", "learned_tol")
EXTRA_VARIANTS = ("learned_tol_every2", "learned_tol_every4")   # reported in the variants table only
# Launch order of each variant on the held-out set, heavy tones first
# (ToneBatch.tone_order): with a guess, sweeps per step rise with a tone's
# frequency, without one they fall.  Measured on a 2048-tone launch (ms,
# longest-first -> this order, identical results): learned 3071 -> 2937 and
# extrapolated 1312 -> 1253 (higher frequency first), plain 1783 -> 1745 (lower
# frequency first).  The reference method (fixed sweeps) keeps longest-first:
# 1145-1156 ms in every order.
TONE_ORDER = {"learned_tol": "high", "learned_tol_every2": "high", "learned_tol_every4": "high",
              "extrapolated_tol": "high", "plain_tol": "low"}
MLP_FLOP_PER_STEP = 2.0 * (769 * 512 + 512 * 512 + 512 * 256)   # forward, one block-step
# HBM bytes per launch on the 256-tone batch from rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE (KiB per dispatch), scaled by what the counters read for a 1 GiB
# device copy (FETCH x 2.0, WRITE x 1.0); profiles/pmc_traffic_c2.md
MEASURED_HBM_BYTES = {
    "learned_tol": (12239.25 * 2.0 + 2258200.0) * 1024,
    "reference_20_2": (6085.375 * 2.0 + 2258260.0) * 1024,
    "ihc_lambda": (1150991.3125 * 2.0 + 4515840.0 + 2288920.5 * 2.0 + 6773760.0) * 1024,
}
PEAK_BF16_TFLOPS = 2500.0    # MI355X dense BF16 MFMA
# VALU instructions per launch on the 256-tone batch (rocprofv3 --pmc SQ_INSTS_VALU;
# profiles/pmc_valu_c2_reference.md, profiles/pmc_mix_c2_learned.md) and the issue
# capacity they are measured against: a wave64 FP32 VALU instruction holds a SIMD
# for 4 cycles, 1024 SIMDs at 2.4 GHz
MEASURED_VALU_INSTS = {"reference_20_2": 1.0995e11, "learned_tol": 1.64154213976e11,
                       "ihc_lambda": 1.40262e9 + 2.68728e9}   # ac_ds_kernel + dc_lambda_kernel
# VALU issue capacity: the rate bm_kernel's own instruction mix reaches at full
# occupancy (SQ_INSTS_VALU / kernel time on the JND sweep's 48300 blocks,
# profiles/pmc_valu_c4_reference.md).  The 4-cycle model (1024 SIMDs x 2.4 GHz
# / 4 cycles per wave64 FP32 instruction = 6.14e11/s) sits below it, so it is
# not a ceiling for this mix.
VALU_ISSUE_PER_S = 7.38e11
VALU_ISSUE_MODEL = ("measured: bm_kernel at full occupancy issues 7.38e11 VALU instructions/s "
                    "(profiles/pmc_valu_c4_reference.md); the 4-cycle wave64 model gives 6.14e11/s")
# L2 -> L1 fill of all CUs: 64 B per clock per CU, 256 CUs at 2.4 GHz
L1_FILL_BYTES_PER_S = 64 * 256 * 2.4e9
# FP32 without packed-FP32 instructions (the build disables them): one FMA per
# lane per cycle, half the published vector peak
PEAK_FP32_UNPACKED_TFLOPS = PEAK_FP32_TFLOPS / 2
# The same counters on one launch of the held-out set (2048 tones, 20480 blocks,
# longest block first; profiles/pmc_traffic_c5_learned.md, profiles/pmc_mix_c5_learned.md)
C5_MEASURED = {
    "tones_per_launch": 2048,
    "hbm_bytes": {"learned_tol": (2918599.625 * 2.0 + 732940369.34375) * 1024,
                  "ihc_lambda": (9206258.75 * 2.0 + 36126720.0 + 18361109.0625 * 2.0 + 54190080.0) * 1024},
    "valu_insts": {"learned_tol": 1.230589350574e12, "plain_tol": 1.269936053668e12,
                   "ihc_lambda": 1.1221170816e10 + 2.1498514912e10},
    "l2_read_bytes": {"learned_tol": 3.90384312554e11 * 128},
    "mfma_insts": {"learned_tol": 4.769789056e10},
}
# L1<-L2 reads of the learned kernel: TCP_TCC_READ_REQ_sum x 128 B (a 1 GiB copy
# reads 8.389e6 requests); profiles/pmc_mix_c2_learned.md
MEASURED_L2_READ_BYTES = {"learned_tol": 6.5295415974e10 * 128}


def load_warm_start_model():
    """Upload the committed warm-start network into the solver."""
    from cochlearmodel_amd import warmstart
    params, c_out, meta = warmstart.load_model()
    warmstart.upload(params, c_out)
    return meta


def sample_blocks(work, n_blocks):
    return work.sample_blocks(n_blocks)


def oracle_items(work, sel):
    from oracle import oracle as O
    it = np.zeros(len(sel), O.ITEM_DTYPE)
    for f in ("start", "mosi", "mesi", "write_start", "tol", "m1tol", "set"):
        it[f] = work.items[f][sel]
    return it


def block_rows(work, sel):
    return work.block_rows(sel)


def exact_reference(work, n_blocks, threads):
    """O2 (exact per-pass pressure solve, FP32 state) on sampled blocks: the
    accuracy yardstick every variant is scored against; also the FP64 rebuild
    of the reference method on the same blocks."""
    from oracle import oracle as O
    sel = sample_blocks(work, n_blocks)
    sc = O.default_scalars(FS)
    sc.jacobi_mode = 2
    inp = work.d_input.cpu().numpy()
    o = O.bm_run(sc, O.section_params(), np.full(SECTIONS, 0.5, np.float32), oracle_items(work, sel), inp, work.bufP1,
                 "f32", threads=threads)
    o64 = O.bm_run(O.default_scalars(FS), O.section_params(), np.full(SECTIONS, 0.5, np.float32),
                   oracle_items(work, sel), inp, work.bufP1, "f64", threads=threads)
    rows = block_rows(work, sel)
    return rows, o["speeds"][rows], o64["speeds"][rows]


def host_cpu():
    """Model name and core count of the host the CPU baseline ran on."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count()}


def cpu_baseline(work, n_blocks: int, threads: int, sc_gpu):
    """The CPU oracle (FP32 port of kernel 17) on every (TB/n_blocks)-th block."""
    from oracle import oracle as O
    inp = work.d_input.cpu().numpy()
    sel = sample_blocks(work, n_blocks)
    stride = max(1, work.TB // n_blocks)
    it = oracle_items(work, sel)
    sc = O.default_scalars(FS)
    for f, _ in O.BMScalars._fields_:
        setattr(sc, f, getattr(sc_gpu, f))
    t0 = time.perf_counter()
    o = O.bm_run(sc, O.section_params(), np.full(SECTIONS, 0.5, np.float32), it, inp, work.bufP1, "f32",
                 threads=threads)
    dt = time.perf_counter() - t0
    steps = int(o["steps"].sum())
    return {"value": SECTIONS * steps / dt, "unit": "section-steps/s", "cores": threads, "kind": "port",
            "sample": f"{len(sel)} of {work.TB} time blocks (every {stride}th), {steps} committed steps, "
                      f"{dt:.2f} s wall, FP32 oracle of the same solver settings with OpenMP over blocks"}, o, sel


def learned_cpu_baseline(work, cpu_blocks: int, threads: int):
    """The headline method on the CPU: the oracle's FP32 solver with the same
    learned predictor-pass guess (the committed network on the host, bfloat16
    operands, float accumulation), on a bounded sample of the launch's blocks,
    beside the GPU's learned run of the same blocks.  The extrapolated guess
    (the fastest CPU variant) is timed on a larger sample for reference."""
    import torch
    from cochlearmodel_amd import warmstart
    from oracle import oracle as O
    params, c_out, _ = warmstart.load_model()
    O.set_ws_net(params, c_out)
    csc = variant_scalars("learned_tol")
    set_network_every("learned_tol")
    work.run_bm(csc, False)
    torch.cuda.synchronize()
    cb, o, sel = cpu_baseline(work, max(1, cpu_blocks // 8), threads, csc)
    cb["solver"] = "learned_tol"
    cb.update(host_cpu())
    cb1, _, _ = cpu_baseline(work, max(1, cpu_blocks // 128), 1, csc)
    cb["single_thread"] = {"value": cb1["value"], "sample": cb1["sample"]}
    rows = work.block_rows(sel)
    cb["parity_rel_l2"] = _rel(work.sampled_speeds(rows), o["speeds"][rows])
    g = work.d_stats.view(work.TB, 3).cpu().numpy()[sel].astype(np.int64)
    cb["parity"] = {"committed_steps": {"gpu": int(g[:, 0].sum()), "cpu": int(o["steps"].sum())},
                    "sweeps_per_step": {"gpu": float(g[:, 2].sum() / g[:, 0].sum()),
                                        "cpu": float(o["sweeps"].sum() / o["steps"].sum())},
                    "note": "the CPU network accumulates in its own order, so guesses differ in the last bits "
                            "and the two runs stop their Jacobi iterations at different iterates within the "
                            "same tolerance"}
    esc = variant_scalars("extrapolated_tol")
    ce, _, _ = cpu_baseline(work, cpu_blocks, threads, esc)
    cb["extrapolated_tol"] = {"value": ce["value"], "sample": ce["sample"]}
    return cb


def c1_main(args):
    """One 1 kHz, 60 dB tone through the reference's VOICE call sequence
    (BMOHCKernel_Init, BMOHCNewKernel, copies): 5 time blocks of 20 ms.  The
    BM time is the kernel's own event timing; end to end adds the host copies."""
    import torch
    from cochlearmodel_amd import solver, workloads
    from cochlearmodel_amd._lib import lib
    lay = solver.voice_layout(FS, 5, 0.02, 0.005)
    sig = workloads.c1_signal()
    for _ in range(args.warmup):
        solver.run_bm_voice(sig, lay, timed=True)
    bm, e2e = [], []
    g = None
    for _ in range(args.steps):
        t0 = time.perf_counter()
        g = solver.run_bm_voice(sig, lay, timed=True)
        e2e.append((time.perf_counter() - t0) * 1e3)
        bm.append(lib().cm_last_kernel_ms())
    steps = int(g["stats"]["steps"].sum())
    bm_avg = float(np.mean(bm))
    out = {"metric": "section-steps/s (one VOICE tone)", "value": SECTIONS * steps / (bm_avg * 1e-3),
           "unit": "section-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": float(np.mean(e2e)), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic",
           "config": {"workload": "1 kHz tone, 60 dB SPL, 100 ms with 10 ms ramps and 220 leading zeros, 44.1 kHz, "
                                  "5 time blocks of 20 ms, reference method", "time_blocks": lay.time_blocks},
           "bm_ms": bm_avg, "committed_steps": steps,
           "note": "5 waves on a 256-CU chip: this case measures latency, not throughput"}
    if not args.no_cpu:
        from oracle import oracle as O
        items = O.make_items(*O.block_decomposition(lay.time_blocks, lay.bufP1, lay.overlap_nodes, 0), 1e-10, 1e-15)
        inp = g["input"]
        res = {}
        for threads in (1, lay.time_blocks):
            t1 = time.perf_counter()
            o = O.bm_run(O.default_scalars(FS), O.section_params(), np.full(SECTIONS, 0.5, np.float32), items, inp,
                         lay.bufP1, "f32", threads=threads)
            res[threads] = (time.perf_counter() - t1, int(o["steps"].sum()), o)
        dt, st, o = res[lay.time_blocks]
        v, vo = g["speeds"][:lay.bufP1], o["speeds"][:lay.bufP1]
        out["cpu_baseline"] = {"value": SECTIONS * st / dt, "unit": "section-steps/s", "cores": lay.time_blocks,
                               "kind": "port", "solver": "reference_20_2",
                               "sample": f"the whole tone ({st} committed steps, {dt:.2f} s wall, one thread per "
                                         "time block)",
                               "single_thread": {"value": SECTIONS * res[1][1] / res[1][0],
                                                 "sample": f"the whole tone, {res[1][0]:.2f} s wall, one thread"},
                               "parity_rel_l2": float(np.linalg.norm(v - vo) / max(np.linalg.norm(vo), 1e-30)),
                               **host_cpu()}
    print(json.dumps(out))


def c3_main(args):
    """Uneven batch: 1024 seeded signals (20-500 ms tones, 3-harmonic complexes
    and noise, 0-100 dB, per-signal OHC spline and IHC), reference method,
    launch shape from the autotuner; one step = BM + IHC/lambda."""
    import torch
    from cochlearmodel_amd import solver, workloads
    from cochlearmodel_amd.uneven import UnevenBatch
    sigs, ohc, ihc, meta = workloads.c3_signals(args.c3_signals)
    ub = UnevenBatch(sigs, ohc, ihc, FS)
    sc = solver.default_scalars(FS)
    print(f"# uneven batch ready: {ub.n} signals, {ub.TB} time blocks, {ub.rows} rows", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    tune = ub.autotune(sc, trial_stride=8)
    t_tune = time.perf_counter() - t0
    print(f"# autotune {t_tune:.1f} s: {tune}", file=sys.stderr, flush=True)
    d_lambda = torch.empty((3, ub.rows, SECTIONS), dtype=torch.float32, device="cuda")
    d_scratch = torch.empty((2, ub.rows, SECTIONS), dtype=torch.float32, device="cuda")
    for _ in range(args.warmup):
        ub.run_bm(sc)
        ub.run_lambda(d_lambda, d_scratch)
    torch.cuda.synchronize()
    bm_ms, lam_ms = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        bm_ms.append(ub.run_bm(sc))
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        ub.run_lambda(d_lambda, d_scratch)
        e1.record()
        e1.synchronize()
        lam_ms.append(e0.elapsed_time(e1))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    steps, passes, sweeps = ub.stats()
    ms_per_step = 1e3 * wall / args.steps
    bm_avg = float(np.mean(bm_ms))
    # the same launch without the tuning: one wave per workgroup, natural order,
    # sweeps in pairs
    ub.clear_order()
    ms_untuned = ub.run_bm(sc, waves_per_block=1)
    flops = flops_per_launch(steps, passes, sweeps)
    lam_bytes = ub.rows * SECTIONS * 16.0
    out = {
        "metric": "section-steps/s (uneven batch)", "value": SECTIONS * steps / (ms_per_step * 1e-3),
        "unit": "section-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{ub.n} seeded uneven signals (20-500 ms; tones 125-10000 Hz, 3-harmonic complexes, "
                               "Gaussian noise; 0-100 dB; per-signal OHC spline in [0, 0.5] and IHC in [5, 8]), "
                               "44.1 kHz, 20 ms time blocks with 5 ms overlap, reference kernel-17 method, "
                               "BM + IHC/lambda per step",
                   "signals": ub.n, "time_blocks": ub.TB, "rows": ub.rows, "samples": int(sum(ub.lengths))},
        "bm_ms": bm_avg, "ihc_lambda_ms": float(np.mean(lam_ms)), "committed_steps": steps,
        "jacobi_iters_per_step": sweeps / steps, "signals_per_s": ub.n / (ms_per_step * 1e-3),
        "autotune": {"sections_per_lane": int(tune["sections_per_lane"]),
                     "waves_per_block": int(tune["waves_per_block"]), "longest_first": bool(tune["order_lpt"]),
                     "jacobi_unroll": int(tune["jacobi_unroll"]),
                     "trial_blocks": int(tune["n_trial_items"]), "trial_ms": [float(x) for x in tune["ms"]],
                     "trial_ms_split": [float(x) for x in tune["ms_split"]],
                     "seconds": t_tune, "bm_ms_untuned": ms_untuned},
        "roofline": {"bound": "valu", "engine": "FP32 VALU", "achieved": flops / (bm_avg * 1e-3) / 1e12,
                     "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": flops / (bm_avg * 1e-3) / 1e12 / PEAK_FP32_TFLOPS, "traffic": None,
                     "kernel": "bm_kernel", "flop_per_launch": flops},
        "ihc_lambda_roofline": {"bound": "hbm", "achieved": lam_bytes / (np.mean(lam_ms) * 1e-3) / 1e9,
                                "peak": PEAK_HBM_GBS, "unit": "GB/s",
                                "frac": lam_bytes / (np.mean(lam_ms) * 1e-3) / 1e9 / PEAK_HBM_GBS},
    }
    if not args.no_cpu:
        from oracle import oracle as O
        threads = min(args.cpu_threads, os.cpu_count() or 1)
        sel = np.arange(0, ub.TB, max(1, ub.TB // args.cpu_blocks))[:args.cpu_blocks]
        # the sampled blocks on a compacted input: block j reads only input
        # [start, mesi + 2) and writes rows [write_start, mesi), so re-basing
        # each block onto its own copied range leaves its arithmetic unchanged
        it = np.zeros(len(sel), O.ITEM_DTYPE)
        pieces, grows, crows, off = [], [], [], 0
        for j, b in enumerate(sel):
            item = ub.items[b]
            a0, a1 = int(item["start"]), int(item["mesi"]) + 2
            pieces.append(ub.input[a0:min(a1, ub.rows)])
            if a1 > ub.rows:
                pieces.append(np.zeros(a1 - ub.rows, np.float32))
            shift = off - a0
            it[j] = (a0 + shift, item["mosi"] + shift, item["mesi"] + shift, item["write_start"] + shift,
                     item["tol"], item["m1tol"], item["set"], 0)
            grows.append(np.arange(item["write_start"], item["mesi"]))
            crows.append(np.arange(item["write_start"] + shift, item["mesi"] + shift))
            off += a1 - a0
        cin = np.concatenate(pieces)
        t1 = time.perf_counter()
        o = O.bm_run(O.default_scalars(FS), O.section_params(), ohc, it, cin, off, "f32", threads=threads)
        dt = time.perf_counter() - t1
        st = int(o["steps"].sum())
        d_rows = torch.from_numpy(np.concatenate(grows)).to("cuda")
        ub.run_bm(sc, timed=False)
        g = ub.d_speeds.index_select(0, d_rows).cpu().numpy()
        ref = o["speeds"][np.concatenate(crows)]
        # one thread on the first sixteenth of the sample: the per-core rate
        n1 = max(1, len(sel) // 16)
        t1 = time.perf_counter()
        o1 = O.bm_run(O.default_scalars(FS), O.section_params(), ohc, it[:n1], cin, off, "f32", threads=1)
        dt1 = time.perf_counter() - t1
        st1 = int(o1["steps"].sum())
        out["cpu_baseline"] = {"value": SECTIONS * st / dt, "unit": "section-steps/s", "cores": threads,
                               "kind": "port", "solver": "reference_20_2",
                               "single_thread": {"value": SECTIONS * st1 / dt1,
                                                 "sample": f"{n1} of the sampled blocks, {st1} committed steps, "
                                                           f"{dt1:.2f} s wall, one thread"},
                               "sample": f"{len(sel)} of {ub.TB} time blocks (every {max(1, ub.TB // args.cpu_blocks)}"
                                         f"th), {st} committed steps, {dt:.2f} s wall, FP32 oracle, OpenMP over "
                                         "blocks",
                               "parity_rel_l2": float(np.linalg.norm(g - ref) / max(np.linalg.norm(ref), 1e-30)),
                               **host_cpu()}
    print(json.dumps(out))


def c4_sweep():
    from cochlearmodel_amd.jnd_sweep import JndSweep
    freqs = np.geomspace(250.0, 8000.0, 16)
    levels = np.arange(10.0, 101.0, 10.0)
    noises = (1111.0, 20.0, 40.0, 60.0, 80.0)        # 1111 = MIN_INF_POWER_LEVEL, no noise
    return JndSweep(freqs, levels, noises, duration=0.3, methods=3)


def c4_oracle_interval(sw, j, threads):
    """CPU oracle of one decoupled interval of the sweep: its blocks re-based onto
    their own input range (BM), then IHC/lambda over its rows."""
    from oracle import oracle as O
    from cochlearmodel_amd import stages
    blk = sw.interval_blocks(j)
    a0 = int(sw.items["start"][blk].min())
    a1 = int(sw.items["mesi"][blk].max()) + 2
    inp = sw.d_input[a0:a1].cpu().numpy()
    it = np.zeros(len(blk), O.ITEM_DTYPE)
    for f in ("start", "mosi", "mesi", "write_start"):
        it[f] = sw.items[f][blk] - a0
    for f in ("tol", "m1tol", "set"):
        it[f] = sw.items[f][blk]
    o = O.bm_run(O.default_scalars(FS), O.section_params(), np.full(SECTIONS, 0.5, np.float32), it, inp, a1 - a0,
                 "f32", threads=threads)
    r0 = j * sw.interval - a0
    speeds = o["speeds"][r0:r0 + sw.interval]
    b, _ = stages.load_ac_filter(FS)
    lam = O.ihc_lambda(speeds, sw.interval, sw.block * sw.TB, b, stages.dc_filter(FS), stages.ihc_damage(8.0),
                       stages.aihc_table(), np.array(stages.SPONT_DEFAULT, np.float32), threads=threads)
    return speeds, lam, int(o["steps"].sum())


def c4_main(args):
    """JND sweep in the Sim_Type 4 layout: 16 tones 250-8000 Hz x 10 levels
    10-100 dB x 5 noise conditions (none, white noise at 20/40/60/80 dB), 300 ms
    intervals, plus one reference interval per noise condition; one step =
    input synthesis + BM + IHC/lambda + JND reductions."""
    import torch
    from oracle import oracle as O
    from cochlearmodel_amd import solver, stages
    sw = c4_sweep()
    sc = solver.default_scalars(FS)
    print(f"# JND sweep ready: {sw.n_prof} intervals, {sw.TB} time blocks, {sw.rows} rows", file=sys.stderr,
          flush=True)
    use_graph = not args.no_graph
    if use_graph:
        sw.build_graph()          # IHC/lambda + JND reductions captured as one HIP graph
    for _ in range(args.warmup):
        sw.step(sc, use_graph=use_graph)
    torch.cuda.synchronize()
    times = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        times.append(sw.step(sc, use_graph=use_graph))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    th0 = time.perf_counter()
    res = sw.finish()
    host_ms = 1e3 * (time.perf_counter() - th0)
    steps, passes, sweeps = sw.stats()
    ms_per_step = 1e3 * wall / args.steps
    avg = {k: float(np.mean([t[k] for t in times])) for k in times[0]}
    if use_graph:
        # one eager step (untimed for the value) gives the per-stage split
        eager = sw.step(sc)
        avg["ihc_ms"], avg["jnd_ms"] = eager["ihc_ms"], eager["jnd_ms"]
        sw.release_graph()
    flops = flops_per_launch(steps, passes, sweeps)
    bm_tf = flops / (avg["bm_ms"] * 1e-3) / 1e12
    lam_bytes = sw.rows * SECTIONS * 16.0
    jnd_bytes = sw.n_prof * sw.length * SECTIONS * 24.0       # 3 lambda planes of each body and of its reference
    out = {
        "metric": "section-steps/s (JND sweep)", "value": SECTIONS * steps / (ms_per_step * 1e-3),
        "unit": "section-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "JND sweep: 16 tones log-spaced 250-8000 Hz x 10 levels 10-100 dB SPL x 5 noise "
                               "conditions (none; seeded white noise at 20/40/60/80 dB SPL), 300 ms intervals at "
                               "44.1 kHz, one reference interval per noise condition, Decouple_Filter=15, "
                               "reference kernel-17 method; input synthesis + BM + IHC/lambda + JND per step "
                               "(IHC/lambda + JND replayed as one captured HIP graph unless --no-graph)",
                   "intervals": sw.n_prof, "time_blocks": sw.TB, "rows": sw.rows},
        "stage_ms": avg, "host_final_and_complex_ms": host_ms, "committed_steps": steps,
        "jacobi_iters_per_step": sweeps / steps, "profiles_per_s": sw.n_prof / (ms_per_step * 1e-3),
        "roofline": {"bound": "valu", "engine": "FP32 VALU", "achieved": bm_tf, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": bm_tf / PEAK_FP32_TFLOPS, "traffic": None, "kernel": "bm_kernel",
                     "flop_per_launch": flops},
        "ihc_lambda_roofline": {"bound": "hbm", "achieved": lam_bytes / (avg["ihc_ms"] * 1e-3) / 1e9,
                                "peak": PEAK_HBM_GBS, "unit": "GB/s",
                                "frac": lam_bytes / (avg["ihc_ms"] * 1e-3) / 1e9 / PEAK_HBM_GBS},
        "jnd_roofline": {"bound": "hbm", "achieved": jnd_bytes / (avg["jnd_ms"] * 1e-3) / 1e9, "peak": PEAK_HBM_GBS,
                         "unit": "GB/s", "frac": jnd_bytes / (avg["jnd_ms"] * 1e-3) / 1e9 / PEAK_HBM_GBS,
                         "bytes_per_launch": jnd_bytes},
        "complex_jnd": {"min_method_db": [float(x) for x in res["complex"][0][0]],
                        "gradient_method_db": [float(x) for x in res["complex"][0][1]],
                        "profiles": "noise-major: noise condition x frequency"},
    }
    if not args.no_cpu:
        # oracle chain on sampled calculated intervals and the reference interval of each
        threads = min(args.cpu_threads, os.cpu_count() or 1)
        n_calc = len(sw.calc)
        pick = sw.calc[np.linspace(0, n_calc - 1, args.c4_intervals).round().astype(int)]
        need = sorted(set(int(j) for j in pick) | set(int(sw.ref_index[j]) for j in pick))
        t1 = time.perf_counter()
        cpu, st_total = {}, 0
        for j in need:
            cpu[j] = c4_oracle_interval(sw, j, threads)
            st_total += cpu[j][2]
        nIHC = np.full(SECTIONS, 30000.0 / SECTIONS)
        W = np.array(stages.W_DEFAULT, np.float32)
        db_cpu = {}
        for j in pick:
            r = int(sw.ref_index[j])
            lam2 = np.concatenate([cpu[int(j)][1], cpu[r][1]], axis=1)
            dA = np.array([sw.profiles["dA"][j], sw.profiles["dA"][r]], np.float32)
            _, fra, fai = O.jnd(lam2, 2, sw.interval, sw.head, sw.length, dA, np.array([1, 1], np.int32), nIHC, W,
                                FS, 0.2 * FS)
            ai, rate = stages.jnd_final(fra.reshape(-1), fai.reshape(-1), 2, 2, np.array([1], np.int32),
                                        np.array([0, 0], np.int32))
            db_cpu[int(j)] = (ai[0], rate[0])
        dt = time.perf_counter() - t1
        # one thread on one calculated interval (BM + IHC/lambda): the per-core rate
        t2 = time.perf_counter()
        st1 = c4_oracle_interval(sw, int(pick[0]), 1)[2]
        dt1 = time.perf_counter() - t2
        # GPU rows of the same intervals (the last step's outputs)
        dv, dl, dmax = 0.0, 0.0, 0.0
        nv = nl = 0.0
        for j in need:
            rs = sw.interval_rows(j)
            g = sw.d_speeds[rs].cpu().numpy()
            dv += float(np.sum((g.astype(np.float64) - cpu[j][0]) ** 2))
            nv += float(np.sum(cpu[j][0].astype(np.float64) ** 2))
            gl = sw.d_lambda[:, rs].cpu().numpy()
            dl += float(np.sum((gl.astype(np.float64) - cpu[j][1]) ** 2))
            nl += float(np.sum(cpu[j][1].astype(np.float64) ** 2))
        for j in pick:
            s = int(sw.serial[j])
            dmax = max(dmax, abs(res["ai_db"][s] - db_cpu[int(j)][0]), abs(res["rate_db"][s] - db_cpu[int(j)][1]))
        out["cpu_baseline"] = {"value": SECTIONS * st_total / dt, "unit": "section-steps/s", "cores": threads,
                               "kind": "port", "solver": "reference_20_2",
                               "sample": f"{len(need)} of {sw.n_prof} intervals ({len(pick)} calculated + their "
                                         f"references), BM + IHC/lambda + JND, {st_total} committed steps, "
                                         f"{dt:.2f} s wall, FP32 oracle, OpenMP over blocks and sections",
                               "parity_rel_l2_bm": (dv / max(nv, 1e-300)) ** 0.5,
                               "parity_rel_l2_lambda": (dl / max(nl, 1e-300)) ** 0.5,
                               "parity_max_jnd_db": dmax,
                               "single_thread": {"value": SECTIONS * st1 / dt1,
                                                 "sample": f"interval {int(pick[0])} (BM + IHC/lambda), {st1} committed "
                                                           f"steps, {dt1:.2f} s wall, one thread"},
                               **host_cpu()}
    print(json.dumps(out))


# Held-out tones per launch (eight seeded draws of 256) and launches: 2 x 2048
# = the 4096 held-out tones.  20480 blocks per launch are 1707 lockstep groups
# of 12 on 256 CUs (6.67 per CU, so the last wave of groups leaves a third of
# the chip idle for 1/7 of the launch, against 1/4 with launches of 1024).
C5_PER_BATCH = 2048
C5_BATCHES = 2


def c5_tone_sets(n_batches: int = C5_BATCHES, per: int = C5_PER_BATCH):
    """The held-out evaluation set: draws of 256 tones from seeds EVAL_SEED + 0,
    1, ... (disjoint from the training seeds), concatenated per launch."""
    from cochlearmodel_amd import batch, warmstart
    sub = per // 256
    sets = []
    for j in range(n_batches):
        draws = [batch.random_tones(256, warmstart.EVAL_SEED + sub * j + i) for i in range(sub)]
        sets.append((np.concatenate([d[0] for d in draws]), np.concatenate([d[1] for d in draws])))
    return sets


def c5_accuracy(tbs, per_batch: int, threads: int):
    """Sampled blocks of every batch through the oracle: O2 (exact per-pass
    solve, FP32 state), O2 on the input scaled by 1 + 2^-20 (the FP32
    reproducibility floor of the method) and the FP64 rebuild of the reference
    method, with the reference's float sample clock and with a double one.
    Returns per batch the sampled rows and the four velocity sets."""
    from oracle import oracle as O
    pert = 1.0 + 2.0 ** -20
    res = []
    for tb in tbs:
        sel = tb.sample_blocks(per_batch)
        it = oracle_items(tb, sel)
        inp = tb.d_input.cpu().numpy()
        rows = tb.block_rows(sel)
        sc = O.default_scalars(FS)
        sc.jacobi_mode = 2
        gam = np.full(SECTIONS, 0.5, np.float32)
        o2 = O.bm_run(sc, O.section_params(), gam, it, inp, tb.bufP1, "f32", threads=threads)["speeds"][rows]
        o2p = O.bm_run(sc, O.section_params(), gam, it, inp * np.float32(pert), tb.bufP1, "f32",
                       threads=threads)["speeds"][rows].astype(np.float64) / pert
        o64 = O.bm_run(O.default_scalars(FS), O.section_params(), gam, it, inp, tb.bufP1, "f64",
                       threads=threads)["speeds"][rows]
        o64c = O.bm_run(O.default_scalars(FS), O.section_params(), gam, it, inp, tb.bufP1, "f64c32",
                        threads=threads)["speeds"][rows]
        res.append((rows, o2.astype(np.float64), o2p, o64.astype(np.float64), o64c.astype(np.float64)))
    return res


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(float(np.linalg.norm(b)), 1e-30))


def c5_main(args):
    """Learned warm start against the plain batched solver on the 4096 held-out
    tones (drawn like the 256-tone grid, seeds disjoint from training), 200 ms
    at 44.1 kHz, 256 sections.  The two launches of 2048 tones stay resident
    (profiles, inputs, BM velocity, lambda high/medium/low: about 185 GB of
    HBM with the shared IHC workspace); one step = input synthesis + BM +
    IHC/lambda of one launch, alternating between the two."""
    import torch
    from cochlearmodel_amd import batch
    sets = c5_tone_sets(C5_BATCHES, C5_PER_BATCH)
    tbs, scratch = [], None
    for f, db in sets:
        tb = batch.ToneBatch(f, db, 0.2, d_scratch=scratch)
        scratch = tb.d_scratch
        tbs.append(tb)
    rows0 = tbs[0].rows
    nb = len(tbs)
    # launch order per variant and batch (each launch installs its own):
    # longest block first, then the variant's heavy tones first
    from cochlearmodel_amd import _lib
    orders = [{None: tb.longest_first_order(), "high": tb.tone_order(True), "low": tb.tone_order(False)}
              for tb in tbs]

    def use_orders(name):
        for tb, o in zip(tbs, orders):
            tb.order = o[TONE_ORDER.get(name)]

    use_orders(args.variant)
    model_meta = load_warm_start_model()
    sc = variant_scalars(args.variant)
    set_network_every(args.variant)
    print(f"# held-out set ready: {nb} x {tbs[0].n} tones, {tbs[0].TB} time blocks per launch, variant "
          f"{args.variant}, {torch.cuda.memory_allocated() / 2**30:.1f} GiB resident", file=sys.stderr, flush=True)
    for k in range(args.warmup):
        tbs[k % nb].step(sc)
    torch.cuda.synchronize()
    bm_ms, ihc_ms, in_ms, order = [], [], [], []
    t0 = time.perf_counter()
    for k in range(args.steps):
        t = tbs[k % nb].step(sc)
        bm_ms.append(t["bm_ms"])
        ihc_ms.append(t["ihc_ms"])
        in_ms.append(t["input_ms"])
        order.append(k % nb)
        print(f"# step {k}: batch {k % nb}, BM {t['bm_ms']:.1f} ms", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # every launch of a batch commits the same steps (the solve is deterministic)
    per_batch = [np.array(tb.stats(), np.int64) for tb in tbs]
    tot = sum(per_batch[j] for j in order)
    steps, passes, sweeps = (int(x) for x in tot)
    ms_per_step = 1e3 * wall / args.steps
    value = SECTIONS * steps / wall
    bm_sum = float(np.sum(bm_ms)) * 1e-3
    flops = flops_per_launch(steps, passes, sweeps)
    achieved = flops / bm_sum / 1e12
    mlp_flop = steps * MLP_FLOP_PER_STEP if args.variant.startswith("learned") else 0.0
    t_peak = flops / (PEAK_FP32_TFLOPS * 1e12) + mlp_flop / (PEAK_BF16_TFLOPS * 1e12)
    t_peak_unpacked = flops / (PEAK_FP32_UNPACKED_TFLOPS * 1e12) + mlp_flop / (PEAK_BF16_TFLOPS * 1e12)
    ihc_bytes = rows0 * SECTIONS * 16.0
    out = {
        "metric": "section-steps/s + Jacobi iters/step: learned warm start vs plain @1e-4 rel err",
        "value": value, "unit": "section-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"learned warm start on the 4096 held-out tones ({nb} launches of {tbs[0].n}; "
                               "frequency log-uniform 250-8000 Hz, level uniform 20-90 dB, seeds disjoint from "
                               "the 2048 training tones), 200 ms at 44.1 kHz, 256 sections, healthy profiles, "
                               "Decouple_Filter=10, all four outputs resident; one step = input synthesis + BM + "
                               "IHC/lambda of one launch",
                   "variant": args.variant,
                   "launch_order": {None: "longest block first",
                                    "high": "longest block first, then higher-frequency tones first",
                                    "low": "longest block first, then lower-frequency tones first"}[
                                        TONE_ORDER.get(args.variant)],
                   "signals": nb * tbs[0].n, "signals_per_step": tbs[0].n,
                   "time_blocks_per_step": tbs[0].TB, "rows_per_step": rows0},
        "jacobi_iters_per_step": sweeps / steps, "passes_per_step": passes / steps, "committed_steps": steps,
        "signals_per_s": tbs[0].n * args.steps / wall,
        "bm_ms": float(np.mean(bm_ms)), "ihc_lambda_ms": float(np.mean(ihc_ms)), "input_ms": float(np.mean(in_ms)),
        "bm_section_steps_per_s": SECTIONS * steps / bm_sum,
        "roofline": {"bound": "valu", "engine": "FP32 VALU" + (" + BF16 MFMA" if mlp_flop else ""),
                     "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None,
                     "kernel": "bm_ws_kernel" if mlp_flop else "bm_kernel", "flop_per_launch": flops / args.steps,
                     "peak_without_packed_fp32": PEAK_FP32_UNPACKED_TFLOPS,
                     "combined_frac": t_peak / bm_sum, "combined_frac_without_packed_fp32": t_peak_unpacked / bm_sum,
                     "combined_note": "time the launch's FP32 work would take at the FP32 peak plus its network "
                                      "FLOP at the BF16 MFMA peak, over the kernel time"},
        "ihc_lambda_roofline": {"bound": "hbm", "achieved": ihc_bytes / (np.mean(ihc_ms) * 1e-3) / 1e9,
                                "peak": PEAK_HBM_GBS, "unit": "GB/s",
                                "frac": ihc_bytes / (np.mean(ihc_ms) * 1e-3) / 1e9 / PEAK_HBM_GBS,
                                "fp64_flop_per_launch": rows0 * SECTIONS * 2.0 * (76 + 88)},
    }
    per_launch_s = bm_sum / args.steps
    # counter figures are per launch: only for the launch size they were measured on
    M = C5_MEASURED if C5_MEASURED["tones_per_launch"] == tbs[0].n else {
        "hbm_bytes": {}, "valu_insts": {}, "l2_read_bytes": {}, "mfma_insts": {}}
    if args.variant in M["hbm_bytes"]:
        out["roofline"]["traffic"] = M["hbm_bytes"][args.variant]
        out["roofline"]["traffic_note"] = ("HBM bytes of one launch (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, calibrated "
                                           "on a 1 GiB copy); algorithmic: input 72 MB + velocity 18.5 GB")
        if args.variant == "learned_tol":
            out["roofline"]["traffic_note"] += (
                "; the learned kernel's groups of 16 (4 waves per SIMD, 128 VGPRs) spill 64 B per lane: the "
                "reloads hit the 4 MB L2 of each XCD (2 MiB of spill slots per XCD), and the spill stores make "
                "almost all of the write traffic")
    if args.variant in M["valu_insts"]:
        out["roofline"]["valu_issue"] = {
            "insts_per_launch": M["valu_insts"][args.variant],
            "frac": M["valu_insts"][args.variant] / per_launch_s / VALU_ISSUE_PER_S,
            "capacity_per_s": VALU_ISSUE_PER_S, "capacity_model": VALU_ISSUE_MODEL}
    if "ihc_lambda" in M["hbm_bytes"]:
        out["ihc_lambda_roofline"]["traffic"] = M["hbm_bytes"]["ihc_lambda"]
        out["ihc_lambda_roofline"]["valu_issue"] = {
            "insts_per_launch": M["valu_insts"]["ihc_lambda"],
            "frac": M["valu_insts"]["ihc_lambda"] / (float(np.mean(ihc_ms)) * 1e-3) / VALU_ISSUE_PER_S,
            "capacity_per_s": VALU_ISSUE_PER_S}
    if mlp_flop:
        mlp = mlp_flop / bm_sum / 1e12
        out["mlp_roofline"] = {"bound": "mfma", "engine": "BF16 MFMA", "achieved": mlp, "peak": PEAK_BF16_TFLOPS,
                               "unit": "TFLOP/s", "frac": mlp / PEAK_BF16_TFLOPS,
                               "flop_per_launch": mlp_flop / args.steps}
        if "learned_tol" in M["mfma_insts"]:
            out["mlp_roofline"]["executed_flop_per_launch"] = M["mfma_insts"]["learned_tol"] * 16 * 16 * 32 * 2
            out["mlp_roofline"]["executed_note"] = ("v_mfma_f32_16x16x32_bf16 count x 16384: groups of 16 blocks fill "
                                                    "the 16-row tile, 769 inputs padded to 800")
            out["roofline"]["l2_read"] = {"bytes_per_launch": M["l2_read_bytes"]["learned_tol"],
                                          "achieved": M["l2_read_bytes"]["learned_tol"] / per_launch_s / 1e12,
                                          "unit": "TB/s",
                                          "note": "L1<-L2 reads (TCP_TCC_READ_REQ x 128 B): the network's 1.606 MB "
                                                  "of BF16 weight fragments streamed once per lockstep round of 16 "
                                                  "blocks, and spill reloads"}
            # the two limits the lockstep round alternates between, back to back:
            # VALU issue (solver) and L2->L1 fill (weight stream)
            valu_s = M["valu_insts"]["learned_tol"] / VALU_ISSUE_PER_S
            fill_s = M["l2_read_bytes"]["learned_tol"] / L1_FILL_BYTES_PER_S
            out["roofline"]["serial_limits"] = {
                "valu_issue_s": valu_s, "l1_fill_s": fill_s, "kernel_s": per_launch_s,
                "frac": (valu_s + fill_s) / per_launch_s,
                "note": "VALU instructions at the measured full-occupancy issue rate plus L1<-L2 bytes at 64 B per "
                        "clock per CU, over the kernel time: a round runs the solver and then the network, so the "
                        "sum is what the two limits allow back to back; it can pass 1 because the waves that "
                        "finish their solver first issue the layer's first weight requests while the slowest "
                        "ones still compute, and the issue rate is the fused kernel's, not a hardware ceiling"}
        out["warm_start_model"] = {k: model_meta[k] for k in ("n_signals", "records", "epochs", "final_mse")}
    threads = min(args.cpu_threads, os.cpu_count() or 1)
    if not args.no_variants:
        acc = c5_accuracy(tbs, max(1, args.accuracy_blocks // nb), threads)
        o2 = np.concatenate([a[1] for a in acc])
        floor = _rel(np.concatenate([a[2] for a in acc]), o2)
        o64 = np.concatenate([a[3] for a in acc])
        o64c = np.concatenate([a[4] for a in acc])
        print(f"# oracle on {len(o2)} sampled rows done; FP32 floor {floor:.2e}", file=sys.stderr, flush=True)
        var = {}
        for name in VARIANTS + EXTRA_VARIANTS:
            vsc = variant_scalars(name)
            set_network_every(name)
            use_orders(name)
            ms, st = 0.0, np.zeros(3, np.int64)
            got = []
            for tb, a in zip(tbs, acc):
                ms += tb.run_bm(vsc, True)
                torch.cuda.synchronize()
                st += np.array(tb.stats(), np.int64)
                got.append(tb.sampled_speeds(a[0]).astype(np.float64))
            g = np.concatenate(got)
            var[name] = {"bm_ms": ms, "bm_section_steps_per_s": SECTIONS * int(st[0]) / (ms * 1e-3),
                         "jacobi_iters_per_step": float(st[2] / st[0]), "passes_per_step": float(st[1] / st[0]),
                         "committed_steps": int(st[0]), "rel_l2_vs_exact_solve": _rel(g, o2),
                         "rel_l2_vs_fp64_rebuild_float_clock": _rel(g, o64c),
                         "rel_l2_vs_fp64_rebuild": _rel(g, o64)}
            print(f"# variant {name}: {ms:.1f} ms over {nb} launches, {st[2] / st[0]:.2f} sweeps/step, "
                  f"rel L2 vs O2 {var[name]['rel_l2_vs_exact_solve']:.2e}", file=sys.stderr, flush=True)
        set_network_every(args.variant)
        use_orders(args.variant)
        out["variants"] = var
        lp, pl = var["learned_tol"], var["plain_tol"]
        out["learned_vs_plain"] = {
            "section_steps_per_s_ratio": lp["bm_section_steps_per_s"] / pl["bm_section_steps_per_s"],
            "jacobi_iters_per_step_ratio": lp["jacobi_iters_per_step"] / pl["jacobi_iters_per_step"],
            "committed_steps_ratio": lp["committed_steps"] / pl["committed_steps"],
            "rel_l2_vs_exact_solve": {"learned": lp["rel_l2_vs_exact_solve"], "plain": pl["rel_l2_vs_exact_solve"],
                                      "fp32_floor": floor},
            "rel_l2_vs_fp64_rebuild_float_clock": {"learned": lp["rel_l2_vs_fp64_rebuild_float_clock"],
                                                   "plain": pl["rel_l2_vs_fp64_rebuild_float_clock"],
                                                   "O2": _rel(o2, o64c)},
            "rel_l2_vs_fp64_rebuild": {"learned": lp["rel_l2_vs_fp64_rebuild"],
                                       "plain": pl["rel_l2_vs_fp64_rebuild"],
                                       "O2": _rel(o2, o64)}}
        out["accuracy_sample"] = (f"BM velocity of {sum(len(a[0]) for a in acc)} rows "
                                  f"({max(1, args.accuracy_blocks // nb)} strided time blocks of each launch) against "
                                  "the CPU oracle's exact per-pass pressure solve O2 (FP32 state) and its FP64 "
                                  "rebuilds of the reference method (FP64 state with the reference's float sample "
                                  "clock, and FP64 throughout); fp32_floor = O2 on the input scaled by "
                                  "1 + 2^-20 against O2, the run-to-run reproducibility of FP32 evaluation")
    if not args.no_cpu:
        out["cpu_baseline"] = learned_cpu_baseline(tbs[0], args.cpu_blocks, threads)
        set_network_every(args.variant)
    for tb in tbs:
        tb.order = None
    _lib.check(_lib.lib().cm_set_bm_order(None, 0), "cm_set_bm_order")
    print(json.dumps(out))



def c2_main(args):
    """The 256-tone grid (32 frequencies x 8 levels), one launch per step."""
    import torch
    work = c2_workload(args.signals)
    model_meta = load_warm_start_model()
    sc = variant_scalars(args.variant)
    print(f"# workload ready: {work.n} signals, {work.TB} time blocks, variant {args.variant}", file=sys.stderr,
          flush=True)
    for _ in range(args.warmup):
        work.step(sc)
    torch.cuda.synchronize()
    bm_ms, ihc_ms, in_ms = [], [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        t = work.step(sc)
        bm_ms.append(t["bm_ms"])
        ihc_ms.append(t["ihc_ms"])
        in_ms.append(t["input_ms"])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    print(f"# timed {args.steps} steps in {wall:.2f} s", file=sys.stderr, flush=True)
    steps, passes, sweeps = work.stats()
    ms_per_step = 1e3 * wall / args.steps
    sec_steps = SECTIONS * steps
    value = sec_steps / (ms_per_step * 1e-3)
    bm_avg = float(np.mean(bm_ms))
    flops = flops_per_launch(steps, passes, sweeps)
    achieved = flops / (bm_avg * 1e-3) / 1e12
    ihc_bytes = work.rows * SECTIONS * 16.0        # read BM, write 3 lambdas
    head_speeds = work.d_speeds.cpu().numpy()
    desc = {"reference_20_2": "reference kernel-17 method (fixed 20/2 Jacobi sweeps)",
            "plain_tol": f"Jacobi to relative residual {TOL:g}, previous-pressure guess",
            "extrapolated_tol": f"Jacobi to relative residual {TOL:g}, extrapolated predictor-pass guess with "
                                "residual gate",
            "learned_tol": f"Jacobi to relative residual {TOL:g}, learned predictor-pass guess (769-512-512-256 "
                           "GELU MLP, BF16 MFMA in the solver) with residual gate against the previous and "
                           "extrapolated pressure"}
    out = {
        "metric": "section-steps/s + Jacobi iters/step: learned warm start vs plain @1e-4 rel err",
        "value": value, "unit": "section-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "256 tones (32 freq 250-8000 Hz x 8 levels 20-90 dB), 200 ms at 44.1 kHz, "
                               "256 sections, Decouple_Filter=10, 2560 time blocks, input synthesis + BM + "
                               "IHC/lambda per step; BM solver: " + desc[args.variant],
                   "variant": args.variant, "signals": work.n, "time_blocks": work.TB, "rows": work.rows,
                   "sections_per_lane": 4 if args.variant == "learned_tol" else args.sections_per_lane},
        "jacobi_iters_per_step": sweeps / steps, "passes_per_step": passes / steps,
        "committed_steps": steps, "signals_per_s": work.n / (ms_per_step * 1e-3),
        "bm_ms": bm_avg, "ihc_lambda_ms": float(np.mean(ihc_ms)), "input_ms": float(np.mean(in_ms)),
        "bm_section_steps_per_s": sec_steps / (bm_avg * 1e-3),
        "roofline": {"bound": "valu", "engine": "FP32 VALU", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
                     "traffic": MEASURED_HBM_BYTES.get(args.variant) if args.signals == 256 else None,
                     "kernel": "bm_kernel", "flop_per_launch": flops},
        "ihc_lambda_roofline": {"bound": "hbm", "achieved": ihc_bytes / (np.mean(ihc_ms) * 1e-3) / 1e9,
                                "peak": PEAK_HBM_GBS, "unit": "GB/s",
                                "frac": ihc_bytes / (np.mean(ihc_ms) * 1e-3) / 1e9 / PEAK_HBM_GBS,
                                "traffic": MEASURED_HBM_BYTES["ihc_lambda"] if args.signals == 256 else None,
                                "fp64_flop_per_launch": work.rows * SECTIONS * 2.0 * (76 + 88)},
    }
    if args.signals == 256:
        # the IHC/lambda FIRs run in FP64: besides HBM, the stage is bound by VALU issue
        out["ihc_lambda_roofline"]["valu_issue"] = {
            "insts_per_launch": MEASURED_VALU_INSTS["ihc_lambda"],
            "frac": MEASURED_VALU_INSTS["ihc_lambda"] / (float(np.mean(ihc_ms)) * 1e-3) / VALU_ISSUE_PER_S,
            "capacity_per_s": VALU_ISSUE_PER_S, "capacity_model": VALU_ISSUE_MODEL}
    if args.signals == 256 and args.variant in MEASURED_VALU_INSTS:
        out["roofline"]["valu_issue"] = {
            "insts_per_launch": MEASURED_VALU_INSTS[args.variant],
            "frac": MEASURED_VALU_INSTS[args.variant] / (bm_avg * 1e-3) / VALU_ISSUE_PER_S,
            "capacity_per_s": VALU_ISSUE_PER_S, "capacity_model": VALU_ISSUE_MODEL}
    if args.variant == "learned_tol":
        out["roofline"]["kernel"] = "bm_ws_kernel"
        if args.signals == 256:
            out["roofline"]["l2_read"] = {"bytes_per_launch": MEASURED_L2_READ_BYTES["learned_tol"],
                                          "achieved": MEASURED_L2_READ_BYTES["learned_tol"] / (bm_avg * 1e-3) / 1e12,
                                          "unit": "TB/s"}
        mlp = steps * MLP_FLOP_PER_STEP / (bm_avg * 1e-3) / 1e12
        out["mlp_roofline"] = {"bound": "mfma", "engine": "BF16 MFMA", "achieved": mlp, "peak": PEAK_BF16_TFLOPS,
                               "unit": "TFLOP/s", "frac": mlp / PEAK_BF16_TFLOPS,
                               "flop_per_launch": steps * MLP_FLOP_PER_STEP}
        out["warm_start_model"] = {k: model_meta[k] for k in ("n_signals", "records", "epochs", "final_mse")}
    threads = min(args.cpu_threads, os.cpu_count() or 1)
    if not args.no_variants:
        # every solver variant on the same resident input: BM kernel time, work and accuracy
        rows, exact, fp64 = exact_reference(work, args.accuracy_blocks, threads)
        nrm64 = max(float(np.linalg.norm(fp64)), 1e-30)
        print(f"# exact-solve reference on {args.accuracy_blocks} blocks done", file=sys.stderr, flush=True)
        nrm = max(float(np.linalg.norm(exact)), 1e-30)
        var = {}
        for name in VARIANTS + EXTRA_VARIANTS:
            vsc = variant_scalars(name)
            set_network_every(name)
            ms = work.run_bm(vsc, True)
            set_network_every("learned_tol")
            torch.cuda.synchronize()
            st, pa, sw = work.stats()
            v = work.d_speeds.cpu().numpy()[rows]
            var[name] = {"bm_ms": ms, "bm_section_steps_per_s": SECTIONS * st / (ms * 1e-3),
                         "jacobi_iters_per_step": sw / st, "passes_per_step": pa / st, "committed_steps": st,
                         "rel_l2_vs_exact_solve": float(np.linalg.norm(v - exact) / nrm),
                         "rel_l2_vs_fp64_rebuild": float(np.linalg.norm(v - fp64) / nrm64)}
            print(f"# variant {name}: {ms:.1f} ms, {sw / st:.2f} sweeps/step", file=sys.stderr, flush=True)
        out["variants"] = var
        lp, pl = var["learned_tol"], var["plain_tol"]
        out["learned_vs_plain"] = {
            "section_steps_per_s_ratio": lp["bm_section_steps_per_s"] / pl["bm_section_steps_per_s"],
            "jacobi_iters_per_step_ratio": lp["jacobi_iters_per_step"] / pl["jacobi_iters_per_step"],
            "rel_l2_vs_exact_solve": {"learned": lp["rel_l2_vs_exact_solve"], "plain": pl["rel_l2_vs_exact_solve"]}}
        out["accuracy_sample"] = (f"BM velocity of {args.accuracy_blocks} sampled time blocks ({len(rows)} rows) "
                                  "against the CPU oracle's exact per-pass pressure solve (FP32 state) and against "
                                  "its FP64 rebuild of the reference method")
    if not args.no_cpu and args.variant == "learned_tol":
        out["cpu_baseline"] = learned_cpu_baseline(work, args.cpu_blocks, threads)
    elif not args.no_cpu:
        # the other variants: the oracle with the same solver settings, bit for
        # bit against the GPU run
        cpu_variant = args.variant
        csc = variant_scalars(cpu_variant)
        gpu_speeds = head_speeds
        cb, o, sel = cpu_baseline(work, args.cpu_blocks, threads, csc)
        cb["solver"] = cpu_variant
        cb.update(host_cpu())
        # one thread on a smaller sample: the per-core rate
        cb1, _, _ = cpu_baseline(work, max(1, args.cpu_blocks // 16), 1, csc)
        cb["single_thread"] = {"value": cb1["value"], "sample": cb1["sample"]}
        out["cpu_baseline"] = cb
        rows = block_rows(work, sel)
        d = np.linalg.norm(gpu_speeds[rows] - o["speeds"][rows]) / max(np.linalg.norm(o["speeds"][rows]), 1e-30)
        out["cpu_baseline"]["parity_rel_l2"] = float(d)
    print(json.dumps(out))



def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--signals", type=int, default=256)
    ap.add_argument("--variant", default="learned_tol", choices=VARIANTS)
    ap.add_argument("--cpu-blocks", type=int, default=512)
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--accuracy-blocks", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-variants", action="store_true")
    ap.add_argument("--workload", default="c5", choices=("c1", "c2", "c3", "c4", "c5"),
                    help="c1: one VOICE tone; c2: the 256-tone grid; c3: 1024 uneven signals; c4: JND sweep; "
                         "c5: learned vs plain on the 4096 held-out tones (headline)")
    ap.add_argument("--c3-signals", type=int, default=1024)
    ap.add_argument("--c4-intervals", type=int, default=8, help="calculated intervals checked on the CPU")
    ap.add_argument("--no-graph", action="store_true", help="c4: launch the post-BM stages eagerly")
    ap.add_argument("--sections-per-lane", type=int, default=4, choices=(4, 2, 1),
                    help="block layout of the fused BM kernels (the learned kernel keeps 4)")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "only one GPU is used: runs on several GPUs are not supported",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch
    torch.cuda.set_device(0)
    if args.copy_probe_mb > 0:
        n = args.copy_probe_mb * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    from cochlearmodel_amd import _lib
    _lib.check(_lib.lib().cm_set_bm_sections_per_lane(args.sections_per_lane), "cm_set_bm_sections_per_lane")
    if args.workload == "c1":
        return c1_main(args)
    if args.workload == "c3":
        return c3_main(args)
    if args.workload == "c4":
        return c4_main(args)
    if args.workload == "c5":
        return c5_main(args)
    return c2_main(args)


if __name__ == "__main__":
    main()
