"""A device-resident JND sweep in the reference's Sim_Type 4 layout, through the
library's batched entry points: profiles (CParams::updateInputProfile,
src/params.cpp:661-782) -> input synthesis (generateInputFromProfile
semantics) -> BM solve over decoupled intervals -> IHC/lambda -> JND
reductions (CudaCalculateJND) -> final dB (calcJNDFinal) -> complex-JND
aggregation per (frequency, noise) profile (calcComplexJND).

Every tested (noise, level, frequency) is one interval of `duration` seconds;
each noise condition adds one reference interval (Generate_One_Ref_Per_Noise_Level).
Intervals are decoupled (Decouple_Filter = interval / 20 ms block), so each
interval's BM blocks start from rest, as in the reference's profile runs.
torch holds device memory and the stream; all compute is the library's.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib, solver, stages, workloads

FS = workloads.FS
SECTIONS = 256
HEAD_S, TAIL_S = 0.014, 0.002            # JND_BLOCK_HEAD / JND_BLOCK_TAIL, include/const.h:348-349
DELTA_ALPHA_TIME_FACTOR = 0.2            # Const_JND_Delta_Alpha_Time_Factor, include/const.h:353
N_IHC = 30000.0 / SECTIONS               # nerve fibres per section (M_tot / SECTIONS)


class JndSweep:
    def __init__(self, freqs, levels, noises, duration: float = 0.3, wn_seed: int = workloads.MASTER_SEED + 4,
                 methods: int = 1, eps: float = 1.0, device: str = "cuda"):
        import torch
        self.torch = torch
        self.L = _lib.lib()
        self.freqs, self.levels, self.noises = list(freqs), list(levels), list(noises)
        self.methods, self.eps = methods, eps
        prof, refs, serial, calc, to_ref = stages.jnd_profiles(self.freqs, self.levels, noises=self.noises)
        self.profiles, self.refs, self.serial, self.calc, self.to_ref = prof, refs, serial, calc, to_ref
        self.n_prof = len(prof)
        self.interval = int(round(FS * duration))                 # JND_Interval_Nodes_Full
        self.block = int(math.ceil(np.float32(0.02) * np.float32(FS)))
        self.DF = self.interval // self.block
        if self.DF * self.block != self.interval:
            raise ValueError("the interval must hold a whole number of 20 ms blocks")
        self.TB = self.n_prof * self.DF
        self.ov = 220
        self.bufP1 = self.ov + self.block * (self.TB + 1)
        self.rows = self.n_prof * self.interval
        self.head = int(round(FS * HEAD_S))                        # JND_Interval_Nodes_Offset
        self.length = int(round(FS * (duration - HEAD_S - TAIL_S)))  # JND_Interval_Nodes_Length
        rng = np.random.default_rng(wn_seed)
        self.WN = rng.standard_normal(self.interval).astype(np.float32)
        e1 = int((duration - TAIL_S) * FS)
        n4 = np.zeros(4, np.float64)
        _lib.check(self.L.cm_input_normalization(_lib.ptr(self.WN), len(self.WN), None, 0, 0, 0, 0, 0.16, 0.0, 0, e1,
                                                 0, e1, C.c_float(FS), _lib.ptr(n4)), "cm_input_normalization")
        self.norms = n4
        dev = device
        self.d_prof = torch.from_numpy(prof.view(np.uint8).copy()).to(dev)
        self.d_WN = torch.from_numpy(self.WN).to(dev)
        self.d_input = torch.zeros(self.bufP1 + 2, dtype=torch.float32, device=dev)
        items = solver.block_table(self.TB, self.bufP1, self.ov, self.DF)
        items["tol"], items["m1tol"] = 1e-10, 1e-15             # setupToleranceProfile, absolute errors
        self.items = items
        self.bm = solver.BatchedBM(np.full(SECTIONS, 0.5, np.float32))
        self.d_items = torch.from_numpy(items.view(np.uint8).copy()).to(dev)
        self.d_speeds = torch.zeros((self.bufP1, SECTIONS), dtype=torch.float32, device=dev)
        self.d_stats = torch.zeros(self.TB * 3, dtype=torch.int64, device=dev)
        self.d_failed_blocks = torch.zeros(self.TB, dtype=torch.int32, device=dev)
        self.ihc = stages.BatchedIHC(FS)
        self.fir_ivs, _ = stages.reference_intervals(self.rows, self.block, self.TB, self.DF)
        self.d_lambda = torch.empty((3, self.rows, SECTIONS), dtype=torch.float32, device=dev)
        self.d_scratch = torch.empty((2, self.rows, SECTIONS), dtype=torch.float32, device=dev)
        # JND tables: every interval is computed, references against themselves
        self.ref_index = stages.reference_index(self.n_prof, refs, serial, to_ref)
        self.d_dA = torch.from_numpy(prof["dA"].astype(np.float32)).to(dev)
        self.d_ref_index = torch.from_numpy(self.ref_index.astype(np.int32)).to(dev)
        self.d_out_index = torch.arange(self.n_prof, dtype=torch.int32, device=dev)
        self.d_nIHC = torch.full((SECTIONS,), N_IHC, dtype=torch.float64, device=dev)
        self.W3 = (C.c_float * 3)(*stages.W_DEFAULT)
        self.d_mean_rate = torch.empty((3, self.n_prof, SECTIONS), dtype=torch.float32, device=dev)
        self.d_F_RA = torch.zeros((3, self.n_prof), dtype=torch.float32, device=dev)
        self.d_FAI = torch.zeros((3, self.n_prof), dtype=torch.float32, device=dev)
        self.stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)

    # -- stages --------------------------------------------------------------
    def generate_input(self):
        L, lib = self.L, _lib
        L.cm_set_stream(self.stream)
        lib.check(L.cm_generate_input_run(lib.ptr(self.d_prof), lib.ptr(self.d_WN), None, lib.ptr(self.d_input),
                                          self.norms[0], self.norms[1], 0, self.norms[2], self.norms[3], 0,
                                          self.n_prof, self.interval, self.head, C.c_float(FS)), "input")

    def run_bm(self, sc, timed: bool = True):
        L, lib = self.L, _lib
        L.cm_set_stream(self.stream)
        lib.check(L.cm_bm_run(C.byref(sc), lib.ptr(self.d_input), self.d_input.numel(), lib.ptr(self.bm.sec),
                              lib.ptr(self.bm.gamma), lib.ptr(self.d_items), self.TB, lib.ptr(self.d_speeds),
                              self.bufP1, None, None, None, None, lib.ptr(self.d_failed_blocks), None, None, None,
                              lib.ptr(self.d_stats), 1, 0, 1 if timed else 0), "bm")
        return L.cm_last_kernel_ms() if timed else None

    def run_lambda(self):
        self.ihc.run(self.d_speeds, self.rows, self.fir_ivs, d_lambda=self.d_lambda, d_scratch=self.d_scratch)

    def run_jnd(self):
        L, lib = self.L, _lib
        L.cm_set_stream(self.stream)
        lib.check(L.cm_jnd_run(lib.ptr(self.d_lambda), self.rows, self.n_prof, self.interval, self.head, self.length,
                               lib.ptr(self.d_dA), lib.ptr(self.d_ref_index), lib.ptr(self.d_out_index), self.n_prof,
                               lib.ptr(self.d_nIHC), self.W3, FS, DELTA_ALPHA_TIME_FACTOR * FS, 1, 1,
                               lib.ptr(self.d_mean_rate), lib.ptr(self.d_F_RA), lib.ptr(self.d_FAI)), "cm_jnd_run")

    # -- the post-BM stages as one HIP graph --------------------------------
    def build_graph(self):
        """Capture IHC/lambda + JND reductions once (cm_post_bm_graph_create)."""
        L, lib = self.L, _lib
        ivs = np.ascontiguousarray(self.fir_ivs, _lib.IHC_INTERVAL_DTYPE)
        h = C.c_void_p()
        lib.check(L.cm_post_bm_graph_create(
            lib.ptr(self.d_speeds), self.rows, lib.ptr(ivs), len(ivs), lib.ptr(self.ihc.b), self.ihc.n_ac,
            lib.ptr(self.ihc.dc), self.ihc.n_dc, lib.ptr(self.ihc.damage), lib.ptr(self.ihc.A), self.ihc.spont,
            stages.LAMBDA_SAT, stages.ETA_AC, stages.ETA_DC, 1.0, lib.ptr(self.d_scratch), lib.ptr(self.d_lambda),
            self.n_prof, self.interval, self.head, self.length, lib.ptr(self.d_dA), lib.ptr(self.d_ref_index),
            lib.ptr(self.d_out_index), self.n_prof, lib.ptr(self.d_nIHC), self.W3, FS, DELTA_ALPHA_TIME_FACTOR * FS,
            1, 1, lib.ptr(self.d_mean_rate), lib.ptr(self.d_F_RA), lib.ptr(self.d_FAI), C.byref(h)),
            "cm_post_bm_graph_create")
        self.graph = h

    def run_graph(self):
        self.L.cm_set_stream(self.stream)
        _lib.check(self.L.cm_post_bm_graph_launch(self.graph), "cm_post_bm_graph_launch")

    def release_graph(self):
        if getattr(self, "graph", None) is not None:
            self.L.cm_post_bm_graph_destroy(self.graph)
            self.graph = None

    def finish(self):
        """Final dB and failed-signal counts per calculated interval (host, as
        calcJNDFinal) and the complex-JND aggregation (rate JND,
        Calculate_From_Mean_Rate)."""
        F_RA = self.d_F_RA.cpu().numpy().reshape(-1)
        FAI = self.d_FAI.cpu().numpy().reshape(-1)
        ai, rate = stages.jnd_final(F_RA, FAI, self.n_prof, self.n_prof, self.refs, self.serial)
        n_calc = len(self.calc)
        failed = stages.jnd_failed_signals(self.d_failed_blocks.cpu().numpy(), self.n_prof, self.refs, self.serial,
                                           self.DF, n_calc)
        final = stages.complex_jnd(rate[:n_calc], len(self.freqs), len(self.levels), len(self.noises), self.methods,
                                   self.eps, failed)
        return {"ai_db": ai[:n_calc], "rate_db": rate[:n_calc], "failed_signals": failed, "complex": final}

    def step(self, sc, timed_bm: bool = True, use_graph: bool = False):
        """Input synthesis + BM + IHC/lambda + JND reductions; per-stage ms.
        With use_graph the post-BM stages are one replay of the captured graph
        (build_graph first)."""
        torch = self.torch
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(5)]
        t = {}
        ev[0].record()
        self.generate_input()
        ev[1].record()
        ms = self.run_bm(sc, timed_bm)
        ev[2].record()
        if use_graph:
            self.run_graph()
            ev[3].record()
        else:
            self.run_lambda()
            ev[3].record()
            self.run_jnd()
        ev[4].record()
        ev[4].synchronize()
        t["input_ms"] = ev[0].elapsed_time(ev[1])
        t["bm_ms"] = ms if timed_bm else ev[1].elapsed_time(ev[2])
        if use_graph:
            t["post_bm_graph_ms"] = ev[2].elapsed_time(ev[4])
        else:
            t["ihc_ms"] = ev[2].elapsed_time(ev[3])
            t["jnd_ms"] = ev[3].elapsed_time(ev[4])
        return t

    def stats(self):
        s = self.d_stats.view(self.TB, 3).cpu().numpy().astype(np.int64)
        return int(s[:, 0].sum()), int(s[:, 1].sum()), int(s[:, 2].sum())

    # -- sampled CPU checks --------------------------------------------------
    def interval_blocks(self, j):
        return np.arange(j * self.DF, (j + 1) * self.DF)

    def interval_rows(self, j):
        return slice(j * self.interval, (j + 1) * self.interval)
