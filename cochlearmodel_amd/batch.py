"""A device-resident batch of tone profiles in the reference's profile layout
(Sim_Type 3: one interval per tone, Decouple_Filter = interval / 20 ms block,
src/cochlea.cu:1148-1176), driven through the C-ABI: input synthesis
(generateInputFromProfile semantics), the BM solve and IHC/lambda.

torch tensors hold device memory and the stream; all compute is the library's.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib, solver, stages, workloads

FS = workloads.FS
SECTIONS = 256


class ToneBatch:
    def __init__(self, freqs, levels, duration: float = 0.2, with_lambda: bool = True, wn_seed_offset: int = 2,
                 d_scratch=None):
        import torch
        self.torch = torch
        self.L = _lib.lib()
        freqs = np.asarray(freqs, np.float64)
        levels = np.asarray(levels, np.float64)
        self.n = len(freqs)
        self.freqs, self.levels = freqs, levels
        self.interval = int(round(FS * duration))                  # JND_Interval_Nodes_Full
        self.block = int(math.ceil(np.float32(0.02) * np.float32(FS)))
        self.DF = self.interval // self.block
        self.TB = self.n * self.DF
        self.ov = 220
        self.bufP1 = self.ov + self.block * (self.TB + 1)
        self.rows = self.n * self.interval
        self.head = int(round(FS * 0.014))
        prof = tone_profiles(freqs, levels)
        self.profiles = prof
        rng = np.random.default_rng(workloads.MASTER_SEED + wn_seed_offset)
        WN = rng.standard_normal(self.interval).astype(np.float32)
        e1 = int((duration - 0.002) * FS)
        n4 = np.zeros(4, np.float64)
        _lib.check(self.L.cm_input_normalization(_lib.ptr(WN), len(WN), None, 0, 0, 0, 0, 0.16, 0.0, 0, e1, 0, e1,
                                                 C.c_float(FS), _lib.ptr(n4)), "normalization")
        self.norms = n4
        dev = "cuda"
        self.d_prof = torch.from_numpy(prof.view(np.uint8).copy()).to(dev)
        self.d_WN = torch.from_numpy(WN).to(dev)
        self.d_input = torch.zeros(self.bufP1 + 2, dtype=torch.float32, device=dev)
        items = solver.block_table(self.TB, self.bufP1, self.ov, self.DF)
        items["tol"], items["m1tol"] = 1e-10, 1e-15          # setupToleranceProfile, absolute errors
        self.items = items
        self.bm = solver.BatchedBM(np.full(SECTIONS, 0.5, np.float32))
        self.d_items = torch.from_numpy(items.view(np.uint8).copy()).to(dev)
        self.d_speeds = torch.zeros((self.bufP1, SECTIONS), dtype=torch.float32, device=dev)
        self.d_stats = torch.zeros(self.TB * 3, dtype=torch.int64, device=dev)
        self.with_lambda = with_lambda
        self.order = None   # launch order installed before each BM launch of this batch (cm_set_bm_order)
        if with_lambda:
            self.ihc = stages.BatchedIHC(FS)
            self.fir_ivs, _ = stages.reference_intervals(self.rows, self.block, self.TB, self.DF)
            self.d_lambda = torch.empty((3, self.rows, SECTIONS), dtype=torch.float32, device=dev)
            # the IHC/lambda AC/dS workspace; batches run one after another may share one
            if d_scratch is None:
                d_scratch = torch.empty((2, self.rows, SECTIONS), dtype=torch.float32, device=dev)
            if d_scratch.numel() < 2 * self.rows * SECTIONS:
                raise ValueError("d_scratch holds fewer than 2 x rows x 256 floats")
            self.d_scratch = d_scratch
        self.stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)

    def device_profiles(self, freqs, levels):
        """Another set of n tones' profiles, resident on the device, for
        generate_input(d_prof=...): the same batch geometry, other signals."""
        assert len(freqs) == self.n and len(levels) == self.n
        return self.torch.from_numpy(tone_profiles(freqs, levels).view(np.uint8).copy()).to("cuda")

    def generate_input(self, d_prof=None):
        L, lib = self.L, _lib
        L.cm_set_stream(self.stream)
        d_prof = self.d_prof if d_prof is None else d_prof
        lib.check(L.cm_generate_input_run(lib.ptr(d_prof), lib.ptr(self.d_WN), None, lib.ptr(self.d_input),
                                          self.norms[0], self.norms[1], 0, self.norms[2], self.norms[3], 0, self.n,
                                          self.interval, self.head, C.c_float(FS)), "input")

    def run_bm(self, sc, timed: bool = True):
        L, lib = self.L, _lib
        L.cm_set_stream(self.stream)
        if self.order is not None:
            lib.check(L.cm_set_bm_order(lib.ptr(self.order), len(self.order)), "order")
        lib.check(L.cm_bm_run(C.byref(sc), lib.ptr(self.d_input), self.d_input.numel(), lib.ptr(self.bm.sec),
                              lib.ptr(self.bm.gamma), lib.ptr(self.d_items), self.TB, lib.ptr(self.d_speeds),
                              self.bufP1, None, None, None, None, None, None, None, None, lib.ptr(self.d_stats), 1, 0,
                              1 if timed else 0), "bm")
        return L.cm_last_kernel_ms() if timed else None

    def step(self, sc, timed_bm: bool = True):
        """Input synthesis + BM + IHC/lambda; per-stage times in ms."""
        torch = self.torch
        t = {}
        e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
        e0.record()
        self.generate_input()
        e1.record()
        ms = self.run_bm(sc, timed_bm)
        if timed_bm:
            t["bm_ms"] = ms
        e2.record()
        if self.with_lambda:
            self.ihc.run(self.d_speeds, self.rows, self.fir_ivs, d_lambda=self.d_lambda, d_scratch=self.d_scratch)
        e3 = torch.cuda.Event(enable_timing=True)
        e3.record()
        e3.synchronize()
        t["input_ms"] = e0.elapsed_time(e1)
        t["ihc_ms"] = e2.elapsed_time(e3)
        return t

    def longest_first_order(self) -> np.ndarray:
        """Launch order of the block table, longest block first (ties in table
        order): the last block of every tone is the shortest (it ends at the
        tone's boundary without the overlap run-out), so in natural order each
        lockstep group mixes lengths and the last wave of work holds
        full-length blocks.  For cm_set_bm_order; results do not depend
        on it."""
        lens = self.items["mesi"].astype(np.int64) - self.items["start"]
        return np.argsort(-lens, kind="stable").astype(np.int32)

    def tone_order(self, high_first: bool = True) -> np.ndarray:
        """Longest block first, then the blocks of higher-frequency tones first
        (low-frequency first with high_first=False; ties in table order).  A
        tone's Jacobi sweeps per step follow its frequency closely: with a
        guess (extrapolated or learned) they rise with frequency, without one
        (plain) they fall.  Taking the heavy tones first deals the heaviest
        work out first, and lockstep groups of neighbouring frequencies wait
        less for their slowest block at each round.  For cm_set_bm_order;
        results do not depend on it."""
        lens = self.items["mesi"].astype(np.int64) - self.items["start"]
        tone = np.arange(self.TB) // self.DF
        key = -self.freqs[tone] if high_first else self.freqs[tone]
        return np.lexsort((np.arange(self.TB), key, -lens)).astype(np.int32)

    def stats(self):
        s = self.d_stats.view(self.TB, 3).cpu().numpy().astype(np.int64)
        return int(s[:, 0].sum()), int(s[:, 1].sum()), int(s[:, 2].sum())

    # -- helpers for sampled CPU checks ----------------------------------
    def sample_blocks(self, n_blocks):
        stride = max(1, self.TB // n_blocks)
        return np.arange(0, self.TB, stride)[:n_blocks]

    def sampled_speeds(self, rows):
        """BM velocity rows (host int array) copied from the device, without
        copying the rest of the batch."""
        idx = self.torch.from_numpy(np.asarray(rows, np.int64)).to(self.d_speeds.device)
        return self.d_speeds.index_select(0, idx).cpu().numpy()

    def block_rows(self, sel):
        ws, me = self.items["write_start"][sel], self.items["mesi"][sel]
        return np.concatenate([np.arange(a, b) for a, b in zip(ws, me)])


def tone_profiles(freqs, levels) -> np.ndarray:
    """device_jnd_params of one tone per interval: peak dA = 10 SPLRef 10^(L/20)
    (Sim_Type 3 profiles without references)."""
    freqs = np.asarray(freqs, np.float64)
    levels = np.asarray(levels, np.float64)
    prof = np.zeros(len(freqs), _lib.PROFILE_DTYPE)
    prof["dA"] = 10.0 * 10.0 ** (levels / 20.0) * 2e-5
    prof["recipdA"] = 1.0 / prof["dA"]
    prof["dBSPLSignal"] = levels
    prof["frequency"] = freqs
    prof["referenceIndex"] = -1
    prof["calculateIndexPosition"] = np.arange(len(freqs))
    return prof


def random_tones(n: int, seed: int, f_lo: float = 250.0, f_hi: float = 8000.0, db_lo: float = 20.0,
                 db_hi: float = 90.0):
    """Tones drawn like the 256-tone grid: frequency log-uniform, level uniform."""
    rng = np.random.default_rng(seed)
    f = np.exp(rng.uniform(np.log(f_lo), np.log(f_hi), n))
    db = rng.uniform(db_lo, db_hi, n)
    return f, db
