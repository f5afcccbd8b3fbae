"""Uneven batches: many signals of different lengths, each with its own OHC
profile (one gamma row) and IHC value (one damage row), integrated in ONE BM
launch and one IHC/lambda launch.

The reference runs such signals one parameter set at a time
(src/solver.cpp:312, a CModel per OHC_Vector/IHC_Vector): each as a VOICE run
with Time_Block_Length 20 ms and Overlap_Time 5 ms (src/solver.cpp:380-681,
src/cochlea.cu:1148-1176).  Here every signal keeps exactly that run's layout -
its own segment of the input buffer (overlap zeros, the signal, zero padding
to totalTimeNodesP1), its own time blocks, its own row range - so the batched
result equals the separate runs, while the GPU sees all blocks at once.  The
launch shape (waves per workgroup, longest-first order) is chosen by the
library's autotuner (cm_bm_autotune); it never changes the results.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib, solver, stages
from ._lib import ITEM_DTYPE, TUNING_DTYPE, check, lib, ptr

SECTIONS = 256


class UnevenBatch:
    def __init__(self, signals, ohc, ihc, fs: float = 44100.0, block_s: float = 0.02, overlap_s: float = 0.005,
                 prepend_overlap: bool = True, device: str = "cuda"):
        import torch
        self.torch = torch
        self.fs = fs
        n = len(signals)
        ohc = np.ascontiguousarray(np.atleast_2d(ohc), np.float32)
        ihc = np.asarray(ihc, np.float64).reshape(-1)
        assert ohc.shape == (n, SECTIONS) and len(ihc) == n
        lay0 = solver.voice_layout(fs, 1, block_s, overlap_s)
        self.block, self.overlap = lay0.block_nodes, lay0.overlap_nodes
        self.layouts, self.base, self.lengths = [], [], []
        items, ivs = [], []
        base = 0
        pieces = []
        for k, sig in enumerate(signals):
            sig = np.asarray(sig, np.float32)
            lead = self.overlap if prepend_overlap else 0
            tb = max(1, int(math.ceil((lead + len(sig) - self.overlap) / self.block)))
            lay = solver.voice_layout(fs, tb, block_s, overlap_s)
            seg = np.zeros(lay.bufP1, np.float32)
            m = min(len(sig), lay.bufP1 - lead)
            seg[lead:lead + m] = sig[:m]
            pieces.append(seg)
            it = solver.block_table(tb, lay.bufP1, self.overlap, 0)
            for f in ("start", "mosi", "mesi", "write_start"):
                it[f] += base
            it["tol"], it["m1tol"], it["set"] = 1e-10, 1e-15, k      # VOICE, absolute errors (Q23)
            items.append(it)
            # one FIR interval per run, as RunIHCKernel sees a VOICE chunk (Decouple_Filter 0)
            ivs.append((base, lay.total_nodes, 0, lay.total_nodes, lay.total_nodes,
                        lay.block_nodes * lay.time_blocks, k))
            self.layouts.append(lay)
            self.base.append(base)
            self.lengths.append(len(sig))
            base += lay.bufP1
        self.n = n
        self.rows = base
        self.items = np.concatenate(items)
        self.TB = len(self.items)
        self.intervals = np.array(ivs, _lib.IHC_INTERVAL_DTYPE)
        self.input = np.concatenate(pieces)
        dev = device
        self.d_input = torch.from_numpy(np.concatenate([self.input, np.zeros(2, np.float32)])).to(dev)
        self.bm = solver.BatchedBM(ohc, device=dev)
        self.d_items = torch.from_numpy(self.items.view(np.uint8).copy()).to(dev)
        self.d_speeds = torch.zeros((self.rows, SECTIONS), dtype=torch.float32, device=dev)
        self.d_stats = torch.zeros(self.TB * 3, dtype=torch.int64, device=dev)
        self.ihc = stages.BatchedIHC(fs, ihc_vectors=list(ihc), device=dev)
        self.waves_per_block = 1
        self.tuning = None
        self.stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)

    def autotune(self, sc, trial_stride: int = 4):
        """Pick the sections per lane, Jacobi unroll, waves per workgroup and
        launch order on a sample of the blocks; the library keeps the section
        layout, unroll and order for the following launches."""
        L = lib()
        L.cm_set_stream(self.stream)
        out = np.zeros(1, TUNING_DTYPE)
        check(L.cm_bm_autotune(C.byref(sc), ptr(self.d_input), self.d_input.numel(), ptr(self.bm.sec),
                               ptr(self.bm.gamma), ptr(self.items), self.TB, ptr(self.d_speeds), self.rows,
                               trial_stride, ptr(out)), "cm_bm_autotune")
        self.waves_per_block = int(out["waves_per_block"][0])
        self.tuning = out[0]
        return out[0]

    def clear_order(self):
        """Back to the untuned launch order, Jacobi unroll and section layout."""
        check(lib().cm_set_bm_order(None, 0), "cm_set_bm_order")
        check(lib().cm_set_bm_jacobi_unroll(2), "cm_set_bm_jacobi_unroll")
        check(lib().cm_set_bm_sections_per_lane(4), "cm_set_bm_sections_per_lane")

    def run_bm(self, sc, timed: bool = True, waves_per_block: int | None = None):
        L = lib()
        L.cm_set_stream(self.stream)
        wpb = self.waves_per_block if waves_per_block is None else waves_per_block
        check(L.cm_bm_run(C.byref(sc), ptr(self.d_input), self.d_input.numel(), ptr(self.bm.sec), ptr(self.bm.gamma),
                          ptr(self.d_items), self.TB, ptr(self.d_speeds), self.rows, None, None, None, None, None,
                          None, None, None, ptr(self.d_stats), wpb, 0, 1 if timed else 0), "cm_bm_run")
        return L.cm_last_kernel_ms() if timed else None

    def run_lambda(self, d_lambda=None, d_scratch=None):
        """lambda [3][rows][256] of every signal (device tensor)."""
        lam, _ = self.ihc.run(self.d_speeds, self.rows, self.intervals, d_lambda=d_lambda, d_scratch=d_scratch)
        return lam

    def stats(self):
        s = self.d_stats.view(self.TB, 3).cpu().numpy().astype(np.int64)
        return int(s[:, 0].sum()), int(s[:, 1].sum()), int(s[:, 2].sum())

    def signal_rows(self, k):
        """Rows of signal k that its own VOICE run emits (overlap discarded)."""
        b, lay = self.base[k], self.layouts[k]
        return slice(b + lay.overlap_nodes, b + lay.total_nodes)
