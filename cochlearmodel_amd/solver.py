"""Host-side mirror of the reference's BM stage (CSolver::Run_Cuda,
src/solver.cpp:234-962) over the C-ABI.

Two ways in:
  * ``run_bm_voice`` follows the reference call sequence for a VOICE input
    (BMOHCKernel_Init -> BMOHCNewKernel -> BMOHCKernel_Copy_Results) with host
    buffers, as src/solver.cpp:380-681 does;
  * ``BatchedBM`` drives ``cm_bm_run`` on device-resident buffers (torch
    tensors as plumbing) for batched and persistent launches.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from ._lib import BMScalars, ITEM_DTYPE, STATS_DTYPE, SECTIONS, check, lib, ptr


def model_vectors() -> np.ndarray:
    out = np.zeros((7, SECTIONS), np.float64)
    check(lib().cm_model_vectors(ptr(out)), "cm_model_vectors")
    return out


def tri_matrix():
    lo, mid, up = (np.zeros(SECTIONS, np.float64) for _ in range(3))
    check(lib().cm_tri_matrix(ptr(lo), ptr(mid), ptr(up)), "cm_tri_matrix")
    return lo, mid, up


def section_params() -> np.ndarray:
    out = np.zeros((9, SECTIONS), np.float32)
    check(lib().cm_pack_section_params(ptr(out)), "cm_pack_section_params")
    return out


def default_scalars(Fs: float = 44100.0, j_fast: int = 2, j_slow: int = 20, max_passes: int = 10) -> BMScalars:
    sc = BMScalars()
    check(lib().cm_default_bm_scalars(C.c_float(Fs), C.byref(sc)), "cm_default_bm_scalars")
    sc.j_fast, sc.j_slow, sc.max_passes = j_fast, j_slow, max_passes
    return sc


def expand_vector(v, size: int = SECTIONS) -> np.ndarray:
    """expandVectorToSize (include/cvector.cpp:609-624): each entry repeated in a block."""
    v = np.atleast_1d(np.asarray(v, np.float32))
    times, rem = divmod(size, len(v))
    out = []
    for i, x in enumerate(v):
        out.extend([x] * (times + (1 if i < rem else 0)))
    return np.asarray(out, np.float32)


@dataclass
class VoiceLayout:
    """Buffer sizes of a VOICE run (include/params.h:277-395, src/solver.cpp:1053-1093)."""
    Fs: float
    time_blocks: int
    block_nodes: int
    overlap_nodes: int

    @property
    def bufP1(self) -> int:          # totalTimeNodesP1
        return self.overlap_nodes + self.block_nodes * (self.time_blocks + 1)

    @property
    def total_nodes(self) -> int:    # totalTimeNodes
        return self.overlap_nodes + self.block_nodes * self.time_blocks

    @property
    def input_buffer_nodes(self) -> int:   # inputBufferTimeNodes
        return self.block_nodes * (self.time_blocks + 1) + 2 * self.overlap_nodes


def voice_layout(Fs: float, time_blocks: int, time_block_length: float, overlap_time: float = 0.005) -> VoiceLayout:
    block = int(math.ceil(np.float32(time_block_length) * np.float32(Fs)))
    overlap_us = int(round(overlap_time * 1000) * 1000)
    overlap = int(overlap_us * Fs / 1000000)
    return VoiceLayout(Fs, time_blocks, block, overlap)


def block_table(time_blocks: int, bufP1: int, overlap: int, decouple_filter: int) -> np.ndarray:
    """Time-block table of BMOHCNewKernel (src/cochlea.cu:1148-1176) as computed
    by the library for a launch; tolerances left at zero."""
    items = np.zeros(time_blocks, ITEM_DTYPE)
    I8 = (bufP1 - overlap) // (time_blocks + 1)
    DF = decouple_filter
    prev_end = 0
    for b in range(time_blocks):
        transient = int(b > 0 and (DF == 0 or (DF != 1 and b % DF != 0)))
        pre = int(DF > 0 and (DF == 1 or (b + 1) % DF == 0) and b != time_blocks - 1)
        post = int(DF > 0 and b % DF == 0)
        in0 = b * I8
        so = in0 + max(1 - transient - post, 0) * overlap
        eo = in0 + I8 + (1 - pre) * overlap
        if b > 0:
            so = max(so, prev_end)
        first = max(in0, so - 1)
        if b > 0 and first < prev_end:
            first = prev_end
        items[b] = (in0, so, eo, first, 0.0, 0.0, 0, 0)
        prev_end = eo
    return items


class BatchedBM:
    """Device-resident batched BM launches through cm_bm_run (torch tensors are
    only the memory/stream plumbing)."""

    def __init__(self, sets_gamma: np.ndarray, device: str = "cuda"):
        import torch
        self.torch = torch
        self.device = device
        self.sec = torch.from_numpy(section_params().reshape(-1)).to(device)
        self.gamma = torch.from_numpy(np.ascontiguousarray(np.atleast_2d(sets_gamma), np.float32).reshape(-1)).to(
            device)

    def run(self, sc: BMScalars, d_input, items: np.ndarray, rows: int, waves_per_block: int = 1,
            persistent: int = 0, timed: bool = False, diagnostics: bool = False, d_speeds=None):
        torch = self.torch
        dev = self.device
        items_d = torch.from_numpy(np.ascontiguousarray(items).view(np.uint8).copy()).to(dev)
        n = len(items)
        if d_speeds is None:
            d_speeds = torch.zeros((rows, SECTIONS), dtype=torch.float32, device=dev)
        stats = torch.zeros(n * 3, dtype=torch.int64, device=dev)
        diag = {}
        if diagnostics:
            diag = dict(failed_nodes=torch.zeros(rows, dtype=torch.int32, device=dev),
                        conv_time=torch.zeros(rows, dtype=torch.float32, device=dev),
                        conv_loops=torch.zeros(rows, dtype=torch.float32, device=dev),
                        conv_jacobi=torch.zeros(rows, dtype=torch.float32, device=dev),
                        failed_blocks=torch.zeros(n, dtype=torch.int32, device=dev),
                        conv_time_blocks=torch.zeros(n, dtype=torch.float32, device=dev),
                        conv_loops_blocks=torch.zeros(n, dtype=torch.float32, device=dev),
                        conv_jacobi_blocks=torch.zeros(n, dtype=torch.float32, device=dev))
        g = lambda k: ptr(diag.get(k))
        L = lib()
        L.cm_set_stream(C.c_void_p(torch.cuda.current_stream().cuda_stream))
        rc = L.cm_bm_run(C.byref(sc), ptr(d_input), d_input.numel(), ptr(self.sec), ptr(self.gamma), ptr(items_d), n,
                         ptr(d_speeds), rows, g("failed_nodes"), g("conv_time"), g("conv_loops"), g("conv_jacobi"),
                         g("failed_blocks"), g("conv_time_blocks"), g("conv_loops_blocks"), g("conv_jacobi_blocks"),
                         ptr(stats), waves_per_block, persistent, 1 if timed else 0)
        check(rc, "cm_bm_run")
        out = {"speeds": d_speeds, "stats": stats.view(n, 3), **diag, "_items": items_d}
        if timed:
            out["kernel_ms"] = L.cm_last_kernel_ms()
        return out


def run_bm_voice(signal: np.ndarray, layout: VoiceLayout, ohc=0.5, max_m1_sp_error=1e-15, max_tolerance=1e-10,
                 relative_error=0, m1_fix=0.0, tol_fix=0.0, spl_ref=2e-5, j_fast=2, j_slow=20, max_passes=10,
                 decouple_filter=0, timed=False, boundary_conditions=1):
    """Reference call sequence for one VOICE chunk; returns speeds [bufP1, 256] and diagnostics.
    boundary_conditions 0 takes the block tolerances from a preceding
    setupToleranceProfile instead of computing them from the input."""
    L = lib()
    sec = section_params()
    lo, mid, up = tri_matrix()
    gamma = expand_vector(ohc)
    in_nodes = layout.input_buffer_nodes
    res_nodes = in_nodes * SECTIONS
    f = lambda a: ptr(np.ascontiguousarray(a, np.float32))
    keep = [np.ascontiguousarray(x, np.float32) for x in (gamma, sec[0], sec[1], up, lo, sec[2], sec[3], sec[4],
                                                          sec[5], sec[6], sec[7])]
    check(L.BMOHCKernel_Init(*[ptr(k) for k in keep], C.c_float(1), C.c_float(0), in_nodes, res_nodes,
                             res_nodes * 3, 1, 0, 0), "BMOHCKernel_Init")
    bufP1 = layout.bufP1
    inp = np.zeros(bufP1, np.float32)
    n = min(len(signal), bufP1)
    inp[:n] = np.asarray(signal[:n], np.float32)
    sc = default_scalars(layout.Fs, j_fast, j_slow, max_passes)
    rc = L.BMOHCNewKernel(ptr(inp), 0, sc.w_ohc, sc.time_step, sc.time_step_out, sc.delta_x, sc.alpha_r, 1, 1, 0,
                          sc.Ts, sc.alpha_l, sc.alpha_s, sc.Gme, sc.a0, sc.a1, sc.a2, sc.sigma_ow, sc.eta1, sc.eta2,
                          layout.time_blocks, bufP1, layout.overlap_nodes,
                          int(round(layout.overlap_nodes / layout.Fs * 1e6)), 1, sc.dt_max, sc.dt_min,
                          decouple_filter, C.c_float(max_m1_sp_error), C.c_float(max_tolerance), relative_error,
                          boundary_conditions,
                          C.c_float(m1_fix), C.c_float(tol_fix), C.c_float(spl_ref), 0, 0, 1 if timed else 0,
                          j_fast, j_slow, max_passes, 17, 2)
    check(rc, "BMOHCNewKernel")
    speeds = np.zeros((bufP1, SECTIONS), np.float32)
    check(L.BMOHCKernel_Copy_Results(ptr(speeds), bufP1 * SECTIONS, 0), "BMOHCKernel_Copy_Results")
    failed = np.zeros(bufP1, np.int32)
    check(L.GeneralKernel_Copy_Results_Bytes(ptr(failed), L.getCudaFailedTimeNodes(), failed.nbytes, 0), "copy")
    conv_loops = np.zeros(bufP1, np.float32)
    check(L.GeneralKernel_Copy_Results_Bytes(ptr(conv_loops), L.getCudaConvergedTimeNodes(), conv_loops.nbytes, 0),
          "copy")
    nb = layout.time_blocks
    blocks = np.zeros(nb, ITEM_DTYPE)
    L.cm_get_last_blocks(ptr(blocks), nb)
    stats = np.zeros(nb, STATS_DTYPE)
    L.cm_get_last_stats(ptr(stats), nb)
    return {"speeds": speeds, "failed_nodes": failed, "conv_loops": conv_loops, "blocks": blocks, "stats": stats,
            "input": inp, "kernel_ms": L.cm_last_kernel_ms() if timed else None}
