"""Host side of the post-BM stages and of input generation over the C-ABI:
AudioGramCreator::setupIHCOHC / runInCuda / calcJND / calcJNDFinal
(src/AudioGramCreator.cpp:261-313, 454-508, 603-974) and the profile input
path of CSolver::Run_Cuda (src/solver.cpp:456-572).
"""
from __future__ import annotations

import ctypes as C
import json
import os

import numpy as np

from ._lib import IHC_INTERVAL_DTYPE, LAMBDA_COUNT, PROFILE_DTYPE, SECTIONS, check, lib, ptr
from .solver import expand_vector

DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")

# include/const.h:268-297, src/params.cpp:51, 71-73, 111, 123-138
AIHC_DEFAULT = (70.0, 50.0, 30.0)
SPONT_DEFAULT = (60.0, 3.0, 0.1)
W_DEFAULT = (0.61, 0.23, 0.16)
LAMBDA_SAT = 500.0
ETA_AC, ETA_DC, TDELTA = 1.0, 100.0, 2e-3
AN_FIBERS = 30000


def decode_filter_mode_2(vec):
    """HFunction::decodeBinFile (src/HFunction.cpp:137-153): [order, isFIR, (a...), b...]."""
    order, is_fir = int(vec[0]), int(vec[1]) == 1
    if is_fir:
        return np.asarray(vec[2:2 + order], np.float64), None
    a = np.asarray(vec[2:2 + order], np.float64)
    b = np.asarray(vec[2 + order:2 + 2 * order], np.float64)
    return b, a


def load_ac_filter(fs: float = 44100.0):
    """The committed default AC filter for this sample rate (FIR: a is None)."""
    path = os.path.join(DATA, f"ac_filter_equiripple_{int(fs)}.json")
    with open(path) as f:
        return decode_filter_mode_2(json.load(f)["filter_mode_2"])


def design_ac_filter(fs: float, function_filter: str):
    """AC low-pass from a Function_Filter string, designed by the library
    (SubModel::analyzeACFilter, src/SubModel.cpp:153-503); returns (b, a) like
    load_ac_filter, a = None for the FIR designs."""
    h = np.zeros(4096, np.float64)
    a = np.zeros(4096, np.float64)
    n = np.zeros(1, np.int32)
    iir = np.zeros(1, np.int32)
    check(lib().cm_design_ac_filter(float(fs), function_filter.encode(), ptr(h), ptr(a), len(h), ptr(n), ptr(iir)),
          "cm_design_ac_filter")
    k = int(n[0])
    return h[:k].copy(), (a[:k].copy() if iir[0] else None)


def dc_filter(fs: float = 44100.0, tdelta: float = TDELTA) -> np.ndarray:
    """Box-car DC filter of setupIHCOHC (src/AudioGramCreator.cpp:806-812)."""
    n = int(round(fs * np.float32(tdelta)))
    v = np.float32(1.0) / np.float32(np.float32(fs) * np.float32(tdelta))
    return np.full(n, float(v), np.float64)


def ihc_damage(ihc_vector) -> np.ndarray:
    """10^IHC_Vector per section (src/AudioGramCreator.cpp:813-820)."""
    return np.power(10.0, expand_vector(ihc_vector).astype(np.float64))


def aihc_table(aihc=AIHC_DEFAULT) -> np.ndarray:
    """model_Aihc: 3 x 256 floats, each fiber group's gain repeated per section."""
    return expand_vector(np.asarray(aihc, np.float32), SECTIONS * LAMBDA_COUNT).reshape(LAMBDA_COUNT, SECTIONS)


def nerves_clusters(spont=SPONT_DEFAULT, W=W_DEFAULT) -> np.ndarray:
    nc = np.zeros(9, np.float32)
    nc[:3] = spont
    nc[6:] = W
    return nc


def run_ihc_reference(calc_time: int, interval_nodes: int, time_blocks: int, decouple_filter: int, fs=44100.0,
                      ihc_vector=8.0, ac=None, backup_stage: int = 0):
    """IHCNewKernel + loadAihc + RunIHCKernel + BMOHCKernel_Copy_Lambda on the
    BM velocity already in the library's buffer.  Returns lambda (3, rows, 256)
    and, for backup_stage 9..14, that stage."""
    L = lib()
    b, a = ac if ac is not None else load_ac_filter(fs)
    dc = dc_filter(fs)
    dmg = ihc_damage(ihc_vector)
    nc = nerves_clusters()
    rc = L.IHCNewKernel(ptr(dmg), ptr(nc), ptr(dc), len(dc), ptr(b), ptr(a), 1 if a is not None else 0, len(b),
                        SECTIONS, time_blocks, 2e-5, None, 0, calc_time, calc_time, calc_time, interval_nodes, 0, 0,
                        LAMBDA_SAT, ETA_AC, ETA_DC, 1, 1, 0, 0, 0, 0, 1.0, 1, 1, 0, decouple_filter, 0)
    check(rc, "IHCNewKernel")
    L.enableloadAihc()
    check(L.loadAihc(ptr(aihc_table())), "loadAihc")
    stage = np.zeros((calc_time, SECTIONS), np.float32) if 9 <= backup_stage <= 14 else None
    check(L.RunIHCKernel(ptr(stage), 0, 1, backup_stage, 0), "RunIHCKernel")
    lam = np.zeros((LAMBDA_COUNT, calc_time, SECTIONS), np.float32)
    check(L.BMOHCKernel_Copy_Lambda(ptr(lam), lam.size, 0), "BMOHCKernel_Copy_Lambda")
    return lam, stage


def reference_intervals(rows: int, interval_nodes: int, time_blocks: int, decouple_filter: int):
    """FIR and IIR interval tables of the reference layout."""
    DF = decouple_filter
    Lw = interval_nodes * DF if DF > 0 else rows
    frdp = interval_nodes * time_blocks
    fir = np.array([(0, rows, 0, Lw, rows, frdp, 0)], IHC_INTERVAL_DTYPE)
    if DF > 0:
        nb = max(1, time_blocks // DF)
        iir = []
        for k in range(nb):
            ln = Lw + (rows - nb * Lw if k + 1 == nb else 0)
            iir.append((k * Lw, ln, 0, ln, ln, frdp, 0))
        iir = np.array(iir, IHC_INTERVAL_DTYPE)
    else:
        iir = np.array([(0, rows, 0, rows, rows, frdp, 0)], IHC_INTERVAL_DTYPE)
    return fir, iir


def signal_intervals(starts, lengths, sets=None):
    """One interval per signal: the FIR restarts at each signal start."""
    n = len(starts)
    out = np.zeros(n, IHC_INTERVAL_DTYPE)
    out["row0"] = starts
    out["len"] = lengths
    out["kbase"] = 0
    out["L"] = lengths
    out["tla"] = lengths
    out["frdp"] = np.asarray(lengths) + 1
    out["set"] = 0 if sets is None else sets
    return out


class BatchedIHC:
    """Device-resident IHC/lambda through cm_ihc_lambda_run."""

    def __init__(self, fs=44100.0, ihc_vectors=None, ac=None, device="cuda"):
        import torch
        self.torch = torch
        b, a = ac if ac is not None else load_ac_filter(fs)
        self.n_ac = len(b)
        self.b = torch.from_numpy(np.ascontiguousarray(b, np.float64)).to(device)
        self.a = None if a is None else torch.from_numpy(np.ascontiguousarray(a, np.float64)).to(device)
        dc = dc_filter(fs)
        self.n_dc = len(dc)
        self.dc = torch.from_numpy(dc).to(device)
        if ihc_vectors is None:
            ihc_vectors = [8.0]
        dm = np.stack([ihc_damage(v) for v in ihc_vectors])
        self.damage = torch.from_numpy(np.ascontiguousarray(dm.reshape(-1))).to(device)
        self.A = torch.from_numpy(aihc_table().reshape(-1).copy()).to(device)
        self.spont = (C.c_float * 3)(*SPONT_DEFAULT)
        self.device = device

    def run(self, d_bm, rows: int, intervals: np.ndarray, iir_intervals=None, d_lambda=None, d_scratch=None,
            stages=False, workspace=True):
        """workspace=False runs the one-pass kernel (no AC/dS workspace)."""
        torch = self.torch
        dev = self.device
        if d_lambda is None:
            d_lambda = torch.empty((LAMBDA_COUNT, rows, SECTIONS), dtype=torch.float32, device=dev)
        if d_scratch is None and workspace:
            d_scratch = torch.empty((2, rows, SECTIONS), dtype=torch.float32, device=dev)
        d_st = torch.zeros((5, rows, SECTIONS), dtype=torch.float32, device=dev) if stages else None
        ivs = np.ascontiguousarray(intervals, IHC_INTERVAL_DTYPE)
        iir = None if iir_intervals is None else np.ascontiguousarray(iir_intervals, IHC_INTERVAL_DTYPE)
        L = lib()
        L.cm_set_stream(C.c_void_p(torch.cuda.current_stream().cuda_stream))
        rc = L.cm_ihc_lambda_run(ptr(d_bm), rows, ptr(ivs), len(ivs), ptr(iir), 0 if iir is None else len(iir),
                                 ptr(self.b), ptr(self.a), self.n_ac, ptr(self.dc), self.n_dc, ptr(self.damage),
                                 ptr(self.A), self.spont, LAMBDA_SAT, ETA_AC, ETA_DC, 1.0, ptr(d_scratch),
                                 ptr(d_lambda), ptr(d_st))
        check(rc, "cm_ihc_lambda_run")
        return d_lambda, d_st


# ------------------------------------------------------------------ JND ---
def jnd_profiles(freqs, levels, noises=(1111.0,), spl_ref=2e-5, one_ref_per_noise=True, calculate_jnd=True):
    """CParams::updateInputProfile (src/params.cpp:661-782) for tone signals:
    profiles, reference positions and the interval maps of
    src/params.cpp:1586-1611."""
    MIN_INF = 1111.0
    tf, tp, tn = len(freqs), len(levels), len(noises)
    add_f = 0 if one_ref_per_noise else 1
    add_n = 1 - add_f
    if not calculate_jnd:
        add_f = add_n = 0
    n_prof = ((tf + add_f) * tp + add_n) * tn
    prof = np.zeros(n_prof, PROFILE_DTYPE)
    spl2pa = lambda db: 0.0 if db == MIN_INF else 10.0 * 10.0 ** (db / 20.0) * spl_ref
    refs = []
    g = 0
    calc = 0
    for n_i, nz in enumerate(noises):
        npow = spl2pa(nz)
        for p_i, lv in enumerate(levels):
            for f_i, fr in enumerate(freqs):
                base = float(np.float32(lv))
                pw = spl2pa(base)
                prof[g]["dBSPLSignal"] = base
                prof[g]["dBSPLNoise"] = nz if nz != MIN_INF else 0.0
                prof[g]["isReference"] = 0
                prof[g]["referenceIndex"] = (tp * tf * (n_i + 1) + n_i) if one_ref_per_noise else (g + tf - f_i)
                prof[g]["calculateIndexPosition"] = calc
                prof[g]["dA"] = pw
                prof[g]["recipdA"] = 1.0 / pw if pw > 0 else 0.0
                prof[g]["Wn"] = npow
                prof[g]["frequency"] = float(np.float32(fr))       # testedFrequencies is float
                calc += 1
                g += 1
            if not one_ref_per_noise and calculate_jnd:
                refs.append(g)
                prof[g] = (0, 0, 0, nz if nz != MIN_INF else 0, npow, 1, 0, -1, g, 0, 0.0)
                g += 1
        if one_ref_per_noise and calculate_jnd:
            refs.append(g)
            prof[g] = (0, 0, 0, nz if nz != MIN_INF else 0, npow, 1, 0, -1, g, 0, 0.0)
            g += 1
    refs = np.asarray(refs, np.int32)
    n_int = n_prof
    serial = np.zeros(n_int, np.int32)
    n_ref = len(refs)
    calc_pos, to_ref = [], []
    for i in range(n_int):
        rr = i * n_ref // n_int
        if i not in set(refs.tolist()) and calculate_jnd:
            serial[i] = len(calc_pos)
            calc_pos.append(i)
            to_ref.append(int(refs[rr]))
        else:
            serial[i] = rr
    return prof, refs, serial, np.asarray(calc_pos, np.int32), np.asarray(to_ref, np.int32)


def reference_index(n_intervals, refs, serial, to_ref, handled=0):
    out = np.zeros(n_intervals, np.int32)
    rs = set(int(r) for r in refs)
    for i in range(n_intervals):
        g = i + handled
        gr = g if (g in rs or len(refs) == 0) else int(to_ref[serial[g]])
        out[i] = gr - handled
    return out


def jnd_final(F_RA, FAI, n_full, n_int, refs, serial, spl_ref=2e-5, max_db=150.0, calc_ai=True, calc_rms=True,
              handled=0):
    ai = np.zeros(n_full, np.float64)
    rate = np.zeros(n_full, np.float64)
    F_RA = np.ascontiguousarray(F_RA, np.float32)
    FAI = np.ascontiguousarray(FAI, np.float32)
    refs = np.ascontiguousarray(refs, np.int32)
    serial = np.ascontiguousarray(serial, np.int32)
    check(lib().cm_jnd_final(ptr(F_RA), ptr(FAI), n_full, n_int, handled, ptr(refs), len(refs), ptr(serial),
                             spl_ref, max_db, int(calc_ai), int(calc_rms), ptr(ai), ptr(rate), None, None),
          "cm_jnd_final")
    return ai, rate


def jnd_failed_signals(failed_blocks, n_int, refs, serial, decouple_filter, n_out, handled=0):
    """Failed-signal counts per calculated interval (calcJNDFinal)."""
    fb = np.ascontiguousarray(failed_blocks, np.int32)
    out = np.zeros(n_out, np.int32)
    refs = np.ascontiguousarray(refs, np.int32)
    serial = np.ascontiguousarray(serial, np.int32)
    check(lib().cm_jnd_failed_signals(ptr(fb), len(fb), n_int, handled, ptr(refs), len(refs), ptr(serial),
                                      decouple_filter, ptr(out)), "cm_jnd_failed_signals")
    return out


# ------------------------------------------------------- complex JND ---
def complex_jnd(raw_db, n_freq, n_levels, n_noises, methods=1, eps=1.0, failed=None):
    """Aggregated JND per (noise, frequency) profile over the tested levels
    (AudioGramCreator::calcComplexJND, src/AudioGramCreator.cpp:548-571) through
    the library's host routine.  Returns (jnd_final, warnings, convergence
    warnings), each [method][profile]."""
    raw = np.ascontiguousarray(raw_db, np.float64)
    n_prof = n_freq * n_noises
    iv = np.zeros(n_prof * n_levels, np.int32)
    check(lib().cm_complex_jnd_intervals(n_freq, n_levels, n_noises, ptr(iv)), "cm_complex_jnd_intervals")
    n_methods = bin(methods & 3).count("1")
    out = np.zeros(n_methods * n_prof, np.float32)
    warn = np.zeros(n_methods * n_prof, np.int32)
    conv = np.zeros(n_methods * n_prof, np.int32)
    fl = None if failed is None else np.ascontiguousarray(failed, np.int32)
    check(lib().cm_complex_jnd(ptr(raw), ptr(fl), len(raw), ptr(iv), n_prof, n_levels, methods, C.c_float(eps),
                               ptr(out), ptr(warn), ptr(conv)), "cm_complex_jnd")
    return out.reshape(n_methods, n_prof), warn.reshape(n_methods, n_prof), conv.reshape(n_methods, n_prof)
