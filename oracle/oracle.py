"""ctypes front end of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Imported by tests/, ``__graft_entry__.smoke()`` and the cpu_baseline leg of
``bench.py``.  The product (``cochlearmodel_amd``) never imports it.  See
``oracle.h`` for the reference lines each function restates.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# CM_ORACLE_LIB points the tests at another build of the same sources (the
# mutant builds of the step-control pins); the default is the in-tree build.
_LIB_PATH = os.environ.get("CM_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
_lib = None

SECTIONS = 256
LAMBDA_COUNT = 3


class BMScalars(C.Structure):
    _fields_ = [(n, C.c_float) for n in (
        "w_ohc", "time_step", "time_step_out", "delta_x", "alpha_r", "Ts", "alpha_l", "alpha_s",
        "Gme", "a0", "a1", "a2", "sigma_ow", "eta1", "eta2", "dt_min", "dt_max")] + [
        ("j_fast", C.c_int32), ("j_slow", C.c_int32), ("max_passes", C.c_int32),
        ("decision_mode", C.c_int32), ("jacobi_mode", C.c_int32), ("check_every", C.c_int32),
        ("max_sweeps", C.c_int32), ("tol_rel", C.c_float), ("warm_start", C.c_int32)]


ITEM_DTYPE = np.dtype([("start", np.int32), ("mosi", np.int32), ("mesi", np.int32),
                       ("write_start", np.int32), ("tol", np.float32), ("m1tol", np.float32),
                       ("set", np.int32), ("pad", np.int32)])


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _lib.orc_tanhf.restype = C.c_float
        _lib.orc_tanhf.argtypes = [C.c_float]
        _lib.orc_expf.restype = C.c_float
        _lib.orc_expf.argtypes = [C.c_float]
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def tanhf(x: float) -> float:
    return lib().orc_tanhf(C.c_float(x))


def expf(x: float) -> float:
    return lib().orc_expf(C.c_float(x))


def model_vectors() -> np.ndarray:
    out = np.zeros((7, SECTIONS), np.float64)
    lib().orc_model_vectors(_p(out))
    return out


def tri_mid_diag() -> np.ndarray:
    out = np.zeros(SECTIONS, np.float64)
    lib().orc_tri_mid_diag(_p(out))
    return out


def default_scalars(Fs: float = 44100.0) -> BMScalars:
    s = BMScalars()
    lib().orc_default_scalars(C.c_float(Fs), C.byref(s))
    return s


def section_params() -> np.ndarray:
    """9 x 256 float32: mass, 1/m_i, R, S, Q, S_ohc, S_tm, R_tm (src/solver.cpp:357-376), m_i."""
    v = model_vectors()
    m = tri_mid_diag()
    out = np.zeros((9, SECTIONS), np.float32)
    out[0] = v[0].astype(np.float32)
    out[1] = np.float32(1.0) / m.astype(np.float32)
    out[2] = v[1].astype(np.float32)
    out[3] = v[2].astype(np.float32)
    out[4] = v[6].astype(np.float32)
    out[5] = v[5].astype(np.float32)
    out[6] = v[4].astype(np.float32)
    out[7] = v[3].astype(np.float32)
    out[8] = (1.0 / out[1].astype(np.float64)).astype(np.float32)   # m_i from the float 1/m_i row
    return out


def set_ws_net(params, c_out: float):
    """Network of the learned predictor-pass guess (warm_start = 2):
    params = [W1 (769 x 512), b1, W2 (512 x 512), b2, W3 (512 x 256), b3]."""
    shapes = [(769, 512), (512,), (512, 512), (512,), (512, 256), (256,)]
    arr = [np.ascontiguousarray(p, np.float32) for p in params]
    assert [a.shape for a in arr] == shapes, [a.shape for a in arr]
    rc = lib().orc_ws_set_net(*[_p(a) for a in arr], C.c_float(c_out))
    assert rc == 0, rc


def ws_forward(x: np.ndarray) -> np.ndarray:
    """The network set by set_ws_net on rows of 769 features."""
    x = np.ascontiguousarray(np.atleast_2d(x), np.float32)
    y = np.zeros((x.shape[0], 256), np.float32)
    rc = lib().orc_ws_forward(_p(x), C.c_int(x.shape[0]), _p(y))
    assert rc == 0, rc
    return y


def block_decomposition(time_blocks: int, bufP1: int, overlap: int, decouple_filter: int):
    a = [np.zeros(time_blocks, np.int32) for _ in range(4)]
    rc = lib().orc_block_decomposition(time_blocks, bufP1, overlap, decouple_filter, *[_p(x) for x in a])
    assert rc == 0
    return a


def make_items(start, mosi, mesi, write_start, tol, m1tol, sets=None) -> np.ndarray:
    n = len(start)
    it = np.zeros(n, ITEM_DTYPE)
    it["start"], it["mosi"], it["mesi"], it["write_start"] = start, mosi, mesi, write_start
    it["tol"], it["m1tol"] = tol, m1tol
    it["set"] = 0 if sets is None else sets
    return it


def bm_run(sc: BMScalars, sec: np.ndarray, gamma: np.ndarray, items: np.ndarray, inp: np.ndarray,
           rows: int, precision: str = "f32", threads: int = 0, diagnostics: bool = False):
    """Run the kernel-17 restatement; returns a dict of outputs."""
    sec = np.ascontiguousarray(sec, np.float32)
    gamma = np.ascontiguousarray(np.atleast_2d(gamma), np.float32)
    items = np.ascontiguousarray(items, ITEM_DTYPE)
    inp = np.ascontiguousarray(inp, np.float32)
    n = len(items)
    steps = np.zeros(n, np.int64)
    passes = np.zeros(n, np.int64)
    sweeps = np.zeros(n, np.int64)
    out = {"steps": steps, "passes": passes, "sweeps": sweeps}
    if precision == "f32":
        speeds = np.zeros((rows, SECTIONS), np.float32)
        fn = np.zeros(rows, np.int32) if diagnostics else None
        ct = np.zeros(rows, np.float32) if diagnostics else None
        cl = np.zeros(rows, np.float32) if diagnostics else None
        cj = np.zeros(rows, np.float32) if diagnostics else None
        fb = np.zeros(n, np.int32) if diagnostics else None
        ctb = np.zeros(n, np.float32) if diagnostics else None
        clb = np.zeros(n, np.float32) if diagnostics else None
        cjb = np.zeros(n, np.float32) if diagnostics else None
        rc = lib().orc_bm_run_f32(C.byref(sc), _p(sec), _p(gamma), _p(items), C.c_int(n), _p(inp),
                                  C.c_int64(inp.size), _p(speeds), C.c_int64(rows), _p(fn), _p(ct), _p(cl),
                                  _p(cj), _p(fb), _p(ctb), _p(clb), _p(cjb), _p(steps), _p(passes),
                                  _p(sweeps), C.c_int(threads))
        out.update(speeds=speeds, failed_nodes=fn, conv_time=ct, conv_loops=cl, conv_jacobi=cj,
                   failed_blocks=fb, conv_time_blocks=ctb, conv_loops_blocks=clb, conv_jacobi_blocks=cjb)
    else:   # "f64": FP64 rebuild; "f64c32": the same with the sample clock in float
        speeds = np.zeros((rows, SECTIONS), np.float64)
        fn = lib().orc_bm_run_f64 if precision == "f64" else lib().orc_bm_run_f64_clock32
        rc = fn(C.byref(sc), _p(sec), _p(gamma), _p(items), C.c_int(n), _p(inp),
                                  C.c_int64(inp.size), _p(speeds), C.c_int64(rows), _p(steps), _p(passes),
                                  _p(sweeps), C.c_int(threads))
        out.update(speeds=speeds)
    assert rc == 0, rc
    return out


PASS_RECORD_DTYPE = np.dtype([("step", np.int32), ("pass_", np.int32), ("in_s", np.int32), ("decided", np.int32),
                              ("t", np.float32), ("dt_in", np.float32), ("dt_out", np.float32), ("L", np.float32),
                              ("F", np.float32), ("e_a", np.float32), ("e_v", np.float32), ("another", np.int32),
                              ("reset", np.int32), ("capped", np.int32), ("sweeps", np.int32)])


def step_control(aref, na, vref, nv, dt: float, tol: float, m1tol: float, dt_min: float = 1e-18,
                 dt_max: float = 1e-6, decision_mode: int = 0) -> dict:
    """One pass's step control (include/cochlea.cuh:14021-14088) on 256-section vectors."""
    vecs = [np.ascontiguousarray(np.broadcast_to(np.asarray(x, np.float32), (SECTIONS,))) for x in (aref, na, vref, nv)]
    out = np.zeros(5, np.float32)
    ar = np.zeros(2, np.int32)
    rc = lib().orc_step_control_f32(*[_p(v) for v in vecs], C.c_float(dt), C.c_float(tol), C.c_float(m1tol),
                                    C.c_float(dt_min), C.c_float(dt_max), C.c_int(decision_mode), _p(out), _p(ar))
    assert rc == 0
    return {"dt": out[0], "L": out[1], "e_a": out[2], "e_v": out[3], "F": out[4], "another": int(ar[0]),
            "reset": int(ar[1])}


def bm_trace(sc: BMScalars, sec: np.ndarray, gamma: np.ndarray, item: np.ndarray, inp: np.ndarray,
             cap: int) -> np.ndarray:
    """Per-pass records of one item's kernel-17 run (stops after cap passes)."""
    rec = np.zeros(cap, PASS_RECORD_DTYPE)
    n = C.c_int(0)
    item = np.ascontiguousarray(np.atleast_1d(item)[:1], ITEM_DTYPE)
    inp = np.ascontiguousarray(inp, np.float32)
    rc = lib().orc_bm_trace_f32(C.byref(sc), _p(np.ascontiguousarray(sec, np.float32)),
                                _p(np.ascontiguousarray(np.atleast_2d(gamma), np.float32)), _p(item), _p(inp),
                                C.c_int64(inp.size), _p(rec), C.c_int(cap), C.byref(n))
    assert rc == 0
    return rec[:n.value]


def jacobi(Y: np.ndarray, rsm: np.ndarray, p0: np.ndarray, sweeps: int) -> np.ndarray:
    p = np.ascontiguousarray(p0, np.float32).copy()
    lib().orc_jacobi(_p(np.ascontiguousarray(Y, np.float32)), _p(np.ascontiguousarray(rsm, np.float32)),
                     _p(p), C.c_int(sweeps))
    return p


def ihc_lambda(bm: np.ndarray, interval_len: int, frdp: int, ac_b: np.ndarray, dc_b: np.ndarray,
               ihc_damage: np.ndarray, A: np.ndarray, spont: np.ndarray, lambda_sat: float = 500.0,
               eta_ac: float = 1.0, eta_dc: float = 100.0, scale_bm: float = 1.0, stages: bool = False,
               threads: int = 0, ac_a: np.ndarray | None = None):
    bm = np.ascontiguousarray(bm, np.float32)
    rows = bm.shape[0]
    lam = np.zeros((LAMBDA_COUNT, rows, SECTIONS), np.float32)
    st = np.zeros((5, rows, SECTIONS), np.float32) if stages else None
    ac_b = np.ascontiguousarray(ac_b, np.float64)
    dc_b = np.ascontiguousarray(dc_b, np.float64)
    ac_a = None if ac_a is None else np.ascontiguousarray(ac_a, np.float64)
    rc = lib().orc_ihc_lambda(_p(bm), C.c_int64(rows), C.c_int(interval_len), C.c_int(frdp), _p(ac_b), _p(ac_a),
                              C.c_int(ac_b.size), _p(dc_b), C.c_int(dc_b.size),
                              _p(np.ascontiguousarray(ihc_damage, np.float64)),
                              _p(np.ascontiguousarray(A, np.float32)), _p(np.ascontiguousarray(spont, np.float32)),
                              C.c_float(lambda_sat), C.c_float(eta_ac), C.c_float(eta_dc), C.c_float(scale_bm),
                              _p(lam), _p(st), C.c_int(threads))
    assert rc == 0
    return (lam, st) if stages else lam


def jnd(lam: np.ndarray, n_intervals: int, interval_nodes: int, head_nodes: int, length_nodes: int,
        dA: np.ndarray, ref_index: np.ndarray, nIHC: np.ndarray, W: np.ndarray, Fs: float,
        delta_alpha_factor: float, calc_ai: bool = True, calc_rms: bool = True, backup_stage: int = 0):
    """Mean rates, F_RA and FisherAISum; with backup_stage 1..6 also that
    stage's buffer (NaN where the stage writes nothing)."""
    lam = np.ascontiguousarray(lam, np.float32)
    rows = lam.shape[1]
    mr = np.zeros((LAMBDA_COUNT, n_intervals, SECTIONS), np.float32)
    fra = np.zeros((LAMBDA_COUNT, n_intervals), np.float32)
    fai = np.zeros((LAMBDA_COUNT, n_intervals), np.float32)
    shape = {1: (LAMBDA_COUNT, n_intervals), 3: lam.shape, 4: lam.shape}.get(
        backup_stage, (LAMBDA_COUNT, n_intervals, SECTIONS))
    buf = np.full(shape, np.nan, np.float32) if backup_stage else None
    rc = lib().orc_jnd(_p(lam), C.c_int64(rows), C.c_int(n_intervals), C.c_int(interval_nodes),
                       C.c_int(head_nodes), C.c_int(length_nodes), _p(np.ascontiguousarray(dA, np.float32)),
                       _p(np.ascontiguousarray(ref_index, np.int32)), _p(np.ascontiguousarray(nIHC, np.float64)),
                       _p(np.ascontiguousarray(W, np.float32)), C.c_double(Fs), C.c_double(delta_alpha_factor),
                       C.c_int(int(calc_ai)), C.c_int(int(calc_rms)), _p(mr), _p(fra), _p(fai),
                       C.c_int(backup_stage), _p(buf))
    assert rc == 0
    return (mr, fra, fai, buf) if backup_stage else (mr, fra, fai)


def generate_input(dA, Wn, freq, WN, wn_dc, wn_norm, sig_norm, interval_len, head, Fs):
    n = len(dA)
    out = np.zeros(n * interval_len, np.float32)
    lib().orc_generate_input(C.c_int(n), _p(np.ascontiguousarray(dA, np.float64)),
                             _p(np.ascontiguousarray(Wn, np.float64)), _p(np.ascontiguousarray(freq, np.float64)),
                             _p(np.ascontiguousarray(WN, np.float32)), C.c_double(wn_dc), C.c_double(wn_norm),
                             C.c_double(sig_norm), C.c_int(interval_len), C.c_int(head), C.c_float(Fs), _p(out))
    return out
