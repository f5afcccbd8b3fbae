"""Learned warm start of the predictor-pass Jacobi solve: capture of training
records from the solver itself, exact targets, training on the GPU with the
library's BF16-MFMA kernels, upload into the solver and evaluation against the
plain and extrapolated guesses.  No reference counterpart.

    python -m cochlearmodel_amd.warmstart train --out cochlearmodel_amd/data/warmstart_mlp.bin
    python -m cochlearmodel_amd.warmstart eval --model cochlearmodel_amd/data/warmstart_mlp.bin
    python -m cochlearmodel_amd.warmstart features --signals 512 --eval-signals 256 --stride 1024 --per-item 32

Model file: float32 W1 (769 x 512), b1, W2 (512 x 512), b2, W3 (512 x 256), b3,
c_out, read by cm_ws_load_file as well; training metadata beside it (.json).

Network input per predictor pass (769 values, each block RMS-normalised):
predictor displacement d, predictor velocity v, the residual r_lin of the
extrapolated pressure p_lin = p + (dt/dt_prev)(p - p_prev), and the input
sample s (scaled by the previous pressure's RMS).  Output y (256): the
correction, applied as p_lin + rms(r_lin) c_out [1,2,1]/4 * y.  Target:
(A^-1 Y - p_lin) / (rms(r_lin) c_out) with A^-1 Y the exact FP64 solve.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

from . import _lib, solver
from ._lib import check, lib, ptr
from .batch import ToneBatch, random_tones
from .mlp import SIZES, WarmStartMLP

HERE = os.path.dirname(os.path.abspath(__file__))
MODEL_PATH = os.path.join(HERE, "data", "warmstart_mlp.bin")
FS = 44100.0
# Batch b of 256 tones is drawn with seed TRAIN_SEED + b (training) or
# EVAL_SEED + b (held out).  The two ranges are a million seeds apart, so no
# batch is ever in both sets (checked by tests/test_warmstart_seeds.py).
TRAIN_SEED = 20261202 + 1_000_000   # training tones
EVAL_SEED = 20261202 + 52           # held-out tones
SEED_RANGE = 1_000_000              # batches available to each set
SEC_M = 8                      # row of the section table holding the pressure matrix diagonal (CM_SEC_M)


def tol_scalars(tau: float = 1e-6, warm_start: int = 1):
    sc = solver.default_scalars(FS)
    sc.jacobi_mode, sc.tol_rel, sc.check_every, sc.max_sweeps, sc.warm_start = 1, tau, 4, 1000, warm_start
    return sc


def capture(tb: ToneBatch, tau: float, stride: int, per_item: int):
    """Run the batch once with the extrapolated guess, recording every
    stride-th step's predictor pass; returns the filled records (device)."""
    import torch
    rec = torch.full((tb.TB * per_item, _lib.WS_RECORD), -1.0, dtype=torch.float32, device="cuda")
    check(lib().cm_ws_set_capture(ptr(rec), rec.shape[0], stride, per_item), "cm_ws_set_capture")
    try:
        tb.run_bm(tol_scalars(tau, 1), timed=False)
    finally:
        lib().cm_ws_set_capture(None, 0, 0, 0)
    torch.cuda.synchronize()
    keep = rec[:, 770] > 0          # slot written, and not the zero-state first step
    return rec[keep].contiguous()


def targets(sec, rec, c_out: float = 1.0):
    import torch
    T = torch.empty((rec.shape[0], 256), dtype=torch.float32, device="cuda")
    check(lib().cm_ws_targets(ptr(sec), ptr(rec), rec.shape[0], C.c_float(c_out), ptr(T)), "cm_ws_targets")
    return T


def upload(params, c_out: float):
    """Model into the solver (cm_ws_load); params = [W1, b1, W2, b2, W3, b3]."""
    p = [np.ascontiguousarray(x, np.float32) for x in params]
    for i, (kin, kout) in enumerate(zip(SIZES[:-1], SIZES[1:])):
        assert p[2 * i].shape == (kin, kout) and p[2 * i + 1].shape == (kout,)
    check(lib().cm_ws_load(ptr(p[0]), ptr(p[1]), ptr(p[2]), ptr(p[3]), ptr(p[4]), ptr(p[5]), C.c_float(c_out)),
          "cm_ws_load")


def save_model(path, params, c_out, meta):
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    flat = [np.ascontiguousarray(p, np.float32).ravel() for p in params] + [np.array([c_out], np.float32)]
    np.concatenate(flat).tofile(path)
    with open(os.path.splitext(path)[0] + ".json", "w") as f:
        json.dump(meta, f, indent=1)


def load_model(path=MODEL_PATH):
    flat = np.fromfile(path, np.float32)
    params, off = [], 0
    for kin, kout in zip(SIZES[:-1], SIZES[1:]):
        params.append(flat[off:off + kin * kout].reshape(kin, kout).copy())
        off += kin * kout
        params.append(flat[off:off + kout].copy())
        off += kout
    if off + 1 != flat.size:
        raise ValueError(f"{path}: {flat.size} floats, expected {off + 1}")
    meta_path = os.path.splitext(path)[0] + ".json"
    meta = json.load(open(meta_path)) if os.path.exists(meta_path) else {}
    return params, float(flat[off]), meta


def build_training_set(n_signals: int, batch_signals: int, stride: int, per_item: int, tau: float, seed: int,
                       log=print):
    import torch
    Xs, Ts = [], []
    for b in range(n_signals // batch_signals):
        f, db = random_tones(batch_signals, seed + b)
        tb = ToneBatch(f, db, 0.2, with_lambda=False)
        tb.generate_input()
        rec = capture(tb, tau, stride, per_item)
        Ts.append(targets(tb.bm.sec, rec))
        Xs.append(rec[:, :SIZES[0]].clone())
        log(f"batch {b}: {rec.shape[0]} records")
        del tb, rec
        torch.cuda.empty_cache()
    return torch.cat(Xs), torch.cat(Ts)


def train(n_signals=2048, batch_signals=256, stride=64, per_item=512, tau=1e-6, epochs=24, lr=1e-3, decay=0.85,
          seed=TRAIN_SEED, log=print):
    """2048 solver trajectories (tones drawn like the 256-tone grid), a record of
    every 64th step of every time block (7.6 M records, 23 GB of features on
    the device), 24 epochs of Adam on the library's BF16-MFMA kernels."""
    import torch
    t0 = time.time()
    X, T = build_training_set(n_signals, batch_signals, stride, per_item, tau, seed, log)
    finite = torch.isfinite(X).all(1) & torch.isfinite(T).all(1)
    X, T = X[finite].contiguous(), T[finite].contiguous()
    c_out = float(torch.sqrt((T.double() ** 2).mean()).item())
    T = (T / c_out).contiguous()
    log(f"{X.shape[0]} records in {time.time() - t0:.1f} s, c_out {c_out:.4e}")
    net = WarmStartMLP(seed=1, device="cuda")
    t1 = time.time()
    hist = net.train(X, T, epochs=epochs, batch=1024, lr=lr, decay=decay, log=log)
    torch.cuda.synchronize()
    meta = {"n_signals": n_signals, "records": int(X.shape[0]), "stride": stride, "per_item": per_item, "tau": tau,
            "epochs": epochs,
            "lr": lr, "decay": decay, "seed": seed, "final_mse": hist[-1], "train_s": time.time() - t1,
            "features": "d, v, r_lin (RMS-normalised), s / rms(p_prev)",
            "target": "(A^-1 Y - p_lin) / (rms(r_lin) c_out)"}
    return net.params_numpy(), c_out, meta, hist


def evaluate(params, c_out, n_signals=256, batch_signals=256, tau=1e-6, seed=EVAL_SEED, log=print):
    """Held-out tones: BM time and sweeps/step of the plain, extrapolated and
    learned guesses on the same resident input."""
    import torch
    upload(params, c_out)
    res = {k: {"bm_ms": 0.0, "steps": 0, "passes": 0, "sweeps": 0} for k in ("plain", "extrapolated", "learned")}
    for b in range(max(1, n_signals // batch_signals)):
        f, db = random_tones(batch_signals, seed + b)
        tb = ToneBatch(f, db, 0.2, with_lambda=False)
        tb.generate_input()
        for name, ws in (("plain", 0), ("extrapolated", 1), ("learned", 2)):
            ms = tb.run_bm(tol_scalars(tau, ws), timed=True)
            torch.cuda.synchronize()
            st, pa, sw = tb.stats()
            r = res[name]
            r["bm_ms"] += ms
            r["steps"] += st
            r["passes"] += pa
            r["sweeps"] += sw
            log(f"batch {b} {name}: {ms:.1f} ms, {sw / st:.2f} sweeps/step")
        del tb
        torch.cuda.empty_cache()
    out = {}
    for k, r in res.items():
        out[k] = {"bm_ms": r["bm_ms"], "section_steps_per_s": 256 * r["steps"] / (r["bm_ms"] * 1e-3),
                  "jacobi_iters_per_step": r["sweeps"] / r["steps"], "passes_per_step": r["passes"] / r["steps"],
                  "committed_steps": r["steps"]}
    return out


def _records(n_signals, batch_signals, stride, per_item, tau, seed, log):
    """Captured predictor-pass records and their raw targets
    (A^-1 Y - p_lin) / rms(r_lin), plus the m row of the pressure matrix."""
    import torch
    R, T, m = [], [], None
    for b in range(max(1, n_signals // batch_signals)):
        f, db = random_tones(batch_signals, seed + b)
        tb = ToneBatch(f, db, 0.2, with_lambda=False)
        tb.generate_input()
        rec = capture(tb, tau, stride, per_item)
        T.append(targets(tb.bm.sec, rec))
        R.append(rec)
        if m is None:
            m = tb.bm.sec.view(-1, 256)[SEC_M].double().clone()
        log(f"records batch {b}: {rec.shape[0]}")
        del tb
        torch.cuda.empty_cache()
    return torch.cat(R), torch.cat(T), m


def _apply_A(m, p):
    """Rows of A p for the solver's pressure matrix: l_i p_{i-1} + m_i p_i + p_{i+1},
    l_i = 1 for 0 < i < 254, p_{256} = 0."""
    import torch
    out = m * p
    out[:, :-1] += p[:, 1:]
    out[:, 1:254] += p[:, 0:253]
    return out


def _smooth(y):
    """[1,2,1]/4 with replicated edges, as the solver applies to the output."""
    import torch
    left = torch.cat([y[:, :1], y[:, :-1]], 1)
    right = torch.cat([y[:, 1:], y[:, -1:]], 1)
    return 0.25 * left + 0.5 * y + 0.25 * right


def study_features(n_train=512, n_eval=256, batch_signals=256, stride=1024, per_item=32, tau=1e-6, epochs=24,
                   lr=1e-3, decay=0.85, log=print):
    """Ablation of the network's inputs and output on the same captured records.

    The configuration names the inputs "displacement, velocity and last
    pressure plus the input sample" and a pressure guess as the output.  Three
    networks of that size (769-512-512-256 GELU) are trained with the same
    schedule and scored on held-out records by how much closer than the
    extrapolated pressure p_lin their guess lands to the exact solve x* = A^-1 Y:

    * residual-correction (the solver's): d, v, r_lin = Y - A p_lin, s -> the
      correction (x* - p_lin) / rms(r_lin), applied smoothed;
    * pressure-correction: d, v, p_lin (the extrapolated last pressure), s -> the
      same correction;
    * pressure-absolute: d, v, p_lin, s -> x* / rms(p_lin), the guess itself.

    Every input vector is divided by its RMS over the sections, as in the
    solver.  Scores per record: e = |g - x*| / |p_lin - x*| and
    rho = |Y - A g| / |Y - A p_lin|; the solver's residual gate takes a guess
    only where rho < 1."""
    import torch
    Rt, Tt, m = _records(n_train, batch_signals, stride, per_item, tau, TRAIN_SEED, log)
    Re, Te, _ = _records(n_eval, batch_signals, stride, per_item, tau, EVAL_SEED, log)

    def rms(x):
        return torch.sqrt((x.double() ** 2).mean(1, keepdim=True)).float() + 1e-30

    def inputs(R, kind):
        if kind == "residual":
            return R[:, :SIZES[0]].contiguous()
        plin = R[:, _lib.WS_REC_PLIN:_lib.WS_REC_PLIN + 256]
        return torch.cat([R[:, :512], plin / rms(plin), R[:, 768:769]], 1).contiguous()

    def target(R, T, kind):
        if kind != "absolute":
            return T
        plin = R[:, _lib.WS_REC_PLIN:_lib.WS_REC_PLIN + 256]
        return (plin + R[:, 769:770] * T) / rms(plin)

    out = {}
    for name, feat, kind in (("residual-correction", "residual", "correction"),
                             ("pressure-correction", "pressure", "correction"),
                             ("pressure-absolute", "pressure", "absolute")):
        X, T = inputs(Rt, feat), target(Rt, Tt, kind)
        ok = torch.isfinite(X).all(1) & torch.isfinite(T).all(1)
        X, T = X[ok].contiguous(), T[ok].contiguous()
        c = float(torch.sqrt((T.double() ** 2).mean()).item())
        net = WarmStartMLP(seed=1, device="cuda")
        hist = net.train(X, (T / c).contiguous(), epochs=epochs, batch=1024, lr=lr, decay=decay)
        Xe = inputs(Re, feat)
        y = torch.cat([net.forward(Xe[i:i + 4096].contiguous()).clone() for i in range(0, Xe.shape[0], 4096)])
        plin = Re[:, _lib.WS_REC_PLIN:_lib.WS_REC_PLIN + 256].double()
        Y = Re[:, _lib.WS_REC_Y:_lib.WS_REC_Y + 256].double()
        s = Re[:, 769:770].double()
        xs = plin + s * Te.double()
        if kind == "correction":
            g = plin + s * c * _smooth(y.double())
        else:
            g = rms(plin).double() * c * y.double()
        e = (g - xs).norm(dim=1) / (plin - xs).norm(dim=1)
        rho = (Y - _apply_A(m, g)).norm(dim=1) / (Y - _apply_A(m, plin)).norm(dim=1)
        good = torch.isfinite(e) & torch.isfinite(rho)
        e, rho = e[good], rho[good]
        out[name] = {"train_mse": hist[-1], "median_e": float(e.median()), "median_rho": float(rho.median()),
                     "frac_rho_below_1": float((rho < 1).double().mean()), "records": int(good.sum())}
        log(f"{name}: {out[name]}")
        del net
        torch.cuda.empty_cache()
    out["train_records"], out["eval_records"] = int(Rt.shape[0]), int(Re.shape[0])
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=("train", "eval", "features"))
    ap.add_argument("--out", default=MODEL_PATH)
    ap.add_argument("--model", default=MODEL_PATH)
    ap.add_argument("--signals", type=int, default=2048)
    ap.add_argument("--eval-signals", type=int, default=256)
    ap.add_argument("--epochs", type=int, default=24)
    ap.add_argument("--stride", type=int, default=64, help="steps between captured records of a block")
    ap.add_argument("--per-item", type=int, default=512, help="record slots per block (>= steps per block / stride)")
    ap.add_argument("--tau", type=float, default=1e-6)
    args = ap.parse_args(argv)
    import torch
    torch.cuda.set_device(0)
    log = lambda m: print(m, flush=True)
    if args.cmd == "train":
        params, c_out, meta, hist = train(n_signals=args.signals, stride=args.stride, per_item=args.per_item,
                                          tau=args.tau, epochs=args.epochs, log=log)
        save_model(args.out, params, c_out, meta)
        log(json.dumps({"saved": args.out, **meta}))
        ev = evaluate(params, c_out, n_signals=args.eval_signals, tau=args.tau, log=log)
    elif args.cmd == "features":
        res = study_features(n_train=args.signals, n_eval=args.eval_signals, stride=args.stride,
                             per_item=args.per_item, tau=args.tau, epochs=args.epochs, log=log)
        print(json.dumps({"feature_study": res}), flush=True)
        return 0
    else:
        params, c_out, meta = load_model(args.model)
        ev = evaluate(params, c_out, n_signals=args.eval_signals, tau=args.tau, log=log)
    print(json.dumps({"evaluation": ev, "held_out_signals": args.eval_signals}), flush=True)


if __name__ == "__main__":
    sys.exit(main())
