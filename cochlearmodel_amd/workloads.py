"""Seeded synthetic inputs for the benchmark configurations (all at
Fs = 44.1 kHz, 256 sections, master seed 20261202).

Amplitude conventions follow the reference: a tone of L dB SPL has peak
10 * SPLRef * 10^(L/20) dyn/cm^2 (CONV_dB_TO_AMP, include/const.h:53).
"""
from __future__ import annotations

import numpy as np

FS = 44100.0
SPL_REF = 2e-5
MASTER_SEED = 20261202


def db_to_peak(db) -> np.ndarray:
    return 10.0 * SPL_REF * 10.0 ** (np.asarray(db, np.float64) / 20.0)


def raised_cosine_ramps(n: int, ramp: int) -> np.ndarray:
    w = np.ones(n)
    if ramp > 0:
        r = 0.5 * (1 - np.cos(np.pi * np.arange(ramp) / ramp))
        w[:ramp] = r
        w[n - ramp:] = r[::-1]
    return w


def tone(freq: float, db: float, duration: float, fs: float = FS, ramp_s: float = 0.010) -> np.ndarray:
    n = int(round(duration * fs))
    t = np.arange(n) / fs
    return (db_to_peak(db) * raised_cosine_ramps(n, int(round(ramp_s * fs))) * np.sin(2 * np.pi * freq * t)).astype(
        np.float32)


def c1_signal(prepend: int = 220) -> np.ndarray:
    """One 1 kHz tone, 60 dB SPL, 100 ms, 10 ms raised-cosine ramps, with
    `prepend` zeros so the reference layout emits the whole tone."""
    return np.concatenate([np.zeros(prepend, np.float32), tone(1000.0, 60.0, 0.1)])


def c2_grid():
    """32 log-spaced frequencies 250-8000 Hz x 8 levels 20-90 dB (256 tones)."""
    freqs = 250.0 * 32.0 ** (np.arange(32) / 31.0)
    levels = np.arange(20, 91, 10, dtype=np.float64)
    f, l = np.meshgrid(freqs, levels, indexing="ij")
    return f.ravel(), l.ravel()


def c2_signals(duration: float = 0.2, n: int | None = None) -> np.ndarray:
    f, l = c2_grid()
    if n is not None:
        f, l = f[:n], l[:n]
    return np.stack([tone(fi, li, duration) for fi, li in zip(f, l)])


def natural_cubic_spline(xk: np.ndarray, yk: np.ndarray, x: np.ndarray) -> np.ndarray:
    from scipy.interpolate import CubicSpline
    return CubicSpline(xk, yk, bc_type="natural")(x)


def c3_signals(n_signals: int = 1024, seed: int = MASTER_SEED + 3, fs: float = FS):
    """Uneven batch: lengths U[20,500] ms; tones (125-10000 Hz log-uniform),
    3-harmonic complexes or Gaussian noise; levels U[0,100] dB; per-signal OHC
    spline in [0,0.5] and IHC in [5,8]."""
    rng = np.random.default_rng(seed)
    sigs, ohc, ihc, meta = [], [], [], []
    knots_x = np.linspace(0, 255, 8)
    xs = np.arange(256)
    for _ in range(n_signals):
        n = int(round(rng.uniform(0.020, 0.500) * fs))
        kind = int(rng.integers(0, 3))
        f = float(np.exp(rng.uniform(np.log(125.0), np.log(10000.0))))
        db = float(rng.uniform(0.0, 100.0))
        t = np.arange(n) / fs
        if kind == 0:
            x = db_to_peak(db) * np.sin(2 * np.pi * f * t)
        elif kind == 1:
            x = db_to_peak(db) / 3.0 * (np.sin(2 * np.pi * f * t) + np.sin(4 * np.pi * f * t)
                                        + np.sin(6 * np.pi * f * t))
        else:
            x = 10.0 * SPL_REF * 10.0 ** (db / 20.0) * rng.standard_normal(n)
        sigs.append(x.astype(np.float32))
        yk = rng.uniform(0.0, 0.5, 8)
        ohc.append(np.clip(natural_cubic_spline(knots_x, yk, xs), 0.0, 0.5).astype(np.float32))
        ihc.append(float(rng.uniform(5.0, 8.0)))
        meta.append((kind, f, db, n))
    return sigs, np.stack(ohc), np.asarray(ihc, np.float32), meta
