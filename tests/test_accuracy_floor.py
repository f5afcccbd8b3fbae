"""Where the error against the exact per-pass solve comes from (CPU).

The benchmark scores every solver variant against O2, the kernel-17 method
with each pass's pressure solved exactly.  These tests pin the decomposition of
that error on stable tones (continuous 50 ms runs from rest):

* FP32 evaluation of the method is reproducible only to 1e-4 .. 3e-4 relative
  at 40-60 dB: a rounding-level change of the input (x (1 + 2^-20)) moves O2
  itself by that much, and so does evaluating it in FP64.  The FP64 evaluation
  moves by less than 1e-5 under the same change, so the floor is the FP32
  state arithmetic, not the pressure solve.  No FP32 variant can be held to
  1e-4 against O2 at these levels.
* The percent-level gap to the FP64 rebuild is the sample clock (t, the 22 us
  time_step_out and the interpolation weight kept in float, Q9): FP64 state
  with a float clock is within the FP32 floor of FP32, and a double clock moves
  it by about 1e-2.
* The step tolerance (MAX_TOLERANCE 1e-10 or 1e-13) does not move the floor.

Run as a script for the table quoted in the documentation:
    python tests/test_accuracy_floor.py
"""
import os
import sys

import numpy as np
import pytest

FS = 44100.0
TONES = [(500.0, 60.0), (1000.0, 40.0), (2000.0, 80.0), (700.0, 20.0)]
N = 2205                     # 50 ms per tone


def _tone(f, db, n):
    t = np.arange(n) / FS
    return (10 * 2e-5 * 10 ** (db / 20) * np.sin(2 * np.pi * f * t)).astype(np.float32)


def _setup(oracle, tones=TONES, n=N):
    inp = np.zeros(len(tones) * n + 4, np.float32)
    for k, (f, db) in enumerate(tones):
        inp[k * n:(k + 1) * n] = _tone(f, db, n)
    starts = [k * n for k in range(len(tones))]
    items = oracle.make_items(starts, starts, [s + n for s in starts], starts, 1e-10, 1e-15)
    return inp, items


def _run(oracle, inp, items, rows, prec="f32", mode=2, tau=1e-6, tol=1e-10, scale=1.0):
    sc = oracle.default_scalars(FS)
    sc.jacobi_mode, sc.tol_rel, sc.check_every, sc.max_sweeps = mode, tau, 4, 1000
    it = items.copy()
    it["tol"] = tol
    o = oracle.bm_run(sc, oracle.section_params(), np.full(256, 0.5, np.float32), it, inp * np.float32(scale),
                      rows, prec, threads=min(8, os.cpu_count() or 1))
    return o["speeds"].astype(np.float64) / scale


def _per_tone(a, b, n=N, k_tones=len(TONES)):
    return np.array([np.linalg.norm(a[k * n:(k + 1) * n] - b[k * n:(k + 1) * n]) /
                     np.linalg.norm(b[k * n:(k + 1) * n]) for k in range(k_tones)])


def decomposition(oracle, tols=(1e-10,)):
    inp, items = _setup(oracle)
    rows = len(TONES) * N
    out = {}
    pert = 1 + 2.0 ** -20
    for tol in tols:
        o2 = _run(oracle, inp, items, rows, tol=tol)
        o2c = _run(oracle, inp, items, rows, "f64c32", tol=tol)
        out[tol] = {
            "O2 FP32, input x(1+2^-20)": _per_tone(_run(oracle, inp, items, rows, tol=tol, scale=pert), o2),
            "O2 FP32 vs FP64 state (float clock)": _per_tone(o2, o2c),
            "O2 FP64 (float clock), input x(1+2^-20)":
                _per_tone(_run(oracle, inp, items, rows, "f64c32", tol=tol, scale=pert), o2c),
            "O2 FP64 float clock vs double clock": _per_tone(o2c, _run(oracle, inp, items, rows, "f64", tol=tol)),
            "reference 20/2 vs O2": _per_tone(_run(oracle, inp, items, rows, mode=0, tol=tol), o2),
            "plain tau 1e-6 vs O2": _per_tone(_run(oracle, inp, items, rows, mode=1, tol=tol), o2),
        }
    return out


@pytest.fixture(scope="module")
def table(oracle):
    return decomposition(oracle, tols=(1e-10, 1e-13))


def test_fp32_floor_and_clock_decomposition(table):
    d = table[1e-10]
    low = np.array([db <= 60 for _, db in TONES])
    # FP64 evaluation is insensitive to a rounding-level input change ...
    assert np.all(d["O2 FP64 (float clock), input x(1+2^-20)"] < 1e-5)
    # ... the FP32 evaluation is not: at 20-60 dB it moves by 1e-4 or more, and
    # by as much as it differs from the FP64 evaluation
    fp32 = d["O2 FP32, input x(1+2^-20)"]
    assert np.all(fp32[low] > 5e-5) and np.all(fp32 < 5e-4)
    gap = d["O2 FP32 vs FP64 state (float clock)"]
    assert np.all(gap < 5e-4) and np.all(gap < 3 * fp32 + 1e-6)
    # the clock's rounding accounts for the percent-level gap to the FP64 rebuild
    assert np.all(d["O2 FP64 float clock vs double clock"] > 5e-3)
    # the Jacobi variants sit within a few floors of O2
    assert np.all(d["plain tau 1e-6 vs O2"] < 5 * fp32 + 1e-4)


def test_step_tolerance_does_not_move_the_floor(table):
    a = table[1e-10]["O2 FP32, input x(1+2^-20)"]
    b = table[1e-13]["O2 FP32, input x(1+2^-20)"]
    low = np.array([db <= 60 for _, db in TONES])
    assert np.all(b[low] > 0.3 * a[low]) and np.all(b[low] < 3 * a[low])


def tau_sweep(oracle, taus=(1e-5, 1e-6, 1e-7, 3e-8)):
    """Plain tolerance-mode Jacobi at each relative residual tau against O2,
    with the sweeps per step it spends (1000-sweep cap per pass)."""
    inp, items = _setup(oracle)
    rows = len(TONES) * N
    o2 = _run(oracle, inp, items, rows)
    out = {}
    for tau in taus:
        sc = oracle.default_scalars(FS)
        sc.jacobi_mode, sc.tol_rel, sc.check_every, sc.max_sweeps = 1, tau, 4, 1000
        o = oracle.bm_run(sc, oracle.section_params(), np.full(256, 0.5, np.float32), items, inp, rows, "f32",
                          threads=min(8, os.cpu_count() or 1))
        out[tau] = (_per_tone(o["speeds"].astype(np.float64), o2), float(o["sweeps"].sum() / o["steps"].sum()))
    return out


if __name__ == "__main__":
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from oracle import oracle as O
    O.build()
    res = decomposition(O, tols=(1e-10, 1e-13))
    print("| quantity (rel L2 per tone) | " + " | ".join(f"{f:g} Hz / {db:g} dB" for f, db in TONES) + " |")
    print("|---|" + "---|" * len(TONES))
    for tol, d in res.items():
        for k, v in d.items():
            print(f"| {k}, MAX_TOLERANCE {tol:g} | " + " | ".join(f"{x:.2e}" for x in v) + " |")
    print()
    print("| plain tau vs O2 | " + " | ".join(f"{f:g} Hz / {db:g} dB" for f, db in TONES) + " | sweeps/step |")
    print("|---|" + "---|" * (len(TONES) + 1))
    for tau, (v, sps) in tau_sweep(O).items():
        print(f"| {tau:g} | " + " | ".join(f"{x:.2e}" for x in v) + f" | {sps:.1f} |")
