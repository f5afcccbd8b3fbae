"""Where the tolerance-mode Jacobi sweeps go (CPU oracle, traced passes).

Single tones (50 ms continuous runs from rest, tau = 1e-6) traced pass by pass:
with the previous pressure as the guess, the predictor pass of each step (the
Euler pass, whose right-hand side moves furthest) takes most of the sweeps at
low and mid levels; the extrapolated guess p + (dt/dt_prev)(p - p_prev) cuts
those predictor sweeps by well over half at those levels (by a third at 90 dB,
where the corrector passes hold as many sweeps) and leaves the corrector
passes alone.  This is why the learned warm start targets the predictor pass only.

Run as a script for the table quoted in the documentation:
    python tests/test_sweep_split.py
"""
import os
import sys

import numpy as np
import pytest

FS = 44100.0
TONES = [(500.0, 30.0), (1000.0, 60.0), (2000.0, 90.0)]
N = 2205


def _tone(f, db, n):
    t = np.arange(n) / FS
    return (10 * 2e-5 * 10 ** (db / 20) * np.sin(2 * np.pi * f * t)).astype(np.float32)


def split(oracle, f, db, warm_start):
    inp = np.zeros(N + 4, np.float32)
    inp[:N] = _tone(f, db, N)
    it = oracle.make_items([0], [0], [N], [0], 1e-10, 1e-15)
    sc = oracle.default_scalars(FS)
    sc.jacobi_mode, sc.tol_rel, sc.check_every, sc.max_sweeps, sc.warm_start = 1, 1e-6, 4, 1000, warm_start
    rec = oracle.bm_trace(sc, oracle.section_params(), np.full(256, 0.5, np.float32), it, inp, 400000)
    steps = len(np.unique(rec["step"]))
    pred = rec["sweeps"][rec["pass_"] == 0].sum()
    corr = rec["sweeps"][rec["pass_"] > 0].sum()
    return pred / steps, corr / steps, len(rec) / steps


@pytest.fixture(scope="module")
def table(oracle):
    return {(f, db): {ws: split(oracle, f, db, ws) for ws in (0, 1)} for f, db in TONES}


def test_predictor_pass_dominates_plain_sweeps(table):
    for (f, db), t in table.items():
        pred, corr, _ = t[0]
        if db <= 60:
            assert pred > 0.6 * (pred + corr), (f, db, pred, corr)
        assert pred > 0.3 * (pred + corr), (f, db, pred, corr)


def test_extrapolation_cuts_predictor_sweeps(table):
    for (f, db), t in table.items():
        p0, c0, n0 = t[0]
        p1, c1, n1 = t[1]
        assert p1 < (0.5 if db <= 60 else 0.75) * p0, (f, db, p0, p1)
        # the corrector passes and the pass count are untouched to within a few percent
        assert abs(c1 - c0) <= 0.1 * c0 + 1.0, (f, db, c0, c1)
        assert abs(n1 - n0) <= 0.02 * n0, (f, db, n0, n1)


if __name__ == "__main__":
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from oracle import oracle as O
    O.build()
    print("| tone | guess | predictor sweeps/step | corrector sweeps/step | passes/step |")
    print("|---|---|---|---|---|")
    for f, db in TONES:
        for ws, name in ((0, "previous pressure"), (1, "extrapolated")):
            p, c, n = split(O, f, db, ws)
            print(f"| {f:g} Hz / {db:g} dB | {name} | {p:.1f} | {c:.1f} | {n:.2f} |")
