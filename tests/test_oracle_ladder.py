"""Pins of the oracle's kernel-17 step control (include/cochlea.cuh:14021-14129,
14212-14230) with hand-derived outcomes.

The GPU kernel equals the oracle bit for bit, so these pins are what fix the
product's step control to the reference's.  Three layers:

* one pass's decision on hand-made section vectors (``oracle.step_control``),
  where each case separates the ladder from one plausible misreading of it;
* whole runs forced into one branch through the item tolerances, whose pass,
  step and dt counts follow from the ladder and the FP32 sample clock alone;
* a traced high-level run that reaches every branch, checked pass by pass
  against the ladder, including the commit-then-halve rule (Q2).

``tests/test_oracle_mutants.py`` rebuilds the oracle with each misreading
compiled in and requires this file to fail on every one of them.
"""
import numpy as np
import pytest

Z = np.zeros(256, np.float32)
DT = np.float32(1e-6)
DT_MIN = np.float32(1e-18)


def _one_hot(i, v):
    x = Z.copy()
    x[i] = v
    return x


def _decide(oracle, aref, na, vref, nv, dt=DT, tol=1e-10, m1tol=1e-15, dt_min=DT_MIN, dt_max=DT, mode=0):
    return oracle.step_control(aref, na, vref, nv, float(dt), float(tol), float(m1tol), float(dt_min),
                               float(dt_max), mode)


# ------------------------------------------------ one pass, hand vectors --
def test_halving_wins_over_fault(oracle):
    """L = 1e4 * 1e-6 / 1e-3 = 10 > 6 and F > 0: the step exits with dt/2
    (L > 6 is tested before F > 0)."""
    d = _decide(oracle, _one_hot(0, 1e4), Z, Z, _one_hot(0, 1e-3))
    assert 9.99 < d["L"] < 10.01 and d["F"] > 0
    assert (d["another"], d["reset"]) == (0, 1)
    assert d["dt"] == np.float32(5e-7)


def test_halving_floors_at_dt_min(oracle):
    """dt = 1.5e-18 with L = 1e16 * 1.5e-18 / 1e-3 = 15: the halved 7.5e-19 is
    floored at dt_min = 1e-18."""
    dt = np.float32(1.5e-18)
    d = _decide(oracle, _one_hot(0, 1e16), Z, Z, _one_hot(0, 1e-3), dt=dt, dt_max=1e-6)
    assert 14.9 < d["L"] < 15.1 and (d["another"], d["reset"]) == (0, 1)
    assert d["dt"] == DT_MIN


def test_middle_band_iterates(oracle):
    """0.5 <= L = 3 <= 6 with F <= 0: iterate, dt unchanged (no middle-band
    accept; the CPU ode::Tester's 2.0 threshold would halve here)."""
    d = _decide(oracle, _one_hot(0, 3e3), Z, Z, _one_hot(0, 1e-3), tol=1.0)
    assert 2.99 < d["L"] < 3.01 and d["F"] < 0
    assert (d["another"], d["reset"]) == (1, 0)
    assert d["dt"] == DT


def test_band_edges_exact(oracle):
    """L exactly 6 and exactly 0.5 both iterate (strict comparisons)."""
    dt = np.float32(2.0 ** -20)
    ev = np.float32(2.0 ** -10)
    for L, ea in ((6.0, 6144.0), (0.5, 512.0)):
        d = _decide(oracle, _one_hot(5, ea), Z, Z, _one_hot(5, ev), dt=dt, tol=1.0, dt_max=1e-6)
        assert d["L"] == np.float32(L)
        assert (d["another"], d["reset"]) == (1, 0), L
        assert d["dt"] == dt


def test_small_L_accepts_and_doubles(oracle):
    """L = 0.4 < 0.5 with F <= 0: exit with 2 dt, capped at dt_max (the
    Tester's 0.25 threshold would keep dt)."""
    dt = np.float32(2.5e-7)
    d = _decide(oracle, _one_hot(0, 4e2 / 0.25), Z, Z, _one_hot(0, 1e-3), dt=dt, tol=1.0)
    assert 0.399 < d["L"] < 0.401
    assert (d["another"], d["reset"]) == (0, 1)
    assert d["dt"] == np.float32(5e-7)
    d = _decide(oracle, _one_hot(0, 4e2), Z, Z, _one_hot(0, 1e-3), dt=DT, tol=1.0)
    assert d["L"] < 0.5 and d["dt"] == DT      # 2e-6 capped at dt_max = 1e-6


def test_small_L_with_fault_iterates(oracle):
    """F > 0 is tested before L < 0.5."""
    d = _decide(oracle, Z, Z, Z, _one_hot(0, 1e-3))
    assert d["L"] == 0 and d["F"] > 0
    assert (d["another"], d["reset"]) == (1, 0) and d["dt"] == DT


def test_zero_fault_accepts(oracle):
    """F == 0 exactly is not a fault (strict F > 0): all-zero vectors with
    tol = 0 give L = 0, F = 0 and accept with 2 dt."""
    d = _decide(oracle, Z, Z, Z, Z, dt=np.float32(2.5e-7), tol=0.0)
    assert d["L"] == 0 and d["F"] == 0
    assert (d["another"], d["reset"]) == (0, 1) and d["dt"] == np.float32(5e-7)


def test_fault_uses_signed_velocity(oracle):
    """Q4: max(tol, v dt) with v signed.  v = -1024 everywhere, |dv| = 2^-7,
    tol = 2^-8, dt = 2^-10: each section adds 2^-7 - max(2^-8, -1) = 2^-8, so
    F = 256 * 2^-8 = 1 > 0 and the pass iterates.  With |v| each section would
    add 2^-7 - 1 < 0 and the pass would accept."""
    nv = np.full(256, -1024.0, np.float32)
    vref = nv - np.float32(2.0 ** -7)
    d = _decide(oracle, Z, Z, vref, nv, dt=np.float32(2.0 ** -10), tol=2.0 ** -8, dt_max=1.0)
    assert d["L"] == 0 and d["e_v"] == 2.0
    assert d["F"] == 1.0
    assert (d["another"], d["reset"]) == (1, 0)


def test_fault_is_a_sum_not_a_count(oracle):
    """Q4: F sums the excess.  One section over the limit by 1e-4 and 255 under
    it by 1e-4 give F = -254e-4 < 0: accept (a count of violators would
    iterate)."""
    d = _decide(oracle, Z, Z, Z, _one_hot(0, 2e-4), tol=1e-4, m1tol=1.0)
    assert d["L"] == 0
    assert d["F"] == pytest.approx(1e-4 - 255e-4, rel=1e-5)
    assert (d["another"], d["reset"]) == (0, 1)


def test_m1tol_is_strict(oracle):
    """Q5: L is computed only when e_v > m1tol.  e_v = 0.5 = m1tol gives L = 0
    and accepts; were e_v >= m1tol enough, L = 1.5e6 * 1e-6 / 0.5 = 3 would
    iterate."""
    d = _decide(oracle, _one_hot(7, 1.5e6), Z, Z, _one_hot(7, 0.5), tol=1.0, m1tol=0.5)
    assert d["e_v"] == np.float32(0.5) and d["L"] == 0
    assert (d["another"], d["reset"]) == (0, 1)
    d = _decide(oracle, _one_hot(7, 1.5e6), Z, Z, _one_hot(7, 0.5), tol=1.0, m1tol=0.4999)
    assert d["L"] == pytest.approx(3.0, rel=1e-6) and (d["another"], d["reset"]) == (1, 0)


def test_acceleration_error_is_a_max(oracle):
    """Q5: e_a = max |a_ref - a|.  |da| = 1 on every section with e_v = 1e-6
    gives L = 1 (iterate); a sum would give L = 256 and halve."""
    d = _decide(oracle, np.ones(256, np.float32), Z, Z, _one_hot(0, 1e-6), tol=1.0)
    assert d["e_a"] == 1.0 and d["L"] == 1.0
    assert (d["another"], d["reset"]) == (1, 0) and d["dt"] == DT


def test_zero_lane_decision(oracle):
    """decision_mode 1 (Q1 zero-partial lane): reductions read as zeros, so
    L = 0, F = 0 and the pass accepts with 2 dt whatever the vectors say."""
    d = _decide(oracle, _one_hot(0, 1e4), Z, Z, _one_hot(0, 1e-3), dt=np.float32(2.5e-7), mode=1)
    assert d["L"] == 0 and d["F"] == 0
    assert (d["another"], d["reset"]) == (0, 1) and d["dt"] == np.float32(5e-7)


# -------------------------------------------- whole runs, forced branches --
def _clock(n_samples, dt, tso):
    """FP32 sample clock of the kernel (Q9, :14145-14151, 14230): a step
    advances the sample when t > tso (then t -= tso), and t += dt after every
    committed step.  Returns committed steps per sample."""
    t = np.float32(0)
    dt, tso = np.float32(dt), np.float32(tso)
    per = np.zeros(n_samples, np.int64)
    s, k = 0, 0
    while s < n_samples:
        k += 1
        if t > tso:
            t = np.float32(t - tso)
            per[s] = k
            k = 0
            s += 1
        t = np.float32(t + dt)
    return per


def _forced_run(oracle, sig, n, tol, m1tol, **kw):
    sc = oracle.default_scalars(44100.0)
    for k, v in kw.items():
        setattr(sc, k, v)
    inp = np.zeros(n + 4, np.float32)
    inp[:len(sig)] = sig
    it = oracle.make_items([0], [0], [n], [0], tol, m1tol)
    out = oracle.bm_run(sc, oracle.section_params(), np.full(256, 0.5, np.float32), it, inp, n, "f32",
                        threads=1, diagnostics=True)
    return sc, out


def test_lipschitz_off_two_passes_at_dt_max(oracle):
    """m1tol = tol = 1e30 force L = 0 and F < 0 on every corrector pass: each
    step is the predictor plus one corrector, dt stays at dt_max, and the
    steps per sample are exactly the sample clock's."""
    n = 300
    t = np.arange(n) / 44100.0
    sig = (0.2 * np.sin(2 * np.pi * 3000 * t)).astype(np.float32)   # 80 dB SPL peak
    sc, out = _forced_run(oracle, sig, n, 1e30, 1e30)
    per = _clock(n, sc.dt_max, sc.time_step_out)
    assert int(out["steps"][0]) == int(per.sum())
    assert int(out["passes"][0]) == 2 * int(per.sum())
    np.testing.assert_array_equal(out["conv_loops"][:n], per.astype(np.float32))
    np.testing.assert_allclose(out["conv_time"][:n], sc.dt_max, rtol=1e-6)
    assert not out["failed_nodes"][:n].any()
    assert np.abs(out["speeds"][:n]).max() > 0


def test_undefined_lipschitz_iterates_to_the_cap(oracle):
    """Silent input with m1tol = -1: e_v = 0 > m1tol, so L = 0 * dt / 0 = NaN,
    which satisfies no comparison of the ladder and falls to 'iterate'.
    Every step then runs Cuda_Outern_Loops passes and is flagged failed, and
    dt never changes."""
    n = 40
    sc, out = _forced_run(oracle, np.zeros(1, np.float32), n, 1e-10, -1.0, max_passes=7)
    per = _clock(n, sc.time_step, sc.time_step_out)
    assert int(out["steps"][0]) == int(per.sum())
    assert int(out["passes"][0]) == 7 * int(per.sum())
    assert np.all(out["failed_nodes"][:n] == 1)
    np.testing.assert_array_equal(out["conv_loops"][:n], per.astype(np.float32))
    assert np.all(out["speeds"][:n] == 0)


# ------------------------------------------------ traced run, every branch --
def test_traced_run_follows_the_ladder(oracle):
    """A 100 dB 4 kHz tone reaches every branch of the ladder.  Pass by pass:
    the decision and dt follow the ladder from the pass's own L and F, and a
    halving pass commits its step (the next pass is the next step's
    predictor) with local time advanced by the halved dt (Q2)."""
    sc = oracle.default_scalars(44100.0)
    n = 400
    t = np.arange(n) / 44100.0
    inp = np.zeros(n + 4, np.float32)
    inp[:n] = (10 * 2e-5 * 10 ** 5 * np.sin(2 * np.pi * 4000 * t)).astype(np.float32)
    it = oracle.make_items([0], [0], [n], [0], 1e-10, 1e-15)
    rec = oracle.bm_trace(sc, oracle.section_params(), np.full(256, 0.5, np.float32), it, inp, 400000)
    tso = np.float32(sc.time_step_out)
    seen = {"halve": 0, "fault": 0, "double": 0, "middle": 0, "cap": 0}
    for i, r in enumerate(rec):
        nxt = rec[i + 1] if i + 1 < len(rec) else None
        if not r["decided"]:
            assert r["pass_"] == 0 and r["another"] == 1 and r["dt_out"] == r["dt_in"]
            continue
        L, F, dti = r["L"], r["F"], r["dt_in"]
        if r["capped"]:
            seen["cap"] += 1
            assert r["another"] == 0
            continue
        if L > 6:
            seen["halve"] += 1
            assert (r["another"], r["reset"]) == (0, 1)
            assert r["dt_out"] == max(np.float32(dti * np.float32(0.5)), np.float32(sc.dt_min))
            assert nxt is not None and nxt["pass_"] == 0 and nxt["step"] == r["step"] + 1
            t_next = np.float32(r["t"] - tso) if r["t"] > tso else r["t"]
            assert nxt["t"] == np.float32(t_next + r["dt_out"])
        elif F > 0:
            seen["fault"] += 1
            assert r["another"] == 1 and r["dt_out"] == dti
        elif L < 0.5:
            seen["double"] += 1
            assert (r["another"], r["reset"]) == (0, 1)
            assert r["dt_out"] == min(np.float32(dti * np.float32(2)), np.float32(sc.dt_max))
            assert nxt is None or (nxt["pass_"] == 0 and nxt["step"] == r["step"] + 1)
        else:
            seen["middle"] += 1
            assert r["another"] == 1 and r["dt_out"] == dti
            assert nxt["step"] == r["step"] and nxt["pass_"] == r["pass_"] + 1
    assert seen["halve"] >= 1 and seen["middle"] >= 10 and seen["fault"] >= 100 and seen["double"] >= 100, seen
    # every committed step advanced local time by the dt its last pass left
    steps = rec[rec["pass_"] == 0]
    assert len(steps) == rec["step"].max() + 1
