"""The oracle against its committed known answers (tests/golden/oracle_c1.npz,
written by tests/golden/make_oracle_fixtures.py): the one-tone VOICE case
through the FP32 kernel-17 method, the exact per-pass solve and the FP64
rebuild, and IHC/lambda of the FP32 velocity.  Any change to the oracle's
arithmetic (an operation reordered, a constant retyped, a branch of the step
control moved) changes a digest here."""
import hashlib
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_oracle_fixtures as G  # noqa: E402

FIXTURE = os.path.join(HERE, "golden", "oracle_c1.npz")


def _sha(a):
    return np.frombuffer(hashlib.sha256(np.ascontiguousarray(a).tobytes()).digest(), np.uint8)


@pytest.fixture(scope="module")
def fx():
    return dict(np.load(FIXTURE, allow_pickle=False))


def _items(oracle, fx):
    return fx["items"].reshape(-1).view(oracle.ITEM_DTYPE).copy()


def test_fixture_input_is_the_c1_case(oracle, fx):
    lay, inp, items, _ = G.case()
    np.testing.assert_array_equal(fx["input"], inp)
    it = _items(oracle, fx)
    for f in ("start", "mosi", "mesi", "write_start"):
        np.testing.assert_array_equal(it[f], items[f])


@pytest.mark.parametrize("name,prec,mode", [("o1", "f32", 0), ("o2", "f32", 2), ("f64", "f64", 0)])
def test_oracle_reproduces_known_answers(oracle, fx, name, prec, mode):
    o = G.run_oracle(oracle, fx["input"], _items(oracle, fx), 5512, prec, mode)
    v = o["speeds"][G.ROW0:G.ROW1]
    np.testing.assert_array_equal(v[::G.STRIDE_V], fx[f"{name}_rows"])
    np.testing.assert_array_equal(np.stack([o["steps"], o["passes"], o["sweeps"]], 1), fx[f"{name}_stats"])
    np.testing.assert_array_equal(_sha(v), fx[f"{name}_sha"])


def test_lambda_known_answers(oracle, fx):
    lay, _, _, ihc = G.case()
    o = G.run_oracle(oracle, fx["input"], _items(oracle, fx), 5512)
    lam = G.lambdas(oracle, o["speeds"], lay, ihc)
    np.testing.assert_array_equal(lam[:, ::G.STRIDE_L], fx["lambda_rows"])
    np.testing.assert_array_equal(_sha(lam), fx["lambda_sha"])


def test_known_answers_hold_the_method_relations(fx):
    """What the stored answers must satisfy whatever produced them: the tone is
    louder than silence everywhere it plays, O1 and O2 differ by the Jacobi
    truncation (well under 1%), the FP64 rebuild by the sample clock
    (percent level), and every block commits more than 22 steps per sample."""
    o1, o2, f64 = fx["o1_rows"].astype(np.float64), fx["o2_rows"].astype(np.float64), fx["f64_rows"]
    assert np.all(np.abs(o1).max(axis=1)[1:] > 0)
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)
    assert rel(o1, o2) < 1e-2
    assert 1e-3 < rel(o1, f64) < 1e-1
    assert np.all(fx["o1_stats"][:, 0] >= 22 * 882)
    # sweeps = 20 + 20 + 2 (passes - 2) per step for the fixed 20/2 method
    st = fx["o1_stats"]
    np.testing.assert_array_equal(st[:, 2], 40 * st[:, 0] + 2 * (st[:, 1] - 2 * st[:, 0]))
