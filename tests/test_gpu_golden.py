"""The library on the GPU against the oracle's committed known answers
(tests/golden/oracle_c1.npz): the one-tone VOICE case through the reference
call sequence, BM velocity and IHC/lambda bit for bit."""
import hashlib
import os
import sys

import numpy as np

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_oracle_fixtures as G  # noqa: E402


def _sha(a):
    return np.frombuffer(hashlib.sha256(np.ascontiguousarray(a).tobytes()).digest(), np.uint8)


@pytest.mark.parametrize("sections_per_lane", [4, 2, 1])
def test_voice_run_matches_known_answers(sections_per_lane):
    """Every block layout (4 sections per lane on one wave; 2 or 1 on a
    workgroup of 2 or 4 waves) reproduces the committed answers bit for bit."""
    from cochlearmodel_amd import _lib, solver, stages
    print(f"case: golden C1 VOICE run, 5 blocks, BM + IHC/lambda, {sections_per_lane} sections per lane", flush=True)
    fx = dict(np.load(os.path.join(HERE, "golden", "oracle_c1.npz"), allow_pickle=False))
    lay = solver.voice_layout(G.FS, 5, 0.02, 0.005)
    L = _lib.lib()
    _lib.check(L.cm_set_bm_sections_per_lane(sections_per_lane))
    try:
        g = solver.run_bm_voice(fx["input"][:lay.bufP1], lay)
    finally:
        _lib.check(L.cm_set_bm_sections_per_lane(4))
    v = g["speeds"][G.ROW0:G.ROW1]
    np.testing.assert_array_equal(v[::G.STRIDE_V], fx["o1_rows"])
    st = g["stats"]
    np.testing.assert_array_equal(np.stack([st["steps"], st["passes"], st["sweeps"]], 1), fx["o1_stats"])
    np.testing.assert_array_equal(_sha(v), fx["o1_sha"])
    rows = lay.total_nodes
    lam, _ = stages.run_ihc_reference(rows, lay.block_nodes, lay.time_blocks, 0)
    np.testing.assert_array_equal(lam[:, ::G.STRIDE_L], fx["lambda_rows"])
    np.testing.assert_array_equal(_sha(lam), fx["lambda_sha"])
