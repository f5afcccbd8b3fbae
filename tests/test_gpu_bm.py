"""GPU parity of the kernel-17 BM solver against the CPU oracle (through the C-ABI)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _flush(msg):
    print(msg, flush=True)


def test_c1_voice_parity(oracle):
    from cochlearmodel_amd import solver, workloads
    _flush("case: C1 voice tone, 5 blocks x 882 + 220 overlap")
    sig = workloads.c1_signal()
    lay = solver.voice_layout(44100.0, 5, 0.02, 0.005)
    assert (lay.block_nodes, lay.overlap_nodes, lay.bufP1) == (882, 220, 5512)
    g = solver.run_bm_voice(sig, lay)
    # oracle on the same inputs
    sc = oracle.default_scalars(44100.0)
    st, mo, me, ws = oracle.block_decomposition(5, lay.bufP1, lay.overlap_nodes, 0)
    np.testing.assert_array_equal(g["blocks"]["start"], st)
    np.testing.assert_array_equal(g["blocks"]["mosi"], mo)
    np.testing.assert_array_equal(g["blocks"]["mesi"], me)
    np.testing.assert_array_equal(g["blocks"]["write_start"], ws)
    items = oracle.make_items(st, mo, me, ws, 1e-10, 1e-15)
    o = oracle.bm_run(sc, oracle.section_params(), np.full(256, 0.5, np.float32), items, g["input"], lay.bufP1,
                      "f32", threads=8, diagnostics=True)
    rows = slice(lay.overlap_nodes, lay.total_nodes)
    vg, vo = g["speeds"][rows], o["speeds"][rows]
    rel = np.linalg.norm(vg - vo) / np.linalg.norm(vo)
    steps_g, steps_o = g["stats"]["steps"].sum(), o["steps"].sum()
    _flush(f"rel L2 {rel:.3e}, exact rows {np.mean(np.all(vg == vo, axis=1)):.4f}, steps {steps_g} vs {steps_o}")
    assert rel <= 1e-4
    assert abs(int(steps_g) - int(steps_o)) <= 0.01 * steps_o
    np.testing.assert_array_equal(g["failed_nodes"][rows], o["failed_nodes"][rows])
