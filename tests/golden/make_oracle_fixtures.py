"""Writes tests/golden/oracle_c1.npz: the CPU oracle's outputs for the one-tone
VOICE case (1 kHz, 60 dB SPL, 100 ms with 10 ms ramps, 5 time blocks of 20 ms
+ 5 ms overlap at 44.1 kHz, OHC 0.5, IHC 8, the committed equiripple AC
filter), as committed known answers.

Stored: the input and block table; per-block step/pass/sweep counts and a
SHA-256 of the velocity rows 220..4629 for the FP32 kernel-17 method (O1), the
exact per-pass solve (O2) and the FP64 rebuild; every 400th of those rows as
values; the SHA-256 of the three lambda planes of O1's velocity and every 800th
row of them.  tests/test_golden_fixtures.py re-runs the oracle against these
(CPU) and tests/test_gpu_golden.py runs the library against them (GPU).

Run: python tests/golden/make_oracle_fixtures.py
"""
import hashlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
OUT = os.path.join(ROOT, "tests", "golden", "oracle_c1.npz")
FS = 44100.0
ROW0, ROW1 = 220, 4630          # rows of the tone: overlap .. overlap + 5 blocks
STRIDE_V, STRIDE_L = 400, 800   # sampled rows of velocity and lambda


def sha(a) -> np.ndarray:
    return np.frombuffer(hashlib.sha256(np.ascontiguousarray(a).tobytes()).digest(), np.uint8).copy()


def case():
    """Input, block table and IHC/lambda parameters of the fixture."""
    from cochlearmodel_amd import solver, stages, workloads
    lay = solver.voice_layout(FS, 5, 0.02, 0.005)
    inp = np.zeros(lay.bufP1 + 2, np.float32)
    sig = workloads.c1_signal()
    inp[:min(len(sig), lay.bufP1)] = sig[:lay.bufP1]
    items = solver.block_table(5, lay.bufP1, lay.overlap_nodes, 0)
    items["tol"], items["m1tol"] = 1e-10, 1e-15
    b, _ = stages.load_ac_filter(FS)
    ihc = {"ac_b": np.asarray(b, np.float64), "dc_b": np.asarray(stages.dc_filter(FS), np.float64),
           "damage": np.asarray(stages.ihc_damage(8.0), np.float64), "A": np.asarray(stages.aihc_table(), np.float32),
           "spont": np.array(stages.SPONT_DEFAULT, np.float32)}
    return lay, inp, items, ihc


def run_oracle(O, inp, items, bufP1, precision="f32", mode=0, threads=8):
    sc = O.default_scalars(FS)
    sc.jacobi_mode = mode
    return O.bm_run(sc, O.section_params(), np.full(256, 0.5, np.float32), items, inp, bufP1, precision,
                    threads=threads)


def lambdas(O, speeds, lay, ihc, threads=8):
    rows = lay.total_nodes
    return O.ihc_lambda(speeds[:rows], rows, lay.block_nodes * lay.time_blocks, ihc["ac_b"], ihc["dc_b"],
                        ihc["damage"], ihc["A"], ihc["spont"], threads=threads)


def main():
    from oracle import oracle as O
    O.build()
    lay, inp, items, ihc = case()
    assert (lay.bufP1, lay.overlap_nodes, lay.total_nodes) == (5512, ROW0, ROW1)
    out = {"input": inp, "items": items.view(np.uint8).reshape(len(items), -1).copy()}
    for name, prec, mode in (("o1", "f32", 0), ("o2", "f32", 2), ("f64", "f64", 0)):
        o = run_oracle(O, inp, items, lay.bufP1, prec, mode)
        v = o["speeds"][ROW0:ROW1]
        out[f"{name}_sha"] = sha(v)
        out[f"{name}_rows"] = v[::STRIDE_V].copy()
        out[f"{name}_stats"] = np.stack([o["steps"], o["passes"], o["sweeps"]], 1).astype(np.int64)
        if name == "o1":
            lam = lambdas(O, o["speeds"], lay, ihc)
            out["lambda_sha"] = sha(lam)
            out["lambda_rows"] = lam[:, ::STRIDE_L].copy()
    np.savez_compressed(OUT, **out)
    print(OUT, os.path.getsize(OUT), "bytes;", {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
