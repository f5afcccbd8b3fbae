"""The C-ABI library loads and exports every function include/cochlea_amd.h declares (no GPU calls)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    src = open(os.path.join(ROOT, "include", "cochlea_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?(?:int|void|float|double|char)\s*\*?\s*([A-Za-z_]\w*)\s*\(", src, re.M)
    return sorted(set(names))


def test_header_declares_reference_entry_points():
    names = header_functions()
    for ref in ("BMOHCKernel_Init", "BMOHCNewKernel", "BMOHCKernel_Copy_Results", "IHCNewKernel", "RunIHCKernel",
                "CudaCalculateJND", "generateInputFromProfile", "setupToleranceProfile", "uploadProfiles",
                "BMOHCKernel_Free", "IHCKernel_Free", "BMOHCKernel_Copy_Lambda", "loadAihc"):
        assert ref in names, ref


def test_library_exports_every_declared_symbol():
    from cochlearmodel_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True, check=True)
    exported = {line.split()[-1] for line in out.stdout.splitlines() if " T " in line}
    missing = [n for n in header_functions() if n not in exported]
    assert not missing, missing
    L = _lib.lib()          # dlopen + every ctypes signature binds
    assert set(_lib.declared_symbols()) <= set(header_functions())
    assert set(header_functions()) <= set(_lib.declared_symbols())


def test_host_model_calls_without_gpu():
    """Host-only entry points work on a machine with no GPU."""
    import numpy as np
    from cochlearmodel_amd import solver, stages
    v = solver.model_vectors()
    assert v.shape == (7, 256) and np.all(v[0] > 0)
    lo, mid, up = solver.tri_matrix()
    assert lo[254] == 0 and lo[255] == 0 and mid[255] == 1.0
    prof, refs, serial, calc, to_ref = stages.jnd_profiles([1000.0, 2000.0], [10.0, 20.0], noises=(1111.0, 40.0))
    assert len(prof) == 10 and list(refs) == [4, 9]
    assert all(to_ref[serial[i]] == (4 if i < 5 else 9) for i in calc)


def test_bm_run_rejects_invalid_arguments_before_any_launch():
    """cm_bm_run validates its arguments on the host: null buffers, negative
    counts, pass and sweep limits, warm-start modes and the disabled
    persistent launch are refused with a status and a message, and an empty
    batch is a no-op.  Every case returns before a device call, so this runs
    without a GPU."""
    import ctypes as C
    from cochlearmodel_amd import _lib, solver
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    L = _lib.lib()
    fake = C.c_void_p(0x1000)   # never dereferenced: every case below returns first

    def run(sc, n_items=1, inp=fake, persistent=0):
        return L.cm_bm_run(C.byref(sc), inp, 16, fake, fake, fake, n_items, fake, 12, None, None, None, None, None,
                           None, None, None, None, 1, persistent, 0)

    def msg():
        return L.cm_last_error().decode()

    sc = solver.default_scalars()
    assert run(sc, inp=None) == -1 and "null argument" in msg()
    assert run(sc, n_items=-1) == -1
    assert run(sc, persistent=1) == -3 and "disabled" in msg()
    for field, value in (("j_fast", -1), ("j_slow", -1), ("max_passes", 0)):
        bad = solver.default_scalars()
        setattr(bad, field, value)
        assert run(bad) == -3 and "Jacobi/pass" in msg(), field
    for ws, check_every in ((3, 4), (-1, 4), (1, 0)):
        bad = solver.default_scalars()
        bad.jacobi_mode, bad.warm_start, bad.check_every, bad.tol_rel, bad.max_sweeps = 1, ws, check_every, 1e-6, 1000
        assert run(bad) == -3 and "warm_start" in msg(), (ws, check_every)
    assert run(sc, n_items=0) == 0
    # launch knobs are host state: bad values are refused, good ones kept
    assert L.cm_set_bm_sections_per_lane(3) == -3 and "1, 2 or 4" in msg()
    assert L.cm_set_bm_jacobi_unroll(3) == -3
    for n in (1, 2, 4):
        assert L.cm_set_bm_sections_per_lane(n) == 0
    assert L.cm_ws_set_group(10) == -3 and "8, 12 or 16" in msg()
    for g in (8, 12, 16, 0):
        assert L.cm_ws_set_group(g) == 0


def test_batched_stage_entry_points_reject_bad_sizes_before_any_launch():
    """The batched IHC/lambda, JND and input-synthesis entry points refuse
    negative sizes (and an empty JND window, Fs <= 0) with -3 and a message,
    and take an empty batch as a no-op; every case returns before a device
    call, so this runs without a GPU."""
    import ctypes as C
    from cochlearmodel_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    L = _lib.lib()
    f = C.c_void_p(0x1000)   # never dereferenced
    sp = (C.c_float * 3)(0.0, 0.0, 0.0)
    W = (C.c_float * 3)(1.0, 1.0, 1.0)
    msg = lambda: L.cm_last_error().decode()

    def ihc(rows, n_ivs, n_iir=0):
        return L.cm_ihc_lambda_run(f, rows, f, n_ivs, None, n_iir, f, None, 76, f, 88, f, f, sp, C.c_float(500),
                                   C.c_float(1), C.c_float(100), C.c_float(1), f, f, None)

    assert ihc(100, -1) == -3 and "negative count" in msg()
    assert ihc(-5, 1) == -3 and "negative count" in msg()
    assert ihc(100, 1, n_iir=-2) == -3
    assert ihc(100, 0) == 0
    assert L.cm_ihc_lambda_run_noalloc(f, 100, f, -1, f, 76, f, 88, f, f, sp, C.c_float(500), C.c_float(1),
                                       C.c_float(100), C.c_float(1), f, None) == -3
    assert L.cm_ihc_lambda_run_noalloc(f, 100, None, 0, f, 76, f, 88, f, f, sp, C.c_float(500), C.c_float(1),
                                       C.c_float(100), C.c_float(1), f, None) == 0

    def jnd(n, nodes, head, length, stride=1, rows=10000, fs=44100.0):
        return L.cm_jnd_run(f, rows, n, nodes, head, length, f, f, f, stride, f, W, fs, 0.2 * fs, 1, 1, f, f, f)

    for args in ((-1, 100, 0, 50), (2, -100, 0, 50), (2, 100, -1, 50), (2, 100, 0, 0), (2, 100, 0, 50, -1)):
        assert jnd(*args) == -3, args
    assert jnd(2, 100, 0, 50, fs=0.0) == -3
    assert jnd(200, 100, 0, 50) == -3 and "exceed" in msg()
    assert jnd(2, 100, 60, 50) == -3 and "exceed" in msg()
    assert jnd(0, 100, 0, 50) == 0

    def gen(start, n, length, fs=44100.0):
        return L.cm_generate_input_run(f, f, None, f, 0.0, 1.0, 0, 0.0, 1.0, start, n, length, 0, C.c_float(fs))

    for args in ((-1, 2, 100), (0, -2, 100), (0, 2, -100)):
        assert gen(*args) == -3 and "negative size" in msg(), args
    assert gen(0, 2, 100, fs=0.0) == -3
    assert gen(0, 0, 100) == 0
