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
