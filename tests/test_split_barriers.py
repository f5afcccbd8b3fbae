"""Barrier placement in the split-layout BM kernels (a time block over 2 or 4
waves), checked on the compiled gfx950 assembly.

Every wave of a block forms the same reduction totals and takes the same
branches, so each wave executes the same sequence of workgroup barriers as
long as no barrier sits in a region the wave would enter with an empty exec
mask.  The compiler skips a divergent region with s_cbranch_execz when it
keeps that branch; this test compiles bm_split_kernel.hip and checks, for
every exec-masked region opened by s_and_saveexec / s_andn2_saveexec that is
not followed by an execz skip, that no s_barrier comes before the region's
restore, and that exec is only ever narrowed by those saveexec forms (no
divergent loops).  A compiler change that broke this would hang a block on
the GPU; here it fails on the CPU.
"""
import os
import re
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(ROOT, "cochlearmodel_amd", "csrc", "bm_split_kernel.hip")


def _assembly(tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    out = tmp_path / "split.s"
    cmd = [hipcc, "--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off", "-Xclang", "-target-feature",
           "-Xclang", "-packed-fp32-ops", "-I", os.path.dirname(SRC), "-I", os.path.join(ROOT, "include"),
           "--cuda-device-only", "-S", SRC, "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    return out.read_text()


def _kernels(asm):
    names = re.findall(r"\n(_ZN2cm15bm_split_kernel\S+):", asm)
    for nm in names:
        i = asm.index("\n" + nm + ":")
        j = asm.index(".Lfunc_end", i)
        lines = [l.strip() for l in asm[i:j].split("\n")]
        yield nm, [l for l in lines if l and not l.startswith(";")]


def test_no_barrier_in_an_unskipped_exec_region(tmp_path):
    asm = _assembly(tmp_path)
    kernels = list(_kernels(asm))
    assert len(kernels) == 18      # N in {2, 1} x 3 modes x unroll {1, 2, 4}
    n_bar = 0
    for name, lines in kernels:
        bars = sum(1 for l in lines if l.startswith("s_barrier"))
        assert bars > 0, name
        n_bar += bars
        for k, l in enumerate(lines):
            # exec is narrowed only by saveexec and widened only by the restore
            if re.match(r"s_\w+_b64 exec,", l):
                assert re.match(r"s_or_b64 exec, exec, s\[\d+:\d+\]", l), (name, l)
            m = re.match(r"s_(and|andn2)_saveexec_b64 (s\[\d+:\d+\])", l)
            if not m:
                continue
            reg = m.group(2)
            n = k + 1
            while lines[n].startswith("s_xor_b64 " + reg):
                n += 1
            if lines[n].startswith("s_cbranch_exec"):
                continue
            for q in range(k + 1, len(lines)):
                assert not lines[q].startswith("s_barrier"), (name, k, q)
                if re.match(r"s_or_b64 exec, exec, " + re.escape(reg), lines[q]) or lines[q].startswith(".LBB"):
                    break
    assert n_bar > 100
