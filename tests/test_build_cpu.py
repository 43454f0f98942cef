"""The package Makefile lets a failing hipcc fail the build: a HIP source that does
not compile produces no object and a non-zero make status, while a good one
builds (hipcc cross-compiles for gfx950 without a GPU)."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
MAKEFILE = ROOT / "ice_crystal_phase_amd" / "Makefile"
HIPCC = Path("/opt/rocm/bin/hipcc")

pytestmark = pytest.mark.skipif(not HIPCC.exists(), reason="hipcc not installed")


def _tree(tmp_path, body):
    pkg = tmp_path / "pkg"
    (pkg / "csrc").mkdir(parents=True)
    (tmp_path / "include").mkdir()
    shutil.copy(MAKEFILE, pkg / "Makefile")
    shutil.copy(ROOT / "include" / "icephase.h", tmp_path / "include" / "icephase.h")
    (pkg / "csrc" / "probe.hip").write_text(body)
    return pkg


def _make_object(pkg):
    obj = pkg / "_build" / "probe.o"
    # a stale object from an earlier good build must not survive a failed compile
    obj.parent.mkdir(parents=True, exist_ok=True)
    obj.write_bytes(b"stale")
    r = subprocess.run(["make", "-s", "-C", str(pkg), "-B", str(obj)], capture_output=True, text=True)
    return r, obj


def test_broken_hip_source_fails_the_build(tmp_path):
    pkg = _tree(tmp_path, "this line is not C++;\n")
    r, obj = _make_object(pkg)
    assert r.returncode != 0, r.stdout + r.stderr
    assert not obj.exists()


def test_good_hip_source_builds(tmp_path):
    pkg = _tree(tmp_path, "#include <hip/hip_runtime.h>\n__global__ void k(int* p) { p[threadIdx.x] = 1; }\n")
    r, obj = _make_object(pkg)
    assert r.returncode == 0, r.stdout + r.stderr
    assert obj.exists() and obj.read_bytes() != b"stale"
