"""CPU tests of the command-line tools' argument handling: every rejection
below happens before a device is touched, with the reference's messages and
exit codes (apps/generate_phase.cpp:111-222, hex_phase/apps/generate_hex_phase.cpp:33-66)."""
from __future__ import annotations

import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
BIN = ROOT / "ice_crystal_phase_amd" / "bin"


def _run(*args, cwd=None):
    return subprocess.run([str(a) for a in args], capture_output=True, text=True, cwd=cwd, timeout=60)


@pytest.fixture(scope="module", autouse=True)
def built():
    if not (BIN / "rainbow_generate").exists() or not (BIN / "ice_generate_phase").exists():
        subprocess.run(["make", "-s", "-C", str(ROOT / "ice_crystal_phase_amd"), "-j8"], check=True)


@pytest.mark.parametrize("args,msg", [
    ([], "--out is required."),
    (["--out", "x"], "--wavelength-nm is required"),
    (["--out", "x", "--wavelength-nm", "550", "--wavelength-nm", "560"], "Duplicate --wavelength-nm."),
    (["--out", "x", "--wavelength-nm", "550", "--storage-gaussian-sigma-deg", "1"], "Global storage blur has been removed"),
    (["--out", "x", "--wavelength-nm", "550", "--diffraction-sigma-deg", "1"], "Manual diffraction width"),
    (["--out", "x", "--wavelength-nm", "550", "--ior", "1.33"], "Do not specify --ior for water"),
    (["--out", "x", "--wavelength-nm", "550", "--material", "constant"], "Constant material requires"),
    (["--out", "x", "--wavelength-nm", "550", "--material", "constant", "--ior", "1.3", "--temperature-c", "4"],
     "Temperature/pressure do not define"),
    (["--out", "x", "--wavelength-nm", "550", "--query-theta", "100", "--cdf-theta", "30"], "CDF dimensions must divide"),
    (["--out", "x", "--wavelength-nm", "550", "--stage", "focal", "--radius-mm", "2"], "Table II covers radius 0.1..1.0 mm"),
    (["--out", "x", "--wavelength-nm", "550", "--radius-mm", "1.5"], "Table II covers radius 0.1..1.0 mm"),
    (["--out", "x", "--wavelength-nm", "550", "--stage", "path", "--max-internal-hits", "3"],
     "apply to --stage incoherent only"),
    (["--out", "x", "--wavelength-nm", "550", "--stage", "bogus"], "Unknown stage."),
    (["--out", "x", "--wavelength-nm", "550", "--focal-offsets", "0,1,2"], "Expected four focal offsets."),
    (["--out", "x", "--wavelength-nm", "550", "--grid", "0"], "Invalid positive integer option."),
    (["--out", "x", "--wavelength-nm", "abc"], "Invalid numeric option."),
    (["--out", "x", "--wavelength-nm", "550", "--radius-mm", "3.5", "--stage", "path"], "Table I range"),
    (["--out", "x", "--wavelength-nm", "550", "--temperature-c", "80"], "liquid branch at 0..60 C"),
    (["--out", "x", "--wavelength-nm", "550", "--bogus", "1"], "Unknown option: --bogus"),
])
def test_rainbow_generate_rejects(args, msg, tmp_path):
    r = _run(BIN / "rainbow_generate", *args, cwd=tmp_path)
    assert r.returncode == 1
    assert r.stderr.startswith("rainbow_generate: ") and msg in r.stderr, r.stderr


def test_rainbow_generate_refuses_existing_part(tmp_path):
    (tmp_path / "x.part").mkdir()
    r = _run(BIN / "rainbow_generate", "--out", "x", "--wavelength-nm", "550", cwd=tmp_path)
    assert r.returncode == 1 and "Previous .part directory exists" in r.stderr


def test_rainbow_generate_default_stage_is_the_references(tmp_path):
    """Without --stage the reference runs its diffraction stage
    (apps/generate_phase.cpp:54), which needs a Table II radius; the tool takes
    the same default (the run itself is a GPU test), and argument errors are
    still reported first, in the reference's order."""
    r = _run(BIN / "rainbow_generate", "--out", "x", "--wavelength-nm", "550", "--radius-mm", "2", cwd=tmp_path)
    assert r.returncode == 1 and "Table II covers radius 0.1..1.0 mm only" in r.stderr, r.stderr
    assert not (tmp_path / "x").exists() and not (tmp_path / "x.part").exists()
    r = _run(BIN / "rainbow_generate", "--out", "x", "--wavelength-nm", "550", "--ior", "1.3", cwd=tmp_path)
    assert r.returncode == 1 and "Do not specify --ior for water" in r.stderr


def test_rainbow_generate_help():
    r = _run(BIN / "rainbow_generate", "--help")
    assert r.returncode == 0 and "--wavelength-nm" in r.stdout


HEX = ["--radius-mm", "0.05", "--length-mm", "0.2", "--ki", "1", "-0.4", "0.3", "--ior", "1.31",
       "--wavelength-nm", "550", "--theta", "90", "--phi", "180"]


@pytest.mark.parametrize("args,msg", [
    (["--out", "x"], "Required option: --orientations"),
    (["--out", "x", "--out", "y"], "Duplicate option: --out"),
    (["--out", "x", "--orientations", "p.csv"] + HEX + ["--cdf-theta", "7"], "coarsening dimensions"),
    (["--out", "x", "--orientations", "p.csv"] + HEX + ["--max-balance-error", "1.5"], "Invalid quality tolerance."),
    (["--out", "x", "--orientations", "p.csv"] + HEX + ["--what", "1"], "Unknown option: --what"),
])
def test_ice_generate_phase_rejects(args, msg, tmp_path):
    r = _run(BIN / "ice_generate_phase", *args, cwd=tmp_path)
    assert r.returncode == 1 and r.stderr.startswith("ice_generate_phase: ") and msg in r.stderr, r.stderr


@pytest.mark.parametrize("csv,msg", [
    ("qw,qx,qy,qz,weight\n1,0,0,0,1\n", "Expected CSV header"),
    ("qw,qx,qy,qz,weight,samples,seed\n1,0,0,0,1,10\n", "line 2: Expected seven CSV fields."),
    ("qw,qx,qy,qz,weight,samples,seed\n0.9,0,0,0,1,10,1\n", "line 2: Quaternion must be unit length"),
    ("qw,qx,qy,qz,weight,samples,seed\n1,0,0,0,0.5,10,1\n", "weights must sum to one"),
    ("qw,qx,qy,qz,weight,samples,seed\n1,0,0,0,1,0,1\n", "Weights and sample counts must be positive."),
    ("# comment\nqw,qx,qy,qz,weight,samples,seed\n1,0,0,0,1,x,1\n", "line 3: Malformed orientation number."),
])
def test_ice_generate_phase_plan_errors(csv, msg, tmp_path):
    (tmp_path / "p.csv").write_text(csv)
    r = _run(BIN / "ice_generate_phase", "--out", "x", "--orientations", "p.csv", *HEX, cwd=tmp_path)
    assert r.returncode == 1 and msg in r.stderr, r.stderr
    assert not (tmp_path / "x.part").exists()
