"""AddressSanitizer + UndefinedBehaviorSanitizer builds of the CPU code run
clean: the oracle's C restatement (every entry point the tests use, on small
cases) and the library's host C++ services (water, shapes, convexity,
projected areas, ice index, over their input ranges and refusals)."""
from __future__ import annotations

import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
SAN = ROOT / "tests" / "sanitize"


def _build():
    subprocess.run(["make", "-s", "-C", str(SAN)], check=True, timeout=600)


def test_oracle_under_sanitizers():
    _build()
    r = subprocess.run([str(SAN / "_build" / "oracle_driver")], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "oracle driver: ok" in r.stdout, r.stdout + r.stderr
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr


def test_host_services_under_sanitizers():
    _build()
    r = subprocess.run([str(SAN / "_build" / "host_driver")], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "host driver: ok" in r.stdout, r.stdout + r.stderr
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
