"""Spectral atlas runner: one deterministic atlas per wavelength, then the
orientation composition of all of them into CDF records.

Drop-in for the reference's hex_atlas/tools/run_spectrum.py (the same plan
file, flags, atlas cache with job receipts, reuse rules, refusals and exit
code 2). Each atlas is made by this package's bin/ice_precompute_atlas (the
area rays traced on the device) with the dispersive Warren-Brandt ice index at
its wavelength; the records are composed by compose.run (device sum and CDF,
the atlases resident in HBM). There is no CPU fallback: the runner stops at
the first failed job and keeps its .part directory.

Run as `python -m ice_crystal_phase_amd.spectrum --plan PLAN.json
--atlas-root DIR --out DIR [--exe PATH] [--reuse] [--dry-run]`.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import math
import subprocess
import sys
import time
from pathlib import Path
from typing import Any

from . import compose

PACKAGE = Path(__file__).resolve().parent
DEFAULT_EXE = PACKAGE / "bin" / "ice_precompute_atlas"

# precompute keys of the plan (run_spectrum.py:23-29): the required ones, the
# optional transport policies, and which of them are counts
REQUIRED = frozenset({"radius_mm", "length_mm", "incident_theta_intervals", "incident_phi_intervals", "theta",
                      "phi", "area_subdivisions", "batch_rays", "max_work_mib"})
OPTIONAL = frozenset({"max_internal_hits", "tail_tolerance", "max_unresolved_fraction", "max_balance_error"})
COUNTS = frozenset({"incident_theta_intervals", "incident_phi_intervals", "theta", "phi", "area_subdivisions",
                    "batch_rays", "max_work_mib", "max_internal_hits"})


def _number(v: Any) -> bool:
    return type(v) in (int, float) and math.isfinite(v)


def read_plan(path: Path) -> dict[str, Any]:
    """The plan, checked before any job starts (run_spectrum.py:32-70):
    precompute settings, increasing unique wavelengths in 200..1100 nm, the
    inclinations and orientation model of the composition, and the per-ray
    work bound the reference tool's buffers imply."""
    plan = json.loads(path.read_text(encoding="utf-8"))
    if not isinstance(plan, dict) or set(plan) != {"wavelengths_nm", "precompute", "composition"}:
        raise ValueError("Plan requires only wavelengths_nm, precompute, composition")
    s = plan["precompute"]
    if not isinstance(s, dict) or not REQUIRED <= set(s) <= REQUIRED | OPTIONAL:
        raise ValueError("Missing/unknown precompute parameter")
    for key, value in s.items():
        if key in COUNTS:
            if type(value) is not int or value <= 0:
                raise ValueError(f"{key} must be a positive integer")
        elif not _number(value) or value < 0:
            raise ValueError(f"{key} must be a finite nonnegative number")
    if not (s["radius_mm"] > 0 and s["length_mm"] > 0):
        raise ValueError("Particle dimensions must be positive")
    waves = plan["wavelengths_nm"]
    if (not isinstance(waves, list) or not waves or len(set(waves)) != len(waves)
            or not all(_number(w) and 200 <= w <= 1100 for w in waves)):
        raise ValueError("Unique vacuum wavelengths in 200..1100 nm are required")
    if any(b < a for a, b in zip(waves, waves[1:])):
        raise ValueError("Wavelengths must be listed in increasing order (stable w-index)")
    c = plan["composition"]
    if not isinstance(c, dict) or set(c) != {"inclinations_deg", "orientation"}:
        raise ValueError("This fixed-condition runner requires inclination angles and one orientation model")
    if len(set(c["inclinations_deg"])) != len(compose.incidents(c)):
        raise ValueError("Duplicate incident conditions")
    compose.orientation_rule(c["orientation"])
    if 16 * s["batch_rays"] * s["theta"] * s["phi"] > s["max_work_mib"] * 1024 ** 2:
        raise ValueError("Per-ray GPU work buffer exceeds max_work_mib; reduce batch_rays explicitly")
    return plan


def atlas_name(wavelength: float) -> str:
    """Cache directory of one wavelength: lambda_<shortest repr, '.' -> 'p'>_nm."""
    return "lambda_" + format(float(wavelength), ".17g").replace(".", "p") + "_nm"


def check_atlas(path: Path, wavelength: float, s: dict[str, Any]) -> dict[str, Any]:
    """A completed atlas of the requested dispersive-ice job (run_spectrum.py:
    77-103); the composer's reader audits every node afterwards."""
    if path.name.endswith(".part"):
        raise ValueError("Incomplete atlas")
    m = json.loads((path / "metadata.json").read_text(encoding="utf-8"))
    if m.get("complete") is not True or m.get("schema") != "ice.single_particle_atlas.v1":
        raise ValueError("Incomplete/unknown cached atlas")
    material = m["material"]
    if (material.get("model") != "ice_ih_warren_brandt_2008_real_only"
            or material.get("automatic_ice_dispersion") is not True
            or material.get("wavelength_nm") != wavelength or material.get("exterior_index") != 1):
        raise ValueError("Cached atlas is not the requested dispersive ice calculation")
    grid = {"incident_theta_intervals": s["incident_theta_intervals"],
            "incident_phi_intervals": s["incident_phi_intervals"], "theta_count": s["theta"], "phi_count": s["phi"]}
    if any(m.get(k) != v for k, v in grid.items()):
        raise ValueError("Cached angular grid differs from request")
    g = m["geometry"]
    if (g["circumradius_mm"] != s["radius_mm"] or g["full_length_mm"] != s["length_mm"]
            or m["area_quadrature"]["subdivisions"] != s["area_subdivisions"]):
        raise ValueError("Cached geometry/area quadrature differs from request")
    if not all((path / name).is_file() for name in ("mass_mm2.npy", "u_edges.npy")):
        raise ValueError("Missing cached atlas array")
    return m


def tool_arguments(s: dict[str, Any]) -> list[str]:
    """The precompute settings as the atlas tool's flags, in sorted key order."""
    out: list[str] = []
    for key in sorted(s):
        out += ["--" + key.replace("_", "-"), str(s[key])]
    return out


def run(exe: Path, plan_path: Path, atlas_root: Path, out: Path, reuse: bool = False, dry_run: bool = False,
        device: int = 0) -> None:
    plan = read_plan(plan_path)
    exe = exe.resolve()
    if not exe.is_file():
        raise ValueError("Precompute executable not found; build CPU or CUDA target explicitly")
    if out.exists() or out.with_name(out.name + ".part").exists():
        raise ValueError("Use a new CDF output directory; existing output/.part is never replaced")
    s, waves = plan["precompute"], plan["wavelengths_nm"]
    raw = 8 * (s["incident_theta_intervals"] + 1) * (s["incident_phi_intervals"] + 1) * s["theta"] * s["phi"]
    print(f"PLAN: {len(waves)} wavelengths x {len(plan['composition']['inclinations_deg'])} incidents")
    print(f"RAW ATLAS: {raw / 1024 ** 3:.3f} GiB/wavelength; {raw * len(waves) / 1024 ** 3:.3f} GiB total "
          "(plus headers/metadata/CDFs)")
    print(f"GPU RAY HISTOGRAM: {16 * s['batch_rays'] * s['theta'] * s['phi'] / 1024 ** 2:.1f} MiB "
          "(additional buffers required)")
    paths = [atlas_root / atlas_name(w) for w in waves]
    commands = [[str(exe), "--material", "ice", "--wavelength-nm", str(w), *tool_arguments(s), "--device",
                 str(device), "--out", str(p)] for w, p in zip(waves, paths)]
    if dry_run:
        for command in commands:
            print(subprocess.list2cmdline(command))
        print("DRY RUN: no files were created and no data were generated")
        return
    # a cached atlas is reused only for the same request made with the same
    # tool and the same device library
    build = {"executable_sha256": compose.sha256_file(exe)}
    library = PACKAGE / "libicephase.so"
    if library.is_file():
        build["device_library_sha256"] = compose.sha256_file(library)
    atlas_root.mkdir(parents=True, exist_ok=True)
    for index, (wave, path, command) in enumerate(zip(waves, paths, commands), start=1):
        receipt = path.with_name(path.name + ".job.json")
        request = {"schema": "ice.spectral_atlas_job.v1", "wavelength_nm": wave, "precompute": s, "material": "ice",
                   "build": build}
        if path.exists():
            if not reuse:
                raise FileExistsError(f"Atlas exists; use --reuse for validated matching jobs: {path}")
            old = json.loads(receipt.read_text(encoding="utf-8"))
            if old.get("request") != request or old.get("complete") is not True:
                raise ValueError("Cached atlas arguments/build differ; use another atlas-root")
            if old.get("metadata_sha256") != compose.sha256_file(path / "metadata.json"):
                raise ValueError("Cached metadata changed since the recorded job")
            check_atlas(path, wave, s)
            print(f"REUSE {index}/{len(paths)}: wavelength_nm={wave}")
            continue
        if receipt.exists() or path.with_name(path.name + ".part").exists():
            raise FileExistsError(f"Unfinished job or stale receipt; inspect it before retrying: {path}")
        log = path.with_name(path.name + ".log")
        start = time.perf_counter()
        with log.open("wb") as f:
            rc = subprocess.run(command, stdout=f, stderr=subprocess.STDOUT, check=False).returncode
        if rc != 0:
            raise RuntimeError(f"Precompute failed (exit={rc}); log={log}; .part retained, no fallback")
        check_atlas(path, wave, s)
        seconds = time.perf_counter() - start
        compose._write_json(receipt, {"complete": True, "request": request,
                                      "metadata_sha256": compose.sha256_file(path / "metadata.json"),
                                      "seconds": seconds})
        print(f"ATLAS {index}/{len(paths)}: wavelength_nm={wave}; seconds={seconds:.3f}", flush=True)
    # the composition config is stored once per content hash beside the atlases
    text = json.dumps(plan["composition"], sort_keys=True, ensure_ascii=True, allow_nan=False).encode("ascii")
    config = atlas_root / f"composition_{hashlib.sha256(text).hexdigest()}.json"
    if config.exists():
        if json.loads(config.read_text(encoding="utf-8")) != plan["composition"]:
            raise ValueError("Composition configuration cache mismatch")
    else:
        compose._write_json(config, plan["composition"])
    compose.run(paths, config, out, output_contract="rainbow", device=device)
    print(f"COMPLETE: {ascii(str(out))}; check_nf_compatibility.py is the downstream acceptance test")


def main(argv: list[str] | None = None) -> None:
    p = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    p.add_argument("--exe", type=Path, default=DEFAULT_EXE,
                   help="atlas tool (default: this package's bin/ice_precompute_atlas)")
    p.add_argument("--plan", type=Path, required=True)
    p.add_argument("--atlas-root", type=Path, required=True)
    p.add_argument("--out", type=Path, required=True)
    p.add_argument("--reuse", action="store_true")
    p.add_argument("--dry-run", action="store_true")
    p.add_argument("--device", type=int, default=0)
    a = p.parse_args(argv)
    try:
        run(a.exe, a.plan, a.atlas_root, a.out, a.reuse, a.dry_run, a.device)
    except Exception as e:
        print("SPECTRAL GENERATION FAILED: " + ascii(str(e)), file=sys.stderr)
        raise SystemExit(2)


if __name__ == "__main__":
    main()
