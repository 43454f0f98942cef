"""Resumable condition-grid generation of water-drop records, traced in-process.

Drop-in for the reference's batch driver (tools/generate_phase_dataset.py:1-431):
the same JSON config (keys, ranges, exact decimal axes, error messages), the
same output tree

    output_root/batch_plan.json     conditions + SHA-256 of the engine files
    output_root/progress.json
    output_root/records/iNNNN/wNNNN incident index outer, wavelength inner
    output_root/receipts/iNNNN_wNNNN.json   metadata hash and NPY sizes
    output_root/logs/iNNNN_wNNNN.log

and the same rules: dry run unless --execute, --resume only against an
identical plan, receipts checked against the record on resume, stop at the
first failure without skipping or deleting anything, a `.batch.lock` that is
never removed for another writer, a free-space reserve.

What differs is the engine. The reference starts one `rainbow_generate`
process per condition; here the conditions are traced in batches of many
records by one persistent launch of libicephase.so (`iph_batch_run`: every
record's rays in one tile queue, then every record's CDF), in this process,
and each record is written with exactly the metadata and arrays
`bin/rainbow_generate` writes for the same condition. The incoherent stage is
the ray histogram, traced in batches; the coherent stages (path, focal,
diffraction) run the patch-model pipeline (iph_coherent_*) one record at a
time on the same context.

Run as `python -m ice_crystal_phase_amd.dataset --config CONFIG [--execute]`.
"""
from __future__ import annotations

import argparse
import ctypes as C
import hashlib
import json
import math
import os
import shutil
import struct
import sys
import tempfile
import time
from concurrent.futures import ThreadPoolExecutor
from contextlib import contextmanager
from fractions import Fraction
from dataclasses import dataclass
from decimal import Decimal, InvalidOperation
from pathlib import Path
from typing import Any, Iterator

import numpy as np

from . import icephase as ip

PACKAGE = Path(__file__).resolve().parent
VERSION = "rainbow.condition_batch.v1"
GENERATOR_VERSION = "icephase.mi355x.drop_histogram.v1"
COHERENT_GENERATOR_VERSION = "icephase.mi355x.drop_patches.v1"
TABLE_II_MESSAGE = ("Table II covers radius 0.1..1.0 mm only. No bandwidth extrapolation is performed. The existing "
                    "focal path also computes diffraction; use incoherent/path for out-of-table reference runs.")


def generator_version(stage: str) -> str:
    return GENERATOR_VERSION if stage == "incoherent" else COHERENT_GENERATOR_VERSION


def table_sigma_degrees(radius: float) -> float | None:
    """Table II primary bandwidth (degrees) of an FP32 radius, linearly
    interpolated between the tabulated radii; None outside 0.1..1 mm
    (include/rainbow/rainbow_diffraction.hpp:17-29)."""
    r = [fp32(x) for x in (0.1, 0.2, 0.3, 0.4, 0.5, 0.6, 0.7, 0.8, 0.9, 1.0)]
    sig = [0.70, 0.45, 0.30, 0.25, 0.22, 0.20, 0.18, 0.17, 0.16, 0.15]
    if not (r[0] <= radius <= r[9]):
        return None
    for j in range(10):
        if radius == r[j]:
            return sig[j]
        if j and radius < r[j]:
            t = (radius - r[j - 1]) / (r[j] - r[j - 1])
            # one rounding of t * (s1 - s0) + s0, as the tool's fma
            return float(Fraction(t) * Fraction(sig[j] - sig[j - 1]) + Fraction(sig[j - 1]))
    return None
ARRAYS = ("phi_cdf.npy", "theta_given_phi_cdf.npy", "u_edges.npy")
CONFIG_KEYS = {"output_root", "generator", "radius_mm", "force_sphere", "temperature_c", "pressure_pa",
               "wavelengths", "inclinations", "solver", "reserve_gib"}
SOLVER_KEYS = {"grid", "query_theta", "query_phi", "cdf_theta", "cdf_phi", "stage", "device", "focal_offsets",
               "allow_underresolved", "maximum_coarsening_tv", "cdf_max_l1", "cdf_max_lost_mass"}
STAGES = ("diffraction", "focal", "path", "incoherent")
# rainbow_generate's defaults for what the config does not set
MAX_INTERNAL_HITS = 8
TAIL_TOLERANCE = 0.0
# device memory one batch may take for its histograms and CDF buffers
BATCH_BYTES = 8 << 30
MAX_BATCH_RECORDS = 256


# ------------------------------------------------------------------ config values
# Numbers in a config are read as the exact decimal they spell (JSON numbers and
# numeric strings alike), so axis endpoints and steps are exact and a plan
# written from them is reproducible. The limits below are the dataset
# contract's (tools/generate_phase_dataset.py).
_SOLVER_INTEGERS = {"grid": (2, 32767), "query_theta": (1, 32767), "query_phi": (1, 32767),
                    "cdf_theta": (1, 32767), "cdf_phi": (1, 32767), "device": (0, 128)}
_SOLVER_FRACTIONS = ("maximum_coarsening_tv", "cdf_max_l1", "cdf_max_lost_mass")   # each in [0, 1]
_CONDITION_SCALARS = {"radius_mm": (1e-9, 3), "temperature_c": (0, 60), "pressure_pa": (50000, 100000000)}
WAVELENGTH_KEYS = frozenset({"start_nm", "stop_nm", "step_nm"})
INCLINATION_KEYS = frozenset({"min_degrees", "max_degrees", "count", "placement"})


def load_json(path: Path, limit: int = 8 << 20) -> dict[str, Any]:
    """A JSON object from a file of at most `limit` bytes (a UTF-8 BOM is
    allowed); NaN and Infinity are refused."""
    if path.stat().st_size > limit:
        raise ValueError(f"JSON too large: {path}")

    def refuse(token: str) -> Any:
        raise ValueError(f"Nonfinite JSON constant: {token}")

    value = json.loads(path.read_text(encoding="utf-8-sig"), parse_constant=refuse)
    if not isinstance(value, dict):
        raise ValueError(f"Expected JSON object: {path}")
    return value


def exact(value: Any) -> Decimal:
    """The exact decimal a JSON number or numeric string denotes."""
    if isinstance(value, bool):
        raise ValueError("Boolean is not a numerical parameter")
    try:
        d = Decimal(str(value))
    except InvalidOperation as exc:
        raise ValueError(f"Invalid decimal: {value}") from exc
    if not d.is_finite():
        raise ValueError("Nonfinite parameter")
    return d


def scalar(value: Any, low: float, high: float) -> float:
    d = exact(value)
    if d < Decimal(str(low)) or d > Decimal(str(high)):
        raise ValueError(f"Parameter {value} is outside [{low}, {high}]")
    return float(d)


def integer(value: Any, low: int, high: int) -> int:
    d = exact(value)
    whole = d == d.to_integral_value()
    if not whole or d < low or d > high:
        raise ValueError(f"Expected integer in [{low}, {high}]")
    return int(d)


def flag(value: Any) -> bool:
    if value is True or value is False:
        return value
    raise ValueError("Expected JSON true or false")


def fp32(x: float) -> float:
    """x rounded to the nearest single-precision value (the solver's precision)."""
    return float(np.float32(x))


def _text(d: Decimal) -> str:
    return format(d, "f")


def wavelengths(spec: dict[str, Any]) -> list[str]:
    """Inclusive exact-step axis start, start + step, ..., stop, as decimal strings."""
    if set(spec) != WAVELENGTH_KEYS:
        raise ValueError("wavelengths needs start_nm, stop_nm, step_nm")
    start, stop, step = exact(spec["start_nm"]), exact(spec["stop_nm"]), exact(spec["step_nm"])
    if step <= 0 or not (380 <= start <= stop <= 830):
        raise ValueError("Require 380 <= start <= stop <= 830 and positive step")
    n_steps, remainder = divmod(stop - start, step)
    if remainder != 0:
        raise ValueError("The wavelength range must contain an integer number of steps; "
                         "no silent endpoint truncation")
    if n_steps + 1 > 100000:
        raise ValueError("Too many wavelengths")
    axis = [_text(start + step * k) for k in range(int(n_steps) + 1)]
    solver_values = np.array([float(v) for v in axis], dtype=np.float32)
    if np.unique(solver_values).size != len(axis):
        raise ValueError("Distinct wavelengths collapse to the same FP32 solver value")
    return axis


def inclinations(spec: dict[str, Any]) -> list[str]:
    """Cell centres of `count` equal angle cells over [min, max], as decimal strings:
    min + (2k + 1) (max - min) / (2 count)."""
    if set(spec) != INCLINATION_KEYS:
        raise ValueError("inclinations needs min_degrees, max_degrees, count, placement")
    if spec["placement"] != "linear_angle_cell_centers":
        raise ValueError("Only explicit linear_angle_cell_centers is supported")
    low, high = exact(spec["min_degrees"]), exact(spec["max_degrees"])
    count = integer(spec["count"], 1, 100000)
    if not (-90 <= low < high <= 90):
        raise ValueError("Canonical inclination domain is [-90,90], NOT [0,180]")
    half_cell = (high - low) / (2 * count)
    return [_text(low + half_cell * (2 * k + 1)) for k in range(count)]


def _solver(s: Any) -> dict[str, Any]:
    if not isinstance(s, dict) or set(s) != SOLVER_KEYS:
        raise ValueError("Invalid solver keys")
    out: dict[str, Any] = {k: integer(s[k], lo, hi) for k, (lo, hi) in _SOLVER_INTEGERS.items() if k != "device"}
    if 4 * out["grid"] * out["grid"] > 0xFFFFFFFF:
        raise ValueError("Source index range exceeds the existing solver contract")
    if any(out["cdf_" + a] > out["query_" + a] or out["query_" + a] % out["cdf_" + a] for a in ("theta", "phi")):
        raise ValueError("CDF grids must divide query grids")
    if s["stage"] not in STAGES:
        raise ValueError("Unknown stage")
    out["stage"] = s["stage"]
    out["device"] = integer(s["device"], *_SOLVER_INTEGERS["device"])
    out["allow_underresolved"] = flag(s["allow_underresolved"])
    for k in _SOLVER_FRACTIONS:
        out[k] = scalar(s[k], 0, 1)
    offsets = s["focal_offsets"]
    if not (isinstance(offsets, list) and len(offsets) == 4):
        raise ValueError("Four focal offsets required")
    out["focal_offsets"] = [integer(v, 0, 3) for v in offsets]
    return out


def normalized_config(raw: dict[str, Any]) -> dict[str, Any]:
    """Validated conditions of a dataset config (the plan's `conditions`)."""
    if set(raw) != CONFIG_KEYS:
        raise ValueError(f"Unexpected/missing config keys: {set(raw) ^ CONFIG_KEYS}")
    solver = _solver(raw["solver"])
    values = {k: scalar(raw[k], lo, hi) for k, (lo, hi) in _CONDITION_SCALARS.items()}
    if solver["stage"] in ("focal", "diffraction") and table_sigma_degrees(fp32(values["radius_mm"])) is None:
        raise ValueError(TABLE_II_MESSAGE)
    return {"radius_mm": values["radius_mm"], "force_sphere": flag(raw["force_sphere"]),
            "temperature_c": values["temperature_c"], "pressure_pa": values["pressure_pa"],
            "wavelength_nm": wavelengths(raw["wavelengths"]), "inclination_degrees": inclinations(raw["inclinations"]),
            "inclination_placement": "linear_angle_cell_centers", "material": "water", "solver": solver}


# ------------------------------------------------------------------ files
def sha256(path: Path) -> str:
    """SHA-256 of a file's bytes, read in 1 MiB pieces."""
    digest = hashlib.sha256()
    with path.open("rb") as f:
        while piece := f.read(1 << 20):
            digest.update(piece)
    return digest.hexdigest()


def object_hash(value: Any) -> str:
    """SHA-256 of the canonical JSON text of `value` (sorted keys, no spaces)."""
    canonical = json.dumps(value, sort_keys=True, separators=(",", ":"), allow_nan=False)
    return hashlib.sha256(canonical.encode("utf-8")).hexdigest()


def write_json(path: Path, value: Any) -> None:
    """Write through a temporary file in the same directory and rename it over
    the target, so a reader sees the old file or the new one, never part of one."""
    path.parent.mkdir(parents=True, exist_ok=True)
    text = json.dumps(value, indent=2, ensure_ascii=False, allow_nan=False) + "\n"
    fd, tmp = tempfile.mkstemp(prefix=path.name + ".", suffix=".tmp", dir=path.parent)
    try:
        with os.fdopen(fd, "w", encoding="utf-8", newline="\n") as f:
            f.write(text)
        os.replace(tmp, path)
    except BaseException:
        Path(tmp).unlink(missing_ok=True)
        raise


@contextmanager
def batch_lock(root: Path) -> Iterator[None]:
    """Exclusive writer lock `root/.batch.lock`, created atomically and removed
    on exit; an existing lock (another writer, or a crashed one) is never taken
    over."""
    lock = root / ".batch.lock"
    try:
        fd = os.open(lock, os.O_CREAT | os.O_EXCL | os.O_WRONLY, 0o644)
    except FileExistsError as exc:
        raise RuntimeError("Batch lock exists. Do not run two writers. After verifying that no batch is running, "
                           "remove the stale .batch.lock manually.") from exc
    with os.fdopen(fd, "w", encoding="ascii") as f:
        f.write(f"pid={os.getpid()}\n")
    try:
        yield
    finally:
        lock.unlink()


def npy_size(path: Path, shape: tuple[int, ...]) -> int:
    """Byte size of a '<f8' C-order NPY file of the given shape, checked from its
    header and the file length only (the data is not read)."""
    from numpy.lib import format as npy
    with path.open("rb") as f:
        try:
            version = npy.read_magic(f)
        except ValueError as exc:
            raise ValueError(f"Not NPY: {path}") from exc
        if version not in ((1, 0), (2, 0)):
            raise ValueError("Unsupported NPY version")
        reader = npy.read_array_header_1_0 if version == (1, 0) else npy.read_array_header_2_0
        try:
            stored_shape, fortran, dtype = reader(f, max_header_size=65536)
        except ValueError as exc:
            raise ValueError(f"Wrong NPY layout: {path}") from exc
        if dtype != np.dtype("<f8") or fortran or tuple(stored_shape) != tuple(shape):
            raise ValueError(f"Wrong NPY layout: {path}")
        expected = f.tell() + 8 * math.prod(shape)
    if path.stat().st_size != expected:
        raise ValueError(f"Truncated or trailing NPY data: {path}")
    return expected


def engine_files() -> dict[str, Path]:
    """Files whose content defines the records: the device library and the
    host code that drives it. Their hashes go into the plan."""
    return {"libicephase.so": ip.LIB_PATH, "dataset.py": PACKAGE / "dataset.py",
            "icephase.py": PACKAGE / "icephase.py"}


# ------------------------------------------------------------------ record checks
def _number(x: Any) -> bool:
    return type(x) in (int, float) and math.isfinite(x)


def _near(x: Any, want: float, tol: float = 1e-12) -> bool:
    return _number(x) and abs(x - want) <= tol * max(1.0, abs(want))


def _record_checks(m: dict[str, Any], c: dict[str, Any], ii: int, ll: int) -> list[tuple[str, bool]]:
    """(label, passed) for every rule a record of condition (ii, ll) obeys."""
    s = c["solver"]
    get = lambda dotted: _dig(m, dotted)   # noqa: E731
    expected = {"complete": True, "schema": "rainbow.phase_cdf.numpy.v2",
                "generator_version": generator_version(s["stage"]),
                "coordinate_contract_id": "rainbow.phase_cdf.coordinates.v1",
                "direction_convention": "physical_propagation", "dtype": "<f8", "order": "C",
                "density_measure": "solid_angle_sr", "cdf_axis_order": ["phi_cell", "theta_edge"],
                "theta_count": s["cdf_theta"], "phi_count": s["cdf_phi"],
                "query_grid": [s["query_theta"], s["query_phi"]], "incident_grid": [s["grid"], s["grid"]],
                "stage": s["stage"], "force_sphere": c["force_sphere"],
                "focal_quarter_turn_offsets": s["focal_offsets"]}
    checks = [(f"mismatched {k}", m.get(k) == v) for k, v in expected.items()]
    alpha = float(c["inclination_degrees"][ii])
    checks += [
        ("radius", _near(m.get("radius_mm"), fp32(c["radius_mm"]))),
        ("wavelength", _near(m.get("wavelength_nm"), fp32(float(c["wavelength_nm"][ll])))),
        ("inclination", _near(m.get("incident_inclination_degrees"), alpha)),
        ("material", get("material.model") == "water_iapws_r9_97"),
        ("temperature", _near(get("material.temperature_kelvin"), c["temperature_c"] + 273.15)),
        ("pressure", _near(get("material.pressure_pascal"), c["pressure_pa"])),
    ]
    frame = m.get("sampling_frame_columns")
    shaped = isinstance(frame, list) and len(frame) == 3 and all(isinstance(r, list) and len(r) == 3 for r in frame)
    checks.append(("frame shape", shaped))
    if shaped:
        finite = all(_number(v) for row in frame for v in row)
        checks.append(("frame values", finite))
        if finite:
            F = np.array(frame, dtype=np.float64)          # columns e0, e1, ki
            checks.append(("frame orthogonality", bool(np.all(np.abs(F.T @ F - np.eye(3)) <= 2e-12))))
            checks.append(("frame handedness", float(np.linalg.det(F)) > 1 - 2e-12))
            a = math.radians(alpha)
            ki = np.array([math.cos(a), -math.sin(a), 0.0])
            checks.append(("incident direction/sign", bool(np.all(np.abs(F[:, 2] - ki) < 2e-6))))
    area = get("geometry.projected_area_mm2")
    checks += [
        ("projected area", _number(area) and area > 0),
        ("area validation", get("geometry.projected_area_converged") is True
         and get("geometry.convexity_verified") is True),
        ("cross-section model", get("cross_sections.model") == "geometric_nonabsorbing"
         and get("cross_sections.wave_scattering_cross_section_computed") is False),
        ("cross sections", get("cross_sections.scattering_mm2") == area and get("cross_sections.extinction_mm2") == area
         and get("cross_sections.absorption_mm2") == 0),
        ("invalid optical data", get("quality.invalid_values") == 0 and get("quality.incomplete_directions") == 0),
        ("underresolution policy", get("quality.allow_underresolved") is s["allow_underresolved"]),
        ("underresolution not authorized", s["allow_underresolved"] or get("quality.underresolved_directions") == 0),
    ]
    for key, limit in (("cdf_l1_mass_error", "cdf_max_l1"), ("cdf_lost_probability_mass", "cdf_max_lost_mass")):
        v = get("quality." + key)
        checks.append((key, _number(v) and 0 <= v <= s[limit]))
    tv = get("storage_processing.coarsening_tv")
    g = get("hg.g")
    checks += [
        ("unexpected global smoothing", get("storage_processing.additional_filter") == "none"),
        ("coarsening error", _number(tv) and 0 <= tv <= s["maximum_coarsening_tv"]),
        ("HG target", get("hg.target") == "saved_cdf" and _number(g) and -1 < g < 1),
    ]
    return checks


def _dig(tree: Any, dotted: str) -> Any:
    for part in dotted.split("."):
        if not isinstance(tree, dict):
            return None
        tree = tree.get(part)
    return tree


def verify_record(path: Path, plan: dict[str, Any], ii: int, ll: int) -> dict[str, Any]:
    """Check a written record against its planned condition (metadata contract,
    frame, policy limits, NPY headers and sizes) and return its receipt body."""
    c = plan["conditions"]
    m = load_json(path / "metadata.json", 1 << 20)
    for label, ok in _record_checks(m, c, ii, ll):
        if not ok:
            raise ValueError(f"{path}: {label}")
    nt, nphi = c["solver"]["cdf_theta"], c["solver"]["cdf_phi"]
    shapes = dict(zip(ARRAYS, ((nphi + 1,), (nphi, nt + 1), (nt + 1,))))
    return {"metadata_sha256": sha256(path / "metadata.json"),
            "npy_sizes": {name: npy_size(path / name, shape) for name, shape in shapes.items()},
            "hg_g": m["hg"]["g"], "projected_area_mm2": m["geometry"]["projected_area_mm2"],
            "underresolved_directions": _dig(m, "quality.underresolved_directions"),
            "coarsening_tv": m["storage_processing"]["coarsening_tv"]}


# ------------------------------------------------------------------ one condition
@dataclass
class Condition:
    """Everything rainbow_generate resolves on the host for one condition."""
    ii: int
    ll: int
    job: str
    path: Path
    receipt: Path
    alpha: float                # incident inclination, degrees
    wavelength_requested: float
    wavelength: float           # FP32 solver wavelength
    radius: float               # FP32 solver radius
    water: Any                  # ip.WaterState at the FP32 wavelength
    n_solver: float             # FP32 interior index
    shape: list[float]          # FP32 Table I coefficients (zeros for a sphere)
    sphere: bool                # traced as the closed-form sphere
    kf: list[float]             # FP32 incident direction
    area: Any                   # ip.Area (certified projected area)


def resolve(c: dict[str, Any], ii: int, ll: int, root: Path) -> Condition:
    """Host-side set-up of one condition, in rainbow_generate's order and
    precision (apps/rainbow_generate.cpp): FP32 radius and wavelength, IAPWS
    water at the FP32 wavelength, Table I shape of the FP32 radius, FP32
    incident direction, certified projected area in FP64."""
    job = f"i{ii:04d}_w{ll:04d}"
    radius = fp32(c["radius_mm"])
    requested = float(c["wavelength_nm"][ll])
    wavelength = fp32(requested)
    water = ip.water_optics(wavelength, c["temperature_c"], c["pressure_pa"])
    shape = ip.drop_shape(radius, c["force_sphere"])
    sphere = c["force_sphere"] or all(v == 0.0 for v in shape)
    alpha = float(c["inclination_degrees"][ii])
    angle = alpha * math.pi / 180.0
    kf = [fp32(math.cos(angle)), fp32(-math.sin(angle)), 0.0]
    area = ip.drop_projected_area(shape, radius, kf)
    return Condition(ii, ll, job, root / f"records/i{ii:04d}/w{ll:04d}", root / "receipts" / (job + ".json"), alpha,
                     requested, wavelength, radius, water, fp32(water.refractive_index), shape, sphere, kf, area)


def device_record(cond: Condition, s: dict[str, Any]) -> ip.Record:
    """The iph_record rainbow_generate traces for this condition."""
    r = ip.Record()
    r.shape = ip.SHAPE_SPHERE if cond.sphere else ip.SHAPE_DROP
    r.max_internal_hits = MAX_INTERNAL_HITS
    r.theta_count, r.phi_count = s["query_theta"], s["query_phi"]
    r.grid_w = r.grid_h = s["grid"]
    r.radius_mm = cond.radius
    r.shape_coefficients = (C.c_double * 8)(*cond.shape)
    r.incident = (C.c_double * 3)(*cond.kf)
    r.interior_index = cond.n_solver
    r.exterior_index = 1.0
    r.wavelength_nm = cond.wavelength
    r.tail_tolerance = TAIL_TOLERANCE
    return r


def policy(s: dict[str, Any]) -> ip.CdfPolicy:
    return ip.CdfPolicy(s["cdf_theta"], s["cdf_phi"], s["cdf_max_l1"], s["cdf_max_lost_mass"],
                        s["maximum_coarsening_tv"])


@dataclass
class Traced:
    """Device results of one condition: histogram (or patch-model) statistics,
    the CDF arrays, the CDF report (status 0 = every audit passed) and, for a
    coherent stage, the density gate and the refusal message if any."""
    stats: Any
    phi: np.ndarray
    theta: np.ndarray
    report: Any
    gate: Any = None
    failure: str | None = None


def _frame(kf: list[float]) -> tuple[list[float], list[float], list[float]]:
    """Binning frame: e0 = normalize(+z - k (k.z)) (+x when k is near the z
    axis), e1 = k x e0, k normalized; FP64 from the FP32 direction."""
    n = math.sqrt(kf[0] * kf[0] + kf[1] * kf[1] + kf[2] * kf[2])
    k = [v / n for v in kf]
    e = [1.0, 0.0, 0.0] if abs(k[2]) >= 0.9 else [0.0, 0.0, 1.0]
    d = e[0] * k[0] + e[1] * k[1] + e[2] * k[2]
    e = [e[i] - d * k[i] for i in range(3)]
    n = math.sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2])
    e = [v / n for v in e]
    f = [k[1] * e[2] - k[2] * e[1], k[2] * e[0] - k[0] * e[2], k[0] * e[1] - k[1] * e[0]]
    return e, f, k


def outer_radius(cond: Condition) -> np.float32:
    """The solver's conservative outer radius in drop units, FP32 rounded
    outward step by step (include/rainbow/raindrop_shape.hpp:45-64)."""
    f32 = np.float32
    ripple = f32(0.0)
    for v in cond.shape[1:]:
        ripple = np.nextafter(f32(ripple + f32(abs(v))), f32(np.inf))
    return np.nextafter(f32(f32(f32(1.0) + f32(cond.shape[0])) + ripple) + f32(2.0 ** -20), f32(np.inf))


def half_extent(cond: Condition) -> float:
    """Half side of the incident grid in drop units: 1.01 x the outer radius (FP32)."""
    return float(np.float32(outer_radius(cond) * np.float32(1.01)))


def reference_distance(cond: Condition) -> float:
    """Distance of both reference planes in drop units: twice the outer radius (FP32)."""
    return float(np.float32(np.float32(2.0) * outer_radius(cond)))


def metadata(cond: Condition, c: dict[str, Any], t: Traced) -> dict[str, Any]:
    """metadata.json of one record: the keys, order and values
    bin/rainbow_generate writes for the same condition."""
    s = c["solver"]
    st, rep, area, w = t.stats, t.report, cond.area, cond.water
    e, f, k = _frame(cond.kf)
    kd = list(cond.kf)
    coherent = s["stage"] != "incoherent"
    sigma = table_sigma_degrees(cond.radius) if s["stage"] in ("focal", "diffraction") else None
    if coherent:
        g = t.gate
        model = {"patch_model": {"method": "bilinear_chord_patches_device_query", "patches": int(st.patches),
                                 "folded_patches": int(st.folded), "candidate_pairs": int(st.candidates),
                                 "hits": int(st.hits)}}
        quality = {"missing_source_cells": int(st.status_counts[4]), "invalid_values": int(g.invalid_values),
                   "incomplete_directions": int(g.incomplete), "underresolved_directions": int(g.underresolved)}
    else:
        q = math.ldexp(1.0, -int(st.qbits)) * st.ray_weight_mm2  # mm^2 per quantum
        model = {"histogram": {
            "method": "fixed_point_u64_ray_histogram",
            "quantum_bits": int(st.qbits),
            "ray_weight_mm2": st.ray_weight_mm2,
            "rays": int(st.rays),
            "rays_missed": int(st.status[8]),
            "outgoing_samples": int(st.samples),
            "max_internal_hits": MAX_INTERNAL_HITS,
            "tail_tolerance_per_ray": TAIL_TOLERANCE,
            "escaped_mm2": float(st.q_escaped) * q,
            "unresolved_mm2": float(st.q_unresolved) * q,
            "forward_axis_mass_mm2": float(st.q_forward) * q,
            "backward_axis_mass_mm2": float(st.q_backward) * q,
            "boundary_snapped_samples": int(st.snapped),
        }}
        quality = {"missing_source_cells": 0, "invalid_values": int(st.invalid), "incomplete_directions": 0,
                   "underresolved_directions": 0}
    return {
        "coordinate_contract_id": "rainbow.phase_cdf.coordinates.v1",
        "direction_convention": "physical_propagation",
        "particle_frame_handedness": "right",
        "particle_up_axis": [0, 1, 0],
        "shape_polar_axis": [0, -1, 0],
        "sampling_frame_layout": "rows_xyz_columns_e0_e1_k",
        "sampling_frame_map": "local_column_to_particle_column",
        "theta_zero": "forward",
        "theta_pi": "backward",
        "azimuth_periodic": True,
        "inclination_definition": "ki=(cos(alpha),-sin(alpha),0); alpha=0 is +x; positive alpha points downward",
        "geometry": {
            "projected_area_mm2": area.area_mm2,
            "projected_area_m2": area.area_mm2 * 1e-6,
            "projected_area_method": ("analytic_sphere" if area.analytic_sphere
                                      else "convex_surface_integral_azimuth_analytic_gk15"),
            "projected_area_shape": "solver_shape_coefficients_promoted_to_fp64",
            "projected_area_converged": True,
            "projected_area_absolute_error_estimate_mm2": area.abs_error_mm2,
            "projected_area_relative_error_estimate": area.rel_error,
            "projected_area_error_is_rigorous_bound": False,
            "convexity_verified": True,
            "convexity_method": ("analytic_sphere" if area.analytic_sphere
                                 else "markov_lipschitz_branch_and_bound_curvature"),
            "convexity_intervals": int(area.convexity_cells),
            "quadrature_evaluations": int(area.evaluations),
            "incident_axis_particle_frame": k,
        },
        "cross_sections": {
            "model": "geometric_nonabsorbing",
            "scattering_mm2": area.area_mm2,
            "extinction_mm2": area.area_mm2,
            "absorption_mm2": 0,
            "scattering_m2": area.area_mm2 * 1e-6,
            "extinction_m2": area.area_mm2 * 1e-6,
            "absorption_m2": 0,
            "assignment": "Csca=Cext=Aproj; Cabs=0 (transport model assumption)",
            "wave_scattering_cross_section_computed": False,
        },
        "material": {
            "model": "water_iapws_r9_97",
            "wavelength_convention": "vacuum_nm",
            "wavelength_requested_nm": cond.wavelength_requested,
            "wavelength_evaluated_nm": cond.wavelength,
            "ambient_model": "unit_absolute_index_air_approximation",
            "absorption_model": "ignored_k_equals_zero",
            "interior_index_evaluated_fp64": w.refractive_index,
            "interior_index_solver_fp32": cond.n_solver,
            "exterior_index_solver": 1.0,
            "refractive_index_release": "IAPWS R9-97",
            "density_model": "IAPWS R6-95(2018) liquid pressure EOS inversion",
            "temperature_kelvin": w.temperature_kelvin,
            "pressure_pascal": w.pressure_pascal,
            "density_kg_m3": w.density_kg_m3,
            "density_pressure_residual_pascal": w.density_pressure_residual_pascal,
            "density_bracket_width_kg_m3": w.density_bracket_width_kg_m3,
        },
        "diffraction_policy": {
            "name": "Sadeghi2012_TableII_local_v1",
            "computed": sigma is not None,
            "selected_output_includes_diffraction": s["stage"] == "diffraction",
            "primary_sigma_degrees": sigma,
            "secondary_sigma_degrees": None if sigma is None else 2 * sigma,
            "bandwidth_source": "paper Table II; secondary sigma x2",
            "between_table_radii": "linear interpolation (project convention)",
            "outside_table_radii": "reject",
            "wavelength_scaling": "none; no rule specified in Table II",
            "support_sigma": 4,
            "support_and_transition_detection_are_project_conventions": True,
            "physical_sigma_depends_on_grid": False,
            "query_dtheta_degrees": 180.0 / s["query_theta"],
            "query_dphi_degrees": 360.0 / s["query_phi"],
        },
        "schema": "rainbow.phase_cdf.numpy.v2",
        "complete": True,
        "dtype": "<f8",
        "order": "C",
        "theta_count": s["cdf_theta"],
        "phi_count": s["cdf_phi"],
        "density_measure": "solid_angle_sr",
        "cell_model": "constant_density_per_spherical_cell",
        "coordinates": "u=(1-cos(theta))/2; v=(phi+pi)/(2*pi)",
        "theta_definition": "angle_between_incident_and_outgoing_propagation_directions",
        "cdf_axis_order": ["phi_cell", "theta_edge"],
        "phi_range_rad": [-math.pi, math.pi],
        "normalization": "integral_p_domega_equals_one",
        "input_polarization": "unpolarized",
        "incident_total_intensity": 1,
        "single_scattering_albedo": 1,
        "collision_model": "geometric_projected_area",
        "projected_area_storage": "metadata.geometry.projected_area_mm2",
        "stage": s["stage"],
        "radius_mm": cond.radius,
        "wavelength_nm": cond.wavelength,
        "exterior_index": 1.0,
        "interior_index": cond.n_solver,
        "force_sphere": c["force_sphere"],
        "incident_grid": [s["grid"], s["grid"]],
        "incident_grid_half_extent": half_extent(cond),
        "incident_grid_placement": ("vertices_over_square_1.01_outer_radius" if coherent
                                    else "cell_centres_over_square_1.01_outer_radius_misses_discarded"),
        "reference_distances": [reference_distance(cond), reference_distance(cond)],
        "query_frame_axis": kd,
        "query_frame_e0": e,
        "sampling_frame_columns": [[e[r], f[r], k[r]] for r in range(3)],
        "shape_coefficients": list(cond.shape),
        "focal_quarter_turn_offsets": list(s["focal_offsets"]),
        "diffraction_primary_sigma_degrees_setting": 0.0,
        "diffraction_zero_setting_selects_Table_II": True,
        "diffraction_transition_contrast": 0.25,
        "folded_branch_model": "finite_branch_area_ratio_v1" if coherent else "none_ray_histogram",
        "folded_relative_tolerance": 1e-8 if coherent else 0.0,
        "source_commit": "unknown",
        "source_dirty_at_configure": True,
        "generator_version": generator_version(s["stage"]),
        "quality": {
            "source_coverage_certified": False,
            "angular_convergence_certified": False,
            **quality,
            "allow_underresolved": s["allow_underresolved"],
            "cdf_l1_mass_error": rep.l1_error,
            "cdf_maximum_cell_mass_error": rep.max_cell_error,
            "cdf_lost_positive_cells": int(rep.lost_cells),
            "cdf_lost_probability_mass": rep.lost_mass,
            "maximum_accepted_l1_mass_error": s["cdf_max_l1"],
            "maximum_accepted_lost_probability_mass": s["cdf_max_lost_mass"],
        },
        "incident_inclination_degrees": cond.alpha,
        "query_grid": [s["query_theta"], s["query_phi"]],
        **model,
        "storage_processing": {
            "additional_filter": "none",
            "additional_filter_is_paper_diffraction": False,
            "sigma_degrees": 0.0,
            "support_sigma": 4.0,
            "gaussian_underresolved": False,
            "aggregation": "nested_solid_angle_mass_preserving",
            "gaussian_integral_relative_change": 0.0,
            "aggregation_integral_relative_change": rep.aggregation_error,
            "coarsening_tv": rep.coarsening_tv,
            "maximum_accepted_coarsening_tv": s["maximum_coarsening_tv"],
        },
        "hg": {
            "g": rep.g_stored,
            "method": "first_moment_of_saved_cell_pdf",
            "target": "saved_cdf",
            "cosine_convention": "dot(incident_propagation,outgoing_propagation)",
            "axis_particle_frame": k,
            "g_source": rep.g_source,
            "g_filtered": rep.g_source,
            "cdf_mass_check": rep.cdf_mass,
        },
    }


# ------------------------------------------------------------------ device engine
def open_context(device: int) -> ip.Context:
    return ip.Context(device)


def trace_conditions(ctx: ip.Context, conds: list[Condition], s: dict[str, Any]) -> list[Traced]:
    """Trace every condition of `conds` in one persistent launch (all records'
    rays in one tile queue, then all CDFs) and read back each record's results.
    Audits are evaluated per record afterwards, so a failing record does not
    hide the ones before it."""
    if s["stage"] != "incoherent":
        return [trace_coherent(ctx, cond, s) for cond in conds]
    records = [device_record(cond, s) for cond in conds]
    with ip.Batch(ctx, records, policy(s)) as b:
        b.run(ip.RUN_TRACE | ip.RUN_CDF | ip.RUN_NO_AUDIT)
        out = []
        for r in range(len(records)):
            phi, theta, rep = b.cdf(r)
            out.append(Traced(b.stats(r), phi, theta, rep))
    return out


def trace_coherent(ctx: ip.Context, cond: Condition, s: dict[str, Any]) -> Traced:
    """One condition through the coherent pipeline: vertex trace, patches,
    query, optics, Table II, gate and CDF; a refused record keeps its message."""
    pol = policy(s)
    with ip.Coherent(ctx, cond.radius, cond.wavelength, cond.n_solver, grid=s["grid"], theta=s["query_theta"],
                     phi=s["query_phi"], inclination_deg=cond.alpha, sphere=bool(cond.sphere and not any(cond.shape)),
                     focal_offsets=tuple(s["focal_offsets"])) as h:
        h.trace()
        h.evaluate()
        try:
            phi, theta, rep, gate = h.cdf(s["stage"], s["allow_underresolved"], pol.out_theta, pol.out_phi,
                                          pol.max_l1, pol.max_lost_mass, pol.max_coarsening_tv)
            return Traced(h.stats(), phi, theta, rep, gate)
        except ip.IcePhaseError as exc:
            empty = np.zeros(0)
            return Traced(h.stats(), empty, empty, ip.CdfReport(), None, str(exc))


def batch_capacity(s: dict[str, Any]) -> int:
    """Records per launch that keep the batch's device buffers under BATCH_BYTES."""
    per = 8 * (s["query_theta"] * s["query_phi"] + 2 * s["cdf_phi"] * (s["cdf_theta"] + 1) +
               s["cdf_theta"] * s["cdf_phi"]) + (1 << 20)
    return max(1, min(MAX_BATCH_RECORDS, BATCH_BYTES // per))


def cli_command(out: Path, c: dict[str, Any], ii: int, ll: int) -> list[str]:
    """The bin/rainbow_generate command line that writes the same record alone."""
    s = c["solver"]
    values = {"out": str(out), "material": "water", "radius-mm": str(c["radius_mm"]),
              "wavelength-nm": c["wavelength_nm"][ll], "inclination-deg": c["inclination_degrees"][ii],
              "temperature-c": str(c["temperature_c"]), "pressure-pa": str(c["pressure_pa"]), "grid": s["grid"],
              "query-theta": s["query_theta"], "query-phi": s["query_phi"], "cdf-theta": s["cdf_theta"],
              "cdf-phi": s["cdf_phi"], "stage": s["stage"], "device": s["device"],
              "max-coarsening-tv": s["maximum_coarsening_tv"], "cdf-max-l1": s["cdf_max_l1"],
              "cdf-max-lost-mass": s["cdf_max_lost_mass"], "focal-offsets": ",".join(map(str, s["focal_offsets"]))}
    args = [str(PACKAGE / "bin" / "rainbow_generate")]
    for k, v in values.items():
        args += ["--" + k, str(v)]
    if c["force_sphere"]:
        args.append("--sphere")
    if s["allow_underresolved"]:
        args.append("--allow-underresolved")
    return args


def _log_lines(cond: Condition, t: Traced | None, meta: dict[str, Any] | None, elapsed: float) -> str:
    lines = [f"Material=water_iapws_r9_97 wavelength_vacuum_nm={cond.wavelength!r} "
             f"interior_index_solver={cond.n_solver!r} exterior_index=1; additional_storage_gaussian=none"]
    if t is not None and meta is not None:
        rep = t.report
        lines.append(f"Saved NumPy CDF dataset: {cond.path}; elapsed_seconds={elapsed!r}; "
                     f"CDF_L1_mass_error={rep.l1_error!r}; lost_probability_mass={rep.lost_mass!r}; "
                     f"saved_grid={meta['theta_count']}x{meta['phi_count']}; hg_g={rep.g_stored!r}; "
                     f"projected_area_mm2={cond.area.area_mm2!r}; coarsening_TV={rep.coarsening_tv!r}")
    return "\n".join(lines) + "\n"


# ------------------------------------------------------------------ driver
def run(argv: list[str] | None = None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--config", type=Path, required=True)
    ap.add_argument("--out-root", type=Path)
    ap.add_argument("--generator", type=Path, help="accepted for command-line compatibility; the engine is "
                    "libicephase.so in this process")
    ap.add_argument("--wavelength-range", nargs=3, metavar=("START_NM", "STOP_NM", "STEP_NM"))
    ap.add_argument("--radius-mm")
    ap.add_argument("--inclination-count", type=int)
    ap.add_argument("--allow-underresolved", action="store_true",
                    help="explicitly permit recorded underresolution warnings")
    ap.add_argument("--resume", action="store_true")
    ap.add_argument("--limit", type=int, help="run at most this many NEW records, without changing the planned "
                    "dataset")
    ap.add_argument("--batch-records", type=int, help="conditions traced per persistent launch (default: as many "
                    "as fit in 8 GiB of device buffers, at most 256)")
    mode = ap.add_mutually_exclusive_group()
    mode.add_argument("--execute", action="store_true", help="actually start the simulations; default is a dry run")
    mode.add_argument("--dry-run", action="store_true")
    args = ap.parse_args(argv)
    if args.limit is not None and args.limit <= 0:
        raise ValueError("limit must be positive")
    if args.batch_records is not None and args.batch_records <= 0:
        raise ValueError("batch-records must be positive")
    raw = load_json(args.config)
    if args.wavelength_range:
        raw["wavelengths"] = dict(zip(("start_nm", "stop_nm", "step_nm"), args.wavelength_range))
    if args.radius_mm is not None:
        raw["radius_mm"] = args.radius_mm
    if args.inclination_count is not None:
        raw["inclinations"]["count"] = args.inclination_count
    if args.allow_underresolved:
        raw["solver"]["allow_underresolved"] = True
    c = normalized_config(raw)
    s = c["solver"]
    root = args.out_root or Path(raw["output_root"])
    # a relative output root is taken from the repository root, as the reference does
    root = (root if root.is_absolute() else PACKAGE.parent / root).resolve()
    reserve = int(scalar(raw["reserve_gib"], 0, 1000000) * (1 << 30))
    nt, np_ = s["cdf_theta"], s["cdf_phi"]
    payload = 8 * ((np_ + 1) + np_ * (nt + 1) + (nt + 1))
    per_record = payload + (128 << 10)  # headers, metadata, receipt; logs extra
    ni, nl = len(c["inclination_degrees"]), len(c["wavelength_nm"])
    per_launch = args.batch_records or batch_capacity(s)
    print(f"conditions={ni} x {nl} = {ni * nl}; CDF={nt}x{np_}; array_bytes_per_record={payload}")
    print(f"array_total_TB={payload * ni * nl / 1e12:.6f}; output={root}")
    print(f"inclination_centers=[{c['inclination_degrees'][0]}, {c['inclination_degrees'][-1]}] degrees")
    print(f"wavelengths=[{c['wavelength_nm'][0]}, {c['wavelength_nm'][-1]}] nm; inclusive exact-step endpoints")
    print(f"In-process: up to {per_launch} conditions per persistent launch on device {s['device']}.")
    if s["allow_underresolved"]:
        print("WARNING: underresolved diffraction is explicitly permitted, not certified accurate.")
    if s["maximum_coarsening_tv"] == 1:
        print("WARNING: maximum_coarsening_tv=1 is not an accuracy acceptance threshold.")
    if not args.execute:
        print("DRY RUN: no simulations or files written. Add --execute to start.")
        print("Equivalent single record: " + " ".join(cli_command(root / "records/i0000/w0000", c, 0, 0)))
        return 0
    files = engine_files()
    for name, path in files.items():
        if not path.is_file():
            raise FileNotFoundError(f"Missing {name}: {path}")
    spec = {"version": VERSION, "conditions": c, "binary_sha256": {n: sha256(p) for n, p in files.items()}}
    signature = object_hash(spec)
    plan = {**spec, "signature": signature,
            "record_layout": "records/i{incident_index:04d}/w{wavelength_index:04d}"}
    root.mkdir(parents=True, exist_ok=True)
    with batch_lock(root):
        manifest = root / "batch_plan.json"
        if manifest.exists():
            if not args.resume:
                raise FileExistsError("Dataset plan exists; use --resume with identical configuration and binaries")
            if load_json(manifest) != plan:
                raise ValueError("Resume rejected: configuration/binaries differ from batch_plan.json")
        else:
            if any(p.name != ".batch.lock" for p in root.iterdir()):
                raise ValueError("Refusing to adopt a nonempty directory without this batch plan")
            write_json(manifest, plan)
        pending: list[tuple[int, int]] = []
        verified = 0
        for ii in range(ni):
            for ll in range(nl):
                job = f"i{ii:04d}_w{ll:04d}"
                path = root / f"records/i{ii:04d}/w{ll:04d}"
                receipt = root / "receipts" / (job + ".json")
                if Path(str(path) + ".part").exists():
                    raise ValueError(f"Incomplete .part exists: {path}.part; inspect and move it manually before "
                                     "retrying")
                if path.exists():
                    if not args.resume or not receipt.exists():
                        raise ValueError(f"Existing record without authorized resume/receipt: {path}")
                    saved = load_json(receipt)
                    actual = verify_record(path, plan, ii, ll)
                    if (saved.get("signature") != signature or saved.get("incident_index") != ii
                            or saved.get("wavelength_index") != ll or saved.get("record") != actual):
                        raise ValueError(f"Resume receipt/content mismatch: {path}")
                    verified += 1
                else:
                    if receipt.exists():
                        raise ValueError(f"Receipt exists but record missing: {path}")
                    pending.append((ii, ll))
        selected = pending if args.limit is None else pending[:args.limit]
        free = shutil.disk_usage(root).free
        print(f"verified_existing={verified}; pending={len(pending)}; this_run={len(selected)}; "
              f"free_TB={free / 1e12:.3f}")
        needed = len(selected) * per_record + reserve
        if needed > free:
            raise RuntimeError(f"Insufficient free space: estimated {needed} bytes including reserve; "
                               f"available {free}")
        start = time.monotonic()
        done = 0
        ctx = open_context(s["device"]) if selected else None
        chunks = [selected[k:k + per_launch] for k in range(0, len(selected), per_launch)]

        def launch(index: int):
            conds = [resolve(c, ii, ll, root) for ii, ll in chunks[index]]
            t0 = time.monotonic()
            traced = trace_conditions(ctx, conds, s)
            return conds, traced, (time.monotonic() - t0) / len(conds)

        # the next launch is traced in a worker thread while this one's records
        # are written (the device calls release the interpreter lock)
        pool = ThreadPoolExecutor(max_workers=1) if len(chunks) > 1 else None
        ahead = None
        try:
            for index in range(len(chunks)):
                conds, traced, share = ahead.result() if ahead is not None else launch(index)
                ahead = pool.submit(launch, index + 1) if pool is not None and index + 1 < len(chunks) else None
                if shutil.disk_usage(root).free < len(conds) * per_record + reserve:
                    raise RuntimeError("Free-space reserve reached; stopping without lowering precision")
                for cond, t in zip(conds, traced):
                    done = _finish(cond, t, c, plan, signature, root, share, done, verified, ni * nl, start,
                                   len(selected))
        finally:
            if ahead is not None:
                try:
                    ahead.result()  # the device finishes before the context goes away; nothing is written
                except Exception:
                    pass
            if pool is not None:
                pool.shutdown(wait=True)
            if ctx is not None:
                ctx.close()
        complete = verified + len(selected) == ni * nl
        write_json(root / "progress.json", {"completed": verified + len(selected), "total": ni * nl,
                                            "complete": complete,
                                            "elapsed_seconds_this_run": time.monotonic() - start})
        print("Dataset complete." if complete else
              "Requested subset finished. Repeat with --resume to continue the same plan.")
    return 0


def _finish(cond: Condition, t: Traced, c: dict[str, Any], plan: dict[str, Any], signature: str, root: Path,
            share: float, done: int, verified: int, total: int, start: float, count: int) -> int:
    """Write one traced condition: log, audits, record, check, receipt, progress."""
    (root / "logs").mkdir(exist_ok=True)
    log = root / "logs" / (cond.job + ".log")
    if log.exists():
        log = root / "logs" / (cond.job + f".{time.time_ns()}.log")
    n = done + 1
    print(f"[{n}/{count}] {cond.job}: alpha={c['inclination_degrees'][cond.ii]} "
          f"lambda={c['wavelength_nm'][cond.ll]}", flush=True)
    failure = t.failure
    if failure is None and t.report.status:
        failure = f"CDF audit failed (status {t.report.status})"
    elif failure is None and c["solver"]["stage"] == "incoherent" and t.stats.invalid:
        failure = "Histogram has invalid outgoing directions."
    if failure is not None:
        with log.open("x", encoding="utf-8") as f:
            f.write(_log_lines(cond, None, None, share) + f"rainbow_generate: {failure}\n")
        raise RuntimeError(f"Generator failed ({failure}); stopped at {cond.job}. See {log}. "
                           "No failed record was skipped or deleted.")
    meta = metadata(cond, c, t)
    cond.path.parent.mkdir(parents=True, exist_ok=True)
    ip.write_record(cond.path, json.dumps(meta, indent=2, allow_nan=False), t.phi, t.theta)
    with log.open("x", encoding="utf-8") as f:
        f.write(_log_lines(cond, t, meta, share))
    checked = verify_record(cond.path, plan, cond.ii, cond.ll)
    write_json(cond.receipt, {"signature": signature, "incident_index": cond.ii, "wavelength_index": cond.ll,
                              "record": checked})
    write_json(root / "progress.json", {"completed": verified + n, "total": total, "last_record": cond.job,
                                        "elapsed_seconds_this_run": time.monotonic() - start})
    print(f"  saved; g={checked['hg_g']:.9g}; Aproj_mm2={checked['projected_area_mm2']:.9g}", flush=True)
    return n


def main() -> None:
    try:
        raise SystemExit(run())
    except KeyboardInterrupt:
        print("Interrupted. No automatic cleanup of records/.part directories.", file=sys.stderr)
        raise SystemExit(130)
    except (OSError, ValueError, RuntimeError, KeyError, TypeError, struct.error) as exc:
        print(f"generate_phase_dataset: {exc}", file=sys.stderr)
        raise SystemExit(1)


if __name__ == "__main__":
    main()
