"""Read-only inspection and exact-cell comparison of saved phase records.

Drop-in for the reference's analysis tool (tools/phase_analysis.py): the same
`inspect` and `compare` commands, the same reports (analysis.json,
theta_profile.json, comparison.json, analysis_complete.json), the same rules:
the saved CDF is never altered, re-normalized, floored or resampled; a
comparison integrates the two saved piecewise-constant densities exactly on
the common refinement of their cells (union of the u edges, the common
rational azimuth lattice), refuses different frames or coordinate contracts,
and needs explicit acknowledgement of differing physical metadata or
orientation plans. Plots need matplotlib, which this image does not carry:
pass --no-plots (a plot request without it fails with the import error, as
in the reference).

Host tooling: it reads record files through `PhaseRecord` and does no device
work.

Run as `python -m ice_crystal_phase_amd.analysis inspect RECORD --out DIR
--no-plots` or `... compare A B --out DIR --no-plots`.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import math
import sys
from dataclasses import dataclass
from functools import cached_property
from pathlib import Path
from typing import Any

import numpy as np

from .record import PhaseRecord

VERSION = "ice.phase_analysis.v1"
FILES = ("metadata.json", "phi_cdf.npy", "theta_given_phi_cdf.npy", "u_edges.npy")
CONTRACTS = ("ice.phase_cdf.ensemble_coordinates.v1", "rainbow.phase_cdf.coordinates.v1")
FRAME_TOLERANCE = 2e-12


def _positive(value: str) -> int:
    n = int(value)
    if n <= 0:
        raise argparse.ArgumentTypeError("must be positive")
    return n


def sha256(path: Path) -> str:
    h = hashlib.sha256()
    with path.open("rb") as f:
        for chunk in iter(lambda: f.read(1 << 20), b""):
            h.update(chunk)
    return h.hexdigest()


def total(values: np.ndarray) -> float:
    """Correctly rounded sum of all entries (one sum whatever the order, so the
    entries go to fsum as one Python list)."""
    return math.fsum(np.ravel(np.asarray(values, dtype=np.float64)).tolist())


def theta_of_u(u: np.ndarray) -> np.ndarray:
    """theta from u = (1 - cos theta)/2, through the arcsine branch that is
    stable near each pole."""
    return np.where(u <= 0.5, 2 * np.arcsin(np.sqrt(u)), np.pi - 2 * np.arcsin(np.sqrt(1 - u)))


@dataclass
class Table:
    """A saved record as cell probabilities [phi_cell, theta_cell]."""
    directory: Path
    metadata: dict[str, Any]
    frame: np.ndarray
    u: np.ndarray
    v: np.ndarray
    mass: np.ndarray
    hashes: dict[str, str]

    @property
    def du(self) -> np.ndarray:
        return np.diff(self.u)

    @property
    def dv(self) -> np.ndarray:
        # equal azimuth cells are exactly 1/Nphi wide by the record contract
        return np.full(self.mass.shape[0], 1 / self.mass.shape[0])

    @property
    def density_uv(self) -> np.ndarray:
        return self.mass / (self.du[None, :] * self.dv[:, None])

    @property
    def density_omega(self) -> np.ndarray:
        return self.density_uv / (4 * np.pi)

    @cached_property
    def theta_mass(self) -> np.ndarray:
        return np.array([math.fsum(col) for col in self.mass.T.tolist()])

    @cached_property
    def g(self) -> float:
        return total(self.mass * (1 - self.u[:-1] - self.u[1:])) / total(self.mass)


def _strict_json(text: str) -> dict[str, Any]:
    def pairs(items):
        out: dict[str, Any] = {}
        for k, v in items:
            if k in out:
                raise ValueError(f"Duplicate metadata key: {k}")
            out[k] = v
        return out

    def constant(token):
        raise ValueError(f"Nonfinite metadata number: {token}")

    value = json.loads(text, object_pairs_hook=pairs, parse_constant=constant)

    def finite(x):
        if isinstance(x, float) and not math.isfinite(x):
            raise ValueError("Nonfinite metadata number")
        if isinstance(x, dict):
            for y in x.values():
                finite(y)
        elif isinstance(x, list):
            for y in x:
                finite(y)

    finite(value)
    return value


def load_table(path: str | Path, max_cells: int = 10_000_000) -> Table:
    """Validate a record directory (PhaseRecord, validate=True) and read its
    cell probabilities; refuses records above max_cells instead of coarsening."""
    p = Path(path).resolve()
    if max_cells <= 0:
        raise ValueError("max_cells must be positive")
    if not p.is_dir():
        raise ValueError("Input must be an extracted record directory, not a ZIP.")
    if (p / "metadata.json").stat().st_size > 1024 * 1024:
        raise ValueError("Metadata exceeds 1 MiB")
    m = _strict_json((p / "metadata.json").read_text(encoding="utf-8"))
    for key, want in (("direction_convention", "physical_propagation"), ("theta_zero", "forward"),
                      ("theta_pi", "backward")):
        if m.get(key) != want:
            raise ValueError(f"Missing or unsupported {key}; no direction convention is inferred.")
    if m.get("coordinate_contract_id") not in CONTRACTS:
        raise ValueError("Unknown coordinate contract")
    nt, nf = m.get("theta_count"), m.get("phi_count")
    if type(nt) is not int or type(nf) is not int or min(nt, nf) < 1:
        raise ValueError("Invalid dimensions")
    if nt * nf > max_cells:
        raise ValueError(f"{nt * nf} cells exceeds --max-cells={max_cells}; no automatic coarsening.")
    with PhaseRecord(p, validate=True) as r:
        a = np.array(r.phi_cdf, dtype=np.float64)
        u = np.array(r.u_edges, dtype=np.float64)
        mass = np.diff(r.theta_cdf, axis=1) * np.diff(a)[:, None]
        frame = np.array(r.frame, dtype=np.float64)
    if not np.isfinite(mass).all() or (mass < 0).any():
        raise ValueError("Invalid reconstructed cell probability")
    t = Table(p, m, frame, u, np.arange(nf + 1, dtype=np.float64) / nf, mass, {n: sha256(p / n) for n in FILES})
    if not np.isfinite(t.density_omega).all():
        raise ValueError("Density cannot be represented; input not modified.")
    for name, column in (("query_frame_axis", 2), ("query_frame_e0", 0)):
        if name in m:
            q = np.asarray(m[name], dtype=float)
            if q.shape != (3,) or not np.allclose(q, frame[:, column], rtol=0, atol=FRAME_TOLERANCE):
                raise ValueError(f"{name} disagrees with sampling frame")
    return t


def cone_mass(t: Table, degrees: float) -> float:
    """Probability inside the cone theta <= degrees about the forward axis,
    from exact overlaps of the cells with u <= sin^2(theta/2)."""
    if not 0 <= degrees <= 180:
        raise ValueError("Cone angle must be in [0,180]")
    edge = math.sin(math.radians(degrees) / 2) ** 2
    overlap = np.maximum(0, np.minimum(t.u[1:], edge) - t.u[:-1])
    return total(t.theta_mass * overlap / t.du)


def angular_summary(t: Table) -> dict[str, Any]:
    m, w = t.metadata, t.mass
    nphi, ntheta = w.shape
    d = t.density_omega
    marginal = t.theta_mass
    theta = np.degrees(theta_of_u(t.u))
    top = np.argsort(-marginal, kind="stable")[:min(10, ntheta)]
    positive = d[d > 0]
    orientation = m.get("orientation", {})
    warnings = []
    if orientation.get("rows", 0) <= 2:
        warnings.append("At most two discrete orientations: not a resolved continuous orientation distribution.")
    tv = m.get("storage_processing", {}).get("coarsening_tv")
    if tv is not None and tv > 0.05:
        warnings.append("Logged coarsening TV exceeds the advisory value 0.05; mass conservation does not preserve "
                        "angular shape.")
    if not m.get("quality", {}).get("angular_convergence_certified", False):
        warnings.append("Angular convergence has not been certified. One record cannot establish "
                        "ray/orientation/grid convergence.")
    if not orientation.get("plan_embedded", False):
        warnings.append("Orientation plan is not embedded; preserve the original CSV and its hash.")
    if m.get("material", {}).get("automatic_ice_dispersion") is False:
        warnings.append("Index was supplied by the caller; wavelength alone does not prove use of an ice dispersion "
                        "model.")
    if m.get("histogram", {}).get("forward_axis_mass_mm2", 0) > 0:
        warnings.append("Forward-axis mass is spread over a finite polar cap by storage; this is not physical "
                        "diffraction.")
    stored_g = m.get("hg", {}).get("g")
    g_rays = m.get("histogram", {}).get("g_rays")
    g_source = m.get("storage_processing", {}).get("g_source")
    audit = m.get("transport_audit", {})
    area = m.get("geometry", {}).get("projected_area_mm2")
    energy = None
    if area and "escaped_mm2" in audit and "unresolved_mm2" in audit:
        energy = (audit["escaped_mm2"] + audit["unresolved_mm2"]) / area - 1
    return {
        "analysis_version": VERSION,
        "input_directory": str(t.directory),
        "sha256": t.hashes,
        "reader_validation": "passed",
        "computed_from_saved_arrays": {
            "shape_phi_theta": [nphi, ntheta],
            "probability_mass": total(w),
            "g": t.g,
            "g_minus_metadata": None if stored_g is None else t.g - stored_g,
            "positive_cells": int(np.count_nonzero(w)),
            "zero_cells": int(np.count_nonzero(w == 0)),
            "zero_cell_fraction": float(np.mean(w == 0)),
            "positive_support_solid_angle_sr": total((w > 0) * 4 * np.pi * t.dv[:, None] * t.du[None, :]),
            "pdf_min_positive_sr_inverse": float(positive.min()),
            "pdf_max_sr_inverse": float(positive.max()),
            "first_polar_band_mass": float(marginal[0]),
            "last_polar_band_mass": float(marginal[-1]),
            "forward_hemisphere_mass": cone_mass(t, 90),
            "cone_probabilities": {str(a): cone_mass(t, a) for a in (1, 2.5, 5, 10, 22, 30, 46, 60, 90, 180)},
            "highest_mass_theta_bands": [{"lower_deg": float(theta[i]), "upper_deg": float(theta[i + 1]),
                                          "mass": float(marginal[i])} for i in top],
        },
        "computed_from_metadata_numbers_not_retraced": {
            "energy_balance_relative": energy,
            "g_source_minus_g_rays": None if g_rays is None or g_source is None else g_source - g_rays,
            "g_saved_minus_g_source": None if g_source is None else t.g - g_source,
            "g_saved_minus_g_rays": None if g_rays is None else t.g - g_rays,
        },
        "producer_metadata_not_independently_recomputed": {
            k: m.get(k) for k in ("backend", "geometry", "material", "orientation", "sampling", "transport_audit",
                                  "histogram", "storage_processing", "quality")},
        "warnings": warnings,
        "scope": "Saved finite-cell probability distribution only. No claim of physical or sampling convergence.",
    }


def check_frames(a: Table, b: Table) -> float:
    if a.metadata.get("coordinate_contract_id") != b.metadata.get("coordinate_contract_id"):
        raise ValueError("Coordinate contracts differ; comparison would require an explicit transformation.")
    if a.metadata.get("sampling_frame_map") != b.metadata.get("sampling_frame_map"):
        raise ValueError("Coordinate spaces differ")
    error = float(np.max(np.abs(a.frame - b.frame)))
    if error > FRAME_TOLERANCE:
        raise ValueError("Sampling frames differ; no automatic rotation/interpolation is performed.")
    return error


def _js_terms(p: np.ndarray, q: np.ndarray) -> np.ndarray:
    """Pointwise Jensen-Shannon integrand m F(|p-q|/(p+q)), m = (p+q)/2,
    F(t) = ((1+t) ln(1+t) + (1-t) ln(1-t))/2, by its series
    sum t^(2n)/(2n(2n-1)) below t = 0.01 so that tiny differences do not cancel."""
    mix = 0.5 * (p + q)
    t = np.zeros_like(p)
    use = mix > 0
    t[use] = np.abs(p[use] - q[use]) / (p[use] + q[use])
    f = np.zeros_like(t)
    small = t < 0.01
    x = t[small] * t[small]
    f[small] = x * (0.5 + x * (1 / 12 + x * (1 / 30 + x * (1 / 56 + x / 90))))
    mid = (t >= 0.01) & (t < 1)
    y = t[mid]
    f[mid] = 0.5 * ((1 + y) * np.log1p(y) + (1 - y) * np.log1p(-y))
    f[t == 1] = math.log(2.0)
    return mix * f


def compare_tables(a: Table, b: Table, max_refined_cells: int = 50_000_000) -> dict[str, Any]:
    """Distances between the two saved densities, integrated exactly over the
    common refinement (every refined cell lies inside one cell of each record,
    where both densities are constant)."""
    frame_error = check_frames(a, b)
    u = np.union1d(a.u, b.u)
    na, nb = a.mass.shape[0], b.mass.shape[0]
    lattice = math.lcm(na, nb)
    if lattice > np.iinfo(np.int64).max:
        raise ValueError("Common phi lattice exceeds int64 range")
    sa, sb = lattice // na, lattice // nb
    v = np.union1d(np.arange(na + 1, dtype=np.int64) * sa, np.arange(nb + 1, dtype=np.int64) * sb)
    cells = (len(u) - 1) * (len(v) - 1)
    if max_refined_cells <= 0 or cells > max_refined_cells:
        raise ValueError(f"{cells} common-refinement cells exceeds limit; no downsampling or edge snapping.")
    # each refined cell is located by its lower edge (a midpoint can round onto
    # an edge when the interval is one ulp wide)
    ia = np.searchsorted(a.u, u[:-1], side="right") - 1
    ib = np.searchsorted(b.u, u[:-1], side="right") - 1
    ja, jb = v[:-1] // sa, v[:-1] // sb
    da, db = a.density_uv, b.density_uv
    du = np.diff(u)
    dv = np.diff(v).astype(np.float64) / lattice
    parts: dict[str, list[float]] = {k: [] for k in ("l1", "l2", "h2", "js", "a_only", "b_only", "mass_a", "mass_b")}
    for row in range(len(v) - 1):
        p, q = da[ja[row], ia], db[jb[row], ib]
        area = du * dv[row]
        diff = p - q
        parts["l1"].append(total(area * np.abs(diff)))
        parts["l2"].append(total(area * diff * diff / (4 * np.pi)))
        parts["h2"].append(total(0.5 * area * (np.sqrt(p) - np.sqrt(q)) ** 2))
        parts["js"].append(total(area * _js_terms(p, q)))
        parts["a_only"].append(total(area * np.where(q == 0, p, 0)))
        parts["b_only"].append(total(area * np.where(p == 0, q, 0)))
        parts["mass_a"].append(total(area * p))
        parts["mass_b"].append(total(area * q))
    s = {k: math.fsum(x) for k, x in parts.items()}
    if not all(math.isfinite(x) for x in s.values()):
        raise ValueError("Nonfinite metric; values were not clamped or re-normalized")
    ma, mb = a.theta_mass / a.du, b.theta_mass / b.du
    return {
        "analysis_version": VERSION,
        "target": "saved_piecewise_constant_solid_angle_densities",
        "integration": "union_of_original_u_edges_and_exact_rational_equal_phi_edges; no bin merging, resampling, "
                       "smoothing, or normalization",
        "common_refinement_shape_phi_theta": [len(v) - 1, len(u) - 1],
        "frame_max_absolute_difference": frame_error,
        "frame_equivalence_absolute_tolerance": FRAME_TOLERANCE,
        "probability_mass_a_integrated": s["mass_a"], "probability_mass_b_integrated": s["mass_b"],
        "l1": s["l1"], "tv": 0.5 * s["l1"],
        "l2_per_sqrt_sr": math.sqrt(s["l2"]),
        "hellinger": math.sqrt(s["h2"]), "jensen_shannon_nats": s["js"],
        "mass_a_where_b_is_zero": s["a_only"], "mass_b_where_a_is_zero": s["b_only"],
        "theta_marginal_tv": 0.5 * total(du * np.abs(ma[ia] - mb[ib])),
        "g_a": a.g, "g_b": b.g, "g_a_minus_b": a.g - b.g,
        "input_a_sha256": a.hashes, "input_b_sha256": b.hashes,
        "scope": "No confidence interval from a single pair; does not certify physical or statistical convergence.",
    }


def physical_differences(a: Table, b: Table) -> list[str]:
    """Metadata that would make a comparison one of two different physical problems."""
    keys = ("wavelength_nm", "interior_index", "exterior_index", "stage", "input_polarization", "material",
            "diffraction_policy", "inter_path_interference", "geometry.shape", "geometry.circumradius_mm",
            "geometry.full_length_mm")

    def at(m: Any, key: str) -> Any:
        for part in key.split("."):
            if not isinstance(m, dict):
                return None
            m = m.get(part)
        return m

    return [k for k in keys if at(a.metadata, k) != at(b.metadata, k)]


def _write_json(path: Path, data: Any) -> None:
    with path.open("x", encoding="utf-8", newline="\n") as f:
        json.dump(data, f, indent=2, ensure_ascii=True, allow_nan=False)
        f.write("\n")


def _new_dir(path: str | Path) -> Path:
    out = Path(path).resolve()
    out.mkdir(parents=True, exist_ok=False)  # an existing directory, even empty, is refused
    return out


def _plot(*_args: Any) -> None:
    import matplotlib  # noqa: F401  (raises ImportError where matplotlib is absent)
    raise ImportError("Plots are not implemented in this package; use --no-plots.")


def parser() -> argparse.ArgumentParser:
    p = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    sub = p.add_subparsers(dest="command", required=True)
    one = sub.add_parser("inspect", help="Validate and summarize one extracted CDF record")
    one.add_argument("record")
    one.add_argument("--label", default="Phase CDF")
    two = sub.add_parser("compare", help="Integrate differences on a common refinement")
    two.add_argument("a")
    two.add_argument("b")
    two.add_argument("--allow-orientation-plan-difference", action="store_true",
                     help="compare records whose orientation CSV hashes differ or are absent")
    two.add_argument("--allow-physical-differences", action="store_true")
    two.add_argument("--max-refined-cells", type=_positive, default=50_000_000)
    for q in (one, two):
        q.add_argument("--out", required=True, help="New output directory; existing paths are refused")
        q.add_argument("--no-plots", action="store_true")
        q.add_argument("--max-cells", type=_positive, default=10_000_000)
    return p


def main(argv: list[str] | None = None) -> int:
    args = parser().parse_args(argv)
    try:
        if args.command == "inspect":
            t = load_table(args.record, args.max_cells)
            summary = angular_summary(t)
            out = _new_dir(args.out)
            _write_json(out / "analysis.json", summary)
            _write_json(out / "theta_profile.json", {
                "theta_edges_deg": np.degrees(theta_of_u(t.u)).tolist(), "u_edges": t.u.tolist(),
                "theta_band_probability": t.theta_mass.tolist(),
                "azimuth_average_pdf_sr_inverse": (t.theta_mass / (4 * np.pi * t.du)).tolist()})
            if not args.no_plots:
                _plot(t, out, args.label)
            print(f"PASS saved-CDF validation: mass={total(t.mass):.17g}, g={t.g:.17g}")
            for warning in summary["warnings"]:
                print("NOTE:", warning)
        else:
            a, b = load_table(args.a, args.max_cells), load_table(args.b, args.max_cells)
            diff = physical_differences(a, b)
            if diff and not args.allow_physical_differences:
                raise ValueError("Physical metadata differ: " + ", ".join(diff) +
                                 "; explicit --allow-physical-differences required")
            ha = a.metadata.get("orientation", {}).get("csv_sha256")
            hb = b.metadata.get("orientation", {}).get("csv_sha256")
            same_plan = ha is not None and ha == hb
            if not same_plan and not args.allow_orientation_plan_difference:
                raise ValueError("Orientation CSV hashes differ or are absent; explicit "
                                 "--allow-orientation-plan-difference required")
            summary = compare_tables(a, b, args.max_refined_cells)
            summary["physical_metadata_differences"] = diff
            summary["orientation_csv_sha256_equal"] = same_plan
            summary["orientation_plan_difference_acknowledged"] = args.allow_orientation_plan_difference
            out = _new_dir(args.out)
            _write_json(out / "comparison.json", summary)
            if not args.no_plots:
                _plot(a, b, out)
            print(f"PASS saved-density comparison: TV={summary['tv']:.17g}, delta_g={summary['g_a_minus_b']:.17g}")
        _write_json(out / "analysis_complete.json", {
            "complete": True, "analysis_version": VERSION, "tool_sha256": sha256(Path(__file__)),
            "python_version": sys.version, "numpy_version": np.__version__, "input_modified": False,
            "plots_created": not args.no_plots})
        return 0
    except (OSError, ValueError, MemoryError, ImportError) as e:
        print("ERROR: " + ascii(str(e)), file=sys.stderr)
        return 2


if __name__ == "__main__":
    raise SystemExit(main())
