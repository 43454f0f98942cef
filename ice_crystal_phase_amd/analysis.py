"""Inspection, exact-cell comparison and plots of saved phase records.

Drop-in for the reference's record tools: the `inspect` and `compare` commands
of tools/phase_analysis.py and the heat map of tools/plot_phase_cdf.py (here
the `plot` command). Their command lines, report files and keys, and refusal
texts are kept, so a caller's scripts run unchanged; the implementation is this
package's own.

What the tools promise, and this one keeps: the saved CDF is read-only and is
never re-normalized, floored, smoothed or resampled. Cell probabilities are
rebuilt from the two CDFs as dPhi_j * dTheta_ij; densities are those
probabilities over the exact cell solid angles 4 pi du_i / Nphi. A comparison
integrates the two saved piecewise-constant densities exactly on the common
refinement of their cells: the union of the two u-edge sets, and an azimuth
lattice of lcm(Nphi_a, Nphi_b) equal cells, on which both densities are
constant; it refuses records in different frames or coordinate contracts, and
records whose physical metadata or orientation plans differ unless told
otherwise.

Host tooling over `PhaseRecord`; no device work.

  python -m ice_crystal_phase_amd.analysis inspect RECORD --out DIR [--label L] [--no-plots]
  python -m ice_crystal_phase_amd.analysis compare A B --out DIR [--allow-...] [--no-plots]
  python -m ice_crystal_phase_amd.analysis plot RECORD --out FIG.png [--theta-range LO HI] ...
"""
from __future__ import annotations

import argparse
import hashlib
import json
import math
import sys
from functools import cached_property
from pathlib import Path
from typing import Any, Callable, Iterable

import numpy as np

from .record import PhaseRecord

VERSION = "ice.phase_analysis.v1"
RECORD_FILES = ("metadata.json", "phi_cdf.npy", "theta_given_phi_cdf.npy", "u_edges.npy")
COORDINATE_CONTRACTS = frozenset({"ice.phase_cdf.ensemble_coordinates.v1", "rainbow.phase_cdf.coordinates.v1"})
DIRECTION_CONVENTION = {"direction_convention": "physical_propagation", "theta_zero": "forward",
                        "theta_pi": "backward"}
FRAME_TOLERANCE = 2e-12
CONE_ANGLES = (1, 2.5, 5, 10, 22, 30, 46, 60, 90, 180)
# metadata paths whose disagreement means two records describe different physics
PHYSICAL_KEYS = ("wavelength_nm", "interior_index", "exterior_index", "stage", "input_polarization", "material",
                 "diffraction_policy", "inter_path_interference", "geometry.shape", "geometry.circumradius_mm",
                 "geometry.full_length_mm")


# ------------------------------------------------------------------ small helpers
def exact_sum(values: Any) -> float:
    """The correctly rounded sum of every entry (independent of order)."""
    return math.fsum(np.asarray(values, dtype=np.float64).ravel().tolist())


def file_digest(path: Path) -> str:
    digest = hashlib.sha256()
    with path.open("rb") as stream:
        while chunk := stream.read(1 << 20):
            digest.update(chunk)
    return digest.hexdigest()


def polar_angle(u: np.ndarray) -> np.ndarray:
    """theta (rad) of u = (1 - cos theta)/2; the half-angle arcsine is taken on
    the side of the nearer pole, where it is well conditioned."""
    u = np.asarray(u, dtype=np.float64)
    near_forward = 2.0 * np.arcsin(np.sqrt(np.clip(u, 0.0, 1.0)))
    near_backward = np.pi - 2.0 * np.arcsin(np.sqrt(np.clip(1.0 - u, 0.0, 1.0)))
    return np.where(u <= 0.5, near_forward, near_backward)


def lookup(tree: Any, dotted: str) -> Any:
    """metadata['a']['b'] for 'a.b'; None where a level is missing."""
    for part in dotted.split("."):
        if not isinstance(tree, dict):
            return None
        tree = tree.get(part)
    return tree


def read_strict_metadata(path: Path) -> dict[str, Any]:
    """metadata.json with the tools' strictness: at most 1 MiB, no duplicate
    keys, no NaN/Infinity anywhere."""
    if path.stat().st_size > 1 << 20:
        raise ValueError("Metadata exceeds 1 MiB")
    def build(pairs: list[tuple[str, Any]]) -> dict[str, Any]:
        obj: dict[str, Any] = {}
        for key, value in pairs:
            if key in obj:
                raise ValueError(f"Duplicate metadata key: {key}")
            obj[key] = value
        return obj

    def refuse_constant(token: str) -> Any:
        raise ValueError(f"Nonfinite metadata number: {token}")

    meta = json.loads(path.read_text(encoding="utf-8"), object_pairs_hook=build, parse_constant=refuse_constant)
    stack: list[Any] = [meta]
    while stack:
        item = stack.pop()
        if isinstance(item, float) and not math.isfinite(item):
            raise ValueError("Nonfinite metadata number")
        if isinstance(item, dict):
            stack.extend(item.values())
        elif isinstance(item, list):
            stack.extend(item)
    return meta


# ------------------------------------------------------------------ one saved record
class SavedDistribution:
    """A record as cell probabilities `cells[j, i]` (azimuth cell j, polar band
    i), with its u edges, sampling frame, metadata and file digests. Derived
    quantities are computed on first use."""

    def __init__(self, directory: Path, metadata: dict[str, Any], frame: np.ndarray, u: np.ndarray,
                 cells: np.ndarray, digests: dict[str, str]):
        self.directory = directory
        self.metadata = metadata
        self.frame = frame
        self.u = u
        self.cells = cells
        self.digests = digests

    @classmethod
    def open(cls, path: str | Path, max_cells: int = 10_000_000) -> "SavedDistribution":
        if max_cells <= 0:
            raise ValueError("max_cells must be positive")
        root = Path(path).resolve()
        if not root.is_dir():
            raise ValueError("Input must be an extracted record directory, not a ZIP.")
        meta = read_strict_metadata(root / "metadata.json")
        for key, expected in DIRECTION_CONVENTION.items():
            if meta.get(key) != expected:
                raise ValueError(f"Missing or unsupported {key}; no direction convention is inferred.")
        if meta.get("coordinate_contract_id") not in COORDINATE_CONTRACTS:
            raise ValueError("Unknown coordinate contract")
        nt, nphi = meta.get("theta_count"), meta.get("phi_count")
        if type(nt) is not int or type(nphi) is not int or nt < 1 or nphi < 1:
            raise ValueError("Invalid dimensions")
        if nt * nphi > max_cells:
            raise ValueError(f"{nt * nphi} cells exceeds --max-cells={max_cells}; no automatic coarsening.")
        with PhaseRecord(root, validate=True) as rec:
            phi_weight = np.diff(np.asarray(rec.phi_cdf, dtype=np.float64))
            cells = np.diff(np.asarray(rec.theta_cdf, dtype=np.float64), axis=1) * phi_weight[:, None]
            u = np.array(rec.u_edges, dtype=np.float64)
            frame = np.array(rec.frame, dtype=np.float64)
        if not (np.isfinite(cells).all() and (cells >= 0).all()):
            raise ValueError("Invalid reconstructed cell probability")
        dist = cls(root, meta, frame, u, cells, {name: file_digest(root / name) for name in RECORD_FILES})
        if not np.isfinite(dist.density).all():
            raise ValueError("Density cannot be represented; input not modified.")
        # frame columns the producer also wrote out by name must be the sampling frame's
        for key, column in (("query_frame_axis", 2), ("query_frame_e0", 0)):
            if key in meta:
                given = np.asarray(meta[key], dtype=np.float64)
                if given.shape != (3,) or np.max(np.abs(given - frame[:, column])) > FRAME_TOLERANCE:
                    raise ValueError(f"{key} disagrees with sampling frame")
        return dist

    # geometry of the cells
    @property
    def nphi(self) -> int:
        return self.cells.shape[0]

    @property
    def ntheta(self) -> int:
        return self.cells.shape[1]

    @cached_property
    def band_width(self) -> np.ndarray:
        """du of each polar band."""
        return np.diff(self.u)

    @cached_property
    def cell_solid_angle(self) -> np.ndarray:
        """Solid angle of one cell of each band: 4 pi du_i / Nphi (sr); every
        azimuth cell is exactly 1/Nphi of the turn."""
        return 4.0 * np.pi * self.band_width / self.nphi

    @cached_property
    def density(self) -> np.ndarray:
        """p(omega) per sr on each cell, [phi, theta]."""
        return self.cells / self.cell_solid_angle[None, :]

    @cached_property
    def band_mass(self) -> np.ndarray:
        """Probability of each polar band (summed over azimuth, correctly rounded)."""
        return np.array([exact_sum(column) for column in self.cells.T])

    @cached_property
    def mean_cosine(self) -> float:
        """g: cell probabilities times the exact cell mean cosine 1 - (u_i + u_{i+1})."""
        cosine = 1.0 - (self.u[:-1] + self.u[1:])
        return exact_sum(self.band_mass * cosine) / exact_sum(self.band_mass)

    @cached_property
    def theta_edges_deg(self) -> np.ndarray:
        return np.degrees(polar_angle(self.u))

    def cone_probability(self, degrees: float) -> float:
        """Probability inside the forward cone of half-angle `degrees`: each band
        contributes its mass times the fraction of its u interval inside the cone
        (the density is constant in u on a band)."""
        if not 0 <= degrees <= 180:
            raise ValueError("Cone angle must be in [0,180]")
        u_cone = math.sin(math.radians(degrees) / 2) ** 2
        inside = np.clip(u_cone - self.u[:-1], 0.0, None)
        inside = np.minimum(inside, self.band_width)
        return exact_sum(self.band_mass * (inside / self.band_width))


# ------------------------------------------------------------------ inspect
# Advisories: (test on the metadata, text). The texts are the report contract.
ADVISORIES: tuple[tuple[Callable[[dict[str, Any]], bool], str], ...] = (
    (lambda m: (lookup(m, "orientation.rows") or 0) <= 2,
     "At most two discrete orientations: not a resolved continuous orientation distribution."),
    (lambda m: (lookup(m, "storage_processing.coarsening_tv") or 0) > 0.05,
     "Logged coarsening TV exceeds the advisory value 0.05; mass conservation does not preserve angular shape."),
    (lambda m: not lookup(m, "quality.angular_convergence_certified"),
     "Angular convergence has not been certified. One record cannot establish ray/orientation/grid convergence."),
    (lambda m: not lookup(m, "orientation.plan_embedded"),
     "Orientation plan is not embedded; preserve the original CSV and its hash."),
    (lambda m: lookup(m, "material.automatic_ice_dispersion") is False,
     "Index was supplied by the caller; wavelength alone does not prove use of an ice dispersion model."),
    (lambda m: (lookup(m, "histogram.forward_axis_mass_mm2") or 0) > 0,
     "Forward-axis mass is spread over a finite polar cap by storage; this is not physical diffraction."),
)


def _difference(a: Any, b: Any) -> float | None:
    return None if a is None or b is None else a - b


def inspect_report(d: SavedDistribution) -> dict[str, Any]:
    """analysis.json of one record: numbers recomputed from the saved arrays,
    numbers derived from metadata values only, the producer's blocks verbatim,
    and advisories."""
    meta, cells = d.metadata, d.cells
    positive = d.density[d.density > 0]
    strongest = np.argsort(-d.band_mass, kind="stable")[:min(10, d.ntheta)]
    g_meta = lookup(meta, "hg.g")
    g_rays = lookup(meta, "histogram.g_rays")
    g_source = lookup(meta, "storage_processing.g_source")
    area = lookup(meta, "geometry.projected_area_mm2")
    escaped, unresolved = lookup(meta, "transport_audit.escaped_mm2"), lookup(meta, "transport_audit.unresolved_mm2")
    balance = (escaped + unresolved) / area - 1 if area and escaped is not None and unresolved is not None else None
    saved = {
        "shape_phi_theta": [d.nphi, d.ntheta],
        "probability_mass": exact_sum(cells),
        "g": d.mean_cosine,
        "g_minus_metadata": _difference(d.mean_cosine, g_meta),
        "positive_cells": int(np.count_nonzero(cells)),
        "zero_cells": int(cells.size - np.count_nonzero(cells)),
        "zero_cell_fraction": float((cells == 0).mean()),
        "positive_support_solid_angle_sr": exact_sum((cells > 0) * d.cell_solid_angle[None, :]),
        "pdf_min_positive_sr_inverse": float(positive.min()),
        "pdf_max_sr_inverse": float(positive.max()),
        "first_polar_band_mass": float(d.band_mass[0]),
        "last_polar_band_mass": float(d.band_mass[-1]),
        "forward_hemisphere_mass": d.cone_probability(90),
        "cone_probabilities": {str(a): d.cone_probability(a) for a in CONE_ANGLES},
        "highest_mass_theta_bands": [{"lower_deg": float(d.theta_edges_deg[i]),
                                      "upper_deg": float(d.theta_edges_deg[i + 1]),
                                      "mass": float(d.band_mass[i])} for i in strongest],
    }
    derived = {
        "energy_balance_relative": balance,
        "g_source_minus_g_rays": _difference(g_source, g_rays),
        "g_saved_minus_g_source": _difference(d.mean_cosine, g_source),
        "g_saved_minus_g_rays": _difference(d.mean_cosine, g_rays),
    }
    producer = {k: meta.get(k) for k in ("backend", "geometry", "material", "orientation", "sampling",
                                         "transport_audit", "histogram", "storage_processing", "quality")}
    return {
        "analysis_version": VERSION,
        "input_directory": str(d.directory),
        "sha256": d.digests,
        "reader_validation": "passed",
        "computed_from_saved_arrays": saved,
        "computed_from_metadata_numbers_not_retraced": derived,
        "producer_metadata_not_independently_recomputed": producer,
        "warnings": [text for test, text in ADVISORIES if test(meta)],
        "scope": "Saved finite-cell probability distribution only. No claim of physical or sampling convergence.",
    }


def theta_profile(d: SavedDistribution) -> dict[str, Any]:
    return {"theta_edges_deg": d.theta_edges_deg.tolist(), "u_edges": d.u.tolist(),
            "theta_band_probability": d.band_mass.tolist(),
            "azimuth_average_pdf_sr_inverse": (d.band_mass / (4 * np.pi * d.band_width)).tolist()}


# ------------------------------------------------------------------ compare
def same_frame(a: SavedDistribution, b: SavedDistribution) -> float:
    """Largest difference of the two sampling frames; refuses anything beyond
    floating-point roundoff (no rotation or interpolation is ever applied)."""
    if a.metadata.get("coordinate_contract_id") != b.metadata.get("coordinate_contract_id"):
        raise ValueError("Coordinate contracts differ; comparison would require an explicit transformation.")
    if a.metadata.get("sampling_frame_map") != b.metadata.get("sampling_frame_map"):
        raise ValueError("Coordinate spaces differ")
    worst = float(np.max(np.abs(a.frame - b.frame)))
    if worst > FRAME_TOLERANCE:
        raise ValueError("Sampling frames differ; no automatic rotation/interpolation is performed.")
    return worst


def js_kernel(t: np.ndarray) -> np.ndarray:
    """F(t) = ((1+t) ln(1+t) + (1-t) ln(1-t)) / 2 for t = (p-q)/(p+q) in [-1, 1],
    so that the Jensen-Shannon integrand is m F(t), m = (p+q)/2. The two
    logarithms cancel to t^2/2 near t = 0, so for |t| < 0.05 the even Taylor
    series sum_n t^(2n) / (2n (2n-1)) is used through t^12 (the first omitted
    term is below 3e-18 relative there); beyond, log1p loses at most ~1e-14."""
    a = np.abs(t)
    out = np.empty_like(a)
    near = a < 0.05
    s = a[near] ** 2
    series = 1.0 / 132.0
    for n in (5, 4, 3, 2, 1):
        series = 1.0 / (2 * n * (2 * n - 1)) + s * series
    out[near] = s * series
    x = a[~near]
    lower = np.zeros_like(x)
    inner = x < 1.0
    lower[inner] = (1.0 - x[inner]) * np.log1p(-x[inner])
    out[~near] = 0.5 * ((1.0 + x) * np.log1p(x) + lower)
    return out


class CommonRefinement:
    """The cells of the union of two records' grids. Polar bands: the union of
    the two u-edge sets; azimuth: L = lcm(Na, Nb) equal cells, whose edges hold
    every edge of both records exactly as integers k/L. Each refined cell lies in
    exactly one cell of each record, located from its lower-left corner."""

    def __init__(self, a: SavedDistribution, b: SavedDistribution, limit: int):
        self.u = np.union1d(a.u, b.u)
        lattice = math.lcm(a.nphi, b.nphi)
        if lattice > np.iinfo(np.int64).max:
            raise ValueError("Common phi lattice exceeds int64 range")
        step_a, step_b = lattice // a.nphi, lattice // b.nphi
        ticks = np.union1d(np.arange(a.nphi + 1, dtype=np.int64) * step_a,
                           np.arange(b.nphi + 1, dtype=np.int64) * step_b)
        self.shape = (len(ticks) - 1, len(self.u) - 1)
        if limit <= 0 or self.shape[0] * self.shape[1] > limit:
            raise ValueError(f"{self.shape[0] * self.shape[1]} common-refinement cells exceeds limit; "
                             "no downsampling or edge snapping.")
        lower = self.u[:-1]
        self.band_a = np.searchsorted(a.u, lower, side="right") - 1
        self.band_b = np.searchsorted(b.u, lower, side="right") - 1
        self.column_a = ticks[:-1] // step_a
        self.column_b = ticks[:-1] // step_b
        self.du = np.diff(self.u)
        self.dv = np.diff(ticks).astype(np.float64) / lattice

    def row_blocks(self, cells_per_block: int = 1 << 22) -> Iterable[slice]:
        rows = max(1, cells_per_block // max(1, self.shape[1]))
        for start in range(0, self.shape[0], rows):
            yield slice(start, min(start + rows, self.shape[0]))


def compare_report(a: SavedDistribution, b: SavedDistribution, max_refined_cells: int = 50_000_000) -> dict[str, Any]:
    """comparison.json: distances between the two saved densities, integrated
    exactly on their common refinement (densities are per unit (u, v), so each
    refined cell weighs du * dv)."""
    frame_error = same_frame(a, b)
    grid = CommonRefinement(a, b, max_refined_cells)
    pa = a.density * (4 * np.pi)   # per unit u * v
    pb = b.density * (4 * np.pi)
    parts: dict[str, list[float]] = {k: [] for k in ("l1", "l2", "h2", "js", "a_only", "b_only", "mass_a",
                                                      "mass_b")}
    for rows in grid.row_blocks():
        p = pa[np.ix_(grid.column_a[rows], grid.band_a)]
        q = pb[np.ix_(grid.column_b[rows], grid.band_b)]
        w = grid.dv[rows][:, None] * grid.du[None, :]
        gap = p - q
        mean = 0.5 * (p + q)
        t = np.divide(gap, p + q, out=np.zeros_like(gap), where=(p + q) > 0)
        parts["l1"].append(exact_sum(w * np.abs(gap)))
        parts["l2"].append(exact_sum(w * gap * gap / (4 * np.pi)))
        parts["h2"].append(exact_sum(0.5 * w * (np.sqrt(p) - np.sqrt(q)) ** 2))
        parts["js"].append(exact_sum(w * mean * js_kernel(t)))
        parts["a_only"].append(exact_sum(w * np.where(q == 0, p, 0.0)))
        parts["b_only"].append(exact_sum(w * np.where(p == 0, q, 0.0)))
        parts["mass_a"].append(exact_sum(w * p))
        parts["mass_b"].append(exact_sum(w * q))
    s = {k: math.fsum(v) for k, v in parts.items()}
    if not all(math.isfinite(v) for v in s.values()):
        raise ValueError("Nonfinite metric; values were not clamped or re-normalized")
    # theta marginals as densities in u, on the refined bands
    marg_a = a.band_mass / a.band_width
    marg_b = b.band_mass / b.band_width
    marginal_tv = 0.5 * exact_sum(grid.du * np.abs(marg_a[grid.band_a] - marg_b[grid.band_b]))
    return {
        "analysis_version": VERSION,
        "target": "saved_piecewise_constant_solid_angle_densities",
        "integration": "union_of_original_u_edges_and_exact_rational_equal_phi_edges; no bin merging, "
                       "resampling, smoothing, or normalization",
        "common_refinement_shape_phi_theta": list(grid.shape),
        "frame_max_absolute_difference": frame_error,
        "frame_equivalence_absolute_tolerance": FRAME_TOLERANCE,
        "probability_mass_a_integrated": s["mass_a"], "probability_mass_b_integrated": s["mass_b"],
        "l1": s["l1"], "tv": 0.5 * s["l1"],
        "l2_per_sqrt_sr": math.sqrt(s["l2"]),
        "hellinger": math.sqrt(s["h2"]), "jensen_shannon_nats": s["js"],
        "mass_a_where_b_is_zero": s["a_only"], "mass_b_where_a_is_zero": s["b_only"],
        "theta_marginal_tv": marginal_tv,
        "g_a": a.mean_cosine, "g_b": b.mean_cosine, "g_a_minus_b": a.mean_cosine - b.mean_cosine,
        "input_a_sha256": a.digests, "input_b_sha256": b.digests,
        "scope": "No confidence interval from a single pair; does not certify physical or statistical convergence.",
    }


def physical_differences(a: SavedDistribution, b: SavedDistribution) -> list[str]:
    return [k for k in PHYSICAL_KEYS if lookup(a.metadata, k) != lookup(b.metadata, k)]


# ------------------------------------------------------------------ figures
def _pyplot():
    import matplotlib
    matplotlib.use("Agg")
    import matplotlib.pyplot as plt
    return plt


def record_figures(d: SavedDistribution, out: Path, label: str) -> list[str]:
    """pdf_linear.png, pdf_log.png (the saved density on the cell grid, no
    interpolation), azimuth_average.png, theta_marginal.png (with the sin theta
    Jacobian, per degree) and cone_cdf.png."""
    plt = _pyplot()
    from matplotlib.colors import LogNorm
    phi_edges = np.linspace(-180.0, 180.0, d.nphi + 1)
    theta = d.theta_edges_deg
    image = d.density.T
    written = []
    for scale in ("linear", "log"):
        fig, ax = plt.subplots(figsize=(10, 5.4), layout="constrained")
        data = np.ma.masked_equal(image, 0) if scale == "log" else image
        mesh = ax.pcolormesh(phi_edges, theta, data, shading="flat", norm=LogNorm() if scale == "log" else None)
        ax.set(xlabel="Azimuth phi [degrees]", ylabel="Scattering polar angle theta [degrees]", xlim=(-180, 180),
               ylim=(180, 0))
        note = "; blank cells = exact zero" if scale == "log" else ""
        ax.set_title(f"{label}: saved PDF ({'logarithmic' if scale == 'log' else 'linear'} scale)\n"
                     f"{d.ntheta} x {d.nphi} cells; no smoothing{note}")
        fig.colorbar(mesh, ax=ax, label="p(omega) [sr^-1]")
        name = f"pdf_{scale}.png"
        fig.savefig(out / name, dpi=180)
        plt.close(fig)
        written.append(name)
    average = d.band_mass / (4 * np.pi * d.band_width)
    fig, ax = plt.subplots(figsize=(10, 4.5), layout="constrained")
    ax.stairs(average, theta, label="Azimuth-averaged PDF")
    ax.set(xlabel="Scattering polar angle theta [degrees]", ylabel="Azimuth-averaged p [sr^-1]", xlim=(0, 180))
    ax.set_title(f"{label}: azimuth average; each step is a stored theta band")
    fig.savefig(out / "azimuth_average.png", dpi=180)
    plt.close(fig)
    written.append("azimuth_average.png")
    # density of theta per degree inside band i: m_i sin(theta) / (2 du_i) * pi/180
    fig, ax = plt.subplots(figsize=(10, 4.5), layout="constrained")
    samples = 16
    xs = np.linspace(theta[:-1], theta[1:], samples, axis=1)
    ys = (d.band_mass / (2 * d.band_width))[:, None] * np.sin(np.radians(xs)) * (np.pi / 180)
    gap = np.full((d.ntheta, 1), np.nan)
    ax.plot(np.hstack([xs, gap]).ravel(), np.hstack([ys, gap]).ravel(), label="Theta marginal")
    ax.set(xlabel="Scattering polar angle theta [degrees]", ylabel="Probability density per degree", xlim=(0, 180))
    ax.set_title(f"{label}: theta marginal (includes spherical Jacobian)")
    fig.savefig(out / "theta_marginal.png", dpi=180)
    plt.close(fig)
    written.append("theta_marginal.png")
    fig, ax = plt.subplots(figsize=(10, 4.5), layout="constrained")
    angles = np.linspace(0.0, 180.0, 1801)
    ax.plot(angles, [d.cone_probability(float(x)) for x in angles])
    ax.set(xlabel="Forward-cone half-angle [degrees]", ylabel="Probability inside cone", xlim=(0, 180), ylim=(0, 1))
    ax.set_title(f"{label}: integrated forward-cone probability")
    fig.savefig(out / "cone_cdf.png", dpi=180)
    plt.close(fig)
    written.append("cone_cdf.png")
    return written


def comparison_figures(a: SavedDistribution, b: SavedDistribution, out: Path) -> list[str]:
    plt = _pyplot()
    fig, ax = plt.subplots(figsize=(10, 4.5), layout="constrained")
    for d, name in ((a, "A"), (b, "B")):
        ax.stairs(d.band_mass / (4 * np.pi * d.band_width), d.theta_edges_deg, label=name)
    ax.set(xlabel="Scattering polar angle theta [degrees]", ylabel="Azimuth-averaged p [sr^-1]", xlim=(0, 180))
    ax.set_title("Saved distributions A and B: azimuth-averaged PDFs")
    ax.legend()
    fig.savefig(out / "comparison_azimuth_average.png", dpi=180)
    plt.close(fig)
    return ["comparison_azimuth_average.png"]


def _view(edges: np.ndarray, low: float, high: float) -> slice:
    """Cells of `edges` (increasing) that overlap [low, high]."""
    first = int(np.searchsorted(edges, low, side="right")) - 1
    last = int(np.searchsorted(edges, high, side="left"))
    first, last = max(first, 0), min(last, len(edges) - 1)
    if last <= first:
        raise ValueError("Requested view contains no cells")
    return slice(first, last)


def heat_map(d: SavedDistribution, out: Path, theta_range=(0.0, 180.0), phi_range=(-180.0, 180.0),
             log_limits: tuple[float, float] | None = None, decades: float | None = None, dpi: int = 180,
             cell_image: Path | None = None) -> dict[str, Any]:
    """The plot_phase_cdf.py figure: log10 of the saved density (per sr) on its
    cells over a theta/phi window, color limits fixed or a number of decades
    below the global maximum; optionally one pixel per cell without axes.
    Returns the summary written by --report."""
    lo_t, hi_t = theta_range
    lo_p, hi_p = phi_range
    if not (0.0 <= lo_t < hi_t <= 180.0):
        raise ValueError("theta range must satisfy 0.0 <= LOW < HIGH <= 180.0")
    if not (-180.0 <= lo_p < hi_p <= 180.0):
        raise ValueError("phi range must satisfy -180.0 <= LOW < HIGH <= 180.0")
    if dpi <= 0:
        raise ValueError("DPI must be positive")
    positive = d.density[d.density > 0]
    if positive.size == 0:
        raise ValueError("No positive finite PDF values")
    with np.errstate(divide="ignore"):
        log_pdf = np.where(d.density > 0, np.log10(np.where(d.density > 0, d.density, 1.0)), -np.inf)
    top = float(np.log10(positive.max()))
    bottom = float(np.log10(positive.min()))
    if log_limits is not None:
        vmin, vmax = map(float, log_limits)
        if not (math.isfinite(vmin) and math.isfinite(vmax) and vmin < vmax):
            raise ValueError("--log-limits requires finite LOW < HIGH")
    elif decades is not None:
        if not (math.isfinite(decades) and decades > 0):
            raise ValueError("--decades must be positive and finite")
        vmin, vmax = top - decades, top
    else:
        vmin, vmax = bottom, top
    phi_edges = np.linspace(-180.0, 180.0, d.nphi + 1)
    rows = _view(d.theta_edges_deg, lo_t, hi_t)
    cols = _view(phi_edges, lo_p, hi_p)
    window = np.ma.masked_invalid(np.where(np.isfinite(log_pdf), log_pdf, np.nan).T[rows, cols])
    plt = _pyplot()
    fig, ax = plt.subplots(figsize=(10, 5.4), layout="constrained")
    mesh = ax.pcolormesh(phi_edges[cols.start:cols.stop + 1], d.theta_edges_deg[rows.start:rows.stop + 1], window,
                         shading="flat", vmin=vmin, vmax=vmax)
    ax.set(xlabel="phi [degrees]", ylabel="theta [degrees]", xlim=(lo_p, hi_p), ylim=(hi_t, lo_t))
    ax.set_title(f"{d.directory.name}: log10 p(omega) / (1 sr^-1); {d.ntheta} x {d.nphi} cells")
    fig.colorbar(mesh, ax=ax, label="log10 p(omega) [sr^-1]")
    fig.savefig(out, dpi=dpi)
    plt.close(fig)
    if cell_image is not None:
        plt.imsave(cell_image, window.filled(np.nan), vmin=vmin, vmax=vmax)
    return {"cdf_reconstructed_mass": exact_sum(d.cells), "g_reconstructed": d.mean_cosine,
            "global_log10_pdf_min_positive": bottom, "global_log10_pdf_max": top,
            "color_limits_log10": [vmin, vmax], "theta_range_deg": [lo_t, hi_t], "phi_range_deg": [lo_p, hi_p],
            "cells_shown_theta_phi": [rows.stop - rows.start, cols.stop - cols.start]}


# ------------------------------------------------------------------ command line
def _write_json(path: Path, data: Any) -> None:
    with path.open("x", encoding="utf-8", newline="\n") as f:
        json.dump(data, f, indent=2, ensure_ascii=True, allow_nan=False)
        f.write("\n")


def _fresh_directory(path: str | Path) -> Path:
    out = Path(path).resolve()
    out.mkdir(parents=True, exist_ok=False)   # an existing directory, even empty, is refused
    return out


def _positive_int(text: str) -> int:
    value = int(text)
    if value <= 0:
        raise argparse.ArgumentTypeError("must be positive")
    return value


def build_parser() -> argparse.ArgumentParser:
    p = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = p.add_subparsers(dest="command", required=True)
    one = sub.add_parser("inspect", help="Validate and plot one extracted CDF record")
    one.add_argument("record")
    one.add_argument("--label", default="Phase CDF")
    two = sub.add_parser("compare", help="Integrate differences on a common refinement")
    two.add_argument("a")
    two.add_argument("b")
    two.add_argument("--allow-orientation-plan-difference", action="store_true",
                     help="Acknowledge different/missing orientation hashes")
    two.add_argument("--allow-physical-differences", action="store_true")
    two.add_argument("--max-refined-cells", type=_positive_int, default=50_000_000)
    for q in (one, two):
        q.add_argument("--out", required=True, help="New output directory; existing paths are refused")
        q.add_argument("--no-plots", action="store_true")
        q.add_argument("--max-cells", type=_positive_int, default=10_000_000)
    fig = sub.add_parser("plot", help="log10 PDF heat map of one record (plot_phase_cdf.py)")
    fig.add_argument("record")
    fig.add_argument("--out", required=True, type=Path, help="output PNG outside the record directory")
    fig.add_argument("--theta-range", nargs=2, type=float, default=(0.0, 180.0), metavar=("LOW", "HIGH"))
    fig.add_argument("--phi-range", nargs=2, type=float, default=(-180.0, 180.0), metavar=("LOW", "HIGH"))
    color = fig.add_mutually_exclusive_group()
    color.add_argument("--log-limits", nargs=2, type=float, metavar=("LOW", "HIGH"))
    color.add_argument("--decades", type=float)
    fig.add_argument("--dpi", type=int, default=180)
    fig.add_argument("--cell-image", type=Path)
    fig.add_argument("--report", type=Path)
    fig.add_argument("--overwrite", action="store_true")
    fig.add_argument("--max-cells", type=_positive_int, default=10_000_000)
    return p


def _run_inspect(args) -> Path:
    d = SavedDistribution.open(args.record, args.max_cells)
    report = inspect_report(d)
    out = _fresh_directory(args.out)
    _write_json(out / "analysis.json", report)
    _write_json(out / "theta_profile.json", theta_profile(d))
    if not args.no_plots:
        record_figures(d, out, args.label)
    print(f"PASS saved-CDF validation: mass={exact_sum(d.cells):.17g}, g={d.mean_cosine:.17g}")
    for text in report["warnings"]:
        print("NOTE:", text)
    return out


def _run_compare(args) -> Path:
    a = SavedDistribution.open(args.a, args.max_cells)
    b = SavedDistribution.open(args.b, args.max_cells)
    differ = physical_differences(a, b)
    if differ and not args.allow_physical_differences:
        raise ValueError("Physical metadata differ: " + ", ".join(differ) +
                         "; explicit --allow-physical-differences required")
    plan_a, plan_b = lookup(a.metadata, "orientation.csv_sha256"), lookup(b.metadata, "orientation.csv_sha256")
    same_plan = plan_a is not None and plan_a == plan_b
    if not same_plan and not args.allow_orientation_plan_difference:
        raise ValueError("Orientation CSV hashes differ or are absent; explicit "
                         "--allow-orientation-plan-difference required")
    report = compare_report(a, b, args.max_refined_cells)
    report["physical_metadata_differences"] = differ
    report["orientation_csv_sha256_equal"] = same_plan
    report["orientation_plan_difference_acknowledged"] = args.allow_orientation_plan_difference
    out = _fresh_directory(args.out)
    _write_json(out / "comparison.json", report)
    if not args.no_plots:
        comparison_figures(a, b, out)
    print(f"PASS saved-density comparison: TV={report['tv']:.17g}, delta_g={report['g_a_minus_b']:.17g}")
    return out


def _run_plot(args) -> int:
    record = Path(args.record).resolve()
    targets = [p for p in (args.out, args.cell_image, args.report) if p is not None]
    if len({Path(p).resolve() for p in targets}) != len(targets):
        raise ValueError("Output paths must be distinct")
    for p in targets:
        p = Path(p).resolve()
        if record == p.parent or record in p.parents:
            raise ValueError("Write debug outputs outside the dataset-record directory")
        if p.exists() and not args.overwrite:
            raise FileExistsError(f"Output exists: {p}; specify --overwrite to replace it")
    if args.out.suffix.lower() != ".png" or (args.cell_image is not None and args.cell_image.suffix.lower() != ".png"):
        raise ValueError("Output must have .png suffix")
    if args.report is not None and args.report.suffix.lower() != ".json":
        raise ValueError("Output must have .json suffix")
    d = SavedDistribution.open(record, args.max_cells)
    summary = heat_map(d, args.out, tuple(args.theta_range), tuple(args.phi_range), args.log_limits, args.decades,
                       args.dpi, args.cell_image)
    if args.report is not None:
        args.report.write_text(json.dumps(summary, indent=2, allow_nan=False) + "\n", encoding="utf-8")
    print(f"[cdf-plot] mass={summary['cdf_reconstructed_mass']:.17g}, g={summary['g_reconstructed']:.17g}, "
          f"log10 PDF range=[{summary['global_log10_pdf_min_positive']:.8g}, {summary['global_log10_pdf_max']:.8g}]")
    return 0


def main(argv: list[str] | None = None) -> int:
    args = build_parser().parse_args(argv)
    if args.command == "plot":
        try:
            return _run_plot(args)
        except (OSError, ValueError, ImportError) as error:
            print(f"plot_phase_cdf: {type(error).__name__}: {error}", file=sys.stderr)
            return 1
    try:
        out = _run_inspect(args) if args.command == "inspect" else _run_compare(args)
        _write_json(out / "analysis_complete.json", {
            "complete": True, "analysis_version": VERSION, "tool_sha256": file_digest(Path(__file__)),
            "python_version": sys.version, "numpy_version": np.__version__, "input_modified": False,
            "plots_created": not args.no_plots})
        return 0
    except (OSError, ValueError, MemoryError, ImportError) as error:
        print("ERROR: " + ascii(str(error)), file=sys.stderr)
        return 2


if __name__ == "__main__":
    raise SystemExit(main())
