"""Record inspection, comparison and plots (host tooling), checked against
closed forms on small synthetic records written to disk:

* the isotropic distribution: density 1/(4 pi) on every cell of any u grid,
  g = 0, forward-cone probability sin^2(a/2), hemisphere 1/2;
* a forward cap (all probability in the first band): g = 1 - u_1, cone
  probability linear in u inside the band;
* comparisons: identity (every distance 0), disjoint support (TV 1, Hellinger 1,
  Jensen-Shannon ln 2), mass moved between two cells (TV = the moved mass),
  two-cell distributions against their closed-form Hellinger, L2 and
  Jensen-Shannon values (the last also for differences of 1e-6, where the
  integrand cancels), and non-nested grids of one density (distance 0 on a
  refinement of the expected shape);
* the command line: report files, plots, refusals and the heat map."""
from __future__ import annotations

import json
import math
from decimal import Decimal, getcontext
from pathlib import Path

import numpy as np
import pytest

from ice_crystal_phase_amd import analysis as A

FOUR_PI = 4 * math.pi


def u_grid(n: int) -> np.ndarray:
    """The records' u edges: sin^2(pi i / 2n)."""
    return np.sin(np.pi * np.arange(n + 1) / (2 * n)) ** 2


def saved_record(root: Path, name: str, probs, u=None, **meta) -> Path:
    """Write a record directory whose cells carry `probs` [nphi, ntheta]
    (summing to 1): the marginal phi CDF, the conditional theta CDFs (an empty
    azimuth cell takes the u edges), u edges, and the contract metadata."""
    p = np.asarray(probs, dtype=np.float64)
    nphi, nt = p.shape
    u = u_grid(nt) if u is None else np.asarray(u, dtype=np.float64)
    rows = p.sum(axis=1)
    phi_cdf = np.concatenate([[0.0], np.cumsum(rows)])
    phi_cdf[-1] = 1.0
    theta_cdf = np.empty((nphi, nt + 1))
    for j in range(nphi):
        theta_cdf[j] = np.concatenate([[0.0], np.cumsum(p[j]) / rows[j]]) if rows[j] > 0 else u
        theta_cdf[j, -1] = 1.0
    saved = np.diff(phi_cdf)[:, None] * np.diff(theta_cdf, axis=1)
    g = float((saved * (1 - u[:-1] - u[1:])).sum() / saved.sum())
    d = root / name
    d.mkdir()
    np.save(d / "phi_cdf.npy", phi_cdf)
    np.save(d / "theta_given_phi_cdf.npy", theta_cdf)
    np.save(d / "u_edges.npy", u)
    m = {"schema": "rainbow.phase_cdf.numpy.v2", "complete": True, "dtype": "<f8", "order": "C",
         "theta_count": nt, "phi_count": nphi, "density_measure": "solid_angle_sr",
         "cell_model": "constant_density_per_spherical_cell", "coordinates": "u=(1-cos(theta))/2; v=(phi+pi)/(2*pi)",
         "cdf_axis_order": ["phi_cell", "theta_edge"], "normalization": "integral_p_domega_equals_one",
         "input_polarization": "unpolarized", "sampling_frame_columns": np.eye(3).tolist(),
         "coordinate_contract_id": "rainbow.phase_cdf.coordinates.v1", "sampling_frame_map": "identity",
         "direction_convention": "physical_propagation", "theta_zero": "forward", "theta_pi": "backward",
         "orientation": {"rows": 1, "csv_sha256": "same-plan"}, "wavelength_nm": 550.0,
         "hg": {"g": g, "method": "first_moment_of_saved_cell_pdf", "target": "saved_cdf",
                "cosine_convention": "dot(incident_propagation,outgoing_propagation)"}}
    m.update(meta)
    (d / "metadata.json").write_text(json.dumps(m), encoding="utf-8")
    return d


def isotropic(nphi: int, u: np.ndarray) -> np.ndarray:
    return np.tile(np.diff(u) / nphi, (nphi, 1))


# ------------------------------------------------------------------ one record
def test_isotropic_distribution(tmp_path):
    u = np.array([0.0, 0.02, 0.25, 0.5, 0.8, 1.0])
    d = A.SavedDistribution.open(saved_record(tmp_path, "iso", isotropic(6, u), u=u))
    np.testing.assert_allclose(d.density, 1 / FOUR_PI, rtol=1e-14)
    assert abs(d.mean_cosine) < 1e-15
    for a in (0.0, 0.5, 17.0, 60.0, 90.0, 133.0, 180.0):
        assert abs(d.cone_probability(a) - math.sin(math.radians(a) / 2) ** 2) < 1e-15, a
    r = A.inspect_report(d)["computed_from_saved_arrays"]
    assert abs(r["forward_hemisphere_mass"] - 0.5) < 1e-15
    assert abs(r["positive_support_solid_angle_sr"] - FOUR_PI) < 1e-13
    assert r["zero_cells"] == 0 and r["shape_phi_theta"] == [6, 5]
    assert abs(r["pdf_max_sr_inverse"] / r["pdf_min_positive_sr_inverse"] - 1) < 1e-14


def test_forward_cap(tmp_path):
    nt, nphi = 12, 8
    u = u_grid(nt)
    p = np.zeros((nphi, nt))
    p[:, 0] = 1 / nphi
    d = A.SavedDistribution.open(saved_record(tmp_path, "cap", p))
    assert d.mean_cosine == pytest.approx(1 - u[1], abs=1e-15)
    edge = math.degrees(A.polar_angle(np.array([u[1]]))[0])
    half = math.degrees(2 * math.asin(math.sqrt(u[1] / 2)))       # the angle where u = u_1 / 2
    assert d.cone_probability(half) == pytest.approx(0.5, abs=1e-13)
    assert d.cone_probability(edge) == pytest.approx(1.0, abs=1e-15)
    assert d.cone_probability(90) == 1.0
    r = A.inspect_report(d)["computed_from_saved_arrays"]
    assert r["first_polar_band_mass"] == pytest.approx(1.0, abs=1e-15) and r["last_polar_band_mass"] == 0
    assert r["zero_cell_fraction"] == pytest.approx((nt - 1) / nt)
    assert r["highest_mass_theta_bands"][0]["lower_deg"] == 0.0


# ------------------------------------------------------------------ comparisons
def _compare(tmp_path, pa, pb, ua=None, ub=None):
    a = A.SavedDistribution.open(saved_record(tmp_path, "a", pa, u=ua))
    b = A.SavedDistribution.open(saved_record(tmp_path, "b", pb, u=ub))
    return A.compare_report(a, b)


def test_identity_has_zero_distance(tmp_path):
    rng = np.random.default_rng(3)
    p = rng.random((10, 9)) ** 3
    p /= p.sum()
    r = _compare(tmp_path, p, p)
    for k in ("l1", "tv", "l2_per_sqrt_sr", "hellinger", "jensen_shannon_nats", "theta_marginal_tv",
              "mass_a_where_b_is_zero", "mass_b_where_a_is_zero", "g_a_minus_b"):
        assert r[k] == 0, k
    assert r["probability_mass_a_integrated"] == pytest.approx(1, abs=1e-15)


def test_disjoint_support(tmp_path):
    pa, pb = np.zeros((4, 6)), np.zeros((4, 6))
    pa[:, 0] = 0.25
    pb[:, -1] = 0.25
    r = _compare(tmp_path, pa, pb)
    assert r["tv"] == pytest.approx(1, abs=1e-15) and r["hellinger"] == pytest.approx(1, abs=1e-15)
    assert r["jensen_shannon_nats"] == pytest.approx(math.log(2), abs=1e-15)
    assert r["mass_a_where_b_is_zero"] == pytest.approx(1) and r["mass_b_where_a_is_zero"] == pytest.approx(1)
    assert r["theta_marginal_tv"] == pytest.approx(1, abs=1e-15)


def test_moved_mass_is_the_total_variation(tmp_path):
    u = u_grid(7)
    pa = isotropic(5, u)
    pb = pa.copy()
    delta = 0.25 * pa[2, 3]
    pb[2, 3] -= delta
    pb[4, 0] += delta
    r = _compare(tmp_path, pa, pb)
    assert r["tv"] == pytest.approx(delta, rel=1e-12)
    assert r["common_refinement_shape_phi_theta"] == [5, 7]


def _two_cell_closed_forms(a0, b0):
    """Hellinger, L2 (per sqrt sr) and Jensen-Shannon of (a0, 1-a0) against
    (b0, 1-b0) on two cells of solid angle 2 pi each, in 50-digit decimals."""
    getcontext().prec = 50
    a = [Decimal(a0), 1 - Decimal(a0)]
    b = [Decimal(b0), 1 - Decimal(b0)]
    h2 = sum((x.sqrt() - y.sqrt()) ** 2 for x, y in zip(a, b)) / 2
    omega = Decimal(2) * Decimal(math.pi)
    l2 = sum((x - y) ** 2 / omega for x, y in zip(a, b))      # integral of (p - q)^2 dOmega
    js = Decimal(0)
    for x, y in zip(a, b):
        m = (x + y) / 2
        js += (x * (x / m).ln() + y * (y / m).ln()) / 2
    return float(h2.sqrt()), float(l2.sqrt()), float(js)


@pytest.mark.parametrize("a0,b0", [(0.3, 0.6), (0.5 + 1e-6, 0.5), (0.999, 0.001)])
def test_two_cell_distances(tmp_path, a0, b0):
    u = np.array([0.0, 0.5, 1.0])     # two bands of 2 pi sr each
    r = _compare(tmp_path, [[a0, 1 - a0]], [[b0, 1 - b0]], ua=u, ub=u)
    h, l2, js = _two_cell_closed_forms(a0, b0)
    assert r["tv"] == pytest.approx(abs(a0 - b0), rel=1e-13)
    assert r["hellinger"] == pytest.approx(h, rel=1e-12)
    assert r["l2_per_sqrt_sr"] == pytest.approx(l2, rel=1e-12)
    assert r["jensen_shannon_nats"] == pytest.approx(js, rel=1e-9)


def test_non_nested_grids_of_one_density(tmp_path):
    ua, ub = u_grid(3), u_grid(5)
    r = _compare(tmp_path, isotropic(4, ua), isotropic(6, ub), ua=ua, ub=ub)
    # azimuth: multiples of 3 and of 2 in twelfths -> 8 cells; u: 4 + 6 edges sharing 0 and 1 -> 7 bands
    assert r["common_refinement_shape_phi_theta"] == [8, 7]
    assert r["tv"] < 1e-15 and r["hellinger"] < 1e-7 and r["jensen_shannon_nats"] < 1e-28


def test_refinement_limit_and_frames(tmp_path):
    a = A.SavedDistribution.open(saved_record(tmp_path, "a", isotropic(4, u_grid(3))))
    b = A.SavedDistribution.open(saved_record(tmp_path, "b", isotropic(6, u_grid(5))))
    with pytest.raises(ValueError, match="common-refinement cells exceeds limit"):
        A.compare_report(a, b, max_refined_cells=55)
    b.frame = b.frame @ np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1.0]])
    with pytest.raises(ValueError, match="Sampling frames differ"):
        A.compare_report(a, b)


# ------------------------------------------------------------------ command line
def test_inspect_writes_reports_and_plots(tmp_path):
    rec = saved_record(tmp_path, "rec", isotropic(8, u_grid(6)))
    out = tmp_path / "report"
    assert A.main(["inspect", str(rec), "--out", str(out), "--label", "iso"]) == 0
    report = json.loads((out / "analysis.json").read_text())
    assert report["reader_validation"] == "passed" and set(report["sha256"]) == set(A.RECORD_FILES)
    assert any("Angular convergence" in w for w in report["warnings"])
    profile = json.loads((out / "theta_profile.json").read_text())
    assert len(profile["theta_edges_deg"]) == 7 and profile["theta_edges_deg"][-1] == 180.0
    done = json.loads((out / "analysis_complete.json").read_text())
    assert done["complete"] and done["plots_created"] and not done["input_modified"]
    for name in ("pdf_linear.png", "pdf_log.png", "azimuth_average.png", "theta_marginal.png", "cone_cdf.png"):
        assert (out / name).read_bytes()[:8] == b"\x89PNG\r\n\x1a\n", name
    # an existing output directory is refused, even for the same input
    assert A.main(["inspect", str(rec), "--out", str(out), "--no-plots"]) == 2


def test_compare_refusals(tmp_path, capsys):
    u = u_grid(4)
    a = saved_record(tmp_path, "a", isotropic(4, u))
    b = saved_record(tmp_path, "b", isotropic(4, u), wavelength_nm=650.0)
    c = saved_record(tmp_path, "c", isotropic(4, u), orientation={"rows": 1})
    assert A.main(["compare", str(a), str(b), "--out", str(tmp_path / "o1"), "--no-plots"]) == 2
    assert "Physical metadata differ: wavelength_nm" in capsys.readouterr().err
    assert A.main(["compare", str(a), str(b), "--out", str(tmp_path / "o2"), "--no-plots",
                   "--allow-physical-differences"]) == 0
    assert A.main(["compare", str(a), str(c), "--out", str(tmp_path / "o3"), "--no-plots"]) == 2
    assert "Orientation CSV hashes differ or are absent" in capsys.readouterr().err
    assert A.main(["compare", str(a), str(c), "--out", str(tmp_path / "o4"),
                   "--allow-orientation-plan-difference"]) == 0
    assert (tmp_path / "o4" / "comparison_azimuth_average.png").exists()
    r = json.loads((tmp_path / "o4" / "comparison.json").read_text())
    assert r["orientation_csv_sha256_equal"] is False and r["orientation_plan_difference_acknowledged"] is True


def test_strict_metadata_and_cell_limit(tmp_path):
    rec = saved_record(tmp_path, "rec", isotropic(4, u_grid(4)))
    with pytest.raises(ValueError, match="exceeds --max-cells=15"):
        A.SavedDistribution.open(rec, max_cells=15)
    text = (rec / "metadata.json").read_text()
    (rec / "metadata.json").write_text(text[:-1] + ', "wavelength_nm": 600}')
    with pytest.raises(ValueError, match="Duplicate metadata key: wavelength_nm"):
        A.SavedDistribution.open(rec)
    (rec / "metadata.json").write_text(text[:-1] + ', "extra": NaN}')
    with pytest.raises(ValueError, match="Nonfinite metadata number"):
        A.SavedDistribution.open(rec)


def test_plot_heat_map(tmp_path, capsys):
    p = isotropic(12, u_grid(9))
    p[:, 0] *= 40
    p /= p.sum()
    rec = saved_record(tmp_path, "rec", p)
    fig, rep = tmp_path / "map.png", tmp_path / "map.json"
    assert A.main(["plot", str(rec), "--out", str(fig), "--report", str(rep), "--decades", "3",
                   "--theta-range", "0", "90", "--cell-image", str(tmp_path / "cells.png")]) == 0
    assert fig.read_bytes()[:4] == b"\x89PNG" and (tmp_path / "cells.png").exists()
    s = json.loads(rep.read_text())
    assert s["cdf_reconstructed_mass"] == pytest.approx(1, abs=1e-15)
    assert s["color_limits_log10"][1] - s["color_limits_log10"][0] == pytest.approx(3)
    assert s["cells_shown_theta_phi"] == [5, 12]     # 9 bands of 20 deg: 0-90 deg spans 5 (the fifth straddles 90)
    # outputs inside the record, or existing ones without --overwrite, are refused
    assert A.main(["plot", str(rec), "--out", str(rec / "x.png")]) == 1
    assert A.main(["plot", str(rec), "--out", str(fig)]) == 1
    assert "Output exists" in capsys.readouterr().err
    assert A.main(["plot", str(rec), "--out", str(fig), "--overwrite"]) == 0
