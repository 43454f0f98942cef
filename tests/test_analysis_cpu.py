"""Record inspection and exact-cell comparison (host tooling), after the
reference's tests of the same tool (tests/test_phase_analysis.py:17-176):
closed forms (uniform measure, forward cap, identity, disjoint support,
isolated-cell coarsening, non-nested grids, one-ulp edges, tiny JS), the
reader round trip and the command line's refusals."""
from __future__ import annotations

import hashlib
import json
import math
from pathlib import Path

import numpy as np
import pytest

from ice_crystal_phase_amd.analysis import Table, angular_summary, compare_tables, cone_mass, load_table, main


def table(mass, u, name="synthetic"):
    w = np.array(mass, dtype=float)
    return Table(Path(name), {"coordinate_contract_id": "ice.phase_cdf.ensemble_coordinates.v1",
                              "sampling_frame_map": "local_column_to_ensemble_reference_column",
                              "direction_convention": "physical_propagation", "theta_zero": "forward",
                              "theta_pi": "backward"},
                 np.eye(3), np.array(u, dtype=float), np.arange(w.shape[0] + 1) / w.shape[0], w, {})


def write_record(path: Path, mass, u):
    t = table(mass, u)
    path.mkdir()
    nt, np_ = t.mass.shape[1], t.mass.shape[0]
    q = t.mass.sum(axis=1)
    a = np.r_[0.0, q.cumsum()]
    a[-1] = 1
    c = np.zeros((np_, nt + 1))
    for j in range(np_):
        c[j] = np.r_[0.0, t.mass[j].cumsum() / q[j]] if q[j] else t.u
    c[:, -1] = 1
    np.save(path / "phi_cdf.npy", a)
    np.save(path / "theta_given_phi_cdf.npy", c)
    np.save(path / "u_edges.npy", t.u)
    rec = np.diff(a)[:, None] * np.diff(c, axis=1)
    g = np.sum(rec * (1 - t.u[:-1] - t.u[1:])) / rec.sum()
    m = dict(t.metadata, schema="rainbow.phase_cdf.numpy.v2", complete=True, dtype="<f8", order="C", theta_count=nt,
             phi_count=np_, density_measure="solid_angle_sr", cell_model="constant_density_per_spherical_cell",
             coordinates="u=(1-cos(theta))/2; v=(phi+pi)/(2*pi)", cdf_axis_order=["phi_cell", "theta_edge"],
             normalization="integral_p_domega_equals_one", input_polarization="unpolarized",
             sampling_frame_columns=np.eye(3).tolist(), orientation={"rows": 3, "csv_sha256": "test"},
             hg={"g": float(g), "method": "first_moment_of_saved_cell_pdf", "target": "saved_cdf",
                 "cosine_convention": "dot(incident_propagation,outgoing_propagation)"})
    (path / "metadata.json").write_text(json.dumps(m), encoding="utf-8")


def test_uniform_measure():
    u = np.array([0, 0.03, 0.3, 0.7, 1.0])
    t = table(np.tile(np.diff(u) / 3, (3, 1)), u)
    np.testing.assert_allclose(t.density_omega, 1 / (4 * np.pi), rtol=2e-15)
    assert abs(t.g) < 1e-15
    for angle in (0, 1, 20, 90, 173, 180):
        assert abs(cone_mass(t, angle) - math.sin(math.radians(angle) / 2) ** 2) < 1e-14


def test_forward_cap():
    t = table([[0.5, 0], [0.5, 0]], [0, 0.1, 1])
    assert abs(t.g - 0.9) < 1e-15
    assert abs(cone_mass(t, math.degrees(2 * math.asin(math.sqrt(0.05)))) - 0.5) < 1e-14


def test_identity():
    t = table([[.1, .2], [.3, .4]], [0, .3, 1])
    s = compare_tables(t, t)
    for key in ("tv", "l1", "l2_per_sqrt_sr", "hellinger", "jensen_shannon_nats", "theta_marginal_tv"):
        assert s[key] == 0, key


def test_disjoint_support():
    a, b = table([[1, 0]], [0, .5, 1]), table([[0, 1]], [0, .5, 1])
    s = compare_tables(a, b)
    assert abs(s["tv"] - 1) < 1e-15 and abs(s["hellinger"] - 1) < 1e-15
    assert abs(s["jensen_shannon_nats"] - math.log(2)) < 1e-15
    assert s["mass_a_where_b_is_zero"] == 1


def test_isolated_cell_coarsening():
    a, b = table([[1, 0], [0, 0]], [0, .5, 1]), table([[1]], [0, 1])
    assert abs(compare_tables(a, b)["tv"] - .75) < 1e-15


def test_common_refinement_preserves_information():
    a = table([[.5, 0, .5, 0]], [0, .25, .5, .75, 1])
    b = table([[0, .5, 0, .5]], [0, .25, .5, .75, 1])
    assert abs(compare_tables(a, b)["tv"] - 1) < 1e-15


def test_non_nested_unequal_phi():
    ua, ub = np.array([0, .1, .5, 1]), np.array([0, .2, .6, .8, 1])
    a = table(np.tile(np.diff(ua) / 3, (3, 1)), ua)
    b = table(np.tile(np.diff(ub) / 7, (7, 1)), ub)
    s = compare_tables(a, b)
    assert s["tv"] < 1e-15
    assert s["common_refinement_shape_phi_theta"] == [9, 6]
    assert abs(s["probability_mass_a_integrated"] - 1) < 1e-15 and abs(s["probability_mass_b_integrated"] - 1) < 1e-15


def test_one_ulp_edge_not_merged():
    x = np.nextafter(.5, 1.)
    a, b = table([[1, 0]], [0, .5, 1]), table([[1, 0]], [0, x, 1])
    s = compare_tables(a, b)
    assert s["common_refinement_shape_phi_theta"] == [1, 3]
    assert s["tv"] > 0 and abs(s["tv"] / (x - .5) - 2) < 1e-12


def test_js_tiny_difference_nonnegative():
    a, b = table([[.5, .5]], [0, .5, 1]), table([[.5 + 1e-10, .5 - 1e-10]], [0, .5, 1])
    js = compare_tables(a, b)["jensen_shannon_nats"]
    assert 0 < js < 1e-18
    # the series equals the closed form t^2/2 + ... for the relative difference t = 1e-10
    assert abs(js / (0.5 * 1e-20) - 1) < 1e-6


def test_frame_refusal():
    a, b = table([[.5, .5]], [0, .5, 1]), table([[.5, .5]], [0, .5, 1])
    b.frame = -np.eye(3)
    with pytest.raises(ValueError, match="frames differ"):
        compare_tables(a, b)


def test_refinement_limit_refusal():
    a = table([[.5, .5]], [0, .5, 1])
    with pytest.raises(ValueError, match="exceeds limit"):
        compare_tables(a, a, max_refined_cells=1)


def test_reader_round_trip_and_no_mutation(tmp_path):
    p = tmp_path / "record"
    write_record(p, [[.25, .75], [0, 0]], [0, .2, 1])
    before = {f.name: hashlib.sha256(f.read_bytes()).hexdigest() for f in p.iterdir()}
    s = angular_summary(load_table(p))
    assert s["reader_validation"] == "passed" and s["computed_from_saved_arrays"]["zero_cells"] == 2
    assert before == {f.name: hashlib.sha256(f.read_bytes()).hexdigest() for f in p.iterdir()}


def test_reject_bad_cdf(tmp_path):
    p = tmp_path / "record"
    write_record(p, [[.25, .75], [0, 0]], [0, .2, 1])
    x = np.load(p / "theta_given_phi_cdf.npy")
    x[0, 1] = -1
    np.save(p / "theta_given_phi_cdf.npy", x)
    with pytest.raises(ValueError):
        load_table(p)


def test_memory_limit_before_dense_load(tmp_path):
    p = tmp_path / "record"
    write_record(p, [[.25, .75], [0, 0]], [0, .2, 1])
    with pytest.raises(ValueError, match="max-cells"):
        load_table(p, max_cells=1)


def test_cli_existing_output_refused(tmp_path, capsys):
    p = tmp_path / "record"
    write_record(p, [[.25, .75], [0, 0]], [0, .2, 1])
    out = tmp_path / "report"
    assert main(["inspect", str(p), "--out", str(out), "--no-plots"]) == 0
    assert json.loads((out / "analysis_complete.json").read_text())["input_modified"] is False
    assert main(["inspect", str(p), "--out", str(out), "--no-plots"]) == 2


def test_cli_changed_plan_requires_acknowledgement(tmp_path):
    a, b = tmp_path / "a", tmp_path / "b"
    for p in (a, b):
        write_record(p, [[.25, .75], [0, 0]], [0, .2, 1])
    m = json.loads((b / "metadata.json").read_text())
    m["orientation"]["csv_sha256"] = "changed"
    (b / "metadata.json").write_text(json.dumps(m))
    args = ["compare", str(a), str(b), "--out", str(tmp_path / "diff"), "--no-plots"]
    assert main(args) == 2
    assert main([*args, "--allow-orientation-plan-difference"]) == 0
    assert json.loads((tmp_path / "diff" / "comparison.json").read_text())["tv"] == 0


def test_cli_plots_need_matplotlib(tmp_path):
    p = tmp_path / "record"
    write_record(p, [[.5, .5]], [0, .5, 1])
    assert main(["inspect", str(p), "--out", str(tmp_path / "plots")]) == 2


def test_part_refusal(tmp_path):
    p = tmp_path / "record.part"
    write_record(p, [[.5, .5]], [0, .5, 1])
    with pytest.raises(ValueError, match=r"\.part"):
        load_table(p)


def test_corrupt_hg_label_refusal(tmp_path):
    p = tmp_path / "record"
    write_record(p, [[.5, .5]], [0, .5, 1])
    m = json.loads((p / "metadata.json").read_text())
    m["hg"]["g"] = .5
    (p / "metadata.json").write_text(json.dumps(m))
    with pytest.raises(ValueError, match="HG label"):
        load_table(p)


def test_nonfinite_metadata_refusal(tmp_path):
    p = tmp_path / "record"
    write_record(p, [[.5, .5]], [0, .5, 1])
    m = json.loads((p / "metadata.json").read_text())
    m["wavelength_nm"] = float("inf")
    (p / "metadata.json").write_text(json.dumps(m))
    with pytest.raises(ValueError, match="Nonfinite"):
        load_table(p)
