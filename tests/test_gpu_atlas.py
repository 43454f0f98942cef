"""GPU tests of the deterministic atlas (iph_atlas_trace): every node's cell
masses against the oracle's restatement of the reference's CPU atlas
(compensated per-ray columns, oracle_hex.c orc_atlas_node) and, where it was
built, against the reference's own CPU atlas tool; node audits; invariance to
how the nodes are split over calls; the command-line tool end to end."""
from __future__ import annotations

import json
import math
import subprocess
from pathlib import Path

import numpy as np
import pytest

import ice_crystal_phase_amd as ip
from ice_crystal_phase_amd import icephase

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
CLI = ROOT / "ice_crystal_phase_amd" / "bin" / "ice_precompute_atlas"
REF_CLI = ROOT / "oracle" / "_ref" / "ice_precompute_atlas_cpu"
# fixed-point quantization (2^-Q of a node's largest ray weight per emission)
# against compensated FP64 sums: relative L1 per node
L1_TOL = 1e-12


def _oracle_settings(oracle, s):
    return oracle.AtlasSettings(s.radius_mm, s.length_mm, s.interior_index, s.exterior_index, s.tail_tolerance,
                                s.incident_theta_intervals, s.incident_phi_intervals, s.area_subdivisions,
                                s.theta_count, s.phi_count, s.max_internal_hits)


@pytest.mark.parametrize("geom", [(0.1, 0.2, 1.31, 3, 2, 12, 24, 2), (0.1, 0.7, 1.31, 4, 3, 36, 72, 3),
                                  (0.05, 0.02, 1.33, 2, 2, 18, 36, 2)], ids=["small", "column", "plate"])
def test_atlas_nodes_equal_oracle(ctx, oracle, geom):
    r, length, n, T, P, nt, np_, S = geom
    s = icephase.atlas_settings(r, length, n, T, P, nt, np_, S)
    print("atlas geometry", geom, flush=True)
    mass, nodes = icephase.atlas_trace(ctx, s)
    os_ = _oracle_settings(oracle, s)
    for q, node in enumerate(nodes):
        want, rep = oracle.atlas_node(os_, node.i, node.j)
        got = mass[q]
        assert np.abs(got - want).sum() / want.sum() < L1_TOL, (node.i, node.j)
        assert list(node.ki) == list(rep.ki) and list(node.e0) == list(rep.e0)
        assert node.rays == rep.rays and node.outputs == rep.outputs
        assert node.projected_area_mm2 == rep.projected_area_mm2
        assert abs(node.escaped_mm2 / rep.escaped_mm2 - 1) < 1e-13
        assert abs(node.unresolved_mm2 - rep.unresolved_mm2) <= 1e-13 * rep.projected_area_mm2
        assert node.histogram_relative_error < 1e-13
        assert node.max_interface_balance_error <= 2048 * 2.3e-16
        assert abs(node.max_balance_error - rep.max_balance_error) < 1e-13


def test_index_matched_delta(ctx):
    """n = 1: the whole projected area in the forward end cap, area/np per cell.
    At phi = 30 deg (j = 2) straight rays from the rule's points leave through
    prism edges: the reference's CPU tool rejects those nodes as ambiguous, and
    so does the device path."""
    s = icephase.atlas_settings(0.1, 0.2, 1.0, 3, 2, 12, 24, 2)
    for i in range(4):
        mass, nodes = icephase.atlas_trace(ctx, s, 3 * i, 2)
        for q, node in enumerate(nodes):
            want = np.zeros((12, 24))
            want[0, :] = node.projected_area_mm2 / 24
            assert np.max(np.abs(mass[q] - want)) < 1e-15, (node.i, node.j)
    with pytest.raises(ip.AuditError, match="status=ambiguous_boundary"):
        icephase.atlas_trace(ctx, s, 5, 1)


def test_split_over_calls_is_bitwise_identical(ctx):
    s = icephase.atlas_settings(0.1, 0.3, 1.31, 5, 3, 24, 48, 2)
    whole, _ = icephase.atlas_trace(ctx, s)
    parts = [icephase.atlas_trace(ctx, s, first, count)[0] for first, count in ((0, 7), (7, 1), (8, 16))]
    assert np.array_equal(whole, np.concatenate(parts))
    again, _ = icephase.atlas_trace(ctx, s)
    assert np.array_equal(whole, again)


def test_interaction_limit_is_rejected(ctx):
    s = icephase.atlas_settings(0.1, 0.2, 1.31, 3, 2, 12, 24, 2, max_internal_hits=1)
    with pytest.raises(ip.AuditError, match="Deterministic area ray rejected"):
        icephase.atlas_trace(ctx, s)


@pytest.mark.skipif(not REF_CLI.exists(), reason="reference atlas tool not built (needs the reference checkout)")
def test_atlas_equals_reference_tool(ctx, tmp_path):
    out = tmp_path / "ref"
    args = [str(REF_CLI), "--out", str(out), "--radius-mm", ".1", "--length-mm", ".4", "--wavelength-nm", "650",
            "--incident-theta-intervals", "4", "--incident-phi-intervals", "2", "--theta", "18", "--phi", "36",
            "--area-subdivisions", "2"]
    print("run:", " ".join(args), flush=True)
    res = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stderr
    ref = np.load(out / "mass_mm2.npy", allow_pickle=False)
    meta = json.loads((out / "metadata.json").read_text())
    s = icephase.atlas_settings(0.1, 0.4, icephase.ice_index(650).n, 4, 2, 18, 36, 2, wavelength_nm=650)
    mass, nodes = icephase.atlas_trace(ctx, s)
    mass = mass.reshape(ref.shape)
    for node, m in zip(nodes, meta["nodes"]):
        want = ref[node.i, node.j]
        assert np.abs(mass[node.i, node.j] - want).sum() / want.sum() < L1_TOL
        assert list(node.ki) == m["ki"] and node.rays == m["rays"] and node.outputs == m["outputs"]


def _run(args, cwd):
    print("run:", " ".join(map(str, args)), flush=True)
    return subprocess.run([str(a) for a in args], capture_output=True, text=True, timeout=300, cwd=cwd)


BASE = ["--radius-mm", ".1", "--length-mm", ".2", "--material", "constant", "--ior", "1.31",
        "--incident-theta-intervals", "3", "--incident-phi-intervals", "2", "--theta", "12", "--phi", "24",
        "--area-subdivisions", "2"]


@pytest.mark.skipif(not CLI.exists(), reason="bin/ice_precompute_atlas not built")
def test_cli_end_to_end(tmp_path, oracle):
    res = _run([CLI, *BASE, "--wavelength-nm", "550", "--out", "atlas550"], tmp_path)
    assert res.returncode == 0, res.stderr
    assert "atlas_complete nodes=12" in res.stdout
    d = tmp_path / "atlas550"
    assert sorted(p.name for p in d.iterdir()) == ["mass_mm2.npy", "metadata.json", "u_edges.npy"]
    mass = np.load(d / "mass_mm2.npy", allow_pickle=False)
    assert mass.shape == (4, 3, 12, 24) and mass.dtype == np.float64
    assert np.array_equal(np.load(d / "u_edges.npy", allow_pickle=False), ip.u_edges(12))
    meta = json.loads((d / "metadata.json").read_text())
    assert meta["schema"] == "ice.single_particle_atlas.v1" and meta["complete"] is True
    assert meta["mass_axes"] == ["incident_theta", "incident_phi", "theta", "phi"]
    assert meta["area_quadrature"] == {"rule": "composite_equal_area_GL2xGL3", "subdivisions": 2,
                                       "randomness": False}
    assert len(meta["nodes"]) == 12 and meta["total_rays"] == sum(n["rays"] for n in meta["nodes"])
    s = icephase.atlas_settings(0.1, 0.2, 1.31, 3, 2, 12, 24, 2)
    os_ = _oracle_settings(oracle, s)
    for n in meta["nodes"]:
        want, _ = oracle.atlas_node(os_, n["i"], n["j"])
        assert np.abs(mass[n["i"], n["j"]] - want).sum() / want.sum() < L1_TOL
    if REF_CLI.exists():
        ref = _run([REF_CLI, *BASE, "--wavelength-nm", "550", "--out", "ref550"], tmp_path)
        assert ref.returncode == 0, ref.stderr
        ref_meta = json.loads((tmp_path / "ref550" / "metadata.json").read_text())
        assert set(ref_meta) == set(meta)
        assert set(ref_meta["nodes"][0]) == set(meta["nodes"][0])
        assert set(ref_meta["material"]) == set(meta["material"])
    # an existing atlas is never replaced; a failing run keeps its .part
    again = _run([CLI, *BASE, "--wavelength-nm", "550", "--out", "atlas550"], tmp_path)
    assert again.returncode == 2 and "already exists" in again.stderr
    fail = _run([CLI, *BASE, "--wavelength-nm", "550", "--out", "fail", "--max-internal-hits", "1"], tmp_path)
    assert fail.returncode == 2 and "Incomplete .part retained" in fail.stderr
    assert not (tmp_path / "fail").exists() and (tmp_path / "fail.part").exists()


@pytest.mark.skipif(not CLI.exists(), reason="bin/ice_precompute_atlas not built")
def test_cli_ice_material(tmp_path):
    res = _run([CLI, "--radius-mm", ".05", "--length-mm", ".1", "--wavelength-nm", "555.5",
                "--incident-theta-intervals", "2", "--incident-phi-intervals", "1", "--theta", "9", "--phi", "18",
                "--area-subdivisions", "1", "--out", "ice"], tmp_path)
    assert res.returncode == 0, res.stderr
    meta = json.loads((tmp_path / "ice" / "metadata.json").read_text())
    st = icephase.ice_index(555.5)
    assert meta["material"]["model"] == "ice_ih_warren_brandt_2008_real_only"
    assert meta["material"]["interior_index"] == st.n and meta["material"]["bracket_nm"] == [550, 560]
    nodes = meta["nodes"]
    assert all(abs((n["escaped_mm2"] + n["unresolved_mm2"]) / n["projected_area_mm2"] - 1) < 1e-10 for n in nodes)
    assert math.isclose(sum(n["projected_area_mm2"] for n in nodes[:1]),
                        icephase.hex_projected_area(0.05, 0.1, (0.0, 1.0, 0.0)), rel_tol=1e-15)
