"""GPU tests of the atlas composition: the device shifted-block sum
(iph_atlas_compose) bitwise against oracle/compose_ref.py, the device CDF of a
composed mass against the oracle's CDF restatement, the reference's CDF test
cases (hex_atlas/tests/test_compose.py:100-116), and the composer end to end
(after hex_atlas/tests/cli_test.py: two atlases from ice_precompute_atlas,
three inclinations, a Gaussian + isotropic orientation mixture), every record
re-derived by the oracles from the atlas files and the coupling weights."""
from __future__ import annotations

import json
import math
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

import ice_crystal_phase_amd as ip
from ice_crystal_phase_amd import compose as cm
from ice_crystal_phase_amd.record import PhaseRecord

sys.path.insert(0, str(Path(__file__).resolve().parent.parent / "oracle"))
from compose_ref import compose_mass_ref  # noqa: E402

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
ATLAS_CLI = ROOT / "ice_crystal_phase_amd" / "bin" / "ice_precompute_atlas"


def _fixture_mass(seed=3, it=6, ip_=3, nt=9, np_=24):
    rng = np.random.default_rng(seed)
    return rng.random((it + 1, ip_ + 1, nt, np_)) * np.exp(rng.normal(0, 3, (it + 1, ip_ + 1, nt, np_)))


def test_device_sum_equals_oracle_bitwise(ctx):
    mass = _fixture_mass()
    rng = np.random.default_rng(7)
    w = rng.random((7, 4, 2, 24))
    w[w < 0.8] = 0.0  # sparse, as real couplings are
    w[0, 0, 1] = rng.random(24)  # one dense group
    print("compose: sparse weights", int((w > 0).sum()), flush=True)
    with ip.AtlasMass(ctx, mass) as dev:
        got = dev.compose(w)
        again = dev.compose(w)
    assert np.array_equal(got, compose_mass_ref(mass, w))
    assert np.array_equal(got, again)


def test_device_sum_wide_azimuth(ctx):
    mass = _fixture_mass(seed=5, it=2, ip_=1, nt=5, np_=1800)
    w = np.zeros((3, 2, 2, 1800))
    w[1, 1, 0, [0, 1, 899, 1799]] = [0.5, 0.25, 1e-3, 2.0]
    w[2, 0, 1, 7] = 3.0
    with ip.AtlasMass(ctx, mass) as dev:
        assert np.array_equal(dev.compose(w), compose_mass_ref(mass, w))


def test_device_sum_rejects_bad_weights(ctx):
    mass = _fixture_mass()
    with ip.AtlasMass(ctx, mass) as dev:
        with pytest.raises(ip.IcePhaseError, match="all zero"):
            dev.compose(np.zeros((7, 4, 2, 24)))
        bad = np.zeros((7, 4, 2, 24))
        bad[0, 0, 0, 0] = -1
        with pytest.raises(ip.IcePhaseError, match="nonnegative"):
            dev.compose(bad)


def test_cdf_uniform_and_zero_rows(ctx):
    edges = (1 - np.cos(np.linspace(0, math.pi, 9))) / 2
    for mass in (np.repeat(np.diff(edges)[:, None] / 12, 12, axis=1), np.eye(8, 12)):
        a, b, r = cm.make_cdf(ctx, mass, edges)
        recovered = np.diff(a)[:, None] * np.diff(b, axis=1)
        np.testing.assert_allclose(recovered.T, mass / mass.sum(), atol=2e-15)
        assert abs(r["cdf_mass"] - 1) < 1e-15
        zero = np.diff(a) == 0
        if zero.any():
            np.testing.assert_array_equal(b[zero], np.repeat(edges[None, :], zero.sum(), axis=0))


def test_cdf_dynamic_range_and_rejection(ctx):
    edges = np.linspace(0, 1, 9)
    mass = np.geomspace(1e-14, 1, 96).reshape(8, 12)
    _, _, r = cm.make_cdf(ctx, mass, edges)
    assert r["cdf_l1_mass_error"] < 1e-12
    with pytest.raises(ValueError):
        cm.make_cdf(ctx, -mass, edges)
    with pytest.raises(ValueError):
        cm.make_cdf(ctx, np.zeros_like(mass), edges)


def test_cdf_equals_oracle_cdf(ctx, oracle):
    mass = _fixture_mass()[2, 1]
    edges = ip.u_edges(9)
    a, b, _ = cm.make_cdf(ctx, mass, edges)
    rc, phi, theta, _ = oracle.build_cdf(mass, 9, 24)
    assert rc == 0 and np.array_equal(a, phi) and np.array_equal(b, theta)


def _run(args, cwd):
    print("run:", " ".join(map(str, args)), flush=True)
    return subprocess.run([str(a) for a in args], capture_output=True, text=True, timeout=300, cwd=cwd)


@pytest.mark.skipif(not ATLAS_CLI.exists(), reason="bin/ice_precompute_atlas not built")
def test_compose_end_to_end(tmp_path, oracle):
    base = [ATLAS_CLI, "--radius-mm", ".1", "--length-mm", ".2", "--material", "constant", "--ior", "1.31",
            "--incident-theta-intervals", "3", "--incident-phi-intervals", "2", "--theta", "12", "--phi", "24",
            "--area-subdivisions", "2"]
    paths = []
    for wave in (550, 650):
        res = _run([*base, "--wavelength-nm", wave, "--out", f"atlas{wave}"], tmp_path)
        assert res.returncode == 0, res.stderr
        paths.append(tmp_path / f"atlas{wave}")
    before = {p: (p / "mass_mm2.npy").read_bytes() for p in paths}
    config = {"inclinations_deg": [0, 35, 80], "orientation": {
        "quadrature": {"azimuth": 4, "tilt": 4, "spin": 3},
        "components": [{"kind": "horizontal_gaussian", "sigma_deg": 10, "fraction": .75},
                       {"kind": "isotropic", "fraction": .25}]}}
    cfg = tmp_path / "compose.json"
    cfg.write_text(json.dumps(config))
    cmd = [sys.executable, "-m", "ice_crystal_phase_amd.compose", "--atlas", *paths, "--config", cfg]
    for out in ("dataset", "repeat"):
        res = subprocess.run([str(a) for a in cmd] + ["--out", str(tmp_path / out)], capture_output=True, text=True,
                             timeout=300, cwd=ROOT)
        assert res.returncode == 0, res.stderr
    plan = json.loads((tmp_path / "dataset" / "batch_plan.json").read_text())
    assert plan["complete"] and len(plan["records"]) == 6
    rule_r, rule_w, rule_meta = cm.orientation_rule(config["orientation"])
    atlases = [cm.Atlas(p) for p in paths]
    gs = []
    for rec in plan["records"]:
        path = tmp_path / "dataset" / rec["path"]
        with PhaseRecord(path, validate=True) as r:
            mass_total, g = r.mass_and_g()
            assert abs(mass_total - 1) < 1e-12
            assert r.metadata["orientation"]["randomness"] is False
            assert r.metadata["atlas"]["no_retracing"] is True
            assert abs(r.metadata["transport_audit"]["balance_error"]) < 2e-10
            gs.append(g)
        for name in ("phi_cdf.npy", "theta_given_phi_cdf.npy", "u_edges.npy"):
            assert (path / name).read_bytes() == (tmp_path / "repeat" / rec["path"] / name).read_bytes()
        assert {p.name for p in path.iterdir()} == {"metadata.json", "phi_cdf.npy", "theta_given_phi_cdf.npy",
                                                    "u_edges.npy"}
        # the oracles re-derive the record: coupling weights -> oracle sum -> oracle CDF
        a = atlases[rec["wavelength_index"]]
        ki = np.asarray(rec["ki"])
        w, _, _, _ = cm.coupling_for_rule(atlases[0], rule_r, rule_w, ki, dict(rule_meta, vertical_uniform_azimuth_policy={}))
        mass = compose_mass_ref(np.asarray(a.mass), w)
        rc, phi, theta, _ = oracle.build_cdf(mass, 12, 24)
        assert rc == 0
        assert np.array_equal(np.load(path / "phi_cdf.npy"), phi)
        assert np.array_equal(np.load(path / "theta_given_phi_cdf.npy"), theta)
    for a in atlases:
        a.close()
    assert max(gs) - min(gs) > .01, "the incident condition is lost in composition"
    assert all((p / "mass_mm2.npy").read_bytes() == before[p] for p in paths)
    again = subprocess.run([str(a) for a in cmd] + ["--out", str(tmp_path / "dataset")], capture_output=True,
                           text=True, timeout=300, cwd=ROOT)
    assert again.returncode == 2 and "compose_atlas failed" in again.stderr
