"""Host side of the atlas composition (no device): D6h folding, the tilt
quadrature, the orientation rule, incident interpolation and the coupling
weights, after the reference's tests (hex_atlas/tests/test_compose.py:26-98,
test_polar_azimuth.py). The device sum and CDF are tested in
test_gpu_compose.py against oracle/compose_ref.py and closed forms."""
from __future__ import annotations

import json
import math
import sys
from pathlib import Path
from types import SimpleNamespace

import numpy as np
import pytest

from ice_crystal_phase_amd import compose as cm

sys.path.insert(0, str(Path(__file__).resolve().parent.parent / "oracle"))
from compose_ref import compose_mass_ref  # noqa: E402


def fixture():
    """Atlas stand-in: node masses = node area x one fixed pattern."""
    it, ip_, nt, np_ = 6, 3, 4, 12
    area = np.zeros((it + 1, ip_ + 1))
    mass = np.zeros((it + 1, ip_ + 1, nt, np_))
    pattern = np.arange(1, nt * np_ + 1, dtype=float).reshape(nt, np_)
    pattern /= pattern.sum()
    for i in range(it + 1):
        for j in range(ip_ + 1):
            area[i, j] = cm.projected_area(.1, .2, cm.canonical_frame(i * math.pi / 2 / it, j * math.pi / 6 / ip_)[:, 2])
            mass[i, j] = area[i, j] * pattern
    return SimpleNamespace(it=it, ip=ip_, nt=nt, np=np_, area=area, mass=mass, residual=np.zeros_like(area),
                           radius=.1, length=.2)


def test_fold_full_sphere():
    for theta in np.linspace(0, math.pi, 21):
        for phi in np.linspace(0, 2 * math.pi, 43):
            k = cm.canonical_frame(theta, phi)[:, 2]
            t, p, q = cm.fold_direction(k)
            assert 0 <= t <= math.pi / 2 and 0 <= p <= math.pi / 6
            np.testing.assert_allclose(q @ cm.canonical_frame(t, p)[:, 2], k, atol=2e-14)
            np.testing.assert_allclose(q.T @ q, np.eye(3), atol=2e-14)


def test_fold_uses_reflections_on_the_mirrored_half():
    # azimuth 40 deg lies past the sector's mirror line (30 deg): an improper operation
    _, p, q = cm.fold_direction(cm.canonical_frame(1.0, math.radians(40))[:, 2])
    assert np.linalg.det(q) < 0 and abs(p - math.radians(20)) < 1e-14
    _, p, q = cm.fold_direction(cm.canonical_frame(1.0, math.radians(70))[:, 2])
    assert np.linalg.det(q) > 0 and abs(p - math.radians(10)) < 1e-14


def test_gaussian_cdf_nodes():
    for sigma in (.1, 3, 10, 90):
        n = 16
        b = cm.gaussian_tilt_nodes(n, sigma)
        s = math.radians(sigma)
        e = math.erf(math.pi / (2 * math.sqrt(2) * s))
        c = np.array([(1 + math.erf((x - math.pi / 2) / (math.sqrt(2) * s)) / e) / 2 for x in b])
        np.testing.assert_allclose(c, (np.arange(n) + .5) / n, atol=6e-14, rtol=0)
        np.testing.assert_allclose(b + b[::-1], math.pi, atol=2e-15)
    np.testing.assert_array_equal(cm.gaussian_tilt_nodes(4, 0), [math.pi / 2] * 4)


def test_rule_no_rng_and_weights():
    config = {"quadrature": {"azimuth": 4, "tilt": 8, "spin": 3}, "components": [
        {"kind": "horizontal_gaussian", "sigma_deg": 3, "fraction": .9}, {"kind": "isotropic", "fraction": .1}]}
    a, w, m = cm.orientation_rule(config)
    b, v, _ = cm.orientation_rule(config)
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(w, v)
    assert m["randomness"] is False
    assert abs(math.fsum(w[:96]) - .9) < 1e-15 and abs(math.fsum(w[96:]) - .1) < 1e-15
    # the first pose: Ry(2 pi/8) Rz(beta_0) Ry(pi/18)
    beta0 = cm.gaussian_tilt_nodes(8, 3)[0]
    np.testing.assert_allclose(a[0], cm.ry(math.pi / 4) @ cm.rz(beta0) @ cm.ry(math.pi / 18), atol=1e-15)


@pytest.mark.parametrize("q", [(1, 1, 1), (3, 2, 5), (72, 12, 24), (5, 9, 31)])
def test_rule_equals_per_pose_products(q):
    """The rule's rotations are the per-pose products of atlas_math.py:143-149
    (azimuth @ Rz(beta), then @ Ry(gamma) for each pose), bit for bit, with the
    weights fraction / (na nb ng) in the same order."""
    na, nb, ng = q
    comps = [{"kind": "horizontal_gaussian", "sigma_deg": 11, "fraction": .7}, {"kind": "isotropic", "fraction": .3}]
    r, w, meta = cm.orientation_rule({"quadrature": {"azimuth": na, "tilt": nb, "spin": ng}, "components": comps})
    mats, weights = [], []
    for comp in comps:
        betas = (cm.gaussian_tilt_nodes(nb, comp["sigma_deg"]) if comp["kind"] == "horizontal_gaussian"
                 else np.arccos(1 - 2 * (np.arange(nb) + 0.5) / nb))
        for ia in range(na):
            azimuth = cm.ry(2 * math.pi * (ia + 0.5) / na)
            for beta in betas:
                tilted = azimuth @ cm.rz(float(beta))
                for ig in range(ng):
                    mats.append(tilted @ cm.ry((math.pi / 3) * (ig + 0.5) / ng))
                    weights.append(comp["fraction"] / (na * nb * ng))
    assert np.array_equal(r, np.asarray(mats)) and np.array_equal(w, np.asarray(weights))
    assert meta["component_nodes"] == [na * nb * ng] * 2


def test_rotation_validation():
    r, w, _ = cm.orientation_rule({"quadrature": {"azimuth": 3, "tilt": 2, "spin": 2},
                                   "components": [{"kind": "isotropic", "fraction": 1.0}]})
    cm.validate_rotations(r, w)
    for bad in (r[3] @ np.diag([1, 1, -1.0]),  # improper: det -1
                r[3] * (1 + 3e-12),            # not orthogonal beyond the 2e-12 tolerance
                r[3] + np.triu(np.full((3, 3), 5e-12), 1)):
        rr = r.copy()
        rr[3] = bad
        with pytest.raises(ValueError, match="proper orthogonal"):
            cm.validate_rotations(rr, w)
    cm.validate_rotations(r * (1 + 5e-13), w)  # within the tolerance
    with pytest.raises(ValueError, match="unnormalized"):
        cm.validate_rotations(r, w * 1.01)
    with pytest.raises(ValueError, match="Malformed"):
        cm.validate_rotations(r[:, :2], w)


def test_bad_rule():
    for sigma in (-1, float("nan")):
        with pytest.raises(ValueError):
            cm.gaussian_tilt_nodes(10, sigma)
    with pytest.raises(ValueError):
        cm.orientation_rule({"quadrature": {"azimuth": 2, "tilt": 2, "spin": 2},
                             "components": [{"kind": "isotropic", "fraction": .5}]})
    with pytest.raises(ValueError):
        cm.orientation_rule({"quadrature": {"azimuth": 2, "tilt": 2, "spin": 2},
                             "components": [{"kind": "fixed", "fraction": 1, "quaternion_wxyz": [1, 0, 0, .1]}]})


def test_interpolation_endpoints_and_partition():
    for i in range(7):
        for j in range(4):
            terms = cm.interpolation(i * math.pi / 12, j * math.pi / 18, 6, 3)
            assert abs(sum(w for _, _, w in terms) - 1) < 1e-15
            assert abs(sum(w for ii, jj, w in terms if (ii, jj) == (i, j)) - 1) < 1e-15
    for t in np.linspace(0, math.pi / 2, 21):
        terms = cm.interpolation(t, .14, 6, 3)
        assert all(w >= 0 for _, _, w in terms) and abs(sum(w for _, _, w in terms) - 1) < 1e-15


def test_no_incident_extrapolation():
    with pytest.raises(ValueError):
        cm.interpolation(-.01, .1, 6, 3)
    with pytest.raises(ValueError):
        cm.interpolation(.1, .6, 6, 3)


@pytest.mark.parametrize("reflected", [False, True])
@pytest.mark.parametrize("shift", [0, 1, 2.3, -.7, 11.999])
def test_pure_shift_pose_couples_two_neighbouring_shifts(reflected, shift):
    """A pose that maps node (2, 1)'s frame to the output frame turned by
    `shift` azimuth cells puts its whole weight on node (2, 1) at shifts
    floor(shift) and floor(shift) + 1, split linearly (test_compose.py:61-77)."""
    a = fixture()
    f = cm.canonical_frame(math.pi / 6, math.pi / 18)
    fw = cm.frame([1, 0, 0])
    q = np.diag([1., 1., -1.]) if reflected else np.eye(3)
    d = 2 * math.pi * shift / a.np
    h = np.array([[math.cos(d), -math.sin(d), 0], [math.sin(d), math.cos(d), 0], [0, 0, 1.]])
    if reflected:
        h = h @ np.diag([1., -1., 1.])
    r = fw @ h @ f.T @ q.T
    w, expected, _ = cm.coupling(a, np.array([r]), np.ones(1), [1, 0, 0])
    lower = math.floor(shift)
    frac = shift - lower
    parity = int(reflected)
    # all but rounding-sized interpolation weight on node (2, 1) with the pose's parity
    assert w[2, 1, parity].sum() > 1 - 1e-12 and w.sum() - w[2, 1, parity].sum() < 1e-12
    assert abs(w[2, 1, parity, lower % 12] - (1 - frac)) < 1e-12
    if frac > 1e-12:
        assert abs(w[2, 1, parity, (lower + 1) % 12] - frac) < 1e-12
    # the composed mass of those weights is the shifted node block (oracle sum)
    out = compose_mass_ref(a.mass, w)
    original = a.mass[2, 1][:, ::-1] if reflected else a.mass[2, 1]
    want = (1 - frac) * np.roll(original, lower, axis=1) + frac * np.roll(original, lower + 1, axis=1)
    np.testing.assert_allclose(out, want, atol=3e-17, rtol=1e-12)
    assert abs(out.sum() - expected) < 1e-15


def test_area_weighted_mixture_weights():
    a = fixture()
    rot = np.array([np.eye(3), cm.rz(math.pi / 2)])
    w, area, _ = cm.coupling(a, rot, np.array([.25, .75]), [1, 0, 0])
    exact = .25 * cm.projected_area(.1, .2, [1, 0, 0]) + .75 * cm.projected_area(.1, .2, [0, -1, 0])
    assert abs(area - exact) < 1e-15
    assert abs(compose_mass_ref(a.mass, w).sum() - exact) < 1e-15


def test_vertical_incidence_integrates_the_azimuth_orbit():
    """k = -y: every isotropic/Gaussian contribution is spread evenly over all
    shifts, so the composed mass is uniform in azimuth (test_polar_azimuth.py)."""
    a = fixture()
    config = {"quadrature": {"azimuth": 8, "tilt": 4, "spin": 3}, "components": [
        {"kind": "horizontal_gaussian", "sigma_deg": 5, "fraction": .5}, {"kind": "isotropic", "fraction": .5}]}
    r, w, meta = cm.orientation_rule(config)
    c, area, _, info = cm.coupling_for_rule(a, r, w, [0, -1, 0], meta)
    assert info["method"] == "analytic_uniform_world_azimuth_at_vertical_incidence"
    assert info["effective_pose_evaluations"] == 2 * 4 * 3
    assert np.all(c == c[..., :1])
    out = compose_mass_ref(a.mass, c)
    np.testing.assert_allclose(out, np.repeat(out.mean(axis=1, keepdims=True), a.np, axis=1), rtol=1e-13)
    plain, area2, _ = cm.coupling(a, r, w, [0, -1, 0])
    assert abs(area / area2 - 1) < 1e-13  # the orbit integral keeps the projected area
    _, _, _, info = cm.coupling_for_rule(a, r, w, [0, -1, 1e-9], meta)
    assert info["method"] == "original_fixed_orientation_quadrature"


# ---- polar incidence: the analytic azimuth orbit (after hex_atlas/tests/test_polar_azimuth.py:13-126)
def _polar_config(na=16):
    return {"quadrature": {"azimuth": na, "tilt": 4, "spin": 3}, "components": [
        {"kind": "horizontal_gaussian", "sigma_deg": 10, "fraction": .9}, {"kind": "isotropic", "fraction": .1}]}


def _evaluate(a, config, ki, analytic=True):
    """Composed mass (oracle sum), unresolved tail, area, frame and method info."""
    r, w, meta = cm.orientation_rule(config)
    if analytic:
        c, area, f, info = cm.coupling_for_rule(a, r, w, ki, meta)
    else:
        (c, area, f), info = cm.coupling(a, r, w, ki), {}
    return compose_mass_ref(a.mass, c), float(np.sum(np.sum(c, axis=(2, 3)) * a.residual)), area, f, info


@pytest.mark.parametrize("sign", [-1, 1])
def test_polar_orbit_is_the_row_mean_of_the_discrete_rule(sign):
    """At k = -y or +y the orbit spreads each row's mass evenly over azimuth:
    each row of the composed mass is the row mean of the discrete quadrature's,
    with the same row totals, tail and projected area."""
    a = fixture()
    a.mass *= .97
    a.residual = .03 * a.area
    old, old_tail, old_area, _, _ = _evaluate(a, _polar_config(), [0, sign, 0], analytic=False)
    new, tail, area, f, info = _evaluate(a, _polar_config(), [0, sign, 0])
    np.testing.assert_allclose(new, np.repeat(old.mean(axis=1, keepdims=True), a.np, axis=1), rtol=1e-12,
                               atol=2e-17)
    np.testing.assert_allclose(new.sum(axis=1), old.sum(axis=1), rtol=1e-12, atol=2e-17)
    assert np.array_equal(f, cm.frame([0, sign, 0]))
    assert abs(tail - old_tail) <= 1e-16 and abs(area - old_area) <= 1e-15
    assert abs(new.sum() + tail - area) <= 1e-15
    assert info["analytic_uniform_azimuth_components"] == [0, 1]
    assert info["effective_pose_evaluations"] == 4 * 3 * 2 and info["output_post_filter"] is False


@pytest.mark.parametrize("sign", [-1, 1])
def test_polar_orbit_equals_a_complete_discrete_cycle(sign):
    """With one azimuth node per output cell the discrete rule is itself a
    complete cyclic average, and agrees with the analytic orbit."""
    a = fixture()
    a.mass[:] = 0
    a.mass[:, :, 1, 2] = a.area
    new = _evaluate(a, _polar_config(3), [0, sign, 0])[0]
    discrete = _evaluate(a, _polar_config(a.np), [0, sign, 0], analytic=False)[0]
    np.testing.assert_allclose(new, discrete, rtol=2e-12, atol=3e-17)


def test_polar_orbit_ignores_the_azimuth_node_count():
    a = fixture()
    first = _evaluate(a, _polar_config(1), [0, -1, 0])[0]
    for na in (2, 7, 16, 23):
        assert np.array_equal(_evaluate(a, _polar_config(na), [0, -1, 0])[0], first)


@pytest.mark.parametrize("ki", [[1, 0, 0], [1, -.7, .2], [1e-13, -1, 0]])
def test_off_pole_directions_use_the_discrete_rule(ki):
    a = fixture()
    new, tail, area, f, info = _evaluate(a, _polar_config(), ki)
    old, old_tail, old_area, old_f, _ = _evaluate(a, _polar_config(), ki, analytic=False)
    assert np.array_equal(new, old) and np.array_equal(f, old_f)
    assert tail == old_tail and area == old_area
    assert info["method"] == "original_fixed_orientation_quadrature"


def test_fixed_pose_at_the_pole_is_not_averaged():
    a = fixture()
    config = dict(_polar_config(), components=[{"kind": "fixed", "fraction": 1., "quaternion_wxyz": [1, 0, 0, 0]}])
    new, *_, info = _evaluate(a, config, [0, -1, 0])
    assert np.array_equal(new, _evaluate(a, config, [0, -1, 0], analytic=False)[0])
    assert float(np.max(np.ptp(new, axis=1))) > 0
    assert info["analytic_uniform_azimuth_components"] == []


def test_mixture_with_a_fixed_pose_averages_only_the_uniform_part():
    a = fixture()
    fixed = {"kind": "fixed", "fraction": .4, "quaternion_wxyz": [1, 0, 0, 0]}
    iso = {"kind": "isotropic", "fraction": .6}
    quadrature = _polar_config()["quadrature"]
    mass, _, area, _, info = _evaluate(a, {"quadrature": quadrature, "components": [fixed, iso]}, [0, -1, 0])
    pieces = [_evaluate(a, {"quadrature": quadrature, "components": [dict(c, fraction=1.)]}, [0, -1, 0])
              for c in (fixed, iso)]
    np.testing.assert_allclose(mass, .4 * pieces[0][0] + .6 * pieces[1][0], rtol=2e-12, atol=2e-17)
    assert abs(area - (.4 * pieces[0][2] + .6 * pieces[1][2])) < 1e-15
    assert info["analytic_uniform_azimuth_components"] == [1]
    assert float(np.max(np.ptp(mass, axis=1))) > 0


def test_polar_orbit_keeps_theta_sparsity(oracle):
    """One populated theta row stays the only populated row (every azimuth of it
    now positive), and its CDF (oracle restatement of the CDF build) gives the
    cell probabilities back."""
    a = fixture()
    a.mass[:] = 0
    a.mass[:, :, 1, 2] = a.area
    mass = _evaluate(a, _polar_config(), [0, -1, 0])[0]
    assert np.all(mass[1] > 0)
    assert np.array_equal(mass[[0, 2, 3]], np.zeros((3, a.np)))
    rc, phi, theta, rep = oracle.build_cdf(mass, a.nt, a.np)
    assert rc == 0
    recovered = (np.diff(phi)[:, None] * np.diff(theta, axis=1)).T
    np.testing.assert_allclose(recovered, mass / mass.sum(), rtol=1e-13, atol=2e-16)
    assert rep.l1_error < 1e-12 and rep.lost_mass == 0


def test_polar_orbit_needs_the_declared_rule():
    a = fixture()
    r, w, meta = cm.orientation_rule(_polar_config())
    with pytest.raises(ValueError, match="declared"):
        cm.coupling_for_rule(a, r, w[::-1].copy(), [0, -1, 0], meta)


def test_residual_not_renormalized_away():
    a = fixture()
    a.mass *= .99
    a.residual = .01 * a.area
    w, area, _ = cm.coupling(a, np.array([cm.ry(.213) @ cm.rz(.7)]), np.ones(1), [1, -.4, .3])
    assert abs(compose_mass_ref(a.mass, w).sum() / area - .99) < 1e-13
    assert abs(float(np.sum(np.sum(w, axis=(2, 3)) * a.residual)) / area - .01) < 1e-13


def test_incidents():
    got = cm.incidents({"inclinations_deg": [0, 35, 90, -90]})
    np.testing.assert_allclose(got[1][0], [math.cos(math.radians(35)), -math.sin(math.radians(35)), 0], atol=1e-16)
    np.testing.assert_array_equal(got[2][0], [0, -1, 0])
    np.testing.assert_array_equal(got[3][0], [0, 1, 0])
    with pytest.raises(ValueError):
        cm.incidents({"inclinations_deg": [91]})
    with pytest.raises(ValueError):
        cm.incidents({"inclinations_deg": [0], "directions": [[1, 0, 0]]})


def test_coupling_stream_is_the_sequence_of_couplings():
    """The composer's stream (next coupling computed in a worker thread) yields
    exactly coupling_for_rule() of each direction, in order; a failing
    direction raises when its turn comes; closing early waits for the worker."""
    a = fixture()
    config = {"quadrature": {"azimuth": 5, "tilt": 4, "spin": 3}, "components": [
        {"kind": "horizontal_gaussian", "sigma_deg": 9, "fraction": .8}, {"kind": "isotropic", "fraction": .2}]}
    r, w, meta = cm.orientation_rule(config)
    kis = [np.array(k, dtype=float) for k in ([0, -1, 0], [1, 0, 0], [.5, -.8, 0], [.2, .9, 0])]
    got = list(cm.coupling_stream(a, r, w, kis, meta))
    assert len(got) == len(kis)
    for k, ((c, area, f, info), seconds) in zip(kis, got):
        c2, area2, f2, info2 = cm.coupling_for_rule(a, r, w, k, meta)
        assert np.array_equal(c, c2) and area == area2 and np.array_equal(f, f2) and info == info2
        assert seconds > 0
    stream = cm.coupling_stream(a, r, w, [kis[1], np.array([np.nan, 1.0, 0.0]), kis[2]], meta)
    next(stream)
    with pytest.raises(ValueError):
        next(stream)
    stream = cm.coupling_stream(a, r, w, kis, meta)
    next(stream)
    stream.close()
    assert list(cm.coupling_stream(a, r, w, [], meta)) == []


@pytest.mark.parametrize("k", [[1, 0, 0], [0.5, -0.8, 0], [1, -.4, .3], [0, -1, 0]])
def test_array_coupling_equals_scalar_rule(k):
    """The array coupling (all poses at once, weights accumulated in the scalar
    order) equals the pose-by-pose rule of atlas_math.py up to last-bit
    differences of the elementary functions."""
    a = fixture()
    config = {"quadrature": {"azimuth": 6, "tilt": 5, "spin": 4}, "components": [
        {"kind": "horizontal_gaussian", "sigma_deg": 7, "fraction": .6}, {"kind": "isotropic", "fraction": .3},
        {"kind": "fixed", "fraction": .1, "quaternion_wxyz": [math.cos(.3), 0, math.sin(.3), 0]}]}
    r, w, _ = cm.orientation_rule(config)
    fast, area_fast, f1 = cm.coupling(a, r, w, k)
    slow, area_slow, f2 = cm.coupling_scalar(a, r, w, k)
    assert np.array_equal(fast > 0, slow > 0)
    assert np.abs(fast - slow).sum() / slow.sum() < 1e-13
    assert abs(area_fast / area_slow - 1) < 1e-14 and np.array_equal(f1, f2)


def _fake_atlas(path, it=2, ip_=1, nt=4, np_=8, wavelength=550.0):
    """A well-formed atlas directory (uniform node masses) for the composer's
    host checks; no device is involved."""
    import json as _json
    path.mkdir()
    nodes, mass = [], np.zeros((it + 1, ip_ + 1, nt, np_))
    for i in range(it + 1):
        for j in range(ip_ + 1):
            f = cm.canonical_frame(i * (math.pi / 2) / it, j * (math.pi / 6) / ip_)
            area = cm.projected_area(.1, .2, f[:, 2])
            mass[i, j] = area / (nt * np_)
            nodes.append({"i": i, "j": j, "ki": f[:, 2].tolist(), "e0": f[:, 0].tolist(), "projected_area_mm2": area,
                          "escaped_mm2": area, "unresolved_mm2": 0.0})
    np.save(path / "mass_mm2.npy", mass)
    np.save(path / "u_edges.npy", (1 - np.cos(np.linspace(0, math.pi, nt + 1))) / 2)
    meta = {"schema": "ice.single_particle_atlas.v1", "complete": True,
            "incident_reconstruction": "bilinear_theta_phi_in_spherical_incident_frame",
            "symmetry": "D6h_O3_full_24_scalar_unpolarized", "incident_theta_intervals": it,
            "incident_phi_intervals": ip_, "theta_count": nt, "phi_count": np_,
            "geometry": {"shape": "regular_hexagonal_prism", "circumradius_mm": .1, "full_length_mm": .2},
            "material": {"model": "constant", "wavelength_nm": wavelength, "interior_index": 1.31,
                         "exterior_index": 1.0},
            "transport": {"max_unresolved_fraction": 1e-10}, "nodes": nodes}
    (path / "metadata.json").write_text(_json.dumps(meta))
    return path


def test_atlas_reader_audits(tmp_path):
    a = cm.Atlas(_fake_atlas(tmp_path / "atlas"))
    assert (a.it, a.ip, a.nt, a.np) == (2, 1, 4, 8) and a.area.shape == (3, 2)
    a.close()
    bad = _fake_atlas(tmp_path / "bad")
    m = json.loads((bad / "metadata.json").read_text())
    m["nodes"][1]["escaped_mm2"] *= 1.01
    (bad / "metadata.json").write_text(json.dumps(m))
    with pytest.raises(ValueError, match="audit"):
        cm.Atlas(bad)
    with pytest.raises(ValueError, match="Incomplete"):
        cm.Atlas(tmp_path / "x.part")


def test_compose_cli_failure_keeps_part_with_reason(tmp_path, monkeypatch):
    """A failure after staging leaves <out>.part with failure.json and exit 2
    (compose_atlas.py:209-212, 225-229); here the device context is refused."""
    atlas = _fake_atlas(tmp_path / "atlas")
    cfg = tmp_path / "c.json"
    cfg.write_text(json.dumps({"inclinations_deg": [0], "orientation": {
        "quadrature": {"azimuth": 2, "tilt": 2, "spin": 2}, "components": [{"kind": "isotropic", "fraction": 1}]}}))

    def no_device(device=0):
        raise cm.ip.IcePhaseError(-2, "no device")

    monkeypatch.setattr(cm.ip, "Context", no_device)
    with pytest.raises(SystemExit) as e:
        cm.main(["--atlas", str(atlas), "--config", str(cfg), "--out", str(tmp_path / "out")])
    assert e.value.code == 2
    assert not (tmp_path / "out").exists()
    assert json.loads((tmp_path / "out.part" / "failure.json").read_text())["complete"] is False
    assert (tmp_path / "out.part" / "orientation_model.json").exists()
    # the staging directory is never adopted or removed
    with pytest.raises(SystemExit):
        cm.main(["--atlas", str(atlas), "--config", str(cfg), "--out", str(tmp_path / "out")])


def test_compose_rejects_noncanonical_incidence_for_the_rainbow_contract(tmp_path):
    atlas = _fake_atlas(tmp_path / "atlas")
    cfg = tmp_path / "c.json"
    cfg.write_text(json.dumps({"directions": [[1, 0, 0.2]], "orientation": {
        "quadrature": {"azimuth": 1, "tilt": 1, "spin": 1}, "components": [{"kind": "isotropic", "fraction": 1}]}}))
    with pytest.raises(ValueError, match="Noncanonical"):
        cm.run([atlas], cfg, tmp_path / "out")
    assert not (tmp_path / "out.part").exists()
