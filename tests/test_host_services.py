"""CPU tests of the host services behind the C-ABI (no device calls): water
index, Table I drop shapes, drop convexity and projected area, hex projected
area; and of the oracle's orientation-node composition and Fraunhofer lobe
restatement, pinned by the compiled reference and by closed forms.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import ice_crystal_phase_amd as ip

GOLDEN = __import__("pathlib").Path(__file__).resolve().parent / "golden"


# ---------------------------------------------------------------- water (A24)
def test_water_matches_golden_reference_values():
    """evaluate_water_optics golden values written from the compiled reference
    (tests/golden/make_golden.py): n to 1e-13, density to 1e-9 kg/m^3."""
    d = np.load(GOLDEN / "materials.npz")
    for lam, c, p, n, rho in d["water"]:
        w = ip.water_optics(lam, c, p)
        assert abs(w.refractive_index - n) < 1e-13, (lam, c, p)
        assert abs(w.density_kg_m3 - rho) < 1e-9
        assert abs(w.density_pressure_residual_pascal) <= 0.02
        assert w.density_bracket_width_kg_m3 <= 2e-10


def test_water_against_compiled_reference_grid(oracle):
    if not oracle.ref_available():
        pytest.skip("compiled reference shim not built (no reference checkout)")
    for lam in (200.0, 380.5, 450.0, 550.0, 633.0, 780.0, 1100.0):
        for c in (0.0, 4.0, 20.0, 37.5, 60.0):
            for p in (5e4, 101325.0, 1e6, 5e7, 1e8):
                rc, n, rho = oracle.ref_water_index(lam, c, p)
                assert rc == 0
                w = ip.water_optics(lam, c, p)
                assert abs(w.refractive_index - n) < 2e-13
                assert abs(w.density_kg_m3 - rho) < 1e-9


def test_water_published_check_values():
    """IAPWS R9-97 check points: n at (0.589 um, 20 C liquid ~998.2 kg/m^3) is
    1.33336 to 5 digits (the sodium D line index of water); the index grows with
    density and falls with wavelength in the visible."""
    assert abs(ip.water_optics(589.0, 20.0).refractive_index - 1.33336) < 2e-5
    n_lo = ip.water_index(550.0, 293.15, 990.0)
    n_hi = ip.water_index(550.0, 293.15, 1000.0)
    assert n_hi > n_lo
    assert ip.water_optics(450.0).refractive_index > ip.water_optics(650.0).refractive_index


@pytest.mark.parametrize("args", [(550.0, -1.0, 101325.0), (550.0, 61.0, 101325.0), (550.0, 20.0, 4e4),
                                  (150.0, 20.0, 101325.0), (1200.0, 20.0, 101325.0)])
def test_water_rejects_out_of_range(args):
    with pytest.raises(ip.IcePhaseError):
        ip.water_optics(*args)


# ---------------------------------------------------------------- Table I (A1)
def test_drop_shape_matches_compiled_reference(oracle):
    if not oracle.ref_available():
        pytest.skip("compiled reference shim not built")
    for r in list(np.linspace(0.05, 3.0, 60)) + [0.4, 1.0, 1.5, 2.0, 2.5, 3.0, 0.4000001]:
        rc, c_ref, _, _ = oracle.ref_drop_shape(float(r))
        assert rc == 0
        c = np.array(ip.drop_shape(float(r)), np.float32)
        assert np.array_equal(c.view(np.uint32), c_ref.view(np.uint32)), r
    rc, c_ref, _, _ = oracle.ref_drop_shape(2.0, True)
    assert np.all(np.array(ip.drop_shape(2.0, True)) == 0) and np.all(c_ref == 0)


def test_drop_shape_table_rows_and_range():
    c = ip.drop_shape(2.0)
    assert np.allclose(c[:4], [-0.0458, -0.0335, -0.1211, -0.0227], atol=1e-7)
    assert all(v == 0 for v in ip.drop_shape(0.3))
    for bad in (0.0, -1.0, 3.01):
        with pytest.raises(ip.IcePhaseError):
            ip.drop_shape(bad)


# ---------------------------------------------------------------- projected area (A25)
def test_drop_projected_area_matches_golden_and_reference(oracle):
    d = np.load(GOLDEN / "materials.npz")
    for row in d["drop_area"]:
        a = ip.drop_projected_area(row[:8], 2.0, row[8:11])
        assert abs(a.area_mm2 / row[11] - 1) < 1e-10
    if not oracle.ref_available():
        return
    for radius in (0.5, 1.0, 2.0, 3.0):
        coeffs = ip.drop_shape(radius)
        for alpha in (-90.0, -61.0, -12.5, 0.0, 33.0, 90.0):
            k = ip.drop_incident(alpha)
            rc, area, sph = oracle.ref_projected_area(coeffs, radius, k)
            assert rc == 0
            a = ip.drop_projected_area(coeffs, radius, k)
            assert abs(a.area_mm2 / area - 1) < 1e-10, (radius, alpha)
            assert a.rel_error < 1e-10


def test_projected_area_closed_forms():
    """Sphere: pi a^2 exactly. Drop viewed along its axis: pi max(r sin chi)^2
    (the silhouette is the largest parallel), by direct maximization."""
    a = ip.drop_projected_area([0.0] * 8, 1.5, (0.3, 0.2, 0.1))
    assert a.analytic_sphere and a.area_mm2 == math.pi * 1.5 * 1.5
    c = ip.drop_shape(2.5)
    chi = np.linspace(0, np.pi, 400001)
    mu = np.cos(chi)
    r = 1 + c[0] + c[1] * mu + sum(c[n] * np.cos(n * chi) for n in range(2, 8))
    rho_max = np.max(r * np.sin(chi))
    a = ip.drop_projected_area(c, 2.5, (0.0, -1.0, 0.0))
    assert abs(a.area_mm2 / (math.pi * (2.5 * rho_max) ** 2) - 1) < 1e-9


def test_nonconvex_drop_rejected():
    with pytest.raises(ip.IcePhaseError, match="convex"):
        ip.drop_projected_area([0, 0, 0.4, 0, 0, 0, 0, 0], 1.0, (1, 0, 0))
    with pytest.raises(ip.IcePhaseError):
        ip.drop_projected_area([0, 0, 0, 0, 0, 0, 0.05, 0], 1.0, (1, 0, 0))


def test_hex_projected_area_closed_forms(oracle):
    """Axial incidence: 3 sqrt3/2 R^2; side-on along a face normal: 2 R L
    (hex_trace/tests/hex_trace_tests.cpp:31-32); side-on along a vertex: sqrt3 R L."""
    R, L = 1.0, 2.0
    assert abs(ip.hex_projected_area(R, L, (0, 1, 0)) - 1.5 * math.sqrt(3) * R * R) < 1e-14
    assert abs(ip.hex_projected_area(R, L, (1, 0, 0)) - 2 * R * L) < 1e-14
    assert abs(ip.hex_projected_area(R, L, (math.sqrt(3) / 2, 0, 0.5)) - math.sqrt(3) * R * L) < 1e-14
    for k in ((1, -0.4, 0.3), (0.2, 0.9, -0.1)):
        ku = np.array(k, float) / np.linalg.norm(k)
        want = oracle.lib().orc_hex_projected_area(0.05, 0.2, oracle.vec3(ku))
        assert abs(ip.hex_projected_area(0.05, 0.2, k) / want - 1) < 1e-15


# ---------------------------------------------------------------- orientation composition (A15)
def _plan_csv(rows):
    return "qw,qx,qy,qz,weight,samples,seed\n" + "".join(",".join(str(v) for v in r) + "\n" for r in rows)


def test_orientation_node_oracle_matches_reference_composition(oracle):
    """Fixed-point composition of three weighted nodes equals the reference
    CLI's FP64 compensated masses (PhaseMassAccumulator with rotation) cell by
    cell to the quantum resolution: a wrong rotation order or transpose moves
    whole samples between cells and fails."""
    if not oracle.ref_available():
        pytest.skip("compiled reference shim not built")
    s2 = math.sqrt(0.5)
    rows = [(1, 0, 0, 0, 0.5, 1500, 1), (s2, 0, s2, 0, 0.25, 1000, 2), (0.9238795325112867, 0.3826834323650898, 0, 0,
                                                                          0.25, 1000, 3)]
    nt, np_ = 36, 72
    k = (1.0, -0.4, 0.3)
    rc, ref_mass, ref_area = oracle.ref_orientation_masses(0.05, 0.2, k, 1.0, 1.31, 64, 1e-12, _plan_csv(rows), nt,
                                                           np_)
    assert rc == 0
    total = sum(r[5] for r in rows)
    qbits = oracle.quantum_bits(total)
    scales = []
    area = 0.0
    for w, x, y, z, wt, n, seed in rows:
        R = ip.quaternion_rotation(w, x, y, z)
        kb = R.T @ (np.array(k) / np.linalg.norm(k))
        a = ip.hex_projected_area(0.05, 0.2, kb)
        scales.append(wt * a / n)
        area += wt * a
    smax = max(scales)
    bins = np.zeros(nt * np_, np.uint64)
    st = oracle.HistStats()
    for (w, x, y, z, wt, n, seed), sc in zip(rows, scales):
        s = oracle.hex_settings(ki=k, seed=seed, total=n, max_hits=64)
        oracle.hex_hist_node(s, ip.quaternion_rotation(w, x, y, z), sc / smax, nt, np_, qbits, bins, st)
    mass = oracle.bins_to_mass(bins, st.q_forward, st.q_backward, nt, np_) * (smax * 2.0 ** -qbits)
    assert abs(area / ref_area - 1) < 1e-14
    quantum = smax * 2.0 ** -qbits
    # each sample rounds to half a quantum; a cell holds at most a few hundred samples
    assert np.max(np.abs(mass - ref_mass)) < 1e3 * quantum
    assert abs(mass.sum() / ref_mass.sum() - 1) < 1e-12


def test_orientation_identity_node_equals_plain_record(oracle):
    """An identity-rotation node with qfactor 1 bins like the plain record, up to
    the extra normalization of the ensemble direction (identical here: the
    incident direction is already a normalized FP64 vector)."""
    k = tuple(np.array([1.0, -0.4, 0.3]) / np.linalg.norm([1.0, -0.4, 0.3]))
    s = oracle.hex_settings(ki=k, total=800)
    b0, st0 = oracle.hex_hist(s, 24, 48)
    b1, st1 = oracle.hex_hist_node(s, np.eye(3), 1.0, 24, 48, oracle.quantum_bits(800))
    assert np.array_equal(b0, b1)
    assert st0.q_escaped == st1.q_escaped


# ---------------------------------------------------------------- Fraunhofer lobe (A7)
def test_aperture_transform_closed_forms(oracle):
    """F(q) of the polygon: rectangle a x b -> A sinc(q1 a/2) sinc(q2 b/2);
    F(q) -> A as q -> 0; a translated polygon picks up exp(-i q.t)."""
    a, b = 0.3, 0.7
    vx = np.array([-a / 2, a / 2, a / 2, -a / 2])
    vy = np.array([-b / 2, -b / 2, b / 2, b / 2])
    for q1, q2 in ((3.0, 5.0), (-11.0, 0.5), (40.0, -17.0), (1e-3, 2e-3)):
        F = oracle.aperture_transform(vx, vy, q1, q2)
        sinc = lambda x: math.sin(x) / x
        want = a * b * sinc(q1 * a / 2) * sinc(q2 * b / 2)
        assert abs(F - want) < 1e-12 * a * b, (q1, q2)
    t = (0.11, -0.05)
    F0 = oracle.aperture_transform(vx, vy, 7.0, 3.0)
    F1 = oracle.aperture_transform(vx + t[0], vy + t[1], 7.0, 3.0)
    assert abs(F1 - F0 * complex(math.cos(-(7 * t[0] + 3 * t[1])), math.sin(-(7 * t[0] + 3 * t[1])))) < 1e-13


def test_aperture_transform_hexagon_brute_force(oracle):
    """Regular hexagon (basal incidence): closed form against a 2D midpoint
    quadrature of exp(-i q.x) over the polygon."""
    vx, vy, _, _ = oracle.hex_aperture(1.0, 2.0, (0.0, 1.0, 0.0))
    assert len(vx) == 6
    area = oracle.polygon_area(vx, vy)
    assert abs(area - 1.5 * math.sqrt(3)) < 1e-12
    n = 1600
    xs = np.linspace(-1, 1, n, endpoint=False) + 1.0 / n
    X, Y = np.meshgrid(xs, xs, indexing="ij")
    inside = np.ones_like(X, bool)
    for i in range(len(vx)):
        j = (i + 1) % len(vx)
        inside &= (vx[j] - vx[i]) * (Y - vy[i]) - (vy[j] - vy[i]) * (X - vx[i]) >= 0
    dA = (2.0 / n) ** 2
    for q1, q2 in ((2.0, 1.0), (-4.5, 3.0)):
        brute = np.sum(np.exp(-1j * (q1 * X + q2 * Y))[inside]) * dA
        F = oracle.aperture_transform(vx, vy, q1, q2)
        assert abs(F - brute) < 5e-3 * area


@pytest.mark.parametrize("k", [(1.0, -0.4, 0.3), (1.0, 0.0, 0.0), (0.2, 0.9, -0.3)])
def test_hex_aperture_area_is_face_sum_area(oracle, k):
    vx, vy, _, _ = oracle.hex_aperture(0.05, 0.2, k)
    assert abs(oracle.polygon_area(vx, vy) / ip.hex_projected_area(0.05, 0.2, k) - 1) < 1e-12


def test_lobe_symmetries_and_energy(oracle):
    """Basal incidence: the regular hexagon's pattern repeats every 60 degrees
    (column j and j + np/6 carry the same mass) and is centro-symmetric; the
    lobe over the leading rows carries the aperture area up to the truncated
    tail (Parseval), and its peak density is A^2/lambda^2."""
    nt, np_, rows = 180, 36, 12
    lam = 550.0
    R, L = 0.05, 0.2
    m, area, dmax = oracle.lobe_cells(R, L, (0.0, 1.0, 0.0), lam, nt, np_, rows)
    assert np.allclose(m, np.roll(m, np_ // 6, axis=1), rtol=1e-9, atol=0)
    assert np.allclose(m, np.roll(m, np_ // 2, axis=1), rtol=1e-9, atol=0)
    captured = m.sum() / area
    assert 0.97 < captured <= 1.0 + 1e-9
    # first row: the cap integral approaches pi (dtheta)^2 A^2 / lambda^2 when the
    # cap is much smaller than the lobe: use a fine grid
    m2, _, _ = oracle.lobe_cells(R, L, (0.0, 1.0, 0.0), lam, 90000, 4, 1)
    lam_mm = lam * 1e-6
    dth = math.pi / 90000
    want = math.pi * dth * dth * area * area / lam_mm ** 2
    assert abs(m2.sum() / want - 1) < 2e-3
