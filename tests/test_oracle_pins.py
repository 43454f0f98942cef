"""Pins of the CPU oracle (runs on CPU; no GPU).

Each oracle function is checked against something independent of its own
formula: the reference's known-answer tests, closed forms, the golden fixtures
written by the compiled reference (tests/golden/make_golden.py) and, when
oracle/_ref/libref.so is built, the compiled reference itself.
"""
import math
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"
C3 = dict(radius_mm=0.05, length_mm=0.2, ki=(1.0, -0.4, 0.3), n_int=1.31, tail=1e-12, seed=12345, max_hits=12)


def g(name):
    return np.load(GOLDEN / name, allow_pickle=False)


# ----------------------------------------------------------------- u edges
@pytest.mark.parametrize("n", [1, 2, 3, 17, 36, 450, 900, 1800])
def test_u_edges_closed_form(oracle, n):
    e = oracle.u_edges(n)
    assert e[0] == 0 and e[-1] == 1 and np.all(np.diff(e) > 0)
    theta = np.arange(n + 1) * math.pi / n
    assert np.max(np.abs(e - (1 - np.cos(theta)) / 2)) < 4e-16
    # the upper half is the binary64 mirror of the lower half: u_{n-i} = 1 - u_i, i < n/2
    for i in range(1, (n + 1) // 2):
        assert e[n - i] == 1 - e[i]


def test_uniform_sphere_mean_cosine_is_zero(oracle):
    """test_phase_storage_cpu.cpp:93-107: sum du (1 - u_lo - u_hi) = 0."""
    for n in (1, 2, 17, 900, 1800):
        e = oracle.u_edges(n)
        gsum = np.sum(np.diff(e).astype(np.longdouble) * (1 - e[:-1] - e[1:]).astype(np.longdouble))
        assert abs(gsum) < 5e-16


def test_saved_cdf_first_moment_kat():
    """test_phase_storage_cpu.cpp:108-118 known answer: g = -0.40125."""
    phi = [0, .25, 1]
    c = [[0, .1, .5, 1], [0, 0, .25, 1]]
    e = [0, .1, .7, 1]
    m = gg = 0.0
    for j in range(2):
        for i in range(3):
            w = (phi[j + 1] - phi[j]) * (c[j][i + 1] - c[j][i])
            m += w
            gg += w * (1 - (e[i] + e[i + 1]))
    assert abs(m - 1) < 2e-16 and abs(gg + 0.40125) < 2e-16


# ----------------------------------------------------------------- phase_bin
def test_phase_bin_reference_kats(oracle):
    """hex_phase/tests/phase_tests.cpp:71-104."""
    k = (1, -0.4, 0.3)
    fr = oracle.make_frame(k)
    e = oracle.u_edges(8)
    ki, e0, e1 = (np.array(v) for v in (fr.ki, fr.e0, fr.e1))
    for i in range(8):
        for j in range(16):
            u = (e[i] + e[i + 1]) / 2
            mu = 1 - 2 * u
            st = 2 * math.sqrt(u * (1 - u))
            ph = -math.pi + (j + .5) * 2 * math.pi / 16
            d = ki * mu + e0 * (st * math.cos(ph)) + e1 * (st * math.sin(ph))
            b = oracle.phase_bin(e, 8, 16, fr, d)
            assert b.valid and not b.pole and b.index == i * 16 + j
    seam = oracle.phase_bin(e, 8, 16, fr, -e0)
    assert seam.valid and seam.index % 16 == 0
    assert oracle.phase_bin(e, 8, 16, fr, ki).pole == 1
    assert oracle.phase_bin(e, 8, 16, fr, -ki).pole == -1
    ax = oracle.make_frame((0, 0, 1))
    ea = oracle.u_edges(20)
    for p in (-1e-14, 0.0, 1e-14):
        c1 = oracle.phase_bin(ea, 20, 40, ax, (1, p, 0))
        assert c1.valid and c1.index == 10 * 40 + 20
        c2 = oracle.phase_bin(ea, 20, 40, ax, (1, 0, p))
        assert c2.valid and c2.index == 10 * 40 + 20
    away = np.array([1, -1e-7, 0.23]) / np.linalg.norm([1, -1e-7, 0.23])
    un = oracle.phase_bin(ea, 20, 40, ax, away)
    assert un.index % 40 == 19 and not un.snapped
    assert not oracle.phase_bin(ea, 20, 40, ax, (0, 0, 0)).valid


def test_phase_bin_golden(oracle):
    z = g("phase_bin.npz")
    for name in ("oblique", "x", "z"):
        k = z[f"{name}_k"]
        fr = oracle.make_frame(k)
        e = oracle.u_edges(900)
        dirs = z[f"{name}_dirs"]
        for t in range(dirs.shape[0]):
            b = oracle.phase_bin(e, 900, 1800, fr, dirs[t])
            assert (b.valid, b.pole, b.snapped) == (z[f"{name}_valid"][t], z[f"{name}_pole"][t],
                                                    z[f"{name}_snapped"][t]), t
            if b.valid and not b.pole:
                assert b.index == z[f"{name}_index"][t], t


# ----------------------------------------------------------------- hex tracer
def test_hex_samples_golden_bitwise(oracle):
    z = g("hex_samples.npz")
    for name in ("c3", "plate", "slab_normal"):
        R, L, kx, ky, kz, n, tail, seed, mh, total = z[f"{name}_settings"]
        s = oracle.hex_settings(radius_mm=R, length_mm=L, ki=(kx, ky, kz), n_int=n, tail=tail, seed=int(seed),
                                max_hits=int(mh), total=int(total))
        samp, aud = oracle.hex_trace(s, 0, int(total))
        assert np.array_equal(samp["dir"], z[f"{name}_dir"]), name
        assert np.array_equal(samp["power"], z[f"{name}_power"]), name
        assert np.array_equal(samp["id"], z[f"{name}_id"]), name
        faces = np.stack([samp["entry_face"], samp["exit_face"], samp["reflections"], samp["kind"]], 1)
        assert np.array_equal(faces, z[f"{name}_faces"]), name
        af = np.stack([aud["escaped"], aud["unresolved"], aud["balance"], aud["max_interface"]], 1)
        assert np.array_equal(af, z[f"{name}_audit_f"]), name
        au = np.stack([aud["outputs"], aud["internal_hits"], aud["tir_count"], aud["status"]], 1)
        assert np.array_equal(au, z[f"{name}_audit_u"]), name


def test_hex_analytic_slab(oracle):
    """hex_trace_tests.cpp:168-191: normal incidence through the basal faces of a
    {1,2} prism: escaped R+(1-R)^2(1+R), residual (1-R)R^2 after two hits."""
    s = oracle.hex_settings(radius_mm=1, length_mm=2, ki=(0, -1, 0), n_int=1.31, tail=1e-15, seed=12345,
                            total=10, max_hits=2)
    samp, aud = oracle.hex_trace(s, 0, 10)
    R = (1 - 1.31) ** 2 / (1 + 1.31) ** 2
    assert np.all(aud["status"] == 2) and np.all(aud["outputs"] == 3)
    assert np.max(np.abs(aud["escaped"] - (R + (1 - R) ** 2 * (1 + R)))) < 3e-14
    assert np.max(np.abs(aud["unresolved"] - (1 - R) * R * R)) < 3e-15
    direct = samp[samp["kind"] == 1]
    assert np.max(np.abs(direct["dir"][::2] - np.array([0, -1, 0]))) < 2e-14
    s.max_hits, s.tail, s.n_int = 256, 0.0, 1.0
    samp, aud = oracle.hex_trace(s, 0, 10)
    assert np.all(aud["outputs"] == 1) and np.max(np.abs(aud["escaped"] - 1)) < 1e-15


def test_hex_normal_incidence_fresnel(oracle):
    """Normal-incidence reflectance ((n-1)/(n+1))^2 (hex_trace_tests.cpp:116-124)."""
    s = oracle.hex_settings(radius_mm=1, length_mm=2, ki=(0, -1, 0), n_int=1.31, tail=0, total=4, max_hits=1)
    samp, aud = oracle.hex_trace(s, 0, 4)
    R = (0.31 / 2.31) ** 2
    refl = samp[samp["kind"] == 0]
    assert np.max(np.abs(refl["power"] - R)) < 2e-15


def test_hex_transparent_limit_and_areas(oracle):
    """n=1: all light forward; pole mass 1/np per column; stored g = 1 - u_1
    (hex_phase/tests/phase_tests.cpp:172-193, phase_cli_test.py:103-110)."""
    nt, np_ = 16, 32
    s = oracle.hex_settings(radius_mm=1, length_mm=2, ki=(0, -1, 0), n_int=1.0, tail=0, total=64, max_hits=8)
    bins, st = oracle.hex_hist(s, nt, np_)
    assert bins.sum() == 0 and st.q_backward == 0 and st.q_forward > 0
    m = oracle.bins_to_mass(bins, st.q_forward, st.q_backward, nt, np_)
    assert np.all(m[np_:] == 0) and np.all(m[:np_] == m[0])
    rc, phi, theta, rep = oracle.build_cdf(m, nt, np_)
    assert rc == 0 and abs(rep.g_stored - (1 - oracle.u_edges(nt)[1])) < 3e-15
    assert abs(oracle.lib().orc_hex_projected_area(1, 2, oracle.vec3((0, 1, 0))) - 3 * math.sqrt(3) / 2) < 1e-14
    assert abs(oracle.lib().orc_hex_projected_area(1, 2, oracle.vec3((1, 0, 0))) - 4) < 1e-14
    # Cauchy: mean projected area over directions = S/4 (hex_trace_tests.cpp:28-45)
    n = 40000
    y = 1 - 2 * (np.arange(n) + 0.5) / n
    ph = 2 * math.pi * np.arange(n) * 0.6180339887498948482
    r = np.sqrt((1 - y) * (1 + y))
    area = np.mean([oracle.lib().orc_hex_projected_area(1, 2, oracle.vec3((r[i] * math.cos(ph[i]), y[i],
                                                                          r[i] * math.sin(ph[i]))))
                    for i in range(n)])
    S = 6 * 2 + 2 * 3 * math.sqrt(3) / 2
    assert abs(area / (S / 4) - 1) < 2e-5


def test_hex_masses_match_golden_and_fixed_point(oracle):
    """The oracle's Kahan masses equal the reference PhaseMassAccumulator
    (golden); the fixed-point histogram reproduces them to the quantum."""
    z = g("cdf.npz")
    nt, np_ = 36, 72
    s = oracle.hex_settings(**dict(C3, total=3000))
    mass = np.zeros(nt * np_)
    fw, bw = oracle.f64(), oracle.f64()
    assert oracle.lib().orc_hex_masses(oracle.C.byref(s), nt, np_, mass.ctypes.data, oracle.C.byref(fw),
                                       oracle.C.byref(bw)) == 0
    assert np.array_equal(mass, z["mass"])
    assert (fw.value, bw.value) == tuple(z["mass_poles"])
    bins, st = oracle.hex_hist(s, nt, np_)
    Q = oracle.quantum_bits(3000)
    k = np.array(C3["ki"]) / np.linalg.norm(C3["ki"])
    area = oracle.lib().orc_hex_projected_area(0.05, 0.2, oracle.vec3(k))
    fixed = oracle.bins_to_mass(bins, st.q_forward, st.q_backward, nt, np_) * (area / 3000) * 2.0 ** -Q
    # per-sample rounding 2^-(Q+1) of a unit ray, times the ray weight
    assert np.max(np.abs(fixed - mass)) <= st.samples * 2.0 ** -(Q + 1) * area / 3000


def test_status_sweep_matches_reference_semantics(oracle):
    """direction/aspect sweep of hex_trace_tests.cpp:253-281: every ray is
    accepted with max 1024 interactions and the energy balance closes."""
    rng = np.random.default_rng(5)
    for c in range(40):
        k = 2 * rng.random(3) - 1
        length = 0.2 + 5 * rng.random()
        s = oracle.hex_settings(radius_mm=1, length_mm=length, ki=tuple(k), n_int=1.31, tail=1e-12, seed=700 + c,
                                total=60, max_hits=1024)
        _, aud = oracle.hex_trace(s, 0, 60)
        assert np.all((aud["status"] == 0) | (aud["status"] == 1))
        assert np.max(np.abs(aud["balance"])) < 1e-10


@pytest.mark.skipif(not __import__("oracle_bind").ref_available(), reason="oracle/_ref not built")
def test_hex_trace_bitwise_vs_compiled_reference(oracle):
    rng = np.random.default_rng(9)
    for c in range(12):
        k = tuple(2 * rng.random(3) - 1)
        s = oracle.hex_settings(radius_mm=0.05 + rng.random(), length_mm=0.05 + 2 * rng.random(), ki=k,
                                n_int=1.2 + 0.2 * rng.random(), tail=1e-12, seed=100 + c, total=400, max_hits=20)
        a, aa = oracle.hex_trace(s, 0, 400)
        b, ba = oracle.ref_hex_trace(s, 0, 400)
        assert np.array_equal(a, b)
        assert np.array_equal(aa, ba)


@pytest.mark.skipif(not __import__("oracle_bind").ref_available(), reason="oracle/_ref not built")
def test_phase_bin_bitwise_vs_compiled_reference(oracle):
    rng = np.random.default_rng(13)
    k = (0.3, 0.9, -0.2)
    fr = oracle.make_frame(k)
    e = oracle.u_edges(450)
    d = rng.normal(size=(5000, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    d = np.array([np.frombuffer(np.array(v).tobytes()) for v in d])
    idx, pole, valid, sn = oracle.ref_phase_bin(450, 900, k, d)
    for t in range(d.shape[0]):
        b = oracle.phase_bin(e, 450, 900, fr, d[t])
        assert (b.valid, b.pole, b.snapped) == (valid[t], pole[t], sn[t])
        if b.valid and not b.pole:
            assert b.index == idx[t]


# ----------------------------------------------------------------- CDF
def test_cdf_golden_bitwise(oracle):
    z = g("cdf.npz")
    for name, (ont, onp) in (("full", (36, 72)), ("coarse", (18, 36))):
        rc, phi, theta, r = oracle.build_cdf(z["mass"], 36, 72, ont, onp)
        assert rc == 0
        assert np.array_equal(phi, z[f"{name}_phi"])
        assert np.array_equal(theta.reshape(-1), z[f"{name}_theta"].reshape(-1))
        rep = z[f"{name}_report"]
        got = [r.g_source, r.g_stored, r.coarsening_tv, r.cdf_mass, r.l1_error, r.lost_mass, r.max_cell_error,
               r.aggregation_error, r.lost_cells]
        assert np.array_equal(np.array(got), rep)
    rc, phi, theta, r = oracle.build_cdf(z["hdr_mass"], 48, 96, 24, 48)
    assert rc == 0 and np.array_equal(phi, z["hdr_phi"]) and np.array_equal(theta.reshape(-1),
                                                                           z["hdr_theta"].reshape(-1))


def test_cdf_reference_kats(oracle):
    """phase_tests.cpp:57-70 and :105-127."""
    nt, np_ = 64, 128
    e = oracle.u_edges(nt)
    m = np.repeat((4 * math.pi / np_) * np.diff(e), np_)
    rc, phi, theta, r = oracle.build_cdf(m, nt, np_)
    assert rc == 0 and abs(r.g_stored) < 1e-15
    assert np.max(np.abs(phi - np.arange(np_ + 1) / np_)) < 1e-14
    assert np.max(np.abs(theta - e[None, :])) < 2e-15
    i = np.arange(48)[:, None]
    j = np.arange(96)[None, :]
    hdr = np.where(j % 7 == 0, 10.0 ** (-12.0 * ((17 * i + 13 * j) % 97) / 96), 0.0).reshape(-1)
    rc, phi, theta, r = oracle.build_cdf(hdr, 48, 96, 24, 48)
    assert rc == 0 and abs(r.cdf_mass - 1) < 2e-15 and r.l1_error < 2e-14
    rc, phi, theta, r = oracle.build_cdf(hdr, 48, 96)
    e48 = oracle.u_edges(48)
    for col in range(96):
        if col % 7:
            assert np.array_equal(theta[col], e48)   # zero column -> u edges
    assert oracle.build_cdf(hdr, 48, 96, 24, 48, max_tv=0.0)[0] != 0
    assert oracle.build_cdf(hdr, 48, 96, 25, 48)[0] != 0
    assert oracle.build_cdf(np.zeros(48 * 96), 48, 96)[0] != 0
    neg = hdr.copy()
    neg[0] = -1
    assert oracle.build_cdf(neg, 48, 96)[0] != 0
    assert oracle.build_cdf(np.array([1, 1e-20, 1e-20, 1e-20]), 2, 2, max_lost=0.0)[0] != 0


@pytest.mark.skipif(not __import__("oracle_bind").ref_available(), reason="oracle/_ref not built")
def test_cdf_bitwise_vs_compiled_reference(oracle):
    rng = np.random.default_rng(3)
    for nt, np_, ont, onp in ((30, 60, 30, 60), (30, 60, 15, 20), (90, 180, 45, 90)):
        m = rng.random(nt * np_) ** 6 * (rng.random(nt * np_) < 0.7)
        rc, phi, theta, r = oracle.build_cdf(m, nt, np_, ont, onp)
        rc2, phi2, theta2, r2 = oracle.ref_build_cdf(m, nt, np_, (0, 1, 0), ont, onp)
        assert rc == rc2 == 0
        assert np.array_equal(phi, phi2) and np.array_equal(theta, theta2)
        assert r.g_stored == r2.g_stored and r.l1_error == r2.l1_error


# ----------------------------------------------------------------- drop tracer
WATER_550 = 1.3346825974540153


@pytest.mark.parametrize("max_hits", [1, 3, 9])
def test_drop_fp64_twin_matches_closed_form_sphere(oracle, max_hits):
    """The traced sphere (3D vectors, Fresnel-Jones splitting) against the
    closed-form single-scattering orders on the same impact parameters."""
    nt, grid = 180, 120
    for sphere in (True, False):   # False: the general Newton intersector with c = 0
        r = oracle.drop_record(sphere=sphere, alpha_deg=0.0, n_int=WATER_550, grid=grid, max_hits=max_hits)
        bins, st = oracle.drop_hist(r, nt, 8, precision=64)
        m = bins.reshape(nt, 8).sum(1).astype(float)
        m[0] += st.q_forward
        m[-1] += st.q_backward
        # the record carries n in single precision; the closed form takes the same value
        ref = oracle.sphere_theta_hist(grid, r.half_extent, float(np.float32(WATER_550)), max_hits, nt)
        Q = oracle.quantum_bits(grid * grid)
        assert abs(m.sum() * 2.0 ** -Q / ref.sum() - 1) < 1e-9
        tv = 0.5 * np.abs(m / m.sum() - ref / ref.sum()).sum()
        assert tv < 1e-6, (sphere, tv)


def test_drop_scaled_sphere_via_c0(oracle):
    """r = 1 + c0 is a sphere of radius 1+c0: the theta marginal equals the unit
    sphere's with impact parameters scaled by 1/(1+c0)."""
    nt, grid, c0 = 120, 96, 0.05
    r = oracle.drop_record(coeffs=(c0, 0, 0, 0, 0, 0, 0, 0), sphere=False, alpha_deg=30.0, n_int=WATER_550,
                           grid=grid, max_hits=4)
    bins, st = oracle.drop_hist(r, nt, 4, precision=64)
    m = bins.reshape(nt, 4).sum(1).astype(float)
    ref = oracle.sphere_theta_hist(grid, np.float32(r.half_extent / (1 + c0)), float(np.float32(WATER_550)), 4, nt)
    tv = 0.5 * np.abs(m / m.sum() - ref / ref.sum()).sum()
    assert tv < 2e-3, tv   # impact parameters 1/(1+c0) apart in float; bins move only at edges


def test_drop_fp32_mirror_vs_fp64_twin(oracle):
    coeffs = (-0.0458, -0.0335, -0.1211, -0.0227, 0, 0, 0, 0)
    for alpha in (0.0, 45.0, -80.0):
        r = oracle.drop_record(coeffs=coeffs, sphere=False, alpha_deg=alpha, n_int=WATER_550, grid=96, max_hits=9)
        b32, s32 = oracle.drop_hist(r, 90, 180, precision=32)
        b64, s64 = oracle.drop_hist(r, 90, 180, precision=64)
        assert s32.rays_hit == s64.rays_hit
        tv = 0.5 * np.abs(b32 / b32.sum() - b64 / b64.sum()).sum()
        assert tv < 1e-3, tv
        Q = oracle.quantum_bits(96 * 96)
        bal = (s32.q_escaped + s32.q_unresolved) / (s32.rays_hit * 2.0 ** Q) - 1
        assert abs(bal) < 1e-6


def test_drop_axis_incidence_is_azimuthally_uniform(oracle):
    """alpha = 90 deg looks along the drop's symmetry axis: the phi marginal of
    the traced record is uniform to sampling accuracy."""
    coeffs = (-0.0458, -0.0335, -0.1211, -0.0227, 0, 0, 0, 0)
    r = oracle.drop_record(coeffs=coeffs, sphere=False, alpha_deg=90.0, n_int=WATER_550, grid=160, max_hits=6)
    bins, st = oracle.drop_hist(r, 30, 36, precision=64)
    col = bins.reshape(30, 36).sum(0).astype(float)
    col /= col.mean()
    assert np.max(np.abs(col - 1)) < 0.02
