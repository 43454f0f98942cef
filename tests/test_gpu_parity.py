"""GPU parity: the HIP path through the C-ABI against the CPU oracle.

Bar: integer histograms and pole/audit counters bitwise equal; CDF arrays
bitwise equal for the same cell masses; report scalars within 1e-12.
"""
import math

import numpy as np
import pytest

import ice_crystal_phase_amd as ip

pytestmark = pytest.mark.gpu

HEX_C3 = dict(radius_mm=0.05, length_mm=0.2, n_int=1.31, incident=(1.0, -0.4, 0.3), seed=12345,
              max_internal_hits=12, tail=1e-12)
WATER_550 = 1.3346825974540153   # IAPWS R9-97 at 550 nm, 20 C, 101325 Pa (reference evaluate_water_optics)
OBLATE = (-0.0458, -0.0335, -0.1211, -0.0227)   # four Chebyshev terms, axis ratio ~0.85


def stats_tuple(st):
    return (st.q_forward, st.q_backward, st.samples, st.invalid, st.snapped, tuple(st.status[:8]),
            st.q_escaped, st.q_unresolved, st.internal_hits)


def oracle_stats_tuple(st):
    return (st.q_forward, st.q_backward, st.samples, st.invalid, st.snapped, tuple(st.status[:8]),
            st.q_escaped, st.q_unresolved, st.internal_hits)


def run_one(ctx, rec, flags=ip.RUN_TRACE | ip.RUN_CDF, policy=None):
    b = ip.Batch(ctx, [rec], policy)
    b.run(flags)
    return b


def test_abi_version(ctx):
    assert ip.lib().iph_abi_version() == ip.icephase.ABI_VERSION


def test_bin_directions_match_oracle(ctx, oracle):
    rng = np.random.default_rng(7)
    nt, np_ = 900, 1800
    edges = oracle.u_edges(nt)
    for k in [(1.0, -0.4, 0.3), (1.0, 0.0, 0.0), (0.0, 0.0, 1.0)]:
        fr = oracle.make_frame(k)
        frame = np.array(list(fr.ki) + list(fr.e0) + list(fr.e1))
        ki, e0, e1 = (np.array(v) for v in (fr.ki, fr.e0, fr.e1))
        dirs = [rng.normal(size=3) for _ in range(20000)]
        # cell centres, cell edges (exact and +-1 ulp), poles, the seam
        for i in range(0, nt, 37):
            for j in range(0, np_, 53):
                for du, dv in ((0.5, 0.5), (0.0, 0.5), (0.5, 0.0), (0.0, 0.0)):
                    u = edges[i] + du * (edges[i + 1] - edges[i])
                    mu = 1 - 2 * u
                    st = 2 * math.sqrt(max(u * (1 - u), 0.0))
                    phi = -math.pi + (j + dv) * 2 * math.pi / np_
                    d = ki * mu + e0 * (st * math.cos(phi)) + e1 * (st * math.sin(phi))
                    dirs.append(d)
                    dirs.append(np.nextafter(d, d + 1))
        dirs += [ki, -ki, -e0, e0, np.nextafter(ki, ki + 1)]
        D = []
        for d in dirs:
            n = np.zeros(3)
            ok = oracle.lib().orc_normalize3(oracle.vec3(d), n.ctypes.data_as(oracle.P(oracle.f64)))
            if ok == 0:
                D.append(n)
        D = np.array(D)
        idx, kind = ctx.bin_directions(nt, np_, frame, D)
        for t in range(D.shape[0]):
            b = oracle.phase_bin(edges, nt, np_, fr, D[t])
            want_kind = 2 if not b.valid else (b.pole if b.pole else 0)
            assert kind[t] == want_kind, (t, D[t])
            if want_kind == 0:
                assert idx[t] == b.index, (t, D[t], idx[t], b.index)


@pytest.mark.parametrize("rays,nt,np_", [(1 << 14, 900, 1800), (1 << 15, 36, 72)])
def test_hex_histogram_bitexact(ctx, oracle, rays, nt, np_):
    rec = ip.hex_record(rays=rays, theta=nt, phi=np_, **HEX_C3)
    b = run_one(ctx, rec, ip.RUN_TRACE)
    s = oracle.hex_settings(radius_mm=0.05, length_mm=0.2, ki=(1.0, -0.4, 0.3), n_int=1.31, tail=1e-12,
                            seed=12345, total=rays, max_hits=12)
    want, st = oracle.hex_hist(s, nt, np_)
    got = b.bins()
    assert np.array_equal(got, want), int(np.count_nonzero(got != want))
    assert stats_tuple(b.stats()) == oracle_stats_tuple(st)


def test_hex_histogram_offset_rays_bitexact(ctx, oracle):
    """Tile cuts are invisible: rays 0..N traced as one record equal the oracle
    traced in two halves (batch-partition invariance, hex_trace_tests.cpp:212-219)."""
    rays, nt, np_ = 50000, 90, 180
    rec = ip.hex_record(rays=rays, theta=nt, phi=np_, **HEX_C3)
    b = ip.Batch(ctx, [rec])
    b.set_schedule(tile_rays=777, blocks=37)
    b.run(ip.RUN_TRACE)
    s = oracle.hex_settings(radius_mm=0.05, length_mm=0.2, ki=(1.0, -0.4, 0.3), n_int=1.31, tail=1e-12,
                            seed=12345, total=rays, max_hits=12)
    a1, _ = oracle.hex_hist(s, nt, np_, first=0, count=20000)
    a2, _ = oracle.hex_hist(s, nt, np_, first=20000, count=rays - 20000)
    assert np.array_equal(b.bins(), a1 + a2)


@pytest.mark.parametrize("alpha,grid", [(0.0, 256), (37.0, 200)])
def test_sphere_histogram_bitexact(ctx, oracle, alpha, grid):
    nt, np_ = 900, 1800
    rec = ip.sphere_record(WATER_550, grid, nt, np_, radius_mm=1.0, inclination_deg=alpha, max_internal_hits=9)
    b = run_one(ctx, rec, ip.RUN_TRACE)
    r = oracle.drop_record(sphere=True, alpha_deg=alpha, n_int=WATER_550, grid=grid, max_hits=9)
    want, st = oracle.drop_hist(r, nt, np_, precision=32)
    got = b.bins()
    assert np.array_equal(got, want), int(np.count_nonzero(got != want))
    g = b.stats()
    assert stats_tuple(g) == oracle_stats_tuple(st)
    assert g.status[8] == grid * grid - st.rays_hit


@pytest.mark.parametrize("alpha", [0.0, -62.0, 85.0])
def test_oblate_drop_histogram_bitexact(ctx, oracle, alpha):
    nt, np_, grid = 900, 1800, 160
    rec = ip.drop_record(OBLATE, WATER_550, grid, nt, np_, radius_mm=2.0, inclination_deg=alpha,
                         max_internal_hits=9)
    b = run_one(ctx, rec, ip.RUN_TRACE)
    r = oracle.drop_record(coeffs=OBLATE + (0,) * 4, sphere=False, alpha_deg=alpha, n_int=WATER_550, grid=grid,
                           max_hits=9)
    want, st = oracle.drop_hist(r, nt, np_, precision=32)
    got = b.bins()
    assert np.array_equal(got, want), int(np.count_nonzero(got != want))
    assert stats_tuple(b.stats()) == oracle_stats_tuple(st)


def test_cdf_from_gpu_histogram_bitexact(ctx, oracle):
    nt, np_ = 900, 1800
    rec = ip.hex_record(rays=1 << 15, theta=nt, phi=np_, **HEX_C3)
    b = run_one(ctx, rec)
    bins = b.bins()
    st = b.stats()
    phi, theta, rep = b.cdf()
    mass = oracle.bins_to_mass(bins, st.q_forward, st.q_backward, nt, np_)
    rc, ophi, otheta, orep = oracle.build_cdf(mass, nt, np_)
    assert rc == 0
    assert np.array_equal(phi, ophi)
    assert np.array_equal(theta, otheta)
    assert rep.status == 0
    for f in ("g_source", "g_stored", "cdf_mass", "l1_error", "lost_mass", "max_cell_error"):
        assert abs(getattr(rep, f) - getattr(orep, f)) <= 1e-12, f
    assert rep.lost_cells == orep.lost_cells


def test_build_cdf_sparse_hdr_and_coarsening(ctx, oracle):
    """Sparse high-dynamic-range masses (hex_phase/tests/phase_tests.cpp:105-127)."""
    nt, np_ = 48, 96
    i = np.arange(nt)[:, None]
    j = np.arange(np_)[None, :]
    mass = np.where(j % 7 == 0, 10.0 ** (-12.0 * ((17 * i + 13 * j) % 97) / 96), 0.0).reshape(-1)
    for ont, onp in ((nt, np_), (24, 48)):
        phi, theta, rep, rc = ctx.build_cdf(mass, nt, np_, ont, onp)
        orc, ophi, otheta, orep = oracle.build_cdf(mass, nt, np_, ont, onp)
        assert rc == 0 and orc == 0
        assert np.array_equal(phi, ophi)
        assert np.array_equal(theta, otheta)
        assert abs(rep.coarsening_tv - orep.coarsening_tv) <= 1e-12
        assert rep.l1_error < 2e-14


def _cdf_masses(nt, np_, seed):
    rng = np.random.default_rng(seed)
    m = rng.random((nt, np_)) ** 6
    m[:, np_ // 3] = 0.0                # an empty column takes the u edges
    m[nt // 2, :] *= 1e4                # a bright row
    m[:, -1] = 0.0
    m[nt // 4, -1] = 1e-20              # a column too faint to move the marginal: d = 0, its mass lost
    m[rng.random((nt, np_)) < 0.2] = 0  # empty cells inside columns: dtheta = 0
    return m.reshape(-1)


@pytest.mark.parametrize("nt,np_", [(1024, 45), (900, 7), (3, 13), (1025, 24), (1100, 40)],
                         ids=["1024x45", "900x7", "3x13", "1025x24", "1100x40"])
def test_cdf_tiles_empty_and_lost_columns_match_the_oracle(ctx, oracle, nt, np_):
    """Row counts on and off the theta pass's 32-row tiles, ragged 64-column
    strips, an empty column (u edges), cells absorbed by the ordered prefix
    (dtheta = 0) and a column too faint to move the marginal (dphi = 0): the
    oracle's CDFs bitwise, its report within 1e-12 and its lost-cell count."""
    print(f"case: CDF {nt}x{np_}", flush=True)
    mass = _cdf_masses(nt, np_, seed=nt * 31 + np_)
    phi, theta, rep, rc = ctx.build_cdf(mass, nt, np_)
    orc, ophi, otheta, orep = oracle.build_cdf(mass, nt, np_)
    assert rc == 0 and orc == 0
    assert np.array_equal(phi, ophi)
    assert np.array_equal(theta, otheta)
    for f in ("g_source", "g_stored", "cdf_mass", "l1_error", "lost_mass", "max_cell_error"):
        assert abs(getattr(rep, f) - getattr(orep, f)) <= 1e-12, f
    assert rep.lost_cells == orep.lost_cells and rep.lost_cells > 0
    assert rep.lost_mass > 0


def test_build_cdf_rejections(ctx):
    nt, np_ = 8, 16
    with pytest.raises(ip.AuditError):
        ctx.build_cdf(np.zeros(nt * np_), nt, np_)
    bad = np.ones(nt * np_)
    bad[3] = -1
    with pytest.raises(ip.AuditError):
        ctx.build_cdf(bad, nt, np_)
    with pytest.raises(ip.IcePhaseError):
        ctx.build_cdf(np.ones(nt * np_), nt, np_, 3, 16)
    hdr = np.array([1, 1e-20, 1e-20, 1e-20])
    with pytest.raises(ip.AuditError):
        ctx.build_cdf(hdr, 2, 2, max_lost=0.0)


def test_graph_replays_and_repeats_are_bitwise_identical(ctx):
    recs = [ip.sphere_record(WATER_550, 128, 90, 180), ip.hex_record(rays=20000, theta=90, phi=180, **HEX_C3),
            ip.drop_record(OBLATE, WATER_550, 96, 90, 180, inclination_deg=20.0)]
    b = ip.Batch(ctx, recs)
    b.run()
    ref = [b.bins(k).copy() for k in range(3)]
    ref_cdf = [b.cdf(k)[1].copy() for k in range(3)]
    for rep in range(3):
        b.run(ip.RUN_TRACE | ip.RUN_CDF | ip.RUN_GRAPH)
        for k in range(3):
            assert np.array_equal(b.bins(k), ref[k]), (rep, k)
            assert np.array_equal(b.cdf(k)[1], ref_cdf[k]), (rep, k)
        b.run()  # eager run between replays
        for k in range(3):
            assert np.array_equal(b.bins(k), ref[k])


def test_mixed_batch_equals_single_records(ctx):
    recs = [ip.sphere_record(WATER_550, 100, 60, 120), ip.hex_record(rays=7000, theta=60, phi=120, **HEX_C3),
            ip.drop_record(OBLATE, 1.33, 90, 60, 120, inclination_deg=-45.0)]
    b = ip.Batch(ctx, recs)
    b.set_schedule(tile_rays=1000)
    b.run()
    for k, r in enumerate(recs):
        single = ip.Batch(ctx, [r])
        single.run()
        assert np.array_equal(b.bins(k), single.bins(0))
        assert np.array_equal(b.cdf(k)[1], single.cdf(0)[1])


def test_full_size_sphere_record_properties(ctx, oracle):
    """Config-1 size (2^22 rays, 900x1800): audit closure, CDF contract, and the
    closed-form single-scattering sphere (orders p=0..9) at the size."""
    nt, np_, grid = 900, 1800, 2048
    rec = ip.sphere_record(WATER_550, grid, nt, np_, radius_mm=1.0, max_internal_hits=9)
    b = run_one(ctx, rec)
    st = b.stats()
    hit = grid * grid - st.status[8]
    Q = st.qbits
    # every hitting ray delivers its unit power as escaped + unresolved quanta; FP32
    # Fresnel powers conserve a unit ray to ~1e-7 relative, quanta round by <= 1
    total = st.q_escaped + st.q_unresolved
    assert abs(total / (hit * float(1 << Q)) - 1) < 1e-6
    assert int(np.sum(b.bins(), dtype=np.uint64)) + st.q_forward + st.q_backward == st.q_escaped
    phi, theta, rep = b.cdf()
    assert rep.status == 0 and abs(rep.cdf_mass - 1) < 1e-12
    assert phi[0] == 0 and phi[-1] == 1 and np.all(np.diff(phi) >= 0)
    assert np.all(theta[:, 0] == 0) and np.all(theta[:, -1] == 1) and np.all(np.diff(theta, axis=1) >= 0)
    # axisymmetric at alpha=0: the phi marginal is uniform once the square
    # incident lattice's azimuthal moire (2048^2 rays vs 1800 columns) is averaged
    # over 10-degree groups
    dphi = np.diff(phi).reshape(36, 50).sum(1) * 36
    assert np.max(np.abs(dphi - 1)) < 0.01, np.max(np.abs(dphi - 1))
    # closed-form theta distribution of the same 2^22 impact parameters (Snell,
    # Fresnel, deviation D_p per order; no ray-surface intersection involved):
    # only FP32 rounding at cell edges may move a sample to a neighbour row.
    m = b.bins().reshape(nt, np_).sum(axis=1).astype(np.float64)
    m[0] += st.q_forward
    m[-1] += st.q_backward
    m /= m.sum()
    ref = oracle.sphere_theta_hist(grid, np.float32(1.0 * 1.01), WATER_550, 9, nt)
    ref /= ref.sum()
    tv = 0.5 * np.abs(m - ref).sum()
    assert tv < 1e-4, tv
