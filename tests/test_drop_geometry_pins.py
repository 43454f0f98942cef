"""Pins of the oracle's non-spherical drop geometry against the reference's own
drop code, compiled from its header-only sources into oracle/_ref/libref.so:

* RaindropShape::implicit_value / gradient (include/rainbow/raindrop_shape.hpp:85-100)
  against the oracle's surface and gradient, at points around the oblate-drop
  workload's shape (4 seeded Chebyshev terms) and the Table I 2 mm and 3 mm shapes;
* a central finite-difference check of the oracle's gradient against its surface;
* RaindropIntersector::intersect (include/rainbow/raindrop_intersection.hpp:337-382)
  against the oracle's entry and exit points and outward normals, fed identical
  ray lines;
* RaindropPathTracer::trace_vertex_unpolarized (include/rainbow/raindrop_paths.hpp:40-107)
  with three internal hits (families R, TT, TRT, TRRT) against the oracle's trace
  of the same vertex origins: status per family, outgoing directions, and powers
  (the reference's field Fresnel responses converted to flux with the textbook
  factor n_t cos_t / (n_i cos_i) at each transmission, from its own geometry).

The HIP drop kernel is bitwise equal to the oracle's FP32 mirror (GPU parity
tests), so these pins carry over to the kernel. Two checks show the comparisons
have power: an up/down flip of the drop (odd Chebyshev terms negated, what a
flipped mu convention would produce) and a sign error in the R'(mu) term of the
gradient both fail them by orders of magnitude."""
from __future__ import annotations

import numpy as np
import pytest

import oracle_bind as ob
from bench import c2_coefficients

pytestmark = pytest.mark.skipif(not ob.ref_available(), reason="oracle/_ref not built (no reference checkout)")

WATER_550 = 1.3346825974540153
ALPHAS = (-62.0, 0.0, 37.0, 85.0)


def _c2():
    return list(c2_coefficients()) + [0.0] * 4


def _table_i(radius_mm):
    rc, c, _, _ = ob.ref_drop_shape(radius_mm)
    assert rc == 0
    return [float(x) for x in c]


SHAPES = {"c2": _c2, "table_i_2mm": lambda: _table_i(2.0), "table_i_3mm": lambda: _table_i(3.0)}


def _record(coeffs, alpha=0.0, max_hits=3):
    return ob.drop_record(coeffs=coeffs, sphere=False, alpha_deg=alpha, n_int=WATER_550, grid=64, max_hits=max_hits)


def _points(n, seed, lo=0.6, hi=1.4):
    """Points at radii lo..hi on random directions, plus the poles and the equator."""
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    pts = d * rng.uniform(lo, hi, size=(n, 1))
    extra = np.array([[0, 1.1, 0], [0, -0.9, 0], [1.05, 0, 0], [0, 0, 0.95], [0.3, -0.8, 0.4]])
    return np.vstack([pts, extra])


def _unit(x):
    return x / np.linalg.norm(x, axis=-1, keepdims=True)


@pytest.mark.parametrize("shape", sorted(SHAPES))
def test_surface_and_gradient_equal_the_references(oracle, shape):
    r = _record(SHAPES[shape]())
    pts = _points(4000, seed=11)
    f, g = ob.drop_surface(r, pts, precision=64)
    fr, gr = ob.ref_drop_surface(r, pts, precision=64)
    # the same formulas; mu = -y/r is formed as -(y * (1/r)) in the oracle: an ulp apart
    assert np.max(np.abs(f - fr)) < 1e-14
    assert np.max(np.abs(g - gr)) < 1e-14
    # the FP32 mirror against the reference's float instantiation
    f32, g32 = ob.drop_surface(r, pts.astype(np.float32).astype(np.float64), precision=32)
    fr32, gr32 = ob.ref_drop_surface(r, pts.astype(np.float32).astype(np.float64), precision=32)
    assert np.max(np.abs(f32 - fr32)) < 2e-6
    assert np.max(np.abs(g32 - gr32)) < 2e-6


@pytest.mark.parametrize("shape", sorted(SHAPES))
def test_gradient_is_the_derivative_of_the_surface(oracle, shape):
    """Central differences of f in FP64 against the oracle's analytic gradient."""
    r = _record(SHAPES[shape]())
    pts = _points(500, seed=12, lo=0.7, hi=1.3)
    _, g = ob.drop_surface(r, pts)
    h = 1e-6
    fd = np.zeros_like(pts)
    for k in range(3):
        e = np.zeros(3)
        e[k] = h
        fp, _ = ob.drop_surface(r, pts + e)
        fm, _ = ob.drop_surface(r, pts - e)
        fd[:, k] = (fp - fm) / (2 * h)
    assert np.max(np.abs(fd - g)) < 1e-8


def test_pins_fail_for_a_flipped_drop_or_a_gradient_sign_error(oracle):
    """The comparisons above would catch the plausible mistakes: the drop turned
    upside down (a flipped mu convention) and a sign error in the R'(mu) term."""
    c = _c2()
    flipped = [(-x if n % 2 else x) for n, x in enumerate(c)]
    pts = _points(2000, seed=13, lo=0.8, hi=1.2)
    fr, gr = ob.ref_drop_surface(_record(c), pts)
    ff, _ = ob.drop_surface(_record(flipped), pts)
    assert np.max(np.abs(ff - fr)) > 1e-2
    # gradient = u - (g - u mu) R'/r; the sign error gives u + (...) = 2u - gradient
    _, g = ob.drop_surface(_record(c), pts)
    u = _unit(pts)
    assert np.max(np.abs((2 * u - g) - gr)) > 1e-2


def _line_origin(r, u, v):
    """The reference's incident line: grid point e0 u + e1 v moved 2R back along ki
    (include/rainbow/raindrop_paths.hpp:56-60), in float."""
    ki, e0, e1 = (np.array(x[:], np.float32) for x in (r.ki, r.e0, r.e1))
    o = np.float32(u) * e0 + np.float32(v) * e1 - np.float32(2.0 * r.outer) * ki
    return o.astype(np.float32), ki


@pytest.mark.parametrize("alpha", ALPHAS)
def test_entry_and_exit_points_equal_the_reference_intersector(oracle, alpha):
    r = _record(_c2(), alpha)
    rng = np.random.default_rng(int(alpha) + 100)
    e0 = np.array(r.e0[:], np.float64)
    hits = 0
    for _ in range(300):
        u, v = (rng.random(2) * 2 - 1) * r.half_extent
        hit, p, n = ob.drop_entry_point(r, u, v)
        o, ki = _line_origin(r, u, v)
        code, pr, nr, _, _ = ob.ref_drop_intersect(r, o, ki)
        assert (code == 1) == hit, (u, v, code)
        if not hit:
            continue
        hits += 1
        cosi = -np.float64(ki) @ n
        # the reference solves in FP32: positions agree to its resolution, scaled by
        # the conditioning of the line-surface intersection (1/cos incidence)
        assert np.max(np.abs(p - pr)) * cosi < 2e-6
        assert np.max(np.abs(n - nr)) * cosi < 2e-6
        # exit of an interior chord from the entry point
        d = _unit(-n + 0.4 * np.cross(n, e0))
        ok, q, nq = ob.drop_exit_point(r, p, d)
        code2, qr, nqr, _, _ = ob.ref_drop_intersect(r, p.astype(np.float32), d.astype(np.float32), 2.0 ** -17, True)
        assert ok and code2 == 1
        assert np.max(np.abs(q - qr)) < 2e-6
        assert np.max(np.abs(nq - nqr)) < 2e-6
    assert hits > 150


@pytest.mark.parametrize("alpha", ALPHAS)
@pytest.mark.parametrize("precision", [32, 64])
def test_path_families_equal_the_reference_trace(oracle, alpha, precision):
    """R, TT, TRT, TRRT of a 33 x 33 vertex grid: the oracle traces the
    reference's vertex origins (u, v computed in float as the reference does)."""
    r = _record(_c2(), alpha, max_hits=3)
    G = 33
    status, d, p, c0, c1 = ob.ref_drop_trace_vertices(r, G, G)
    ki = np.array(r.ki[:], np.float64)
    n = float(r.n_int)
    h = np.float32(r.half_extent)
    counts = np.zeros(4, int)
    for idx in range(G * G):
        ix, iy = idx % G, idx // G
        u = (np.float32(2.0) * np.float32(ix) / np.float32(G - 1) - np.float32(1.0)) * h
        v = (np.float32(2.0) * np.float32(iy) / np.float32(G - 1) - np.float32(1.0)) * h
        samples, audit = ob.drop_trace_origin(r, float(u), float(v), precision)
        family = {0 if s["kind"] == 0 else int(s["reflections"]): s for s in samples}
        assert (audit["status"] == 8) == (status[0, idx] == 2), (idx, audit["status"], status[0, idx])  # miss
        for k in range(4):
            valid = status[k, idx] == 1
            assert valid == (k in family), (alpha, idx, k, status[k, idx])
            if not valid:
                continue
            counts[k] += 1
            assert np.max(np.abs(d[k, idx] - family[k]["dir"])) < 5e-5
            field_power = 0.5 * (np.sum(c0[k, idx].astype(np.float64) ** 2) + np.sum(c1[k, idx].astype(np.float64) ** 2))
            if k > 0:
                # field -> flux: n_t cos_t / (n_i cos_i) at entry and at exit
                _, g1 = ob.ref_drop_surface(r, p[0, idx][None].astype(np.float64))
                n1 = _unit(g1[0])
                ci1 = -ki @ n1
                ct1 = np.sqrt(1 - (1 - ci1 * ci1) / (n * n))
                _, g2 = ob.ref_drop_surface(r, p[k, idx][None].astype(np.float64))
                n2 = _unit(g2[0])
                ct2 = d[k, idx].astype(np.float64) @ n2
                ci2 = np.sqrt(1 - (1 - ct2 * ct2) / (n * n))
                field_power *= (ct1 * ct2) / (ci1 * ci2)
            assert abs(family[k]["power"] - field_power) <= 1e-3 * field_power + 1e-6, (alpha, idx, k)
    assert np.all(counts > 400), counts
