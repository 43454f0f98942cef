"""The drop kernel's approximate cell decision (bin_fast_approx in
csrc/phase_bin.hpp) is only allowed to bin an emission where the FP64 phase_bin
rule gives the same cell. Its guarantee rests on an error bound; these tests
pin the pieces of that bound and the decision itself on the CPU:

  * the atan polynomial (kAtanPoly, read from the header) is within 1.5e-7 rad
    of atan on [0, 1] when evaluated with FP32 FMA steps;
  * atan2_approx, with the hardware reciprocal's 1-ulp error taken at both
    extremes, is within 7e-7 rad of atan2;
  * an FP32 emulation of bin_fast_approx, on directions crowded at cell edges,
    the poles and the phi seam, never returns a cell that differs from the
    oracle's FP64 phase_bin (tests/oracle_bind.py, the rule of
    hex_phase/include/ice_crystal/phase_math.hpp:37-81).
"""
from __future__ import annotations

import re
from pathlib import Path

import numpy as np

import oracle_bind as ob

HEADER = Path(__file__).resolve().parent.parent / "ice_crystal_phase_amd" / "csrc" / "phase_bin.hpp"
F32 = np.float32


def coefficients() -> np.ndarray:
    text = HEADER.read_text()
    body = re.search(r"kAtanPoly\[8\]\s*=\s*\{([^}]*)\}", text).group(1)
    vals = [float(v.strip().rstrip("f")) for v in body.split(",") if v.strip()]
    assert len(vals) == 8
    return np.array(vals, dtype=np.float32)


def fma32(a, b, c):
    # one rounding: the product of two FP32 values is exact in FP64
    return (a.astype(np.float64) * b.astype(np.float64) + c.astype(np.float64)).astype(np.float32)


def atan01(t: np.ndarray, c: np.ndarray) -> np.ndarray:
    t2 = (t * t).astype(F32)
    p = np.full_like(t, c[7])
    for k in range(6, -1, -1):
        p = fma32(p, t2, np.full_like(t, c[k]))
    return (t * p).astype(F32)


def rcp(x: np.ndarray, ulps: int) -> np.ndarray:
    r = (F32(1) / x).astype(F32)
    for _ in range(abs(ulps)):
        r = np.nextafter(r, F32(np.inf) if ulps > 0 else F32(0)).astype(F32)
    return r


def atan2_approx(y: np.ndarray, x: np.ndarray, c: np.ndarray, ulps: int) -> np.ndarray:
    ax, ay = np.abs(x), np.abs(y)
    with np.errstate(divide="ignore", invalid="ignore"):
        t = (np.minimum(ax, ay) * rcp(np.maximum(ax, ay), ulps)).astype(F32)
    a = atan01(t, c)
    a = np.where(ay > ax, (F32(1.57079637) - a).astype(F32), a)
    a = np.where(x < 0, (F32(3.14159274) - a).astype(F32), a)
    return np.copysign(a, y).astype(F32)


def test_atan_polynomial_bound():
    c = coefficients()
    t = np.linspace(0, 1, 2_000_001, dtype=np.float32)
    err = np.abs(atan01(t, c).astype(np.float64) - np.arctan(t.astype(np.float64)))
    assert err.max() < 1.5e-7, err.max()


def test_atan2_approx_bound():
    c = coefficients()
    rng = np.random.default_rng(7)
    ang = rng.uniform(-np.pi, np.pi, 400_000)
    rad = 10.0 ** rng.uniform(-6, 1, ang.size)
    y = (rad * np.sin(ang)).astype(F32)
    x = (rad * np.cos(ang)).astype(F32)
    exact = np.arctan2(y.astype(np.float64), x.astype(np.float64))
    for ulps in (-1, 0, 1):
        got = atan2_approx(y, x, c, ulps).astype(np.float64)
        d = np.abs(got - exact)
        d = np.minimum(d, 2 * np.pi - d)  # +-pi at the seam
        assert d.max() < 7e-7, (ulps, d.max())


def fast_cell(nt: int, np_: int, frame, d: np.ndarray, c: np.ndarray, ulps: int):
    """FP32 emulation of bin_fast_approx; -1 where it leaves the cell open."""
    e0, e1, ki = (np.array(v[:]) for v in (frame.e0, frame.e1, frame.ki))
    x, y, z = (d @ e0).astype(F32), (d @ e1).astype(F32), (d @ ki).astype(F32)
    rho = np.sqrt(fma32(y, y, (x * x).astype(F32)).astype(np.float64)).astype(F32)
    rho = np.nextafter(rho, F32(np.inf) if ulps > 0 else F32(0)).astype(F32) if ulps else rho  # 1-ulp sqrt
    ri = (atan2_approx(rho, z, c, ulps) * F32(F32(nt) * F32(1.0 / np.pi))).astype(F32)
    ci = ((atan2_approx(y, x, c, ulps) + F32(np.pi)).astype(F32) * F32(F32(np_) * F32(0.5 / np.pi))).astype(F32)
    rfl, cfl = np.floor(ri), np.floor(ci)
    rfr, cfr = (ri - rfl).astype(F32), (ci - cfl).astype(F32)
    m = F32(1e-2)
    ok = (rfr > m) & (rfr < 1 - m) & (cfr > m) & (cfr < 1 - m) & (rfl >= 0) & (rfl < nt) & (cfl >= 0) & (cfl < np_)
    return np.where(ok, rfl.astype(np.int64) * np_ + cfl.astype(np.int64), -1)


def test_fast_decision_agrees_with_the_fp64_rule():
    c = coefficients()
    rng = np.random.default_rng(11)
    for nt, np_ in ((900, 1800), (36, 72), (450, 900)):
        frame = ob.drop_frame(np.array([0.93, -0.31, 0.2]) / np.linalg.norm([0.93, -0.31, 0.2]))
        edges = ob.u_edges(nt)
        e0, e1, ki = (np.array(v[:]) for v in (frame.e0, frame.e1, frame.ki))
        # theta, phi near cell edges (within +-0.03 cells), uniform ones, the poles and the seam
        n = 6000
        ti = rng.integers(0, nt + 1, n) + rng.uniform(-0.03, 0.03, n)
        pj = rng.integers(0, np_ + 1, n) + rng.uniform(-0.03, 0.03, n)
        th = np.concatenate([ti * np.pi / nt, rng.uniform(0, np.pi, n), rng.uniform(0, 1e-3, 200),
                             np.pi - rng.uniform(0, 1e-3, 200)])
        ph = np.concatenate([pj * 2 * np.pi / np_ - np.pi, rng.uniform(-np.pi, np.pi, n),
                             rng.uniform(-np.pi, np.pi, 400)])
        th = np.clip(th, 0, np.pi)
        d = (np.sin(th)[:, None] * (np.cos(ph)[:, None] * e0 + np.sin(ph)[:, None] * e1)
             + np.cos(th)[:, None] * ki)
        d = d.astype(np.float32).astype(np.float64)  # the kernel bins FP32 directions
        # the drop rule divides the projections by their norm; the oracle's
        # phase_bin wants a unit vector, and away from edges both give one cell
        unit = d / np.linalg.norm(d, axis=1)[:, None]
        exact = np.array([ob.phase_bin(edges, nt, np_, frame, v) for v in unit], dtype=object)
        for ulps in (-1, 0, 1):
            fast = fast_cell(nt, np_, frame, d, c, ulps)
            taken = fast >= 0
            assert 0.5 < taken.mean() < 1.0
            for k in np.flatnonzero(taken):
                b = exact[k]
                assert b.valid and b.pole == 0 and not b.snapped and int(b.index) == int(fast[k]), (nt, np_, k)
