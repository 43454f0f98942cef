"""Schedule autotuner of the persistent trace launch (tile length in rays and
workgroups per kind), and the uneven-batch workload it is tuned on.

The trace kernels pull tiles of consecutive rays from an atomic queue, so the
schedule trades per-tile overhead (short tiles) against tail imbalance (long
tiles, few records) and occupancy against queue contention (workgroups). The
tuner times the trace stage of a proxy batch - the same records with their
ray counts divided by `shrink` - for each candidate, then applies the fastest
schedule, scaled back, to the real batch. The choice is returned with every
timing so it can be recorded beside the measurement.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import icephase as icp


@dataclass
class Choice:
    tile_rays: int
    blocks: int
    ms: float


@dataclass
class UnevenSpec:
    """One record of the uneven batch: kind 0 sphere, 1 Table I drop, 2 hex prism."""
    kind: int
    radius_mm: float
    wavelength_nm: float
    alpha_deg: float
    rays: int
    aspect: float
    seed: int


def uneven_specs(n: int = 512, seed: int = 4, log2_rays=(20.0, 26.0)) -> list[UnevenSpec]:
    """The uneven batch's draws: n seeded records, one third spheres, one third
    oblate drops (Table I shape of the radius), one third hexagonal prisms (aspect
    L/(2R) log-uniform in [0.2, 4], n = 1.31); radius U[0.1, 3] mm, wavelength
    U[400, 700] nm, inclination U[-90, 90] deg, rays floor(2^U[20, 26])."""
    rng = np.random.default_rng(seed)
    kind = rng.permutation(np.arange(n) % 3)
    r = rng.uniform(0.1, 3.0, n)
    lam = rng.uniform(400.0, 700.0, n)
    alpha = rng.uniform(-90.0, 90.0, n)
    rays = np.floor(2.0 ** rng.uniform(*log2_rays, n)).astype(np.int64)
    aspect = np.exp(rng.uniform(math.log(0.2), math.log(4.0), n))
    return [UnevenSpec(int(kind[k]), float(r[k]), float(lam[k]), float(alpha[k]), int(rays[k]), float(aspect[k]), k)
            for k in range(n)]


def spec_record(s: UnevenSpec, theta: int = 900, phi: int = 1800):
    """The library record of one spec: hex prisms trace `rays` random entries;
    drops and spheres the nearest square grid (IAPWS water index at the FP32
    wavelength), at most 9 internal hits; hex at most 12, tail 1e-12."""
    if s.kind == 2:
        a = math.radians(s.alpha_deg)
        return icp.hex_record(s.radius_mm, 2 * s.radius_mm * s.aspect, 1.31, s.rays,
                              incident=(math.cos(a), -math.sin(a), 0.0), seed=s.seed, theta=theta, phi=phi,
                              max_internal_hits=12, tail=1e-12, wavelength_nm=s.wavelength_nm)
    grid = max(1, int(round(math.sqrt(float(s.rays)))))
    nw = icp.water_optics(float(np.float32(s.wavelength_nm))).refractive_index
    if s.kind == 0:
        return icp.sphere_record(nw, grid, theta, phi, radius_mm=s.radius_mm, inclination_deg=s.alpha_deg,
                                 max_internal_hits=9, wavelength_nm=s.wavelength_nm)
    c = icp.drop_shape(float(np.float32(s.radius_mm)))
    return icp.drop_record(c, nw, grid, theta, phi, radius_mm=s.radius_mm, inclination_deg=s.alpha_deg,
                           max_internal_hits=9, wavelength_nm=s.wavelength_nm)


def uneven_records(n: int = 512, seed: int = 4, theta: int = 900, phi: int = 1800, log2_rays=(20.0, 26.0)):
    """The uneven batch (see uneven_specs) as library records."""
    return [spec_record(s, theta, phi) for s in uneven_specs(n, seed, log2_rays)]


def shrink_records(recs, shrink: int):
    """Copies with ray counts divided by `shrink` (drops: grid sides by sqrt)."""
    out = []
    side = max(1, int(round(math.sqrt(shrink))))
    for r in recs:
        c = icp.Record.from_buffer_copy(r)
        if c.shape == icp.SHAPE_HEX:
            c.rays = max(1, int(c.rays) // shrink)
        else:
            c.grid_w = max(1, int(c.grid_w) // side)
            c.grid_h = max(1, int(c.grid_h) // side)
        out.append(c)
    return out


def candidates(cus: int = 256):
    tiles = (0, 2048, 8192, 32768, 131072)
    blocks = (cus * 4, cus * 8, cus * 16)
    return [(t, b) for t in tiles for b in blocks]


def tune(ctx: icp.Context, recs, shrink: int = 16, repeats: int = 2, cands=None, log=None) -> Choice:
    """Fastest (tile_rays, blocks) for the trace stage of `recs`, timed on the
    proxy batch. tile_rays 0 keeps the library's per-record default."""
    proxy = shrink_records(recs, shrink)
    best = None
    with icp.Batch(ctx, proxy) as b:
        b.run(icp.RUN_TRACE | icp.RUN_NO_AUDIT)  # warm
        for t, blocks in (cands or candidates()):
            tp = 0 if t == 0 else max(256, t // shrink)
            b.set_schedule(tp, blocks)
            ms = []
            for _ in range(repeats):
                b.run(icp.RUN_TRACE | icp.RUN_GRAPH | icp.RUN_NO_AUDIT)
                ms.append(b.timing()[0])
            m = min(ms)
            if log:
                log(f"tile {t} blocks {blocks}: {m:.3f} ms")
            if best is None or m < best.ms:
                best = Choice(t, blocks, m)
    return best
