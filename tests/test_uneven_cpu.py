"""CPU checks of the uneven-batch workload definition (no device calls)."""
from __future__ import annotations

import math

import numpy as np

import ice_crystal_phase_amd as ip
from ice_crystal_phase_amd import autotune


def test_uneven_records_mix_and_ranges():
    recs = autotune.uneven_records(60, seed=4, theta=36, phi=72)
    kinds = [r.shape for r in recs]
    assert kinds.count(ip.SHAPE_SPHERE) == kinds.count(ip.SHAPE_DROP) == kinds.count(ip.SHAPE_HEX) == 20
    for r in recs:
        n = ip.record_rays(r)
        assert 2 ** 19.9 <= n <= 2 ** 26.05
        assert 0.1 <= r.radius_mm <= 3.0 and 400 <= r.wavelength_nm <= 700
        if r.shape == ip.SHAPE_HEX:
            assert 0.2 - 1e-12 <= r.length_mm / (2 * r.radius_mm) <= 4 + 1e-12
            assert r.interior_index == 1.31
        else:
            assert r.grid_w == r.grid_h and abs(r.interior_index - 1.334) < 0.01


def test_shrink_keeps_the_mix():
    recs = autotune.uneven_records(12, seed=1, theta=36, phi=72)
    small = autotune.shrink_records(recs, 16)
    for a, b in zip(recs, small):
        assert a.shape == b.shape
        ra, rb = ip.record_rays(a), ip.record_rays(b)
        assert 8 <= ra / rb <= 32
