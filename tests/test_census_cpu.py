"""The drop kernel's operation census (bench.CENSUS, the numerator of its
roofline) is reproducible from committed code: the counting build of the oracle
(oracle/census.cpp) recomputes every constant on a subsample of the workload,
its event counts equal the oracle's own trace of the same rays, and its
counting rules give the closed-form count of a single known operation chain."""
from __future__ import annotations

import math
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT / "oracle"))

import bench  # noqa: E402
import census  # noqa: E402  (oracle/census.py)
import oracle_bind as ob  # noqa: E402


@pytest.fixture(scope="module")
def workload():
    recs, _, _ = bench.workload(192, 26, 900, 1800)
    return recs


def test_census_constants_are_reproduced(workload):
    got = census.census_of(workload, 900, 1800, rows=8, every=24)
    want = bench.CENSUS["drop"]
    for event in ("entry", "hit", "emit"):
        assert got[event] == pytest.approx(want[event], rel=0.03), (event, got[event], want[event])
    # misses cost 12 to ~400 operations depending on how far inside the bounding
    # circle they pass; their mean moves more with the sample (1 % of the total)
    assert got["miss"] == pytest.approx(want["miss"], rel=0.15)


def test_sphere_census_is_exact(workload):
    import ice_crystal_phase_amd as ip
    n = ip.water_optics(550.0).refractive_index
    rec = ip.sphere_record(n, 2048, 900, 1800, radius_mm=1.0, inclination_deg=0.0, max_internal_hits=9)
    got = census.census_of([rec], 900, 1800, rows=4)
    want = bench.CENSUS["sphere"]
    # closed-form entry and exit: the same operations for every ray and encounter
    for event in ("miss", "entry", "hit"):
        assert got[event] == want[event]


def test_counted_trace_sees_the_oracles_events(workload):
    rec = workload[37]
    r = ob.drop_record(coeffs=tuple(rec.shape_coefficients), sphere=False,
                       alpha_deg=math.degrees(math.atan2(-rec.incident[1], rec.incident[0])),
                       n_int=rec.interior_index, grid=rec.grid_w, max_hits=rec.max_internal_hits)
    first, count = 1000 * rec.grid_w + 3000, 5000
    c = ob.drop_census(r, 900, 1800, first, count)
    _, st = ob.drop_hist(r, 900, 1800, precision=32, qbits=ob.quantum_bits(rec.grid_w * rec.grid_h), first=first,
                         count=count)
    assert (c.entries, c.hits, c.emissions) == (st.rays_hit, st.internal_hits, st.samples)
    assert c.misses + c.entries == count


def test_binning_count_matches_its_operation_chain():
    """The counting rules give the closed-form length of the phase_bin chain.
    A direction on the grid's axis (a pole) stops after the frame projection
    (3 dot products, 15), its norm (5 + sqrt) and mu, s (2 divides + hypot): 24.
    A generic direction adds u (4 in the forward hemisphere, 5 in the backward
    one), two candidate edges (widths 2, tolerances 2, |u - edge| 1: 10),
    atan2, v (2), the column index (1), floor(index + 1/2) (2) and the column
    snap test (2): 46 or 47."""
    r = ob.drop_record(sphere=True, alpha_deg=0.0, n_int=1.33, grid=1, max_hits=3)
    c = ob.drop_census(r, 900, 1800, 0, 1)   # the one-ray grid: the centre ray, whose exits are all axial
    assert c.emissions >= 2 and c.ops_emit == 24 * c.emissions
    r = ob.drop_record(sphere=True, alpha_deg=0.0, n_int=1.33, grid=64, max_hits=3)
    c = ob.drop_census(r, 900, 1800, 0, 64 * 64)
    assert 45.5 <= c.ops_emit / c.emissions <= 47.0
