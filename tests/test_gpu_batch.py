"""GPU tests of batch scheduling: the histograms of an uneven mixed batch are
bitwise independent of the tile/workgroup schedule (integer fixed point), and
the autotuner returns a valid schedule."""
from __future__ import annotations

import numpy as np
import pytest

import ice_crystal_phase_amd as ip
from ice_crystal_phase_amd import autotune

pytestmark = pytest.mark.gpu


def test_schedule_invariance_and_autotune(ctx):
    print("case: uneven batch 24 records 36x72, schedules", flush=True)
    recs = autotune.uneven_records(24, seed=4, theta=36, phi=72, log2_rays=(12.0, 16.0))
    with ip.Batch(ctx, recs) as b:
        b.run(ip.RUN_TRACE)
        ref = [b.bins(i).copy() for i in range(len(recs))]
        for tile, blocks in ((256, 64), (4096, 2048), (1 << 16, 256)):
            b.set_schedule(tile, blocks)
            b.run(ip.RUN_TRACE | ip.RUN_GRAPH)
            for i in range(len(recs)):
                assert np.array_equal(b.bins(i), ref[i]), (tile, blocks, i)
    choice = autotune.tune(ctx, recs, shrink=4, repeats=1,
                           cands=[(0, 1024), (2048, 1024), (8192, 2048)])
    assert choice.ms > 0 and choice.blocks in (1024, 2048)
