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


def test_schedules_that_wrap_the_tile_queue_are_refused(ctx):
    """The persistent launch hands out tiles from a 32-bit queue counter and
    counts a tile's rays in 32 bits: schedules whose tiles (plus one exit ticket
    per workgroup) exceed 2^32 - 1, or whose tiles exceed 2^31 rays, are refused
    before anything is launched, and the batch keeps its previous schedule."""
    print("case: tile-queue limits, 64 hex records of 2^26 rays (not run)", flush=True)
    big = [ip.hex_record(0.05, 0.2, 1.31, 1 << 26, seed=k, theta=8, phi=16) for k in range(64)]
    with ip.Batch(ctx, big) as b:
        with pytest.raises(ip.IcePhaseError, match="32-bit tile queue"):
            b.set_schedule(1, 0)            # 2^32 tiles
        with pytest.raises(ip.IcePhaseError, match="32-bit tile queue"):
            b.set_schedule(2, 1 << 31)      # 2^31 tiles + 2^31 exit tickets
        b.set_schedule(4, 1024)             # 2^30 tiles + 1024: fits
        with pytest.raises(ip.IcePhaseError, match="at most 2\\^31"):
            b.set_schedule((1 << 31) + 1, 0)
    print("case: tile-queue limits, small batch runs after a refusal", flush=True)
    small = [ip.hex_record(0.05, 0.2, 1.31, 4096, seed=k, theta=18, phi=36) for k in range(2)]
    with ip.Batch(ctx, small) as b:
        b.run(ip.RUN_TRACE)
        ref = [b.bins(i).copy() for i in range(2)]
        with pytest.raises(ip.IcePhaseError):
            b.set_schedule((1 << 31) + 1, 0)
        b.run(ip.RUN_TRACE)
        for i in range(2):
            assert np.array_equal(b.bins(i), ref[i])


def test_hex_cell_table_and_direct_atomics_give_the_same_histogram(ctx, oracle):
    """The hex kernels add their quanta into a workgroup LDS cell table flushed
    once per tile, or (switched off) every quantum straight to HBM, the path a
    cell that finds no slot takes: both give the oracle's histogram bit for bit,
    eagerly and from the captured graph, and the direct kernel counts one HBM
    atomic per binned emission."""
    print("case: hex column 2^15 rays 900x1800, cell table on / off", flush=True)
    nt, np_, rays = 900, 1800, 1 << 15
    rec = ip.hex_record(0.05, 0.2, 1.31, rays, incident=(1.0, -0.4, 0.3), seed=12345, theta=nt, phi=np_,
                        max_internal_hits=12, tail=1e-12)
    s = oracle.hex_settings(radius_mm=0.05, length_mm=0.2, ki=(1.0, -0.4, 0.3), n_int=1.31, tail=1e-12,
                            seed=12345, total=rays, max_hits=12)
    want, _ = oracle.hex_hist(s, nt, np_)
    atomics = {}
    with ip.Batch(ctx, [rec]) as b:
        for enable, flags in ((True, ip.RUN_TRACE), (False, ip.RUN_TRACE), (False, ip.RUN_TRACE | ip.RUN_GRAPH),
                              (True, ip.RUN_TRACE | ip.RUN_GRAPH)):
            b.set_cell_table(enable)
            b.run(flags)
            got = b.bins()
            assert np.array_equal(got, want), (enable, flags, int(np.count_nonzero(got != want)))
            atomics[enable] = int(b.stats().cell_atomics)
        st = b.stats()
    print(f"HBM cell atomics, table on / off: {atomics[True]} / {atomics[False]}", flush=True)
    assert 0 < atomics[True] < atomics[False] <= st.samples


def test_atomic_probe_counts_every_increment(ctx):
    """The atomic calibration lands every increment (the library checks the slab
    sum against the count and fails otherwise) and refuses bad arguments."""
    print("case: atomic probe random 2^20 cells, hot cell, coalesced, refusals", flush=True)
    assert ctx.atomic_probe(1 << 20, 1 << 24) > 0
    assert ctx.atomic_probe(1, 1 << 20, "hot") > 0
    assert ctx.atomic_probe(3, 1000) > 0
    # the wrapping counter of the coalesced pattern: a cell count that divides
    # the grid's stride, one that does not, and a single cell
    for cells in (1 << 16, 1_620_001, 1):
        assert ctx.atomic_probe(cells, (1 << 22) + 7, "coalesced") > 0
    for cells, n in ((0, 10), (1 << 32, 10), (10, 0), (10, (1 << 40) + 1)):
        with pytest.raises(ip.IcePhaseError):
            ctx.atomic_probe(cells, n)
