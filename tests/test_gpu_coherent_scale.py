"""Coherent drop stages at the bench's full size: 1025 x 1025 vertices per family,
900 x 1800 query cells. Here a patch covers a handful of query cells (57.2 M
candidate pairs over 3.16 M patches, many with none or one), so the 64
consecutive candidates a wave of the patch query takes at a time often cross 64
or more patches: the case the small parity records, whose patches span many
cells, never reach. The counts are those of the patch search that was checked hit
for hit against the reference's brute-force query at the parity size (one
patch search per candidate, no per-wave shortcut); a wrong patch for a candidate
changes the hit count.
"""
import numpy as np
import pytest

import ice_crystal_phase_amd as ip

pytestmark = pytest.mark.gpu


@pytest.mark.skip(reason="a GPU memory fault ended the box during this test's first run (full-size record "
                         "after the parity-size records in one process, with a per-lane patch search by lane "
                         "shuffles that has since been taken out); its cause is not found, so the case stays "
                         "off the GPU until it is")
def test_full_size_query_counts(ctx):
    print("case: coherent 1025x1025 vertices, 900x1800 cells", flush=True)
    n = float(np.float32(ip.water_optics(550.0).refractive_index))
    with ip.Coherent(ctx, 0.5, 550.0, n, grid=1025, theta=900, phi=1800, inclination_deg=20.0) as c:
        c.trace()
        c.evaluate()
        st = c.stats()
    assert (st.patches, st.folded, st.candidates) == (3155098, 2350, 57224542)
    assert st.hits == 4072150
