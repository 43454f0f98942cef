"""The folded grid of the surrogate rests on a property of the drop records:
an axisymmetric drop lit in a vertical plane scatters symmetrically about that
plane, so in the record frame cell (i, j) and (i, (np/2 - 1 - j) mod np) carry
the same mass. Checked here on the CPU oracle's fixed-point histograms (the
GPU histograms equal them bit for bit, tests/test_gpu_parity.py): exactly on
power-of-two ray grids, whose FP32 coordinates mirror exactly, and to ~1e-7
in total variation on other grids, where a few mirrored rays round into the
neighbouring cell."""
from __future__ import annotations

import numpy as np
import pytest

import ice_crystal_phase_amd as ip

from oracle_bind import drop_hist, drop_record


def _hist(alpha, radius, np_, grid, nt=45):
    c = ip.drop_shape(radius)
    r = drop_record(coeffs=c, sphere=not any(c), alpha_deg=alpha, n_int=1.3335, grid=grid, max_hits=9)
    bins, _ = drop_hist(r, nt, np_)
    return np.asarray(bins).reshape(nt, np_), (np_ // 2 - 1 - np.arange(np_)) % np_


@pytest.mark.parametrize("alpha,radius,np_", [(0.0, 1.0, 90), (35.0, 2.4, 64), (-72.0, 0.8, 90), (89.0, 2.9, 36)])
def test_drop_histogram_is_mirror_symmetric(alpha, radius, np_):
    h, mirror = _hist(alpha, radius, np_, 256)
    assert h.sum() > 0
    assert np.array_equal(h, h[:, mirror])


def test_mirror_symmetry_on_other_grids_to_rounding():
    h, mirror = _hist(30.0, 2.0, 90, 192)
    hf = h.astype(np.float64)
    tv = 0.5 * np.abs(hf - hf[:, mirror]).sum() / hf.sum()
    assert 0 < tv < 1e-6
