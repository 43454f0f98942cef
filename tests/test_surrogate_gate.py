"""Host logic of the surrogate's validated domain (the gate that sends records
outside it to the tracer): bin lookup at the edges, persistence."""
from __future__ import annotations

import numpy as np

from ice_crystal_phase_amd import surrogate as S


def _gate(accept):
    return S.Gate(1e-2, 0.9, list(S.RADIUS_EDGES), accept, [0.0] * 11, [0.0] * 11, [10] * 11)


def test_bins_and_edges():
    g = _gate([True, False] * 5 + [True])
    r = np.array([0.25, 0.4999, 0.5, 0.74, 0.75, 2.76, 2.999, 3.0, 0.2, 3.1])
    conds = np.stack([np.zeros_like(r), np.full_like(r, 550.0), r, np.ones_like(r)], axis=1)
    # bin k covers [0.25 + 0.25 k, 0.5 + 0.25 k); 3.0 and beyond, and below 0.25, are outside
    assert g.accepts(conds).tolist() == [True, True, False, False, True, True, True, False, False, False]


def test_round_trip(tmp_path):
    g = _gate([False] * 3 + [True] * 8)
    g.validation = {"records": 256, "grid": 16384}
    p = tmp_path / "gate.json"
    g.save(p)
    h = S.Gate.load(p)
    assert h == g
    conds = S.drop_conditions(32, seed=3)
    assert np.array_equal(h.accepts(conds), conds[:, 2] >= 1.0)
