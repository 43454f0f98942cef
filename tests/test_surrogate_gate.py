"""Host logic of the surrogate's validated domain (the gate that sends records
outside it to the tracer): bin lookup at the edges, persistence."""
from __future__ import annotations

import numpy as np

from ice_crystal_phase_amd import surrogate as S


def _gate(accept):
    """A gate over radius bins only (the flat form: one alpha bin)."""
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
    assert h.accept == g.accept and h.alpha_edges == g.alpha_edges and h.validation == g.validation
    conds = S.drop_conditions(32, seed=3)
    assert np.array_equal(h.accepts(conds), conds[:, 2] >= 1.0)


def test_radius_and_inclination_cells():
    acc = [[k < 3, k < 6, False] for k in range(11)]   # small drops: all |alpha| < 60
    g = S.Gate(1e-2, 0.9, list(S.RADIUS_EDGES), acc, alpha_edges=list(S.ALPHA_EDGES))
    conds = np.array([[10.0, 550, 0.3, 1], [-45.0, 550, 0.3, 1], [-75.0, 550, 0.3, 1], [29.9, 550, 1.1, 1],
                      [-30.0, 550, 1.1, 1], [59.0, 550, 1.6, 1], [90.0, 550, 0.3, 1]])
    # radius bins: 0.3 -> 0, 1.1 -> 3, 1.6 -> 5; |alpha| bins: [0, 30), [30, 60), [60, 90]
    assert g.accepts(conds).tolist() == [True, True, False, False, True, True, False]
