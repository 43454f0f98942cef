"""Test infrastructure only (never imported by the product): NumPy
restatement of the atlas composition sum, compose_mass of
hex_atlas/python/atlas_math.py:394-411, the oracle of the device
composition (iph_atlas_compose).

    out = 0; for i, j ascending, parity 0 then 1, shift s ascending with w > 0:
        out += w * roll(cell_parity(i, j), s, axis=phi)

cell_1 is the node block with its azimuth cells reversed (q -> np-1-q). Each
term is formed as a rounded product, then added, as NumPy does. Pinned by the
closed-form fractional shift / reflection case and by linearity in the tests.
"""
from __future__ import annotations

import numpy as np


def compose_mass_ref(mass: np.ndarray, weights: np.ndarray) -> np.ndarray:
    ni, nj, nt, np_ = mass.shape
    assert weights.shape == (ni, nj, 2, np_)
    out = np.zeros((nt, np_), dtype=np.float64)
    for i in range(ni):
        for j in range(nj):
            for parity in range(2):
                cell = np.asarray(mass[i, j])
                if parity:
                    cell = cell[:, ::-1]
                for shift, value in enumerate(weights[i, j, parity]):
                    if value > 0:
                        out += float(value) * np.roll(cell, shift, axis=1)
    return out
