"""The compiled-reference checker of the coherent stages runs on the CPU and
its records have the library's layouts (vertex, field, patch, folded record,
hit, per-direction result, config). Checks of the checker itself, from what
the geometry fixes: a forced sphere's R family leaves by specular reflection
and its optical path to the reference planes is exact; patch areas tile the
incident grid; the four stages agree where they must (one hit per direction:
path = incoherent; zero focal offsets and no transition: diffraction = focal)."""
from __future__ import annotations

import numpy as np
import pytest

import coherent_ref as cr

pytestmark = pytest.mark.skipif(not cr.available(), reason="oracle/_ref/libref_coherent.so is not built")


def test_record_layouts():
    cr.lib()  # raises if any reference record size differs from the library's dtypes


def test_sphere_reflection_family():
    cfg = cr.config(radius_mm=0.3, grid=17, theta=18, phi=36, sphere=True)
    v, f = cr.trace(cfg)
    n = cfg.grid * cfg.grid
    tc = np.zeros(16, np.float32)
    assert cr.lib().refc_trace_config(cfg, tc.ctypes.data) == 0
    w = tc[0:3].astype(np.float64)
    R = v[:n]
    ok = R["status"] == 1
    assert ok.sum() > 100
    p = R["pos"][ok].astype(np.float64)
    nrm = p / np.linalg.norm(p, axis=1)[:, None]
    spec = w - 2 * (nrm @ w)[:, None] * nrm
    assert np.abs(R["dir"][ok] - spec).max() < 2e-6
    # optical path of R: (2 R_out + w.p) + (2 R_out - d.p), in cycles of radius/wavelength
    ref_d = float(tc[7])
    geo = (ref_d + p @ w) + (ref_d - np.einsum("ij,ij->i", p, R["dir"][ok].astype(np.float64)))
    cyc = geo * (0.3e6 / 550.0)
    got = R["turns"][ok] + R["frac"][ok].astype(np.float64)
    assert np.abs(got - cyc).max() < 2e-3


def test_patch_areas_tile_the_grid():
    cfg = cr.config(radius_mm=0.5, grid=17, theta=18, phi=36)
    p, fr, counts = cr.patches(cfg)
    assert counts.sum() == 4 * 16 * 16
    tc = np.zeros(16, np.float32)
    cr.lib().refc_trace_config(cfg, tc.ctypes.data)
    du = np.float32(2.0) * tc[6] / np.float32(cfg.grid - 1)
    assert np.all(p["area"] == du * du)
    assert len(fr) == counts[3]


def test_stage_identities():
    cfg = cr.config(radius_mm=0.5, grid=25, theta=36, phi=72)
    d = cr.pipeline(cfg)
    single = d["hit_count"] == 1
    assert single.sum() > 50
    inc = cr.stage_density(d, "incoherent")
    path = cr.stage_density(d, "path")
    assert np.allclose(path[single], inc[single], rtol=1e-6, atol=0)
    plain = d["transition_kind"] == 0
    assert np.array_equal(cr.stage_density(d, "diffraction")[plain], cr.stage_density(d, "focal")[plain])
