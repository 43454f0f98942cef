"""Conformance check of one record directory against the on-disk contract
(README.md:153-210 and python/rainbow_dataset/record.py:19-138 of the
reference describe it): metadata keys and fixed values, NPY v1.0 '<f8'
C-order arrays of the documented shapes with no trailing bytes, strictly
increasing u edges from 0 to 1, monotone CDFs with exact 0/1 endpoints, a
right-handed orthonormal sampling frame, and the HG label equal to the first
moment of the saved cell pdf.

Test infrastructure: written for this repository's tests, it loads arrays with
numpy.load(allow_pickle=False) only.
"""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np

FIXED = {
    "dtype": "<f8", "order": "C", "density_measure": "solid_angle_sr",
    "cell_model": "constant_density_per_spherical_cell",
    "coordinates": "u=(1-cos(theta))/2; v=(phi+pi)/(2*pi)",
    "normalization": "integral_p_domega_equals_one", "input_polarization": "unpolarized",
}


def _npy(path: Path, shape) -> np.ndarray:
    raw = path.read_bytes()
    assert raw[:6] == b"\x93NUMPY" and raw[6] == 1 and raw[7] == 0, "NPY v1.0 magic"
    hlen = raw[8] | (raw[9] << 8)
    assert (10 + hlen) % 64 == 0, "header padded to 64 bytes"
    header = raw[10:10 + hlen].decode("latin1")
    assert header.endswith("\n") and "'descr': '<f8'" in header and "'fortran_order': False" in header
    a = np.load(path, allow_pickle=False)
    assert a.dtype.str == "<f8" and a.shape == tuple(shape) and a.flags.c_contiguous
    assert len(raw) == 10 + hlen + a.nbytes, "no trailing data"
    return a


def check_record(directory: str | Path) -> dict:
    d = Path(directory)
    assert not d.name.endswith(".part")
    meta_path = d / "metadata.json"
    assert meta_path.stat().st_size <= 1024 * 1024
    m = json.loads(meta_path.read_text(encoding="utf-8"))
    assert m.get("schema") in ("rainbow.phase_cdf.numpy.v1", "rainbow.phase_cdf.numpy.v2")
    assert m.get("complete") is True
    for k, v in FIXED.items():
        assert m.get(k) == v, k
    assert m.get("cdf_axis_order") == ["phi_cell", "theta_edge"]
    nt, np_ = m["theta_count"], m["phi_count"]
    assert type(nt) is int and type(np_) is int and 1 <= nt <= 0x7FFFFFFF and 1 <= np_ <= 0x7FFFFFFF
    phi = _npy(d / "phi_cdf.npy", (np_ + 1,))
    theta = _npy(d / "theta_given_phi_cdf.npy", (np_, nt + 1))
    u = _npy(d / "u_edges.npy", (nt + 1,))
    frame = np.asarray(m["sampling_frame_columns"], np.float64)
    assert frame.shape == (3, 3) and np.isfinite(frame).all()
    assert np.allclose(frame.T @ frame, np.eye(3), rtol=0, atol=2e-12)
    assert abs(np.linalg.det(frame) - 1.0) <= 2e-12
    assert np.isfinite(u).all() and u[0] == 0 and u[-1] == 1 and np.all(np.diff(u) > 0)
    assert np.isfinite(phi).all() and phi[0] == 0 and phi[-1] == 1 and np.all(np.diff(phi) >= 0)
    assert np.isfinite(theta).all() and np.all(theta >= 0) and np.all(theta <= 1)
    assert np.all(theta[:, 0] == 0) and np.all(theta[:, -1] == 1)
    assert np.all(np.diff(theta, axis=1) >= 0)
    if m["schema"].endswith(".v2"):
        hg = m["hg"]
        assert hg["method"] == "first_moment_of_saved_cell_pdf" and hg["target"] == "saved_cdf"
        assert hg["cosine_convention"] == "dot(incident_propagation,outgoing_propagation)"
        g = hg["g"]
        assert isinstance(g, (int, float)) and -1 < g < 1
        # cell probability = dphi * dtheta_cdf; cell mean cosine = 1 - (u_i + u_{i+1})
        p = np.diff(phi)[:, None] * np.diff(theta, axis=1)
        mass = p.sum()
        g_cells = float((p * (1 - (u[:-1] + u[1:]))[None, :]).sum())
        assert abs(mass - 1) <= 1e-10 and abs(g_cells - g) <= 5e-12, (mass, g_cells, g)
    return m


def cell_probabilities(directory: str | Path) -> np.ndarray:
    """Saved cell probabilities [theta, phi] of a record."""
    d = Path(directory)
    phi = np.load(d / "phi_cdf.npy", allow_pickle=False)
    theta = np.load(d / "theta_given_phi_cdf.npy", allow_pickle=False)
    return (np.diff(phi)[:, None] * np.diff(theta, axis=1)).T


def key_paths(obj, prefix="") -> set[str]:
    out = set()
    if isinstance(obj, dict):
        for k, v in obj.items():
            out.add(prefix + k)
            out |= key_paths(v, prefix + k + ".")
    return out
