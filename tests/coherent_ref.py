"""ctypes binding of oracle/_ref/libref_coherent.so: the reference's coherent
drop pipeline compiled from its own headers (oracle/ref_coherent.cpp), the
checker of the HIP coherent stages. Test infrastructure only."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

import ice_crystal_phase_amd as ip
from ice_crystal_phase_amd.icephase import (DIRECTION_DTYPE, FIELD_DTYPE, FOLDED_DTYPE, HIT_DTYPE, PATCH_DTYPE,
                                            VERTEX_DTYPE, CoherentConfig)

ROOT = Path(__file__).resolve().parent.parent
LIB = ROOT / "oracle" / "_ref" / "libref_coherent.so"
_lib = None


def available() -> bool:
    return LIB.exists()


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        L = C.CDLL(str(LIB))
        P = C.POINTER
        L.refc_sizes.argtypes = [C.c_uint32]
        L.refc_sizes.restype = C.c_uint32
        L.refc_trace_config.argtypes = [P(CoherentConfig), C.c_void_p]
        L.refc_trace.argtypes = [P(CoherentConfig), C.c_void_p, C.c_void_p]
        L.refc_patches.argtypes = [P(CoherentConfig), C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, P(C.c_uint32),
                                   C.c_void_p, C.c_uint32, P(C.c_uint32), C.c_void_p]
        L.refc_query.argtypes = [P(CoherentConfig), C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p,
                                 C.c_uint64, C.c_void_p, C.c_void_p]
        L.refc_pipeline.argtypes = [P(CoherentConfig), C.c_void_p, C.c_void_p, C.c_void_p]
        L.refc_table_sigma.argtypes = [C.c_float]
        L.refc_table_sigma.restype = C.c_double
        L.refc_directions.argtypes = [P(CoherentConfig), C.c_void_p]
        sizes = [L.refc_sizes(i) for i in range(7)]
        want = [VERTEX_DTYPE.itemsize, FIELD_DTYPE.itemsize, PATCH_DTYPE.itemsize, FOLDED_DTYPE.itemsize,
                HIT_DTYPE.itemsize, DIRECTION_DTYPE.itemsize, C.sizeof(CoherentConfig)]
        if sizes != want:
            raise RuntimeError(f"reference record sizes {sizes} differ from the library's {want}")
        _lib = L
    return _lib


def config(radius_mm=0.5, wavelength_nm=550.0, interior_index=None, grid=33, theta=36, phi=72, inclination_deg=20.0,
           sphere=False, focal_offsets=(0, 0, 0, 0), primary_sigma_deg=0.0, transition_contrast=0.25):
    if interior_index is None:
        interior_index = float(np.float32(ip.water_optics(float(np.float32(wavelength_nm))).refractive_index))
    return CoherentConfig(radius_mm, wavelength_nm, interior_index, 0.0, inclination_deg, grid, 1 if sphere else 0,
                          theta, phi, (C.c_uint32 * 4)(*focal_offsets), primary_sigma_deg, transition_contrast)


def _nv(cfg) -> int:
    return 4 * cfg.grid * cfg.grid


def trace(cfg):
    v = np.zeros(_nv(cfg), VERTEX_DTYPE)
    f = np.zeros(_nv(cfg), FIELD_DTYPE)
    rc = lib().refc_trace(C.byref(cfg), v.ctypes.data, f.ctypes.data)
    if rc:
        raise RuntimeError(f"refc_trace: {rc}")
    return v, f


def patches(cfg, vertices=None, second=None):
    cells = 4 * (cfg.grid - 1) ** 2
    p = np.zeros(cells, PATCH_DTYPE)
    fr = np.zeros(cells, FOLDED_DTYPE)
    n, nf = C.c_uint32(), C.c_uint32()
    counts = np.zeros(8, np.uint32)
    rc = lib().refc_patches(C.byref(cfg), None if vertices is None else vertices.ctypes.data,
                            None if second is None else second.ctypes.data, p.ctypes.data, cells, C.byref(n),
                            fr.ctypes.data, cells, C.byref(nf), counts.ctypes.data)
    if rc:
        raise RuntimeError(f"refc_patches: {rc}")
    return p[:n.value], fr[:nf.value], counts


def query(cfg, first, count, vertices=None, second=None, cap=1 << 20):
    h = np.zeros(cap, HIT_DTYPE)
    counts = np.zeros(count, np.uint32)
    flags = np.zeros(count, np.uint32)
    rc = lib().refc_query(C.byref(cfg), None if vertices is None else vertices.ctypes.data,
                          None if second is None else second.ctypes.data, first, count, h.ctypes.data, cap,
                          counts.ctypes.data, flags.ctypes.data)
    if rc:
        raise RuntimeError(f"refc_query: {rc}")
    return h[:int(counts.sum())], counts, flags


def pipeline(cfg, vertices=None, second=None):
    out = np.zeros(cfg.theta_count * cfg.phi_count, DIRECTION_DTYPE)
    rc = lib().refc_pipeline(C.byref(cfg), None if vertices is None else vertices.ctypes.data,
                             None if second is None else second.ctypes.data, out.ctypes.data)
    if rc:
        raise RuntimeError(f"refc_pipeline: {rc}")
    return out


def directions(cfg):
    out = np.zeros((cfg.theta_count * cfg.phi_count, 3), np.float32)
    rc = lib().refc_directions(C.byref(cfg), out.ctypes.data)
    if rc:
        raise RuntimeError(f"refc_directions: {rc}")
    return out


def stage_density(d: np.ndarray, stage: str) -> np.ndarray:
    key = {"incoherent": "incoherent", "path": "path", "focal": "focal", "diffraction": "diffraction"}[stage]
    return d[key + "_s"] + d[key + "_p"]


def stage_mass(d: np.ndarray, stage: str, theta: int, phi: int) -> np.ndarray:
    """Normalized cell masses of a stage: w = d du per cell (the CDF weight), / total."""
    e = ip.u_edges(theta)
    w = stage_density(d, stage).reshape(theta, phi) * np.diff(e)[:, None]
    return w / w.sum()
