"""ctypes binding of libicephase.so (C-ABI declared in include/icephase.h).

This is the Python face of the same boundary the C++ command-line tools use.
It raises when the library is missing or no HIP device is usable: there is no
CPU fallback anywhere in the product.
"""
from __future__ import annotations

import ctypes as C
import math
import weakref
from dataclasses import dataclass, field
from pathlib import Path
from typing import Sequence

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libicephase.so"

u32, u64, i32, f64 = C.c_uint32, C.c_uint64, C.c_int32, C.c_double
P = C.POINTER

SHAPE_SPHERE, SHAPE_DROP, SHAPE_HEX = 0, 1, 2
REC_DIFFRACTION = 1
ABI_VERSION = 3
RUN_TRACE, RUN_CDF, RUN_GRAPH, RUN_NO_AUDIT = 1, 2, 4, 8

STATUS_NAMES = ("exhausted", "tail_tolerance_reached", "interaction_limit", "ambiguous_boundary",
                "invalid_input", "numerical_failure", "output_overflow", "replay_mismatch", "missed")


class IcePhaseError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"icephase error {code}: {message}")
        self.code = code


class AuditError(IcePhaseError):
    pass


class Record(C.Structure):
    _fields_ = [("shape", u32), ("max_internal_hits", u32), ("theta_count", u32), ("phi_count", u32),
                ("grid_w", u32), ("grid_h", u32), ("rays", u64), ("seed", u64), ("radius_mm", f64),
                ("length_mm", f64), ("shape_coefficients", f64 * 8), ("incident", f64 * 3),
                ("interior_index", f64), ("exterior_index", f64), ("wavelength_nm", f64),
                ("tail_tolerance", f64), ("rotation", f64 * 9), ("weight", f64), ("histogram_of", u32),
                ("flags", u32)]


class HistStats(C.Structure):
    _fields_ = [("q_forward", u64), ("q_backward", u64), ("samples", u64), ("invalid", u64), ("snapped", u64),
                ("status", u64 * 9), ("q_escaped", u64), ("q_unresolved", u64), ("internal_hits", u64),
                ("rays", u64), ("qbits", u32), ("pad", u32), ("ray_weight_mm2", f64),
                ("projected_area_mm2", f64), ("q_diffraction", u64), ("diffraction_captured", f64),
                ("cell_atomics", u64)]


class WaterState(C.Structure):
    _fields_ = [("temperature_kelvin", f64), ("pressure_pascal", f64), ("density_kg_m3", f64),
                ("density_pressure_residual_pascal", f64), ("density_bracket_width_kg_m3", f64),
                ("refractive_index", f64)]


class Area(C.Structure):
    _fields_ = [("area_mm2", f64), ("abs_error_mm2", f64), ("rel_error", f64), ("evaluations", u32),
                ("convexity_cells", u32), ("analytic_sphere", i32), ("pad", i32)]


class SurConfig(C.Structure):
    _fields_ = [("theta_count", u32), ("phi_count", u32), ("max_records", u32), ("dataset", u32), ("seed", u64)]


class CdfPolicy(C.Structure):
    _fields_ = [("out_theta", u32), ("out_phi", u32), ("max_l1", f64), ("max_lost_mass", f64),
                ("max_coarsening_tv", f64)]


class CdfReport(C.Structure):
    _fields_ = [("g_source", f64), ("g_stored", f64), ("coarsening_tv", f64), ("cdf_mass", f64),
                ("l1_error", f64), ("lost_mass", f64), ("max_cell_error", f64), ("aggregation_error", f64),
                ("fine_mass", f64), ("stored_mass", f64), ("lost_cells", u64), ("malformed", u64),
                ("weight_underflow", u64), ("negative", u64), ("status", i32), ("pad", i32)]


class IceState(C.Structure):
    _fields_ = [("wavelength_nm", f64), ("n", f64), ("kappa", f64), ("lower_nm", f64), ("upper_nm", f64),
                ("log_wavelength_fraction", f64)]


class AtlasSettings(C.Structure):
    _fields_ = [("radius_mm", f64), ("length_mm", f64), ("interior_index", f64), ("exterior_index", f64),
                ("wavelength_nm", f64), ("incident_theta_intervals", u32), ("incident_phi_intervals", u32),
                ("theta_count", u32), ("phi_count", u32), ("area_subdivisions", u32), ("max_internal_hits", u32),
                ("tail_tolerance", f64), ("max_unresolved_fraction", f64), ("max_balance_error", f64)]


class AtlasNode(C.Structure):
    _fields_ = [("i", u32), ("j", u32), ("theta", f64), ("phi", f64), ("ki", f64 * 3), ("e0", f64 * 3),
                ("projected_area_mm2", f64), ("escaped_mm2", f64), ("unresolved_mm2", f64),
                ("histogram_relative_error", f64), ("rays", u64), ("outputs", u64), ("max_balance_error", f64),
                ("max_interface_balance_error", f64)]


_lib: C.CDLL | None = None


class CoherentConfig(C.Structure):
    _fields_ = [("radius_mm", C.c_float), ("wavelength_nm", C.c_float), ("interior_index", C.c_float),
                ("reserved", C.c_float), ("inclination_deg", f64), ("grid", u32), ("sphere", u32),
                ("theta_count", u32), ("phi_count", u32), ("focal_offsets", u32 * 4), ("primary_sigma_deg", f64),
                ("transition_contrast", f64)]


class CoherentStats(C.Structure):
    _fields_ = [("status_counts", u32 * 8), ("patches", u32), ("folded", u32), ("bad_vertices", u32),
                ("first_bad_vertex", u32), ("candidates", u64), ("hits", u64), ("ms_trace", f64),
                ("ms_patches", f64), ("ms_folded", f64), ("ms_query", f64), ("ms_optics", f64),
                ("ms_diffraction", f64)]


class CoherentGate(C.Structure):
    _fields_ = [("invalid_values", u32), ("incomplete", u32), ("underresolved", u32), ("first_problem", u32),
                ("maximum", f64)]


# Record layouts shared with the reference's drop pipeline (OutgoingVertex,
# Field32, OutgoingPatch, FoldedPatchRecord, PatchQueryHit) and the per-direction
# result (iph_coherent_direction).
VERTEX_DTYPE = np.dtype([("pos", "<f4", 3), ("dir", "<f4", 3), ("bx", "<f4", 3), ("field", "<f4", 4),
                         ("turns", "<i4"), ("frac", "<f4"), ("status", "<u4"), ("diag", "<u4")])
FIELD_DTYPE = np.dtype([("field", "<f4", 4)])
PATCH_DTYPE = np.dtype([("v", "<u4", 4), ("id", "<u4"), ("status", "<u4"), ("omega", "<f4"), ("area", "<f4")])
BRANCH_DTYPE = np.dtype([("u0", "<f8"), ("u1", "<f8"), ("area", "<f8"), ("omega", "<f8"), ("ru", "<f8"),
                         ("rv", "<f8"), ("sign", "<i4"), ("focal", "<u4"), ("qturns", "<u4"), ("extra", "<u4")])
FOLDED_DTYPE = np.dtype([("compact", "<u4"), ("id", "<u4"), ("flags", "<u4"), ("nb", "<u4"), ("jac", "<f8", 4),
                         ("b", BRANCH_DTYPE, 4), ("err", "<f8"), ("panels", "<u4"), ("pad", "<u4")])
HIT_DTYPE = np.dtype([("compact", "<u4"), ("id", "<u4"), ("root", "<u4"), ("flags", "<u4"), ("u", "<f4"),
                      ("v", "<f4"), ("t", "<f4"), ("res", "<f4")])
DIRECTION_DTYPE = np.dtype([("incoherent_s", "<f8"), ("incoherent_p", "<f8"), ("path_s", "<f8"), ("path_p", "<f8"),
                            ("focal_s", "<f8"), ("focal_p", "<f8"), ("diffraction_s", "<f8"),
                            ("diffraction_p", "<f8"), ("family_incoherent", "<f8", 4), ("sigma_rad", "<f8"),
                            ("blend", "<f8"), ("path_field", "<f8", 8), ("focal_field", "<f8", 8),
                            ("hit_count", "<u4"), ("evaluated_hits", "<u4"), ("rejected_hits", "<u4"),
                            ("optical_flags", "<u4"), ("query_flags", "<u4"), ("focal_flags", "<u4"),
                            ("diffraction_flags", "<u4"), ("transition_flags", "<u4"), ("family_hits", "<u4", 4),
                            ("transition_kind", "<u4"), ("folded_hits", "<u4"), ("regular_hits", "<u4"),
                            ("refinement_hits", "<u4")])
STAGES = {"incoherent": 1, "path": 2, "focal": 3, "diffraction": 4}


def lib() -> C.CDLL:
    """Load libicephase.so; raise (never fall back) when it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise IcePhaseError(-2, f"{LIB_PATH} is not built; run __graft_entry__.build() "
                                "(make -C ice_crystal_phase_amd)")
    L = C.CDLL(str(LIB_PATH))
    L.iph_abi_version.restype = C.c_int
    L.iph_struct_size.argtypes = [C.c_int]
    L.iph_struct_size.restype = u32
    L.iph_last_error.argtypes = [C.c_void_p]
    L.iph_last_error.restype = C.c_char_p
    L.iph_ctx_create.argtypes = [C.c_int, P(C.c_void_p)]
    L.iph_ctx_destroy.argtypes = [C.c_void_p]
    L.iph_ctx_destroy.restype = None
    L.iph_u_edges.argtypes = [u32, P(f64)]
    L.iph_bin_directions.argtypes = [C.c_void_p, u32, u32, P(f64), u64, C.c_void_p, C.c_void_p, C.c_void_p]
    L.iph_batch_create.argtypes = [C.c_void_p, P(Record), u32, P(CdfPolicy), P(C.c_void_p)]
    L.iph_batch_destroy.argtypes = [C.c_void_p]
    L.iph_batch_destroy.restype = None
    L.iph_batch_run.argtypes = [C.c_void_p, u32]
    L.iph_batch_set_schedule.argtypes = [C.c_void_p, u64, u32]
    L.iph_batch_set_cell_table.argtypes = [C.c_void_p, u32]
    L.iph_batch_timing.argtypes = [C.c_void_p, P(f64)]
    L.iph_batch_stats.argtypes = [C.c_void_p, u32, P(HistStats)]
    L.iph_batch_bins.argtypes = [C.c_void_p, u32, C.c_void_p]
    L.iph_batch_cdf.argtypes = [C.c_void_p, u32, C.c_void_p, C.c_void_p, P(CdfReport)]
    L.iph_build_cdf.argtypes = [C.c_void_p, C.c_void_p, u32, u32, P(CdfPolicy), C.c_void_p, C.c_void_p,
                                P(CdfReport)]
    L.iph_build_cdf_noalloc.argtypes = [C.c_void_p, C.c_void_p, u32, u32, P(CdfPolicy), C.c_void_p, C.c_void_p,
                                        P(CdfReport)]
    L.iph_atomic_probe.argtypes = [C.c_void_p, u32, u64, u64, P(f64)]
    L.iph_write_record.argtypes = [C.c_char_p, C.c_char_p, C.c_void_p, C.c_void_p, u32, u32]
    L.iph_water_optics.argtypes = [f64, f64, f64, P(WaterState)]
    L.iph_water_index.argtypes = [f64, f64, f64, P(f64)]
    L.iph_drop_shape.argtypes = [C.c_float, C.c_int, P(C.c_float)]
    L.iph_drop_projected_area.argtypes = [P(f64), f64, P(f64), P(Area)]
    L.iph_hex_projected_area.argtypes = [f64, f64, P(f64), P(f64)]
    L.iph_batch_probabilities.argtypes = [C.c_void_p, u32, C.c_void_p]
    L.iph_sur_create.argtypes = [C.c_void_p, P(SurConfig), P(C.c_void_p)]
    L.iph_sur_destroy.argtypes = [C.c_void_p]
    L.iph_sur_destroy.restype = None
    L.iph_sur_param_count.restype = u32
    L.iph_sur_get_params.argtypes = [C.c_void_p, C.c_void_p]
    L.iph_sur_set_params.argtypes = [C.c_void_p, C.c_void_p]
    L.iph_sur_set_record.argtypes = [C.c_void_p, u32, C.c_void_p, u32, P(C.c_float)]
    L.iph_sur_set_target.argtypes = [C.c_void_p, u32, C.c_void_p, P(C.c_float)]
    L.iph_sur_set_loss.argtypes = [C.c_void_p, C.c_float]
    L.iph_sur_gradient.argtypes = [C.c_void_p, u32, P(C.c_float), C.c_void_p]
    L.iph_sur_train_step.argtypes = [C.c_void_p, C.c_void_p, u32, C.c_float, C.c_void_p]
    L.iph_sur_predict.argtypes = [C.c_void_p, C.c_void_p, u32, C.c_void_p]
    L.iph_sur_compare.argtypes = [C.c_void_p, C.c_void_p, u32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.iph_sur_records.argtypes = [C.c_void_p, C.c_void_p, u32, P(CdfPolicy), C.c_void_p, C.c_void_p, C.c_void_p]
    L.iph_sur_timing.argtypes = [C.c_void_p, P(f64)]
    L.iph_ice_index.argtypes = [f64, P(IceState)]
    L.iph_atlas_node_rays.argtypes = [P(AtlasSettings), u32, u32, C.c_void_p, u64, P(u64), P(AtlasNode)]
    L.iph_atlas_trace.argtypes = [C.c_void_p, P(AtlasSettings), u32, u32, C.c_void_p, C.c_void_p]
    L.iph_write_npy.argtypes = [C.c_char_p, P(u64), u32, C.c_void_p]
    L.iph_atlas_mass_create.argtypes = [C.c_void_p, C.c_void_p, u32, u32, u32, u32, P(C.c_void_p)]
    L.iph_atlas_mass_destroy.argtypes = [C.c_void_p]
    L.iph_atlas_mass_destroy.restype = None
    L.iph_atlas_compose.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    L.iph_coherent_create.argtypes = [C.c_void_p, P(CoherentConfig), P(C.c_void_p)]
    L.iph_coherent_destroy.argtypes = [C.c_void_p]
    L.iph_coherent_destroy.restype = None
    L.iph_coherent_trace.argtypes = [C.c_void_p]
    L.iph_coherent_set_vertices.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    L.iph_coherent_get_vertices.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    L.iph_coherent_evaluate.argtypes = [C.c_void_p]
    L.iph_coherent_stats_get.argtypes = [C.c_void_p, P(CoherentStats)]
    L.iph_coherent_get_patches.argtypes = [C.c_void_p, C.c_void_p, u32, P(u32), C.c_void_p, u32, P(u32)]
    L.iph_coherent_get_hits.argtypes = [C.c_void_p, u32, u32, C.c_void_p, u64, C.c_void_p, C.c_void_p]
    L.iph_coherent_get_directions.argtypes = [C.c_void_p, C.c_void_p]
    L.iph_coherent_density.argtypes = [C.c_void_p, u32, C.c_void_p, P(CoherentGate)]
    L.iph_coherent_cdf.argtypes = [C.c_void_p, u32, u32, P(CdfPolicy), C.c_void_p, C.c_void_p, P(CdfReport),
                                   P(CoherentGate)]
    for which, (name, st) in enumerate((("Record", Record), ("HistStats", HistStats), ("CdfReport", CdfReport),
                                        ("CdfPolicy", CdfPolicy), ("WaterState", WaterState), ("Area", Area),
                                        ("IceState", IceState), ("AtlasSettings", AtlasSettings),
                                        ("AtlasNode", AtlasNode), ("CoherentConfig", CoherentConfig),
                                        ("CoherentStats", CoherentStats), ("CoherentDirection", None),
                                        ("CoherentGate", CoherentGate))):
        if st is None:
            if L.iph_struct_size(which) != DIRECTION_DTYPE.itemsize:
                raise IcePhaseError(-1, f"ABI mismatch for {name}")
            continue
        if L.iph_struct_size(which) != C.sizeof(st):
            raise IcePhaseError(-1, f"ABI mismatch for {name}: {L.iph_struct_size(which)} != {C.sizeof(st)}")
    if L.iph_abi_version() != ABI_VERSION:
        raise IcePhaseError(-1, "Unsupported icephase ABI version")
    _lib = L
    return L


def _check(rc: int, ctx_ptr=None) -> None:
    if rc != 0:
        msg = lib().iph_last_error(ctx_ptr).decode(errors="replace")
        if rc == -3:
            raise AuditError(rc, msg)
        raise IcePhaseError(rc, msg)


def _live(handle: C.c_void_p, what: str) -> C.c_void_p:
    """The handle of an open library object; a closed one raises here, not in
    the library (whose message for a NULL handle is not the context's)."""
    if not handle:
        raise IcePhaseError(-1, f"{what} is closed")
    return handle


EXPORTED = ("iph_abi_version", "iph_struct_size", "iph_ctx_create", "iph_ctx_destroy", "iph_last_error",
            "iph_u_edges", "iph_bin_directions", "iph_batch_create", "iph_batch_destroy", "iph_batch_run",
            "iph_batch_set_schedule", "iph_batch_set_cell_table", "iph_batch_timing", "iph_batch_stats", "iph_batch_bins", "iph_batch_cdf",
            "iph_build_cdf", "iph_build_cdf_noalloc", "iph_atomic_probe", "iph_write_record", "iph_water_optics", "iph_water_index", "iph_drop_shape",
            "iph_drop_projected_area", "iph_hex_projected_area", "iph_batch_probabilities", "iph_sur_create",
            "iph_sur_destroy", "iph_sur_param_count", "iph_sur_get_params", "iph_sur_set_params",
            "iph_sur_set_record", "iph_sur_set_target", "iph_sur_set_loss", "iph_sur_gradient", "iph_sur_train_step", "iph_sur_predict",
            "iph_sur_compare", "iph_sur_records", "iph_sur_timing", "iph_ice_index", "iph_atlas_node_rays",
            "iph_atlas_trace", "iph_write_npy", "iph_atlas_mass_create", "iph_atlas_mass_destroy",
            "iph_atlas_compose", "iph_coherent_create", "iph_coherent_destroy", "iph_coherent_trace",
            "iph_coherent_set_vertices", "iph_coherent_get_vertices", "iph_coherent_evaluate", "iph_coherent_stats_get",
            "iph_coherent_get_patches", "iph_coherent_get_hits", "iph_coherent_get_directions",
            "iph_coherent_density", "iph_coherent_cdf")


def u_edges(nt: int) -> np.ndarray:
    out = np.empty(nt + 1, np.float64)
    _check(lib().iph_u_edges(nt, out.ctypes.data_as(P(f64))))
    return out


# ------------------------------------------------------------------ host services
def water_optics(wavelength_nm: float, celsius: float = 20.0, pascal: float = 101325.0) -> WaterState:
    """IAPWS-95 liquid density + IAPWS R9-97 index (iph_water_optics)."""
    w = WaterState()
    _check(lib().iph_water_optics(wavelength_nm, celsius, pascal, C.byref(w)))
    return w


def water_index(wavelength_nm: float, kelvin: float, density: float) -> float:
    n = f64()
    _check(lib().iph_water_index(wavelength_nm, kelvin, density, C.byref(n)))
    return n.value


def drop_shape(radius_mm: float, sphere: bool = False) -> list[float]:
    """Table I coefficients (single precision values, as floats)."""
    c = (C.c_float * 8)()
    _check(lib().iph_drop_shape(radius_mm, 1 if sphere else 0, c))
    return [float(v) for v in c]


def drop_projected_area(coeffs: Sequence[float], radius_mm: float, incident) -> Area:
    c = (f64 * 8)(*(list(coeffs) + [0.0] * (8 - len(coeffs)))[:8])
    k = (f64 * 3)(*incident)
    a = Area()
    _check(lib().iph_drop_projected_area(c, radius_mm, k, C.byref(a)))
    return a


def hex_projected_area(radius_mm: float, length_mm: float, incident) -> float:
    k = (f64 * 3)(*incident)
    a = f64()
    _check(lib().iph_hex_projected_area(radius_mm, length_mm, k, C.byref(a)))
    return a.value


def ice_index(wavelength_nm: float) -> IceState:
    """Warren & Brandt (2008) ice Ih index (iph_ice_index)."""
    st = IceState()
    _check(lib().iph_ice_index(wavelength_nm, C.byref(st)))
    return st


def atlas_settings(radius_mm: float, length_mm: float, n_int: float, theta_intervals: int, phi_intervals: int,
                   nt: int, np_: int, subdivisions: int, wavelength_nm: float = 550.0, n_ext: float = 1.0,
                   max_internal_hits: int = 2048, tail: float = 1e-12, max_unresolved: float = 1e-10,
                   max_balance: float = 1e-10) -> AtlasSettings:
    return AtlasSettings(radius_mm, length_mm, n_int, n_ext, wavelength_nm, theta_intervals, phi_intervals, nt, np_,
                         subdivisions, max_internal_hits, tail, max_unresolved, max_balance)


def atlas_node_rays(s: AtlasSettings, i: int, j: int):
    """Area rays of atlas node (i, j): rows (x, y, z, face, weight mm^2), and the node."""
    n = u64()
    node = AtlasNode()
    _check(lib().iph_atlas_node_rays(C.byref(s), i, j, None, 0, C.byref(n), C.byref(node)))
    out = np.zeros((n.value, 5), np.float64)
    _check(lib().iph_atlas_node_rays(C.byref(s), i, j, out.ctypes.data, n.value, C.byref(n), C.byref(node)))
    return out, node


# ------------------------------------------------------------------ records
def drop_incident(inclination_deg: float) -> tuple[float, float, float]:
    """ki = (cos a, -sin a, 0): apps/generate_phase.cpp:217 convention."""
    a = inclination_deg * math.pi / 180.0
    return (math.cos(a), -math.sin(a), 0.0)


def sphere_record(n_int: float, grid: int, theta: int = 900, phi: int = 1800, radius_mm: float = 1.0,
                  inclination_deg: float = 0.0, max_internal_hits: int = 9, wavelength_nm: float = 550.0,
                  tail: float = 0.0, grid_h: int | None = None) -> Record:
    r = Record()
    r.shape, r.max_internal_hits, r.theta_count, r.phi_count = SHAPE_SPHERE, max_internal_hits, theta, phi
    r.grid_w, r.grid_h = grid, grid if grid_h is None else grid_h
    r.radius_mm = radius_mm
    r.incident = (f64 * 3)(*drop_incident(inclination_deg))
    r.interior_index, r.exterior_index, r.wavelength_nm, r.tail_tolerance = n_int, 1.0, wavelength_nm, tail
    return r


def drop_record(coeffs: Sequence[float], n_int: float, grid: int, theta: int = 900, phi: int = 1800,
                radius_mm: float = 2.0, inclination_deg: float = 0.0, max_internal_hits: int = 9,
                wavelength_nm: float = 550.0, tail: float = 0.0) -> Record:
    r = sphere_record(n_int, grid, theta, phi, radius_mm, inclination_deg, max_internal_hits, wavelength_nm, tail)
    r.shape = SHAPE_DROP
    c = list(coeffs) + [0.0] * (8 - len(coeffs))
    r.shape_coefficients = (f64 * 8)(*c[:8])
    return r


def hex_record(radius_mm: float, length_mm: float, n_int: float, rays: int, incident=(1.0, -0.4, 0.3),
               seed: int = 12345, theta: int = 900, phi: int = 1800, max_internal_hits: int = 12,
               tail: float = 1e-12, wavelength_nm: float = 550.0, n_ext: float = 1.0, rotation=None,
               weight: float = 0.0, histogram_of: int = 0, diffraction: bool = False) -> Record:
    """Hex record; rotation (3x3 body -> ensemble, row-major) makes it an
    orientation node traced in the body frame and binned in the ensemble frame;
    histogram_of = k > 0 adds it into record k-1's histogram."""
    r = Record()
    r.shape, r.max_internal_hits, r.theta_count, r.phi_count = SHAPE_HEX, max_internal_hits, theta, phi
    r.rays, r.seed = rays, seed
    r.radius_mm, r.length_mm = radius_mm, length_mm
    r.incident = (f64 * 3)(*incident)
    r.interior_index, r.exterior_index, r.wavelength_nm, r.tail_tolerance = n_int, n_ext, wavelength_nm, tail
    if rotation is not None:
        r.rotation = (f64 * 9)(*np.asarray(rotation, np.float64).reshape(9))
    r.weight, r.histogram_of = weight, histogram_of
    r.flags = REC_DIFFRACTION if diffraction else 0
    return r


def quaternion_rotation(w: float, x: float, y: float, z: float) -> np.ndarray:
    """Active Hamilton quaternion (w,x,y,z) -> row-major body -> ensemble matrix,
    columns the images of the body axes (hex_phase/src/phase_table.cpp:11-21)."""
    n = w * w + x * x + y * y + z * z
    if not math.isfinite(n) or abs(n - 1) > 2e-12:
        raise ValueError("Quaternion must be unit length (w,x,y,z).")
    k = 1 / math.sqrt(n)
    w, x, y, z = w * k, x * k, y * k, z * k
    cx = (1 - 2 * (y * y + z * z), 2 * (x * y + w * z), 2 * (x * z - w * y))
    cy = (2 * (x * y - w * z), 1 - 2 * (x * x + z * z), 2 * (y * z + w * x))
    cz = (2 * (x * z + w * y), 2 * (y * z - w * x), 1 - 2 * (x * x + y * y))
    return np.array([[cx[r], cy[r], cz[r]] for r in range(3)], np.float64)


def record_rays(r: Record) -> int:
    return int(r.rays) if r.shape == SHAPE_HEX else int(r.grid_w) * int(r.grid_h)


# ------------------------------------------------------------------ context / batch
class Context:
    """One HIP device, one stream (iph_ctx)."""

    def __init__(self, device: int = 0):
        self._p = C.c_void_p()
        self._batches = weakref.WeakSet()
        _check(lib().iph_ctx_create(device, C.byref(self._p)))

    @property
    def ptr(self):
        return self._p

    def close(self) -> None:
        """Destroy the context; live batches are closed first."""
        if self._p:
            for b in list(self._batches):
                b.close()
            lib().iph_ctx_destroy(self._p)
            self._p = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def bin_directions(self, nt: int, np_: int, frame: np.ndarray, dirs: np.ndarray):
        frame = np.ascontiguousarray(frame, np.float64).reshape(9)
        dirs = np.ascontiguousarray(dirs, np.float64).reshape(-1, 3)
        n = dirs.shape[0]
        index = np.zeros(n, np.uint32)
        kind = np.zeros(n, np.int32)
        _check(lib().iph_bin_directions(self._p, nt, np_, frame.ctypes.data_as(P(f64)), n, dirs.ctypes.data,
                                        index.ctypes.data, kind.ctypes.data), self._p)
        return index, kind

    ATOMIC_PATTERNS = {"random": 0, "hot": 1, "coalesced": 2}

    def atomic_probe(self, cells: int, atomics: int, pattern: str = "random") -> float:
        """Device ms of `atomics` u64 atomicAdds over a slab of `cells` u64:
        random cells, one hot cell or consecutive cells (iph_atomic_probe)."""
        if pattern not in self.ATOMIC_PATTERNS:
            raise ValueError(f"atomic probe pattern must be one of {sorted(self.ATOMIC_PATTERNS)}")
        ms = f64()
        _check(lib().iph_atomic_probe(self._p, self.ATOMIC_PATTERNS[pattern], cells, atomics, C.byref(ms)), self._p)
        return ms.value

    def build_cdf(self, mass: np.ndarray, nt: int, np_: int, out_theta: int = 0, out_phi: int = 0,
                  max_l1: float = 1e-10, max_lost: float = 1e-12, max_tv: float = 1.0, check: bool = True):
        mass = np.ascontiguousarray(mass, np.float64).reshape(-1)
        pol = CdfPolicy(out_theta, out_phi, max_l1, max_lost, max_tv)
        ont = out_theta or nt
        onp = out_phi or np_
        phi = np.zeros(onp + 1, np.float64)
        theta = np.zeros(onp * (ont + 1), np.float64)
        rep = CdfReport()
        rc = lib().iph_build_cdf(self._p, mass.ctypes.data, nt, np_, C.byref(pol), phi.ctypes.data,
                                 theta.ctypes.data, C.byref(rep))
        if check:
            _check(rc, self._p)
        return phi, theta.reshape(onp, ont + 1), rep, rc


def atlas_trace(ctx: "Context", s: AtlasSettings, first: int = 0, count: int | None = None):
    """Cell masses [count, nt, np] (mm^2) and node reports of atlas nodes
    first..first+count-1 (iph_atlas_trace)."""
    total = (s.incident_theta_intervals + 1) * (s.incident_phi_intervals + 1)
    count = total - first if count is None else count
    mass = np.zeros((count, s.theta_count, s.phi_count), np.float64)
    nodes = (AtlasNode * count)()
    _check(lib().iph_atlas_trace(ctx.ptr, C.byref(s), first, count, mass.ctypes.data, nodes), ctx.ptr)
    return mass, list(nodes)


class AtlasMass:
    """Device copy of an atlas mass array [ni, nj, nt, np] for composition."""

    def __init__(self, ctx: "Context", mass: np.ndarray):
        m = np.ascontiguousarray(mass, np.float64)
        if m.ndim != 4:
            raise ValueError("atlas mass must be [incident_theta, incident_phi, theta, phi]")
        self.ctx, self.shape = ctx, m.shape
        self._p = C.c_void_p()
        _check(lib().iph_atlas_mass_create(ctx.ptr, m.ctypes.data, *m.shape, C.byref(self._p)), ctx.ptr)
        ctx._batches.add(self)

    def compose(self, weights: np.ndarray) -> np.ndarray:
        """Composed cell mass [nt, np] of coupling weights [ni, nj, 2, np]."""
        _live(self._p, "atlas mass")
        ni, nj, nt, np_ = self.shape
        w = np.ascontiguousarray(weights, np.float64)
        if w.shape != (ni, nj, 2, np_):
            raise ValueError(f"weights must have shape {(ni, nj, 2, np_)}")
        out = np.zeros((nt, np_), np.float64)
        _check(lib().iph_atlas_compose(self._p, w.ctypes.data, out.ctypes.data), self.ctx.ptr)
        return out

    def close(self) -> None:
        if self._p and self.ctx.ptr:
            lib().iph_atlas_mass_destroy(self._p)
        self._p = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Coherent:
    """One drop record on the patch model (iph_coherent): the coherent stages
    path, focal and diffraction of rainbow_generate, and the patch-model
    incoherent stage."""

    def __init__(self, ctx: "Context", radius_mm: float, wavelength_nm: float, interior_index: float,
                 grid: int = 129, theta: int = 360, phi: int = 720, inclination_deg: float = 20.0,
                 sphere: bool = False, focal_offsets=(0, 0, 0, 0), primary_sigma_deg: float = 0.0,
                 transition_contrast: float = 0.25):
        self.ctx = ctx
        self.config = CoherentConfig(radius_mm, wavelength_nm, interior_index, 0.0, inclination_deg, grid,
                                     1 if sphere else 0, theta, phi, (u32 * 4)(*focal_offsets), primary_sigma_deg,
                                     transition_contrast)
        self.grid, self.theta, self.phi = grid, theta, phi
        self.nv = 4 * grid * grid
        self._p = C.c_void_p()
        _check(lib().iph_coherent_create(ctx.ptr, C.byref(self.config), C.byref(self._p)), ctx.ptr)
        ctx._batches.add(self)

    def trace(self) -> None:
        _check(lib().iph_coherent_trace(_live(self._p, "coherent record")), self.ctx.ptr)

    def set_vertices(self, vertices: np.ndarray, second: np.ndarray) -> None:
        v = np.ascontiguousarray(vertices)
        f = np.ascontiguousarray(second)
        if v.nbytes != self.nv * VERTEX_DTYPE.itemsize or f.nbytes != self.nv * FIELD_DTYPE.itemsize:
            raise ValueError("vertex arrays must hold 4 * grid^2 records")
        _check(lib().iph_coherent_set_vertices(_live(self._p, "coherent record"), v.ctypes.data, f.ctypes.data),
               self.ctx.ptr)

    def vertices(self):
        v = np.zeros(self.nv, VERTEX_DTYPE)
        f = np.zeros(self.nv, FIELD_DTYPE)
        _check(lib().iph_coherent_get_vertices(_live(self._p, "coherent record"), v.ctypes.data, f.ctypes.data),
               self.ctx.ptr)
        return v, f

    def evaluate(self) -> None:
        _check(lib().iph_coherent_evaluate(_live(self._p, "coherent record")), self.ctx.ptr)

    def stats(self) -> CoherentStats:
        st = CoherentStats()
        _check(lib().iph_coherent_stats_get(_live(self._p, "coherent record"), C.byref(st)), self.ctx.ptr)
        return st

    def patches(self):
        n, nf = u32(), u32()
        h = _live(self._p, "coherent record")
        _check(lib().iph_coherent_get_patches(h, None, 0, C.byref(n), None, 0, C.byref(nf)), self.ctx.ptr)
        p = np.zeros(n.value, PATCH_DTYPE)
        f = np.zeros(nf.value, FOLDED_DTYPE)
        _check(lib().iph_coherent_get_patches(h, p.ctypes.data, n.value, C.byref(n), f.ctypes.data, nf.value,
                                              C.byref(nf)), self.ctx.ptr)
        return p, f

    def hits(self, first: int, count: int, cap: int = 1 << 20):
        h = np.zeros(cap, HIT_DTYPE)
        counts = np.zeros(count, np.uint32)
        flags = np.zeros(count, np.uint32)
        _check(lib().iph_coherent_get_hits(_live(self._p, "coherent record"), first, count, h.ctypes.data, cap,
                                           counts.ctypes.data, flags.ctypes.data), self.ctx.ptr)
        return h[:int(counts.sum())], counts, flags

    def directions(self) -> np.ndarray:
        out = np.zeros(self.theta * self.phi, DIRECTION_DTYPE)
        _check(lib().iph_coherent_get_directions(_live(self._p, "coherent record"), out.ctypes.data), self.ctx.ptr)
        return out

    def density(self, stage: str):
        out = np.zeros(self.theta * self.phi, np.float64)
        g = CoherentGate()
        _check(lib().iph_coherent_density(_live(self._p, "coherent record"), STAGES[stage], out.ctypes.data,
                                          C.byref(g)), self.ctx.ptr)
        return out.reshape(self.theta, self.phi), g

    def cdf(self, stage: str, allow_underresolved: bool = False, out_theta: int = 0, out_phi: int = 0,
            max_l1: float = 1e-10, max_lost: float = 1e-12, max_tv: float = 1.0):
        pol = CdfPolicy(out_theta, out_phi, max_l1, max_lost, max_tv)
        ont, onp = out_theta or self.theta, out_phi or self.phi
        phi = np.zeros(onp + 1, np.float64)
        theta = np.zeros(onp * (ont + 1), np.float64)
        rep, g = CdfReport(), CoherentGate()
        _check(lib().iph_coherent_cdf(_live(self._p, "coherent record"), STAGES[stage], 1 if allow_underresolved else 0,
                                      C.byref(pol), phi.ctypes.data, theta.ctypes.data, C.byref(rep), C.byref(g)),
               self.ctx.ptr)
        return phi, theta.reshape(onp, ont + 1), rep, g

    def close(self) -> None:
        if self._p and self.ctx.ptr:
            lib().iph_coherent_destroy(self._p)
        self._p = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass(eq=False)
class Batch:
    """Device state of several records (iph_batch)."""

    ctx: Context
    records: list
    policy: CdfPolicy | None = None
    _p: C.c_void_p = field(default_factory=C.c_void_p)

    def __post_init__(self):
        arr = (Record * len(self.records))(*self.records)
        self._arr = arr
        pol = C.byref(self.policy) if self.policy is not None else None
        _check(lib().iph_batch_create(self.ctx.ptr, arr, len(self.records), pol, C.byref(self._p)), self.ctx.ptr)
        self.ctx._batches.add(self)

    def close(self):
        if self._p and self.ctx.ptr:
            lib().iph_batch_destroy(self._p)
        self._p = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, flags: int = RUN_TRACE | RUN_CDF) -> None:
        _check(lib().iph_batch_run(_live(self._p, "batch"), flags), self.ctx.ptr)

    def set_schedule(self, tile_rays: int = 0, blocks: int = 0) -> None:
        _check(lib().iph_batch_set_schedule(_live(self._p, "batch"), tile_rays, blocks), self.ctx.ptr)

    def set_cell_table(self, enable: bool) -> None:
        """Hex kernels: workgroup LDS cell table (True, default) or every quantum straight to HBM."""
        _check(lib().iph_batch_set_cell_table(_live(self._p, "batch"), int(enable)), self.ctx.ptr)

    def timing(self) -> tuple[float, float, float]:
        ms = (f64 * 3)()
        _check(lib().iph_batch_timing(_live(self._p, "batch"), ms))
        return ms[0], ms[1], ms[2]

    def stats(self, rec: int = 0) -> HistStats:
        st = HistStats()
        _check(lib().iph_batch_stats(_live(self._p, "batch"), rec, C.byref(st)), self.ctx.ptr)
        return st

    def bins(self, rec: int = 0) -> np.ndarray:
        r = self.records[rec]
        out = np.zeros(int(r.theta_count) * int(r.phi_count), np.uint64)
        _check(lib().iph_batch_bins(_live(self._p, "batch"), rec, out.ctypes.data), self.ctx.ptr)
        return out

    @property
    def ptr(self):
        return self._p

    def cdf(self, rec: int = 0):
        r = self.records[rec]
        ont = (self.policy.out_theta if self.policy is not None and self.policy.out_theta else r.theta_count)
        onp = (self.policy.out_phi if self.policy is not None and self.policy.out_phi else r.phi_count)
        phi = np.zeros(onp + 1, np.float64)
        theta = np.zeros(onp * (ont + 1), np.float64)
        rep = CdfReport()
        _check(lib().iph_batch_cdf(_live(self._p, "batch"), rec, phi.ctypes.data, theta.ctypes.data, C.byref(rep)),
               self.ctx.ptr)
        return phi, theta.reshape(onp, ont + 1), rep


def write_record(out_dir: str | Path, metadata_json: str, phi: np.ndarray, theta: np.ndarray) -> None:
    theta = np.ascontiguousarray(theta, np.float64)
    phi = np.ascontiguousarray(phi, np.float64)
    np_, nt1 = theta.shape
    _check(lib().iph_write_record(str(out_dir).encode(), metadata_json.encode(), phi.ctypes.data,
                                  theta.ctypes.data, nt1 - 1, np_))
