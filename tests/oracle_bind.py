"""ctypes bindings to the CPU oracle (oracle/liboracle.so) and, when it was
built, the compiled-reference shim (oracle/_ref/libref.so).

Test infrastructure only: the product never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
ORACLE_SO = ROOT / "oracle" / "liboracle.so"
REF_SO = ROOT / "oracle" / "_ref" / "libref.so"

u32, u64, i32, i64, f64, f32 = C.c_uint32, C.c_uint64, C.c_int32, C.c_int64, C.c_double, C.c_float
P = C.POINTER


class Frame(C.Structure):
    _fields_ = [("ki", f64 * 3), ("e0", f64 * 3), ("e1", f64 * 3)]


class Bin(C.Structure):
    _fields_ = [("index", u64), ("mu", f64), ("pole", i32), ("valid", i32), ("snapped", i32), ("pad", i32)]


class HexSettings(C.Structure):
    _fields_ = [("radius_mm", f64), ("length_mm", f64), ("ki", f64 * 3), ("n_ext", f64), ("n_int", f64),
                ("tail", f64), ("seed", u64), ("total", u64), ("max_hits", u32), ("pad", u32)]


class RayAudit(C.Structure):
    _fields_ = [("escaped", f64), ("unresolved", f64), ("balance", f64), ("max_interface", f64),
                ("outputs", u64), ("internal_hits", u32), ("tir_count", u32), ("status", u32), ("pad", u32)]


class Sample(C.Structure):
    _fields_ = [("dir", f64 * 3), ("power", f64), ("length_mm", f64), ("id", u64), ("entry_face", u32),
                ("exit_face", u32), ("reflections", u32), ("kind", u32)]


class HistStats(C.Structure):
    _fields_ = [("q_forward", u64), ("q_backward", u64), ("samples", u64), ("invalid", u64), ("snapped", u64),
                ("status", u64 * 8), ("q_escaped", u64), ("q_unresolved", u64), ("rays_hit", u64),
                ("internal_hits", u64)]


class CdfReport(C.Structure):
    _fields_ = [("g_source", f64), ("g_stored", f64), ("coarsening_tv", f64), ("cdf_mass", f64),
                ("l1_error", f64), ("lost_mass", f64), ("max_cell_error", f64), ("aggregation_error", f64),
                ("lost_cells", u64), ("status", i32), ("pad", i32)]


class RefCdfReport(C.Structure):
    _fields_ = [("g_source", f64), ("g_stored", f64), ("coarsening_tv", f64), ("cdf_mass", f64),
                ("l1_error", f64), ("lost_mass", f64), ("max_cell_error", f64), ("aggregation_error", f64),
                ("lost_cells", u64)]


class AtlasSettings(C.Structure):
    _fields_ = [("radius_mm", f64), ("length_mm", f64), ("n_int", f64), ("n_ext", f64), ("tail", f64),
                ("theta_intervals", u32), ("phi_intervals", u32), ("subdivisions", u32), ("theta_count", u32),
                ("phi_count", u32), ("max_hits", u32)]


class AtlasReport(C.Structure):
    _fields_ = [("theta", f64), ("phi", f64), ("ki", f64 * 3), ("e0", f64 * 3), ("projected_area_mm2", f64),
                ("escaped_mm2", f64), ("unresolved_mm2", f64), ("max_balance_error", f64),
                ("max_interface_balance_error", f64), ("rays", u64), ("outputs", u64)]


class DropRecord(C.Structure):
    _fields_ = [("coeffs", f32 * 8), ("inner", f32), ("outer", f32), ("rmax", f32), ("ki", f32 * 3), ("e0", f32 * 3),
                ("e1", f32 * 3), ("n_int", f32), ("n_ext", f32), ("half_extent", f32), ("grid_w", u32),
                ("grid_h", u32), ("max_hits", u32), ("sphere", u32), ("tail", f32)]


SAMPLE_DTYPE = np.dtype([("dir", "<f8", 3), ("power", "<f8"), ("length_mm", "<f8"), ("id", "<u8"),
                         ("entry_face", "<u4"), ("exit_face", "<u4"), ("reflections", "<u4"), ("kind", "<u4")])
AUDIT_DTYPE = np.dtype([("escaped", "<f8"), ("unresolved", "<f8"), ("balance", "<f8"), ("max_interface", "<f8"),
                        ("outputs", "<u8"), ("internal_hits", "<u4"), ("tir_count", "<u4"), ("status", "<u4"),
                        ("pad", "<u4")])
assert SAMPLE_DTYPE.itemsize == C.sizeof(Sample)
assert AUDIT_DTYPE.itemsize == C.sizeof(RayAudit)


def _ptr(a: np.ndarray, ctype=C.c_void_p):
    return a.ctypes.data_as(ctype)


def build_oracle() -> None:
    """Build the oracle (and _ref when the reference checkout exists)."""
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], check=True)


_lib = None
_ref = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not ORACLE_SO.exists():
            build_oracle()
        L = C.CDLL(str(ORACLE_SO))
        L.orc_u_edges.argtypes = [u32, P(f64)]
        L.orc_make_frame.argtypes = [P(f64), P(Frame)]
        L.orc_drop_frame.argtypes = [P(f64), P(Frame)]
        L.orc_phase_bin.argtypes = [P(f64), u32, u32, P(Frame), P(f64)]
        L.orc_phase_bin.restype = Bin
        L.orc_hex_trace.argtypes = [P(HexSettings), u64, u64, C.c_void_p, u64, C.c_void_p]
        L.orc_hex_trace.restype = i64
        L.orc_hex_projected_area.argtypes = [f64, f64, P(f64)]
        L.orc_hex_projected_area.restype = f64
        L.orc_quantum_bits.argtypes = [u64]
        L.orc_quantum_bits.restype = u32
        L.orc_hex_hist.argtypes = [P(HexSettings), u32, u32, u32, u64, u64, C.c_void_p, P(HistStats)]
        L.orc_hex_masses.argtypes = [P(HexSettings), u32, u32, C.c_void_p, P(f64), P(f64)]
        L.orc_bins_to_mass.argtypes = [C.c_void_p, u64, u64, u32, u32, C.c_void_p]
        L.orc_bins_to_mass.restype = None
        L.orc_build_cdf.argtypes = [C.c_void_p, u32, u32, u32, u32, f64, f64, f64, C.c_void_p, C.c_void_p,
                                    P(CdfReport)]
        L.orc_drop_make.argtypes = [P(f64), C.c_int, f64, f64, u32, u32, u32, f64, P(DropRecord)]
        L.orc_drop_trace.argtypes = [P(DropRecord), C.c_int, u64, u64, C.c_void_p, u64, C.c_void_p]
        L.orc_drop_trace.restype = i64
        L.orc_drop_hist.argtypes = [P(DropRecord), C.c_int, u32, u32, u32, u64, u64, C.c_void_p, P(HistStats)]
        L.orc_drop_surface.argtypes = [P(DropRecord), C.c_int, u64, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_drop_entry_point.argtypes = [P(DropRecord), C.c_int, f64, f64, P(f64), P(f64)]
        L.orc_drop_exit_point.argtypes = [P(DropRecord), C.c_int, P(f64), P(f64), P(f64), P(f64)]
        L.orc_drop_trace_origin.argtypes = [P(DropRecord), C.c_int, f64, f64, C.c_void_p, u64, C.c_void_p]
        L.orc_drop_trace_origin.restype = i64
        L.orc_sphere_order.argtypes = [f64, f64, u32, P(f64), P(f64), P(f64)]
        L.orc_sphere_order.restype = None
        L.orc_normalize3.argtypes = [P(f64), P(f64)]
        L.orc_sphere_theta_hist.argtypes = [u32, f32, f64, u32, u32, C.c_void_p]
        L.orc_sphere_theta_hist.restype = None
        L.orc_hex_hist_node.argtypes = [P(HexSettings), P(f64), f64, u32, u32, u32, C.c_void_p, P(HistStats)]
        L.orc_hex_node_area.argtypes = [P(HexSettings), P(f64)]
        L.orc_hex_node_area.restype = f64
        L.orc_aperture_transform.argtypes = [P(f64), P(f64), u32, f64, f64, P(f64), P(f64)]
        L.orc_aperture_transform.restype = None
        L.orc_hex_aperture.argtypes = [f64, f64, P(f64), P(f64), P(f64), P(f64), P(f64)]
        L.orc_hex_aperture.restype = u32
        L.orc_lobe_cell.argtypes = [P(f64), P(f64), u32, f64, f64, P(f64), P(f64), f64, P(Frame), u32, u32, u32,
                                    u32]
        L.orc_lobe_cell.restype = f64
        L.orc_atlas_rule.argtypes = [P(AtlasSettings), u32, u32, C.c_void_p, u64]
        L.orc_atlas_rule.restype = i64
        L.orc_atlas_node.argtypes = [P(AtlasSettings), u32, u32, C.c_void_p, P(AtlasReport)]
        _lib = L
    return _lib


REF_ATLAS_CLI = ROOT / "oracle" / "_ref" / "ice_precompute_atlas_cpu"


def atlas_rule(st: AtlasSettings, i: int, j: int) -> np.ndarray:
    """Oracle area rule of node (i, j): rows (x, y, z, face, weight)."""
    n = lib().orc_atlas_rule(C.byref(st), i, j, None, 0)
    assert n > 0, n
    out = np.zeros((n, 5), np.float64)
    assert lib().orc_atlas_rule(C.byref(st), i, j, out.ctypes.data, n) == n
    return out


def atlas_node(st: AtlasSettings, i: int, j: int):
    """Oracle cell masses [nt, np] (mm^2) and report of node (i, j)."""
    mass = np.zeros(st.theta_count * st.phi_count, np.float64)
    rep = AtlasReport()
    rc = lib().orc_atlas_node(C.byref(st), i, j, mass.ctypes.data, C.byref(rep))
    assert rc == 0, rc
    return mass.reshape(st.theta_count, st.phi_count), rep


def ref_available() -> bool:
    return REF_SO.exists()


def ref() -> C.CDLL:
    global _ref
    if _ref is None:
        R = C.CDLL(str(REF_SO))
        R.ref_u_edges.argtypes = [u32, P(f64)]
        R.ref_make_frame.argtypes = [P(f64), P(f64), P(f64), P(f64)]
        R.ref_phase_bin.argtypes = [u32, u32, P(f64), u64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_void_p]
        R.ref_hex_trace.argtypes = [f64, f64, P(f64), f64, f64, u64, u64, u32, f64, u64, u64, C.c_void_p, u64,
                                    C.c_void_p]
        R.ref_hex_trace.restype = i64
        R.ref_hex_masses.argtypes = [f64, f64, P(f64), f64, f64, u64, u64, u32, f64, u32, u32, C.c_void_p,
                                     P(f64), P(f64)]
        R.ref_build_host_cdf.argtypes = [C.c_void_p, u32, u32, P(f64), u32, u32, f64, f64, f64, C.c_void_p,
                                         C.c_void_p, P(RefCdfReport)]
        R.ref_water_index.argtypes = [f64, f64, f64, P(f64), P(f64)]
        R.ref_projected_area.argtypes = [P(f64), f64, P(f64), P(f64), P(C.c_int)]
        R.ref_drop_shape.argtypes = [f32, C.c_int, P(f32), P(f32), P(f32)]
        R.ref_drop_surface.argtypes = [P(f32), f32, f32, C.c_int, u64, C.c_void_p, C.c_void_p, C.c_void_p]
        R.ref_drop_intersect.argtypes = [P(f32), f32, f32, P(f32), P(f32), f32, C.c_int, P(f32), P(f32), P(f32),
                                         P(C.c_int)]
        R.ref_drop_solve64.argtypes = [P(f32), f32, f32, P(f32), P(f32), f32, C.c_int, P(f64)]
        R.ref_drop_trace_vertices.argtypes = [P(f32), f32, f32, P(f32), P(f32), f32, f32, f32, u32, u32,
                                              C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        R.ref_orientation_masses.argtypes = [f64, f64, P(f64), f64, f64, u32, f64, C.c_char_p, u32, u32,
                                             C.c_void_p, P(f64)]
        _ref = R
    return _ref


def vec3(x) -> "C.Array":
    return (f64 * 3)(*[float(v) for v in x])


# ---------------------------------------------------------------- oracle helpers
def u_edges(n: int) -> np.ndarray:
    out = np.empty(n + 1, np.float64)
    assert lib().orc_u_edges(n, _ptr(out, P(f64))) == 0
    return out


def make_frame(k) -> Frame:
    f = Frame()
    assert lib().orc_make_frame(vec3(k), C.byref(f)) == 0
    return f


def drop_frame(k) -> Frame:
    f = Frame()
    assert lib().orc_drop_frame(vec3(k), C.byref(f)) == 0
    return f


def phase_bin(edges: np.ndarray, nt: int, np_: int, frame: Frame, d) -> Bin:
    return lib().orc_phase_bin(_ptr(edges, P(f64)), nt, np_, C.byref(frame), vec3(d))


def hex_settings(radius_mm=0.05, length_mm=0.2, ki=(1, -0.4, 0.3), n_int=1.31, n_ext=1.0, tail=1e-12,
                 seed=12345, total=1024, max_hits=12) -> HexSettings:
    s = HexSettings()
    s.radius_mm, s.length_mm = radius_mm, length_mm
    s.ki = (f64 * 3)(*ki)
    s.n_ext, s.n_int, s.tail, s.seed, s.total, s.max_hits = n_ext, n_int, tail, seed, total, max_hits
    return s


def hex_trace(s: HexSettings, first: int, count: int, cap: int = 1 << 20):
    samples = np.zeros(cap, SAMPLE_DTYPE)
    audits = np.zeros(count, AUDIT_DTYPE)
    n = lib().orc_hex_trace(C.byref(s), first, count, samples.ctypes.data, cap, audits.ctypes.data)
    assert n >= 0, n
    return samples[:n], audits


def hex_hist(s: HexSettings, nt: int, np_: int, qbits: int | None = None, first: int = 0,
             count: int | None = None):
    count = s.total if count is None else count
    qbits = lib().orc_quantum_bits(s.total) if qbits is None else qbits
    bins = np.zeros(nt * np_, np.uint64)
    st = HistStats()
    assert lib().orc_hex_hist(C.byref(s), nt, np_, qbits, first, count, bins.ctypes.data, C.byref(st)) == 0
    return bins, st


def bins_to_mass(bins: np.ndarray, qf: int, qb: int, nt: int, np_: int) -> np.ndarray:
    m = np.empty(nt * np_, np.float64)
    lib().orc_bins_to_mass(np.ascontiguousarray(bins, np.uint64).ctypes.data, qf, qb, nt, np_, m.ctypes.data)
    return m


def build_cdf(mass: np.ndarray, nt: int, np_: int, ont: int | None = None, onp: int | None = None,
              max_l1=1e-10, max_lost=1e-12, max_tv=1.0):
    ont = nt if ont is None else ont
    onp = np_ if onp is None else onp
    mass = np.ascontiguousarray(mass, np.float64)
    phi = np.zeros(onp + 1, np.float64)
    theta = np.zeros(onp * (ont + 1), np.float64)
    r = CdfReport()
    rc = lib().orc_build_cdf(mass.ctypes.data, nt, np_, ont, onp, max_l1, max_lost, max_tv, phi.ctypes.data,
                             theta.ctypes.data, C.byref(r))
    return rc, phi, theta.reshape(onp, ont + 1), r


def drop_record(coeffs=(0,) * 8, sphere=True, alpha_deg=0.0, n_int=1.3330, grid=64, max_hits=9,
                tail=0.0) -> DropRecord:
    r = DropRecord()
    c = (f64 * 8)(*[float(v) for v in coeffs])
    assert lib().orc_drop_make(c, 1 if sphere else 0, alpha_deg, n_int, grid, grid, max_hits, tail,
                               C.byref(r)) == 0
    return r


def drop_trace(r: DropRecord, precision: int = 32, first: int = 0, count: int | None = None, cap=1 << 22):
    count = r.grid_w * r.grid_h if count is None else count
    samples = np.zeros(cap, SAMPLE_DTYPE)
    audits = np.zeros(count, AUDIT_DTYPE)
    n = lib().orc_drop_trace(C.byref(r), precision, first, count, samples.ctypes.data, cap, audits.ctypes.data)
    assert n >= 0
    return samples[:n], audits


def drop_hist(r: DropRecord, nt: int, np_: int, precision: int = 32, qbits: int | None = None, first=0,
              count=None):
    total = r.grid_w * r.grid_h
    count = total if count is None else count
    qbits = lib().orc_quantum_bits(total) if qbits is None else qbits
    bins = np.zeros(nt * np_, np.uint64)
    st = HistStats()
    assert lib().orc_drop_hist(C.byref(r), precision, nt, np_, qbits, first, count, bins.ctypes.data,
                               C.byref(st)) == 0
    return bins, st


def sphere_theta_hist(grid: int, half_extent: float, n: float, max_hits: int, nt: int) -> np.ndarray:
    out = np.zeros(nt, np.float64)
    lib().orc_sphere_theta_hist(grid, half_extent, n, max_hits, nt, out.ctypes.data)
    return out


def quantum_bits(rays: int) -> int:
    return int(lib().orc_quantum_bits(rays))


def sphere_order(b: float, n: float, p: int):
    psi, ps, pp = f64(), f64(), f64()
    lib().orc_sphere_order(b, n, p, C.byref(psi), C.byref(ps), C.byref(pp))
    return psi.value, ps.value, pp.value


# ---------------------------------------------------------------- reference shim helpers
def ref_hex_trace(s: HexSettings, first: int, count: int, cap: int = 1 << 20):
    samples = np.zeros(cap, SAMPLE_DTYPE)
    audits = np.zeros(count, AUDIT_DTYPE)
    n = ref().ref_hex_trace(s.radius_mm, s.length_mm, s.ki, s.n_ext, s.n_int, s.seed, s.total, s.max_hits,
                            s.tail, first, count, samples.ctypes.data, cap, audits.ctypes.data)
    assert n >= 0, n
    return samples[:n], audits


def ref_phase_bin(nt: int, np_: int, k, dirs: np.ndarray):
    dirs = np.ascontiguousarray(dirs, np.float64).reshape(-1, 3)
    n = dirs.shape[0]
    index = np.zeros(n, np.uint64)
    pole = np.zeros(n, np.int32)
    valid = np.zeros(n, np.int32)
    snapped = np.zeros(n, np.int32)
    assert ref().ref_phase_bin(nt, np_, vec3(k), n, dirs.ctypes.data, index.ctypes.data, pole.ctypes.data,
                               valid.ctypes.data, snapped.ctypes.data) == 0
    return index, pole, valid, snapped


def ref_build_cdf(mass: np.ndarray, nt: int, np_: int, k, ont=None, onp=None, max_l1=1e-10, max_lost=1e-12,
                  max_tv=1.0):
    ont = nt if ont is None else ont
    onp = np_ if onp is None else onp
    mass = np.ascontiguousarray(mass, np.float64)
    phi = np.zeros(onp + 1, np.float64)
    theta = np.zeros(onp * (ont + 1), np.float64)
    r = RefCdfReport()
    rc = ref().ref_build_host_cdf(mass.ctypes.data, nt, np_, vec3(k), ont, onp, max_l1, max_lost, max_tv,
                                  phi.ctypes.data, theta.ctypes.data, C.byref(r))
    return rc, phi, theta.reshape(onp, ont + 1), r


def ref_hex_masses(s: HexSettings, nt: int, np_: int):
    mass = np.zeros(nt * np_, np.float64)
    fw, bw = f64(), f64()
    assert ref().ref_hex_masses(s.radius_mm, s.length_mm, s.ki, s.n_ext, s.n_int, s.seed, s.total, s.max_hits,
                                s.tail, nt, np_, mass.ctypes.data, C.byref(fw), C.byref(bw)) == 0
    return mass, fw.value, bw.value


def ref_water_index(lam_nm: float, celsius: float = 20.0, pressure: float = 101325.0):
    n, rho = f64(), f64()
    rc = ref().ref_water_index(lam_nm, celsius, pressure, C.byref(n), C.byref(rho))
    return rc, n.value, rho.value


def ref_projected_area(coeffs, radius_mm, k):
    area = f64()
    sph = C.c_int()
    rc = ref().ref_projected_area((f64 * 8)(*coeffs), radius_mm, vec3(k), C.byref(area), C.byref(sph))
    return rc, area.value, bool(sph.value)


def hex_hist_node(s: HexSettings, rot: np.ndarray, qfactor: float, nt: int, np_: int, qbits: int,
                  bins: np.ndarray | None = None, st: HistStats | None = None):
    """Accumulate one orientation node (rot: 3x3 body -> ensemble) into bins/st."""
    bins = np.zeros(nt * np_, np.uint64) if bins is None else bins
    st = HistStats() if st is None else st
    r = (f64 * 9)(*np.asarray(rot, np.float64).reshape(9))
    assert lib().orc_hex_hist_node(C.byref(s), r, qfactor, nt, np_, qbits, bins.ctypes.data, C.byref(st)) == 0
    return bins, st


def hex_node_area(s: HexSettings, rot) -> float:
    r = (f64 * 9)(*np.asarray(rot, np.float64).reshape(9))
    return float(lib().orc_hex_node_area(C.byref(s), r))


def aperture_transform(vx, vy, q1: float, q2: float) -> complex:
    vx = np.ascontiguousarray(vx, np.float64)
    vy = np.ascontiguousarray(vy, np.float64)
    re, im = f64(), f64()
    lib().orc_aperture_transform(_ptr(vx, P(f64)), _ptr(vy, P(f64)), len(vx), q1, q2, C.byref(re), C.byref(im))
    return complex(re.value, im.value)


def hex_aperture(radius_mm: float, length_mm: float, k):
    vx = np.zeros(8, np.float64)
    vy = np.zeros(8, np.float64)
    p1, p2 = (f64 * 3)(), (f64 * 3)()
    nv = lib().orc_hex_aperture(radius_mm, length_mm, vec3(k), _ptr(vx, P(f64)), _ptr(vy, P(f64)), p1, p2)
    assert nv >= 3
    return vx[:nv].copy(), vy[:nv].copy(), np.array(p1[:]), np.array(p2[:])


def polygon_area(vx, vy) -> float:
    return 0.5 * float(np.sum(vx * np.roll(vy, -1) - np.roll(vx, -1) * vy))


def lobe_cells(radius_mm, length_mm, k, lambda_nm, nt, np_, rows, frame: Frame | None = None):
    """Oracle lobe masses (mm^2) of the leading `rows` rows, frame = make_frame(k) by default."""
    vx, vy, p1, p2 = hex_aperture(radius_mm, length_mm, k)
    area = polygon_area(vx, vy)
    dmax = max(np.hypot(vx[i] - vx[j], vy[i] - vy[j]) for i in range(len(vx)) for j in range(len(vx)))
    g = make_frame(k) if frame is None else frame
    vxa, vya = np.ascontiguousarray(vx), np.ascontiguousarray(vy)
    out = np.zeros((rows, np_), np.float64)
    for i in range(rows):
        for j in range(np_):
            out[i, j] = lib().orc_lobe_cell(_ptr(vxa, P(f64)), _ptr(vya, P(f64)), len(vx), area, dmax, vec3(p1),
                                            vec3(p2), lambda_nm * 1e-6, C.byref(g), nt, np_, i, j)
    return out, area, dmax


def ref_drop_shape(radius_mm: float, sphere: bool = False):
    c = (f32 * 8)()
    inner, outer = f32(), f32()
    rc = ref().ref_drop_shape(radius_mm, 1 if sphere else 0, c, C.byref(inner), C.byref(outer))
    return rc, np.array(c[:], np.float32), inner.value, outer.value


def ref_orientation_masses(radius_mm, length_mm, k, n_ext, n_int, max_hits, tail, csv: str, nt, np_):
    mass = np.zeros(nt * np_, np.float64)
    area = f64()
    rc = ref().ref_orientation_masses(radius_mm, length_mm, vec3(k), n_ext, n_int, max_hits, tail, csv.encode(), nt,
                                      np_, mass.ctypes.data, C.byref(area))
    return rc, mass, area.value


# ---------------------------------------------------------------- drop geometry (oracle and reference)
def drop_surface(r: DropRecord, pts: np.ndarray, precision: int = 64):
    """Oracle f = r - R(mu) and outward gradient at points [n, 3]."""
    pts = np.ascontiguousarray(pts, np.float64)
    n = len(pts)
    f, g = np.zeros(n), np.zeros((n, 3))
    assert lib().orc_drop_surface(C.byref(r), precision, n, pts.ctypes.data, f.ctypes.data, g.ctypes.data) == 0
    return f, g


def drop_entry_point(r: DropRecord, u: float, v: float, precision: int = 64):
    p, n = (f64 * 3)(), (f64 * 3)()
    hit = lib().orc_drop_entry_point(C.byref(r), precision, u, v, p, n)
    return bool(hit), np.array(p[:]), np.array(n[:])


def drop_exit_point(r: DropRecord, p, d, precision: int = 64):
    q, n = (f64 * 3)(), (f64 * 3)()
    ok = lib().orc_drop_exit_point(C.byref(r), precision, vec3(p), vec3(d), q, n)
    return bool(ok), np.array(q[:]), np.array(n[:])


def drop_trace_origin(r: DropRecord, u: float, v: float, precision: int = 32, cap: int = 256):
    samples = np.zeros(cap, SAMPLE_DTYPE)
    audit = np.zeros(1, AUDIT_DTYPE)
    n = lib().orc_drop_trace_origin(C.byref(r), precision, u, v, samples.ctypes.data, cap, audit.ctypes.data)
    assert n >= 0
    return samples[:n], audit[0]


def _ref_shape(r: DropRecord):
    return (f32 * 8)(*r.coeffs[:]), r.inner, r.outer


def ref_drop_surface(r: DropRecord, pts: np.ndarray, precision: int = 64):
    """The reference's RaindropShape::implicit_value / gradient at points [n, 3]."""
    pts = np.ascontiguousarray(pts, np.float64)
    n = len(pts)
    f, g = np.zeros(n), np.zeros((n, 3))
    c, inner, outer = _ref_shape(r)
    assert ref().ref_drop_surface(c, inner, outer, precision, n, pts.ctypes.data, f.ctypes.data, g.ctypes.data) == 0
    return f, g


def ref_drop_intersect(r: DropRecord, origin, direction, start: float = 0.0, inside: bool = False):
    """The reference's RaindropIntersector::intersect: (code, pos, normal, distance, fp64 fallback)."""
    c, inner, outer = _ref_shape(r)
    pos, nrm, dist, fb = (f32 * 3)(), (f32 * 3)(), f32(), C.c_int()
    code = ref().ref_drop_intersect(c, inner, outer, (f32 * 3)(*origin), (f32 * 3)(*direction), start,
                                    1 if inside else 0, pos, nrm, C.byref(dist), C.byref(fb))
    return code, np.array(pos[:], np.float64), np.array(nrm[:], np.float64), dist.value, bool(fb.value)


def ref_drop_solve64(r: DropRecord, origin, direction, start: float = 0.0, inside: bool = False):
    c, inner, outer = _ref_shape(r)
    d = f64()
    code = ref().ref_drop_solve64(c, inner, outer, (f32 * 3)(*origin), (f32 * 3)(*direction), start,
                                  1 if inside else 0, C.byref(d))
    return code, d.value


def ref_drop_trace_vertices(r: DropRecord, w: int, h: int, reference_distance: float | None = None):
    """The reference's RaindropPathTracer::trace_vertex_unpolarized over a w x h
    vertex grid (families R, TT, TRT, TRRT, path-major)."""
    c, inner, outer = _ref_shape(r)
    n = 4 * w * h
    status = np.zeros(n, np.uint32)
    d, p = np.zeros((n, 3), np.float32), np.zeros((n, 3), np.float32)
    c0, c1 = np.zeros((n, 4), np.float32), np.zeros((n, 4), np.float32)
    dist = 2.0 * outer if reference_distance is None else reference_distance
    assert ref().ref_drop_trace_vertices(c, inner, outer, (f32 * 3)(*r.ki[:]), (f32 * 3)(*r.e0[:]), r.n_int,
                                         r.half_extent, dist, w, h, status.ctypes.data, d.ctypes.data, p.ctypes.data,
                                         c0.ctypes.data, c1.ctypes.data) == 0
    return (status.reshape(4, h * w), d.reshape(4, h * w, 3), p.reshape(4, h * w, 3), c0.reshape(4, h * w, 4),
            c1.reshape(4, h * w, 4))


# ---------------------------------------------------------------- operation census
CENSUS_SO = ROOT / "oracle" / "libcensus.so"


class Census(C.Structure):
    _fields_ = [("misses", u64), ("entries", u64), ("hits", u64), ("emissions", u64), ("ops_miss", u64),
                ("ops_entry", u64), ("ops_hit", u64), ("ops_emit", u64)]


_census = None


def drop_census(r: DropRecord, nt: int, np_: int, first: int, count: int) -> Census:
    """Counted operations and events of rays [first, first + count) of a drop
    record (oracle/census.cpp)."""
    global _census
    if _census is None:
        if not CENSUS_SO.exists():
            build_oracle()
        lib()
        _census = C.CDLL(str(CENSUS_SO))
        _census.orc_drop_census.argtypes = [P(DropRecord), u32, u32, u64, u64, P(Census)]
    out = Census()
    assert _census.orc_drop_census(C.byref(r), nt, np_, first, count, C.byref(out)) == 0
    return out


def oracle_env_ok() -> bool:
    return ORACLE_SO.exists() or (ROOT / "oracle" / "Makefile").exists()


os.environ.setdefault("ICEPHASE_ROOT", str(ROOT))
