"""The allocation-free CDF entry point (iph_build_cdf_noalloc) against the
allocating one (iph_build_cdf), through the C-ABI: the same phi and theta
arrays and the same report, bit for bit, on calls in a row whose second input is
smaller than the first (so it runs over scratch the first call left behind),
with and without a coarsened stored grid. Requests beyond the fixed scratch and
host pointers are refused, and a good call still works after each refusal.

Device arrays are allocated here through the HIP runtime libicephase.so is
linked against (not torch's bundled copy), so the pointers belong to the same
runtime instance as the library's context."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import ice_crystal_phase_amd as ip
from ice_crystal_phase_amd import icephase

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2
IPH_E_INVALID = -1


class _Hip:
    def __init__(self):
        ip.lib()  # loads libamdhip64.so.7 as the library's dependency
        self.h = C.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.live: list[int] = []

    def alloc(self, nbytes: int) -> int:
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), nbytes) == 0
        self.live.append(p.value)
        return p.value

    def upload(self, arr: np.ndarray) -> int:
        p = self.alloc(arr.nbytes)
        assert self.h.hipMemcpy(p, arr.ctypes.data, arr.nbytes, H2D) == 0
        return p

    def download(self, p: int, shape) -> np.ndarray:
        out = np.empty(shape, np.float64)
        assert self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0
        return out

    def free_all(self):
        for p in self.live:
            self.h.hipFree(p)
        self.live.clear()


@pytest.fixture(scope="module")
def hip():
    h = _Hip()
    yield h
    h.free_all()


def _masses(nt, np_, seed):
    rng = np.random.default_rng(seed)
    m = rng.random((nt, np_)) ** 4
    m[:, np_ // 3] = 0.0          # an empty column takes the caller's edges
    m[nt // 2, :] *= 1e3          # a bright row
    return m


def _noalloc(ctx, hip, mass, nt, np_, ont=0, onp=0):
    ont_, onp_ = ont or nt, onp or np_
    d_mass = hip.upload(np.ascontiguousarray(mass, np.float64).reshape(-1))
    d_phi = hip.alloc((onp_ + 1) * 8)
    d_theta = hip.alloc(onp_ * (ont_ + 1) * 8)
    pol = icephase.CdfPolicy(ont, onp, 1e-10, 1e-12, 1.0)
    rep = icephase.CdfReport()
    rc = ip.lib().iph_build_cdf_noalloc(ctx.ptr, d_mass, nt, np_, C.byref(pol), d_phi, d_theta, C.byref(rep))
    icephase._check(rc, ctx.ptr)
    return hip.download(d_phi, onp_ + 1), hip.download(d_theta, (onp_, ont_ + 1)), rep


def _same_report(a, b):
    for name, _ in icephase.CdfReport._fields_:
        va, vb = getattr(a, name), getattr(b, name)
        assert va == vb or (va != va and vb != vb), name


@pytest.mark.parametrize("cases", [
    # (nt, np, out_theta, out_phi) in call order: each second call is smaller
    [(180, 360, 0, 0), (90, 180, 0, 0)],
    [(360, 720, 120, 240), (90, 180, 45, 90)],
    [(900, 1800, 0, 0), (36, 72, 18, 24)],
], ids=["plain", "coarsened", "full-then-small"])
def test_noalloc_cdf_equals_build_cdf(ctx, hip, cases):
    for k, (nt, np_, ont, onp) in enumerate(cases):
        print(f"case: noalloc CDF call {k + 1}: {nt}x{np_} -> {ont or nt}x{onp or np_}", flush=True)
        mass = _masses(nt, np_, seed=nt + 7 * k)
        phi, theta, rep = _noalloc(ctx, hip, mass, nt, np_, ont, onp)
        want_phi, want_theta, want_rep, rc = ctx.build_cdf(mass, nt, np_, ont, onp)
        assert rc == 0 and rep.status == 0
        assert np.array_equal(phi, want_phi)
        assert np.array_equal(theta, want_theta)
        _same_report(rep, want_rep)
    hip.free_all()


def test_noalloc_cdf_refusals(ctx, hip):
    print("case: noalloc CDF refusals (oversize scratch, host pointers)", flush=True)
    nt, np_ = 1800, 3600
    d_mass = hip.upload(_masses(nt, np_, 3).reshape(-1))
    d_phi = hip.alloc((np_ + 1) * 8)
    d_theta = hip.alloc(np_ * (nt + 1) * 8)
    rep = icephase.CdfReport()
    # a coarsened 1800 x 1800 stored grid needs 25.9 MB of scratch: refused before any launch
    pol = icephase.CdfPolicy(1800, 1800, 1e-10, 1e-12, 1.0)
    rc = ip.lib().iph_build_cdf_noalloc(ctx.ptr, d_mass, nt, np_, C.byref(pol), d_phi, d_theta, C.byref(rep))
    assert rc == IPH_E_INVALID and b"bytes of scratch" in ip.lib().iph_last_error(ctx.ptr)
    # more than 4096 theta rows: refused
    rc = ip.lib().iph_build_cdf_noalloc(ctx.ptr, d_mass, 4098, 2, None, d_phi, d_theta, C.byref(rep))
    assert rc == IPH_E_INVALID
    # host arrays are refused (the kernels would read host addresses)
    host = np.zeros(90 * 180)
    out_phi, out_theta = np.zeros(181), np.zeros(180 * 91)
    rc = ip.lib().iph_build_cdf_noalloc(ctx.ptr, host.ctypes.data, 90, 180, None, out_phi.ctypes.data,
                                        out_theta.ctypes.data, C.byref(rep))
    assert rc == IPH_E_INVALID and b"device memory" in ip.lib().iph_last_error(ctx.ptr)
    # the same full-size grid without coarsening fits, and equals build_cdf
    mass = _masses(nt, np_, 3)
    phi, theta, rep = _noalloc(ctx, hip, mass, nt, np_)
    want_phi, want_theta, want_rep, rc = ctx.build_cdf(mass, nt, np_)
    assert rc == 0 and np.array_equal(phi, want_phi) and np.array_equal(theta, want_theta)
    _same_report(rep, want_rep)
    hip.free_all()
