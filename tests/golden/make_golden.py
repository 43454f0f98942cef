"""Write the golden fixtures from the compiled reference (oracle/_ref/libref.so,
built by `make -C oracle` from the reference's own CPU sources).

Each fixture holds inputs and the reference's outputs for them; the tests pin
the CPU oracle against these files, so the oracle stays pinned where the
reference checkout (and hence oracle/_ref) is absent.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import math
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
import oracle_bind as ob  # noqa: E402

C3 = dict(radius_mm=0.05, length_mm=0.2, ki=(1.0, -0.4, 0.3), n_int=1.31, tail=1e-12, seed=12345, max_hits=12)


def hex_samples():
    """trace_batch_cpu samples (bitwise) for three geometries."""
    out = {}
    cases = [("c3", dict(C3, total=384)),
             ("plate", dict(radius_mm=1.0, length_mm=0.4, ki=(0.2, -1.0, 0.7), n_int=1.31, tail=1e-12, seed=7,
                            max_hits=64, total=256)),
             ("slab_normal", dict(radius_mm=1.0, length_mm=2.0, ki=(0.0, -1.0, 0.0), n_int=1.31, tail=1e-15, seed=3,
                                  max_hits=2, total=16))]
    for name, kw in cases:
        s = ob.hex_settings(**kw)
        samp, aud = ob.ref_hex_trace(s, 0, kw["total"])
        out[f"{name}_settings"] = np.array([kw["radius_mm"], kw["length_mm"], *kw["ki"], kw["n_int"], kw["tail"],
                                            kw["seed"], kw["max_hits"], kw["total"]], np.float64)
        out[f"{name}_dir"] = samp["dir"].copy()
        out[f"{name}_power"] = samp["power"].copy()
        out[f"{name}_id"] = samp["id"].copy()
        out[f"{name}_faces"] = np.stack([samp["entry_face"], samp["exit_face"], samp["reflections"], samp["kind"]],
                                        1).astype(np.uint32)
        out[f"{name}_audit_f"] = np.stack([aud["escaped"], aud["unresolved"], aud["balance"], aud["max_interface"]], 1)
        out[f"{name}_audit_u"] = np.stack([aud["outputs"], aud["internal_hits"], aud["tir_count"], aud["status"]],
                                          1).astype(np.uint64)
    return out


def bins():
    """phase_bin on random and boundary directions (900 x 1800 grid)."""
    rng = np.random.default_rng(11)
    nt, np_ = 900, 1800
    out = {}
    for name, k in (("oblique", (1.0, -0.4, 0.3)), ("x", (1.0, 0.0, 0.0)), ("z", (0.0, 0.0, 1.0))):
        fr = ob.make_frame(k)
        ki, e0, e1 = (np.array(v) for v in (fr.ki, fr.e0, fr.e1))
        e = ob.u_edges(nt)
        d = list(rng.normal(size=(1500, 3)))
        for i in list(range(0, nt, 61)) + [1, nt - 1]:
            for j in list(range(0, np_, 97)) + [0, np_ - 1]:
                mu = 1 - 2 * e[i]
                st = 2 * math.sqrt(max(e[i] * (1 - e[i]), 0.0))
                phi = -math.pi + j * 2 * math.pi / np_
                v = ki * mu + e0 * (st * math.cos(phi)) + e1 * (st * math.sin(phi))
                d += [v, np.nextafter(v, v + 1), np.nextafter(v, v - 1)]
        d += [ki, -ki, -e0, e0]
        dirs = []
        for v in d:
            n = np.zeros(3)
            if ob.lib().orc_normalize3(ob.vec3(v), n.ctypes.data_as(ob.P(ob.f64))) == 0:
                dirs.append(n)
        dirs = np.array(dirs)
        idx, pole, valid, snapped = ob.ref_phase_bin(nt, np_, k, dirs)
        out[f"{name}_k"] = np.array(k, np.float64)
        out[f"{name}_dirs"] = dirs
        out[f"{name}_index"] = idx
        out[f"{name}_pole"] = pole
        out[f"{name}_valid"] = valid
        out[f"{name}_snapped"] = snapped
    return out


def cdf():
    """PhaseMassAccumulator masses and build_host_cdf arrays of a small record."""
    out = {}
    nt, np_ = 36, 72
    s = ob.hex_settings(**dict(C3, total=3000))
    mass, fw, bw = ob.ref_hex_masses(s, nt, np_)
    out["mass"] = mass
    out["mass_poles"] = np.array([fw, bw])
    for name, (ont, onp) in (("full", (nt, np_)), ("coarse", (18, 36))):
        rc, phi, theta, r = ob.ref_build_cdf(mass, nt, np_, (1.0, -0.4, 0.3), ont, onp)
        assert rc == 0
        out[f"{name}_phi"] = phi
        out[f"{name}_theta"] = theta
        out[f"{name}_report"] = np.array([r.g_source, r.g_stored, r.coarsening_tv, r.cdf_mass, r.l1_error,
                                          r.lost_mass, r.max_cell_error, r.aggregation_error, r.lost_cells])
    # sparse high-dynamic-range masses (hex_phase/tests/phase_tests.cpp:105-127)
    i = np.arange(48)[:, None]
    j = np.arange(96)[None, :]
    hdr = np.where(j % 7 == 0, 10.0 ** (-12.0 * ((17 * i + 13 * j) % 97) / 96), 0.0).reshape(-1)
    rc, phi, theta, r = ob.ref_build_cdf(hdr, 48, 96, (0.0, -1.0, 0.0), 24, 48)
    assert rc == 0
    out["hdr_mass"] = hdr
    out["hdr_phi"] = phi
    out["hdr_theta"] = theta
    return out


def materials():
    """evaluate_water_optics and compute_projected_area values."""
    rows = []
    for lam in (226.5, 400.0, 450.0, 550.0, 589.0, 650.0, 700.0, 1013.98):
        for t, p in ((0.0, 1e5), (20.0, 101325.0), (40.0, 1e7)):
            rc, n, rho = ob.ref_water_index(lam, t, p)
            assert rc == 0
            rows.append((lam, t, p, n, rho))
    areas = []
    for coeffs, k in (((0,) * 8, (1, 0, 0)), ((-0.0458, -0.0335, -0.1211, -0.0227, 0, 0, 0, 0), (1, 0, 0)),
                      ((-0.0458, -0.0335, -0.1211, -0.0227, 0, 0, 0, 0), (0, -1, 0)),
                      ((-0.0458, -0.0335, -0.1211, -0.0227, 0, 0, 0, 0), (0.6, -0.8, 0))):
        rc, a, sph = ob.ref_projected_area(coeffs, 2.0, k)
        assert rc == 0
        areas.append(list(coeffs) + list(k) + [a])
    return {"water": np.array(rows), "drop_area": np.array(areas)}


def main() -> None:
    if not ob.ref_available():
        ob.build_oracle()
    if not ob.ref_available():
        raise SystemExit("oracle/_ref/libref.so is not built (needs the reference checkout)")
    np.savez_compressed(HERE / "hex_samples.npz", **hex_samples())
    np.savez_compressed(HERE / "phase_bin.npz", **bins())
    np.savez_compressed(HERE / "cdf.npz", **cdf())
    np.savez_compressed(HERE / "materials.npz", **materials())
    for f in sorted(HERE.glob("*.npz")):
        print(f.name, f.stat().st_size)


if __name__ == "__main__":
    main()
