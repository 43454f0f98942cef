"""Orientation composition of deterministic single-particle atlases into phase
records, with no new ray tracing.

Drop-in for the reference's composer (hex_atlas/python/compose_atlas.py and
hex_atlas/python/atlas_math.py): the same composition config, the same
deterministic orientation quadrature (azimuth x Gaussian or isotropic tilt x
spin, fixed poses), D6h folding of each pose's incident direction into the
atlas's fundamental domain, bilinear incident interpolation of A-normalized
node masses, exact azimuth shift / reflection of the piecewise-constant cells,
the analytic uniform-azimuth orbit at vertical incidence, and the same output
tree (records/iNNNN/wNNNN with the record contract, orientation model files,
generation_plan.json and batch_plan.json, staged in <out>.part).

The coupling weights are assembled on the host, all poses at once as arrays
(the pose-by-pose rule is kept as coupling_scalar, the reference form it is tested against), the
next incident's in a worker thread while the current one is on the device. The two
array operations run on the device through libicephase.so: the fixed-order
shifted-block sum over every atlas node (iph_atlas_compose, bitwise the
reference's order and roundings) with the atlas resident in HBM across
conditions, and the FP64 CDF build (iph_build_cdf, the same ordered scans as
the reference's make_cdf). There is no CPU fallback.

Run as `python -m ice_crystal_phase_amd.compose --atlas DIR [DIR ...] --config
CONFIG --out DIR`.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import math
import sys
import time
from concurrent.futures import ThreadPoolExecutor
from contextlib import closing
from pathlib import Path
from typing import Any, Iterator

import numpy as np

from . import icephase as ip

PI = math.pi
EPS = float(np.finfo(np.float64).eps)
PACKAGE = Path(__file__).resolve().parent
GENERATOR_VERSION = "icephase.mi355x.atlas_composition.v1"


# ------------------------------------------------------------------ small helpers
def positive_int(x: Any, name: str, maximum: int = 1_000_000) -> int:
    """x itself when it is a JSON integer in [1, maximum]."""
    if type(x) is int and 1 <= x <= maximum:
        return x
    raise ValueError(f"{name} must be an integer in [1,{maximum}]")


def sha256_file(path: Path) -> str:
    return hashlib.sha256(path.read_bytes()).hexdigest()


def unit(v: Any) -> np.ndarray:
    """v / |v| for a finite nonzero three-vector; v is first divided by its
    largest magnitude, so the norm neither overflows nor underflows."""
    a = np.array(v, dtype=np.float64)
    if a.shape != (3,) or not np.isfinite(a).all():
        raise ValueError("Expected finite nonzero three-vector")
    big = np.abs(a).max()
    if big == 0:
        raise ValueError("Expected finite nonzero three-vector")
    a /= big
    return a / math.sqrt(float(a @ a))


def frame(ki: Any) -> np.ndarray:
    """Right-handed columns [e0 e1 k]: e0 is the world axis least aligned with
    k, made orthogonal to k; e1 = k x e0."""
    k = unit(ki)
    i = int(np.argmin(np.abs(k)))
    axis = np.zeros(3)
    axis[i] = 1.0
    e0 = unit(axis - k * k[i])
    return np.stack([e0, unit(np.cross(k, e0)), k], axis=1)


def _axis_rotation(axis: int, angle: float) -> np.ndarray:
    """Active right-handed rotation by `angle` about world axis 0 (x), 1 (y) or 2 (z)."""
    c, s = math.cos(angle), math.sin(angle)
    j, k = (axis + 1) % 3, (axis + 2) % 3      # the plane the rotation turns, in cyclic order
    m = np.eye(3)
    m[j, j] = m[k, k] = c
    m[k, j], m[j, k] = s, -s
    return m


def ry(a: float) -> np.ndarray:
    return _axis_rotation(1, a)


def rz(a: float) -> np.ndarray:
    return _axis_rotation(2, a)


def quaternion_matrix(q: Any) -> np.ndarray:
    """Active rotation of a unit quaternion (w, x, y, z):
    R = I + 2 w [v]x + 2 [v]x^2, v = (x, y, z)."""
    q = np.asarray(q, dtype=np.float64)
    if q.shape != (4,) or not np.isfinite(q).all() or abs(float(q @ q) - 1) > 2e-12:
        raise ValueError("Fixed quaternion must be unit wxyz")
    q = q / math.sqrt(float(q @ q))
    w, v = q[0], q[1:]
    cross = np.array([[0.0, -v[2], v[1]], [v[2], 0.0, -v[0]], [-v[1], v[0], 0.0]])
    return np.eye(3) + 2.0 * w * cross + 2.0 * (cross @ cross)


# ------------------------------------------------------------------ orientation quadrature
def gaussian_tilt_nodes(count: int, sigma_deg: float) -> np.ndarray:
    """Tilt nodes of a normal density N(pi/2, sigma^2) truncated to [0, pi]
    (measure d beta, no sin(beta) factor): the midpoints in probability of
    `count` equal strata, found by 64 bisection steps on the erf CDF. sigma = 0
    is a delta at pi/2 (atlas_math.py:70-104)."""
    positive_int(count, "tilt count")
    sigma = math.radians(float(sigma_deg))
    if not math.isfinite(sigma) or sigma < 0:
        raise ValueError("sigma_deg must be finite and nonnegative")
    if sigma == 0:
        return np.full(count, PI / 2)
    scale = math.sqrt(2) * sigma
    norm = math.erf(PI / (2 * math.sqrt(2) * sigma))
    if not norm > 0:
        raise ValueError("Gaussian normalization is not representable")
    nodes = []
    for k in range(count):
        target = 2 * ((k + 0.5) / count) - 1
        lo, hi = 0.0, PI
        for _ in range(64):
            mid = (lo + hi) / 2
            if math.erf((mid - PI / 2) / scale) / norm < target:
                lo = mid
            else:
                hi = mid
        nodes.append((lo + hi) / 2)
    out = np.asarray(nodes)
    if count > 1 and np.any(np.diff(out) <= 0):
        raise ValueError("Positive Gaussian width is below angular resolution; use explicit sigma_deg=0 for a delta")
    return out


def validate_rotations(r: np.ndarray, w: np.ndarray) -> None:
    if r.ndim != 3 or r.shape[1:] != (3, 3) or w.shape != (len(r),) or len(r) == 0:
        raise ValueError("Malformed orientation quadrature")
    if not np.isfinite(r).all() or not np.isfinite(w).all() or np.any(w <= 0) or abs(math.fsum(w) - 1) > 2e-12:
        raise ValueError("Nonfinite / nonpositive / unnormalized orientation weights")
    # R^T R = I and det R = 1 to 2e-12 (atlas_math.py:166), with the Gram
    # matrix and the cofactor determinant written out over all poses at once
    c = [r[:, :, i] for i in range(3)]
    gram_ok = all(np.max(np.abs(np.einsum("nk,nk->n", c[i], c[j]) - (i == j))) <= 2e-12
                  for i in range(3) for j in range(i, 3))
    det = (r[:, 0, 0] * (r[:, 1, 1] * r[:, 2, 2] - r[:, 1, 2] * r[:, 2, 1])
           - r[:, 0, 1] * (r[:, 1, 0] * r[:, 2, 2] - r[:, 1, 2] * r[:, 2, 0])
           + r[:, 0, 2] * (r[:, 1, 0] * r[:, 2, 1] - r[:, 1, 1] * r[:, 2, 0]))
    if not (gram_ok and np.max(np.abs(det - 1)) <= 2e-12):
        raise ValueError("Orientation nodes must be proper orthogonal rotations")


def orientation_rule(config: dict[str, Any]) -> tuple[np.ndarray, np.ndarray, dict[str, Any]]:
    """Body-to-ensemble rotations R = Ry(alpha) Rz(beta) Ry(gamma) (body axis
    +y) and NUMBER weights of every component (atlas_math.py:105-160): alpha at
    the midpoints of [0, 2 pi), beta isotropic (arccos midpoints) or Gaussian
    tilt nodes, gamma at the midpoints of [0, pi/3) (uniform spin of a regular
    hexagonal prism); a fixed component is one pose. Each component keeps its
    fraction."""
    if set(config) != {"quadrature", "components"}:
        raise ValueError("orientation keys must be quadrature and components")
    q = config["quadrature"]
    if set(q) != {"azimuth", "tilt", "spin"}:
        raise ValueError("quadrature keys must be azimuth, tilt, spin")
    na, nb, ng = (positive_int(q[k], k, 4096) for k in ("azimuth", "tilt", "spin"))
    components = config["components"]
    if not isinstance(components, list) or not components:
        raise ValueError("Nonempty orientation components required")
    fractions = [float(c["fraction"]) for c in components]
    if any(not math.isfinite(f) or not 0 < f <= 1 for f in fractions) or abs(math.fsum(fractions) - 1) > 1e-12:
        raise ValueError("Positive NUMBER fractions must sum to one")
    mats: list[np.ndarray] = []
    weights: list[float] = []
    counts: list[int] = []
    for comp, fraction in zip(components, fractions):
        kind = comp["kind"]
        extra = {"sigma_deg"} if kind == "horizontal_gaussian" else {"quaternion_wxyz"} if kind == "fixed" else set()
        if set(comp) != {"kind", "fraction"} | extra:
            raise ValueError(f"Unknown/missing component keys for {kind}")
        if kind == "fixed":
            mats.append(quaternion_matrix(comp["quaternion_wxyz"]))
            weights.append(fraction)
            counts.append(1)
            continue
        if kind == "isotropic":
            betas = np.arccos(1 - 2 * (np.arange(nb) + 0.5) / nb)
        elif kind == "horizontal_gaussian":
            betas = gaussian_tilt_nodes(nb, comp["sigma_deg"])
        else:
            raise ValueError("Supported models: isotropic, horizontal_gaussian, fixed")
        if na * nb * ng > 2_000_000:
            raise ValueError("Too many explicit quadrature nodes")
        share = fraction / (na * nb * ng)
        # the spin factors are the same for every (alpha, beta): one stacked
        # product per tilted frame gives the same matrices as one product per pose
        spins = np.stack([ry((PI / 3) * (ig + 0.5) / ng) for ig in range(ng)])
        for ia in range(na):
            azimuth = ry(2 * PI * (ia + 0.5) / na)
            for beta in betas:
                mats.extend(azimuth @ rz(float(beta)) @ spins)
        weights.extend([share] * (na * nb * ng))
        counts.append(na * nb * ng)
    r, w = np.asarray(mats), np.asarray(weights)
    validate_rotations(r, w)
    meta = {"kind": "deterministic_probability_midpoint_product_quadrature", "config": config,
            "component_nodes": counts, "rotation": "Ry(alpha)_Rz(beta)_Ry(gamma)_body_to_ensemble",
            "horizontal_density_measure": "d_beta; truncated normal on [0,pi]", "uniform_spin_range": [0, PI / 3],
            "randomness": False, "cdf_weights_analytic": True, "optical_integral_analytic": False}
    return r, w, meta


# ------------------------------------------------------------------ symmetry and interpolation
def canonical_frame(theta: float, phi: float) -> np.ndarray:
    """Columns [e0 e1 k] of the atlas's incident spherical frame at (theta, phi)."""
    st, ct, sp, cp = math.sin(theta), math.cos(theta), math.sin(phi), math.cos(phi)
    k = np.array([st * cp, ct, st * sp])
    e = np.array([ct * cp, -st, ct * sp])
    return np.column_stack((e, np.cross(k, e), k))


def fold_direction(direction: Any) -> tuple[float, float, np.ndarray]:
    """Canonical (theta, phi) in [0, pi/2] x [0, pi/6] and the D6h operation Q
    (possibly improper) with Q k(theta, phi) = direction: the sixfold sector,
    the mirror inside it, and the basal mirror (atlas_math.py:177-200)."""
    k = unit(direction)
    theta = math.atan2(math.hypot(k[0], k[2]), abs(k[1]))
    az = math.atan2(k[2], k[0]) % (2 * PI)
    sector = min(5, int(az / (PI / 3)))
    local = az - sector * (PI / 3)
    mirror = local > PI / 6
    phi = PI / 3 - local if mirror else local
    if phi < -64 * EPS or phi > PI / 6 + 64 * EPS:
        raise ValueError("Canonical azimuth outside fundamental region")
    phi = min(PI / 6, max(0.0, phi))
    angle = (sector + int(mirror)) * (PI / 3)
    q = ry(-angle) @ np.diag([1.0, -1.0 if k[1] < 0 else 1.0, -1.0 if mirror else 1.0])
    if np.max(np.abs(q @ canonical_frame(theta, phi)[:, 2] - k)) > 2e-12:
        raise ValueError("Symmetry unfolding is inconsistent")
    return theta, phi, q


def interpolation(t: float, p: float, ntheta: int, nphi: int) -> list[tuple[int, int, float]]:
    """Bilinear weights of the endpoint-inclusive node grid at (t, p); no
    extrapolation outside the fundamental domain (atlas_math.py:201-211)."""
    inside = (math.isfinite(t) and math.isfinite(p) and -64 * EPS <= t <= PI / 2 + 64 * EPS
              and -64 * EPS <= p <= PI / 6 + 64 * EPS)
    if not inside:
        raise ValueError("Incident interpolation outside canonical domain")
    x = min(float(ntheta), max(0.0, t / (PI / 2) * ntheta))
    y = min(float(nphi), max(0.0, p / (PI / 6) * nphi))
    i, j = min(ntheta - 1, int(x)), min(nphi - 1, int(y))
    a, b = x - i, y - j
    return [(i + di, j + dj, u * v) for di, u in ((0, 1 - a), (1, a)) for dj, v in ((0, 1 - b), (1, b)) if u * v > 0]


def projected_area(radius: float, length: float, k: Any) -> float:
    """Projected area of the prism along k (faces' areas times max(0, -n.k))."""
    k = unit(k)
    s3 = math.sqrt(3)
    normals = np.array([[1, 0, 0], [0.5, 0, s3 / 2], [-0.5, 0, s3 / 2], [-1, 0, 0], [-0.5, 0, -s3 / 2],
                        [0.5, 0, -s3 / 2], [0, 1, 0], [0, -1, 0]])
    areas = np.array([radius * length] * 6 + [3 * s3 * radius * radius / 2] * 2)
    return float(np.dot(areas, np.maximum(0, -normals @ k)))


# ------------------------------------------------------------------ atlas reader
class Atlas:
    """A complete atlas directory (mass_mm2.npy memory-mapped) with the node
    audits of atlas_math.py:222-270; `device(ctx)` uploads the masses once."""

    def __init__(self, directory: str | Path):
        self.directory = Path(directory)
        if self.directory.name.endswith(".part"):
            raise ValueError("Incomplete atlas directory")
        self.meta = json.loads((self.directory / "metadata.json").read_text(encoding="utf-8"))
        m = self.meta
        if m.get("schema") != "ice.single_particle_atlas.v1" or m.get("complete") is not True:
            raise ValueError("Incomplete or unknown atlas schema")
        if (m.get("incident_reconstruction") != "bilinear_theta_phi_in_spherical_incident_frame"
                or m.get("symmetry") != "D6h_O3_full_24_scalar_unpolarized"):
            raise ValueError("Unsupported atlas frame/interpolation/symmetry")
        self.it, self.ip = (positive_int(m[k], k, 4096) for k in ("incident_theta_intervals", "incident_phi_intervals"))
        self.nt, self.np = (positive_int(m[k], k) for k in ("theta_count", "phi_count"))
        self.mass = np.load(self.directory / "mass_mm2.npy", mmap_mode="r", allow_pickle=False)
        self.edges = np.load(self.directory / "u_edges.npy", allow_pickle=False)
        if (self.mass.shape != (self.it + 1, self.ip + 1, self.nt, self.np) or self.mass.dtype.str != "<f8"
                or not self.mass.flags.c_contiguous):
            raise ValueError("Atlas mass layout does not match metadata")
        e = self.edges
        if (e.shape != (self.nt + 1,) or e.dtype.str != "<f8" or not np.isfinite(e).all() or e[0] != 0 or e[-1] != 1
                or np.any(np.diff(e) <= 0)):
            raise ValueError("Invalid atlas u edges")
        if len(m["nodes"]) != (self.it + 1) * (self.ip + 1):
            raise ValueError("Missing incident atlas nodes")
        g = m["geometry"]
        self.radius, self.length = float(g["circumradius_mm"]), float(g["full_length_mm"])
        if g["shape"] != "regular_hexagonal_prism" or not (
                math.isfinite(self.radius) and math.isfinite(self.length) and self.radius > 0 and self.length > 0):
            raise ValueError("Unsupported geometry")
        self.area = np.zeros((self.it + 1, self.ip + 1))
        self.residual = np.zeros_like(self.area)
        for index, n in enumerate(m["nodes"]):
            i, j = divmod(index, self.ip + 1)
            if (n["i"], n["j"]) != (i, j):
                raise ValueError("Node order mismatch")
            f = canonical_frame(i * (PI / 2) / self.it, j * (PI / 6) / self.ip)
            if (not np.allclose(n["ki"], f[:, 2], rtol=0, atol=2e-12)
                    or not np.allclose(n["e0"], f[:, 0], rtol=0, atol=2e-12)):
                raise ValueError("Node frame mismatch")
            cell = self.mass[i, j]
            if not np.isfinite(cell).all() or np.any(cell < 0):
                raise ValueError("Invalid atlas cell mass")
            area, residual, escaped = (float(n[k]) for k in ("projected_area_mm2", "unresolved_mm2", "escaped_mm2"))
            if (not all(math.isfinite(v) for v in (area, residual, escaped)) or area <= 0 or escaped <= 0
                    or residual < 0):
                raise ValueError("Invalid node accounting")
            if (abs((escaped + residual) / area - 1) > 2e-10 or abs(float(np.sum(cell)) / escaped - 1) > 2e-10
                    or abs(area / projected_area(self.radius, self.length, n["ki"]) - 1) > 2e-12):
                raise ValueError("Node area / mass / residual audit failed")
            self.area[i, j], self.residual[i, j] = area, residual
        self.hashes = {name: sha256_file(self.directory / name)
                       for name in ("metadata.json", "mass_mm2.npy", "u_edges.npy")}
        self._device: ip.AtlasMass | None = None

    def device(self, ctx: ip.Context) -> ip.AtlasMass:
        if self._device is None or self._device.ctx is not ctx:
            self._device = ip.AtlasMass(ctx, np.asarray(self.mass))
        return self._device

    def close(self) -> None:
        if self._device is not None:
            self._device.close()
            self._device = None
        if isinstance(self.mass, np.memmap):
            self.mass._mmap.close()


# ------------------------------------------------------------------ coupling
_NORMALS = np.array([[1, 0, 0], [0.5, 0, math.sqrt(3) / 2], [-0.5, 0, math.sqrt(3) / 2], [-1, 0, 0],
                     [-0.5, 0, -math.sqrt(3) / 2], [0.5, 0, -math.sqrt(3) / 2], [0, 1, 0], [0, -1, 0]])


def coupling(atlas: Atlas, rotations: np.ndarray, number_weights: np.ndarray,
             ki: Any) -> tuple[np.ndarray, float, np.ndarray]:
    """Convolution weights [node_theta, node_phi, parity, shift] that multiply
    the RAW node masses: for each pose, the body-frame incident direction is
    folded into the fundamental domain, the A-normalized node masses are
    interpolated and scaled by the exact projected area A(s), and the output
    azimuth offset of the folded frame is split between the two neighbouring
    whole-cell shifts (atlas_math.py:272-307).

    All poses are processed as arrays; the weights are accumulated pose by
    pose in the scalar rule's order (np.add.at applies its terms in sequence),
    so the result equals coupling_scalar() up to the last-bit differences of
    array versus scalar elementary functions (checked to 1e-13)."""
    validate_rotations(rotations, number_weights)
    out_frame = frame(ki)
    k = out_frame[:, 2]
    n = len(rotations)
    # body-frame incident directions, normalized as unit() does
    body = np.einsum("nji,j->ni", rotations, k)
    body = body / np.max(np.abs(body), axis=1, keepdims=True)
    body = body / np.linalg.norm(body, axis=1, keepdims=True)
    # D6h fold (fold_direction, vectorized)
    theta = np.arctan2(np.hypot(body[:, 0], body[:, 2]), np.abs(body[:, 1]))
    az = np.mod(np.arctan2(body[:, 2], body[:, 0]), 2 * PI)
    sector = np.minimum(5, np.floor(az / (PI / 3)).astype(np.int64))
    local = az - sector * (PI / 3)
    mirror = local > PI / 6
    phi = np.where(mirror, PI / 3 - local, local)
    if np.any(phi < -64 * EPS) or np.any(phi > PI / 6 + 64 * EPS):
        raise ValueError("Canonical azimuth outside fundamental region")
    phi = np.minimum(PI / 6, np.maximum(0.0, phi))
    angle = (sector + mirror.astype(np.int64)) * (PI / 3)
    c, s_ = np.cos(-angle), np.sin(-angle)
    sy = np.where(body[:, 1] < 0, -1.0, 1.0)
    sz = np.where(mirror, -1.0, 1.0)
    # q = Ry(-angle) diag(1, sy, sz)
    q = np.zeros((n, 3, 3))
    q[:, 0, 0], q[:, 0, 2] = c, s_ * sz
    q[:, 1, 1] = sy
    q[:, 2, 0], q[:, 2, 2] = -s_, c * sz
    st, ct, sp, cp = np.sin(theta), np.cos(theta), np.sin(phi), np.cos(phi)
    kc = np.stack([st * cp, ct, st * sp], axis=1)
    ec = np.stack([ct * cp, -st, ct * sp], axis=1)
    if np.max(np.abs(np.einsum("nij,nj->ni", q, kc) - body)) > 2e-12:
        raise ValueError("Symmetry unfolding is inconsistent")
    e = np.einsum("nij,nj->ni", rotations, np.einsum("nij,nj->ni", q, ec))
    shift = np.arctan2(e @ out_frame[:, 1], e @ out_frame[:, 0]) / (2 * PI) * atlas.np
    nearest = np.round(shift)
    shift = np.where(np.abs(shift - nearest) <= 64 * EPS * atlas.np, nearest, shift)
    lower = np.floor(shift)
    frac = shift - lower
    lower = lower.astype(np.int64)
    parity = (sy * sz < 0).astype(np.int64)
    face_areas = np.array([atlas.radius * atlas.length] * 6 + [3 * math.sqrt(3) * atlas.radius * atlas.radius / 2] * 2)
    areas = np.maximum(0, -body @ _NORMALS.T) @ face_areas  # projected_area() of every pose
    nw = np.asarray(number_weights, dtype=np.float64)
    expected_area = math.fsum(map(float, nw * areas))
    # bilinear incident interpolation (interpolation(), vectorized)
    if (np.any(~np.isfinite(theta)) or np.any(theta < -64 * EPS) or np.any(theta > PI / 2 + 64 * EPS)):
        raise ValueError("Incident interpolation outside canonical domain")
    x = np.minimum(float(atlas.it), np.maximum(0.0, theta / (PI / 2) * atlas.it))
    y = np.minimum(float(atlas.ip), np.maximum(0.0, phi / (PI / 6) * atlas.ip))
    i0 = np.minimum(atlas.it - 1, x.astype(np.int64))
    j0 = np.minimum(atlas.ip - 1, y.astype(np.int64))
    ax, by = x - i0, y - j0
    w = np.zeros((atlas.it + 1, atlas.ip + 1, 2, atlas.np), dtype=np.float64)
    # terms in the scalar order: pose, (di, dj) in (0,0), (0,1), (1,0), (1,1), then lower, lower + 1,
    # as flat indices into w; np.add.at applies them one after another in that order. A corner the
    # scalar rule skips (zero bilinear weight) adds an exact +0.0 here, which leaves every sum as it is.
    index = np.empty((n, 4, 2), dtype=np.int64)
    value = np.empty((n, 4, 2), dtype=np.float64)
    shifts = (lower % atlas.np, (lower + 1) % atlas.np)
    for corner, ((di, u_), (dj, v_)) in enumerate((a, b) for a in ((0, 1 - ax), (1, ax))
                                                  for b in ((0, 1 - by), (1, by))):
        mix = u_ * v_
        ii, jj = i0 + di, j0 + dj
        cval = np.where(mix > 0, nw * areas * mix / atlas.area[ii, jj], 0.0)
        block = ((ii * (atlas.ip + 1) + jj) * 2 + parity) * atlas.np
        index[:, corner, 0] = block + shifts[0]
        index[:, corner, 1] = block + shifts[1]
        value[:, corner, 0] = cval * (1 - frac)
        value[:, corner, 1] = cval * frac
    np.add.at(w.reshape(-1), index.reshape(-1), value.reshape(-1))
    represented = float(np.sum(np.sum(w, axis=(2, 3)) * atlas.area))
    if abs(represented / expected_area - 1) > 1e-11 or np.any(w < 0) or not np.isfinite(w).all():
        raise ValueError("Orientation-to-convolution area audit failed")
    return w, expected_area, out_frame


def coupling_scalar(atlas: Atlas, rotations: np.ndarray, number_weights: np.ndarray,
                    ki: Any) -> tuple[np.ndarray, float, np.ndarray]:
    """coupling() one pose at a time with the scalar operations of
    atlas_math.py:272-307 (the reference form the array version is tested against)."""
    validate_rotations(rotations, number_weights)
    out_frame = frame(ki)
    k = out_frame[:, 2]
    w = np.zeros((atlas.it + 1, atlas.ip + 1, 2, atlas.np), dtype=np.float64)
    expected = []
    for r, nw in zip(rotations, number_weights):
        body = r.T @ k
        theta, phi, q = fold_direction(body)
        e = r @ q @ canonical_frame(theta, phi)[:, 0]
        shift = math.atan2(float(e @ out_frame[:, 1]), float(e @ out_frame[:, 0])) / (2 * PI) * atlas.np
        nearest = round(shift)
        if abs(shift - nearest) <= 64 * EPS * atlas.np:
            shift = float(nearest)  # a rounding-sized offset is a whole-cell shift
        lower = math.floor(shift)
        frac = shift - lower
        parity = int(np.linalg.det(q) < 0)
        a = projected_area(atlas.radius, atlas.length, body)
        expected.append(float(nw) * a)
        for i, j, mix in interpolation(theta, phi, atlas.it, atlas.ip):
            c = float(nw) * a * mix / atlas.area[i, j]
            w[i, j, parity, lower % atlas.np] += c * (1 - frac)
            w[i, j, parity, (lower + 1) % atlas.np] += c * frac
    expected_area = math.fsum(expected)
    represented = float(np.sum(np.sum(w, axis=(2, 3)) * atlas.area))
    if abs(represented / expected_area - 1) > 1e-11 or np.any(w < 0) or not np.isfinite(w).all():
        raise ValueError("Orientation-to-convolution area audit failed")
    return w, expected_area, out_frame


def coupling_for_rule(atlas: Atlas, rotations: np.ndarray, number_weights: np.ndarray, ki: Any,
                      rule_meta: dict[str, Any]) -> tuple[np.ndarray, float, np.ndarray, dict[str, Any]]:
    """coupling(), except at exactly vertical incidence (k = +-y), where each
    isotropic or Gaussian component's uniform world azimuth only rotates the
    output azimuth: its orbit is integrated exactly by spreading every
    tilt/spin contribution evenly over all shifts (atlas_math.py:308-393).
    Fixed poses and non-vertical directions use the plain rule."""
    k = unit(ki)
    vertical = bool(k[0] == 0.0 and k[2] == 0.0)
    config = rule_meta["config"]
    components = config["components"]
    eligible = [c["kind"] in ("isotropic", "horizontal_gaussian") for c in components]
    if not vertical or not any(eligible):
        w, area, f = coupling(atlas, rotations, number_weights, k)
        return w, area, f, {"method": "original_fixed_orientation_quadrature", "exact_vertical_incidence": vertical,
                            "analytic_uniform_azimuth_components": [], "effective_pose_evaluations": len(rotations),
                            "symmetry_axis_ensemble": [0, 1, 0], "direction_snapped": False,
                            "output_post_filter": False}
    ref_r, ref_w, _ = orientation_rule(config)
    if not np.array_equal(rotations, ref_r) or not np.array_equal(number_weights, ref_w):
        raise ValueError("Analytic azimuth requires the declared orientation_rule arrays")
    result = np.zeros((atlas.it + 1, atlas.ip + 1, 2, atlas.np), dtype=np.float64)
    areas, modes = [], []
    evaluations = 0
    for comp, analytic in zip(components, eligible):
        fraction = float(comp["fraction"])
        quadrature = dict(config["quadrature"])
        if analytic:
            quadrature["azimuth"] = 1  # one representative alpha: the orbit is integrated below
        r, nw, _ = orientation_rule({"quadrature": quadrature, "components": [dict(comp, fraction=1.0)]})
        wc, ac, out_frame = coupling(atlas, r, nw, k)
        if analytic:
            wc = np.broadcast_to(np.sum(wc, axis=-1, keepdims=True, dtype=np.float64) / atlas.np, wc.shape)
        result += fraction * wc
        areas.append(fraction * ac)
        evaluations += len(r)
        modes.append("analytic_uniform_alpha" if analytic else "fixed_pose")
    area = math.fsum(areas)
    represented = float(np.sum(np.sum(result, axis=(2, 3)) * atlas.area))
    if not area > 0 or not np.isfinite(result).all() or np.any(result < 0) or abs(represented / area - 1) > 1e-11:
        raise ValueError("Analytic azimuth coupling failed area audit")
    return result, area, out_frame, {
        "method": "analytic_uniform_world_azimuth_at_vertical_incidence", "exact_vertical_incidence": True,
        "analytic_uniform_azimuth_components": [i for i, x in enumerate(eligible) if x],
        "component_methods": modes, "reference_pose_count": len(rotations),
        "effective_pose_evaluations": evaluations, "symmetry_axis_ensemble": [0, 1, 0],
        "direction_snapped": False, "output_post_filter": False,
        "tilt_spin_and_atlas_interpolation_remain_numerical": True}


def coupling_stream(atlas: Atlas, rotations: np.ndarray, number_weights: np.ndarray, kis: list[np.ndarray],
                    rule_meta: dict[str, Any]) -> Iterator[tuple[tuple[np.ndarray, float, np.ndarray, dict[str, Any]],
                                                               float]]:
    """coupling_for_rule() of each incident direction in order, each with the
    seconds it took, the next one computed in a worker thread while the caller
    works on the current one (the device sum, the CDF and the record files).
    Closing the stream waits for the coupling in flight; an error of a coupling
    is raised when its result is taken."""
    def timed(k: np.ndarray) -> tuple[tuple[np.ndarray, float, np.ndarray, dict[str, Any]], float]:
        start = time.perf_counter()
        result = coupling_for_rule(atlas, rotations, number_weights, k, rule_meta)
        return result, time.perf_counter() - start

    with ThreadPoolExecutor(max_workers=1) as pool:
        ahead = pool.submit(timed, kis[0]) if kis else None
        for index in range(len(kis)):
            current = ahead.result()
            ahead = pool.submit(timed, kis[index + 1]) if index + 1 < len(kis) else None
            yield current


# ------------------------------------------------------------------ device operations
def compose_mass(ctx: ip.Context, atlas: Atlas, weights: np.ndarray) -> tuple[np.ndarray, float]:
    """Composed cell mass [nt, np] (iph_atlas_compose) and the composed
    unresolved area (sum of node weights times node residuals)."""
    out = atlas.device(ctx).compose(weights)
    residual = float(np.sum(np.sum(weights, axis=(2, 3)) * atlas.residual))
    if not np.isfinite(out).all() or np.any(out < 0):
        raise ValueError("Invalid composed mass")
    return out, residual


def make_cdf(ctx: ip.Context, mass: np.ndarray, edges: np.ndarray, max_l1: float = 1e-10,
             max_lost: float = 1e-12) -> tuple[np.ndarray, np.ndarray, dict[str, Any]]:
    """phi marginal [np+1] and theta|phi CDF [np, nt+1] of a cell-mass array
    [nt, np] (iph_build_cdf: ordered FP64 scans of mass/max, audited), with the
    report of atlas_math.py:412-454."""
    if mass.ndim != 2 or not np.isfinite(mass).all() or np.any(mass < 0) or not float(np.max(mass)) > 0:
        raise ValueError("CDF requires finite nonnegative nonempty mass")
    nt, np_ = mass.shape
    if (edges.shape != (nt + 1,) or not np.isfinite(edges).all() or edges[0] != 0 or edges[-1] != 1
            or np.any(np.diff(edges) <= 0)):
        raise ValueError("Invalid CDF edges")
    phi, theta, rep, rc = ctx.build_cdf(mass, nt, np_, max_l1=max_l1, max_lost=max_lost, check=False)
    if rc != 0:
        raise ValueError(f"CDF audit rejected l1={rep.l1_error} lost={rep.lost_mass} g={rep.g_stored}")
    # an empty azimuth column's conditional CDF is the edge list itself: the
    # caller's (the device fills the standard u edges, equal for an atlas grid)
    empty = ~(mass > 0).any(axis=0)
    theta[empty] = edges
    return phi, theta, {"cdf_mass": rep.cdf_mass, "cdf_l1_mass_error": rep.l1_error,
                        "cdf_lost_probability_mass": rep.lost_mass,
                        "cdf_maximum_cell_mass_error": rep.max_cell_error, "g": rep.g_stored,
                        "resolved_mass_mm2": math.fsum(mass.flat)}


# ------------------------------------------------------------------ records
def _inclination_direction(alpha: float) -> np.ndarray:
    """ki = (cos a, -sin a, 0), with the exact axis at a = +-90 deg."""
    if not (math.isfinite(alpha) and -90 <= alpha <= 90):
        raise ValueError("Inclination must be in [-90,90] degrees")
    if abs(alpha) == 90:
        return np.array([0.0, -math.copysign(1.0, alpha), 0.0])
    r = math.radians(alpha)
    return unit([math.cos(r), -math.sin(r), 0.0])


def incidents(config: dict[str, Any]) -> list[tuple[np.ndarray, dict[str, Any]]]:
    """The incident conditions of a composition config: `inclinations_deg`
    (canonical directions in the x-y plane) or explicit `directions`."""
    given = [k for k in ("inclinations_deg", "directions") if k in config]
    if len(given) != 1:
        raise ValueError("Specify exactly one of inclinations_deg or directions")
    if given[0] == "inclinations_deg":
        out = [(_inclination_direction(float(a)), {"inclination_deg": float(a)}) for a in config["inclinations_deg"]]
    else:
        out = [(unit(k), {}) for k in config["directions"]]
    if not 1 <= len(out) <= 10000:
        raise ValueError("Require 1..10000 incident conditions")
    return out


def _canonical_incidence(k: np.ndarray) -> bool:
    """The rainbow record contract's incident directions: (cos a, -sin a, 0), a in [-90, 90]."""
    return abs(k[2]) <= 2e-14 and k[0] >= -2e-14


def _record_contract_keys(a: Atlas) -> dict[str, Any]:
    return dict(schema="rainbow.phase_cdf.numpy.v2", complete=True, generator_version=GENERATOR_VERSION,
                backend="hip_fixed_order_atlas_convolution", dtype="<f8", order="C", theta_count=a.nt,
                phi_count=a.np, density_measure="solid_angle_sr", cell_model="constant_density_per_spherical_cell",
                coordinates="u=(1-cos(theta))/2; v=(phi+pi)/(2*pi)", cdf_axis_order=["phi_cell", "theta_edge"],
                normalization="integral_p_domega_equals_one", input_polarization="unpolarized",
                direction_convention="physical_propagation", theta_zero="forward", theta_pi="backward")


def _frame_keys(f: np.ndarray, contract: str) -> dict[str, Any]:
    rainbow = contract == "rainbow"
    keys = dict(source_repository="ice_crystal_phase_amd", source_base_commit="unknown", azimuth_periodic=True,
                particle_frame_handedness="right", particle_up_axis=[0, 1, 0], shape_polar_axis=[0, -1, 0],
                phi_range_rad=[-math.pi, math.pi],
                transport_reference_frame="ensemble_reference; not an individual crystal body frame",
                shape_polar_axis_semantics="oriented transport conditioning axis -y; body c-axis is "
                                           "geometry.body_long_axis",
                coordinate_contract_id="rainbow.phase_cdf.coordinates.v1" if rainbow
                else "ice.phase_cdf.ensemble_coordinates.v1",
                sampling_frame_layout="rows_xyz_columns_e0_e1_k",
                sampling_frame_map="local_column_to_particle_column" if rainbow
                else "local_column_to_ensemble_reference_column",
                sampling_frame_columns=f.tolist(), query_frame_axis=f[:, 2].tolist(), query_frame_e0=f[:, 0].tolist(),
                nf_condition_axes=["wavelength_nm", "dot([0,-1,0],incident_propagation)"],
                geometry_and_orientation_fixed_within_dataset=True)
    if rainbow:
        k = f[:, 2]
        if not _canonical_incidence(k):
            raise ValueError("RainbowReference requires canonical ki=(cos(alpha),-sin(alpha),0); no implicit "
                             "rotation of a non-axisymmetric ensemble")
        keys["incident_inclination_degrees"] = math.degrees(math.atan2(-k[1], max(0.0, k[0])))
    return keys


def _physics_keys(a: Atlas, expected: float) -> dict[str, Any]:
    material = {"wavelength_convention": "vacuum_nm", **a.meta["material"]}
    geometry = {**a.meta["geometry"], "projected_area_mm2": expected, "projected_area_m2": expected * 1e-6,
                "projected_area_method": "deterministic_orientation_quadrature_of_exact_facet_projection"}
    mm2 = {"scattering": expected, "extinction": expected, "absorption": 0}
    sections = {f"{k}_mm2": v for k, v in mm2.items()}
    sections.update({f"{k}_m2": (v * 1e-6 if v else 0) for k, v in mm2.items()})
    return dict(wavelength_nm=material["wavelength_nm"], interior_index=material["interior_index"],
                exterior_index=material["exterior_index"], geometry=geometry, material=material,
                cross_sections=dict(model="geometric_nonabsorbing", **sections,
                                    wave_scattering_cross_section_computed=False),
                single_scattering_albedo=1, stage="incoherent",
                diffraction_policy=dict(computed=False, selected_output_includes_diffraction=False),
                inter_path_interference=False, tir_relative_phase_retained=True)


def _provenance_keys(a: Atlas, rule_meta: dict[str, Any], plan_hash: str, source_hash: str,
                     azimuth_integration: dict[str, Any]) -> dict[str, Any]:
    return dict(
        source_sha256=source_hash,
        orientation=dict(representation="deterministic_SO3_number_quadrature", rule=rule_meta,
                         model_parameter_policy="explicit_user_parameters_not_aerodynamic_prediction",
                         plan_sha256=plan_hash, randomness=False, plan_embedded=False, plan_location="dataset_root",
                         azimuth_integration=azimuth_integration),
        atlas=dict(files_sha256=a.hashes, incident_theta_intervals=a.it, incident_phi_intervals=a.ip,
                   angular_reconstruction="A_normalized_bilinear_theta_phi_in_incident_spherical_frame",
                   symmetry="D6h_including_reflections", transform="phi_shift_or_phi_reflection_then_shift",
                   azimuth_remap="exact_overlap_of_piecewise_constant_equal_width_cells",
                   analytical_average_of_continuous_optics=False, no_retracing=True))


def _result_keys(expected: float, residual: float, report: dict[str, Any], f: np.ndarray) -> dict[str, Any]:
    g, escaped = report["g"], report["resolved_mass_mm2"]
    quality = {k: v for k, v in report.items() if k not in ("g", "resolved_mass_mm2")}
    certified = ("source_coverage", "angular_convergence", "orientation_convergence", "incident_atlas_convergence",
                 "continuous_wave_solution")
    quality.update({f"{k}_certified": False for k in certified})
    quality.update(maximum_accepted_l1_mass_error=1e-10, maximum_accepted_lost_probability_mass=1e-12)
    return dict(
        transport_audit=dict(escaped_mm2=escaped, unresolved_mm2=residual, unresolved_fraction=residual / expected,
                             balance_error=(escaped + residual) / expected - 1,
                             cdf_normalized_to_resolved_escaped_mass=True,
                             residual_estimator="same_covariant_incident_interpolation_as_response"),
        storage_processing=dict(gaussian_sigma_degrees=0, coarsening_tv=0, mass_conservative_aggregation=False,
                                g_source=g, g_stored=g),
        hg=dict(g=g, axis_particle_frame=f[:, 2].tolist(), method="first_moment_of_saved_cell_pdf",
                target="saved_cdf", cosine_convention="dot(incident_propagation,outgoing_propagation)"),
        quality=quality)


def metadata(a: Atlas, expected: float, residual: float, report: dict[str, Any], f: np.ndarray, plan_hash: str,
             rule_meta: dict[str, Any], condition: dict[str, Any], source_hash: str,
             azimuth_integration: dict[str, Any], output_contract: str = "rainbow") -> dict[str, Any]:
    """metadata.json of a composed record (the key set of the reference
    composer's records, hex_atlas/python/compose_atlas.py:63-122), assembled
    from its sections: record contract, frame, physics, provenance, results."""
    meta = _frame_keys(f, output_contract)
    meta.update(_record_contract_keys(a))
    meta.update(_physics_keys(a, expected))
    meta.update(_provenance_keys(a, rule_meta, plan_hash, source_hash, azimuth_integration))
    meta.update(_result_keys(expected, residual, report, f))
    meta["condition"] = condition
    return meta


def _check_batch(atlases: list[Atlas]) -> list[float]:
    """Atlases composed in one run must differ only in wavelength (and the
    index it implies): same geometry, grids, areas, exterior medium, material
    model and transport policy. Returns their wavelengths."""
    if not atlases:
        raise ValueError("At least one atlas is required")
    ref = atlases[0]
    for a in atlases[1:]:
        same_grid = (a.it, a.ip, a.nt, a.np) == (ref.it, ref.ip, ref.nt, ref.np)
        if not (same_grid and a.meta["geometry"] == ref.meta["geometry"] and np.array_equal(a.edges, ref.edges)
                and np.array_equal(a.area, ref.area)):
            raise ValueError("Batch atlases must share geometry, angular grids, and areas")
        if any(a.meta["material"].get(k) != ref.meta["material"].get(k)
               for k in ("model", "exterior_index", "absorption", "birefringence")):
            raise ValueError("A spectral batch must use one material model and exterior medium")
        if a.meta["transport"] != ref.meta["transport"]:
            raise ValueError("A spectral batch must use one transport/audit policy")
    waves = [float(a.meta["material"]["wavelength_nm"]) for a in atlases]
    if len(set(waves)) != len(waves) or not all(math.isfinite(w) and w > 0 for w in waves):
        raise ValueError("Unique positive wavelength per atlas required")
    if (ref.it + 1) * (ref.ip + 1) * 2 * ref.np * 8 > 512 * 1024 ** 2:
        raise ValueError("Coupling buffer would exceed the documented 512 MiB bound")
    return waves


def _write_json(path: Path, value: Any) -> None:
    with path.open("x", encoding="utf-8", newline="\n") as f:
        json.dump(value, f, ensure_ascii=True, allow_nan=False, indent=2)
        f.write("\n")


def _write_npy(path: Path, value: np.ndarray) -> None:
    with path.open("xb") as f:
        np.save(f, np.ascontiguousarray(value, dtype="<f8"), allow_pickle=False)


def _staging(out: Path) -> Path:
    if out.name in ("", ".", "..") or out.name.endswith(".part") or out.exists():
        raise ValueError("Output must be a new named directory, not .part")
    part = out.with_name(out.name + ".part")
    out.parent.mkdir(parents=True, exist_ok=True)
    part.mkdir()  # exclusive: an incomplete run is neither removed nor adopted
    return part


def run(atlas_paths: list[Path], config_path: Path, out: Path, output_contract: str = "rainbow",
        device: int = 0) -> None:
    config = json.loads(config_path.read_text(encoding="utf-8"))
    if "orientation" not in config or not set(config) <= {"orientation", "inclinations_deg", "directions"}:
        raise ValueError("Unknown/missing composition configuration keys")
    incident_list = incidents(config)
    atlases: list[Atlas] = []
    part = None
    ctx = None
    try:
        atlases = [Atlas(path) for path in atlas_paths]
        first = atlases[0] if atlases else None
        rotations, weights, rule_meta = orientation_rule(config["orientation"])
        rule_meta = {**rule_meta, "vertical_uniform_azimuth_policy": {
            "version": "analytic_uniform_alpha_at_exact_vertical_v1",
            "reference_arrays": "discrete_reference_not_literal_nodes_at_analytic_conditions",
            "eligible_components": ["horizontal_gaussian", "isotropic"], "fixed_components_unchanged": True,
            "nonvertical_conditions_unchanged": True}}
        if output_contract not in ("rainbow", "ensemble"):
            raise ValueError("Unknown output contract")
        if output_contract == "rainbow" and not all(_canonical_incidence(k) for k, _ in incident_list):
            raise ValueError("Noncanonical incident direction is not RainbowReference-compatible; select "
                             "--output-contract ensemble explicitly")
        wavelengths = _check_batch(atlases)
        part = _staging(out)
        _write_json(part / "orientation_model.json", rule_meta)
        _write_npy(part / "orientation_rotations.npy", rotations)
        _write_npy(part / "orientation_weights.npy", weights)
        plan_files = {n: sha256_file(part / n)
                      for n in ("orientation_model.json", "orientation_rotations.npy", "orientation_weights.npy")}
        plan_hash = hashlib.sha256(json.dumps(plan_files, sort_keys=True).encode("ascii")).hexdigest()
        source_hash = hashlib.sha256("".join(sha256_file(PACKAGE / n) for n in ("compose.py", "icephase.py"))
                                     .encode("ascii")).hexdigest()
        plan: dict[str, Any] = dict(
            schema="ice.deterministic_phase_dataset.v1", complete=False,
            nf_condition_order=["wavelength_nm", "dot([0,-1,0],incident_propagation)"],
            output_contract=output_contract, per_record_schema="rainbow.phase_cdf.numpy.v2",
            condition_axes=["incident", "wavelength"], incident_directions=[k.tolist() for k, _ in incident_list],
            wavelengths_nm=wavelengths, orientation_plan_sha256=plan_hash, orientation_files_sha256=plan_files,
            input_config=config, atlases=[a.hashes for a in atlases], source_sha256=source_hash,
            computation="deterministic_offline_mass_convolution_no_raytracing", records=[])
        _write_json(part / "generation_plan.json", plan)
        ctx = ip.Context(device)
        # the coupling of the next incident is computed in a worker thread while
        # this one's records are composed on the device and written
        with closing(coupling_stream(first, rotations, weights, [k for k, _ in incident_list],
                                     rule_meta)) as couplings:
            for i, ((ki, condition), ((c, expected, fr, azimuth), seconds)) in enumerate(zip(incident_list,
                                                                                           couplings)):
                start = time.perf_counter() - seconds
                for w, a in enumerate(atlases):
                    mass, residual = compose_mass(ctx, a, c)
                    total = math.fsum(mass.flat)
                    if (abs((total + residual) / expected - 1) > 2e-10
                            or residual / expected > a.meta["transport"]["max_unresolved_fraction"] + 2e-15):
                        raise ValueError("Composed mass and unresolved area fail the transport audit")
                    phi, theta, report = make_cdf(ctx, mass, a.edges)
                    relative = Path("records") / f"i{i:04d}" / f"w{w:04d}"
                    record = part / relative
                    record.mkdir(parents=True)
                    _write_npy(record / "phi_cdf.npy", phi)
                    _write_npy(record / "theta_given_phi_cdf.npy", theta)
                    _write_npy(record / "u_edges.npy", a.edges)
                    _write_json(record / "metadata.json", metadata(a, expected, residual, report, fr, plan_hash,
                                                                  rule_meta, condition, source_hash, azimuth,
                                                                  output_contract))
                    files = {n: sha256_file(record / n)
                             for n in ("metadata.json", "phi_cdf.npy", "theta_given_phi_cdf.npy", "u_edges.npy")}
                    plan["records"].append(dict(incident_index=i, wavelength_index=w, path=relative.as_posix(),
                                                condition=[wavelengths[w], -float(ki[1])], ki=ki.tolist(),
                                                wavelength_nm=wavelengths[w], g=report["g"], files_sha256=files,
                                                azimuth_integration=azimuth))
                    print(f"record i={i} w={w} mass=1 g={report['g']:.17g} A={expected:.17g}", flush=True)
                print(f"incident_complete i={i} offline_seconds={time.perf_counter() - start:.6f}", flush=True)
        plan["complete"] = True
        _write_json(part / "batch_plan.json", plan)
        if out.exists():
            raise FileExistsError("Destination appeared while composing")
        part.rename(out)
    except Exception as e:
        if part is not None and part.exists() and not (part / "failure.json").exists():
            _write_json(part / "failure.json", dict(complete=False, error=str(e)))
        raise
    finally:
        for a in atlases:
            a.close()
        if ctx is not None:
            ctx.close()


def main(argv: list[str] | None = None) -> None:
    p = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    p.add_argument("--atlas", nargs="+", type=Path, required=True)
    p.add_argument("--config", type=Path, required=True)
    p.add_argument("--out", type=Path, required=True)
    p.add_argument("--output-contract", choices=["rainbow", "ensemble"], default="rainbow",
                   help="rainbow: strict canonical incidents; ensemble: explicit arbitrary-vector records")
    p.add_argument("--device", type=int, default=0)
    args = p.parse_args(argv)
    try:
        run(args.atlas, args.config, args.out, args.output_contract, args.device)
    except Exception as e:
        print(f"compose_atlas failed: {e!a}", file=sys.stderr)
        raise SystemExit(2)


if __name__ == "__main__":
    main()
