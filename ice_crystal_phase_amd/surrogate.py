"""Learned surrogate of the phase-function records (host side of the
iph_sur_* C-ABI in include/icephase.h).

A conditional density MLP over the (u, phi) cells, conditioned on a record's
(inclination deg, wavelength nm, equal-volume radius mm, aspect ratio), trained
on the GPU by the library's BF16 MFMA kernels on the cell probabilities of
traced records (copied device to device into the surrogate's training set),
and served as FP64 CDF records by the same CDF kernels as traced records. The
plain tracer stays the reference and the fallback: `compare` measures total
variation and |g| error against traced records on the device, `fit_gate`
turns such measurements on a validation set into the domain the network may
serve, and `serve` traces every record outside it.
"""
from __future__ import annotations

import ctypes as C
import json
import math
import time
from dataclasses import asdict, dataclass, field

import numpy as np

from . import icephase as icp


# ---------------------------------------------------------------- drop family
def drop_aspect(coeffs) -> float:
    """Axis ratio of the drop r(chi): vertical semi-extent (r(mu=1) + r(mu=-1))/2
    over the horizontal semi-extent max r sin chi."""
    c = list(coeffs) + [0.0] * (8 - len(coeffs))
    chi = np.linspace(0.0, np.pi, 20001)
    r = 1 + c[0] + sum(c[n] * np.cos(n * chi) for n in range(1, 8))
    vert = 0.5 * (r[0] + r[-1])
    return float(vert / np.max(r * np.sin(chi)))


def drop_conditions(n: int, seed: int, alpha=(-90.0, 90.0), wavelength=(400.0, 700.0),
                    radius=(0.25, 3.0)) -> np.ndarray:
    """n seeded records of the water-drop family: (alpha, lambda, radius, aspect),
    the aspect that of the Table I shape of the radius."""
    rng = np.random.default_rng(seed)
    a = rng.uniform(*alpha, n)
    lam = rng.uniform(*wavelength, n)
    r = rng.uniform(*radius, n)
    asp = np.array([drop_aspect(icp.drop_shape(float(np.float32(x)))) for x in r])
    return np.stack([a, lam, r, asp], axis=1).astype(np.float32)


def drop_records(conds: np.ndarray, grid: int, theta: int, phi: int, max_internal_hits: int = 9):
    """Tracer records of the family (Table I shape, IAPWS water index at 20 C)."""
    recs = []
    for a, lam, r, _ in np.asarray(conds, np.float64):
        n = icp.water_optics(float(np.float32(lam))).refractive_index
        coeffs = icp.drop_shape(float(np.float32(r)))
        if all(v == 0 for v in coeffs):
            rec = icp.sphere_record(n, grid, theta, phi, radius_mm=float(r), inclination_deg=float(a),
                                    max_internal_hits=max_internal_hits, wavelength_nm=float(lam))
        else:
            rec = icp.drop_record(coeffs, n, grid, theta, phi, radius_mm=float(r), inclination_deg=float(a),
                                  max_internal_hits=max_internal_hits, wavelength_nm=float(lam))
        recs.append(rec)
    return recs


# ---------------------------------------------------------------- network
class Surrogate:
    """Device state of one network and its training set (iph_sur)."""

    def __init__(self, ctx: icp.Context, theta: int = 450, phi: int = 900, max_records: int = 64,
                 dataset: int = 0, seed: int = 5):
        self.ctx = ctx
        self.theta, self.phi, self.max_records, self.dataset = theta, phi, max_records, dataset
        self.M = theta * phi
        self._p = C.c_void_p()
        cfg = icp.SurConfig(theta, phi, max_records, dataset, seed)
        icp._check(icp.lib().iph_sur_create(ctx.ptr, C.byref(cfg), C.byref(self._p)), ctx.ptr)

    def close(self):
        if self._p and self.ctx.ptr:
            icp.lib().iph_sur_destroy(self._p)
        self._p = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @staticmethod
    def param_count() -> int:
        return int(icp.lib().iph_sur_param_count())

    def params(self) -> np.ndarray:
        out = np.zeros(self.param_count(), np.float32)
        icp._check(icp.lib().iph_sur_get_params(self._p, out.ctypes.data), self.ctx.ptr)
        return out

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32).reshape(-1)
        if p.size != self.param_count():
            raise ValueError(f"expected {self.param_count()} parameters (this network's layout), got {p.size}")
        icp._check(icp.lib().iph_sur_set_params(self._p, p.ctypes.data), self.ctx.ptr)

    def set_record(self, slot: int, batch: icp.Batch, rec: int, cond) -> None:
        c = (C.c_float * 4)(*[float(x) for x in cond])
        icp._check(icp.lib().iph_sur_set_record(self._p, slot, batch.ptr, rec, c), self.ctx.ptr)

    def set_target(self, slot: int, probs: np.ndarray, cond) -> None:
        probs = np.ascontiguousarray(probs, np.float32).reshape(-1)
        if probs.size != self.M:
            raise ValueError(f"expected {self.M} cell probabilities, got {probs.size}")
        c = (C.c_float * 4)(*[float(x) for x in cond])
        icp._check(icp.lib().iph_sur_set_target(self._p, slot, probs.ctypes.data, c), self.ctx.ptr)

    def set_loss(self, tv_weight: float) -> None:
        """Train on cross entropy + tv_weight * total variation (0: cross entropy alone)."""
        icp._check(icp.lib().iph_sur_set_loss(self._p, float(tv_weight)), self.ctx.ptr)

    def gradient(self, slot: int):
        """Loss and packed gradient of one training slot."""
        loss = C.c_float()
        g = np.zeros(self.param_count(), np.float32)
        icp._check(icp.lib().iph_sur_gradient(self._p, slot, C.byref(loss), g.ctypes.data), self.ctx.ptr)
        return float(loss.value), g

    def train_step(self, slots, lr: float) -> np.ndarray:
        slots = np.ascontiguousarray(slots, np.uint32)
        loss = np.zeros(slots.size, np.float32)
        icp._check(icp.lib().iph_sur_train_step(self._p, slots.ctypes.data, slots.size, lr, loss.ctypes.data),
                   self.ctx.ptr)
        return loss

    def predict(self, conds: np.ndarray) -> np.ndarray:
        """Cell probabilities [n, M] of n <= max_records records (host)."""
        conds = np.ascontiguousarray(conds, np.float32).reshape(-1, 4)
        out = np.zeros((conds.shape[0], self.M), np.float32)
        icp._check(icp.lib().iph_sur_predict(self._p, conds.ctypes.data, conds.shape[0], out.ctypes.data),
                   self.ctx.ptr)
        return out

    def compare(self, conds: np.ndarray, batch: icp.Batch, recs) -> tuple[np.ndarray, np.ndarray]:
        """(TV, |dg|) of the predictions against traced records of a batch."""
        conds = np.ascontiguousarray(conds, np.float32).reshape(-1, 4)
        n = conds.shape[0]
        r = np.ascontiguousarray(recs, np.uint32)
        tv = np.zeros(n)
        dg = np.zeros(n)
        icp._check(icp.lib().iph_sur_compare(self._p, conds.ctypes.data, n, batch.ptr, r.ctypes.data, tv.ctypes.data,
                                             dg.ctypes.data), self.ctx.ptr)
        return tv, dg

    def records(self, conds: np.ndarray, out_theta: int = 0, out_phi: int = 0):
        """Served records: FP64 CDF arrays (phi [n, onp+1], theta [n, onp, ont+1])."""
        conds = np.ascontiguousarray(conds, np.float32).reshape(-1, 4)
        n = conds.shape[0]
        ont, onp = out_theta or self.theta, out_phi or self.phi
        phi = np.zeros((n, onp + 1), np.float64)
        theta = np.zeros((n, onp, ont + 1), np.float64)
        reps = (icp.CdfReport * n)()
        pol = icp.CdfPolicy(out_theta, out_phi, 1e-10, 1e-12, 1.0)
        icp._check(icp.lib().iph_sur_records(self._p, conds.ctypes.data, n, C.byref(pol), phi.ctypes.data,
                                             theta.ctypes.data, reps), self.ctx.ptr)
        return phi, theta, list(reps)

    def timing(self):
        ms = (icp.f64 * 3)()
        icp._check(icp.lib().iph_sur_timing(self._p, ms))
        return ms[0], ms[1], ms[2]


def fill_training_set(sur: Surrogate, conds: np.ndarray, grid: int, chunk: int = 64,
                      max_internal_hits: int = 9, log=None) -> float:
    """Trace every record of `conds` and copy its cell probabilities into the
    surrogate's training slots 0..n-1. Returns the device trace time (ms)."""
    ms = 0.0
    for k0 in range(0, len(conds), chunk):
        if log and k0 and k0 % 1024 == 0:
            log(f"traced {k0}/{len(conds)} training records ({ms / 1e3:.1f} s device)")
        recs = drop_records(conds[k0:k0 + chunk], grid, sur.theta, sur.phi, max_internal_hits)
        with icp.Batch(sur.ctx, recs) as b:
            b.run(icp.RUN_TRACE | icp.RUN_GRAPH)
            ms += b.timing()[0]
            for i in range(len(recs)):
                sur.set_record(k0 + i, b, i, conds[k0 + i])
    return ms


def evaluate(sur: Surrogate, conds: np.ndarray, grid: int, max_internal_hits: int = 9):
    """TV and |dg| of the surrogate against fresh traced records; also the
    device time the tracer spent on them (ms)."""
    tv, dg = [], []
    ms = 0.0
    for k0 in range(0, len(conds), sur.max_records):
        c = conds[k0:k0 + sur.max_records]
        recs = drop_records(c, grid, sur.theta, sur.phi, max_internal_hits)
        with icp.Batch(sur.ctx, recs) as b:
            b.run(icp.RUN_TRACE | icp.RUN_GRAPH)
            ms += b.timing()[0]
            t, g = sur.compare(c, b, list(range(len(recs))))
        tv.append(t)
        dg.append(g)
    return np.concatenate(tv), np.concatenate(dg), ms


@dataclass
class TrainReport:
    epochs: int
    steps: int
    seconds: float
    final_loss: float
    losses: list


def train(sur: Surrogate, n_records: int, epochs: int, batch: int = 8, lr: float = 2e-3, lr_min: float = 2e-5,
          seed: int = 0, time_limit: float | None = None, log=None) -> TrainReport:
    """Adam with a cosine learning-rate schedule over the shuffled training
    slots 0..n_records-1; with a time limit the schedule follows the elapsed
    fraction of it (whichever of steps and time is further along)."""
    steps_per_epoch = max(1, n_records // batch)
    total = epochs * steps_per_epoch
    rng = np.random.default_rng(seed)
    t0 = time.time()
    losses = []
    step = 0
    lr_t = lr
    for ep in range(epochs):
        order = rng.permutation(n_records).astype(np.uint32)
        ep_loss = 0.0
        for s in range(steps_per_epoch):
            frac = step / max(1, total - 1)
            if time_limit:
                frac = max(frac, min(1.0, (time.time() - t0) / time_limit))
            lr_t = lr_min + 0.5 * (lr - lr_min) * (1 + math.cos(math.pi * frac))
            l = sur.train_step(order[s * batch:(s + 1) * batch], lr_t)
            ep_loss += float(l.mean())
            step += 1
        losses.append(ep_loss / steps_per_epoch)
        if log:
            log(f"epoch {ep + 1}/{epochs} loss {losses[-1]:.6f} lr {lr_t:.2e} t {time.time() - t0:.1f}s")
        if time_limit is not None and time.time() - t0 > time_limit:
            break
    return TrainReport(len(losses), step, time.time() - t0, losses[-1], losses)


# ---------------------------------------------------------------- validated domain and fallback
RADIUS_EDGES = tuple(0.25 + 0.25 * k for k in range(12))  # 0.25 .. 3.0 mm, 11 bins
ALPHA_EDGES = (0.0, 30.0, 60.0, 90.0 + 1e-6)              # |inclination| deg, 3 bins


@dataclass
class Gate:
    """Validated domain of a trained network: the (radius, |inclination|) cells
    whose validation records reached total variation <= tv_limit at the given
    quantile. Records in the other cells (or outside the edges) are traced.
    accept, tv_quantile, tv_mean and counts are [radius bin][alpha bin]."""
    tv_limit: float
    quantile: float
    radius_edges: list
    accept: list
    tv_quantile: list = field(default_factory=list)
    tv_mean: list = field(default_factory=list)
    counts: list = field(default_factory=list)
    validation: dict = field(default_factory=dict)
    alpha_edges: list = field(default_factory=lambda: [0.0, 90.0 + 1e-6])

    def __post_init__(self):
        # a gate over radius bins only: one alpha bin
        if self.accept and not isinstance(self.accept[0], list):
            self.accept = [[bool(a)] for a in self.accept]
            self.tv_quantile = [[x] for x in self.tv_quantile]
            self.tv_mean = [[x] for x in self.tv_mean]
            self.counts = [[x] for x in self.counts]

    def cells(self, conds: np.ndarray):
        c = np.asarray(conds, np.float64).reshape(-1, 4)
        kr = np.searchsorted(np.asarray(self.radius_edges), c[:, 2], side="right") - 1
        ka = np.searchsorted(np.asarray(self.alpha_edges), np.abs(c[:, 0]), side="right") - 1
        ok = (kr >= 0) & (kr < len(self.radius_edges) - 1) & (ka >= 0) & (ka < len(self.alpha_edges) - 1)
        return kr, ka, ok

    def accepts(self, conds: np.ndarray) -> np.ndarray:
        kr, ka, ok = self.cells(conds)
        acc = np.asarray(self.accept, bool)
        out = np.zeros(kr.shape, bool)
        out[ok] = acc[kr[ok], ka[ok]]
        return out

    def save(self, path) -> None:
        with open(path, "w") as f:
            json.dump(asdict(self), f, indent=1)

    @staticmethod
    def load(path) -> "Gate":
        with open(path) as f:
            return Gate(**json.load(f))


def fit_gate(sur: Surrogate, conds: np.ndarray, grid: int, tv_limit: float = 1e-2, quantile: float = 0.9,
             radius_edges=RADIUS_EDGES, alpha_edges=ALPHA_EDGES, min_count: int = 8) -> Gate:
    """Measure the network against traced validation records (grid^2 rays) and
    accept the (radius, |inclination|) cells holding at least min_count records
    whose TV quantile is within tv_limit."""
    tv, dg, _ = evaluate(sur, conds, grid)
    g = Gate(tv_limit, quantile, list(radius_edges), [[False] * (len(alpha_edges) - 1)] * (len(radius_edges) - 1),
             alpha_edges=list(alpha_edges))
    kr, ka, ok = g.cells(conds)
    accept, tq, tm, n = [], [], [], []
    for i in range(len(radius_edges) - 1):
        row = [[], [], [], []]
        for j in range(len(alpha_edges) - 1):
            t = tv[ok & (kr == i) & (ka == j)]
            q = float(np.quantile(t, quantile)) if t.size else float("nan")
            row[0].append(bool(t.size >= min_count and q <= tv_limit))
            row[1].append(q)
            row[2].append(float(t.mean()) if t.size else float("nan"))
            row[3].append(int(t.size))
        accept.append(row[0]), tq.append(row[1]), tm.append(row[2]), n.append(row[3])
    c = np.asarray(conds, np.float64)
    g.accept, g.tv_quantile, g.tv_mean, g.counts = accept, tq, tm, n
    g.validation = {"records": int(len(conds)), "grid": int(grid), "tv_mean": float(tv.mean()),
                    "dg_mean": float(dg.mean()),
                    "per_record": [[round(float(a), 3), round(float(l), 2), round(float(r), 4), round(float(t), 6)]
                                   for (a, l, r, _), t in zip(c, tv)]}
    return g


@dataclass
class Served:
    phi: np.ndarray          # [n, np + 1]
    theta: np.ndarray        # [n, np, nt + 1]
    by_surrogate: np.ndarray  # [n] bool
    surrogate_ms: float      # device time of the network + CDF records
    tracer_ms: float         # device time of the fallback trace + CDF


def serve(sur: Surrogate, gate: Gate, conds: np.ndarray, grid: int, max_internal_hits: int = 9) -> Served:
    """FP64 CDF records of `conds`: the network for records inside the gate's
    domain, the tracer (grid^2 rays) for the others."""
    conds = np.ascontiguousarray(conds, np.float32).reshape(-1, 4)
    n = conds.shape[0]
    use = gate.accepts(conds)
    phi = np.zeros((n, sur.phi + 1))
    theta = np.zeros((n, sur.phi, sur.theta + 1))
    s_ms = t_ms = 0.0
    idx = np.flatnonzero(use)
    for k0 in range(0, idx.size, sur.max_records):
        part = idx[k0:k0 + sur.max_records]
        p, t, _ = sur.records(conds[part])
        phi[part], theta[part] = p, t
        s_ms += sur.timing()[1]
    idx = np.flatnonzero(~use)
    for k0 in range(0, idx.size, 64):
        part = idx[k0:k0 + 64]
        recs = drop_records(conds[part], grid, sur.theta, sur.phi, max_internal_hits)
        with icp.Batch(sur.ctx, recs) as b:
            b.run(icp.RUN_TRACE | icp.RUN_CDF | icp.RUN_GRAPH)
            tr, cdf, _ = b.timing()
            t_ms += tr + cdf
            for i, r in enumerate(part):
                phi[r], theta[r], _ = b.cdf(i)
    return Served(phi, theta, use, s_ms, t_ms)
