"""Train the learned surrogate on tracer records and evaluate it against fresh
traced records.

    python -m ice_crystal_phase_amd.surrogate_train --records 4096 --grid 8192 \
        --time-limit 900 --out surrogate.npy --gate 256 --gate-out surrogate.gate.json
    python -m ice_crystal_phase_amd.surrogate_train --epochs 0 --init surrogate.npy \
        --gate 256 --gate-out surrogate.gate.json          # evaluation and gate only

Training set: `--records` water-drop records (seed 5) at theta x phi cells,
traced with grid^2 rays each and copied device to device into the
surrogate's training slots. Evaluation: `--eval` fresh records (seed 6) traced
with eval_grid^2 rays, total variation and |g| error of the network's
prediction against each. Prints one JSON line with the losses, TV and timing.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import icephase as icp
from . import surrogate as S


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--records", type=int, default=4096)
    ap.add_argument("--grid", type=int, default=2048, help="rays per axis of a training record")
    ap.add_argument("--theta", type=int, default=450)
    ap.add_argument("--phi", type=int, default=900)
    ap.add_argument("--epochs", type=int, default=200)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--lr-min", type=float, default=2e-5)
    ap.add_argument("--time-limit", type=float, default=None, help="seconds of training")
    ap.add_argument("--eval", type=int, default=64)
    ap.add_argument("--eval-grid", type=int, default=8192)
    ap.add_argument("--init", default=None, help="start from these parameters (.npy)")
    ap.add_argument("--out", default=None, help="write the trained parameters (.npy)")
    ap.add_argument("--tv-weight", type=float, default=0.0,
                    help="train on cross entropy + this weight x total variation (0: cross entropy alone)")
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--gate", type=int, default=0,
                    help="validation records (seed + 2) for the served domain; 0 = no gate")
    ap.add_argument("--gate-out", default=None, help="write the gate (.json)")
    ap.add_argument("--tv-limit", type=float, default=1e-2)
    ap.add_argument("--gate-alpha", action="store_true",
                    help="gate over radius x |inclination| cells (default: radius bins only)")
    a = ap.parse_args(argv)

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    conds = S.drop_conditions(a.records, seed=a.seed)
    fresh = S.drop_conditions(a.eval, seed=a.seed + 1)
    train = a.epochs > 0
    if not train and not a.init:
        ap.error("--epochs 0 (evaluation only) needs --init")
    with icp.Context(0) as ctx, S.Surrogate(ctx, a.theta, a.phi, max_records=64, dataset=a.records if train else 1,
                                            seed=a.seed) as sur:
        if a.init:
            sur.set_params(np.load(a.init, allow_pickle=False))
        if a.tv_weight:
            sur.set_loss(a.tv_weight)
        out = {"records": a.records, "grid": a.grid, "cells": [a.theta, a.phi], "tv_weight": a.tv_weight}
        if train:
            t0 = time.time()
            trace_ms = S.fill_training_set(sur, conds, a.grid, log=log)
            log(f"training set: {a.records} records traced in {trace_ms / 1e3:.2f} s device "
                f"({time.time() - t0:.1f} s wall)")
            # the epoch count is capped by --time-limit; the cosine schedule spans --epochs
            rep = S.train(sur, a.records, a.epochs, a.batch, a.lr, a.lr_min, seed=a.seed,
                          time_limit=a.time_limit, log=log)
            out.update({"epochs": rep.epochs, "steps": rep.steps, "train_seconds": rep.seconds,
                        "loss_first": rep.losses[0], "loss_final": rep.final_loss,
                        "step_ms_last": sur.timing()[2]})
            if a.out:
                np.save(a.out, sur.params())
        gate = None
        if a.gate:
            gate = S.fit_gate(sur, S.drop_conditions(a.gate, seed=a.seed + 2), a.eval_grid, tv_limit=a.tv_limit,
                              alpha_edges=S.ALPHA_EDGES if a.gate_alpha else (0.0, 90.0 + 1e-6))
            log(f"gate: accept {gate.accept} tv_q90 {[[round(x, 4) for x in r] for r in gate.tv_quantile]}")
            if a.gate_out:
                gate.save(a.gate_out)
        tv, dg, eval_ms = S.evaluate(sur, fresh, a.eval_grid)
        out.update({"eval_records": a.eval, "eval_grid": a.eval_grid, "tv_mean": float(tv.mean()),
                    "tv_max": float(tv.max()), "tv_p90": float(np.quantile(tv, 0.9)), "dg_mean": float(dg.mean()),
                    "dg_max": float(dg.max()), "eval_trace_ms": eval_ms})
        if gate is not None:
            inside = gate.accepts(fresh)
            out["gate_accept"] = gate.accept
            out["fresh_inside"] = int(inside.sum())
            if inside.any():
                out["tv_inside_mean"] = float(tv[inside].mean())
                out["tv_inside_max"] = float(tv[inside].max())
        print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
