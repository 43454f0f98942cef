"""Derivation of the drop kernel's operation census (test infrastructure only).

Runs the counting build of the oracle (oracle/census.cpp, the FP32 mirror of
the kernel and the FP64 phase_bin rule with counting number types) over rows of
rays spread evenly across every record of a workload, and prints the mean
operations per event: miss, entry (set-up, entry solve, first split, the
external reflection's power and the ray's last partial encounter), hit (one
internal encounter with its transmitted branch's power) and emit (the binning
of one emitted direction). bench.py's CENSUS holds the values this prints for
the two drop workloads:

    python oracle/census.py                   # oblate drops (192 records) and the sphere record
    python oracle/census.py --rows 8 --every 24   # a quicker subsample

The workload definitions are bench.py's (workload(), the sphere leg's record).
"""
from __future__ import annotations

import argparse
import json
import math
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))


def census_of(records, nt: int, np_: int, rows: int = 8, every: int = 1) -> dict[str, float]:
    """Mean operations per event over `rows` full grid rows per record (rows at
    the centres of `rows` equal bands of the grid), for every `every`-th record."""
    import oracle_bind as ob
    import ice_crystal_phase_amd as ip
    tot = np.zeros(8)
    for rec in records[::every]:
        alpha = math.degrees(math.atan2(-rec.incident[1], rec.incident[0]))
        sphere = rec.shape == ip.SHAPE_SPHERE
        r = ob.drop_record(coeffs=tuple(rec.shape_coefficients), sphere=sphere, alpha_deg=alpha,
                           n_int=rec.interior_index, grid=rec.grid_w, max_hits=rec.max_internal_hits)
        for j in range(rows):
            iy = int((2 * j + 1) * rec.grid_h // (2 * rows))
            c = ob.drop_census(r, nt, np_, iy * rec.grid_w, rec.grid_w)
            tot += np.array([c.misses, c.entries, c.hits, c.emissions, c.ops_miss, c.ops_entry, c.ops_hit,
                             c.ops_emit], dtype=np.float64)
    counts = dict(zip(("misses", "entries", "hits", "emissions"), tot[:4].astype(int).tolist()))
    means = {"miss": tot[4] / max(tot[0], 1), "entry": tot[5] / tot[1], "hit": tot[6] / tot[2],
             "emit": tot[7] / tot[3]}
    return {**{k: round(v, 1) for k, v in means.items()}, "events": counts}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--rows", type=int, default=8, help="grid rows per record")
    ap.add_argument("--every", type=int, default=1, help="use every k-th record")
    args = ap.parse_args(argv)
    import bench
    import ice_crystal_phase_amd as ip
    recs, _, _ = bench.workload(192, 26, 900, 1800)
    n = ip.water_optics(550.0).refractive_index
    sphere = ip.sphere_record(n, 2048, 900, 1800, radius_mm=1.0, inclination_deg=0.0, max_internal_hits=9)
    out = {"drop": census_of(recs, 900, 1800, args.rows, args.every),
           "sphere": census_of([sphere], 900, 1800, max(args.rows, 16), 1)}
    print(json.dumps(out, indent=2))
    return 0


if __name__ == "__main__":
    sys.exit(main())
