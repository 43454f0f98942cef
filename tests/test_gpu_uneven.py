"""GPU parity of the uneven batch: one mixed launch of spheres, Table I drops
and hexagonal prisms at 900 x 1800 cells (the uneven workload's draws with
reduced ray counts, plus forced extremes: hex aspect 0.2 and 4, radius 0.1 and
3 mm, inclination +-89 deg) against the CPU oracle record by record: integer
histograms and audit counters bitwise (the snap diagnostic to 1 % of the
samples, see below), CDF arrays bitwise for the same masses, and conservation
of every record's quanta (escaped + unresolved = hitting rays x 2^Q to the
power balance of the FP32 drop and FP64 hex chains)."""
from __future__ import annotations

import math

import numpy as np
import pytest

import ice_crystal_phase_amd as ip
from ice_crystal_phase_amd import autotune

pytestmark = pytest.mark.gpu

NT, NP = 900, 1800


def _specs():
    specs = autotune.uneven_specs(24, seed=4, log2_rays=(12.0, 15.0))
    S = autotune.UnevenSpec
    specs += [
        S(2, 0.1, 450.0, 89.0, 1 << 13, 0.2, 1001),
        S(2, 3.0, 650.0, -89.0, 1 << 13, 4.0, 1002),
        S(2, 3.0, 550.0, 0.0, 1 << 13, 0.2, 1003),
        S(2, 0.1, 400.0, -45.0, 1 << 13, 4.0, 1004),
        S(1, 0.1, 550.0, 89.0, 1 << 13, 1.0, 1005),    # Table I sphere row: all terms zero
        S(1, 3.0, 700.0, -89.0, 1 << 13, 1.0, 1006),   # the flattest Table I drop
        S(1, 3.0, 400.0, 89.0, 1 << 13, 1.0, 1007),
        S(0, 3.0, 500.0, -89.0, 1 << 13, 1.0, 1008),
    ]
    return specs


def _oracle_hist(oracle, s, rec):
    if s.kind == 2:
        st = oracle.hex_settings(radius_mm=s.radius_mm, length_mm=2 * s.radius_mm * s.aspect,
                                 ki=tuple(rec.incident), n_int=1.31, tail=1e-12, seed=s.seed, total=s.rays,
                                 max_hits=12)
        return oracle.hex_hist(st, NT, NP)
    coeffs = [0.0] * 8 if s.kind == 0 else [float(c) for c in rec.shape_coefficients]
    r = oracle.drop_record(coeffs=coeffs, sphere=s.kind == 0, alpha_deg=s.alpha_deg, n_int=rec.interior_index,
                           grid=rec.grid_w, max_hits=9)
    return oracle.drop_hist(r, NT, NP, precision=32)


def test_uneven_mixed_batch_equals_oracle(ctx, oracle):
    specs = _specs()
    recs = [autotune.spec_record(s, NT, NP) for s in specs]
    kinds = [s.kind for s in specs]
    print(f"case: uneven mixed batch {len(recs)} records ({kinds.count(0)} spheres, {kinds.count(1)} drops, "
          f"{kinds.count(2)} hex), {NT}x{NP}", flush=True)
    with ip.Batch(ctx, recs) as b:
        b.run(ip.RUN_TRACE | ip.RUN_CDF | ip.RUN_GRAPH)
        for i, (s, rec) in enumerate(zip(specs, recs)):
            want, ost = _oracle_hist(oracle, s, rec)
            got, st = b.bins(i), b.stats(i)
            assert np.array_equal(got, want), (i, s, int(np.count_nonzero(got != want)))
            assert (st.q_forward, st.q_backward, st.samples, st.invalid, tuple(st.status[:8]),
                    st.q_escaped, st.q_unresolved, st.internal_hits) == \
                   (ost.q_forward, ost.q_backward, ost.samples, ost.invalid, tuple(ost.status[:8]),
                    ost.q_escaped, ost.q_unresolved, ost.internal_hits), (i, s)
            # The snap counter is a diagnostic of phase_bin: a sample counts when its
            # FP64 column or row index lies within the snap tolerance of an edge but
            # not exactly on it. Prisms lit in their mirror plane (x-y here) emit many
            # directions a hair off the plane, whose atan2 lands within one ulp of the
            # phi = +-pi/2 column edges; device atan2 (not correctly rounded) and the
            # host's differ in that last ulp, so exactly-on-edge and near-edge trade
            # places. The snap moves both to the same column, so the bins above are
            # bitwise equal; the counter agrees to 1 % of the samples.
            assert abs(int(st.snapped) - int(ost.snapped)) <= max(16, st.samples // 100), (i, s, st.snapped,
                                                                                           ost.snapped)
            # conservation: every hitting ray carries 2^Q quanta, split into escaped
            # and unresolved quanta (FP32 drop chains conserve power to ~1e-7, the
            # FP64 hex chain to its per-interface balance audit, 2048 eps)
            hit = ip.record_rays(rec) - st.status[8]
            total = float(hit) * 2.0 ** st.qbits
            tol = 1e-12 if s.kind == 2 else 1e-5
            rounding = 0.5 * float(st.samples + hit)   # half a quantum per emission and per remainder
            assert abs(float(st.q_escaped + st.q_unresolved) - total) <= tol * total + rounding, (i, s)
            # the device CDF of this record equals the oracle CDF of the same masses
            phi, theta, rep = b.cdf(i)
            rc, ophi, otheta, _ = oracle.build_cdf(oracle.bins_to_mass(got, st.q_forward, st.q_backward, NT, NP),
                                                   NT, NP)
            assert rc == 0 and rep.status == 0
            assert np.array_equal(phi, ophi) and np.array_equal(theta, otheta), (i, s)
