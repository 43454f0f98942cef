/* Runs the oracle's C restatement (oracle/oracle_hex.c, oracle_drop.c) under
 * AddressSanitizer and UndefinedBehaviorSanitizer on small cases of every
 * entry point the tests use: u edges, frames, phase_bin over random
 * directions, the hex trace, histogram and masses, the drop trace and
 * histogram in FP32 and FP64, drop geometry, the atlas rule and node, the
 * aperture lobe, the CDF and the sphere closed form. Exit status 0 and no
 * sanitizer report is the pass. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../oracle/oracle.h"

static uint64_t rng = 0x9e3779b97f4a7c15ull;
static double uniform(void)
{
    rng ^= rng << 13;
    rng ^= rng >> 7;
    rng ^= rng << 17;
    return (double)(rng >> 11) / 9007199254740992.0;
}

#define CHECK(x)                                                   \
    do {                                                           \
        if (!(x)) {                                                \
            fprintf(stderr, "check failed: %s (line %d)\n", #x, __LINE__); \
            return 1;                                              \
        }                                                          \
    } while (0)

int main(void)
{
    const uint32_t nt = 36, np = 72;
    double* edges = malloc((nt + 1) * sizeof(double));
    CHECK(orc_u_edges(nt, edges) == 0);
    orc_frame f;
    const double k[3] = {1.0, -0.4, 0.3};
    CHECK(orc_make_frame(k, &f) == 0);
    for (int i = 0; i < 20000; ++i) {
        double d[3] = {2 * uniform() - 1, 2 * uniform() - 1, 2 * uniform() - 1};
        if (i % 97 == 0) d[0] = d[1] = 0, d[2] = i % 2 ? 1 : -1; /* poles */
        const double n = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        if (!(n > 0)) continue;
        for (int a = 0; a < 3; ++a) d[a] /= n;
        const orc_bin b = orc_phase_bin_raw(edges, nt, np, &f, d);
        (void)b;
    }
    /* hex: trace, histogram, masses, aperture and one lobe cell */
    orc_hex_settings hs;
    memset(&hs, 0, sizeof hs);
    hs.radius_mm = 0.05;
    hs.length_mm = 0.2;
    hs.ki[0] = 1.0; hs.ki[1] = -0.4; hs.ki[2] = 0.3;
    hs.n_int = 1.31;
    hs.n_ext = 1.0;
    hs.total = 4096;
    hs.seed = 12345;
    hs.max_hits = 12;
    hs.tail = 1e-12;
    orc_sample* samples = malloc(4096 * 16 * sizeof(orc_sample));
    orc_ray_audit* audits = malloc(4096 * sizeof(orc_ray_audit));
    CHECK(orc_hex_trace(&hs, 0, 4096, samples, 4096 * 16, audits) > 0);
    uint64_t* bins = calloc((size_t)nt * np, sizeof(uint64_t));
    orc_hist_stats st;
    CHECK(orc_hex_hist(&hs, nt, np, orc_quantum_bits(4096), 0, 4096, bins, &st) == 0);
    double* mass = calloc((size_t)nt * np, sizeof(double));
    double fwd = 0, bwd = 0;
    CHECK(orc_hex_masses(&hs, nt, np, mass, &fwd, &bwd) == 0);
    double vx[16], vy[16], p1[3], p2[3];
    const uint32_t nv = orc_hex_aperture(0.05, 0.2, k, vx, vy, p1, p2);
    CHECK(nv >= 4 && nv <= 16);
    /* CDF of the hex masses */
    double* phi = malloc((np + 1) * sizeof(double));
    double* theta = malloc((size_t)np * (nt + 1) * sizeof(double));
    orc_cdf_report rep;
    CHECK(orc_build_cdf(mass, nt, np, nt, np, 1e-10, 1e-12, 1.0, phi, theta, &rep) == 0);
    /* drops: a Table I-like oblate drop and a sphere, FP32 and FP64 */
    const double coeffs[8] = {-0.044708, -0.032825, -0.124905, -0.021774, 0, 0, 0, 0};
    for (int sphere = 0; sphere < 2; ++sphere) {
        orc_drop_record r;
        CHECK(orc_drop_make(coeffs, sphere, 37.0, 1.333, 32, 32, 9, 0.0, &r) == 0);
        for (int prec = 32; prec <= 64; prec += 32) {
            memset(bins, 0, (size_t)nt * np * sizeof(uint64_t));
            CHECK(orc_drop_hist(&r, prec, nt, np, orc_quantum_bits(1024), 0, 1024, bins, &st) == 0);
            CHECK(orc_drop_trace(&r, prec, 0, 1024, samples, 4096 * 16, audits) >= 0);
            double p[3], nrm[3];
            (void)orc_drop_entry_point(&r, prec, 0.1, -0.2, p, nrm);
            const double pts[6] = {0.5, 0.1, 0.2, 1.0, 0.0, 0.0};
            double fv[2], g[6];
            CHECK(orc_drop_surface(&r, prec, 2, pts, fv, g) == 0);
        }
    }
    /* atlas: the area rule and one node on a small grid */
    orc_atlas_settings as;
    memset(&as, 0, sizeof as);
    as.radius_mm = 0.1;
    as.length_mm = 0.2;
    as.n_int = 1.31;
    as.theta_intervals = 4;
    as.phi_intervals = 3;
    as.theta_count = 18;
    as.phi_count = 36;
    as.subdivisions = 2;
    as.n_ext = 1.0;
    as.max_hits = 64;
    as.tail = 1e-12;
    double* rule = malloc(1 << 20);
    CHECK(orc_atlas_rule(&as, 1, 1, rule, (1 << 20) / sizeof(double)) > 0);
    double* node = calloc(18 * 36, sizeof(double));
    orc_atlas_report ar;
    CHECK(orc_atlas_node(&as, 1, 1, node, &ar) == 0);
    free(edges); free(samples); free(audits); free(bins); free(mass); free(phi); free(theta); free(rule); free(node);
    puts("oracle driver: ok");
    return 0;
}
