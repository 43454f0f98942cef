/*
 * CPU oracle for the phase-function path — TEST INFRASTRUCTURE ONLY.
 *
 * This library checks the product and is never part of it. Only tests/, the smoke
 * check in __graft_entry__.py and the cpu_baseline leg of bench.py may load
 * it. The product path (libicephase.so) never links, loads or calls it.
 *
 * Each function restates the reference algorithm it cites (paths are relative
 * to the reference checkout). Two families:
 *   - FP64 restatements of the reference CPU path (hex tracer, phase_bin,
 *     PhaseMassAccumulator, build_host_cdf, u edges). They follow the
 *     reference operation order so that their outputs are bitwise equal to the
 *     compiled reference where libm agrees.
 *   - An FP32 mirror of the drop tracer the HIP kernel runs (the reference has
 *     no CPU drop histogram; its drop density is an OptiX patch solver), plus
 *     an FP64 twin of the same algorithm and a closed-form sphere oracle.
 *
 * Pinning (what checks each function; see tests/test_oracle_*.py):
 *   orc_u_edges        closed form sin^2(pi i/2n), strict increase, golden file
 *   orc_phase_bin      reference KATs (hex_phase/tests/phase_tests.cpp:71-104)
 *                      + golden vectors from the compiled reference
 *   orc_hex_*          analytic slab R+(1-R)^2(1+R), transparent n=1 limit,
 *                      Fresnel closed forms, golden samples from the compiled
 *                      reference (bitwise)
 *   orc_build_cdf      uniform-sphere KAT, sparse HDR KAT, golden arrays from
 *                      the compiled reference (bitwise)
 *   orc_drop_*         sphere closed-form deflection D_p and Fresnel powers
 *                      (orc_sphere_analytic_*), FP32 vs FP64 twin agreement;
 *                      non-spherical drops against the reference's own drop
 *                      code compiled into oracle/_ref: surface and gradient
 *                      (RaindropShape, FP64 to 1e-14), entry/exit points and
 *                      normals (RaindropIntersector), per-family directions,
 *                      statuses and flux powers of R/TT/TRT/TRRT
 *                      (RaindropPathTracer), and central differences of the
 *                      gradient (tests/test_drop_geometry_pins.py)
 *   orc_atlas_rule     closed form: the ray areas add to the projected area;
 *                      ray origins on the lit faces; node frames and ray
 *                      counts equal the compiled reference tool's metadata
 *   orc_atlas_node     index-matched n = 1 delta, energy balance per node,
 *                      mirror symmetry of the phi = 0 node, masses of the
 *                      compiled reference tool (relative L1 < 1e-14)
 *   orc_water_*        not present: the water index is host material code,
 *                      checked directly against the compiled reference and
 *                      golden values (tests/test_host_services.py)
 */
#ifndef ICEPHASE_ORACLE_H
#define ICEPHASE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- grid ---------------------------------------------------------------- */
/* phase_cdf_math.hpp:16-32. Returns 0 on success. */
int orc_u_edges(uint32_t n, double* out /* n+1 */);

typedef struct orc_frame { double ki[3], e0[3], e1[3]; } orc_frame;
/* hex_trace/include/ice_crystal/math.hpp:61-69 (make_frame). */
int orc_make_frame(const double k_in[3], orc_frame* f);
/* The drop contract frame: e0 = normalize(axis - k (k.axis)) with axis=+z unless
 * |k.z|>=0.9 (then +x); e1 = k x e0. include/rainbow/vec3.hpp:56-65 in FP64. */
int orc_drop_frame(const double k_in[3], orc_frame* f);

typedef struct orc_bin {
    uint64_t index;
    double mu;
    int32_t pole;     /* +1 forward, -1 backward, 0 ordinary */
    int32_t valid;
    int32_t snapped;
    int32_t pad;
} orc_bin;
/* Drop histogram binning: the phase_bin rules applied to d/|d| of an
 * unnormalized direction (no unit-length precondition; mu and s are formed from
 * x, y, z and n = |(x,y,z)| exactly as phase_bin forms them). */
orc_bin orc_phase_bin_raw(const double* u_edges, uint32_t nt, uint32_t np, const orc_frame* f, const double dir[3]);
/* hex_phase/include/ice_crystal/phase_math.hpp:37-81. */
orc_bin orc_phase_bin(const double* u_edges, uint32_t nt, uint32_t np,
                      const orc_frame* f, const double dir[3]);

/* ---- hex tracer (FP64) ----------------------------------------------------- */
typedef struct orc_hex_settings {
    double radius_mm, length_mm;     /* circumradius, FULL length */
    double ki[3];                    /* incident direction (normalized inside) */
    double n_ext, n_int;
    double tail;                     /* residual_power_tolerance */
    uint64_t seed, total;            /* RNG seed, total incident samples */
    uint32_t max_hits;               /* max_internal_hits */
    uint32_t pad;
} orc_hex_settings;

typedef struct orc_ray_audit {
    double escaped, unresolved, balance, max_interface;
    uint64_t outputs;
    uint32_t internal_hits, tir_count, status, pad;
} orc_ray_audit;

typedef struct orc_sample {
    double dir[3];
    double power;
    double length_mm;
    uint64_t id;
    uint32_t entry_face, exit_face, reflections, kind;
} orc_sample;

/* Trace rays [first, first+count). samples may be NULL (count only); cap is its
 * capacity. Returns the number of samples written, or -1 on overflow.
 * hex_trace/include/ice_crystal/hex_trace_core.hpp:80-170 (trace_sample). */
int64_t orc_hex_trace(const orc_hex_settings* s, uint64_t first, uint64_t count,
                      orc_sample* samples, uint64_t cap, orc_ray_audit* audits);
double orc_hex_projected_area(double radius_mm, double length_mm, const double k[3]);

/* ---- deterministic atlas (hex_atlas/src/atlas.cpp) -------------------------- */
typedef struct orc_atlas_settings {
    double radius_mm, length_mm, n_int, n_ext, tail;
    uint32_t theta_intervals, phi_intervals, subdivisions, theta_count, phi_count, max_hits;
} orc_atlas_settings;
typedef struct orc_atlas_report {
    double theta, phi, ki[3], e0[3];
    double projected_area_mm2, escaped_mm2, unresolved_mm2, max_balance_error, max_interface_balance_error;
    uint64_t rays, outputs;
} orc_atlas_report;
/* Area rule of node (i, j): rows (x, y, z, face, weight mm^2); out may be NULL
 * to count. Returns the ray count, -1 when cap is too small, < -1 on failure. */
int64_t orc_atlas_rule(const orc_atlas_settings* s, uint32_t i, uint32_t j, double* out, uint64_t cap);
/* Cell masses (mm^2, nt*np) of node (i, j); 0 on success. */
int orc_atlas_node(const orc_atlas_settings* s, uint32_t i, uint32_t j, double* mass, orc_atlas_report* rep);

/* ---- fixed-point histogram -------------------------------------------------- */
typedef struct orc_hist_stats {
    uint64_t q_forward, q_backward;   /* pole quanta */
    uint64_t samples, invalid, snapped;
    uint64_t status[8];
    uint64_t q_escaped, q_unresolved; /* per-ray audit quanta (rint(p*2^Q)) */
    uint64_t rays_hit;                /* rays that produced an entry */
    uint64_t internal_hits;
} orc_hist_stats;

/* Quantum exponent used for a record of n rays: Q = 62 - ceil(log2(n)), capped
 * to [8, 52]. Same rule as the product kernels (documented there). */
uint32_t orc_quantum_bits(uint64_t rays);

/* Trace + bin + fixed-point accumulate the hex record. bins: nt*np u64 (zeroed
 * by the callee). Returns 0 on success. */
int orc_hex_hist(const orc_hex_settings* s, uint32_t nt, uint32_t np, uint32_t qbits,
                 uint64_t first, uint64_t count, uint64_t* bins, orc_hist_stats* st);

/* Orientation node of a composed record: trace normalize(R^T k) in the body
 * frame, bin R d in make_frame(k), quanta rint(power*qfactor*2^qbits); bins and
 * st are accumulated (hex_phase/src/phase_generate.cpp:44-53,
 * hex_phase/src/phase_table.cpp:52-62). rot is row-major body -> ensemble. */
int orc_hex_hist_node(const orc_hex_settings* s, const double rot[9], double qfactor, uint32_t nt, uint32_t np,
                      uint32_t qbits, uint64_t* bins, orc_hist_stats* st);

/* Projected area along the node's body direction normalize(R^T k). */
double orc_hex_node_area(const orc_hex_settings* s, const double rot[9]);

/* ---- Fraunhofer lobe (no reference counterpart: hex_trace/README.md:20-21) -- */
/* Closed-form transform of a counter-clockwise polygon aperture. */
void orc_aperture_transform(const double* vx, const double* vy, uint32_t nv, double q1, double q2, double* re,
                            double* im);
/* Convex aperture of the prism seen along k, in make_frame(k)'s (e0, e1). */
uint32_t orc_hex_aperture(double radius_mm, double length_mm, const double k[3], double* vx, double* vy,
                          double p1[3], double p2[3]);
/* Lobe mass of one cell (mm^2), |F|^2/lambda^2 integrated over the cell. */
double orc_lobe_cell(const double* vx, const double* vy, uint32_t nv, double area, double dmax, const double p1[3],
                     const double p2[3], double lambda_mm, const orc_frame* g, uint32_t nt, uint32_t np, uint32_t i,
                     uint32_t j);

/* Reference PhaseMassAccumulator semantics: per-bin Kahan FP64 of w=scale*power,
 * poles spread by /np into the end rows. hex_phase/src/phase_table.cpp:39-76. */
int orc_hex_masses(const orc_hex_settings* s, uint32_t nt, uint32_t np,
                   double* mass, double* forward, double* backward);

/* Integer bins -> FP64 cell masses in quanta (common factor dropped):
 * m = (double)q + [row 0] (double)qf/np + [row nt-1] (double)qb/np. */
void orc_bins_to_mass(const uint64_t* bins, uint64_t qf, uint64_t qb,
                      uint32_t nt, uint32_t np, double* mass);

/* ---- CDF -------------------------------------------------------------------- */
typedef struct orc_cdf_report {
    double g_source, g_stored, coarsening_tv, cdf_mass;
    double l1_error, lost_mass, max_cell_error, aggregation_error;
    uint64_t lost_cells;
    int32_t status;   /* 0 ok; >0 the audit that failed */
    int32_t pad;
} orc_cdf_report;
/* hex_phase/src/phase_table.cpp:111-187 (build_host_cdf), policy defaults
 * max_l1=1e-10, max_lost=1e-12, max_tv=1. */
int orc_build_cdf(const double* mass, uint32_t nt, uint32_t np,
                  uint32_t out_nt, uint32_t out_np,
                  double max_l1, double max_lost, double max_tv,
                  double* phi /* out_np+1 */, double* theta /* out_np*(out_nt+1) */,
                  orc_cdf_report* r);

/* ---- drop tracer (FP32 mirror of the HIP kernel, FP64 twin) ------------------- */
typedef struct orc_drop_record {
    float coeffs[8];        /* r(mu)=1+c0+c1 mu+sum c_n T_n(mu), mu=-y/r */
    float inner, outer;     /* conservative radii (outer sets the incident grid extent) */
    float rmax;             /* bounding radius of the intersector: max r(chi) rounded up to 2^-20 */
    float ki[3], e0[3], e1[3];
    float n_int, n_ext;
    float half_extent;      /* grid covers [-h,h]^2 in the (e0,e1) plane */
    uint32_t grid_w, grid_h;
    uint32_t max_hits;      /* internal interactions traced */
    uint32_t sphere;        /* 1: closed-form unit sphere */
    float tail;             /* stop when remaining power <= tail */
} orc_drop_record;

/* Fill an orc_drop_record from physical inputs the same way the product's
 * host code does (ice_crystal_phase_amd/csrc/records.cpp). */
int orc_drop_make(const double coeffs[8], int sphere, double alpha_deg, double n_int,
                  uint32_t grid_w, uint32_t grid_h, uint32_t max_hits, double tail,
                  orc_drop_record* out);

/* Per emitted sample callback-free interface: trace rays [first,first+count) of
 * the grid (ray id = iy*grid_w+ix) and write up to cap samples. precision 32 or
 * 64 (64 runs the same algorithm in double). */
int64_t orc_drop_trace(const orc_drop_record* r, int precision, uint64_t first, uint64_t count,
                       orc_sample* samples, uint64_t cap, orc_ray_audit* audits);

int orc_drop_hist(const orc_drop_record* r, int precision, uint32_t nt, uint32_t np,
                  uint32_t qbits, uint64_t first, uint64_t count,
                  uint64_t* bins, orc_hist_stats* st);

/* Geometry of the drop intersector, for pinning it against the reference's
 * RaindropShape / RaindropIntersector (include/rainbow/raindrop_shape.hpp:85-100,
 * include/rainbow/raindrop_intersection.hpp:337-382):
 *   orc_drop_surface: f = r - R(mu) and the outward (unnormalized) gradient at n
 *     points (x,y,z triples);
 *   orc_drop_entry_point: the first surface point of the line e0 u + e1 v + t ki
 *     and the unit outward normal there (0 = miss);
 *   orc_drop_exit_point: the exit of the interior ray from surface point p along
 *     d and the unit outward normal there (0 = unresolved);
 *   orc_drop_trace_origin: the full trace of the ray through e0 u + e1 v (samples
 *     as orc_drop_trace, in emission order: R, then T at internal hit 1, 2, ...). */
int orc_drop_surface(const orc_drop_record* r, int precision, uint64_t n, const double* pts, double* f,
                     double* grad);
int orc_drop_entry_point(const orc_drop_record* r, int precision, double u, double v, double p[3], double nrm[3]);
int orc_drop_exit_point(const orc_drop_record* r, int precision, const double p[3], const double d[3], double q[3],
                        double nrm[3]);
int64_t orc_drop_trace_origin(const orc_drop_record* r, int precision, double u, double v, orc_sample* samples,
                              uint64_t cap, orc_ray_audit* audit);

/* Closed-form sphere: impact parameter b in [0,1), index n. Order p=0 is the
 * external reflection, p>=1 transmits after p-1 internal reflections. Returns
 * psi, the signed rotation of the direction in the (ki, impact vector) plane
 * (out = cos(psi) ki + sin(psi) rho_hat), and the s and p power fractions.
 * Textbook Fresnel; deviation D_p=(p-1)pi+2i-2p r for p>=1 (psi=-D_p),
 * D_0=pi-2i (psi=+D_0) (tests/raindrop_validation.hpp:155-157 in the reference). */
void orc_sphere_order(double b, double n, uint32_t p, double* psi, double* ps, double* pp);

/* Closed-form theta marginal of a sphere record: every cell-centre ray of the
 * grid x grid incident grid over [-h,h]^2 with impact parameter b<1, orders
 * p=0..max_hits, power 0.5(Ps+Pp), binned by u=(1-cos psi)/2 on the u edges. */
void orc_sphere_theta_hist(uint32_t grid, float half_extent, double n, uint32_t max_hits, uint32_t nt,
                           double* hist);

/* FP64 normalize by max-abs scaling (hex_trace/include/ice_crystal/math.hpp:38-47). */
int orc_normalize3(const double in[3], double out[3]);

#ifdef __cplusplus
}
#endif
#endif
