/*
 * icephase — C-ABI of the MI355X phase-function generator (libicephase.so).
 *
 * This library replaces the reference's device-side path for one record or
 * a batch of records: trace, histogram, FP64 CDF, record files. Every entry point
 * below names the reference interface it stands in for (paths relative to the
 * reference checkout).
 *
 * Conventions (the reference's, include/rainbow/cuda_driver.hpp and
 * src/cuda_driver.cpp:148-213):
 *   - one HIP device and one stream per context; the context is not
 *     thread-safe; every call is synchronous (returns after the device work it
 *     started has finished);
 *   - no exceptions or C++ types cross the ABI: functions return IPH_OK (0) or
 *     a negative IPH_E_* code, and iph_last_error() gives the message
 *     (the reference throws std::runtime_error with the same texts);
 *   - the caller owns every host array it passes; the library owns the device
 *     memory of a batch until iph_batch_destroy.
 * The library fails loudly when no HIP device is usable: there is no CPU
 * fallback.
 */
#ifndef ICEPHASE_H
#define ICEPHASE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IPH_ABI_VERSION 3

enum iph_status {
    IPH_OK = 0,
    IPH_E_INVALID = -1,   /* invalid argument / unsupported configuration */
    IPH_E_DEVICE = -2,    /* HIP runtime error */
    IPH_E_AUDIT = -3,     /* a trace or CDF quality audit rejected the record */
    IPH_E_IO = -4,        /* record files could not be written */
    IPH_E_EXISTS = -5,    /* output directory or its .part staging exists */
    IPH_E_NOMEM = -6
};

enum iph_shape {
    IPH_SHAPE_SPHERE = 0, /* unit sphere (equal-volume radius) */
    IPH_SHAPE_DROP = 1,   /* axisymmetric Chebyshev drop r(mu)=1+c0+c1 mu+sum c_n T_n(mu), mu=-y/r */
    IPH_SHAPE_HEX = 2     /* regular hexagonal prism, axis +y */
};

/* One phase-function record. Units: mm, nm, degrees.
 * Drop records (sphere/drop): rays = grid_w*grid_h cell-centre rays over the
 * square [-h,h]^2 of the plane normal to `incident`, h = 1.01*outer radius
 * (include/rainbow/raindrop_settings.hpp:41-42 uses the same 1.01 factor), FP32
 * tracing. Hex records: `rays` samples drawn by the reference's counter RNG
 * (hex_trace/include/ice_crystal/math.hpp:136-146) with `seed`, FP64 tracing. */
typedef struct iph_record {
    uint32_t shape;                /* enum iph_shape */
    uint32_t max_internal_hits;    /* internal encounters traced (each emits T, continues R) */
    uint32_t theta_count, phi_count;  /* histogram grid (cells uniform in theta) */
    uint32_t grid_w, grid_h;       /* drop incident grid; ignored for hex */
    uint64_t rays;                 /* hex samples; drops: 0 or grid_w*grid_h */
    uint64_t seed;                 /* hex RNG seed */
    double radius_mm;              /* drop equal-volume radius / hex circumradius */
    double length_mm;              /* hex full length (c = length/2, aspect c/a = length/(2 radius)) */
    double shape_coefficients[8];  /* drop c0..c7 (must be convex); ignored for sphere/hex */
    double incident[3];            /* incident propagation direction (normalized by the library) */
    double interior_index, exterior_index;
    double wavelength_nm;          /* hex diffraction lobe wavelength; otherwise recorded only */
    double tail_tolerance;         /* stop a ray when its remaining power <= this */
    /* Orientation node (hex only; hex_phase/src/phase_generate.cpp:44-53 and
     * hex_phase/include/ice_crystal/phase_math.hpp:6-20): body -> ensemble
     * rotation, row-major, ko_ensemble = R ko_body. All zero = no rotation:
     * the record is traced and binned in its own frame. With a rotation, the
     * ensemble direction `incident` is normalized, traced in the body frame as
     * normalize(R^T k) and every outgoing direction is binned as R ko in the
     * ensemble frame (make_frame of k), as ice_generate_phase does. */
    double rotation[9];
    /* Node cross-section weight (number probability of the orientation); the
     * sample weight is weight * projected_area / rays (0 = 1). */
    double weight;
    /* 0: the record owns its histogram. k > 0: the record adds its samples to
     * the histogram of record k-1 (an orientation node of that record; same
     * grid, same ensemble incident direction, k-1 < this record's index). Only
     * owners get CDFs. */
    uint32_t histogram_of;
    uint32_t flags;                /* IPH_REC_* */
} iph_record;

enum iph_record_flags {
    /* hex: add the Fraunhofer diffraction lobe of the projected aperture
     * (closed-form polygon transform, total mass = projected area). */
    IPH_REC_DIFFRACTION = 1
};

/* Per-record trace/histogram audit. Power quanta are rint(p * 2^qbits) of a
 * unit-power incident ray; the mass a quantum represents is
 * ray_weight_mm2 * 2^-qbits. */
typedef struct iph_hist_stats {
    uint64_t q_forward, q_backward;  /* pole quanta (spread over the end rows) */
    uint64_t samples, invalid, snapped;
    uint64_t status[9];              /* trace statuses 0..7 (hex_trace_core.hpp:8-13), [8] = missed */
    uint64_t q_escaped, q_unresolved;
    uint64_t internal_hits;
    uint64_t rays;
    uint32_t qbits, pad;
    double ray_weight_mm2;           /* cross-section carried by one ray */
    double projected_area_mm2;       /* hex: exact face sum (weighted over nodes); drops: GK15 */
    uint64_t q_diffraction;          /* quanta of the Fraunhofer lobe (IPH_REC_DIFFRACTION) */
    double diffraction_captured;     /* lobe power inside the computed rows / projected area,
                                        before the lobe is scaled to carry exactly the area */
    uint64_t cell_atomics;           /* u64 atomicAdds the trace kernels issued into histogram
                                        cells (drop records merge equal cells of adjacent lanes,
                                        so this is at most the number of binned samples) */
} iph_hist_stats;

typedef struct iph_cdf_policy {
    uint32_t out_theta, out_phi;     /* stored grid; 0 = histogram grid; must divide it */
    double max_l1;                   /* default 1e-10 (phase_table.hpp CdfPolicy) */
    double max_lost_mass;            /* default 1e-12 */
    double max_coarsening_tv;        /* default 1 */
} iph_cdf_policy;

typedef struct iph_cdf_report {
    double g_source, g_stored, coarsening_tv, cdf_mass;
    double l1_error, lost_mass, max_cell_error, aggregation_error;
    double fine_mass, stored_mass;
    uint64_t lost_cells, malformed, weight_underflow, negative;
    int32_t status;                  /* 0 accepted; >0 which audit failed */
    int32_t pad;
} iph_cdf_report;

typedef struct iph_ctx iph_ctx;
typedef struct iph_batch iph_batch;
typedef struct iph_sur iph_sur;

/* Replaces the ABI sentinel kernels and sizeof fingerprints
 * (device/phase_cdf.cu:65,240; hex_trace/device/hex_trace.cu:6-11;
 * hex_trace/src/hex_trace_cuda.cpp:39-44). Returns IPH_ABI_VERSION. */
int iph_abi_version(void);
/* sizeof of the C-ABI structs for binding checks: 0 iph_record, 1
 * iph_hist_stats, 2 iph_cdf_report, 3 iph_cdf_policy, 4 iph_water_state,
 * 5 iph_area, 6 iph_ice_state, 7 iph_atlas_settings, 8 iph_atlas_node. */
uint32_t iph_struct_size(int which);

/* Replaces rainbow::CudaContext(device) (src/cuda_driver.cpp:148-213).
 * iph_ctx_destroy also destroys every batch of the context that is still
 * alive; do not use those batch handles afterwards. */
int iph_ctx_create(int device, iph_ctx** out);
void iph_ctx_destroy(iph_ctx* ctx);
/* Replaces the exception texts of RAINBOW_CUDA_CHECK (include/rainbow/cuda_error.hpp:11-30).
 * ctx may be NULL for the context-free functions (thread-local message). */
const char* iph_last_error(const iph_ctx* ctx);

/* Replaces rainbow::make_phase_u_edges (include/rainbow/phase_cdf_math.hpp:16-32). */
int iph_u_edges(uint32_t theta_count, double* out /* theta_count+1 */);

/* Bins n FP64 directions on the device with the histogram's own binning
 * (hex_phase/include/ice_crystal/phase_math.hpp:37-81, called per sample by
 * hex_phase/device/phase_accumulate.cu:10-28). frame: ki, e0, e1 (9 doubles).
 * kind[k]: 0 cell (index[k] = i*phi_count+j), +1 forward pole, -1 backward pole,
 * 2 invalid. */
int iph_bin_directions(iph_ctx* ctx, uint32_t theta_count, uint32_t phi_count, const double* frame,
                       uint64_t n, const double* dirs, uint32_t* index, int32_t* kind);

/* Build the device state for nrec records: validates every record (convexity
 * of drops, prism validity, index ranges), allocates the histogram slabs and
 * CDF buffers. Replaces the per-record construction of
 * RaindropTracer (include/rainbow/raindrop_tracer.hpp:27-39),
 * HexPrismTracer (hex_trace/include/ice_crystal/hex_trace_cuda.hpp:20-28) and
 * PhaseMassAccumulatorCuda (hex_phase/include/ice_crystal/phase_cuda.hpp:17).
 * policy may be NULL (defaults). */
int iph_batch_create(iph_ctx* ctx, const iph_record* recs, uint32_t nrec, const iph_cdf_policy* policy,
                     iph_batch** out);
void iph_batch_destroy(iph_batch* b);

enum iph_run_flags {
    IPH_RUN_TRACE = 1,   /* zero + trace + histogram (one persistent launch per record kind) */
    IPH_RUN_CDF = 2,     /* FP64 CDFs of every record from the current histograms */
    IPH_RUN_GRAPH = 4,   /* replay the steps as one captured hipGraph */
    IPH_RUN_NO_AUDIT = 8 /* skip the host-side policy check (timing loops) */
};
/* Runs the batch on the device: replaces HexPrismTracer::trace + download_audits
 * (hex_trace/src/hex_trace_cuda.cpp:47-69), PhaseMassAccumulatorCuda::add/finish
 * (hex_phase/src/phase_cuda.cpp:37-66), RaindropTracer::trace_unpolarized
 * (src/raindrop_tracer.cpp:144-177) and PhaseCdf::build (src/phase_cdf.cpp:275-472).
 * Returns IPH_E_AUDIT if a CDF audit rejects a record (see iph_batch_cdf). */
int iph_batch_run(iph_batch* b, uint32_t flags);
/* Tile size in rays and workgroups of the trace launch (0 = library choice).
 * The autotuner in the package sets these. */
int iph_batch_set_schedule(iph_batch* b, uint64_t tile_rays, uint32_t blocks);
/* enable = 1 (default): the hex kernels add their quanta into a
 * workgroup-private LDS cell table (2048 slots) flushed into the histogram once
 * per tile, and a cell that finds no slot within 16 probes goes to the
 * histogram directly; 0: every quantum goes to the histogram directly (u64 HBM
 * atomics). The histogram is the same bits either way. No reference
 * counterpart (the reference accumulates samples after tracing,
 * hex_phase/device/phase_accumulate.cu:10-28). */
int iph_batch_set_cell_table(iph_batch* b, uint32_t enable);
/* Device time of the last run, from HIP events on the batch stream:
 * ms[0] zero+trace launches, ms[1] CDF kernels, ms[2] whole run. */
int iph_batch_timing(const iph_batch* b, double* ms /* 3 */);

/* Replaces summarize()/TraceSummary (hex_trace/src/hex_trace.cpp:72-88) and
 * PhaseMassAccumulator::statistics (hex_phase/src/phase_table.cpp:77-78). */
int iph_batch_stats(iph_batch* b, uint32_t rec, iph_hist_stats* out);
/* Copies record rec's integer histogram (theta-major, theta_count*phi_count). */
int iph_batch_bins(iph_batch* b, uint32_t rec, uint64_t* out);
/* Copies record rec's CDF arrays and audit: replaces PhaseCdf::write_arrays'
 * sources and audit_statistics/storage_statistics (src/phase_cdf.cpp:482-493). */
int iph_batch_cdf(iph_batch* b, uint32_t rec, double* phi /* onp+1 */, double* theta /* onp*(ont+1) */,
                  iph_cdf_report* report);

/* FP64 CDF of caller masses (theta-major, nt*np, nonnegative): replaces
 * PhaseCdf::build with a Scalar density view (include/rainbow/phase_cdf.hpp:29-31,
 * used by hex_phase/apps/generate_hex_phase.cpp:125-134) and its host twin
 * build_host_cdf (hex_phase/src/phase_table.cpp:111-187). */
int iph_build_cdf(iph_ctx* ctx, const double* mass, uint32_t nt, uint32_t np, const iph_cdf_policy* policy,
                  double* phi, double* theta, iph_cdf_report* report);

/* Measurement only, no counterpart in the reference: the device time (ms) of
 * `atomics` u64 atomicAdd(+1) operations on a zeroed slab of `cells` u64, to
 * cells drawn by a counter hash (mode 0), all to one cell (mode 1), or to
 * consecutive cells in increment order, wrapping (mode 2: the ceiling), launched
 * like the trace kernels (8 workgroups of 256 per CU). The slab is summed
 * afterwards and IPH_E_AUDIT returned unless the sum equals `atomics`. It
 * calibrates the histogram atomics' throughput against what the device
 * sustains for the same footprint. 1 <= cells < 2^32, 1 <= atomics <= 2^40. */
int iph_atomic_probe(iph_ctx* ctx, uint32_t mode, uint64_t cells, uint64_t atomics, double* ms);

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc): the same CDF as iph_build_cdf, bitwise, from caller DEVICE arrays
 * (d_mass nt*np, d_phi onp+1, d_theta onp*(ont+1), all on the context's device)
 * with no allocation. Its scratch (u edges, column sums, reduction partials,
 * report scalars, the coarsened grid) is a fixed 9 MiB device array per GPU,
 * held until the process ends: theta_count <= 4096, and a coarsened stored grid
 * of at most about 1.1 M cells; larger requests are refused with IPH_E_INVALID.
 * Calls on one GPU share that array, so calls that overlap on one GPU (two
 * contexts' streams, two threads) race; each call leaves its scratch contents
 * behind and clears what must start at zero itself. Synchronous. */
int iph_build_cdf_noalloc(iph_ctx* ctx, const double* d_mass, uint32_t nt, uint32_t np, const iph_cdf_policy* policy,
                          double* d_phi, double* d_theta, iph_cdf_report* report);

/* Writes <out_dir>.part/{phi_cdf,theta_given_phi_cdf,u_edges}.npy and
 * metadata.json, then renames to out_dir: replaces DatasetDirectory +
 * NpyFloat64Writer (src/npy_writer.cpp:18-121) and PhaseCdf::write_arrays
 * (src/phase_cdf.cpp:482-493). Refuses an existing out_dir or .part. */
int iph_write_record(const char* out_dir, const char* metadata_json, const double* phi, const double* theta,
                     uint32_t nt, uint32_t np);

/* Device export of record rec's cell probabilities (FP32, theta-major
 * theta_count*phi_count, pole quanta spread over the end rows, normalized to
 * one) into a device buffer of the same context (the surrogate's training
 * set uses it). No reference counterpart; the probabilities are the cell masses
 * PhaseCdf::build stores, normalized. */
int iph_batch_probabilities(iph_batch* b, uint32_t rec, float* d_out);

/* ---- learned surrogate (no reference counterpart: the reference has no
 * learned part, README.md:31) ------------------------------------------------
 * A conditional density MLP over the (u, phi) cells of a theta_count x
 * phi_count grid, conditioned on one record's (inclination deg, wavelength nm,
 * equal-volume radius mm, aspect ratio): 47 inputs (the row coordinate t, 16
 * Fourier features of the azimuth, even about v = 1/4, 13 Fourier pairs of t,
 * 4 conditions), 4 hidden layers of 256 (SiLU), one log density per cell,
 * softmax over the cells. Water-drop records are mirror symmetric in the
 * azimuth (cell j and (phi_count/2 - 1 - j) mod phi_count hold the same
 * mass), and so is the network: phi_count must be even. BF16 MFMA kernels with FP32
 * accumulation for training (forward, backward, Adam) and serving. Training
 * records live in the surrogate's own device memory, filled from traced
 * records; the tracer stays the reference (iph_sur_compare). */
typedef struct iph_sur_config {
    uint32_t theta_count, phi_count;
    uint32_t max_records;      /* records per predict / records / compare call */
    uint32_t dataset;          /* training-set slots (device memory) */
    uint64_t seed;             /* parameter initialization */
} iph_sur_config;
int iph_sur_create(iph_ctx* ctx, const iph_sur_config* cfg, iph_sur** out);
void iph_sur_destroy(iph_sur* s);
/* Packed FP32 parameters: W1 [256][47], b1 [256], (W [256][256], b [256]) x 3,
 * w5 [256], b5 [1] (row-major [out][in]). */
uint32_t iph_sur_param_count(void);
int iph_sur_get_params(iph_sur* s, float* packed);
/* Replaces the parameters and starts a new optimization (Adam moments and step
 * count reset). */
int iph_sur_set_params(iph_sur* s, const float* packed);
/* Training-set slot from a traced record of a batch of the same context, or
 * from host probabilities (theta_count*phi_count, summing to one). cond: alpha
 * deg, wavelength nm, radius mm, aspect. */
int iph_sur_set_record(iph_sur* s, uint32_t slot, iph_batch* b, uint32_t rec, const float cond[4]);
int iph_sur_set_target(iph_sur* s, uint32_t slot, const float* host_probs, const float cond[4]);
/* Training loss of the gradients below: cross entropy + tv_weight * total
 * variation 1/2 sum |p - P| (default 0: cross entropy alone; the records the
 * network learns are mirror symmetric, which the folded total variation
 * assumes). No reference counterpart (the reference has no learned part). */
int iph_sur_set_loss(iph_sur* s, float tv_weight);
/* Cross-entropy loss of one slot and its parameter gradient (packed, host;
 * may be NULL); with a total-variation weight the gradient is that of the
 * weighted sum, the loss stays the cross entropy. */
int iph_sur_gradient(iph_sur* s, uint32_t slot, float* loss, float* packed_grad);
/* One Adam step (beta 0.9/0.999) on the mean gradient of n slots; loss[n]
 * receives their losses (may be NULL). */
int iph_sur_train_step(iph_sur* s, const uint32_t* slots, uint32_t n, float lr, float* loss);
/* Cell probabilities of n <= max_records records (host array n*theta*phi). */
int iph_sur_predict(iph_sur* s, const float* conds, uint32_t n, float* probs);
/* Predictions of n records against traced records recs[] of batch b on the
 * device: tv[k] = 0.5 sum |p - q|, dg[k] = |g(p) - g(q)| with the exact cell
 * mean cosine 1 - (u_i + u_{i+1}). */
int iph_sur_compare(iph_sur* s, const float* conds, uint32_t n, iph_batch* b, const uint32_t* recs, double* tv,
                    double* dg);
/* Records served by the surrogate: probabilities, then the FP64 CDF arrays of
 * each record on the device (the same CDF kernels as traced records), copied to
 * host arrays phi [n][onp+1], theta [n][onp][ont+1]; reports[n] may be NULL. */
int iph_sur_records(iph_sur* s, const float* conds, uint32_t n, const iph_cdf_policy* policy, double* phi,
                    double* theta, iph_cdf_report* reports);
/* Device time of the last calls: ms[0] network part of predict/records, ms[1]
 * whole iph_sur_records (network + CDF), ms[2] last training step. */
int iph_sur_timing(const iph_sur* s, double* ms /* 3 */);

/* ---- host material and geometry services (no device work) ---------------- */

typedef struct iph_water_state {
    double temperature_kelvin, pressure_pascal, density_kg_m3;
    double density_pressure_residual_pascal, density_bracket_width_kg_m3;
    double refractive_index;
} iph_water_state;
/* Replaces rainbow::evaluate_water_optics (src/water_refractive_index.cpp:157-186):
 * IAPWS-95 liquid density at (celsius, pascal), then the IAPWS R9-97 index at
 * the vacuum wavelength. Rejects states outside 0..60 C, 50 kPa..100 MPa. */
int iph_water_optics(double wavelength_nm, double celsius, double pascal, iph_water_state* out);
/* Replaces rainbow::water_refractive_index_r9_97 (src/water_refractive_index.cpp:139-155). */
int iph_water_index(double wavelength_nm, double kelvin, double density_kg_m3, double* n);

/* Replaces RaindropShape::try_make (include/rainbow/raindrop_shape.hpp:22-69):
 * Table I coefficients (single precision) for an equal-volume radius in (0,3] mm. */
int iph_drop_shape(float radius_mm, int sphere, float coefficients[8]);

typedef struct iph_area {
    double area_mm2, abs_error_mm2, rel_error;
    uint32_t evaluations, convexity_cells;
    int32_t analytic_sphere, pad;
} iph_area;
/* Replaces rainbow::compute_projected_area (src/projected_area.cpp:163-228):
 * certified strict convexity, then the projected area of the drop of
 * equal-volume radius radius_mm along incident (particle frame). */
int iph_drop_projected_area(const double coefficients[8], double radius_mm, const double incident[3],
                            iph_area* out);
/* Replaces HexPrism::projected_area_mm2 (hex_trace/include/ice_crystal/hex_prism.hpp:45-51). */
int iph_hex_projected_area(double radius_mm, double length_mm, const double incident[3], double* area_mm2);

typedef struct iph_ice_state {
    double wavelength_nm, n, kappa, lower_nm, upper_nm, log_wavelength_fraction;
} iph_ice_state;
/* Replaces iceCrystal::optics::evaluate_ice_wb2008 (hex_atlas/src/ice_optics.cpp:14-30):
 * Warren & Brandt (2008) ice Ih index, n linear and log(kappa) linear in
 * log(wavelength), knots exact, 200..1100 nm (IPH_E_INVALID outside). */
int iph_ice_index(double wavelength_nm, iph_ice_state* out);

/* ---- deterministic single-particle atlas --------------------------------- */

/* Replaces iceCrystal::atlas::Settings (hex_atlas/include/ice_crystal/atlas.hpp:22-30)
 * without the per-ray column work limits (batch_rays, max_work_bytes), which
 * this implementation does not need. Incident nodes (i, j), i = 0..T, j = 0..P,
 * cover theta = (pi/2) i/T, phi = (pi/6) j/P: the D6h fundamental domain. */
typedef struct iph_atlas_settings {
    double radius_mm, length_mm;  /* circumradius, full length */
    double interior_index, exterior_index, wavelength_nm;
    uint32_t incident_theta_intervals, incident_phi_intervals;
    uint32_t theta_count, phi_count, area_subdivisions, max_internal_hits;
    double tail_tolerance, max_unresolved_fraction, max_balance_error;
} iph_atlas_settings;

/* One incident node: its direction and frame (k = (sin t cos p, cos t, sin t
 * sin p), e0 = (cos t cos p, -sin t, cos t sin p) normalized, e1 = k x e0) and
 * the audits apps/precompute_atlas.cpp:148-153 reports per node. */
typedef struct iph_atlas_node {
    uint32_t i, j;
    double theta, phi, ki[3], e0[3];
    double projected_area_mm2, escaped_mm2, unresolved_mm2, histogram_relative_error;
    uint64_t rays, outputs;
    double max_balance_error, max_interface_balance_error;
} iph_atlas_node;

/* Replaces atlas::make_area_rule + make_node (hex_atlas/src/atlas.cpp:11-88),
 * host only: the deterministic area rays of node (i, j), GL 2 x 3 points in
 * each of area_subdivisions^2 equal-area cells of every lit face (6 triangular
 * sectors on a basal face). rays: capacity x 5 doubles (x, y, z in
 * circumradius units, face, projected weight mm^2), may be NULL to query the
 * count. node: its direction, frame and projected area (audits zero). */
int iph_atlas_node_rays(const iph_atlas_settings* s, uint32_t i, uint32_t j, double* rays, uint64_t capacity,
                        uint64_t* count, iph_atlas_node* node);

/* Replaces atlas::CudaTracer::trace / trace_cpu over a range of nodes
 * (hex_atlas/src/atlas_cuda.cpp:28-57, src/atlas.cpp:139-154; the node loop of
 * apps/precompute_atlas.cpp:139-156): nodes n = first .. first+count-1 in
 * (i, j) order, n = i (P+1) + j, traced together in one persistent launch.
 * Every area ray follows the deterministic splitting chain and bins its
 * weighted emissions into the node's u64 fixed-point histogram (order-free);
 * pole emissions are spread uniformly over the end-cap row. mass_mm2: count x
 * theta_count x phi_count FP64 cell masses. Returns IPH_E_AUDIT with the
 * reference's messages when a ray ends in a rejected status or a node fails its
 * area / escaped / residual conservation audit. */
int iph_atlas_trace(iph_ctx* ctx, const iph_atlas_settings* s, uint32_t first, uint32_t count, double* mass_mm2,
                    iph_atlas_node* nodes);

/* Device-resident copy of an atlas mass array [ni][nj][nt][np] (mm^2; ni, nj
 * incident nodes, phi_count <= 8192), for composing many conditions from one
 * upload. Destroyed with its context if not before. */
typedef struct iph_atlas_mass iph_atlas_mass;
int iph_atlas_mass_create(iph_ctx* ctx, const double* mass, uint32_t ni, uint32_t nj, uint32_t nt, uint32_t np,
                          iph_atlas_mass** out);
void iph_atlas_mass_destroy(iph_atlas_mass* m);
/* Replaces compose_mass (hex_atlas/python/atlas_math.py:394-411): out [nt][np]
 * = sum over nodes (i, j) ascending, parity 0 then 1, shift s ascending of
 * weights[i][j][parity][s] (> 0 only) times the node's cell block shifted by s
 * azimuth cells (parity 1: azimuth-reflected first, cell q -> np-1-q), in that
 * order with the product rounded before each addition: bitwise the reference's
 * sum. weights: ni x nj x 2 x np, finite and nonnegative. */
int iph_atlas_compose(iph_atlas_mass* m, const double* weights, double* out);

/* NPY v1.0 '<f8' C-order file of the given shape (src/npy_writer.cpp:18-95). */
int iph_write_npy(const char* path, const uint64_t* shape, uint32_t ndim, const double* data);

/* ---- Coherent drop stages (rainbow_generate --stage path|focal|diffraction) ----
 * One drop record on the patch model of the reference's default pipeline
 * (apps/generate_phase.cpp:426-440): a grid x grid vertex trace through R, TT,
 * TRT and TRRT with two unit-input Jones response columns and the optical path
 * in whole turns plus a fraction, bilinear chord patches in direction space
 * (folded branches where the map changes orientation), the patches that
 * contain each query direction, and per direction the incoherent, path-
 * coherent and focal (quarter-turn) intensities, then the Table II local
 * Gaussian across rainbow transitions. Destroyed with its context if not before. */
typedef struct iph_coherent iph_coherent;
enum { IPH_STAGE_INCOHERENT = 1, IPH_STAGE_PATH = 2, IPH_STAGE_FOCAL = 3, IPH_STAGE_DIFFRACTION = 4 };

/* RaindropSettings + WaveOptionsSettings of the reference
 * (include/rainbow/raindrop_settings.hpp:10-44, include/rainbow/wave_optics_data.hpp:93-99):
 * incident direction (cos a, -sin a, 0) for inclination a; grid vertices per
 * axis; query grid theta x phi cell centres. primary_sigma_deg 0 selects Table II
 * (radius 0.1..1 mm); transition_contrast is 0.25 in the reference. */
typedef struct iph_coherent_config {
    float radius_mm, wavelength_nm, interior_index, reserved;
    double inclination_deg;
    uint32_t grid, sphere, theta_count, phi_count;
    uint32_t focal_offsets[4];
    double primary_sigma_deg;
    double transition_contrast;
} iph_coherent_config;

/* Counters and stage times (ms, HIP events) of the last evaluation. */
typedef struct iph_coherent_stats {
    uint32_t status_counts[8]; /* patch cells by PatchCellStatus (include/rainbow/outgoing_patch.hpp:15-25) */
    uint32_t patches, folded, bad_vertices, first_bad_vertex;
    uint64_t candidates, hits;
    double ms_trace, ms_patches, ms_folded, ms_query, ms_optics, ms_diffraction;
} iph_coherent_stats;

/* One query direction (theta-major): the fields of the reference's
 * PatchOpticalResult, FocalOpticalResult and DiffractionResult that the stages
 * and their audits read (include/rainbow/patch_optics_data.hpp:46-82,
 * include/rainbow/wave_optics_data.hpp:31-83). Fields: s re, s im, p re, p im
 * of response column 0, then column 1. */
typedef struct iph_coherent_direction {
    double incoherent_s, incoherent_p, path_s, path_p, focal_s, focal_p, diffraction_s, diffraction_p;
    double family_incoherent[4];
    double sigma_rad, blend;
    double path_field[8];
    double focal_field[8];
    uint32_t hit_count, evaluated_hits, rejected_hits, optical_flags;
    uint32_t query_flags, focal_flags, diffraction_flags, transition_flags;
    uint32_t family_hits[4];
    uint32_t transition_kind, folded_hits, regular_hits, refinement_hits;
} iph_coherent_direction;

/* The CDF input gate of a stage (device/phase_cdf.cu:93-111): directions with a
 * non-finite or negative density, incomplete directions (rejected or pending
 * hits, invalid focal or diffraction results), underresolved Table II filters,
 * the first such direction, and the largest density. */
typedef struct iph_coherent_gate {
    uint32_t invalid_values, incomplete, underresolved, first_problem;
    double maximum;
} iph_coherent_gate;

int iph_coherent_create(iph_ctx* ctx, const iph_coherent_config* config, iph_coherent** out);
void iph_coherent_destroy(iph_coherent* h);
/* Replaces RaindropTracer::trace_unpolarized (src/raindrop_tracer.cpp:144-177). */
int iph_coherent_trace(iph_coherent* h);
/* Vertex arrays in the reference's layouts: 4 * grid^2 OutgoingVertex (68 bytes,
 * path-major, include/rainbow/raindrop_trace_data.hpp:41-53) and as many
 * second-column Field32 (16 bytes) - the input PatchAccel::build takes
 * (src/patch_accel.cpp:62-75). set replaces the trace; get reads them back. */
int iph_coherent_set_vertices(iph_coherent* h, const void* vertices, const void* second);
int iph_coherent_get_vertices(iph_coherent* h, void* vertices, void* second);
/* Replaces PatchAccel::build, PatchOptics::prepare_folded, PatchQuery::query_grid
 * and PatchOptics::evaluate_wave_device (src/patch_accel.cpp:62-140,
 * src/patch_optics.cpp:264-291, src/patch_query.cpp:137-212,
 * src/patch_optics.cpp:54-234). IPH_E_AUDIT for failed vertices or invalid
 * patch sources, as the reference refuses them. */
int iph_coherent_evaluate(iph_coherent* h);
int iph_coherent_stats_get(iph_coherent* h, iph_coherent_stats* out);
/* Compact patches (32-byte OutgoingPatch, patch_id order) and folded records
 * (320-byte FoldedPatchRecord, compact order); NULL arrays return counts only. */
int iph_coherent_get_patches(iph_coherent* h, void* patches, uint32_t cap, uint32_t* count, void* folded,
                             uint32_t folded_cap, uint32_t* folded_count);
/* Ordered hit lists (32-byte PatchQueryHit) of directions [first, first+count). */
int iph_coherent_get_hits(iph_coherent* h, uint32_t first, uint32_t count, void* hits, uint64_t cap,
                          uint32_t* counts, uint32_t* flags);
int iph_coherent_get_directions(iph_coherent* h, iph_coherent_direction* out);
/* Density of a stage per direction (out: theta_count * phi_count, may be NULL) and its gate. */
int iph_coherent_density(iph_coherent* h, uint32_t stage, double* out, iph_coherent_gate* gate);
/* Replaces PhaseCdf::build for the drop stages (src/phase_cdf.cpp:275-472):
 * the gate (IPH_E_AUDIT on invalid or incomplete directions, a zero maximum, or
 * underresolved directions unless allowed), weights (d / max) du, and the CDF
 * of iph_build_cdf. phi: phi_out + 1, theta: phi_out x (theta_out + 1). */
int iph_coherent_cdf(iph_coherent* h, uint32_t stage, uint32_t allow_underresolved, const iph_cdf_policy* policy,
                     double* phi, double* theta, iph_cdf_report* report, iph_coherent_gate* gate);

#ifdef __cplusplus
}
#endif
#endif
