// Deterministic single-particle atlas of a hexagonal prism (C-ABI in
// include/icephase.h): for every incident node of the D6h fundamental domain,
// a fixed area quadrature over the lit faces replaces random entry points, and
// each area ray runs the hex tracer's deterministic splitting chain
// (hex::trace_from). Every node is one hex record of the persistent trace
// launch whose rays read their entries from a table, so all nodes of a call
// share one tile queue; emissions land in the node's u64 fixed-point histogram
// (integer sums: the result is independent of scheduling), scaled to mm^2 on
// readout.
//
// Reference: hex_atlas/src/atlas.cpp (area rule :11-51, node :69-88, audits
// :89-138), hex_atlas/src/atlas_cuda.cpp:28-57, apps/precompute_atlas.cpp.
#include "host_records.hpp"
#include "runtime_internal.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

namespace icephase {
hipError_t launch_zero(const DevRecord* d_recs, uint32_t nrec, uint32_t* d_queue, hipStream_t s);
hipError_t launch_trace_area(bool wide, const DevRecord* d_recs, TileMap map, uint32_t* d_queue, uint32_t blocks,
                             uint32_t threads, hipStream_t s);
}  // namespace icephase

using namespace icephase;

namespace {

// Cell masses of every node from its fixed-point bins (quanta x the node's mm^2
// per quantum; the pole quanta spread evenly over the end-cap rows, forward
// then backward) and the exact integer total of the node's binned quanta.
__global__ void __launch_bounds__(256)
    icephase_atlas_mass(const unsigned long long* __restrict__ bins, const unsigned long long* __restrict__ stats,
                        const double* __restrict__ scale, uint32_t count, uint32_t nt, uint32_t np,
                        double* __restrict__ mass, unsigned long long* __restrict__ totals)
{
    __shared__ unsigned long long s_part[4];
    const uint64_t cells = (uint64_t)nt * np;
    for (uint32_t q = blockIdx.y; q < count; q += gridDim.y) {
        const unsigned long long* B = bins + (uint64_t)q * cells;
        const unsigned long long* S = stats + (uint64_t)q * kStatCount;
        const double sc = scale[q];
        const double fwd = (double)S[kStatForward] * sc / np, bwd = (double)S[kStatBackward] * sc / np;
        double* M = mass + (uint64_t)q * cells;
        unsigned long long part = 0;
        for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < cells;
             c += (uint64_t)gridDim.x * blockDim.x) {
            const unsigned long long b = B[c];
            part += b;
            double m = (double)b * sc;
            if (c < np) m += fwd;
            if (c >= cells - np) m += bwd;
            M[c] = m;
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off, 64);
        if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = part;
        __syncthreads();
        if (threadIdx.x == 0) {
            const unsigned long long t = s_part[0] + s_part[1] + s_part[2] + s_part[3];
            if (t) atomicAdd(&totals[q], t);
        }
        __syncthreads();
    }
}

constexpr double kPi = 3.141592653589793238462643383279502884;
constexpr uint64_t kAtlasTileRays = 256;  // one workgroup's lanes per tile

struct HostRay {
    d3 pos;
    double weight;  // projected area, mm^2
    uint32_t face;
};

// Compensated sum, the recurrence of the reference's Sum.
struct Kahan {
    double v = 0, c = 0;
    void add(double x)
    {
        const double y = x - c, t = v + y;
        c = (t - v) - y;
        v = t;
    }
};

bool check_settings(const iph_atlas_settings* s, std::string& err)
{
    if (!s) {
        err = "Invalid atlas settings.";
        return false;
    }
    hex::Prism p{s->radius_mm, s->length_mm};
    const bool ok = p.valid() && s->interior_index > 0 && s->exterior_index > 0 && s->wavelength_nm > 0 &&
                    std::isfinite(s->interior_index) && std::isfinite(s->exterior_index) &&
                    std::isfinite(s->wavelength_nm) && s->incident_theta_intervals && s->incident_phi_intervals &&
                    s->incident_theta_intervals <= 4096 && s->incident_phi_intervals <= 4096 && s->theta_count &&
                    s->phi_count && (uint64_t)s->theta_count * s->phi_count <= 100000000ull && s->area_subdivisions &&
                    (uint64_t)s->area_subdivisions * s->area_subdivisions <= 1000000ull && s->max_internal_hits;
    if (!ok) {
        err = "Invalid atlas settings.";
        return false;
    }
    for (double v : {s->tail_tolerance, s->max_unresolved_fraction, s->max_balance_error})
        if (!std::isfinite(v) || v < 0 || v >= 1) {
            err = "Invalid audit tolerance.";
            return false;
        }
    if (s->tail_tolerance > s->max_unresolved_fraction) {
        err = "Ray tail exceeds global unresolved budget.";
        return false;
    }
    return true;
}

// Incident direction and frame of node (i, j): the endpoints of both axes
// use exact sines and cosines; k is normalized twice, as make_node and then
// make_trace_settings do (hex_atlas/src/atlas.cpp:73-84, hex_trace.cpp:20).
bool node_geometry(const iph_atlas_settings& s, uint32_t i, uint32_t j, iph_atlas_node& n, hex::Basis& f, d3& k,
                   std::string& err)
{
    const uint32_t T = s.incident_theta_intervals, P = s.incident_phi_intervals;
    if (i > T || j > P) {
        err = "Incident node index.";
        return false;
    }
    std::memset(&n, 0, sizeof n);
    n.i = i;
    n.j = j;
    n.theta = (kPi / 2) * i / T;
    n.phi = (kPi / 6) * j / P;
    const double st = i == 0 ? 0 : (i == T ? 1 : std::sin(n.theta));
    const double ct = i == 0 ? 1 : (i == T ? 0 : std::cos(n.theta));
    const double sp = j == 0 ? 0 : (j == P ? 0.5 : std::sin(n.phi));
    const double cp = j == 0 ? 1 : (j == P ? kSqrt3 / 2 : std::cos(n.phi));
    d3 k1, e;
    if (!unit_of(D3(st * cp, ct, st * sp), k1) || !unit_of(k1, k) || !unit_of(D3(ct * cp, -st, ct * sp), e)) {
        err = "Invalid incident spherical frame.";
        return false;
    }
    f.e0 = e;
    f.e1 = cross(k, e);
    if (!hex::basis_ok(f, k)) {
        err = "Invalid incident spherical frame.";
        return false;
    }
    n.ki[0] = k.x, n.ki[1] = k.y, n.ki[2] = k.z;
    n.e0[0] = e.x, n.e0[1] = e.y, n.e0[2] = e.z;
    return true;
}

// Composite Gauss-Legendre 2 x 3 rule on area_subdivisions^2 equal-area cells
// of every lit face (hex_atlas/src/atlas.cpp:11-51): side faces are unit-wide
// rectangles, basal faces six triangles mapped by the square root of the
// radial parameter. Weight = projected face area / sectors / n^2 x GL weight.
bool area_rule(const hex::Prism& h, d3 k, uint32_t n, std::vector<HostRay>& rays, std::string& err)
{
    rays.clear();
    const double ga = 1 / std::sqrt(3.0), gb = std::sqrt(3.0 / 5.0);
    const double u[2] = {(1 - ga) / 2, (1 + ga) / 2};
    const double v[3] = {(1 - gb) / 2, 0.5, (1 + gb) / 2};
    const double vw[3] = {5.0 / 18, 4.0 / 9, 5.0 / 18};
    Kahan total;
    for (unsigned f = 0; f < 8; ++f) {
        const double projected = h.face_area(f) * gt_max(0, -dot(hex::Prism::normal(f), k));
        if (projected == 0) continue;
        const unsigned sectors = f < 6 ? 1 : 6;
        const d3 normal = hex::Prism::normal(f);
        const d3 tangent = D3(-normal.z, 0, normal.x);
        for (unsigned sector = 0; sector < sectors; ++sector)
            for (uint32_t a = 0; a < n; ++a)
                for (uint32_t b = 0; b < n; ++b)
                    for (unsigned gi = 0; gi < 2; ++gi)
                        for (unsigned gj = 0; gj < 3; ++gj) {
                            const double x = (a + u[gi]) / n, y = (b + v[gj]) / n;
                            HostRay r;
                            r.face = f;
                            if (f < 6) {
                                r.pos = (normal * (kSqrt3 / 2) + tangent * (x - 0.5)) + D3(0, (2 * y - 1) * h.half(), 0);
                            } else {
                                r.pos = (hex::Prism::vertex(sector) * (1 - y) + hex::Prism::vertex(sector + 1) * y) *
                                        std::sqrt(x);
                                r.pos.y = (f == 6 ? 1 : -1) * h.half();
                            }
                            r.weight = (projected / sectors / double(n) / n) * 0.5 * vw[gj];
                            if (!(r.weight > 0) || !std::isfinite(r.weight) || !h.on_face(r.pos, f)) {
                                err = "Area quadrature is not representable / lies on ambiguous boundary.";
                                return false;
                            }
                            rays.push_back(r);
                            total.add(r.weight);
                        }
    }
    if (rays.empty() || std::fabs(total.v / h.projected_area(k) - 1) > 1e-12) {
        err = "Projected-area quadrature failed conservation audit.";
        return false;
    }
    return true;
}

// Fixed-point exponent of an atlas node, Q = 61 - ceil(log2(rays)) in [8, 60]:
// the node's emitted plus unresolved quanta (at most 2^Q per ray, plus
// rounding) stay below 2^63. A node has few rays, so Q is far above the 52-bit
// cap of the random-ray records.
uint32_t atlas_bits(uint64_t rays)
{
    uint32_t lg = 0;
    while (lg < 64 && (uint64_t(1) << lg) < rays) ++lg;
    int q = 61 - (int)lg;
    q = q > 60 ? 60 : (q < 8 ? 8 : q);
    return (uint32_t)q;
}

const char* status_text(uint32_t s)
{
    static const char* names[9] = {"exhausted", "tail_tolerance_reached", "interaction_limit", "ambiguous_boundary",
                                   "invalid_input", "numerical_failure", "output_overflow", "replay_mismatch",
                                   "missed"};
    return s < 9 ? names[s] : "unknown";
}

}  // namespace

extern "C" {

int iph_atlas_node_rays(const iph_atlas_settings* s, uint32_t i, uint32_t j, double* rays, uint64_t capacity,
                        uint64_t* count, iph_atlas_node* node)
{
    std::string err;
    if (!check_settings(s, err)) return fail(nullptr, IPH_E_INVALID, err);
    if (!count || !node) return fail(nullptr, IPH_E_INVALID, "Invalid arguments.");
    hex::Basis f;
    d3 k;
    if (!node_geometry(*s, i, j, *node, f, k, err)) return fail(nullptr, IPH_E_INVALID, err);
    const hex::Prism h{s->radius_mm, s->length_mm};
    std::vector<HostRay> rule;
    if (!area_rule(h, k, s->area_subdivisions, rule, err)) return fail(nullptr, IPH_E_INVALID, err);
    node->projected_area_mm2 = h.projected_area(k);
    node->rays = rule.size();
    *count = rule.size();
    if (rays) {
        if (capacity < rule.size()) return fail(nullptr, IPH_E_INVALID, "Ray buffer too small.");
        for (size_t r = 0; r < rule.size(); ++r) {
            double* o = rays + 5 * r;
            o[0] = rule[r].pos.x, o[1] = rule[r].pos.y, o[2] = rule[r].pos.z;
            o[3] = (double)rule[r].face;
            o[4] = rule[r].weight;
        }
    }
    return IPH_OK;
}

int iph_atlas_trace(iph_ctx* ctx, const iph_atlas_settings* s, uint32_t first, uint32_t count, double* mass_mm2,
                    iph_atlas_node* nodes)
{
    if (!ctx) return fail(nullptr, IPH_E_INVALID, "ctx must not be NULL.");
    std::string err;
    if (!check_settings(s, err)) return fail(ctx, IPH_E_INVALID, err);
    const uint64_t nj = (uint64_t)s->incident_phi_intervals + 1;
    const uint64_t total_nodes = ((uint64_t)s->incident_theta_intervals + 1) * nj;
    if (!mass_mm2 || !nodes || !count || (uint64_t)first + count > total_nodes)
        return fail(ctx, IPH_E_INVALID, "Node range outside the atlas, or NULL outputs.");
    const uint32_t nt = s->theta_count, np = s->phi_count;
    const uint64_t cells = (uint64_t)nt * np;
    IPH_HIP(ctx, hipSetDevice(ctx->device));

    // host set-up: node geometry, area rules, device records
    const hex::Prism prism{s->radius_mm, s->length_mm};
    std::vector<DevRecord> recs(count);
    std::vector<double> scale(count);  // mm^2 per quantum
    std::vector<AreaRay> all_rays;
    std::vector<size_t> ray_off(count);
    std::vector<double> edges;
    if (!make_u_edges(nt, edges, err)) return fail(ctx, IPH_E_INVALID, err);
    std::vector<HostRay> rule;
    for (uint32_t q = 0; q < count; ++q) {
        const uint64_t n = (uint64_t)first + q;
        hex::Basis f;
        d3 k;
        if (!node_geometry(*s, (uint32_t)(n / nj), (uint32_t)(n % nj), nodes[q], f, k, err))
            return fail(ctx, IPH_E_INVALID, err);
        if (!area_rule(prism, k, s->area_subdivisions, rule, err)) return fail(ctx, IPH_E_INVALID, err);
        double wmax = 0;
        for (const HostRay& r : rule) wmax = gt_max(wmax, r.weight);
        ray_off[q] = all_rays.size();
        for (const HostRay& r : rule) all_rays.push_back(AreaRay{r.pos, r.weight / wmax, r.face, 0});
        DevRecord& R = recs[q];
        std::memset(&R, 0, sizeof R);
        R.kind = kKindHex;
        R.nt = nt;
        R.np = np;
        R.rays = rule.size();
        R.tile_rays = kAtlasTileRays;
        R.qbits = atlas_bits(R.rays);
        R.qfactor = 1.0;
        R.owns = 1;
        R.frame.ki = k;
        R.frame.e0 = f.e0;
        R.frame.e1 = f.e1;
        hex::Config& c = R.hex;
        c.prism = prism;
        c.k = k;
        c.f = f;
        c.ni = s->exterior_index;
        c.nt = s->interior_index;
        c.tail = s->tail_tolerance;
        c.seed = 0;
        c.max_hits = s->max_internal_hits;
        c.area = prism.projected_area(k);
        scale[q] = std::ldexp(wmax, -(int)R.qbits);
        nodes[q].projected_area_mm2 = c.area;
        nodes[q].rays = rule.size();
    }

    // device buffers
    DevBuf<DevRecord> d_recs;
    DevBuf<AreaRay> d_rays;
    DevBuf<unsigned long long> d_bins, d_stats;
    DevBuf<double> d_edges;
    DevBuf<uint32_t> d_queue, d_kind_recs;
    DevBuf<uint64_t> d_tile_start;
    IPH_HIP(ctx, d_recs.alloc(count));
    IPH_HIP(ctx, d_rays.alloc(all_rays.size()));
    IPH_HIP(ctx, d_bins.alloc(count * cells));
    IPH_HIP(ctx, d_stats.alloc((size_t)count * kStatCount));
    IPH_HIP(ctx, d_edges.alloc(edges.size()));
    IPH_HIP(ctx, d_queue.alloc(4));
    IPH_HIP(ctx, d_kind_recs.alloc(count));
    IPH_HIP(ctx, d_tile_start.alloc((size_t)count + 1));
    std::vector<uint32_t> kind_recs(count);
    std::vector<uint64_t> tile_start(count + 1, 0);
    for (uint32_t q = 0; q < count; ++q) {
        DevRecord& R = recs[q];
        R.bins = d_bins.p + (size_t)q * cells;
        R.stats = d_stats.p + (size_t)q * kStatCount;
        R.edges = d_edges.p;
        R.area = d_rays.p + ray_off[q];
        kind_recs[q] = q;
        tile_start[q + 1] = tile_start[q] + (R.rays + R.tile_rays - 1) / R.tile_rays;
    }
    hipStream_t st = ctx->stream;
    IPH_HIP(ctx, hipMemcpyAsync(d_rays.p, all_rays.data(), all_rays.size() * sizeof(AreaRay), hipMemcpyHostToDevice, st));
    IPH_HIP(ctx, hipMemcpyAsync(d_edges.p, edges.data(), edges.size() * sizeof(double), hipMemcpyHostToDevice, st));
    IPH_HIP(ctx, hipMemcpyAsync(d_recs.p, recs.data(), count * sizeof(DevRecord), hipMemcpyHostToDevice, st));
    IPH_HIP(ctx, hipMemcpyAsync(d_kind_recs.p, kind_recs.data(), count * sizeof(uint32_t), hipMemcpyHostToDevice, st));
    IPH_HIP(ctx, hipMemcpyAsync(d_tile_start.p, tile_start.data(), (count + 1) * sizeof(uint64_t),
                                hipMemcpyHostToDevice, st));
    TileMap map{};
    map.recs_of_kind = d_kind_recs.p;
    map.tile_start = d_tile_start.p;
    map.n = count;
    map.total = tile_start[count];
    const uint32_t blocks = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)ctx->cus * 8, map.total));
    if (map.total + blocks > kMaxQueueTickets)
        return fail(ctx, IPH_E_INVALID, "too many area rays for one call's 32-bit tile queue; trace fewer nodes per call.");
    IPH_HIP(ctx, launch_zero(d_recs.p, count, d_queue.p, st));
    IPH_HIP(ctx, launch_trace_area(nt + 1 > 2048, d_recs.p, map, d_queue.p + 1, blocks, 256, st));
    // FP64 masses and exact quanta totals on the device, straight into the caller's array
    DevBuf<double> d_scale, d_mass;
    DevBuf<unsigned long long> d_totals;
    IPH_HIP(ctx, d_scale.alloc(count));
    IPH_HIP(ctx, d_mass.alloc(count * cells));
    IPH_HIP(ctx, d_totals.alloc(count));
    IPH_HIP(ctx, hipMemcpyAsync(d_scale.p, scale.data(), count * sizeof(double), hipMemcpyHostToDevice, st));
    IPH_HIP(ctx, hipMemsetAsync(d_totals.p, 0, count * sizeof(unsigned long long), st));
    {
        const uint32_t gx = (uint32_t)std::min<uint64_t>((cells + 255) / 256, 64);
        const uint32_t gy = std::min<uint32_t>(count, 65535);
        hipLaunchKernelGGL(icephase_atlas_mass, dim3(gx, gy), dim3(256), 0, st, d_bins.p, d_stats.p, d_scale.p, count,
                           nt, np, d_mass.p, d_totals.p);
        IPH_HIP(ctx, hipGetLastError());
    }
    std::vector<unsigned long long> stats((size_t)count * kStatCount), totals(count);
    IPH_HIP(ctx, hipMemcpyAsync(mass_mm2, d_mass.p, count * cells * sizeof(double), hipMemcpyDeviceToHost, st));
    IPH_HIP(ctx, hipMemcpyAsync(stats.data(), d_stats.p, stats.size() * sizeof(unsigned long long),
                                hipMemcpyDeviceToHost, st));
    IPH_HIP(ctx, hipMemcpyAsync(totals.data(), d_totals.p, count * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                st));
    IPH_HIP(ctx, hipStreamSynchronize(st));

    // readout and audits (hex_atlas/src/atlas.cpp:89-138)
    for (uint32_t q = 0; q < count; ++q) {
        const unsigned long long* S = stats.data() + (size_t)q * kStatCount;
        iph_atlas_node& node = nodes[q];
        const std::string where =
            "node (" + std::to_string(node.i) + "," + std::to_string(node.j) + "): ";
        for (uint32_t k = 2; k < 9; ++k)
            if (S[kStatStatus0 + k])
                return fail(ctx, IPH_E_AUDIT,
                            where + "Deterministic area ray rejected: " + std::to_string(S[kStatStatus0 + k]) +
                                " rays with status=" + status_text(k) +
                                ". No jitter, dropped ray, or renormalized omission is applied.");
        if (S[kStatInvalid])
            return fail(ctx, IPH_E_AUDIT, where + "Invalid outgoing direction in the atlas histogram.");
        double ray_balance, face_balance;
        std::memcpy(&ray_balance, &S[kStatRayBalance], sizeof ray_balance);
        std::memcpy(&face_balance, &S[kStatFaceBalance], sizeof face_balance);
        node.max_balance_error = ray_balance;
        node.max_interface_balance_error = face_balance;
        if (!(ray_balance <= s->max_balance_error) || !(face_balance <= s->max_balance_error))
            return fail(ctx, IPH_E_AUDIT, where + "Ray power balance exceeds the tolerance.");
        const double sc = scale[q];
        node.escaped_mm2 = (double)S[kStatEscaped] * sc;
        node.unresolved_mm2 = (double)S[kStatUnresolved] * sc;
        node.outputs = S[kStatSamples];
        // the cell masses are nonnegative multiples of the quantum; their total
        // is exact in integers (binned + pole quanta), which equals the escaped
        // quanta when no emission was lost
        const unsigned long long binned = totals[q] + S[kStatForward] + S[kStatBackward];
        if (!(binned > 0) || !(node.escaped_mm2 > 0) || !std::isfinite(sc))
            return fail(ctx, IPH_E_AUDIT, where + "Incomplete atlas node.");
        node.histogram_relative_error = std::fabs((double)binned / (double)S[kStatEscaped] - 1);
        const double area = node.projected_area_mm2;
        if (node.histogram_relative_error > s->max_balance_error ||
            std::fabs((node.escaped_mm2 + node.unresolved_mm2) / area - 1) > s->max_balance_error ||
            node.unresolved_mm2 / area > s->max_unresolved_fraction)
            return fail(ctx, IPH_E_AUDIT, where + "Atlas area / escaped / residual conservation audit failed.");
    }
    return IPH_OK;
}

int iph_ice_index(double wavelength_nm, iph_ice_state* out)
{
    if (!out) return fail(nullptr, IPH_E_INVALID, "Invalid arguments.");
    IceState st;
    std::string err;
    if (!ice_index_wb2008(wavelength_nm, st, err)) return fail(nullptr, IPH_E_INVALID, err);
    out->wavelength_nm = st.wavelength_nm;
    out->n = st.n;
    out->kappa = st.kappa;
    out->lower_nm = st.lower_nm;
    out->upper_nm = st.upper_nm;
    out->log_wavelength_fraction = st.log_wavelength_fraction;
    return IPH_OK;
}

int iph_write_npy(const char* path, const uint64_t* shape, uint32_t ndim, const double* data)
{
    if (!path || !shape || !ndim || ndim > 32 || !data) return fail(nullptr, IPH_E_INVALID, "Invalid NPY arguments.");
    std::string err;
    if (!write_npy(path, std::vector<uint64_t>(shape, shape + ndim), data, err)) return fail(nullptr, IPH_E_IO, err);
    return IPH_OK;
}

}  // extern "C"
