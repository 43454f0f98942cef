// C-ABI runtime of libicephase.so (declarations and reference citations in
// include/icephase.h). One HIP device and one default-priority stream per
// context; every entry point is synchronous; errors are returned as codes with
// the message kept per context (or per thread for context-free calls).
#include "host_records.hpp"
#include "runtime_internal.hpp"
#include "surrogate.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

namespace icephase {
hipError_t launch_zero(const DevRecord* d_recs, uint32_t nrec, uint32_t* d_queue, hipStream_t s);
hipError_t launch_trace(int kind, bool wide, const DevRecord* d_recs, TileMap map, uint32_t* d_queue,
                        uint32_t blocks, uint32_t threads, hipStream_t s, bool direct);
hipError_t launch_bin_dirs(const BinFrame& g, uint32_t nt, uint32_t np, const double* d_edges, uint64_t n,
                           const double* d_dirs, uint32_t* d_index, int32_t* d_kind, hipStream_t s);
hipError_t launch_cdf(const CdfJob* d_jobs, uint32_t njobs, uint32_t max_onp, bool any_coarse, uint32_t reduce_blocks,
                      hipStream_t s);
uint32_t cdf_partial_blocks(uint32_t reduce_blocks, uint32_t max_onp);
hipError_t launch_lobe(const DevRecord* d_recs, uint32_t rec, uint32_t rows, uint32_t np, hipStream_t s);
uint32_t lobe_blocks(uint32_t rows, uint32_t np);
}  // namespace icephase

using namespace icephase;

namespace {

struct CdfSetup {
    uint32_t ont, onp;
    double max_l1, max_lost, max_tv;
};

bool resolve_policy(const iph_cdf_policy* p, uint32_t nt, uint32_t np, CdfSetup& s, std::string& err)
{
    s.ont = p && p->out_theta ? p->out_theta : nt;
    s.onp = p && p->out_phi ? p->out_phi : np;
    s.max_l1 = p ? p->max_l1 : 1e-10;
    s.max_lost = p ? p->max_lost_mass : 1e-12;
    s.max_tv = p ? p->max_coarsening_tv : 1.0;
    if (!s.ont || !s.onp || s.ont > nt || s.onp > np || nt % s.ont || np % s.onp) {
        err = "Storage dimensions must divide histogram dimensions.";
        return false;
    }
    if (!(std::isfinite(s.max_l1) && s.max_l1 >= 0 && s.max_l1 <= 2 && std::isfinite(s.max_lost) &&
          s.max_lost >= 0 && s.max_lost <= 1 && std::isfinite(s.max_tv) && s.max_tv >= 0 && s.max_tv <= 1)) {
        err = "Invalid CDF quality policy.";
        return false;
    }
    return true;
}

// Acceptance rules of build_host_cdf (hex_phase/src/phase_table.cpp:111-187) and
// PhaseCdf::build (src/phase_cdf.cpp:448-468) applied to the device scalars.
iph_cdf_report make_report(const double* sc, const CdfSetup& s, bool coarse)
{
    iph_cdf_report r;
    std::memset(&r, 0, sizeof r);
    r.fine_mass = sc[kCdfMassFine];
    r.stored_mass = sc[kCdfMassStored];
    r.g_source = sc[kCdfAxialFine] / sc[kCdfMassFine];
    r.aggregation_error = coarse ? std::fabs(sc[kCdfMassStored] / sc[kCdfMassFine] - 1) : 0.0;
    r.coarsening_tv = sc[kCdfTv];
    r.l1_error = sc[kCdfL1];
    r.lost_mass = sc[kCdfLost];
    r.max_cell_error = sc[kCdfMaxCell];
    r.cdf_mass = sc[kCdfRecovered];
    r.g_stored = sc[kCdfAxialCdf] / sc[kCdfRecovered];
    r.lost_cells = (uint64_t)sc[kCdfLostCells];
    r.malformed = (uint64_t)sc[kCdfMalformed];
    r.weight_underflow = (uint64_t)sc[kCdfUnderflow];
    r.negative = (uint64_t)sc[kCdfNegative];
    if (r.negative)
        r.status = 1;  // negative / non-finite mass
    else if (!(sc[kCdfMassFine] > 0) || !std::isfinite(sc[kCdfMassFine]))
        r.status = 2;  // empty histogram
    else if (!std::isfinite(r.coarsening_tv) || r.coarsening_tv > s.max_tv || r.aggregation_error > 1e-12)
        r.status = 3;  // coarsening audit
    else if (r.weight_underflow)
        r.status = 4;
    else if (r.malformed || !(sc[kCdfTotal] > 0))
        r.status = 5;
    else if (r.l1_error > s.max_l1 || r.lost_mass > s.max_lost || std::fabs(r.cdf_mass - 1) > 1e-10 ||
             !std::isfinite(r.g_stored) || !(r.g_stored > -1 && r.g_stored < 1))
        r.status = 6;  // CDF mass / HG audit
    return r;
}

const char* report_reason(int status)
{
    switch (status) {
    case 1: return "Negative/nonfinite mass.";
    case 2: return "Empty/nonfinite histogram.";
    case 3: return "Conservative coarsening audit failed.";
    case 4: return "CDF weight underflow.";
    case 5: return "Malformed CDF; refusing repair.";
    case 6: return "CDF mass / HG audit failed.";
    default: return "ok";
    }
}

}  // namespace

// Fixed device scratch of iph_build_cdf_noalloc: one copy per GPU, held until
// the process ends. In host code `scratch` only names the symbol; its device
// address is looked up once per device (hipGetSymbolAddress, with the device
// current) when a context is created, kept in g_scratch and in the context, and
// only that address is passed to kernels and copies.
constexpr size_t kNoallocScratchBytes = 9u << 20;
constexpr uint32_t kNoallocMaxTheta = 4096;
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

namespace {

std::mutex g_scratch_mu;
std::vector<unsigned char*> g_scratch;  // per device ordinal

// Device address of the current device's scratch (looked up on first use).
int device_scratch(int device, int count, unsigned char** out, std::string& err)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch.size() < (size_t)count) g_scratch.resize(count, nullptr);
    if (!g_scratch[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return IPH_E_DEVICE;
        }
        g_scratch[device] = static_cast<unsigned char*>(p);
    }
    *out = g_scratch[device];
    return IPH_OK;
}

// Byte offsets of one allocation-free CDF inside the scratch (256-byte aligned).
struct NoallocLayout {
    size_t job, edges, cedges, cols, scalars, partials, coarse, bytes;
};

NoallocLayout noalloc_layout(uint32_t nt, const CdfSetup& s, uint32_t pblocks, bool coarse)
{
    auto up = [](size_t x) { return (x + 255) & ~size_t(255); };
    NoallocLayout L{};
    size_t o = 0;
    L.job = o; o = up(o + sizeof(CdfJob));
    L.edges = o; o = up(o + (size_t)(nt + 1) * sizeof(double));
    L.cedges = o; o = up(o + (size_t)(s.ont + 1) * sizeof(double));
    L.cols = o; o = up(o + (size_t)s.onp * sizeof(double));
    L.scalars = o; o = up(o + kCdfScalars * sizeof(double));
    L.partials = o; o = up(o + (size_t)pblocks * 8 * sizeof(double));
    L.coarse = o; o = up(o + (coarse ? (size_t)s.ont * s.onp * sizeof(double) : 0));
    L.bytes = o;
    return L;
}

// True when p is device (or managed) memory of the given device.
bool device_memory_of(const void* p, int device)
{
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();  // an unregistered host pointer: clear the query's error
        return false;
    }
    return (a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged) && a.device == device;
}

}  // namespace

struct iph_batch {
    iph_ctx* ctx = nullptr;
    uint32_t nrec = 0;
    std::vector<iph_record> recs;
    std::vector<DevRecord> host;
    std::vector<CdfSetup> cdf;
    std::vector<double> weight, area;
    std::vector<std::vector<double>> edges_host, cedges_host;
    // device
    DevBuf<DevRecord> d_recs;
    DevBuf<unsigned long long> d_bins, d_stats;
    DevBuf<double> d_edges;
    DevBuf<uint32_t> d_queue;
    DevBuf<uint32_t> d_kind_recs;
    DevBuf<uint64_t> d_tile_start;
    DevBuf<CdfJob> d_jobs;
    DevBuf<double> d_cdf;  // phi, theta, cols, scalars, partials, coarse
    std::vector<size_t> off_phi, off_theta;
    std::vector<size_t> off_bins;
    std::vector<CdfJob> jobs;           // one per histogram owner
    std::vector<int32_t> job_of;        // record -> job index, -1 for orientation nodes
    std::vector<uint32_t> owner;        // record -> owner record
    std::vector<uint64_t> group_rays;   // owner: rays of every record adding into it
    std::vector<double> group_area;     // owner: sum of node weight * projected area
    DevBuf<double> d_lobe;              // lobe cell masses and partial sums
    std::vector<uint32_t> lobe_recs;    // records with a diffraction lobe
    TileMap map[2]{};
    bool wide[2] = {false, false};
    uint32_t max_onp = 1, reduce_blocks = 512;
    bool any_coarse = false;
    uint32_t blocks[2] = {0, 0};
    uint64_t tile_rays = 0;
    bool cell_table = true;  // hex kernels: workgroup LDS cell table (false: every quantum straight to HBM)
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    // second stream (default priority) for the hex kernel of a mixed batch, with
    // the fork / join events that order it after the zero kernel and before the CDF
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    hipGraphExec_t graphs[2] = {nullptr, nullptr};  // [0] zero + trace, [1] CDF
    double ms[3] = {0, 0, 0};
    ~iph_batch()
    {
        if (side) {
            (void)hipStreamSynchronize(side);
            (void)hipStreamDestroy(side);
        }
        if (fork) (void)hipEventDestroy(fork);
        if (join) (void)hipEventDestroy(join);
        for (auto& g : graphs)
            if (g) (void)hipGraphExecDestroy(g);
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
    }
};

namespace icephase {
bool batch_record_grid(const iph_batch* b, uint32_t rec, uint32_t& nt, uint32_t& np)
{
    if (!b || rec >= b->nrec) return false;
    nt = b->host[rec].nt;
    np = b->host[rec].np;
    return true;
}
}  // namespace icephase

extern "C" {

int iph_abi_version(void) { return IPH_ABI_VERSION; }

uint32_t iph_struct_size(int which)
{
    switch (which) {
    case 0: return (uint32_t)sizeof(iph_record);
    case 1: return (uint32_t)sizeof(iph_hist_stats);
    case 2: return (uint32_t)sizeof(iph_cdf_report);
    case 3: return (uint32_t)sizeof(iph_cdf_policy);
    case 4: return (uint32_t)sizeof(iph_water_state);
    case 5: return (uint32_t)sizeof(iph_area);
    case 6: return (uint32_t)sizeof(iph_ice_state);
    case 7: return (uint32_t)sizeof(iph_atlas_settings);
    case 8: return (uint32_t)sizeof(iph_atlas_node);
    case 9: return (uint32_t)sizeof(iph_coherent_config);
    case 10: return (uint32_t)sizeof(iph_coherent_stats);
    case 11: return (uint32_t)sizeof(iph_coherent_direction);
    case 12: return (uint32_t)sizeof(iph_coherent_gate);
    default: return 0;
    }
}

const char* iph_last_error(const iph_ctx* ctx) { return ctx ? ctx->err.c_str() : thread_error().c_str(); }

int iph_ctx_create(int device, iph_ctx** out)
{
    if (!out) return fail(nullptr, IPH_E_INVALID, "out must not be NULL.");
    *out = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0)
        return fail(nullptr, IPH_E_DEVICE, std::string("No HIP device available: ") + hipGetErrorString(e));
    if (device < 0 || device >= count) return fail(nullptr, IPH_E_INVALID, "Invalid device ordinal.");
    auto c = std::make_unique<iph_ctx>();
    c->device = device;
    e = hipSetDevice(device);
    if (e != hipSuccess) return fail(nullptr, IPH_E_DEVICE, std::string("hipSetDevice: ") + hipGetErrorString(e));
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess)
        return fail(nullptr, IPH_E_DEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
        c->cus = prop.multiProcessorCount;
    std::string err;
    if (device_scratch(device, count, &c->cdf_scratch, err) != IPH_OK) {
        (void)hipStreamDestroy(c->stream);
        return fail(nullptr, IPH_E_DEVICE, err);
    }
    *out = c.release();
    return IPH_OK;
}

void iph_ctx_destroy(iph_ctx* ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    // batches still alive are owned by the context: release them first
    const std::vector<iph_batch*> live = ctx->batches;
    for (iph_batch* b : live) iph_batch_destroy(b);
    const std::vector<iph_sur*> sur_live = ctx->surrogates;
    for (iph_sur* s : sur_live) iph_sur_destroy(s);
    const std::vector<iph_atlas_mass*> atlas_live = ctx->atlases;
    for (iph_atlas_mass* m : atlas_live) iph_atlas_mass_destroy(m);
    const std::vector<iph_coherent*> coherent_live = ctx->coherents;
    for (iph_coherent* c : coherent_live) iph_coherent_destroy(c);
    if (ctx->stream) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipStreamDestroy(ctx->stream);
    }
    delete ctx;
}

int iph_u_edges(uint32_t nt, double* out)
{
    if (!out) return fail(nullptr, IPH_E_INVALID, "out must not be NULL.");
    std::vector<double> e;
    std::string err;
    if (!make_u_edges(nt, e, err)) return fail(nullptr, IPH_E_INVALID, err);
    std::memcpy(out, e.data(), e.size() * sizeof(double));
    return IPH_OK;
}

int iph_bin_directions(iph_ctx* ctx, uint32_t nt, uint32_t np, const double* frame, uint64_t n, const double* dirs,
                       uint32_t* index, int32_t* kind)
{
    if (!ctx) return fail(nullptr, IPH_E_INVALID, "ctx must not be NULL.");
    if (!frame || (n && (!dirs || !index || !kind)) || !nt || !np || (uint64_t)nt * np > 0xffffffffull)
        return fail(ctx, IPH_E_INVALID, "Invalid binning arguments.");
    if (!n) return IPH_OK;
    std::vector<double> e;
    std::string err;
    if (!make_u_edges(nt, e, err)) return fail(ctx, IPH_E_INVALID, err);
    IPH_HIP(ctx, hipSetDevice(ctx->device));
    BinFrame g;
    g.ki = D3(frame[0], frame[1], frame[2]);
    g.e0 = D3(frame[3], frame[4], frame[5]);
    g.e1 = D3(frame[6], frame[7], frame[8]);
    DevBuf<double> de, dd;
    DevBuf<uint32_t> di;
    DevBuf<int32_t> dk;
    IPH_HIP(ctx, de.alloc(e.size()));
    IPH_HIP(ctx, dd.alloc(3 * n));
    IPH_HIP(ctx, di.alloc(n));
    IPH_HIP(ctx, dk.alloc(n));
    IPH_HIP(ctx, hipMemcpyAsync(de.p, e.data(), e.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    IPH_HIP(ctx, hipMemcpyAsync(dd.p, dirs, 3 * n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    IPH_HIP(ctx, launch_bin_dirs(g, nt, np, de.p, n, dd.p, di.p, dk.p, ctx->stream));
    IPH_HIP(ctx, hipMemcpyAsync(index, di.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    IPH_HIP(ctx, hipMemcpyAsync(kind, dk.p, n * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    IPH_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return IPH_OK;
}

int iph_batch_create(iph_ctx* ctx, const iph_record* recs, uint32_t nrec, const iph_cdf_policy* policy,
                     iph_batch** out)
{
    if (!ctx) return fail(nullptr, IPH_E_INVALID, "ctx must not be NULL.");
    if (!out || !recs || nrec == 0) return fail(ctx, IPH_E_INVALID, "Need at least one record.");
    *out = nullptr;
    IPH_HIP(ctx, hipSetDevice(ctx->device));
    auto b = std::make_unique<iph_batch>();
    b->ctx = ctx;
    b->nrec = nrec;
    b->recs.assign(recs, recs + nrec);
    b->host.resize(nrec);
    b->cdf.resize(nrec);
    b->weight.resize(nrec);
    b->area.resize(nrec);
    b->edges_host.resize(nrec);
    b->cedges_host.resize(nrec);
    std::string err;
    size_t bins_total = 0, edges_total = 0;
    b->off_bins.resize(nrec);
    std::vector<size_t> off_edges(nrec), off_cedges(nrec);
    b->owner.resize(nrec);
    b->group_rays.assign(nrec, 0);
    b->group_area.assign(nrec, 0.0);
    for (uint32_t r = 0; r < nrec; ++r) {
        if (!prepare_record(recs[r], 0, b->host[r], b->weight[r], b->area[r], err))
            return fail(ctx, IPH_E_INVALID, "record " + std::to_string(r) + ": " + err);
        b->owner[r] = r;
        if (recs[r].histogram_of) {
            const uint32_t o = recs[r].histogram_of - 1;
            if (o >= r || recs[o].histogram_of != 0 || recs[o].shape != IPH_SHAPE_HEX || recs[r].shape != IPH_SHAPE_HEX ||
                recs[o].theta_count != recs[r].theta_count || recs[o].phi_count != recs[r].phi_count ||
                std::memcmp(recs[o].incident, recs[r].incident, sizeof recs[r].incident) != 0 ||
                !b->host[o].rotated || !b->host[r].rotated)
                return fail(ctx, IPH_E_INVALID,
                            "record " + std::to_string(r) +
                                ": an orientation node must follow its owner (a rotated hex record with the same grid "
                                "and ensemble incident direction).");
            b->owner[r] = o;
        }
        b->group_rays[b->owner[r]] += b->host[r].rays;
        b->group_area[b->owner[r]] += b->area[r];
    }
    // composition: every member quantizes with the group's exponent, weighted
    // relative to the group's largest sample weight
    {
        std::vector<double> max_scale(nrec, 0.0);
        for (uint32_t r = 0; r < nrec; ++r) max_scale[b->owner[r]] = std::max(max_scale[b->owner[r]], b->weight[r]);
        for (uint32_t r = 0; r < nrec; ++r) {
            const uint32_t o = b->owner[r];
            if (o == r && b->group_rays[r] == b->host[r].rays) continue;  // alone: unchanged
            b->host[r].qfactor = b->weight[r] / max_scale[o];
            b->host[r].qbits = quantum_bits(b->group_rays[o]);
        }
        for (uint32_t r = 0; r < nrec; ++r)
            if (b->owner[r] == r) b->weight[r] = max_scale[r];
    }
    for (uint32_t r = 0; r < nrec; ++r) {
        if (!resolve_policy(policy, recs[r].theta_count, recs[r].phi_count, b->cdf[r], err))
            return fail(ctx, IPH_E_INVALID, "record " + std::to_string(r) + ": " + err);
        if (!make_u_edges(recs[r].theta_count, b->edges_host[r], err)) return fail(ctx, IPH_E_INVALID, err);
        const uint32_t dt = recs[r].theta_count / b->cdf[r].ont;
        b->cedges_host[r].resize(b->cdf[r].ont + 1);
        for (uint32_t i = 0; i <= b->cdf[r].ont; ++i) b->cedges_host[r][i] = b->edges_host[r][(size_t)i * dt];
        b->off_bins[r] = b->owner[r] == r ? bins_total : b->off_bins[b->owner[r]];
        if (b->owner[r] == r) bins_total += (size_t)recs[r].theta_count * recs[r].phi_count;
        off_edges[r] = edges_total;
        edges_total += b->edges_host[r].size();
        off_cedges[r] = edges_total;
        edges_total += b->cedges_host[r].size();
        if (b->host[r].nt + 1 > 2048) b->wide[b->host[r].kind] = true;
    }
    IPH_HIP(ctx, b->d_recs.alloc(nrec));
    IPH_HIP(ctx, b->d_bins.alloc(bins_total));
    IPH_HIP(ctx, b->d_stats.alloc((size_t)nrec * kStatCount));
    IPH_HIP(ctx, b->d_edges.alloc(edges_total));
    IPH_HIP(ctx, b->d_queue.alloc(4));
    {
        std::vector<double> all(edges_total);
        for (uint32_t r = 0; r < nrec; ++r) {
            std::copy(b->edges_host[r].begin(), b->edges_host[r].end(), all.begin() + off_edges[r]);
            std::copy(b->cedges_host[r].begin(), b->cedges_host[r].end(), all.begin() + off_cedges[r]);
        }
        IPH_HIP(ctx, hipMemcpy(b->d_edges.p, all.data(), all.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    for (uint32_t r = 0; r < nrec; ++r) {
        b->host[r].bins = b->d_bins.p + b->off_bins[r];
        b->host[r].stats = b->d_stats.p + (size_t)b->owner[r] * kStatCount;
        b->host[r].edges = b->d_edges.p + off_edges[r];
    }
    // diffraction lobes: cell masses + partial sums (+ total, captured fraction)
    {
        size_t lobe_total = 0;
        std::vector<size_t> off_lobe(nrec, 0);
        for (uint32_t r = 0; r < nrec; ++r)
            if (b->host[r].lobe_on) {
                b->lobe_recs.push_back(r);
                off_lobe[r] = lobe_total;
                lobe_total += (size_t)b->host[r].lobe.rows * b->host[r].np +
                              lobe_blocks(b->host[r].lobe.rows, b->host[r].np) + 2;
            }
        IPH_HIP(ctx, b->d_lobe.alloc(lobe_total));
        for (uint32_t r : b->lobe_recs) {
            Lobe& L = b->host[r].lobe;
            L.mass = b->d_lobe.p + off_lobe[r];
            L.partial = L.mass + (size_t)L.rows * b->host[r].np;
        }
    }
    // CDF buffers
    size_t cdf_total = 0;
    b->off_phi.resize(nrec);
    b->off_theta.resize(nrec);
    std::vector<size_t> off_cols(nrec), off_sc(nrec), off_part(nrec), off_coarse(nrec);
    for (uint32_t r = 0; r < nrec; ++r) b->max_onp = std::max(b->max_onp, b->cdf[r].onp);
    const uint32_t pblocks = cdf_partial_blocks(b->reduce_blocks, b->max_onp);
    for (uint32_t r = 0; r < nrec; ++r) {
        if (b->owner[r] != r) continue;
        const CdfSetup& s = b->cdf[r];
        const bool coarse = s.ont != recs[r].theta_count || s.onp != recs[r].phi_count;
        b->any_coarse = b->any_coarse || coarse;
        b->off_phi[r] = cdf_total; cdf_total += s.onp + 1;
        b->off_theta[r] = cdf_total; cdf_total += (size_t)s.onp * (s.ont + 1);
        off_cols[r] = cdf_total; cdf_total += s.onp;
        off_sc[r] = cdf_total; cdf_total += kCdfScalars;
        off_part[r] = cdf_total; cdf_total += (size_t)pblocks * 8;
        off_coarse[r] = cdf_total; cdf_total += coarse ? (size_t)s.ont * s.onp : 0;
    }
    IPH_HIP(ctx, b->d_cdf.alloc(cdf_total));
    b->job_of.assign(nrec, -1);
    for (uint32_t r = 0; r < nrec; ++r) {
        if (b->owner[r] != r) continue;
        const CdfSetup& s = b->cdf[r];
        const bool coarse = s.ont != recs[r].theta_count || s.onp != recs[r].phi_count;
        b->job_of[r] = (int32_t)b->jobs.size();
        b->jobs.emplace_back();
        CdfJob& J = b->jobs.back();
        J.nt = recs[r].theta_count; J.np = recs[r].phi_count; J.ont = s.ont; J.onp = s.onp;
        J.bins = b->host[r].bins; J.mass = nullptr; J.stats = b->host[r].stats;
        J.edges = b->d_edges.p + off_edges[r]; J.cedges = b->d_edges.p + off_cedges[r];
        J.coarse = coarse ? b->d_cdf.p + off_coarse[r] : nullptr;
        J.phi = b->d_cdf.p + b->off_phi[r]; J.theta = b->d_cdf.p + b->off_theta[r];
        J.cols = b->d_cdf.p + off_cols[r]; J.scalars = b->d_cdf.p + off_sc[r];
        J.partials = b->d_cdf.p + off_part[r];
    }
    IPH_HIP(ctx, b->d_jobs.alloc(b->jobs.size()));
    IPH_HIP(ctx, hipMemcpy(b->d_jobs.p, b->jobs.data(), b->jobs.size() * sizeof(CdfJob), hipMemcpyHostToDevice));
    for (auto& e : b->ev) IPH_HIP(ctx, hipEventCreate(&e));
    IPH_HIP(ctx, hipStreamCreate(&b->side));
    IPH_HIP(ctx, hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
    IPH_HIP(ctx, hipEventCreateWithFlags(&b->join, hipEventDisableTiming));
    iph_batch* raw = b.release();
    const int rc = iph_batch_set_schedule(raw, 0, 0);
    if (rc != IPH_OK) {
        delete raw;
        return rc;
    }
    ctx->batches.push_back(raw);
    *out = raw;
    return IPH_OK;
}

void iph_batch_destroy(iph_batch* b)
{
    if (!b) return;
    iph_ctx* ctx = b->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    auto it = std::find(ctx->batches.begin(), ctx->batches.end(), b);
    if (it != ctx->batches.end()) ctx->batches.erase(it);
    delete b;
}

int iph_batch_set_schedule(iph_batch* b, uint64_t tile_rays, uint32_t blocks)
{
    if (!b) return fail(nullptr, IPH_E_INVALID, "batch must not be NULL.");
    iph_ctx* ctx = b->ctx;
    // The tile queue and a tile's ray counter are 32-bit device counters: a
    // workgroup takes one more queue ticket than there are tiles before it sees
    // the end, and a tile's counter runs past its ray count by at most one
    // ticket per wave. Schedules that could wrap either are refused here.
    if (tile_rays > kMaxTileRays)
        return fail(ctx, IPH_E_INVALID, "tile_rays must be at most 2^31.");
    {
        uint64_t tiles = 0;
        for (uint32_t r = 0; r < b->nrec; ++r) {
            uint64_t t = tile_rays ? tile_rays : default_tile_rays(b->host[r].rays);
            tiles += (b->host[r].rays + t - 1) / t;
        }
        const uint64_t want = blocks ? blocks : (uint64_t)ctx->cus * 8;
        if (tiles + want > kMaxQueueTickets)
            return fail(ctx, IPH_E_INVALID, "schedule has more tiles than the 32-bit tile queue holds; "
                                            "use larger tiles or fewer rays per batch.");
    }
    IPH_HIP(ctx, hipSetDevice(ctx->device));
    b->tile_rays = tile_rays;
    std::vector<uint32_t> kind_recs[2];
    std::vector<uint64_t> starts[2];
    for (int k = 0; k < 2; ++k) starts[k].push_back(0);
    for (uint32_t r = 0; r < b->nrec; ++r) {
        DevRecord& R = b->host[r];
        const uint64_t t = tile_rays ? tile_rays : default_tile_rays(R.rays);
        R.tile_rays = t;
        kind_recs[R.kind].push_back(r);
        starts[R.kind].push_back(starts[R.kind].back() + (R.rays + t - 1) / t);
    }
    IPH_HIP(ctx, hipMemcpy(b->d_recs.p, b->host.data(), b->nrec * sizeof(DevRecord), hipMemcpyHostToDevice));
    size_t n_kind = kind_recs[0].size() + kind_recs[1].size();
    IPH_HIP(ctx, b->d_kind_recs.alloc(n_kind));
    IPH_HIP(ctx, b->d_tile_start.alloc(n_kind + 2));
    size_t o_rec = 0, o_start = 0;
    for (int k = 0; k < 2; ++k) {
        TileMap& m = b->map[k];
        m.n = (uint32_t)kind_recs[k].size();
        m.total = starts[k].back();
        m.recs_of_kind = b->d_kind_recs.p + o_rec;
        m.tile_start = b->d_tile_start.p + o_start;
        if (m.n) {
            IPH_HIP(ctx, hipMemcpy(b->d_kind_recs.p + o_rec, kind_recs[k].data(), m.n * sizeof(uint32_t),
                                   hipMemcpyHostToDevice));
            IPH_HIP(ctx, hipMemcpy(b->d_tile_start.p + o_start, starts[k].data(), (m.n + 1) * sizeof(uint64_t),
                                   hipMemcpyHostToDevice));
        }
        o_rec += m.n;
        o_start += m.n + 1;
        const uint64_t want = blocks ? blocks : (uint64_t)ctx->cus * 8;
        b->blocks[k] = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(want, m.total ? m.total : 1));
    }
    for (auto& g : b->graphs)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
    return IPH_OK;
}

int iph_batch_set_cell_table(iph_batch* b, uint32_t enable)
{
    if (!b) return fail(nullptr, IPH_E_INVALID, "batch must not be NULL.");
    if (enable > 1) return fail(b->ctx, IPH_E_INVALID, "cell table switch must be 0 or 1.");
    b->cell_table = enable != 0;
    // the captured graphs hold the kernel of the previous choice
    for (auto& g : b->graphs)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
    return IPH_OK;
}

static int enqueue(iph_batch* b, uint32_t flags, bool timed)
{
    iph_ctx* ctx = b->ctx;
    hipStream_t s = ctx->stream;
    if (timed) IPH_HIP(ctx, hipEventRecord(b->ev[0], s));
    if (flags & IPH_RUN_TRACE) {
        IPH_HIP(ctx, launch_zero(b->d_recs.p, b->nrec, b->d_queue.p, s));
        // A mixed batch runs its two kinds at once: the FP64 hex kernel holds one
        // wave per SIMD, the drop waves fill the rest (both add order-free into
        // their own records' slabs, so the histograms are the same bits).
        const bool both = b->map[0].total && b->map[1].total && b->side;
        if (both) {
            IPH_HIP(ctx, hipEventRecord(b->fork, s));
            IPH_HIP(ctx, hipStreamWaitEvent(b->side, b->fork, 0));
            IPH_HIP(ctx, launch_trace(1, b->wide[1], b->d_recs.p, b->map[1], b->d_queue.p + 1, b->blocks[1], 256,
                                      b->side, !b->cell_table));
            IPH_HIP(ctx, launch_trace(0, b->wide[0], b->d_recs.p, b->map[0], b->d_queue.p, b->blocks[0], 256, s,
                                      false));
            IPH_HIP(ctx, hipEventRecord(b->join, b->side));
            IPH_HIP(ctx, hipStreamWaitEvent(s, b->join, 0));
        } else {
            for (int k = 0; k < 2; ++k)
                if (b->map[k].total)
                    IPH_HIP(ctx, launch_trace(k, b->wide[k], b->d_recs.p, b->map[k], b->d_queue.p + k, b->blocks[k],
                                              256, s, k == 1 && !b->cell_table));
        }
        for (uint32_t r : b->lobe_recs)
            IPH_HIP(ctx, launch_lobe(b->d_recs.p, r, b->host[r].lobe.rows, b->host[r].np, s));
    }
    if (timed) IPH_HIP(ctx, hipEventRecord(b->ev[1], s));
    if (flags & IPH_RUN_CDF)
        IPH_HIP(ctx, launch_cdf(b->d_jobs.p, (uint32_t)b->jobs.size(), b->max_onp, b->any_coarse, b->reduce_blocks, s));
    if (timed) IPH_HIP(ctx, hipEventRecord(b->ev[2], s));
    return IPH_OK;
}

static int check_cdf(iph_batch* b)
{
    iph_ctx* ctx = b->ctx;
    for (uint32_t r = 0; r < b->nrec; ++r) {
        if (b->job_of[r] < 0) continue;
        const CdfJob& J = b->jobs[b->job_of[r]];
        double sc[kCdfScalars];
        IPH_HIP(ctx, hipMemcpy(sc, J.scalars, sizeof sc, hipMemcpyDeviceToHost));
        const bool coarse = J.coarse != nullptr;
        const iph_cdf_report rep = make_report(sc, b->cdf[r], coarse);
        if (rep.status)
            return fail(ctx, IPH_E_AUDIT, "record " + std::to_string(r) + ": " + report_reason(rep.status));
    }
    return IPH_OK;
}

int iph_batch_run(iph_batch* b, uint32_t flags)
{
    if (!b) return fail(nullptr, IPH_E_INVALID, "batch must not be NULL.");
    iph_ctx* ctx = b->ctx;
    IPH_HIP(ctx, hipSetDevice(ctx->device));
    const uint32_t work = flags & (IPH_RUN_TRACE | IPH_RUN_CDF);
    if (!work) return fail(ctx, IPH_E_INVALID, "Nothing to run (flags).");
    if (flags & IPH_RUN_GRAPH) {
        // Two captured graphs (zero + trace, CDF) replayed back to back with
        // events between them: per-stage device times of every replay.
        for (int stage = 0; stage < 2; ++stage) {
            const uint32_t part = stage == 0 ? IPH_RUN_TRACE : IPH_RUN_CDF;
            if (!(work & part) || b->graphs[stage]) continue;
            hipGraph_t g = nullptr;
            IPH_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            const int rc = enqueue(b, part, false);
            const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
            if (rc != IPH_OK) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (e != hipSuccess) return fail(ctx, IPH_E_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
            const hipError_t ie = hipGraphInstantiate(&b->graphs[stage], g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ie != hipSuccess)
                return fail(ctx, IPH_E_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
        }
        IPH_HIP(ctx, hipEventRecord(b->ev[0], ctx->stream));
        if (work & IPH_RUN_TRACE) IPH_HIP(ctx, hipGraphLaunch(b->graphs[0], ctx->stream));
        IPH_HIP(ctx, hipEventRecord(b->ev[1], ctx->stream));
        if (work & IPH_RUN_CDF) IPH_HIP(ctx, hipGraphLaunch(b->graphs[1], ctx->stream));
        IPH_HIP(ctx, hipEventRecord(b->ev[2], ctx->stream));
        IPH_HIP(ctx, hipStreamSynchronize(ctx->stream));
        float t0 = 0, t1 = 0;
        IPH_HIP(ctx, hipEventElapsedTime(&t0, b->ev[0], b->ev[1]));
        IPH_HIP(ctx, hipEventElapsedTime(&t1, b->ev[1], b->ev[2]));
        b->ms[0] = t0;
        b->ms[1] = t1;
        b->ms[2] = (double)t0 + t1;
    } else {
        const int rc = enqueue(b, work, true);
        if (rc != IPH_OK) return rc;
        IPH_HIP(ctx, hipStreamSynchronize(ctx->stream));
        float t0 = 0, t1 = 0;
        IPH_HIP(ctx, hipEventElapsedTime(&t0, b->ev[0], b->ev[1]));
        IPH_HIP(ctx, hipEventElapsedTime(&t1, b->ev[1], b->ev[2]));
        b->ms[0] = t0;
        b->ms[1] = t1;
        b->ms[2] = (double)t0 + t1;
    }
    if ((work & IPH_RUN_CDF) && !(flags & IPH_RUN_NO_AUDIT)) return check_cdf(b);
    return IPH_OK;
}

int iph_batch_timing(const iph_batch* b, double* ms)
{
    if (!b || !ms) return fail(nullptr, IPH_E_INVALID, "Invalid arguments.");
    std::memcpy(ms, b->ms, sizeof b->ms);
    return IPH_OK;
}

static int node_error(iph_batch* b, uint32_t rec)
{
    return fail(b->ctx, IPH_E_INVALID,
                "record " + std::to_string(rec) + " is an orientation node of record " + std::to_string(b->owner[rec]) +
                    "; read the owner's histogram.");
}

int iph_batch_stats(iph_batch* b, uint32_t rec, iph_hist_stats* out)
{
    if (!b || !out || rec >= b->nrec) return fail(b ? b->ctx : nullptr, IPH_E_INVALID, "Invalid record index.");
    if (b->owner[rec] != rec) return node_error(b, rec);
    iph_ctx* ctx = b->ctx;
    IPH_HIP(ctx, hipSetDevice(ctx->device));
    unsigned long long st[kStatCount];
    IPH_HIP(ctx, hipMemcpy(st, b->host[rec].stats, sizeof st, hipMemcpyDeviceToHost));
    std::memset(out, 0, sizeof *out);
    out->q_forward = st[kStatForward];
    out->q_backward = st[kStatBackward];
    out->samples = st[kStatSamples];
    out->invalid = st[kStatInvalid];
    out->snapped = st[kStatSnapped];
    for (int k = 0; k < 9; ++k) out->status[k] = st[kStatStatus0 + k];
    out->q_escaped = st[kStatEscaped];
    out->q_unresolved = st[kStatUnresolved];
    out->internal_hits = st[kStatInternalHits];
    out->rays = b->group_rays[rec];
    out->qbits = b->host[rec].qbits;
    out->ray_weight_mm2 = b->weight[rec];
    out->projected_area_mm2 = b->group_area[rec];
    out->q_diffraction = st[kStatDiffraction];
    out->cell_atomics = st[kStatCellAtomics];
    out->diffraction_captured = 0;
    bool first = true;
    for (uint32_t r : b->lobe_recs) {
        if (b->owner[r] != rec) continue;
        const Lobe& L = b->host[r].lobe;
        double cap = 0;
        IPH_HIP(ctx, hipMemcpy(&cap, L.partial + lobe_blocks(L.rows, b->host[r].np) + 1, sizeof cap,
                               hipMemcpyDeviceToHost));
        out->diffraction_captured = first ? cap : std::min(out->diffraction_captured, cap);
        first = false;
    }
    return IPH_OK;
}

int iph_batch_bins(iph_batch* b, uint32_t rec, uint64_t* out)
{
    if (!b || !out || rec >= b->nrec) return fail(b ? b->ctx : nullptr, IPH_E_INVALID, "Invalid record index.");
    if (b->owner[rec] != rec) return node_error(b, rec);
    iph_ctx* ctx = b->ctx;
    IPH_HIP(ctx, hipSetDevice(ctx->device));
    const size_t n = (size_t)b->host[rec].nt * b->host[rec].np;
    IPH_HIP(ctx, hipMemcpy(out, b->host[rec].bins, n * sizeof(uint64_t), hipMemcpyDeviceToHost));
    return IPH_OK;
}

int iph_batch_cdf(iph_batch* b, uint32_t rec, double* phi, double* theta, iph_cdf_report* report)
{
    if (!b || rec >= b->nrec) return fail(b ? b->ctx : nullptr, IPH_E_INVALID, "Invalid record index.");
    if (b->owner[rec] != rec) return node_error(b, rec);
    iph_ctx* ctx = b->ctx;
    IPH_HIP(ctx, hipSetDevice(ctx->device));
    const CdfJob& J = b->jobs[b->job_of[rec]];
    if (phi) IPH_HIP(ctx, hipMemcpy(phi, J.phi, (J.onp + 1) * sizeof(double), hipMemcpyDeviceToHost));
    if (theta)
        IPH_HIP(ctx, hipMemcpy(theta, J.theta, (size_t)J.onp * (J.ont + 1) * sizeof(double), hipMemcpyDeviceToHost));
    if (report) {
        double sc[kCdfScalars];
        IPH_HIP(ctx, hipMemcpy(sc, J.scalars, sizeof sc, hipMemcpyDeviceToHost));
        *report = make_report(sc, b->cdf[rec], J.coarse != nullptr);
    }
    return IPH_OK;
}

}  // extern "C"

namespace icephase {
namespace {
// The CDF of a host or device mass array (device_mass: already on the context's device).
int cdf_of_mass(iph_ctx* ctx, const double* mass, bool device_mass, uint32_t nt, uint32_t np,
                const iph_cdf_policy* policy, double* phi, double* theta, iph_cdf_report* report)
{
    CdfSetup s;
    std::string err;
    if (!resolve_policy(policy, nt, np, s, err)) return fail(ctx, IPH_E_INVALID, err);
    std::vector<double> e;
    if (!make_u_edges(nt, e, err)) return fail(ctx, IPH_E_INVALID, err);
    const uint32_t dt = nt / s.ont;
    std::vector<double> ce(s.ont + 1);
    for (uint32_t i = 0; i <= s.ont; ++i) ce[i] = e[(size_t)i * dt];
    IPH_HIP(ctx, hipSetDevice(ctx->device));
    const bool coarse = s.ont != nt || s.onp != np;
    const uint32_t red = 512;
    const uint32_t pblocks = cdf_partial_blocks(red, s.onp);
    const size_t n = (size_t)nt * np;
    const size_t sz_theta = (size_t)s.onp * (s.ont + 1);
    DevBuf<double> buf;
    const size_t total = (device_mass ? 0 : n) + e.size() + ce.size() + (s.onp + 1) + sz_theta + s.onp + kCdfScalars +
                         (size_t)pblocks * 8 + (coarse ? (size_t)s.ont * s.onp : 0);
    IPH_HIP(ctx, buf.alloc(total));
    double* p = buf.p;
    CdfJob J;
    std::memset(&J, 0, sizeof J);
    J.nt = nt; J.np = np; J.ont = s.ont; J.onp = s.onp;
    const double* d_mass = mass;
    if (!device_mass) {
        IPH_HIP(ctx, hipMemcpyAsync(p, mass, n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        d_mass = p;
        p += n;
    }
    double* d_e = p; p += e.size();
    double* d_ce = p; p += ce.size();
    J.mass = d_mass; J.edges = d_e; J.cedges = d_ce;
    J.phi = p; p += s.onp + 1;
    J.theta = p; p += sz_theta;
    J.cols = p; p += s.onp;
    J.scalars = p; p += kCdfScalars;
    J.partials = p; p += (size_t)pblocks * 8;
    J.coarse = coarse ? p : nullptr;
    IPH_HIP(ctx, hipMemcpyAsync(d_e, e.data(), e.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    IPH_HIP(ctx, hipMemcpyAsync(d_ce, ce.data(), ce.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    DevBuf<CdfJob> dj;
    IPH_HIP(ctx, dj.alloc(1));
    IPH_HIP(ctx, hipMemcpyAsync(dj.p, &J, sizeof J, hipMemcpyHostToDevice, ctx->stream));
    IPH_HIP(ctx, launch_cdf(dj.p, 1, s.onp, coarse, red, ctx->stream));
    double sc[kCdfScalars];
    IPH_HIP(ctx, hipMemcpyAsync(phi, J.phi, (s.onp + 1) * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    IPH_HIP(ctx, hipMemcpyAsync(theta, J.theta, sz_theta * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    IPH_HIP(ctx, hipMemcpyAsync(sc, J.scalars, sizeof sc, hipMemcpyDeviceToHost, ctx->stream));
    IPH_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const iph_cdf_report rep = make_report(sc, s, coarse);
    if (report) *report = rep;
    if (rep.status) return fail(ctx, IPH_E_AUDIT, report_reason(rep.status));
    return IPH_OK;
}
}  // namespace

int build_cdf_from_device(iph_ctx* ctx, const double* d_mass, uint32_t nt, uint32_t np, const iph_cdf_policy* policy,
                          double* phi, double* theta, iph_cdf_report* report)
{
    return cdf_of_mass(ctx, d_mass, true, nt, np, policy, phi, theta, report);
}
}  // namespace icephase

extern "C" {

int iph_build_cdf(iph_ctx* ctx, const double* mass, uint32_t nt, uint32_t np, const iph_cdf_policy* policy,
                  double* phi, double* theta, iph_cdf_report* report)
{
    if (!ctx) return fail(nullptr, IPH_E_INVALID, "ctx must not be NULL.");
    if (!mass || !phi || !theta || !nt || !np || (uint64_t)nt * np > 0xffffffffull)
        return fail(ctx, IPH_E_INVALID, "Invalid CDF arguments.");
    return cdf_of_mass(ctx, mass, false, nt, np, policy, phi, theta, report);
}

int iph_build_cdf_noalloc(iph_ctx* ctx, const double* d_mass, uint32_t nt, uint32_t np, const iph_cdf_policy* policy,
                          double* d_phi, double* d_theta, iph_cdf_report* report)
{
    if (!ctx) return fail(nullptr, IPH_E_INVALID, "ctx must not be NULL.");
    if (!d_mass || !d_phi || !d_theta || !nt || !np || (uint64_t)nt * np > 0xffffffffull)
        return fail(ctx, IPH_E_INVALID, "Invalid CDF arguments.");
    if (nt > kNoallocMaxTheta)
        return fail(ctx, IPH_E_INVALID, "theta_count above 4096, the allocation-free CDF's limit.");
    if (!ctx->cdf_scratch) return fail(ctx, IPH_E_DEVICE, "The context has no device scratch.");
    CdfSetup s;
    std::string err;
    if (!resolve_policy(policy, nt, np, s, err)) return fail(ctx, IPH_E_INVALID, err);
    const bool coarse = s.ont != nt || s.onp != np;
    const uint32_t red = 512;
    const uint32_t pblocks = cdf_partial_blocks(red, s.onp);
    const NoallocLayout L = noalloc_layout(nt, s, pblocks, coarse);
    if (L.bytes > kNoallocScratchBytes)
        return fail(ctx, IPH_E_INVALID,
                    "This CDF needs " + std::to_string(L.bytes) + " bytes of scratch; the allocation-free entry point "
                    "holds " + std::to_string(kNoallocScratchBytes) + " (a coarsened grid of at most about 1.1 M cells). "
                    "Use iph_build_cdf.");
    IPH_HIP(ctx, hipSetDevice(ctx->device));
    if (!device_memory_of(d_mass, ctx->device) || !device_memory_of(d_phi, ctx->device) ||
        !device_memory_of(d_theta, ctx->device))
        return fail(ctx, IPH_E_INVALID, "mass, phi and theta must be device memory of the context's device.");
    double e[kNoallocMaxTheta + 1], ce[kNoallocMaxTheta + 1];
    if (!make_u_edges(nt, e, err)) return fail(ctx, IPH_E_INVALID, err);
    const uint32_t dt = nt / s.ont;
    for (uint32_t i = 0; i <= s.ont; ++i) ce[i] = e[(size_t)i * dt];
    unsigned char* base = ctx->cdf_scratch;
    CdfJob J;
    std::memset(&J, 0, sizeof J);
    J.nt = nt; J.np = np; J.ont = s.ont; J.onp = s.onp;
    J.mass = d_mass;
    J.edges = reinterpret_cast<double*>(base + L.edges);
    J.cedges = reinterpret_cast<double*>(base + L.cedges);
    J.phi = d_phi;
    J.theta = d_theta;
    J.cols = reinterpret_cast<double*>(base + L.cols);
    J.scalars = reinterpret_cast<double*>(base + L.scalars);  // cleared by the first CDF kernel
    J.partials = reinterpret_cast<double*>(base + L.partials);
    J.coarse = coarse ? reinterpret_cast<double*>(base + L.coarse) : nullptr;
    CdfJob* d_job = reinterpret_cast<CdfJob*>(base + L.job);
    hipStream_t st = ctx->stream;
    IPH_HIP(ctx, hipMemcpyAsync(base + L.edges, e, (size_t)(nt + 1) * sizeof(double), hipMemcpyHostToDevice, st));
    IPH_HIP(ctx, hipMemcpyAsync(base + L.cedges, ce, (size_t)(s.ont + 1) * sizeof(double), hipMemcpyHostToDevice, st));
    IPH_HIP(ctx, hipMemcpyAsync(d_job, &J, sizeof J, hipMemcpyHostToDevice, st));
    IPH_HIP(ctx, launch_cdf(d_job, 1, s.onp, coarse, red, st));
    double sc[kCdfScalars];
    IPH_HIP(ctx, hipMemcpyAsync(sc, J.scalars, sizeof sc, hipMemcpyDeviceToHost, st));
    // the host arrays above stay alive until the copies have run
    IPH_HIP(ctx, hipStreamSynchronize(st));
    const iph_cdf_report rep = make_report(sc, s, coarse);
    if (report) *report = rep;
    if (rep.status) return fail(ctx, IPH_E_AUDIT, report_reason(rep.status));
    return IPH_OK;
}

int iph_write_record(const char* out_dir, const char* metadata_json, const double* phi, const double* theta,
                     uint32_t nt, uint32_t np)
{
    if (!out_dir || !metadata_json || !phi || !theta || !nt || !np)
        return fail(nullptr, IPH_E_INVALID, "Invalid record arguments.");
    std::vector<double> e;
    std::string err;
    if (!make_u_edges(nt, e, err)) return fail(nullptr, IPH_E_INVALID, err);
    const int rc = write_record_dir(out_dir, metadata_json, phi, theta, e, nt, np, err);
    if (rc != IPH_OK) return fail(nullptr, rc, err);
    return IPH_OK;
}

int iph_batch_probabilities(iph_batch* b, uint32_t rec, float* d_out)
{
    if (!b || !d_out || rec >= b->nrec) return fail(b ? b->ctx : nullptr, IPH_E_INVALID, "Invalid arguments.");
    if (b->owner[rec] != rec) return node_error(b, rec);
    iph_ctx* ctx = b->ctx;
    IPH_HIP(ctx, hipSetDevice(ctx->device));
    unsigned long long st[kStatCount];
    IPH_HIP(ctx, hipMemcpy(st, b->host[rec].stats, sizeof st, hipMemcpyDeviceToHost));
    IPH_HIP(ctx, sur::export_probs(b->host[rec].bins, st[kStatForward], st[kStatBackward], b->host[rec].nt,
                                   b->host[rec].np, b->host[rec].stats + kStatScratch, d_out, ctx->stream));
    IPH_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return IPH_OK;
}

int iph_water_optics(double wavelength_nm, double celsius, double pascal, iph_water_state* out)
{
    if (!out) return fail(nullptr, IPH_E_INVALID, "out must not be NULL.");
    WaterState w;
    std::string err;
    if (!water_optics(wavelength_nm, celsius, pascal, w, err)) return fail(nullptr, IPH_E_INVALID, err);
    out->temperature_kelvin = w.temperature_kelvin;
    out->pressure_pascal = w.pressure_pascal;
    out->density_kg_m3 = w.density_kg_m3;
    out->density_pressure_residual_pascal = w.density_pressure_residual_pascal;
    out->density_bracket_width_kg_m3 = w.density_bracket_width_kg_m3;
    out->refractive_index = w.refractive_index;
    return IPH_OK;
}

int iph_water_index(double wavelength_nm, double kelvin, double density_kg_m3, double* n)
{
    if (!n) return fail(nullptr, IPH_E_INVALID, "n must not be NULL.");
    std::string err;
    if (!water_index_r9_97(wavelength_nm, kelvin, density_kg_m3, *n, err)) return fail(nullptr, IPH_E_INVALID, err);
    return IPH_OK;
}

int iph_drop_shape(float radius_mm, int sphere, float coefficients[8])
{
    if (!coefficients) return fail(nullptr, IPH_E_INVALID, "coefficients must not be NULL.");
    std::string err;
    if (!drop_shape_table(radius_mm, sphere != 0, coefficients, err)) return fail(nullptr, IPH_E_INVALID, err);
    return IPH_OK;
}

int iph_drop_projected_area(const double coefficients[8], double radius_mm, const double incident[3], iph_area* out)
{
    if (!coefficients || !incident || !out) return fail(nullptr, IPH_E_INVALID, "Invalid arguments.");
    AreaResult a;
    std::string err;
    if (!drop_projected_area(coefficients, radius_mm, incident, a, err)) return fail(nullptr, IPH_E_INVALID, err);
    std::memset(out, 0, sizeof *out);
    out->area_mm2 = a.area_mm2;
    out->abs_error_mm2 = a.abs_error_mm2;
    out->rel_error = a.rel_error;
    out->evaluations = a.evaluations;
    out->convexity_cells = a.convexity_cells;
    out->analytic_sphere = a.analytic_sphere ? 1 : 0;
    return IPH_OK;
}

int iph_hex_projected_area(double radius_mm, double length_mm, const double incident[3], double* area_mm2)
{
    if (!incident || !area_mm2) return fail(nullptr, IPH_E_INVALID, "Invalid arguments.");
    hex::Prism p;
    p.R = radius_mm;
    p.L = length_mm;
    if (!p.valid()) return fail(nullptr, IPH_E_INVALID, "Invalid hex prism geometry.");
    d3 k;
    if (!unit_of(D3(incident[0], incident[1], incident[2]), k))
        return fail(nullptr, IPH_E_INVALID, "Incident direction must be finite and nonzero.");
    *area_mm2 = p.projected_area(k);
    return IPH_OK;
}

}  // extern "C"
