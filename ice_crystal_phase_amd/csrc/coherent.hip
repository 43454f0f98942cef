// Coherent drop stages on gfx950: vertex trace -> patches -> folded branches ->
// patch query -> per-direction optics (incoherent, path, focal) -> Table II
// diffraction -> density -> CDF. One record at a time, everything resident in
// HBM between stages; only counters cross to the host.
//
// Reference pipeline replaced (apps/generate_phase.cpp:426-440):
//   RaindropTracer::trace_unpolarized   -> icephase_coh_trace (own intersector, no BVH)
//   PatchAccel::build                   -> icephase_coh_classify + scan + icephase_coh_compact
//   PatchOptics::prepare_folded         -> icephase_coh_fold (slots from a scan, not an atomic ticket)
//   PatchQuery::query_grid (OptiX)      -> icephase_coh_extent + scan + icephase_coh_query (count / fill)
//                                          + icephase_coh_finish (per-direction order and summary)
//   PatchOptics::evaluate_wave_device   -> icephase_coh_optics, icephase_coh_detect, icephase_coh_filter
//   PhaseCdf::build (density gate, w = (d / max) du, ordered scans) -> icephase_coh_density + the CDF kernels
//
// The query replaces the OptiX BVH with a direct enumeration: each patch's
// axis-aligned box (the reference's padded PatchAabb) bounds a cone of
// directions, the cone a (theta, phi) window of query cells; every cell in the
// window whose direction ray meets the box is tested with the bilinear solver.
// Hits are collected in two passes (count, then fill at atomically reserved
// slots) and ordered per direction by (patch id, compact index, root), so the
// result does not depend on the order in which cells and patches are visited.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "coherent_device.hpp"
#include "host_records.hpp"
#include "kernels.hpp"
#include "runtime_internal.hpp"

using namespace icephase;
using namespace icephase::coh;

namespace {

constexpr int kThreads = 256;

struct TraceParams {
    drop::Shape shape;
    v3 w, bx, e1;
    float half, ref_in, ref_out, radius_mm, wavelength_nm, n_ext, n_int;
    uint32_t gw, gh;
};

// ---------------------------------------------------------------- own intersector
drop::f3 F(v3 a) { return drop::f3{a.x, a.y, a.z}; }
__device__ inline drop::f3 Fd(v3 a) { return drop::f3{a.x, a.y, a.z}; }
__device__ inline v3 Vd(drop::f3 a) { return v3{a.x, a.y, a.z}; }

enum Code : uint32_t { kHit = 1, kMissed = 0, kFailed = 2 };

// Entry from outside along w through the point o of the plane through the
// origin (rho2 = |o|^2): the convex-body bracket of the histogram tracer.
__device__ inline uint32_t entry_hit(const drop::Shape& S, v3 o, v3 w, float rho2, v3& p, v3& n)
{
    using namespace drop;
    const float rmax2 = S.sphere ? 1.0f : S.rmax * S.rmax;
    if (!(rho2 < rmax2)) return kMissed;
    const f3 of = Fd(o), wf = Fd(w);
    float t;
    if (S.sphere) {
        t = -sqrtf(1.0f - rho2);
    } else {
        const float ta = -sqrtf(rmax2 - rho2);
        float tm = 0.0f, g0;
        const float fm = surface_slope(S, of, wf, g0);
        bool inside = fm < 0.0f;
        if (!inside) {
            float t0 = 0.0f, t1 = -g0 * sqrtf(rho2);
            for (int it = 0; it < 8; ++it) {
                t1 = t1 < ta ? ta : t1;
                t1 = t1 > -ta ? -ta : t1;
                float g1;
                const float f1 = surface_slope(S, axpy(wf, t1, of), wf, g1);
                if (f1 < 0.0f) {
                    tm = t1;
                    inside = true;
                    break;
                }
                if (g1 == g0) break;
                const float t2 = t1 - g1 * (t1 - t0) / (g1 - g0);
                t0 = t1;
                g0 = g1;
                if (fabsf(t2 - t1) <= 0x1p-20f) break;
                t1 = t2;
            }
        }
        if (!inside) return kMissed;
        const float rm = 1.0f + S.c[0];
        const float g2 = fmaf(rm, rm, -rho2);
        const float x0 = g2 > 0.0f ? -sqrtf(g2) : fmaf(tm - ta, 0.5f, ta);
        t = newton(S, of, wf, ta, tm, true, x0);
    }
    const f3 pf = axpy(wf, t, of);
    p = Vd(pf);
    n = Vd(outward_normal(S, pf));
    return kHit;
}

// Exit from inside: the chord from p along d to the surface.
__device__ inline uint32_t exit_hit(const drop::Shape& S, v3 p, v3 d, v3& q, v3& n)
{
    using namespace drop;
    const f3 pf = Fd(p), df = Fd(d);
    const float b = drop::dot(pf, df);
    float tx;
    if (S.sphere) {
        const float cc = drop::dot(pf, pf) - 1.0f;
        float disc = fmaf(b, b, -cc);
        if (disc < 0.0f) disc = 0.0f;
        tx = -b + sqrtf(disc);
    } else {
        const float cb = drop::dot(pf, pf) - S.rmax * S.rmax;
        float disc = fmaf(b, b, -cb);
        if (disc < 0.0f) disc = 0.0f;
        const float tb = -b + sqrtf(disc);
        const float ts = 0x1p-16f;
        if (!(ts < tb) || !(surface(S, axpy(df, ts, pf)) < 0.0f)) return kFailed;
        tx = newton(S, pf, df, ts, tb, false, -2.0f * b);
    }
    const f3 qf = axpy(df, tx, pf);
    q = Vd(qf);
    n = Vd(outward_normal(S, qf));
    return kHit;
}

__device__ inline void mark_from(Vertex* V, uint32_t n, uint32_t i, unsigned first, uint32_t status, uint32_t diag)
{
    for (unsigned f = first; f < 4; ++f) {
        V[f * n + i].status = status;
        V[f * n + i].diag = diag;
    }
}

__device__ inline void emit(Vertex& x, Field& second, const PRay& out, v3 point, Path path, const TraceParams& T,
                            uint32_t diag)
{
    const Pair after = Pair{T.ref_out, 0.0f} - pair_dot(point, out.dir);
    x.diag = diag;
    if (!finite(out.r.c[0]) || !finite(out.r.c[1])) {
        x.status = kBadInterface;
        return;
    }
    if (!add_distance(path, after, T.n_ext)) {
        x.status = kPhaseOverflow;
        return;
    }
    x.pos = point;
    x.dir = out.dir;
    x.bx = out.fr.e0;
    x.f = out.r.c[0];
    second = out.r.c[1];
    x.q = path.phase;
    x.status = kValid;
}

}  // namespace

// One thread per grid vertex: R at entry, T at each of the next three internal
// encounters (TT, TRT, TRRT), reflected branch continued, unpolarized response.
extern "C" __global__ void __launch_bounds__(kThreads)
    icephase_coh_trace(TraceParams T, Vertex* __restrict__ V, Field* __restrict__ F2)
{
    const uint32_t n = T.gw * T.gh;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (unsigned f = 0; f < 4; ++f) {
        const Vertex z{};
        V[f * n + i] = z;
        F2[f * n + i] = Field{cf{0, 0}, cf{0, 0}};
    }
    const uint32_t ix = i % T.gw, iy = i / T.gw;
    const float u = (2.0f * (float)ix / (float)(T.gw - 1) - 1.0f) * T.half;
    const float v = (2.0f * (float)iy / (float)(T.gh - 1) - 1.0f) * T.half;
    const v3 o = T.bx * u + T.e1 * v;
    v3 p, nrm;
    const uint32_t code = entry_hit(T.shape, o, T.w, u * u + v * v, p, nrm);
    uint32_t diag = 0;
    if (code != kHit) {
        mark_from(V, n, i, 0, code == kMissed ? kMiss : kUnresolved, diag);
        return;
    }
    Path path = make_path(T.radius_mm, T.wavelength_nm);
    const Pair before = Pair{T.ref_in, 0.0f} + pair_dot(p, T.w);
    if (!add_distance(path, before, T.n_ext)) {
        mark_from(V, n, i, 0, kPhaseOverflow, diag);
        return;
    }
    PRay active;
    active.dir = T.w;
    active.fr = Frame{T.bx, T.e1};
    active.r.c[0] = Field{cf{1.0f, 0.0f}, cf{0.0f, 0.0f}};
    active.r.c[1] = Field{cf{0.0f, 0.0f}, cf{1.0f, 0.0f}};
    const Branches b = interface_split(active, nrm, T.n_ext, T.n_int);
    diag |= b.diag;
    if (!b.ok) {
        mark_from(V, n, i, 0, kBadInterface, diag);
        return;
    }
    emit(V[i], F2[i], b.refl, p, path, T, diag);
    if (!b.has_t) {
        mark_from(V, n, i, 1, kTir, diag);
        return;
    }
    active = b.trans;
    v3 prev = p;
    for (unsigned fam = 1; fam < 4; ++fam) {
        v3 q, nq;
        const uint32_t hc = exit_hit(T.shape, prev, active.dir, q, nq);
        if (hc != kHit) {
            mark_from(V, n, i, fam, kUnresolved, diag);
            return;
        }
        if (!add_distance(path, pair_distance(prev, q), T.n_int)) {
            mark_from(V, n, i, fam, kPhaseOverflow, diag);
            return;
        }
        const Branches s = interface_split(active, -nq, T.n_int, T.n_ext);
        diag |= s.diag;
        if (!s.ok) {
            mark_from(V, n, i, fam, kBadInterface, diag);
            return;
        }
        if (s.has_t) {
            emit(V[fam * n + i], F2[fam * n + i], s.trans, q, path, T, diag);
        } else {
            V[fam * n + i].status = kTir;
            V[fam * n + i].diag = diag;
        }
        active = s.refl;
        prev = q;
    }
}

// Vertices that are neither valid, missed nor totally reflected, or valid with
// non-finite data, fail the record (the reference's trace validation).
extern "C" __global__ void __launch_bounds__(kThreads)
    icephase_coh_validate(const Vertex* __restrict__ V, const Field* __restrict__ F2, uint32_t count,
                          unsigned int* __restrict__ bad, unsigned int* __restrict__ first)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const Vertex& x = V[i];
    bool b = x.status != kValid && x.status != kMiss && x.status != kTir;
    if (x.status == kValid)
        b = b || !finite(x.pos) || !finite(x.dir) || !finite(x.bx) || !valid(x.q) || !finite(x.f) || !finite(F2[i]);
    if (b) {
        atomicAdd(bad, 1u);
        atomicMin(first, i);
    }
}

// Pass 1 of the patch build: the status of every logical cell (four families)
// and the per-status counts.
extern "C" __global__ void __launch_bounds__(kThreads)
    icephase_coh_classify(const Vertex* __restrict__ V, Layout L, uint32_t* __restrict__ flag,
                          unsigned int* __restrict__ counts)
{
    __shared__ unsigned int hist[kPatchStates];
    if (threadIdx.x < kPatchStates) hist[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t id = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id < L.cells) {
        const Patch P = make_patch(V, L, (uint32_t)id);
        flag[id] = has_geometry(P.status) ? 1u : 0u;
        atomicAdd(&hist[P.status], 1u);
    }
    __syncthreads();
    if (threadIdx.x < kPatchStates && hist[threadIdx.x]) atomicAdd(&counts[threadIdx.x], hist[threadIdx.x]);
}

// Pass 2: patches with geometry in patch_id order at their scanned offsets;
// the folded flag of each compact patch for the next scan.
extern "C" __global__ void __launch_bounds__(kThreads)
    icephase_coh_compact(const Vertex* __restrict__ V, Layout L, const uint32_t* __restrict__ flag,
                         const uint64_t* __restrict__ offset, Patch* __restrict__ out, uint32_t* __restrict__ fold_flag)
{
    const uint64_t id = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= L.cells || !flag[id]) return;
    const Patch P = make_patch(V, L, (uint32_t)id);
    const uint64_t k = offset[id];
    out[k] = P;
    fold_flag[k] = P.status == kNeedsRefinement ? 1u : 0u;
}

// Folded records of NeedsRefinement patches, one thread each; lookup[k] is the
// record slot of compact patch k (or ~0).
extern "C" __global__ void __launch_bounds__(64)
    icephase_coh_fold(const Vertex* __restrict__ V, uint32_t nv, const Patch* __restrict__ P, uint32_t np_,
                      const uint32_t* __restrict__ fold_flag, const uint64_t* __restrict__ slot, v3 inc,
                      FocalConfig focal, uint32_t with_focal, FoldConfig cfg, uint32_t* __restrict__ lookup,
                      Folded* __restrict__ rec)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= np_) return;
    if (!fold_flag[k]) {
        lookup[k] = 0xffffffffu;
        return;
    }
    const uint32_t s = (uint32_t)slot[k];
    lookup[k] = s;
    rec[s] = fold_prepare(V, nv, P[k], k, inc, with_focal ? &focal : nullptr, cfg);
}

// ---------------------------------------------------------------- query
struct QueryGrid {
    v3 axis, e0, e1;
    uint32_t nt, np;
};
struct Extent {
    uint32_t row0, rows, col0, cols;
};
// What the query reads of a compact patch: corner directions, padded box, id.
struct PatchGeom {
    v3 c[4];
    v3 lo, hi;
    uint32_t id, pad[3];
};

__device__ inline void patch_corners(const Vertex* V, const Patch& P, v3 (&c)[4])
{
    for (int i = 0; i < 4; ++i) c[i] = V[P.v[i]].dir;
}
// the reference's padded box of the four corners
__device__ inline void patch_box(const v3 (&c)[4], v3& lo, v3& hi)
{
    lo = c[0];
    hi = c[0];
    for (int i = 1; i < 4; ++i) {
        lo = v3{fminf(lo.x, c[i].x), fminf(lo.y, c[i].y), fminf(lo.z, c[i].z)};
        hi = v3{fmaxf(hi.x, c[i].x), fmaxf(hi.y, c[i].y), fmaxf(hi.z, c[i].z)};
    }
    const float pad = 0x1p-19f;
    lo = v3{nextafterf(lo.x - pad, -INFINITY), nextafterf(lo.y - pad, -INFINITY), nextafterf(lo.z - pad, -INFINITY)};
    hi = v3{nextafterf(hi.x + pad, INFINITY), nextafterf(hi.y + pad, INFINITY), nextafterf(hi.z + pad, INFINITY)};
}
// does the segment t w, t in [0, 4], meet the closed box? (FP64: conservative)
__device__ inline bool ray_meets_box(v3 w, v3 lo, v3 hi)
{
    double t0 = 0.0, t1 = 4.0;
    const double d[3] = {w.x, w.y, w.z}, l[3] = {lo.x, lo.y, lo.z}, h[3] = {hi.x, hi.y, hi.z};
    for (int k = 0; k < 3; ++k) {
        if (d[k] == 0.0) {
            if (l[k] > 0.0 || h[k] < 0.0) return false;
            continue;
        }
        double a = l[k] / d[k], b = h[k] / d[k];
        if (a > b) {
            const double t = a;
            a = b;
            b = t;
        }
        t0 = fmax(t0, a * (1 - 1e-12));
        t1 = fmin(t1, b * (1 + 1e-12));
    }
    return t0 <= t1;
}

// Candidate window of each compact patch: the query cells whose centre lies in
// the cone about the box centre that holds the patch's box (the largest angle
// to a box vertex bounds the angle to every point of the box), widened by
// 1e-6 rad for the rounding of the stored grid directions.
extern "C" __global__ void __launch_bounds__(kThreads)
    icephase_coh_extent(const Vertex* __restrict__ V, const Patch* __restrict__ P, uint32_t np_, QueryGrid G,
                        Extent* __restrict__ ext, uint32_t* __restrict__ count, PatchGeom* __restrict__ geom)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= np_) return;
    v3 c[4], lo, hi;
    const Patch Pk = P[k];
    patch_corners(V, Pk, c);
    patch_box(c, lo, hi);
    PatchGeom g;
    for (int i = 0; i < 4; ++i) g.c[i] = c[i];
    g.lo = lo;
    g.hi = hi;
    g.id = Pk.id;
    g.pad[0] = g.pad[1] = g.pad[2] = 0;
    geom[k] = g;
    const coh::d3 mid = coh::d3{0.5 * ((double)lo.x + hi.x), 0.5 * ((double)lo.y + hi.y), 0.5 * ((double)lo.z + hi.z)};
    const double ml = length(mid);
    Extent e{0, G.nt, 0, G.np};
    bool all = !(ml > 0.25);
    if (!all) {
        const coh::d3 cd = mid / ml;
        double rho = 0.0;
        for (int m = 0; m < 8; ++m) {
            const coh::d3 q = coh::d3{(m & 1) ? hi.x : lo.x, (m & 2) ? hi.y : lo.y, (m & 4) ? hi.z : lo.z};
            const double cq = dot(cd, q) / length(q);
            rho = fmax(rho, acos(fmin(1.0, fmax(-1.0, cq))));
        }
        rho += 1e-6;
        if (rho > 1.2) all = true;
        else {
            const coh::d3 ax = wide(G.axis), a0 = wide(G.e0), a1 = wide(G.e1);
            const double th = acos(fmin(1.0, fmax(-1.0, dot(cd, ax))));
            const double ph = atan2(dot(cd, a1), dot(cd, a0));
            const double rstep = coh::kPi / G.nt, cstep = 2.0 * coh::kPi / G.np;
            const long r0 = (long)ceil((th - rho) / rstep - 0.5), r1 = (long)floor((th + rho) / rstep - 0.5);
            const long rlo = r0 < 0 ? 0 : r0, rhi = r1 > (long)G.nt - 1 ? (long)G.nt - 1 : r1;
            e.row0 = (uint32_t)(rlo < (long)G.nt ? rlo : (long)G.nt - 1);
            e.rows = rhi >= rlo ? (uint32_t)(rhi - rlo + 1) : 0u;
            if (th - rho <= 0.0 || th + rho >= coh::kPi || sin(rho) >= sin(th)) {
                e.col0 = 0;
                e.cols = G.np;
            } else {
                const double half = asin(sin(rho) / sin(th)) + 1e-6;
                const long c0 = (long)ceil((ph - half + coh::kPi) / cstep - 0.5);
                const long c1 = (long)floor((ph + half + coh::kPi) / cstep - 0.5);
                if (c1 - c0 + 1 >= (long)G.np) {
                    e.col0 = 0;
                    e.cols = G.np;
                } else if (c1 < c0) {
                    e.col0 = 0;
                    e.cols = 0;
                } else {
                    e.col0 = (uint32_t)(((c0 % (long)G.np) + (long)G.np) % (long)G.np);
                    e.cols = (uint32_t)(c1 - c0 + 1);
                }
            }
        }
    }
    if (all) e = Extent{0, G.nt, 0, G.np};
    ext[k] = e;
    const uint64_t n = (uint64_t)e.rows * e.cols;
    count[k] = n > 0x7fffffffull ? 0x7fffffffu : (uint32_t)n;
}

struct QueryPass {
    const PatchGeom* G;
    const Extent* ext;
    const uint64_t* coff;   // candidate offsets, np_ + 1
    const v3* dirs;         // nt * np query directions
    uint32_t npatch, np;
    uint64_t total;
    unsigned int* hits;     // count pass: hits per direction
    unsigned int* notes;    // count pass: OR of query notes per direction
    unsigned int* first;    // count pass: lowest problem patch id per direction
    const uint64_t* hoff;   // fill pass: hit offsets per direction
    unsigned int* cursor;   // fill pass: per-direction fill counters
    QHit* out;              // fill pass: hit storage
    struct Staged* stage;   // count pass: hits in per-wave runs of slots, gaps marked (up to cap)
    unsigned long long* staged;
    uint64_t cap;
    uint32_t fill;
};

// A hit found by the count pass with its direction, kept so that the hits need
// not be recomputed to be stored; dir == kStageGap marks a slot left unused.
struct Staged {
    uint32_t dir;
    QHit hit;
};
constexpr uint32_t kStageGap = 0xffffffffu;
constexpr uint32_t kStageChunk = 256;  // staging slots a wave reserves at a time

extern "C" __global__ void __launch_bounds__(kThreads) icephase_coh_query(QueryPass Q)
{
    const uint32_t lane = threadIdx.x & 63;
    // each wave walks one contiguous run of candidates, 64 at a time: the patch of
    // its first candidate is searched once (by lane 0, broadcast), after which the
    // patch cursor only moves forward (from the last lane's patch of the previous
    // step), so no step repeats the search over all patches
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t per = ((Q.total + nwaves - 1) / nwaves + 63) & ~63ull;
    const uint64_t begin = wave * per;
    const uint64_t end = begin + per < Q.total ? begin + per : Q.total;
    if (begin >= end) return;  // the whole wave
    uint32_t base = 0;
    if (lane == 0) {
        uint32_t lo = 0, hi = Q.npatch;  // the patch whose candidate range holds begin
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (Q.coff[mid] <= begin)
                lo = mid;
            else
                hi = mid;
        }
        base = lo;
    }
    base = __shfl(base, 0, 64);
    // count pass: the wave stages its hits in a run of slots it reserved with one
    // atomic (kStageChunk or more at a time; one counter shared by every wave would
    // take an atomic on one address per step with a hit); the slots it leaves
    // unused are marked as gaps, which the scatter skips
    unsigned long long chunk = 0;
    uint32_t used = 0, size = 0;
    for (uint64_t wk = begin; wk < end; wk += 64) {
        const uint64_t k = wk + lane;
        uint32_t p = base;  // consecutive candidates: step forward to the patch of k
        if (k < end)
            while (p + 1 < Q.npatch && Q.coff[p + 1] <= k) ++p;
        // lane 63 holds the step's last patch; if its k is past the end, this is the last step
        base = __shfl(p, 63, 64);
        PHits h;
        h.count = 0;
        h.flags = 0;
        uint64_t di = 0;
        uint32_t pid = 0;
        if (k < end) {
            const Extent e = Q.ext[p];
            const uint32_t r = (uint32_t)(k - Q.coff[p]);  // below 2^31: extent counts are clamped there
            const uint32_t row = e.row0 + r / e.cols;
            uint32_t col = e.col0 + r % e.cols;
            if (col >= Q.np) col -= Q.np;
            di = (uint64_t)row * Q.np + col;
            const v3 w = Q.dirs[di];
            const PatchGeom& g = Q.G[p];
            if (ray_meets_box(w, g.lo, g.hi)) {
                v3 c[4];
                for (int i = 0; i < 4; ++i) c[i] = g.c[i];
                h = bilinear_intersect(c, w);
                pid = g.id;
            }
        }
        if (!Q.fill) {
            if (h.flags) atomicOr(&Q.notes[di], h.flags);
            if (h.flags & (kQueryUnresolved | kQueryInvalid)) atomicMin(&Q.first[di], pid);
            if (h.count) atomicAdd(&Q.hits[di], h.count);
            // the wave's inclusive prefix of hit counts
            uint32_t incl = h.count;
            for (uint32_t off = 1; off < 64; off <<= 1) {
                const uint32_t t = __shfl_up(incl, off, 64);
                if (lane >= off) incl += t;
            }
            const uint32_t wave_hits = __shfl(incl, 63, 64);
            if (wave_hits) {
                if (used + wave_hits > size) {
                    for (uint32_t z = used + lane; z < size; z += 64)
                        if (chunk + z < Q.cap) Q.stage[chunk + z].dir = kStageGap;
                    const uint32_t want = wave_hits > kStageChunk ? wave_hits : kStageChunk;
                    unsigned long long first = 0;
                    if (lane == 0) first = atomicAdd(Q.staged, (unsigned long long)want);
                    chunk = __shfl(first, 0, 64);
                    used = 0;
                    size = want;
                }
                const unsigned long long slot = chunk + used + (incl - h.count);
                for (uint32_t j = 0; j < h.count; ++j) {
                    if (slot + j >= Q.cap) break;  // over capacity: the host recomputes with the fill pass
                    Staged& z = Q.stage[slot + j];
                    z.dir = (uint32_t)di;
                    z.hit = QHit{p, pid, j, h.h[j].flags, h.h[j].u, h.h[j].v, h.h[j].t, h.h[j].res};
                }
                used += wave_hits;
            }
        } else {
            for (uint32_t j = 0; j < h.count; ++j) {
                const uint32_t at = atomicAdd(&Q.cursor[di], 1u);
                if (at >= Q.hits[di]) continue;  // more than counted: the finish pass flags the mismatch
                QHit q;
                q.compact = p;
                q.id = pid;
                q.root = j;
                q.flags = h.h[j].flags;
                q.u = h.h[j].u;
                q.v = h.h[j].v;
                q.t = h.h[j].t;
                q.res = h.h[j].res;
                Q.out[Q.hoff[di] + at] = q;
            }
        }
    }
    if (!Q.fill)
        for (uint32_t z = used + lane; z < size; z += 64)
            if (chunk + z < Q.cap) Q.stage[chunk + z].dir = kStageGap;
}

// Staged hits to their direction's segment (slots by per-direction counters;
// the finish pass restores the order).
extern "C" __global__ void __launch_bounds__(kThreads)
    icephase_coh_scatter(const Staged* __restrict__ stage, uint64_t n, const uint64_t* __restrict__ hoff,
                         const unsigned int* __restrict__ cnt, unsigned int* __restrict__ cursor, QHit* __restrict__ out)
{
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const Staged z = stage[k];
    if (z.dir == kStageGap) return;
    const uint32_t at = atomicAdd(&cursor[z.dir], 1u);
    if (at >= cnt[z.dir]) return;  // the finish pass flags the mismatch
    out[hoff[z.dir] + at] = z.hit;
}

struct Summary {
    uint32_t hits, flags, first, regular, refinement, boundary, candidates, pad;
};

__device__ inline bool hit_before(const QHit& a, const QHit& b)
{
    if (a.id != b.id) return a.id < b.id;
    if (a.compact != b.compact) return a.compact < b.compact;
    return a.root < b.root;
}

// Per direction: order the hits, drop repeated (patch, root) reports and summarize.
extern "C" __global__ void __launch_bounds__(kThreads)
    icephase_coh_finish(QHit* __restrict__ H, const uint64_t* __restrict__ hoff, const unsigned int* __restrict__ cnt,
                        const unsigned int* __restrict__ cursor, const unsigned int* __restrict__ notes,
                        const unsigned int* __restrict__ first, const Patch* __restrict__ P, uint32_t npatch,
                        uint64_t ndir, Summary* __restrict__ S)
{
    const uint64_t d = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= ndir) return;
    Summary s{0, 0, first[d], 0, 0, 0, cnt[d], 0};
    const uint32_t nn = notes[d];
    if (nn & kQueryUnresolved) s.flags |= kQUnresolved;
    if (nn & kQueryInvalid) s.flags |= kQInvalid;
    if (nn & kQueryFp64) s.flags |= kQFp64;
    if (cursor[d] != cnt[d]) s.flags |= kQCountMismatch;
    QHit* h = H + hoff[d];
    const uint32_t n = cursor[d] < cnt[d] ? cursor[d] : cnt[d];
    for (uint32_t i = 1; i < n; ++i) {
        const QHit x = h[i];
        uint32_t j = i;
        while (j && hit_before(x, h[j - 1])) {
            h[j] = h[j - 1];
            --j;
        }
        h[j] = x;
    }
    uint32_t kept = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const QHit x = h[i];
        if (kept && x.compact == h[kept - 1].compact && x.root == h[kept - 1].root) {
            s.flags |= kQDuplicate;
            continue;
        }
        if (x.compact >= npatch) {
            s.flags |= kQInvalid;
            continue;
        }
        h[kept++] = x;
        const uint32_t st = P[x.compact].status;
        if (st == kNeedsRefinement) {
            ++s.refinement;
            s.flags |= kQRefinement;
        } else if (st == kRegularPos || st == kRegularNeg) {
            ++s.regular;
        } else {
            s.flags |= kQInvalid;
        }
        if (x.flags & kHitBoundary) {
            ++s.boundary;
            s.flags |= kQBoundary;
        }
        if (x.flags & kHitSingular) s.flags |= kQSingular;
    }
    s.hits = kept;
    S[d] = s;
}

// ---------------------------------------------------------------- optics
struct OpticsPass {
    const Vertex* V;
    const Field* F2;
    uint32_t nv;
    const Patch* P;
    uint32_t npatch;
    const uint32_t* lookup;
    const Folded* rec;
    uint32_t nrec;
    const v3* dirs;
    const uint64_t* hoff;
    const QHit* H;
    const Summary* S;
    v3 inc, bx;
    FocalConfig focal;
    uint64_t ndir;
    Optics* out;
};

__device__ inline void nan_optics(Optics& r)
{
    const double nan = __builtin_nan("");
    r.inc_s = r.inc_p = r.path_s = r.path_p = nan;
    r.path_field[0] = r.path_field[1] = F64{nan, nan, nan, nan};
}
__device__ inline void nan_focal(Optics& r)
{
    const double nan = __builtin_nan("");
    r.focal_s = r.focal_p = nan;
    r.focal_field[0] = r.focal_field[1] = F64{nan, nan, nan, nan};
}

// One thread per query direction: every hit in order, compensated sums of the
// two response columns (unpolarized input: intensities average the columns).
extern "C" __global__ void __launch_bounds__(128) icephase_coh_optics(OpticsPass O)
{
    const uint64_t d = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= O.ndir) return;
    Optics r{};
    r.first_problem = 0xffffffffu;
    r.focal_flags = kFocalInputError;
    const Summary s = O.S[d];
    r.hits = s.hits;
    r.qflags = s.flags;
    r.regular_hits = s.regular;
    r.refinement_hits = s.refinement;
    if (s.flags & kQueryErrorMask) {
        r.flags = kOptQueryError;
        r.first_problem = s.first;
        r.rejected = s.hits;
        nan_optics(r);
        nan_focal(r);
        O.out[d] = r;
        return;
    }
    const bool folded_on = O.lookup && O.rec;
    if ((s.flags & kQRefinement) && !folded_on) r.flags |= kOptRefinePending;
    if (s.flags & kQBoundary) r.flags |= kOptBoundaryPending;
    if (s.flags & kQSingular) r.flags |= kOptSingularPending;
    const v3 w = O.dirs[d];
    Frame tf;
    if (!scattering_frame(O.inc, O.bx, w, tf)) {
        r.flags |= kOptFrame;
        r.rejected = s.hits;
        nan_optics(r);
        nan_focal(r);
        O.out[d] = r;
        return;
    }
    Neum sr[2] = {}, si[2] = {}, pr[2] = {}, pi[2] = {}, is{}, ip{};
    Neum fsr[2] = {}, fsi[2] = {}, fpr[2] = {}, fpi[2] = {}, env[4] = {};
    uint32_t focal_flags = focal_layout_valid(O.focal) ? 0u : (uint32_t)kFocalGeometry;
    uint32_t focal_first = 0xffffffffu, recorded_nonregular = 0;
    const QHit* H = O.H + O.hoff[d];
    for (uint32_t i = 0; i < s.hits; ++i) {
        const QHit h = H[i];
        uint32_t flags = 0;
        Contribution c;
        const Branch* used = nullptr;
        if (i != 0u && !hit_before(H[i - 1], h))
            flags = kOptInvalid;
        else if (h.compact >= O.npatch)
            flags = kOptInvalid;
        else {
            const Patch P = O.P[h.compact];
            recorded_nonregular += P.status == kNeedsRefinement;
            const Folded* fr = nullptr;
            if (folded_on) {
                const uint32_t slot = O.lookup[h.compact];
                if (slot < O.nrec && O.rec[slot].compact == h.compact) fr = &O.rec[slot];
            }
            flags = evaluate_hit(O.V, O.F2, O.nv, P, h, w, tf, c, fr, &used);
        }
        if (flags) {
            r.flags |= flags;
            ++r.rejected;
            if (h.id < r.first_problem) r.first_problem = h.id;
            continue;
        }
        ++r.evaluated;
        r.folded_hits += used != nullptr;
        for (int j = 0; j < 2; ++j) {
            sr[j].add(c.field[j].sr);
            si[j].add(c.field[j].si);
            pr[j].add(c.field[j].pr);
            pi[j].add(c.field[j].pi);
        }
        is.add(c.inc_s);
        ip.add(c.inc_p);
        // focal quarter turns: the branch's (folded) or the cell centre's
        FocalEstimate e{0, 0, 0, 0};
        uint32_t code = kFocalNone;
        const Patch P = O.P[h.compact];
        if (used) {
            code = used->focal;
            if (focal_layout_valid(O.focal)) {
                const uint32_t cells = (O.focal.w - 1u) * (O.focal.h - 1u);
                e.family = h.id / cells;
                if (e.family >= 4) code |= kFocalGeometry;
                e.qturns = used->qturns;
                e.extra = used->extra;
            } else {
                code |= kFocalGeometry;
            }
        } else {
            code = focal_estimate(O.V, O.nv, P, O.inc, O.focal, 0.5, 0.5, true, e);
        }
        if (code != kFocalNone) {
            focal_flags |= code;
            if (h.id < focal_first) focal_first = h.id;
        } else {
            for (int j = 0; j < 2; ++j) {
                const F64 f = quarter_turns(c.field[j], e.qturns);
                fsr[j].add(f.sr);
                fsi[j].add(f.si);
                fpr[j].add(f.pr);
                fpi[j].add(f.pi);
            }
            env[e.family].add(c.inc_s + c.inc_p);
            ++r.fam_hits[e.family];
            ++r.corrected;
            r.extra += e.extra;
        }
    }
    if (folded_on && (s.flags & kQRefinement)) {
        if (recorded_nonregular == 0u) r.flags |= kOptRefinePending;
        if (recorded_nonregular != s.refinement) r.flags |= kOptInvalid;
    }
    for (int j = 0; j < 2; ++j) r.path_field[j] = F64{sr[j].value(), si[j].value(), pr[j].value(), pi[j].value()};
    r.inc_s = is.value();
    r.inc_p = ip.value();
    r.path_s = 0.0;
    r.path_p = 0.0;
    for (int j = 0; j < 2; ++j) {
        r.path_s += 0.5 * fma(sr[j].value(), sr[j].value(), si[j].value() * si[j].value());
        r.path_p += 0.5 * fma(pr[j].value(), pr[j].value(), pi[j].value() * pi[j].value());
    }
    if (!finite(r.path_field[0]) || !finite(r.path_field[1]) || !finite(r.inc_s) || !finite(r.inc_p) ||
        !finite(r.path_s) || !finite(r.path_p))
        r.flags |= kOptArith;
    if (r.flags & kOptErrorMask) nan_optics(r);
    if (r.flags & kOptErrorMask) focal_flags |= kFocalInputError;
    if (r.flags & kOptPendingMask) focal_flags |= kFocalInputPending;
    if (r.first_problem < focal_first) focal_first = r.first_problem;
    for (int j = 0; j < 2; ++j) r.focal_field[j] = F64{fsr[j].value(), fsi[j].value(), fpr[j].value(), fpi[j].value()};
    r.focal_s = 0.0;
    r.focal_p = 0.0;
    for (int j = 0; j < 2; ++j) {
        const F64& f = r.focal_field[j];
        r.focal_s += 0.5 * fma(f.sr, f.sr, f.si * f.si);
        r.focal_p += 0.5 * fma(f.pr, f.pr, f.pi * f.pi);
    }
    for (int f = 0; f < 4; ++f) r.fam_inc[f] = env[f].value();
    if (!finite(r.focal_field[0]) || !finite(r.focal_field[1]) || !finite(r.focal_s + r.focal_p))
        focal_flags |= kFocalArithmetic;
    r.focal_flags = focal_flags;
    if (focal_flags) {
        nan_focal(r);
        r.first_problem = r.first_problem < focal_first ? r.first_problem : focal_first;
    }
    O.out[d] = r;
}

// ---------------------------------------------------------------- Table II diffraction
// Rainbow transitions between a row and the next at fixed phi.
extern "C" __global__ void __launch_bounds__(kThreads)
    icephase_coh_detect(const Optics* __restrict__ R, DiffConfig C, uint32_t* __restrict__ trans)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)C.nt * C.np) return;
    uint32_t f = 0;
    if (!diff_valid(C))
        f = kTransUnknown;
    else if (i / C.np + 1u != C.nt) {
        const Optics& a = R[i];
        const Optics& b = R[i + C.np];
        if (!focal_sample_valid(a) || !focal_sample_valid(b))
            f = kTransUnknown;
        else {
            const double x = a.focal_s + a.focal_p, y = b.focal_s + b.focal_p;
            const double peak = fmax(x, y);
            if (peak > 0 && finite(peak) && !(fabs(y - x) / peak < C.contrast)) {
                if (a.fam_hits[2] == 0 && b.fam_hits[2] > 0 && b.fam_inc[2] > 0) f |= kTransPrimary;
                if (a.fam_hits[3] > 0 && b.fam_hits[3] == 0 && a.fam_inc[3] > 0) f |= kTransSecondary;
            }
        }
    }
    trans[i] = f;
}

__device__ inline void nan_diff(Diffracted& o)
{
    o.s = o.p = __builtin_nan("");
}

// Local spherical Gaussian of the nearest transition's bandwidth on the total
// focal intensity, tapered between the inner blend and the support radius.
// Rows holding at least one transition or unknown flag, counted per row.
extern "C" __global__ void __launch_bounds__(kThreads)
    icephase_coh_rowflags(const uint32_t* __restrict__ trans, uint32_t nt, uint32_t np, uint32_t* __restrict__ rows)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)nt * np || trans[i] == 0u) return;
    atomicAdd(&rows[i / np], 1u);
}

extern "C" __global__ void __launch_bounds__(128)
    icephase_coh_filter(const Optics* __restrict__ R, const uint32_t* __restrict__ trans,
                        const uint64_t* __restrict__ rowoff, DiffConfig C, Diffracted* __restrict__ out)
{
    const uint64_t index = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (index >= (uint64_t)C.nt * C.np) return;
    Diffracted o{0, 0, 0, 0, 0, 0, trans[index], 0};
    if (!diff_valid(C)) {
        o.flags = kDiffInvalid;
        nan_diff(o);
        out[index] = o;
        return;
    }
    const Optics& centre = R[index];
    if (!focal_sample_valid(centre)) {
        o.flags = kDiffInputUnavailable;
        nan_diff(o);
        out[index] = o;
        return;
    }
    o.s = centre.focal_s;
    o.p = centre.focal_p;
    const int nr = (int)C.nt, nc = (int)C.np, row = (int)(index / C.np), col = (int)(index % C.np);
    const double dt = coh::kPi / (double)nr, dp = 2 * coh::kPi / (double)nc, t = (row + 0.5) * dt;
    const double support = C.support;
    const double search = 2 * C.sigma * support;
    const int lo = imax(0, (int)floor((t - search) / dt) - 1);
    const int hi = imin(nr - 2, (int)ceil((t + search) / dt));
    // no flag anywhere in the searched rows: the search would find nothing
    if (hi < lo || rowoff[hi + 1] == rowoff[lo]) {
        out[index] = o;
        return;
    }
    double best = support + 1;
    for (int rr = lo; rr <= hi; ++rr) {
        const double tr = (rr + 1.0) * dt;
        const int width = phi_radius(t, tr, search, dp, nc);
        if (width < 0) continue;
        const int count = imin(nc, 2 * width + 1), start = count == nc ? 0 : col - width;
        for (int j = 0; j < count; ++j) {
            const int cc = iwrap(start + j, nc);
            const double dist = angular_distance(t, tr, (cc - col) * dp);
            if (dist > search + kAngTol) continue;
            const uint32_t flag = trans[(uint64_t)rr * C.np + (uint32_t)cc];
            if (flag & kTransUnknown) o.flags |= kDiffDetectionUnavailable;
            for (uint32_t kind = kTransPrimary; kind <= kTransSecondary; kind *= 2u) {
                if (!(flag & kind)) continue;
                const double sigma = C.sigma * (kind == kTransSecondary ? 2 : 1);
                const double dn = dist / sigma;
                if (dn < support && (dn < best - 1e-12 || (fabs(dn - best) <= 1e-12 && kind < o.kind))) {
                    best = dn;
                    o.sigma = sigma;
                    o.kind = kind;
                }
            }
        }
    }
    if (o.flags & kDiffDetectionUnavailable) {
        nan_diff(o);
        out[index] = o;
        return;
    }
    if (o.kind == 0) {
        out[index] = o;
        return;
    }
    const double sigma = o.sigma;
    if (dt > 0.5 * sigma || sin(t) * dp > 0.5 * sigma) o.flags |= kDiffUnderresolved;
    o.blend = 1;
    if (best > C.inner) {
        const double x = (best - C.inner) / (support - C.inner);
        o.blend = 1 - x * x * (3 - 2 * x);
    }
    const double radius = support * sigma;
    const int lower = imax(0, (int)floor((t - radius) / dt) - 1);
    const int upper = imin(nr - 1, (int)ceil((t + radius) / dt));
    Neum ss{}, sp{}, ws{};
    for (int rr = lower; rr <= upper; ++rr) {
        const double tr = (rr + 0.5) * dt;
        const int width = phi_radius(t, tr, radius, dp, nc);
        if (width < 0) continue;
        const int count = imin(nc, 2 * width + 1), start = count == nc ? 0 : col - width;
        const double area = 2 * sin(tr) * sin(0.5 * dt) * dp;
        for (int j = 0; j < count; ++j) {
            const int cc = iwrap(start + j, nc);
            const double dist = angular_distance(t, tr, (cc - col) * dp);
            if (dist > radius + kAngTol) continue;
            const Optics& smp = R[(uint64_t)rr * C.np + (uint32_t)cc];
            if (!focal_sample_valid(smp)) {
                o.flags |= kDiffStencilUnavailable;
                continue;
            }
            const double wgt = exp(-0.5 * (dist / sigma) * (dist / sigma)) * area;
            ws.add(wgt);
            ss.add(wgt * smp.focal_s);
            sp.add(wgt * smp.focal_p);
        }
    }
    if (o.flags & kDiffStencilUnavailable) {
        nan_diff(o);
        out[index] = o;
        return;
    }
    const double wsum = ws.value();
    if (!(wsum > 0) || !finite(wsum)) {
        o.flags |= kDiffInvalid;
        nan_diff(o);
        out[index] = o;
        return;
    }
    o.s = fma(o.blend, ss.value() / wsum - centre.focal_s, centre.focal_s);
    o.p = fma(o.blend, sp.value() / wsum - centre.focal_p, centre.focal_p);
    if (!(o.s >= 0 && o.p >= 0) || !finite(o.s + o.p)) {
        o.flags |= kDiffInvalid;
        nan_diff(o);
        out[index] = o;
        return;
    }
    o.flags |= kDiffFiltered;
    out[index] = o;
}

// ---------------------------------------------------------------- density and CDF input
// Density of a stage per direction and the gate's counters: [0] invalid,
// [1] incomplete, [2] underresolved, [3] first problem; max density as u64 bits.
extern "C" __global__ void __launch_bounds__(kThreads)
    icephase_coh_density(const Optics* __restrict__ R, const Diffracted* __restrict__ D, uint32_t stage, uint64_t n,
                         double* __restrict__ out, unsigned int* __restrict__ gate,
                         unsigned long long* __restrict__ maxbits)
{
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const Optics& r = R[k];
    bool incomplete = !((r.flags & (kOptErrorMask | kOptPendingMask)) == 0u && r.rejected == 0u &&
                        r.evaluated == r.hits);
    bool under = false;
    double d;
    if (stage == 1)
        d = r.inc_s + r.inc_p;
    else if (stage == 2)
        d = r.path_s + r.path_p;
    else {
        incomplete = incomplete || r.focal_flags != 0u;
        if (stage == 3)
            d = r.focal_s + r.focal_p;
        else {
            incomplete = incomplete || (D[k].flags & kDiffUnavailableMask) != 0u;
            under = (D[k].flags & kDiffUnderresolved) != 0u;
            d = D[k].s + D[k].p;
        }
    }
    const bool bad = !(d == d) || !finite(d) || d < 0;
    out[k] = d;
    if (bad) atomicAdd(&gate[0], 1u);
    if (incomplete) atomicAdd(&gate[1], 1u);
    if (under) atomicAdd(&gate[2], 1u);
    if (bad || incomplete || under) atomicMin(&gate[3], (unsigned int)k);
    if (!bad) atomicMax(maxbits, (unsigned long long)__double_as_longlong(d));  // d >= 0: order of the bits
}

// w = (d / max) du of the query row (theta-major), the reference's CDF weight.
extern "C" __global__ void __launch_bounds__(kThreads)
    icephase_coh_weights(const double* __restrict__ d, const double* __restrict__ edges, uint32_t np,
                         uint64_t n, const unsigned long long* __restrict__ maxbits, double* __restrict__ w)
{
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double mx = __longlong_as_double((long long)*maxbits);
    const uint64_t i = k / np;
    w[k] = (d[k] / mx) * (edges[i + 1] - edges[i]);
}

// ---------------------------------------------------------------- scans
// Exclusive prefix sums of u32 counts into u64 offsets (n + 1 entries): per
// block sums, one block over the sums, then per-block apply.
extern "C" __global__ void __launch_bounds__(kThreads)
    icephase_scan_blocks(const uint32_t* __restrict__ in, uint64_t n, uint64_t* __restrict__ sums)
{
    __shared__ uint64_t part[kThreads / 64];
    const uint64_t base = (uint64_t)blockIdx.x * kThreads * 4;
    uint64_t s = 0;
    for (int j = 0; j < 4; ++j) {
        const uint64_t k = base + (uint64_t)j * kThreads + threadIdx.x;
        if (k < n) s += in[k];
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t t = 0;
        for (int w = 0; w < kThreads / 64; ++w) t += part[w];
        sums[blockIdx.x] = t;
    }
}
extern "C" __global__ void __launch_bounds__(64) icephase_scan_sums(uint64_t* __restrict__ sums, uint32_t nb)
{
    // one wave: lane l owns a contiguous chunk of the block sums
    const uint32_t l = threadIdx.x, chunk = (nb + 63) / 64;
    const uint32_t b0 = l * chunk < nb ? l * chunk : nb, b1 = b0 + chunk < nb ? b0 + chunk : nb;
    uint64_t own = 0;
    for (uint32_t b = b0; b < b1; ++b) own += sums[b];
    uint64_t incl = own;
    for (int off = 1; off < 64; off <<= 1) {
        const uint64_t up = __shfl_up(incl, off, 64);
        if ((int)l >= off) incl += up;
    }
    uint64_t run = incl - own;
    for (uint32_t b = b0; b < b1; ++b) {
        const uint64_t v = sums[b];
        sums[b] = run;
        run += v;
    }
    if (l == 63) sums[nb] = incl;
}
extern "C" __global__ void __launch_bounds__(kThreads)
    icephase_scan_apply(const uint32_t* __restrict__ in, uint64_t n, const uint64_t* __restrict__ sums,
                        uint64_t* __restrict__ out)
{
    __shared__ uint64_t lane_tot[kThreads];
    const uint64_t base = (uint64_t)blockIdx.x * kThreads * 4;
    // thread t owns items base + 4t .. 4t+3 (contiguous)
    uint32_t v[4];
    uint64_t s = 0;
    for (int j = 0; j < 4; ++j) {
        const uint64_t k = base + 4ull * threadIdx.x + j;
        v[j] = k < n ? in[k] : 0u;
        s += v[j];
    }
    lane_tot[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < kThreads; off <<= 1) {
        const uint64_t add = threadIdx.x >= (unsigned)off ? lane_tot[threadIdx.x - off] : 0ull;
        __syncthreads();
        lane_tot[threadIdx.x] += add;
        __syncthreads();
    }
    uint64_t run = sums[blockIdx.x] + lane_tot[threadIdx.x] - s;
    for (int j = 0; j < 4; ++j) {
        const uint64_t k = base + 4ull * threadIdx.x + j;
        if (k < n) out[k] = run;
        run += v[j];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == kThreads - 1) out[n] = sums[gridDim.x];
}

extern "C" __global__ void __launch_bounds__(kThreads) icephase_coh_fill_u32(uint32_t* p, uint64_t n, uint32_t v)
{
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (uint64_t)gridDim.x * blockDim.x)
        p[k] = v;
}

// ================================================================ host object
struct iph_coherent {
    iph_ctx* ctx = nullptr;
    iph_coherent_config cfg{};
    TraceParams tp{};
    Layout L{};
    FocalConfig focal{};
    FoldConfig fold{1e-8, 2048, 20};
    DiffConfig diff{};
    QueryGrid grid{};
    uint32_t nv = 0;  // 4 * grid^2
    bool traced = false, evaluated = false;
    // device arrays
    DevBuf<Vertex> V;
    DevBuf<Field> F2;
    DevBuf<v3> dirs;
    DevBuf<uint32_t> flag, fold_flag, lookup, cand_count;
    DevBuf<uint64_t> offsets, fold_slot, coff, hoff, sums;
    DevBuf<Patch> P;
    DevBuf<Folded> rec;
    DevBuf<Extent> ext;
    DevBuf<PatchGeom> geom;
    DevBuf<unsigned int> qcnt, qnotes, qfirst, qcursor, counters;
    DevBuf<QHit> H;
    DevBuf<Staged> stage;
    DevBuf<unsigned long long> staged;
    DevBuf<Summary> S;
    DevBuf<Optics> R;
    DevBuf<uint32_t> trans, rowcnt;
    DevBuf<uint64_t> rowoff;
    DevBuf<Diffracted> D;
    DevBuf<double> density, weights, edges;
    DevBuf<unsigned long long> maxbits;
    // host results
    iph_coherent_stats stats{};
    uint32_t npatch = 0, nfold = 0;
    uint64_t candidates = 0, nhits = 0;
    std::vector<uint32_t> status_counts;
};

namespace {

uint32_t blocks_for(uint64_t n, uint32_t t) { return (uint32_t)((n + t - 1) / t); }

// Reallocate only when the buffer is smaller than n (records of one size reuse their buffers).
template <class T>
hipError_t grow(DevBuf<T>& b, size_t n)
{
    if (b.p && b.n >= n) return hipSuccess;
    return b.alloc(n);
}

// exclusive scan of n u32 values into out[0..n] (u64), on the context stream
hipError_t scan(iph_coherent* h, const uint32_t* in, uint64_t n, uint64_t* out)
{
    const uint64_t per = (uint64_t)kThreads * 4;
    const uint32_t nb = (uint32_t)((n + per - 1) / per);
    const uint32_t nbx = nb ? nb : 1;
    const hipError_t ge = grow(h->sums, (size_t)nbx + 1);
    if (ge != hipSuccess) return ge;
    hipStream_t s = h->ctx->stream;
    if (n == 0) {
        const uint64_t z = 0;
        return hipMemcpyAsync(out, &z, sizeof z, hipMemcpyHostToDevice, s);
    }
    hipLaunchKernelGGL(icephase_scan_blocks, dim3(nb), dim3(kThreads), 0, s, in, n, h->sums.p);
    hipLaunchKernelGGL(icephase_scan_sums, dim3(1), dim3(64), 0, s, h->sums.p, nb);
    hipLaunchKernelGGL(icephase_scan_apply, dim3(nb), dim3(kThreads), 0, s, in, n, h->sums.p, out);
    return hipGetLastError();
}

hipError_t fill_u32(iph_coherent* h, uint32_t* p, uint64_t n, uint32_t v)
{
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(icephase_coh_fill_u32, dim3(std::min<uint64_t>(blocks_for(n, kThreads), 4096)), dim3(kThreads),
                       0, h->ctx->stream, p, n, v);
    return hipGetLastError();
}

struct Timer {
    hipEvent_t a = nullptr, b = nullptr;
    hipStream_t s;
    explicit Timer(hipStream_t st) : s(st)
    {
        (void)hipEventCreate(&a);
        (void)hipEventCreate(&b);
        (void)hipEventRecord(a, s);
    }
    double stop()
    {
        float ms = 0;
        (void)hipEventRecord(b, s);
        (void)hipEventSynchronize(b);
        (void)hipEventElapsedTime(&ms, a, b);
        return ms;
    }
    ~Timer()
    {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

// Table II of Sadeghi et al. (2012): primary bandwidth (degrees) against drop
// radius (mm), linearly interpolated (include/rainbow/rainbow_diffraction.hpp:17-29).
bool table_sigma_degrees(float radius, double& out)
{
    static const float r[10] = {0.1f, 0.2f, 0.3f, 0.4f, 0.5f, 0.6f, 0.7f, 0.8f, 0.9f, 1.0f};
    static const double s[10] = {0.70, 0.45, 0.30, 0.25, 0.22, 0.20, 0.18, 0.17, 0.16, 0.15};
    if (!(radius >= r[0] && radius <= r[9])) return false;
    for (unsigned j = 0; j < 10; ++j) {
        if (radius == r[j]) {
            out = s[j];
            return true;
        }
        if (j && radius < r[j]) {
            out = fma(((double)radius - r[j - 1]) / ((double)r[j] - r[j - 1]), s[j] - s[j - 1], s[j - 1]);
            return true;
        }
    }
    return false;
}

// The query grid's directions: cell centres in (theta, phi) about the incident
// axis, evaluated in FP64 and rounded (include/rainbow/query_direction_grid.hpp:34-66).
std::vector<v3> grid_directions(const QueryGrid& G)
{
    std::vector<v3> out((size_t)G.nt * G.np);
    for (uint32_t row = 0; row < G.nt; ++row) {
        const double t = coh::kPi * ((double)row + 0.5) / (double)G.nt, st = std::sin(t), ct = std::cos(t);
        for (uint32_t col = 0; col < G.np; ++col) {
            const double p = -coh::kPi + 2.0 * coh::kPi * ((double)col + 0.5) / (double)G.np;
            const double x = st * std::cos(p), y = st * std::sin(p);
            const v3 w{(float)((double)G.axis.x * ct + (double)G.e0.x * x + (double)G.e1.x * y),
                       (float)((double)G.axis.y * ct + (double)G.e0.y * x + (double)G.e1.y * y),
                       (float)((double)G.axis.z * ct + (double)G.e0.z * x + (double)G.e1.z * y)};
            out[(size_t)row * G.np + col] = normalized(w);
        }
    }
    return out;
}

int invalid(iph_ctx* ctx, const std::string& m) { return fail(ctx, IPH_E_INVALID, m); }

// The reference's conservative outer radius, FP32 and rounded outward at each
// step (include/rainbow/raindrop_shape.hpp:45-64): the incident grid spans
// 1.01 of it and both reference planes lie at twice it.
float reference_outer_radius(const float c[8])
{
    float sum = 0.0f;
    for (int k = 1; k < 8; ++k) sum = std::nextafter(sum + std::fabs(c[k]), INFINITY);
    const float centre = 1.0f + c[0];
    return std::nextafter(centre + sum + 0x1p-20f, INFINITY);
}

}  // namespace

extern "C" {

int iph_coherent_create(iph_ctx* ctx, const iph_coherent_config* c, iph_coherent** out)
{
    if (!ctx || !c || !out) return fail(ctx, IPH_E_INVALID, "ctx, config and out must not be NULL.");
    *out = nullptr;
    if (!(c->radius_mm > 0.0f && c->radius_mm <= 3.0f))
        return invalid(ctx, "radius_mm must be in (0,3]; no coefficient extrapolation is implemented.");
    if (!(c->wavelength_nm >= 380.0f && c->wavelength_nm <= 830.0f))
        return invalid(ctx, "wavelength_nm must be in [380,830] for this implementation.");
    if (!(c->interior_index >= 1.0f && c->interior_index <= 2.0f))
        return invalid(ctx, "interior_index must be in [1,2]; exterior_index is 1.");
    if (c->grid < 2 || (uint64_t)c->grid * c->grid > 0xffffffffull / 4)
        return invalid(ctx, "Invalid grid dimensions or 32-bit path-major index overflow.");
    if (!c->theta_count || !c->phi_count || (uint64_t)c->theta_count * c->phi_count > 0xffffffffull)
        return invalid(ctx, "Invalid query grid dimensions.");
    if (!std::isfinite(c->inclination_deg)) return invalid(ctx, "inclination_deg must be finite.");
    for (int j = 0; j < 4; ++j)
        if (c->focal_offsets[j] > 3) return invalid(ctx, "Focal offsets must be 0..3.");
    auto* h = new iph_coherent;
    h->ctx = ctx;
    h->cfg = *c;
    // shape (Table I): the intersector's bounds, and the reference's FP32 outer
    // radius, which fixes the grid extent and the reference planes
    float coef[8];
    if (iph_drop_shape(c->radius_mm, c->sphere ? 1 : 0, coef) != IPH_OK) {
        delete h;
        return invalid(ctx, "radius_mm must be in (0,3]; no coefficient extrapolation is implemented.");
    }
    if (c->sphere)
        for (int m = 0; m < 8; ++m) coef[m] = 0.0f;
    bool round = true;
    for (int m = 0; m < 8; ++m) round = round && coef[m] == 0.0f;
    std::string err;
    if (!icephase::drop_shape_bounds(coef, round, h->tp.shape, err)) {
        delete h;
        return invalid(ctx, err);
    }
    // incident frame (the reference's settings: w = (cos a, -sin a, 0), e0 from the +z / +x axis)
    const double a = c->inclination_deg * coh::kPi / 180.0;
    const v3 w0{(float)std::cos(a), (float)-std::sin(a), 0.0f};
    const v3 w = normalized(w0);
    const v3 axis = std::fabs(w.z) < 0.9f ? v3{0, 0, 1} : v3{1, 0, 0};
    const v3 bx = normalized(axis - w * dot(w, axis));
    const v3 e1 = normalized(cross(w, bx));
    const float outer = reference_outer_radius(coef);
    h->tp.w = w;
    h->tp.bx = bx;
    h->tp.e1 = e1;
    h->tp.half = outer * 1.01f;
    h->tp.ref_in = h->tp.ref_out = 2.0f * outer;
    h->tp.radius_mm = c->radius_mm;
    h->tp.wavelength_nm = c->wavelength_nm;
    h->tp.n_ext = 1.0f;
    h->tp.n_int = c->interior_index;
    h->tp.gw = h->tp.gh = c->grid;
    const uint32_t gw = c->grid;
    const float du = 2.0f * h->tp.half / (float)(gw - 1);
    h->L = Layout{gw, gw, gw * gw, (gw - 1) * (gw - 1), 4 * (gw - 1) * (gw - 1), du * du};
    h->nv = 4 * gw * gw;
    h->focal = FocalConfig{gw, gw, h->tp.half, 0, {c->focal_offsets[0], c->focal_offsets[1], c->focal_offsets[2],
                                                   c->focal_offsets[3]}};
    double sigma = c->primary_sigma_deg;
    h->diff = DiffConfig{0, 4.0, c->transition_contrast, 3.0, c->theta_count, c->phi_count};
    if (sigma > 0 || table_sigma_degrees(c->radius_mm, sigma)) h->diff.sigma = sigma * coh::kPi / 180.0;
    h->grid = QueryGrid{w, bx, e1, c->theta_count, c->phi_count};
    const hipError_t e0 = hipSetDevice(ctx->device);
    if (e0 != hipSuccess) {
        delete h;
        return fail(ctx, IPH_E_DEVICE, hipGetErrorString(e0));
    }
    const uint64_t ndir = (uint64_t)c->theta_count * c->phi_count;
    hipError_t e = h->V.alloc(h->nv);
    if (e == hipSuccess) e = h->F2.alloc(h->nv);
    if (e == hipSuccess) e = h->dirs.alloc(ndir);
    if (e == hipSuccess) e = h->counters.alloc(16);
    if (e == hipSuccess) {
        const std::vector<v3> d = grid_directions(h->grid);
        e = hipMemcpy(h->dirs.p, d.data(), d.size() * sizeof(v3), hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) {
        std::vector<double> ed;
        if (!make_u_edges(c->theta_count, ed, err)) {
            delete h;
            return invalid(ctx, err);
        }
        e = h->edges.alloc(ed.size());
        if (e == hipSuccess) e = hipMemcpy(h->edges.p, ed.data(), ed.size() * sizeof(double), hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
        delete h;
        return fail(ctx, IPH_E_DEVICE, std::string("coherent record allocation: ") + hipGetErrorString(e));
    }
    ctx->coherents.push_back(h);
    *out = h;
    return IPH_OK;
}

void iph_coherent_destroy(iph_coherent* h)
{
    if (!h) return;
    iph_ctx* ctx = h->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    auto it = std::find(ctx->coherents.begin(), ctx->coherents.end(), h);
    if (it != ctx->coherents.end()) ctx->coherents.erase(it);
    delete h;
}

int iph_coherent_trace(iph_coherent* h)
{
    if (!h) return fail(nullptr, IPH_E_INVALID, "handle must not be NULL.");
    iph_ctx* ctx = h->ctx;
    IPH_HIP(ctx, hipSetDevice(ctx->device));
    Timer t(ctx->stream);
    const uint32_t n = h->tp.gw * h->tp.gh;
    hipLaunchKernelGGL(icephase_coh_trace, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0, ctx->stream, h->tp,
                       h->V.p, h->F2.p);
    IPH_HIP(ctx, hipGetLastError());
    h->stats.ms_trace = t.stop();
    IPH_HIP(ctx, hipStreamSynchronize(ctx->stream));
    h->traced = true;
    h->evaluated = false;
    return IPH_OK;
}

int iph_coherent_set_vertices(iph_coherent* h, const void* vertices, const void* second)
{
    if (!h || !vertices || !second) return fail(h ? h->ctx : nullptr, IPH_E_INVALID, "Invalid arguments.");
    iph_ctx* ctx = h->ctx;
    IPH_HIP(ctx, hipSetDevice(ctx->device));
    IPH_HIP(ctx, hipMemcpy(h->V.p, vertices, (size_t)h->nv * sizeof(Vertex), hipMemcpyHostToDevice));
    IPH_HIP(ctx, hipMemcpy(h->F2.p, second, (size_t)h->nv * sizeof(Field), hipMemcpyHostToDevice));
    h->traced = true;
    h->evaluated = false;
    h->stats.ms_trace = 0;
    return IPH_OK;
}

int iph_coherent_get_vertices(iph_coherent* h, void* vertices, void* second)
{
    if (!h || !vertices || !second) return fail(h ? h->ctx : nullptr, IPH_E_INVALID, "Invalid arguments.");
    if (!h->traced) return fail(h->ctx, IPH_E_INVALID, "No vertices: trace or set them first.");
    iph_ctx* ctx = h->ctx;
    IPH_HIP(ctx, hipSetDevice(ctx->device));
    IPH_HIP(ctx, hipMemcpy(vertices, h->V.p, (size_t)h->nv * sizeof(Vertex), hipMemcpyDeviceToHost));
    IPH_HIP(ctx, hipMemcpy(second, h->F2.p, (size_t)h->nv * sizeof(Field), hipMemcpyDeviceToHost));
    return IPH_OK;
}

int iph_coherent_evaluate(iph_coherent* h)
{
    if (!h) return fail(nullptr, IPH_E_INVALID, "handle must not be NULL.");
    iph_ctx* ctx = h->ctx;
    if (!h->traced) return fail(ctx, IPH_E_INVALID, "No vertices: trace or set them first.");
    IPH_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    h->evaluated = false;
    iph_coherent_stats& st = h->stats;
    const double ms_trace = st.ms_trace;
    std::memset(&st, 0, sizeof st);
    st.ms_trace = ms_trace;
    unsigned int cnt[16];
    // trace validation and patch classification
    {
        Timer t(s);
        IPH_HIP(ctx, hipMemsetAsync(h->counters.p, 0, 16 * sizeof(unsigned int), s));
        IPH_HIP(ctx, fill_u32(h, h->counters.p + 9, 1, 0xffffffffu));
        hipLaunchKernelGGL(icephase_coh_validate, dim3(blocks_for(h->nv, kThreads)), dim3(kThreads), 0, s, h->V.p,
                           h->F2.p, h->nv, h->counters.p + 8, h->counters.p + 9);
        IPH_HIP(ctx, grow(h->flag, h->L.cells));
        IPH_HIP(ctx, grow(h->offsets, (size_t)h->L.cells + 1));
        hipLaunchKernelGGL(icephase_coh_classify, dim3(blocks_for(h->L.cells, kThreads)), dim3(kThreads), 0, s, h->V.p,
                           h->L, h->flag.p, h->counters.p);
        IPH_HIP(ctx, hipGetLastError());
        IPH_HIP(ctx, scan(h, h->flag.p, h->L.cells, h->offsets.p));
        uint64_t np64 = 0;
        IPH_HIP(ctx, hipMemcpyAsync(&np64, h->offsets.p + h->L.cells, sizeof np64, hipMemcpyDeviceToHost, s));
        IPH_HIP(ctx, hipMemcpyAsync(cnt, h->counters.p, sizeof cnt, hipMemcpyDeviceToHost, s));
        IPH_HIP(ctx, hipStreamSynchronize(s));
        for (int k = 0; k < 8; ++k) st.status_counts[k] = cnt[k];
        st.bad_vertices = cnt[8];
        if (cnt[8]) {
            st.first_bad_vertex = cnt[9];
            return fail(ctx, IPH_E_AUDIT, "Trace contains failed vertices (first at " + std::to_string(cnt[9]) + ").");
        }
        if (cnt[kInvalidVertex] || cnt[kInvalidGeometry])
            return fail(ctx, IPH_E_AUDIT, "Invalid source vertices: the patches were not built.");
        h->npatch = (uint32_t)np64;
        st.patches = h->npatch;
        const uint32_t npx = h->npatch ? h->npatch : 1;
        IPH_HIP(ctx, grow(h->P, npx));
        IPH_HIP(ctx, grow(h->fold_flag, npx));
        IPH_HIP(ctx, grow(h->fold_slot, (size_t)npx + 1));
        IPH_HIP(ctx, grow(h->lookup, npx));
        hipLaunchKernelGGL(icephase_coh_compact, dim3(blocks_for(h->L.cells, kThreads)), dim3(kThreads), 0, s, h->V.p,
                           h->L, h->flag.p, h->offsets.p, h->P.p, h->fold_flag.p);
        IPH_HIP(ctx, hipGetLastError());
        st.ms_patches = t.stop();
    }
    // folded records
    {
        Timer t(s);
        IPH_HIP(ctx, scan(h, h->fold_flag.p, h->npatch, h->fold_slot.p));
        uint64_t nf = 0;
        IPH_HIP(ctx, hipMemcpyAsync(&nf, h->fold_slot.p + h->npatch, sizeof nf, hipMemcpyDeviceToHost, s));
        IPH_HIP(ctx, hipStreamSynchronize(s));
        h->nfold = (uint32_t)nf;
        st.folded = h->nfold;
        IPH_HIP(ctx, grow(h->rec, h->nfold ? h->nfold : 1));
        if (h->npatch)
            hipLaunchKernelGGL(icephase_coh_fold, dim3(blocks_for(h->npatch, 64)), dim3(64), 0, s, h->V.p, h->nv,
                               h->P.p, h->npatch, h->fold_flag.p, h->fold_slot.p, h->tp.w, h->focal, 1u, h->fold,
                               h->lookup.p, h->rec.p);
        IPH_HIP(ctx, hipGetLastError());
        st.ms_folded = t.stop();
    }
    // query
    const uint64_t ndir = (uint64_t)h->grid.nt * h->grid.np;
    {
        Timer t(s);
        const uint32_t npx = h->npatch ? h->npatch : 1;
        IPH_HIP(ctx, grow(h->ext, npx));
        IPH_HIP(ctx, grow(h->geom, npx));
        IPH_HIP(ctx, grow(h->cand_count, npx));
        IPH_HIP(ctx, grow(h->coff, (size_t)npx + 1));
        if (h->npatch)
            hipLaunchKernelGGL(icephase_coh_extent, dim3(blocks_for(h->npatch, kThreads)), dim3(kThreads), 0, s,
                               h->V.p, h->P.p, h->npatch, h->grid, h->ext.p, h->cand_count.p, h->geom.p);
        IPH_HIP(ctx, scan(h, h->cand_count.p, h->npatch, h->coff.p));
        uint64_t total = 0;
        IPH_HIP(ctx, hipMemcpyAsync(&total, h->coff.p + h->npatch, sizeof total, hipMemcpyDeviceToHost, s));
        IPH_HIP(ctx, hipStreamSynchronize(s));
        h->candidates = total;
        st.candidates = total;
        IPH_HIP(ctx, grow(h->qcnt, ndir));
        IPH_HIP(ctx, grow(h->qnotes, ndir));
        IPH_HIP(ctx, grow(h->qfirst, ndir));
        IPH_HIP(ctx, grow(h->qcursor, ndir));
        IPH_HIP(ctx, grow(h->hoff, ndir + 1));
        IPH_HIP(ctx, fill_u32(h, h->qcnt.p, ndir, 0));
        IPH_HIP(ctx, fill_u32(h, h->qnotes.p, ndir, 0));
        IPH_HIP(ctx, fill_u32(h, h->qfirst.p, ndir, 0xffffffffu));
        IPH_HIP(ctx, fill_u32(h, h->qcursor.p, ndir, 0));
        QueryPass Q{};
        Q.G = h->geom.p;
        Q.ext = h->ext.p;
        Q.coff = h->coff.p;
        Q.dirs = h->dirs.p;
        Q.npatch = h->npatch;
        Q.np = h->grid.np;
        Q.total = total;
        Q.hits = h->qcnt.p;
        Q.notes = h->qnotes.p;
        Q.first = h->qfirst.p;
        const uint32_t qblocks = (uint32_t)std::min<uint64_t>(blocks_for(total ? total : 1, kThreads),
                                                              (uint64_t)ctx->cus * 16);
        // staging for the count pass: four hits per direction to start with,
        // grown to the last record's count, plus one chunk per wave (the most its
        // gaps can take)
        const uint64_t gaps = (uint64_t)qblocks * (kThreads / 64) * kStageChunk;
        IPH_HIP(ctx, grow(h->stage, std::max<uint64_t>(std::max<uint64_t>(4 * ndir, h->nhits), 1u << 20) + gaps));
        IPH_HIP(ctx, grow(h->staged, 1));
        IPH_HIP(ctx, hipMemsetAsync(h->staged.p, 0, sizeof(unsigned long long), s));
        Q.stage = h->stage.p;
        Q.staged = h->staged.p;
        Q.cap = h->stage.n;
        if (total) hipLaunchKernelGGL(icephase_coh_query, dim3(qblocks), dim3(kThreads), 0, s, Q);
        IPH_HIP(ctx, hipGetLastError());
        IPH_HIP(ctx, scan(h, h->qcnt.p, ndir, h->hoff.p));
        uint64_t nh = 0;
        unsigned long long reserved = 0;  // staging slots the waves reserved (hits and gaps)
        IPH_HIP(ctx, hipMemcpyAsync(&nh, h->hoff.p + ndir, sizeof nh, hipMemcpyDeviceToHost, s));
        IPH_HIP(ctx, hipMemcpyAsync(&reserved, h->staged.p, sizeof reserved, hipMemcpyDeviceToHost, s));
        IPH_HIP(ctx, hipStreamSynchronize(s));
        h->nhits = nh;
        st.hits = nh;
        IPH_HIP(ctx, grow(h->H, nh ? nh : 1));
        if (reserved <= Q.cap) {
            if (reserved)
                hipLaunchKernelGGL(icephase_coh_scatter, dim3(blocks_for(reserved, kThreads)), dim3(kThreads), 0, s,
                                   h->stage.p, (uint64_t)reserved, h->hoff.p, h->qcnt.p, h->qcursor.p, h->H.p);
        } else {
            Q.fill = 1;
            Q.hoff = h->hoff.p;
            Q.cursor = h->qcursor.p;
            Q.out = h->H.p;
            if (total) hipLaunchKernelGGL(icephase_coh_query, dim3(qblocks), dim3(kThreads), 0, s, Q);
        }
        IPH_HIP(ctx, hipGetLastError());
        IPH_HIP(ctx, grow(h->S, ndir));
        hipLaunchKernelGGL(icephase_coh_finish, dim3(blocks_for(ndir, kThreads)), dim3(kThreads), 0, s, h->H.p,
                           h->hoff.p, h->qcnt.p, h->qcursor.p, h->qnotes.p, h->qfirst.p, h->P.p, h->npatch, ndir,
                           h->S.p);
        IPH_HIP(ctx, hipGetLastError());
        st.ms_query = t.stop();
    }
    // optics
    {
        Timer t(s);
        IPH_HIP(ctx, grow(h->R, ndir));
        OpticsPass O{};
        O.V = h->V.p;
        O.F2 = h->F2.p;
        O.nv = h->nv;
        O.P = h->P.p;
        O.npatch = h->npatch;
        O.lookup = h->lookup.p;
        O.rec = h->rec.p;
        O.nrec = h->nfold;
        O.dirs = h->dirs.p;
        O.hoff = h->hoff.p;
        O.H = h->H.p;
        O.S = h->S.p;
        O.inc = h->tp.w;
        O.bx = h->tp.bx;
        O.focal = h->focal;
        O.ndir = ndir;
        O.out = h->R.p;
        hipLaunchKernelGGL(icephase_coh_optics, dim3(blocks_for(ndir, 128)), dim3(128), 0, s, O);
        IPH_HIP(ctx, hipGetLastError());
        st.ms_optics = t.stop();
    }
    // diffraction (needs a sigma: Table II range or an explicit one)
    {
        Timer t(s);
        IPH_HIP(ctx, grow(h->trans, ndir));
        IPH_HIP(ctx, grow(h->D, ndir));
        IPH_HIP(ctx, grow(h->rowcnt, h->grid.nt));
        IPH_HIP(ctx, grow(h->rowoff, (size_t)h->grid.nt + 1));
        hipLaunchKernelGGL(icephase_coh_detect, dim3(blocks_for(ndir, kThreads)), dim3(kThreads), 0, s, h->R.p,
                           h->diff, h->trans.p);
        IPH_HIP(ctx, fill_u32(h, h->rowcnt.p, h->grid.nt, 0));
        hipLaunchKernelGGL(icephase_coh_rowflags, dim3(blocks_for(ndir, kThreads)), dim3(kThreads), 0, s,
                           h->trans.p, h->grid.nt, h->grid.np, h->rowcnt.p);
        IPH_HIP(ctx, scan(h, h->rowcnt.p, h->grid.nt, h->rowoff.p));
        hipLaunchKernelGGL(icephase_coh_filter, dim3(blocks_for(ndir, 128)), dim3(128), 0, s, h->R.p, h->trans.p,
                           h->rowoff.p, h->diff, h->D.p);
        IPH_HIP(ctx, hipGetLastError());
        st.ms_diffraction = t.stop();
    }
    IPH_HIP(ctx, hipStreamSynchronize(s));
    h->evaluated = true;
    return IPH_OK;
}

int iph_coherent_stats_get(iph_coherent* h, iph_coherent_stats* out)
{
    if (!h || !out) return fail(h ? h->ctx : nullptr, IPH_E_INVALID, "Invalid arguments.");
    *out = h->stats;
    return IPH_OK;
}

int iph_coherent_get_patches(iph_coherent* h, void* patches, uint32_t cap, uint32_t* count, void* folded,
                             uint32_t fcap, uint32_t* fcount)
{
    if (!h || !count || !fcount) return fail(h ? h->ctx : nullptr, IPH_E_INVALID, "Invalid arguments.");
    if (!h->evaluated) return fail(h->ctx, IPH_E_INVALID, "Evaluate the record first.");
    iph_ctx* ctx = h->ctx;
    *count = h->npatch;
    *fcount = h->nfold;
    if ((patches && cap < h->npatch) || (folded && fcap < h->nfold))
        return fail(ctx, IPH_E_INVALID, "Output capacity below the patch or folded-record count.");
    IPH_HIP(ctx, hipSetDevice(ctx->device));
    if (patches && h->npatch)
        IPH_HIP(ctx, hipMemcpy(patches, h->P.p, (size_t)h->npatch * sizeof(Patch), hipMemcpyDeviceToHost));
    if (folded && h->nfold)
        IPH_HIP(ctx, hipMemcpy(folded, h->rec.p, (size_t)h->nfold * sizeof(Folded), hipMemcpyDeviceToHost));
    return IPH_OK;
}

int iph_coherent_get_hits(iph_coherent* h, uint32_t first, uint32_t count, void* hits, uint64_t cap,
                          uint32_t* counts, uint32_t* flags)
{
    if (!h || !hits || !counts || !flags) return fail(h ? h->ctx : nullptr, IPH_E_INVALID, "Invalid arguments.");
    if (!h->evaluated) return fail(h->ctx, IPH_E_INVALID, "Evaluate the record first.");
    const uint64_t ndir = (uint64_t)h->grid.nt * h->grid.np;
    if ((uint64_t)first + count > ndir) return fail(h->ctx, IPH_E_INVALID, "Direction range beyond the query grid.");
    iph_ctx* ctx = h->ctx;
    IPH_HIP(ctx, hipSetDevice(ctx->device));
    std::vector<Summary> S(count);
    std::vector<uint64_t> off(count + 1);
    if (count) {
        IPH_HIP(ctx, hipMemcpy(S.data(), h->S.p + first, count * sizeof(Summary), hipMemcpyDeviceToHost));
        IPH_HIP(ctx, hipMemcpy(off.data(), h->hoff.p + first, (count + 1) * sizeof(uint64_t), hipMemcpyDeviceToHost));
    }
    uint64_t used = 0;
    for (uint32_t k = 0; k < count; ++k) {
        counts[k] = S[k].hits;
        flags[k] = S[k].flags;
        if (used + S[k].hits > cap) return fail(ctx, IPH_E_INVALID, "Hit capacity exceeded.");
        if (S[k].hits)
            IPH_HIP(ctx, hipMemcpy(static_cast<QHit*>(hits) + used, h->H.p + off[k], S[k].hits * sizeof(QHit),
                                   hipMemcpyDeviceToHost));
        used += S[k].hits;
    }
    return IPH_OK;
}

int iph_coherent_get_directions(iph_coherent* h, iph_coherent_direction* out)
{
    if (!h || !out) return fail(h ? h->ctx : nullptr, IPH_E_INVALID, "Invalid arguments.");
    if (!h->evaluated) return fail(h->ctx, IPH_E_INVALID, "Evaluate the record first.");
    iph_ctx* ctx = h->ctx;
    const uint64_t ndir = (uint64_t)h->grid.nt * h->grid.np;
    IPH_HIP(ctx, hipSetDevice(ctx->device));
    std::vector<Optics> R(ndir);
    std::vector<Diffracted> D(ndir);
    IPH_HIP(ctx, hipMemcpy(R.data(), h->R.p, ndir * sizeof(Optics), hipMemcpyDeviceToHost));
    IPH_HIP(ctx, hipMemcpy(D.data(), h->D.p, ndir * sizeof(Diffracted), hipMemcpyDeviceToHost));
    for (uint64_t k = 0; k < ndir; ++k) {
        const Optics& r = R[k];
        iph_coherent_direction& o = out[k];
        std::memset(&o, 0, sizeof o);
        o.incoherent_s = r.inc_s;
        o.incoherent_p = r.inc_p;
        o.path_s = r.path_s;
        o.path_p = r.path_p;
        o.focal_s = r.focal_s;
        o.focal_p = r.focal_p;
        o.diffraction_s = D[k].s;
        o.diffraction_p = D[k].p;
        for (int f = 0; f < 4; ++f) {
            o.family_incoherent[f] = r.fam_inc[f];
            o.family_hits[f] = r.fam_hits[f];
        }
        o.sigma_rad = D[k].sigma;
        o.blend = D[k].blend;
        const F64* pf = r.path_field;
        const F64* ff = r.focal_field;
        for (int j = 0; j < 2; ++j) {
            o.path_field[4 * j + 0] = pf[j].sr;
            o.path_field[4 * j + 1] = pf[j].si;
            o.path_field[4 * j + 2] = pf[j].pr;
            o.path_field[4 * j + 3] = pf[j].pi;
            o.focal_field[4 * j + 0] = ff[j].sr;
            o.focal_field[4 * j + 1] = ff[j].si;
            o.focal_field[4 * j + 2] = ff[j].pr;
            o.focal_field[4 * j + 3] = ff[j].pi;
        }
        o.hit_count = r.hits;
        o.evaluated_hits = r.evaluated;
        o.rejected_hits = r.rejected;
        o.optical_flags = r.flags;
        o.query_flags = r.qflags;
        o.focal_flags = r.focal_flags;
        o.diffraction_flags = D[k].flags;
        o.transition_flags = D[k].transition;
        o.transition_kind = D[k].kind;
        o.folded_hits = r.folded_hits;
        o.regular_hits = r.regular_hits;
        o.refinement_hits = r.refinement_hits;
    }
    return IPH_OK;
}

int iph_coherent_density(iph_coherent* h, uint32_t stage, double* out, iph_coherent_gate* gate)
{
    if (!h || !gate) return fail(h ? h->ctx : nullptr, IPH_E_INVALID, "Invalid arguments.");
    if (stage < IPH_STAGE_INCOHERENT || stage > IPH_STAGE_DIFFRACTION)
        return fail(h->ctx, IPH_E_INVALID, "stage must be incoherent (1), path (2), focal (3) or diffraction (4).");
    if (!h->evaluated) return fail(h->ctx, IPH_E_INVALID, "Evaluate the record first.");
    if (stage >= IPH_STAGE_FOCAL && !(h->diff.sigma > 0))
        return fail(h->ctx, IPH_E_INVALID,
                    "Table II covers radius 0.1..1.0 mm only; focal and diffraction need a primary sigma.");
    iph_ctx* ctx = h->ctx;
    IPH_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    const uint64_t n = (uint64_t)h->grid.nt * h->grid.np;
    IPH_HIP(ctx, grow(h->density, n));
    IPH_HIP(ctx, grow(h->maxbits, 1));
    IPH_HIP(ctx, hipMemsetAsync(h->counters.p, 0, 16 * sizeof(unsigned int), s));
    IPH_HIP(ctx, fill_u32(h, h->counters.p + 3, 1, 0xffffffffu));
    IPH_HIP(ctx, hipMemsetAsync(h->maxbits.p, 0, sizeof(unsigned long long), s));
    hipLaunchKernelGGL(icephase_coh_density, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0, s, h->R.p, h->D.p,
                       stage, n, h->density.p, h->counters.p, h->maxbits.p);
    IPH_HIP(ctx, hipGetLastError());
    unsigned int g[4];
    unsigned long long mb = 0;
    IPH_HIP(ctx, hipMemcpyAsync(g, h->counters.p, sizeof g, hipMemcpyDeviceToHost, s));
    IPH_HIP(ctx, hipMemcpyAsync(&mb, h->maxbits.p, sizeof mb, hipMemcpyDeviceToHost, s));
    if (out) IPH_HIP(ctx, hipMemcpyAsync(out, h->density.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
    IPH_HIP(ctx, hipStreamSynchronize(s));
    gate->invalid_values = g[0];
    gate->incomplete = g[1];
    gate->underresolved = g[2];
    gate->first_problem = g[3];
    double mx;
    std::memcpy(&mx, &mb, sizeof mx);
    gate->maximum = mx;
    return IPH_OK;
}

int iph_coherent_cdf(iph_coherent* h, uint32_t stage, uint32_t allow_underresolved, const iph_cdf_policy* policy,
                     double* phi, double* theta, iph_cdf_report* report, iph_coherent_gate* gate)
{
    if (!h || !phi || !theta) return fail(h ? h->ctx : nullptr, IPH_E_INVALID, "Invalid arguments.");
    iph_coherent_gate g{};
    const int rc = iph_coherent_density(h, stage, nullptr, &g);
    if (gate) *gate = g;
    if (rc) return rc;
    iph_ctx* ctx = h->ctx;
    if (g.invalid_values || g.incomplete || !(g.maximum > 0) || (!allow_underresolved && g.underresolved))
        return fail(ctx, IPH_E_AUDIT,
                    "CDF input rejected: invalid_values=" + std::to_string(g.invalid_values) +
                        ", incomplete=" + std::to_string(g.incomplete) + ", underresolved=" +
                        std::to_string(g.underresolved) + ", max_intensity=" + std::to_string(g.maximum) +
                        (g.underresolved && !allow_underresolved ? " (diffraction is underresolved on this query grid)"
                                                                 : ""));
    const uint64_t n = (uint64_t)h->grid.nt * h->grid.np;
    IPH_HIP(ctx, grow(h->weights, n));
    hipLaunchKernelGGL(icephase_coh_weights, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0, ctx->stream,
                       h->density.p, h->edges.p, h->grid.np, n, h->maxbits.p, h->weights.p);
    IPH_HIP(ctx, hipGetLastError());
    return icephase::build_cdf_from_device(ctx, h->weights.p, h->grid.nt, h->grid.np, policy, phi, theta, report);
}

}  // extern "C"
