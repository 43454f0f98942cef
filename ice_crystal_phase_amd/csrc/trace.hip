// Fused trace + bin kernels for gfx950.
//
// One persistent launch covers every record of a batch. The host cuts each
// record into tiles of consecutive ray ids; workgroups pull tiles from an
// atomic queue until the queue is empty (the exit condition every wave
// reaches), so uneven records and uneven ray lengths balance across the 256
// CUs. Each lane traces whole rays and sends every emission to the record's
// histogram as a u64 fixed-point quantum (rint(power * 2^Q)): integer sums are
// associative, so the histogram is bitwise independent of tile order, wave
// scheduling and atomic arrival order. Pole quanta and audit counters stay in
// lane registers and are reduced per wave before one atomic per counter.
//
// The record's u edges are staged in LDS when they fit (nt <= kLdsEdges-1) so
// the row lookup (phase_bin.hpp) reads LDS, not L2.
#include "kernels.hpp"

namespace icephase {

constexpr int kLdsEdges = 2048;

__device__ inline unsigned long long wave_sum(unsigned long long v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

struct LaneCounters {
    unsigned long long fwd, bwd, samples, invalid, snapped, escaped, unresolved, hits;
    uint32_t status[9];
};

__device__ inline void flush_counters(LaneCounters& c, unsigned long long* stats)
{
    const unsigned long long v[8] = {c.fwd, c.bwd, c.samples, c.invalid, c.snapped, c.escaped, c.unresolved, c.hits};
    const uint32_t slot[8] = {kStatForward, kStatBackward, kStatSamples, kStatInvalid, kStatSnapped, kStatEscaped,
                              kStatUnresolved, kStatInternalHits};
    const bool lead = (threadIdx.x & 63) == 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const unsigned long long s = wave_sum(v[k]);
        if (lead && s) atomicAdd(&stats[slot[k]], s);
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const unsigned long long s = wave_sum((unsigned long long)c.status[k]);
        if (lead && s) atomicAdd(&stats[kStatStatus0 + k], s);
    }
}

// Predicated adds keep the status array in registers (no dynamic indexing).
__device__ inline void count_status(LaneCounters& c, uint32_t s)
{
    s = s <= 8 ? s : 7;
#pragma unroll
    for (uint32_t k = 0; k < 9; ++k) c.status[k] += (s == k) ? 1u : 0u;
}

template <class EdgeLoad>
__device__ inline void deposit(const DevRecord& R, d3 dir, double power, EdgeLoad edge, LaneCounters& c,
                               unsigned long long& ray_quanta)
{
    const unsigned long long q = quantize(power, R.qbits);
    ray_quanta += q;
    ++c.samples;
    const CellHit h = bin_direction(R.frame, R.nt, R.np, dir, edge);
    if (h.kind == 2) {
        ++c.invalid;
        return;
    }
    c.snapped += h.snapped;
    if (h.kind == 1) {
        c.fwd += q;
    } else if (h.kind == -1) {
        c.bwd += q;
    } else if (q) {
        atomicAdd(&R.bins[h.index], q);
    }
}

template <int KIND, class EdgeLoad>
__device__ inline void trace_tile(const DevRecord& R, uint64_t first, uint64_t count, EdgeLoad edge,
                                  LaneCounters& c)
{
    if constexpr (KIND == kKindHex) {
        for (uint64_t i = threadIdx.x; i < count; i += blockDim.x) {
            unsigned long long rq = 0;
            auto sink = [&](d3 dir, double pw) { deposit(R, dir, pw, edge, c, rq); };
            const hex::RayResult r = hex::trace(R.hex, first + i, sink);
            c.escaped += rq;
            c.unresolved += quantize(r.unresolved, R.qbits);
            c.hits += r.internal_hits;
            count_status(c, r.status);
        }
    } else {
        for (uint64_t i = threadIdx.x; i < count; i += blockDim.x) {
            unsigned long long rq = 0;
            auto sink = [&](drop::f3 dir, float pw) {
                d3 d;
                if (!unit_of(D3((double)dir.x, (double)dir.y, (double)dir.z), d)) {
                    ++c.samples;
                    ++c.invalid;
                    rq += quantize((double)pw, R.qbits);
                    return;
                }
                deposit(R, d, (double)pw, edge, c, rq);
            };
            const drop::RayResult r = drop::trace(R.drop, first + i, sink);
            c.escaped += rq;
            c.unresolved += quantize((double)r.unresolved, R.qbits);
            c.hits += r.internal_hits;
            count_status(c, r.status);
        }
    }
}

// Persistent tile loop of one record kind. LDS_EDGES: the record's u edges are
// staged in LDS (nt < kLdsEdges); otherwise they are read from global memory.
template <int KIND, bool LDS_EDGES>
__device__ inline void trace_loop(const DevRecord* __restrict__ recs, TileMap map, uint32_t* __restrict__ queue)
{
    __shared__ double s_edges[LDS_EDGES ? kLdsEdges : 1];
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_cached;
    if (threadIdx.x == 0) s_cached = 0xffffffffu;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) s_tile = atomicAdd(queue, 1u);
        __syncthreads();
        const uint64_t t = __builtin_amdgcn_readfirstlane(s_tile);
        if (t >= map.total) break;
        // record of tile t: last k with tile_start[k] <= t (uniform scalar search)
        uint32_t lo = 0, hi = map.n;
        while (lo + 1 < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (map.tile_start[mid] <= t)
                lo = mid;
            else
                hi = mid;
        }
        lo = __builtin_amdgcn_readfirstlane(lo);
        const uint32_t rec = __builtin_amdgcn_readfirstlane(map.recs_of_kind[lo]);
        const DevRecord& R = recs[rec];
        const uint64_t first = (t - map.tile_start[lo]) * R.tile_rays;
        const uint64_t count = R.rays - first < R.tile_rays ? R.rays - first : R.tile_rays;
        LaneCounters c{};
        if constexpr (LDS_EDGES) {
            if (s_cached != rec) {
                __syncthreads();
                for (uint32_t k = threadIdx.x; k <= R.nt; k += blockDim.x) s_edges[k] = R.edges[k];
                __syncthreads();
                if (threadIdx.x == 0) s_cached = rec;
                __syncthreads();
            }
            trace_tile<KIND>(R, first, count, [&](int k) { return s_edges[k]; }, c);
        } else {
            const double* e = R.edges;
            trace_tile<KIND>(R, first, count, [&](int k) { return e[k]; }, c);
        }
        flush_counters(c, R.stats);
    }
}

extern "C" __global__ void __launch_bounds__(256)
    icephase_trace_drop(const DevRecord* __restrict__ recs, TileMap map, uint32_t* __restrict__ queue)
{
    trace_loop<kKindDrop, true>(recs, map, queue);
}
extern "C" __global__ void __launch_bounds__(256)
    icephase_trace_hex(const DevRecord* __restrict__ recs, TileMap map, uint32_t* __restrict__ queue)
{
    trace_loop<kKindHex, true>(recs, map, queue);
}
extern "C" __global__ void __launch_bounds__(256)
    icephase_trace_drop_wide(const DevRecord* __restrict__ recs, TileMap map, uint32_t* __restrict__ queue)
{
    trace_loop<kKindDrop, false>(recs, map, queue);
}
extern "C" __global__ void __launch_bounds__(256)
    icephase_trace_hex_wide(const DevRecord* __restrict__ recs, TileMap map, uint32_t* __restrict__ queue)
{
    trace_loop<kKindHex, false>(recs, map, queue);
}

// Bin caller directions (C-ABI iph_bin_directions).
extern "C" __global__ void __launch_bounds__(256)
    icephase_bin_dirs(BinFrame g, uint32_t nt, uint32_t np, const double* __restrict__ edges, uint64_t n,
                      const double* __restrict__ dirs, uint32_t* __restrict__ index, int32_t* __restrict__ kind)
{
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (uint64_t)gridDim.x * blockDim.x) {
        const CellHit h = bin_direction(g, nt, np, D3(dirs[3 * k], dirs[3 * k + 1], dirs[3 * k + 2]),
                                        [&](int e) { return edges[e]; });
        index[k] = h.index;
        kind[k] = h.kind;
    }
}

// Zero every record's histogram and counters, and the tile queue. A kernel, not
// hipMemsetAsync, so that the whole batch can live in a captured graph.
extern "C" __global__ void __launch_bounds__(256)
    icephase_zero(const DevRecord* __restrict__ recs, uint32_t nrec, uint32_t* __restrict__ queue)
{
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x < 4) queue[threadIdx.x] = 0u;
    for (uint32_t r = blockIdx.y; r < nrec; r += gridDim.y) {
        const DevRecord& R = recs[r];
        const uint64_t n = (uint64_t)R.nt * R.np;
        for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (uint64_t)gridDim.x * blockDim.x)
            R.bins[k] = 0ull;
        if (blockIdx.x == 0)
            for (uint32_t k = threadIdx.x; k < kStatCount; k += blockDim.x) R.stats[k] = 0ull;
    }
}

}  // namespace icephase

namespace icephase {

hipError_t launch_zero(const DevRecord* d_recs, uint32_t nrec, uint32_t* d_queue, hipStream_t s)
{
    hipLaunchKernelGGL(icephase_zero, dim3(256, nrec < 65535 ? nrec : 65535), dim3(256), 0, s, d_recs, nrec, d_queue);
    return hipGetLastError();
}

// kind: kKindDrop or kKindHex; wide: edges too large for LDS.
hipError_t launch_trace(int kind, bool wide, const DevRecord* d_recs, TileMap map, uint32_t* d_queue, uint32_t blocks,
                        uint32_t threads, hipStream_t s)
{
    if (kind == kKindHex) {
        if (wide)
            hipLaunchKernelGGL(icephase_trace_hex_wide, dim3(blocks), dim3(threads), 0, s, d_recs, map, d_queue);
        else
            hipLaunchKernelGGL(icephase_trace_hex, dim3(blocks), dim3(threads), 0, s, d_recs, map, d_queue);
    } else {
        if (wide)
            hipLaunchKernelGGL(icephase_trace_drop_wide, dim3(blocks), dim3(threads), 0, s, d_recs, map, d_queue);
        else
            hipLaunchKernelGGL(icephase_trace_drop, dim3(blocks), dim3(threads), 0, s, d_recs, map, d_queue);
    }
    return hipGetLastError();
}

hipError_t launch_bin_dirs(const BinFrame& g, uint32_t nt, uint32_t np, const double* d_edges, uint64_t n,
                           const double* d_dirs, uint32_t* d_index, int32_t* d_kind, hipStream_t s)
{
    uint64_t blocks = (n + 255) / 256;
    blocks = blocks < 1 ? 1 : (blocks > 4096 ? 4096 : blocks);
    hipLaunchKernelGGL(icephase_bin_dirs, dim3((uint32_t)blocks), dim3(256), 0, s, g, nt, np, d_edges, n, d_dirs, d_index,
                       d_kind);
    return hipGetLastError();
}

}  // namespace icephase
