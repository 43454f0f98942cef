// FP64 CDF construction for gfx950, batched over records (blockIdx.y).
//
// Semantics follow the reference host CDF (hex_phase/src/phase_table.cpp:111-187,
// build_host_cdf) and its GPU counterpart's audit (device/phase_cdf.cu:164-236):
//   w = m / largest; per phi column an ORDERED binary64 prefix over theta (exact
//   plateaus, monotone); ordered marginal over phi; zero column -> u edges;
//   endpoints exactly 0 and 1; audit of (d phi)(d theta) against m / total.
// Every stored CDF value is produced by the same sequence of IEEE operations as
// the reference's ordered loops, so the arrays are bitwise the reference's for
// the same cell masses. Reductions that only feed the report (total mass, g,
// L1, TV) run in parallel with compensated per-lane sums.
//
// Layout: masses are theta-major [nt][np]; a column scan reads one row per
// step with consecutive lanes on consecutive phi (coalesced). theta is
// phi-major [np][nt+1]; each wave stages a 64 x 32 tile in LDS and writes it
// back as contiguous 256-byte column segments.
#include "kernels.hpp"

namespace icephase {

// Every array of a CdfJob lives in device global memory; its generic pointer is
// cast to the global address space so that reads and writes are global (not
// flat) instructions: flat accesses also count against the LDS / scalar-load
// counter, so each LDS or scalar wait would wait for the prefetched rows too.
template <class T>
using gptr = __attribute__((address_space(1))) T*;
template <class T>
__device__ inline gptr<T> G(T* p)
{
    return (gptr<T>)p;
}

constexpr int kRedThreads = 256;
constexpr uint32_t kSumChunk = 16;    // rows a column-sum lane has in flight
constexpr uint32_t kMomentBatch = 8;  // cells a moments thread has in flight

struct Neumaier {
    double s, c;
    __device__ void add(double x)
    {
        const double t = s + x;
        c += fabs(s) >= fabs(x) ? (s - t) + x : (x - t) + s;
        s = t;
    }
    __device__ double value() const { return s + c; }
};

// Cell mass of the integer source: (double)q + pole share in the end rows
// (same operation order as the oracle's orc_bins_to_mass).
__device__ inline double cell_mass(const CdfJob& J, uint32_t i, uint32_t j)
{
    if (!J.bins) return G(J.mass)[(uint64_t)i * J.np + j];
    double m = (double)G(J.bins)[(uint64_t)i * J.np + j];
    if (i == 0) m = m + (double)G(J.stats)[kStatForward] / (double)J.np;
    if (i + 1 == J.nt) m = m + (double)G(J.stats)[kStatBackward] / (double)J.np;
    return m;
}

// Mass of a stored cell: the fine mass, or the coarse buffer after aggregation.
__device__ inline double stored_mass(const CdfJob& J, uint32_t i, uint32_t j)
{
    if (J.coarse) return G(J.coarse)[(uint64_t)i * J.onp + j];
    return cell_mass(J, i, j);
}

// Reader of a job's stored cells with the source resolved once: FP64 rows (the
// coarse buffer, or the FP64 masses) or u64 quanta with the pole shares added in
// the end rows (the operations of cell_mass, with each share computed once).
struct StoredReader {
    gptr<const double> f;
    gptr<const unsigned long long> q;
    uint64_t stride;
    uint32_t last;  // last fine row
    double first_share, last_share;

    __device__ static StoredReader of(const CdfJob& J)
    {
        StoredReader r{nullptr, nullptr, J.np, J.nt - 1, 0.0, 0.0};
        if (J.coarse) {
            r.f = G((const double*)J.coarse);
            r.stride = J.onp;
        } else if (!J.bins) {
            r.f = G(J.mass);
        } else {
            r.q = G(J.bins);
            r.first_share = (double)G(J.stats)[kStatForward] / (double)J.np;
            r.last_share = (double)G(J.stats)[kStatBackward] / (double)J.np;
        }
        return r;
    }
    // The fine cells (cell_mass) even where the job stores a coarsened grid.
    __device__ static StoredReader fine(const CdfJob& J)
    {
        StoredReader r{nullptr, nullptr, J.np, J.nt - 1, 0.0, 0.0};
        if (!J.bins) {
            r.f = G(J.mass);
        } else {
            r.q = G(J.bins);
            r.first_share = (double)G(J.stats)[kStatForward] / (double)J.np;
            r.last_share = (double)G(J.stats)[kStatBackward] / (double)J.np;
        }
        return r;
    }
    // The cell's 64-bit word, whatever the source. Reading a batch of words and
    // converting them afterwards keeps every load of the batch in flight at once;
    // a conversion per load (the source test and the u64 -> double between the
    // loads) makes the compiler wait for each load before issuing the next.
    __device__ unsigned long long word(uint32_t i, uint32_t j) const
    {
        const gptr<const unsigned long long> w = q ? q : (gptr<const unsigned long long>)f;
        return w[(uint64_t)i * stride + j];
    }
    // The mass of row i from its word (the operations of cell_mass).
    __device__ double mass(unsigned long long w, uint32_t i) const
    {
        if (f) return __longlong_as_double((long long)w);
        double m = (double)w;
        if (i == 0) m = m + first_share;
        if (i == last) m = m + last_share;
        return m;
    }
};

__device__ inline double block_sum(double v, double* sh)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    __syncthreads();
    if (l == 0) sh[w] = v;
    __syncthreads();
    double t = 0;
    if (threadIdx.x == 0)
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) t += sh[k];
    return t;
}
__device__ inline double block_max(double v, double* sh)
{
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    __syncthreads();
    if (l == 0) sh[w] = v;
    __syncthreads();
    double t = 0;
    if (threadIdx.x == 0)
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) t = fmax(t, sh[k]);
    return t;
}

// Conservative aggregation: one thread per stored cell, Kahan in (row, column)
// order over its fine cells (phase_table.cpp:128-131 with math.hpp:120-128).
extern "C" __global__ void __launch_bounds__(kRedThreads) icephase_cdf_coarsen(const CdfJob* __restrict__ jobs)
{
    const CdfJob J = jobs[blockIdx.y];
    if (!J.coarse) return;
    const uint32_t dt = J.nt / J.ont, dp = J.np / J.onp;
    const uint64_t n = (uint64_t)J.ont * J.onp;
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t ci = (uint32_t)(k / J.onp), cj = (uint32_t)(k % J.onp);
        double v = 0, c = 0;
        for (uint32_t i = ci * dt; i < (ci + 1) * dt; ++i)
            for (uint32_t j = cj * dp; j < (cj + 1) * dp; ++j) {
                const double y = cell_mass(J, i, j) - c, t = v + y;
                c = (t - v) - y;
                v = t;
            }
        G(J.coarse)[k] = v;
    }
}

// Pass 1: fine moments (mass, axial), stored moments, largest stored mass,
// negative / non-finite count. partials: 8 doubles per block.
extern "C" __global__ void __launch_bounds__(kRedThreads) icephase_cdf_moments(const CdfJob* __restrict__ jobs)
{
    __shared__ double sh[kRedThreads / 64];
    const CdfJob J = jobs[blockIdx.y];
    Neumaier mf{0, 0}, af{0, 0}, ms{0, 0}, as{0, 0};
    double neg = 0, largest = 0;
    // rows over blocks, columns over threads: one row's reads are contiguous, and
    // each thread has kMomentBatch of them in flight
    const StoredReader fine = StoredReader::fine(J);
    for (uint32_t i = blockIdx.x; i < J.nt; i += gridDim.x) {
        const double axial = 1 - G(J.edges)[i] - G(J.edges)[i + 1];
        for (uint32_t j0 = threadIdx.x; j0 < J.np; j0 += blockDim.x * kMomentBatch) {
            unsigned long long wv[kMomentBatch];
#pragma unroll
            for (uint32_t b = 0; b < kMomentBatch; ++b) {
                const uint32_t j = j0 + b * blockDim.x;
                wv[b] = j < J.np ? fine.word(i, j) : 0ull;
            }
#pragma unroll
            for (uint32_t b = 0; b < kMomentBatch; ++b) {
                if (j0 + b * blockDim.x >= J.np) continue;
                const double m = fine.mass(wv[b], i);
                if (!(m >= 0 && m <= DBL_MAX)) {
                    neg += 1;
                    continue;
                }
                mf.add(m);
                af.add(m * axial);
                largest = fmax(largest, m);
            }
        }
    }
    if (!J.coarse) {
        ms = mf;
        as = af;
    }
    if (J.coarse) {
        largest = 0;
        for (uint32_t i = blockIdx.x; i < J.ont; i += gridDim.x) {
            const double axial = 1 - G(J.cedges)[i] - G(J.cedges)[i + 1];
            for (uint32_t j = threadIdx.x; j < J.onp; j += blockDim.x) {
                const double m = stored_mass(J, i, j);
                if (!(m >= 0 && m <= DBL_MAX)) continue;
                ms.add(m);
                as.add(m * axial);
                largest = fmax(largest, m);
            }
        }
    }
    const double v0 = block_max(largest, sh);
    const double v1 = block_sum(mf.value(), sh);
    const double v2 = block_sum(af.value(), sh);
    const double v3 = block_sum(ms.value(), sh);
    const double v4 = block_sum(as.value(), sh);
    const double v5 = block_sum(neg, sh);
    if (threadIdx.x == 0) {
        auto p = G(J.partials) + (uint64_t)blockIdx.x * 8;
        p[0] = v0; p[1] = v1; p[2] = v2; p[3] = v3; p[4] = v4; p[5] = v5; p[6] = 0; p[7] = 0;
    }
}

extern "C" __global__ void __launch_bounds__(kRedThreads)
    icephase_cdf_moments_final(const CdfJob* __restrict__ jobs, uint32_t blocks)
{
    __shared__ double sh[kRedThreads / 64];
    const CdfJob J = jobs[blockIdx.x];
    double mx = 0;
    Neumaier s1{0, 0}, s2{0, 0}, s3{0, 0}, s4{0, 0};
    double neg = 0;
    for (uint32_t b = threadIdx.x; b < blocks; b += blockDim.x) {
        const auto p = G(J.partials) + (uint64_t)b * 8;
        mx = fmax(mx, p[0]);
        s1.add(p[1]); s2.add(p[2]); s3.add(p[3]); s4.add(p[4]);
        neg += p[5];
    }
    const double v0 = block_max(mx, sh);
    const double v1 = block_sum(s1.value(), sh);
    const double v2 = block_sum(s2.value(), sh);
    const double v3 = block_sum(s3.value(), sh);
    const double v4 = block_sum(s4.value(), sh);
    const double v5 = block_sum(neg, sh);
    if (threadIdx.x == 0) {
        G(J.scalars)[kCdfLargest] = v0;
        G(J.scalars)[kCdfMassFine] = v1;
        G(J.scalars)[kCdfAxialFine] = v2;
        G(J.scalars)[kCdfMassStored] = v3;
        G(J.scalars)[kCdfAxialStored] = v4;
        G(J.scalars)[kCdfNegative] = v5;
    }
}

// Coarsening total variation (report only; phase_table.cpp:136-143).
extern "C" __global__ void __launch_bounds__(kRedThreads) icephase_cdf_tv(const CdfJob* __restrict__ jobs)
{
    __shared__ double sh[kRedThreads / 64];
    const CdfJob J = jobs[blockIdx.y];
    if (!J.coarse) return;
    const uint32_t dt = J.nt / J.ont, dp = J.np / J.onp;
    const double M = G(J.scalars)[kCdfMassFine], Mc = G(J.scalars)[kCdfMassStored];
    Neumaier tv{0, 0};
    const uint64_t n = (uint64_t)J.nt * J.np;
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t i = (uint32_t)(k / J.np), j = (uint32_t)(k % J.np);
        const uint32_t ci = i / dt, cj = j / dp;
        const double fraction = (G(J.edges)[i + 1] - G(J.edges)[i]) / (G(J.cedges)[ci + 1] - G(J.cedges)[ci]) / dp;
        tv.add(fabs(cell_mass(J, i, j) / M - G(J.coarse)[(uint64_t)ci * J.onp + cj] / Mc * fraction));
    }
    const double v = block_sum(tv.value(), sh);
    if (threadIdx.x == 0) G(J.partials)[(uint64_t)blockIdx.x * 8 + 6] = v;
}

extern "C" __global__ void __launch_bounds__(kRedThreads)
    icephase_cdf_tv_final(const CdfJob* __restrict__ jobs, uint32_t blocks)
{
    __shared__ double sh[kRedThreads / 64];
    const CdfJob J = jobs[blockIdx.x];
    if (!J.coarse) {
        if (threadIdx.x == 0) G(J.scalars)[kCdfTv] = 0;
        return;
    }
    Neumaier s{0, 0};
    for (uint32_t b = threadIdx.x; b < blocks; b += blockDim.x) s.add(G(J.partials)[(uint64_t)b * 8 + 6]);
    const double v = block_sum(s.value(), sh);
    if (threadIdx.x == 0) G(J.scalars)[kCdfTv] = 0.5 * v;
}

// Pass 2: per stored phi column, the ordered sum of w = m / largest.
extern "C" __global__ void __launch_bounds__(kRedThreads) icephase_cdf_colsum(const CdfJob* __restrict__ jobs)
{
    const CdfJob J = jobs[blockIdx.y];
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= J.onp) return;
    const double largest = G(J.scalars)[kCdfLargest];
    const StoredReader src = StoredReader::of(J);
    double s = 0;
    uint32_t under = 0;
    for (uint32_t i0 = 0; i0 < J.ont; i0 += kSumChunk) {
        const uint32_t rows = J.ont - i0 < kSumChunk ? J.ont - i0 : kSumChunk;
        // the chunk's reads first, all in flight together; then the ordered sum
        unsigned long long wv[kSumChunk];
#pragma unroll
        for (uint32_t r = 0; r < kSumChunk; ++r) wv[r] = r < rows ? src.word(i0 + r, j) : 0ull;
#pragma unroll
        for (uint32_t r = 0; r < kSumChunk; ++r) {
            if (r < rows) {
                const double m = src.mass(wv[r], i0 + r);
                const double w = m / largest;
                under += (m > 0 && !(w > 0)) ? 1u : 0u;
                s += w;
            }
        }
    }
    G(J.cols)[j] = s;
    if (under) atomicAdd(&J.scalars[kCdfUnderflow], (double)under);
}

// Pass 3: ordered marginal over phi and its normalization (phase_table.cpp:162-176).
// The column sums are staged in LDS by the block (kMarginalLds of them at a
// time), one lane runs the ordered prefix, and the block writes and normalizes.
constexpr uint32_t kMarginalLds = 4096;

extern "C" __global__ void __launch_bounds__(256) icephase_cdf_marginal(const CdfJob* __restrict__ jobs)
{
    __shared__ double s_pre[kMarginalLds];
    __shared__ double s_acc;
    const CdfJob J = jobs[blockIdx.x];
    if (threadIdx.x == 0) s_acc = 0;
    for (uint32_t j0 = 0; j0 < J.onp; j0 += kMarginalLds) {
        const uint32_t n = J.onp - j0 < kMarginalLds ? J.onp - j0 : kMarginalLds;
        __syncthreads();
        for (uint32_t k = threadIdx.x; k < n; k += blockDim.x) s_pre[k] = G(J.cols)[j0 + k];
        __syncthreads();
        if (threadIdx.x == 0) {
            double acc = s_acc;
            for (uint32_t k = 0; k < n; ++k) {
                acc = acc + s_pre[k];
                s_pre[k] = acc;
            }
            s_acc = acc;
        }
        __syncthreads();
        for (uint32_t k = threadIdx.x; k < n; k += blockDim.x) G(J.phi)[j0 + k + 1] = s_pre[k];
    }
    __syncthreads();
    const double total = s_acc;
    if (threadIdx.x == 0) {
        G(J.scalars)[kCdfTotal] = total;
        G(J.phi)[0] = 0;
        G(J.phi)[J.onp] = 1;
    }
    // phi[j] / total for 1 <= j < onp; each thread divides the values it wrote
    for (uint32_t j0 = 0; j0 < J.onp; j0 += kMarginalLds) {
        const uint32_t n = J.onp - j0 < kMarginalLds ? J.onp - j0 : kMarginalLds;
        for (uint32_t k = threadIdx.x; k < n; k += blockDim.x) {
            const uint32_t j = j0 + k + 1;
            if (j < J.onp) G(J.phi)[j] = G(J.phi)[j] / total;
        }
    }
}

// Pass 4: the ordered prefix again (same operations -> same values), normalized
// into phi-major theta, with the audit of each cell. A block owns a strip of 64
// stored columns and walks it in tiles of 32 rows staged in LDS: its 256 threads
// read the tile row by row (coalesced) and form w = m / largest; one wave runs
// the ordered prefix down each column (the only serial step); all threads form
// the normalized values and the audit terms; the tile leaves as contiguous
// 256-byte column segments. The divisions thus run at full width instead of
// one lane per column.
constexpr uint32_t kThetaCols = 64;
constexpr uint32_t kThetaRows = 32;
constexpr uint32_t kThetaThreads = 256;
constexpr uint32_t kThetaPhases = kThetaThreads / kThetaCols;  // rows r = phase + 4k of a tile per thread

extern "C" __global__ void __launch_bounds__(kThetaThreads) icephase_cdf_theta(const CdfJob* __restrict__ jobs)
{
    __shared__ double sv[kThetaRows][kThetaCols + 1];  // w, then the prefix, then the normalized value
    __shared__ double s_carry[kThetaCols];               // prefix after the previous tile
    __shared__ double n_carry[2][kThetaCols];            // normalized value of the previous tile's last row
    __shared__ double s_red[kThetaPhases][8];
    __shared__ double s_ce[kThetaRows + 1];  // the tile's stored u edges
    const CdfJob J = jobs[blockIdx.y];
    const StoredReader src = StoredReader::of(J);
    const uint32_t c = threadIdx.x % kThetaCols, ph = threadIdx.x / kThetaCols;
    const uint32_t j0 = blockIdx.x * kThetaCols;
    const uint32_t j = j0 + c;
    const bool active = j < J.onp;
    const uint32_t nt = J.ont;
    const uint64_t stride = (uint64_t)nt + 1;
    const double largest = G(J.scalars)[kCdfLargest];
    const double total = G(J.scalars)[kCdfTotal];
    const double Mc = G(J.scalars)[kCdfMassStored];
    const double col = active ? G(J.cols)[j] : 0.0;
    const double a0 = active ? G(J.phi)[j] : 0.0, a1 = active ? G(J.phi)[j + 1] : 0.0;
    const double da = a1 - a0;
    // audit sums (report only): plain FP64 per thread over at most a few hundred
    // cells, then fixed-order wave and block sums
    double l1 = 0, lost = 0, rec = 0, ax = 0;
    double maxcell = 0, lostcells = 0, malformed = 0;
    if (ph == 0) {
        s_carry[c] = 0;
        n_carry[0][c] = 0;
        if (active) {
            G(J.theta)[(uint64_t)j * stride] = 0.0;
            malformed += (!(a1 >= a0) || !(total > 0) || !(total <= DBL_MAX)) ? 1 : 0;
        }
    }
    uint32_t tile = 0;
    for (uint32_t i0 = 0; i0 < nt; i0 += kThetaRows, ++tile) {
        const uint32_t rows = nt - i0 < kThetaRows ? nt - i0 : kThetaRows;
        if (threadIdx.x <= rows) s_ce[threadIdx.x] = G(J.cedges)[i0 + threadIdx.x];
        // stored masses of the tile, w = m / largest
        double mv[kThetaRows / kThetaPhases];
        unsigned long long wv[kThetaRows / kThetaPhases];
#pragma unroll
        for (uint32_t k = 0; k < kThetaRows / kThetaPhases; ++k) {
            const uint32_t r = ph + kThetaPhases * k;
            wv[k] = (active && r < rows) ? src.word(i0 + r, j) : 0ull;
        }
#pragma unroll
        for (uint32_t k = 0; k < kThetaRows / kThetaPhases; ++k) {
            const uint32_t r = ph + kThetaPhases * k;
            mv[k] = (active && r < rows) ? src.mass(wv[k], i0 + r) : 0.0;
        }
#pragma unroll
        for (uint32_t k = 0; k < kThetaRows / kThetaPhases; ++k) sv[ph + kThetaPhases * k][c] = mv[k] / largest;
        __syncthreads();
        // the ordered prefix down each column: s += w, row by row
        if (ph == 0) {
            double s = s_carry[c];
            for (uint32_t r = 0; r < rows; ++r) {
                s += sv[r][c];
                sv[r][c] = s;
            }
            s_carry[c] = s;
        }
        __syncthreads();
        // normalized values
#pragma unroll
        for (uint32_t k = 0; k < kThetaRows / kThetaPhases; ++k) {
            const uint32_t r = ph + kThetaPhases * k;
            if (r < rows) {
                const uint32_t i = i0 + r;
                double next;
                if (i + 1 == nt)
                    next = 1.0;
                else
                    next = col > 0 ? sv[r][c] / col : s_ce[r + 1];
                sv[r][c] = next;
            }
        }
        __syncthreads();
        // audit of each cell (dphi)(dtheta) against m / Mc
        if (active) {
#pragma unroll
            for (uint32_t k = 0; k < kThetaRows / kThetaPhases; ++k) {
                const uint32_t r = ph + kThetaPhases * k;
                if (r < rows) {
                    const double next = sv[r][c];
                    const double prev = r ? sv[r - 1][c] : n_carry[tile & 1][c];
                    const double q = da * (next - prev);
                    const double p = mv[k] / Mc;
                    malformed += (!(q >= 0 && q <= DBL_MAX) || next < prev) ? 1 : 0;
                    const double err = fabs(p - q);
                    l1 += err;
                    maxcell = fmax(maxcell, err);
                    if (p > 0 && q == 0) {
                        lost += p;
                        lostcells += 1;
                    }
                    rec += q;
                    ax += q * (1 - s_ce[r] - s_ce[r + 1]);
                }
            }
        }
        if (ph == (rows - 1) % kThetaPhases) n_carry[(tile + 1) & 1][c] = sv[rows - 1][c];
        // transposed store: per instruction lanes 0-31 write 32 rows of one column and
        // lanes 32-63 of the next; wave ph owns columns 16 ph .. 16 ph + 15
        {
            const uint32_t lane = threadIdx.x & 63, r = lane & 31, half = lane >> 5;
#pragma unroll
            for (uint32_t cp = 0; cp < kThetaCols / (2 * kThetaPhases); ++cp) {
                const uint32_t cc = ph * (kThetaCols / kThetaPhases) + 2 * cp + half;
                const uint32_t jj = j0 + cc;
                if (jj < J.onp && r < rows) G(J.theta)[(uint64_t)jj * stride + i0 + r + 1] = sv[r][cc];
            }
        }
        __syncthreads();
    }
    // block reduction of the audit terms
    double v[8] = {l1, lost, rec, ax, maxcell, lostcells, malformed, 0};
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        for (int off = 32; off > 0; off >>= 1) {
            const double o = __shfl_xor(v[k], off, 64);
            v[k] = k == 4 ? fmax(v[k], o) : v[k] + o;
        }
    }
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int k = 0; k < 8; ++k) s_red[ph][k] = v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        auto p = G(J.partials) + (uint64_t)blockIdx.x * 8;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            double t = s_red[0][k];
            for (uint32_t w = 1; w < kThetaPhases; ++w) t = k == 4 ? fmax(t, s_red[w][k]) : t + s_red[w][k];
            p[k] = t;
        }
    }
}

extern "C" __global__ void __launch_bounds__(kRedThreads)
    icephase_cdf_audit_final(const CdfJob* __restrict__ jobs, uint32_t blocks)
{
    __shared__ double sh[kRedThreads / 64];
    const CdfJob J = jobs[blockIdx.x];
    Neumaier s[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
    double mx = 0, lc = 0, mal = 0;
    for (uint32_t b = threadIdx.x; b < blocks; b += blockDim.x) {
        const auto p = G(J.partials) + (uint64_t)b * 8;
        s[0].add(p[0]); s[1].add(p[1]); s[2].add(p[2]); s[3].add(p[3]);
        mx = fmax(mx, p[4]);
        lc += p[5];
        mal += p[6];
    }
    const double v0 = block_sum(s[0].value(), sh);
    const double v1 = block_sum(s[1].value(), sh);
    const double v2 = block_sum(s[2].value(), sh);
    const double v3 = block_sum(s[3].value(), sh);
    const double v4 = block_max(mx, sh);
    const double v5 = block_sum(lc, sh);
    const double v6 = block_sum(mal, sh);
    if (threadIdx.x == 0) {
        G(J.scalars)[kCdfL1] = v0;
        G(J.scalars)[kCdfLost] = v1;
        G(J.scalars)[kCdfRecovered] = v2;
        G(J.scalars)[kCdfAxialCdf] = v3;
        G(J.scalars)[kCdfMaxCell] = v4;
        G(J.scalars)[kCdfLostCells] = v5;
        G(J.scalars)[kCdfMalformed] = v6;
    }
}

extern "C" __global__ void __launch_bounds__(64) icephase_cdf_clear(const CdfJob* __restrict__ jobs)
{
    const CdfJob J = jobs[blockIdx.x];
    for (uint32_t k = threadIdx.x; k < kCdfScalars; k += blockDim.x) G(J.scalars)[k] = 0;
}

}  // namespace icephase

namespace icephase {

// The whole CDF sequence for njobs records on one stream. reduce_blocks is the
// x-extent of the reduction grids (partials hold 8 doubles per block).
hipError_t launch_cdf(const CdfJob* d_jobs, uint32_t njobs, uint32_t max_onp, bool any_coarse, uint32_t reduce_blocks,
                      hipStream_t s)
{
    const dim3 red(reduce_blocks, njobs);
    hipLaunchKernelGGL(icephase_cdf_clear, dim3(njobs), dim3(64), 0, s, d_jobs);
    if (any_coarse) hipLaunchKernelGGL(icephase_cdf_coarsen, red, dim3(kRedThreads), 0, s, d_jobs);
    hipLaunchKernelGGL(icephase_cdf_moments, red, dim3(kRedThreads), 0, s, d_jobs);
    hipLaunchKernelGGL(icephase_cdf_moments_final, dim3(njobs), dim3(kRedThreads), 0, s, d_jobs, reduce_blocks);
    if (any_coarse) {
        hipLaunchKernelGGL(icephase_cdf_tv, red, dim3(kRedThreads), 0, s, d_jobs);
        hipLaunchKernelGGL(icephase_cdf_tv_final, dim3(njobs), dim3(kRedThreads), 0, s, d_jobs, reduce_blocks);
    }
    hipLaunchKernelGGL(icephase_cdf_colsum, dim3((max_onp + kRedThreads - 1) / kRedThreads, njobs), dim3(kRedThreads),
                       0, s, d_jobs);
    hipLaunchKernelGGL(icephase_cdf_marginal, dim3(njobs), dim3(256), 0, s, d_jobs);
    const uint32_t col_blocks = (max_onp + kThetaCols - 1) / kThetaCols;
    hipLaunchKernelGGL(icephase_cdf_theta, dim3(col_blocks, njobs), dim3(kThetaThreads), 0, s, d_jobs);
    hipLaunchKernelGGL(icephase_cdf_audit_final, dim3(njobs), dim3(kRedThreads), 0, s, d_jobs, col_blocks);
    return hipGetLastError();
}

uint32_t cdf_partial_blocks(uint32_t reduce_blocks, uint32_t max_onp)
{
    const uint32_t col_blocks = (max_onp + kThetaCols - 1) / kThetaCols;
    return reduce_blocks > col_blocks ? reduce_blocks : col_blocks;
}

}  // namespace icephase
