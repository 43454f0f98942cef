// Learned surrogate of the phase-function records: a conditional density MLP
// over the (u, phi) cells, BF16 MFMA (v_mfma_f32_16x16x32_bf16) with FP32
// accumulation, trained and served by the kernels in this file.
//
// Network (per cell of an nt x np grid, conditioned on one record):
//   x = [t, cos(2 pi k (v - 1/4)) (k = 1..16),
//        sin(pi f t), cos(pi f t) (f in kThetaFreq, 13), c0..c3]   (47 inputs)
//       t = (i + 1/2) / nt of the cell's theta row (the grid's uniform
//       coordinate of u), v = cell-centre (phi + pi) / 2 pi, c = normalized
//       (alpha, lambda, radius, aspect);
//   h1 = silu(W1 x + b1), h2..h4 = silu(W h + b)       (4 x 256)
//   f = w5 . h4 + b5,  p(cell) = softmax over cells of (f + log du_i)
// so f is a log density in u (cells carry du_i / np of the u-v square).
// The first layer is separable over the grid (the t inputs depend on the theta
// row, the v features on the phi column, c on the record) and runs in fp32 on
// the VALU: z1 = T[i] + (b1 + W1_c c) + P[j], with the theta table T = W1_t
// x_t(t_i) and the phi table P = W1_v x_v(v_j) rebuilt whenever the weights
// change, so neighbouring theta rows stay distinct. Hidden activations are
// bf16, every MFMA product accumulates in fp32. The weight gradient of layer 1
// splits x into bf16 hi + lo parts (two MFMA passes). The phi features are even
// about v = 1/4, so the network is mirror symmetric like the drop records, and
// every kernel runs over the folded grid (surrogate.hpp, Static).
//
// Training loss per record: cross entropy -sum P log p against the traced
// record's cell probabilities P. Kernels:
//   sur_forward     fused 4-layer forward per 128-row tile (activations in LDS),
//                   logits, and the pre-activations Z_l when training;
//   sur_lse_*       log-sum-exp per record (fixed-order two-level reduction);
//   sur_dlogits     dL/df = p - P (+ the weighted total-variation term, whose
//                   record sum sur_tv_partial forms first) and the loss (over
//                   the folded rows);
//   sur_probs/fold  folded rows <-> full-grid cells;
//   sur_backward    fused data gradients layers 4 -> 1 per tile (dZ_l to HBM);
//   sur_wgrad       weight gradients dW_l = dZ_l^T H_{l-1}, split over rows,
//                   partial tiles reduced in fixed order (deterministic);
//   sur_adam/pack   Adam on the fp32 master weights, bf16 copies for MFMA,
//                   and the layer-1 theta and phi tables.
#include "surrogate.hpp"

#include <hip/hip_runtime.h>

namespace icephase {
namespace sur {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Rows of the activation tile of one workgroup (4 waves x 64 output columns).
constexpr int TM = kTileRows;
constexpr int RB = TM / 16;
// forward tile rows and resident forward workgroups per CU (its LDS tile is kFwdRows x 512 bytes;
// 96-row tiles at 3 workgroups per CU measured 9% slower than 128 at 2)
constexpr int kFwdRows = 128;
constexpr int kFwdBlocks = 2;
static_assert(128 % kFwdRows == 0 && 128 % TM == 0, "Mpad, a multiple of 128, holds whole tiles");

// LDS index of element (row, col) of a TM x 256 bf16 tile: 8-element chunks
// are XOR-swizzled by row so the 16 rows an MFMA fragment reads fall in
// different banks (64 KB per tile).
__device__ inline int sw(int row, int col) { return row * 256 + (col ^ ((row & 7) << 3)); }

__device__ inline float bf2f(__bf16 b)
{
    return __builtin_bit_cast(float, (uint32_t)__builtin_bit_cast(uint16_t, b) << 16);
}
// round to nearest even (v_cvt_pk_bf16_f32)
__device__ inline __bf16 f2bf(float x) { return (__bf16)x; }
// the activation needs no IEEE division: exp and the hardware reciprocal
// (1 ulp) keep the layer epilogues short next to the MFMA work
__device__ inline float sigmoid(float z) { return __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }

__device__ inline float silu(float z) { return z * sigmoid(z); }

// four fp32 values as four bf16 in 8 bytes (two v_cvt_pk_bf16_f32)
__device__ inline uint2 pack4(float a, float b, float c, float d)
{
    const bf16x2 lo = {f2bf(a), f2bf(b)}, hi = {f2bf(c), f2bf(d)};
    return make_uint2(__builtin_bit_cast(uint32_t, lo), __builtin_bit_cast(uint32_t, hi));
}
__device__ inline float4 unpack4(uint2 v)
{
    return make_float4(__builtin_bit_cast(float, v.x << 16), __builtin_bit_cast(float, v.x & 0xffff0000u),
                       __builtin_bit_cast(float, v.y << 16), __builtin_bit_cast(float, v.y & 0xffff0000u));
}

// x plus the value of the lane CTRL selects within its 16-lane row (0 where
// that lane is outside the row)
template <int CTRL>
__device__ inline float add_dpp(float x)
{
    return x + __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), CTRL, 0xF, 0xF, true));
}
// sum over the 16 lanes of each row, in a fixed order; lane 15 of the row holds it
__device__ inline float row16_sum(float x)
{
    x = add_dpp<0x111>(x);  // row_shr:1
    x = add_dpp<0x112>(x);  // row_shr:2
    x = add_dpp<0x114>(x);  // row_shr:4
    return add_dpp<0x118>(x);  // row_shr:8
}

__device__ inline f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// This wave's 64 output columns of Z = H W^T over the TM-row tile H in LDS,
// computed transposed (W is the MFMA A operand) so that every lane holds four
// consecutive columns of one row:
//   acc[rb][cb][i] = Z[16 rb + (lane & 15)][n0 + 16 cb + 4 (lane >> 4) + i].
// W is a global [256][256] bf16 matrix, streamed in 32-wide k slices with the
// next slice loaded while the current one feeds the MFMAs.
template <int R = RB>
__device__ inline void slab_mma(const __bf16* Hs, const __bf16* __restrict__ W, int n0, f32x4 acc[R][4])
{
    const int lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4;
    const __bf16* wrow = W + (size_t)(n0 + r) * 256 + 8 * q;
    bf16x8 w[4];
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) w[cb] = *reinterpret_cast<const bf16x8*>(wrow + 16 * 256 * cb);
#pragma unroll 1
    for (int k0 = 0; k0 < 256; k0 += 32) {
        // next slice (the last one again at the end: in bounds, unused)
        const int kn = k0 + 32 < 256 ? k0 + 32 : k0;
        bf16x8 wn[4];
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) wn[cb] = *reinterpret_cast<const bf16x8*>(wrow + 16 * 256 * cb + kn);
#pragma unroll
        for (int rb = 0; rb < R; ++rb) {
            const bf16x8 h = *reinterpret_cast<const bf16x8*>(Hs + sw(16 * rb + r, k0 + 8 * q));
#pragma unroll
            for (int cb = 0; cb < 4; ++cb) acc[rb][cb] = mfma(w[cb], h, acc[rb][cb]);
        }
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) w[cb] = wn[cb];
    }
}

template <int R = RB>
__device__ inline void zero(f32x4 acc[R][4])
{
#pragma unroll
    for (int rb = 0; rb < R; ++rb)
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) acc[rb][cb] = f32x4{0, 0, 0, 0};
}

// Forward of one TMF-row tile of one record; grid = (Mpad / TMF, records).
// Layer 1 in fp32 (separable input), layers 2-4 bf16 MFMA with the
// activations in one LDS tile; the output layer is folded into the layer-4
// epilogue. TRAIN stores the pre-activations Z[l][row][n] (bf16) of record 0.
template <bool TRAIN, int TMF>
__global__ void __launch_bounds__(256, kFwdBlocks) sur_forward(Net net, Static st, const float* __restrict__ conds,
                                                   float* __restrict__ logits, __bf16* __restrict__ Z)
{
    __shared__ __align__(16) __bf16 Hs[TMF * 256];
    constexpr int RBF = TMF / 16;
    const uint32_t rec = blockIdx.y, row0 = blockIdx.x * TMF, Mpad = st.Mpad;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, q = lane >> 4, c16 = lane & 15;
    const int n0 = 64 * wave + 4 * q;  // first of the lane's columns in each 16-column block
    f32x4 acc[RBF][4];
    {
        // layer 1: z = (T[i] + (b1 + W1_c c)) + P[j]
        const float* c = conds + 4 * rec;
        const float c0 = c[0], c1 = c[1], c2 = c[2], c3 = c[3];
        float base[4][4];
#pragma unroll
        for (int cb = 0; cb < 4; ++cb)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int col = n0 + 16 * cb + i;
                const float* w = net.w1 + col * kInPad + kColCond;
                base[cb][i] = fmaf(w[3], c3, fmaf(w[2], c2, fmaf(w[1], c1, fmaf(w[0], c0, net.b[0][col]))));
            }
#pragma unroll
        for (int rb = 0; rb < RBF; ++rb) {
            const uint32_t row = row0 + 16 * rb + c16;
            uint32_t ti = row / st.nq;
            const uint32_t j = row - ti * st.nq;  // canonical column
            ti = ti < st.nt ? ti : st.nt - 1;     // padding rows (their logits are -inf)
            const float* T = net.ptheta + (size_t)ti * 256 + n0;
            const float* P = net.pphi + (size_t)j * 256 + n0;
#pragma unroll
            for (int cb = 0; cb < 4; ++cb) {
                const float4 tv = *reinterpret_cast<const float4*>(T + 16 * cb);
                const float4 p = *reinterpret_cast<const float4*>(P + 16 * cb);
                acc[rb][cb][0] = (tv.x + base[cb][0]) + p.x;
                acc[rb][cb][1] = (tv.y + base[cb][1]) + p.y;
                acc[rb][cb][2] = (tv.z + base[cb][2]) + p.z;
                acc[rb][cb][3] = (tv.w + base[cb][3]) + p.w;
            }
        }
    }
    for (int layer = 0;; ++layer) {
        if (layer > 0) {
            zero<RBF>(acc);
            slab_mma<RBF>(Hs, net.w[layer - 1], 64 * wave, acc);
            __syncthreads();  // every wave has read the input tile
            const float* bias = net.b[layer] + n0;
#pragma unroll
            for (int cb = 0; cb < 4; ++cb) {
                const float4 b = *reinterpret_cast<const float4*>(bias + 16 * cb);
#pragma unroll
                for (int rb = 0; rb < RBF; ++rb) {
                    acc[rb][cb][0] += b.x;
                    acc[rb][cb][1] += b.y;
                    acc[rb][cb][2] += b.z;
                    acc[rb][cb][3] += b.w;
                }
            }
        }
        if (TRAIN) {
#pragma unroll
            for (int rb = 0; rb < RBF; ++rb)
#pragma unroll
                for (int cb = 0; cb < 4; ++cb) {
                    const f32x4 z = acc[rb][cb];
                    *reinterpret_cast<uint2*>(Z + ((size_t)layer * Mpad + row0 + 16 * rb + c16) * 256 + n0 + 16 * cb) =
                        pack4(z[0], z[1], z[2], z[3]);
                }
        }
        if (layer == 3) break;
#pragma unroll
        for (int rb = 0; rb < RBF; ++rb)
#pragma unroll
            for (int cb = 0; cb < 4; ++cb) {
                const f32x4 z = acc[rb][cb];
                *reinterpret_cast<uint2*>(Hs + sw(16 * rb + c16, n0 + 16 * cb)) =
                    pack4(silu(z[0]), silu(z[1]), silu(z[2]), silu(z[3]));
            }
        __syncthreads();
    }
    // output layer: f = w5 . silu(z4) + b5 + log du. Lane partials over its 16
    // columns, then over the four lane groups, then over the four waves (the
    // LDS tile is free after the last barrier), all in a fixed order.
    float w5[4][4];
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) {
        const float4 w = *reinterpret_cast<const float4*>(net.w5 + n0 + 16 * cb);
        w5[cb][0] = w.x;
        w5[cb][1] = w.y;
        w5[cb][2] = w.z;
        w5[cb][3] = w.w;
    }
    float* part = reinterpret_cast<float*>(Hs);
#pragma unroll
    for (int rb = 0; rb < RBF; ++rb) {
        float a = 0;
#pragma unroll
        for (int cb = 0; cb < 4; ++cb)
#pragma unroll
            for (int i = 0; i < 4; ++i) a = fmaf(silu(acc[rb][cb][i]), w5[cb][i], a);
        a += __shfl_xor(a, 16, 64);
        a += __shfl_xor(a, 32, 64);
        if (q == 0) part[wave * TMF + 16 * rb + c16] = a;
    }
    __syncthreads();
    if (threadIdx.x < TMF) {
        const int row = threadIdx.x;
        const float f = (part[row] + part[TMF + row]) + (part[2 * TMF + row] + part[3 * TMF + row]);
        logits[(size_t)rec * Mpad + row0 + row] = (f + net.b5[0]) + st.logw[row0 + row];
    }
}

// log-sum-exp over the rows of each record: per-block (max, sum exp) partials,
// then one block per record combines them in order.
__global__ void __launch_bounds__(256) sur_lse_partial(const float* __restrict__ logits, uint32_t Mpad,
                                                       float2* __restrict__ part)
{
    __shared__ float2 sh[256];
    const uint32_t rec = blockIdx.y;
    const float* f = logits + (size_t)rec * Mpad;
    float m = -INFINITY, s = 0;
    for (uint32_t r = blockIdx.x * 256 + threadIdx.x; r < Mpad; r += gridDim.x * 256) {
        const float x = f[r];
        if (x > m) {
            s = s * __expf(m - x) + 1.0f;
            m = x;
        } else if (x > -INFINITY) {
            s += __expf(x - m);
        }
    }
    sh[threadIdx.x] = make_float2(m, s);
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            const float2 a = sh[threadIdx.x], b = sh[threadIdx.x + w];
            const float mm = fmaxf(a.x, b.x);
            const float ss = (a.x > -INFINITY ? a.y * __expf(a.x - mm) : 0.0f) + (b.x > -INFINITY ? b.y * __expf(b.x - mm) : 0.0f);
            sh[threadIdx.x] = make_float2(mm, ss);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(size_t)rec * gridDim.x + blockIdx.x] = sh[0];
}

// lane k combines parts k, k + 64, ...; the 64 lane sums are added in lane
// order (FP64), so the result does not depend on scheduling
__global__ void __launch_bounds__(64) sur_lse_final(const float2* __restrict__ part, uint32_t nparts,
                                                    float* __restrict__ lse)
{
    __shared__ double sh[64];
    const float2* p = part + (size_t)blockIdx.x * nparts;
    float m = -INFINITY;
    for (uint32_t k = threadIdx.x; k < nparts; k += 64) m = fmaxf(m, p[k].x);
#pragma unroll
    for (int w = 32; w > 0; w >>= 1) m = fmaxf(m, __shfl_xor(m, w, 64));
    double s = 0;
    for (uint32_t k = threadIdx.x; k < nparts; k += 64)
        if (p[k].x > -INFINITY) s += (double)p[k].y * (double)__expf(p[k].x - m);
    sh[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int k = 0; k < 64; ++k) t += sh[k];
        lse[blockIdx.x] = m + (float)log(t);
    }
}

// log of the number of full-grid cells a folded row stands for (1 or 2)
__device__ inline float log_mult(const Static& st, uint32_t q)
{
    return st.canon[q] == st.mirr[q] ? 0.0f : 0.69314718f;
}

// Half the sign of p - P: the subgradient of the total variation
// 1/2 sum |p - P| with respect to p.
__device__ inline float tv_sign(float p, float P) { return p > P ? 0.5f : (p < P ? -0.5f : 0.0f); }

// Total-variation term of the loss, first pass: per block, the sum over folded
// rows of g p with g = tv_sign(p, P); the record's sum G enters every row's
// gradient p (g - G). A folded row stands for a mirror pair of cells whose
// targets are equal (the tracer's records are mirror symmetric), so its p, P
// and g are the pair's, and these sums are the full grid's.
__global__ void __launch_bounds__(256) sur_tv_partial(Static st, const float* __restrict__ logits,
                                                      const float* __restrict__ lse, const float* __restrict__ target,
                                                      float* __restrict__ part)
{
    __shared__ float sh[256];
    const uint32_t r = blockIdx.x * 256 + threadIdx.x;
    float v = 0;
    if (r < st.Mpad) {
        const float f = logits[r];
        const float P = r < st.Mh ? target[r] : 0.0f;
        const float p = f > -INFINITY ? __expf(f - lse[0]) : 0.0f;
        v = tv_sign(p, P) * p;
    }
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// dL/df over the folded rows (record 0) of L = CE + tvw TV: p - P, plus
// tvw p (g - G) when tvw > 0 (G from sur_tv_partial, summed). Loss partials
// sum P (lse - f + log mult), the full-grid cross entropy -sum P_cell log p_cell.
__global__ void __launch_bounds__(256) sur_dlogits(Static st, const float* __restrict__ logits,
                                                   const float* __restrict__ lse, const float* __restrict__ target,
                                                   float tvw, const float* __restrict__ tvG, float* __restrict__ dF,
                                                   float* __restrict__ loss_part)
{
    __shared__ float sh[256];
    const uint32_t r = blockIdx.x * 256 + threadIdx.x;
    float l = 0;
    if (r < st.Mpad) {
        const float f = logits[r];
        const float P = r < st.Mh ? target[r] : 0.0f;
        const float p = f > -INFINITY ? __expf(f - lse[0]) : 0.0f;
        float d = p - P;
        if (tvw > 0.0f) d = fmaf(tvw * p, tv_sign(p, P) - tvG[0], d);
        dF[r] = d;
        if (P > 0) l = P * ((lse[0] - f) + log_mult(st, r % st.nq));
    }
    sh[threadIdx.x] = l;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss_part[blockIdx.x] = sh[0];
}

// full-grid cell probabilities of every record, theta-major: a folded row's
// p = exp(f - lse - log mult) goes to both cells of its pair. FP64 (CDF
// input) and FP32 copies.
__global__ void __launch_bounds__(256) sur_probs(Static st, const float* __restrict__ logits,
                                                 const float* __restrict__ lse, double* __restrict__ mass,
                                                 float* __restrict__ probs)
{
    const uint32_t rec = blockIdx.y;
    for (uint32_t r = blockIdx.x * 256 + threadIdx.x; r < st.Mh; r += gridDim.x * 256) {
        const uint32_t i = r / st.nq, q = r - i * st.nq;
        const float f = logits[(size_t)rec * st.Mpad + r];
        const float p = f > -INFINITY ? __expf((f - lse[rec]) - log_mult(st, q)) : 0.0f;
        const size_t c0 = (size_t)rec * st.M + (size_t)i * st.np + st.canon[q];
        const size_t c1 = (size_t)rec * st.M + (size_t)i * st.np + st.mirr[q];
        if (mass) {
            mass[c0] = (double)p;
            mass[c1] = (double)p;
        }
        if (probs) {
            probs[c0] = p;
            probs[c1] = p;
        }
    }
}

// folded training target: row (i, q) = P(i, canon q) + P(i, mirror q)
__global__ void __launch_bounds__(256) sur_fold(Static st, const float* __restrict__ full, float* __restrict__ folded)
{
    for (uint32_t r = blockIdx.x * 256 + threadIdx.x; r < st.Mh; r += gridDim.x * 256) {
        const uint32_t i = r / st.nq, q = r - i * st.nq;
        const float* row = full + (size_t)i * st.np;
        const uint32_t a = st.canon[q], b = st.mirr[q];
        folded[r] = a == b ? row[a] : row[a] + row[b];
    }
}

// silu'(z) = s (1 + z (1 - s)), s = sigmoid(z), as two FMAs
__device__ inline float silu_grad(float z)
{
    const float s = sigmoid(z);
    return fmaf(z, fmaf(-s, s, s), s);
}

// Data gradients of one TM-row tile, layers 4 -> 1: dZ_l to HBM; output
// layer partials (dw5 per column, db5) per tile.
__global__ void __launch_bounds__(256, 2) sur_backward(Net net, const float* __restrict__ dF,
                                                   const __bf16* __restrict__ Z, uint32_t Mpad,
                                                   __bf16* __restrict__ dZ, float* __restrict__ part_w5,
                                                   float* __restrict__ part_b5, float* __restrict__ part_db,
                                                   int accumulate)
{
    // part_db: [4 layers][tiles][256] column sums of the fp32 dZ (bias gradients)
    const uint32_t tiles = gridDim.x;
    __shared__ __align__(16) __bf16 Gs[TM * 256];
    const uint32_t row0 = blockIdx.x * TM;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, q = lane >> 4, c16 = lane & 15;
    const int n0 = 64 * wave;
    const int n0q = n0 + 4 * q;  // first of the lane's columns in each 16-column block
    {
        // layer 4: dZ4 = dF w5 silu'(Z4); partial dw5 = sum_rows dF silu(Z4),
        // db4 = sum_rows dZ4, db5 = sum_rows dF. Lanes hold four consecutive
        // columns of one row, as in the hidden-layer epilogues below.
        const __bf16* z4 = Z + ((size_t)3 * Mpad + row0) * 256;
        __bf16* dz4 = dZ + ((size_t)3 * Mpad + row0) * 256;
        float d[RB];
#pragma unroll
        for (int rb = 0; rb < RB; ++rb) d[rb] = dF[row0 + 16 * rb + c16];
        float* gw_out = part_w5 + (size_t)blockIdx.x * 256 + n0q;
        float* gz_out = part_db + ((size_t)3 * tiles + blockIdx.x) * 256 + n0q;
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) {
            const int col = n0q + 16 * cb;
            const float4 w5 = *reinterpret_cast<const float4*>(net.w5 + col);
            const float wv[4] = {w5.x, w5.y, w5.z, w5.w};
            float gw[4] = {0, 0, 0, 0}, gz[4] = {0, 0, 0, 0};
#pragma unroll
            for (int rb = 0; rb < RB; ++rb) {
                const int row = 16 * rb + c16;
                const float4 z = unpack4(*reinterpret_cast<const uint2*>(z4 + (size_t)row * 256 + col));
                const float zv[4] = {z.x, z.y, z.z, z.w};
                float dz[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const float sg = sigmoid(zv[i]);
                    gw[i] = fmaf(d[rb], zv[i] * sg, gw[i]);
                    dz[i] = (d[rb] * wv[i]) * fmaf(zv[i], fmaf(-sg, sg, sg), sg);
                    gz[i] += dz[i];
                }
                const uint2 g = pack4(dz[0], dz[1], dz[2], dz[3]);
                *reinterpret_cast<uint2*>(dz4 + (size_t)row * 256 + col) = g;
                *reinterpret_cast<uint2*>(Gs + sw(row, col)) = g;
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                gw[i] = row16_sum(gw[i]);
                gz[i] = row16_sum(gz[i]);
            }
            if (c16 == 15) {
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    gw_out[16 * cb + i] = accumulate ? gw_out[16 * cb + i] + gw[i] : gw[i];
                    gz_out[16 * cb + i] = accumulate ? gz_out[16 * cb + i] + gz[i] : gz[i];
                }
            }
        }
        if (threadIdx.x < 16) {
            float gb = 0;
#pragma unroll
            for (int rb = 0; rb < RB; ++rb) gb += d[rb];
            gb = row16_sum(gb);
            if (threadIdx.x == 15) part_b5[blockIdx.x] = accumulate ? part_b5[blockIdx.x] + gb : gb;
        }
    }
    __syncthreads();
    f32x4 acc[RB][4];
    for (int layer = 3; layer >= 1; --layer) {
        // dH_{layer-1} = dZ_layer W_layer  (W^T stored [in][out] is the A operand);
        // the pre-activations the epilogue needs are loaded before the MFMAs
        const __bf16* zp = Z + ((size_t)(layer - 1) * Mpad + row0) * 256;
        __bf16* dzp = dZ + ((size_t)(layer - 1) * Mpad + row0) * 256;
        uint2 zr[RB][4];
#pragma unroll
        for (int rb = 0; rb < RB; ++rb)
#pragma unroll
            for (int cb = 0; cb < 4; ++cb)
                zr[rb][cb] = *reinterpret_cast<const uint2*>(zp + (size_t)(16 * rb + c16) * 256 + n0q + 16 * cb);
        zero(acc);
        slab_mma(Gs, net.wt[layer - 1], n0, acc);
        __syncthreads();  // every wave has read dZ_layer
        float* db = part_db + ((size_t)(layer - 1) * tiles + blockIdx.x) * 256 + n0q;
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) {
            const int col = n0q + 16 * cb;
            float cs0 = 0, cs1 = 0, cs2 = 0, cs3 = 0;
#pragma unroll
            for (int rb = 0; rb < RB; ++rb) {
                const int row = 16 * rb + c16;
                const float4 z = unpack4(zr[rb][cb]);
                const float d0 = acc[rb][cb][0] * silu_grad(z.x), d1 = acc[rb][cb][1] * silu_grad(z.y);
                const float d2 = acc[rb][cb][2] * silu_grad(z.z), d3 = acc[rb][cb][3] * silu_grad(z.w);
                cs0 += d0;
                cs1 += d1;
                cs2 += d2;
                cs3 += d3;
                const uint2 g = pack4(d0, d1, d2, d3);
                *reinterpret_cast<uint2*>(dzp + (size_t)row * 256 + col) = g;
                *reinterpret_cast<uint2*>(Gs + sw(row, col)) = g;  // unused after layer 1
            }
            // column sums over the tile: the 16 lanes of a row hold the other rows
            cs0 = row16_sum(cs0);
            cs1 = row16_sum(cs1);
            cs2 = row16_sum(cs2);
            cs3 = row16_sum(cs3);
            if (c16 == 15) {
                db[16 * cb + 0] = accumulate ? db[16 * cb + 0] + cs0 : cs0;
                db[16 * cb + 1] = accumulate ? db[16 * cb + 1] + cs1 : cs1;
                db[16 * cb + 2] = accumulate ? db[16 * cb + 2] + cs2 : cs2;
                db[16 * cb + 3] = accumulate ? db[16 * cb + 3] + cs3 : cs3;
            }
        }
        __syncthreads();
    }
}

// Four rows x four bf16 columns (v[k] = row k) written column-major into an
// LDS tile T[col][row]: one 8-byte store per column.
__device__ inline void store_cols(__bf16* T, int stride, const uint2 v[4])
{
    uint2 c[4];
    c[0] = make_uint2((v[0].x & 0xffffu) | (v[1].x << 16), (v[2].x & 0xffffu) | (v[3].x << 16));
    c[1] = make_uint2((v[0].x >> 16) | (v[1].x & 0xffff0000u), (v[2].x >> 16) | (v[3].x & 0xffff0000u));
    c[2] = make_uint2((v[0].y & 0xffffu) | (v[1].y << 16), (v[2].y & 0xffffu) | (v[3].y << 16));
    c[3] = make_uint2((v[0].y >> 16) | (v[1].y & 0xffff0000u), (v[2].y >> 16) | (v[3].y & 0xffff0000u));
#pragma unroll
    for (int j = 0; j < 4; ++j) *reinterpret_cast<uint2*>(T + j * stride) = c[j];
}

// Weight gradients of one layer: dW[o][i] = sum_rows dZ[row][o] H[row][i] over
// the rows of split s, for all 256 outputs and IN inputs. H = silu(Zprev) for
// hidden layers; for the input layer H = x_hi + x_lo (two passes). Waves:
// (IN/64) x 4, each 64 outputs x 64 inputs. Rows go in chunks of 32 through
// two column-major LDS buffers (the MFMA reads 8 consecutive rows of one
// column): while the MFMAs consume one chunk, every thread loads its 4 x 4
// blocks of the next chunk's dZ and H and stages them (silu included) into the
// other buffer, so the staging work fills the MFMA pipeline's shadow.
template <int IN>
__global__ void __launch_bounds__(IN * 4) sur_wgrad(const __bf16* __restrict__ dZ, const __bf16* __restrict__ Zprev,
                                                    const __bf16* __restrict__ Xhi, const __bf16* __restrict__ Xlo,
                                                    const float* __restrict__ cond, uint32_t rows_per_split,
                                                    uint32_t Mpad, float* __restrict__ part, int accumulate)
{
    constexpr int T = IN * 4;             // threads
    constexpr int S = 40;                 // LDS stride (bf16) of a 32-row column
    constexpr int NH = IN == 64 ? 2 : 1;  // input layer: hi and lo tables
    constexpr int NT = 512 + 8 * (IN / 4) * NH;  // 4 x 4 blocks per chunk: dZ, then H
    constexpr int PER = NT / T;                  // blocks per thread
    static_assert(NT % T == 0, "whole blocks per thread");
    __shared__ __align__(16) __bf16 dZT[2][256 * S];
    __shared__ __align__(16) __bf16 HT[2][NH][IN * S];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, q = lane >> 4, c16 = lane & 15;
    const int o0 = 64 * (wave & 3), i0 = 64 * (wave >> 2);
    const uint32_t r_begin = blockIdx.x * rows_per_split;
    const uint32_t r_end = min(Mpad, r_begin + rows_per_split);
    f32x4 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = f32x4{0, 0, 0, 0};
    // block p of this thread: task = threadIdx.x + p T; rows 4 (task & 7) ..,
    // columns 4 (task >> 3) .. of dZ (task < 512) or of H
    uint2 v[PER][4], vn[PER][4];
    auto load = [&](uint2 (&v)[PER][4], uint32_t r0) {
#pragma unroll
        for (int p = 0; p < PER; ++p) {
            const int task = threadIdx.x + p * T, rg = task & 7;
            const bool is_dz = task < 512;
            const int c = 4 * ((is_dz ? task : task - 512) >> 3);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const size_t row = r0 + 4 * rg + k;
                const __bf16* src;
                if (is_dz)
                    src = dZ + row * 256 + c;
                else if (IN == 64)
                    src = (c < 64 ? Xhi : Xlo) + row * 64 + (c & 63);
                else
                    src = Zprev + row * 256 + c;
                v[p][k] = *reinterpret_cast<const uint2*>(src);
            }
        }
    };
    float ch[4], cl[4];  // condition inputs of the input layer, hi and lo
    if (IN == 64) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            ch[k] = bf2f(f2bf(cond[k]));
            cl[k] = cond[k] - ch[k];
        }
    }
    auto stage = [&](int buf) {
#pragma unroll
        for (int p = 0; p < PER; ++p) {
            const int task = threadIdx.x + p * T, rg = task & 7;
            if (task < 512) {
                store_cols(dZT[buf] + 4 * (task >> 3) * S + 4 * rg, S, v[p]);
            } else if (IN == 64) {
                const int c = 4 * ((task - 512) >> 3), table = c >> 6, cc = c & 63;
                uint2 w[4] = {v[p][0], v[p][1], v[p][2], v[p][3]};
                if (cc + 3 >= kColCond && cc < kColCond + 4) {
                    // substitute the record's condition for its four columns
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        float4 x = unpack4(w[k]);
                        float xs[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            const int col = cc + j;
                            if (col >= kColCond && col < kColCond + 4)
                                xs[j] = table ? cl[col - kColCond] : ch[col - kColCond];
                        }
                        w[k] = pack4(xs[0], xs[1], xs[2], xs[3]);
                    }
                }
                store_cols(&HT[buf][table][cc * S + 4 * rg], S, w);
            } else {
                const int c = 4 * ((task - 512) >> 3);
                float h[4][4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const float4 z = unpack4(v[p][k]);
                    h[k][0] = silu(z.x);
                    h[k][1] = silu(z.y);
                    h[k][2] = silu(z.z);
                    h[k][3] = silu(z.w);
                }
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    *reinterpret_cast<uint2*>(&HT[buf][0][(c + j) * S + 4 * rg]) =
                        pack4(h[0][j], h[1][j], h[2][j], h[3][j]);
            }
        }
    };
    // two chunks in flight: v holds chunk i + 1 while chunk i + 2 loads into vn
    // (past the split's end the last chunk is loaded and staged again, unused)
    auto clamp = [&](uint32_t r) { return r < r_end ? r : r_end - 32; };
    load(v, r_begin);
    stage(0);
    load(v, clamp(r_begin + 32));
    __syncthreads();
    int cur = 0;
    for (uint32_t r0 = r_begin; r0 < r_end; r0 += 32) {
        load(vn, clamp(r0 + 64));
        bf16x8 a[4];
#pragma unroll
        for (int ob = 0; ob < 4; ++ob)
            a[ob] = *reinterpret_cast<const bf16x8*>(dZT[cur] + (o0 + 16 * ob + c16) * S + 8 * q);
#pragma unroll
        for (int h = 0; h < NH; ++h)
#pragma unroll
            for (int ib = 0; ib < 4; ++ib) {
                const bf16x8 b = *reinterpret_cast<const bf16x8*>(&HT[cur][h][(i0 + 16 * ib + c16) * S + 8 * q]);
#pragma unroll
                for (int ob = 0; ob < 4; ++ob) acc[ob][ib] = mfma(a[ob], b, acc[ob][ib]);
            }
        stage(cur ^ 1);
        __syncthreads();
        cur ^= 1;
#pragma unroll
        for (int p = 0; p < PER; ++p)
#pragma unroll
            for (int k = 0; k < 4; ++k) v[p][k] = vn[p][k];
    }
    float* P = part + (size_t)blockIdx.x * 256 * IN;
#pragma unroll
    for (int ob = 0; ob < 4; ++ob)
#pragma unroll
        for (int ib = 0; ib < 4; ++ib)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float* d = P + (size_t)(o0 + 16 * ob + 4 * q + i) * IN + i0 + 16 * ib + c16;
                *d = accumulate ? *d + acc[ob][ib][i] : acc[ob][ib][i];
            }
}

// Two-pass fixed-order reduction over splits: out[e] (+)= sum_s part[s * n + e].
// Pass 1: block (x, g) sums splits [g C, (g+1) C) of elements [64 x, 64 x + 64)
// in four interleaved-free runs (thread group t sums its quarter in order, the
// quarters are combined in order); pass 2 adds the G group sums in order.
__global__ void __launch_bounds__(256) sur_reduce1(const float* __restrict__ part, uint32_t splits, uint32_t n,
                                                   uint32_t chunk, float* __restrict__ part2)
{
    __shared__ float sh[4][64];
    const uint32_t e = blockIdx.x * 64 + (threadIdx.x & 63), t = threadIdx.x >> 6;
    const uint32_t s0 = blockIdx.y * chunk, s1 = min(splits, s0 + chunk);
    const uint32_t q = (chunk + 3) / 4, a = min(s1, s0 + t * q), b = min(s1, s0 + (t + 1) * q);
    float s = 0;
    if (e < n) {
#pragma unroll 4
        for (uint32_t k = a; k < b; ++k) s += part[(size_t)k * n + e];
    }
    sh[t][threadIdx.x & 63] = s;
    __syncthreads();
    if (t == 0 && e < n)
        part2[(size_t)blockIdx.y * n + e] = ((sh[0][threadIdx.x] + sh[1][threadIdx.x]) + sh[2][threadIdx.x]) +
                                            sh[3][threadIdx.x];
}

__global__ void __launch_bounds__(256) sur_reduce2(const float* __restrict__ part2, uint32_t groups, uint32_t n,
                                                   float* __restrict__ out, int accumulate)
{
    const uint32_t e = blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    float s = 0;
    for (uint32_t k = 0; k < groups; ++k) s += part2[(size_t)k * n + e];
    out[e] = accumulate ? out[e] + s : s;
}

__global__ void __launch_bounds__(256) sur_adam(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                                                float* __restrict__ g, uint32_t n, float lr, float inv_batch, float c1,
                                                float c2)
{
    const uint32_t e = blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    const float gr = g[e] * inv_batch;
    const float mm = 0.9f * m[e] + 0.1f * gr;
    const float vv = 0.999f * v[e] + 0.001f * gr * gr;
    m[e] = mm;
    v[e] = vv;
    w[e] -= lr * (mm * c1) / (sqrtf(vv * c2) + 1e-8f);
    g[e] = 0.0f;
}

// bf16 operand copies (W and W^T) of the fp32 master weights of layers 2-4.
__global__ void __launch_bounds__(256) sur_pack(const float* __restrict__ p, __bf16* __restrict__ w,
                                                __bf16* __restrict__ wt)
{
    const uint32_t e = blockIdx.x * 256 + threadIdx.x;
    if (e < 3 * 65536) {
        const uint32_t l = e >> 16, k = e & 65535, o = k >> 8, i = k & 255;
        const __bf16 b = f2bf(p[kOffW2 + l * kLayerStride + k]);
        w[e] = b;
        wt[(size_t)l * 65536 + i * 256 + o] = b;
    }
}

// Layer-1 tables, fp32; block j, thread o:
//   phi table   P[q][o] = sum_k W1[o][kColPhi + k] feat[q][k]     (16 phi features of canonical column q)
//   theta table T[i][o] = W1[o][0] t_i + sum_k W1[o][kColTheta + k] tfeat[i][1 + k]
__global__ void __launch_bounds__(256) sur_phi_table(const float* __restrict__ p, const float* __restrict__ feat,
                                                     float* __restrict__ P)
{
    const uint32_t j = blockIdx.x, o = threadIdx.x;
    const float* w = p + kOffW1 + o * kInPad + kColPhi;
    const float* f = feat + (size_t)j * kFourier;
    float s = 0;
    for (int k = 0; k < kFourier; ++k) s = fmaf(w[k], f[k], s);
    P[(size_t)j * 256 + o] = s;
}

__global__ void __launch_bounds__(256) sur_theta_table(const float* __restrict__ p, const float* __restrict__ tfeat,
                                                       float* __restrict__ T)
{
    const uint32_t i = blockIdx.x, o = threadIdx.x;
    const float* w = p + kOffW1 + o * kInPad;
    const float* f = tfeat + (size_t)i * kThetaIn;
    float s = w[0] * f[0];
    for (int k = 0; k < 2 * kThetaFreqs; ++k) s = fmaf(w[kColTheta + k], f[1 + k], s);
    T[(size_t)i * 256 + o] = s;
}

// out[0] = sum of x[0..n) in FP64: 256 strided lane sums added in lane order
__global__ void __launch_bounds__(256) sur_sum(const float* __restrict__ x, uint32_t n, float* __restrict__ out)
{
    __shared__ double sh[256];
    double s = 0;
    for (uint32_t k = threadIdx.x; k < n; k += 256) s += x[k];
    sh[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int k = 0; k < 256; ++k) t += sh[k];
        out[0] = (float)t;
    }
}

__global__ void sur_clear_u64(unsigned long long* p)
{
    if (threadIdx.x == 0) *p = 0ull;
}

// Cell probabilities of a traced record: (bins + pole shares) / total quanta.
__global__ void __launch_bounds__(256) sur_export_total(const unsigned long long* __restrict__ bins, uint64_t n,
                                                        unsigned long long* __restrict__ total)
{
    __shared__ unsigned long long sh[256];
    unsigned long long s = 0;
    for (uint64_t k = (uint64_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (uint64_t)gridDim.x * 256) s += bins[k];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0 && sh[0]) atomicAdd(total, sh[0]);
}

__global__ void __launch_bounds__(256) sur_export_probs(const unsigned long long* __restrict__ bins,
                                                        unsigned long long qf, unsigned long long qb, uint32_t nt,
                                                        uint32_t np, const unsigned long long* __restrict__ total,
                                                        float* __restrict__ out)
{
    const double all = (double)(*total + qf + qb);
    const double f = (double)qf / np, b = (double)qb / np;
    const uint64_t n = (uint64_t)nt * np;
    for (uint64_t k = (uint64_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (uint64_t)gridDim.x * 256) {
        const uint64_t i = k / np;
        double m = (double)bins[k];
        if (i == 0) m += f;
        if (i + 1 == nt) m += b;
        out[k] = all > 0 ? (float)(m / all) : 0.0f;
    }
}

hipError_t export_probs(const unsigned long long* bins, unsigned long long qf, unsigned long long qb, uint32_t nt,
                        uint32_t np, unsigned long long* scratch, float* out, hipStream_t s)
{
    hipLaunchKernelGGL(sur_clear_u64, dim3(1), dim3(64), 0, s, scratch);
    hipLaunchKernelGGL(sur_export_total, dim3(512), dim3(256), 0, s, bins, (uint64_t)nt * np, scratch);
    hipLaunchKernelGGL(sur_export_probs, dim3(1024), dim3(256), 0, s, bins, qf, qb, nt, np, scratch, out);
    return hipGetLastError();
}

__global__ void __launch_bounds__(256) sur_compare(const float* __restrict__ p, const float* __restrict__ q,
                                                   uint32_t M, uint32_t np, const double* __restrict__ row_cos,
                                                   double* __restrict__ part)
{
    __shared__ double sa[256], sg[256];
    double a = 0, g = 0;
    for (uint32_t k = blockIdx.x * 256 + threadIdx.x; k < M; k += gridDim.x * 256) {
        const double d = (double)p[k] - (double)q[k];
        a += fabs(d);
        g += d * row_cos[k / np];
    }
    sa[threadIdx.x] = a;
    sg[threadIdx.x] = g;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            sa[threadIdx.x] += sa[threadIdx.x + w];
            sg[threadIdx.x] += sg[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = sa[0];
        part[2 * blockIdx.x + 1] = sg[0];
    }
}

hipError_t compare(const float* p, const float* q, uint32_t M, uint32_t np, const double* row_cos, double* part,
                   hipStream_t s)
{
    hipLaunchKernelGGL(sur_compare, dim3(kCompareBlocks), dim3(256), 0, s, p, q, M, np, row_cos, part);
    return hipGetLastError();
}

// ---------------------------------------------------------------- launchers
hipError_t forward(const Net& net, const Static& st, const float* d_conds, uint32_t nrec, bool train, float* logits,
                   __bf16* Z, hipStream_t s)
{
    const dim3 grid(st.Mpad / kFwdRows, nrec);
    if (train)
        hipLaunchKernelGGL((sur_forward<true, kFwdRows>), grid, dim3(256), 0, s, net, st, d_conds, logits, Z);
    else
        hipLaunchKernelGGL((sur_forward<false, kFwdRows>), grid, dim3(256), 0, s, net, st, d_conds, logits, Z);
    return hipGetLastError();
}

hipError_t lse(const Static& st, const float* logits, uint32_t nrec, float2* part, float* out, hipStream_t s)
{
    hipLaunchKernelGGL(sur_lse_partial, dim3(kLseParts, nrec), dim3(256), 0, s, logits, st.Mpad, part);
    hipLaunchKernelGGL(sur_lse_final, dim3(nrec), dim3(64), 0, s, part, (uint32_t)kLseParts, out);
    return hipGetLastError();
}

hipError_t dlogits(const Static& st, const float* logits, const float* lse_v, const float* target, float tvw,
                   float* tv_part, float* tv_G, float* dF, float* loss_part, float* loss, hipStream_t s)
{
    const uint32_t blocks = (st.Mpad + 255) / 256;
    if (tvw > 0.0f) {
        hipLaunchKernelGGL(sur_tv_partial, dim3(blocks), dim3(256), 0, s, st, logits, lse_v, target, tv_part);
        hipLaunchKernelGGL(sur_sum, dim3(1), dim3(256), 0, s, tv_part, blocks, tv_G);
    }
    hipLaunchKernelGGL(sur_dlogits, dim3(blocks), dim3(256), 0, s, st, logits, lse_v, target, tvw, tv_G, dF,
                       loss_part);
    hipLaunchKernelGGL(sur_sum, dim3(1), dim3(256), 0, s, loss_part, blocks, loss);
    return hipGetLastError();
}

hipError_t probs(const Static& st, const float* logits, const float* lse_v, uint32_t nrec, double* mass, float* p,
                 hipStream_t s)
{
    hipLaunchKernelGGL(sur_probs, dim3(256, nrec), dim3(256), 0, s, st, logits, lse_v, mass, p);
    return hipGetLastError();
}

hipError_t fold(const Static& st, const float* full, float* folded, hipStream_t s)
{
    hipLaunchKernelGGL(sur_fold, dim3(256), dim3(256), 0, s, st, full, folded);
    return hipGetLastError();
}

hipError_t backward(const Net& net, const Static& st, const float* dF, const __bf16* Z, __bf16* dZ, float* part_w5,
                    float* part_b5, float* part_db, bool accumulate, hipStream_t s)
{
    hipLaunchKernelGGL(sur_backward, dim3(st.Mpad / TM), dim3(256), 0, s, net, dF, Z, st.Mpad, dZ, part_w5, part_b5,
                       part_db, accumulate ? 1 : 0);
    return hipGetLastError();
}

hipError_t wgrad(const Static& st, int layer, const __bf16* dZl, const __bf16* Zprev, const float* d_cond,
                 float* part, bool accumulate, hipStream_t s)
{
    const int a = accumulate ? 1 : 0;
    if (layer == 0)
        hipLaunchKernelGGL(sur_wgrad<64>, dim3(st.splits), dim3(256), 0, s, dZl, nullptr, st.xhi, st.xlo, d_cond,
                           st.rows_per_split, st.Mpad, part, a);
    else
        hipLaunchKernelGGL(sur_wgrad<256>, dim3(st.splits), dim3(1024), 0, s, dZl, Zprev, nullptr, nullptr, nullptr,
                           st.rows_per_split, st.Mpad, part, a);
    return hipGetLastError();
}

hipError_t reduce(const float* part, uint32_t splits, uint32_t n, float* out, bool accumulate, float* part2,
                  hipStream_t s)
{
    const uint32_t groups = splits < kReduceGroups ? splits : kReduceGroups;
    const uint32_t chunk = (splits + groups - 1) / groups;
    hipLaunchKernelGGL(sur_reduce1, dim3((n + 63) / 64, groups), dim3(256), 0, s, part, splits, n, chunk, part2);
    hipLaunchKernelGGL(sur_reduce2, dim3((n + 255) / 256), dim3(256), 0, s, part2, groups, n, out, accumulate ? 1 : 0);
    return hipGetLastError();
}

hipError_t adam(float* w, float* m, float* v, float* g, uint32_t n, float lr, float inv_batch, float c1, float c2,
                hipStream_t s)
{
    hipLaunchKernelGGL(sur_adam, dim3((n + 255) / 256), dim3(256), 0, s, w, m, v, g, n, lr, inv_batch, c1, c2);
    return hipGetLastError();
}

hipError_t pack(const float* p, __bf16* w, __bf16* wt, const Static& st, float* pphi, float* ptheta, hipStream_t s)
{
    hipLaunchKernelGGL(sur_pack, dim3(3 * 65536 / 256), dim3(256), 0, s, p, w, wt);
    hipLaunchKernelGGL(sur_phi_table, dim3(st.nq), dim3(256), 0, s, p, st.feat, pphi);
    hipLaunchKernelGGL(sur_theta_table, dim3(st.nt), dim3(256), 0, s, p, st.tfeat, ptheta);
    return hipGetLastError();
}

}  // namespace sur
}  // namespace icephase
