// Learned surrogate of the phase-function records: a conditional density MLP
// over the (u, phi) cells, BF16 MFMA (v_mfma_f32_16x16x32_bf16) with FP32
// accumulation, trained and served by the kernels in this file.
//
// Network (per cell of an nt x np grid, conditioned on one record):
//   x = [t, sin(2 pi k v), cos(2 pi k v) (k = 1..16), c0..c3]   (37 inputs)
//       t = cell-centre theta / pi (the grid's uniform coordinate of u),
//       v = cell-centre (phi + pi) / 2 pi, c = normalized (alpha, lambda,
//       radius, aspect);
//   h1 = silu(W1 x + b1), h2..h4 = silu(W h + b)       (4 x 256)
//   f = w5 . h4 + b5,  p(cell) = softmax over cells of (f + log du_i)
// so f is a log density in u (cells carry du_i / np of the u-v square).
// The first layer splits both x and W1 into bf16 hi + lo parts (three MFMA
// passes), so neighbouring theta rows stay distinct; hidden activations are
// bf16, every product accumulates in fp32.
//
// Training loss per record: cross entropy -sum P log p against the traced
// record's cell probabilities P. Kernels:
//   sur_forward     fused 4-layer forward per 128-row tile (activations in LDS),
//                   logits, and the pre-activations Z_l when training;
//   sur_lse_*       log-sum-exp per record (fixed-order two-level reduction);
//   sur_dlogits     dL/df = p - P and the loss;
//   sur_backward    fused data gradients layers 4 -> 1 per tile (dZ_l to HBM);
//   sur_wgrad       weight gradients dW_l = dZ_l^T H_{l-1}, split over rows,
//                   partial tiles reduced in fixed order (deterministic);
//   sur_adam/pack   Adam on the fp32 master weights, bf16 copies for MFMA.
#include "surrogate.hpp"

#include <hip/hip_runtime.h>

namespace icephase {
namespace sur {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Rows of the activation tile of one workgroup (4 waves x 64 output columns).
constexpr int TM = 64;
constexpr int RB = TM / 16;

// LDS index of element (row, col) of a TM x 256 bf16 tile: 8-element chunks
// are XOR-swizzled by row so the 16 rows an MFMA fragment reads fall in
// different banks (32 KB per tile).
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

__device__ inline f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// One TM-row x 64-column output slab of C = A B^T for this wave: A is the
// LDS tile (columns [a_off, a_off + K)), B is a global
// [N][K] bf16 matrix (row stride K), output columns [n0, n0 + 64).
template <int K>
__device__ inline void slab_mma(const __bf16* As, int a_off, const __bf16* __restrict__ B, int n0, f32x4 acc[RB][4])
{
    const int lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4;
#pragma unroll 2
    for (int k0 = 0; k0 < K; k0 += 32) {
        bf16x8 b[4];
#pragma unroll
        for (int cb = 0; cb < 4; ++cb)
            b[cb] = *reinterpret_cast<const bf16x8*>(B + (size_t)(n0 + 16 * cb + r) * K + k0 + 8 * q);
#pragma unroll
        for (int rb = 0; rb < RB; ++rb) {
            const bf16x8 a = *reinterpret_cast<const bf16x8*>(As + sw(16 * rb + r, a_off + k0 + 8 * q));
#pragma unroll
            for (int cb = 0; cb < 4; ++cb) acc[rb][cb] = mfma(a, b[cb], acc[rb][cb]);
        }
    }
}

__device__ inline void zero(f32x4 acc[RB][4])
{
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) acc[rb][cb] = f32x4{0, 0, 0, 0};
}

// Static input features of tile rows into LDS: hi parts in columns [0, 64),
// lo parts in [64, 128); the record's 4 condition values in columns 33..36.
__device__ inline void load_inputs(__bf16* Hs, const __bf16* __restrict__ Xhi, const __bf16* __restrict__ Xlo,
                                   uint32_t row0, const float* cond)
{
    // TM rows x 64 columns x 2 tables of 16-byte vectors
    for (int v = threadIdx.x; v < TM * 16; v += 256) {
        const int table = v / (TM * 8), e = v % (TM * 8), row = e >> 3, c8 = (e & 7) * 8;
        const __bf16* src = (table ? Xlo : Xhi) + (size_t)(row0 + row) * 64 + c8;
        *reinterpret_cast<bf16x8*>(Hs + sw(row, table * 64 + c8)) = *reinterpret_cast<const bf16x8*>(src);
    }
    __syncthreads();
    if (threadIdx.x < TM) {
        const int row = threadIdx.x;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const __bf16 hi = f2bf(cond[c]);
            Hs[sw(row, 33 + c)] = hi;
            Hs[sw(row, 64 + 33 + c)] = f2bf(cond[c] - bf2f(hi));
        }
    }
    __syncthreads();
}

// Forward. grid = (Mpad / TM, records). TRAIN stores the pre-activations
// Z[l][row][n] (bf16) of record 0.
template <bool TRAIN>
__global__ void __launch_bounds__(256, 2) sur_forward(Net net, const __bf16* __restrict__ Xhi, const __bf16* __restrict__ Xlo,
                                                   const float* __restrict__ conds, const float* __restrict__ logw,
                                                   uint32_t Mpad, float* __restrict__ logits, __bf16* __restrict__ Z)
{
    __shared__ __align__(16) __bf16 Hs[TM * 256];
    const uint32_t rec = blockIdx.y, row0 = blockIdx.x * TM;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, q = lane >> 4, c16 = lane & 15;
    const int n0 = 64 * wave;
    load_inputs(Hs, Xhi, Xlo, row0, conds + 4 * rec);
    f32x4 acc[RB][4];
    for (int layer = 0; layer < 4; ++layer) {
        zero(acc);
        if (layer == 0) {
            slab_mma<64>(Hs, 0, net.w1hi, n0, acc);   // x_hi W1_hi
            slab_mma<64>(Hs, 64, net.w1hi, n0, acc);  // x_lo W1_hi
            slab_mma<64>(Hs, 0, net.w1lo, n0, acc);   // x_hi W1_lo
        } else {
            slab_mma<256>(Hs, 0, net.w[layer - 1], n0, acc);
        }
        __syncthreads();  // every wave has read the input tile
        const float* bias = net.b[layer];
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) {
            const int col = n0 + 16 * cb + c16;
            const float bc = bias[col];
#pragma unroll
            for (int rb = 0; rb < RB; ++rb)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int row = 16 * rb + 4 * q + i;
                    const float z = acc[rb][cb][i] + bc;
                    if (TRAIN) Z[((size_t)layer * Mpad + row0 + row) * 256 + col] = f2bf(z);
                    Hs[sw(row, col)] = f2bf(z * sigmoid(z));
                }
        }
        __syncthreads();
    }
    // output layer: four threads per row, 64 products each
    const int row = threadIdx.x >> 2, part = threadIdx.x & 3;
    float s = 0;
    for (int k = 0; k < 64; ++k) s += bf2f(Hs[sw(row, 64 * part + k)]) * net.w5[64 * part + k];
    s += __shfl_xor(s, 1, 64);
    s += __shfl_xor(s, 2, 64);
    if (part == 0) logits[(size_t)rec * Mpad + row0 + row] = s + net.b5[0] + logw[row0 + row];
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

__global__ void __launch_bounds__(64) sur_lse_final(const float2* __restrict__ part, uint32_t nparts,
                                                    float* __restrict__ lse)
{
    if (threadIdx.x != 0) return;
    const float2* p = part + (size_t)blockIdx.x * nparts;
    float m = -INFINITY;
    for (uint32_t k = 0; k < nparts; ++k) m = fmaxf(m, p[k].x);
    double s = 0;
    for (uint32_t k = 0; k < nparts; ++k)
        if (p[k].x > -INFINITY) s += (double)p[k].y * (double)__expf(p[k].x - m);
    lse[blockIdx.x] = m + (float)log(s);
}

// dL/df = p - P (record 0), loss partials sum P (lse - f).
__global__ void __launch_bounds__(256) sur_dlogits(const float* __restrict__ logits, const float* __restrict__ lse,
                                                   const float* __restrict__ target, uint32_t M, uint32_t Mpad,
                                                   float* __restrict__ dF, float* __restrict__ loss_part)
{
    __shared__ float sh[256];
    const uint32_t r = blockIdx.x * 256 + threadIdx.x;
    float l = 0;
    if (r < Mpad) {
        const float f = logits[r];
        const float P = r < M ? target[r] : 0.0f;
        const float p = f > -INFINITY ? __expf(f - lse[0]) : 0.0f;
        dF[r] = p - P;
        if (P > 0) l = P * (lse[0] - f);
    }
    sh[threadIdx.x] = l;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss_part[blockIdx.x] = sh[0];
}

// probabilities p = exp(f - lse) of every record, theta-major, FP64 (CDF input)
// and FP32 copies.
__global__ void __launch_bounds__(256) sur_probs(const float* __restrict__ logits, const float* __restrict__ lse,
                                                 uint32_t M, uint32_t Mpad, double* __restrict__ mass,
                                                 float* __restrict__ probs)
{
    const uint32_t rec = blockIdx.y;
    for (uint32_t r = blockIdx.x * 256 + threadIdx.x; r < M; r += gridDim.x * 256) {
        const float f = logits[(size_t)rec * Mpad + r];
        const float p = f > -INFINITY ? __expf(f - lse[rec]) : 0.0f;
        if (mass) mass[(size_t)rec * M + r] = (double)p;
        if (probs) probs[(size_t)rec * M + r] = p;
    }
}

__device__ inline float silu_grad(float z)
{
    const float s = sigmoid(z);
    return s * (1.0f + z * (1.0f - s));
}

// Data gradients of one TM-row tile, layers 4 -> 1: dZ_l to HBM; output
// layer partials (dw5 per column, db5) per tile.
__global__ void __launch_bounds__(256, 2) sur_backward(Net net, const float* __restrict__ dF,
                                                   const __bf16* __restrict__ Z, uint32_t Mpad,
                                                   __bf16* __restrict__ dZ, float* __restrict__ part_w5,
                                                   float* __restrict__ part_b5, float* __restrict__ part_db)
{
    // part_db: [4 layers][tiles][256] column sums of the fp32 dZ (bias gradients)
    const uint32_t tiles = gridDim.x;
    __shared__ __align__(16) __bf16 Gs[TM * 256];
    const uint32_t row0 = blockIdx.x * TM;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, q = lane >> 4, c16 = lane & 15;
    const int n0 = 64 * wave;
    {
        // thread = column n of layer 4
        const int n = threadIdx.x;
        const float w5 = net.w5[n];
        const __bf16* z4 = Z + ((size_t)3 * Mpad + row0) * 256;
        __bf16* dz4 = dZ + ((size_t)3 * Mpad + row0) * 256;
        float gw = 0, gb = 0, gz = 0;
        for (int row = 0; row < TM; ++row) {
            const float d = dF[row0 + row];
            const float z = bf2f(z4[(size_t)row * 256 + n]);
            const float s = sigmoid(z);
            gw += d * (z * s);
            gb += d;
            const float dz = d * w5 * (s * (1.0f + z * (1.0f - s)));
            gz += dz;
            const __bf16 g = f2bf(dz);
            dz4[(size_t)row * 256 + n] = g;
            Gs[sw(row, n)] = g;
        }
        part_w5[(size_t)blockIdx.x * 256 + n] = gw;
        part_db[((size_t)3 * tiles + blockIdx.x) * 256 + n] = gz;
        if (n == 0) part_b5[blockIdx.x] = gb;
    }
    __syncthreads();
    f32x4 acc[RB][4];
    for (int layer = 3; layer >= 1; --layer) {
        // dH_{layer-1} = dZ_layer W_layer  (B = W^T stored [in][out])
        zero(acc);
        slab_mma<256>(Gs, 0, net.wt[layer - 1], n0, acc);
        __syncthreads();
        const __bf16* zp = Z + ((size_t)(layer - 1) * Mpad + row0) * 256;
        __bf16* dzp = dZ + ((size_t)(layer - 1) * Mpad + row0) * 256;
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) {
            const int col = n0 + 16 * cb + c16;
            float colsum = 0;
#pragma unroll
            for (int rb = 0; rb < RB; ++rb)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int row = 16 * rb + 4 * q + i;
                    const float z = bf2f(zp[(size_t)row * 256 + col]);
                    const float dz = acc[rb][cb][i] * silu_grad(z);
                    colsum += dz;
                    const __bf16 g = f2bf(dz);
                    dzp[(size_t)row * 256 + col] = g;
                    Gs[sw(row, col)] = g;
                }
            // the four lane groups hold the other rows of this column
            colsum += __shfl_xor(colsum, 16, 64);
            colsum += __shfl_xor(colsum, 32, 64);
            if (q == 0) part_db[((size_t)(layer - 1) * tiles + blockIdx.x) * 256 + col] = colsum;
        }
        __syncthreads();
    }
}

// Weight gradients of one layer: dW[o][i] = sum_rows dZ[row][o] H[row][i] over
// the rows of split s, for all 256 outputs and IN inputs. H = silu(Zprev) for
// hidden layers; for the input layer H = x_hi + x_lo (two passes). Waves:
// (IN/64) x 4, each 64 outputs x 64 inputs.
template <int IN>
__global__ void __launch_bounds__(IN * 4) sur_wgrad(const __bf16* __restrict__ dZ, const __bf16* __restrict__ Zprev,
                                                    const __bf16* __restrict__ Xhi, const __bf16* __restrict__ Xlo,
                                                    const float* __restrict__ cond, uint32_t rows_per_split,
                                                    uint32_t Mpad, float* __restrict__ part)
{
    constexpr int T = IN * 4;          // threads
    constexpr int S = 40;              // LDS stride (bf16) of a 32-row transposed chunk
    constexpr int NH = IN == 64 ? 2 : 1;  // input layer: hi and lo tables
    __shared__ __align__(16) __bf16 dZT[256 * S];
    __shared__ __align__(16) __bf16 HT[NH][IN * S];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, q = lane >> 4, c16 = lane & 15;
    const int o0 = 64 * (wave & 3), i0 = 64 * (wave >> 2);
    const uint32_t r_begin = blockIdx.x * rows_per_split;
    const uint32_t r_end = min(Mpad, r_begin + rows_per_split);
    f32x4 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = f32x4{0, 0, 0, 0};
    for (uint32_t r0 = r_begin; r0 < r_end; r0 += 32) {
        // stage the dZ chunk (32 rows x 256) transposed
        for (int e = threadIdx.x; e < 32 * 32; e += T) {
            const int row = e >> 5, c8 = (e & 31) * 8;
            const bf16x8 v = *reinterpret_cast<const bf16x8*>(dZ + (size_t)(r0 + row) * 256 + c8);
#pragma unroll
            for (int j = 0; j < 8; ++j) dZT[(c8 + j) * S + row] = v[j];
        }
        // stage the H chunk (32 rows x IN) transposed
        for (int e = threadIdx.x; e < 32 * (IN / 8); e += T) {
            const int row = e / (IN / 8), c8 = (e % (IN / 8)) * 8;
            if (IN == 64) {
                bf16x8 hi = *reinterpret_cast<const bf16x8*>(Xhi + (size_t)(r0 + row) * 64 + c8);
                bf16x8 lo = *reinterpret_cast<const bf16x8*>(Xlo + (size_t)(r0 + row) * 64 + c8);
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const int c = c8 + j;
                    if (c >= 33 && c < 37) {
                        const __bf16 h = f2bf(cond[c - 33]);
                        hi[j] = h;
                        lo[j] = f2bf(cond[c - 33] - bf2f(h));
                    }
                    HT[0][c * S + row] = hi[j];
                    HT[NH - 1][c * S + row] = lo[j];
                }
            } else {
                const bf16x8 z = *reinterpret_cast<const bf16x8*>(Zprev + (size_t)(r0 + row) * 256 + c8);
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const float zz = bf2f(z[j]);
                    HT[0][(c8 + j) * S + row] = f2bf(zz * sigmoid(zz));
                }
            }
        }
        __syncthreads();
        bf16x8 a[4];
#pragma unroll
        for (int ob = 0; ob < 4; ++ob) a[ob] = *reinterpret_cast<const bf16x8*>(dZT + (o0 + 16 * ob + c16) * S + 8 * q);
#pragma unroll
        for (int h = 0; h < NH; ++h)
#pragma unroll
            for (int ib = 0; ib < 4; ++ib) {
                const bf16x8 b = *reinterpret_cast<const bf16x8*>(&HT[h][(i0 + 16 * ib + c16) * S + 8 * q]);
#pragma unroll
                for (int ob = 0; ob < 4; ++ob) acc[ob][ib] = mfma(a[ob], b, acc[ob][ib]);
            }
        __syncthreads();
    }
    float* P = part + (size_t)blockIdx.x * 256 * IN;
#pragma unroll
    for (int ob = 0; ob < 4; ++ob)
#pragma unroll
        for (int ib = 0; ib < 4; ++ib)
#pragma unroll
            for (int i = 0; i < 4; ++i) P[(size_t)(o0 + 16 * ob + 4 * q + i) * IN + i0 + 16 * ib + c16] = acc[ob][ib][i];
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

// bf16 operand copies of the fp32 master weights.
__global__ void __launch_bounds__(256) sur_pack(const float* __restrict__ p, __bf16* __restrict__ w1hi,
                                                __bf16* __restrict__ w1lo, __bf16* __restrict__ w, __bf16* __restrict__ wt)
{
    const uint32_t e = blockIdx.x * 256 + threadIdx.x;
    if (e < 256 * 64) {
        const float x = p[kOffW1 + e];
        const __bf16 h = f2bf(x);
        w1hi[e] = h;
        w1lo[e] = f2bf(x - bf2f(h));
    }
    if (e < 3 * 65536) {
        const uint32_t l = e >> 16, k = e & 65535, o = k >> 8, i = k & 255;
        const __bf16 b = f2bf(p[kOffW2 + l * (65536 + 256) + k]);
        w[e] = b;
        wt[(size_t)l * 65536 + i * 256 + o] = b;
    }
}

__global__ void __launch_bounds__(256) sur_sum(const float* __restrict__ x, uint32_t n, float* __restrict__ out)
{
    if (threadIdx.x || blockIdx.x) return;
    double s = 0;
    for (uint32_t k = 0; k < n; ++k) s += x[k];
    out[0] = (float)s;
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
    const dim3 grid(st.Mpad / TM, nrec);
    if (train)
        hipLaunchKernelGGL(sur_forward<true>, grid, dim3(256), 0, s, net, st.xhi, st.xlo, d_conds, st.logw, st.Mpad,
                           logits, Z);
    else
        hipLaunchKernelGGL(sur_forward<false>, grid, dim3(256), 0, s, net, st.xhi, st.xlo, d_conds, st.logw, st.Mpad,
                           logits, Z);
    return hipGetLastError();
}

hipError_t lse(const Static& st, const float* logits, uint32_t nrec, float2* part, float* out, hipStream_t s)
{
    hipLaunchKernelGGL(sur_lse_partial, dim3(kLseParts, nrec), dim3(256), 0, s, logits, st.Mpad, part);
    hipLaunchKernelGGL(sur_lse_final, dim3(nrec), dim3(64), 0, s, part, (uint32_t)kLseParts, out);
    return hipGetLastError();
}

hipError_t dlogits(const Static& st, const float* logits, const float* lse_v, const float* target, float* dF,
                   float* loss_part, float* loss, hipStream_t s)
{
    const uint32_t blocks = (st.Mpad + 255) / 256;
    hipLaunchKernelGGL(sur_dlogits, dim3(blocks), dim3(256), 0, s, logits, lse_v, target, st.M, st.Mpad, dF, loss_part);
    hipLaunchKernelGGL(sur_sum, dim3(1), dim3(64), 0, s, loss_part, blocks, loss);
    return hipGetLastError();
}

hipError_t probs(const Static& st, const float* logits, const float* lse_v, uint32_t nrec, double* mass, float* p,
                 hipStream_t s)
{
    hipLaunchKernelGGL(sur_probs, dim3(256, nrec), dim3(256), 0, s, logits, lse_v, st.M, st.Mpad, mass, p);
    return hipGetLastError();
}

hipError_t backward(const Net& net, const Static& st, const float* dF, const __bf16* Z, __bf16* dZ, float* part_w5,
                    float* part_b5, float* part_db, hipStream_t s)
{
    hipLaunchKernelGGL(sur_backward, dim3(st.Mpad / TM), dim3(256), 0, s, net, dF, Z, st.Mpad, dZ, part_w5, part_b5,
                       part_db);
    return hipGetLastError();
}

hipError_t wgrad(const Static& st, int layer, const __bf16* dZl, const __bf16* Zprev, const float* d_cond,
                 float* part, hipStream_t s)
{
    if (layer == 0)
        hipLaunchKernelGGL(sur_wgrad<64>, dim3(st.splits), dim3(256), 0, s, dZl, nullptr, st.xhi, st.xlo, d_cond,
                           st.rows_per_split, st.Mpad, part);
    else
        hipLaunchKernelGGL(sur_wgrad<256>, dim3(st.splits), dim3(1024), 0, s, dZl, Zprev, nullptr, nullptr, nullptr,
                           st.rows_per_split, st.Mpad, part);
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

hipError_t pack(const float* p, __bf16* w1hi, __bf16* w1lo, __bf16* w, __bf16* wt, hipStream_t s)
{
    hipLaunchKernelGGL(sur_pack, dim3(3 * 65536 / 256), dim3(256), 0, s, p, w1hi, w1lo, w, wt);
    return hipGetLastError();
}

}  // namespace sur
}  // namespace icephase
