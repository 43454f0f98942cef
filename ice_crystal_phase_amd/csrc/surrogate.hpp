// Learned surrogate: device-side layout shared by the kernels (surrogate.hip)
// and the C-ABI host code (surrogate_api.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace icephase {
namespace sur {

constexpr int kHidden = 256;
constexpr int kFourier = 16;     // phi features cos(2 pi k (v - 1/4)), k = 1..16
constexpr int kThetaFreqs = 13;  // theta frequencies f (kThetaFreq): sin, cos (pi f t)
// Input columns: [0] t, [1, 17) phi features, [17, 43) theta features, [43, 47)
// the record's condition.
constexpr int kColPhi = 1;
constexpr int kColTheta = kColPhi + kFourier;          // 17
constexpr int kColCond = kColTheta + 2 * kThetaFreqs;  // 43
constexpr int kInputs = kColCond + 4;                  // 47
constexpr int kStatic = kColCond;                      // 43 record-independent inputs
constexpr int kThetaIn = 1 + 2 * kThetaFreqs;          // inputs that depend on the theta row (27)
constexpr int kInPad = 64;                             // input columns padded for MFMA
constexpr double kThetaFreq[kThetaFreqs] = {1, 2, 3, 4, 6, 8, 12, 16, 24, 32, 48, 64, 96};

// fp32 parameters in device layout: W1 [256][64] (columns >= kInputs stay zero),
// b1, then (W [256][256], b [256]) for layers 2..4, w5 [256], b5.
constexpr uint32_t kOffW1 = 0;
constexpr uint32_t kOffB1 = 256 * 64;
constexpr uint32_t kOffW2 = kOffB1 + 256;
constexpr uint32_t kLayerStride = 65536 + 256;
constexpr uint32_t kOffW5 = kOffW2 + 3 * kLayerStride;
constexpr uint32_t kOffB5 = kOffW5 + 256;
constexpr uint32_t kParams = kOffB5 + 1;
constexpr uint32_t kParamsPad = (kParams + 255) / 256 * 256;
// packed (C-ABI) parameters: W1 [256][kInputs] without the padding
constexpr uint32_t kPacked = kParams - 256 * (kInPad - kInputs);

constexpr uint32_t kLseParts = 256;
constexpr uint32_t kCompareBlocks = 512;
constexpr uint32_t kReduceGroups = 32;
constexpr uint32_t kTileRows = 128;  // rows of a forward/backward workgroup tile

inline uint32_t off_w(int layer) { return layer == 0 ? kOffW1 : kOffW2 + (layer - 1) * kLayerStride; }
inline uint32_t off_b(int layer) { return layer == 0 ? kOffB1 : kOffW2 + (layer - 1) * kLayerStride + 65536; }

struct Net {
    const float* w1;      // fp32 [256][64] master weights of layer 1
    const float* pphi;    // [nq][256] layer-1 phi table (phi-feature part of W1 x)
    const float* ptheta;  // [nt][256] layer-1 theta table (t and theta-feature part of W1 x)
    const __bf16* w[3];   // layers 2..4, [out][in]
    const __bf16* wt[3];  // transposes, [in][out]
    const float* b[4];
    const float* w5;
    const float* b5;
};

// The drop records are mirror symmetric in phi: cell (i, j) and (i, mirror(j)),
// mirror(j) = (np/2 - 1 - j) mod np (v -> 1/2 - v), hold the same mass, and the
// phi features are even about v = 1/4, so the network is symmetric too. Kernels
// therefore run over the folded grid: rows r = i nq + q for the nq canonical
// columns canon[q] (canon[q] <= mirr[q]); a row carries the mass of both cells
// of its pair (one if canon[q] == mirr[q]).
struct Static {
    uint32_t nt, np, nq;
    uint32_t M;     // cells of the full grid, nt np
    uint32_t Mh;    // rows of the folded grid, nt nq
    uint32_t Mpad;  // Mh padded to a multiple of 128
    const uint32_t* canon;  // [nq] full-grid column of canonical column q
    const uint32_t* mirr;   // [nq] its mirror (== canon[q] for a self-paired column)
    const __bf16* xhi;  // [Mpad][64] static inputs of the folded rows, hi part
    const __bf16* xlo;  //                                            lo part
    const float* logw;  // [Mpad] log du_i + log(cells of the pair), -inf for padding rows
    const float* feat;   // [nq][16] phi features cos(2 pi k (v - 1/4)) of canonical columns (fp32)
    const float* tfeat;  // [nt][27] theta inputs t_i, sin(pi f t_i), cos(pi f t_i) (fp32)
    uint32_t rows_per_split, splits;
};

hipError_t forward(const Net& net, const Static& st, const float* d_conds, uint32_t nrec, bool train, float* logits,
                   __bf16* Z, hipStream_t s);
hipError_t lse(const Static& st, const float* logits, uint32_t nrec, float2* part, float* out, hipStream_t s);
// target: folded-grid probabilities [Mh]; the loss is the full-grid cross entropy
hipError_t dlogits(const Static& st, const float* logits, const float* lse, const float* target, float tvw,
                   float* tv_part, float* tv_G, float* dF, float* loss_part, float* loss, hipStream_t s);
// full-grid cell probabilities [nrec][M] (FP64 and/or FP32) from folded logits
hipError_t probs(const Static& st, const float* logits, const float* lse, uint32_t nrec, double* mass, float* p,
                 hipStream_t s);
// folded training target [Mh] from full-grid probabilities [M]
hipError_t fold(const Static& st, const float* full, float* folded, hipStream_t s);
// part_db: [4][Mpad/kTileRows][256] fp32 column sums of dZ (bias gradients).
// accumulate: add into the partial buffers (records of one step, in order)
// instead of overwriting them.
hipError_t backward(const Net& net, const Static& st, const float* dF, const __bf16* Z, __bf16* dZ, float* part_w5,
                    float* part_b5, float* part_db, bool accumulate, hipStream_t s);
// part: [splits][256][in] weight-gradient partials of one layer
hipError_t wgrad(const Static& st, int layer, const __bf16* dZl, const __bf16* Zprev, const float* d_cond, float* part,
                 bool accumulate, hipStream_t s);
// part2: kReduceGroups * n floats of scratch
hipError_t reduce(const float* part, uint32_t splits, uint32_t n, float* out, bool accumulate, float* part2,
                  hipStream_t s);
hipError_t adam(float* w, float* m, float* v, float* g, uint32_t n, float lr, float inv_batch, float c1, float c2,
                hipStream_t s);
// bf16 W, W^T of layers 2-4 and the layer-1 phi [nq][256] and theta [nt][256]
// tables from the fp32 parameters
hipError_t pack(const float* p, __bf16* w, __bf16* wt, const Static& st, float* pphi, float* ptheta, hipStream_t s);
// Partial sums of |p - q| and (p - q) * row_cos[row] over kCompareBlocks blocks.
hipError_t compare(const float* p, const float* q, uint32_t M, uint32_t np, const double* row_cos, double* part,
                   hipStream_t s);
// FP32 cell probabilities of a traced record (u64 bins + pole quanta); scratch: one u64.
hipError_t export_probs(const unsigned long long* bins, unsigned long long qf, unsigned long long qb, uint32_t nt,
                        uint32_t np, unsigned long long* scratch, float* out, hipStream_t s);

}  // namespace sur
}  // namespace icephase
