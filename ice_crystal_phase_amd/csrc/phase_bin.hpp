// Direction -> (u, phi) cell binning in FP64 and the fixed-point quantizer,
// shared by every trace kernel. Device code (host-callable for the C-ABI's
// scalar helpers).
//
// phase_bin follows hex_phase/include/ice_crystal/phase_math.hpp:37-81: FP64
// frame projection, pole test s <= 256 eps, stable u from (mu, s), half-open u
// edges with the 64-eps / 1e-7-cell snap, phi = atan2(y, x) with the periodic
// seam and the same snap rule. Three changes of method keep the result and cut
// the cost on the GPU:
//   * a direction whose FP32 (theta, phi) lies well inside a cell (bin_fast)
//     takes that cell directly; only the ~4 % near cell edges run the FP64 rules
//     (bin_exact);
//   * the u row is guessed from u = sin^2(pi i / 2n) in FP32 and then corrected
//     against the FP64 edges, which returns the same row as the reference's
//     binary search for any u in [0,1];
//   * the phi column is taken from an FP32 atan2 when the FP32 column index is
//     more than PHI_MARGIN away from every integer (its error is below 1e-3
//     columns there, so floor() and the no-snap outcome are certain) and from
//     the FP64 atan2 otherwise.
#pragma once

#include <hip/hip_runtime.h>
#include <cfloat>
#include <cstdint>

namespace icephase {

constexpr double kPi = 3.141592653589793238462643383279502884;
constexpr double kSqrt3 = 1.732050807568877293527446341505872367;

struct d3 {
    double x, y, z;
};
__host__ __device__ inline d3 D3(double x, double y, double z) { return d3{x, y, z}; }
__host__ __device__ inline d3 operator+(d3 a, d3 b) { return D3(a.x + b.x, a.y + b.y, a.z + b.z); }
__host__ __device__ inline d3 operator-(d3 a, d3 b) { return D3(a.x - b.x, a.y - b.y, a.z - b.z); }
__host__ __device__ inline d3 operator-(d3 a) { return D3(-a.x, -a.y, -a.z); }
__host__ __device__ inline d3 operator*(d3 a, double s) { return D3(a.x * s, a.y * s, a.z * s); }
__host__ __device__ inline d3 operator/(d3 a, double s) { return D3(a.x / s, a.y / s, a.z / s); }
__host__ __device__ inline double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ inline d3 cross(d3 a, d3 b)
{
    return D3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__host__ __device__ inline bool finite64(double v) { return v >= -DBL_MAX && v <= DBL_MAX; }
__host__ __device__ inline bool finite64(d3 a) { return finite64(a.x) && finite64(a.y) && finite64(a.z); }
__host__ __device__ inline double gt_max(double a, double b) { return a > b ? a : b; }
__host__ __device__ inline double max_abs(d3 a) { return gt_max(fabs(a.x), gt_max(fabs(a.y), fabs(a.z))); }

// Max-abs scaled normalization (hex_trace/include/ice_crystal/math.hpp:38-47).
__host__ __device__ inline bool unit_of(d3 a, d3& out)
{
    const double s = max_abs(a);
    if (!(s > 0) || !finite64(s)) return false;
    a = a / s;
    const double n = sqrt(dot(a, a));
    if (!(n > 0) || !finite64(n)) return false;
    out = a / n;
    return finite64(out);
}
__host__ __device__ inline bool is_unit64(d3 a)
{
    return finite64(a) && fabs(dot(a, a) - 1.0) <= 128 * DBL_EPSILON;
}

// Binning frame of one record: (e0, e1, ki), right handed, FP64.
struct BinFrame {
    d3 ki, e0, e1;
};

struct CellHit {
    uint32_t index;  // i*np + j when kind == 0
    int32_t kind;    // 0 cell, +1 forward pole, -1 backward pole, 2 invalid
    uint32_t snapped;
};

constexpr double kPhiMargin = 2e-3;
constexpr float kCellMargin = 1e-2f;

constexpr int32_t kUndecided = 3;

// Fast cell decision: theta and phi in FP32 from the FP64 projections (x, y, z)
// rounded to FP32. That rounding is a relative error of 6e-8, so theta and phi
// carry at most ~1.2e-6 rad of error; when both fractional cell positions are at
// least kCellMargin (1e-2 cells = 3.5e-5 rad for theta, 3.5e-5 rad for phi at
// 1800 columns) from an edge, the FP64 rules (bin_exact) give this same cell with
// no snap and no pole. Returns that cell (kind 0) or kind kUndecided.
__host__ __device__ inline CellHit bin_fast(uint32_t nt, uint32_t np, double x, double y, double z)
{
    const float xf = (float)x, yf = (float)y, zf = (float)z;
    const float th = atan2f(sqrtf(xf * xf + yf * yf), zf);
    const float ri = th * ((float)nt * (float)(1.0 / kPi));
    const float ph = atan2f(yf, xf);
    const float ci = (ph + (float)kPi) * ((float)np * (float)(0.5 / kPi));
    const float rfl = floorf(ri), cfl = floorf(ci);
    const float rfr = ri - rfl, cfr = ci - cfl;
    if (rfr > kCellMargin && rfr < 1.0f - kCellMargin && cfr > kCellMargin && cfr < 1.0f - kCellMargin &&
        rfl >= 0.0f && rfl < (float)nt && cfl >= 0.0f && cfl < (float)np)
        return CellHit{(uint32_t)rfl * np + (uint32_t)cfl, 0, 0};
    return CellHit{0, kUndecided, 0};
}

// atan(t) for t in [0, 1] as t P(t^2), P of degree 7 in t^2: a minimax fit
// (iterated reweighted least squares on Chebyshev nodes), |error| < 1.5e-7 rad
// with FP32 Horner steps.
constexpr float kAtanPoly[8] = {0.9999993443489075f,  -0.33329859375953674f, 0.19946566224098206f,
                                -0.1390862911939621f, 0.0964219719171524f,   -0.0559123158454895f,
                                0.021862948313355446f, -0.00405456405133009f};

__device__ inline float atan01(float t)
{
    const float t2 = t * t;
    float p = kAtanPoly[7];
#pragma unroll
    for (int k = 6; k >= 0; --k) p = fmaf(p, t2, kAtanPoly[k]);
    return t * p;
}

// atan2(y, x) in [-pi, pi] to within 7e-7 rad: octant reduction, the hardware
// reciprocal (1 ulp), atan01. NaN when x = y = 0 or an input is NaN.
__device__ inline float atan2_approx(float y, float x)
{
    const float ax = fabsf(x), ay = fabsf(y);
    const float t = fminf(ax, ay) * __builtin_amdgcn_rcpf(fmaxf(ax, ay));
    float a = atan01(t);
    a = ay > ax ? 1.57079637f - a : a;
    a = x < 0.0f ? 3.14159274f - a : a;
    return copysignf(a, y);
}

// bin_fast on the device with cheaper arithmetic and the same guarantee: the
// projections by FP64 FMA dot products rounded to FP32 (relative error 6e-8),
// rho by the hardware square root (1 ulp), theta = atan2(rho, z) and phi =
// atan2(y, x) by atan2_approx. Together theta and phi are within 1.2e-6 rad, so
// the fractional cell positions are within 5e-7 n cells (n = nt or np) of the
// exact ones: below half the kCellMargin of 1e-2 cells up to 8192 rows and
// columns (larger grids always take the FP64 rules). Inside the margins the FP64
// rules give this same cell with no snap and no pole.
__device__ inline CellHit bin_fast_approx(uint32_t nt, uint32_t np, const BinFrame& g, d3 d)
{
    if (nt > 8192 || np > 8192) return CellHit{0, kUndecided, 0};
    const float x = (float)fma(g.e0.z, d.z, fma(g.e0.y, d.y, g.e0.x * d.x));
    const float y = (float)fma(g.e1.z, d.z, fma(g.e1.y, d.y, g.e1.x * d.x));
    const float z = (float)fma(g.ki.z, d.z, fma(g.ki.y, d.y, g.ki.x * d.x));
    const float rho = __builtin_amdgcn_sqrtf(fmaf(y, y, x * x));
    const float ri = atan2_approx(rho, z) * ((float)nt * (float)(1.0 / kPi));
    const float ci = (atan2_approx(y, x) + (float)kPi) * ((float)np * (float)(0.5 / kPi));
    const float rfl = floorf(ri), cfl = floorf(ci);
    const float rfr = ri - rfl, cfr = ci - cfl;
    if (rfr > kCellMargin && rfr < 1.0f - kCellMargin && cfr > kCellMargin && cfr < 1.0f - kCellMargin &&
        rfl >= 0.0f && rfl < (float)nt && cfl >= 0.0f && cfl < (float)np)
        return CellHit{(uint32_t)rfl * np + (uint32_t)cfl, 0, 0};
    return CellHit{0, kUndecided, 0};
}

// The phase_bin rules after the unit-length checks, in FP64: (x, y, z) are the
// frame projections of the direction, n their norm. edge(k) returns the FP64 u
// edge k.
template <class EdgeLoad>
__host__ __device__ inline CellHit bin_exact(uint32_t nt, uint32_t np, double x, double y, double z, double n,
                                             EdgeLoad edge)
{
    CellHit h{0, 2, 0};
    const double mu = z / n, s = hypot(x, y) / n;
    if (s <= 256 * DBL_EPSILON) {
        h.kind = mu >= 0 ? 1 : -1;
        return h;
    }
    const double u = mu >= 0 ? (0.5 * s * s) / (1 + mu) : 1 - (0.5 * s * s) / (1 - mu);
    if (!(u >= 0 && u <= 1)) return h;
    // Row: FP32 guess from u = sin^2(pi i/2n), corrected against the FP64 edges.
    const float uf = (float)u;
    int lo = (int)(asinf(sqrtf(uf)) * (float)(2.0 / kPi) * (float)nt);
    lo = lo < 0 ? 0 : (lo > (int)nt - 1 ? (int)nt - 1 : lo);
    while (lo > 0 && u < edge(lo)) --lo;
    while (lo + 1 < (int)nt && u >= edge(lo + 1)) ++lo;
    uint32_t snapped = 0;
    for (int cand = 0; cand < 2; ++cand) {
        const int e = lo + cand;
        if (e == 0 || e >= (int)nt) continue;
        const double ee = edge(e);
        const double width = fmin(ee - edge(e - 1), edge(e + 1) - ee);
        const double tol = fmin(64 * DBL_EPSILON * fmax(fabs(u), fabs(ee)), 1e-7 * width);
        if (fabs(u - ee) <= tol) {
            snapped = (u != ee);
            lo = e;
            break;
        }
    }
    // Column.
    uint32_t j;
    {
        const float phif = atan2f((float)y, (float)x);
        const float idxf = (phif + (float)kPi) * (float)(1.0 / (2.0 * kPi)) * (float)np;
        const float fl = floorf(idxf);
        const float fr = idxf - fl;
        if (fr > (float)kPhiMargin && fr < 1.0f - (float)kPhiMargin && fl >= 0.0f && fl < (float)np) {
            j = (uint32_t)fl;
        } else {
            double phi = atan2(y, x);
            if (phi >= kPi) phi -= 2 * kPi;
            const double v = (phi + kPi) / (2 * kPi);
            double index = v * np;
            const double nearest = floor(index + 0.5);
            if (fabs(index - nearest) <= fmin(64 * DBL_EPSILON * np, 1e-7)) {
                snapped = snapped || (index != nearest);
                index = nearest;
            }
            j = (uint32_t)index;
            if (j == np) j = 0;
            if (j >= np) return h;
        }
    }
    h.index = (uint32_t)lo * np + j;
    h.kind = 0;
    h.snapped = snapped;
    return h;
}

// The fast decision, then the FP64 rules where it leaves the cell undecided.
template <class EdgeLoad>
__host__ __device__ inline CellHit bin_core(uint32_t nt, uint32_t np, double x, double y, double z, double n,
                                            EdgeLoad edge)
{
    const CellHit f = bin_fast(nt, np, x, y, z);
    return f.kind == 0 ? f : bin_exact(nt, np, x, y, z, n, edge);
}

// phase_bin of a unit FP64 direction (hex records): the reference's unit checks first.
template <class EdgeLoad>
__host__ __device__ inline CellHit bin_direction(const BinFrame& g, uint32_t nt, uint32_t np, d3 d, EdgeLoad edge)
{
    CellHit h{0, 2, 0};
    if (!is_unit64(d)) return h;
    const double x = dot(g.e0, d), y = dot(g.e1, d), z = dot(g.ki, d);
    const double n = sqrt(x * x + y * y + z * z);
    if (!(n > 0) || !finite64(n) || fabs(n - 1) > 512 * DBL_EPSILON) return h;
    return bin_core(nt, np, x, y, z, n, edge);
}

// Drop records: the same rules applied to d/|d| of an FP32 direction promoted to
// FP64 (mu and s are ratios to n, so no separate normalization is needed).
template <class EdgeLoad>
__host__ __device__ inline CellHit bin_raw(const BinFrame& g, uint32_t nt, uint32_t np, d3 d, EdgeLoad edge)
{
    CellHit h{0, 2, 0};
    if (!finite64(d)) return h;
    const double x = dot(g.e0, d), y = dot(g.e1, d), z = dot(g.ki, d);
    const double n = sqrt(x * x + y * y + z * z);
    if (!(n > 0) || !finite64(n)) return h;
    return bin_core(nt, np, x, y, z, n, edge);
}

// Fixed-point quantum: rint(p * 2^qbits) for a power fraction p in [0, 1].
__host__ __device__ inline uint64_t quantize(double p, uint32_t qbits)
{
    return (uint64_t)rint(ldexp(p, (int)qbits));
}

// Q = 62 - ceil(log2(rays)), clamped to [8, 52]: the sum over all rays of a record
// (each ray's emitted quanta total at most 2^Q plus rounding) stays below 2^63.
__host__ __device__ inline uint32_t quantum_bits(uint64_t rays)
{
    uint32_t lg = 0;
    while (lg < 64 && (uint64_t(1) << lg) < rays) ++lg;
    int q = 62 - (int)lg;
    q = q > 52 ? 52 : q;
    q = q < 8 ? 8 : q;
    return (uint32_t)q;
}

}  // namespace icephase
