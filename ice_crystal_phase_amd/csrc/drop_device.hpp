// Axisymmetric water-drop tracer, FP32, device side.
//
// Geometry: r(mu) = 1 + c0 + c1 mu + sum_{n>=2} c_n T_n(mu), mu = -y/r, in units
// of the equal-volume radius (include/rainbow/raindrop_shape.hpp:70-100). The
// reference intersects this surface with an interval-bracketed Newton solver
// inside an OptiX custom primitive (raindrop_intersection.hpp:116-382); this
// kernel instead uses what convexity gives for free, with no BVH:
//   * sphere: closed-form entry from the grid coordinates, closed-form chord;
//   * drop: the inside of a convex body along a line is one interval, so the
//     entry root is bracketed by the bounding-sphere entry and any interior
//     point (the foot of the perpendicular, else the minimum of f along the
//     line found by a secant search on f'), and the exit root by a small self
//     offset and the bounding-sphere exit; each bracket is solved by Newton
//     with bisection fallback. The per-ray control flow (entry, encounters,
//     emission) is in trace.hip.
// Splitting is per ray, FP32, two-column Jones (the Mueller M11 row for
// unpolarized input), flux-normalized (t_flux = sqrt(nt ct / ni ci) t_E, as in
// hex_trace/include/ice_crystal/flux_interface.hpp:26-88). Emission follows the
// linear chain of hex_trace/include/ice_crystal/hex_trace_core.hpp:101-163:
// external R at entry, T at every internal encounter, continue with R.
//
// Every operation is one correctly rounded IEEE single-precision operation:
// +, -, *, / and sqrt, and fused multiply-adds (fmaf, one rounding) written out
// at fixed places (dot and cross products, the Chebyshev recurrence, points on
// a ray, the Fresnel and Jones algebra). The build has -ffp-contract=off, so
// the compiler fuses nothing else, and the CPU mirror in the oracle, which
// calls fmaf at the same places, reproduces each emitted direction and power
// bitwise.
#pragma once

#include "phase_bin.hpp"

namespace icephase {
namespace drop {

struct f3 {
    float x, y, z;
};
__device__ inline f3 F3(float x, float y, float z) { return f3{x, y, z}; }
__device__ inline f3 operator+(f3 a, f3 b) { return F3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ inline f3 operator-(f3 a, f3 b) { return F3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ inline f3 operator*(f3 a, float s) { return F3(a.x * s, a.y * s, a.z * s); }
__device__ inline f3 operator/(f3 a, float s) { return F3(a.x / s, a.y / s, a.z / s); }
// x-term product first, then the y and z terms fused in
__device__ inline float dot(f3 a, f3 b) { return fmaf(a.z, b.z, fmaf(a.y, b.y, a.x * b.x)); }
__device__ inline f3 cross(f3 a, f3 b)
{
    return F3(fmaf(a.y, b.z, -(a.z * b.y)), fmaf(a.z, b.x, -(a.x * b.z)), fmaf(a.x, b.y, -(a.y * b.x)));
}
// o + d t, one rounding per component
__device__ inline f3 axpy(f3 d, float t, f3 o) { return F3(fmaf(d.x, t, o.x), fmaf(d.y, t, o.y), fmaf(d.z, t, o.z)); }
__device__ inline f3 unit(f3 a) { return a * (1.0f / sqrtf(dot(a, a))); }

struct Shape {
    float c[8];
    float inner, outer;
    float rmax;  // bounding radius of the intersector (max r(chi), rounded up to 2^-20)
    uint32_t sphere;
    uint32_t terms;  // 1 + index of the last nonzero coefficient, at least 2
};

// r(mu) and r'(mu) by the Chebyshev recurrence, stopped after the last nonzero
// coefficient: the skipped terms would add c_n T_n = +-0 to finite sums, which
// leaves every value unchanged. `terms` is the same for every lane of a wave
// (one record per tile), so the early exit is a scalar branch.
__device__ inline float cheb_radius(const float* c, uint32_t terms, float mu, float& der)
{
    float prev = 1.0f, cur = mu, dprev = 0.0f, dcur = 1.0f;
    const float mu2 = 2.0f * mu;
    float rad = fmaf(c[1], mu, 1.0f + c[0]);
    float d = c[1];
#pragma unroll
    for (int n = 2; n < 8; ++n) {
        if ((uint32_t)n >= terms) break;
        const float next = fmaf(mu2, cur, -prev);                 // T_{n}
        const float dnext = fmaf(mu2, dcur, 2.0f * cur) - dprev;  // T'_{n}
        rad = fmaf(c[n], next, rad);
        d = fmaf(c[n], dnext, d);
        prev = cur;
        cur = next;
        dprev = dcur;
        dcur = dnext;
    }
    der = d;
    return rad;
}

// Outward (unnormalized) gradient u - (g - u mu) k, g = (0,-1,0), k = R'(mu)/r:
// components u_i + (u_i mu) k and u_y + (1 + u_y mu) k.
__device__ inline f3 outward(f3 u, float mu, float k)
{
    return F3(fmaf(u.x * mu, k, u.x), fmaf(fmaf(u.y, mu, 1.0f), k, u.y), fmaf(u.z * mu, k, u.z));
}

__device__ inline float surface(const Shape& s, f3 p)
{
    const float r = sqrtf(dot(p, p));
    if (s.sphere) return r - 1.0f;
    if (r < s.inner * 0.25f) return r - s.inner;
    const float mu = -(p.y * (1.0f / r));
    float der;
    return r - cheb_radius(s.c, s.terms, mu, der);
}

__device__ inline f3 gradient(const Shape& s, f3 p)
{
    const float r = sqrtf(dot(p, p));
    const float ir = 1.0f / r;
    const f3 u = p * ir;
    if (s.sphere) return u;
    const float mu = -u.y;
    float der;
    (void)cheb_radius(s.c, s.terms, mu, der);
    return outward(u, mu, der * ir);
}

// surface(s, p) and dot(gradient(s, p), d) from one evaluation of r, mu and the
// Chebyshev recurrence: the same IEEE operations, so the same values.
__device__ inline float surface_slope(const Shape& s, f3 p, f3 d, float& slope)
{
    const float r = sqrtf(dot(p, p));
    const float ir = 1.0f / r;
    const f3 u = p * ir;
    if (s.sphere) {
        slope = dot(u, d);
        return r - 1.0f;
    }
    const float mu = -u.y;
    float der;
    const float rad = cheb_radius(s.c, s.terms, mu, der);
    slope = dot(outward(u, mu, der * ir), d);
    return r < s.inner * 0.25f ? r - s.inner : r - rad;
}

// Newton with bisection fallback on f(o + t d), t in [a,b]; f(a) > 0 iff a_pos.
// A Newton step below the tolerance ends the solve whether or not its
// candidate lies strictly inside the bracket: once x has converged, rounding
// can leave a bracket end on the root, and a candidate on or past it must not
// send the solve back to bisecting the rest of the bracket (that cost up to 20
// iterations on a few percent of exits, and a wave waits for its slowest lane).
// Two Newton steps in a row that contract quadratically predict the next one:
// about K step^2 with K = step / prev^2. When that prediction is below an
// eighth of the tolerance the candidate is accepted without the confirming
// evaluation (on the headline oblate drops this saves about one iteration per
// solve; directions stay as close to an FP64 trace as with the confirmation).
__device__ inline float newton(const Shape& s, f3 o, f3 d, float a, float b, bool a_pos, float x)
{
    const float tol = 0x1p-21f;
    float prev = 0.0f;  // length of the previous Newton step, 0 after a bisection
    if (!(x > a && x < b)) x = fmaf(b - a, 0.5f, a);
    for (int it = 0; it < 48; ++it) {
        const f3 p = axpy(d, x, o);
        float slope;
        const float fx = surface_slope(s, p, d, slope);
        if (fx == 0.0f) return x;
        if ((fx > 0.0f) == a_pos)
            a = x;
        else
            b = x;
        const float scale = fabsf(x) > 1.0f ? fabsf(x) : 1.0f;
        if (b - a <= tol * scale) return x;
        const float cand = x - fx / slope;
        const bool inside = cand > a && cand < b;
        const float step = fabsf(cand - x);
        if (step <= tol * scale) return inside ? cand : x;
        if (inside && prev > 0.0f && step <= prev * 0.25f &&
            step * step * step <= 0.125f * tol * scale * prev * prev)
            return cand;
        prev = inside ? step : 0.0f;
        x = inside ? cand : fmaf(b - a, 0.5f, a);
    }
    return x;
}

struct c32 {
    float re, im;
};
__device__ inline c32 cmul(c32 a, c32 b)
{
    return c32{fmaf(a.re, b.re, -(a.im * b.im)), fmaf(a.re, b.im, a.im * b.re)};
}
__device__ inline c32 cscale(c32 a, float b) { return c32{a.re * b, a.im * b}; }
// x a + y b
__device__ inline c32 clin(c32 x, float a, c32 y, float b)
{
    return c32{fmaf(y.re, b, x.re * a), fmaf(y.im, b, x.im * a)};
}
__device__ inline float cn2(c32 a) { return fmaf(a.im, a.im, a.re * a.re); }

struct Jones {
    c32 c0x, c0y, c1x, c1y;
    __device__ float power() const { return 0.5f * (((cn2(c0x) + cn2(c0y)) + cn2(c1x)) + cn2(c1y)); }
};

struct Ray {
    f3 dir, e0, e1;
    Jones j;
};
struct Split {
    Ray r, t;
    bool ok, has_t, tir;
};

__device__ inline c32 tir_coefficient(float a, float b)
{
    const float s = fabsf(a) > fabsf(b) ? fabsf(a) : fabsf(b);
    const float x = a / s, y = b / s, den = fmaf(y, y, x * x);
    return c32{fmaf(x, x, -(y * y)) / den, (-2.0f * x * y) / den};
}

// n points into the incident medium.
__device__ inline Split split(const Ray& in, f3 n, float ni, float nt)
{
    Split o;
    o.ok = false;
    o.has_t = false;
    o.tir = false;
    const f3 d = in.dir;
    float ci = -dot(d, n);
    if (!(ci >= -1e-3f && ni > 0.0f && nt > 0.0f)) return o;
    if (ci < 0.0f) ci = 0.0f;
    if (ci > 1.0f) ci = 1.0f;
    f3 s = cross(d, n);
    const float s2 = dot(s, s);
    const float sine = sqrtf(s2);
    if (s2 > 0x1p-40f)
        s = s * (1.0f / sine);
    else
        s = unit(axpy(d, -dot(d, in.e0), in.e0));
    const f3 pin = cross(d, s);  // d, s unit and orthogonal
    const float a = dot(in.e0, s), b = dot(in.e1, s);
    const float cc = dot(in.e0, pin), dd = dot(in.e1, pin);
    Jones j;
    j.c0x = clin(in.j.c0x, a, in.j.c0y, b);
    j.c0y = clin(in.j.c0x, cc, in.j.c0y, dd);
    j.c1x = clin(in.j.c1x, a, in.j.c1y, b);
    j.c1y = clin(in.j.c1x, cc, in.j.c1y, dd);
    const float eta = ni / nt;
    const float st = eta * sine;
    o.r.dir = unit(axpy(n, 2.0f * ci, d));
    o.r.e0 = s;
    o.r.e1 = cross(o.r.dir, s);
    c32 rs, rp;
    if (st > 1.0f) {
        const float beta = sqrtf((st - 1.0f) * (st + 1.0f));
        rs = tir_coefficient(ni * ci, nt * beta);
        rp = tir_coefficient(nt * ci, ni * beta);
        o.tir = true;
    } else {
        const float ct = sqrtf((1.0f - st) * (1.0f + st));
        const float ds = fmaf(nt, ct, ni * ci), dp = fmaf(ni, ct, nt * ci);
        if (!(ds > 0.0f && dp > 0.0f)) return o;
        rs = c32{fmaf(ni, ci, -(nt * ct)) / ds, 0.0f};
        rp = c32{fmaf(nt, ci, -(ni * ct)) / dp, 0.0f};
        const float num = (2.0f * sqrtf(ni * ci)) * sqrtf(nt * ct);
        const float ts = num / ds, tp = num / dp;
        o.t.dir = unit(axpy(n, -ct, axpy(n, ci, d) * eta));
        o.t.e0 = s;
        o.t.e1 = cross(o.t.dir, s);
        o.t.j.c0x = cscale(j.c0x, ts);
        o.t.j.c0y = cscale(j.c0y, tp);
        o.t.j.c1x = cscale(j.c1x, ts);
        o.t.j.c1y = cscale(j.c1y, tp);
        o.has_t = true;
    }
    o.r.j.c0x = cmul(j.c0x, rs);
    o.r.j.c0y = cmul(j.c0y, rp);
    o.r.j.c1x = cmul(j.c1x, rs);
    o.r.j.c1y = cmul(j.c1y, rp);
    o.ok = true;
    return o;
}

__device__ inline f3 outward_normal(const Shape& s, f3 p) { return unit(gradient(s, p)); }

struct Config {
    Shape shape;
    f3 ki, e0, e1;
    float n_int, n_ext, h, tail;
    uint32_t w, hgt, max_hits;
};

enum Status : uint32_t {
    kExhausted = 0,
    kTail = 1,
    kLimit = 2,
    kUnresolved = 3,
    kNumeric = 5,
    kMiss = 8,
};

}  // namespace drop
}  // namespace icephase
