// Coherent drop stages (path, focal, diffraction), per-element arithmetic for gfx950.
//
// The reference evaluates wave optics on a patch model of the outgoing ray
// field (Sadeghi et al. 2012, Sec. 4): a square vertex grid on the incident
// plane is traced through R, TT, TRT and TRRT with two FP32 Jones response
// columns and an integer + fraction optical path; adjacent vertices form
// bilinear chord patches in direction space; each query direction collects
// the patches whose spherical image contains it, and sums the transported,
// interpolated fields with the interpolated propagation phase (path), plus the
// focal-line quarter turns (focal), then a local Table II Gaussian across the
// rainbow transitions (diffraction). The arithmetic below reproduces the
// reference's operations one for one wherever the result is fixed by IEEE
// rounding (compensated optical path sums, interval certificates, Jacobians,
// Gauss quadrature, root selection, transport, Neumaier sums), so that the
// same inputs give the same bits; transcendental calls (atan2, sin, cos, asin,
// exp) come from the device math library and may differ from the host's in
// the last place. Sources, by stage:
//   optical path  include/rainbow/optical_path.hpp:13-91, phase_cycles.hpp:15-144
//   interface     include/rainbow/dielectric_interface.hpp:34-133
//   vertex trace  include/rainbow/raindrop_paths.hpp:40-144 (own intersector)
//   patches       include/rainbow/patch_geometry.hpp:10-207, outgoing_patch.hpp:90-106
//   folded        include/rainbow/folded_patch_geometry.hpp:9-140, folded_patch_builder.hpp:10-214
//   query         include/rainbow/bilinear_patch_intersection.hpp:87-322, patch_query_data.hpp:60-129
//   evaluator     include/rainbow/patch_optical_evaluator.hpp:27-300, polarization_transport.hpp:26-214,
//                 phase_interpolation.hpp:20-149
//   focal         include/rainbow/focal_phase.hpp:16-104
//   diffraction   include/rainbow/rainbow_diffraction.hpp:10-175
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "drop_device.hpp"

#define CDEV __host__ __device__ inline

namespace icephase {
namespace coh {

// ---------------------------------------------------------------- vectors
struct v3 {
    float x, y, z;
};
CDEV v3 operator+(v3 a, v3 b) { return v3{a.x + b.x, a.y + b.y, a.z + b.z}; }
CDEV v3 operator-(v3 a, v3 b) { return v3{a.x - b.x, a.y - b.y, a.z - b.z}; }
CDEV v3 operator-(v3 a) { return v3{-a.x, -a.y, -a.z}; }
CDEV v3 operator*(v3 a, float s) { return v3{a.x * s, a.y * s, a.z * s}; }
CDEV v3 operator/(v3 a, float s) { return v3{a.x / s, a.y / s, a.z / s}; }
// dot products are two fused multiply-adds, cross products symmetric differences
CDEV float dot(v3 a, v3 b) { return fmaf(a.x, b.x, fmaf(a.y, b.y, a.z * b.z)); }
CDEV v3 cross(v3 a, v3 b) { return v3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
CDEV float length(v3 a) { return sqrtf(dot(a, a)); }
CDEV v3 normalized(v3 a) { return a / length(a); }
CDEV v3 along(v3 p, v3 d, float t) { return v3{fmaf(t, d.x, p.x), fmaf(t, d.y, p.y), fmaf(t, d.z, p.z)}; }
CDEV bool finite(float x) { return x >= -0x1.fffffep127f && x <= 0x1.fffffep127f; }
CDEV bool finite(v3 a) { return finite(a.x) && finite(a.y) && finite(a.z); }
CDEV bool finite(double x) { return x >= -0x1.fffffffffffffp1023 && x <= 0x1.fffffffffffffp1023; }

struct d3 {
    double x, y, z;
};
CDEV d3 operator+(d3 a, d3 b) { return d3{a.x + b.x, a.y + b.y, a.z + b.z}; }
CDEV d3 operator-(d3 a, d3 b) { return d3{a.x - b.x, a.y - b.y, a.z - b.z}; }
CDEV d3 operator*(d3 a, double s) { return d3{a.x * s, a.y * s, a.z * s}; }
CDEV d3 operator/(d3 a, double s) { return d3{a.x / s, a.y / s, a.z / s}; }
CDEV double dot(d3 a, d3 b) { return fma(a.x, b.x, fma(a.y, b.y, a.z * b.z)); }
CDEV d3 cross(d3 a, d3 b) { return d3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
CDEV double length(d3 a) { return sqrt(dot(a, a)); }
CDEV d3 normalized(d3 a) { return a / length(a); }
CDEV d3 wide(v3 a) { return d3{(double)a.x, (double)a.y, (double)a.z}; }

// ---------------------------------------------------------------- complex fields
struct cf {
    float re, im;
};
CDEV cf operator+(cf a, cf b) { return cf{a.re + b.re, a.im + b.im}; }
CDEV cf operator*(cf a, float s) { return cf{a.re * s, a.im * s}; }
CDEV cf operator*(cf a, cf b) { return cf{fmaf(a.re, b.re, -(a.im * b.im)), fmaf(a.re, b.im, a.im * b.re)}; }

struct Field {  // components on (e0, e1) of the ray's transverse frame
    cf x, y;
};
CDEV bool finite(const Field& f) { return finite(f.x.re) && finite(f.x.im) && finite(f.y.re) && finite(f.y.im); }

// ---------------------------------------------------------------- optical path
// Whole turns plus a fraction in [0, 1).
struct Cycles {
    int32_t turns;
    float frac;
};
CDEV bool valid(Cycles c) { return c.frac >= 0.0f && c.frac < 1.0f; }

CDEV bool store_cycles(Cycles& c, int64_t whole, float hi, float lo)
{
    const float integral = floorf(hi);
    whole += (int64_t)integral;
    float f = (hi - integral) + lo;
    if (f < 0.0f) {
        --whole;
        f += 1.0f;
    }
    if (f >= 1.0f) {
        ++whole;
        f -= 1.0f;
    }
    if (whole < -2147483648LL || whole > 2147483647LL) return false;
    c.turns = (int32_t)whole;
    c.frac = f == 0.0f ? 0.0f : f;
    return true;
}
CDEV bool set_cycles(Cycles& c, int32_t whole, float residual)
{
    if (!(residual > -0x1p32f && residual < 0x1p32f)) return false;
    return store_cycles(c, (int64_t)whole, residual, 0.0f);
}
CDEV bool add_cycles(Cycles& c, Cycles r)
{
    if (!valid(c) || !valid(r)) return false;
    const float s = c.frac + r.frac;
    const float rv = s - c.frac, cv = s - rv;
    const float e = (c.frac - cv) + (r.frac - rv);
    return store_cycles(c, (int64_t)c.turns + (int64_t)r.turns, s, e);
}

// Unevaluated FP32 sum hi + lo.
struct Pair {
    float hi, lo;
};
CDEV Pair two_sum(float a, float b)
{
    const float s = a + b, v = s - a;
    return Pair{s, (a - (s - v)) + (b - v)};
}
CDEV Pair two_prod(float a, float b)
{
    const float p = a * b;
    return Pair{p, fmaf(a, b, -p)};
}
CDEV Pair operator+(Pair a, Pair b)
{
    const Pair s = two_sum(a.hi, b.hi), t = two_sum(a.lo, b.lo);
    const Pair v = two_sum(s.lo, t.hi), w = two_sum(s.hi, v.hi);
    return two_sum(w.hi, (w.lo + v.lo) + t.lo);
}
CDEV Pair operator-(Pair a) { return Pair{-a.hi, -a.lo}; }
CDEV Pair operator-(Pair a, Pair b) { return a + (-b); }
CDEV Pair operator*(Pair a, Pair b)
{
    const Pair p = two_prod(a.hi, b.hi);
    return two_sum(p.hi, fmaf(a.hi, b.lo, fmaf(a.lo, b.hi, p.lo)) + a.lo * b.lo);
}
CDEV Pair divided(Pair a, float d)
{
    const float q = a.hi / d;
    const float r = fmaf(-q, d, a.hi) + a.lo;
    return two_sum(q, r / d);
}
CDEV Pair root(Pair a)
{
    const float r = sqrtf(a.hi);
    if (r == 0.0f) return Pair{0.0f, 0.0f};
    return two_sum(r, (fmaf(-r, r, a.hi) + a.lo) / (2.0f * r));
}
CDEV Pair pair_dot(v3 a, v3 b) { return (two_prod(a.x, b.x) + two_prod(a.y, b.y)) + two_prod(a.z, b.z); }
CDEV Pair pair_distance(v3 a, v3 b)
{
    const Pair x = two_sum(a.x, -b.x), y = two_sum(a.y, -b.y), z = two_sum(a.z, -b.z);
    return root((x * x + y * y) + z * z);
}

// Accumulated phase of one ray: cycles per drop unit = radius_mm * 10^6 / wavelength_nm.
struct Path {
    Cycles phase;
    Pair per_unit;
};
CDEV Path make_path(float radius_mm, float wavelength_nm)
{
    Path p;
    p.phase = Cycles{0, 0.0f};
    p.per_unit = divided(two_prod(radius_mm, 1000000.0f), wavelength_nm);
    return p;
}
CDEV bool add_distance(Path& p, Pair distance, float index)
{
    const Pair q = (p.per_unit * Pair{index, 0.0f}) * distance;
    if (!(q.hi >= 0.0f && q.hi < 0x1p31f && fabsf(q.lo) < 0x1p31f)) return false;
    const float integral = floorf(q.hi);
    Cycles inc, corr, cand = p.phase;
    if (!set_cycles(inc, (int32_t)integral, q.hi - integral) || !set_cycles(corr, 0, q.lo) || !add_cycles(inc, corr) ||
        !add_cycles(cand, inc))
        return false;
    p.phase = cand;
    return true;
}

// ---------------------------------------------------------------- vertex records
// The reference's OutgoingVertex layout (68 bytes, include/rainbow/raindrop_trace_data.hpp:41-53)
// so vertex arrays pass across the boundary unchanged; column 1 of the
// unpolarized response lives in a parallel Field array.
enum VertexState : uint32_t { kUnwritten = 0, kValid = 1, kMiss = 2, kTir = 3, kUnresolved = 4, kBadInterface = 5,
                              kPhaseOverflow = 6 };
enum TraceNote : uint32_t { kIntersectionFallback = 1, kSnellFallback = 2 };
struct Vertex {
    v3 pos, dir, bx;
    Field f;
    Cycles q;
    uint32_t status, diag;
};
static_assert(sizeof(Vertex) == 68, "vertex record layout");

struct Frame {
    v3 e0, e1;
};
struct Response {  // two unit-input columns
    Field c[2];
};
struct PRay {
    v3 dir;
    Frame fr;
    Response r;
};
struct Branches {
    PRay refl, trans;
    bool has_t, ok;
    uint32_t diag;
};

CDEV cf tir_phase(float a, float b)
{
    const float s = fmaxf(fabsf(a), fabsf(b));
    const float x = a / s, y = b / s, den = fmaf(x, x, y * y);
    return cf{fmaf(x, x, -y * y) / den, (-2.0f * x * y) / den};
}

// Field amplitude split at an interface; n points into the incident medium.
CDEV Branches interface_split(const PRay& in, v3 n, float ni, float nt)
{
    Branches o;
    o.has_t = false;
    o.ok = false;
    o.diag = 0;
    const v3 wi = in.dir;
    float c = -dot(wi, n);
    if (!(c >= -0x1p-19f && c <= 1.0f + 0x1p-19f && ni > 0.0f && nt > 0.0f)) return o;
    c = fminf(1.0f, fmaxf(0.0f, c));
    v3 s = cross(wi, n);
    const float sin2 = dot(s, s);
    if (sin2 > 0x1p-40f)
        s = s / sqrtf(sin2);
    else
        s = normalized(in.fr.e0 - wi * dot(wi, in.fr.e0));
    const v3 pi = normalized(cross(wi, s));
    const float a = dot(in.fr.e0, s), b = dot(in.fr.e1, s), cc = dot(in.fr.e0, pi), dd = dot(in.fr.e1, pi);
    Response loc;
    for (int j = 0; j < 2; ++j)
        loc.c[j] = Field{in.r.c[j].x * a + in.r.c[j].y * b, in.r.c[j].x * cc + in.r.c[j].y * dd};
    const float eta = ni / nt;
    float disc = fmaf(-(eta * eta), sin2, 1.0f);
    if (fabsf(disc) < 0x1p-15f) {  // near the critical angle the branch is decided in FP64
        const d3 wd = normalized(wide(wi)), nd = normalized(wide(n));
        const double ratio = (double)ni / (double)nt;
        const d3 sd = cross(wd, nd);
        disc = (float)fma(-ratio * ratio, dot(sd, sd), 1.0);
        c = (float)(-dot(wd, nd));
        o.diag |= kSnellFallback;
    }
    const v3 wr = normalized(along(wi, n, 2.0f * c));
    o.refl.dir = wr;
    o.refl.fr = Frame{s, normalized(cross(wr, s))};
    cf rs, rp;
    if (disc < 0.0f) {
        const float beta = sqrtf(-disc);
        rs = tir_phase(ni * c, nt * beta);
        rp = tir_phase(nt * c, ni * beta);
    } else {
        const float ct = sqrtf(disc);
        const float ds = ni * c + nt * ct, dp = nt * c + ni * ct;
        if (!(ds > 0.0f && dp > 0.0f)) return o;
        rs = cf{(ni * c - nt * ct) / ds, 0.0f};
        rp = cf{(nt * c - ni * ct) / dp, 0.0f};
        const float ts = 2.0f * ni * c / ds, tp = 2.0f * ni * c / dp;
        const v3 wt = normalized(wi * eta + n * (eta * c - ct));
        o.trans.dir = wt;
        o.trans.fr = Frame{s, normalized(cross(wt, s))};
        for (int j = 0; j < 2; ++j) o.trans.r.c[j] = Field{loc.c[j].x * ts, loc.c[j].y * tp};
        o.has_t = true;
    }
    for (int j = 0; j < 2; ++j) o.refl.r.c[j] = Field{loc.c[j].x * rs, loc.c[j].y * rp};
    o.ok = finite(wr) && (!o.has_t || finite(o.trans.dir));
    return o;
}

// ---------------------------------------------------------------- patches
enum PatchState : uint32_t { kPatchUnwritten = 0, kRegularPos = 1, kRegularNeg = 2, kNeedsRefinement = 3,
                             kMissingCorners = 4, kNoOutgoing = 5, kInvalidVertex = 6, kInvalidGeometry = 7 };
constexpr uint32_t kPatchStates = 8;
CDEV bool has_geometry(uint32_t s) { return s == kRegularPos || s == kRegularNeg || s == kNeedsRefinement; }

// The reference's OutgoingPatch layout (32 bytes); corners in tensor order 00,10,01,11.
struct Patch {
    uint32_t v[4];
    uint32_t id;
    uint32_t status;
    float omega;  // signed solid angle (NaN unless regular)
    float area;   // incident cell area in drop units^2
};
static_assert(sizeof(Patch) == 32, "patch record layout");

struct Layout {
    uint32_t w, h, per_path, cells_per_path, cells;
    float area;
};
CDEV void corners_of(const Layout& L, uint32_t id, uint32_t (&out)[4])
{
    const uint32_t fam = id / L.cells_per_path, cell = id % L.cells_per_path;
    const uint32_t ix = cell % (L.w - 1), iy = cell / (L.w - 1);
    const uint32_t first = fam * L.per_path + iy * L.w + ix;
    out[0] = first;
    out[1] = first + 1;
    out[2] = first + L.w;
    out[3] = first + L.w + 1;
}

// Outward-rounded FP32 intervals (certify the sign of the radial Jacobian).
struct Iv {
    float lo, hi;
};
CDEV Iv iv_add(Iv a, Iv b) { return Iv{nextafterf(a.lo + b.lo, -INFINITY), nextafterf(a.hi + b.hi, INFINITY)}; }
CDEV Iv iv_sub(Iv a, Iv b) { return Iv{nextafterf(a.lo - b.hi, -INFINITY), nextafterf(a.hi - b.lo, INFINITY)}; }
CDEV Iv iv_mul(Iv a, Iv b)
{
    const float p0 = a.lo * b.lo, p1 = a.lo * b.hi, p2 = a.hi * b.lo, p3 = a.hi * b.hi;
    return Iv{nextafterf(fminf(fminf(p0, p1), fminf(p2, p3)), -INFINITY),
              nextafterf(fmaxf(fmaxf(p0, p1), fmaxf(p2, p3)), INFINITY)};
}
struct Iv3 {
    Iv x, y, z;
};
CDEV Iv3 iv3(v3 p) { return Iv3{Iv{p.x, p.x}, Iv{p.y, p.y}, Iv{p.z, p.z}}; }
CDEV Iv3 iv_sub(Iv3 a, Iv3 b) { return Iv3{iv_sub(a.x, b.x), iv_sub(a.y, b.y), iv_sub(a.z, b.z)}; }
CDEV Iv3 iv_cross(Iv3 a, Iv3 b)
{
    return Iv3{iv_sub(iv_mul(a.y, b.z), iv_mul(a.z, b.y)), iv_sub(iv_mul(a.z, b.x), iv_mul(a.x, b.z)),
               iv_sub(iv_mul(a.x, b.y), iv_mul(a.y, b.x))};
}
CDEV Iv iv_dot(Iv3 a, Iv3 b) { return iv_add(iv_add(iv_mul(a.x, b.x), iv_mul(a.y, b.y)), iv_mul(a.z, b.z)); }

CDEV float diff_of_products(float x, float y, float z, float w)
{
    const float zw = z * w;
    const float residual = fmaf(-z, w, zw);
    return fmaf(x, y, -zw) + residual;
}
CDEV float det3(v3 a, v3 b, v3 c)
{
    const v3 u = b - a, v = c - a;
    return fmaf(a.x, diff_of_products(u.y, v.z, u.z, v.y),
                fmaf(a.y, diff_of_products(u.z, v.x, u.x, v.z), a.z * diff_of_products(u.x, v.y, u.y, v.x)));
}
// solid angle of a triangle seen from the origin (lengths need not be exactly 1)
CDEV float tri_angle(v3 a, v3 b, v3 c)
{
    const float la = length(a), lb = length(b), lc = length(c);
    const float den = fmaf(la, lb * lc, fmaf(dot(a, b), lc, fmaf(dot(b, c), la, dot(c, a) * lb)));
    return 2.0f * atan2f(det3(a, b, c), den);
}

// +1 / -1 when the orientation of the chord patch is certified and its corners
// share an open hemisphere, else 0.
CDEV int certified_orientation(const v3 (&p)[4])
{
    const Iv3 p0 = iv3(p[0]), p1 = iv3(p[1]), p2 = iv3(p[2]), p3 = iv3(p[3]);
    const Iv d0 = iv_dot(p0, iv_cross(iv_sub(p1, p0), iv_sub(p2, p0)));
    const Iv d1 = iv_dot(p1, iv_cross(iv_sub(p1, p0), iv_sub(p3, p1)));
    const Iv d2 = iv_dot(p2, iv_cross(iv_sub(p3, p2), iv_sub(p2, p0)));
    const Iv d3v = iv_dot(p3, iv_cross(iv_sub(p3, p2), iv_sub(p3, p1)));
    const bool pos = d0.lo > 0 && d1.lo > 0 && d2.lo > 0 && d3v.lo > 0;
    const bool neg = d0.hi < 0 && d1.hi < 0 && d2.hi < 0 && d3v.hi < 0;
    if (!pos && !neg) return 0;
    const v3 axis = (p[0] + p[1]) + (p[2] + p[3]);
    const Iv3 a = iv3(axis);
    if (!(iv_dot(a, p0).lo > 0 && iv_dot(a, p1).lo > 0 && iv_dot(a, p2).lo > 0 && iv_dot(a, p3).lo > 0)) return 0;
    return pos ? 1 : -1;
}

CDEV bool usable(const Vertex& v)
{
    const float n2 = dot(v.dir, v.dir);
    return finite(v.pos) && finite(v.dir) && finite(v.bx) && fabsf(n2 - 1.0f) <= 0x1p-16f && valid(v.q) && finite(v.f);
}

// Status, geometry and solid angle of logical cell id (all four families).
CDEV Patch make_patch(const Vertex* V, const Layout& L, uint32_t id)
{
    Patch P;
    P.id = id;
    P.area = L.area;
    P.omega = __builtin_nanf("");
    corners_of(L, id, P.v);
    unsigned good = 0;
    bool bad_vertex = false, bad_geometry = false;
    v3 c[4];
    for (int i = 0; i < 4; ++i) {
        const Vertex& x = V[P.v[i]];
        c[i] = x.dir;
        if (x.status == kValid) {
            ++good;
            bad_geometry |= !usable(x);
        } else if (x.status != kMiss && x.status != kTir) {
            bad_vertex = true;
        }
    }
    if (bad_vertex)
        P.status = kInvalidVertex;
    else if (bad_geometry)
        P.status = kInvalidGeometry;
    else if (good == 0)
        P.status = kNoOutgoing;
    else if (good < 4)
        P.status = kMissingCorners;
    else {
        const int o = certified_orientation(c);
        P.status = kNeedsRefinement;
        if (o != 0) {
            const float ang = tri_angle(c[0], c[1], c[3]) + tri_angle(c[0], c[3], c[2]);
            if (finite(ang) && ang * (float)o > 0.0f) {
                P.status = o > 0 ? kRegularPos : kRegularNeg;
                P.omega = ang;
            }
        }
    }
    return P;
}

// ---------------------------------------------------------------- folded patches
enum FoldFlags : uint32_t { kFoldReady = 0, kFoldInvalidInput = 1, kFoldHemisphere = 2, kFoldNoSignChange = 4,
                            kFoldTopology = 8, kFoldQuadrature = 16, kFoldMetric = 32 };
enum FocalFlags : uint32_t { kFocalNone = 0, kFocalInputError = 1, kFocalInputPending = 2, kFocalDerivativePending = 4,
                             kFocalGeometry = 8, kFocalArithmetic = 16 };
constexpr uint32_t kFocalErrorMask = kFocalInputError | kFocalGeometry | kFocalArithmetic;

// One connected component of sign * J > 0 (layout of the reference's FoldedBranch, 64 bytes).
struct Branch {
    double u0, u1, area, omega, ru, rv;
    int32_t sign;
    uint32_t focal, qturns, extra;
};
// Folded record (layout of the reference's FoldedPatchRecord, 320 bytes).
struct Folded {
    uint32_t compact, id, flags, nb;
    double jac[4];
    Branch b[4];
    double err;
    uint32_t panels, pad;
};
static_assert(sizeof(Branch) == 64 && sizeof(Folded) == 320, "folded record layout");

struct FoldConfig {
    double tol;
    uint32_t max_panels, max_depth;
};
struct FocalConfig {
    uint32_t w, h;
    float half;
    uint32_t pad;
    uint32_t offsets[4];
};

CDEV double mixd(double a, double b, double t) { return fma(t, b - a, a); }
CDEV double bilin(const double* a, double u, double v) { return mixd(mixd(a[0], a[1], u), mixd(a[2], a[3], u), v); }

struct FoldGeom {
    d3 c[4];
    double j[4], scale;
};
CDEV d3 fold_point(const FoldGeom& g, double u, double v)
{
    return (g.c[0] * (1 - u) + g.c[1] * u) * (1 - v) + (g.c[2] * (1 - u) + g.c[3] * u) * v;
}
CDEV double fold_det(d3 a, d3 b, d3 c)
{
    const double yz = b.z * c.y, zx = b.x * c.z, xy = b.y * c.x;
    const double x = fma(b.y, c.z, -yz) + fma(-b.z, c.y, yz);
    const double y = fma(b.z, c.x, -zx) + fma(-b.x, c.z, zx);
    const double z = fma(b.x, c.y, -xy) + fma(-b.y, c.x, xy);
    return fma(a.x, x, fma(a.y, y, a.z * z));
}
CDEV uint32_t fold_init(FoldGeom& g, const v3 (&in)[4])
{
    for (int i = 0; i < 4; ++i) {
        if (!finite(in[i])) return kFoldInvalidInput;
        g.c[i] = wide(in[i]);
        const double q = dot(g.c[i], g.c[i]);
        if (!(q > 0 && q <= 4)) return kFoldInvalidInput;
    }
    const d3 axis = (g.c[0] + g.c[1]) + (g.c[2] + g.c[3]);
    for (int i = 0; i < 4; ++i) {
        double bound = 0;
        for (int k = 0; k < 4; ++k)
            bound += fabs(g.c[k].x * g.c[i].x) + fabs(g.c[k].y * g.c[i].y) + fabs(g.c[k].z * g.c[i].z);
        if (!(dot(axis, g.c[i]) > 0x1p-46 * bound)) return kFoldHemisphere;
    }
    const d3 eu = g.c[1] - g.c[0], ev = g.c[2] - g.c[0];
    g.j[0] = fold_det(g.c[0], eu, ev);
    g.j[1] = fold_det(g.c[1], eu, g.c[3] - g.c[1]);
    g.j[2] = fold_det(g.c[2], g.c[3] - g.c[2], ev);
    g.j[3] = fold_det(g.c[3], g.c[3] - g.c[2], g.c[3] - g.c[1]);
    g.scale = 0;
    bool pos = false, neg = false;
    for (int i = 0; i < 4; ++i) {
        if (!finite(g.j[i])) return kFoldInvalidInput;
        g.scale = fmax(g.scale, fabs(g.j[i]));
        pos |= g.j[i] > 0;
        neg |= g.j[i] < 0;
    }
    if (!(pos && neg && g.scale > 0)) return kFoldNoSignChange;
    for (int i = 0; i < 4; ++i) g.j[i] /= g.scale;
    return kFoldReady;
}
// v interval of sign * J(u, v) > 0 at fixed u
CDEV bool fold_section(const double* a, int sign, double u, double& lo, double& hi)
{
    const double b = (double)sign * mixd(a[0], a[1], u);
    const double t = (double)sign * mixd(a[2], a[3], u);
    if (b <= 0 && t <= 0) return false;
    lo = 0;
    hi = 1;
    if (b > 0 && t > 0) return true;
    const double r = fabs(b) / (fabs(b) + fabs(t));
    if (b > 0)
        hi = r;
    else
        lo = r;
    return hi > lo;
}
CDEV unsigned fold_cuts(const FoldGeom& g, double* x)
{
    unsigned n = 2;
    x[0] = 0;
    x[1] = 1;
    for (int e = 0; e < 2; ++e) {
        const double a = g.j[2 * e], b = g.j[2 * e + 1];
        if ((a < 0 && b > 0) || (a > 0 && b < 0)) x[n++] = fabs(a) / (fabs(a) + fabs(b));
    }
    for (unsigned i = 1; i < n; ++i) {
        const double t = x[i];
        unsigned k = i;
        while (k && x[k - 1] > t) {
            x[k] = x[k - 1];
            --k;
        }
        x[k] = t;
    }
    unsigned m = 1;
    for (unsigned i = 1; i < n; ++i)
        if (x[i] != x[m - 1]) x[m++] = x[i];
    return m;
}
CDEV bool fold_components(const FoldGeom& g, Folded& r)
{
    double x[4] = {0, 0, 0, 0};
    const unsigned n = fold_cuts(g, x);
    r.nb = 0;
    for (int sign = -1; sign <= 1; sign += 2) {
        int active = -1;
        for (unsigned i = 0; i + 1 < n; ++i) {
            const double mid = (x[i] + x[i + 1]) * 0.5;
            double lo = 0, hi = 0;
            if (!(mid > x[i] && mid < x[i + 1])) return false;
            if (!fold_section(g.j, sign, mid, lo, hi)) {
                active = -1;
                continue;
            }
            double bl = 0, bh = 0;
            const bool joined = active >= 0 && fold_section(g.j, sign, x[i], bl, bh);
            if (!joined) {
                if (r.nb >= 4) return false;
                active = (int)(r.nb++);
                r.b[active].sign = sign;
                r.b[active].u0 = x[i];
            }
            r.b[active].u1 = x[i + 1];
        }
    }
    return r.nb >= 2;
}
CDEV int fold_branch_at(const Folded& r, double u, double v)
{
    if (r.flags != kFoldReady || !(u >= 0 && u <= 1 && v >= 0 && v <= 1)) return -1;
    const double b = mixd(r.jac[0], r.jac[1], u), t = mixd(r.jac[2], r.jac[3], u);
    const double jac = mixd(b, t, v);
    const double bound = (1 - u) * (1 - v) * fabs(r.jac[0]) + u * (1 - v) * fabs(r.jac[1]) +
                         (1 - u) * v * fabs(r.jac[2]) + u * v * fabs(r.jac[3]);
    if (fabs(jac) <= 0x1p-46 * bound) return -1;
    const int sign = jac > 0 ? 1 : -1;
    for (unsigned i = 0; i < r.nb; ++i)
        if (r.b[i].sign == sign && u >= r.b[i].u0 && u <= r.b[i].u1) return (int)i;
    return -1;
}

// Gauss-Legendre nodes and weights of orders 8 and 16 on [-1, 1] (the
// published tables, to the digits the reference's quadrature uses).
__constant__ const double kGx8[8] = {-0.96028985649753618, -0.79666647741362673, -0.52553240991632899,
                                     -0.18343464249564978, 0.18343464249564978,  0.52553240991632899,
                                     0.79666647741362673,  0.96028985649753618};
__constant__ const double kGw8[8] = {0.10122853629037706, 0.22238103445337443, 0.31370664587788688,
                                     0.36268378337836166, 0.36268378337836166, 0.31370664587788688,
                                     0.22238103445337443, 0.10122853629037706};
__constant__ const double kGx16[16] = {-0.98940093499164994, -0.9445750230732326,  -0.86563120238783176,
                                       -0.755404408355003,   -0.61787624440264377, -0.45801677765722737,
                                       -0.28160355077925892, -0.095012509837637441, 0.095012509837637441,
                                       0.28160355077925892,  0.45801677765722737,  0.61787624440264377,
                                       0.755404408355003,    0.86563120238783176,  0.9445750230732326,
                                       0.98940093499164994};
__constant__ const double kGw16[16] = {0.027152459411754176, 0.062253523938647456, 0.095158511682492605,
                                       0.12462897125553407,  0.14959598881657671,  0.16915651939500265,
                                       0.18260341504492364,  0.18945061045506864,  0.18945061045506864,
                                       0.18260341504492364,  0.16915651939500265,  0.14959598881657671,
                                       0.12462897125553407,  0.095158511682492605, 0.062253523938647456,
                                       0.027152459411754176};

struct Quad {
    double area, mu, mv, omega;
};
struct Panel {
    double u0, u1, z0, z1;
    unsigned depth;
};
__device__ inline Quad fold_quadrature(const FoldGeom& g, int sign, const Panel& p, bool high)
{
    const unsigned n = high ? 16u : 8u;
    const double* xs = high ? kGx16 : kGx8;
    const double* ws = high ? kGw16 : kGw8;
    const double uh = (p.u1 - p.u0) * 0.5, uc = (p.u0 + p.u1) * 0.5;
    const double zh = (p.z1 - p.z0) * 0.5, zc = (p.z0 + p.z1) * 0.5;
    Quad t{0, 0, 0, 0};
    for (unsigned i = 0; i < n; ++i) {
        const double u = fma(uh, xs[i], uc);
        double vl = 0, vh = 0;
        if (!fold_section(g.j, sign, u, vl, vh)) continue;
        const double width = vh - vl;
        for (unsigned k = 0; k < n; ++k) {
            const double z = fma(zh, xs[k], zc), v = fma(z, width, vl);
            const double w = uh * zh * ws[i] * ws[k] * width;
            const d3 pt = fold_point(g, u, v);
            const double q = dot(pt, pt);
            const double J = fabs(bilin(g.j, u, v)) * g.scale;
            t.area += w;
            t.mu += w * u;
            t.mv += w * v;
            t.omega += w * J / (q * sqrt(q));
        }
    }
    return t;
}
__device__ inline bool fold_integrate(const FoldGeom& g, Branch& b, const FoldConfig& c, Folded& rec)
{
    double cuts[4] = {0, 0, 0, 0};
    const unsigned n = fold_cuts(g, cuts);
    Quad sum{0, 0, 0, 0};
    for (unsigned k = 0; k + 1 < n; ++k) {
        const double lo = fmax(cuts[k], b.u0), hi = fmin(cuts[k + 1], b.u1);
        if (!(hi > lo)) continue;
        Panel stack[26];
        unsigned size = 1;
        stack[0] = Panel{lo, hi, 0, 1, 0};
        while (size) {
            const Panel p = stack[--size];
            if (++rec.panels > c.max_panels) return false;
            const Quad low = fold_quadrature(g, b.sign, p, false), high = fold_quadrature(g, b.sign, p, true);
            if (!(high.area > 0 && high.omega > 0 && finite(high.area) && finite(high.omega))) return false;
            const double ea = fabs(high.area - low.area) / high.area;
            const double es = fabs(high.omega - low.omega) / high.omega;
            const double emu = fabs(high.mu - low.mu) / high.area;
            const double emv = fabs(high.mv - low.mv) / high.area;
            const double error = fmax(fmax(ea, es), fmax(emu, emv));
            if (error <= c.tol) {
                rec.err = fmax(rec.err, error);
                sum.area += high.area;
                sum.mu += high.mu;
                sum.mv += high.mv;
                sum.omega += high.omega;
                continue;
            }
            if (p.depth >= c.max_depth || size + 2 > 26) return false;
            if ((p.depth & 1u) == 0) {
                const double m = (p.u0 + p.u1) * 0.5;
                if (!(m > p.u0 && m < p.u1)) return false;
                stack[size++] = Panel{m, p.u1, p.z0, p.z1, p.depth + 1};
                stack[size++] = Panel{p.u0, m, p.z0, p.z1, p.depth + 1};
            } else {
                const double m = (p.z0 + p.z1) * 0.5;
                if (!(m > p.z0 && m < p.z1)) return false;
                stack[size++] = Panel{p.u0, p.u1, m, p.z1, p.depth + 1};
                stack[size++] = Panel{p.u0, p.u1, p.z0, m, p.depth + 1};
            }
        }
    }
    if (!(sum.area > 0 && sum.omega > 0 && finite(sum.omega))) return false;
    b.area = sum.area;
    b.omega = sum.omega;
    b.ru = sum.mu / sum.area;
    b.rv = sum.mv / sum.area;
    double vl = 0, vh = 0;
    if (!(b.ru > b.u0 && b.ru < b.u1) || !fold_section(g.j, b.sign, b.ru, vl, vh)) return false;
    if (!(b.rv > vl && b.rv < vh)) b.rv = (vl + vh) * 0.5;
    return b.rv > vl && b.rv < vh;
}

// ---------------------------------------------------------------- focal phase
struct FocalEstimate {
    double radial;
    uint32_t family, extra, qturns;
};
CDEV bool focal_layout_valid(const FocalConfig& c)
{
    if (c.w < 2 || c.h < 2 || !(c.half > 0) || !(c.half <= 0x1.fffffep127f)) return false;
    if ((uint64_t)c.w * c.h > 0xffffffffull / 4u || (uint64_t)(c.w - 1u) * (c.h - 1u) > 0xffffffffull / 4u)
        return false;
    for (int j = 0; j < 4; ++j)
        if (c.offsets[j] > 3) return false;
    return true;
}
// radial derivative of the corner scattering angles on the emitting plane:
// at the cell centre (regular patches) or at a branch representative (folded)
__device__ inline uint32_t focal_estimate(const Vertex* V, uint32_t nv, const Patch& P, v3 inc, const FocalConfig& c,
                                          double u, double v, bool centre, FocalEstimate& out)
{
    if (!(u >= 0 && u <= 1 && v >= 0 && v <= 1) || !V || !focal_layout_valid(c) || !finite(inc)) return kFocalGeometry;
    const uint32_t n = c.w * c.h, nx = c.w - 1u, cells = nx * (c.h - 1u);
    if (P.id >= 4u * cells || nv < 4u * n) return kFocalGeometry;
    const uint32_t fam = P.id / cells, cell = P.id % cells;
    const uint32_t ix = cell % nx, iy = cell / nx;
    const uint32_t first = fam * n + iy * c.w + ix;
    const uint32_t expect[4] = {first, first + 1u, first + c.w, first + c.w + 1u};
    const d3 wi = wide(inc);
    if (!(dot(wi, wi) > 0)) return kFocalGeometry;
    double th[4];
    for (int j = 0; j < 4; ++j) {
        if (P.v[j] != expect[j]) return kFocalGeometry;
        const Vertex& a = V[expect[j]];
        if (a.status != kValid || !finite(a.dir)) return kFocalGeometry;
        const d3 wo = wide(a.dir);
        if (!(dot(wo, wo) > 0)) return kFocalGeometry;
        th[j] = atan2(length(cross(wi, wo)), dot(wi, wo));
        if (!finite(th[j])) return kFocalGeometry;
    }
    const double e = (double)c.half;
    const double du = 2 * e / (double)(c.w - 1u), dv = 2 * e / (double)(c.h - 1u);
    const double U = (2 * ((double)ix + u) / (double)(c.w - 1u) - 1) * e;
    const double Vv = (2 * ((double)iy + v) / (double)(c.h - 1u) - 1) * e;
    const double r2 = fma(U, U, Vv * Vv);
    if (!(r2 > 0)) return kFocalDerivativePending;
    const double gu = centre ? ((th[1] - th[0]) + (th[3] - th[2])) / (2 * du)
                             : ((1 - v) * (th[1] - th[0]) + v * (th[3] - th[2])) / du;
    const double gv = centre ? ((th[2] - th[0]) + (th[3] - th[1])) / (2 * dv)
                             : ((1 - u) * (th[2] - th[0]) + u * (th[3] - th[1])) / dv;
    const double num = fma(U, gu, Vv * gv);
    const double guard = 0x1p-46 * (fabs(U * gu) + fabs(Vv * gv) + 1);
    if (!finite(num) || !finite(r2)) return kFocalGeometry;
    if (fabs(num) <= guard) return kFocalDerivativePending;
    FocalEstimate a;
    a.family = fam;
    a.radial = num / (2 * r2);
    a.extra = num > 0 ? 1u : 0u;
    a.qturns = (c.offsets[fam] + a.extra) & 3u;
    if (!finite(a.radial)) return kFocalGeometry;
    out = a;
    return kFocalNone;
}

// Folded record of a NeedsRefinement patch (branch areas, solid angles,
// representatives, focal quarter turns at the representatives).
__device__ inline Folded fold_prepare(const Vertex* V, uint32_t nv, const Patch& P, uint32_t compact, v3 inc,
                                      const FocalConfig* focal, const FoldConfig& cfg)
{
    Folded r;
    for (int i = 0; i < 4; ++i) r.jac[i] = 0;
    for (int i = 0; i < 4; ++i) r.b[i] = Branch{0, 0, 0, 0, 0, 0, 0, kFocalInputPending, 0, 0};
    r.compact = compact;
    r.id = P.id;
    r.flags = kFoldInvalidInput;
    r.nb = 0;
    r.err = 0;
    r.panels = 0;
    r.pad = 0;
    if (!V || P.status != kNeedsRefinement || !(P.area > 0) || !finite(P.area)) return r;
    v3 c[4];
    for (int i = 0; i < 4; ++i) {
        const uint32_t id = P.v[i];
        if (id >= nv || V[id].status != kValid || !usable(V[id])) return r;
        c[i] = V[id].dir;
    }
    FoldGeom g;
    r.flags = fold_init(g, c);
    if (r.flags != kFoldReady) return r;
    for (int i = 0; i < 4; ++i) r.jac[i] = g.j[i];
    if (!fold_components(g, r)) {
        r.flags = kFoldTopology;
        return r;
    }
    double area = 0;
    for (unsigned i = 0; i < r.nb; ++i) {
        if (!fold_integrate(g, r.b[i], cfg, r)) {
            r.flags = kFoldQuadrature;
            return r;
        }
        area += r.b[i].area;
    }
    if (!finite(area) || fabs(area - 1) > 8 * cfg.tol) {
        r.flags = kFoldMetric;
        return r;
    }
    if (focal)
        for (unsigned i = 0; i < r.nb; ++i) {
            Branch& b = r.b[i];
            FocalEstimate e{0, 0, 0, 0};
            b.focal = focal_estimate(V, nv, P, inc, *focal, b.ru, b.rv, false, e);
            if (b.focal == kFocalNone) {
                b.qturns = e.qturns;
                b.extra = e.extra;
            }
        }
    return r;
}

// ---------------------------------------------------------------- bilinear patch query
enum HitFlags : uint32_t { kHitBoundary = 1, kHitSingular = 2, kHitFp64 = 4 };
enum QueryNotes : uint32_t { kQueryUnresolved = 1, kQueryInvalid = 2, kQueryFp64 = 4 };
struct PHit {
    float u, v, t, res;
    uint32_t flags;
};
struct PHits {
    PHit h[2];
    uint32_t count, flags;
};

template <class T>
CDEV T tabs(T x) { return x < T(0) ? -x : x; }
template <class T>
CDEV T tmax(T a, T b) { return a > b ? a : b; }
template <class T>
CDEV T tfma(T a, T b, T c);
template <>
CDEV float tfma<float>(float a, float b, float c) { return fmaf(a, b, c); }
template <>
CDEV double tfma<double>(double a, double b, double c) { return fma(a, b, c); }
template <class T>
CDEV T tsqrt(T x);
template <>
CDEV float tsqrt<float>(float x) { return sqrtf(x); }
template <>
CDEV double tsqrt<double>(double x) { return sqrt(x); }
template <class T>
CDEV T tcopysign(T x, T y);
template <>
CDEV float tcopysign<float>(float x, float y) { return copysignf(x, y); }
template <>
CDEV double tcopysign<double>(double x, double y) { return copysign(x, y); }
template <class T>
constexpr T eps_of();
template <>
constexpr float eps_of<float>() { return 0x1p-23f; }
template <>
constexpr double eps_of<double>() { return 0x1p-52; }

template <class T>
struct V3 {
    T x, y, z;
};
template <class T>
CDEV V3<T> vcast(v3 a) { return V3<T>{(T)a.x, (T)a.y, (T)a.z}; }
template <class T>
CDEV V3<T> vsub(V3<T> a, V3<T> b) { return V3<T>{a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T>
CDEV V3<T> vadd(V3<T> a, V3<T> b) { return V3<T>{a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T>
CDEV V3<T> vmul(V3<T> a, T s) { return V3<T>{a.x * s, a.y * s, a.z * s}; }
template <class T>
CDEV T comp(V3<T> p, unsigned k) { return k == 0 ? p.x : (k == 1 ? p.y : p.z); }

template <class T>
CDEV T dprod(T x, T y, T z, T w)
{
    const T zw = z * w;
    return tfma(x, y, -zw) + tfma(-z, w, zw);
}
template <class T>
struct P2 {
    T x, y;
};
template <class T>
CDEV P2<T> p2add(P2<T> a, P2<T> b) { return P2<T>{a.x + b.x, a.y + b.y}; }
template <class T>
CDEV P2<T> p2mul(P2<T> a, T s) { return P2<T>{a.x * s, a.y * s}; }
template <class T>
CDEV T p2big(P2<T> a) { return tmax(tabs(a.x), tabs(a.y)); }
template <class T>
CDEV T p2cross(P2<T> a, P2<T> b) { return dprod(a.x, b.y, a.y, b.x); }
template <class T>
CDEV P2<T> project(V3<T> p, V3<T> w, unsigned kz)
{
    const unsigned kx = (kz + 1) % 3, ky = (kz + 2) % 3;
    const T z = comp(w, kz);
    return P2<T>{dprod(z, comp(p, kx), comp(w, kx), comp(p, kz)), dprod(z, comp(p, ky), comp(w, ky), comp(p, kz))};
}
CDEV v3 patch_eval(const v3 (&c)[4], float u, float v)
{
    const v3 lo = c[0] * (1.0f - u) + c[1] * u;
    const v3 up = c[2] * (1.0f - u) + c[3] * u;
    return lo * (1.0f - v) + up * v;
}

// Both isolated roots of P(u, v) = t w on the unnormalized chord patch, in
// precision T; retry = the FP32 pass could not decide, re-solve in FP64.
template <class T>
__device__ inline PHits bilinear_solve(const v3 (&c)[4], v3 dir, bool& retry)
{
    PHits out;
    out.count = 0;
    out.flags = 0;
    retry = false;
    const V3<T> w = vcast<T>(dir);
    unsigned kz = 0;
    if (tabs(w.y) > tabs(w.x)) kz = 1;
    if (tabs(w.z) > tabs(comp(w, kz))) kz = 2;
    const V3<T> p0 = vcast<T>(c[0]), p1 = vcast<T>(c[1]), p2 = vcast<T>(c[2]), p3 = vcast<T>(c[3]);
    const V3<T> eu = vsub(p1, p0), ev = vsub(p2, p0);
    const V3<T> tw = vsub(vsub(p3, p2), eu);
    P2<T> a = project(p0, w, kz), b = project(eu, w, kz), cc = project(ev, w, kz), d = project(tw, w, kz);
    const T scale = tmax(tmax(p2big(a), p2big(b)), tmax(p2big(cc), p2big(d)));
    if (!(scale > T(0))) {
        retry = true;
        return out;
    }
    a = p2mul(a, T(1) / scale);
    b = p2mul(b, T(1) / scale);
    cc = p2mul(cc, T(1) / scale);
    d = p2mul(d, T(1) / scale);
    const T eps = eps_of<T>();
    const P2<T> q[4] = {a, p2add(a, b), p2add(a, cc), p2add(p2add(p2add(a, b), cc), d)};
    T xlo = q[0].x, xhi = xlo, ylo = q[0].y, yhi = ylo;
    for (int i = 1; i < 4; ++i) {
        xlo = xlo < q[i].x ? xlo : q[i].x;
        xhi = tmax(xhi, q[i].x);
        ylo = ylo < q[i].y ? ylo : q[i].y;
        yhi = tmax(yhi, q[i].y);
    }
    const T guard = T(64) * eps;
    if (xlo > guard || xhi < -guard || ylo > guard || yhi < -guard) return out;
    const T A = p2cross(b, d), B = p2cross(a, d) + p2cross(b, cc), C = p2cross(a, cc);
    const T mag = tmax(tabs(A), tmax(tabs(B), tabs(C)));
    if (!(mag > T(0))) {
        retry = true;
        return out;
    }
    const T qa = A / mag, qb = B / mag, qc = C / mag;
    if (sizeof(T) == 4 && A != T(0) && tabs(qa) < T(64) * eps) {
        retry = true;
        return out;
    }
    T roots[2] = {T(0), T(0)};
    unsigned nroots = 0;
    bool repeated = false;
    if (A == T(0)) {
        if (B == T(0)) {
            if (tabs(qc) <= guard) retry = true;
            return out;
        }
        roots[nroots++] = -qc / qb;
    } else {
        const T disc = dprod(qb, qb, T(4) * qa, qc);
        const T err = T(64) * eps * (qb * qb + tabs(T(4) * qa * qc));
        if (disc < T(0)) {
            if (disc >= -err) retry = true;
            return out;
        }
        if (disc == T(0)) {
            if (sizeof(T) == 4) {
                retry = true;
                return out;
            }
            roots[nroots++] = -qb / (T(2) * qa);
            repeated = true;
        } else {
            if (disc <= err) {
                retry = true;
                return out;
            }
            const T qq = -T(0.5) * (qb + tcopysign(tsqrt(disc), qb));
            roots[0] = qq / qa;
            roots[1] = qc / qq;
            if (roots[1] < roots[0]) {
                const T tmp = roots[0];
                roots[0] = roots[1];
                roots[1] = tmp;
            }
            nroots = 2;
        }
    }
    for (unsigned k = 0; k < nroots; ++k) {
        T u = roots[k];
        if (!(u == u)) {
            retry = true;
            continue;
        }
        if (u < -guard || u > T(1) + guard) continue;
        const P2<T> num = p2add(a, p2mul(b, u)), den = p2add(cc, p2mul(d, u));
        if (p2big(den) <= guard) {
            retry = true;
            continue;
        }
        T v = tabs(den.x) >= tabs(den.y) ? -num.x / den.x : -num.y / den.y;
        if (!(v == v)) {
            retry = true;
            continue;
        }
        if (v < -guard || v > T(1) + guard) continue;
        T cond = T(1);
        for (int it = 0; it < 3; ++it) {
            const P2<T> fu = p2add(b, p2mul(d, v)), fv = p2add(cc, p2mul(d, u));
            const T det = p2cross(fu, fv);
            const P2<T> res = p2add(p2add(p2add(a, p2mul(b, u)), p2mul(cc, v)), p2mul(d, u * v));
            if (tabs(det) <= T(64) * eps * p2big(fu) * p2big(fv)) {
                if (!repeated) retry = true;
                break;
            }
            cond = tmax(T(1), tmax(p2big(fu), p2big(fv)) / tabs(det));
            const T du = p2cross(res, fv) / det;
            const T dv = p2cross(fu, res) / det;
            u -= du;
            v -= dv;
        }
        const T perr = T(32) * eps * cond;
        if (sizeof(T) == 4) {
            if (perr > T(2e-5) || u <= perr || u >= T(1) - perr || v <= perr || v >= T(1) - perr) {
                retry = true;
                continue;
            }
        }
        if (perr > T(1e-5) && !repeated) {
            retry = true;
            continue;
        }
        if (u < -perr || u > T(1) + perr || v < -perr || v > T(1) + perr) continue;
        const bool boundary = u <= perr || u >= T(1) - perr || v <= perr || v >= T(1) - perr;
        if (u < T(0) || u > T(1)) u = u < T(0) ? T(0) : T(1);
        if (v < T(0) || v > T(1)) v = v < T(0) ? T(0) : T(1);
        const P2<T> res = p2add(p2add(p2add(a, p2mul(b, u)), p2mul(cc, v)), p2mul(d, u * v));
        if (p2big(res) > T(256) * eps) {
            retry = true;
            continue;
        }
        const V3<T> lo = vadd(vmul(p0, T(1) - u), vmul(p1, u));
        const V3<T> up = vadd(vmul(p2, T(1) - u), vmul(p3, u));
        const V3<T> pt = vadd(vmul(lo, T(1) - v), vmul(up, v));
        const T t = comp(pt, kz) / comp(w, kz);
        if (!(t > T(0))) {
            if (tabs(t) <= T(64) * eps) retry = true;
            continue;
        }
        PHit h;
        h.u = (float)u;
        h.v = (float)v;
        h.t = (float)t;
        h.flags = 0;
        if (boundary) h.flags |= kHitBoundary;
        if (repeated) h.flags |= kHitSingular;
        if (sizeof(T) == 8) h.flags |= kHitFp64;
        const v3 e3 = patch_eval(c, h.u, h.v) - dir * h.t;
        h.res = fmaxf(fabsf(e3.x), fmaxf(fabsf(e3.y), fabsf(e3.z)));
        if (!(h.res <= 0x1p-21f && h.t < 4.0f)) {
            retry = true;
            continue;
        }
        if (out.count < 2)
            out.h[out.count++] = h;
        else
            retry = true;
    }
    return out;
}

__device__ inline PHits bilinear_intersect(const v3 (&c)[4], v3 dir)
{
    const float n2 = dot(dir, dir);
    PHits bad;
    bad.count = 0;
    bad.flags = kQueryInvalid;
    if (!finite(dir) || !(n2 > 0.99f && n2 < 1.01f)) return bad;
    for (int i = 0; i < 4; ++i)
        if (!finite(c[i]) || !(dot(c[i], c[i]) <= 4.0f)) return bad;
    bool retry = false;
    PHits r = bilinear_solve<float>(c, dir, retry);
    if (retry) {
        bool again = false;
        r = bilinear_solve<double>(c, dir, again);
        r.flags |= kQueryFp64;
        if (again) r.flags |= kQueryUnresolved;
    }
    if (r.count == 2) {
        PHit& a = r.h[0];
        PHit& b = r.h[1];
        if (b.u < a.u || (b.u == a.u && b.v < a.v)) {
            const PHit t = a;
            a = b;
            b = t;
        }
    }
    return r;
}

// One stored query hit (the reference's PatchQueryHit layout, 32 bytes).
struct QHit {
    uint32_t compact, id, root, flags;
    float u, v, t, res;
};
static_assert(sizeof(QHit) == 32, "query hit layout");

// ---------------------------------------------------------------- evaluator
struct F64 {  // s and p components, FP64
    double sr, si, pr, pi;
};
struct Neum {
    double s, c;
    CDEV void add(double x)
    {
        const double n = s + x;
        c += fabs(s) >= fabs(x) ? (s - n) + x : (x - n) + s;
        s = n;
    }
    CDEV double value() const { return s + c; }
};

enum OpticalFlags : uint32_t { kOptQueryError = 1, kOptInvalid = 2, kOptFrame = 4, kOptPhase = 8, kOptArith = 16,
                               kOptRefinePending = 32, kOptBoundaryPending = 64, kOptSingularPending = 128 };
constexpr uint32_t kOptErrorMask = kOptQueryError | kOptInvalid | kOptFrame | kOptPhase | kOptArith;
constexpr uint32_t kOptPendingMask = kOptRefinePending | kOptBoundaryPending | kOptSingularPending;
enum QueryFlags : uint32_t { kQUnresolved = 1, kQInvalid = 2, kQCountMismatch = 4, kQRefinement = 8, kQBoundary = 16,
                             kQFp64 = 32, kQDuplicate = 64, kQSingular = 128, kQOverflow = 256 };
constexpr uint32_t kQueryErrorMask = kQUnresolved | kQInvalid | kQCountMismatch | kQOverflow;

CDEV bool unit_vec(v3 a) { return finite(a) && fabsf(dot(a, a) - 1.0f) <= 0x1p-10f; }
CDEV bool is_frame(v3 d, const Frame& f)
{
    if (!unit_vec(f.e0) || !unit_vec(f.e1)) return false;
    if (fabsf(dot(d, f.e0)) > 0x1p-10f || fabsf(dot(d, f.e1)) > 0x1p-10f || fabsf(dot(f.e0, f.e1)) > 0x1p-10f)
        return false;
    return dot(cross(f.e0, f.e1), d) >= 1.0f - 4.0f * 0x1p-10f;
}
CDEV bool scattering_frame(v3 inc, v3 bx, v3 out, Frame& f)
{
    if (!unit_vec(inc) || !unit_vec(out) || !unit_vec(bx) || fabsf(dot(inc, bx)) > 0x1p-10f) return false;
    v3 e0 = cross(inc, out);
    float n2 = dot(e0, e0);
    if (n2 > 0x1p-30f)
        e0 = e0 / sqrtf(n2);
    else {
        e0 = bx - out * dot(out, bx);
        n2 = dot(e0, e0);
        if (!(n2 > 0x1p-30f)) return false;
        e0 = e0 / sqrtf(n2);
    }
    v3 e1 = cross(out, e0);
    const float m2 = dot(e1, e1);
    if (!(m2 > 0x1p-30f)) return false;
    e1 = e1 / sqrtf(m2);
    const Frame cand{e0, e1};
    if (!is_frame(out, cand)) return false;
    f = cand;
    return true;
}
CDEV d3 min_rotation(d3 x, d3 a, d3 b, double cosine)
{
    const d3 v = cross(a, b);
    const d3 vx = cross(v, x);
    return (x + vx) + cross(v, vx) / (1.0 + cosine);
}
CDEV bool fits_float(double x) { return x >= -0x1.fffffep127 && x <= 0x1.fffffep127; }
// Parallel transport of both response columns to the target direction and frame.
CDEV bool transport(v3 sd, const Frame& sf, const Response& src, v3 td, const Frame& tf, Response& out)
{
    if (!finite(src.c[0]) || !finite(src.c[1]) || !unit_vec(sd) || !unit_vec(td) || !is_frame(sd, sf) ||
        !is_frame(td, tf))
        return false;
    const d3 sw = normalized(wide(sd)), tw = normalized(wide(td));
    const double cosine = dot(sw, tw);
    if (!(1.0 + cosine > 0x1p-30)) return false;
    const d3 s0 = wide(sf.e0), s1 = wide(sf.e1), t0 = wide(tf.e0), t1 = wide(tf.e1);
    Response cand;
    for (int j = 0; j < 2; ++j) {
        const Field& f = src.c[j];
        const d3 re = s0 * (double)f.x.re + s1 * (double)f.y.re;
        const d3 im = s0 * (double)f.x.im + s1 * (double)f.y.im;
        const d3 rt = min_rotation(re, sw, tw, cosine), it = min_rotation(im, sw, tw, cosine);
        const double xr = dot(rt, t0), xi = dot(it, t0), yr = dot(rt, t1), yi = dot(it, t1);
        if (!fits_float(xr) || !fits_float(xi) || !fits_float(yr) || !fits_float(yi)) return false;
        cand.c[j] = Field{cf{(float)xr, (float)xi}, cf{(float)yr, (float)yi}};
    }
    out = cand;
    return true;
}

CDEV double cycles_diff(Cycles a, Cycles b)
{
    const int64_t whole = (int64_t)a.turns - (int64_t)b.turns;
    return (double)whole + ((double)a.frac - (double)b.frac);
}
CDEV bool interp_cycles(Cycles q00, Cycles q10, Cycles q01, Cycles q11, float u, float v, Cycles& out)
{
    if (!valid(q00) || !valid(q10) || !valid(q01) || !valid(q11) || !(u >= 0.0f && u <= 1.0f) ||
        !(v >= 0.0f && v <= 1.0f))
        return false;
    if (u == 0.0f && v == 0.0f) {
        out = q00;
        return true;
    }
    if (u == 1.0f && v == 0.0f) {
        out = q10;
        return true;
    }
    if (u == 0.0f && v == 1.0f) {
        out = q01;
        return true;
    }
    if (u == 1.0f && v == 1.0f) {
        out = q11;
        return true;
    }
    const double du0 = cycles_diff(q10, q00), du1 = cycles_diff(q11, q01), dv0 = cycles_diff(q01, q00);
    const double ud = (double)u, vd = (double)v;
    const double row = fma(ud, du1 - du0, dv0);
    const double delta = fma(vd, row, ud * du0);
    if (!(delta > -0x1p33 && delta < 0x1p33)) return false;
    const double ip = floor(delta);
    double f = (double)q00.frac + (delta - ip);
    int64_t turns = (int64_t)q00.turns + (int64_t)ip;
    if (f < 0.0) {
        --turns;
        f += 1.0;
    }
    if (f >= 1.0) {
        ++turns;
        f -= 1.0;
    }
    if (turns < -2147483648LL || turns > 2147483647LL) return false;
    float sf = (float)f;
    if (sf >= 1.0f) {
        ++turns;
        sf = 0.0f;
        if (turns > 2147483647LL) return false;
    } else if (sf < 0.0f) {
        --turns;
        sf += 1.0f;
        if (turns < -2147483648LL) return false;
    }
    out.turns = (int32_t)turns;
    out.frac = sf == 0.0f ? 0.0f : sf;
    return valid(out);
}
// exp(+i 2 pi q): exact quarter turns, else cos/sin of 2 pi (fraction reduced to [-1/2, 1/2])
__device__ inline bool unit_phasor(Cycles q, cf& out)
{
    if (!valid(q)) return false;
    if (q.frac == 0.0f) { out = cf{1.0f, 0.0f}; return true; }
    if (q.frac == 0.25f) { out = cf{0.0f, 1.0f}; return true; }
    if (q.frac == 0.5f) { out = cf{-1.0f, 0.0f}; return true; }
    if (q.frac == 0.75f) { out = cf{0.0f, -1.0f}; return true; }
    const float red = q.frac > 0.5f ? q.frac - 1.0f : q.frac;
    const float ang = 0x1.921fb6p+2f * red;
    out = cf{cosf(ang), sinf(ang)};
    return true;
}
CDEV double lerp2(double a, double b, double c, double d, double u, double v)
{
    const double lo = fma(u, b - a, a), up = fma(u, d - c, c);
    return fma(v, up - lo, lo);
}
CDEV F64 quarter_turns(F64 a, uint32_t q)
{
    switch (q & 3u) {
    case 0: return a;
    case 1: return F64{-a.si, a.sr, -a.pi, a.pr};
    case 2: return F64{-a.sr, -a.si, -a.pr, -a.pi};
    default: return F64{a.si, -a.sr, a.pi, -a.pr};
    }
}
CDEV bool finite(const F64& f) { return finite(f.sr) && finite(f.si) && finite(f.pr) && finite(f.pi); }

struct Contribution {
    F64 field[2];
    double inc_s, inc_p;
};

// One hit's transported, interpolated, area-weighted field (unpolarized: two columns).
__device__ inline uint32_t evaluate_hit(const Vertex* V, const Field* second, uint32_t nv, const Patch& P,
                                        const QHit& h, v3 dir, const Frame& tf, Contribution& out,
                                        const Folded* folded, const Branch** used)
{
    *used = nullptr;
    if ((h.flags & ~(uint32_t)(kHitBoundary | kHitSingular | kHitFp64)) != 0u || h.id != P.id || h.root > 1u ||
        !(h.u >= 0.0f && h.u <= 1.0f && h.v >= 0.0f && h.v <= 1.0f) || !(h.t > 0.0f && finite(h.t)) ||
        !(h.res >= 0.0f && finite(h.res)))
        return kOptInvalid;
    uint32_t pending = 0;
    const Branch* br = nullptr;
    if (P.status == kNeedsRefinement) {
        if (!folded || folded->id != P.id || folded->compact != h.compact || folded->flags != kFoldReady)
            pending |= kOptRefinePending;
        else {
            const int b = fold_branch_at(*folded, (double)h.u, (double)h.v);
            if (b < 0)
                pending |= kOptSingularPending;
            else
                br = &folded->b[b];
        }
    } else if (!(P.status == kRegularPos || P.status == kRegularNeg)) {
        return kOptInvalid;
    }
    if (h.flags & kHitBoundary) pending |= kOptBoundaryPending;
    if (h.flags & kHitSingular) pending |= kOptSingularPending;
    if (pending) return pending;
    const double area = br ? (double)P.area * br->area : (double)P.area;
    const double omega = br ? br->omega : fabs((double)P.omega);
    if (!(area > 0.0 && finite(area) && omega > 0.0 && finite(omega))) return kOptInvalid;
    Field tr[2][4];
    Cycles q[4];
    for (int k = 0; k < 4; ++k) {
        const uint32_t idx = P.v[k];
        if (idx >= nv) return kOptInvalid;
        const Vertex& x = V[idx];
        if (x.status != kValid || !finite(x.pos) || !finite(x.dir) || !finite(x.bx) || !finite(x.f))
            return kOptInvalid;
        if (!valid(x.q)) return kOptPhase;
        const v3 e1 = cross(x.dir, x.bx);
        const d3 e1d = wide(e1);
        const double l2 = dot(e1d, e1d);
        if (!(l2 > 0.0 && finite(l2))) return kOptFrame;
        const Frame sf{x.bx, normalized(e1)};
        const Response src{{x.f, second[idx]}};
        if (!finite(src.c[1])) return kOptInvalid;
        Response t;
        if (!transport(x.dir, sf, src, dir, tf, t)) return kOptFrame;
        tr[0][k] = t.c[0];
        tr[1][k] = t.c[1];
        q[k] = x.q;
    }
    Cycles qi;
    if (!interp_cycles(q[0], q[1], q[2], q[3], h.u, h.v, qi)) return kOptPhase;
    cf ph;
    if (!unit_phasor(qi, ph)) return kOptPhase;
    const double amp = sqrt(area / omega);
    const double c = (double)ph.re, s = (double)ph.im;
    Contribution cand;
    cand.inc_s = 0.0;
    cand.inc_p = 0.0;
    for (int j = 0; j < 2; ++j) {
        const Field* f = tr[j];
        const double sr = amp * lerp2(f[0].x.re, f[1].x.re, f[2].x.re, f[3].x.re, h.u, h.v);
        const double si = amp * lerp2(f[0].x.im, f[1].x.im, f[2].x.im, f[3].x.im, h.u, h.v);
        const double pr = amp * lerp2(f[0].y.re, f[1].y.re, f[2].y.re, f[3].y.re, h.u, h.v);
        const double pim = amp * lerp2(f[0].y.im, f[1].y.im, f[2].y.im, f[3].y.im, h.u, h.v);
        cand.inc_s += 0.5 * fma(sr, sr, si * si);
        cand.inc_p += 0.5 * fma(pr, pr, pim * pim);
        cand.field[j] = F64{fma(sr, c, -si * s), fma(sr, s, si * c), fma(pr, c, -pim * s), fma(pr, s, pim * c)};
    }
    if (!finite(cand.inc_s) || !finite(cand.inc_p) || !finite(cand.field[0]) || !finite(cand.field[1]))
        return kOptArith;
    out = cand;
    *used = br;
    return 0;
}

// ---------------------------------------------------------------- per-direction results
struct Optics {
    double inc_s, inc_p, path_s, path_p;       // incoherent and path-coherent intensities
    double focal_s, focal_p;                    // with focal quarter turns (NaN when invalid)
    double fam_inc[4];                          // per-family incoherent envelope (focal evaluation)
    F64 path_field[2], focal_field[2];
    uint32_t hits, evaluated, rejected, flags;  // optical flags
    uint32_t qflags, focal_flags, first_problem, folded_hits;
    uint32_t fam_hits[4];
    uint32_t regular_hits, refinement_hits, corrected, extra;
};

// Diffraction result per direction.
struct Diffracted {
    double s, p, sigma, blend;
    uint32_t flags, kind, transition, pad;
};
enum DiffFlags : uint32_t { kDiffFiltered = 1, kDiffUnderresolved = 2, kDiffInputUnavailable = 4,
                            kDiffStencilUnavailable = 8, kDiffDetectionUnavailable = 16, kDiffInvalid = 32 };
constexpr uint32_t kDiffUnavailableMask = kDiffInputUnavailable | kDiffStencilUnavailable | kDiffDetectionUnavailable |
                                          kDiffInvalid;
enum TransitionFlags : uint32_t { kTransPrimary = 1, kTransSecondary = 2, kTransUnknown = 4 };

struct DiffConfig {
    double sigma;         // primary sigma (rad); secondary = 2 sigma
    double support;       // 4
    double contrast;      // 0.25
    double inner;         // 3
    uint32_t nt, np;
};
constexpr double kPi = 3.141592653589793238462643383279502884;
constexpr double kAngTol = 0x1p-44;

CDEV bool diff_valid(const DiffConfig& c)
{
    return c.nt >= 2u && c.np >= 1u && c.nt <= 32767u && c.np <= 32767u && c.sigma > 0 && finite(c.sigma) &&
           c.support >= 2 && c.support <= 8 && c.inner >= 0 && c.inner < c.support && 2 * c.sigma * c.support < kPi &&
           c.contrast >= 0 && c.contrast < 1;
}
CDEV bool focal_sample_valid(const Optics& r)
{
    return r.focal_flags == 0u && r.focal_s >= 0 && r.focal_p >= 0 && finite(r.focal_s + r.focal_p);
}
__device__ inline double angular_distance(double a, double b, double phi)
{
    const double u = sin(0.5 * (a - b)), v = sin(0.5 * phi);
    const double h = fma(sin(a) * sin(b), v * v, u * u);
    return 2 * asin(sqrt(fmin(1.0, fmax(0.0, h))));
}
CDEV int imax(int a, int b) { return a > b ? a : b; }
CDEV int imin(int a, int b) { return a < b ? a : b; }
CDEV int iwrap(int a, int n)
{
    const int r = a % n;
    return r < 0 ? r + n : r;
}
__device__ inline int phi_radius(double t, double r, double radius, double dp, int nc)
{
    if (fabs(t - r) > radius + kAngTol) return -1;
    const double a = sin(t) * sin(r), b = cos(t) * cos(r);
    if (!(a > 0)) return nc;
    const double x = (cos(radius) - b) / a;
    if (x <= -1) return nc;
    if (x >= 1) return 1;
    return imin(nc, (int)ceil(acos(x) / dp) + 1);
}

}  // namespace coh
}  // namespace icephase
