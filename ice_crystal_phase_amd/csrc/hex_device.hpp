// Hexagonal-prism tracer, FP64, device side.
//
// The arithmetic follows the reference CPU tracer operation for operation so
// that each emitted direction and power fraction is bitwise the reference's:
//   prism geometry, sampling      hex_trace/include/ice_crystal/hex_prism.hpp:10-176
//   flux Fresnel split            hex_trace/include/ice_crystal/flux_interface.hpp:14-88
//   linear emission chain, audit  hex_trace/include/ice_crystal/hex_trace_core.hpp:80-170
//   counter RNG                   hex_trace/include/ice_crystal/math.hpp:136-146
// The chain is fused with binning: every emission goes straight to the
// histogram instead of a CSR sample buffer, so each ray is traced once (the
// reference traces it twice, count then replay).
#pragma once

#include "phase_bin.hpp"

namespace icephase {
namespace hex {

struct cplx {
    double re, im;
};
__device__ inline cplx cadd(cplx a, cplx b) { return cplx{a.re + b.re, a.im + b.im}; }
__device__ inline cplx cscale(cplx a, double b) { return cplx{a.re * b, a.im * b}; }
__device__ inline cplx cmul(cplx a, cplx b) { return cplx{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ inline double cnorm2(cplx a) { return a.re * a.re + a.im * a.im; }

// Two-column flux response: column k answers unit input field k.
struct Jones {
    cplx c0x, c0y, c1x, c1y;
    __device__ double power() const
    {
        return 0.5 * (cnorm2(c0x) + cnorm2(c0y) + cnorm2(c1x) + cnorm2(c1y));
    }
    __device__ bool finite() const
    {
        return finite64(c0x.re) && finite64(c0x.im) && finite64(c0y.re) && finite64(c0y.im) &&
               finite64(c1x.re) && finite64(c1x.im) && finite64(c1y.re) && finite64(c1y.im);
    }
    __device__ bool nonzero() const
    {
        return c0x.re != 0 || c0x.im != 0 || c0y.re != 0 || c0y.im != 0 || c1x.re != 0 || c1x.im != 0 ||
               c1y.re != 0 || c1y.im != 0;
    }
};

struct Basis {
    d3 e0, e1;
};

__host__ __device__ inline bool basis_ok(const Basis& f, d3 k)
{
    return is_unit64(k) && is_unit64(f.e0) && is_unit64(f.e1) && fabs(dot(k, f.e0)) <= 128 * DBL_EPSILON &&
           fabs(dot(k, f.e1)) <= 128 * DBL_EPSILON && max_abs(cross(k, f.e0) - f.e1) <= 256 * DBL_EPSILON;
}

__host__ __device__ inline bool frame_of(d3 k, Basis& f)
{
    if (!is_unit64(k)) return false;
    const double ax = fabs(k.x), ay = fabs(k.y), az = fabs(k.z);
    const d3 a = (ax <= ay && ax <= az) ? D3(1, 0, 0) : (ay <= az ? D3(0, 1, 0) : D3(0, 0, 1));
    if (!unit_of(a - k * dot(k, a), f.e0)) return false;
    if (!unit_of(cross(k, f.e0), f.e1)) return false;
    return basis_ok(f, k);
}

// Prism in units of the circumradius; axis +y; faces 0..5 sides, 6 = +y, 7 = -y.
struct Prism {
    double R, L;  // mm
    __host__ __device__ double half() const { return 0.5 * (L / R); }
    __host__ __device__ bool valid() const
    {
        return finite64(R) && finite64(L) && R > 0 && L > 0 && finite64(half()) && half() > 0 &&
               finite64(R * R) && R * R > 0 && finite64(R * L) && R * L > 0;
    }
    __host__ __device__ static d3 normal(unsigned f)
    {
        switch (f) {
        case 0: return D3(1, 0, 0);
        case 1: return D3(0.5, 0, kSqrt3 / 2);
        case 2: return D3(-0.5, 0, kSqrt3 / 2);
        case 3: return D3(-1, 0, 0);
        case 4: return D3(-0.5, 0, -kSqrt3 / 2);
        case 5: return D3(0.5, 0, -kSqrt3 / 2);
        case 6: return D3(0, 1, 0);
        case 7: return D3(0, -1, 0);
        default: return D3(0, 0, 0);
        }
    }
    __host__ __device__ static d3 vertex(unsigned f)
    {
        switch (f % 6) {
        case 0: return D3(kSqrt3 / 2, 0, 0.5);
        case 1: return D3(0, 0, 1);
        case 2: return D3(-kSqrt3 / 2, 0, 0.5);
        case 3: return D3(-kSqrt3 / 2, 0, -0.5);
        case 4: return D3(0, 0, -1);
        default: return D3(kSqrt3 / 2, 0, -0.5);
        }
    }
    __host__ __device__ double offset(unsigned f) const { return f < 6 ? kSqrt3 / 2 : half(); }
    __host__ __device__ double face_area(unsigned f) const { return f < 6 ? R * L : (3 * kSqrt3 / 2) * R * R; }
    __host__ __device__ double projected_area(d3 k) const
    {
        double v = 0, c = 0;  // compensated, same recurrence as the reference Sum
        for (unsigned f = 0; f < 8; ++f) {
            const double x = face_area(f) * gt_max(0, -dot(normal(f), k));
            const double y = x - c, t = v + y;
            c = (t - v) - y;
            v = t;
        }
        return v;
    }
    __host__ __device__ double tolerance(d3 p, double t) const
    {
        return 128 * DBL_EPSILON * gt_max(1, gt_max(half(), gt_max(max_abs(p), fabs(t))));
    }
    __host__ __device__ bool on_face(d3 p, unsigned face) const
    {
        if (face >= 8 || !finite64(p)) return false;
        const double tol = tolerance(p, 0);
        for (unsigned f = 0; f < 8; ++f) {
            const double gap = offset(f) - dot(normal(f), p);
            if (f == face) {
                if (fabs(gap) > tol) return false;
            } else if (!(gap > tol)) {
                return false;
            }
        }
        return true;
    }
};

__device__ inline uint64_t splitmix_finalize(uint64_t x)
{
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}
__device__ inline double open_uniform(uint64_t seed, uint64_t id, unsigned dim)
{
    const uint64_t x = splitmix_finalize(seed + 0x9e3779b97f4a7c15ull * (3 * id + dim + 1));
    return ((double)(x >> 12) + 0.5) * 0x1p-52;
}

struct Entry {
    d3 pos;
    unsigned face;
    bool ok;
};

__device__ inline Entry pick_entry(const Prism& h, d3 k, double area, double u0, double u1, double u2)
{
    Entry s{D3(0, 0, 0), 8, false};
    if (!(area > 0) || !finite64(area)) return s;
    const double target = u0 * area;
    double acc = 0;
    unsigned last = 8;
    for (unsigned f = 0; f < 8; ++f) {
        const double w = h.face_area(f) * gt_max(0, -dot(Prism::normal(f), k));
        if (!(w > 0)) continue;
        last = f;
        acc += w;
        if (target < acc) {
            s.face = f;
            break;
        }
    }
    if (s.face == 8) s.face = last;
    if (s.face == 8) return s;
    if (s.face < 6) {
        const d3 n = Prism::normal(s.face);
        const d3 tangent = D3(-n.z, 0, n.x);
        s.pos = (n * (kSqrt3 / 2) + tangent * (u1 - 0.5)) + D3(0, (2 * u2 - 1) * h.half(), 0);
    } else {
        const double sector = 6 * u1;
        const unsigned j = (unsigned)sector;
        const double r = sqrt(sector - (double)j);
        s.pos = (Prism::vertex(j) * (1 - u2) + Prism::vertex(j + 1) * u2) * r;
        s.pos.y = (s.face == 6 ? 1 : -1) * h.half();
    }
    s.ok = finite64(s.pos);
    return s;
}

struct Exit {
    d3 pos;
    double dist;
    unsigned face;
    int status;  // 0 regular, 2 ambiguous, 3 invalid
};

// CHAIN: (p, start, d) come from a link of the chain that already passed these
// input checks on the same values (the entry point's on_face in hex::enter or
// the previous next_face's status 0; the direction's unit check inside the
// split that produced it), so their outcome is known and they are skipped.
template <bool CHAIN = false>
__device__ inline Exit next_face(const Prism& h, d3 p, d3 d, unsigned start)
{
    Exit hit{D3(0, 0, 0), 0, 8, 3};
    if constexpr (!CHAIN) {
        if (!is_unit64(d) || start >= 8) return hit;
        if (!h.on_face(p, start)) {
            hit.status = 2;
            return hit;
        }
    }
    if (dot(Prism::normal(start), d) > 128 * DBL_EPSILON) return hit;
    double best = DBL_MAX, second = DBL_MAX;
    for (unsigned f = 0; f < 8; ++f) {
        if (f == start) continue;
        const double cosine = dot(Prism::normal(f), d);
        if (!(cosine > 0)) continue;
        const double t = (h.offset(f) - dot(Prism::normal(f), p)) / cosine;
        if (!(t > 0) || !finite64(t)) continue;
        if (t < best) {
            second = best;
            best = t;
            hit.face = f;
        } else if (t < second) {
            second = t;
        }
    }
    if (hit.face == 8) return hit;
    hit.dist = best;
    hit.pos = p + d * best;
    const double tol = h.tolerance(hit.pos, best);
    hit.status = (second - best <= tol || !h.on_face(hit.pos, hit.face)) ? 2 : 0;
    return hit;
}

struct Ray {
    d3 dir;
    Basis f;
    Jones j;
};
struct Split {
    Ray r, t;
    double balance;
    bool ok, has_t, tir;
};

__device__ inline cplx tir_phase(double a, double b)
{
    const double s = gt_max(fabs(a), fabs(b));
    a /= s;
    b /= s;
    const double den = a * a + b * b;
    return cplx{(a * a - b * b) / den, -2 * a * b / den};
}

__device__ inline Jones rotate_into(const Jones& j, const Basis& o, const Basis& n)
{
    const double a = dot(o.e0, n.e0), b = dot(o.e1, n.e0), c = dot(o.e0, n.e1), d = dot(o.e1, n.e1);
    Jones r;
    r.c0x = cadd(cscale(j.c0x, a), cscale(j.c0y, b));
    r.c0y = cadd(cscale(j.c0x, c), cscale(j.c0y, d));
    r.c1x = cadd(cscale(j.c1x, a), cscale(j.c1y, b));
    r.c1y = cadd(cscale(j.c1x, c), cscale(j.c1y, d));
    return r;
}
__device__ inline Jones rows(Jones j, cplx s, cplx p)
{
    j.c0x = cmul(j.c0x, s);
    j.c0y = cmul(j.c0y, p);
    j.c1x = cmul(j.c1x, s);
    j.c1y = cmul(j.c1y, p);
    return j;
}

// n points into the incident medium; flux-normalized transmission. CHAIN: the
// incoming ray is the output of an earlier split, whose success required
// basis_ok and finite Jones entries of these same values, and n is a face
// normal (unit, checked once per record for all eight faces), so those input
// checks are skipped; the refractive indices are still checked.
template <bool CHAIN = false>
__device__ inline Split split(const Ray& in, d3 n, double ni, double nt)
{
    Split b;
    b.ok = false;
    b.has_t = false;
    b.tir = false;
    b.balance = 0;
    if constexpr (!CHAIN) {
        if (!basis_ok(in.f, in.dir) || !is_unit64(n) || !in.j.finite()) return b;
    }
    if (!(ni > 0 && nt > 0 && finite64(ni) && finite64(nt))) return b;
    double ci = -dot(in.dir, n);
    if (ci < 0 || ci > 1 + 128 * DBL_EPSILON) return b;
    if (ci > 1) ci = 1;
    if (ni == nt) {
        b.t = in;
        b.has_t = true;
        if (!unit_of(in.dir + n * (2 * ci), b.r.dir)) return b;
        if (!frame_of(b.r.dir, b.r.f)) return b;
        b.r.j = Jones{};
        b.ok = true;
        return b;
    }
    d3 s = cross(in.dir, n);
    const double scale = max_abs(s);
    const d3 scaled = scale > 0 ? s / scale : D3(0, 0, 0);
    const double sine = scale * sqrt(dot(scaled, scaled));
    if (sine > 0) {
        if (!unit_of(s, s)) return b;
    } else {
        s = in.f.e0;
    }
    d3 pin;
    if (!unit_of(cross(in.dir, s), pin)) return b;
    const Jones j = rotate_into(in.j, in.f, Basis{s, pin});
    const double ns = gt_max(ni, nt), a = ni / ns, c = nt / ns, eta = a / c;
    if (!finite64(eta) || !(a > 0 && c > 0)) return b;
    const double st = eta * sine;
    if (!finite64(st)) return b;
    if (!unit_of(in.dir + n * (2 * ci), b.r.dir)) return b;
    b.r.f.e0 = s;
    if (!unit_of(cross(b.r.dir, s), b.r.f.e1)) return b;
    cplx rs{0, 0}, rp{0, 0};
    if (st > 1) {
        const double inverse = 1 / st;
        const double beta = st * sqrt((1 - inverse) * (1 + inverse));
        rs = tir_phase(a * ci, c * beta);
        rp = tir_phase(c * ci, a * beta);
        b.tir = true;
    } else {
        const double ct = sqrt((1 - st) * (1 + st));
        const double ds = a * ci + c * ct, dp = c * ci + a * ct;
        if (!(ds > 0 && dp > 0)) return b;
        rs = cplx{(a * ci - c * ct) / ds, 0};
        rp = cplx{(c * ci - a * ct) / dp, 0};
        const double numerator = 2 * sqrt(a * ci) * sqrt(c * ct);
        const double ts = numerator / ds, tp = numerator / dp;
        b.t.j = rows(j, cplx{ts, 0}, cplx{tp, 0});
        if (!unit_of((in.dir + n * ci) * eta - n * ct, b.t.dir)) return b;
        b.t.f.e0 = s;
        if (!unit_of(cross(b.t.dir, s), b.t.f.e1)) return b;
        b.has_t = true;
    }
    b.r.j = rows(j, rs, rp);
    const double before = in.j.power();
    const double after = b.r.j.power() + (b.has_t ? b.t.j.power() : 0);
    b.balance = after - before;
    b.ok = basis_ok(b.r.f, b.r.dir) && b.r.j.finite() &&
           (!b.has_t || (basis_ok(b.t.f, b.t.dir) && b.t.j.finite())) && finite64(b.balance) &&
           fabs(b.balance) <= 2048 * DBL_EPSILON * gt_max(before, DBL_MIN);
    return b;
}

enum Status : uint32_t {
    kExhausted = 0,
    kTail = 1,
    kLimit = 2,
    kAmbiguous = 3,
    kInvalid = 4,
    kNumeric = 5,
};

struct Config {
    Prism prism;
    d3 k;
    Basis f;
    double ni, nt, tail, area;
    uint64_t seed;
    uint32_t max_hits;
};

struct RayResult {
    double escaped;  // plain sum (audits use the quanta)
    double unresolved;
    double max_balance;  // largest |R + T - incident| power of the ray's interfaces
    uint32_t status, internal_hits, tir, outputs;
};

// One ray's splitting chain (hex_trace/include/ice_crystal/hex_trace_core.hpp:102-162:
// emit the exterior reflection, then the transmission at every internal
// encounter, continuing the internal reflection), advanced one interface per
// call so that a wave can run its rays' chains in lockstep and bin their
// emissions together. Each call leaves the emission of its interface, if any,
// in (dir, pw) with `sink` set; the operations applied to a ray, and their
// order, are those of the reference chain.
struct Chain {
    Ray ray;  // the internally reflected ray being followed
    d3 p;     // its point on face `cur`
    unsigned cur;
    RayResult out;
    bool alive;  // more encounters to trace
    bool sink;   // (dir, pw) holds an emission of the last call
    d3 dir;
    double pw;
};

// The escaping branch r of an interface: a non-finite, negative, or zero power
// with a nonzero field ends the ray (numeric); a zero power emits nothing.
__device__ inline bool take_emission(Chain& ch, const Ray& r)
{
    const double pw = r.j.power();
    if (!finite64(pw) || pw < 0 || (!pw && r.j.nonzero())) {
        ch.out.status = kNumeric;
        ch.alive = false;
        return false;
    }
    if (pw > 0) {
        ch.sink = true;
        ch.dir = r.dir;
        ch.pw = pw;
        ch.out.escaped += pw;
        ++ch.out.outputs;
    }
    return true;
}

// Entry at a point on a lit face: the exterior reflection is this call's emission.
__device__ inline void enter(Chain& ch, const Config& c, const Entry& e)
{
    ch.out = RayResult{0, 1, 0, kInvalid, 0, 0, 0};
    ch.alive = false;
    ch.sink = false;
    if (!e.ok || e.face >= 8) return;
    if (!c.prism.on_face(e.pos, e.face)) {
        ch.out.status = kAmbiguous;
        return;
    }
    if (!(dot(Prism::normal(e.face), c.k) < 0)) return;
    const Ray inc{c.k, c.f, Jones{cplx{1, 0}, cplx{0, 0}, cplx{0, 0}, cplx{1, 0}}};
    const Split sp = split(inc, Prism::normal(e.face), c.ni, c.nt);
    if (!sp.ok) {
        ch.out.status = kNumeric;
        return;
    }
    ch.out.max_balance = fabs(sp.balance);
    if (!take_emission(ch, sp.r)) return;
    if (!sp.has_t) {
        ch.out.tir = 1;
        ch.out.unresolved = 0;
        ch.out.status = kExhausted;
        return;
    }
    ch.ray = sp.t;
    ch.p = e.pos;
    ch.cur = e.face;
    ch.alive = true;
}

// One internal encounter of a live chain: its transmission is this call's emission.
__device__ inline void encounter(Chain& ch, const Config& c)
{
    ch.sink = false;
    const double power = ch.ray.j.power();
    ch.out.unresolved = power;
    if (!finite64(power) || power < 0 || (!power && ch.ray.j.nonzero())) {
        ch.out.status = kNumeric;
        ch.alive = false;
        return;
    }
    if (power == 0) {
        ch.out.status = kExhausted;
        ch.alive = false;
        return;
    }
    if (power <= c.tail) {
        ch.out.status = kTail;
        ch.alive = false;
        return;
    }
    const Exit hit = next_face<true>(c.prism, ch.p, ch.ray.dir, ch.cur);
    if (hit.status != 0) {
        ch.out.status = hit.status == 2 ? kAmbiguous : kNumeric;
        ch.alive = false;
        return;
    }
    const Split sp = split<true>(ch.ray, -Prism::normal(hit.face), c.nt, c.ni);
    if (!sp.ok) {
        ch.out.status = kNumeric;
        ch.alive = false;
        return;
    }
    ++ch.out.internal_hits;
    if (sp.tir) ++ch.out.tir;
    ch.out.max_balance = gt_max(ch.out.max_balance, fabs(sp.balance));
    if (sp.has_t && !take_emission(ch, sp.t)) return;
    ch.ray = sp.r;
    ch.p = hit.pos;
    ch.cur = hit.face;
}

// A chain still live after max_hits encounters: the remainder's status.
__device__ inline void finish(Chain& ch, const Config& c)
{
    const double rem = ch.ray.j.power();
    ch.out.unresolved = rem;
    if (!finite64(rem) || rem < 0 || (!rem && ch.ray.j.nonzero()))
        ch.out.status = kNumeric;
    else
        ch.out.status = rem == 0 ? kExhausted : (rem <= c.tail ? kTail : kLimit);
    ch.alive = false;
}

// Entry point of sample `id`, drawn from the counter RNG over the lit faces.
__device__ inline Entry sample_entry(const Config& c, uint64_t id)
{
    return pick_entry(c.prism, c.k, c.area, open_uniform(c.seed, id, 0), open_uniform(c.seed, id, 1),
                      open_uniform(c.seed, id, 2));
}

}  // namespace hex
}  // namespace icephase
