/*
 * Oracle (test infrastructure only): the drop tracer algorithm written once for
 * REAL = float (the FP32 mirror of the HIP kernel's arithmetic, operation for
 * operation) and REAL = double (its FP64 twin). Included twice by oracle_drop.c
 * with SFX naming the instantiation. Every operation is one correctly rounded
 * IEEE operation, as in the kernel (built -ffp-contract=off): +, -, *, /,
 * SQRT, and the fused multiply-adds FMA(a, b, c) = a b + c with one rounding
 * (fmaf / fma) at exactly the kernel's places (dot and cross products, the
 * Chebyshev recurrence, points on a ray, the Fresnel and Jones algebra).
 *
 * The reference has no CPU drop histogram (its drop density is an OptiX patch
 * solver). The geometry follows include/rainbow/raindrop_shape.hpp:70-100
 * (r(mu)=1+c0+c1 mu+sum c_n T_n(mu), mu=-y/r, outward gradient), the splitting
 * follows the flux convention of hex_trace/include/ice_crystal/flux_interface.hpp:26-88
 * in single precision, and the emission rule (R at entry, T at every internal
 * encounter, continue R) follows hex_trace/include/ice_crystal/hex_trace_core.hpp:101-163.
 */
/* Event marks for the operation census (census.cpp); empty in the oracle build:
 * 0 ray start, 1 miss, 2 entry done (first split and the external reflection),
 * 3 one internal encounter done, 4 ray end. */
#ifndef CENSUS_MARK
#define CENSUS_MARK(event)
#endif

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define V CAT(v3, SFX)
#define C CAT(cx, SFX)
#define J CAT(jn, SFX)
#define FN(name) CAT(name, SFX)

typedef struct { REAL x, y, z; } V;
typedef struct { REAL re, im; } C;
typedef struct { C c0x, c0y, c1x, c1y; } J;
typedef struct { V dir, e0, e1; J j; } FN(ray_);
typedef struct { FN(ray_) r, t; int valid, has_t, tir; } FN(split_);

static inline V FN(mk_)(REAL x, REAL y, REAL z) { V r = {x, y, z}; return r; }
static inline V FN(scl_)(V a, REAL s) { return FN(mk_)(a.x * s, a.y * s, a.z * s); }
static inline V FN(div_)(V a, REAL s) { return FN(mk_)(a.x / s, a.y / s, a.z / s); }
static inline REAL FN(dot_)(V a, V b) { return FMA(a.z, b.z, FMA(a.y, b.y, a.x * b.x)); }
static inline V FN(crs_)(V a, V b)
{ return FN(mk_)(FMA(a.y, b.z, -(a.z * b.y)), FMA(a.z, b.x, -(a.x * b.z)), FMA(a.x, b.y, -(a.y * b.x))); }
/* o + d t, one rounding per component */
static inline V FN(axpy_)(V d, REAL t, V o) { return FN(mk_)(FMA(d.x, t, o.x), FMA(d.y, t, o.y), FMA(d.z, t, o.z)); }
static inline V FN(unit_)(V a) { return FN(scl_)(a, (REAL)1 / SQRT(FN(dot_)(a, a))); }

/* Chebyshev radius and slope, raindrop_shape.hpp:71-84 recurrence order,
 * stopped after the last nonzero coefficient (terms = its index + 1, at least
 * 2): the omitted terms add c_n T_n = +-0 to finite sums, which changes no value. */
static inline REAL FN(radius_)(const REAL* c, uint32_t terms, REAL mu, REAL* der)
{
    REAL prev = (REAL)1, cur = mu, dprev = (REAL)0, dcur = (REAL)1;
    const REAL mu2 = (REAL)2 * mu;
    REAL rad = FMA(c[1], mu, (REAL)1 + c[0]);
    REAL d = c[1];
    for (int n = 2; n < 8 && (uint32_t)n < terms; ++n) {
        const REAL next = FMA(mu2, cur, -prev);
        const REAL dnext = FMA(mu2, dcur, (REAL)2 * cur) - dprev;
        rad = FMA(c[n], next, rad);
        d = FMA(c[n], dnext, d);
        prev = cur; cur = next; dprev = dcur; dcur = dnext;
    }
    *der = d;
    return rad;
}

typedef struct {
    REAL c[8];
    REAL inner, outer, rmax;
    V ki, e0, e1;
    REAL n_int, n_ext, h, tail;
    uint32_t w, hgt, max_hits, sphere, terms;
} FN(rec_);

/* outward (unnormalized) gradient u - (g - u mu) k with g = (0,-1,0), k = R'(mu)/r:
 * components u_i + (u_i mu) k and u_y + (1 + u_y mu) k */
static inline V FN(outward_)(V u, REAL mu, REAL k)
{ return FN(mk_)(FMA(u.x * mu, k, u.x), FMA(FMA(u.y, mu, (REAL)1), k, u.y), FMA(u.z * mu, k, u.z)); }

static inline REAL FN(surf_)(const FN(rec_)* R, V p)
{
    const REAL r = SQRT(FN(dot_)(p, p));
    if (R->sphere) return r - (REAL)1;
    if (r < R->inner * (REAL)0.25) return r - R->inner;
    const REAL mu = -(p.y * ((REAL)1 / r));
    REAL der;
    return r - FN(radius_)(R->c, R->terms, mu, &der);
}
/* outward (unnormalized) gradient of f at p */
static inline V FN(grad_)(const FN(rec_)* R, V p)
{
    const REAL r = SQRT(FN(dot_)(p, p));
    const REAL ir = (REAL)1 / r;
    const V u = FN(scl_)(p, ir);
    if (R->sphere) return u;
    const REAL mu = -u.y;
    REAL der;
    (void)FN(radius_)(R->c, R->terms, mu, &der);
    return FN(outward_)(u, mu, der * ir);
}

/* f(p) and grad f(p) . d from one evaluation of r, mu and the Chebyshev
 * recurrence (the kernel's surface_slope): the same operations as surf_ and
 * grad_ above, so the same values. */
static inline REAL FN(surf_slope_)(const FN(rec_)* R, V p, V d, REAL* slope)
{
    const REAL r = SQRT(FN(dot_)(p, p));
    const REAL ir = (REAL)1 / r;
    const V u = FN(scl_)(p, ir);
    if (R->sphere) {
        *slope = FN(dot_)(u, d);
        return r - (REAL)1;
    }
    const REAL mu = -u.y;
    REAL der;
    const REAL rad = FN(radius_)(R->c, R->terms, mu, &der);
    *slope = FN(dot_)(FN(outward_)(u, mu, der * ir), d);
    return r < R->inner * (REAL)0.25 ? r - R->inner : r - rad;
}

/* Safeguarded Newton on f(o + t d) over [a,b] where f(a) has sign sa (+1/-1)
 * and f(b) the opposite sign. Returns t. 48 iterations cap. */
static REAL FN(newton_)(const FN(rec_)* R, V o, V d, REAL a, REAL b, int a_pos, REAL x)
{
    const REAL tol = TOL;
    REAL prev = (REAL)0; /* length of the previous Newton step, 0 after a bisection */
    if (!(x > a && x < b)) x = FMA(b - a, (REAL)0.5, a);
    for (int it = 0; it < 48; ++it) {
        const V p = FN(axpy_)(d, x, o);
        REAL slope;
        const REAL fx = FN(surf_slope_)(R, p, d, &slope);
        if (fx == (REAL)0) return x;
        if ((fx > (REAL)0) == (a_pos != 0)) a = x; else b = x;
        const REAL scale = FABS(x) > (REAL)1 ? FABS(x) : (REAL)1;
        if (b - a <= tol * scale) return x;
        const REAL cand = x - fx / slope;
        const int inside = cand > a && cand < b;
        const REAL step = FABS(cand - x);
        /* a step below the tolerance ends the solve even when rounding left a
         * bracket end on the root and the candidate is not strictly inside */
        if (step <= tol * scale) return inside ? cand : x;
        /* two quadratically contracting steps predict the next, K step^2 with
         * K = step / prev^2: accept the candidate when that is below tol / 8 */
        if (inside && prev > (REAL)0 && step <= prev * (REAL)0.25 &&
            step * step * step <= (REAL)0.125 * tol * scale * prev * prev)
            return cand;
        prev = inside ? step : (REAL)0;
        x = inside ? cand : FMA(b - a, (REAL)0.5, a);
    }
    return x;
}

/* Entry: ray line o + t ki with o = e0 u + e1 v (closest approach at t=0).
 * Returns 1 and writes t on a hit, 0 on a miss. */
static int FN(entry_)(const FN(rec_)* R, V o, REAL u, REAL v, REAL* tout)
{
    const REAL rho2 = FMA(v, v, u * u);
    if (R->sphere) {
        if (!(rho2 < (REAL)1)) return 0;
        *tout = -SQRT((REAL)1 - rho2);
        return 1;
    }
    const REAL out2 = R->rmax * R->rmax;
    if (!(rho2 < out2)) return 0;
    const REAL ta = -SQRT(out2 - rho2);
    REAL tm = (REAL)0;
    REAL g0;
    const REAL fm = FN(surf_slope_)(R, o, R->ki, &g0);
    if (!(fm < (REAL)0)) {
        /* foot of the perpendicular is outside: minimize f along the line by the
         * secant method on f'(t) = grad f . ki from the foot (the inside of a convex
         * body along a line is the single interval around that minimum); any
         * iterate with f < 0 is an interior point, convergence with f >= 0 is a miss */
        REAL t0 = (REAL)0;
        REAL t1 = -g0 * SQRT(rho2);
        int found = 0;
        for (int it = 0; it < 8; ++it) {
            if (t1 < ta) t1 = ta;
            if (t1 > -ta) t1 = -ta;
            const V p1 = FN(axpy_)(R->ki, t1, o);
            REAL g1;
            const REAL f1 = FN(surf_slope_)(R, p1, R->ki, &g1);
            if (f1 < (REAL)0) { tm = t1; found = 1; break; }
            if (g1 == g0) break;
            const REAL t2 = t1 - g1 * (t1 - t0) / (g1 - g0);
            t0 = t1; g0 = g1;
            if (FABS(t2 - t1) <= MIN_STEP) break;
            t1 = t2;
        }
        if (!found) return 0;
    }
    /* sphere-like first guess from the mean radius */
    const REAL rm = (REAL)1 + R->c[0];
    const REAL g2 = FMA(rm, rm, -rho2);
    const REAL x0 = g2 > (REAL)0 ? -SQRT(g2) : FMA(tm - ta, (REAL)0.5, ta);
    *tout = FN(newton_)(R, o, R->ki, ta, tm, 1, x0);
    return 1;
}

/* Exit from surface point p along inward d. Returns 0 when unresolved. */
static int FN(exit_)(const FN(rec_)* R, V p, V d, REAL* tout)
{
    const REAL b = FN(dot_)(p, d);
    if (R->sphere) {
        const REAL c = FN(dot_)(p, p) - (REAL)1;
        REAL disc = FMA(b, b, -c);
        if (disc < (REAL)0) disc = (REAL)0;
        *tout = -b + SQRT(disc);
        return 1;
    }
    const REAL cb = FN(dot_)(p, p) - R->rmax * R->rmax;
    REAL disc = FMA(b, b, -cb);
    if (disc < (REAL)0) disc = (REAL)0;
    const REAL tb = -b + SQRT(disc);
    const REAL ts = SELF_OFFSET;
    if (!(ts < tb)) return 0;
    if (!(FN(surf_)(R, FN(axpy_)(d, ts, p)) < (REAL)0)) return 0;
    *tout = FN(newton_)(R, p, d, ts, tb, 0, (REAL)-2 * b);
    return 1;
}

static inline C FN(cm_)(C a, C b) { C r = {FMA(a.re, b.re, -(a.im * b.im)), FMA(a.re, b.im, a.im * b.re)}; return r; }
static inline C FN(cr_)(C a, REAL b) { C r = {a.re * b, a.im * b}; return r; }
/* x a + y b */
static inline C FN(cl_)(C x, REAL a, C y, REAL b) { C r = {FMA(y.re, b, x.re * a), FMA(y.im, b, x.im * a)}; return r; }
static inline REAL FN(n2_)(C a) { return FMA(a.im, a.im, a.re * a.re); }
static inline REAL FN(pow_)(const J* j)
{ return (REAL)0.5 * (((FN(n2_)(j->c0x) + FN(n2_)(j->c0y)) + FN(n2_)(j->c1x)) + FN(n2_)(j->c1y)); }
static inline C FN(tir_)(REAL a, REAL b)
{
    const REAL s = FABS(a) > FABS(b) ? FABS(a) : FABS(b);
    const REAL x = a / s, y = b / s, den = FMA(y, y, x * x);
    C r = {FMA(x, x, -(y * y)) / den, ((REAL)-2 * x * y) / den};
    return r;
}

/* Flux-normalized Fresnel split; n points into the incident medium. */
static FN(split_) FN(split_fn_)(const FN(ray_)* in, V n, REAL ni, REAL nt)
{
    FN(split_) o;
    memset(&o, 0, sizeof o);
    const V d = in->dir;
    REAL ci = -FN(dot_)(d, n);
    if (!(ci >= (REAL)-1e-3 && ni > (REAL)0 && nt > (REAL)0)) return o;
    if (ci < (REAL)0) ci = (REAL)0;
    if (ci > (REAL)1) ci = (REAL)1;
    V s = FN(crs_)(d, n);
    const REAL s2 = FN(dot_)(s, s);
    const REAL sine = SQRT(s2);
    if (s2 > TINY_S2) s = FN(scl_)(s, (REAL)1 / sine);
    else s = FN(unit_)(FN(axpy_)(d, -FN(dot_)(d, in->e0), in->e0));
    const V pin = FN(crs_)(d, s); /* d, s unit and orthogonal */
    const REAL a = FN(dot_)(in->e0, s), b = FN(dot_)(in->e1, s);
    const REAL cc = FN(dot_)(in->e0, pin), dd = FN(dot_)(in->e1, pin);
    J j;
    j.c0x = FN(cl_)(in->j.c0x, a, in->j.c0y, b);
    j.c0y = FN(cl_)(in->j.c0x, cc, in->j.c0y, dd);
    j.c1x = FN(cl_)(in->j.c1x, a, in->j.c1y, b);
    j.c1y = FN(cl_)(in->j.c1x, cc, in->j.c1y, dd);
    const REAL eta = ni / nt;
    const REAL st = eta * sine;
    o.r.dir = FN(unit_)(FN(axpy_)(n, (REAL)2 * ci, d));
    o.r.e0 = s;
    o.r.e1 = FN(crs_)(o.r.dir, s);
    C rs, rp;
    if (st > (REAL)1) {
        const REAL beta = SQRT((st - (REAL)1) * (st + (REAL)1));
        rs = FN(tir_)(ni * ci, nt * beta);
        rp = FN(tir_)(nt * ci, ni * beta);
        o.tir = 1;
    } else {
        const REAL ct = SQRT(((REAL)1 - st) * ((REAL)1 + st));
        const REAL ds = FMA(nt, ct, ni * ci), dp = FMA(ni, ct, nt * ci);
        if (!(ds > (REAL)0 && dp > (REAL)0)) return o;
        rs.re = FMA(ni, ci, -(nt * ct)) / ds; rs.im = (REAL)0;
        rp.re = FMA(nt, ci, -(ni * ct)) / dp; rp.im = (REAL)0;
        const REAL num = ((REAL)2 * SQRT(ni * ci)) * SQRT(nt * ct);
        const REAL ts = num / ds, tp = num / dp;
        o.t.dir = FN(unit_)(FN(axpy_)(n, -ct, FN(scl_)(FN(axpy_)(n, ci, d), eta)));
        o.t.e0 = s;
        o.t.e1 = FN(crs_)(o.t.dir, s);
        o.t.j.c0x = FN(cr_)(j.c0x, ts); o.t.j.c0y = FN(cr_)(j.c0y, tp);
        o.t.j.c1x = FN(cr_)(j.c1x, ts); o.t.j.c1y = FN(cr_)(j.c1y, tp);
        o.has_t = 1;
    }
    o.r.j.c0x = FN(cm_)(j.c0x, rs); o.r.j.c0y = FN(cm_)(j.c0y, rp);
    o.r.j.c1x = FN(cm_)(j.c1x, rs); o.r.j.c1y = FN(cm_)(j.c1y, rp);
    o.valid = 1;
    return o;
}

static inline V FN(outward_normal_)(const FN(rec_)* R, V p) { return FN(unit_)(FN(grad_)(R, p)); }

/* Trace the ray through o = e0 u + e1 v along ki; samples carry id.
 * emit(dir, power) for each escaping branch. */
static orc_ray_audit FN(trace_origin_)(const FN(rec_)* R, uint64_t id, REAL u, REAL v, orc_sample* out,
                                         uint64_t cap, uint64_t* n_out, int* overflow)
{
    orc_ray_audit a;
    memset(&a, 0, sizeof a);
    a.status = 8; /* miss */
    CENSUS_MARK(0);
    const V o = FN(axpy_)(R->e1, v, FN(scl_)(R->e0, u));
    REAL t;
    if (!FN(entry_)(R, o, u, v, &t)) { a.unresolved = 0; CENSUS_MARK(1); return a; }
    V p = FN(axpy_)(R->ki, t, o);
    FN(ray_) ray;
    ray.dir = R->ki; ray.e0 = R->e0; ray.e1 = R->e1;
    memset(&ray.j, 0, sizeof ray.j);
    ray.j.c0x.re = (REAL)1; ray.j.c1y.re = (REAL)1;
    double escaped = 0;
    FN(split_) sp = FN(split_fn_)(&ray, FN(outward_normal_)(R, p), R->n_ext, R->n_int);
    if (!sp.valid) { a.status = ST_NUMERIC_D; a.unresolved = 1; return a; }
#define EMIT(RAY, KIND)                                                          \
    do {                                                                          \
        const REAL pw = FN(pow_)(&(RAY).j);                                      \
        if (pw > (REAL)0) {                                                       \
            if (out) {                                                            \
                if (*n_out >= cap) { *overflow = 1; return a; }                   \
                orc_sample* smp = &out[*n_out];                                   \
                smp->dir[0] = (double)(RAY).dir.x; smp->dir[1] = (double)(RAY).dir.y; \
                smp->dir[2] = (double)(RAY).dir.z; smp->power = (double)pw;       \
                smp->length_mm = 0; smp->id = id; smp->kind = (KIND);             \
                smp->entry_face = smp->exit_face = 8; smp->reflections = a.internal_hits; \
            }                                                                     \
            ++*n_out; ++a.outputs; escaped += (double)pw;                          \
        }                                                                         \
    } while (0)
    EMIT(sp.r, 0u);
    CENSUS_MARK(2);
    if (!sp.has_t) { a.tir_count = 1; a.status = 0; a.escaped = escaped; a.unresolved = 0; CENSUS_MARK(4); return a; }
    ray = sp.t;
    a.status = 2; /* interaction limit unless something ends it first */
    for (uint32_t hit = 0; hit < R->max_hits; ++hit) {
        const REAL pw = FN(pow_)(&ray.j);
        if (pw == (REAL)0) { a.status = 0; break; }
        if (pw <= R->tail) { a.status = 1; break; }
        REAL tx;
        if (!FN(exit_)(R, p, ray.dir, &tx)) { a.status = 3; break; }
        const V q = FN(axpy_)(ray.dir, tx, p);
        const V nout = FN(outward_normal_)(R, q);
        sp = FN(split_fn_)(&ray, FN(scl_)(nout, (REAL)-1), R->n_int, R->n_ext);
        if (!sp.valid) { a.status = ST_NUMERIC_D; break; }
        ++a.internal_hits;
        if (sp.tir) ++a.tir_count;
        if (sp.has_t) EMIT(sp.t, 1u);
        ray = sp.r;
        p = q;
        CENSUS_MARK(3);
    }
#undef EMIT
    const REAL rem = FN(pow_)(&ray.j);
    if (a.status == 2 && rem == (REAL)0) a.status = 0;
    a.escaped = escaped;
    a.unresolved = (double)rem;
    CENSUS_MARK(4);
    return a;
}

/* Trace grid ray id: the cell centre (ix, iy) of the w x hgt grid over [-h,h]^2. */
static orc_ray_audit FN(trace_ray_)(const FN(rec_)* R, uint64_t id, orc_sample* out, uint64_t cap,
                                      uint64_t* n_out, int* overflow)
{
    const uint32_t ix = (uint32_t)(id % R->w), iy = (uint32_t)(id / R->w);
    const REAL u = ((REAL)(2u * ix + 1u) / (REAL)R->w - (REAL)1) * R->h;
    const REAL v = ((REAL)(2u * iy + 1u) / (REAL)R->hgt - (REAL)1) * R->h;
    return FN(trace_origin_)(R, id, u, v, out, cap, n_out, overflow);
}

#undef V
#undef C
#undef J
#undef FN
#undef CAT
#undef CAT2
