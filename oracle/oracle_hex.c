/*
 * Oracle (test infrastructure only; see oracle.h): FP64 restatement of the
 * reference hexagonal-prism tracer, phase binning, fixed-point histogram and
 * host CDF. Operation order follows the cited reference lines so that results
 * agree bitwise with the compiled reference wherever libm agrees.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define PI_D 3.141592653589793238462643383279502884
#define SQRT3_D 1.732050807568877293527446341505872367

typedef struct { double x, y, z; } v3;
typedef struct { double re, im; } cpx;
typedef struct { cpx x, y; } fld;
typedef struct { fld c[2]; } jones;
typedef struct { v3 e0, e1; } frm;

static inline v3 mk(double x, double y, double z) { v3 r = {x, y, z}; return r; }
static inline v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 neg(v3 a) { return mk(-a.x, -a.y, -a.z); }
static inline v3 scl(v3 a, double s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline v3 dvd(v3 a, double s) { return mk(a.x / s, a.y / s, a.z / s); }
static inline double dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 crs(v3 a, v3 b)
{ return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static inline int fin(double x) { return x >= -DBL_MAX && x <= DBL_MAX; }
static inline double mx(double a, double b) { return a > b ? a : b; }
static inline int vfin(v3 a) { return fin(a.x) && fin(a.y) && fin(a.z); }
static inline double maxabs(v3 a) { return mx(fabs(a.x), mx(fabs(a.y), fabs(a.z))); }

/* math.hpp:38-47 */
static int nrm(v3 a, v3* r)
{
    const double s = maxabs(a);
    if (!(s > 0) || !fin(s)) return 0;
    a = dvd(a, s);
    const double n = sqrt(dot(a, a));
    if (!(n > 0) || !fin(n)) return 0;
    *r = dvd(a, n);
    return vfin(*r);
}
static int is_unit(v3 a) { return vfin(a) && fabs(dot(a, a) - 1.0) <= 128 * DBL_EPSILON; }
/* math.hpp:51-60 */
static int frame_valid(frm f, v3 k)
{
    return is_unit(k) && is_unit(f.e0) && is_unit(f.e1) &&
           fabs(dot(k, f.e0)) <= 128 * DBL_EPSILON && fabs(dot(k, f.e1)) <= 128 * DBL_EPSILON &&
           maxabs(sub(crs(k, f.e0), f.e1)) <= 256 * DBL_EPSILON;
}
/* math.hpp:61-69 */
static int make_frame(v3 k, frm* f)
{
    if (!is_unit(k)) return 0;
    const double ax = fabs(k.x), ay = fabs(k.y), az = fabs(k.z);
    const v3 a = (ax <= ay && ax <= az) ? mk(1, 0, 0) : (ay <= az ? mk(0, 1, 0) : mk(0, 0, 1));
    if (!nrm(sub(a, scl(k, dot(k, a))), &f->e0)) return 0;
    if (!nrm(crs(k, f->e0), &f->e1)) return 0;
    return frame_valid(*f, k);
}

static inline cpx cadd(cpx a, cpx b) { cpx r = {a.re + b.re, a.im + b.im}; return r; }
static inline cpx cmulr(cpx a, double b) { cpx r = {a.re * b, a.im * b}; return r; }
static inline cpx cmul(cpx a, cpx b) { cpx r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; return r; }
static inline double cn2(cpx a) { return a.re * a.re + a.im * a.im; }
static jones jid(void) { jones j; memset(&j, 0, sizeof j); j.c[0].x.re = 1; j.c[1].y.re = 1; return j; }
static double jpow(const jones* j)
{ return 0.5 * (cn2(j->c[0].x) + cn2(j->c[0].y) + cn2(j->c[1].x) + cn2(j->c[1].y)); }
static int jfin(const jones* j)
{
    for (int i = 0; i < 2; ++i)
        if (!fin(j->c[i].x.re) || !fin(j->c[i].x.im) || !fin(j->c[i].y.re) || !fin(j->c[i].y.im)) return 0;
    return 1;
}
static int jnonzero(const jones* j)
{
    for (int i = 0; i < 2; ++i)
        if (j->c[i].x.re != 0 || j->c[i].x.im != 0 || j->c[i].y.re != 0 || j->c[i].y.im != 0) return 1;
    return 0;
}
/* math.hpp:109-118 */
static jones change_frame(const jones* j, frm o, frm n)
{
    jones r;
    const double a = dot(o.e0, n.e0), b = dot(o.e1, n.e0), c = dot(o.e0, n.e1), d = dot(o.e1, n.e1);
    for (int k = 0; k < 2; ++k) {
        r.c[k].x = cadd(cmulr(j->c[k].x, a), cmulr(j->c[k].y, b));
        r.c[k].y = cadd(cmulr(j->c[k].x, c), cmulr(j->c[k].y, d));
    }
    return r;
}
static jones scale_rows(jones j, cpx s, cpx p)
{
    for (int k = 0; k < 2; ++k) { j.c[k].x = cmul(j.c[k].x, s); j.c[k].y = cmul(j.c[k].y, p); }
    return j;
}

typedef struct { double value, correction; } kahan;
static inline void kadd(kahan* s, double x)
{
    const double y = x - s->correction, t = s->value + y;
    s->correction = (t - s->value) - y;
    s->value = t;
}

/* ---- prism (hex_prism.hpp:10-51) ---- */
typedef struct { double R, L; } prism;
static inline double half_len(prism h) { return 0.5 * (h.L / h.R); }
static int prism_valid(prism h)
{
    return fin(h.R) && fin(h.L) && h.R > 0 && h.L > 0 && fin(half_len(h)) && half_len(h) > 0 &&
           fin(h.R * h.R) && h.R * h.R > 0 && fin(h.R * h.L) && h.R * h.L > 0;
}
static v3 pnormal(unsigned f)
{
    switch (f) {
    case 0: return mk(1, 0, 0);
    case 1: return mk(0.5, 0, SQRT3_D / 2);
    case 2: return mk(-0.5, 0, SQRT3_D / 2);
    case 3: return mk(-1, 0, 0);
    case 4: return mk(-0.5, 0, -SQRT3_D / 2);
    case 5: return mk(0.5, 0, -SQRT3_D / 2);
    case 6: return mk(0, 1, 0);
    case 7: return mk(0, -1, 0);
    default: return mk(0, 0, 0);
    }
}
static v3 pvertex(unsigned f)
{
    switch (f % 6) {
    case 0: return mk(SQRT3_D / 2, 0, 0.5);
    case 1: return mk(0, 0, 1);
    case 2: return mk(-SQRT3_D / 2, 0, 0.5);
    case 3: return mk(-SQRT3_D / 2, 0, -0.5);
    case 4: return mk(0, 0, -1);
    default: return mk(SQRT3_D / 2, 0, -0.5);
    }
}
static inline double poffset(prism h, unsigned f) { return f < 6 ? SQRT3_D / 2 : half_len(h); }
static inline double face_area(prism h, unsigned f)
{ return f < 6 ? h.R * h.L : (3 * SQRT3_D / 2) * h.R * h.R; }
static double proj_area(prism h, v3 k)
{
    kahan a = {0, 0};
    for (unsigned f = 0; f < 8; ++f) kadd(&a, face_area(h, f) * mx(0, -dot(pnormal(f), k)));
    return a.value;
}
double orc_hex_projected_area(double radius_mm, double length_mm, const double k[3])
{
    prism h = {radius_mm, length_mm};
    return proj_area(h, mk(k[0], k[1], k[2]));
}
static inline double geom_tol(prism h, v3 p, double t)
{ return 128 * DBL_EPSILON * mx(1, mx(half_len(h), mx(maxabs(p), fabs(t)))); }

typedef struct { v3 pos; unsigned face; int valid; } entry_s;
/* hex_prism.hpp:64-98 */
static entry_s sample_entry(prism h, v3 k, double u0, double u1, double u2)
{
    entry_s s; s.pos = mk(0, 0, 0); s.face = 8; s.valid = 0;
    if (!prism_valid(h) || !is_unit(k) || !(u0 >= 0 && u0 < 1 && u1 >= 0 && u1 < 1 && u2 >= 0 && u2 < 1)) return s;
    const double a = proj_area(h, k);
    if (!(a > 0) || !fin(a)) return s;
    const double target = u0 * a;
    double acc = 0;
    unsigned last = 8;
    for (unsigned f = 0; f < 8; ++f) {
        const double w = face_area(h, f) * mx(0, -dot(pnormal(f), k));
        if (!(w > 0)) continue;
        last = f; acc += w;
        if (target < acc) { s.face = f; break; }
    }
    if (s.face == 8) s.face = last;
    if (s.face == 8) return s;
    if (s.face < 6) {
        const v3 n = pnormal(s.face);
        const v3 tangent = mk(-n.z, 0, n.x);
        s.pos = add(add(scl(n, SQRT3_D / 2), scl(tangent, u1 - 0.5)), mk(0, (2 * u2 - 1) * half_len(h), 0));
    } else {
        const double sector = 6 * u1;
        const unsigned j = (unsigned)sector;
        const double r = sqrt(sector - (double)j);
        s.pos = scl(add(scl(pvertex(j), 1 - u2), scl(pvertex(j + 1), u2)), r);
        s.pos.y = (s.face == 6 ? 1 : -1) * half_len(h);
    }
    s.valid = vfin(s.pos);
    return s;
}

enum { HIT_REGULAR = 0, HIT_MISS = 1, HIT_AMBIG = 2, HIT_INVALID = 3 };
typedef struct { v3 pos; double dist; unsigned face; int status; } bhit;

/* hex_prism.hpp:108-120 */
static int strict_face_point(prism h, v3 p, unsigned face)
{
    if (face >= 8 || !vfin(p)) return 0;
    const double tol = geom_tol(h, p, 0);
    for (unsigned f = 0; f < 8; ++f) {
        const double gap = poffset(h, f) - dot(pnormal(f), p);
        if (f == face) { if (fabs(gap) > tol) return 0; }
        else if (!(gap > tol)) return 0;
    }
    return 1;
}
/* hex_prism.hpp:124-148 */
static bhit next_boundary(prism h, v3 p, v3 d, unsigned start)
{
    bhit hit; hit.pos = mk(0, 0, 0); hit.dist = 0; hit.face = 8; hit.status = HIT_INVALID;
    if (!prism_valid(h) || !is_unit(d) || start >= 8) return hit;
    if (!strict_face_point(h, p, start)) { hit.status = HIT_AMBIG; return hit; }
    if (dot(pnormal(start), d) > 128 * DBL_EPSILON) return hit;
    double best = DBL_MAX, second = DBL_MAX;
    for (unsigned f = 0; f < 8; ++f) {
        if (f == start) continue;
        const double cosine = dot(pnormal(f), d);
        if (!(cosine > 0)) continue;
        const double t = (poffset(h, f) - dot(pnormal(f), p)) / cosine;
        if (!(t > 0) || !fin(t)) continue;
        if (t < best) { second = best; best = t; hit.face = f; }
        else if (t < second) second = t;
    }
    if (hit.face == 8) return hit;
    hit.dist = best; hit.pos = add(p, scl(d, best));
    const double tol = geom_tol(h, hit.pos, best);
    if (second - best <= tol || !strict_face_point(h, hit.pos, hit.face)) hit.status = HIT_AMBIG;
    else hit.status = HIT_REGULAR;
    return hit;
}

/* ---- flux Fresnel (flux_interface.hpp:14-88) ---- */
typedef struct { v3 dir; frm f; jones j; } fray;
typedef struct { fray refl, trans; double balance; int valid, has_t, tir; } branches;

static cpx tir_phase(double a, double b)
{
    const double s = mx(fabs(a), fabs(b));
    a /= s; b /= s;
    const double den = a * a + b * b;
    cpx r = {(a * a - b * b) / den, -2 * a * b / den};
    return r;
}
static branches split_flux(const fray* ray, v3 n, double ni, double nt)
{
    branches b;
    memset(&b, 0, sizeof b);
    if (!frame_valid(ray->f, ray->dir) || !is_unit(n) || !jfin(&ray->j) ||
        !(ni > 0 && nt > 0 && fin(ni) && fin(nt))) return b;
    double ci = -dot(ray->dir, n);
    if (ci < 0 || ci > 1 + 128 * DBL_EPSILON) return b;
    if (ci > 1) ci = 1;
    if (ni == nt) {
        b.trans = *ray; b.has_t = 1;
        b.refl.dir = add(ray->dir, scl(n, 2 * ci));
        if (!nrm(b.refl.dir, &b.refl.dir)) return b;
        if (!make_frame(b.refl.dir, &b.refl.f)) return b;
        b.valid = 1;
        return b;
    }
    v3 s = crs(ray->dir, n);
    const double scale = maxabs(s);
    const v3 scaled = scale > 0 ? dvd(s, scale) : mk(0, 0, 0);
    const double sine = scale * sqrt(dot(scaled, scaled));
    if (sine > 0) { if (!nrm(s, &s)) return b; }
    else s = ray->f.e0;
    v3 pin;
    if (!nrm(crs(ray->dir, s), &pin)) return b;
    frm sp = {s, pin};
    const jones j = change_frame(&ray->j, ray->f, sp);
    const double ns = mx(ni, nt), a = ni / ns, c = nt / ns, eta = a / c;
    if (!fin(eta) || !(a > 0 && c > 0)) return b;
    const double st = eta * sine;
    if (!fin(st)) return b;
    if (!nrm(add(ray->dir, scl(n, 2 * ci)), &b.refl.dir)) return b;
    b.refl.f.e0 = s;
    if (!nrm(crs(b.refl.dir, s), &b.refl.f.e1)) return b;
    cpx rs = {0, 0}, rp = {0, 0};
    if (st > 1) {
        const double inverse = 1 / st;
        const double beta = st * sqrt((1 - inverse) * (1 + inverse));
        rs = tir_phase(a * ci, c * beta);
        rp = tir_phase(c * ci, a * beta);
        b.tir = 1;
    } else {
        const double ct = sqrt((1 - st) * (1 + st));
        const double ds = a * ci + c * ct, dp = c * ci + a * ct;
        if (!(ds > 0 && dp > 0)) return b;
        rs.re = (a * ci - c * ct) / ds; rs.im = 0;
        rp.re = (c * ci - a * ct) / dp; rp.im = 0;
        const double numerator = 2 * sqrt(a * ci) * sqrt(c * ct);
        const double ts = numerator / ds, tp = numerator / dp;
        const cpx cts = {ts, 0}, ctp = {tp, 0};
        b.trans.j = scale_rows(j, cts, ctp);
        if (!nrm(sub(scl(add(ray->dir, scl(n, ci)), eta), scl(n, ct)), &b.trans.dir)) return b;
        b.trans.f.e0 = s;
        if (!nrm(crs(b.trans.dir, s), &b.trans.f.e1)) return b;
        b.has_t = 1;
    }
    b.refl.j = scale_rows(j, rs, rp);
    const double before = jpow(&ray->j);
    const double after = jpow(&b.refl.j) + (b.has_t ? jpow(&b.trans.j) : 0);
    b.balance = after - before;
    b.valid = frame_valid(b.refl.f, b.refl.dir) && jfin(&b.refl.j) &&
              (!b.has_t || (frame_valid(b.trans.f, b.trans.dir) && jfin(&b.trans.j))) &&
              fin(b.balance) && fabs(b.balance) <= 2048 * DBL_EPSILON * mx(before, DBL_MIN);
    return b;
}

/* ---- RNG (math.hpp:136-146) ---- */
static inline uint64_t mix64(uint64_t x)
{
    x = (x ^ (x >> 30)) * UINT64_C(0xbf58476d1ce4e5b9);
    x = (x ^ (x >> 27)) * UINT64_C(0x94d049bb133111eb);
    return x ^ (x >> 31);
}
static inline double uni(uint64_t seed, uint64_t id, unsigned dim)
{
    const uint64_t x = mix64(seed + UINT64_C(0x9e3779b97f4a7c15) * (3 * id + dim + 1));
    return ((double)(x >> 12) + 0.5) * 0x1p-52;
}

/* ---- trace (hex_trace_core.hpp:8-170) ---- */
enum { ST_EXHAUSTED = 0, ST_TAIL = 1, ST_LIMIT = 2, ST_AMBIG = 3, ST_INVALID = 4, ST_NUMERIC = 5, ST_OVERFLOW = 6 };

typedef struct {
    prism h; v3 k; frm f; double ni, nt, tail; uint64_t seed, total; uint32_t max_hits;
} hexcfg;

static int hex_setup(const orc_hex_settings* s, hexcfg* c)
{
    c->h.R = s->radius_mm; c->h.L = s->length_mm;
    if (!nrm(mk(s->ki[0], s->ki[1], s->ki[2]), &c->k) || !make_frame(c->k, &c->f)) return -1;
    c->ni = s->n_ext; c->nt = s->n_int; c->tail = s->tail;
    c->seed = s->seed; c->total = s->total; c->max_hits = s->max_hits;
    const double a = proj_area(c->h, c->k);
    int ok = prism_valid(c->h) && frame_valid(c->f, c->k) && fin(c->ni) && fin(c->nt) && c->ni > 0 &&
             c->nt > 0 && fin(c->tail) && c->tail >= 0 && c->tail < 1 && c->total > 0 &&
             c->max_hits > 0 && fin(a) && a > 0;
    return ok ? 0 : -1;
}

typedef struct { orc_sample* buf; uint64_t cap, n; int overflow; } sink;

static int emit(sink* sk, orc_ray_audit* a, kahan* esc, uint64_t id, const fray* ray, unsigned entry,
                unsigned exitf, unsigned refl, unsigned kind, double length)
{
    const double power = jpow(&ray->j);
    if (!fin(power) || power < 0 || (!power && jnonzero(&ray->j))) return ST_NUMERIC;
    if (power == 0) return ST_EXHAUSTED;
    if (sk->buf) {
        if (sk->n >= sk->cap) { sk->overflow = 1; return ST_OVERFLOW; }
        orc_sample* o = &sk->buf[sk->n];
        o->dir[0] = ray->dir.x; o->dir[1] = ray->dir.y; o->dir[2] = ray->dir.z;
        o->power = power; o->length_mm = length; o->id = id;
        o->entry_face = entry; o->exit_face = exitf; o->reflections = refl; o->kind = kind;
    }
    sk->n++;
    kadd(esc, power);
    a->outputs++;
    return ST_EXHAUSTED;
}
static orc_ray_audit finish(orc_ray_audit a, kahan esc, double residual, int status)
{
    a.escaped = esc.value; a.unresolved = residual;
    a.balance = esc.value + residual - 1; a.status = (uint32_t)status;
    return a;
}
/* hex_trace_core.hpp:102-162 (trace_from_entry) */
static orc_ray_audit trace_entry(const hexcfg* c, uint64_t id, entry_s e, sink* sk)
{
    orc_ray_audit a; memset(&a, 0, sizeof a); a.unresolved = 1; a.status = ST_INVALID;
    kahan esc = {0, 0};
    if (id >= c->total || !e.valid || e.face >= 8) return finish(a, esc, 1, ST_INVALID);
    if (!strict_face_point(c->h, e.pos, e.face)) return finish(a, esc, 1, ST_AMBIG);
    if (!(dot(pnormal(e.face), c->k) < 0)) return finish(a, esc, 1, ST_INVALID);
    fray inc; inc.dir = c->k; inc.f = c->f; inc.j = jid();
    branches sp = split_flux(&inc, pnormal(e.face), c->ni, c->nt);
    if (!sp.valid) return finish(a, esc, 1, ST_NUMERIC);
    a.max_interface = fabs(sp.balance);
    int em = emit(sk, &a, &esc, id, &sp.refl, e.face, e.face, 0, 0, 0);
    if (em != ST_EXHAUSTED) return finish(a, esc, 1, em);
    if (!sp.has_t) { a.tir_count = 1; return finish(a, esc, 0, ST_EXHAUSTED); }
    fray ray = sp.trans;
    v3 p = e.pos;
    unsigned cur = e.face, refl = 0;
    kahan length = {0, 0};
    for (unsigned enc = 0; enc < c->max_hits; ++enc) {
        const double power = jpow(&ray.j);
        if (!fin(power) || power < 0 || (!power && jnonzero(&ray.j))) return finish(a, esc, power, ST_NUMERIC);
        if (power == 0) return finish(a, esc, 0, ST_EXHAUSTED);
        if (power <= c->tail) return finish(a, esc, power, ST_TAIL);
        const bhit hit = next_boundary(c->h, p, ray.dir, cur);
        if (hit.status != HIT_REGULAR)
            return finish(a, esc, power, hit.status == HIT_AMBIG ? ST_AMBIG : ST_NUMERIC);
        const double segment = hit.dist * c->h.R;
        if (!fin(segment) || segment < 0) return finish(a, esc, power, ST_NUMERIC);
        kadd(&length, segment);
        if (!fin(length.value)) return finish(a, esc, power, ST_NUMERIC);
        sp = split_flux(&ray, neg(pnormal(hit.face)), c->nt, c->ni);
        if (!sp.valid) return finish(a, esc, power, ST_NUMERIC);
        ++a.internal_hits;
        if (sp.tir) ++a.tir_count;
        a.max_interface = mx(a.max_interface, fabs(sp.balance));
        if (sp.has_t) {
            em = emit(sk, &a, &esc, id, &sp.trans, e.face, hit.face, refl, 1, length.value);
            if (em != ST_EXHAUSTED) return finish(a, esc, power, em);
        }
        ray = sp.refl; p = hit.pos; cur = hit.face; ++refl;
    }
    const double rem = jpow(&ray.j);
    if (!fin(rem) || rem < 0 || (!rem && jnonzero(&ray.j))) return finish(a, esc, rem, ST_NUMERIC);
    return finish(a, esc, rem, rem == 0 ? ST_EXHAUSTED : (rem <= c->tail ? ST_TAIL : ST_LIMIT));
}
/* hex_trace_core.hpp:163-170 (trace_sample) */
static orc_ray_audit trace_one(const hexcfg* c, uint64_t id, sink* sk)
{
    const entry_s e = sample_entry(c->h, c->k, uni(c->seed, id, 0), uni(c->seed, id, 1), uni(c->seed, id, 2));
    return trace_entry(c, id, e, sk);
}

int64_t orc_hex_trace(const orc_hex_settings* s, uint64_t first, uint64_t count,
                      orc_sample* samples, uint64_t cap, orc_ray_audit* audits)
{
    hexcfg c;
    if (hex_setup(s, &c)) return -2;
    sink sk = {samples, cap, 0, 0};
    for (uint64_t i = 0; i < count; ++i) {
        orc_ray_audit a = trace_one(&c, first + i, &sk);
        if (audits) audits[i] = a;
        if (sk.overflow) return -1;
    }
    return (int64_t)sk.n;
}

/* ---- grid & bins ---- */
int orc_u_edges(uint32_t n, double* e)
{
    if (n == 0) return -1;
    e[0] = 0; e[n] = 1;
    for (uint32_t i = 1; i < n; ++i) {
        const uint32_t k = i < n - i ? i : n - i;
        const double s = sin((0.5 * PI_D) * (double)k / (double)n);
        e[i] = i <= n / 2u ? s * s : 1.0 - s * s;
    }
    for (uint32_t i = 0; i < n; ++i) if (!(e[i + 1] > e[i])) return -2;
    return 0;
}

int orc_normalize3(const double in[3], double out[3])
{
    v3 r;
    if (!nrm(mk(in[0], in[1], in[2]), &r)) return -1;
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
    return 0;
}

int orc_make_frame(const double k_in[3], orc_frame* f)
{
    v3 k; frm fr;
    if (!nrm(mk(k_in[0], k_in[1], k_in[2]), &k) || !make_frame(k, &fr)) return -1;
    f->ki[0] = k.x; f->ki[1] = k.y; f->ki[2] = k.z;
    f->e0[0] = fr.e0.x; f->e0[1] = fr.e0.y; f->e0[2] = fr.e0.z;
    f->e1[0] = fr.e1.x; f->e1[1] = fr.e1.y; f->e1[2] = fr.e1.z;
    return 0;
}

int orc_drop_frame(const double k_in[3], orc_frame* f)
{
    v3 k;
    if (!nrm(mk(k_in[0], k_in[1], k_in[2]), &k)) return -1;
    const v3 axis = fabs(k.z) < 0.9 ? mk(0, 0, 1) : mk(1, 0, 0);
    v3 e0, e1;
    if (!nrm(sub(axis, scl(k, dot(k, axis))), &e0)) return -1;
    if (!nrm(crs(k, e0), &e1)) return -1;
    f->ki[0] = k.x; f->ki[1] = k.y; f->ki[2] = k.z;
    f->e0[0] = e0.x; f->e0[1] = e0.y; f->e0[2] = e0.z;
    f->e1[0] = e1.x; f->e1[1] = e1.y; f->e1[2] = e1.z;
    return 0;
}

/* phase_math.hpp:37-81: everything after the unit-length checks. */
/* the phase_bin rule (phase_math.hpp:37-81) in double: oracle_bin_rule.h */
#define BREAL double
#define BFN(name) name##d
#define BVAL(x) (x)
#define BHYPOT hypot
#define BATAN2 atan2
#define BFLOOR floor
#define BSQRT sqrt
#define BFABS fabs
#define BFMIN fmin
#define BFMAX fmax
#include "oracle_bin_rule.h"
#undef BREAL
#undef BFN
#undef BVAL
#undef BHYPOT
#undef BATAN2
#undef BFLOOR
#undef BSQRT
#undef BFABS
#undef BFMIN
#undef BFMAX

orc_bin orc_phase_bin(const double* ue, uint32_t nt, uint32_t np, const orc_frame* g, const double dir[3])
{
    orc_bin b; memset(&b, 0, sizeof b);
    const v3 d = mk(dir[0], dir[1], dir[2]);
    if (!ue || !nt || !np || !is_unit(d)) return b;
    const v3 e0 = mk(g->e0[0], g->e0[1], g->e0[2]), e1 = mk(g->e1[0], g->e1[1], g->e1[2]);
    const v3 ki = mk(g->ki[0], g->ki[1], g->ki[2]);
    const double x = dot(e0, d), y = dot(e1, d), z = dot(ki, d);
    const double n = sqrt(x * x + y * y + z * z);
    if (!(n > 0) || !fin(n) || fabs(n - 1) > 512 * DBL_EPSILON) return b;
    return bin_core_d(ue, nt, np, x, y, z, n);
}

orc_bin orc_phase_bin_raw(const double* ue, uint32_t nt, uint32_t np, const orc_frame* g, const double dir[3])
{
    return bin_raw_d(ue, nt, np, g, dir);
}

uint32_t orc_quantum_bits(uint64_t rays)
{
    uint32_t lg = 0;
    while (lg < 64 && (UINT64_C(1) << lg) < rays) ++lg;
    int q = 62 - (int)lg;
    if (q > 52) q = 52;
    if (q < 8) q = 8;
    return (uint32_t)q;
}

static inline uint64_t quantize(double p, uint32_t qbits) { return (uint64_t)llrint(ldexp(p, (int)qbits)); }

int orc_hex_hist(const orc_hex_settings* s, uint32_t nt, uint32_t np, uint32_t qbits,
                 uint64_t first, uint64_t count, uint64_t* bins, orc_hist_stats* st)
{
    hexcfg c;
    if (hex_setup(s, &c)) return -2;
    double* ue = (double*)malloc(sizeof(double) * (nt + 1));
    if (orc_u_edges(nt, ue)) { free(ue); return -3; }
    orc_frame g;
    /* single orientation: the histogram frame is the tracer's own frame */
    g.ki[0] = c.k.x; g.ki[1] = c.k.y; g.ki[2] = c.k.z;
    g.e0[0] = c.f.e0.x; g.e0[1] = c.f.e0.y; g.e0[2] = c.f.e0.z;
    g.e1[0] = c.f.e1.x; g.e1[1] = c.f.e1.y; g.e1[2] = c.f.e1.z;
    memset(bins, 0, sizeof(uint64_t) * (size_t)nt * np);
    memset(st, 0, sizeof *st);
    orc_sample buf[4096];
    for (uint64_t i = 0; i < count; ++i) {
        sink sk = {buf, 4096, 0, 0};
        const orc_ray_audit a = trace_one(&c, first + i, &sk);
        if (sk.overflow) { free(ue); return -1; }
        st->status[a.status & 7]++;
        st->internal_hits += a.internal_hits;
        if (a.status != ST_INVALID && a.status != ST_AMBIG) st->rays_hit++;
        uint64_t qe = 0;
        for (uint64_t k = 0; k < sk.n; ++k) {
            const uint64_t q = quantize(buf[k].power, qbits);
            qe += q;
            const orc_bin b = orc_phase_bin(ue, nt, np, &g, buf[k].dir);
            st->samples++;
            if (!b.valid) { st->invalid++; continue; }
            st->snapped += (uint64_t)b.snapped;
            if (b.pole > 0) st->q_forward += q;
            else if (b.pole < 0) st->q_backward += q;
            else bins[b.index] += q;
        }
        st->q_escaped += qe;
        st->q_unresolved += quantize(a.unresolved, qbits);
    }
    free(ue);
    return 0;
}

/* One orientation node of a composed record (hex_phase/src/phase_generate.cpp:44-53,
 * hex_phase/src/phase_table.cpp:52-62): the ensemble direction s->ki is
 * normalized, the node is traced in the body frame as normalize(R^T k), every
 * outgoing direction is binned as R d in make_frame(k), and the emission quanta
 * are rint(power * qfactor * 2^qbits). bins/st are accumulated, not zeroed. */
int orc_hex_hist_node(const orc_hex_settings* s, const double rot[9], double qfactor, uint32_t nt, uint32_t np,
                      uint32_t qbits, uint64_t* bins, orc_hist_stats* st)
{
    v3 ke;
    if (!nrm(mk(s->ki[0], s->ki[1], s->ki[2]), &ke)) return -2;
    const double* m = rot;
    orc_hex_settings body = *s;
    body.ki[0] = (m[0] * ke.x + m[3] * ke.y) + m[6] * ke.z;
    body.ki[1] = (m[1] * ke.x + m[4] * ke.y) + m[7] * ke.z;
    body.ki[2] = (m[2] * ke.x + m[5] * ke.y) + m[8] * ke.z;
    hexcfg c;
    if (hex_setup(&body, &c)) return -2;
    frm fe;
    if (!make_frame(ke, &fe)) return -2;
    orc_frame g;
    g.ki[0] = ke.x; g.ki[1] = ke.y; g.ki[2] = ke.z;
    g.e0[0] = fe.e0.x; g.e0[1] = fe.e0.y; g.e0[2] = fe.e0.z;
    g.e1[0] = fe.e1.x; g.e1[1] = fe.e1.y; g.e1[2] = fe.e1.z;
    double* ue = (double*)malloc(sizeof(double) * (nt + 1));
    if (orc_u_edges(nt, ue)) { free(ue); return -3; }
    orc_sample buf[4096];
    for (uint64_t i = 0; i < c.total; ++i) {
        sink sk = {buf, 4096, 0, 0};
        const orc_ray_audit a = trace_one(&c, i, &sk);
        if (sk.overflow) { free(ue); return -1; }
        st->status[a.status & 7]++;
        st->internal_hits += a.internal_hits;
        if (a.status != ST_INVALID && a.status != ST_AMBIG) st->rays_hit++;
        uint64_t qe = 0;
        for (uint64_t k = 0; k < sk.n; ++k) {
            const uint64_t q = quantize(buf[k].power * qfactor, qbits);
            qe += q;
            const double* d = buf[k].dir;
            const double r[3] = {(m[0] * d[0] + m[1] * d[1]) + m[2] * d[2], (m[3] * d[0] + m[4] * d[1]) + m[5] * d[2],
                                 (m[6] * d[0] + m[7] * d[1]) + m[8] * d[2]};
            const orc_bin b = orc_phase_bin(ue, nt, np, &g, r);
            st->samples++;
            if (!b.valid) { st->invalid++; continue; }
            st->snapped += (uint64_t)b.snapped;
            if (b.pole > 0) st->q_forward += q;
            else if (b.pole < 0) st->q_backward += q;
            else bins[b.index] += q;
        }
        st->q_escaped += qe;
        st->q_unresolved += quantize(a.unresolved * qfactor, qbits);
    }
    free(ue);
    return 0;
}

/* Projected area of the prism along the node's body direction normalize(R^T k)
 * (the operations orc_hex_hist_node traces with). */
double orc_hex_node_area(const orc_hex_settings* s, const double rot[9])
{
    v3 ke;
    if (!nrm(mk(s->ki[0], s->ki[1], s->ki[2]), &ke)) return -1;
    const double* m = rot;
    orc_hex_settings body = *s;
    body.ki[0] = (m[0] * ke.x + m[3] * ke.y) + m[6] * ke.z;
    body.ki[1] = (m[1] * ke.x + m[4] * ke.y) + m[7] * ke.z;
    body.ki[2] = (m[2] * ke.x + m[5] * ke.y) + m[8] * ke.z;
    hexcfg c;
    if (hex_setup(&body, &c)) return -1;
    return proj_area(c.h, c.k);
}

/* ---- Fraunhofer lobe of the projected aperture ---- */
/* F(q) of a counter-clockwise polygon by the divergence theorem over its edges:
 * F = (i/|q|^2) sum (q1 e_y - q2 e_x) exp(-i q.m) sinc(q.e/2). */
void orc_aperture_transform(const double* vx, const double* vy, uint32_t nv, double q1, double q2, double* re,
                            double* im)
{
    double sr = 0, si = 0;
    for (uint32_t k = 0; k < nv; ++k) {
        const uint32_t k1 = k + 1 == nv ? 0 : k + 1;
        const double ex = vx[k1] - vx[k], ey = vy[k1] - vy[k];
        const double mx_ = 0.5 * (vx[k1] + vx[k]), my_ = 0.5 * (vy[k1] + vy[k]);
        const double qn = q1 * ey - q2 * ex;
        const double h = 0.5 * (q1 * ex + q2 * ey);
        const double sinc = fabs(h) < 1e-8 ? 1.0 - h * h / 6.0 : sin(h) / h;
        const double ph = -(q1 * mx_ + q2 * my_);
        const double w = qn * sinc;
        sr += w * cos(ph);
        si += w * sin(ph);
    }
    const double qq = q1 * q1 + q2 * q2;
    *re = -si / qq;
    *im = sr / qq;
}

static int cmp_pt(const void* a, const void* b)
{
    const double* p = (const double*)a;
    const double* q = (const double*)b;
    if (p[0] != q[0]) return p[0] < q[0] ? -1 : 1;
    if (p[1] != q[1]) return p[1] < q[1] ? -1 : 1;
    return 0;
}

/* Aperture of the prism seen along k (normalized inside): convex hull of the 12
 * projected vertices in make_frame(k)'s (e0, e1), counter-clockwise. Returns the
 * vertex count (0 on failure); p1/p2 receive e0/e1. */
uint32_t orc_hex_aperture(double radius_mm, double length_mm, const double k_in[3], double* vx, double* vy,
                          double p1[3], double p2[3])
{
    v3 k;
    frm f;
    if (!nrm(mk(k_in[0], k_in[1], k_in[2]), &k) || !make_frame(k, &f)) return 0;
    double pts[12][2];
    int n = 0;
    for (unsigned fc = 0; fc < 6; ++fc)
        for (int sg = -1; sg <= 1; sg += 2) {
            const v3 v = pvertex(fc);
            const v3 p = mk(radius_mm * v.x, 0.5 * sg * length_mm, radius_mm * v.z);
            pts[n][0] = dot(p, f.e0);
            pts[n][1] = dot(p, f.e1);
            ++n;
        }
    qsort(pts, 12, sizeof pts[0], cmp_pt);
    const double sc = radius_mm > length_mm ? radius_mm : length_mm;
    const double tol = 1e-12 * sc * sc;
    double hull[24][2];
    int h = 0;
#define TURN(o, a, b) (((a)[0] - (o)[0]) * ((b)[1] - (o)[1]) - ((a)[1] - (o)[1]) * ((b)[0] - (o)[0]))
    for (int i = 0; i < 12; ++i) {
        while (h >= 2 && TURN(hull[h - 2], hull[h - 1], pts[i]) <= tol) --h;
        hull[h][0] = pts[i][0]; hull[h][1] = pts[i][1]; ++h;
    }
    for (int i = 10, t = h + 1; i >= 0; --i) {
        while (h >= t && TURN(hull[h - 2], hull[h - 1], pts[i]) <= tol) --h;
        hull[h][0] = pts[i][0]; hull[h][1] = pts[i][1]; ++h;
    }
#undef TURN
    const int nv = h - 1;
    if (nv < 3 || nv > 8) return 0;
    for (int i = 0; i < nv; ++i) { vx[i] = hull[i][0]; vy[i] = hull[i][1]; }
    p1[0] = f.e0.x; p1[1] = f.e0.y; p1[2] = f.e0.z;
    p2[0] = f.e1.x; p2[1] = f.e1.y; p2[2] = f.e1.z;
    return (uint32_t)nv;
}

static const double LGX[4] = {-0.8611363115940526, -0.3399810435848563, 0.3399810435848563, 0.8611363115940526};
static const double LGW[4] = {0.3478548451374538, 0.6521451548625461, 0.6521451548625461, 0.3478548451374538};

static int panels(double x)
{
    const double c = ceil(x);
    return c < 1 ? 1 : (c > 64 ? 64 : (int)c);
}

/* Lobe mass (mm^2) in cell (i, j) of an nt x np grid with binning frame g:
 * int |F(q)|^2 / lambda^2 dOmega over the cell, q = k0 (d.p1, d.p2), by 4-point
 * Gauss-Legendre in theta and phi on sub-panels no wider than half a fringe. */
double orc_lobe_cell(const double* vx, const double* vy, uint32_t nv, double area, double dmax, const double p1[3],
                     const double p2[3], double lambda_mm, const orc_frame* g, uint32_t nt, uint32_t np, uint32_t i,
                     uint32_t j)
{
    const double k0 = 2 * PI_D / lambda_mm;
    const double t0 = PI_D * i / nt, t1 = PI_D * (i + 1) / nt;
    const double f0 = -PI_D + 2 * PI_D * j / np, f1 = -PI_D + 2 * PI_D * (j + 1) / np;
    const double fringe = lambda_mm / dmax;
    const double thi = t1 < 0.5 * PI_D ? t1 : 0.5 * PI_D;
    const int nts = panels(2 * (t1 - t0) / fringe), nps = panels(2 * sin(thi) * (f1 - f0) / fringe);
    const double ht = (t1 - t0) / nts, hp = (f1 - f0) / nps;
    double total = 0;
    for (int a = 0; a < nts; ++a) {
        double rowsum = 0;
        for (int gt = 0; gt < 4; ++gt) {
            const double th = t0 + ht * a + 0.5 * ht * (1 + LGX[gt]);
            const double st = sin(th), ct = cos(th);
            double colsum = 0;
            for (int b = 0; b < nps; ++b)
                for (int gp = 0; gp < 4; ++gp) {
                    const double ph = f0 + hp * b + 0.5 * hp * (1 + LGX[gp]);
                    const double cx = st * cos(ph), cy = st * sin(ph);
                    double d[3];
                    for (int c = 0; c < 3; ++c) d[c] = cx * g->e0[c] + cy * g->e1[c] + ct * g->ki[c];
                    const double q1 = k0 * (d[0] * p1[0] + d[1] * p1[1] + d[2] * p1[2]);
                    const double q2 = k0 * (d[0] * p2[0] + d[1] * p2[1] + d[2] * p2[2]);
                    double dens;
                    if (sqrt(q1 * q1 + q2 * q2) * dmax < 1e-4) {
                        dens = area * area;
                    } else {
                        double re, im;
                        orc_aperture_transform(vx, vy, nv, q1, q2, &re, &im);
                        dens = re * re + im * im;
                    }
                    colsum += LGW[gp] * dens / (lambda_mm * lambda_mm);
                }
            rowsum += LGW[gt] * st * colsum;
        }
        total += rowsum;
    }
    return total * (0.5 * ht) * (0.5 * hp);
}

int orc_hex_masses(const orc_hex_settings* s, uint32_t nt, uint32_t np,
                   double* mass, double* forward, double* backward)
{
    hexcfg c;
    if (hex_setup(s, &c)) return -2;
    double* ue = (double*)malloc(sizeof(double) * (nt + 1));
    if (orc_u_edges(nt, ue)) { free(ue); return -3; }
    orc_frame g;
    g.ki[0] = c.k.x; g.ki[1] = c.k.y; g.ki[2] = c.k.z;
    g.e0[0] = c.f.e0.x; g.e0[1] = c.f.e0.y; g.e0[2] = c.f.e0.z;
    g.e1[0] = c.f.e1.x; g.e1[1] = c.f.e1.y; g.e1[2] = c.f.e1.z;
    const size_t n = (size_t)nt * np;
    kahan* sum = (kahan*)calloc(n, sizeof(kahan));
    kahan fw = {0, 0}, bw = {0, 0};
    const double scale = proj_area(c.h, c.k) / (double)c.total;
    orc_sample buf[4096];
    for (uint64_t i = 0; i < c.total; ++i) {
        sink sk = {buf, 4096, 0, 0};
        (void)trace_one(&c, i, &sk);
        if (sk.overflow) { free(ue); free(sum); return -1; }
        for (uint64_t k = 0; k < sk.n; ++k) {
            const double w = scale * buf[k].power;
            const orc_bin b = orc_phase_bin(ue, nt, np, &g, buf[k].dir);
            if (!b.valid) { free(ue); free(sum); return -4; }
            if (b.pole > 0) kadd(&fw, w);
            else if (b.pole < 0) kadd(&bw, w);
            else kadd(&sum[b.index], w);
        }
    }
    const double f = fw.value / np, bk = bw.value / np;
    for (size_t k = 0; k < n; ++k) {
        kahan v = sum[k];
        const size_t i = k / np;
        if (i == 0) kadd(&v, f);
        if (i + 1 == nt) kadd(&v, bk);
        mass[k] = v.value;
    }
    if (forward) *forward = fw.value;
    if (backward) *backward = bw.value;
    free(ue); free(sum);
    return 0;
}

void orc_bins_to_mass(const uint64_t* bins, uint64_t qf, uint64_t qb, uint32_t nt, uint32_t np, double* mass)
{
    const double fs = (double)qf / (double)np, bs = (double)qb / (double)np;
    for (uint32_t i = 0; i < nt; ++i)
        for (uint32_t j = 0; j < np; ++j) {
            double m = (double)bins[(size_t)i * np + j];
            if (i == 0) m = m + fs;
            if (i + 1 == nt) m = m + bs;
            mass[(size_t)i * np + j] = m;
        }
}

/* ---- CDF (phase_table.cpp:111-187) ---- */
int orc_build_cdf(const double* mass, uint32_t nt, uint32_t np, uint32_t ont, uint32_t onp,
                  double max_l1, double max_lost, double max_tv,
                  double* phi, double* theta, orc_cdf_report* r)
{
    memset(r, 0, sizeof *r);
    if (!nt || !np || !ont || !onp || ont > nt || onp > np || nt % ont || np % onp) { r->status = 1; return -1; }
    double* e = (double*)malloc(sizeof(double) * (nt + 1));
    double* ce = (double*)malloc(sizeof(double) * (ont + 1));
    if (orc_u_edges(nt, e)) { free(e); free(ce); r->status = 1; return -1; }
    const size_t n = (size_t)nt * np, on = (size_t)ont * onp;
    /* measure_mass (phase_table.cpp:83-97) */
    kahan tot = {0, 0}, ax = {0, 0};
    for (size_t k = 0; k < n; ++k) {
        const double w = mass[k];
        if (!fin(w) || w < 0) { free(e); free(ce); r->status = 2; return -1; }
        const size_t i = k / np;
        kadd(&tot, w); kadd(&ax, w * (1 - e[i] - e[i + 1]));
    }
    if (!(tot.value > 0) || !fin(tot.value) || !fin(ax.value)) { free(e); free(ce); r->status = 3; return -1; }
    const uint32_t dt = nt / ont, dp = np / onp;
    for (uint32_t i = 0; i <= ont; ++i) ce[i] = e[(size_t)i * dt];
    kahan* sums = (kahan*)calloc(on, sizeof(kahan));
    for (uint32_t i = 0; i < nt; ++i)
        for (uint32_t j = 0; j < np; ++j) kadd(&sums[(size_t)(i / dt) * onp + j / dp], mass[(size_t)i * np + j]);
    double* coarse = (double*)malloc(sizeof(double) * on);
    for (size_t k = 0; k < on; ++k) coarse[k] = sums[k].value;
    free(sums);
    kahan ctot = {0, 0}, cax = {0, 0};
    for (size_t k = 0; k < on; ++k) {
        const size_t i = k / onp;
        kadd(&ctot, coarse[k]); kadd(&cax, coarse[k] * (1 - ce[i] - ce[i + 1]));
    }
    r->g_source = ax.value / tot.value;
    r->aggregation_error = fabs(ctot.value / tot.value - 1);
    kahan tv = {0, 0};
    for (uint32_t i = 0; i < nt; ++i)
        for (uint32_t j = 0; j < np; ++j) {
            const uint32_t ci = i / dt, cj = j / dp;
            const double fraction = (e[i + 1] - e[i]) / (ce[ci + 1] - ce[ci]) / dp;
            kadd(&tv, fabs(mass[(size_t)i * np + j] / tot.value - coarse[(size_t)ci * onp + cj] / ctot.value * fraction));
        }
    r->coarsening_tv = 0.5 * tv.value;
    if (!fin(r->coarsening_tv) || r->coarsening_tv > max_tv || r->aggregation_error > 1e-12) {
        free(e); free(ce); free(coarse); r->status = 4; return -1;
    }
    double largest = coarse[0];
    for (size_t k = 1; k < on; ++k) if (coarse[k] > largest) largest = coarse[k];
    double* cols = (double*)malloc(sizeof(double) * onp);
    memset(phi, 0, sizeof(double) * (onp + 1));
    memset(theta, 0, sizeof(double) * (size_t)onp * (ont + 1));
    for (uint32_t j = 0; j < onp; ++j) {
        const size_t base = (size_t)j * (ont + 1);
        double s = 0;
        for (uint32_t i = 0; i < ont; ++i) {
            const double w = coarse[(size_t)i * onp + j] / largest;
            if (coarse[(size_t)i * onp + j] > 0 && !(w > 0)) { r->status = 5; goto fail; }
            s += w;
            theta[base + i + 1] = s;
        }
        cols[j] = s;
        phi[j + 1] = phi[j] + s;
    }
    {
        const double total = phi[onp];
        if (!(total > 0) || !fin(total)) { r->status = 6; goto fail; }
        for (uint32_t j = 0; j < onp; ++j) {
            const size_t base = (size_t)j * (ont + 1);
            for (uint32_t i = 1; i < ont; ++i)
                theta[base + i] = cols[j] > 0 ? theta[base + i] / cols[j] : ce[i];
            theta[base + ont] = 1;
        }
        for (uint32_t j = 1; j < onp; ++j) phi[j] /= total;
        phi[onp] = 1;
        kahan l1 = {0, 0}, lost = {0, 0}, rec = {0, 0}, axial = {0, 0};
        for (uint32_t j = 0; j < onp; ++j)
            for (uint32_t i = 0; i < ont; ++i) {
                const size_t base = (size_t)j * (ont + 1) + i;
                const double q = (phi[j + 1] - phi[j]) * (theta[base + 1] - theta[base]);
                const double p = coarse[(size_t)i * onp + j] / ctot.value;
                if (!fin(q) || q < 0 || phi[j + 1] < phi[j] || theta[base + 1] < theta[base]) { r->status = 7; goto fail; }
                if (coarse[(size_t)i * onp + j] > 0 && !(p > 0)) { r->status = 8; goto fail; }
                const double err = fabs(p - q);
                kadd(&l1, err);
                if (err > r->max_cell_error) r->max_cell_error = err;
                if (p > 0 && q == 0) { kadd(&lost, p); r->lost_cells++; }
                kadd(&rec, q);
                kadd(&axial, q * (1 - ce[i] - ce[i + 1]));
            }
        r->l1_error = l1.value; r->lost_mass = lost.value; r->cdf_mass = rec.value;
        r->g_stored = axial.value / rec.value;
        if (r->l1_error > max_l1 || r->lost_mass > max_lost || fabs(r->cdf_mass - 1) > 1e-10 ||
            !fin(r->g_stored) || !(r->g_stored > -1 && r->g_stored < 1)) { r->status = 9; goto fail; }
    }
    free(e); free(ce); free(coarse); free(cols);
    return 0;
fail:
    free(e); free(ce); free(coarse); free(cols);
    return -1;
}

/* ---- deterministic atlas node (hex_atlas/src/atlas.cpp) ---------------------
 * Node direction and frame (:69-88), the GL 2x3 area rule (:11-51), and the
 * CPU trace (:139-154): each area ray is traced from its entry, its weighted
 * emissions are binned into a per-ray compensated column (pole emissions spread
 * as w/np over the end-cap row, atlas.hpp:62-73), and the columns are added to
 * the node's compensated cells in ray order. */
static int atlas_node_geometry(uint32_t T, uint32_t P, uint32_t i, uint32_t j, v3* k, frm* f, double* theta,
                               double* phi)
{
    const double pi = 3.141592653589793238462643383279502884;
    *theta = (pi / 2) * i / T;
    *phi = (pi / 6) * j / P;
    const double st = i == 0 ? 0 : (i == T ? 1 : sin(*theta));
    const double ct = i == 0 ? 1 : (i == T ? 0 : cos(*theta));
    const double sp = j == 0 ? 0 : (j == P ? 0.5 : sin(*phi));
    const double cp = j == 0 ? 1 : (j == P ? SQRT3_D / 2 : cos(*phi));
    v3 k1, e;
    if (!nrm(mk(st * cp, ct, st * sp), &k1) || !nrm(k1, k)) return -1;
    if (!nrm(mk(ct * cp, -st, ct * sp), &e)) return -1;
    f->e0 = e;
    f->e1 = crs(*k, e);
    return frame_valid(*f, *k) ? 0 : -1;
}

int64_t orc_atlas_rule(const orc_atlas_settings* s, uint32_t i, uint32_t j, double* out, uint64_t cap)
{
    prism h = {s->radius_mm, s->length_mm};
    v3 k; frm f; double theta, phi;
    if (!prism_valid(h) || i > s->theta_intervals || j > s->phi_intervals ||
        atlas_node_geometry(s->theta_intervals, s->phi_intervals, i, j, &k, &f, &theta, &phi))
        return -2;
    const uint32_t n = s->subdivisions;
    const double ga = 1 / sqrt(3.0), gb = sqrt(3.0 / 5.0);
    const double u[2] = {(1 - ga) / 2, (1 + ga) / 2};
    const double v[3] = {(1 - gb) / 2, 0.5, (1 + gb) / 2};
    const double vw[3] = {5.0 / 18, 4.0 / 9, 5.0 / 18};
    kahan total = {0, 0};
    uint64_t m = 0;
    for (unsigned face = 0; face < 8; ++face) {
        const double projected = face_area(h, face) * mx(0, -dot(pnormal(face), k));
        if (projected == 0) continue;
        const unsigned sectors = face < 6 ? 1 : 6;
        for (unsigned sector = 0; sector < sectors; ++sector)
            for (uint32_t a = 0; a < n; ++a)
                for (uint32_t b = 0; b < n; ++b)
                    for (unsigned gi = 0; gi < 2; ++gi)
                        for (unsigned gj = 0; gj < 3; ++gj) {
                            const double x = (a + u[gi]) / n, y = (b + v[gj]) / n;
                            v3 p;
                            if (face < 6) {
                                const v3 nf = pnormal(face);
                                const v3 t = mk(-nf.z, 0, nf.x);
                                p = add(add(scl(nf, SQRT3_D / 2), scl(t, x - 0.5)), mk(0, (2 * y - 1) * half_len(h), 0));
                            } else {
                                p = scl(add(scl(pvertex(sector), 1 - y), scl(pvertex(sector + 1), y)), sqrt(x));
                                p.y = (face == 6 ? 1 : -1) * half_len(h);
                            }
                            const double w = (projected / sectors / (double)n / n) * 0.5 * vw[gj];
                            if (!(w > 0) || !fin(w) || !strict_face_point(h, p, face)) return -3;
                            if (out) {
                                if (m >= cap) return -1;
                                double* o = out + 5 * m;
                                o[0] = p.x; o[1] = p.y; o[2] = p.z; o[3] = (double)face; o[4] = w;
                            }
                            kadd(&total, w);
                            ++m;
                        }
    }
    if (!m || fabs(total.value / proj_area(h, k) - 1) > 1e-12) return -4;
    return (int64_t)m;
}

int orc_atlas_node(const orc_atlas_settings* s, uint32_t i, uint32_t j, double* mass, orc_atlas_report* rep)
{
    prism h = {s->radius_mm, s->length_mm};
    v3 k; frm f; double theta, phi;
    if (atlas_node_geometry(s->theta_intervals, s->phi_intervals, i, j, &k, &f, &theta, &phi)) return -2;
    const int64_t m = orc_atlas_rule(s, i, j, NULL, 0);
    if (m <= 0) return -3;
    double* rule = (double*)malloc((size_t)m * 5 * sizeof(double));
    const uint32_t nt = s->theta_count, np = s->phi_count;
    const size_t cells = (size_t)nt * np;
    kahan* acc = (kahan*)calloc(cells, sizeof(kahan));
    kahan* col = (kahan*)calloc(cells, sizeof(kahan));
    double* edges = (double*)malloc((nt + 1) * sizeof(double));
    const uint64_t cap = 2 * (uint64_t)s->max_hits + 4;
    orc_sample* smp = (orc_sample*)malloc(cap * sizeof(orc_sample));
    int rc = 0;
    if (!rule || !acc || !col || !edges || !smp || orc_atlas_rule(s, i, j, rule, (uint64_t)m) != m ||
        orc_u_edges(nt, edges)) {
        rc = -4;
        goto done;
    }
    hexcfg c;
    c.h = h; c.k = k; c.f = f; c.ni = s->n_ext; c.nt = s->n_int; c.tail = s->tail;
    c.seed = 0; c.total = (uint64_t)m; c.max_hits = s->max_hits;
    orc_frame g;
    g.ki[0] = k.x; g.ki[1] = k.y; g.ki[2] = k.z;
    g.e0[0] = f.e0.x; g.e0[1] = f.e0.y; g.e0[2] = f.e0.z;
    g.e1[0] = f.e1.x; g.e1[1] = f.e1.y; g.e1[2] = f.e1.z;
    memset(rep, 0, sizeof *rep);
    rep->ki[0] = k.x; rep->ki[1] = k.y; rep->ki[2] = k.z;
    rep->e0[0] = f.e0.x; rep->e0[1] = f.e0.y; rep->e0[2] = f.e0.z;
    rep->theta = theta; rep->phi = phi;
    rep->projected_area_mm2 = proj_area(h, k);
    rep->rays = (uint64_t)m;
    kahan escaped = {0, 0}, unresolved = {0, 0};
    for (int64_t r = 0; r < m; ++r) {
        const double* q = rule + 5 * r;
        entry_s e; e.pos = mk(q[0], q[1], q[2]); e.face = (unsigned)q[3]; e.valid = 1;
        const double w = q[4];
        sink sk = {smp, cap, 0, 0};
        const orc_ray_audit a = trace_entry(&c, (uint64_t)r, e, &sk);
        if (sk.overflow || !(a.status == ST_EXHAUSTED || a.status == ST_TAIL)) { rc = -5; goto done; }
        rep->outputs += a.outputs;
        rep->max_balance_error = mx(rep->max_balance_error, fabs(a.balance));
        rep->max_interface_balance_error = mx(rep->max_interface_balance_error, a.max_interface);
        kadd(&escaped, w * a.escaped);
        kadd(&unresolved, w * a.unresolved);
        memset(col, 0, cells * sizeof(kahan));
        for (uint64_t t = 0; t < sk.n; ++t) {
            const orc_bin b = orc_phase_bin(edges, nt, np, &g, smp[t].dir);
            const double wp = w * smp[t].power;
            if (!b.valid || !(wp > 0) || !fin(wp)) { rc = -6; goto done; }
            if (!b.pole) {
                kadd(&col[b.index], wp);
            } else {
                const size_t row = b.pole > 0 ? 0 : nt - 1;
                for (uint32_t jj = 0; jj < np; ++jj) kadd(&col[row * np + jj], wp / np);
            }
        }
        for (size_t q2 = 0; q2 < cells; ++q2)
            if (col[q2].value > 0) kadd(&acc[q2], col[q2].value);
    }
    for (size_t q2 = 0; q2 < cells; ++q2) mass[q2] = acc[q2].value;
    rep->escaped_mm2 = escaped.value;
    rep->unresolved_mm2 = unresolved.value;
done:
    free(rule); free(acc); free(col); free(edges); free(smp);
    return rc;
}
