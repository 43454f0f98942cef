/*
 * Oracle (test infrastructure only; see oracle.h): drop tracer mirror and the
 * closed-form sphere. The FP32 instantiation mirrors the HIP kernel's arithmetic
 * operation for operation; the FP64 instantiation runs the same algorithm in
 * double for statistical comparison.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define PI_D 3.141592653589793238462643383279502884
#define ST_NUMERIC_D 5

/* ---- FP32 instantiation ---- */
#define REAL float
#define SFX f
#define SQRT sqrtf
#define FABS fabsf
#define FMA fmaf
#define TOL 0x1p-21f
#define TINY_S2 0x1p-40f
#define SELF_OFFSET 0x1p-16f
#define MIN_STEP 0x1p-20f
#include "oracle_drop_impl.h"
#undef REAL
#undef SFX
#undef SQRT
#undef FABS
#undef FMA
#undef TOL
#undef TINY_S2
#undef SELF_OFFSET
#undef MIN_STEP

/* ---- FP64 instantiation ---- */
#define REAL double
#define SFX d
#define SQRT sqrt
#define FABS fabs
#define FMA fma
#define TOL 0x1p-44
#define TINY_S2 0x1p-90
#define SELF_OFFSET 0x1p-30
#define MIN_STEP 0x1p-40
#include "oracle_drop_impl.h"
#undef REAL
#undef SFX
#undef SQRT
#undef FABS
#undef FMA
#undef TOL
#undef TINY_S2
#undef SELF_OFFSET
#undef MIN_STEP

int orc_normalize3(const double in[3], double out[3]);

int orc_drop_make(const double coeffs[8], int sphere, double alpha_deg, double n_int,
                  uint32_t grid_w, uint32_t grid_h, uint32_t max_hits, double tail,
                  orc_drop_record* r)
{
    memset(r, 0, sizeof *r);
    if (!(grid_w > 0 && grid_h > 0 && n_int > 0)) return -1;
    double s = 0;
    for (int n = 0; n < 8; ++n) r->coeffs[n] = sphere ? 0.0f : (float)coeffs[n];
    for (int n = 1; n < 8; ++n) s += fabs((double)r->coeffs[n]);
    const double center = 1.0 + (double)r->coeffs[0];
    r->inner = nextafterf((float)(center - s - 0x1p-20), -INFINITY);
    r->outer = nextafterf((float)(center + s + 0x1p-20), INFINITY);
    if (sphere) { r->inner = 1.0f; r->outer = 1.0f; }
    /* bounding radius: max over chi of r(chi) = 1 + sum c_n cos(n chi) on a dense
     * grid, +1e-6, rounded up to a multiple of 2^-20 (exact in float) */
    {
        double rm = 0;
        for (int k = 0; k <= 16384; ++k) {
            const double chi = PI_D * k / 16384;
            double rr = 1 + (double)r->coeffs[0];
            for (int m = 1; m < 8; ++m) rr += (double)r->coeffs[m] * cos(m * chi);
            if (rr > rm) rm = rr;
        }
        r->rmax = sphere ? 1.0f : (float)(ceil((rm + 1e-6) * 1048576.0) / 1048576.0);
    }
    const double a = alpha_deg * PI_D / 180.0;
    const double k[3] = {cos(a), -sin(a), 0.0};
    orc_frame f;
    if (orc_drop_frame(k, &f)) return -1;
    for (int i = 0; i < 3; ++i) {
        r->ki[i] = (float)f.ki[i];
        r->e0[i] = (float)f.e0[i];
        r->e1[i] = (float)f.e1[i];
    }
    r->n_int = (float)n_int;
    r->n_ext = 1.0f;
    r->half_extent = (float)((double)r->outer * 1.01);
    r->grid_w = grid_w; r->grid_h = grid_h;
    r->max_hits = max_hits;
    r->sphere = sphere ? 1u : 0u;
    r->tail = (float)tail;
    return 0;
}

/* 1 + the index of the last nonzero shape coefficient, at least 2 */
static uint32_t orc_drop_terms(const orc_drop_record* r)
{
    uint32_t t = 2;
    for (uint32_t m = 2; m < 8; ++m)
        if (r->coeffs[m] != 0.0f) t = m + 1;
    return t;
}

static void load_f(const orc_drop_record* r, rec_f* R)
{
    for (int n = 0; n < 8; ++n) R->c[n] = r->coeffs[n];
    R->inner = r->inner; R->outer = r->outer; R->rmax = r->rmax;
    R->ki = mk_f(r->ki[0], r->ki[1], r->ki[2]);
    R->e0 = mk_f(r->e0[0], r->e0[1], r->e0[2]);
    R->e1 = mk_f(r->e1[0], r->e1[1], r->e1[2]);
    R->n_int = r->n_int; R->n_ext = r->n_ext; R->h = r->half_extent; R->tail = r->tail;
    R->w = r->grid_w; R->hgt = r->grid_h; R->max_hits = r->max_hits; R->sphere = r->sphere;
    R->terms = orc_drop_terms(r);
}
static void load_d(const orc_drop_record* r, rec_d* R)
{
    for (int n = 0; n < 8; ++n) R->c[n] = r->coeffs[n];
    R->inner = r->inner; R->outer = r->outer; R->rmax = r->rmax;
    /* the FP64 twin uses the same rounded inputs, promoted and renormalized */
    double k[3] = {r->ki[0], r->ki[1], r->ki[2]}, e0[3] = {r->e0[0], r->e0[1], r->e0[2]};
    double nk[3], ne0[3];
    orc_normalize3(k, nk);
    const double kd = nk[0] * e0[0] + nk[1] * e0[1] + nk[2] * e0[2];
    double t[3] = {e0[0] - nk[0] * kd, e0[1] - nk[1] * kd, e0[2] - nk[2] * kd};
    orc_normalize3(t, ne0);
    R->ki = mk_d(nk[0], nk[1], nk[2]);
    R->e0 = mk_d(ne0[0], ne0[1], ne0[2]);
    R->e1 = unit_d(crs_d(R->ki, R->e0));
    R->n_int = r->n_int; R->n_ext = r->n_ext; R->h = r->half_extent; R->tail = r->tail;
    R->w = r->grid_w; R->hgt = r->grid_h; R->max_hits = r->max_hits; R->sphere = r->sphere;
    R->terms = orc_drop_terms(r);
}

int64_t orc_drop_trace(const orc_drop_record* r, int precision, uint64_t first, uint64_t count,
                       orc_sample* samples, uint64_t cap, orc_ray_audit* audits)
{
    uint64_t n = 0;
    int overflow = 0;
    if (precision == 32) {
        rec_f R; load_f(r, &R);
        for (uint64_t i = 0; i < count; ++i) {
            orc_ray_audit a = trace_ray_f(&R, first + i, samples, cap, &n, &overflow);
            if (audits) audits[i] = a;
            if (overflow) return -1;
        }
    } else {
        rec_d R; load_d(r, &R);
        for (uint64_t i = 0; i < count; ++i) {
            orc_ray_audit a = trace_ray_d(&R, first + i, samples, cap, &n, &overflow);
            if (audits) audits[i] = a;
            if (overflow) return -1;
        }
    }
    return (int64_t)n;
}

int orc_drop_surface(const orc_drop_record* r, int precision, uint64_t n, const double* pts, double* f,
                     double* grad)
{
    if (precision == 32) {
        rec_f R; load_f(r, &R);
        for (uint64_t i = 0; i < n; ++i) {
            const v3f p = mk_f((float)pts[3 * i], (float)pts[3 * i + 1], (float)pts[3 * i + 2]);
            f[i] = (double)surf_f(&R, p);
            const v3f g = grad_f(&R, p);
            grad[3 * i] = g.x; grad[3 * i + 1] = g.y; grad[3 * i + 2] = g.z;
        }
    } else {
        rec_d R; load_d(r, &R);
        for (uint64_t i = 0; i < n; ++i) {
            const v3d p = mk_d(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
            f[i] = surf_d(&R, p);
            const v3d g = grad_d(&R, p);
            grad[3 * i] = g.x; grad[3 * i + 1] = g.y; grad[3 * i + 2] = g.z;
        }
    }
    return 0;
}

int orc_drop_entry_point(const orc_drop_record* r, int precision, double u, double v, double p[3], double nrm[3])
{
    if (precision == 32) {
        rec_f R; load_f(r, &R);
        const float uf = (float)u, vf = (float)v;
        const v3f o = axpy_f(R.e1, vf, scl_f(R.e0, uf));
        float t;
        if (!entry_f(&R, o, uf, vf, &t)) return 0;
        const v3f q = axpy_f(R.ki, t, o);
        const v3f n = outward_normal_f(&R, q);
        p[0] = q.x; p[1] = q.y; p[2] = q.z;
        nrm[0] = n.x; nrm[1] = n.y; nrm[2] = n.z;
    } else {
        rec_d R; load_d(r, &R);
        const v3d o = axpy_d(R.e1, v, scl_d(R.e0, u));
        double t;
        if (!entry_d(&R, o, u, v, &t)) return 0;
        const v3d q = axpy_d(R.ki, t, o);
        const v3d n = outward_normal_d(&R, q);
        p[0] = q.x; p[1] = q.y; p[2] = q.z;
        nrm[0] = n.x; nrm[1] = n.y; nrm[2] = n.z;
    }
    return 1;
}

int orc_drop_exit_point(const orc_drop_record* r, int precision, const double p[3], const double d[3], double q[3],
                        double nrm[3])
{
    if (precision == 32) {
        rec_f R; load_f(r, &R);
        const v3f pf = mk_f((float)p[0], (float)p[1], (float)p[2]);
        const v3f df = mk_f((float)d[0], (float)d[1], (float)d[2]);
        float t;
        if (!exit_f(&R, pf, df, &t)) return 0;
        const v3f x = axpy_f(df, t, pf);
        const v3f n = outward_normal_f(&R, x);
        q[0] = x.x; q[1] = x.y; q[2] = x.z;
        nrm[0] = n.x; nrm[1] = n.y; nrm[2] = n.z;
    } else {
        rec_d R; load_d(r, &R);
        const v3d pd = mk_d(p[0], p[1], p[2]), dd = mk_d(d[0], d[1], d[2]);
        double t;
        if (!exit_d(&R, pd, dd, &t)) return 0;
        const v3d x = axpy_d(dd, t, pd);
        const v3d n = outward_normal_d(&R, x);
        q[0] = x.x; q[1] = x.y; q[2] = x.z;
        nrm[0] = n.x; nrm[1] = n.y; nrm[2] = n.z;
    }
    return 1;
}

int64_t orc_drop_trace_origin(const orc_drop_record* r, int precision, double u, double v, orc_sample* samples,
                              uint64_t cap, orc_ray_audit* audit)
{
    uint64_t n = 0;
    int overflow = 0;
    orc_ray_audit a;
    if (precision == 32) {
        rec_f R; load_f(r, &R);
        a = trace_origin_f(&R, 0, (float)u, (float)v, samples, cap, &n, &overflow);
    } else {
        rec_d R; load_d(r, &R);
        a = trace_origin_d(&R, 0, u, v, samples, cap, &n, &overflow);
    }
    if (audit) *audit = a;
    return overflow ? -1 : (int64_t)n;
}

int orc_drop_hist(const orc_drop_record* r, int precision, uint32_t nt, uint32_t np,
                  uint32_t qbits, uint64_t first, uint64_t count, uint64_t* bins, orc_hist_stats* st)
{
    double* ue = (double*)malloc(sizeof(double) * (nt + 1));
    if (orc_u_edges(nt, ue)) { free(ue); return -3; }
    const double k[3] = {r->ki[0], r->ki[1], r->ki[2]};
    orc_frame g;
    /* binning frame: the drop contract frame of the traced (FP32-rounded) direction */
    if (orc_drop_frame(k, &g)) { free(ue); return -2; }
    memset(bins, 0, sizeof(uint64_t) * (size_t)nt * np);
    memset(st, 0, sizeof *st);
    rec_f Rf; rec_d Rd;
    load_f(r, &Rf); load_d(r, &Rd);
    orc_sample buf[64];
    for (uint64_t i = 0; i < count; ++i) {
        uint64_t n = 0;
        int overflow = 0;
        orc_ray_audit a = precision == 32 ? trace_ray_f(&Rf, first + i, buf, 64, &n, &overflow)
                                          : trace_ray_d(&Rd, first + i, buf, 64, &n, &overflow);
        if (overflow) { free(ue); return -1; }
        if (a.status != 8) st->rays_hit++;
        st->status[a.status & 7] += a.status == 8 ? 0 : 1;
        st->internal_hits += a.internal_hits;
        uint64_t qe = 0;
        for (uint64_t s = 0; s < n; ++s) {
            const uint64_t q = (uint64_t)llrint(ldexp(buf[s].power, (int)qbits));
            qe += q;
            st->samples++;
            const orc_bin b = orc_phase_bin_raw(ue, nt, np, &g, buf[s].dir);
            if (!b.valid) { st->invalid++; continue; }
            st->snapped += (uint64_t)b.snapped;
            if (b.pole > 0) st->q_forward += q;
            else if (b.pole < 0) st->q_backward += q;
            else bins[b.index] += q;
        }
        st->q_escaped += qe;
        st->q_unresolved += (uint64_t)llrint(ldexp(a.unresolved, (int)qbits));
    }
    free(ue);
    return 0;
}

void orc_sphere_theta_hist(uint32_t grid, float half_extent, double n, uint32_t max_hits, uint32_t nt,
                           double* hist)
{
    double* ue = (double*)malloc(sizeof(double) * (nt + 1));
    orc_u_edges(nt, ue);
    memset(hist, 0, sizeof(double) * nt);
    for (uint32_t iy = 0; iy < grid; ++iy)
        for (uint32_t ix = 0; ix < grid; ++ix) {
            const double u = ((2.0 * ix + 1) / grid - 1) * (double)half_extent;
            const double v = ((2.0 * iy + 1) / grid - 1) * (double)half_extent;
            const double b = sqrt(u * u + v * v);
            if (!(b < 1)) continue;
            for (uint32_t p = 0; p <= max_hits; ++p) {
                double psi, ps, pp;
                orc_sphere_order(b, n, p, &psi, &ps, &pp);
                const double uu = 0.5 * (1 - cos(psi));
                uint32_t lo = 0, hi = nt;
                while (lo + 1 < hi) {
                    const uint32_t mid = lo + (hi - lo) / 2;
                    if (uu >= ue[mid]) lo = mid; else hi = mid;
                }
                hist[lo] += 0.5 * (ps + pp);
            }
        }
    free(ue);
}

void orc_sphere_order(double b, double n, uint32_t p, double* theta, double* ps, double* pp)
{
    const double si = b, ci = sqrt((1 - b) * (1 + b));
    const double sr = b / n, cr = sqrt((1 - sr) * (1 + sr));
    const double i = asin(b), r = asin(sr);
    const double rs = (ci - n * cr) / (ci + n * cr);
    const double rp = (n * ci - cr) / (n * ci + cr);
    const double Rs = rs * rs, Rp = rp * rp;
    (void)si;
    double psi;
    if (p == 0) {
        psi = PI_D - 2 * i;
        *ps = Rs; *pp = Rp;
    } else {
        psi = -((double)(p - 1) * PI_D + 2 * i - 2 * (double)p * r);
        *ps = (1 - Rs) * (1 - Rs) * pow(Rs, (double)(p - 1));
        *pp = (1 - Rp) * (1 - Rp) * pow(Rp, (double)(p - 1));
    }
    *theta = psi; /* signed in-plane rotation from ki toward the impact vector */
}
