// Host-side particle geometry: Table I drop shapes, the drop convexity
// certificate and the drop projected area (see materials.hpp).
#include "materials.hpp"

#include <algorithm>
#include <cfenv>
#include <cmath>
#include <cstring>
#include <vector>

namespace icephase {
namespace {
constexpr double kPiD = 3.141592653589793238462643383279502884;
}

// ---------------------------------------------------------------------------
// Table I (Sadeghi et al. 2012, p. 5): Chebyshev coefficients c0..c7 of the
// equal-volume-normalized meridian by radius. The table's values are data of
// the published paper and must be alike; rows at a = 0.4, 1.0, ..., 3.0 mm.
bool drop_shape_table(float radius_mm, bool sphere, float c[8], std::string& err)
{
    static const float kRadius[6] = {0.4f, 1.0f, 1.5f, 2.0f, 2.5f, 3.0f};
    static const float kRow[6][8] = {
        {0, 0, 0, 0, 0, 0, 0, 0},
        {-0.0131f, -0.0120f, -0.0376f, -0.0096f, -0.0004f, 0.0015f, 0.0005f, 0},
        {-0.0282f, -0.0230f, -0.0779f, -0.0175f, 0.0021f, 0.0046f, 0.0011f, -0.0006f},
        {-0.0458f, -0.0335f, -0.1211f, -0.0227f, 0.0083f, 0.0089f, 0.0012f, -0.0021f},
        {-0.0644f, -0.0416f, -0.1629f, -0.0246f, 0.0176f, 0.0131f, 0.0002f, -0.0044f},
        {-0.0840f, -0.0480f, -0.2034f, -0.0237f, 0.0297f, 0.0166f, -0.0021f, -0.0072f}};
    if (!(radius_mm > 0.0f && radius_mm <= 3.0f)) {
        err = "Drop radius must lie in (0, 3] mm (Table I range; no extrapolation).";
        return false;
    }
    for (int k = 0; k < 8; ++k) c[k] = 0.0f;
    if (sphere || radius_mm <= kRadius[0]) return true;
    int hi = 1;
    while (hi < 5 && radius_mm > kRadius[hi]) ++hi;
    // single-precision interpolation with one fused rounding per coefficient
    const float t = (radius_mm - kRadius[hi - 1]) / (kRadius[hi] - kRadius[hi - 1]);
    for (int k = 0; k < 8; ++k) c[k] = std::fmaf(t, kRow[hi][k] - kRow[hi - 1][k], kRow[hi - 1][k]);
    return true;
}

// ---------------------------------------------------------------------------
// Convexity certificate.
namespace {

struct Meridian {
    double c[8];
    // r, r', r'' in mu by the Chebyshev recurrences
    void eval(double mu, double& r, double& r1, double& r2) const
    {
        double t0 = 1, t1 = mu, p0 = 0, p1 = 1, s0 = 0, s1 = 0;
        r = 1 + c[0] + c[1] * mu;
        r1 = c[1];
        r2 = 0;
        for (int n = 2; n < 8; ++n) {
            const double t = 2 * mu * t1 - t0;
            const double p = 2 * t1 + 2 * mu * p1 - p0;
            const double s = 4 * p1 + 2 * mu * s1 - s0;
            r += c[n] * t;
            r1 += c[n] * p;
            r2 += c[n] * s;
            t0 = t1, t1 = t, p0 = p1, p1 = p, s0 = s1, s1 = s;
        }
    }
    // the three certified quantities at mu
    void conditions(double mu, double v[3]) const
    {
        double r, r1, r2;
        eval(mu, r, r1, r2);
        const double s2 = 1 - mu * mu;
        v[0] = r;
        v[1] = r + mu * r1;
        v[2] = r * r + 2 * s2 * r1 * r1 - r * (s2 * r2 - mu * r1);
    }
};

}  // namespace

bool drop_convexity(const double c[8], uint32_t& cells, std::string& err)
{
    Meridian m;
    for (int k = 0; k < 8; ++k) {
        if (!std::isfinite(c[k])) {
            err = "Nonfinite shape coefficient.";
            return false;
        }
        m.c[k] = c[k];
    }
    // Bounds on [-1, 1]: |T_n| <= 1, |T_n'| <= n^2, |T_n''| <= n^2 (n^2 - 1)/3.
    double B0 = std::fabs(1 + c[0]) + std::fabs(c[1]), B1 = std::fabs(c[1]), B2 = 0;
    for (int n = 2; n < 8; ++n) {
        B0 += std::fabs(c[n]);
        B1 += n * n * std::fabs(c[n]);
        B2 += n * n * (n * n - 1) / 3.0 * std::fabs(c[n]);
    }
    // sup norms of r (degree 7), G (degree 7), Q (degree <= 14) and, by Markov's
    // inequality, Lipschitz constants degree^2 * sup norm.
    const double sup[3] = {B0, B0 + B1, B0 * B0 + 2 * B1 * B1 + B0 * (B2 + B1)};
    const double lip[3] = {49 * sup[0], 49 * sup[1], 196 * sup[2]};
    // evaluation error allowance (a few hundred ulps of the sup norm)
    const double slack[3] = {1e-13 * sup[0], 1e-13 * sup[1], 1e-13 * sup[2]};
    struct Cell {
        double lo, hi;
        int depth;
    };
    std::vector<Cell> stack;
    const int seeds = 64;
    for (int k = seeds - 1; k >= 0; --k) stack.push_back({-1.0 + 2.0 * k / seeds, -1.0 + 2.0 * (k + 1) / seeds, 0});
    cells = 0;
    while (!stack.empty()) {
        const Cell cl = stack.back();
        stack.pop_back();
        if (++cells > (1u << 22)) {
            err = "Drop convexity could not be certified within the work limit.";
            return false;
        }
        const double mid = 0.5 * (cl.lo + cl.hi), half = 0.5 * (cl.hi - cl.lo);
        double v[3];
        m.conditions(mid, v);
        bool certified = true;
        for (int q = 0; q < 3; ++q) {
            if (!std::isfinite(v[q])) {
                err = "Nonfinite drop curvature.";
                return false;
            }
            if (v[q] <= 0) {
                err = q == 0 ? "Drop radius r(mu) must stay positive."
                             : "Drop shape is not convex (a principal curvature changes sign); the analytic convex "
                               "intersection requires a strictly convex particle.";
                return false;
            }
            if (v[q] - half * lip[q] - slack[q] <= 0) certified = false;
        }
        if (certified) continue;
        if (cl.depth >= 40) {
            err = "Drop convexity could not be certified (curvature too close to zero).";
            return false;
        }
        stack.push_back({mid, cl.hi, cl.depth + 1});
        stack.push_back({cl.lo, mid, cl.depth + 1});
    }
    return true;
}

// ---------------------------------------------------------------------------
// Projected area.
namespace {

// Half the azimuth integral of |n.k| dA over the parallel at polar angle chi,
// divided by d(chi): with k split into a transverse part kt and an axial part ka
// (relative to the symmetry axis) and n dA = r sin(chi) [h e_rho + v e_axis],
// h = sin(chi) (r + mu r_mu), v = mu r - sin^2(chi) r_mu, the azimuth integral
// int_0^{2pi} |A cos(phi) + B| dphi, A = kt h, B = ka v, is 2 pi |B| when
// |A| <= |B| and 4 (sqrt(A^2 - B^2) + |B| asin(|B|/|A|)) otherwise.
struct Integrand {
    Meridian m;
    double kt, ka;
    void parts(double chi, double& r, double& A, double& B) const
    {
        const double mu = std::cos(chi), sn = std::sin(chi);
        double r1, r2;
        m.eval(mu, r, r1, r2);
        A = std::fabs(kt * sn * (r + mu * r1));
        B = std::fabs(ka * (mu * r - sn * sn * r1));
    }
    double operator()(double chi) const
    {
        double r, A, B;
        parts(chi, r, A, B);
        const double sn = std::sin(chi);
        double j;
        if (A <= B)
            j = 2 * kPiD * B;
        else
            j = 4 * (std::sqrt((A - B) * (A + B)) + B * std::asin(B / A));
        return 0.5 * r * sn * j;
    }
    // A - B: its sign changes mark where the azimuth integral changes form
    double kink(double chi) const
    {
        double r, A, B;
        parts(chi, r, A, B);
        return A - B;
    }
};

// Gauss-Kronrod (7, 15) nodes and weights on [-1, 1] (the published rule's
// constants): Kronrod nodes x[0..7] (x[7] = 0), Kronrod weights, Gauss weights
// of the 7-point rule on the odd Kronrod nodes.
constexpr double kX[8] = {0.991455371120812639206854697526329, 0.949107912342758524526189684047851,
                          0.864864423359769072789712788640926, 0.741531185599394439863864773280788,
                          0.586087235467691130294144838258730, 0.405845151377397166906606412076961,
                          0.207784955007898467600689403773245, 0.0};
constexpr double kWK[8] = {0.022935322010529224963732008058970, 0.063092092629978553290700663189204,
                           0.104790010322250183839876322541518, 0.140653259715525918745189590510238,
                           0.169004726639267902826583426598550, 0.190350578064785409913256402421014,
                           0.204432940075298892414161999234649, 0.209482141084727828012999174891714};
constexpr double kWG[4] = {0.129484966168869693270611432679082, 0.279705391489276667901467771423780,
                           0.381830050505118944950369775488975, 0.417959183673469387755102040816327};

struct Estimate {
    double value, error;
};

Estimate gk15(const Integrand& f, double a, double b)
{
    const double c = 0.5 * (a + b), h = 0.5 * (b - a);
    const double fc = f(c);
    double k = kWK[7] * fc, g = kWG[3] * fc;
    for (int i = 0; i < 7; ++i) {
        const double s = f(c - h * kX[i]) + f(c + h * kX[i]);
        k += kWK[i] * s;
        if (i & 1) g += kWG[i / 2] * s;
    }
    return {k * h, std::fabs((k - g) * h)};
}

// depth-first adaptive refinement with the tolerance split in proportion to
// the panel width
void adapt(const Integrand& f, double a, double b, double tol, int depth, double& sum, double& comp, double& err,
           uint32_t& evals, bool& ok)
{
    const Estimate e = gk15(f, a, b);
    evals += 15;
    const double m = 0.5 * (a + b);
    if (e.error <= tol || depth >= 48 || m <= a || m >= b) {
        if (e.error > tol) ok = false;
        const double y = e.value - comp, t = sum + y;
        comp = (t - sum) - y;
        sum = t;
        err += e.error;
        return;
    }
    adapt(f, a, m, 0.5 * tol, depth + 1, sum, comp, err, evals, ok);
    adapt(f, m, b, 0.5 * tol, depth + 1, sum, comp, err, evals, ok);
}

}  // namespace

bool drop_projected_area(const double c[8], double radius_mm, const double k[3], AreaResult& out, std::string& err)
{
    out = AreaResult{};
    if (std::fegetround() != FE_TONEAREST) {
        err = "Projected area requires round-to-nearest arithmetic.";
        return false;
    }
    if (!(std::isfinite(radius_mm) && radius_mm > 0)) {
        err = "Invalid projected-area radius.";
        return false;
    }
    for (int i = 0; i < 3; ++i)
        if (!std::isfinite(k[i])) {
            err = "Nonfinite incident direction.";
            return false;
        }
    const double norm = std::hypot(std::hypot(k[0], k[1]), k[2]);
    if (!(norm > 0 && std::isfinite(norm))) {
        err = "Invalid incident direction.";
        return false;
    }
    bool sphere = true;
    for (int n = 1; n < 8; ++n) sphere = sphere && c[n] == 0;
    const double r2 = radius_mm * radius_mm;
    if (sphere) {
        const double r = 1 + c[0];
        if (!(r > 0 && std::isfinite(r))) {
            err = "Invalid sphere radius.";
            return false;
        }
        out.area_mm2 = kPiD * r * r * r2;
        out.rel_error = 8 * 2.220446049250313e-16;
        out.abs_error_mm2 = out.area_mm2 * out.rel_error;
        out.analytic_sphere = true;
        return true;
    }
    if (!drop_convexity(c, out.convexity_cells, err)) return false;
    Integrand f;
    for (int n = 0; n < 8; ++n) f.m.c[n] = c[n];
    // symmetry axis is -y: axial part |k.y|, transverse part |(k.x, k.z)|
    f.kt = std::hypot(k[0], k[2]) / norm;
    f.ka = std::fabs(k[1]) / norm;
    // break points: sign changes of A - B on a fine polar grid, refined by bisection
    std::vector<double> cuts{0.0};
    const int grid = 512;
    double prev = f.kink(0.0);
    for (int s = 1; s <= grid; ++s) {
        const double x = kPiD * s / grid;
        const double cur = f.kink(x);
        if ((prev > 0) != (cur > 0)) {
            double lo = kPiD * (s - 1) / grid, hi = x;
            const bool lo_pos = prev > 0;
            for (int it = 0; it < 80 && hi - lo > 1e-15; ++it) {
                const double m = 0.5 * (lo + hi);
                if ((f.kink(m) > 0) == lo_pos)
                    lo = m;
                else
                    hi = m;
            }
            cuts.push_back(0.5 * (lo + hi));
        }
        prev = cur;
    }
    cuts.push_back(kPiD);
    // first pass for the magnitude, then the adaptive pass at 1e-12 relative
    double rough = 0;
    for (size_t i = 0; i + 1 < cuts.size(); ++i) rough += gk15(f, cuts[i], cuts[i + 1]).value;
    if (!(rough > 0) || !std::isfinite(rough)) {
        err = "Projected-area integrand is invalid.";
        return false;
    }
    const double tol = 1e-12 * rough;
    double sum = 0, comp = 0, e = 0;
    bool ok = true;
    for (size_t i = 0; i + 1 < cuts.size(); ++i)
        adapt(f, cuts[i], cuts[i + 1], tol * (cuts[i + 1] - cuts[i]) / kPiD, 0, sum, comp, e, out.evaluations, ok);
    if (!ok || !(sum > 0) || !std::isfinite(sum)) {
        err = "Projected-area quadrature did not converge.";
        return false;
    }
    out.area_mm2 = sum * r2;
    out.abs_error_mm2 = e * r2;
    out.rel_error = e / sum;
    return true;
}

}  // namespace icephase
