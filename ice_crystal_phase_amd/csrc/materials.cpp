// Water refractive index: IAPWS R9-97 (refractive index of ordinary water,
// Eq. 1 with the Table 1 coefficients) at the liquid density of IAPWS-95
// (R6-95(2018), residual Helmholtz energy, Tables 2-3, pressure from Eq. 6 of
// Table 3). The reference evaluates the same two releases in
// src/water_refractive_index.cpp:139-187; the coefficients are the releases'
// published values and must be alike, the evaluation and the density solve are
// this file's own (term groups evaluated as delta * d(phi_r)/d(delta) directly,
// density by Illinois regula falsi on the liquid bracket).
#include "materials.hpp"

#include <cmath>
#include <stdexcept>

namespace icephase {
namespace {

struct Compensated {
    double s = 0, c = 0;
    void add(double x)
    {
        const double t = s + x;
        c += std::fabs(s) >= std::fabs(x) ? (s - t) + x : (x - t) + s;
        s = t;
    }
    double value() const { return s + c; }
};

// R6-95 Table 2, residual part, terms 1..51: n_i, d_i, t_i, c_i (c_i = 0 for the
// seven polynomial terms).
constexpr int kTerms = 51;
constexpr double kN[kTerms] = {
    0.12533547935523e-1,  0.78957634722828e1,   -0.87803203303561e1,  0.31802509345418,
    -0.26145533859358,    -0.78199751687981e-2, 0.88089493102134e-2,  -0.66856572307965,
    0.20433810950965,     -0.66212605039687e-4, -0.19232721156002,    -0.25709043003438,
    0.16074868486251,     -0.40092828925807e-1, 0.39343422603254e-6,  -0.75941377088144e-5,
    0.56250979351888e-3,  -0.15608652257135e-4, 0.11537996422951e-8,  0.36582165144204e-6,
    -0.13251180074668e-11, -0.62639586912454e-9, -0.10793600908932,   0.17611491008752e-1,
    0.22132295167546,     -0.40247669763528,    0.58083399985759,     0.49969146990806e-2,
    -0.31358700712549e-1, -0.74315929710341,    0.47807329915480,     0.20527940895948e-1,
    -0.13636435110343,    0.14180634400617e-1,  0.83326504880713e-2,  -0.29052336009585e-1,
    0.38615085574206e-1,  -0.20393486513704e-1, -0.16554050063734e-2, 0.19955571979541e-2,
    0.15870308324157e-3,  -0.16388568342530e-4, 0.43613615723811e-1,  0.34994005463765e-1,
    -0.76788197844621e-1, 0.22446277332006e-1,  -0.62689710414685e-4, -0.55711118565645e-9,
    -0.19905718354408,    0.31777497330738,     -0.11841182425981};
constexpr unsigned char kD[kTerms] = {1, 1, 1, 2, 2, 3, 4, 1, 1, 1, 2, 2, 3, 4, 4, 5, 7,
                                      9, 10, 11, 13, 15, 1, 2, 2, 2, 3, 4, 4, 4, 5, 6, 6, 7,
                                      9, 9, 9, 9, 9, 10, 10, 12, 3, 4, 4, 5, 14, 3, 6, 6, 6};
constexpr double kT[kTerms] = {-0.5, 0.875, 1, 0.5, 0.75, 0.375, 1, 4, 6, 12, 1, 5, 4, 2, 13, 9, 3,
                               4, 11, 4, 13, 1, 7, 1, 9, 10, 10, 3, 7, 10, 10, 6, 10, 10,
                               1, 2, 3, 4, 8, 6, 9, 8, 16, 22, 23, 23, 10, 50, 44, 46, 50};
constexpr unsigned char kC[kTerms] = {0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1,
                                      1, 1, 1, 1, 1, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2,
                                      2, 2, 2, 2, 2, 2, 2, 2, 3, 3, 3, 3, 4, 6, 6, 6, 6};
// Terms 52..54 (Gaussian bell): n, t, alpha, beta, gamma (d = 3, epsilon = 1).
constexpr double kGauss[3][5] = {{-0.31306260323435e2, 0, 20, 150, 1.21},
                                 {0.31546140237781e2, 1, 20, 150, 1.21},
                                 {-0.25213154341695e4, 4, 20, 250, 1.25}};
// Terms 55..56 (nonanalytic): n, b, C, D with a = 3.5, beta = 0.3, A = 0.32, B = 0.2.
constexpr double kNonAnalytic[2][4] = {{-0.14874640856724, 0.85, 28, 700},
                                       {0.31806110878444, 0.95, 32, 800}};

double ipow(double x, unsigned k)
{
    double r = 1;
    for (unsigned i = 0; i < k; ++i) r *= x;
    return r;
}

// delta * d(phi_r)/d(delta) of IAPWS-95.
double delta_phi_r_delta(double delta, double tau)
{
    Compensated sum;
    for (int i = 0; i < kTerms; ++i) {
        const double base = kN[i] * ipow(delta, kD[i]) * std::pow(tau, kT[i]);
        if (kC[i] == 0) {
            sum.add(base * kD[i]);
        } else {
            const double dc = ipow(delta, kC[i]);
            sum.add(base * std::exp(-dc) * (double(kD[i]) - double(kC[i]) * dc));
        }
    }
    const double x = delta - 1.0;
    for (const auto& g : kGauss) {
        const double y = tau - g[4];
        const double e = std::exp(-g[2] * x * x - g[3] * y * y);
        sum.add(g[0] * ipow(delta, 3) * std::pow(tau, g[1]) * e * (3.0 - 2.0 * g[2] * delta * x));
    }
    // Nonanalytic terms: Delta^b delta psi, with the delta-derivative taken in
    // closed form. The liquid bracket keeps |delta - 1| >= 2, far from the
    // critical point, so no limit forms are needed.
    const double A = 0.32, B = 0.2, a = 3.5, beta = 0.3;
    const double x2 = x * x;
    const double theta = (1.0 - tau) + A * std::pow(x2, 1.0 / (2.0 * beta));
    const double dtheta = (A / beta) * x * std::pow(x2, 1.0 / (2.0 * beta) - 1.0);
    const double Delta = theta * theta + B * std::pow(x2, a);
    const double dDelta = 2.0 * theta * dtheta + 2.0 * B * a * x * std::pow(x2, a - 1.0);
    for (const auto& v : kNonAnalytic) {
        const double psi = std::exp(-v[2] * x2 - v[3] * (tau - 1.0) * (tau - 1.0));
        const double Db = std::pow(Delta, v[1]);
        // delta * d/d(delta) [Delta^b delta psi] = Delta^b delta psi (1 + delta (b Delta'/Delta - 2 C x))
        sum.add(v[0] * Db * delta * psi * (1.0 + delta * (v[1] * dDelta / Delta - 2.0 * v[2] * x)));
    }
    return sum.value();
}

double iapws95_pressure(double T, double rho)
{
    const double R = 461.51805, Tc = 647.096, rhoc = 322.0;  // the release's own constants
    return rho * R * T * (1.0 + delta_phi_r_delta(rho / rhoc, Tc / T));
}

}  // namespace

bool water_index_r9_97(double lambda_nm, double T, double rho, double& n, std::string& err)
{
    if (!(std::isfinite(lambda_nm) && lambda_nm >= 200.0 && lambda_nm <= 1100.0 && std::isfinite(T) &&
          T >= 261.15 && T <= 773.15 && std::isfinite(rho) && rho >= 0.0 && rho <= 1060.0)) {
        err = "IAPWS R9-97: wavelength/state outside supported range.";
        return false;
    }
    // Reduced variables (reference values 1000 kg/m^3, 273.15 K, 0.589 um).
    const double rr = rho / 1000.0, tr = T / 273.15, lr = lambda_nm / 589.0, l2 = lr * lr;
    const double a0 = 0.244257733, a1 = 0.974634476e-2, a2 = -0.373234996e-2, a3 = 0.268678472e-3,
                 a4 = 0.158920570e-2, a5 = 0.245934259e-2, a6 = 0.900704920, a7 = -0.166626219e-1;
    const double luv = 0.2292020, lir = 5.432937;
    const double rhs = a0 + a1 * rr + a2 * tr + a3 * l2 * tr + a4 / l2 + a5 / (l2 - luv * luv) +
                       a6 / (l2 - lir * lir) + a7 * rr * rr;
    // (n^2 - 1)/(n^2 + 2) = rr * rhs  =>  n^2 = (1 + 2 A)/(1 - A), A = rr * rhs
    const double A = rr * rhs;
    const double n2 = (1.0 + 2.0 * A) / (1.0 - A);
    if (!(n2 > 0.0 && std::isfinite(n2))) {
        err = "IAPWS R9-97 returned invalid index.";
        return false;
    }
    n = std::sqrt(n2);
    return true;
}

bool water_optics(double lambda_nm, double celsius, double pascal, WaterState& out, std::string& err)
{
    if (!(std::isfinite(celsius) && celsius >= 0.0 && celsius <= 60.0 && std::isfinite(pascal) &&
          pascal >= 50000.0 && pascal <= 100000000.0)) {
        err = "Water material supports liquid branch at 0..60 C and 50 kPa..100 MPa.";
        return false;
    }
    const double T = celsius + 273.15;
    double n = 0;
    if (!water_index_r9_97(lambda_nm, T, 1000.0, n, err)) return false;
    // Liquid density: root of p(T, rho) - pascal on [970, 1060] kg/m^3 (the liquid
    // branch for 0..60 C, 50 kPa..100 MPa), Illinois regula falsi.
    double lo = 970.0, hi = 1060.0;
    double flo = iapws95_pressure(T, lo) - pascal, fhi = iapws95_pressure(T, hi) - pascal;
    if (!(std::isfinite(flo) && std::isfinite(fhi) && flo < 0.0 && fhi > 0.0)) {
        err = "IAPWS-95: liquid density bracket unavailable.";
        return false;
    }
    int side = 0;
    for (int it = 0; it < 200; ++it) {
        double x = hi - fhi * (hi - lo) / (fhi - flo);
        if (!(x > lo && x < hi)) x = 0.5 * (lo + hi);
        const double fx = iapws95_pressure(T, x) - pascal;
        if (!std::isfinite(fx)) {
            err = "IAPWS-95: nonfinite pressure.";
            return false;
        }
        if (fx < 0.0) {
            lo = x;
            flo = fx;
            if (side == -1) fhi *= 0.5;
            side = -1;
        } else {
            hi = x;
            fhi = fx;
            if (side == 1) flo *= 0.5;
            side = 1;
        }
        if (fx == 0.0 || hi - lo <= 2e-10) break;
    }
    const double rl = iapws95_pressure(T, lo) - pascal, rh = iapws95_pressure(T, hi) - pascal;
    const bool use_lo = std::fabs(rl) < std::fabs(rh);
    out.temperature_kelvin = T;
    out.pressure_pascal = pascal;
    out.density_kg_m3 = use_lo ? lo : hi;
    out.density_pressure_residual_pascal = use_lo ? rl : rh;
    out.density_bracket_width_kg_m3 = hi - lo;
    if (std::fabs(out.density_pressure_residual_pascal) > 0.02) {
        err = "IAPWS-95: density pressure residual is too large.";
        return false;
    }
    return water_index_r9_97(lambda_nm, T, out.density_kg_m3, out.refractive_index, err);
}

}  // namespace icephase
