// Host-side material and particle geometry services (no device code):
// water index, drop shape table, drop convexity certificate and projected area.
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

namespace icephase {

struct WaterState {
    double temperature_kelvin = 0, pressure_pascal = 0, density_kg_m3 = 0;
    double density_pressure_residual_pascal = 0, density_bracket_width_kg_m3 = 0;
    double refractive_index = 0;
};

// IAPWS R9-97 real index at (vacuum wavelength nm, K, kg/m^3)
// (reference: src/water_refractive_index.cpp:139-155).
bool water_index_r9_97(double lambda_nm, double kelvin, double rho, double& n, std::string& err);
// Liquid density from IAPWS-95 at (C, Pa), then R9-97
// (reference: src/water_refractive_index.cpp:157-186).
bool water_optics(double lambda_nm, double celsius, double pascal, WaterState& out, std::string& err);

// Sadeghi et al. 2012 Table I drop shape by equal-volume radius (0, 3] mm,
// linear in radius between the table rows, sphere row for radius <= 0.4 mm
// (reference: include/rainbow/raindrop_shape.hpp:22-69). Single precision, as
// the tracer uses the coefficients.
bool drop_shape_table(float radius_mm, bool sphere, float coefficients[8], std::string& err);

// Strict convexity of r(mu) = 1 + c0 + c1 mu + sum_{n>=2} c_n T_n(mu) on [-1, 1]:
// r > 0, G = r + mu r' > 0 and Q = r^2 + 2(1-mu^2) r'^2 - r((1-mu^2) r'' - mu r') > 0,
// certified by a Lipschitz branch and bound (Markov's inequality bounds each
// polynomial's slope by its degree squared times a coefficient bound).
// `cells` returns the number of intervals examined.
bool drop_convexity(const double c[8], uint32_t& cells, std::string& err);

struct AreaResult {
    double area_mm2 = 0, abs_error_mm2 = 0, rel_error = 0;
    uint32_t evaluations = 0, convexity_cells = 0;
    bool analytic_sphere = false;
};
// Projected area of the convex drop r(chi) a along the direction k (particle
// frame, polar axis -y): half the surface integral of |n.k|, the azimuth
// integral in closed form, polar angle by adaptive Gauss-Kronrod (7,15) seeded
// with the polar angles where the azimuthal integrand changes form
// (reference: src/projected_area.cpp:163-228).
bool drop_projected_area(const double c[8], double radius_mm, const double k[3], AreaResult& out, std::string& err);

// Warren & Brandt (2008) ice index at a vacuum wavelength in [200, 1100] nm:
// n (used by the tracer) and kappa (reported, not applied), with the bracket.
struct IceState {
    double wavelength_nm, n, kappa, lower_nm, upper_nm, log_wavelength_fraction;
};
bool ice_index_wb2008(double wavelength_nm, IceState& out, std::string& err);

}  // namespace icephase
