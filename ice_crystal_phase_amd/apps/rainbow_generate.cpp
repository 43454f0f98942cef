// rainbow_generate: one water-drop phase-function record, drop-in for the
// reference tool of the same name (apps/generate_phase.cpp): same flags, the
// same on-disk contract (metadata.json, phi_cdf.npy [Nphi+1],
// theta_given_phi_cdf.npy [Nphi, Ntheta+1], u_edges.npy, staged in <out>.part
// and renamed), the same error convention ("rainbow_generate: <message>" on
// stderr, exit 1).
//
// --stage path|focal|diffraction (the default, as in the reference) run the
// patch-model pipeline on the device (iph_coherent_*: vertex trace, patches,
// folded branches, patch query, coherent optics, focal quarter turns, Table II
// filter, CDF), --grid counting incident vertices per axis (default 129).
// --stage incoherent is the ray histogram of the MI355X kernels, the same
// estimand as the reference's incoherent patch stage, --grid counting rays per
// axis (default 513). Flags added by this tool, for the incoherent stage only:
// --shape-coefficients c0,..,c7 (overrides Table I), --max-internal-hits K,
// --tail-tolerance E.
#include <chrono>
#include <cmath>
#include <cstring>
#include <filesystem>
#include <iomanip>
#include <iostream>
#include <limits>
#include <string>
#include <vector>

#include "cli_support.hpp"

namespace {

constexpr double kPi = 3.141592653589793238462643383279502884;

struct Command {
    std::string out;
    float radius_mm = 0.4f;  // RaindropSettings default (include/rainbow/raindrop_settings.hpp:12)
    float wavelength_nm = 0;
    double wavelength_requested = 0;
    bool wavelength_given = false, sphere = false, allow_underresolved = false;
    bool water = true, index_given = false, state_given = false;
    double ior = 0, celsius = 20.0, pascal = 101325.0;
    uint32_t grid = 0, theta = 360, phi = 720, cdf_theta = 0, cdf_phi = 0;
    double inclination = 20.0, max_tv = 1.0, max_l1 = 1e-10, max_lost = 1e-12, transition_contrast = 0.25;
    bool histogram_options = false;  // --shape-coefficients / --max-internal-hits / --tail-tolerance given
    double paper_sigma = 0;          // Table II primary sigma (degrees) for focal / diffraction
    int device = 0;
    std::string stage = "diffraction";  // the reference default (apps/generate_phase.cpp:54)
    unsigned focal[4] = {0, 0, 0, 0};
    bool coefficients_given = false;
    double coefficients[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    uint32_t max_hits = 8;
    double tail = 0.0;
    // resolved
    iph_water_state water_state{};
    float interior_index = 0;
    float shape[8] = {0, 0, 0, 0, 0, 0, 0, 0};
};

// Table II of Sadeghi et al. (2012): primary bandwidth (degrees) by drop radius
// (mm), linearly interpolated; 0 outside 0.1..1 mm (include/rainbow/rainbow_diffraction.hpp:17-29).
double table_sigma_degrees(float radius)
{
    static const float r[10] = {0.1f, 0.2f, 0.3f, 0.4f, 0.5f, 0.6f, 0.7f, 0.8f, 0.9f, 1.0f};
    static const double s[10] = {0.70, 0.45, 0.30, 0.25, 0.22, 0.20, 0.18, 0.17, 0.16, 0.15};
    if (!(radius >= r[0] && radius <= r[9])) return 0;
    for (unsigned j = 0; j < 10; ++j) {
        if (radius == r[j]) return s[j];
        if (j && radius < r[j]) return std::fma(((double)radius - r[j - 1]) / ((double)r[j] - r[j - 1]), s[j] - s[j - 1], s[j - 1]);
    }
    return 0;
}

// The solver's conservative outer radius in drop units, FP32 rounded outward
// (include/rainbow/raindrop_shape.hpp:45-64): the incident grid spans 1.01 of
// it and the reference planes lie at twice it.
float outer_radius(const float* c)
{
    float sum = 0.0f;
    for (int k = 1; k < 8; ++k) sum = std::nextafter(sum + std::fabs(c[k]), INFINITY);
    return std::nextafter(1.0f + c[0] + sum + 0x1p-20f, INFINITY);
}

void usage()
{
    std::cout << "rainbow_generate --out DIRECTORY [--modules DIRECTORY] [--sphere]\n"
                 "  --radius-mm R --wavelength-nm L (required, vacuum nm) --grid VERTICES_PER_AXIS\n"
                 "  --inclination-deg A --query-theta T --query-phi P\n"
                 "  --stage diffraction|focal|incoherent|path (default diffraction)\n"
                 "  --material water (default) --temperature-c T --pressure-pa P\n"
                 "  --material constant --ior N (explicit nondispersive reference)\n"
                 "  --cdf-max-l1 E --cdf-max-lost-mass E --device ORDINAL\n"
                 "  --cdf-theta T --cdf-phi P (integer divisors of query dimensions)\n"
                 "  --max-coarsening-tv E (quality policy, not a blur width)\n"
                 "  --focal-offsets 0,0,0,0\n"
                 "  diffraction width: Table II radius lookup, secondary = 2*primary\n"
                 "  --allow-underresolved  (explicitly accept flagged angular underresolution)\n"
                 "  --stage incoherent only (ray histogram, --grid = rays per axis, default 513):\n"
                 "  --shape-coefficients c0,c1,...  (up to 8; default Table I by radius)\n"
                 "  --max-internal-hits K (default 8) --tail-tolerance E (default 0)\n"
                 "  No additional/global storage Gaussian is applied.\n";
}

bool parse(int argc, char** argv, Command& c)
{
    for (int i = 1; i < argc; ++i) {
        const std::string option = argv[i];
        if (option == "--sphere") {
            c.sphere = true;
            continue;
        }
        if (option == "--allow-underresolved") {
            c.allow_underresolved = true;
            continue;
        }
        if (option == "--help") {
            usage();
            return false;
        }
        if (++i >= argc) throw std::invalid_argument("Missing option value.");
        const std::string value = argv[i];
        if (option == "--out")
            c.out = value;
        else if (option == "--modules")
            ;  // the kernels are in libicephase.so; accepted for command-line compatibility
        else if (option == "--radius-mm")
            c.radius_mm = (float)cli::number(value);
        else if (option == "--wavelength-nm") {
            if (c.wavelength_given) throw std::invalid_argument("Duplicate --wavelength-nm.");
            c.wavelength_requested = cli::number(value);
            c.wavelength_given = true;
            if (!(c.wavelength_requested > 0 && c.wavelength_requested <= (std::numeric_limits<float>::max)()))
                throw std::invalid_argument("Wavelength must be positive and representable by the solver.");
            c.wavelength_nm = (float)c.wavelength_requested;
            if (!(c.wavelength_nm > 0)) throw std::invalid_argument("Wavelength underflows solver precision.");
        } else if (option == "--material") {
            if (value == "water")
                c.water = true;
            else if (value == "constant")
                c.water = false;
            else
                throw std::invalid_argument("Material must be water or constant.");
        } else if (option == "--temperature-c") {
            c.celsius = cli::number(value);
            c.state_given = true;
        } else if (option == "--pressure-pa") {
            c.pascal = cli::number(value);
            c.state_given = true;
        } else if (option == "--ior") {
            c.ior = cli::number(value);
            c.index_given = true;
        } else if (option == "--grid")
            c.grid = cli::positive_u32(value);
        else if (option == "--query-theta")
            c.theta = cli::positive_u32(value);
        else if (option == "--query-phi")
            c.phi = cli::positive_u32(value);
        else if (option == "--inclination-deg")
            c.inclination = cli::number(value);
        else if (option == "--cdf-theta")
            c.cdf_theta = cli::positive_u32(value);
        else if (option == "--cdf-phi")
            c.cdf_phi = cli::positive_u32(value);
        else if (option == "--storage-gaussian-sigma-deg" || option == "--storage-gaussian-support")
            throw std::invalid_argument(
                "Global storage blur has been removed from rainbow_generate. Remove this option; only local Table-II "
                "diffraction and cell-mass aggregation are applied.");
        else if (option == "--max-coarsening-tv")
            c.max_tv = cli::number(value);
        else if (option == "--diffraction-sigma-deg")
            throw std::invalid_argument(
                "Manual diffraction width is not supported by rainbow_generate. Remove this option; Table II is "
                "selected automatically.");
        else if (option == "--transition-contrast")
            c.transition_contrast = cli::number(value);
        else if (option == "--cdf-max-l1")
            c.max_l1 = cli::number(value);
        else if (option == "--cdf-max-lost-mass")
            c.max_lost = cli::number(value);
        else if (option == "--device") {
            c.device = cli::integer(value);
            if (c.device < 0) throw std::invalid_argument("Invalid device ordinal.");
        } else if (option == "--stage") {
            if (value != "incoherent" && value != "path" && value != "focal" && value != "diffraction")
                throw std::invalid_argument("Unknown stage.");
            c.stage = value;
        } else if (option == "--focal-offsets") {
            size_t begin = 0;
            for (unsigned k = 0; k < 4; ++k) {
                const size_t end = value.find(',', begin);
                if ((k < 3 && end == std::string::npos) || (k == 3 && end != std::string::npos))
                    throw std::invalid_argument("Expected four focal offsets.");
                const std::string part = value.substr(begin, end == std::string::npos ? std::string::npos : end - begin);
                const uint64_t x = cli::u64(part);
                if (x > 3) throw std::invalid_argument("Focal offsets must be 0..3.");
                c.focal[k] = (unsigned)x;
                begin = end == std::string::npos ? value.size() : end + 1;
            }
        } else if (option == "--shape-coefficients") {
            size_t begin = 0;
            unsigned k = 0;
            while (begin <= value.size()) {
                const size_t end = value.find(',', begin);
                if (k >= 8) throw std::invalid_argument("At most 8 shape coefficients.");
                c.coefficients[k++] =
                    cli::number(value.substr(begin, end == std::string::npos ? std::string::npos : end - begin));
                if (end == std::string::npos) break;
                begin = end + 1;
            }
            c.coefficients_given = true;
            c.histogram_options = true;
        } else if (option == "--max-internal-hits") {
            c.max_hits = cli::positive_u32(value);
            c.histogram_options = true;
        } else if (option == "--tail-tolerance") {
            c.tail = cli::number(value);
            c.histogram_options = true;
        } else
            throw std::invalid_argument("Unknown option: " + option);
    }
    if (c.out.empty()) throw std::invalid_argument("--out is required.");
    if (!c.wavelength_given)
        throw std::invalid_argument("--wavelength-nm is required; there is no implicit wavelength.");
    if (c.water) {
        if (c.index_given)
            throw std::invalid_argument(
                "Do not specify --ior for water. It is evaluated from wavelength and state; use --material constant "
                "only for explicit reference calculations.");
        cli::check(iph_water_optics((double)c.wavelength_nm, c.celsius, c.pascal, &c.water_state));
        c.interior_index = (float)c.water_state.refractive_index;
    } else {
        if (!c.index_given || !(c.ior > 0 && c.ior <= (std::numeric_limits<float>::max)()))
            throw std::invalid_argument("Constant material requires a finite positive --ior.");
        if (c.state_given) throw std::invalid_argument("Temperature/pressure do not define a constant-index material.");
        c.interior_index = (float)c.ior;
        if (!(c.interior_index > 0)) throw std::invalid_argument("Index underflows solver precision.");
    }
    const bool coherent = c.stage != "incoherent";
    if (coherent && c.histogram_options)
        throw std::invalid_argument(
            "--shape-coefficients, --max-internal-hits and --tail-tolerance apply to --stage incoherent only.");
    if (!c.grid) c.grid = coherent ? 129u : 513u;
    if (coherent && (c.grid < 2 || (uint64_t)c.grid * c.grid > 0xffffffffull / 4))
        throw std::invalid_argument("Invalid grid dimensions or 32-bit path-major index overflow.");
    if (c.stage == "focal" || c.stage == "diffraction") {
        c.paper_sigma = table_sigma_degrees(c.radius_mm);
        if (!(c.paper_sigma > 0))
            throw std::invalid_argument(
                "Table II covers radius 0.1..1.0 mm only. No bandwidth extrapolation is performed. The existing focal "
                "path also computes diffraction; use incoherent/path for out-of-table reference runs.");
    }
    if (!c.cdf_theta) c.cdf_theta = c.theta;
    if (!c.cdf_phi) c.cdf_phi = c.phi;
    if (c.cdf_theta > c.theta || c.cdf_phi > c.phi || c.theta % c.cdf_theta || c.phi % c.cdf_phi)
        throw std::invalid_argument("CDF dimensions must divide the query dimensions without upsampling.");
    if (!(c.max_tv >= 0 && c.max_tv <= 1)) throw std::invalid_argument("Coarsening TV limit must lie in [0,1].");
    if (!(c.tail >= 0 && c.tail < 1)) throw std::invalid_argument("Tail tolerance must lie in [0,1).");
    if (c.coefficients_given) {
        if (c.sphere) throw std::invalid_argument("--sphere and --shape-coefficients are exclusive.");
        for (int k = 0; k < 8; ++k) c.shape[k] = (float)c.coefficients[k];
        if (!(c.radius_mm > 0.0f)) throw std::invalid_argument("Radius must be positive.");
    } else {
        cli::check(iph_drop_shape(c.radius_mm, c.sphere ? 1 : 0, c.shape));
    }
    namespace fs = std::filesystem;
    if (fs::exists(c.out)) throw std::runtime_error("Output already exists.");
    if (fs::exists(c.out + ".part"))
        throw std::runtime_error("Previous .part directory exists; inspect it before recomputing.");
    return true;
}

int run(int argc, char** argv)
{
    Command c;
    if (!parse(argc, argv, c)) return 0;
    const double angle = c.inclination * kPi / 180.0;
    const float kf[3] = {(float)std::cos(angle), (float)(-std::sin(angle)), 0.0f};
    bool sphere = c.sphere;
    if (!sphere) {
        bool zero = true;
        for (float v : c.shape) zero = zero && v == 0.0f;
        sphere = zero;  // Table I sphere row: the closed-form sphere is the same surface
    }
    // independent CPU geometry first: convexity certificate + projected area
    double coeffs[8];
    for (int k = 0; k < 8; ++k) coeffs[k] = (double)c.shape[k];
    const double kd[3] = {(double)kf[0], (double)kf[1], (double)kf[2]};
    iph_area area{};
    cli::check(iph_drop_projected_area(coeffs, (double)c.radius_mm, kd, &area));
    std::cout << std::setprecision(17) << "Material=" << (c.water ? "water_iapws_r9_97" : "constant")
              << " wavelength_vacuum_nm=" << c.wavelength_nm << " interior_index_solver=" << c.interior_index
              << " exterior_index=1; additional_storage_gaussian=none\n";
    const bool coherent = c.stage != "incoherent";
    if (coherent && c.paper_sigma > 0)
        std::cout << "Local diffraction sigma_primary_deg=" << c.paper_sigma
                  << " sigma_secondary_deg=" << 2 * c.paper_sigma << " query_step_theta_deg=" << 180.0 / c.theta
                  << " query_step_phi_deg=" << 360.0 / c.phi << " (phi arc step additionally varies with theta)\n";
    const auto start = std::chrono::steady_clock::now();

    iph_cdf_policy pol{c.cdf_theta, c.cdf_phi, c.max_l1, c.max_lost, c.max_tv};
    iph_ctx* ctx = nullptr;
    cli::check(iph_ctx_create(c.device, &ctx));
    struct Guard {
        iph_ctx* c;
        ~Guard() { iph_ctx_destroy(c); }
    } guard{ctx};
    std::vector<double> phi(c.cdf_phi + 1), theta((size_t)c.cdf_phi * (c.cdf_theta + 1));
    iph_cdf_report rep{};
    iph_hist_stats st{};
    iph_coherent_stats cst{};
    iph_coherent_gate gate{};
    iph_record r;
    std::memset(&r, 0, sizeof r);
    r.radius_mm = (double)c.radius_mm;
    if (coherent) {
        iph_coherent_config cc{};
        cc.radius_mm = c.radius_mm;
        cc.wavelength_nm = c.wavelength_nm;
        cc.interior_index = c.interior_index;
        cc.inclination_deg = c.inclination;
        cc.grid = c.grid;
        cc.sphere = c.sphere ? 1u : 0u;
        cc.theta_count = c.theta;
        cc.phi_count = c.phi;
        for (int k = 0; k < 4; ++k) cc.focal_offsets[k] = c.focal[k];
        cc.primary_sigma_deg = 0.0;  // Table II
        cc.transition_contrast = c.transition_contrast;
        iph_coherent* h = nullptr;
        cli::check(iph_coherent_create(ctx, &cc, &h), ctx);
        cli::check(iph_coherent_trace(h), ctx);
        cli::check(iph_coherent_evaluate(h), ctx);
        const uint32_t code = c.stage == "path" ? IPH_STAGE_PATH : (c.stage == "focal" ? IPH_STAGE_FOCAL
                                                                                        : IPH_STAGE_DIFFRACTION);
        const int rc = iph_coherent_cdf(h, code, c.allow_underresolved ? 1u : 0u, &pol, phi.data(), theta.data(),
                                        &rep, &gate);
        cli::check(iph_coherent_stats_get(h, &cst), ctx);
        cli::check(rc, ctx);
    } else {
        r.shape = sphere ? IPH_SHAPE_SPHERE : IPH_SHAPE_DROP;
        r.max_internal_hits = c.max_hits;
        r.theta_count = c.theta;
        r.phi_count = c.phi;
        r.grid_w = r.grid_h = c.grid;
        for (int k = 0; k < 8; ++k) r.shape_coefficients[k] = (double)c.shape[k];
        for (int k = 0; k < 3; ++k) r.incident[k] = (double)kf[k];
        r.interior_index = (double)c.interior_index;
        r.exterior_index = 1.0;
        r.wavelength_nm = (double)c.wavelength_nm;
        r.tail_tolerance = c.tail;
        iph_batch* b = nullptr;
        cli::check(iph_batch_create(ctx, &r, 1, &pol, &b), ctx);
        const int rc = iph_batch_run(b, IPH_RUN_TRACE | IPH_RUN_CDF);
        cli::check(iph_batch_stats(b, 0, &st), ctx);
        cli::check(iph_batch_cdf(b, 0, phi.data(), theta.data(), &rep), ctx);
        cli::check(rc, ctx);
        if (st.invalid) throw std::runtime_error("Histogram has invalid outgoing directions.");
    }

    // frame columns of the binning frame: e0 = normalize(+z - k (k.z)), e1 = k x e0
    double k[3] = {kd[0], kd[1], kd[2]};
    {
        const double n = std::sqrt(k[0] * k[0] + k[1] * k[1] + k[2] * k[2]);
        for (double& v : k) v /= n;
    }
    const double ax[3] = {0, 0, std::fabs(k[2]) < 0.9 ? 1.0 : 0.0};
    double e[3] = {ax[0], ax[1], ax[2]};
    if (std::fabs(k[2]) >= 0.9) e[0] = 1, e[2] = 0;
    {
        const double d = e[0] * k[0] + e[1] * k[1] + e[2] * k[2];
        for (int i = 0; i < 3; ++i) e[i] -= d * k[i];
        const double n = std::sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
        for (double& v : e) v /= n;
    }
    const double f[3] = {k[1] * e[2] - k[2] * e[1], k[2] * e[0] - k[0] * e[2], k[0] * e[1] - k[1] * e[0]};
    const double q = std::ldexp(1.0, -(int)st.qbits) * st.ray_weight_mm2;  // mm^2 per quantum
    // incident grid half extent and reference distances in drop units, as the solver sets them
    float solver_shape[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (!c.sphere)
        for (int k = 0; k < 8; ++k) solver_shape[k] = c.shape[k];
    const float outer = outer_radius(solver_shape);
    const double half_extent = (double)(outer * 1.01f);
    const double reference_distance = (double)(2.0f * outer);

    cli::Json j;
    j.kv("coordinate_contract_id", "rainbow.phase_cdf.coordinates.v1");
    j.kv("direction_convention", "physical_propagation");
    j.kv("particle_frame_handedness", "right");
    j.key("particle_up_axis").open_array().value(0).value(1).value(0).close_array();
    j.key("shape_polar_axis").open_array().value(0).value(-1).value(0).close_array();
    j.kv("sampling_frame_layout", "rows_xyz_columns_e0_e1_k");
    j.kv("sampling_frame_map", "local_column_to_particle_column");
    j.kv("theta_zero", "forward");
    j.kv("theta_pi", "backward");
    j.kv("azimuth_periodic", true);
    j.kv("inclination_definition", "ki=(cos(alpha),-sin(alpha),0); alpha=0 is +x; positive alpha points downward");
    j.key("geometry").open_object();
    j.kv("projected_area_mm2", area.area_mm2);
    j.kv("projected_area_m2", area.area_mm2 * 1e-6);
    j.kv("projected_area_method",
         area.analytic_sphere ? "analytic_sphere" : "convex_surface_integral_azimuth_analytic_gk15");
    j.kv("projected_area_shape", "solver_shape_coefficients_promoted_to_fp64");
    j.kv("projected_area_converged", true);
    j.kv("projected_area_absolute_error_estimate_mm2", area.abs_error_mm2);
    j.kv("projected_area_relative_error_estimate", area.rel_error);
    j.kv("projected_area_error_is_rigorous_bound", false);
    j.kv("convexity_verified", true);
    j.kv("convexity_method", area.analytic_sphere ? "analytic_sphere" : "markov_lipschitz_branch_and_bound_curvature");
    j.kv("convexity_intervals", area.convexity_cells);
    j.kv("quadrature_evaluations", area.evaluations);
    j.key("incident_axis_particle_frame").array(k, 3);
    j.close_object();
    j.key("cross_sections").open_object();
    j.kv("model", "geometric_nonabsorbing");
    j.kv("scattering_mm2", area.area_mm2);
    j.kv("extinction_mm2", area.area_mm2);
    j.kv("absorption_mm2", 0);
    j.kv("scattering_m2", area.area_mm2 * 1e-6);
    j.kv("extinction_m2", area.area_mm2 * 1e-6);
    j.kv("absorption_m2", 0);
    j.kv("assignment", "Csca=Cext=Aproj; Cabs=0 (transport model assumption)");
    j.kv("wave_scattering_cross_section_computed", false);
    j.close_object();
    j.key("material").open_object();
    j.kv("model", c.water ? "water_iapws_r9_97" : "constant");
    j.kv("wavelength_convention", "vacuum_nm");
    j.kv("wavelength_requested_nm", c.wavelength_requested);
    j.kv("wavelength_evaluated_nm", (double)c.wavelength_nm);
    j.kv("ambient_model", "unit_absolute_index_air_approximation");
    j.kv("absorption_model", "ignored_k_equals_zero");
    j.kv("interior_index_evaluated_fp64", c.water ? c.water_state.refractive_index : c.ior);
    j.kv("interior_index_solver_fp32", (double)c.interior_index);
    j.kv("exterior_index_solver", 1.0);
    if (c.water) {
        j.kv("refractive_index_release", "IAPWS R9-97");
        j.kv("density_model", "IAPWS R6-95(2018) liquid pressure EOS inversion");
        j.kv("temperature_kelvin", c.water_state.temperature_kelvin);
        j.kv("pressure_pascal", c.water_state.pressure_pascal);
        j.kv("density_kg_m3", c.water_state.density_kg_m3);
        j.kv("density_pressure_residual_pascal", c.water_state.density_pressure_residual_pascal);
        j.kv("density_bracket_width_kg_m3", c.water_state.density_bracket_width_kg_m3);
    }
    j.close_object();
    j.key("diffraction_policy").open_object();
    j.kv("name", "Sadeghi2012_TableII_local_v1");
    j.kv("computed", c.paper_sigma > 0);
    j.kv("selected_output_includes_diffraction", c.stage == "diffraction");
    if (c.paper_sigma > 0) {
        j.kv("primary_sigma_degrees", c.paper_sigma);
        j.kv("secondary_sigma_degrees", 2 * c.paper_sigma);
    } else {
        j.key("primary_sigma_degrees").null();
        j.key("secondary_sigma_degrees").null();
    }
    j.kv("bandwidth_source", "paper Table II; secondary sigma x2");
    j.kv("between_table_radii", "linear interpolation (project convention)");
    j.kv("outside_table_radii", "reject");
    j.kv("wavelength_scaling", "none; no rule specified in Table II");
    j.kv("support_sigma", 4);
    j.kv("support_and_transition_detection_are_project_conventions", true);
    j.kv("physical_sigma_depends_on_grid", false);
    j.kv("query_dtheta_degrees", 180.0 / c.theta);
    j.kv("query_dphi_degrees", 360.0 / c.phi);
    j.close_object();
    j.kv("schema", "rainbow.phase_cdf.numpy.v2");
    j.kv("complete", true);
    j.kv("dtype", "<f8");
    j.kv("order", "C");
    j.kv("theta_count", c.cdf_theta);
    j.kv("phi_count", c.cdf_phi);
    j.kv("density_measure", "solid_angle_sr");
    j.kv("cell_model", "constant_density_per_spherical_cell");
    j.kv("coordinates", "u=(1-cos(theta))/2; v=(phi+pi)/(2*pi)");
    j.kv("theta_definition", "angle_between_incident_and_outgoing_propagation_directions");
    j.key("cdf_axis_order").array(std::vector<std::string>{"phi_cell", "theta_edge"});
    const double pr[2] = {-kPi, kPi};
    j.key("phi_range_rad").array(pr, 2);
    j.kv("normalization", "integral_p_domega_equals_one");
    j.kv("input_polarization", "unpolarized");
    j.kv("incident_total_intensity", 1);
    j.kv("single_scattering_albedo", 1);
    j.kv("collision_model", "geometric_projected_area");
    j.kv("projected_area_storage", "metadata.geometry.projected_area_mm2");
    j.kv("stage", c.stage);
    j.kv("radius_mm", (double)c.radius_mm);
    j.kv("wavelength_nm", (double)c.wavelength_nm);
    j.kv("exterior_index", 1.0);
    j.kv("interior_index", (double)c.interior_index);
    j.kv("force_sphere", c.sphere);
    j.key("incident_grid").open_array().value(c.grid).value(c.grid).close_array();
    j.kv("incident_grid_half_extent", half_extent);
    j.kv("incident_grid_placement", coherent ? "vertices_over_square_1.01_outer_radius"
                                             : "cell_centres_over_square_1.01_outer_radius_misses_discarded");
    j.key("reference_distances").open_array().value(reference_distance).value(reference_distance).close_array();
    j.key("query_frame_axis").array(kd, 3);
    j.key("query_frame_e0").array(e, 3);
    j.key("sampling_frame_columns").open_array();
    for (int row = 0; row < 3; ++row) {
        const double col[3] = {e[row], f[row], k[row]};
        j.array(col, 3);
    }
    j.close_array();
    {
        const double sc[8] = {c.shape[0], c.shape[1], c.shape[2], c.shape[3],
                              c.shape[4], c.shape[5], c.shape[6], c.shape[7]};
        j.key("shape_coefficients").array(sc, 8);
    }
    j.key("focal_quarter_turn_offsets").open_array();
    for (unsigned v : c.focal) j.value(v);
    j.close_array();
    j.kv("diffraction_primary_sigma_degrees_setting", 0.0);
    j.kv("diffraction_zero_setting_selects_Table_II", true);
    j.kv("diffraction_transition_contrast", c.transition_contrast);
    j.kv("folded_branch_model", coherent ? "finite_branch_area_ratio_v1" : "none_ray_histogram");
    j.kv("folded_relative_tolerance", coherent ? 1e-8 : 0.0);
    j.kv("source_commit", "unknown");
    j.kv("source_dirty_at_configure", true);
    j.kv("generator_version", coherent ? "icephase.mi355x.drop_patches.v1" : "icephase.mi355x.drop_histogram.v1");
    j.key("quality").open_object();
    j.kv("source_coverage_certified", false);
    j.kv("angular_convergence_certified", false);
    j.kv("missing_source_cells", coherent ? cst.status_counts[4] : 0u);
    j.kv("invalid_values", coherent ? (uint64_t)gate.invalid_values : st.invalid);
    j.kv("incomplete_directions", coherent ? gate.incomplete : 0u);
    j.kv("underresolved_directions", coherent ? gate.underresolved : 0u);
    j.kv("allow_underresolved", c.allow_underresolved);
    j.kv("cdf_l1_mass_error", rep.l1_error);
    j.kv("cdf_maximum_cell_mass_error", rep.max_cell_error);
    j.kv("cdf_lost_positive_cells", rep.lost_cells);
    j.kv("cdf_lost_probability_mass", rep.lost_mass);
    j.kv("maximum_accepted_l1_mass_error", c.max_l1);
    j.kv("maximum_accepted_lost_probability_mass", c.max_lost);
    j.close_object();
    j.kv("incident_inclination_degrees", c.inclination);
    j.key("query_grid").open_array().value(c.theta).value(c.phi).close_array();
    if (coherent) {
        j.key("patch_model").open_object();
        j.kv("method", "bilinear_chord_patches_device_query");
        j.kv("patches", cst.patches);
        j.kv("folded_patches", cst.folded);
        j.kv("candidate_pairs", cst.candidates);
        j.kv("hits", cst.hits);
        j.close_object();
    } else {
        j.key("histogram").open_object();
        j.kv("method", "fixed_point_u64_ray_histogram");
        j.kv("quantum_bits", st.qbits);
        j.kv("ray_weight_mm2", st.ray_weight_mm2);
        j.kv("rays", st.rays);
        j.kv("rays_missed", st.status[8]);
        j.kv("outgoing_samples", st.samples);
        j.kv("max_internal_hits", c.max_hits);
        j.kv("tail_tolerance_per_ray", c.tail);
        j.kv("escaped_mm2", (double)st.q_escaped * q);
        j.kv("unresolved_mm2", (double)st.q_unresolved * q);
        j.kv("forward_axis_mass_mm2", (double)st.q_forward * q);
        j.kv("backward_axis_mass_mm2", (double)st.q_backward * q);
        j.kv("boundary_snapped_samples", st.snapped);
        j.close_object();
    }
    j.key("storage_processing").open_object();
    j.kv("additional_filter", "none");
    j.kv("additional_filter_is_paper_diffraction", false);
    j.kv("sigma_degrees", 0.0);
    j.kv("support_sigma", 4.0);
    j.kv("gaussian_underresolved", false);
    j.kv("aggregation", "nested_solid_angle_mass_preserving");
    j.kv("gaussian_integral_relative_change", 0.0);
    j.kv("aggregation_integral_relative_change", rep.aggregation_error);
    j.kv("coarsening_tv", rep.coarsening_tv);
    j.kv("maximum_accepted_coarsening_tv", c.max_tv);
    j.close_object();
    j.key("hg").open_object();
    j.kv("g", rep.g_stored);
    j.kv("method", "first_moment_of_saved_cell_pdf");
    j.kv("target", "saved_cdf");
    j.kv("cosine_convention", "dot(incident_propagation,outgoing_propagation)");
    j.key("axis_particle_frame").array(k, 3);
    j.kv("g_source", rep.g_source);
    j.kv("g_filtered", rep.g_source);
    j.kv("cdf_mass_check", rep.cdf_mass);
    j.close_object();
    const std::string meta = j.text();
    cli::check(iph_write_record(c.out.c_str(), meta.c_str(), phi.data(), theta.data(), c.cdf_theta, c.cdf_phi));
    const double elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
    std::cout << "Saved NumPy CDF dataset: " << c.out << "; elapsed_seconds=" << elapsed
              << "; CDF_L1_mass_error=" << rep.l1_error << "; lost_probability_mass=" << rep.lost_mass
              << "; saved_grid=" << c.cdf_theta << 'x' << c.cdf_phi << "; hg_g=" << rep.g_stored
              << "; projected_area_mm2=" << area.area_mm2 << "; coarsening_TV=" << rep.coarsening_tv << '\n';
    return 0;
}

}  // namespace

int main(int argc, char** argv)
{
    try {
        return run(argc, argv);
    } catch (const std::exception& e) {
        std::cerr << "rainbow_generate: " << e.what() << '\n';
        return 1;
    }
}
