// ice_generate_phase: hexagonal-prism phase-function record over a weighted
// orientation plan, drop-in for the reference tool of the same name
// (hex_phase/apps/generate_hex_phase.cpp, hex_phase/src/phase_generate.cpp):
// same flags and orientation CSV, the same on-disk contract and metadata keys,
// the same failure convention (option errors: "ice_generate_phase: <message>",
// exit 1; generation failures: <out>.part/failure.json kept, exit 2).
//
// Every orientation node is one record of one batch on the GPU; the nodes add
// their samples into the first node's fixed-point histogram (bitwise
// order-independent), then the FP64 CDF is built on the device. Flag added by
// this tool: --diffraction (add the Fraunhofer lobe of each node's projected
// aperture, carrying the node's projected area).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <set>
#include <string>
#include <vector>

#include "cli_support.hpp"

namespace {

constexpr double kPi = 3.141592653589793238462643383279502884;

struct Options {
    std::string out, orientations;
    double radius = 0, length = 0, ki[3] = {0, 0, 0}, ior = 0, external_ior = 1, wavelength = 0;
    uint32_t theta = 0, phi = 0, cdf_theta = 0, cdf_phi = 0;
    uint64_t batch_size = 4096, max_output_records = 4 * 1024 * 1024;
    uint32_t max_hits = 1024;
    double tail = 1e-12, max_unresolved = 1e-10, max_balance = 1e-10, max_hist_rel = 1e-10;
    double max_l1 = 1e-10, max_lost = 1e-12, max_tv = 1.0;
    int device = 0;
    bool diffraction = false;
};

const char* kStatusNames[8] = {"exhausted", "tail_tolerance_reached", "interaction_limit", "ambiguous_boundary",
                               "invalid_input", "numerical_failure", "output_overflow", "replay_mismatch"};

Options parse(int argc, char** argv)
{
    Options o;
    std::set<std::string> seen;
    for (int i = 1; i < argc; ++i) {
        const std::string key = argv[i];
        if (!seen.insert(key).second) throw std::invalid_argument("Duplicate option: " + key);
        auto next = [&]() -> std::string {
            if (++i >= argc) throw std::invalid_argument("Missing value: " + key);
            return argv[i];
        };
        if (key == "--out")
            o.out = next();
        else if (key == "--orientations")
            o.orientations = next();
        else if (key == "--modules")
            (void)next();  // kernels are in libicephase.so
        else if (key == "--radius-mm")
            o.radius = cli::number(next());
        else if (key == "--length-mm")
            o.length = cli::number(next());
        else if (key == "--ki") {
            for (double& v : o.ki) v = cli::number(next());
        } else if (key == "--ior")
            o.ior = cli::number(next());
        else if (key == "--external-ior")
            o.external_ior = cli::number(next());
        else if (key == "--wavelength-nm")
            o.wavelength = cli::number(next());
        else if (key == "--theta")
            o.theta = (uint32_t)cli::u64(next());
        else if (key == "--phi")
            o.phi = (uint32_t)cli::u64(next());
        else if (key == "--cdf-theta")
            o.cdf_theta = (uint32_t)cli::u64(next());
        else if (key == "--cdf-phi")
            o.cdf_phi = (uint32_t)cli::u64(next());
        else if (key == "--batch-size")
            o.batch_size = cli::u64(next());
        else if (key == "--max-output-records")
            o.max_output_records = cli::u64(next());
        else if (key == "--max-internal-hits")
            o.max_hits = (uint32_t)cli::u64(next());
        else if (key == "--tail-tolerance")
            o.tail = cli::number(next());
        else if (key == "--max-unresolved-fraction")
            o.max_unresolved = cli::number(next());
        else if (key == "--max-balance-error")
            o.max_balance = cli::number(next());
        else if (key == "--max-histogram-relative-error")
            o.max_hist_rel = cli::number(next());
        else if (key == "--cdf-max-l1")
            o.max_l1 = cli::number(next());
        else if (key == "--cdf-max-lost-mass")
            o.max_lost = cli::number(next());
        else if (key == "--max-coarsening-tv")
            o.max_tv = cli::number(next());
        else if (key == "--device")
            o.device = cli::integer(next());
        else if (key == "--diffraction") {
            o.diffraction = true;
        } else
            throw std::invalid_argument("Unknown option: " + key);
    }
    for (const char* k : {"--out", "--orientations", "--radius-mm", "--length-mm", "--ki", "--ior", "--wavelength-nm",
                          "--theta", "--phi"})
        if (!seen.count(k)) throw std::invalid_argument(std::string("Required option: ") + k);
    if (o.device < 0) throw std::invalid_argument("Negative device ordinal.");
    if (!seen.count("--cdf-theta")) o.cdf_theta = o.theta;
    if (!seen.count("--cdf-phi")) o.cdf_phi = o.phi;
    // validate_generation_settings (hex_phase/src/phase_generate.cpp:29-38)
    if (!o.theta || !o.phi || (uint64_t)o.theta * o.phi > 0xffffffffull)
        throw std::invalid_argument("Unsupported phase grid dimensions.");
    if (!(std::isfinite(o.max_l1) && o.max_l1 >= 0 && o.max_l1 <= 2 && std::isfinite(o.max_lost) && o.max_lost >= 0 &&
          o.max_lost <= 1 && std::isfinite(o.max_tv) && o.max_tv >= 0 && o.max_tv <= 1))
        throw std::invalid_argument("Invalid CDF quality policy.");
    if (!(o.wavelength > 0) || !std::isfinite(o.wavelength) || !o.batch_size || !o.max_output_records ||
        !o.cdf_theta || !o.cdf_phi || o.cdf_theta > o.theta || o.cdf_phi > o.phi || o.theta % o.cdf_theta ||
        o.phi % o.cdf_phi)
        throw std::invalid_argument("Invalid wavelength, capacity, or coarsening dimensions.");
    for (double t : {o.max_unresolved, o.max_balance, o.max_hist_rel})
        if (!std::isfinite(t) || t < 0 || t >= 1) throw std::invalid_argument("Invalid quality tolerance.");
    if (!(std::isfinite(o.ior) && o.ior > 0 && std::isfinite(o.external_ior) && o.external_ior > 0))
        throw std::invalid_argument("Refractive indices must be finite and positive.");
    if (!o.max_hits) throw std::invalid_argument("max_internal_hits must be positive.");
    if (!(std::isfinite(o.tail) && o.tail >= 0 && o.tail < 1))
        throw std::invalid_argument("tail_tolerance must lie in [0,1).");
    double a = 0;
    cli::check(iph_hex_projected_area(o.radius, o.length, o.ki, &a));
    return o;
}

void write_failure(const std::string& dir, const std::string& reason, const iph_hist_stats* st)
{
    cli::Json j;
    j.kv("complete", false);
    j.kv("error", reason);
    j.kv("processed_rays", st ? st->rays : 0);
    j.key("status_counts").open_object();
    for (int i = 0; i < 8; ++i) j.kv(kStatusNames[i], st ? st->status[i] : 0);
    j.close_object();
    std::ofstream f(dir + "/failure.json", std::ios::binary | std::ios::trunc);
    f << j.text();
}

int run(const Options& o)
{
    const cli::OrientationPlan plan = cli::read_orientations(o.orientations);
    // records: one per orientation node; node k > 0 adds into node 0
    std::vector<iph_record> recs(plan.nodes.size());
    double expected_area = 0;
    for (size_t k = 0; k < plan.nodes.size(); ++k) {
        const cli::OrientationNode& n = plan.nodes[k];
        iph_record& r = recs[k];
        std::memset(&r, 0, sizeof r);
        r.shape = IPH_SHAPE_HEX;
        r.max_internal_hits = o.max_hits;
        r.theta_count = o.theta;
        r.phi_count = o.phi;
        r.rays = n.samples;
        r.seed = n.seed;
        r.radius_mm = o.radius;
        r.length_mm = o.length;
        for (int i = 0; i < 3; ++i) r.incident[i] = o.ki[i];
        r.interior_index = o.ior;
        r.exterior_index = o.external_ior;
        r.wavelength_nm = o.wavelength;
        r.tail_tolerance = o.tail;
        std::memcpy(r.rotation, n.rotation, sizeof r.rotation);
        r.weight = n.weight;
        r.histogram_of = k == 0 ? 0 : 1;
        r.flags = o.diffraction ? IPH_REC_DIFFRACTION : 0;
    }
    namespace fs = std::filesystem;
    const std::string staging = o.out + ".part";
    if (fs::exists(o.out)) throw std::runtime_error("Dataset already exists; refusing to overwrite.");
    if (!fs::create_directory(staging))
        throw std::runtime_error("Staging directory exists; inspect/retire the previous .part directory.");
    iph_hist_stats st{};
    bool have_stats = false;
    try {
        iph_ctx* ctx = nullptr;
        cli::check(iph_ctx_create(o.device, &ctx));
        struct Guard {
            iph_ctx* c;
            ~Guard() { iph_ctx_destroy(c); }
        } guard{ctx};
        iph_cdf_policy pol{o.cdf_theta, o.cdf_phi, o.max_l1, o.max_lost, o.max_tv};
        iph_batch* b = nullptr;
        cli::check(iph_batch_create(ctx, recs.data(), (uint32_t)recs.size(), &pol, &b), ctx);
        const int rc = iph_batch_run(b, IPH_RUN_TRACE | IPH_RUN_CDF);
        cli::check(iph_batch_stats(b, 0, &st), ctx);
        have_stats = true;
        std::vector<double> phi(o.cdf_phi + 1), theta((size_t)o.cdf_phi * (o.cdf_theta + 1));
        iph_cdf_report rep;
        cli::check(iph_batch_cdf(b, 0, phi.data(), theta.data(), &rep), ctx);
        expected_area = st.projected_area_mm2;
        std::cout << "orientation=" << plan.nodes.size() << '/' << plan.nodes.size() << " rays=" << st.rays << '\n';
        // composition audits (validate_composition, hex_phase/src/phase_generate.cpp:64-82)
        // accepted statuses are exhausted and tail_tolerance_reached only
        // (hex_trace/include/ice_crystal/hex_trace_core.hpp:64-65, hex_trace.cpp:102-105)
        for (int s = 2; s < 8; ++s)
            if (st.status[s])
                throw std::runtime_error(
                    "Hex trace is NOT accepted. Inspect per-ray statuses and unresolved power; do not renormalize it "
                    "away.");
        if (st.rays != plan.total_samples || st.invalid) throw std::runtime_error("Incomplete composition / output-count mismatch.");
        const double q = std::ldexp(1.0, -(int)st.qbits) * st.ray_weight_mm2;  // mm^2 per quantum
        const double escaped = (double)st.q_escaped * q, unresolved = (double)st.q_unresolved * q;
        if (!(expected_area > 0) || !(escaped > 0) || !std::isfinite(escaped) || !(unresolved >= 0))
            throw std::runtime_error("Invalid integrated cross sections.");
        if (unresolved / expected_area > o.max_unresolved)
            throw std::runtime_error("Unresolved light exceeds explicit global budget.");
        const double balance = std::fabs((escaped + unresolved) / expected_area - 1);
        if (balance > o.max_balance) throw std::runtime_error("Orientation-weighted energy balance failed.");
        // grid mass vs the independent per-ray audit: the histogram's quanta (cells
        // + poles, without the lobe) against the escaped quanta, exact in integers
        std::vector<uint64_t> bins((size_t)o.theta * o.phi);
        cli::check(iph_batch_bins(b, 0, bins.data()), ctx);
        unsigned long long grid_q = st.q_forward + st.q_backward;
        for (uint64_t v : bins) grid_q += v;
        grid_q -= st.q_diffraction;
        const double hist_rel = std::fabs((double)grid_q / (double)st.q_escaped - 1);
        if (hist_rel > o.max_hist_rel) throw std::runtime_error("Histogram mass differs from independent escaped-ray audit.");
        cli::check(rc, ctx);

        const double lobe_mm2 = (double)st.q_diffraction * q;
        const double total_area = expected_area + (o.diffraction ? expected_area : 0.0);
        // frame of the ensemble direction (make_frame: the world axis of smallest |k_i|)
        double k[3] = {o.ki[0], o.ki[1], o.ki[2]};
        {
            const double s = std::max(std::fabs(k[0]), std::max(std::fabs(k[1]), std::fabs(k[2])));
            for (double& v : k) v /= s;
            const double n = std::sqrt(k[0] * k[0] + k[1] * k[1] + k[2] * k[2]);
            for (double& v : k) v /= n;
        }
        int ax = 0;
        for (int i = 1; i < 3; ++i)
            if (std::fabs(k[i]) < std::fabs(k[ax])) ax = i;
        double e0[3] = {0, 0, 0};
        e0[ax] = 1;
        {
            const double d = e0[0] * k[0] + e0[1] * k[1] + e0[2] * k[2];
            for (int i = 0; i < 3; ++i) e0[i] -= k[i] * d;
            const double n = std::sqrt(e0[0] * e0[0] + e0[1] * e0[1] + e0[2] * e0[2]);
            for (double& v : e0) v /= n;
        }
        const double e1[3] = {k[1] * e0[2] - k[2] * e0[1], k[2] * e0[0] - k[0] * e0[2], k[0] * e0[1] - k[1] * e0[0]};

        cli::Json j;
        j.kv("schema", "rainbow.phase_cdf.numpy.v2");
        j.kv("complete", true);
        j.kv("generator_version", "icephase.mi355x.hex_phase.v1");
        j.kv("backend", "hip_gfx950_fixed_point_histogram_device_fp64_cdf");
        j.kv("base_revision", "none");
        j.kv("source_fingerprint_at_configure", "unknown");
        j.kv("dtype", "<f8");
        j.kv("order", "C");
        j.kv("theta_count", o.cdf_theta);
        j.kv("phi_count", o.cdf_phi);
        j.kv("density_measure", "solid_angle_sr");
        j.kv("cell_model", "constant_density_per_spherical_cell");
        j.kv("coordinates", "u=(1-cos(theta))/2; v=(phi+pi)/(2*pi)");
        j.key("cdf_axis_order").array(std::vector<std::string>{"phi_cell", "theta_edge"});
        j.kv("normalization", "integral_p_domega_equals_one");
        j.kv("direction_convention", "physical_propagation");
        j.kv("coordinate_contract_id", "ice.phase_cdf.ensemble_coordinates.v1");
        j.kv("sampling_frame_layout", "rows_xyz_columns_e0_e1_k");
        j.kv("sampling_frame_map", "local_column_to_ensemble_reference_column");
        j.kv("theta_zero", "forward");
        j.kv("theta_pi", "backward");
        j.kv("azimuth_periodic", true);
        j.kv("input_polarization", "unpolarized");
        j.kv("incident_total_intensity", 1);
        j.kv("single_scattering_albedo", 1);
        j.kv("stage", "incoherent");
        j.kv("wavelength_nm", o.wavelength);
        j.kv("interior_index", o.ior);
        j.kv("exterior_index", o.external_ior);
        j.key("query_frame_axis").array(k, 3);
        j.key("query_frame_e0").array(e0, 3);
        j.key("sampling_frame_columns").open_array();
        for (int r = 0; r < 3; ++r) {
            const double col[3] = {e0[r], e1[r], k[r]};
            j.array(col, 3);
        }
        j.close_array();
        j.key("geometry").open_object();
        j.kv("shape", "regular_hexagonal_prism");
        j.key("body_long_axis").open_array().value(0).value(1).value(0).close_array();
        j.kv("circumradius_mm", o.radius);
        j.kv("full_length_mm", o.length);
        j.kv("projected_area_mm2", expected_area);
        j.kv("projected_area_m2", expected_area * 1e-6);
        j.kv("projected_area_method", "number_weighted_mean_of_convex_face_projections");
        j.close_object();
        j.key("cross_sections").open_object();
        j.kv("model", o.diffraction ? "geometric_plus_fraunhofer_nonabsorbing" : "geometric_nonabsorbing");
        j.kv("scattering_mm2", total_area);
        j.kv("extinction_mm2", total_area);
        j.kv("absorption_mm2", 0);
        j.kv("scattering_m2", total_area * 1e-6);
        j.kv("extinction_m2", total_area * 1e-6);
        j.kv("absorption_m2", 0);
        j.kv("wave_scattering_cross_section_computed", false);
        j.kv("assignment", o.diffraction ? "Csca=Cext=2*number_mean_Aproj (rays + diffraction); Cabs=0"
                                         : "Csca=Cext=number_mean_Aproj; Cabs=0 (transport model)");
        j.close_object();
        j.key("material").open_object();
        j.kv("model", "caller_supplied_real_index_at_wavelength");
        j.kv("wavelength_convention", "vacuum_nm");
        j.kv("absorption_model", "ignored_k_equals_zero");
        j.kv("birefringence", false);
        j.kv("automatic_ice_dispersion", false);
        j.close_object();
        j.key("diffraction_policy").open_object();
        j.kv("computed", o.diffraction);
        j.kv("selected_output_includes_diffraction", o.diffraction);
        if (o.diffraction) {
            j.kv("model", "scalar_fraunhofer_projected_aperture_closed_form_polygon_transform");
            j.kv("energy", "projected_area_per_node (extinction paradox)");
            j.kv("lobe_mm2", lobe_mm2);
            j.kv("captured_before_normalization", st.diffraction_captured);
        }
        j.close_object();
        j.kv("inter_path_interference", false);
        j.kv("tir_relative_phase_retained", true);
        j.key("orientation").open_object();
        j.kv("representation", "weighted_discrete_SO3_number_measure");
        j.kv("rotation", "active_Hamilton_quaternion_wxyz_body_to_ensemble");
        j.kv("rows", (uint64_t)plan.nodes.size());
        j.kv("csv_sha256", plan.sha256);
        j.kv("input_weight_sum", plan.input_weight_sum);
        j.kv("roundoff_weight_normalization", 1 / plan.input_weight_sum);
        j.kv("plan_embedded", false);
        j.kv("retain_input_csv_for_reproduction", true);
        j.kv("symmetry_reduction", false);
        j.close_object();
        j.key("sampling").open_object();
        j.kv("generator", "counter_splitmix64_open52_v1");
        j.kv("planned_rays", plan.total_samples);
        j.kv("processed_rays", st.rays);
        j.kv("outgoing_samples", st.samples);
        j.kv("batch_size", o.batch_size);
        j.kv("max_output_records_per_batch", o.max_output_records);
        j.kv("weight", "orientation_number_mass * Aproj / node_samples * path_power");
        j.close_object();
        j.key("transport_audit").open_object();
        j.kv("escaped_mm2", escaped);
        j.kv("unresolved_mm2", unresolved);
        j.kv("unresolved_fraction", unresolved / expected_area);
        j.kv("maximum_balance_error", balance);
        j.key("maximum_interface_balance_error").null();
        j.kv("max_internal_hits", o.max_hits);
        j.kv("tail_tolerance_per_ray", o.tail);
        j.kv("max_unresolved_fraction", o.max_unresolved);
        j.kv("max_balance_error", o.max_balance);
        j.kv("cdf_normalized_to_resolved_escaped_mass", true);
        j.key("status_counts").open_object();
        for (int s = 0; s < 8; ++s) j.kv(kStatusNames[s], st.status[s]);
        j.close_object();
        j.close_object();
        j.key("histogram").open_object();
        j.kv("theta_count", o.theta);
        j.kv("phi_count", o.phi);
        j.kv("mass_mm2", (double)grid_q * q);
        j.kv("relative_mass_audit_error", hist_rel);
        j.kv("maximum_accepted_relative_mass_error", o.max_hist_rel);
        j.kv("g_rays", rep.g_source);
        j.kv("forward_axis_mass_mm2", (double)st.q_forward * q);
        j.kv("backward_axis_mass_mm2", (double)st.q_backward * q);
        j.kv("pole_policy", "axis_mass_uniform_in_azimuth_in_histogram_end_cap");
        j.kv("pole_roundoff_tolerance", 256 * 2.220446049250313e-16);
        j.kv("atomic_absorbed_addends", 0);
        j.kv("atomic_absorbed_addend_mass_mm2", 0.0);
        j.kv("boundary_snapped_samples", st.snapped);
        j.kv("boundary_snap_rule", "64eps_coordinate_scaled_capped_at_1e-7_cell");
        j.kv("atomic_diagnostic_is_error_bound", false);
        j.kv("accumulation", "u64_fixed_point_quanta_order_independent");
        j.kv("quantum_bits", st.qbits);
        j.kv("diffraction_mm2", lobe_mm2);
        j.close_object();
        j.key("storage_processing").open_object();
        j.kv("gaussian_sigma_degrees", 0);
        j.kv("mass_conservative_aggregation", true);
        j.kv("coarsening_tv", rep.coarsening_tv);
        j.kv("maximum_accepted_coarsening_tv", o.max_tv);
        j.kv("aggregation_integral_relative_change", rep.aggregation_error);
        j.kv("g_source", rep.g_source);
        j.kv("g_stored", rep.g_stored);
        j.close_object();
        j.key("hg").open_object();
        j.kv("g", rep.g_stored);
        j.kv("method", "first_moment_of_saved_cell_pdf");
        j.kv("target", "saved_cdf");
        j.kv("cosine_convention", "dot(incident_propagation,outgoing_propagation)");
        j.close_object();
        j.key("quality").open_object();
        j.kv("source_coverage_certified", false);
        j.kv("angular_convergence_certified", false);
        j.kv("orientation_convergence_certified", false);
        j.kv("continuous_wave_solution_certified", false);
        j.kv("cdf_l1_mass_error", rep.l1_error);
        j.kv("cdf_maximum_cell_mass_error", rep.max_cell_error);
        j.kv("cdf_lost_positive_cells", rep.lost_cells);
        j.kv("cdf_lost_probability_mass", rep.lost_mass);
        j.kv("maximum_accepted_l1_mass_error", o.max_l1);
        j.kv("maximum_accepted_lost_probability_mass", o.max_lost);
        j.kv("cdf_mass", rep.cdf_mass);
        j.close_object();
        const std::string meta = j.text();
        fs::remove(staging);  // iph_write_record stages the record itself
        cli::check(iph_write_record(o.out.c_str(), meta.c_str(), phi.data(), theta.data(), o.cdf_theta, o.cdf_phi));
        std::cout.precision(17);
        std::cout << "saved=" << o.out << " g=" << rep.g_stored << " Csca_mm2=" << expected_area
                  << " unresolved_fraction=" << unresolved / expected_area << " histogram_relative_error=" << hist_rel
                  << " cdf_l1=" << rep.l1_error << '\n';
        return 0;
    } catch (const std::exception& e) {
        if (!fs::exists(staging)) fs::create_directory(staging);
        write_failure(staging, e.what(), have_stats ? &st : nullptr);
        std::cerr << "Generation failed; retained " << staging << ": " << e.what() << '\n';
        return 2;
    }
}

}  // namespace

int main(int argc, char** argv)
{
    try {
        if (argc == 2 && std::string(argv[1]) == "--help") {
            std::cout << "ice_generate_phase --out DIR --orientations CSV\n"
                         " --radius-mm R --length-mm L --ki X Y Z --ior N --wavelength-nm NM\n"
                         " --theta NT --phi NP [--cdf-theta CT --cdf-phi CP]\n"
                         " [--batch-size N --max-output-records N --max-internal-hits N --tail-tolerance E]\n"
                         " [--max-unresolved-fraction E --max-balance-error E --max-histogram-relative-error E]\n"
                         " [--cdf-max-l1 E --cdf-max-lost-mass E --max-coarsening-tv E]\n"
                         " [--device N --modules DIR (accepted, unused)] [--diffraction]\n"
                         "No implicit material dispersion, smoothing, symmetry folding, or CPU fallback.\n";
            return 0;
        }
        return run(parse(argc, argv));
    } catch (const std::exception& e) {
        std::cerr << "ice_generate_phase: " << e.what() << '\n';
        return 1;
    }
}
