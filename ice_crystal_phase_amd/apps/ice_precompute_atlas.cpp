// ice_precompute_atlas: deterministic single-particle atlas of a hexagonal ice
// prism, drop-in for the reference tool of the same name
// (hex_atlas/apps/precompute_atlas.cpp): same flags, the same output directory
// (mass_mm2.npy [incident_theta, incident_phi, theta, phi] FP64, u_edges.npy,
// metadata.json with the same keys, staged in <out>.part and renamed), the same
// failure convention ("ice_precompute_atlas: <message>", exit 2, an incomplete
// .part kept).
//
// The nodes are traced on the GPU through iph_atlas_trace, many nodes per
// persistent launch; the mass array is streamed to disk chunk by chunk.
// --batch-rays and --max-work-mib are accepted for compatibility: they bound
// the reference's per-ray histogram columns, which this implementation does
// not use.
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

namespace fs = std::filesystem;

// device histogram budget of one iph_atlas_trace call (u64 cells)
constexpr uint64_t kChunkCells = 1ull << 27;

struct Options {
    iph_atlas_settings s{};
    std::string out;
    int device = 0;
    bool ice = true;
    iph_ice_state optics{};
};

Options parse(int argc, char** argv)
{
    Options o;
    iph_atlas_settings& s = o.s;
    s.exterior_index = 1;
    s.max_internal_hits = 2048;
    s.tail_tolerance = 1e-12;
    s.max_unresolved_fraction = 1e-10;
    s.max_balance_error = 1e-10;
    std::set<std::string> seen;
    for (int i = 1; i < argc; ++i) {
        const std::string key = argv[i];
        if (!seen.insert(key).second || i + 1 >= argc)
            throw std::invalid_argument("Duplicate option or missing argument.");
        const std::string value = argv[++i];
        if (key == "--out")
            o.out = value;
        else if (key == "--modules")
            ;  // the kernels are in libicephase.so
        else if (key == "--device") {
            const uint64_t v = cli::u64(value);
            if (v > 2147483647u) throw std::invalid_argument("Device range.");
            o.device = (int)v;
        } else if (key == "--radius-mm")
            s.radius_mm = cli::number(value);
        else if (key == "--length-mm")
            s.length_mm = cli::number(value);
        else if (key == "--material") {
            if (value == "ice")
                o.ice = true;
            else if (value == "constant")
                o.ice = false;
            else
                throw std::invalid_argument("Material must be ice or constant.");
        } else if (key == "--ior")
            s.interior_index = cli::number(value);
        else if (key == "--exterior-ior")
            s.exterior_index = cli::number(value);
        else if (key == "--wavelength-nm")
            s.wavelength_nm = cli::number(value);
        else if (key == "--incident-theta-intervals")
            s.incident_theta_intervals = (uint32_t)cli::u64(value);
        else if (key == "--incident-phi-intervals")
            s.incident_phi_intervals = (uint32_t)cli::u64(value);
        else if (key == "--theta")
            s.theta_count = (uint32_t)cli::u64(value);
        else if (key == "--phi")
            s.phi_count = (uint32_t)cli::u64(value);
        else if (key == "--area-subdivisions")
            s.area_subdivisions = (uint32_t)cli::u64(value);
        else if (key == "--batch-rays" || key == "--max-work-mib")
            (void)cli::u64(value);  // per-ray column limits of the reference; no such columns here
        else if (key == "--max-internal-hits")
            s.max_internal_hits = (uint32_t)cli::u64(value);
        else if (key == "--tail-tolerance")
            s.tail_tolerance = cli::number(value);
        else if (key == "--max-unresolved-fraction")
            s.max_unresolved_fraction = cli::number(value);
        else if (key == "--max-balance-error")
            s.max_balance_error = cli::number(value);
        else
            throw std::invalid_argument("Unknown option: " + key);
    }
    if (o.out.empty()) throw std::invalid_argument("--out required.");
    if (!seen.count("--wavelength-nm")) throw std::invalid_argument("--wavelength-nm is required (vacuum nm).");
    if (o.ice) {
        if (seen.count("--ior"))
            throw std::invalid_argument(
                "Ice index is evaluated from wavelength; --ior requires explicit --material constant.");
        cli::check(iph_ice_index(s.wavelength_nm, &o.optics));
        s.interior_index = o.optics.n;
    } else {
        if (!seen.count("--ior")) throw std::invalid_argument("Constant material requires --ior.");
        o.optics = iph_ice_state{s.wavelength_nm, s.interior_index, 0, 0, 0, 0};
    }
    // settings validation without tracing (node 0's rule is built and checked)
    uint64_t count = 0;
    iph_atlas_node probe{};
    cli::check(iph_atlas_node_rays(&s, 0, 0, nullptr, 0, &count, &probe));
    return o;
}

void material_json(cli::Json& j, const Options& o)
{
    const iph_ice_state& m = o.optics;
    j.key("material").open_object();
    j.kv("model", o.ice ? "ice_ih_warren_brandt_2008_real_only" : "constant");
    j.kv("wavelength_convention", "vacuum_nm");
    j.kv("wavelength_nm", m.wavelength_nm);
    j.kv("wavelength_requested_nm", m.wavelength_nm);
    j.kv("wavelength_evaluated_nm", m.wavelength_nm);
    j.kv("interior_index", m.n);
    j.kv("exterior_index", o.s.exterior_index);
    j.kv("interior_index_evaluated_fp64", m.n);
    j.kv("interior_index_solver_fp64", m.n);
    j.kv("automatic_ice_dispersion", o.ice);
    j.kv("absorption", false);
    j.kv("birefringence", false);
    j.kv("absorption_model", "ignored_k_equals_zero");
    j.kv("kappa_solver", 0);
    if (o.ice) {
        j.kv("source_doi", "10.1029/2007JD009744");
        j.kv("reference_state", "ice_Ih_near_melting_compilation");
        j.kv("source_url", "https://www.atmos.washington.edu/ice_optical_constants/IOP_2008_ASCIItable.dat");
        // SHA-256 of the published rows 0.199..1.100 um as distributed with the table
        j.kv("table_subset_sha256", "71235dd08b3f65fc5fdf638101577d4da12af9cd725da21fa1677c0f91bfddcb");
        j.key("supported_nm").open_array().value(200).value(1100).close_array();
        j.kv("n_interpolation", "linear_n_in_log_vacuum_wavelength");
        j.kv("kappa_interpolation", "linear_log_kappa_in_log_vacuum_wavelength");
        j.kv("temperature_dependence_computed", false);
        j.kv("kappa_evaluated_not_used", m.kappa);
        j.key("bracket_nm").open_array().value(m.lower_nm).value(m.upper_nm).close_array();
        j.kv("log_wavelength_fraction", m.log_wavelength_fraction);
    }
    j.close_object();
}

// NPY v1.0 header of a '<f8' C-order array (the rule of the record writer).
std::string npy_header(const std::vector<uint64_t>& shape)
{
    std::string dims = "(";
    for (uint64_t n : shape) dims += std::to_string(n) + ", ";
    dims += ")";
    std::string h = "{'descr': '<f8', 'fortran_order': False, 'shape': " + dims + ", }";
    const size_t pad = (64u - ((10u + h.size() + 1u) % 64u)) % 64u;
    h.append(pad, ' ');
    h.push_back('\n');
    const unsigned char prefix[10] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0, (unsigned char)(h.size() & 255u),
                                      (unsigned char)((h.size() >> 8) & 255u)};
    return std::string((const char*)prefix, 10) + h;
}

int run(const Options& o)
{
    const iph_atlas_settings& s = o.s;
    std::cout.precision(17);
    std::cout << "material=" << (o.ice ? "ice_wb2008_real_only" : "constant") << " wavelength_nm=" << s.wavelength_nm
              << " n=" << s.interior_index << " kappa_evaluated_not_used=" << o.optics.kappa << '\n';
    const fs::path dest(o.out);
    if (dest.empty() || dest.filename().empty() || dest.filename() == "." || dest.filename() == ".." ||
        dest.extension() == ".part")
        throw std::invalid_argument("Invalid dataset destination.");
    fs::path staging = dest;
    staging += ".part";
    if (fs::exists(dest)) throw std::runtime_error("Dataset already exists; refusing to overwrite.");
    if (!dest.parent_path().empty()) fs::create_directories(dest.parent_path());
    if (!fs::create_directory(staging))
        throw std::runtime_error("Staging directory exists; inspect/retire the previous .part directory.");
    try {
        const uint64_t ni = (uint64_t)s.incident_theta_intervals + 1, nj = (uint64_t)s.incident_phi_intervals + 1;
        const uint64_t nodes = ni * nj, cells = (uint64_t)s.theta_count * s.phi_count;
        const uint64_t per_call = std::max<uint64_t>(1, std::min<uint64_t>(nodes, kChunkCells / cells));
        std::ofstream mass(staging / "mass_mm2.npy", std::ios::binary | std::ios::trunc);
        mass.exceptions(std::ios::badbit | std::ios::failbit);
        const std::string header = npy_header({ni, nj, s.theta_count, s.phi_count});
        mass.write(header.data(), (std::streamsize)header.size());
        std::vector<double> edges(s.theta_count + 1);
        cli::check(iph_u_edges(s.theta_count, edges.data()));
        {
            const uint64_t shape[1] = {edges.size()};
            cli::check(iph_write_npy((staging / "u_edges.npy").string().c_str(), shape, 1, edges.data()));
        }
        iph_ctx* ctx = nullptr;
        cli::check(iph_ctx_create(o.device, &ctx));
        struct Guard {
            iph_ctx* c;
            ~Guard() { iph_ctx_destroy(c); }
        } guard{ctx};
        std::vector<iph_atlas_node> info(nodes);
        std::vector<double> buf(per_call * cells);
        uint64_t total_rays = 0;
        for (uint64_t first = 0; first < nodes; first += per_call) {
            const uint32_t n = (uint32_t)std::min<uint64_t>(per_call, nodes - first);
            cli::check(iph_atlas_trace(ctx, &s, (uint32_t)first, n, buf.data(), info.data() + first), ctx);
            mass.write((const char*)buf.data(), (std::streamsize)(n * cells * sizeof(double)));
            for (uint32_t q = 0; q < n; ++q) {
                const iph_atlas_node& a = info[first + q];
                total_rays += a.rays;
                std::cout << "atlas_node=" << a.i << ',' << a.j << " rays=" << a.rays
                          << " residual=" << a.unresolved_mm2 / a.projected_area_mm2 << '\n';
            }
        }
        mass.close();
        cli::Json j;
        j.kv("schema", "ice.single_particle_atlas.v1");
        j.kv("complete", true);
        j.kv("backend", "hip_u64_fixed_point_histogram");
        j.kv("source_fingerprint_at_configure", "unknown");
        j.kv("dtype", "<f8");
        j.kv("order", "C");
        j.key("mass_axes").array(std::vector<std::string>{"incident_theta", "incident_phi", "theta", "phi"});
        j.kv("incident_theta_intervals", s.incident_theta_intervals);
        j.kv("incident_phi_intervals", s.incident_phi_intervals);
        j.kv("theta_count", s.theta_count);
        j.kv("phi_count", s.phi_count);
        j.key("geometry").open_object();
        j.kv("shape", "regular_hexagonal_prism");
        j.kv("circumradius_mm", s.radius_mm);
        j.kv("full_length_mm", s.length_mm);
        j.key("body_long_axis").open_array().value(0).value(1).value(0).close_array();
        j.close_object();
        material_json(j, o);
        j.kv("symmetry", "D6h_O3_full_24_scalar_unpolarized");
        j.kv("incident_reconstruction", "bilinear_theta_phi_in_spherical_incident_frame");
        j.kv("incident_formula",
             "k=(sin(t)cos(p),cos(t),sin(t)sin(p));e0=(cos(t)cos(p),-sin(t),cos(t)sin(p));e1=k_cross_e0");
        j.key("area_quadrature").open_object();
        j.kv("rule", "composite_equal_area_GL2xGL3");
        j.kv("subdivisions", s.area_subdivisions);
        j.kv("randomness", false);
        j.close_object();
        j.key("transport").open_object();
        j.kv("path_rule", "R_and_T_R_power_m_T");
        j.kv("polarization", "two_input_JonesFlux");
        j.kv("interference", false);
        j.kv("diffraction", false);
        j.kv("tail_tolerance", s.tail_tolerance);
        j.kv("max_internal_hits", s.max_internal_hits);
        j.kv("max_unresolved_fraction", s.max_unresolved_fraction);
        j.kv("max_balance_error", s.max_balance_error);
        j.close_object();
        j.kv("pole_policy", "axis_mass_uniform_in_azimuth_in_histogram_end_cap");
        j.key("nodes").open_array();
        for (const iph_atlas_node& a : info) {
            j.open_object();
            j.kv("i", a.i);
            j.kv("j", a.j);
            j.kv("theta", a.theta);
            j.kv("phi", a.phi);
            j.key("ki").array(a.ki, 3);
            j.key("e0").array(a.e0, 3);
            j.kv("projected_area_mm2", a.projected_area_mm2);
            j.kv("escaped_mm2", a.escaped_mm2);
            j.kv("unresolved_mm2", a.unresolved_mm2);
            j.kv("histogram_relative_error", a.histogram_relative_error);
            j.kv("rays", a.rays);
            j.kv("outputs", a.outputs);
            j.kv("max_balance_error", a.max_balance_error);
            j.kv("max_interface_balance_error", a.max_interface_balance_error);
            j.close_object();
        }
        j.close_array();
        j.kv("total_rays", total_rays);
        j.kv("numerical_convergence_certified", false);
        {
            std::ofstream meta(staging / "metadata.json", std::ios::binary | std::ios::trunc);
            meta.exceptions(std::ios::badbit | std::ios::failbit);
            meta << j.text();
        }
        if (fs::exists(dest)) throw std::runtime_error("Destination appeared while writing; refusing to replace it.");
        fs::rename(staging, dest);
        std::cout << "atlas_complete nodes=" << nodes << " rays=" << total_rays << '\n';
        return 0;
    } catch (...) {
        std::cerr << "Atlas generation failed. Incomplete .part retained; no ray omission or fallback.\n";
        throw;
    }
}

}  // namespace

int main(int argc, char** argv)
{
    try {
        if (argc == 2 && std::string(argv[1]) == "--help") {
            std::cout << "ice_precompute_atlas --out DIR --radius-mm R --length-mm L --wavelength-nm NM\n"
                         " --material ice (default, WB2008 n(lambda), 200..1100nm) | --material constant --ior N\n"
                         " --incident-theta-intervals T --incident-phi-intervals P --theta NT --phi NP"
                         " --area-subdivisions S\n"
                         " [--max-internal-hits 2048 --tail-tolerance 1e-12 --batch-rays 64 --max-work-mib 512]\n"
                         " [--max-unresolved-fraction 1e-10 --max-balance-error 1e-10 --exterior-ior 1]\n"
                         " [--device 0 --modules DIR (accepted, unused)]. No random seeds, smoothing, or CPU"
                         " fallback.\n";
            return 0;
        }
        return run(parse(argc, argv));
    } catch (const std::exception& e) {
        std::cerr << "ice_precompute_atlas: " << e.what() << '\n';
        return 2;
    }
}
