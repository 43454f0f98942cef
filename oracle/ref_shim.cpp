// Test infrastructure only. A thin C API over the reference's own CPU sources,
// compiled from where they lie in the reference checkout (see oracle/Makefile)
// into oracle/_ref/libref.so. It lets the tests pin the oracle restatement
// against the compiled reference and write golden vectors. Nothing in the
// product links or loads it.
#include <ice_crystal/hex_trace.hpp>
#include <ice_crystal/phase_table.hpp>
#include <rainbow/phase_cdf_math.hpp>
#include <rainbow/water_refractive_index.hpp>
#include <rainbow/projected_area.hpp>
#include <rainbow/raindrop_shape.hpp>
#include <rainbow/raindrop_intersection.hpp>
#include <rainbow/raindrop_paths.hpp>
#include <ice_crystal/phase_generate.hpp>
#include <ice_crystal/orientation_plan.hpp>

#include <cstdint>
#include <cstring>
#include <exception>
#include <vector>

using namespace iceCrystal;

extern "C" {

struct ref_sample { double dir[3]; double power; double length_mm; uint64_t id;
                    uint32_t entry_face, exit_face, reflections, kind; };
struct ref_audit { double escaped, unresolved, balance, max_interface; uint64_t outputs;
                   uint32_t internal_hits, tir_count, status, pad; };

int ref_u_edges(uint32_t n, double* out)
{
    try { auto e = rainbow::make_phase_u_edges(n); std::memcpy(out, e.data(), e.size() * sizeof(double)); return 0; }
    catch (const std::exception&) { return -1; }
}

int ref_make_frame(const double k[3], double ki[3], double e0[3], double e1[3])
{
    try {
        const auto g = make_phase_grid(1, 1, {k[0], k[1], k[2]});
        ki[0] = g.ki.x; ki[1] = g.ki.y; ki[2] = g.ki.z;
        e0[0] = g.frame.e0.x; e0[1] = g.frame.e0.y; e0[2] = g.frame.e0.z;
        e1[0] = g.frame.e1.x; e1[1] = g.frame.e1.y; e1[2] = g.frame.e1.z;
        return 0;
    } catch (const std::exception&) { return -1; }
}

// Bins n directions on the grid of make_phase_grid(nt, np, ki).
int ref_phase_bin(uint32_t nt, uint32_t np, const double k[3], uint64_t n, const double* dirs,
                  uint64_t* index, int32_t* pole, int32_t* valid, int32_t* snapped)
{
    try {
        const auto g = make_phase_grid(nt, np, {k[0], k[1], k[2]});
        for (uint64_t i = 0; i < n; ++i) {
            const auto b = phase_bin(g.view(), {dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]});
            index[i] = b.index; pole[i] = b.pole; valid[i] = b.valid; snapped[i] = b.boundary_snapped;
        }
        return 0;
    } catch (const std::exception&) { return -1; }
}

// trace_batch_cpu over [first, first+count). Returns the sample count or <0.
int64_t ref_hex_trace(double R, double L, const double k[3], double ni, double nt, uint64_t seed,
                      uint64_t total, uint32_t max_hits, double tail, uint64_t first, uint64_t count,
                      ref_sample* out, uint64_t cap, ref_audit* audits)
{
    try {
        const auto s = make_trace_settings({R, L}, {k[0], k[1], k[2]}, ni, nt, total, seed, max_hits, tail);
        const auto b = trace_batch_cpu(s, first, count);
        if (b.outgoing.size() > cap) return -1;
        for (std::size_t i = 0; i < b.outgoing.size(); ++i) {
            const auto& o = b.outgoing[i];
            ref_sample& r = out[i];
            r.dir[0] = o.direction.x; r.dir[1] = o.direction.y; r.dir[2] = o.direction.z;
            r.power = o.power_fraction; r.length_mm = o.internal_length_mm; r.id = o.incident_sample_id;
            r.entry_face = o.entry_face; r.exit_face = o.exit_face; r.reflections = o.internal_reflections;
            r.kind = static_cast<uint32_t>(o.kind);
        }
        for (std::size_t i = 0; i < b.audits.size() && audits; ++i) {
            const auto& a = b.audits[i];
            audits[i] = {a.escaped_power, a.unresolved_power, a.balance_error, a.max_interface_balance_error,
                         a.output_count, a.internal_hits, a.tir_count, static_cast<uint32_t>(a.status), 0};
        }
        return static_cast<int64_t>(b.outgoing.size());
    } catch (const std::exception&) { return -2; }
}

// Single-orientation PhaseMassAccumulator over all rays of the record.
int ref_hex_masses(double R, double L, const double k[3], double ni, double nt, uint64_t seed,
                   uint64_t total, uint32_t max_hits, double tail, uint32_t gnt, uint32_t gnp,
                   double* mass, double* forward, double* backward)
{
    try {
        const auto s = make_trace_settings({R, L}, {k[0], k[1], k[2]}, ni, nt, total, seed, max_hits, tail);
        PhaseMassAccumulator h(make_phase_grid(gnt, gnp, s.incident_direction));
        const double scale = s.prism.projected_area_mm2(s.incident_direction) / double(total);
        for (uint64_t first = 0; first < total;) {
            const uint64_t count = std::min<uint64_t>(4096, total - first);
            const auto b = trace_batch_cpu(s, first, count);
            h.add(b.outgoing, Rotation{}, scale);
            first += count;
        }
        const auto m = h.masses();
        std::memcpy(mass, m.data(), m.size() * sizeof(double));
        const auto st = h.statistics();
        *forward = st.forward_mass; *backward = st.backward_mass;
        return 0;
    } catch (const std::exception&) { return -1; }
}

struct ref_cdf_report { double g_source, g_stored, coarsening_tv, cdf_mass, l1_error, lost_mass,
                        max_cell_error, aggregation_error; uint64_t lost_cells; };

int ref_build_host_cdf(const double* mass, uint32_t nt, uint32_t np, const double k[3], uint32_t ont,
                       uint32_t onp, double max_l1, double max_lost, double max_tv,
                       double* phi, double* theta, ref_cdf_report* rep)
{
    try {
        const auto g = make_phase_grid(nt, np, {k[0], k[1], k[2]});
        CdfPolicy p; p.max_l1_error = max_l1; p.max_lost_mass = max_lost; p.max_coarsening_tv = max_tv;
        const auto c = build_host_cdf(g, {mass, std::size_t(nt) * np}, ont, onp, p);
        std::memcpy(phi, c.phi.data(), c.phi.size() * sizeof(double));
        std::memcpy(theta, c.theta.data(), c.theta.size() * sizeof(double));
        const auto& r = c.report;
        *rep = {r.g_source, r.g_stored, r.coarsening_tv, r.cdf_mass, r.l1_error, r.lost_mass,
                r.maximum_cell_error, r.aggregation_relative_error, r.lost_cells};
        return 0;
    } catch (const std::exception&) { return -1; }
}

int ref_water_index(double lambda_nm, double celsius, double pressure, double* n, double* rho)
{
    try {
        const auto s = rainbow::evaluate_water_optics(lambda_nm, celsius, pressure);
        *n = s.refractive_index; *rho = s.density_kg_m3;
        return 0;
    } catch (const std::exception&) { return -1; }
}

int ref_projected_area(const double c[8], double radius_mm, const double k[3], double* area, int* sphere)
{
    try {
        std::array<double, 8> coeffs{};
        for (int i = 0; i < 8; ++i) coeffs[i] = c[i];
        const auto r = rainbow::compute_projected_area(coeffs, radius_mm, {k[0], k[1], k[2]});
        *area = r.area_mm2; *sphere = r.analytic_sphere;
        return 0;
    } catch (const std::exception&) { return -1; }
}

int ref_drop_shape(float radius_mm, int sphere, float c[8], float* inner, float* outer)
{
    rainbow::RaindropShape s{};
    if (!rainbow::RaindropShape::try_make(radius_mm, sphere != 0, s)) return -1;
    for (int i = 0; i < 8; ++i) c[i] = s.coefficients[i];
    *inner = s.inner_radius; *outer = s.outer_radius;
    return 0;
}

// The reference CLI's composition (hex_phase/apps/generate_hex_phase.cpp:73-118,
// CPU branch): every node traced with node_trace_settings and accumulated with
// PhaseMassAccumulator::add(outgoing, rotation, weight*area/samples).
int ref_orientation_masses(double R, double L, const double k[3], double ni, double nti, uint32_t max_hits,
                           double tail, const char* csv, uint32_t nt, uint32_t np, double* mass, double* area_out)
{
    try {
        PhaseGenerationSettings s{};
        s.prism.circumradius_mm = R; s.prism.length_mm = L;
        s.grid = make_phase_grid(nt, np, Vec3{k[0], k[1], k[2]});
        s.interior_index = nti; s.exterior_index = ni; s.wavelength_nm = 550;
        s.max_internal_hits = max_hits; s.tail_tolerance = tail;
        s.out_nt = nt; s.out_np = np;
        const auto plan = parse_orientation_plan(csv);
        PhaseMassAccumulator acc(s.grid);
        double area = 0;
        for (const auto& node : plan.nodes) {
            const auto c = node_trace_settings(s, node);
            const double scale = node.weight * node_area(s, node) / double(node.samples);
            area += node.weight * node_area(s, node);
            const auto batch = trace_batch_cpu(c, 0, node.samples, 64u * 1024 * 1024);
            acc.add(batch.outgoing, node.rotation, scale);
        }
        const auto m = acc.masses();
        std::memcpy(mass, m.data(), m.size() * sizeof(double));
        *area_out = area;
        return 0;
    } catch (const std::exception&) { return -1; }
}

// ---- the reference's drop geometry and path trace (header-only), for pinning
// the oracle's drop intersector against it

static rainbow::RaindropShape drop_shape_of(const float c[8], float inner, float outer)
{
    rainbow::RaindropShape s{};
    for (int i = 0; i < 8; ++i) s.coefficients[i] = c[i];
    s.inner_radius = inner; s.outer_radius = outer;
    return s;
}

// RaindropShape::implicit_value / gradient (include/rainbow/raindrop_shape.hpp:85-100)
// in double (precision 64) or float (32) at n points.
int ref_drop_surface(const float c[8], float inner, float outer, int precision, uint64_t n, const double* pts,
                     double* f, double* grad)
{
    const auto s = drop_shape_of(c, inner, outer);
    for (uint64_t i = 0; i < n; ++i) {
        if (precision == 64) {
            const rainbow::Vec3T<double> p{pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
            f[i] = s.implicit_value(p);
            const auto g = s.gradient(p);
            grad[3 * i] = g.x; grad[3 * i + 1] = g.y; grad[3 * i + 2] = g.z;
        } else {
            const rainbow::Vec3 p{float(pts[3 * i]), float(pts[3 * i + 1]), float(pts[3 * i + 2])};
            f[i] = s.implicit_value(p);
            const auto g = s.gradient(p);
            grad[3 * i] = g.x; grad[3 * i + 1] = g.y; grad[3 * i + 2] = g.z;
        }
    }
    return 0;
}

// RaindropIntersector::intersect (include/rainbow/raindrop_intersection.hpp:337-382):
// FP32 solve, FP64 re-solve when unresolved or grazing. Returns the BoundaryCode
// (0 miss, 1 hit, 2 unresolved) and, on a hit, position, unit outward normal,
// distance and the fallback flag.
int ref_drop_intersect(const float c[8], float inner, float outer, const float origin[3], const float dir[3],
                       float start, int inside, float pos[3], float nrm[3], float* dist, int* fallback)
{
    const auto s = drop_shape_of(c, inner, outer);
    const rainbow::RaindropIntersector is{s};
    rainbow::BoundaryRay ray{};
    ray.physical_origin = {origin[0], origin[1], origin[2]};
    ray.direction = {dir[0], dir[1], dir[2]};
    ray.start_distance = start;
    ray.starts_inside = inside != 0;
    const auto h = is.intersect(ray);
    pos[0] = h.position.x; pos[1] = h.position.y; pos[2] = h.position.z;
    nrm[0] = h.outward_normal.x; nrm[1] = h.outward_normal.y; nrm[2] = h.outward_normal.z;
    *dist = h.distance;
    *fallback = (h.diagnostics & rainbow::IntersectionFallback) != 0;
    return static_cast<int>(h.code);
}

// detail::solve_surface<double> alone: the FP64 root distance (0x1p-44 relative
// bracket width) of the same ray.
int ref_drop_solve64(const float c[8], float inner, float outer, const float origin[3], const float dir[3],
                     float start, int inside, double* dist)
{
    const auto s = drop_shape_of(c, inner, outer);
    rainbow::BoundaryRay ray{};
    ray.physical_origin = {origin[0], origin[1], origin[2]};
    ray.direction = {dir[0], dir[1], dir[2]};
    ray.start_distance = start;
    ray.starts_inside = inside != 0;
    const auto r = rainbow::detail::solve_surface<double>(s, ray);
    *dist = r.distance;
    return static_cast<int>(r.code);
}

// RaindropPathTracer::trace_vertex_unpolarized over a w x h vertex grid with the
// reference's RaindropIntersector (include/rainbow/raindrop_paths.hpp:40-107):
// per family (R, TT, TRT, TRRT; path-major, 4*w*h entries) the status, outgoing
// direction, surface point and the two unit-input response columns (s/p field
// components in the ray frame: x.re, x.im, y.re, y.im per column).
int ref_drop_trace_vertices(const float c[8], float inner, float outer, const float ki[3], const float basis_x[3],
                            float n_int, float half_extent, float reference_distance, uint32_t w, uint32_t h,
                            uint32_t* status, float* dir, float* pos, float* col0, float* col1)
{
    rainbow::RaindropTraceConfig cfg{};
    cfg.shape = drop_shape_of(c, inner, outer);
    cfg.incident_direction = {ki[0], ki[1], ki[2]};
    cfg.incident_basis_x = {basis_x[0], basis_x[1], basis_x[2]};
    cfg.radius_mm = 2.0f;
    cfg.wavelength_nm = 550.0f;
    cfg.exterior_index = 1.0f;
    cfg.interior_index = n_int;
    cfg.grid_half_extent = half_extent;
    cfg.reference_distance = cfg.outgoing_reference_distance = reference_distance;
    cfg.grid_width = w; cfg.grid_height = h;
    const rainbow::RaindropIntersector is{cfg.shape};
    std::vector<rainbow::OutgoingVertex> v(std::size_t(4) * w * h);
    std::vector<rainbow::Field32> second(v.size());
    for (uint32_t i = 0; i < w * h; ++i)
        rainbow::RaindropPathTracer::trace_vertex_unpolarized(cfg, i, is, v.data(), second.data());
    for (std::size_t i = 0; i < v.size(); ++i) {
        status[i] = static_cast<uint32_t>(v[i].status);
        dir[3 * i] = v[i].direction_drop.x; dir[3 * i + 1] = v[i].direction_drop.y; dir[3 * i + 2] = v[i].direction_drop.z;
        pos[3 * i] = v[i].position_drop.x; pos[3 * i + 1] = v[i].position_drop.y; pos[3 * i + 2] = v[i].position_drop.z;
        const rainbow::Field32 f0 = v[i].field, f1 = second[i];
        col0[4 * i] = f0.x.real; col0[4 * i + 1] = f0.x.imag; col0[4 * i + 2] = f0.y.real; col0[4 * i + 3] = f0.y.imag;
        col1[4 * i] = f1.x.real; col1[4 * i + 1] = f1.x.imag; col1[4 * i + 2] = f1.y.real; col1[4 * i + 3] = f1.y.imag;
    }
    return 0;
}

} // extern "C"
