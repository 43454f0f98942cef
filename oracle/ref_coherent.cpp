// Test infrastructure only: the reference's coherent drop pipeline, compiled
// from its own headers where they lie (oracle/Makefile), as the checker of the
// HIP coherent stages. Nothing here is shipped or called by the product.
//
// What is wrapped (all reference code, driven the way its own CPU tests drive it):
//   * the unpolarized vertex trace: RaindropPathTracer::trace_vertex_unpolarized
//     with the CPU RaindropIntersector (include/rainbow/raindrop_paths.hpp:24-37,
//     tests/unpolarized_test_support.hpp:66-78);
//   * the patch build: PatchConstruction::make per logical cell, compacted in
//     patch_id order, and FoldedPatchBuilder::prepare for NeedsRefinement cells
//     with the production focal layout (include/rainbow/patch_geometry.hpp:160-207,
//     include/rainbow/folded_patch_builder.hpp:186-214, src/patch_optics.cpp:264-291);
//   * the brute-force patch query of every query direction (count and fill
//     passes over every patch, PatchHitCollector::finish:
//     tests/unpolarized_test_support.hpp:141-161);
//   * PatchOpticalEvaluator::evaluate_direction with focal phase and the folded
//     view (include/rainbow/patch_optical_evaluator.hpp:140-300);
//   * RainbowDiffraction::detect / filter over the whole grid
//     (include/rainbow/rainbow_diffraction.hpp:42-139), Table II sigma.
// Grid directions come from QueryDirectionGrid (include/rainbow/query_direction_grid.hpp).
#include <rainbow/raindrop_paths.hpp>
#include <rainbow/raindrop_settings.hpp>
#include <rainbow/patch_optical_evaluator.hpp>
#include <rainbow/folded_patch_builder.hpp>
#include <rainbow/query_direction_grid.hpp>
#include <rainbow/rainbow_diffraction.hpp>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <exception>
#include <vector>

using namespace rainbow;

namespace {

struct Config {
    float radius_mm, wavelength_nm, interior_index, pad0;
    double inclination_deg;
    uint32_t grid, sphere, theta, phi;
    uint32_t focal_offsets[4];
    double primary_sigma_deg;     // 0: Table II
    double transition_contrast;
};

// One query direction's outputs (a flat record for ctypes).
struct Direction {
    double inc_s, inc_p, path_s, path_p, focal_s, focal_p, diff_s, diff_p;
    double family_incoherent[4];
    double sigma_rad, blend;
    double field[8];          // path field, column 0 then column 1 (s re, s im, p re, p im)
    double focal_field[8];
    uint32_t hit_count, evaluated_hits, rejected_hits, optical_flags;
    uint32_t query_flags, focal_flags, diffraction_flags, transition_flags;
    uint32_t family_hits[4];
    uint32_t transition_kind, folded_hits, regular_hits, refinement_hits;
};

RaindropSettings settings_of(const Config& c)
{
    RaindropSettings s;
    s.radius_mm = c.radius_mm;
    s.wavelength_nm = c.wavelength_nm;
    s.interior_index = c.interior_index;
    s.grid_width = s.grid_height = c.grid;
    s.force_sphere = c.sphere != 0;
    const double a = c.inclination_deg * QueryDirectionGrid::pi / 180.0;
    s.incident_direction = {static_cast<float>(std::cos(a)), static_cast<float>(-std::sin(a)), 0.0f};
    return s;
}

struct Built {
    RaindropTraceConfig config{};
    std::vector<OutgoingVertex> vertices;
    std::vector<Field32> second;
    std::vector<OutgoingPatch> patches;
    std::vector<uint32_t> lookup;
    std::vector<FoldedPatchRecord> folded;
    FocalPhaseConfig focal{};
    uint32_t status_counts[patch_status_count]{};
};

void trace_into(Built& b)
{
    const uint32_t n = b.config.vertex_count();
    b.vertices.assign(std::size_t(n) * 4, OutgoingVertex{});
    b.second.assign(b.vertices.size(), Field32{});
    RaindropIntersector intersector{b.config.shape};
    for (uint32_t i = 0; i < n; ++i)
        RaindropPathTracer::trace_vertex_unpolarized(b.config, i, intersector, b.vertices.data(), b.second.data());
}

bool build_patches(Built& b)
{
    PatchBuildLayout layout{};
    if (!PatchBuildLayout::try_make(b.config, layout)) return false;
    for (uint32_t id = 0; id < layout.cell_count; ++id) {
        const auto made = PatchConstruction::make(b.vertices.data(), layout, id);
        ++b.status_counts[static_cast<uint32_t>(made.patch.status)];
        if (has_patch_geometry(made.patch.status)) b.patches.push_back(made.patch);
    }
    b.lookup.assign(b.patches.size(), 0xffffffffu);
    for (uint32_t i = 0; i < b.patches.size(); ++i) {
        if (b.patches[i].status != PatchCellStatus::NeedsRefinement) continue;
        b.lookup[i] = static_cast<uint32_t>(b.folded.size());
        b.folded.push_back(FoldedPatchBuilder::prepare(b.vertices.data(), static_cast<uint32_t>(b.vertices.size()),
                                                       b.patches[i], i, b.config.incident_direction, &b.focal));
    }
    return true;
}

void query_one(const Built& b, const Vec3 direction, std::vector<PatchQueryHit>& hits, PatchQuerySummary& summary)
{
    PatchHitCollector counter{};
    auto geometry_of = [&](uint32_t i) {
        BilinearPatchGeometry g{};
        for (unsigned k = 0; k < 4; ++k) g.corners[k] = b.vertices[b.patches[i].vertex_indices[k]].direction_drop;
        return g;
    };
    for (uint32_t i = 0; i < b.patches.size(); ++i)
        counter.consume(b.patches[i], i, BilinearPatchIntersector::intersect(geometry_of(i), direction));
    hits.assign(counter.count, PatchQueryHit{});
    PatchHitCollector collector{};
    collector.storage = hits.data();
    collector.capacity = counter.count;
    collector.writing = true;
    for (uint32_t i = 0; i < b.patches.size(); ++i)
        collector.consume(b.patches[i], i, BilinearPatchIntersector::intersect(geometry_of(i), direction));
    summary = collector.finish(b.patches.data(), static_cast<uint32_t>(b.patches.size()));
}

int make(const Config* c, const void* vertices, const void* second, Built& b)
{
    b.config = settings_of(*c).make_config();
    b.focal = FocalPhaseConfig{b.config.grid_width, b.config.grid_height, b.config.grid_half_extent, 0,
                               {c->focal_offsets[0], c->focal_offsets[1], c->focal_offsets[2], c->focal_offsets[3]}};
    if (vertices && second) {
        const std::size_t n = std::size_t(b.config.vertex_count()) * 4;
        b.vertices.resize(n);
        b.second.resize(n);
        std::memcpy(b.vertices.data(), vertices, n * sizeof(OutgoingVertex));
        std::memcpy(b.second.data(), second, n * sizeof(Field32));
    } else {
        trace_into(b);
    }
    return build_patches(b) ? 0 : -2;
}

void put(double* out, const OpticalField64& f)
{
    out[0] = f.s_real;
    out[1] = f.s_imag;
    out[2] = f.p_real;
    out[3] = f.p_imag;
}

}  // namespace

extern "C" {

uint32_t refc_sizes(uint32_t which)
{
    switch (which) {
    case 0: return sizeof(OutgoingVertex);
    case 1: return sizeof(Field32);
    case 2: return sizeof(OutgoingPatch);
    case 3: return sizeof(FoldedPatchRecord);
    case 4: return sizeof(PatchQueryHit);
    case 5: return sizeof(Direction);
    case 6: return sizeof(Config);
    default: return 0;
    }
}

// The trace config the settings make: incident direction, basis x, grid half
// extent, reference distances, shape coefficients, outer radius (floats).
int refc_trace_config(const Config* c, float* out16)
{
    try {
        const auto t = settings_of(*c).make_config();
        const float v[16] = {t.incident_direction.x, t.incident_direction.y, t.incident_direction.z,
                             t.incident_basis_x.x, t.incident_basis_x.y, t.incident_basis_x.z,
                             t.grid_half_extent, t.reference_distance, t.outgoing_reference_distance,
                             t.shape.outer_radius, t.exterior_index, t.interior_index, t.radius_mm,
                             t.wavelength_nm, 0.0f, 0.0f};
        std::memcpy(out16, v, sizeof v);
        return 0;
    } catch (const std::exception&) {
        return -1;
    }
}

// 4 * grid^2 OutgoingVertex (path-major) and as many second-column fields.
int refc_trace(const Config* c, void* vertices, void* second)
{
    try {
        Built b;
        b.config = settings_of(*c).make_config();
        trace_into(b);
        std::memcpy(vertices, b.vertices.data(), b.vertices.size() * sizeof(OutgoingVertex));
        std::memcpy(second, b.second.data(), b.second.size() * sizeof(Field32));
        return 0;
    } catch (const std::exception&) {
        return -1;
    }
}

// Compact patches (patch_id order) and folded records built from the given
// vertices (or from the reference's own trace when vertices is null).
int refc_patches(const Config* c, const void* vertices, const void* second, void* patches, uint32_t patch_cap,
                 uint32_t* patch_count, void* folded, uint32_t folded_cap, uint32_t* folded_count,
                 uint32_t* status_counts)
{
    try {
        Built b;
        const int rc = make(c, vertices, second, b);
        if (rc) return rc;
        *patch_count = static_cast<uint32_t>(b.patches.size());
        *folded_count = static_cast<uint32_t>(b.folded.size());
        for (unsigned s = 0; s < patch_status_count; ++s) status_counts[s] = b.status_counts[s];
        if (b.patches.size() > patch_cap || b.folded.size() > folded_cap) return -3;
        std::memcpy(patches, b.patches.data(), b.patches.size() * sizeof(OutgoingPatch));
        std::memcpy(folded, b.folded.data(), b.folded.size() * sizeof(FoldedPatchRecord));
        return 0;
    } catch (const std::exception&) {
        return -1;
    }
}

// The hits of directions [first, first + count) of the query grid: hit lists
// concatenated (cap entries at most), per-direction hit counts and flags.
int refc_query(const Config* c, const void* vertices, const void* second, uint32_t first, uint32_t count,
               void* hits, uint64_t cap, uint32_t* counts, uint32_t* flags)
{
    try {
        Built b;
        const int rc = make(c, vertices, second, b);
        if (rc) return rc;
        const auto grid = QueryDirectionGrid::make(b.config, c->theta, c->phi);
        const auto dirs = grid.directions();
        uint64_t used = 0;
        std::vector<PatchQueryHit> h;
        for (uint32_t k = 0; k < count; ++k) {
            PatchQuerySummary s{};
            query_one(b, dirs[first + k], h, s);
            counts[k] = s.hit_count;
            flags[k] = s.flags;
            if (used + s.hit_count > cap) return -3;
            std::memcpy(static_cast<PatchQueryHit*>(hits) + used, h.data(), s.hit_count * sizeof(PatchQueryHit));
            used += s.hit_count;
        }
        return 0;
    } catch (const std::exception&) {
        return -1;
    }
}

// The whole grid: per-direction optics, focal phase, transitions and the
// Table II filter. vertices/second may be null (the reference traces them).
int refc_pipeline(const Config* c, const void* vertices, const void* second, Direction* out)
{
    try {
        Built b;
        const int rc = make(c, vertices, second, b);
        if (rc) return rc;
        const auto grid = QueryDirectionGrid::make(b.config, c->theta, c->phi);
        const auto dirs = grid.directions();
        const std::size_t n = dirs.size();
        std::vector<FocalOpticalResult> focal(n);
        const FoldedPatchView view{b.lookup.data(), b.folded.data(), static_cast<uint32_t>(b.lookup.size()),
                                   static_cast<uint32_t>(b.folded.size())};
        std::vector<PatchQueryHit> hits;
        for (std::size_t k = 0; k < n; ++k) {
            PatchQuerySummary summary{};
            query_one(b, dirs[k], hits, summary);
            const uint64_t offsets[2] = {0, hits.size()};
            PatchOpticsParams p{};
            p.vertices = b.vertices.data();
            p.vertex_count = static_cast<uint32_t>(b.vertices.size());
            p.patches = b.patches.data();
            p.patch_count = static_cast<uint32_t>(b.patches.size());
            p.directions = &dirs[k];
            p.direction_count = 1;
            p.offsets = offsets;
            p.hits = hits.data();
            p.summaries = &summary;
            p.hit_storage_count = hits.size();
            p.incident_direction = b.config.incident_direction;
            p.incident_basis_x = b.config.incident_basis_x;
            p.input_polarization = IncidentPolarization::Unpolarized;
            p.second_input_fields = b.second.data();
            p.second_input_count = p.vertex_count;
            const auto r = PatchOpticalEvaluator::evaluate_direction(p, 0, &b.focal, &focal[k], view);
            Direction& d = out[k];
            d = Direction{};
            d.inc_s = r.regular_partial_incoherent_s;
            d.inc_p = r.regular_partial_incoherent_p;
            d.path_s = r.regular_partial_path_s;
            d.path_p = r.regular_partial_path_p;
            put(d.field, r.regular_partial_path_field);
            put(d.field + 4, r.regular_partial_path_field_second);
            put(d.focal_field, focal[k].field);
            put(d.focal_field + 4, focal[k].field_second);
            d.focal_s = focal[k].intensity_s;
            d.focal_p = focal[k].intensity_p;
            for (unsigned f = 0; f < 4; ++f) {
                d.family_incoherent[f] = focal[k].family_incoherent[f];
                d.family_hits[f] = focal[k].family_hits[f];
            }
            d.hit_count = r.hit_count;
            d.evaluated_hits = r.evaluated_hits;
            d.rejected_hits = r.rejected_hits;
            d.optical_flags = r.flags;
            d.query_flags = summary.flags;
            d.focal_flags = focal[k].flags;
            d.folded_hits = r.folded_evaluated_hits();
            d.regular_hits = summary.regular_hits;
            d.refinement_hits = summary.refinement_hits;
        }
        DiffractionParams dp{};
        std::vector<RainbowTransition> transitions(n);
        std::vector<DiffractionResult> diffraction(n);
        dp.focal = focal.data();
        dp.transitions = transitions.data();
        dp.results = diffraction.data();
        dp.theta_count = c->theta;
        dp.phi_count = c->phi;
        double sigma = c->primary_sigma_deg;
        if (sigma == 0.0 && !RainbowDiffraction::table_sigma_degrees(c->radius_mm, sigma)) return -4;
        dp.config.primary_sigma_rad = sigma * RainbowDiffraction::pi / 180.0;
        dp.config.transition_contrast = c->transition_contrast;
        if (!RainbowDiffraction::valid_config(dp)) return -5;
        for (uint32_t i = 0; i < n; ++i) transitions[i] = RainbowDiffraction::detect(dp, i);
        for (uint32_t i = 0; i < n; ++i) diffraction[i] = RainbowDiffraction::filter(dp, i);
        for (std::size_t k = 0; k < n; ++k) {
            out[k].diff_s = diffraction[k].intensity_s;
            out[k].diff_p = diffraction[k].intensity_p;
            out[k].sigma_rad = diffraction[k].sigma_rad;
            out[k].blend = diffraction[k].blend;
            out[k].diffraction_flags = diffraction[k].flags;
            out[k].transition_kind = diffraction[k].transition_kind;
            out[k].transition_flags = transitions[k].flags;
        }
        return 0;
    } catch (const std::exception&) {
        return -1;
    }
}

// Table II sigma (degrees) for a radius; -1 outside 0.1..1 mm.
double refc_table_sigma(float radius_mm)
{
    double s = 0;
    return RainbowDiffraction::table_sigma_degrees(radius_mm, s) ? s : -1.0;
}

// The query grid's unit directions (float, theta-major) for a config.
int refc_directions(const Config* c, float* out)
{
    try {
        const auto t = settings_of(*c).make_config();
        const auto d = QueryDirectionGrid::make(t, c->theta, c->phi).directions();
        std::memcpy(out, d.data(), d.size() * sizeof(Vec3));
        return 0;
    } catch (const std::exception&) {
        return -1;
    }
}

}  // extern "C"
