// Host-side record preparation (see host_records.hpp).
#include "host_records.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <sstream>
#include <algorithm>
#include <cfloat>
#include <utility>

namespace icephase {

bool make_u_edges(uint32_t n, std::vector<double>& e, std::string& err)
{
    if (n == 0) {
        err = "theta_count must be positive.";
        return false;
    }
    e.assign((size_t)n + 1, 0.0);
    return make_u_edges(n, e.data(), err);
}

bool make_u_edges(uint32_t n, double* e, std::string& err)
{
    if (n == 0) {
        err = "theta_count must be positive.";
        return false;
    }
    e[0] = 0;
    e[n] = 1;
    for (uint32_t i = 1; i < n; ++i) {
        const uint32_t k = i < n - i ? i : n - i;
        const double s = std::sin((0.5 * kPi) * double(k) / double(n));
        e[i] = i <= n / 2u ? s * s : 1.0 - s * s;
    }
    for (uint32_t i = 0; i < n; ++i)
        if (!(e[i + 1] > e[i])) {
            err = "Angular cells cannot be represented in binary64.";
            return false;
        }
    return true;
}

bool drop_frame(d3 k_in, BinFrame& f)
{
    d3 k;
    if (!unit_of(k_in, k)) return false;
    const d3 axis = std::fabs(k.z) < 0.9 ? D3(0, 0, 1) : D3(1, 0, 0);
    d3 e0, e1;
    if (!unit_of(axis - k * dot(k, axis), e0)) return false;
    if (!unit_of(cross(k, e0), e1)) return false;
    f.ki = k;
    f.e0 = e0;
    f.e1 = e1;
    return true;
}

bool drop_convex(const double c[8], std::string& err)
{
    uint32_t cells = 0;
    return drop_convexity(c, cells, err);
}

bool rotation_valid(const double m[9])
{
    // columns are the images of the body axes (Rotation::valid,
    // hex_phase/include/ice_crystal/phase_math.hpp:12-18)
    const d3 x = D3(m[0], m[3], m[6]), y = D3(m[1], m[4], m[7]), z = D3(m[2], m[5], m[8]);
    return is_unit64(x) && is_unit64(y) && is_unit64(z) && std::fabs(dot(x, y)) <= 128 * DBL_EPSILON &&
           max_abs(cross(x, y) - z) <= 256 * DBL_EPSILON;
}

// Aperture polygon of the prism seen along c.k, in the trace basis (e0, e1) of
// c.k, and the lobe's sampling parameters.
static bool setup_lobe(const hex::Config& c, DevRecord& R, double wavelength_nm, std::string& err)
{
    if (!(std::isfinite(wavelength_nm) && wavelength_nm > 0)) {
        err = "Diffraction needs a finite positive wavelength.";
        return false;
    }
    Lobe& L = R.lobe;
    std::vector<std::pair<double, double>> pts;
    for (int f = 0; f < 6; ++f)
        for (int s = -1; s <= 1; s += 2) {
            const d3 v = hex::Prism::vertex(f);
            const d3 p = D3(c.prism.R * v.x, 0.5 * s * c.prism.L, c.prism.R * v.z);
            pts.emplace_back(dot(p, c.f.e0), dot(p, c.f.e1));
        }
    // convex hull (monotone chain), counter-clockwise, collinear points dropped
    std::sort(pts.begin(), pts.end());
    const double scale = c.prism.R > c.prism.L ? c.prism.R : c.prism.L;
    auto turn = [&](const std::pair<double, double>& o, const std::pair<double, double>& a,
                    const std::pair<double, double>& b) {
        return (a.first - o.first) * (b.second - o.second) - (a.second - o.second) * (b.first - o.first);
    };
    std::vector<std::pair<double, double>> hull(2 * pts.size());
    size_t k = 0;
    for (size_t i = 0; i < pts.size(); ++i) {
        while (k >= 2 && turn(hull[k - 2], hull[k - 1], pts[i]) <= 1e-12 * scale * scale) --k;
        hull[k++] = pts[i];
    }
    for (size_t i = pts.size() - 1, t = k + 1; i > 0; --i) {
        while (k >= t && turn(hull[k - 2], hull[k - 1], pts[i - 1]) <= 1e-12 * scale * scale) --k;
        hull[k++] = pts[i - 1];
    }
    hull.resize(k > 1 ? k - 1 : k);
    if (hull.size() < 3 || hull.size() > 8) {
        err = "Projected aperture polygon is degenerate.";
        return false;
    }
    L.nv = (uint32_t)hull.size();
    double a2 = 0, dmax = 0;
    for (uint32_t i = 0; i < L.nv; ++i) {
        L.vx[i] = hull[i].first;
        L.vy[i] = hull[i].second;
        const auto& q = hull[(i + 1) % L.nv];
        a2 += hull[i].first * q.second - q.first * hull[i].second;
        for (uint32_t j = 0; j < L.nv; ++j)
            dmax = std::fmax(dmax, std::hypot(hull[i].first - hull[j].first, hull[i].second - hull[j].second));
    }
    L.area = 0.5 * a2;
    L.dmax = dmax;
    if (!(L.area > 0) || std::fabs(L.area / c.area - 1) > 1e-9) {
        err = "Projected aperture polygon disagrees with the face-sum projected area.";
        return false;
    }
    const double lambda = wavelength_nm * 1e-6;  // mm
    L.k0 = 2 * kPi / lambda;
    L.inv_lambda2 = 1 / (lambda * lambda);
    // aperture basis in binning-frame (world) coordinates
    if (R.rotated) {
        const double* m = R.rot;
        auto rot = [&](d3 v) {
            return D3((m[0] * v.x + m[1] * v.y) + m[2] * v.z, (m[3] * v.x + m[4] * v.y) + m[5] * v.z,
                      (m[6] * v.x + m[7] * v.y) + m[8] * v.z);
        };
        L.p1 = rot(c.f.e0);
        L.p2 = rot(c.f.e1);
    } else {
        L.p1 = c.f.e0;
        L.p2 = c.f.e1;
    }
    // rows within min(0.5 rad, 200 fringes) of the forward direction
    const double cut = std::fmin(0.5, 200 * lambda / dmax);
    uint32_t rows = (uint32_t)std::ceil(cut / (kPi / R.nt));
    rows = rows < 1 ? 1 : (rows > R.nt ? R.nt : rows);
    L.rows = rows;
    R.lobe_on = 1;
    return true;
}

static bool finite3(const double* v)
{
    return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]);
}

bool drop_shape_bounds(const float coef[8], bool sphere, drop::Shape& shape, std::string& err)
{
    shape.sphere = sphere ? 1u : 0u;
    for (int m = 0; m < 8; ++m) shape.c[m] = sphere ? 0.0f : coef[m];
    if (!sphere) {
        double cd[8];
        for (int m = 0; m < 8; ++m) cd[m] = (double)shape.c[m];
        if (!drop_convex(cd, err)) return false;
    }
    double s = 0;
    for (int m = 1; m < 8; ++m) s += std::fabs((double)shape.c[m]);
    const double center = 1.0 + (double)shape.c[0];
    shape.inner = std::nextafter((float)(center - s - 0x1p-20), -INFINITY);
    shape.outer = std::nextafter((float)(center + s + 0x1p-20), INFINITY);
    if (sphere) {
        shape.inner = 1.0f;
        shape.outer = 1.0f;
    }
    // bounding radius of the intersector: max over chi of r(chi) on a dense grid,
    // + 1e-6, rounded up to a multiple of 2^-20 (exact in single precision)
    double rm = 0;
    for (int k = 0; k <= 16384; ++k) {
        const double chi = kPi * k / 16384;
        double rr = 1 + (double)shape.c[0];
        for (int m = 1; m < 8; ++m) rr += (double)shape.c[m] * std::cos(m * chi);
        if (rr > rm) rm = rr;
    }
    shape.rmax = sphere ? 1.0f : (float)(std::ceil((rm + 1e-6) * 1048576.0) / 1048576.0);
    shape.terms = 2;
    for (uint32_t m = 2; m < 8; ++m)
        if (shape.c[m] != 0.0f) shape.terms = m + 1;
    return true;
}

bool prepare_record(const iph_record& in, uint64_t tile_rays, DevRecord& R, double& weight, double& area,
                    std::string& err)
{
    std::memset(&R, 0, sizeof R);
    R.qfactor = 1.0;
    R.owns = in.histogram_of == 0 ? 1u : 0u;
    if (in.flags & ~(uint32_t)IPH_REC_DIFFRACTION) {
        err = "Unknown record flags.";
        return false;
    }
    {
        bool any = false;
        for (int k = 0; k < 9; ++k) {
            if (!std::isfinite(in.rotation[k])) {
                err = "Rotation must be finite.";
                return false;
            }
            any = any || in.rotation[k] != 0;
        }
        if (any) {
            if (in.shape != IPH_SHAPE_HEX) {
                err = "Orientation rotations apply to hex records only.";
                return false;
            }
            if (!rotation_valid(in.rotation)) {
                err = "Rotation is not orthonormal and right-handed.";
                return false;
            }
            std::memcpy(R.rot, in.rotation, sizeof R.rot);
            R.rotated = 1;
        }
    }
    if (in.shape != IPH_SHAPE_HEX && (in.histogram_of != 0 || in.weight != 0 || in.flags != 0)) {
        err = "Orientation nodes, node weights and diffraction apply to hex records only.";
        return false;
    }
    if (in.theta_count == 0 || in.phi_count == 0 || (uint64_t)in.theta_count * in.phi_count > 0xffffffffull ||
        in.phi_count > 0x7fffffffu) {
        err = "Unsupported phase grid dimensions.";
        return false;
    }
    if (in.max_internal_hits == 0) {
        err = "max_internal_hits must be positive.";
        return false;
    }
    if (!(std::isfinite(in.interior_index) && in.interior_index > 0 && std::isfinite(in.exterior_index) &&
          in.exterior_index > 0)) {
        err = "Refractive indices must be finite and positive.";
        return false;
    }
    if (!(std::isfinite(in.tail_tolerance) && in.tail_tolerance >= 0 && in.tail_tolerance < 1)) {
        err = "tail_tolerance must lie in [0,1).";
        return false;
    }
    if (!finite3(in.incident)) {
        err = "Incident direction must be finite and nonzero.";
        return false;
    }
    R.nt = in.theta_count;
    R.np = in.phi_count;
    const d3 k_in = D3(in.incident[0], in.incident[1], in.incident[2]);
    if (in.shape == IPH_SHAPE_HEX) {
        R.kind = kKindHex;
        if (in.rays == 0) {
            err = "Hex record needs a positive ray count.";
            return false;
        }
        R.rays = in.rays;
        hex::Config& c = R.hex;
        c.prism.R = in.radius_mm;
        c.prism.L = in.length_mm;
        if (!c.prism.valid()) {
            err = "Invalid hex prism geometry.";
            return false;
        }
        // the splitting chain skips the unit check of the (negated) face normals:
        // they hold for all eight faces, which this confirms with the same operations
        for (unsigned f = 0; f < 8; ++f)
            if (!is_unit64(hex::Prism::normal(f)) || !is_unit64(-hex::Prism::normal(f))) {
                err = "Internal error: a prism face normal fails its unit check.";
                return false;
            }
        d3 k_ens;
        if (!unit_of(k_in, k_ens)) {
            err = "Incident direction must be finite and nonzero.";
            return false;
        }
        if (R.rotated) {
            // trace in the body frame: normalize(R^T k)
            // (hex_phase/src/phase_generate.cpp:44-49, phase_math.hpp:11)
            const double* m = R.rot;
            const d3 kb = D3((m[0] * k_ens.x + m[3] * k_ens.y) + m[6] * k_ens.z,
                             (m[1] * k_ens.x + m[4] * k_ens.y) + m[7] * k_ens.z,
                             (m[2] * k_ens.x + m[5] * k_ens.y) + m[8] * k_ens.z);
            if (!unit_of(kb, c.k) || !hex::frame_of(c.k, c.f)) {
                err = "Incident direction must be finite and nonzero.";
                return false;
            }
        } else {
            c.k = k_ens;
            if (!hex::frame_of(c.k, c.f)) {
                err = "Incident direction must be finite and nonzero.";
                return false;
            }
        }
        c.ni = in.exterior_index;
        c.nt = in.interior_index;
        c.tail = in.tail_tolerance;
        c.seed = in.seed;
        c.max_hits = in.max_internal_hits;
        c.area = c.prism.projected_area(c.k);
        if (!(c.area > 0) || !std::isfinite(c.area) || !(c.area / double(in.rays) > 0)) {
            err = "Per-incident-sample area underflows FP64.";
            return false;
        }
        if (R.rotated) {
            // bin in the ensemble frame (make_frame of the ensemble direction)
            hex::Basis fe;
            if (!hex::frame_of(k_ens, fe)) {
                err = "Incident direction must be finite and nonzero.";
                return false;
            }
            R.frame.ki = k_ens;
            R.frame.e0 = fe.e0;
            R.frame.e1 = fe.e1;
        } else {
            R.frame.ki = c.k;
            R.frame.e0 = c.f.e0;
            R.frame.e1 = c.f.e1;
        }
        const double node_weight = in.weight == 0 ? 1.0 : in.weight;
        if (!(std::isfinite(node_weight) && node_weight > 0)) {
            err = "Orientation weight must be finite and positive.";
            return false;
        }
        weight = node_weight * c.area / double(in.rays);
        area = node_weight * c.area;
        if (!(weight > 0) || !std::isfinite(weight)) {
            err = "Orientation area weight over/underflow.";
            return false;
        }
        if (in.flags & IPH_REC_DIFFRACTION) {
            if (!setup_lobe(c, R, in.wavelength_nm, err)) return false;
        }
    } else if (in.shape == IPH_SHAPE_SPHERE || in.shape == IPH_SHAPE_DROP) {
        R.kind = kKindDrop;
        if (in.grid_w == 0 || in.grid_h == 0 || (uint64_t)in.grid_w * in.grid_h > (1ull << 40) ||
            in.grid_w > (1u << 23) || in.grid_h > (1u << 23)) {
            err = "Invalid drop incident grid.";
            return false;
        }
        const uint64_t n = (uint64_t)in.grid_w * in.grid_h;
        if (in.rays != 0 && in.rays != n) {
            err = "Drop record: rays must be 0 or grid_w*grid_h.";
            return false;
        }
        if (!(std::isfinite(in.radius_mm) && in.radius_mm > 0)) {
            err = "radius_mm must be finite and positive.";
            return false;
        }
        R.rays = n;
        drop::Config& c = R.drop;
        const bool sphere = in.shape == IPH_SHAPE_SPHERE;
        float coef[8];
        for (int m = 0; m < 8; ++m) {
            if (!std::isfinite(in.shape_coefficients[m])) {
                err = "Shape coefficients must be finite.";
                return false;
            }
            coef[m] = (float)in.shape_coefficients[m];
        }
        if (!drop_shape_bounds(coef, sphere, c.shape, err)) return false;
        BinFrame f;
        if (!drop_frame(k_in, f)) {
            err = "Incident direction must be finite and nonzero.";
            return false;
        }
        c.ki = drop::f3{(float)f.ki.x, (float)f.ki.y, (float)f.ki.z};
        c.e0 = drop::f3{(float)f.e0.x, (float)f.e0.y, (float)f.e0.z};
        c.e1 = drop::f3{(float)f.e1.x, (float)f.e1.y, (float)f.e1.z};
        c.n_int = (float)in.interior_index;
        c.n_ext = (float)in.exterior_index;
        c.h = (float)((double)c.shape.outer * 1.01);
        c.tail = (float)in.tail_tolerance;
        c.w = in.grid_w;
        c.hgt = in.grid_h;
        c.max_hits = in.max_internal_hits;
        // binning frame: the contract frame of the traced (single-precision) direction
        if (!drop_frame(D3((double)c.ki.x, (double)c.ki.y, (double)c.ki.z), R.frame)) {
            err = "Incident direction must be finite and nonzero.";
            return false;
        }
        const double cell = (2.0 * (double)c.h / in.grid_w) * (2.0 * (double)c.h / in.grid_h);
        weight = cell * in.radius_mm * in.radius_mm;
        {
            // projected area of the traced shape along the traced direction
            double cd[8];
            for (int m = 0; m < 8; ++m) cd[m] = (double)c.shape.c[m];
            const double kd[3] = {(double)c.ki.x, (double)c.ki.y, (double)c.ki.z};
            AreaResult ar;
            if (!drop_projected_area(cd, in.radius_mm, kd, ar, err)) return false;
            area = ar.area_mm2;
        }
    } else {
        err = "Unknown shape.";
        return false;
    }
    R.qbits = quantum_bits(R.rays);
    R.tile_rays = tile_rays ? tile_rays : 4096;
    return true;
}

bool write_npy(const std::string& path, const std::vector<uint64_t>& shape, const double* data, std::string& err)
{
    uint64_t count = 1;
    std::ostringstream dims;
    dims << '(';
    for (uint64_t n : shape) {
        count *= n;
        dims << n << ", ";
    }
    dims << ')';
    std::string header = "{'descr': '<f8', 'fortran_order': False, 'shape': " + dims.str() + ", }";
    const size_t pad = (64u - ((10u + header.size() + 1u) % 64u)) % 64u;
    header.append(pad, ' ');
    header.push_back('\n');
    std::ofstream out(path, std::ios::binary | std::ios::trunc);
    if (!out) {
        err = "NPY: cannot open output file.";
        return false;
    }
    const unsigned char prefix[10] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0, (unsigned char)(header.size() & 255u),
                                      (unsigned char)((header.size() >> 8) & 255u)};
    out.write((const char*)prefix, 10);
    out.write(header.data(), (std::streamsize)header.size());
    out.write((const char*)data, (std::streamsize)(count * sizeof(double)));  // little-endian host
    out.flush();
    if (!out) {
        err = "NPY: write failed (check free disk space).";
        return false;
    }
    out.close();
    return !out.fail();
}

int write_record_dir(const std::string& out_dir, const std::string& metadata, const double* phi, const double* theta,
                     const std::vector<double>& edges, uint32_t nt, uint32_t np, std::string& err)
{
    namespace fs = std::filesystem;
    const fs::path dest(out_dir);
    if (dest.empty() || dest.filename().empty() || dest.filename() == "." || dest.filename() == ".." ||
        dest.extension() == ".part") {
        err = "Invalid dataset destination.";
        return IPH_E_INVALID;
    }
    fs::path staging = dest;
    staging += ".part";
    std::error_code ec;
    if (fs::exists(dest, ec)) {
        err = "Dataset already exists; refusing to overwrite.";
        return IPH_E_EXISTS;
    }
    if (!dest.parent_path().empty()) fs::create_directories(dest.parent_path(), ec);
    if (!fs::create_directory(staging, ec)) {
        err = "Staging directory exists; inspect/retire the previous .part directory.";
        return IPH_E_EXISTS;
    }
    if (!write_npy((staging / "phi_cdf.npy").string(), {(uint64_t)np + 1}, phi, err) ||
        !write_npy((staging / "theta_given_phi_cdf.npy").string(), {(uint64_t)np, (uint64_t)nt + 1}, theta, err) ||
        !write_npy((staging / "u_edges.npy").string(), {(uint64_t)nt + 1}, edges.data(), err))
        return IPH_E_IO;
    {
        std::ofstream m(staging / "metadata.json", std::ios::binary | std::ios::trunc);
        m << metadata;
        m.flush();
        if (!m) {
            err = "Metadata write failed.";
            return IPH_E_IO;
        }
    }
    if (fs::exists(dest, ec)) {
        err = "Destination appeared while writing; refusing to replace it.";
        return IPH_E_EXISTS;
    }
    fs::rename(staging, dest, ec);
    if (ec) {
        err = "Rename of the staging directory failed: " + ec.message();
        return IPH_E_IO;
    }
    return IPH_OK;
}

}  // namespace icephase
