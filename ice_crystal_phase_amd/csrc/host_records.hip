// Host-side record preparation (see host_records.hpp).
#include "host_records.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <sstream>

namespace icephase {

bool make_u_edges(uint32_t n, std::vector<double>& e, std::string& err)
{
    if (n == 0) {
        err = "theta_count must be positive.";
        return false;
    }
    e.assign((size_t)n + 1, 0.0);
    e.front() = 0;
    e.back() = 1;
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
    const int n = 16384;
    double min_r = 1e300, min_k = 1e300;
    for (int s = 0; s <= n; ++s) {
        const double chi = kPi * s / n;
        double r = 1 + c[0], r1 = 0, r2 = 0;
        for (int m = 1; m < 8; ++m) {
            r += c[m] * std::cos(m * chi);
            r1 -= m * c[m] * std::sin(m * chi);
            r2 -= m * m * c[m] * std::cos(m * chi);
        }
        min_r = std::fmin(min_r, r);
        min_k = std::fmin(min_k, r * r + 2 * r1 * r1 - r * r2);
    }
    if (!(min_r > 0)) {
        err = "Drop radius r(mu) must stay positive.";
        return false;
    }
    if (!(min_k > 1e-9)) {
        err = "Drop shape is not convex (meridian curvature changes sign); the analytic convex intersection "
              "requires a convex particle.";
        return false;
    }
    return true;
}

static bool finite3(const double* v)
{
    return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]);
}

bool prepare_record(const iph_record& in, uint64_t tile_rays, DevRecord& R, double& weight, double& area,
                    std::string& err)
{
    std::memset(&R, 0, sizeof R);
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
        if (!unit_of(k_in, c.k) || !hex::frame_of(c.k, c.f)) {
            err = "Incident direction must be finite and nonzero.";
            return false;
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
        R.frame.ki = c.k;
        R.frame.e0 = c.f.e0;
        R.frame.e1 = c.f.e1;
        weight = c.area / double(in.rays);
        area = c.area;
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
        c.shape.sphere = sphere ? 1u : 0u;
        double s = 0;
        for (int m = 0; m < 8; ++m) {
            if (!std::isfinite(in.shape_coefficients[m])) {
                err = "Shape coefficients must be finite.";
                return false;
            }
            c.shape.c[m] = sphere ? 0.0f : (float)in.shape_coefficients[m];
        }
        if (!sphere) {
            double cd[8];
            for (int m = 0; m < 8; ++m) cd[m] = (double)c.shape.c[m];
            if (!drop_convex(cd, err)) return false;
        }
        for (int m = 1; m < 8; ++m) s += std::fabs((double)c.shape.c[m]);
        const double center = 1.0 + (double)c.shape.c[0];
        c.shape.inner = std::nextafter((float)(center - s - 0x1p-20), -INFINITY);
        c.shape.outer = std::nextafter((float)(center + s + 0x1p-20), INFINITY);
        if (sphere) {
            c.shape.inner = 1.0f;
            c.shape.outer = 1.0f;
        }
        // bounding radius of the intersector: max over chi of r(chi) on a dense grid,
        // + 1e-6, rounded up to a multiple of 2^-20 (exact in single precision)
        double rm = 0;
        for (int k = 0; k <= 16384; ++k) {
            const double chi = kPi * k / 16384;
            double rr = 1 + (double)c.shape.c[0];
            for (int m = 1; m < 8; ++m) rr += (double)c.shape.c[m] * std::cos(m * chi);
            if (rr > rm) rm = rr;
        }
        c.shape.rmax = sphere ? 1.0f : (float)(std::ceil((rm + 1e-6) * 1048576.0) / 1048576.0);
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
        area = sphere ? kPi * in.radius_mm * in.radius_mm : 0.0;
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
