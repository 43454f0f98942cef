// Host-side record preparation shared by the C-ABI runtime: validation,
// per-record device parameters, u edges, and the record files.
#pragma once

#include "kernels.hpp"
#include "materials.hpp"

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/icephase.h"

namespace icephase {

// rainbow::make_phase_u_edges (include/rainbow/phase_cdf_math.hpp:16-32):
// u_i = sin^2(pi min(i, n-i) / 2n), mirrored about 1/2, exact 0 and 1 ends.
bool make_u_edges(uint32_t n, std::vector<double>& out, std::string& err);
// The same edges into a caller array of n + 1 doubles (no allocation).
bool make_u_edges(uint32_t n, double* out, std::string& err);

// Drop contract frame: e0 = normalize(axis - k (k.axis)), axis = +z unless
// |k.z| >= 0.9 (then +x), e1 = k x e0 (include/rainbow/vec3.hpp:56-65, FP64).
bool drop_frame(d3 k, BinFrame& f);

// Certified strict convexity of the drop (materials.hpp drop_convexity; the
// reference's interval check is src/projected_area.cpp:50-87).
bool drop_convex(const double c[8], std::string& err);

// The tracer's view of a drop: coefficients (zero for a sphere), the
// conservative inner / outer radii, the intersector's bounding radius and the
// number of Chebyshev terms to evaluate. False (err set) for a non-convex shape.
bool drop_shape_bounds(const float coef[8], bool sphere, drop::Shape& shape, std::string& err);

// Proper rotation check of a row-major body -> ensemble matrix.
bool rotation_valid(const double m[9]);

// Fills everything but the device pointers. tile_rays: requested tile length.
bool prepare_record(const iph_record& in, uint64_t tile_rays, DevRecord& out, double& ray_weight_mm2,
                    double& projected_area_mm2, std::string& err);

// NPY v1.0 '<f8' C-order writer (src/npy_writer.cpp:18-95).
bool write_npy(const std::string& path, const std::vector<uint64_t>& shape, const double* data, std::string& err);

// <out>.part staging + rename (src/npy_writer.cpp:97-121).
int write_record_dir(const std::string& out_dir, const std::string& metadata, const double* phi, const double* theta,
                     const std::vector<double>& edges, uint32_t nt, uint32_t np, std::string& err);

}  // namespace icephase
