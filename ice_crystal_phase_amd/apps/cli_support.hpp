// Support code of the command-line tools (rainbow_generate, ice_generate_phase):
// strict option parsing, an ordered JSON writer, orientation plans and SHA-256.
// The tools call the device path only through the C-ABI (include/icephase.h).
#pragma once

#include <cstdint>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/icephase.h"

namespace cli {

// std::from_chars-strict numbers: the whole token, finite (reference:
// apps/generate_phase.cpp:57-72, hex_phase/apps/generate_hex_phase.cpp:21-26).
double number(const std::string& s);
uint32_t positive_u32(const std::string& s);
uint64_t u64(const std::string& s);
int integer(const std::string& s);

// Ordered JSON object writer, 17 significant digits, classic locale.
class Json {
public:
    Json& key(const std::string& k);
    Json& value(double v);
    Json& value(uint64_t v);
    Json& value(int64_t v);
    Json& value(uint32_t v) { return value((uint64_t)v); }
    Json& value(int v) { return value((int64_t)v); }
    Json& value(bool v);
    Json& value(const char* v) { return str(v); }
    Json& value(const std::string& v) { return str(v); }
    Json& null();
    Json& str(const std::string& v);
    Json& array(const double* v, size_t n);
    Json& array(const std::vector<std::string>& v);
    Json& open_object();
    Json& close_object();
    Json& open_array();
    Json& close_array();
    template <class T>
    Json& kv(const std::string& k, const T& v)
    {
        key(k);
        return value(v);
    }
    // the finished document (closes the top-level object)
    std::string text()
    {
        close_object();
        return out_.str() + "\n";
    }
    Json();

private:
    void sep();
    void raw(const std::string& v);
    std::ostringstream out_;
    std::vector<bool> first_;
    std::vector<bool> is_array_;
    bool after_key_ = false;
    int depth_ = 0;
};

struct OrientationNode {
    double q[4];         // w, x, y, z
    double rotation[9];  // row-major body -> ensemble
    double weight;       // renormalized number probability
    uint64_t samples, seed;
};
struct OrientationPlan {
    std::vector<OrientationNode> nodes;
    uint64_t total_samples = 0;
    double input_weight_sum = 0;
    std::string sha256;
};
// qw,qx,qy,qz,weight,samples,seed CSV (reference:
// hex_phase/src/orientation_plan.cpp:68-117): header line, '#' comments,
// unit quaternions (2e-12), positive weights summing to one within 2e-12
// (renormalized), positive sample counts. Throws std::invalid_argument.
OrientationPlan parse_orientations(const std::string& csv);
OrientationPlan read_orientations(const std::string& path);
// Active Hamilton quaternion -> rotation (hex_phase/src/phase_table.cpp:11-21).
void rotation_of_quaternion(double w, double x, double y, double z, double R[9]);
// FIPS 180-4 SHA-256, lowercase hex.
std::string sha256_hex(const std::string& bytes);

// Throws std::runtime_error with iph_last_error() when rc != IPH_OK.
void check(int rc, const iph_ctx* ctx = nullptr);

}  // namespace cli
