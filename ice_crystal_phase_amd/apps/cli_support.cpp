// Support code of the command-line tools (see cli_support.hpp).
#include "cli_support.hpp"

#include <charconv>
#include <cmath>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <locale>

namespace cli {

double number(const std::string& s)
{
    double x = 0;
    const auto r = std::from_chars(s.data(), s.data() + s.size(), x);
    if (s.empty() || r.ec != std::errc{} || r.ptr != s.data() + s.size() || !std::isfinite(x))
        throw std::invalid_argument("Invalid numeric option.");
    return x;
}

uint32_t positive_u32(const std::string& s)
{
    uint32_t x = 0;
    const auto r = std::from_chars(s.data(), s.data() + s.size(), x);
    if (s.empty() || r.ec != std::errc{} || r.ptr != s.data() + s.size() || x == 0)
        throw std::invalid_argument("Invalid positive integer option.");
    return x;
}

uint64_t u64(const std::string& s)
{
    uint64_t x = 0;
    const auto r = std::from_chars(s.data(), s.data() + s.size(), x);
    if (s.empty() || r.ec != std::errc{} || r.ptr != s.data() + s.size())
        throw std::invalid_argument("Malformed numeric option.");
    return x;
}

int integer(const std::string& s)
{
    int x = 0;
    const auto r = std::from_chars(s.data(), s.data() + s.size(), x);
    if (s.empty() || r.ec != std::errc{} || r.ptr != s.data() + s.size())
        throw std::invalid_argument("Malformed numeric option.");
    return x;
}

void check(int rc, const iph_ctx* ctx)
{
    if (rc != IPH_OK) throw std::runtime_error(iph_last_error(ctx));
}

// ---------------------------------------------------------------- JSON
Json::Json()
{
    out_.imbue(std::locale::classic());
    out_ << std::setprecision(17);
    out_ << '{';
    first_.push_back(true);
    is_array_.push_back(false);
    depth_ = 1;
}

void Json::sep()
{
    if (after_key_) {
        after_key_ = false;
        return;
    }
    if (!first_.back()) out_ << ',';
    first_.back() = false;
    if (!is_array_.back()) {
        out_ << '\n';
        for (int i = 0; i < depth_; ++i) out_ << "  ";
    }
}

Json& Json::key(const std::string& k)
{
    sep();
    raw(k);
    out_ << ": ";
    after_key_ = true;
    return *this;
}

Json& Json::value(double v)
{
    sep();
    if (!std::isfinite(v))
        out_ << "null";
    else
        out_ << v;
    return *this;
}
Json& Json::value(uint64_t v)
{
    sep();
    out_ << v;
    return *this;
}
Json& Json::value(int64_t v)
{
    sep();
    out_ << v;
    return *this;
}
Json& Json::value(bool v)
{
    sep();
    out_ << (v ? "true" : "false");
    return *this;
}
Json& Json::null()
{
    sep();
    out_ << "null";
    return *this;
}
Json& Json::str(const std::string& v)
{
    sep();
    raw(v);
    return *this;
}

void Json::raw(const std::string& v)
{
    static const char* hex = "0123456789abcdef";
    std::string o = "\"";
    for (unsigned char c : v) {
        if (c == '"' || c == '\\') {
            o += '\\';
            o += (char)c;
        } else if (c < 32) {
            o += "\\u00";
            o += hex[c >> 4];
            o += hex[c & 15];
        } else {
            o += (char)c;
        }
    }
    o += '"';
    out_ << o;
}
Json& Json::array(const double* v, size_t n)
{
    open_array();
    for (size_t i = 0; i < n; ++i) value(v[i]);
    return close_array();
}
Json& Json::array(const std::vector<std::string>& v)
{
    open_array();
    for (const auto& s : v) str(s);
    return close_array();
}
Json& Json::open_object()
{
    sep();
    out_ << '{';
    first_.push_back(true);
    is_array_.push_back(false);
    ++depth_;
    return *this;
}
Json& Json::close_object()
{
    --depth_;
    const bool empty = first_.back();
    first_.pop_back();
    is_array_.pop_back();
    if (!empty) {
        out_ << '\n';
        for (int i = 0; i < depth_; ++i) out_ << "  ";
    }
    out_ << '}';
    return *this;
}
Json& Json::open_array()
{
    sep();
    out_ << '[';
    first_.push_back(true);
    is_array_.push_back(true);
    ++depth_;
    return *this;
}
Json& Json::close_array()
{
    --depth_;
    first_.pop_back();
    is_array_.pop_back();
    out_ << ']';
    return *this;
}

// ---------------------------------------------------------------- orientation plans
void rotation_of_quaternion(double w, double x, double y, double z, double R[9])
{
    const double n = w * w + x * x + y * y + z * z;
    if (!std::isfinite(n) || std::fabs(n - 1) > 2e-12) {
        throw std::invalid_argument("Quaternion must be unit length (w,x,y,z).");
    }
    const double s = 1 / std::sqrt(n);
    w *= s, x *= s, y *= s, z *= s;
    // columns: images of the body axes
    const double cx[3] = {1 - 2 * (y * y + z * z), 2 * (x * y + w * z), 2 * (x * z - w * y)};
    const double cy[3] = {2 * (x * y - w * z), 1 - 2 * (x * x + z * z), 2 * (y * z + w * x)};
    const double cz[3] = {2 * (x * z + w * y), 2 * (y * z - w * x), 1 - 2 * (x * x + y * y)};
    for (int r = 0; r < 3; ++r) {
        R[3 * r + 0] = cx[r];
        R[3 * r + 1] = cy[r];
        R[3 * r + 2] = cz[r];
    }
    auto dot = [](const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
    const double eps = 2.220446049250313e-16;
    bool ok = std::fabs(dot(cx, cx) - 1) <= 128 * eps && std::fabs(dot(cy, cy) - 1) <= 128 * eps &&
              std::fabs(dot(cz, cz) - 1) <= 128 * eps && std::fabs(dot(cx, cy)) <= 128 * eps;
    const double cr[3] = {cx[1] * cy[2] - cx[2] * cy[1], cx[2] * cy[0] - cx[0] * cy[2], cx[0] * cy[1] - cx[1] * cy[0]};
    for (int i = 0; i < 3; ++i) ok = ok && std::fabs(cr[i] - cz[i]) <= 256 * eps;
    if (!ok) throw std::invalid_argument("Quaternion rotation failed orthonormality check.");
}

namespace {
std::string trim(const std::string& s)
{
    const size_t a = s.find_first_not_of(" \t\r");
    if (a == std::string::npos) return {};
    return s.substr(a, s.find_last_not_of(" \t\r") - a + 1);
}
template <class T>
bool parse_number(const std::string& field, T& v)
{
    const std::string s = trim(field);
    if (s.empty()) return false;
    const auto r = std::from_chars(s.data(), s.data() + s.size(), v);
    return r.ec == std::errc{} && r.ptr == s.data() + s.size();
}
}  // namespace

OrientationPlan parse_orientations(const std::string& csv)
{
    OrientationPlan plan;
    plan.sha256 = sha256_hex(csv);
    bool header = false;
    double wsum = 0, wcomp = 0;
    size_t start = 0, line_no = 0;
    while (start < csv.size()) {
        size_t end = csv.find('\n', start);
        if (end == std::string::npos) end = csv.size();
        const std::string line = trim(csv.substr(start, end - start));
        start = end + 1;
        ++line_no;
        if (line.empty() || line[0] == '#') continue;
        auto fail = [&](const std::string& why) -> OrientationPlan {
            throw std::invalid_argument("Orientation CSV line " + std::to_string(line_no) + ": " + why);
        };
        if (!header) {
            if (line != "qw,qx,qy,qz,weight,samples,seed")
                return fail("Expected CSV header qw,qx,qy,qz,weight,samples,seed.");
            header = true;
            continue;
        }
        std::vector<std::string> cols;
        size_t p = 0;
        while (true) {
            const size_t q = line.find(',', p);
            cols.push_back(line.substr(p, q == std::string::npos ? std::string::npos : q - p));
            if (q == std::string::npos) break;
            p = q + 1;
        }
        if (cols.size() != 7) return fail("Expected seven CSV fields.");
        OrientationNode n{};
        for (int i = 0; i < 4; ++i)
            if (!parse_number(cols[i], n.q[i])) return fail("Malformed orientation number.");
        if (!parse_number(cols[4], n.weight) || !parse_number(cols[5], n.samples) || !parse_number(cols[6], n.seed))
            return fail("Malformed orientation number.");
        try {
            rotation_of_quaternion(n.q[0], n.q[1], n.q[2], n.q[3], n.rotation);
        } catch (const std::exception& e) {
            return fail(e.what());
        }
        if (!std::isfinite(n.weight) || !(n.weight > 0) || n.samples == 0)
            return fail("Weights and sample counts must be positive.");
        if (n.samples > UINT64_MAX - plan.total_samples) return fail("Planned sample count overflows.");
        plan.total_samples += n.samples;
        {
            const double t = wsum + n.weight;
            wcomp += std::fabs(wsum) >= std::fabs(n.weight) ? (wsum - t) + n.weight : (n.weight - t) + wsum;
            wsum = t;
        }
        plan.nodes.push_back(n);
    }
    const double total = wsum + wcomp;
    if (!header || plan.nodes.empty() || !std::isfinite(total) || std::fabs(total - 1) > 2e-12)
        throw std::invalid_argument("Orientation number-probability weights must sum to one (within 2e-12).");
    plan.input_weight_sum = total;
    for (auto& n : plan.nodes) n.weight /= total;
    return plan;
}

// SHA-256 (FIPS 180-4).
std::string sha256_hex(const std::string& bytes)
{
    static const uint32_t K[64] = {
        0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5,
        0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174,
        0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
        0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967,
        0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85,
        0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
        0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
        0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2};
    uint32_t H[8] = {0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a, 0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19};
    std::string msg = bytes;
    const uint64_t bits = (uint64_t)bytes.size() * 8;
    msg.push_back((char)0x80);
    while (msg.size() % 64 != 56) msg.push_back('\0');
    for (int i = 7; i >= 0; --i) msg.push_back((char)((bits >> (8 * i)) & 0xff));
    auto rotr = [](uint32_t v, int s) { return (v >> s) | (v << (32 - s)); };
    for (size_t off = 0; off < msg.size(); off += 64) {
        uint32_t w[64];
        for (int t = 0; t < 16; ++t)
            w[t] = (uint32_t)(unsigned char)msg[off + 4 * t] << 24 | (uint32_t)(unsigned char)msg[off + 4 * t + 1] << 16 |
                   (uint32_t)(unsigned char)msg[off + 4 * t + 2] << 8 | (uint32_t)(unsigned char)msg[off + 4 * t + 3];
        for (int t = 16; t < 64; ++t) {
            const uint32_t s0 = rotr(w[t - 15], 7) ^ rotr(w[t - 15], 18) ^ (w[t - 15] >> 3);
            const uint32_t s1 = rotr(w[t - 2], 17) ^ rotr(w[t - 2], 19) ^ (w[t - 2] >> 10);
            w[t] = w[t - 16] + s0 + w[t - 7] + s1;
        }
        uint32_t a = H[0], b = H[1], c = H[2], d = H[3], e = H[4], f = H[5], g = H[6], h = H[7];
        for (int t = 0; t < 64; ++t) {
            const uint32_t S1 = rotr(e, 6) ^ rotr(e, 11) ^ rotr(e, 25);
            const uint32_t ch = (e & f) ^ (~e & g);
            const uint32_t t1 = h + S1 + ch + K[t] + w[t];
            const uint32_t S0 = rotr(a, 2) ^ rotr(a, 13) ^ rotr(a, 22);
            const uint32_t mj = (a & b) ^ (a & c) ^ (b & c);
            const uint32_t t2 = S0 + mj;
            h = g, g = f, f = e, e = d + t1, d = c, c = b, b = a, a = t1 + t2;
        }
        H[0] += a, H[1] += b, H[2] += c, H[3] += d, H[4] += e, H[5] += f, H[6] += g, H[7] += h;
    }
    static const char* hex = "0123456789abcdef";
    std::string out;
    for (uint32_t v : H)
        for (int s = 28; s >= 0; s -= 4) out.push_back(hex[(v >> s) & 15]);
    return out;
}


OrientationPlan read_orientations(const std::string& path)
{
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("Cannot read orientation plan.");
    in.seekg(0, std::ios::end);
    const std::streamoff size = in.tellg();
    if (size < 0 || size > (std::streamoff)256 * 1024 * 1024)
        throw std::length_error("Orientation plan exceeds explicit 256 MiB input limit.");
    in.seekg(0);
    std::string data((size_t)size, '\0');
    in.read(&data[0], size);
    if (in.gcount() != size) throw std::runtime_error("Short/changed orientation file.");
    return parse_orientations(data);
}

}  // namespace cli
