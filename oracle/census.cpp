// Test infrastructure only: the operation census of the drop tracer, the
// per-event algorithmic work that the drop kernel's roofline counts. Nothing in
// the product links or loads it.
//
// The oracle's drop trace (oracle_drop_impl.h, the FP32 mirror of the kernel)
// and its phase_bin rule (oracle_bin_rule.h) are instantiated here with
// counting number types. One operation is counted for every +, -, *, / and
// square root, two for every fused multiply-add (a product and a sum, the
// convention of the FP32 peak rate), and one for every libm call the algorithm
// makes (hypot, atan2, floor); comparisons, absolute values, min/max,
// negations and conversions are not counted. The trace marks its events (oracle_drop_impl.h, CENSUS_MARK), so
// the operations between marks are attributed to them:
//   miss      a ray that misses the drop: ray set-up and the entry search;
//   entry     a hitting ray: set-up, entry solve, first split and its external
//             reflection's power, plus the ray's last partial encounter;
//   hit       one internal encounter: exit solve, normal, split, and the
//             transmitted branch's power;
//   emit      the binning of one emitted direction (the phase_bin rule in FP64).
#include "oracle.h"

#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace census {

uint64_t ops = 0;

template <class T>
struct Num {
    T v;
    Num() = default;
    Num(double x) : v((T)x) {}
    Num(float x) : v((T)x) {}
    Num(int x) : v((T)x) {}
    Num(unsigned x) : v((T)x) {}
    explicit operator double() const { return (double)v; }
    explicit operator float() const { return (float)v; }
    explicit operator uint32_t() const { return (uint32_t)v; }
    static Num of(T x) { Num r; r.v = x; return r; }
    // hidden friends: either operand may be converted (literals, integers)
    friend Num operator+(Num a, Num b) { ++ops; return of(a.v + b.v); }
    friend Num operator-(Num a, Num b) { ++ops; return of(a.v - b.v); }
    friend Num operator*(Num a, Num b) { ++ops; return of(a.v * b.v); }
    friend Num operator/(Num a, Num b) { ++ops; return of(a.v / b.v); }
    friend Num operator-(Num a) { return of(-a.v); }
    friend bool operator<(Num a, Num b) { return a.v < b.v; }
    friend bool operator>(Num a, Num b) { return a.v > b.v; }
    friend bool operator<=(Num a, Num b) { return a.v <= b.v; }
    friend bool operator>=(Num a, Num b) { return a.v >= b.v; }
    friend bool operator==(Num a, Num b) { return a.v == b.v; }
    friend bool operator!=(Num a, Num b) { return a.v != b.v; }
};

using F = Num<float>;
using D = Num<double>;

inline F sqrt_f(F a) { ++ops; return F::of(std::sqrt(a.v)); }
inline F fabs_f(F a) { return F::of(std::fabs(a.v)); }
inline F fma_f(F a, F b, F c) { ops += 2; return F::of(std::fma(a.v, b.v, c.v)); }
inline D sqrt_d(D a) { ++ops; return D::of(std::sqrt(a.v)); }
inline D hypot_d(D a, D b) { ++ops; return D::of(std::hypot(a.v, b.v)); }
inline D atan2_d(D a, D b) { ++ops; return D::of(std::atan2(a.v, b.v)); }
inline D floor_d(D a) { ++ops; return D::of(std::floor(a.v)); }
inline D fabs_d(D a) { return D::of(std::fabs(a.v)); }
inline D fmin_d(D a, D b) { return D::of(std::fmin(a.v, b.v)); }
inline D fmax_d(D a, D b) { return D::of(std::fmax(a.v, b.v)); }

// event marks of the current ray: (event, operations so far)
struct Mark {
    int event;
    uint64_t ops;
};
std::vector<Mark> marks;

}  // namespace census

using census::D;
using census::F;

#define PI_D 3.141592653589793238462643383279502884
#define ST_NUMERIC_D 5
#define CENSUS_MARK(event) census::marks.push_back(census::Mark{(event), census::ops})

// the drop trace, FP32 mirror, counted
#define REAL F
#define SFX c
#define SQRT census::sqrt_f
#define FABS census::fabs_f
#define FMA census::fma_f
#define TOL 0x1p-21f
#define TINY_S2 0x1p-40f
#define SELF_OFFSET 0x1p-16f
#define MIN_STEP 0x1p-20f
#include "oracle_drop_impl.h"
#undef REAL
#undef SFX
#undef SQRT
#undef FABS
#undef FMA

// the phase_bin rule in FP64, counted
#define BREAL D
#define BFN(name) name##c
#define BVAL(x) ((double)(x))
#define BHYPOT census::hypot_d
#define BATAN2 census::atan2_d
#define BFLOOR census::floor_d
#define BSQRT census::sqrt_d
#define BFABS census::fabs_d
#define BFMIN census::fmin_d
#define BFMAX census::fmax_d
#include "oracle_bin_rule.h"

extern "C" {

// Operation totals and event counts of rays [first, first + count) of a drop
// record (ray id = iy * grid_w + ix), traced by the counted FP32 mirror and
// binned by the counted phase_bin rule on the record's frame and nt x np grid.
struct orc_census {
    uint64_t misses, entries, hits, emissions;
    uint64_t ops_miss, ops_entry, ops_hit, ops_emit;
};

int orc_drop_census(const orc_drop_record* r, uint32_t nt, uint32_t np, uint64_t first, uint64_t count,
                    orc_census* out)
{
    std::memset(out, 0, sizeof *out);
    std::vector<double> ue(nt + 1);
    if (orc_u_edges(nt, ue.data())) return -1;
    orc_frame g;
    const double k[3] = {r->ki[0], r->ki[1], r->ki[2]};
    if (orc_drop_frame(k, &g)) return -2;
    rec_c R;
    for (int n = 0; n < 8; ++n) R.c[n] = r->coeffs[n];
    R.inner = r->inner; R.outer = r->outer; R.rmax = r->rmax;
    R.ki = mk_c(r->ki[0], r->ki[1], r->ki[2]);
    R.e0 = mk_c(r->e0[0], r->e0[1], r->e0[2]);
    R.e1 = mk_c(r->e1[0], r->e1[1], r->e1[2]);
    R.n_int = r->n_int; R.n_ext = r->n_ext; R.h = r->half_extent; R.tail = r->tail;
    R.w = r->grid_w; R.hgt = r->grid_h; R.max_hits = r->max_hits; R.sphere = r->sphere;
    R.terms = 2;
    for (uint32_t m = 2; m < 8; ++m)
        if (r->coeffs[m] != 0.0f) R.terms = m + 1;
    orc_sample buf[64];
    for (uint64_t i = 0; i < count; ++i) {
        census::marks.clear();
        uint64_t n = 0;
        int overflow = 0;
        (void)trace_ray_c(&R, first + i, buf, 64, &n, &overflow);
        if (overflow) return -3;
        const auto& m = census::marks;
        uint64_t last = m.front().ops;
        for (size_t j = 1; j < m.size(); ++j) {
            const uint64_t d = m[j].ops - last;
            switch (m[j].event) {
            case 1: out->misses++; out->ops_miss += d; break;
            case 2: out->entries++; out->ops_entry += d; break;
            case 3: out->hits++; out->ops_hit += d; break;
            case 4: out->ops_entry += d; break;   // the ray's last partial encounter
            default: break;
            }
            last = m[j].ops;
        }
        for (uint64_t s = 0; s < n; ++s) {
            const uint64_t before = census::ops;
            (void)bin_raw_c(ue.data(), nt, np, &g, buf[s].dir);
            out->emissions++;
            out->ops_emit += census::ops - before;
        }
    }
    return 0;
}

}  // extern "C"
