// Orientation composition of a deterministic atlas on the device (C-ABI in
// include/icephase.h): the composed cell mass of one incident condition is a
// fixed-order sum of azimuth-shifted (and, for improper symmetry operations,
// azimuth-reflected) atlas node blocks,
//
//   out[t][p] = sum over (i, j) ascending, parity 0 then 1, shift s ascending,
//               w[i][j][parity][s] > 0, of  w * cell_parity(i, j)[t][(p - s) mod np],
//   cell_0 = mass[i][j],  cell_1[t][q] = mass[i][j][t][np - 1 - q],
//
// which is compose_mass of hex_atlas/python/atlas_math.py:394-411 evaluated in
// the same order with the same roundings (product rounded, then added; no
// fused multiply-add), so the result is bitwise the reference's.
//
// One wave per (row t, 64-wide azimuth chunk), about one wave per SIMD for a
// 180 x 360 grid: the wave stages the node's row (reversed for parity 1) in
// LDS (dynamic), then every lane accumulates its output cell over
// the group's terms. The sum of a cell is one serial chain of adds (its order
// is the result), so the terms are read eight at a time (one scalar load of
// 8 x 16 bytes) and their LDS reads and products issued ahead of the adds.
// The atlas mass stays resident on the device across conditions and
// wavelengths.
#include "runtime_internal.hpp"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

namespace {

constexpr uint32_t kMaxPhi = 8192;  // LDS row: at most 64 KiB of doubles
constexpr uint32_t kLanes = 64;
constexpr uint32_t kUnroll = 8;

struct Group {
    uint64_t cell;    // offset of the node's [nt][np] block
    uint32_t parity;  // 1: azimuth-reflected block
    uint32_t first;   // first term
    uint32_t count;
    uint32_t pad;
};

struct Term {
    uint32_t shift, pad;
    double weight;
};

// WRAP: the row is staged twice in a row (2 np doubles), so the wrapped read
// (p - s) mod np is the plain index p + np - s (np <= 4096); otherwise the
// wrap is a compare and select.
template <bool WRAP>
__global__ void __launch_bounds__(64) icephase_compose(const double* __restrict__ mass, uint32_t np,
                                                       const Group* __restrict__ groups, uint32_t ngroups,
                                                       const Term* __restrict__ terms, double* __restrict__ out)
{
    extern __shared__ double row[];
    const uint32_t t = blockIdx.y;
    const uint32_t p = blockIdx.x * kLanes + threadIdx.x;
    const uint32_t pn = p + np;
    double acc = 0.0;
    for (uint32_t g = 0; g < ngroups; ++g) {
        const Group G = groups[g];
        const double* src = mass + G.cell + (uint64_t)t * np;
        __syncthreads();
        for (uint32_t q = threadIdx.x; q < np; q += kLanes) {
            const double v = src[G.parity ? np - 1 - q : q];
            row[q] = v;
            if (WRAP) row[q + np] = v;
        }
        __syncthreads();
        if (p < np) {
            const Term* T = terms + G.first;
            uint32_t k = 0;
            for (; k + kUnroll <= G.count; k += kUnroll) {
                double prod[kUnroll];
#pragma unroll
                for (uint32_t u = 0; u < kUnroll; ++u) {
                    const uint32_t s = T[k + u].shift;
                    const uint32_t q = WRAP ? pn - s : (p >= s ? p - s : pn - s);
                    prod[u] = T[k + u].weight * row[q];
                }
#pragma unroll
                for (uint32_t u = 0; u < kUnroll; ++u) acc = acc + prod[u];
            }
            for (; k < G.count; ++k) {
                const uint32_t s = T[k].shift;
                const uint32_t q = WRAP ? pn - s : (p >= s ? p - s : pn - s);
                acc = acc + T[k].weight * row[q];
            }
        }
    }
    if (p < np) out[(uint64_t)t * np + p] = acc;
}

}  // namespace

using namespace icephase;

struct iph_atlas_mass {
    iph_ctx* ctx = nullptr;
    uint32_t ni = 0, nj = 0, nt = 0, np = 0;
    DevBuf<double> mass;
};

extern "C" {

int iph_atlas_mass_create(iph_ctx* ctx, const double* mass, uint32_t ni, uint32_t nj, uint32_t nt, uint32_t np,
                          iph_atlas_mass** out)
{
    if (!ctx) return fail(nullptr, IPH_E_INVALID, "ctx must not be NULL.");
    if (!mass || !out || !ni || !nj || !nt || !np || np > kMaxPhi || nt > 65535 ||
        (uint64_t)ni * nj * nt * np > (1ull << 36))
        return fail(ctx, IPH_E_INVALID, "Invalid atlas mass dimensions (phi_count <= 8192, theta_count <= 65535).");
    *out = nullptr;
    IPH_HIP(ctx, hipSetDevice(ctx->device));
    auto* m = new iph_atlas_mass;
    m->ctx = ctx;
    m->ni = ni, m->nj = nj, m->nt = nt, m->np = np;
    const size_t n = (size_t)ni * nj * nt * np;
    hipError_t e = m->mass.alloc(n);
    if (e == hipSuccess) e = hipMemcpy(m->mass.p, mass, n * sizeof(double), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        delete m;
        return fail(ctx, IPH_E_DEVICE, std::string("atlas mass upload: ") + hipGetErrorString(e));
    }
    ctx->atlases.push_back(m);
    *out = m;
    return IPH_OK;
}

void iph_atlas_mass_destroy(iph_atlas_mass* m)
{
    if (!m) return;
    iph_ctx* ctx = m->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (size_t k = 0; k < ctx->atlases.size(); ++k)
        if (ctx->atlases[k] == m) {
            ctx->atlases.erase(ctx->atlases.begin() + (long)k);
            break;
        }
    delete m;
}

int iph_atlas_compose(iph_atlas_mass* m, const double* weights, double* out)
{
    if (!m) return fail(nullptr, IPH_E_INVALID, "atlas mass must not be NULL.");
    iph_ctx* ctx = m->ctx;
    if (!weights || !out) return fail(ctx, IPH_E_INVALID, "Invalid composition arguments.");
    const uint32_t np = m->np, nt = m->nt;
    const size_t cells = (size_t)nt * np;
    std::vector<Group> groups;
    std::vector<Term> terms;
    for (uint32_t i = 0; i < m->ni; ++i)
        for (uint32_t j = 0; j < m->nj; ++j)
            for (uint32_t parity = 0; parity < 2; ++parity) {
                const double* w = weights + (((size_t)i * m->nj + j) * 2 + parity) * np;
                Group g{((size_t)i * m->nj + j) * cells, parity, (uint32_t)terms.size(), 0, 0};
                for (uint32_t s = 0; s < np; ++s) {
                    if (!std::isfinite(w[s]) || w[s] < 0)
                        return fail(ctx, IPH_E_INVALID, "Composition weights must be finite and nonnegative.");
                    if (w[s] > 0) {
                        terms.push_back(Term{s, 0, w[s]});
                        ++g.count;
                    }
                }
                if (g.count) groups.push_back(g);
            }
    if (groups.empty()) return fail(ctx, IPH_E_INVALID, "Composition weights are all zero.");
    IPH_HIP(ctx, hipSetDevice(ctx->device));
    DevBuf<Group> d_groups;
    DevBuf<Term> d_terms;
    DevBuf<double> d_out;
    IPH_HIP(ctx, d_groups.alloc(groups.size()));
    IPH_HIP(ctx, d_terms.alloc(terms.size()));
    IPH_HIP(ctx, d_out.alloc(cells));
    hipStream_t s = ctx->stream;
    IPH_HIP(ctx, hipMemcpyAsync(d_groups.p, groups.data(), groups.size() * sizeof(Group), hipMemcpyHostToDevice, s));
    IPH_HIP(ctx, hipMemcpyAsync(d_terms.p, terms.data(), terms.size() * sizeof(Term), hipMemcpyHostToDevice, s));
    const dim3 grid((np + kLanes - 1) / kLanes, nt);
    if (np <= kMaxPhi / 2)
        hipLaunchKernelGGL(icephase_compose<true>, grid, dim3(kLanes), 2 * np * sizeof(double), s, m->mass.p, np,
                           d_groups.p, (uint32_t)groups.size(), d_terms.p, d_out.p);
    else
        hipLaunchKernelGGL(icephase_compose<false>, grid, dim3(kLanes), np * sizeof(double), s, m->mass.p, np,
                           d_groups.p, (uint32_t)groups.size(), d_terms.p, d_out.p);
    IPH_HIP(ctx, hipGetLastError());
    IPH_HIP(ctx, hipMemcpyAsync(out, d_out.p, cells * sizeof(double), hipMemcpyDeviceToHost, s));
    IPH_HIP(ctx, hipStreamSynchronize(s));
    return IPH_OK;
}

}  // extern "C"
