// Internals shared by the C-ABI translation units (runtime.hip,
// surrogate_api.hip): the context, error reporting and device buffers.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/icephase.h"

struct iph_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int cus = 256;
    std::string err;
    std::vector<iph_batch*> batches;  // live batches; destroyed with the context
    std::vector<iph_sur*> surrogates; // live surrogates; destroyed with the context
    std::vector<iph_atlas_mass*> atlases;  // live device atlas masses; destroyed with the context
    std::vector<iph_coherent*> coherents;  // live coherent drop records; destroyed with the context
    unsigned char* cdf_scratch = nullptr;  // device address of this device's fixed CDF scratch (runtime.hip)
};

namespace icephase {

// message of the last context-free failure on this thread
inline std::string& thread_error()
{
    thread_local std::string s;
    return s;
}

inline int fail(iph_ctx* c, int code, const std::string& msg)
{
    if (c)
        c->err = msg;
    else
        thread_error() = msg;
    return code;
}

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(size_t count)
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = count;
        if (!count) return hipSuccess;
        return hipMalloc(&p, count * sizeof(T));
    }
};

// Limits of the persistent tile queue (a 32-bit ticket counter per launch) and
// of a tile's 32-bit ray counter.
constexpr uint64_t kMaxTileRays = 1ull << 31;
constexpr uint64_t kMaxQueueTickets = 0xffffffffull;

// Library default tile: about 64 tiles per record, 2048..65536 rays per tile.
inline uint64_t default_tile_rays(uint64_t rays)
{
    const uint64_t t = rays / 64;
    return t < 2048 ? 2048 : (t > 65536 ? 65536 : t);
}

// Grid (theta_count, phi_count) of record rec of a batch; false for an invalid
// record index.
bool batch_record_grid(const iph_batch* b, uint32_t rec, uint32_t& nt, uint32_t& np);

// The CDF of FP64 cell masses already in device memory (theta-major nt x np),
// as iph_build_cdf does for a host array; phi and theta are host outputs.
int build_cdf_from_device(iph_ctx* ctx, const double* d_mass, uint32_t nt, uint32_t np, const iph_cdf_policy* policy,
                          double* phi, double* theta, iph_cdf_report* report);

}  // namespace icephase

#define IPH_HIP(ctx, call)                                                                                  \
    do {                                                                                                    \
        const hipError_t e_ = (call);                                                                       \
        if (e_ != hipSuccess)                                                                               \
            return ::icephase::fail((ctx), IPH_E_DEVICE,                                                    \
                                    std::string(#call) + " failed: " + hipGetErrorString(e_) + " (" +       \
                                        __FILE__ + ":" + std::to_string(__LINE__) + ")");                   \
    } while (0)
