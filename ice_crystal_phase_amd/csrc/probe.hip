// Atomic-throughput calibration (measurement only; no reference counterpart).
//
// The trace kernels bin every emission with one u64 atomicAdd into a record's
// fixed-point histogram slab. This kernel measures what the hardware sustains
// for the same operation on its own: `atomics` increments of 1 to cells drawn
// by a counter hash from a slab of `cells` u64 (the histogram footprint of a
// batch), to consecutive cells in the order of the increments (a wave's 64
// lanes update 64 adjacent cells: the device's ceiling for the operation), or
// all to one cell (the worst contention). The slab is summed afterwards; the
// sum must equal the number of increments exactly.
#include "runtime_internal.hpp"

#include <hip/hip_runtime.h>

namespace {

// 32-bit integer hash (multiply-xorshift): a few VALU instructions per
// increment, so the probe measures the atomics rather than the address arithmetic.
__device__ inline uint32_t mix32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    return x ^ (x >> 16);
}

__global__ void __launch_bounds__(256) probe_zero(unsigned long long* __restrict__ slab, uint64_t cells)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < cells; i += (uint64_t)gridDim.x * blockDim.x)
        slab[i] = 0ull;
}

// mode 0: random cells (a 32-bit hash of the increment's index scaled to the
// cell count by a 32 x 32 -> 64 multiply, no modulo); mode 1: cell 0 only;
// mode 2: cell = index mod cells, kept by a wrapping counter (stride mod cells
// added per step, so no 64-bit modulo inside the loop).
__global__ void __launch_bounds__(256) probe_atomics(unsigned long long* __restrict__ slab, uint64_t cells,
                                                     uint64_t atomics, uint32_t mode)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint32_t n = (uint32_t)cells;
    const uint64_t first = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t step = (uint32_t)(stride % cells);
    uint32_t next = (uint32_t)(first % cells);
    for (uint64_t i = first; i < atomics; i += stride) {
        uint64_t cell = 0;
        if (mode == 0) {
            const uint32_t h = mix32((uint32_t)i ^ (uint32_t)(i >> 32) * 0x9e3779b9u);
            cell = ((uint64_t)h * n) >> 32;
        } else if (mode == 2) {
            cell = next;
            next = next >= n - step ? next - (n - step) : next + step;
        }
        atomicAdd(&slab[cell], 1ull);
    }
}

__global__ void __launch_bounds__(256) probe_sum(const unsigned long long* __restrict__ slab, uint64_t cells,
                                                 unsigned long long* __restrict__ total)
{
    unsigned long long s = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < cells; i += (uint64_t)gridDim.x * blockDim.x)
        s += slab[i];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(total, s);
}

}  // namespace

using namespace icephase;

extern "C" int iph_atomic_probe(iph_ctx* ctx, uint32_t mode, uint64_t cells, uint64_t atomics, double* ms)
{
    if (!ctx || !ms || mode > 2 || cells == 0 || cells > 0xffffffffull || atomics == 0 || atomics > (1ull << 40))
        return fail(ctx, IPH_E_INVALID, "Invalid atomic probe arguments (mode 0|1|2, 1 <= cells < 2^32, "
                                        "1 <= atomics <= 2^40).");
    IPH_HIP(ctx, hipSetDevice(ctx->device));
    DevBuf<unsigned long long> slab, total;
    IPH_HIP(ctx, slab.alloc(cells));
    IPH_HIP(ctx, total.alloc(1));
    hipStream_t s = ctx->stream;
    const uint32_t blocks = (uint32_t)ctx->cus * 8;
    hipLaunchKernelGGL(probe_zero, dim3(blocks), dim3(256), 0, s, slab.p, cells);
    hipLaunchKernelGGL(probe_zero, dim3(1), dim3(64), 0, s, total.p, (uint64_t)1);
    hipEvent_t ev[2];
    IPH_HIP(ctx, hipEventCreate(&ev[0]));
    const hipError_t e1 = hipEventCreate(&ev[1]);
    if (e1 != hipSuccess) {
        (void)hipEventDestroy(ev[0]);
        return fail(ctx, IPH_E_DEVICE, "hipEventCreate failed");
    }
    hipError_t e = hipEventRecord(ev[0], s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(probe_atomics, dim3(blocks), dim3(256), 0, s, slab.p, cells, atomics, mode);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(ev[1], s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(probe_sum, dim3(blocks), dim3(256), 0, s, slab.p, cells, total.p);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    float t = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, ev[0], ev[1]);
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    if (e != hipSuccess) return fail(ctx, IPH_E_DEVICE, std::string("atomic probe: ") + hipGetErrorString(e));
    unsigned long long sum = 0;
    IPH_HIP(ctx, hipMemcpy(&sum, total.p, sizeof sum, hipMemcpyDeviceToHost));
    if (sum != atomics) return fail(ctx, IPH_E_AUDIT, "atomic probe: the slab sum differs from the increments.");
    *ms = t;
    return IPH_OK;
}
