// Fraunhofer diffraction lobe of a hex record's projected aperture.
//
// The reference hex path carries no diffraction (hex_trace/README.md:20-21);
// this lobe is the closed-form scalar Fraunhofer pattern of the polygon the
// prism projects onto the plane normal to the incident direction:
//   dsigma/dOmega(q) = |F(q)|^2 / lambda^2,   F(q) = int_P exp(-i q.x) dA,
//   q = (2 pi / lambda) (d.p1, d.p2),
// with F in closed form from the divergence theorem over the polygon edges
// (edge a->b, e = b - a, midpoint m):
//   F(q) = (i / |q|^2) sum_edges (q1 e_y - q2 e_x) exp(-i q.m) sinc(q.e / 2).
// By Parseval the lobe integrates to the aperture area over the small-angle
// sphere, so it carries exactly the projected area (the extinction paradox).
//
// Cell masses are integrated over each (theta, phi) cell in theta and phi with
// 4-point Gauss-Legendre on sub-panels no wider than half a fringe
// (lambda / (2 dmax)), for the leading `rows` theta rows. Three kernels:
// masses (+ per-workgroup partial sums), the ordered total, and the deposit
// of rint(mass / total * rays * qfactor * 2^Q) quanta into the histogram, so
// the lobe carries the record's whole incident power. All sums are in fixed
// order: the lobe is bitwise reproducible.
#include "kernels.hpp"

namespace icephase {

namespace {

constexpr double kGx[4] = {-0.8611363115940526, -0.3399810435848563, 0.3399810435848563, 0.8611363115940526};
constexpr double kGw[4] = {0.3478548451374538, 0.6521451548625461, 0.6521451548625461, 0.3478548451374538};

__host__ __device__ inline double lobe_density(const Lobe& L, double q1, double q2)
{
    const double qq = q1 * q1 + q2 * q2;
    if (sqrt(qq) * L.dmax < 1e-4) return L.area * L.area * L.inv_lambda2;  // centred aperture: F = A - O(q^2)
    double re = 0, im = 0;
    for (uint32_t k = 0; k < L.nv; ++k) {
        const uint32_t k1 = k + 1 == L.nv ? 0 : k + 1;
        const double ex = L.vx[k1] - L.vx[k], ey = L.vy[k1] - L.vy[k];
        const double mx = 0.5 * (L.vx[k1] + L.vx[k]), my = 0.5 * (L.vy[k1] + L.vy[k]);
        const double qn = q1 * ey - q2 * ex;
        const double s = 0.5 * (q1 * ex + q2 * ey);
        const double sinc = fabs(s) < 1e-8 ? 1.0 - s * s / 6.0 : sin(s) / s;
        const double ph = -(q1 * mx + q2 * my);
        const double w = qn * sinc;
        re += w * cos(ph);
        im += w * sin(ph);
    }
    // F = i (re + i im) / |q|^2
    return (re * re + im * im) / (qq * qq) * L.inv_lambda2;
}

__host__ __device__ inline int clamp_panels(double x)
{
    const double c = ceil(x);
    return c < 1 ? 1 : (c > 64 ? 64 : (int)c);
}

}  // namespace

// Mass (mm^2) of the lobe in cell (i, j) of an nt x np grid whose binning frame
// is g. Shared by the kernel and the host (the C-ABI's scalar test hook).
__host__ __device__ inline double lobe_cell_mass(const Lobe& L, const BinFrame& g, uint32_t nt, uint32_t np, uint32_t i,
                                                 uint32_t j)
{
    const double lambda = 2 * kPi / L.k0;
    const double t0 = kPi * i / nt, t1 = kPi * (i + 1) / nt;
    const double f0 = -kPi + 2 * kPi * j / np, f1 = -kPi + 2 * kPi * (j + 1) / np;
    const double fringe = lambda / L.dmax;
    const double thi = t1 < 0.5 * kPi ? t1 : 0.5 * kPi;
    const int nts = clamp_panels(2 * (t1 - t0) / fringe);
    const int nps = clamp_panels(2 * sin(thi) * (f1 - f0) / fringe);
    const double ht = (t1 - t0) / nts, hp = (f1 - f0) / nps;
    double total = 0;
    for (int a = 0; a < nts; ++a) {
        const double ta = t0 + ht * a;
        double rowsum = 0;
        for (int gt = 0; gt < 4; ++gt) {
            const double th = ta + 0.5 * ht * (1 + kGx[gt]);
            const double st = sin(th), ct = cos(th);
            double colsum = 0;
            for (int b = 0; b < nps; ++b) {
                const double fb = f0 + hp * b;
                for (int gp = 0; gp < 4; ++gp) {
                    const double ph = fb + 0.5 * hp * (1 + kGx[gp]);
                    const double cx = st * cos(ph), cy = st * sin(ph);
                    const d3 d = D3(cx * g.e0.x + cy * g.e1.x + ct * g.ki.x, cx * g.e0.y + cy * g.e1.y + ct * g.ki.y,
                                    cx * g.e0.z + cy * g.e1.z + ct * g.ki.z);
                    colsum += kGw[gp] * lobe_density(L, L.k0 * dot(d, L.p1), L.k0 * dot(d, L.p2));
                }
            }
            rowsum += kGw[gt] * st * colsum;
        }
        total += rowsum;
    }
    return total * (0.5 * ht) * (0.5 * hp);
}

__device__ inline double block_sum256(double v, double* sh)
{
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

extern "C" __global__ void __launch_bounds__(256) icephase_lobe_cells(const DevRecord* __restrict__ recs, uint32_t rec)
{
    __shared__ double sh[256];
    const DevRecord& R = recs[rec];
    const uint64_t n = (uint64_t)R.lobe.rows * R.np;
    const uint64_t cell = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    double m = 0;
    if (cell < n) {
        m = lobe_cell_mass(R.lobe, R.frame, R.nt, R.np, (uint32_t)(cell / R.np), (uint32_t)(cell % R.np));
        R.lobe.mass[cell] = m;
    }
    const double s = block_sum256(m, sh);
    if (threadIdx.x == 0) R.lobe.partial[blockIdx.x] = s;
}

extern "C" __global__ void __launch_bounds__(256)
    icephase_lobe_total(const DevRecord* __restrict__ recs, uint32_t rec, uint32_t nblocks)
{
    __shared__ double sh[256];
    const DevRecord& R = recs[rec];
    double v = 0;
    for (uint32_t k = threadIdx.x; k < nblocks; k += 256) v += R.lobe.partial[k];
    const double total = block_sum256(v, sh);
    if (threadIdx.x == 0) {
        R.lobe.partial[nblocks] = total;
        R.lobe.partial[nblocks + 1] = total / R.lobe.area;  // captured fraction, read by the host
    }
}

extern "C" __global__ void __launch_bounds__(256)
    icephase_lobe_deposit(const DevRecord* __restrict__ recs, uint32_t rec, uint32_t nblocks)
{
    const DevRecord& R = recs[rec];
    const uint64_t n = (uint64_t)R.lobe.rows * R.np;
    const uint64_t cell = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    const double total = R.lobe.partial[nblocks];
    unsigned long long q = 0;
    if (cell < n && total > 0) {
        q = quantize(R.lobe.mass[cell] / total * ((double)R.rays * R.qfactor), R.qbits);
        if (q) atomicAdd(&R.bins[cell], q);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) q += __shfl_xor(q, off, 64);
    if ((threadIdx.x & 63) == 0 && q) atomicAdd(&R.stats[kStatDiffraction], q);
}

hipError_t launch_lobe(const DevRecord* d_recs, uint32_t rec, uint32_t rows, uint32_t np, hipStream_t s)
{
    const uint64_t n = (uint64_t)rows * np;
    const uint32_t blocks = (uint32_t)((n + 255) / 256);
    hipLaunchKernelGGL(icephase_lobe_cells, dim3(blocks), dim3(256), 0, s, d_recs, rec);
    hipLaunchKernelGGL(icephase_lobe_total, dim3(1), dim3(256), 0, s, d_recs, rec, blocks);
    hipLaunchKernelGGL(icephase_lobe_deposit, dim3(blocks), dim3(256), 0, s, d_recs, rec, blocks);
    return hipGetLastError();
}

uint32_t lobe_blocks(uint32_t rows, uint32_t np) { return (uint32_t)(((uint64_t)rows * np + 255) / 256); }

double host_lobe_cell_mass(const Lobe& L, const BinFrame& g, uint32_t nt, uint32_t np, uint32_t i, uint32_t j)
{
    return lobe_cell_mass(L, g, nt, np, i, j);
}

}  // namespace icephase
