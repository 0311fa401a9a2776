// grid.hip -- cavity-bias insertion grid on the GPU (gcmc/python/dataset.py:21-139).
//
// Cells are int((x - start)/dx) % n per axis (truncation toward zero, then a
// non-negative remainder, in FP64 as the reference's Python does); a cell is
// occupied when any atom of the initial structure falls in it.  The grid is
// the list of corners start + i*dx (FP64, stored as FP32) of the free cells,
// in linear order (x slowest, z fastest); the reference lists the same
// cells in CPython set order, a permutation of these.
//
// Three kernels, all bound by HBM traffic of integer work:
//   k_grid_mark     one thread per atom: FP64 cell index, atomic OR into a
//                   bitmap of occupied cells (32 cells per word)
//   k_grid_count    one thread per bitmap word: free cells per word, wave
//                   and workgroup scans -> per-workgroup totals
//   k_grid_write    each word writes the corners of its free cells at its
//                   workgroup's offset (exclusive scan of the workgroup totals,
//                   done on the host, which needs the count anyway)
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/pygcmc_hip.h"

namespace gcmc_grid {

constexpr int WG = 256;

__global__ void k_fill_u32(uint32_t *p, size_t n, uint32_t v)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__device__ inline long long cell_of(double x, double start, double dx, int n)
{
    // Python int() truncates toward zero; % with n > 0 is non-negative
    const long long q = (long long)trunc((x - start) / dx);
    long long m = q % n;
    return m < 0 ? m + n : m;
}

__global__ void k_grid_mark(const double *xyz, int natom, double s0, double s1, double s2, double dx, int n0,
                            int n1, int n2, uint32_t *bits)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= natom) return;
    const long long i = cell_of(xyz[3 * (size_t)a + 0], s0, dx, n0);
    const long long j = cell_of(xyz[3 * (size_t)a + 1], s1, dx, n1);
    const long long k = cell_of(xyz[3 * (size_t)a + 2], s2, dx, n2);
    const long long lin = (i * n1 + j) * n2 + k;
    atomicOr(&bits[lin >> 5], 1u << (unsigned)(lin & 31));
}

// free cells of word w (bits beyond the last cell do not count)
__device__ inline uint32_t free_mask(const uint32_t *bits, long long w, long long total)
{
    uint32_t f = ~bits[w];
    const long long lo = w << 5;
    if (lo + 32 > total) f &= (1u << (unsigned)(total - lo)) - 1u;
    return f;
}

__device__ inline int block_exclusive_scan(int v, int *tmp)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) tmp[wave] = x;
    __syncthreads();
    int base = 0;
    for (int w = 0; w < wave; ++w) base += tmp[w];
    __syncthreads();
    return base + x - v;
}

__global__ __launch_bounds__(WG) void k_grid_count(const uint32_t *bits, long long nword, long long total,
                                                   int *block_tot)
{
    __shared__ int tmp[WG / 64];
    const long long w = (long long)blockIdx.x * WG + threadIdx.x;
    const int c = w < nword ? __popc(free_mask(bits, w, total)) : 0;
    const int ex = block_exclusive_scan(c, tmp);
    if (threadIdx.x == WG - 1) block_tot[blockIdx.x] = ex + c;
}

__global__ __launch_bounds__(WG) void k_grid_write(const uint32_t *bits, long long nword, long long total,
                                                   const long long *block_off, double s0, double s1, double s2,
                                                   double dx, int n1, int n2, float *grid)
{
    __shared__ int tmp[WG / 64];
    const long long w = (long long)blockIdx.x * WG + threadIdx.x;
    uint32_t f = w < nword ? free_mask(bits, w, total) : 0u;
    const int ex = block_exclusive_scan(__popc(f), tmp);
    long long o = block_off[blockIdx.x] + ex;
    while (f) {
        const int b = __ffs(f) - 1;
        f &= f - 1u;
        const long long lin = (w << 5) + b;
        const long long k = lin % n2;
        const long long j = (lin / n2) % n1;
        const long long i = lin / ((long long)n1 * n2);
        grid[3 * o + 0] = (float)((double)i * dx + s0);
        grid[3 * o + 1] = (float)((double)j * dx + s1);
        grid[3 * o + 2] = (float)((double)k * dx + s2);
        ++o;
    }
}

}  // namespace gcmc_grid

namespace gcmc_internal {
int set_error(int code, const std::string &msg);
}

extern "C" {

int gcmc_cavity_grid(int natom, const double *xyz, const double *startxyz, double grid_dx, const int *grid_n,
                     int max_points, float *grid, int *npoints)
{
    using namespace gcmc_grid;
    if (natom < 0 || (natom > 0 && !xyz) || !startxyz || !grid_n || !npoints)
        return gcmc_internal::set_error(GCMC_E_ARG, "bad arguments");
    if (!(grid_dx > 0.0)) return gcmc_internal::set_error(GCMC_E_ARG, "grid_dx must be positive");
    for (int i = 0; i < 3; ++i)
        if (grid_n[i] < 1) return gcmc_internal::set_error(GCMC_E_ARG, "grid_n must be positive");
    const long long total = (long long)grid_n[0] * grid_n[1] * grid_n[2];
    if (total > (1ll << 40)) return gcmc_internal::set_error(GCMC_E_ARG, "grid too large");
    const long long nword = (total + 31) >> 5;
    const long long nblock = (nword + WG - 1) / WG;
    if (nblock > (1ll << 30)) return gcmc_internal::set_error(GCMC_E_ARG, "grid too large");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return gcmc_internal::set_error(GCMC_E_DEVICE, "no HIP device");
    double *d_xyz = nullptr;
    uint32_t *d_bits = nullptr;
    int *d_tot = nullptr;
    long long *d_off = nullptr;
    float *d_grid = nullptr;
    auto done = [&](int rc, const std::string &m) {
        if (d_xyz) (void)hipFree(d_xyz);
        if (d_bits) (void)hipFree(d_bits);
        if (d_tot) (void)hipFree(d_tot);
        if (d_off) (void)hipFree(d_off);
        if (d_grid) (void)hipFree(d_grid);
        return rc ? gcmc_internal::set_error(rc, m) : 0;
    };
    hipError_t e = hipMalloc(&d_bits, nword * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&d_tot, nblock * sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&d_off, nblock * sizeof(long long));
    if (e == hipSuccess && natom > 0) e = hipMalloc(&d_xyz, 3 * (size_t)natom * sizeof(double));
    if (e == hipSuccess && natom > 0)
        e = hipMemcpy(d_xyz, xyz, 3 * (size_t)natom * sizeof(double), hipMemcpyHostToDevice);
    if (e != hipSuccess) return done(GCMC_E_HIP, std::string("grid: ") + hipGetErrorString(e));
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((nword + 255) / 256)), dim3(256), 0, 0, d_bits, (size_t)nword,
                       0u);
    if (natom > 0)
        hipLaunchKernelGGL(k_grid_mark, dim3((natom + 255) / 256), dim3(256), 0, 0, (const double *)d_xyz, natom,
                           startxyz[0], startxyz[1], startxyz[2], grid_dx, grid_n[0], grid_n[1], grid_n[2],
                           d_bits);
    hipLaunchKernelGGL(k_grid_count, dim3((unsigned)nblock), dim3(WG), 0, 0, (const uint32_t *)d_bits, nword, total,
                       d_tot);
    std::vector<int> tot(nblock);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpy(tot.data(), d_tot, nblock * sizeof(int), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return done(GCMC_E_HIP, std::string("grid: ") + hipGetErrorString(e));
    // exclusive offsets of the workgroups' free cells
    std::vector<long long> off(nblock);
    long long nfree = 0;
    for (long long b = 0; b < nblock; ++b) {
        off[b] = nfree;
        nfree += tot[b];
    }
    if (nfree > 0x7fffffffll) return done(GCMC_E_ARG, "more than 2^31 free grid points");
    *npoints = (int)nfree;
    if (!grid) return done(0, "");
    if (nfree > max_points) return done(GCMC_E_ARG, "grid buffer too small for the free cells");
    if (nfree > 0) {
        e = hipMalloc(&d_grid, 3 * (size_t)nfree * sizeof(float));
        if (e == hipSuccess) e = hipMemcpy(d_off, off.data(), nblock * sizeof(long long), hipMemcpyHostToDevice);
        if (e != hipSuccess) return done(GCMC_E_HIP, std::string("grid: ") + hipGetErrorString(e));
        hipLaunchKernelGGL(k_grid_write, dim3((unsigned)nblock), dim3(WG), 0, 0, (const uint32_t *)d_bits, nword,
                           total, (const long long *)d_off, startxyz[0], startxyz[1], startxyz[2], grid_dx,
                           grid_n[1], grid_n[2], d_grid);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpy(grid, d_grid, 3 * (size_t)nfree * sizeof(float), hipMemcpyDeviceToHost);
        if (e != hipSuccess) return done(GCMC_E_HIP, std::string("grid: ") + hipGetErrorString(e));
    }
    return done(0, "");
}

}  // extern "C"

// ---- counter calibration ------------------------------------------------------
// A plain device-to-device copy of a known size, for scaling the FETCH_SIZE /
// WRITE_SIZE counters of a profiler run (bytes read and written are exactly
// `bytes` per copy).
namespace gcmc_grid {
__global__ void k_copy_probe(const float4 *a, float4 *b, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}
}  // namespace gcmc_grid

extern "C" int gcmc_copy_probe(long long bytes, int copies)
{
    using namespace gcmc_grid;
    if (bytes < 16 || bytes % 16 != 0 || copies < 1) return gcmc_internal::set_error(GCMC_E_ARG, "bad copy size");
    float4 *a = nullptr, *b = nullptr;
    hipError_t e = hipMalloc(&a, (size_t)bytes);
    if (e == hipSuccess) e = hipMalloc(&b, (size_t)bytes);
    if (e == hipSuccess) e = hipMemset(a, 1, (size_t)bytes);
    const size_t n = (size_t)bytes / 16;
    for (int c = 0; c < copies && e == hipSuccess; ++c) {
        hipLaunchKernelGGL(k_copy_probe, dim3(4096), dim3(256), 0, 0, (const float4 *)a, b, n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return e == hipSuccess ? 0 : gcmc_internal::set_error(GCMC_E_HIP, std::string("copy probe: ") + hipGetErrorString(e));
}
