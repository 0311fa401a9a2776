// pme.hip -- smooth particle-mesh Ewald, reciprocal part (Essmann et al.,
// J. Chem. Phys. 103, 8577 (1995)), for an orthorhombic box.  The reference
// declares a PME flag (gcmc/cpp/gcmc.h:91) that no kernel reads; these
// kernels give the engine's optional Ewald electrostatics.
//
// Units as calcEnergy (K = 1388.431112 kJ/mol A e^-2):
//   Q(k)   = sum_i q_i W_i(k)        order-4 cardinal B-spline weights W
//   G(m)   = K/(pi V) exp(-pi^2 |m~|^2/beta^2)/|m~|^2 B(m),  G(0) = 0
//            m~ = (m1'/L1, m2'/L2, m3'/L3), m' the signed frequency,
//            B(m) = prod_i 1/|sum_{k=0}^{2} M4(k+1) e^{2 pi i m_i k/K_i}|^2
//   Phi    = IFFT(G . FFT(Q))         (unnormalised inverse)
//   E_rec  = 1/2 sum_k Q(k) Phi(k) = 1/2 sum_m G(m) |FFT(Q)(m)|^2
//   phi(r) = sum_k W_r(k) Phi(k)      reciprocal potential at r (per unit charge)
//
// The charge grid accumulates in 32.32 fixed point (int64 atomics): sums are
// exact and order independent, so a run is reproducible and removing a
// charge (spreading -q at the same point) cancels it exactly.
//
// Kernels (HBM / atomic bound; no MFMA -- not a dense contraction):
//   k_pme_spread    one thread per charge, 64 integer atomic adds to Q
//   k_pme_tofloat   fixed-point grid -> FP32 grid for the FFT
//   k_pme_convolve  one thread per half-spectrum entry: energy term
//                   (FP64 block sums) and X *= G
//   k_pme_gather    one thread per point, 64 reads of Phi
// The FFTs are hipFFT R2C / C2R plans on the caller's stream.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <cmath>
#include <string>
#include <vector>

#include "../../include/pygcmc_hip.h"
#include "pme_spline.h"

namespace gcmc_internal {
int set_error(int code, const std::string &msg);
}

namespace gcmc_pme_k {

constexpr double KC = 1388.431112;

struct Geo {
    int K[3];
    float L[3];
};

using gcmc_spline::spline4;
using gcmc_spline::scaled;

constexpr double FIX = 4294967296.0;   // 2^32

// n charges, or *count of them when count is not null (n = capacity)
__global__ void k_pme_spread(int n, const int *count, const float4 *xyzq, Geo g, unsigned long long *Q)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || (count && i >= *count)) return;
    const float4 a = xyzq[i];
    if (a.w == 0.0f) return;
    int f0, f1, f2;
    float w0[4], w1[4], w2[4];
    spline4(scaled(a.x, g.L[0], g.K[0]), g.K[0], f0, w0);
    spline4(scaled(a.y, g.L[1], g.K[1]), g.K[1], f1, w1);
    spline4(scaled(a.z, g.L[2], g.K[2]), g.K[2], f2, w2);
    for (int i0 = 0; i0 < 4; ++i0) {
        int k0 = f0 - i0;
        k0 += k0 < 0 ? g.K[0] : 0;
        for (int i1 = 0; i1 < 4; ++i1) {
            int k1 = f1 - i1;
            k1 += k1 < 0 ? g.K[1] : 0;
            const float w01 = a.w * w0[i0] * w1[i1];
            unsigned long long *row = Q + ((size_t)k0 * g.K[1] + k1) * g.K[2];
            for (int i2 = 0; i2 < 4; ++i2) {
                int k2 = f2 - i2;
                k2 += k2 < 0 ? g.K[2] : 0;
                const long long v = __double2ll_rn((double)(w01 * w2[i2]) * FIX);
                atomicAdd(row + k2, (unsigned long long)v);
            }
        }
    }
}

__global__ void k_pme_tofloat(const unsigned long long *Qi, float *Q, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) Q[i] = (float)((double)(long long)Qi[i] * (1.0 / FIX));
}

__global__ void k_pme_zero64(unsigned long long *p, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

__global__ void k_pme_zero_count(int *c)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *c = 0;
}

// X: K0 x K1 x (K2/2+1) half spectrum; G same layout (real)
__global__ __launch_bounds__(256) void k_pme_convolve(size_t nh, int K2h, int K2, const float *G, hipfftComplex *X,
                                                      double *energy)
{
    __shared__ double red[4];
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    double e = 0.0;
    if (i < nh) {
        const float gm = G[i];
        hipfftComplex x = X[i];
        const int m2 = (int)(i % (size_t)K2h);
        // entries 0 < m2 < K2/2 stand for two conjugate frequencies
        const double mult = (m2 == 0 || 2 * m2 == K2) ? 1.0 : 2.0;
        e = 0.5 * mult * (double)gm * ((double)x.x * x.x + (double)x.y * x.y);
        x.x *= gm;
        x.y *= gm;
        X[i] = x;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) e += __shfl_down(e, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = e;
    __syncthreads();
    if (threadIdx.x == 0 && energy) atomicAdd(energy, red[0] + red[1] + red[2] + red[3]);
}

__global__ void k_pme_gather(int n, const float *xyz, Geo g, const float *Phi, float *phi)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    phi[i] = gcmc_spline::gather(Phi, g.K, g.L, xyz[3 * i + 0], xyz[3 * i + 1], xyz[3 * i + 2]);
}

__global__ void k_zero_d(double *p)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0.0;
}

}  // namespace gcmc_pme_k

struct gcmc_pme {
    gcmc_pme_k::Geo geo{};
    float beta = 0.0f;
    size_t ngrid = 0, nhalf = 0;
    unsigned long long *Qi = nullptr;
    float *Q = nullptr, *Phi = nullptr, *G = nullptr;
    float *Gk = nullptr;   // real-space grid kernel IFFT(G): Phi = Gk (*) Q
    hipfftComplex *X = nullptr;
    hipfftHandle fwd = 0, inv = 0;
    bool have_fwd = false, have_inv = false;
};

namespace {

using namespace gcmc_pme_k;

// |b(m)|^2 of the order-4 spline for frequency m of K points
double bmod2(int m, int K)
{
    const double M4[3] = {1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0};
    double re = 0.0, im = 0.0;
    for (int k = 0; k < 3; ++k) {
        const double a = 2.0 * M_PI * (double)m * k / K;
        re += M4[k] * cos(a);
        im += M4[k] * sin(a);
    }
    return 1.0 / (re * re + im * im);
}

int pme_hip(hipError_t e, const char *what)
{
    return gcmc_internal::set_error(GCMC_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

int pme_fft(hipfftResult r, const char *what)
{
    return gcmc_internal::set_error(GCMC_E_HIP, std::string(what) + ": hipFFT status " + std::to_string((int)r));
}

}  // namespace

extern "C" {

int gcmc_pme_destroy(gcmc_pme *p)
{
    if (!p) return 0;
    if (p->have_fwd) hipfftDestroy(p->fwd);
    if (p->have_inv) hipfftDestroy(p->inv);
    if (p->Qi) (void)hipFree(p->Qi);
    if (p->Q) (void)hipFree(p->Q);
    if (p->Gk) (void)hipFree(p->Gk);
    if (p->Phi) (void)hipFree(p->Phi);
    if (p->G) (void)hipFree(p->G);
    if (p->X) (void)hipFree(p->X);
    delete p;
    return 0;
}

int gcmc_pme_create(gcmc_pme **out, const int *grid, const float *box, float beta)
{
    if (!out || !grid || !box) return gcmc_internal::set_error(GCMC_E_ARG, "null argument");
    *out = nullptr;
    for (int i = 0; i < 3; ++i) {
        if (grid[i] < 4 || grid[i] > 4096) return gcmc_internal::set_error(GCMC_E_ARG, "PME grid must be 4..4096");
        if (!(box[i] > 0.0f)) return gcmc_internal::set_error(GCMC_E_ARG, "box edges must be positive");
    }
    if (!(beta > 0.0f)) return gcmc_internal::set_error(GCMC_E_ARG, "beta must be positive");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return gcmc_internal::set_error(GCMC_E_DEVICE, "no HIP device");
    gcmc_pme *p = new gcmc_pme;
    for (int i = 0; i < 3; ++i) {
        p->geo.K[i] = grid[i];
        p->geo.L[i] = box[i];
    }
    p->beta = beta;
    const int K0 = grid[0], K1 = grid[1], K2 = grid[2], K2h = K2 / 2 + 1;
    p->ngrid = (size_t)K0 * K1 * K2;
    p->nhalf = (size_t)K0 * K1 * K2h;
    hipError_t e = hipMalloc(&p->Qi, p->ngrid * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMalloc(&p->Q, p->ngrid * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&p->Gk, p->ngrid * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&p->Phi, p->ngrid * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&p->G, p->nhalf * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&p->X, p->nhalf * sizeof(hipfftComplex));
    if (e != hipSuccess) {
        gcmc_pme_destroy(p);
        return pme_hip(e, "PME buffers");
    }
    // influence function on the half spectrum (host, FP64, once)
    std::vector<float> G(p->nhalf);
    const double V = (double)box[0] * box[1] * box[2];
    const double b2 = (double)beta * beta;
    std::vector<double> B0(K0), B1(K1), B2(K2h);
    for (int m = 0; m < K0; ++m) B0[m] = bmod2(m, K0);
    for (int m = 0; m < K1; ++m) B1[m] = bmod2(m, K1);
    for (int m = 0; m < K2h; ++m) B2[m] = bmod2(m, K2);
    for (int a = 0; a < K0; ++a) {
        const double ma = (a <= K0 / 2 ? a : a - K0) / (double)box[0];
        for (int b = 0; b < K1; ++b) {
            const double mb = (b <= K1 / 2 ? b : b - K1) / (double)box[1];
            for (int c = 0; c < K2h; ++c) {
                const double mc = c / (double)box[2];
                const double m2 = ma * ma + mb * mb + mc * mc;
                double g = 0.0;
                if (a || b || c) g = KC / (M_PI * V) * exp(-M_PI * M_PI * m2 / b2) / m2 * B0[a] * B1[b] * B2[c];
                G[((size_t)a * K1 + b) * K2h + c] = (float)g;
            }
        }
    }
    e = hipMemcpy(p->G, G.data(), p->nhalf * sizeof(float), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        gcmc_pme_destroy(p);
        return pme_hip(e, "PME influence function");
    }
    hipfftResult r = hipfftPlan3d(&p->fwd, K0, K1, K2, HIPFFT_R2C);
    if (r != HIPFFT_SUCCESS) {
        gcmc_pme_destroy(p);
        return pme_fft(r, "PME forward plan");
    }
    p->have_fwd = true;
    r = hipfftPlan3d(&p->inv, K0, K1, K2, HIPFFT_C2R);
    if (r != HIPFFT_SUCCESS) {
        gcmc_pme_destroy(p);
        return pme_fft(r, "PME inverse plan");
    }
    p->have_inv = true;
    // grid kernel: the potential grid of a unit charge at grid point 0
    std::vector<hipfftComplex> Gc(p->nhalf);
    for (size_t i = 0; i < p->nhalf; ++i) {
        Gc[i].x = G[i];
        Gc[i].y = 0.0f;
    }
    e = hipMemcpy(p->X, Gc.data(), p->nhalf * sizeof(hipfftComplex), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        gcmc_pme_destroy(p);
        return pme_hip(e, "PME grid kernel");
    }
    r = hipfftExecC2R(p->inv, p->X, p->Gk);
    if (r == HIPFFT_SUCCESS && hipDeviceSynchronize() != hipSuccess) r = HIPFFT_EXEC_FAILED;
    if (r != HIPFFT_SUCCESS) {
        gcmc_pme_destroy(p);
        return pme_fft(r, "PME grid kernel");
    }
    *out = p;
    return 0;
}

int gcmc_pme_clear(gcmc_pme *p, void *stream)
{
    if (!p) return gcmc_internal::set_error(GCMC_E_ARG, "null PME plan");
    hipLaunchKernelGGL(k_pme_zero64, dim3((unsigned)((p->ngrid + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       p->Qi, p->ngrid);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : pme_hip(e, "PME clear");
}

int gcmc_pme_spread(gcmc_pme *p, int n, const float *xyzq, void *stream)
{
    if (!p || n < 0 || (n > 0 && !xyzq)) return gcmc_internal::set_error(GCMC_E_ARG, "bad PME spread arguments");
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_pme_spread, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, n, (const int *)nullptr,
                       (const float4 *)xyzq, p->geo, p->Qi);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : pme_hip(e, "PME spread");
}

int gcmc_pme_solve(gcmc_pme *p, double *energy, void *stream)
{
    if (!p) return gcmc_internal::set_error(GCMC_E_ARG, "null PME plan");
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_pme_tofloat, dim3((unsigned)((p->ngrid + 255) / 256)), dim3(256), 0, s,
                       (const unsigned long long *)p->Qi, p->Q, p->ngrid);
    hipfftResult r = hipfftSetStream(p->fwd, s);
    if (r == HIPFFT_SUCCESS) r = hipfftSetStream(p->inv, s);
    if (r == HIPFFT_SUCCESS) r = hipfftExecR2C(p->fwd, p->Q, p->X);
    if (r != HIPFFT_SUCCESS) return pme_fft(r, "PME forward FFT");
    if (energy) hipLaunchKernelGGL(k_zero_d, dim3(1), dim3(64), 0, s, energy);
    const int K2h = p->geo.K[2] / 2 + 1;
    hipLaunchKernelGGL(k_pme_convolve, dim3((unsigned)((p->nhalf + 255) / 256)), dim3(256), 0, s, p->nhalf, K2h,
                       p->geo.K[2], (const float *)p->G, p->X, energy);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return pme_hip(e, "PME convolve");
    r = hipfftExecC2R(p->inv, p->X, p->Phi);
    if (r != HIPFFT_SUCCESS) return pme_fft(r, "PME inverse FFT");
    return 0;
}

int gcmc_pme_gather(gcmc_pme *p, int n, const float *xyz, float *phi, void *stream)
{
    if (!p || n < 0 || (n > 0 && (!xyz || !phi))) return gcmc_internal::set_error(GCMC_E_ARG, "bad PME gather arguments");
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_pme_gather, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, n, xyz, p->geo,
                       (const float *)p->Phi, phi);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : pme_hip(e, "PME gather");
}

}  // extern "C"

namespace gcmc_internal {
// for the engine's Ewald mode: potential grid, grid shape, spread of a
// device-counted list, counter reset
const float *pme_phi(const gcmc_pme *p) { return p->Phi; }
const float *pme_kernel(const gcmc_pme *p) { return p->Gk; }
void pme_grid(const gcmc_pme *p, int K[3])
{
    for (int i = 0; i < 3; ++i) K[i] = p->geo.K[i];
}
int pme_spread_counted(gcmc_pme *p, int cap, const int *count, const float4 *xyzq, int *reset, hipStream_t s)
{
    if (cap > 0)
        hipLaunchKernelGGL(gcmc_pme_k::k_pme_spread, dim3((cap + 255) / 256), dim3(256), 0, s, cap, count, xyzq, p->geo,
                           p->Qi);
    if (reset) hipLaunchKernelGGL(gcmc_pme_k::k_pme_zero_count, dim3(1), dim3(64), 0, s, reset);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : pme_hip(e, "PME spread");
}
}  // namespace gcmc_internal
