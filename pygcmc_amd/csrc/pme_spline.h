// pme_spline.h -- order-4 cardinal B-spline weights of the PME grid, shared by
// the PME kernels (spread / gather) and the engine's trial energies, so that
// both interpolate the same reciprocal potential.
#pragma once

#include <hip/hip_runtime.h>

namespace gcmc_spline {

// weights of the 4 grid points f, f-1, f-2, f-3 of scaled coordinate u in [0, K)
__device__ inline void spline4(float u, int K, int &f, float w[4])
{
    const float fl = floorf(u);
    f = (int)fl;
    const float t = u - fl;  // [0, 1)
    const float t2 = t * t, t3 = t2 * t;
    const float s = 1.0f - t;
    w[0] = t3 * (1.0f / 6.0f);                                         // M4(t)
    w[1] = (-3.0f * t3 + 3.0f * t2 + 3.0f * t + 1.0f) * (1.0f / 6.0f);  // M4(t+1)
    w[2] = (3.0f * t3 - 6.0f * t2 + 4.0f) * (1.0f / 6.0f);              // M4(t+2)
    w[3] = s * s * s * (1.0f / 6.0f);                                   // M4(t+3)
    f = ((f % K) + K) % K;
}

// scaled coordinate K * frac(x / L) in [0, K)
__device__ inline float scaled(float x, float L, int K)
{
    float fr = x / L;
    fr -= floorf(fr);
    const float u = fr * (float)K;
    return u >= (float)K ? 0.0f : u;
}

// interpolated potential of grid Phi [K0][K1][K2] at (x, y, z)
__device__ inline float gather(const float *Phi, const int K[3], const float L[3], float x, float y, float z)
{
    int f0, f1, f2;
    float w0[4], w1[4], w2[4];
    spline4(scaled(x, L[0], K[0]), K[0], f0, w0);
    spline4(scaled(y, L[1], K[1]), K[1], f1, w1);
    spline4(scaled(z, L[2], K[2]), K[2], f2, w2);
    float s = 0.0f;
    for (int i0 = 0; i0 < 4; ++i0) {
        int k0 = f0 - i0;
        k0 += k0 < 0 ? K[0] : 0;
        for (int i1 = 0; i1 < 4; ++i1) {
            int k1 = f1 - i1;
            k1 += k1 < 0 ? K[1] : 0;
            const float *row = Phi + ((size_t)k0 * K[1] + k1) * K[2];
            float r = 0.0f;
            for (int i2 = 0; i2 < 4; ++i2) {
                int k2 = f2 - i2;
                k2 += k2 < 0 ? K[2] : 0;
                r = fmaf(w2[i2], row[k2], r);
            }
            s = fmaf(w0[i0] * w1[i1], r, s);
        }
    }
    return s;
}

// Potential at x of a unit charge at y as the grid represents it:
// sum_{k, k'} W_x(k) W_y(k') Gk(k - k'), with Gk = IFFT(G) the periodic grid
// kernel.  The weights factor per axis, so the 64 x 64 terms collapse to
// 7 x 7 x 7 offsets d = (f_x - f_y) - (j - 3) with coefficients c_j =
// sum_{i - i' = j - 3} w_x[i] w_y[i'].  `j0` selects one slab of 49 offsets
// along the first axis (callers split the 7 slabs over threads).
__device__ inline float pair_kernel_slab(const float *Gk, const int K[3], const float L[3], const float x[3],
                                         const float y[3], int j0)
{
    int fx[3], fy[3];
    float wx[3][4], wy[3][4], c[3][7];
    for (int a = 0; a < 3; ++a) {
        spline4(scaled(x[a], L[a], K[a]), K[a], fx[a], wx[a]);
        spline4(scaled(y[a], L[a], K[a]), K[a], fy[a], wy[a]);
        for (int j = 0; j < 7; ++j) c[a][j] = 0.0f;
        for (int i = 0; i < 4; ++i)
            for (int ii = 0; ii < 4; ++ii) c[a][i - ii + 3] += wx[a][i] * wy[a][ii];
    }
    int d0 = (fx[0] - fy[0]) - (j0 - 3);
    d0 = ((d0 % K[0]) + K[0]) % K[0];
    float s = 0.0f;
    for (int j1 = 0; j1 < 7; ++j1) {
        int d1 = (fx[1] - fy[1]) - (j1 - 3);
        d1 = ((d1 % K[1]) + K[1]) % K[1];
        const float *row = Gk + ((size_t)d0 * K[1] + d1) * K[2];
        float r = 0.0f;
        for (int j2 = 0; j2 < 7; ++j2) {
            int d2 = (fx[2] - fy[2]) - (j2 - 3);
            d2 = ((d2 % K[2]) + K[2]) % K[2];
            r = fmaf(c[2][j2], row[d2], r);
        }
        s = fmaf(c[1][j1], r, s);
    }
    return c[0][j0] * s;
}

}  // namespace gcmc_spline
