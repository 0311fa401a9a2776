// gcmc_math.h -- arithmetic definitions shared by host and device code of the
// MI355X GCMC engine: counter-based variates, the rotation of a rigid
// fragment, and the FP32 group-inclusion test of pyGCMC.
//
// Every function here is a fixed sequence of IEEE FP32 operations (explicit
// fmaf, no contraction: the library is compiled with -ffp-contract=off), so
// the host and every GPU produce identical bits for a given input.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define GC_HD __host__ __device__ __forceinline__

namespace gcmc {

// ---- Philox4x32-10 (Salmon, Moraes, Dror, Shaw; SC'11) -------------------
struct u32x4 { uint32_t x, y, z, w; };

GC_HD uint32_t mulhi32(uint32_t a, uint32_t b)
{
#ifdef __HIP_DEVICE_COMPILE__
    return __umulhi(a, b);
#else
    return (uint32_t)(((uint64_t)a * b) >> 32);
#endif
}

GC_HD u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1)
{
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0 = mulhi32(M0, c.x), lo0 = M0 * c.x;
        uint32_t hi1 = mulhi32(M1, c.z), lo1 = M1 * c.z;
        u32x4 n = {hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
        c = n;
        k0 += W0;
        k1 += W1;
    }
    return c;
}

// Variates of one system: key (seed, 0), counter (step, index, purpose, 0).
enum Purpose : uint32_t { P_JITTER = 0, P_ROTATION = 1, P_PICK = 2, P_DECIDE = 3 };

GC_HD u32x4 draw(uint32_t seed, uint32_t step, uint32_t index, uint32_t purpose)
{
    return philox4x32_10(u32x4{step, index, purpose, 0u}, seed, 0u);
}

// uniform on (0,1] with 24 bits (the range of curand_uniform)
GC_HD float u01f(uint32_t x) { return (float)((x >> 8) + 1u) * 0x1p-24f; }
// uniform on (0,1) with 52 bits
GC_HD double u01d(uint32_t a, uint32_t b)
{
    uint64_t m = ((uint64_t)a << 20) | (uint64_t)(b >> 12);
    return (double)(2u * m + 1u) * 0x1p-53;
}
// integer in [0, n)
GC_HD uint32_t below(uint32_t x, uint32_t n) { return mulhi32(x, n); }

// ---- sin/cos of the rotation angle ---------------------------------------
// Cody-Waite reduction by pi/2 then Taylor polynomials to r^9 / r^10 on
// |r| <= pi/4 (truncation < 3e-9 relative); a few ulp, fully deterministic.
GC_HD void sincos_f(float x, float* s, float* c)
{
    float k = rintf(x * 0x1.45f306p-1f);
    int q = (int)k;
    float r = fmaf(-k, 0x1.921fb6p+0f, x);
    r = fmaf(-k, -0x1.777a5cp-25f, r);
    float r2 = r * r;
    float ps = fmaf(r2, 0x1.71de3ap-19f, -0x1.a01a02p-13f);
    ps = fmaf(r2, ps, 0x1.111112p-7f);
    ps = fmaf(r2, ps, -0x1.555556p-3f);
    float sr = fmaf(r * r2, ps, r);
    float pc = fmaf(r2, -0x1.27e4fcp-22f, 0x1.a01a02p-16f);
    pc = fmaf(r2, pc, -0x1.6c16c2p-10f);
    pc = fmaf(r2, pc, 0x1.555556p-5f);
    pc = fmaf(r2, pc, -0.5f);
    float cr = fmaf(r2, pc, 1.0f);
    switch (q & 3) {
    case 0: *s = sr; *c = cr; break;
    case 1: *s = cr; *c = -sr; break;
    case 2: *s = -sr; *c = -cr; break;
    default: *s = -cr; *c = sr; break;
    }
}

// Rodrigues matrix of a (not normalised) axis and an angle, row major.
GC_HD void rotation_matrix(float a0, float a1, float a2, float phi, float R[9])
{
    float nrm = sqrtf(fmaf(a2, a2, fmaf(a1, a1, a0 * a0)));
    float n0 = a0 / nrm, n1 = a1 / nrm, n2 = a2 / nrm;
    float s, c;
    sincos_f(phi, &s, &c);
    float C = 1.0f - c;
    R[0] = fmaf(n0 * n0, C, c);
    R[1] = fmaf(n0 * n1, C, -(n2 * s));
    R[2] = fmaf(n0 * n2, C, n1 * s);
    R[3] = fmaf(n1 * n0, C, n2 * s);
    R[4] = fmaf(n1 * n1, C, c);
    R[5] = fmaf(n1 * n2, C, -(n0 * s));
    R[6] = fmaf(n2 * n0, C, -(n1 * s));
    R[7] = fmaf(n2 * n1, C, n0 * s);
    R[8] = fmaf(n2 * n2, C, c);
}

GC_HD void rotate(const float R[9], float x0, float x1, float x2, float* y)
{
    y[0] = fmaf(x2, R[2], fmaf(x1, R[1], x0 * R[0]));
    y[1] = fmaf(x2, R[5], fmaf(x1, R[4], x0 * R[3]));
    y[2] = fmaf(x2, R[8], fmaf(x1, R[7], x0 * R[6]));
}

// ---- group inclusion test -------------------------------------------------
// pyGCMC keeps a group when distanceP(centre, group centre) <= cutoff in FP32
// with image shift fast_round(d/L) = (int)(q +- 0.5f).  Per axis the shift is
// a step function of |d|: thresholds t1 < t2 < t3 (host-computed by bisection
// over floats) give the same shift without the IEEE division, and r2 <= r2max
// (largest float whose sqrtf is <= cutoff) replaces the square root.
struct AxisCut {
    float L;       // box edge
    float t[3];    // |d| >= t[0]: shift 1, |d| >= t[1]: shift 2; t[2]: end of table
};

// reference semantics, literally (gcmc_energy.h:110-112)
GC_HD float fast_round_ref(float a) { return a >= 0.0f ? (float)(int)(a + 0.5f) : (float)(int)(a - 0.5f); }

// d - fast_round(d/L)*L without the division for |d| < t[2]
GC_HD float axis_delta(float x, float y, const AxisCut& a)
{
    float d = x - y;
    float m = fabsf(d);
    float r;
    if (m < a.t[2]) {
        r = (m >= a.t[0] ? 1.0f : 0.0f) + (m >= a.t[1] ? 1.0f : 0.0f);
        r = d < 0.0f ? -r : r;
    } else {
        r = fast_round_ref(d / a.L);
    }
    float t = r * a.L;
    return d - t;
}

GC_HD float dist2_p(float x0, float x1, float x2, float y0, float y1, float y2,
                    const AxisCut* ax)
{
    float d0 = axis_delta(x0, y0, ax[0]);
    float d1 = axis_delta(x1, y1, ax[1]);
    float d2 = axis_delta(x2, y2, ax[2]);
    return fmaf(d2, d2, fmaf(d1, d1, d0 * d0));
}

GC_HD bool within(float x0, float x1, float x2, float y0, float y1, float y2,
                  const AxisCut* ax, float r2max)
{
    return dist2_p(x0, x1, x2, y0, y1, y2, ax) <= r2max;
}

// order-independent hash of an included-group set, and a pose fingerprint
GC_HD uint32_t group_hash(uint32_t g)
{
    uint32_t h = (g + 1u) * 0x9E3779B1u;
    h ^= h >> 15;
    h *= 0x85EBCA77u;
    h ^= h >> 13;
    return h;
}

}  // namespace gcmc
