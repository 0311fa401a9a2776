// adjoint_kernels.h -- interaction energy of given fragment poses with the
// environment of a replica, and its derivatives with respect to the fragment
// atom coordinates and to per-type parameters.  No counterpart in the
// reference (which has no gradients); the energy is calcEnergy's
// (gcmc/cpp/gcmc_energy.h:142-204: group-centre cutoff, LJ + Coulomb).
//
// Parameters per atom type t (all types of the pair table's column list):
//   eps[t]   CHARMM epsilon, kcal/mol   -> epsilon_t = 4.184 |eps[t]| kJ/mol
//   rmin[t]  CHARMM Rmin/2, A           -> sigma_t = rmin[t] 2^(5/6) / 10 nm
//   dq[t]    charge shift added to every atom of type t (fragment and
//            environment), e
// combined per pair as sigma = (sigma_a + sigma_b)/2, epsilon =
// sqrt(epsilon_a epsilon_b) (gcmc/python/simulation.py:63-72).  With dq = 0
// and the parameters the pair table was built from, the energy equals the
// engine's trial energy.
//
// One workgroup per pose: the candidate scan is eval_energy's (cell list,
// exact FP32 group test); the pair pass keeps per-fragment-atom gradient
// sums in registers (one lane per environment atom, fragment atoms
// unrolled), adds environment-type terms once per environment atom to LDS
// accumulators, and writes, per pose, dE/dx [nf][3] and the per-type
// derivatives [3][Y] (dq, eps, rmin).  FP32 arithmetic and accumulation;
// the energy keeps the engine's compensated sum.
#pragma once

#include "engine_kernels.h"

namespace gcmc {

constexpr double SIGMA_PER_RMINHALF_D = 0.17817974362806785;  // 2^(5/6) / 10
constexpr float SIGMA_PER_RMINHALF = (float)SIGMA_PER_RMINHALF_D;

struct AdjLds {
    float4 *tab;    // [X*Y] (100 sigma^2 hi, 4 epsilon, 100 sigma^2 lo, 0) from the per-type parameters
    float *dq;      // [Y]
    float *inv2e;   // [Y] 1 / (2 |eps|), 0 where eps = 0
    float *gq, *ge, *gr;  // [Y] per-type derivative sums of this pose
    float *red;     // [6*MAXF] per-fragment-atom sums
};

__host__ __device__ inline int adj_lds_bytes(int X, int Y)
{
    auto up = [](int x) { return (x + 15) & ~15; };
    return up(16 * X * Y) + 5 * up(4 * Y) + up(4 * 6 * MAXF);
}

__device__ inline AdjLds adj_carve(char *b, int X, int Y)
{
    auto up = [](int x) { return (x + 15) & ~15; };
    AdjLds A;
    A.tab = (float4 *)b; b += up(16 * X * Y);
    A.dq = (float *)b; b += up(4 * Y);
    A.inv2e = (float *)b; b += up(4 * Y);
    A.gq = (float *)b; b += up(4 * Y);
    A.ge = (float *)b; b += up(4 * Y);
    A.gr = (float *)b; b += up(4 * Y);
    A.red = (float *)b;
    return A;
}

template <int NF>
struct AdjPass {
    AdjLds A;
    float gx[NF], gy[NF], gz[NF], sr[NF], se[NF], sq[NF];

    template <int NT, bool STAGED>
    __device__ __attribute__((always_inline)) int run(CRep &C, const Lds &L, int n, int nf, float &acc,
                                                      float &cmp, double &acc64)
    {
        const int At = wg_scan<NT>(L.incn, n, L.misc);
        const float Lx = C.L[0], Ly = C.L[1], Lz = C.L[2];
        const float ix = C.invL[0], iy = C.invL[1], iz = C.invL[2];
        for (int t = threadIdx.x; t < At; t += NT) {
            const int g = bsearch_le(L.incn, n, t);
            const int a = L.incs[g] + (t - L.incn[g]);
            const float4 at = C.atom[a];
            const int ty = C.atype[a];
            const float qa = at.w + A.dq[ty];
            float tr = 0.f, te = 0.f, tq = 0.f;
#pragma unroll
            for (int f = 0; f < NF; ++f) {
                if (f < nf) {
                    const float4 fr = L.frag[f];
                    const float4 pp = A.tab[L.frow[f] + ty];
                    float ex, ey, ez;
                    float dx = two_diff(at.x, fr.x, ex);
                    float dy = two_diff(at.y, fr.y, ey);
                    float dz = two_diff(at.z, fr.z, ez);
                    dx = fmaf(-rintf(dx * ix), Lx, dx);
                    dy = fmaf(-rintf(dy * iy), Ly, dy);
                    dz = fmaf(-rintf(dz * iz), Lz, dz);
                    const float r2h = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                    const float r2c = fmaf(dz, ez, fmaf(dy, ey, dx * ex));
                    const float r2 = fmaf(2.0f, r2c, r2h);
                    float ir = __builtin_amdgcn_rsqf(r2);
                    ir = ir * fmaf(-0.5f * r2 * ir, ir, 1.5f);
                    const float ir2 = ir * ir;
                    const float rr = __builtin_amdgcn_rcpf(r2);
                    float s2 = pp.x * rr;
                    s2 = fmaf(fmaf(-s2, r2, pp.x) + pp.z, rr, s2);
                    const float s6 = s2 * s2 * s2;
                    const float s12 = s6 * s6;
                    const float lj = pp.y * (s12 - s6);
                    const float el = (fr.w * qa) * ir;
                    const float term = lj + el;
                    const float s = acc + term;
                    const float bp = s - acc;
                    const float err = (acc - (s - bp)) + (term - bp);
                    acc = s;
                    cmp += err;
                    // dE/d(r^2); with d = x_a - x_f, dE/dx_f = -2 dE/d(r^2) d
                    const float dr2 = (pp.y * fmaf(-6.0f, s12, 3.0f * s6) - 0.5f * el) * ir2;
                    const float gs = -2.0f * dr2;
                    gx[f] = fmaf(gs, dx + ex, gx[f]);
                    gy[f] = fmaf(gs, dy + ey, gy[f]);
                    gz[f] = fmaf(gs, dz + ez, gz[f]);
                    // dE/dsigma_pair = 4 eps (12 s^12 - 6 s^6) / sigma_pair, 1/sigma = 10/sqrt(pp.x);
                    // dsigma_pair/drmin = 2^(5/6)/20 on each side
                    const float isg = pp.x > 0.0f ? 10.0f * __builtin_amdgcn_rsqf(pp.x) : 0.0f;
                    const float ds = pp.y * fmaf(12.0f, s12, -6.0f * s6) * isg * (0.5f * SIGMA_PER_RMINHALF);
                    sr[f] += ds;
                    tr += ds;
                    // dE/deps[t] = lj / (2 |eps[t]|) on each side
                    se[f] += lj;
                    te += lj;
                    // dE/dq: K q_a / r to the fragment side (K applied later), K q_f / r here
                    sq[f] = fmaf(qa, ir, sq[f]);
                    tq = fmaf(fr.w, ir, tq);
                }
            }
            atomicAdd(&A.gr[ty], tr);
            atomicAdd(&A.ge[ty], te * A.inv2e[ty]);
            atomicAdd(&A.gq[ty], tq);
        }
        __syncthreads();
        return At;
    }
};

// poses [npose][nf][3], centres [npose][3]; outputs E [npose] (FP64),
// dEdx [npose][nf][3], partial [npose][3][Y] (dq, eps, rmin)
template <int NT, int NF>
__global__ __launch_bounds__(NT) void k_adjoint(const Rep *reps, int r, int spi, const float *centres,
                                               const float *poses, const float *eps_kcal,
                                               const float *rmin_half, const float *dq, double *E,
                                               float *dEdx, float *partial, LdsLayout lay, int adj_off)
{
    extern __shared__ __align__(16) char smem[];
    const Lds L = carve(smem, lay);
    CRep &C = rep_of(reps, r);
    const int X = C.X, Y = C.Y;
    AdjPass<NF> P;
    P.A = adj_carve(smem + adj_off, X, Y);
    CSpecies &F = species_of(C, spi);
    const int nf = F.nf, i = blockIdx.x, tid = threadIdx.x, lane = tid & 63;
    for (int t = tid; t < Y; t += NT) {
        const float ek = fabsf(eps_kcal[t]);
        P.A.inv2e[t] = ek > 0.0f ? 0.5f / ek : 0.0f;
        P.A.dq[t] = dq ? dq[t] : 0.0f;
    }
    for (int q = tid; q < 6 * MAXF; q += NT) P.A.red[q] = 0.0f;
    load_species_table(C, L);
    // pair entries in FP64, rounded once to FP32 (as a host-built table is)
    for (int q = tid; q < X * Y; q += NT) {
        const int x = q / Y, y = q - x * Y;
        const double sg = 0.5 * ((double)rmin_half[x] + (double)rmin_half[y]) * (double)SIGMA_PER_RMINHALF_D;
        const double ee = 4.184 * 4.184 * fabs((double)eps_kcal[x]) * fabs((double)eps_kcal[y]);
        const double s2 = 100.0 * sg * sg;
        const float hi = (float)s2;
        P.A.tab[q] = make_float4(hi, (float)(4.0 * sqrt(ee)), (float)(s2 - (double)hi), 0.0f);
    }
    if (tid < nf) {
        const float *p = poses + ((size_t)i * nf + tid) * 3;
        const float kq = dq ? KCOUL_F * (F.tmpl[tid].w + dq[F.ttype[tid]]) : F.kq[tid];
        L.frag[tid] = make_float4(p[0], p[1], p[2], kq);
        L.frow[tid] = F.trow[tid];
    }
    __syncthreads();
    for (int t = tid; t < Y; t += NT) {
        P.A.gq[t] = 0.0f;
        P.A.ge[t] = 0.0f;
        P.A.gr[t] = 0.0f;
    }
#pragma unroll
    for (int f = 0; f < NF; ++f) {
        P.gx[f] = P.gy[f] = P.gz[f] = 0.0f;
        P.sr[f] = P.se[f] = P.sq[f] = 0.0f;
    }
    __syncthreads();
    double e;
    int n, na;
    uint32_t h;
    eval_energy_pass<NT, false>(C, centres[3 * i], centres[3 * i + 1], centres[3 * i + 2], nf, -1, L, P, &e, &n, &h,
                                &na, nullptr);
#pragma unroll
    for (int f = 0; f < NF; ++f) {
        if (f < nf) {
            float v[6] = {P.gx[f], P.gy[f], P.gz[f], P.sr[f], P.se[f], P.sq[f]};
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                float x = v[k];
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
                if (lane == 0) atomicAdd(&P.A.red[6 * f + k], x);
            }
        }
    }
    __syncthreads();
    if (tid < nf) {
        const int tf = F.ttype[tid];
        const float *rd = P.A.red + 6 * tid;
        float *o = dEdx + ((size_t)i * nf + tid) * 3;
        o[0] = rd[0];
        o[1] = rd[1];
        o[2] = rd[2];
        atomicAdd(&P.A.gr[tf], rd[3]);
        atomicAdd(&P.A.ge[tf], rd[4] * P.A.inv2e[tf]);
        atomicAdd(&P.A.gq[tf], rd[5] * KCOUL_F);
    }
    __syncthreads();
    float *pp = partial + (size_t)i * 3 * Y;
    for (int t = tid; t < Y; t += NT) {
        pp[t] = P.A.gq[t];
        pp[Y + t] = P.A.ge[t];
        pp[2 * Y + t] = P.A.gr[t];
    }
    if (tid == 0) E[i] = e;
}

// tmp[c][j] = sum over poses p of chunk c of w[p] partial[p][j] (FP64);
// block (64 columns x 16 pose lanes), grid (column blocks, chunks)
__global__ __launch_bounds__(1024) void k_adjoint_colsum(int npose, int m, const float *partial, const float *w,
                                                        double *tmp)
{
    __shared__ double red[16][64];
    const int j = blockIdx.x * 64 + (threadIdx.x & 63);
    const int pl = threadIdx.x >> 6;
    const int nch = gridDim.y;
    const int per = (npose + nch - 1) / nch;
    const int p0 = blockIdx.y * per, p1 = min(npose, p0 + per);
    double s = 0.0;
    if (j < m)
        for (int p = p0 + pl; p < p1; p += 16) s += (double)(w ? w[p] : 1.0f) * (double)partial[(size_t)p * m + j];
    red[pl][threadIdx.x & 63] = s;
    __syncthreads();
    if (pl == 0 && j < m) {
        double t = 0.0;
        for (int q = 0; q < 16; ++q) t += red[q][threadIdx.x & 63];
        tmp[(size_t)blockIdx.y * m + j] = t;
    }
}

__global__ void k_adjoint_chunksum(int nch, int m, const double *tmp, float *out)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m) return;
    double t = 0.0;
    for (int c = 0; c < nch; ++c) t += tmp[(size_t)c * m + j];
    out[j] = (float)t;
}

// out[p][k] = w[p] x[p][k]
__global__ void k_scale_rows(int npose, int per, const float *x, const float *w, float *out)
{
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (size_t)npose * per) return;
    const int p = (int)(q / per);
    out[q] = (w ? w[p] : 1.0f) * x[q];
}

}  // namespace gcmc
