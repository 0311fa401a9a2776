// engine_kernels.h -- device phases of one GCMC move and the kernels that
// schedule them.
//
// A move of one replica is three phases, each run by one workgroup of NT
// threads (NT = 256, 512 or 1024, a tuning choice):
//   prep    decode moveArray[step]; pick slots (partial Fisher-Yates);
//           trial centres                      (gcmc_move_*.h host side)
//   item    one trial evaluation: build the pose, scan the cell list,
//           group test, pair energies          (Gmove_*, calcEnergy)
//   accept  grouping of trials, selection, Metropolis in FP64, commit
//           (move_add/del/trn/rot host logic, Gupdate*)
// The plain form launches item and accept kernels per move; the batched form
// runs every replica in one persistent kernel where workgroups of an XCD
// claim items of that XCD's replicas and the workgroup finishing a replica's
// last item runs its accept and prep phases (no grid-wide barrier, replicas
// advance independently).
#pragma once

#include "engine_types.h"
#include "pme_spline.h"

namespace gcmc {

// The buffers a Rep points to are all in device global memory, but a pointer
// read out of a struct is generic to the compiler, and generic (flat) loads
// count against the LDS counter as well: a wait for an LDS read then also
// waits for every outstanding flat load.  gbl() restates the address space so
// the hot loads are global_load / global_store.  (The host pass only parses
// device functions; address space 1 means nothing there.)
#if defined(__HIP_DEVICE_COMPILE__)
template <class T>
using gptr = __attribute__((address_space(1))) T *;
#else
template <class T>
using gptr = T *;
#endif
template <class T>
__device__ __forceinline__ gptr<T> gbl(T *p)
{
    return (gptr<T>)p;
}

// A replica's Rep and species table are never written while kernels run.
// Read through the constant address space, a field at a wave-uniform address
// is a scalar load (scalar data cache) rather than a vector load through the
// texture path that the evaluations' gathers keep busy.
#if defined(__HIP_DEVICE_COMPILE__)
using CRep = const __attribute__((address_space(4))) Rep;
using CSpecies = const __attribute__((address_space(4))) Species;
#else
using CRep = const Rep;
using CSpecies = const Species;
#endif
__device__ __forceinline__ CRep &rep_of(const Rep *reps, int r) { return ((CRep *)reps)[r]; }
__device__ __forceinline__ CSpecies &species_of(CRep &C, int s) { return ((CSpecies *)C.sp)[s]; }

// scratch ints of a workgroup (Lds::misc)
enum Misc : int {
    M_SCAN = 0,      // [0, NW]   scan partials, total at NW
    M_NIN = 40,      // [40, 56)  per-wave included count
    M_HASH = 56,     // [56, 72)  per-wave set hash
    M_ACC = 80,      // accept: decision / count of accepted
    M_NG = 81,       //         groups
    M_N = 82,        //         species count after the move
    M_CNT = 83,      // batched: items claimed (consecutive, from M_ITEM)
    M_INV = 84,      //          the claim needs an L1 invalidate (cu_acquire_needed)
    M_FOUND = 88,    // batched: replica claimed (-1 none, -2 exit)
    M_ITEM = 89,     //          item claimed
    M_LAST = 90,     //          this workgroup finished the move's last item
    M_NX = 91,       // [91, 97)  next move, loaded during the accept phase:
                     //           end step, move, species count, conf, maxNum, startRes
};

struct Lds {
    int *smob, *sfst, *sfix, *soff, *incs, *incn;
    float4 *frag;
    int *frow;
    double4 *fragd;   // fragment pose minus the trial centre (x, y, z), K q (w), FP64
    float4 *pair;
    int4 *spt;
    int *misc;
    double *dred;
    int *members;
    double *w, *val;
    int2 *accl;
    uint32_t *adjl;   // adjacency rows staged for the accept phase (null: read C.adj)
    double *rsel, *racc, *ea;  // accept: selection / acceptance draw per group, exp term per trial
    int *J;
    int inc_cap;
    uint16_t *cand, *amap;  // index maps expanded by wg_scan (valid when the total <= map_cap)
    int map_cap;
};

__device__ inline Lds carve(char *b, const LdsLayout &l)
{
    Lds s;
    s.smob = (int *)(b + l.smob);
    s.sfst = (int *)(b + l.sfst);
    s.sfix = (int *)(b + l.sfix);
    s.soff = (int *)(b + l.soff);
    s.incs = (int *)(b + l.incs);
    s.incn = (int *)(b + l.incn);
    s.frag = (float4 *)(b + l.frag);
    s.frow = (int *)(b + l.frow);
    s.fragd = (double4 *)(b + l.fragd);
    s.pair = (float4 *)(b + (l.pair >= 0 ? l.pair : 0));
    s.spt = (int4 *)(b + l.spt);
    s.misc = (int *)(b + l.misc);
    s.dred = (double *)(b + l.dred);
    s.members = (int *)(b + l.members);
    s.w = (double *)(b + l.w);
    s.val = (double *)(b + l.val);
    s.accl = (int2 *)(b + l.accl);
    s.adjl = l.adjl >= 0 ? (uint32_t *)(b + l.adjl) : nullptr;
    s.rsel = (double *)(b + l.rsel);
    s.racc = (double *)(b + l.racc);
    s.ea = (double *)(b + l.ea);
    s.J = (int *)(b + l.J);
    s.inc_cap = l.inc_cap;
    s.cand = (uint16_t *)(b + l.cand);
    s.amap = (uint16_t *)(b + l.amap);
    s.map_cap = l.map_cap;
    return s;
}

__device__ inline void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// ---- wave-wide scans on DPP (every lane of the wave active) -----------------
// Inclusive scan: Hillis-Steele within each row of 16 lanes (row_shr 1, 2, 4,
// 8), then row 0's last lane into row 1 and row 2's into row 3 (row_bcast:15),
// then lane 31 into rows 2 and 3 (row_bcast:31).  Lanes without a source keep
// the identity.  Register moves between lanes instead of ds_bpermute round
// trips through the LDS crossbar.
enum : int { DPP_SHR1 = 0x111, DPP_SHR2 = 0x112, DPP_SHR4 = 0x114, DPP_SHR8 = 0x118, DPP_BC15 = 0x142, DPP_BC31 = 0x143 };

template <int CTRL, int ROWS>
__device__ __forceinline__ int dpp_i(int identity, int v)
{
    return __builtin_amdgcn_update_dpp(identity, v, CTRL, ROWS, 0xf, false);
}

template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_d(double identity, double v)
{
    const long long b = __double_as_longlong(v), o = __double_as_longlong(identity);
    const int lo = __builtin_amdgcn_update_dpp((int)o, (int)b, CTRL, ROWS, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(b >> 32), CTRL, ROWS, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

template <class T, class Op>
__device__ __forceinline__ T wave_scan(T x, T id, Op op)
{
    if constexpr (sizeof(T) == 8) {
        x = op(x, dpp_d<DPP_SHR1, 0xf>(id, x));
        x = op(x, dpp_d<DPP_SHR2, 0xf>(id, x));
        x = op(x, dpp_d<DPP_SHR4, 0xf>(id, x));
        x = op(x, dpp_d<DPP_SHR8, 0xf>(id, x));
        x = op(x, dpp_d<DPP_BC15, 0xa>(id, x));
        x = op(x, dpp_d<DPP_BC31, 0xc>(id, x));
    } else {
        x = op(x, (T)dpp_i<DPP_SHR1, 0xf>((int)id, (int)x));
        x = op(x, (T)dpp_i<DPP_SHR2, 0xf>((int)id, (int)x));
        x = op(x, (T)dpp_i<DPP_SHR4, 0xf>((int)id, (int)x));
        x = op(x, (T)dpp_i<DPP_SHR8, 0xf>((int)id, (int)x));
        x = op(x, (T)dpp_i<DPP_BC15, 0xa>((int)id, (int)x));
        x = op(x, (T)dpp_i<DPP_BC31, 0xc>((int)id, (int)x));
    }
    return x;
}

__device__ __forceinline__ int wave_scan_add(int x)
{
    return wave_scan(x, 0, [](int a, int b) { return a + b; });
}
__device__ __forceinline__ uint32_t wave_scan_add(uint32_t x)
{
    return wave_scan(x, 0u, [](uint32_t a, uint32_t b) { return a + b; });
}
__device__ __forceinline__ double wave_scan_add(double x)
{
    return wave_scan(x, 0.0, [](double a, double b) { return a + b; });
}
// minimum over the wave, in every lane
__device__ __forceinline__ double wave_min(double x)
{
    const double m = wave_scan(x, __builtin_huge_val(), [](double a, double b) { return b < a ? b : a; });
    const long long bm = __double_as_longlong(m);
    const int lo = __builtin_amdgcn_readlane((int)bm, 63), hi = __builtin_amdgcn_readlane((int)(bm >> 32), 63);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ unsigned wave_min(unsigned x)
{
    const unsigned m = wave_scan(x, 0xFFFFFFFFu, [](unsigned a, unsigned b) { return b < a ? b : a; });
    return (unsigned)__builtin_amdgcn_readlane((int)m, 63);
}

// 64-bit wave values by lane index (idx uniform across the wave)
__device__ inline double lane_d(double v, int idx)
{
    const int i = __builtin_amdgcn_readfirstlane(idx);
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), i);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), i);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ inline unsigned long long lane_u64(unsigned long long v, int idx)
{
    const int i = __builtin_amdgcn_readfirstlane(idx);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)v, i);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(v >> 32), i);
    return ((unsigned long long)hi << 32) | lo;
}
__device__ inline int lane_i(int v, int idx)
{
    return __builtin_amdgcn_readlane(v, __builtin_amdgcn_readfirstlane(idx));
}
// lane r receives v of the lane whose dst is r (dst a permutation of 0..63)
__device__ inline double permute_d(int dst, double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_ds_permute(dst * 4, (int)(b & 0xffffffffll));
    const int hi = __builtin_amdgcn_ds_permute(dst * 4, (int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// exclusive scan of a[0..n) in LDS, a[n] = total; returns total
template <int NT>
// When map is given and the total is at most cap, also writes map[j] = i for
// every j in [a'[i], a'[i+1]) (a' the scanned array): the entry each item
// belongs to, so a thread finds it with one read instead of a binary search.
__device__ inline int wg_scan(int *a, int n, int *misc, uint16_t *map = nullptr, int cap = 0)
{
    constexpr int NW = NT / 64;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int per = (n + NT - 1) / NT;
    const int b = tid * per;
    const int e = min(n, b + per);
    int sum = 0;
    for (int i = b; i < e; ++i) sum += a[i];
    const int x = wave_scan_add(sum);
    if (lane == 63) misc[M_SCAN + wave] = x;
    __syncthreads();
    if (tid == 0) {
        int t = 0;
        for (int w = 0; w < NW; ++w) {
            int v = misc[M_SCAN + w];
            misc[M_SCAN + w] = t;
            t += v;
        }
        misc[M_SCAN + NW] = t;
    }
    __syncthreads();
    int base = misc[M_SCAN + wave] + x - sum;
    const int total = misc[M_SCAN + NW];
    const bool expand = map != nullptr && total <= cap;
    for (int i = b; i < e; ++i) {
        int v = a[i];
        a[i] = base;
        if (expand)
            for (int j = 0; j < v; ++j) map[base + j] = (uint16_t)i;
        base += v;
    }
    if (tid == 0) a[n] = total;
    __syncthreads();
    return total;
}

// largest e in [0, n) with a[e] <= m (a[0] = 0 <= m < a[n])
__device__ inline int bsearch_le(const int *a, int n, int m)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (a[mid] <= m) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// bsearch_le for two keys at once (independent LDS reads in flight together)
__device__ inline void bsearch_le2(const int *a, int n, int m0, int m1, int &g0, int &g1)
{
    int lo0 = 0, hi0 = n - 1, lo1 = 0, hi1 = n - 1;
    while (lo0 < hi0 || lo1 < hi1) {
        const int mid0 = (lo0 + hi0 + 1) >> 1, mid1 = (lo1 + hi1 + 1) >> 1;
        const int v0 = a[mid0], v1 = a[mid1];
        if (lo0 < hi0) {
            if (v0 <= m0) lo0 = mid0;
            else hi0 = mid0 - 1;
        }
        if (lo1 < hi1) {
            if (v1 <= m1) lo1 = mid1;
            else hi1 = mid1 - 1;
        }
    }
    g0 = lo0;
    g1 = lo1;
}

// a - b = s + e exactly (Knuth TwoSum of a and -b); returns s
__device__ inline float two_diff(float a, float b, float &e)
{
    const float s = a - b;
    const float bv = s - a;
    const float av = s - bv;
    e = (a - av) - (b + bv);
    return s;
}

// ---- pair energies of the buffered included groups ------------------------
// lane <-> environment atom (float4 load, coalesced within a group), loop
// over the fragment atoms held in LDS; FP64 terms into acc64 (default), or
// FP32 terms with TwoSum into (acc, cmp) (fp32_coulomb and Ewald modes).
// (cxd, cyd, czd): the trial centre; wrap: image shift per pair.
template <int NT, bool STAGED, bool EW>
__device__ inline int pair_pass(CRep &C, const Lds &L, int n, int nf, float &acc, float &cmp, double &acc64,
                                double cxd, double cyd, double czd, bool wrap)
{
    const int A = wg_scan<NT>(L.incn, n, L.misc, L.amap, L.map_cap);
    const bool mapped = A <= L.map_cap;
    const auto atom = gbl(C.atom);
    const auto atype = gbl(C.atype);
    const auto GPT = gbl(C.pair);
    const float Lx = C.L[0], Ly = C.L[1], Lz = C.L[2];
    const float ix = C.invL[0], iy = C.invL[1], iz = C.invL[2];
    const float eb = C.ew_beta;
    const bool c64 = !EW && C.coul64 != 0;
    double accd = 0.0;
    // one environment atom against every fragment atom; a thread takes atoms
    // t, t + NT, t + 2 NT, ... in that order (the order of its sums)
    auto pair_row = [&](const int f, const int ty) {
        if constexpr (STAGED)
            return L.pair[L.frow[f] + ty];
        else
            return GPT[L.frow[f] + ty];
    };
    // Default mode: every pair term in FP64.  The environment atom's
    // position relative to the trial centre, image-shifted once per atom, and
    // the fragment atom's (L.fragd) are exact differences of FP32 positions
    // (to 2^-53); their difference needs no further image shift when
    // Rep::nowrap holds, else it takes one per pair.  1/r: the FP32 estimate
    // (relative error ~2^-22) and one FP64 Newton step (~1.5 2^-44).
    const double Lxd = Lx, Lyd = Ly, Lzd = Lz, ixd = ix, iyd = iy, izd = iz;
    auto atom_terms64 = [&](const float4 at, const int ty) {
        double ax = (double)at.x - cxd, ay = (double)at.y - cyd, az = (double)at.z - czd;
        ax = fma(-rint(ax * ixd), Lxd, ax);
        ay = fma(-rint(ay * iyd), Lyd, ay);
        az = fma(-rint(az * izd), Lzd, az);
        const double qa = (double)at.w;
        // the table entry of fragment atom f + 1 is fetched while atom f's
        // terms are computed (the last fetch repeats atom nf - 1's)
        float4 pn = pair_row(0, ty);
        for (int f = 0; f < nf; ++f) {
            const double4 F = L.fragd[f];
            const float4 pp = pn;
            pn = pair_row(min(f + 1, nf - 1), ty);
            double ux = ax - F.x, uy = ay - F.y, uz = az - F.z;
            if (wrap) {
                ux = fma(-rint(ux * ixd), Lxd, ux);
                uy = fma(-rint(uy * iyd), Lyd, uy);
                uz = fma(-rint(uz * izd), Lzd, uz);
            }
            const double r2 = fma(uz, uz, fma(uy, uy, ux * ux));
            double y = (double)__builtin_amdgcn_rsqf((float)r2);
            y = y * fma(-0.5 * r2 * y, y, 1.5);
            const double s2 = ((double)pp.x + (double)pp.z) * (y * y);   // 100 sigma^2 / r^2
            const double s6 = s2 * s2 * s2;
            accd = fma((double)pp.y, fma(s6, s6, -s6), accd);            // 4 eps (s^12 - s^6)
            accd = fma(F.w * qa, y, accd);                                // K q_f q_a / r
        }
    };
    // fp32_coulomb and Ewald modes: FP32 terms, TwoSum into (acc, cmp)
    auto atom_terms = [&](const float4 at, const int ty) {
        // the table entry of fragment atom f + 1 is fetched while atom f's
        // terms are computed (the last fetch repeats atom nf - 1's)
        float4 pn = pair_row(0, ty);
        for (int f = 0; f < nf; ++f) {
            const float4 fr = L.frag[f];
            const float4 pp = pn;
            pn = pair_row(min(f + 1, nf - 1), ty);
            // displacement as an exact sum s + e (TwoSum); the image shift
            // s - n L is exact for |n| <= 1 (Sterbenz)
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
            ir = ir * fmaf(-0.5f * r2 * ir, ir, 1.5f);     // one Newton step
            // s^2 = 100 sigma^2 / r^2 to ~0.5 ulp: 100 sigma^2 held exactly as
            // pp.x + pp.z, quotient with one residual correction.  s^12
            // multiplies the relative error of s^2 by six, and a rounded
            // table value would err alike on every pair of a type pair.
            const float rr = __builtin_amdgcn_rcpf(r2);
            float s2 = pp.x * rr;
            s2 = fmaf(fmaf(-s2, r2, pp.x) + pp.z, rr, s2);
            const float s6 = s2 * s2 * s2;
            float lj = pp.y * fmaf(s6, s6, -s6);           // 4 eps (s^12 - s^6)
            if (s2 > 1.0f) {
                // inside sigma s^12 dominates and carries six times the
                // relative error of r^2: this rare pair's LJ term in FP64,
                // from the exact displacement
                const double ux = (double)dx + (double)ex, uy = (double)dy + (double)ey,
                             uz = (double)dz + (double)ez;
                const double r2d = fma(uz, uz, fma(uy, uy, ux * ux));
                double q = __builtin_amdgcn_rcp(r2d);
                q = q * fma(-r2d, q, 2.0);
                q = q * fma(-r2d, q, 2.0);
                const double s2d = ((double)pp.x + (double)pp.z) * q;
                const double s6d = s2d * s2d * s2d;
                accd = fma((double)pp.y, fma(s6d, s6d, -s6d), accd);
                lj = 0.0f;
            }
            float el = (fr.w * at.w) * ir;                  // K q_f q_a / r
            if (EW) el *= erfcf(eb * (r2 * ir));            // Ewald real-space part
            const float term = lj + el;
            const float s = acc + term;                     // TwoSum
            const float bp = s - acc;
            const float err = (acc - (s - bp)) + (term - bp);
            acc = s;
            cmp += err;
        }
    };
    // U atoms per thread per round, all loads issued before any is used: a
    // thread takes atoms t, t + NT, t + 2 NT, ... in that order (the order of
    // its sums, whatever U is)
    constexpr int U = 3;
    for (int t0 = threadIdx.x; t0 < A; t0 += U * NT) {
        int a[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int tu = t0 + u * NT;
            const int tq = tu < A ? tu : t0;
            int g;
            if (mapped) {
                g = L.amap[tq];
            } else {
                int gx;
                bsearch_le2(L.incn, n, tq, tq, g, gx);
            }
            a[u] = L.incs[g] + (tq - L.incn[g]);
        }
        float4 at[U];
        int ty[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            at[u] = atom[a[u]];
            ty[u] = atype[a[u]];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (t0 + u * NT < A) {
                if (c64) atom_terms64(at[u], ty[u]);
                else atom_terms(at[u], ty[u]);
            }
        }
    }
    __syncthreads();
    acc64 += accd;
    return A;
}

// the energy pass of eval_energy
template <bool EW>
struct EnergyPass {
    double cx, cy, cz;   // trial centre
    bool wrap;           // image shift per pair (else per environment atom)
    template <int NT, bool STAGED>
    __device__ int run(CRep &C, const Lds &L, int n, int nf, float &acc, float &cmp, double &acc64)
    {
        return pair_pass<NT, STAGED, EW>(C, L, n, nf, acc, cmp, acc64, cx, cy, cz, wrap);
    }
};

// species table in LDS (mobile group id -> atom range)
__device__ inline void load_species_table(CRep &C, const Lds &L)
{
    for (int s = threadIdx.x; s < C.T; s += blockDim.x) {
        CSpecies &F = species_of(C, s);
        L.spt[s] = make_int4(F.startRes, F.maxNum, F.atomBase, F.nf);
    }
}

// the replica's pair table (X x Y entries) into LDS, straight from global
// memory (global_load_lds: no registers); the caller waits (vmcnt) and
// synchronises before the first read
template <int NT>
__device__ inline void stage_pair_table(CRep &C, const Lds &L)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int n = C.X * C.Y;
    const auto P = gbl(C.pair);
    for (int b = wave * 64; b < n; b += NT)
        if (b + lane < n) __builtin_amdgcn_global_load_lds((const void *)(P + b + lane), (void *)&L.pair[b], 16, 0, 0);
}

// ---- calcEnergy over the cell list -----------------------------------------
// Candidates: fixed and mobile groups of the stencil cells around the test
// centre plus the overflow bucket; kept when gid != exclude and the FP32
// distanceP test passes (gcmc_energy.h:152-191).  Needs L.spt loaded.
// Thread 0 receives the results.
template <int NT, bool STAGED, class Pass>
__device__ __attribute__((always_inline)) inline void eval_energy_pass(CRep &C, float cx, float cy, float cz,
                                                                       int nf, int exclude, const Lds &L, Pass &P,
                                                                       double *outE, int *outN, uint32_t *outH,
                                                                       int *outA, unsigned long long *ticks)
{
    unsigned long long tk0 = wall_clock64(), tpair = 0;
    constexpr int NW = NT / 64;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const auto stencil = gbl(C.stencil);
    const auto cellv = gbl(C.cell);
    const auto fix = gbl(C.fix);
    const auto mob_rec = gbl(C.mob_rec);
    const int cc = cell_index(C.L, C.start, C.inv_h, C.nc, cx, cy, cz);
    const int c2 = cc % C.nc[2];
    const int c1 = (cc / C.nc[2]) % C.nc[1];
    const int c0 = cc / (C.nc[1] * C.nc[2]);
    // the centre's position inside its cell
    float u[3];
    {
        const int ci[3] = {c0, c1, c2};
        cell_offset(C.L, C.start, C.cell_h, ci, cx, cy, cz, u);
    }
    const int ne = C.nst + 1;
    for (int e = tid; e < ne; e += NT) {
        int cell, fs, fc, mc;
        if (e < C.nst) {
            const int code = stencil[e];
            const int d[3] = {(code & 255) - 128, ((code >> 8) & 255) - 128, ((code >> 16) & 255) - 128};
            // cells whose box lies beyond rc + slack from the centre hold no
            // group centre that passes the test: no loads for them
            if (!stencil_cell_near(u, d, C.cell_h, C.prune, C.prune_r2)) {
                fs = 0;
                fc = 0;
                mc = 0;
                L.smob[e] = 0;
            } else {
                // stencil offsets span at most one box per axis: wrap by one add
                int i0 = c0 + d[0];
                int i1 = c1 + d[1];
                int i2 = c2 + d[2];
                i0 += i0 < 0 ? C.nc[0] : (i0 >= C.nc[0] ? -C.nc[0] : 0);
                i1 += i1 < 0 ? C.nc[1] : (i1 >= C.nc[1] ? -C.nc[1] : 0);
                i2 += i2 < 0 ? C.nc[2] : (i2 >= C.nc[2] ? -C.nc[2] : 0);
                cell = (i0 * C.nc[1] + i1) * C.nc[2] + i2;
                const int2 ci = cellv[cell];
                fs = ci.x;
                fc = cellv[cell + 1].x - fs;
                mc = ci.y;
                L.smob[e] = cell * C.cap;
            }
        } else {
            fs = 0;
            fc = 0;
            mc = cellv[C.ncell].y;
            L.smob[e] = C.ovf_base;
        }
        L.sfst[e] = fs;
        L.sfix[e] = fc;
        L.soff[e] = fc + mc;
    }
    __syncthreads();
    const int M = wg_scan<NT>(L.soff, ne, L.misc, L.cand, L.map_cap);
    const bool mapped = M <= L.map_cap;
    const unsigned long long tk1 = wall_clock64();
    float acc = 0.0f, cmp = 0.0f;
    double acc64 = 0.0;  // FP64 pair terms (Coulomb)
    int nin = 0, natom = 0;
    uint32_t hs = 0;
    // Candidates in chunks of R rounds of NT.  A thread first issues the
    // record loads of its R candidates of the chunk (one memory round trip
    // for all of them), then tests them in order; each wave appends the
    // groups it includes to its own segment of SEG = 64 R slots, so no
    // barrier is needed within a chunk; after a chunk the unused slots get
    // zero atoms and one pair pass covers all segments (groups in segment
    // order).
    constexpr int R = INC_CAP / NT, SEG = 64 * R;
    for (int base = 0; base < M; base += R * NT) {
        float4 rec[R];
        int2 rng[R];
        bool fixed[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int m = base + r * NT + tid;
            rec[r] = make_float4(0.f, 0.f, 0.f, 0.f);
            rng[r] = make_int2(0, 0);
            fixed[r] = false;
            if (m < M) {
                const int e = mapped ? (int)L.cand[m] : bsearch_le(L.soff, ne, m);
                const int local = m - L.soff[e];
                const int fc = L.sfix[e];
                if (local < fc) {
                    const int q = L.sfst[e] + local;
                    rec[r] = fix[q].p;
                    rng[r] = fix[q].rng;
                    fixed[r] = true;
                } else {
                    rec[r] = mob_rec[L.smob[e] + (local - fc)];
                }
            }
        }
        int cnt = 0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int m = base + r * NT + tid;
            bool inc = false;
            int2 rr = rng[r];
            if (m < M) {
                const float4 p = rec[r];
                int gid;
                if (fixed[r]) {
                    gid = __float_as_int(p.w);
                } else {
                    const int word = __float_as_int(p.w);
                    gid = word & GID_MASK;
                    const int4 sp = L.spt[word >> GID_BITS];   // (startRes, maxNum, atomBase, nf)
                    rr = make_int2(sp.z + (gid - sp.x) * sp.w, sp.w);
                }
                inc = (gid != exclude) && within(cx, cy, cz, p.x, p.y, p.z, C.ax, C.r2max);
                if (inc) {
                    nin += 1;
                    hs += group_hash((uint32_t)gid);
                }
            }
            const unsigned long long bal = __ballot(inc);
            if (inc) {
                const int pos = wave * SEG + cnt + __popcll(bal & ((1ull << lane) - 1ull));
                L.incs[pos] = rr.x;
                L.incn[pos] = rr.y;
            }
            cnt += __popcll(bal);
        }
        for (int q = cnt + lane; q < SEG; q += 64) L.incn[wave * SEG + q] = 0;
        // a staged pair table was loaded into LDS asynchronously (eval_item):
        // every wave's copies have landed once the barrier passes
        if (STAGED) asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
        const unsigned long long tp0 = wall_clock64();
        natom += P.template run<NT, STAGED>(C, L, INC_CAP, nf, acc, cmp, acc64);
        tpair += wall_clock64() - tp0;
    }
    const double v = wave_scan_add((double)acc + (double)cmp + acc64);
    nin = wave_scan_add(nin);
    hs = wave_scan_add(hs);
    if (lane == 63) {
        L.dred[wave] = v;
        L.misc[M_NIN + wave] = nin;
        L.misc[M_HASH + wave] = (int)hs;
    }
    __syncthreads();
    if (tid == 0) {
        double E = 0.0;
        int N = 0;
        uint32_t H = 0;
        for (int w = 0; w < NW; ++w) {
            E += L.dred[w];
            N += L.misc[M_NIN + w];
            H += (uint32_t)L.misc[M_HASH + w];
        }
        *outE = E;
        *outN = N;
        *outH = H;
        *outA = natom;
        if (ticks) {
            const unsigned long long tk2 = wall_clock64();
            ticks[1] = tk1 - tk0;
            ticks[2] = (tk2 - tk1) - tpair;
            ticks[3] = tpair;
        }
    }
    __syncthreads();
}

// pose_near: the pose in L.frag was generated from the species template or
// the molecule's atoms (within the species' extent of the centre, the bound
// Rep::nowrap rests on); poses given by a caller take the per-pair image shift.
// Thread f reads L.frag[f] (written by thread f, or before a barrier) and
// writes L.fragd[f]; eval_energy_pass has a barrier before its first pair
// pass.
template <int NT, bool STAGED, bool EW = false>
__device__ __attribute__((always_inline)) inline void eval_energy(CRep &C, float cx, float cy, float cz, int nf,
                                                                  int exclude, const Lds &L, double *outE, int *outN,
                                                                  uint32_t *outH, int *outA,
                                                                  unsigned long long *ticks = nullptr,
                                                                  bool pose_near = false)
{
    EnergyPass<EW> P{(double)cx, (double)cy, (double)cz, !(pose_near && C.nowrap != 0)};
    const int tid = threadIdx.x;
    if (tid < nf) {
        const float4 q = L.frag[tid];
        L.fragd[tid] = make_double4((double)q.x - P.cx, (double)q.y - P.cy, (double)q.z - P.cz, (double)q.w);
    }
    eval_energy_pass<NT, STAGED>(C, cx, cy, cz, nf, exclude, L, P, outE, outN, outH, outA, ticks);
}

__device__ inline uint32_t pose_hash(const float4 *p, int nf)
{
    uint32_t h = 2166136261u;
    for (int a = 0; a < nf; ++a) {
        h = (h ^ __float_as_uint(p[a].x)) * 16777619u;
        h = (h ^ __float_as_uint(p[a].y)) * 16777619u;
        h = (h ^ __float_as_uint(p[a].z)) * 16777619u;
    }
    return h;
}

__device__ inline bool tracing(CRep &C, int step, int &row)
{
    row = step - C.tr.first;
    return C.tr.k != nullptr && row >= 0 && row < C.tr.max_steps;
}

// Replica data is read and written by workgroups of one XCD (batched form)
// or by consecutive kernels (plain form).  Counters shared by workgroups use
// L2 atomics of that XCD (LOCAL) or device-wide atomics.
template <bool LOCAL, class T>
__device__ inline void count_add(T *p, T v)
{
    if (LOCAL) __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    else __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// statistics of one evaluation (counters of the replica's State)
struct ItemStats {
    unsigned long long ticks[4];
    unsigned long long nat, nin, nf;
};

template <bool LOCAL>
__device__ inline void add_item_stats(State &W, const ItemStats &s)
{
    for (int i = 0; i < 4; ++i) count_add<LOCAL>(&W.t_sub[i], s.ticks[i]);
    count_add<LOCAL>(&W.n_eval, 1ull);
    count_add<LOCAL>(&W.n_pair, s.nat * s.nf);
    count_add<LOCAL>(&W.n_atom, s.nat);
    count_add<LOCAL>(&W.n_group, s.nin);
    count_add<LOCAL>(&W.n_fatom, s.nf);
}

// ---- Ewald reciprocal term of a pose ------------------------------------------
// sum_f q_f phi(x_f) over the pose held in L.frag, phi the PME potential of the
// current state.  When the molecule itself is in that state (base >= 0: its
// atoms y_g at C.atom[base..]), its own share of phi -- what the grid holds
// for its charges, periodic images and neutralising background included --
// is removed: sum_{f,g} q_f q_g Kgrid(x_f, y_g) (pme_spline.h).  What remains
// is the reciprocal interaction with every other charge; for a rigid
// molecule the Ewald self, intramolecular and self-image terms are constants
// of the species, left out alike on insertion and deletion.  Called by every
// thread; thread 0 gets the sum.
template <int NT>
__device__ inline double ewald_recip(CRep &C, CSpecies &F, const Lds &L, int nf, int base)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    double v = 0.0;
    if (tid < nf) {
        const float4 x = L.frag[tid];
        v = (double)F.tmpl[tid].w * (double)gcmc_spline::gather(C.pme_phi, C.pk, C.L, x.x, x.y, x.z);
    }
    if (base >= 0) {
        const int nw = nf * nf * 7;
        for (int w = tid; w < nw; w += NT) {
            const int f = w / (7 * nf), g = (w / 7) % nf, j0 = w % 7;
            const float4 xf = L.frag[f];
            const float4 yg = C.atom[base + g];
            const float xv[3] = {xf.x, xf.y, xf.z}, yv[3] = {yg.x, yg.y, yg.z};
            const float kv = gcmc_spline::pair_kernel_slab(C.pme_gk, C.pk, C.L, xv, yv, j0);
            v -= (double)F.tmpl[f].w * (double)F.tmpl[g].w * (double)kv;
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if (lane == 0) L.dred[wave] = v;
    __syncthreads();
    double t = 0.0;
    if (tid == 0)
        for (int q = 0; q < NT / 64; ++q) t += L.dred[q];
    __syncthreads();
    return t;
}

// ---- one trial evaluation ----------------------------------------------------
// item < k: trial item (insert/delete) or E_old (translate/rotate);
// item >= k: E_new of trial item-k (translate/rotate).
template <int NT, bool STAGED, bool LOCAL, bool EW>
// Thread 0 adds the statistics to S, or leaves them in *defer (thread 0's
// copy) for the caller to add once the item's completion is published.
__device__ inline void eval_item(CRep &C, const State &S, int item, const Lds &L, bool stage_rep = true,
                                 bool stage_tab = true, ItemStats *defer = nullptr)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const unsigned long long tq0 = wall_clock64();
    const int kind = S.kind, spi = S.sp, k = S.k, step = S.step;
    CSpecies &F = species_of(C, spi);
    const int nf = F.nf;
    const auto tc = gbl(C.tc);
    const auto atom = gbl(C.atom);
    // translate / rotate: item 2t is trial t's old pose, 2t + 1 its new pose;
    // a claim of consecutive items evaluates both, the second reading the
    // neighbourhood the first left in L1
    int t = item, phase = 0;
    if (kind >= K_TRN) {
        t = item >> 1;
        phase = item & 1;
    }
    // the species table stays in LDS while a workgroup keeps evaluating
    // items of one replica (stage_rep = false), the pair table while it
    // keeps evaluating items of replicas with that table (stage_tab = false)
    if (stage_rep) load_species_table(C, L);
    // the pair table is copied without a wait here: the pose, stencil and
    // candidate work runs meanwhile, and the wait is before the first pair
    // pass (eval_energy_pass)
    if (STAGED && stage_tab) stage_pair_table<NT>(C, L);
    const float4 tcv = tc[t];
    const int slot = (kind == K_INS) ? -1 : gbl(C.tslot)[t];
    const int gid = (kind == K_INS) ? -1 : F.startRes + slot;
    const float4 cold = (gid >= 0) ? gbl(C.res_pos)[gid] : make_float4(0.f, 0.f, 0.f, 0.f);
    float cx, cy, cz;
    if (kind == K_INS || phase == 1 || kind == K_DEL) {
        cx = tcv.x; cy = tcv.y; cz = tcv.z;
    } else {
        cx = cold.x; cy = cold.y; cz = cold.z;
    }
    const bool newpose = (kind == K_INS) || (kind == K_DEL) || (phase == 1);
    // the first block of trial centres for the adjacency row below, loaded
    // now so the load overlaps the pose work
    const bool adjrow = kind != K_DEL && phase == 0;
    float4 cj0 = make_float4(0.f, 0.f, 0.f, 0.f);
    if (adjrow && tid < k) cj0 = tc[tid];
    if (tid < nf) {
        const int a = tid;
        const float4 tm = F.tmpl[a];
        float px, py, pz;
        if (kind == K_INS) {
            const u32x4 y = draw(C.seed, (uint32_t)step, (uint32_t)t, P_ROTATION);
            float R[9], v[3];
            rotation_matrix(u01f(y.x), u01f(y.y), u01f(y.z), u01f(y.w) * 0x1.921fb6p+2f, R);
            rotate(R, tm.x, tm.y, tm.z, v);
            px = v[0] + cx; py = v[1] + cy; pz = v[2] + cz;
        } else {
            const int base = F.atomBase + slot * nf;
            const float4 cur = atom[base + a];
            if (kind == K_DEL || phase == 0) {
                px = cur.x; py = cur.y; pz = cur.z;
            } else if (kind == K_TRN) {
                const u32x4 x = draw(C.seed, (uint32_t)step, (uint32_t)t, P_JITTER);
                const float u[3] = {u01f(x.x), u01f(x.y), u01f(x.z)};
                const float co[3] = {cold.x, cold.y, cold.z};
                float r[3];
                for (int i = 0; i < 3; ++i) {
                    float d = u[i] * C.dx;
                    if (co[i] + d > C.L[i] + C.start[i]) d = d - C.L[i];
                    if (co[i] + d < C.start[i]) d = d + C.L[i];
                    r[i] = d;
                }
                px = cur.x + r[0]; py = cur.y + r[1]; pz = cur.z + r[2];
            } else {
                float cen0 = 0.f, cen1 = 0.f, cen2 = 0.f;
                const float fn = (float)nf;
                for (int b = 0; b < nf; ++b) {
                    const float4 q = atom[base + b];
                    cen0 = cen0 + q.x / fn;
                    cen1 = cen1 + q.y / fn;
                    cen2 = cen2 + q.z / fn;
                }
                const u32x4 y = draw(C.seed, (uint32_t)step, (uint32_t)t, P_ROTATION);
                float R[9], v[3];
                rotation_matrix(u01f(y.x), u01f(y.y), u01f(y.z), u01f(y.w) * 0x1.921fb6p+2f, R);
                rotate(R, cur.x - cen0, cur.y - cen1, cur.z - cen2, v);
                px = v[0] + cen0; py = v[1] + cen1; pz = v[2] + cen2;
            }
        }
        L.frag[a] = make_float4(px, py, pz, F.kq[a]);
        L.frow[a] = F.trow[a];
        if (kind == K_INS || phase == 1) gbl(C.pose)[t * MAXF + a] = make_float4(px, py, pz, 0.f);
    }
    __syncthreads();

    // adjacency row of trial t over the trial centres (grouping of the accept phase)
    if (adjrow) {
        for (int j0 = 0; j0 < k; j0 += NT) {
            const int j = j0 + tid;
            bool b = false;
            if (j < k && j != t) {
                const float4 cj = (j0 == 0) ? cj0 : tc[j];
                b = within(tcv.x, tcv.y, tcv.z, cj.x, cj.y, cj.z, C.ax, C.r2max);
            }
            const unsigned long long bal = __ballot(b);
            if (lane == 0) {
                const int wi = (j0 + wave * 64) >> 5;
                if (wi < ADJW) gbl(C.adj)[t * ADJW + wi] = (uint32_t)bal;
                if (wi + 1 < ADJW) gbl(C.adj)[t * ADJW + wi + 1] = (uint32_t)(bal >> 32);
            }
        }
    }

    double E;
    int nin, nat;
    uint32_t hh;
    unsigned long long ticks[4];
    ticks[0] = wall_clock64() - tq0;
    eval_energy<NT, STAGED, EW>(C, cx, cy, cz, nf, gid, L, &E, &nin, &hh, &nat, ticks, true);
    if (EW) {
        // the moving molecule (delete, translate, rotate) is in the PME grid
        const double er = ewald_recip<NT>(C, F, L, nf, kind == K_INS ? -1 : F.atomBase + slot * nf);
        if (tid == 0) E += er;
    }
    if (tid == 0) {
        gbl(C.E)[phase * k + t] = E;
        ItemStats st;
        for (int i = 0; i < 4; ++i) st.ticks[i] = ticks[i];
        st.nat = (unsigned long long)nat;
        st.nin = (unsigned long long)nin;
        st.nf = (unsigned long long)nf;
        if (defer) *defer = st;
        else add_item_stats<LOCAL>(const_cast<State &>(S), st);
        int row;
        if (tracing(C, step, row) && t < C.tr.max_k) {
            const int q = row * C.tr.max_k + t;
            C.tr.e[2 * q + phase] = E;
            C.tr.ninc[2 * q + phase] = nin;
            C.tr.inchash[2 * q + phase] = hh;
            if (newpose) {
                C.tr.posehash[q] = pose_hash(L.frag, nf);
                C.tr.centre[3 * q + 0] = tcv.x;
                C.tr.centre[3 * q + 1] = tcv.y;
                C.tr.centre[3 * q + 2] = tcv.z;
                C.tr.idx[q] = slot;
            }
        }
    }
}

// ---- cell-list maintenance (single thread) ----------------------------------
__device__ inline void cell_insert(CRep &C, int gid, int spi, float x, float y, float z)
{
    const int c = cell_index(C.L, C.start, C.inv_h, C.nc, x, y, z);
    const int n = C.cell[c].y;
    int pos;
    if (n < C.cap) {
        pos = c * C.cap + n;
        C.cell[c].y = n + 1;
    } else {
        const int o = C.cell[C.ncell].y;
        pos = C.ovf_base + o;
        C.cell[C.ncell].y = o + 1;
    }
    C.mob_slot[pos] = gid;
    C.mob_rec[pos] = make_float4(x, y, z, __int_as_float(mob_word(gid, spi)));
    C.mob_loc[gid] = pos;
}

__device__ inline void cell_remove(CRep &C, int gid)
{
    const int pos = C.mob_loc[gid];
    int c, base;
    if (pos >= C.ovf_base) {
        c = C.ncell;
        base = C.ovf_base;
    } else {
        c = pos / C.cap;
        base = c * C.cap;
    }
    const int n = C.cell[c].y - 1;
    const int moved = C.mob_slot[base + n];
    C.mob_slot[pos] = moved;
    C.mob_rec[pos] = C.mob_rec[base + n];
    C.mob_loc[moved] = pos;
    C.cell[c].y = n;
    C.mob_loc[gid] = -1;
}

// group `from` now lives in slot `to` (its centre and species are unchanged)
__device__ inline void cell_rename(CRep &C, int from, int to)
{
    const int pos = C.mob_loc[from];
    C.mob_slot[pos] = to;
    float4 r = C.mob_rec[pos];
    r.w = __int_as_float((__float_as_int(r.w) & ~GID_MASK) | to);
    C.mob_rec[pos] = r;
    C.mob_loc[to] = pos;
    C.mob_loc[from] = -1;
}

// ---- selection and Metropolis (one wave) --------------------------------------
// Greedy grouping in ascending trial order: leader = first unused trial,
// members = unused trials within the cutoff of the leader; Boltzmann weights
// relative to the group minimum; cumulative pick; acceptance
//   insert:  min(1, cav/(m P)/(N+1) exp(B - beta E))   (gcmc_move_add.h:164-174)
//   move:    min(1, 1/(m P) exp(-beta dE))             (gcmc_move_trn.h:215-222)
// On accept the unused trials within the cutoff of the pick are dropped.
// Trial values are staged in L.val, adjacency rows in L.adjl (when the layout
// has room; else read from C.adj).  The weight sum and the cumulative pick
// are wave-parallel: lane l sums a contiguous chunk of the members in order,
// the chunk sums are scanned across lanes, and the lane whose chunk holds
// the crossing walks it -- one fixed FP64 summation order.
__device__ inline uint32_t adj_word(CRep &C, const Lds &L, int nw, int t, int w)
{
    return L.adjl ? L.adjl[t * nw + w] : C.adj[t * ADJW + w];
}

// A move's descriptor and the replica values its accept phase reads, loaded
// once at its start: every later phase would otherwise reload them after a
// barrier, one dependent L2 round trip each on the replica's serial path.
struct MoveCtx {
    int kind, spi, k, step, nf, N, maxN, startRes, atomBase;
    double B, energy;
};

__device__ inline MoveCtx move_ctx(CRep &C, const State &S)
{
    MoveCtx m;
    m.kind = S.kind;
    m.spi = S.sp;
    m.k = S.k;
    m.step = S.step;
    m.N = S.totalNum[m.spi];
    m.energy = S.energy;
    CSpecies &F = species_of(C, m.spi);
    m.nf = F.nf;
    m.maxN = F.maxNum;
    m.startRes = F.startRes;
    m.atomBase = F.atomBase;
    m.B = F.B;
    return m;
}

// the next move's descriptor into L.misc[M_NX..] (one thread; the accept
// phase's selection wave is busy meanwhile).  The species count is the one
// before the current move: prep takes the current move's result for its own
// species.
__device__ inline void load_next_move(CRep &C, const State &S, const Lds &L, const MoveCtx &m)
{
    const int end = S.end, step = m.step + 1;
    const int mv = step < end ? C.moves[step] : 0;
    CSpecies &F = species_of(C, mv >> 2);
    L.misc[M_NX + 0] = end;
    L.misc[M_NX + 1] = mv;
    L.misc[M_NX + 2] = S.totalNum[mv >> 2];
    L.misc[M_NX + 3] = F.conf;
    L.misc[M_NX + 4] = F.maxNum;
    L.misc[M_NX + 5] = F.startRes;
}

__device__ inline void greedy_select(CRep &C, State &S, const Lds &L, int row, bool trace, const MoveCtx &mc)
{
    const int lane = threadIdx.x & 63;
    const int kind = mc.kind, k = mc.k;
    const bool ins = kind == K_INS;
    const int nw = (k + 31) >> 5;
    uint32_t u = 0u;
    if (lane < nw) {
        const int hi = k - lane * 32;
        u = hi >= 32 ? 0xffffffffu : ((1u << hi) - 1u);
    }
    int N = mc.N;
    const int maxN = mc.maxN;
    const double cav = (double)C.cav;
    int g = 0, na = 0;
    double et = mc.energy;
    while (true) {
        const unsigned long long bal = __ballot(u != 0u);
        if (bal == 0ull) break;
        const int f = __ffsll((long long)bal) - 1;
        const uint32_t wf = __shfl(u, f, 64);
        const int it = f * 32 + __ffs(wf) - 1;
        uint32_t m = (lane < nw) ? (adj_word(C, L, nw, it, lane) & u) : 0u;
        if (lane == (it >> 5)) m |= 1u << (it & 31);
        u &= ~m;
        const int cnt = __popc(m);
        const int x = wave_scan_add(cnt);
        const int mtot = __builtin_amdgcn_readlane(x, 63);
        int o = x - cnt;
        uint32_t mm = m;
        while (mm) {
            const int b = __ffs(mm) - 1;
            mm &= mm - 1u;
            L.members[o++] = lane * 32 + b;
        }
        wave_sync();
        const double ran_sel = L.rsel[g];
        const double ran_acc = L.racc[g];
        int pick = it, pq = 0;
        double sumP = 1.0, P = 1.0;
        if (mtot > 1) {
            double emin = __builtin_huge_val();
            for (int q = lane; q < mtot; q += 64) {
                const double v = L.val[L.members[q]];
                emin = v < emin ? v : emin;
            }
            emin = wave_min(emin);
            for (int q = lane; q < mtot; q += 64) L.w[q] = exp(-BETA * (L.val[L.members[q]] - emin));
            wave_sync();
            // chunk sums in member order, scanned across lanes
            const int per = (mtot + 63) >> 6;
            const int q0 = min(lane * per, mtot), q1 = min(q0 + per, mtot);
            double part = 0.0;
            for (int q = q0; q < q1; ++q) part += L.w[q];
            const double incl = wave_scan_add(part);
            sumP = lane_d(incl, 63);
            if (sumP != 0.0) {
                // first member q with sum_{j<=q} w_j > ran_sel * sumP (the last one
                // when rounding leaves the target at the total)
                const double target = ran_sel * sumP;
                const double excl = incl - part;
                const unsigned long long hit = __ballot(q1 > q0 && excl <= target && target < incl);
                int sel = mtot - 1;
                if (hit) {
                    const int l = __ffsll((long long)hit) - 1;
                    int found = q1 - 1;
                    if (lane == l) {
                        double c = excl;
                        for (int q = q0; q < q1; ++q) {
                            c += L.w[q];
                            if (target < c) {
                                found = q;
                                break;
                            }
                        }
                    }
                    sel = __shfl(found, l, 64);
                }
                pq = sel;
                pick = L.members[sel];
            }
            P = sumP != 0.0 ? L.w[pq] / sumP : 1.0;
        }
        // (a single-member group has weight exp(0) = 1: pick = leader, P = 1)
        const double vp = L.val[pick];
        double pp;
        if (ins)
            pp = (cav / ((double)mtot * P)) / ((double)N + 1.0) * L.ea[pick];
        else
            pp = (1.0 / ((double)mtot * P)) * L.ea[pick];
        const double p = pp < 1.0 ? pp : 1.0;
        const bool acc = ran_acc < p;
        int rec = acc ? 1 : 0;
        if (acc) {
            const uint32_t ap = (lane < nw) ? adj_word(C, L, nw, pick, lane) : 0u;
            u &= ~ap;
            if (ins) {
                if (N < maxN) {
                    if (lane == 0) L.accl[na] = make_int2(pick, N);
                    N += 1;
                    na += 1;
                    et += vp;
                } else {
                    rec = 2;
                }
            } else {
                if (lane == 0) L.accl[na] = make_int2(pick, C.tslot[pick]);
                na += 1;
                et += vp;
            }
        }
        if (trace && lane == 0 && g < C.tr.max_k) {
            C.tr.gpick[row * C.tr.max_k + g] = pick;
            C.tr.gacc[row * C.tr.max_k + g] = rec;
        }
        g += 1;
        wave_sync();
    }
    if (lane == 0) {
        L.misc[M_ACC] = na;
        L.misc[M_NG] = g;
        L.misc[M_N] = N;
        S.energy = et;
        // counters only: an atomic add without return keeps the load of
        // the old value off the serial path
        if (na) count_add<false>(&S.acc[mc.spi * 4 + kind], (long long)na);
    }
}

// greedy_select for k <= 64 with the per-trial values in registers: lane t
// holds trial t's value, exp term, adjacency row (bit j: trial j within the
// cutoff) and slot, and group t's two variates.  Group members are compacted
// into lanes 0..m-1 in trial order (ds_permute), so the weights, their sums
// and the decisions are those of greedy_select, bit for bit.  Wave 0 only.
__device__ inline void greedy_select64(CRep &C, State &S, const Lds &L, int row, bool trace, const MoveCtx &mc,
                                       double val, double ea, double rsel, double racc, unsigned long long adj,
                                       int tslot)
{
    const int lane = threadIdx.x & 63;
    const int kind = mc.kind, k = mc.k;
    const bool ins = kind == K_INS;
    const unsigned long long below = (1ull << lane) - 1ull;
    unsigned long long U = k >= 64 ? ~0ull : ((1ull << k) - 1ull);
    int N = mc.N;
    const int maxN = mc.maxN;
    const double cav = (double)C.cav;
    int g = 0, na = 0;
    double et = mc.energy;
    while (U != 0ull) {
        const int it = __ffsll((long long)U) - 1;
        const unsigned long long M = (lane_u64(adj, it) & U) | (1ull << it);
        U &= ~M;
        const int mtot = __popcll(M);
        const double ran_sel = lane_d(rsel, g);
        const double ran_acc = lane_d(racc, g);
        int pick = it;
        double P = 1.0;
        if (mtot > 1) {
            const bool in = ((M >> lane) & 1ull) != 0ull;
            const double emin = wave_min(in ? val : __builtin_huge_val());
            const double wl = in ? exp(-BETA * (val - emin)) : 0.0;
            const int dst = in ? __popcll(M & below) : mtot + __popcll(~M & below);
            const double w = permute_d(dst, wl);                      // lane q: member q's weight
            const int tix = __builtin_amdgcn_ds_permute(dst * 4, lane);  // lane q: member q's trial
            const double part = lane < mtot ? w : 0.0;
            const double incl = wave_scan_add(part);
            const double sumP = lane_d(incl, 63);
            int pq = 0;
            if (sumP != 0.0) {
                const double target = ran_sel * sumP;
                const double excl = incl - part;
                const unsigned long long hit = __ballot(lane < mtot && excl <= target && target < incl);
                pq = hit ? __ffsll((long long)hit) - 1 : mtot - 1;
                pick = lane_i(tix, pq);
                P = lane_d(w, pq) / sumP;
            }
        }
        const double vp = lane_d(val, pick);
        const double eap = lane_d(ea, pick);
        double pp;
        if (ins)
            pp = (cav / ((double)mtot * P)) / ((double)N + 1.0) * eap;
        else
            pp = (1.0 / ((double)mtot * P)) * eap;
        const double p = pp < 1.0 ? pp : 1.0;
        const bool acc = ran_acc < p;
        int rec = acc ? 1 : 0;
        if (acc) {
            U &= ~lane_u64(adj, pick);
            if (ins) {
                if (N < maxN) {
                    if (lane == 0) L.accl[na] = make_int2(pick, N);
                    N += 1;
                    na += 1;
                    et += vp;
                } else {
                    rec = 2;
                }
            } else {
                const int sl = lane_i(tslot, pick);
                if (lane == 0) L.accl[na] = make_int2(pick, sl);
                na += 1;
                et += vp;
            }
        }
        if (trace && lane == 0 && g < C.tr.max_k) {
            C.tr.gpick[row * C.tr.max_k + g] = pick;
            C.tr.gacc[row * C.tr.max_k + g] = rec;
        }
        g += 1;
    }
    if (lane == 0) {
        L.misc[M_ACC] = na;
        L.misc[M_NG] = g;
        L.misc[M_N] = N;
        S.energy = et;
        // counters only: an atomic add without return keeps the load of
        // the old value off the serial path
        if (na) count_add<false>(&S.acc[mc.spi * 4 + kind], (long long)na);
    }
}

// ---- commit of the accepted trials (L.accl[0..na): trial, slot) ------------------
// GupdateAdd / GupdateTrn / GupdateRot: copy committed poses, then the group
// centres and cell-list entries in acceptance order (thread 0).
template <int NT, bool EW>
__device__ inline void commit_accepted(CRep &C, State &S, const Lds &L, int row, bool trace, const MoveCtx &mc)
{
    const int tid = threadIdx.x;
    const int kind = mc.kind, spi = mc.spi;
    CSpecies &F = species_of(C, spi);
    const int nf = mc.nf;
    const int na = L.misc[M_ACC];
    // GupdateAdd / GupdateTrn / GupdateRot: copy committed poses
    for (int idx = tid; idx < na * nf; idx += NT) {
        const int q = idx / nf, a = idx - q * nf;
        const int2 ts = L.accl[q];
        const int base = mc.atomBase + ts.y * nf;
        const float4 p = C.pose[ts.x * MAXF + a];
        float4 dst;
        if (kind == K_INS) {
            dst = make_float4(p.x, p.y, p.z, F.tmpl[a].w);
            C.atype[base + a] = F.ttype[a];
        } else {
            dst = C.atom[base + a];
            if (EW) C.pme_delta[2 * idx] = make_float4(dst.x, dst.y, dst.z, -F.tmpl[a].w);
            dst.x = p.x; dst.y = p.y; dst.z = p.z;
        }
        if (EW) {
            if (kind == K_INS) C.pme_delta[2 * idx] = make_float4(0.f, 0.f, 0.f, 0.f);
            C.pme_delta[2 * idx + 1] = make_float4(p.x, p.y, p.z, F.tmpl[a].w);
        }
        C.atom[base + a] = dst;
    }
    if (EW && tid == 0) *C.pme_ndelta = 2 * na * nf;
    if (tid == 0) {
        for (int q = 0; q < na; ++q) {
            const int2 ts = L.accl[q];
            const int gid = mc.startRes + ts.y;
            const float4 c = C.tc[ts.x];
            if (kind == K_INS) {
                C.res_pos[gid] = make_float4(c.x, c.y, c.z, 0.f);
                cell_insert(C, gid, spi, c.x, c.y, c.z);
                if (trace && ts.x < C.tr.max_k) C.tr.idx[row * C.tr.max_k + ts.x] = ts.y;
            } else if (kind == K_TRN) {
                const float4 old = C.res_pos[gid];
                C.res_pos[gid] = make_float4(c.x, c.y, c.z, old.w);
                cell_remove(C, gid);
                cell_insert(C, gid, spi, c.x, c.y, c.z);
            }
        }
        if (kind == K_INS) S.totalNum[spi] = L.misc[M_N];
    }
}

// ---- accept phase: decisions and commit ------------------------------------------
template <int NT, bool EW>
__device__ inline void accept_move(CRep &C, State &S, const Lds &L, const MoveCtx &mc, bool prefetch = false)
{
    const int tid = threadIdx.x;
    const int kind = mc.kind, spi = mc.spi, k = mc.k, step = mc.step;
    CSpecies &F = species_of(C, spi);
    const int nf = mc.nf;
    int row;
    const bool trace = tracing(C, step, row);
    // stage the trial values (E for insert/delete, E_new - E_old otherwise);
    // with at most 64 trials of an insert or move they stay in registers
    const bool in_regs = k <= 64 && kind != K_DEL;
    if (!in_regs)
        for (int t = tid; t < k; t += NT) L.val[t] = (kind == K_TRN || kind == K_ROT) ? C.E[k + t] - C.E[t] : C.E[t];
    if (tid == 0) {
        count_add<false>(&S.att[spi * 4 + kind], 1ll);
        L.misc[M_ACC] = 0;
        L.misc[M_NG] = 0;
        L.misc[M_N] = mc.N;
    }
    __syncthreads();
    if (prefetch && tid == 64) load_next_move(C, S, L, mc);
    if (kind == K_DEL && k == 1) {
        const int N = mc.N;
        if (tid == 0) {
            const double E = L.val[0];
            const u32x4 x = draw(C.seed, (uint32_t)step, 0u, P_DECIDE);
            const double ran = u01d(x.z, x.w);
            const double pp = (double)N * exp(-mc.B - BETA * (-E));
            const double p = pp < 1.0 ? pp : 1.0;
            const int acc = ran < p;
            L.misc[M_ACC] = acc;
            L.misc[M_NG] = 1;
            if (trace) {
                C.tr.gpick[row * C.tr.max_k] = 0;
                C.tr.gacc[row * C.tr.max_k] = acc;
            }
            if (acc) {
                S.energy = mc.energy - E;
                count_add<false>(&S.acc[spi * 4 + kind], 1ll);
            }
        }
        __syncthreads();
        if (L.misc[M_ACC]) {
            // GupdateDel (gcmc_move_del.h:55-94): swap-remove with the last slot
            const int newN = N - 1;
            const int hole_slot = C.tslot[0];
            const int hole = F.startRes + hole_slot, last = F.startRes + newN;
            if (EW) {
                // the deleted molecule's charges leave the PME grid (the moved
                // last molecule keeps its coordinates)
                for (int a = tid; a < nf; a += NT) {
                    const float4 o = C.atom[F.atomBase + hole_slot * nf + a];
                    C.pme_delta[a] = make_float4(o.x, o.y, o.z, -F.tmpl[a].w);
                }
                if (tid == 0) *C.pme_ndelta = nf;
                __syncthreads();
            }
            if (newN > 0 && hole != last) {
                const int bh = F.atomBase + hole_slot * nf, bl = F.atomBase + newN * nf;
                for (int a = tid; a < nf; a += NT) {
                    const float4 src = C.atom[bl + a];
                    float4 dst = C.atom[bh + a];
                    dst.x = src.x; dst.y = src.y; dst.z = src.z;
                    C.atom[bh + a] = dst;
                }
            }
            if (tid == 0) {
                cell_remove(C, hole);
                if (newN > 0 && hole != last) {
                    const float4 pl = C.res_pos[last];
                    float4 ph = C.res_pos[hole];
                    ph.x = pl.x; ph.y = pl.y; ph.z = pl.z;
                    C.res_pos[hole] = ph;
                    cell_rename(C, last, hole);
                }
                S.totalNum[spi] = newN;
                L.misc[M_N] = newN;
            }
        }
    } else if (k > 0 && in_regs) {
        // lane t of wave 0 holds trial t's values (greedy_select64)
        if ((tid >> 6) == 0) {
            const int t = tid;
            double val = 0.0, ea = 0.0, rsel = 0.0, racc = 0.0;
            unsigned long long adj = 0ull;
            int sl = 0;
            if (t < k) {
                const auto E = gbl(C.E);
                val = (kind == K_TRN || kind == K_ROT) ? E[k + t] - E[t] : E[t];
                const auto A = gbl(C.adj);
                adj = (unsigned long long)A[t * ADJW];
                if (k > 32) adj |= (unsigned long long)A[t * ADJW + 1] << 32;
                if (kind != K_INS) sl = gbl(C.tslot)[t];
                const u32x4 xr = draw(C.seed, (uint32_t)step, (uint32_t)t, P_DECIDE);
                rsel = u01d(xr.x, xr.y);
                racc = u01d(xr.z, xr.w);
                ea = exp(((kind == K_INS) ? mc.B : 0.0) - BETA * val);
            }
            greedy_select64(C, S, L, row, trace, mc, val, ea, rsel, racc, adj, sl);
        }
        __syncthreads();
        commit_accepted<NT, EW>(C, S, L, row, trace, mc);
    } else if (k > 0 && kind != K_DEL) {
        if (L.adjl) {
            const int nw = (k + 31) >> 5;
            for (int q = tid; q < k * nw; q += NT) {
                const int t = q / nw, w = q - t * nw;
                L.adjl[q] = C.adj[t * ADJW + w];
            }
        }
        // everything of the serial selection that does not depend on earlier
        // groups: the variates of group g (at most k groups) and each trial's
        // exp(B - beta E) (insert) or exp(-beta dE) (move)
        {
            const double Bx = (kind == K_INS) ? mc.B : 0.0;
            for (int t = tid; t < k; t += NT) {
                const u32x4 xr = draw(C.seed, (uint32_t)step, (uint32_t)t, P_DECIDE);
                L.rsel[t] = u01d(xr.x, xr.y);
                L.racc[t] = u01d(xr.z, xr.w);
                L.ea[t] = exp(Bx - BETA * L.val[t]);
            }
        }
        __syncthreads();
        if ((tid >> 6) == 0) greedy_select(C, S, L, row, trace, mc);
        __syncthreads();
        commit_accepted<NT, EW>(C, S, L, row, trace, mc);
    }
    if (tid == 0) {
        if (trace) {
            C.tr.k[row] = k;
            C.tr.ngroups[row] = L.misc[M_NG];
            C.tr.naccept[row] = L.misc[M_ACC];
            C.tr.n_after[row] = S.totalNum[spi];
            C.tr.etot[row] = S.energy;
        }
        S.step = step + 1;
    }
    __syncthreads();
}

template <int NT, bool EW>
__device__ inline void accept_step(CRep &C, State &S, const Lds &L)
{
    const MoveCtx mc = move_ctx(C, S);
    accept_move<NT, EW>(C, S, L, mc);
}

// ---- prep phase: picks and trial centres of S.step ------------------------------
// pre: the move's descriptor loaded with the previous move (for the same
// species, the count that move's accept phase left in L.misc[M_N]), or null
// use_pre: the move after prev (species prev_spi) was loaded into L.misc[M_NX..]
template <int NT>
__device__ inline int prep_step(CRep &C, State &S, const Lds &L, bool use_pre, int prev_step, int prev_spi)
{
    const int tid = threadIdx.x;
    int step, mv, N, conf, maxNum, startRes;
    if (use_pre) {
        step = prev_step + 1;
        mv = L.misc[M_NX + 1];
        N = ((mv >> 2) == prev_spi) ? L.misc[M_N] : L.misc[M_NX + 2];
        conf = L.misc[M_NX + 3];
        maxNum = L.misc[M_NX + 4];
        startRes = L.misc[M_NX + 5];
    } else {
        step = S.step;
        mv = C.moves[step];
        CSpecies &F0 = species_of(C, mv >> 2);
        N = S.totalNum[mv >> 2];
        conf = F0.conf;
        maxNum = F0.maxNum;
        startRes = F0.startRes;
    }
    const int spi = mv >> 2, kind = mv & 3;
    int k;
    if (kind == K_INS) k = (N >= maxNum) ? 0 : conf;
    else if (kind == K_DEL) k = (N == 0) ? 0 : 1;
    else k = conf < N ? conf : N;
    if (k > C.maxk) k = C.maxk;
    if (kind != K_INS && k > 0) {
        // partial Fisher-Yates (gcmc_move_trn.h:135-149): pick i takes the value at
        // position J[i] after the swaps 0..i-1, found by tracing the swaps back
        for (int i = tid; i < k; i += NT) {
            const u32x4 x = draw(C.seed, (uint32_t)step, (uint32_t)i, P_PICK);
            L.J[i] = i + (int)below(x.x, (uint32_t)(N - i));
        }
        __syncthreads();
        for (int i = tid; i < k; i += NT) {
            int p = L.J[i];
            for (int m = i - 1; m >= 0; --m) {
                const int jm = L.J[m];
                p = (p == m) ? jm : ((p == jm) ? m : p);
            }
            C.tslot[i] = p;
        }
    }
    __syncthreads();
    for (int t = tid; t < k; t += NT) {
        float c[3];
        if (kind == K_INS) {
            const u32x4 x = draw(C.seed, (uint32_t)step, (uint32_t)t, P_JITTER);
            const uint32_t g = below(x.w, (uint32_t)C.G);
            const float u[3] = {u01f(x.x), u01f(x.y), u01f(x.z)};
            for (int i = 0; i < 3; ++i) c[i] = C.grid[3 * g + i] + u[i] * C.dx;
        } else {
            const float4 r = C.res_pos[startRes + C.tslot[t]];
            c[0] = r.x; c[1] = r.y; c[2] = r.z;
            if (kind == K_TRN) {
                const u32x4 x = draw(C.seed, (uint32_t)step, (uint32_t)t, P_JITTER);
                const float u[3] = {u01f(x.x), u01f(x.y), u01f(x.z)};
                for (int i = 0; i < 3; ++i) {
                    float d = u[i] * C.dx;
                    if (c[i] + d > C.L[i] + C.start[i]) d = d - C.L[i];
                    if (c[i] + d < C.start[i]) d = d + C.L[i];
                    c[i] = c[i] + d;
                }
            }
        }
        C.tc[t] = make_float4(c[0], c[1], c[2], 0.f);
    }
    const int nitems = (kind == K_TRN || kind == K_ROT) ? 2 * k : k;
    if (tid == 0) {
        S.kind = kind;
        S.sp = spi;
        S.k = k;
        S.nitems = nitems;
        S.done = 0;
    }
    __syncthreads();
    return nitems;
}

template <int NT>
__device__ inline int prep_step(CRep &C, State &S, const Lds &L)
{
    return prep_step<NT>(C, S, L, false, 0, 0);
}

// Items a workgroup takes with one claim of a move of `unc` unclaimed items
// in size class sz (0: at most 4, 1: 5-16, 2: more).  Consecutive items of one
// move share the claim's round trips and acquire (an L1 invalidate), and the
// workgroup's second and later evaluations find the move's data in L1: with
// eight systems per XCD, a pass of the 64-system workload takes 700 instead of
// 1150 ms with 4 items per claim of the largest moves and 2 of the middle
// ones.  With few systems per XCD (nc < 4) the workgroups would idle while a
// few of them work through a move's items one after another, so a claim then
// takes no more than a fair share, ceil(unc / workgroups of the XCD).  Small
// moves, a replica's critical path, take one item per claim.
__device__ inline int claim_take(int sz, int unc, int nc, int nwx)
{
    const int want = sz == 2 ? 4 : (sz == 1 ? 2 : 1);
    if (nc >= 4) return want;
    const int share = (unc + nwx - 1) / nwx;
    return min(want, max(share, 1));
}

// XCD-local ordering.  Every workgroup touching a replica in the batched form
// runs on the same XCD, whose L2 is shared and whose L1s are write-through: a
// release waits until this thread's memory operations reached L2; an acquire
// invalidates the CU's L1 at device scope (sc1) so later loads read L2 -- a
// workgroup-scope invalidate (sc0) leaves L1 as it is, since one workgroup's
// waves share that L1.  No L2 write-back: replica data never leaves the XCD.
__device__ inline void xcd_release() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }
__device__ inline void xcd_acquire() { asm volatile("buffer_inv sc1" ::: "memory"); }

// claim-word priority of a replica's published move: lower is claimed first,
// the replica with the most estimated work left first -- the run lasts as long
// as its slowest chain
__device__ inline int claim_prio(CRep &C, int step, int end)
{
    const uint32_t left = C.wcum[end] - C.wcum[step];
    return 0xFFFFFF - (int)(left < 0xFFFFFFu ? left : 0xFFFFFFu);
}

// advance past empty moves and publish the next move with work (batched form)
// use_pre: the move after prev_step (species prev_spi) is in L.misc[M_NX..]
template <int NT>
__device__ inline void advance_and_publish(CRep &C, State &S, Sched *sch, const Lds &L, bool use_pre,
                                           int prev_step, int prev_spi)
{
    while (true) {
        if (use_pre ? prev_step + 1 >= L.misc[M_NX + 0] : S.step >= S.end) {
            xcd_release();
            __syncthreads();
            if (threadIdx.x == 0) {
                __hip_atomic_store(&S.claim, claim_word(0xFFFFFF, 0, 0), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_fetch_add(&sch->finished, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            break;
        }
        int step, end;
        if (use_pre) {
            step = prev_step + 1;
            end = L.misc[M_NX + 0];
        } else {
            step = S.step;
            end = S.end;
        }
        const int nitems = prep_step<NT>(C, S, L, use_pre, prev_step, prev_spi);
        use_pre = false;
        if (nitems > 0) {
            const int prio = claim_prio(C, step, end);
            if (threadIdx.x == 0) {
                S.t_pub = wall_clock64();
                __hip_atomic_store(&S.done, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            xcd_release();
            __syncthreads();
            if (threadIdx.x == 0) {
                // every thread's writes have completed: a CU whose L1 was
                // invalidated after this time holds no older copy of them
                S.t_data = wall_clock64();
                xcd_release();
                __hip_atomic_store(&S.claim, claim_word(prio, nitems, 0), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            break;
        }
        accept_step<NT, false>(C, S, L);
    }
    __syncthreads();
}

template <int NT>
__device__ inline void advance_and_publish(CRep &C, State &S, Sched *sch, const Lds &L)
{
    advance_and_publish<NT>(C, S, sch, L, false, 0, 0);
}

// ---- kernels ---------------------------------------------------------------------

// start of a run: set the end step of every replica and reset the scheduler
__global__ void k_begin(State *st, const Rep *reps, int r0, int nrep, int nsteps, Sched *sch)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        sch->finished = 0;
        sch->abort = 0;
        sch->progress = 0;
    }
    if (i < nrep) {
        State &S = st[r0 + i];
        const long long e = (long long)S.step + nsteps;
        S.end = (int)(e < reps[r0 + i].nmoves ? e : reps[r0 + i].nmoves);
    }
}

// prepare the current move of replicas r0 + blockIdx.x (batched: publish it)
template <int NT, bool STAGED>
__global__ __launch_bounds__(NT) void k_start(const Rep *reps, State *st, Sched *sch, LdsLayout lay,
                                             int r0, int batched)
{
    extern __shared__ __align__(16) char smem[];
    const Lds L = carve(smem, lay);
    const int r = r0 + blockIdx.x;
    if (batched) advance_and_publish<NT>(rep_of(reps, r), st[r], sch, L);
    else if (st[r].step < st[r].end) prep_step<NT>(rep_of(reps, r), st[r], L);
}

template <int NT, bool STAGED, bool EW>
__global__ __launch_bounds__(NT) void k_eval_plain(const Rep *reps, const State *st, int r, LdsLayout lay)
{
    extern __shared__ __align__(16) char smem[];
    const State &S = st[r];
    if (S.step >= S.end || (int)blockIdx.x >= S.nitems) return;
    const Lds L = carve(smem, lay);
    eval_item<NT, STAGED, false, EW>(rep_of(reps, r), S, blockIdx.x, L);
}

template <int NT, bool STAGED, bool EW>
__global__ __launch_bounds__(NT) void k_accept_plain(const Rep *reps, State *st, int r, LdsLayout lay)
{
    extern __shared__ __align__(16) char smem[];
    State &S = st[r];
    if (S.step >= S.end) return;
    const Lds L = carve(smem, lay);
    accept_step<NT, EW>(rep_of(reps, r), S, L);
    if (S.step < S.end) prep_step<NT>(rep_of(reps, r), S, L);
}

// Which replicas a workgroup serves: the XCD it runs on (XCC_ID hardware
// register) selects a class; class c owns replicas cls_reps[cls_start[c] ..
// cls_start[c+1]).  xcc_class maps XCC ids to classes (-1: none).
struct XcdMap {
    int nclass;
    int xcc_class[16];
    const int *cls_start;
    const int *cls_reps;
    unsigned long long *cu_inv;   // [CU_SLOTS] time of each CU's last L1 invalidate (100 MHz ticks)
};

__device__ inline int my_xcc() { return (int)(__builtin_amdgcn_s_getreg((20) | (0 << 6) | (15 << 11)) & 0xf); }

// This CU's slot in XcdMap::cu_inv: XCC id, shader engine, shader array and
// CU id from the XCC_ID and HW_ID registers (distinct for distinct CUs).
constexpr int CU_SLOTS = 4096;
__device__ inline int my_cu_slot()
{
    const unsigned hw = __builtin_amdgcn_s_getreg((4) | (0 << 6) | (31 << 11));
    const unsigned cu = (hw >> 8) & 15u, sh = (hw >> 12) & 1u, se = (hw >> 13) & 7u;
    return (int)(((unsigned)my_xcc() << 8) | (se << 5) | (sh << 4) | cu);
}

// The acquire before evaluating a claimed move's items is a device-scope L1
// invalidate, which also empties the L1 of the CU's other workgroups.  It is
// needed only when the CU's L1 may hold lines older than the move's data:
// not when this CU's last invalidate came after the move's data was complete
// (State::t_data).  One thread asks; both times are read at device scope.
__device__ inline int cu_acquire_needed(const State &S, const unsigned long long *cu_inv, int slot)
{
    const unsigned long long td = __hip_atomic_load(&S.t_data, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned long long last = __hip_atomic_load(&cu_inv[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return last >= td ? 0 : 1;
}

// persistent form: workgroups of an XCD claim evaluations from that XCD's
// replicas; the workgroup completing a replica's last evaluation runs its
// accept and prep phases and publishes the next move.  Exits when every
// replica finished, or when no replica advanced during spin_limit polls
// (then sch->abort is set).
template <int NT, bool STAGED>
__global__ __launch_bounds__(NT, 768 / NT) void k_batched(const Rep *reps, State *st, int R, Sched *sch,
                                               LdsLayout lay, int spin_limit, XcdMap xm, int aff)
{
    extern __shared__ __align__(16) char smem[];
    const Lds L = carve(smem, lay);
    const int tid = threadIdx.x;
    const int cls = xm.xcc_class[my_xcc()];
    if (cls < 0) return;
    const int c0 = xm.cls_start[cls], nc = xm.cls_start[cls + 1] - c0;
    if (nc <= 0) return;
    const int rot = (int)(blockIdx.x / (unsigned)xm.nclass) % nc;
    const int nwx = max((int)(gridDim.x / (unsigned)xm.nclass), 1);   // workgroups per XCD
    const int slot = my_cu_slot();
    int idle = 0, seen = -1;
    // replica whose species table, and pair table (Rep::tab) that are in LDS
    int cached_rep = -1, cached_tab = -1;
    while (true) {
        if (tid < 64) {
            // wave 0 polls every replica of this XCD at once, one claim word
            // per lane.  Claim words are device-scope: a workgroup-scope load
            // may hit this CU's L1, which only a claiming workgroup invalidates,
            // so idle workgroups would keep seeing an old word.  Polls are
            // loads; an atomic is issued only to claim (read-modify-writes of
            // one address serialize).  The scheduler's words (progress,
            // finished, abort) are read only when nothing could be claimed:
            // a workgroup that finds work goes from its last item to its next
            // with two memory round trips (the claim words, the claim).
            const int lane = tid;
            int found = -1, item = 0, cnt = 1;
            for (int b0 = 0; b0 < nc && found < 0; b0 += 64) {
                int r = -1;
                unsigned long long w = 0ull;
                if (b0 + lane < nc) {
                    r = xm.cls_reps[c0 + (rot + b0 + lane) % nc];
                    w = __hip_atomic_load(&st[r].claim, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
                bool avail = r >= 0 && (w & 0xFFFFFull) < ((w >> 20) & 0xFFFFFull);
                // Moves with few unclaimed items first (up to 4, then up to
                // 16): a replica waiting on one or a few evaluations is on its
                // own critical chain.  Within those classes and among the
                // larger moves, the replica with the most estimated work left
                // first (claim_prio): the run ends with its last replica, and
                // a large move occupies the XCD's workgroups for several
                // rounds whichever order they come in.  (Fewest unclaimed
                // items first among the large moves too: 5 % slower.)
                const unsigned long long unc = ((w >> 20) & 0xFFFFFull) - (w & 0xFFFFFull);
                unsigned size = unc <= 4ull ? 0u : (unc <= 16ull ? 1u : 2u);
                // affinity (aff > 0): a move of at least aff unclaimed items
                // of the replica whose tables this workgroup holds in LDS
                // comes before the other such moves (its species table and
                // staged pair table are reused)
                if (aff > 0 && unc >= (unsigned long long)aff) size = (r == cached_rep) ? 2u : 3u;
                const unsigned prio = (unsigned)(w >> 40) >> 8;
                const int take_lane = claim_take((int)size, (int)unc, nc, nwx);
                while (__ballot(avail) != 0ull) {
                    const unsigned key =
                        wave_min(avail ? ((size << 22) | (prio << 6) | (unsigned)lane) : 0xFFFFFFFFu);
                    const int l = (int)(key & 63u);
                    const int rr = __shfl(r, l, 64);
                    const int take = __shfl(take_lane, l, 64);
                    unsigned long long old = 0ull;
                    if (lane == l)
                        old = __hip_atomic_fetch_add(&st[rr].claim, (unsigned long long)take, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT);
                    old = __shfl(old, l, 64);
                    const int c = (int)(old & 0xFFFFFull), n = (int)((old >> 20) & 0xFFFFFull);
                    if (c < n) {
                        found = rr;
                        item = c;
                        cnt = min(take, n - c);
                        break;
                    }
                    if (lane == l) avail = false;
                }
            }
            if (found == -1) {
                int ctl[3] = {0, 0, 0};
                if (lane == 0) {
                    ctl[0] = __hip_atomic_load(&sch->progress, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    ctl[1] = __hip_atomic_load(&sch->finished, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    ctl[2] = __hip_atomic_load(&sch->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
                const int prog = __shfl(ctl[0], 0, 64), fin = __shfl(ctl[1], 0, 64), ab = __shfl(ctl[2], 0, 64);
                if (prog != seen) {
                    seen = prog;
                    idle = 0;
                }
                if (fin >= R || ab != 0) {
                    found = -2;
                } else if (++idle > spin_limit) {
                    // no replica advanced during spin_limit polls: stop instead of hanging
                    if (lane == 0) __hip_atomic_store(&sch->abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    found = -2;
                }
            }
            if (lane == 0) {
                L.misc[M_FOUND] = found;
                L.misc[M_ITEM] = item;
                L.misc[M_CNT] = cnt;
                L.misc[M_INV] = found >= 0 ? cu_acquire_needed(st[found], xm.cu_inv, slot) : 1;
            }
        }
        __syncthreads();
        // uniform across the workgroup: the replica's fields become scalar loads
        const int found = __builtin_amdgcn_readfirstlane(L.misc[M_FOUND]);
        const int item = __builtin_amdgcn_readfirstlane(L.misc[M_ITEM]);
        const int cnt = __builtin_amdgcn_readfirstlane(L.misc[M_CNT]);
        const int need_inv = __builtin_amdgcn_readfirstlane(L.misc[M_INV]);
        __syncthreads();
        if (found == -2) break;
        if (found == -1) {
            // about 0.9 us between polls of an idle workgroup (1.5 % faster
            // than 0.2 us on short runs of the 64-replica workload, equal to
            // 0.45-1.8 us on full runs; evaluation times do not change)
            __builtin_amdgcn_s_sleep(32);
            continue;
        }
        if (need_inv) {
            const unsigned long long tinv = wall_clock64();
            xcd_acquire();
            if (tid == 0) {
                // recorded once this wave's invalidate has executed
                xcd_release();
                __hip_atomic_store(&xm.cu_inv[slot], tinv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        CRep &C = rep_of(reps, found);
        State &S = st[found];
        const unsigned long long te0 = wall_clock64();
        const int tab = C.tab;
        ItemStats ist;
        // the claimed items one after another; the last one's statistics are
        // added after the completion count (the others' before it)
        for (int j = 0; j < cnt; ++j) {
            if (j > 0) __syncthreads();
            eval_item<NT, STAGED, true, false>(C, S, item + j, L, j == 0 && found != cached_rep,
                                               j == 0 && tab != cached_tab, j + 1 < cnt ? nullptr : &ist);
        }
        cached_rep = found;
        cached_tab = tab;
        const unsigned long long te1 = wall_clock64();
        xcd_release();
        __syncthreads();
        if (tid == 0) {
            const int old = __hip_atomic_fetch_add(&S.done, cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            L.misc[M_LAST] = (old + cnt == S.nitems) ? 1 : 0;
            // statistics after the completion count: the release above does
            // not wait for them
            add_item_stats<true>(S, ist);
            count_add<true>(&S.t_eval, te1 - te0);
        }
        __syncthreads();
        if (L.misc[M_LAST]) {
            // accept and prep are this replica's serial critical path: issue
            // ahead of the evaluation waves sharing the SIMDs
            __builtin_amdgcn_s_setprio(3);
            xcd_acquire();
            const unsigned long long ta0 = wall_clock64();
            const unsigned long long tpub = S.t_pub;
            const int step0 = S.step;
            const MoveCtx mc = move_ctx(C, S);
            accept_move<NT, false>(C, S, L, mc, true);
            const unsigned long long ta1 = wall_clock64();
            advance_and_publish<NT>(C, S, sch, L, true, mc.step, mc.spi);
            if (tid == 0) {
                const unsigned long long ta2 = wall_clock64();
                // timers after the next move is published, off its path; the
                // workgroup that finishes that move may already be adding its
                // own, so these are atomic adds too
                count_add<true>(&S.t_step, ta0 - tpub);
                count_add<true>(&S.t_accept, ta1 - ta0);
                count_add<true>(&S.t_prep, ta2 - ta1);
                int row;
                if (tracing(C, step0, row)) {
                    C.tr.tick[4 * row + 0] = tpub;
                    C.tr.tick[4 * row + 1] = ta0;
                    C.tr.tick[4 * row + 2] = ta1;
                    C.tr.tick[4 * row + 3] = ta2;
                }
                __hip_atomic_fetch_add(&sch->progress, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            __builtin_amdgcn_s_setprio(0);
            cached_rep = cached_tab = -1;  // accept and prep reuse the evaluation region
        }
        __syncthreads();
    }
}

// XCC id of each workgroup (engine set-up: which XCDs exist)
__global__ void k_xcc_probe(int *out)
{
    if (threadIdx.x == 0) out[blockIdx.x] = my_xcc();
}

// energies of given poses (tests and the adjoint's forward)
template <int NT, bool STAGED, bool EW>
__global__ __launch_bounds__(NT) void k_eval_poses(const Rep *reps, int r, int spi, const float *centres,
                                                  const float *poses, const int *excl, double *E, int *ninc,
                                                  uint32_t *hash, LdsLayout lay)
{
    extern __shared__ __align__(16) char smem[];
    const Lds L = carve(smem, lay);
    CRep &C = rep_of(reps, r);
    CSpecies &F = species_of(C, spi);
    const int nf = F.nf, i = blockIdx.x, tid = threadIdx.x;
    if (STAGED) stage_pair_table<NT>(C, L);
    load_species_table(C, L);
    if (tid < nf) {
        const float *p = poses + ((size_t)i * nf + tid) * 3;
        L.frag[tid] = make_float4(p[0], p[1], p[2], F.kq[tid]);
        L.frow[tid] = F.trow[tid];
    }
    __syncthreads();
    double e;
    int n, na;
    uint32_t h;
    eval_energy<NT, STAGED, EW>(C, centres[3 * i], centres[3 * i + 1], centres[3 * i + 2], nf, excl[i], L, &e, &n,
                                &h, &na);
    if (EW) {
        const int x = excl[i];
        const bool own = x >= F.startRes && x < F.startRes + F.maxNum;
        const double er = ewald_recip<NT>(C, F, L, nf, own ? F.atomBase + (x - F.startRes) * nf : -1);
        if (tid == 0) e += er;
    }
    if (tid == 0) {
        E[i] = e;
        ninc[i] = n;
        hash[i] = h;
    }
}

}  // namespace gcmc
