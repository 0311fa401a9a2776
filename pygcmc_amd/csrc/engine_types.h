// engine_types.h -- device-resident layout of the GCMC engine.
//
// One "replica" is one independent GCMC system (pyGCMC runs exactly one per
// call, gcmc/cpp/gcmc.cu:21).  Its data lives in HBM as structure-of-arrays:
//   res_pos  float4[R]   group centres (x, y, z, -)           gcmc.h residue
//   res_rng  int2[R]     (atomStart, atomNum)
//   atom     float4[A]   (x, y, z, charge)                     gcmc.h Atom
//   atype    int[A]      LJ type (column of the pair table)
//   pair     float2[X*Y] (100*sigma^2, 4*epsilon) per (fragment type, type)
// plus a cell list over group centres: fixed groups in a static CSR
// (cell / fix sorted by cell), mobile groups (fragment slots) in
// per-cell buckets of `cap` slots with one overflow bucket.
#pragma once

#include <stdint.h>

#include "gcmc_math.h"

namespace gcmc {

constexpr int MAXWAVE = 16;          // waves of the largest workgroup (1024 threads)
constexpr int MAXT = 32;             // species per system
constexpr int MAXF = 20;             // atoms per fragment (AtomArray.atoms[20])
constexpr int MAXK = 2048;           // trials per move (confBias)
constexpr int ADJW = MAXK / 32;      // words of one adjacency row
constexpr int MISC_INTS = 104;
constexpr int INC_CAP = 1024;         // included-group slots per pair pass (a multiple of the workgroup size)        // scratch ints of a workgroup (see engine_kernels.h)
// a mobile group's cell record carries its group id and species in one word
constexpr int GID_BITS = 24;         // group ids below 2^24
constexpr int GID_MASK = (1 << GID_BITS) - 1;
__host__ __device__ inline int mob_word(int gid, int spi) { return gid | (spi << GID_BITS); }
constexpr double BETA = 1.0 / (0.0083115 * 300.0);  // gcmc/cpp/gcmc.h:40-41
constexpr double KCOUL = 1388.431112;                // gcmc/cpp/gcmc_energy.h:137
constexpr float KCOUL_F = 1388.431112f;

enum Kind { K_INS = 0, K_DEL = 1, K_TRN = 2, K_ROT = 3 };

struct Species {
    int startRes, maxNum, nf, conf;
    int atomBase;          // first atom of slot 0 (slot j starts at atomBase + j*nf)
    int pad[3];
    double B;              // beta*muex + ln(L0 L1 L2 conc 6.023e-4)
    float4 tmpl[MAXF];     // centred template (x, y, z, charge)
    float kq[MAXF];        // 1388.431112 * charge, rounded once
    int trow[MAXF];        // type * Y: row of the pair table
    int ttype[MAXF];       // type
};

struct DevTrace {
    int max_steps, max_k, first;
    int *k, *ngroups, *naccept, *n_after;
    double *etot;
    float *centre;
    double *e;
    int *ninc;
    uint32_t *inchash, *posehash;
    int *idx, *gpick, *gacc;
    unsigned long long *tick;  // [S][4] publish, accept start, accept end, prep end (batched form)
};

// a fixed group of the cell list: centre and id, atom range (one 32 B
// record, so a candidate's two loads fall in one cache line)
struct __align__(16) FixRec {
    float4 p;                // (x, y, z, gid bits)
    int2 rng;                // (atomStart, atomNum)
    int2 pad;
};
static_assert(sizeof(FixRec) == 32, "FixRec layout");

struct Rep {
    AxisCut ax[3];
    float L[3], start[3], invL[3];
    float rc, r2max, dx, cav;
    float inv_h[3];
    int nc[3], ncell;
    int G, T, X, Y;
    int nfix, nres, natom, maxk;
    int cap, ovf_base, nst, nmoves;
    int coul64, prune;       // Coulomb pair terms in FP64; bit i: stencil cells may be
                             // pruned along axis i by their distance from the trial centre
    float cell_h[3];         // cell edges L / nc
    float prune_r2;          // (rc + stencil slack)^2
    uint32_t seed;
    int nowrap;              // pair displacements need no per-pair image shift: every environment
                             // atom of an included group and every fragment atom lie, per axis,
                             // less than L/2 apart after one image shift of the atom around the
                             // trial centre (rc + group extents < L/2, host-checked)
    int tab;                 // replicas with equal tab have equal pair tables (one copy in LDS serves them)
    int pad1;
    const float *grid;
    const int *moves;
    const uint32_t *wcum;    // [nmoves+1] estimated critical-path cost of moves [0, s), scaled (scheduling)
    float4 *res_pos;
    int2 *res_rng;
    float4 *atom;
    int *atype;
    const float4 *pair;      // (100 sigma^2 hi, 4 eps, 100 sigma^2 lo, 0) per (fragment type, type)
    const Species *sp;
    // per cell c in [0, ncell]: (first fixed record, mobile count); the fixed
    // records of cell c are [cell[c].x, cell[c + 1].x), and cell[ncell].y
    // counts the overflow bucket
    int2 *cell;
    const FixRec *fix;       // fixed groups sorted by cell
    int *mob_slot, *mob_loc;
    float4 *mob_rec;         // bucket entries (x, y, z, mob_word(gid, species) bits), parallel to mob_slot
    const int *stencil;      // packed (dx+128) | (dy+128)<<8 | (dz+128)<<16
    // per-step buffers
    float4 *tc;              // trial centres
    int *tslot;              // picked slot (delete/translate/rotate)
    double *E;               // [2*maxk] E (phase 0), E_new (phase 1)
    float4 *pose;            // [maxk*MAXF] new poses
    uint32_t *adj;           // [maxk*ADJW] trial-centre adjacency rows
    // Ewald electrostatics (ew != 0): pair terms K q q' erfc(beta r)/r and the
    // reciprocal potential pme_phi [pk[0]][pk[1]][pk[2]] of the current state
    int ew, pk[3];
    float ew_beta;
    int pme_cap;             // capacity of pme_delta
    const float *pme_phi;
    const float *pme_gk;     // grid kernel: potential grid of a unit charge at point 0
    float4 *pme_delta;       // charge changes of the last move (x, y, z, dq)
    int *pme_ndelta;
    DevTrace tr;
};

struct State {
    int step, end;                 // current step, exclusive end of this run
    int kind, sp, k, nitems;       // descriptor of the current step
    int done;                      // evaluations finished in the current step
    int err;
    unsigned long long claim;      // (epoch:24 | nitems:20 | claimed:20)
    double energy;                 // running FP64 sum of accepted changes
    int totalNum[MAXT];
    long long att[MAXT * 4], acc[MAXT * 4];
    // work counters: evaluations, fragment-environment pairs, included
    // environment atoms and groups, fragment atoms (summed over evaluations)
    unsigned long long n_eval, n_pair, n_atom, n_group, n_fatom;
    // 100 MHz wall-clock ticks spent in evaluations, accept and prep phases,
    // and between a move's publication and its last evaluation finishing
    unsigned long long t_eval, t_accept, t_prep, t_step;
    unsigned long long t_pub;
    // 100 MHz time by which every write the current move's evaluations read
    // (prep, the previous move's commit) had reached L2; set when it is published
    unsigned long long t_data;
    // sub-phases of evaluations (ticks, summed): pose+adjacency, stencil
    // count+scan, candidate tests, pair passes
    unsigned long long t_sub[4];
};

struct Sched {
    int finished;                  // replicas at the end of their run
    int abort;                     // set on a spin-limit expiry
    int progress;                  // completed moves (any replica)
    int pad;
};

// byte offsets of the dynamic LDS carve-up (host computes, device applies)
struct LdsLayout {
    int smob, sfst, sfix, soff;    // stencil entries (int[nst+2]): mobile-slot base, fixed start, fixed count, offset
    int incs, incn;                // included groups: first atom, atom count (int[inc_cap+1])
    int frag, frow;                // fragment pose (float4[MAXF], int[MAXF])
    int fragd;                     // fragment pose relative to the trial centre, FP64 (double4[MAXF])
    int pair;                      // float4[X*Y]: the pair table held in LDS, or -1
    int spt;                       // species table int4[MAXT] (startRes, maxNum, atomBase, nf)
    int misc, dred;                // int[MISC_INTS], double[MAXWAVE]
    int members, w, val, accl;     // accept: int[maxk], double[maxk] x2, int2[maxk]
    int adjl;                      // accept: adjacency rows uint32[maxk * ceil(maxk/32)] or -1
    int rsel, racc, ea;            // accept: double[maxk] each -- draws per group, exp terms per trial
    int J;                         // prep: int[maxk]
    int inc_cap;                   // included groups buffered before a pair pass
    int cand, amap, map_cap;       // uint16[map_cap] each: candidate -> stencil entry, pair-pass atom -> group
    int total;
};

// claim word of a replica: scheduling priority (24 bits), items published
// (20 bits), items claimed (20 bits)
__host__ __device__ inline unsigned long long claim_word(int prio, int nitems, int claimed)
{
    return ((unsigned long long)(uint32_t)(prio & 0xFFFFFF) << 40) |
           ((unsigned long long)(uint32_t)(nitems & 0xFFFFF) << 20) |
           (unsigned long long)(uint32_t)(claimed & 0xFFFFF);
}

// cell of a position (wrapped into the box), same float code on host and GPU
__host__ __device__ inline int cell_index(const float L[3], const float start[3],
                                          const float inv_h[3], const int nc[3],
                                          float x, float y, float z)
{
    float p[3] = {x, y, z};
    int c[3];
    for (int i = 0; i < 3; ++i) {
        float w = p[i] - start[i];
        w = w - floorf(w / L[i]) * L[i];
        int ci = (int)(w * inv_h[i]);
        c[i] = ci < 0 ? 0 : (ci >= nc[i] ? nc[i] - 1 : ci);
    }
    return (c[0] * nc[1] + c[1]) * nc[2] + c[2];
}

// position of (x, y, z) inside its cell c[3] (cell_index's wrap), per axis
__host__ __device__ inline void cell_offset(const float L[3], const float start[3], const float h[3],
                                            const int c[3], float x, float y, float z, float u[3])
{
    const float p[3] = {x, y, z};
    for (int i = 0; i < 3; ++i) {
        float w = p[i] - start[i];
        w = w - floorf(w / L[i]) * L[i];
        u[i] = w - (float)c[i] * h[i];
    }
}

// Whether the cell at stencil offset d[3] from a centre at position u[3]
// inside its cell may hold a group centre within the cutoff: false when the
// cell's box lies farther than sqrt(r2) (rc + slack) from the centre along
// the axes whose bit is set in `prune`, true otherwise.
__host__ __device__ inline bool stencil_cell_near(const float u[3], const int d[3], const float h[3], int prune,
                                                  float r2)
{
    float g2 = 0.0f;
    for (int i = 0; i < 3; ++i) {
        if (!((prune >> i) & 1)) continue;
        const float g = d[i] > 0 ? (float)d[i] * h[i] - u[i] : (d[i] < 0 ? u[i] - (float)(d[i] + 1) * h[i] : 0.0f);
        g2 = fmaf(g, g, g2);
    }
    return g2 <= r2;
}

}  // namespace gcmc
