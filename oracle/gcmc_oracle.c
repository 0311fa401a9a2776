/*
 * gcmc_oracle.c -- CPU FP64 restatement of pyGCMC's GCMC hot path.
 *
 * TEST INFRASTRUCTURE. This file is the checker, never the product: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it
 * (through oracle/oracle.py).  Nothing under pygcmc_amd/ links or calls it.
 *
 * What it restates (reference paths relative to the pyGCMC tree):
 *   struct ABI ............ gcmc/cpp/gcmc.h:47-100
 *   constants ............. gcmc/cpp/gcmc.h:25-41 (beta, MOLES_TO_MOLECULES)
 *   fast_round/distanceP .. gcmc/cpp/gcmc_energy.h:110-122
 *   LJ / Coulomb pair ..... gcmc/cpp/gcmc_energy.h:126-138
 *   calcEnergy ............ gcmc/cpp/gcmc_energy.h:142-204 (group-centre cutoff,
 *                           fixed groups [0, frag[0].startRes), then each
 *                           species' active slots minus the moving residue)
 *   Rodrigues rotation .... gcmc/cpp/gcmc_energy.h:29-72
 *   insertion pose ........ gcmc/cpp/gcmc_energy.h:75-107 (randomFragment)
 *   translate ............. gcmc/cpp/gcmc_move_trn.h:15-89
 *   rotate (centroid) ..... gcmc/cpp/gcmc_move_rot.h:15-115
 *   delete ................ gcmc/cpp/gcmc_move_del.h:15-52, 97-187
 *   selection/Metropolis .. gcmc/cpp/gcmc_move_add.h:86-201,
 *                           gcmc_move_trn.h:125-253, gcmc_move_rot.h:151-275
 *   update semantics ...... GupdateAdd/Del/Trn/Rot (gcmc_move_*.h)
 *   driver loop ........... gcmc/cpp/gcmc.cu:77-98 (move = type*4 + kind)
 *   pair-table rule ....... gcmc/python/simulation.py:63-72 (sigma mean,
 *                           epsilon geometric mean) -- used by orc_adjoint,
 *                           whose derivatives have no reference counterpart
 *
 * Declared deviations, identical in the product (both follow one definition):
 *   - variates come from Philox4x32-10 (Salmon et al., SC'11) keyed by
 *     (seed, 0) with counter (step, trial, purpose, 0) instead of cuRAND
 *     XORWOW + glibc rand(); the reference streams cannot be reproduced.
 *   - trial selection order is ascending trial index (the reference iterates
 *     a libstdc++ unordered_set, an implementation artefact).
 *   - sin/cos of the rotation angle use a Cody-Waite reduced Taylor
 *     polynomial evaluated with explicit fmaf, so poses are bit-identical
 *     between this file and the GPU.
 *   - acceptance arithmetic is FP64 throughout (the reference mixes float
 *     and double); a delete on an empty species is rejected (the reference
 *     divides by zero, gcmc_move_del.h:113).
 *
 * Energies are FP64 from the FP32 coordinates/parameters; the group inclusion
 * mask is the reference's FP32 distanceP test, bit for bit.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ---- struct ABI (gcmc/cpp/gcmc.h:47-100) ---------------------------- */
typedef struct { float position[3]; float charge; int type; } orc_atom;
typedef struct {
    char name[4]; int startRes; float muex; float conc; int confBias;
    float mcTime; int totalNum; int maxNum; int num_atoms; orc_atom atoms[20];
} orc_frag;
typedef struct {
    int mcsteps; float cutoff; float grid_dx; float startxyz[3]; float cryst[3];
    int showInfo; float cavityFactor; int fragTypeNum; int totalGridNum;
    int totalResNum; int totalAtomNum; int ffXNum; int ffYNum; int PME;
    unsigned int seed;
} orc_info;
typedef struct { float position[3]; int atomNum; int atomStart; } orc_res;

_Static_assert(sizeof(orc_atom) == 20, "Atom is 20 bytes");
_Static_assert(sizeof(orc_frag) == 436, "AtomArray is 436 bytes");
_Static_assert(sizeof(orc_info) == 76, "InfoStruct is 76 bytes");
_Static_assert(sizeof(orc_res) == 20, "residue is 20 bytes");

/* gcmc/cpp/gcmc.h:38-41 */
#define ORC_MOLES_TO_MOLECULES 0.0006023
#define ORC_BETA (1.0 / (0.0083115 * 300.0))
#define ORC_KCOUL 1388.431112 /* gcmc_energy.h:137 */

/* ---- Philox4x32-10 (Salmon, Moraes, Dror, Shaw; SC'11) ------------------ */
#define PH_M0 0xD2511F53u
#define PH_M1 0xCD9E8D57u
#define PH_W0 0x9E3779B9u
#define PH_W1 0xBB67AE85u

void orc_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int round = 0; round < 10; ++round) {
        if (round > 0) { k0 += PH_W0; k1 += PH_W1; }
        uint64_t p0 = (uint64_t)PH_M0 * c0;
        uint64_t p1 = (uint64_t)PH_M1 * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static void orc_draw(uint32_t seed, uint32_t step, uint32_t t, uint32_t purpose, uint32_t out[4])
{
    uint32_t ctr[4] = {step, t, purpose, 0u};
    uint32_t key[2] = {seed, 0u};
    orc_philox4x32_10(ctr, key, out);
}

/* (0,1] with 24 bits, the range of curand_uniform */
float orc_u01f(uint32_t x) { return (float)((x >> 8) + 1u) * 0x1p-24f; }
/* (0,1) with 52 bits: (2m+1) 2^-53 */
double orc_u01d(uint32_t a, uint32_t b)
{
    uint64_t m = ((uint64_t)a << 20) | (uint64_t)(b >> 12);
    return (double)(2u * m + 1u) * 0x1p-53;
}
/* integer in [0, n) by multiply-high */
uint32_t orc_below(uint32_t x, uint32_t n) { return (uint32_t)(((uint64_t)x * n) >> 32); }

/* ---- sin/cos: Cody-Waite reduction by pi/2 + Taylor polynomials -------- */
void orc_sincosf(float x, float *s, float *c)
{
    float k = rintf(x * 0x1.45f306p-1f);            /* 2/pi */
    int q = (int)k;
    float r = fmaf(-k, 0x1.921fb6p+0f, x);          /* pi/2, high part */
    r = fmaf(-k, -0x1.777a5cp-25f, r);              /* pi/2, low part  */
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

/* ---- FP32 distance test, gcmc_energy.h:110-122 -------------------------- */
static float orc_fast_round(float a)
{
    return a >= 0.0f ? (float)(int)(a + 0.5f) : (float)(int)(a - 0.5f);
}

float orc_distanceP32(const float x[3], const float y[3], const float L[3])
{
    float d[3];
    for (int i = 0; i < 3; ++i) {
        d[i] = x[i] - y[i];
        float t = orc_fast_round(d[i] / L[i]) * L[i];
        d[i] = d[i] - t;
    }
    float r2 = fmaf(d[2], d[2], fmaf(d[1], d[1], d[0] * d[0]));
    return sqrtf(r2);
}

static int orc_within(const float x[3], const float y[3], const float L[3], float rc)
{
    return orc_distanceP32(x, y, L) <= rc;
}

/* ---- rotation, gcmc_energy.h:29-72 (Rodrigues, positive-octant axis) --- */
void orc_rotmat(const float a[3], float phi, float R[9])
{
    float nrm = sqrtf(fmaf(a[2], a[2], fmaf(a[1], a[1], a[0] * a[0])));
    float n0 = a[0] / nrm, n1 = a[1] / nrm, n2 = a[2] / nrm;
    float s, c;
    orc_sincosf(phi, &s, &c);
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

static void orc_apply(const float R[9], const float x[3], float y[3])
{
    for (int j = 0; j < 3; ++j)
        y[j] = fmaf(x[2], R[3 * j + 2], fmaf(x[1], R[3 * j + 1], x[0] * R[3 * j + 0]));
}

/* Error codes of orc_run / orc_run_c (oracle.py raises on any of them). */
#define ORC_E_STEPS (-1)  /* [first, first + nsteps) outside the move schedule */
#define ORC_E_ALLOC (-2)  /* trial buffers could not be allocated */
#define ORC_E_MOVE (-3)   /* a move code names a species outside fragTypeNum */
#define ORC_E_K (-4)      /* a move's trial count exceeds ORC_MAXK or ORC_MAXF atoms */
#define ORC_E_FOLLOW (-5) /* the followed trace is shorter or narrower than the run */
#define ORC_E_STATE (-6)  /* a species count outside [0, maxNum] */
#define ORC_E_ALLOC_CELLS (-7) /* the cell list could not be allocated */

/* ---- system view -------------------------------------------------------- */
/* Optional cell list over the centres of the active groups (a CPU timing
   variant; the all-groups scan of calcEnergy is the default).  Cells are at
   least rc + 0.05 A wide, so every group whose centre passes the FP32 test
   lies in the 3 x 3 x 3 block of cells around the trial centre's cell (the
   whole axis when it has fewer than 3 cells). */
typedef struct {
    int nc[3], ncell;
    float h[3];
    int *start;   /* [ncell + 1] */
    int *group;   /* group ids sorted by cell */
    int cap;      /* capacity of group */
} orc_cells;

typedef struct {
    const orc_info *info;
    orc_frag *frag;
    orc_res *res;
    orc_atom *atom;
    const float *grid;
    const float *ff;
    const int *moves;
    orc_cells *cells;   /* NULL: scan every group */
} orc_sys;

typedef struct {
    int max_steps, max_k;
    int *k, *ngroups, *naccept, *n_after;
    double *etot;
    float *centre;        /* [S][K][3] */
    double *e;            /* [S][K][2] */
    int *ninc;            /* [S][K][2] */
    uint32_t *inchash;    /* [S][K][2] */
    uint32_t *posehash;   /* [S][K]    */
    int *idx;             /* [S][K]    */
    int *gpick;           /* [S][K]    */
    int *gacc;            /* [S][K]    */
} orc_trace;

/* hash of an included-group set (order independent) */
static uint32_t orc_mix(uint32_t g)
{
    uint32_t h = (g + 1u) * 0x9E3779B1u;
    h ^= h >> 15;
    h *= 0x85EBCA77u;
    h ^= h >> 13;
    return h;
}

static uint32_t orc_posehash(int nf, const float (*p)[3])
{
    uint32_t h = 2166136261u;
    for (int i = 0; i < nf; ++i)
        for (int j = 0; j < 3; ++j) {
            uint32_t b;
            memcpy(&b, &p[i][j], 4);
            h = (h ^ b) * 16777619u;
        }
    return h;
}

/* one pair, FP64 from FP32 inputs: gcmc_energy.h:126-138, 159-162 */
static inline double orc_pair(const orc_sys *S, const orc_atom *a, const float f[3],
                              float qf, int tf)
{
    const float *L = S->info->cryst;
    double r2 = 0.0;
    for (int i = 0; i < 3; ++i) {
        double Li = (double)L[i];
        double di = (double)a->position[i] - (double)f[i];
        double q = di / Li;
        double rr = q >= 0.0 ? floor(q + 0.5) : ceil(q - 0.5);
        di -= rr * Li;
        r2 += di * di;
    }
    int Y = S->info->ffYNum;
    int ti = tf * Y + a->type;
    double sig = (double)S->ff[2 * ti], eps = (double)S->ff[2 * ti + 1];
    double s2 = sig * sig * 100.0 / r2;
    double s6 = s2 * s2 * s2;
    double lj = 4.0 * eps * (s6 * s6 - s6);
    double el = ORC_KCOUL * ((double)qf * (double)a->charge) / sqrt(r2);
    return lj + el;
}

static void orc_group_energy(const orc_sys *S, int g, const float c[3], int nf,
                             const float (*pose)[3], const orc_atom *tmpl,
                             double *e, int *ninc, uint32_t *hash)
{
    const orc_res *r = &S->res[g];
    if (!orc_within(c, r->position, S->info->cryst, S->info->cutoff)) return;
    *ninc += 1;
    *hash += orc_mix((uint32_t)g);
    for (int a = r->atomStart; a < r->atomStart + r->atomNum; ++a)
        for (int f = 0; f < nf; ++f)
            *e += orc_pair(S, &S->atom[a], pose[f], tmpl[f].charge, tmpl[f].type);
}

static int orc_cell_of(const orc_cells *CL, const orc_info *info, const float p[3], int ci[3])
{
    for (int i = 0; i < 3; ++i) {
        double w = (double)p[i] - (double)info->startxyz[i];
        const double L = (double)info->cryst[i];
        w -= floor(w / L) * L;
        int c = (int)(w / (double)CL->h[i]);
        ci[i] = c < 0 ? 0 : (c >= CL->nc[i] ? CL->nc[i] - 1 : c);
    }
    return (ci[0] * CL->nc[1] + ci[1]) * CL->nc[2] + ci[2];
}

/* bins the fixed groups and every species' active slots (a cell list of the
   current state) */
static int orc_cells_build(const orc_sys *S)
{
    orc_cells *CL = S->cells;
    const orc_info *info = S->info;
    const int nfix = S->frag[0].startRes;
    int n = nfix;
    for (int s = 0; s < info->fragTypeNum; ++s) n += S->frag[s].totalNum;
    if (n > CL->cap) {
        free(CL->group);
        CL->group = (int *)malloc(sizeof(int) * (size_t)n);
        if (!CL->group) { CL->cap = 0; return ORC_E_ALLOC_CELLS; }
        CL->cap = n;
    }
    int *cellof = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int *gid = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    if (!cellof || !gid) { free(cellof); free(gid); return ORC_E_ALLOC_CELLS; }
    int m = 0;
    for (int g = 0; g < nfix; ++g) gid[m++] = g;
    for (int s = 0; s < info->fragTypeNum; ++s)
        for (int j = 0; j < S->frag[s].totalNum; ++j) gid[m++] = S->frag[s].startRes + j;
    memset(CL->start, 0, sizeof(int) * (size_t)(CL->ncell + 1));
    for (int q = 0; q < m; ++q) {
        int ci[3];
        cellof[q] = orc_cell_of(CL, info, S->res[gid[q]].position, ci);
        CL->start[cellof[q] + 1] += 1;
    }
    for (int c = 0; c < CL->ncell; ++c) CL->start[c + 1] += CL->start[c];
    int *fill = (int *)malloc(sizeof(int) * (size_t)CL->ncell);
    if (!fill) { free(cellof); free(gid); return ORC_E_ALLOC_CELLS; }
    memcpy(fill, CL->start, sizeof(int) * (size_t)CL->ncell);
    for (int q = 0; q < m; ++q) CL->group[fill[cellof[q]]++] = gid[q];
    free(fill);
    free(cellof);
    free(gid);
    return 0;
}

/* calcEnergy, gcmc_energy.h:142-204; exclude = moving residue or -1 */
double orc_energy(const orc_sys *S, const float c[3], int species,
                  const float (*pose)[3], int exclude, int *ninc, uint32_t *hash)
{
    const orc_frag *F = &S->frag[species];
    double e = 0.0;
    int n = 0;
    uint32_t h = 0;
    if (S->cells) {
        /* timing variant: the groups of the cells around c, same test */
        const orc_cells *CL = S->cells;
        int ci[3], lst[3][3], cnt[3];
        orc_cell_of(CL, S->info, c, ci);
        for (int i = 0; i < 3; ++i) {
            if (CL->nc[i] < 3) {
                cnt[i] = CL->nc[i];
                for (int j = 0; j < CL->nc[i]; ++j) lst[i][j] = j;
            } else {
                cnt[i] = 3;
                for (int j = 0; j < 3; ++j) lst[i][j] = (ci[i] + j - 1 + CL->nc[i]) % CL->nc[i];
            }
        }
        for (int a = 0; a < cnt[0]; ++a)
            for (int b = 0; b < cnt[1]; ++b)
                for (int d = 0; d < cnt[2]; ++d) {
                    const int cell = (lst[0][a] * CL->nc[1] + lst[1][b]) * CL->nc[2] + lst[2][d];
                    for (int q = CL->start[cell]; q < CL->start[cell + 1]; ++q) {
                        const int g = CL->group[q];
                        if (g != exclude) orc_group_energy(S, g, c, F->num_atoms, pose, F->atoms, &e, &n, &h);
                    }
                }
        if (ninc) *ninc = n;
        if (hash) *hash = h;
        return e;
    }
    int nfix = S->frag[0].startRes;
    for (int g = 0; g < nfix; ++g)
        orc_group_energy(S, g, c, F->num_atoms, pose, F->atoms, &e, &n, &h);
    for (int s = 0; s < S->info->fragTypeNum; ++s) {
        int g0 = S->frag[s].startRes, g1 = g0 + S->frag[s].totalNum;
        for (int g = g0; g < g1; ++g)
            if (g != exclude)
                orc_group_energy(S, g, c, F->num_atoms, pose, F->atoms, &e, &n, &h);
    }
    if (ninc) *ninc = n;
    if (hash) *hash = h;
    return e;
}

/* ---- trial poses -------------------------------------------------------- */
/* randomFragment, gcmc_energy.h:75-107 */
void orc_insert_pose(const orc_sys *S, uint32_t seed, int step, int t, int species,
                     float c[3], float (*pose)[3])
{
    const orc_frag *F = &S->frag[species];
    uint32_t x[4], y[4];
    orc_draw(seed, (uint32_t)step, (uint32_t)t, 0u, x);
    orc_draw(seed, (uint32_t)step, (uint32_t)t, 1u, y);
    uint32_t g = orc_below(x[3], (uint32_t)S->info->totalGridNum);
    for (int i = 0; i < 3; ++i) {
        float jit = orc_u01f(x[i]) * S->info->grid_dx;
        c[i] = S->grid[3 * g + i] + jit;
    }
    float axis[3] = {orc_u01f(y[0]), orc_u01f(y[1]), orc_u01f(y[2])};
    float phi = orc_u01f(y[3]) * 0x1.921fb6p+2f;
    float R[9];
    orc_rotmat(axis, phi, R);
    for (int a = 0; a < F->num_atoms; ++a) {
        float v[3];
        orc_apply(R, F->atoms[a].position, v);
        for (int i = 0; i < 3; ++i) pose[a][i] = v[i] + c[i];
    }
}

/* Gmove_trn translation, gcmc_move_trn.h:54-75 */
void orc_trn_pose(const orc_sys *S, uint32_t seed, int step, int t, int species, int slot,
                  float c[3], float (*pose)[3])
{
    const orc_frag *F = &S->frag[species];
    const orc_res *r = &S->res[F->startRes + slot];
    const float *L = S->info->cryst, *st = S->info->startxyz;
    uint32_t x[4];
    orc_draw(seed, (uint32_t)step, (uint32_t)t, 0u, x);
    float rr[3];
    for (int i = 0; i < 3; ++i) {
        float ci = r->position[i];
        float d = orc_u01f(x[i]) * S->info->grid_dx;
        if (ci + d > L[i] + st[i]) d = d - L[i];
        if (ci + d < st[i]) d = d + L[i];
        rr[i] = d;
        c[i] = ci + d;
    }
    for (int a = 0; a < F->num_atoms; ++a)
        for (int i = 0; i < 3; ++i)
            pose[a][i] = S->atom[r->atomStart + a].position[i] + rr[i];
}

/* Gmove_rot rotation about the centroid, gcmc_move_rot.h:317-355 */
void orc_rot_pose(const orc_sys *S, uint32_t seed, int step, int t, int species, int slot,
                  float c[3], float (*pose)[3])
{
    const orc_frag *F = &S->frag[species];
    const orc_res *r = &S->res[F->startRes + slot];
    int nf = F->num_atoms;
    uint32_t y[4];
    orc_draw(seed, (uint32_t)step, (uint32_t)t, 1u, y);
    float axis[3] = {orc_u01f(y[0]), orc_u01f(y[1]), orc_u01f(y[2])};
    float phi = orc_u01f(y[3]) * 0x1.921fb6p+2f;
    float cen[3] = {0.0f, 0.0f, 0.0f};
    for (int i = 0; i < 3; ++i)
        for (int a = 0; a < nf; ++a)
            cen[i] = cen[i] + S->atom[r->atomStart + a].position[i] / (float)nf;
    float R[9];
    orc_rotmat(axis, phi, R);
    for (int a = 0; a < nf; ++a) {
        float u[3], v[3];
        for (int i = 0; i < 3; ++i) u[i] = S->atom[r->atomStart + a].position[i] - cen[i];
        orc_apply(R, u, v);
        for (int i = 0; i < 3; ++i) pose[a][i] = v[i] + cen[i];
    }
    for (int i = 0; i < 3; ++i) c[i] = r->position[i];
}

/* partial Fisher-Yates of n distinct slots out of N (gcmc_move_trn.h:135-149) */
void orc_picks(uint32_t seed, int step, int N, int n, int *out)
{
    int *nums = (int *)malloc(sizeof(int) * (size_t)(N > 0 ? N : 1));
    for (int i = 0; i < N; ++i) nums[i] = i;
    for (int i = 0; i < n; ++i) {
        uint32_t x[4];
        orc_draw(seed, (uint32_t)step, (uint32_t)i, 2u, x);
        int j = i + (int)orc_below(x[0], (uint32_t)(N - i));
        int tmp = nums[i]; nums[i] = nums[j]; nums[j] = tmp;
        out[i] = nums[i];
    }
    free(nums);
}

/* B = beta*muex + ln(nbar), gcmc_move_add.h:103-105 (FP64) */
double orc_species_B(const orc_info *info, const orc_frag *F)
{
    double nbar = (double)info->cryst[0] * (double)info->cryst[1] * (double)info->cryst[2]
                  * (double)F->conc * ORC_MOLES_TO_MOLECULES;
    return ORC_BETA * (double)F->muex + log(nbar);
}

/* ---- state updates (GupdateAdd/Del/Trn/Rot) ----------------------------- */
static void orc_write_pose(orc_sys *S, int species, int slot, const float c[3],
                           const float (*pose)[3], int set_qt)
{
    orc_frag *F = &S->frag[species];
    orc_res *r = &S->res[F->startRes + slot];
    for (int i = 0; i < 3; ++i) r->position[i] = c[i];
    for (int a = 0; a < F->num_atoms; ++a) {
        orc_atom *A = &S->atom[r->atomStart + a];
        for (int i = 0; i < 3; ++i) A->position[i] = pose[a][i];
        if (set_qt) { A->charge = F->atoms[a].charge; A->type = F->atoms[a].type; }
    }
}

static void orc_delete_slot(orc_sys *S, int species, int hole)
{
    orc_frag *F = &S->frag[species];
    F->totalNum -= 1;
    int last = F->totalNum;
    if (last == 0) return; /* GupdateDel returns early when the species empties */
    orc_res *h = &S->res[F->startRes + hole];
    orc_res *l = &S->res[F->startRes + last];
    for (int i = 0; i < 3; ++i) h->position[i] = l->position[i];
    for (int a = 0; a < h->atomNum; ++a)
        for (int i = 0; i < 3; ++i)
            S->atom[h->atomStart + a].position[i] = S->atom[l->atomStart + a].position[i];
}

/* ---- one MC step ------------------------------------------------------- */
#define ORC_MAXK 4096
#define ORC_MAXF 20

typedef struct {
    double e_tot;
    long long attempts[64][4];
    long long accepts[64][4];
    long long evals;      /* trial energy evaluations (translate/rotate count two) */
} orc_stats;

static inline int T_IDX(const orc_trace *T, int s, int k) { return s * T->max_k + k; }

/*
 * Runs steps [first, first+nsteps).  If follow != NULL the per-group picks
 * and accept flags are taken from follow (another engine's trace of the same
 * chain) while this engine records its own values in out; that keeps both
 * states in lock-step after a rare disagreement.  Steps are indexed in the
 * trace arrays from 0 (= first).
 */


/* Checks every argument the chain reads before any state changes: the steps
   against the schedule length, each move code against the species count,
   each species' count, atom count and trial count against the buffers, and a
   followed trace against the steps and trials it must cover. */
static int orc_check(const orc_sys *S, int nmoves, int first, int nsteps, const orc_trace *follow)
{
    const orc_info *info = S->info;
    if (first < 0 || nsteps < 0 || nmoves < 0 || (long long)first + nsteps > (long long)nmoves) return ORC_E_STEPS;
    if (info->fragTypeNum < 1 || info->fragTypeNum > 64) return ORC_E_MOVE;
    for (int s = 0; s < info->fragTypeNum; ++s) {
        const orc_frag *F = &S->frag[s];
        if (F->totalNum < 0 || F->totalNum > F->maxNum) return ORC_E_STATE;
        if (F->num_atoms < 1 || F->num_atoms > ORC_MAXF || F->confBias < 0 || F->confBias > ORC_MAXK)
            return ORC_E_K;
    }
    for (int i = first; i < first + nsteps; ++i) {
        const int mv = S->moves[i];
        if (mv < 0 || mv / 4 >= info->fragTypeNum) return ORC_E_MOVE;
    }
    if (follow) {
        if (follow->max_steps < nsteps) return ORC_E_FOLLOW;
        for (int i = first; i < first + nsteps; ++i)
            if (S->frag[S->moves[i] / 4].confBias > follow->max_k) return ORC_E_FOLLOW;
    }
    return 0;
}

int orc_run(orc_sys *S, uint32_t seed, int first, int nsteps, orc_trace *out,
            const orc_trace *follow, orc_stats *st, int nthreads)
{
    const orc_info *info = S->info;
    const float *L = info->cryst;
    float rc = info->cutoff;
    /* per-call trial buffers, shared by the OpenMP workers of the trial loop
       (each trial writes only its own entries) */
    typedef struct {
        float centres[ORC_MAXK][3];
        float poses[ORC_MAXK][ORC_MAXF][3];
        double e0[ORC_MAXK], e1[ORC_MAXK], w[ORC_MAXK], dtmp[ORC_MAXK];
        int ninc0[ORC_MAXK], ninc1[ORC_MAXK], slots[ORC_MAXK], members[ORC_MAXK];
        uint32_t ih0[ORC_MAXK], ih1[ORC_MAXK];
        unsigned char unused[ORC_MAXK];
    } orc_buf;
    orc_buf *bf = (orc_buf *)malloc(sizeof(orc_buf));
    if (!bf) return ORC_E_ALLOC;
    float (*centres)[3] = bf->centres;
    float (*poses)[ORC_MAXF][3] = bf->poses;
    double *e0 = bf->e0, *e1 = bf->e1, *w = bf->w;
    int *ninc0 = bf->ninc0, *ninc1 = bf->ninc1, *slots = bf->slots, *members = bf->members;
    uint32_t *ih0 = bf->ih0, *ih1 = bf->ih1;
    unsigned char *unused = bf->unused;
    int cells_dirty = 1;

    for (int si = 0; si < nsteps; ++si) {
        int step = first + si;
        int mv = S->moves[step];
        int sp = mv / 4, kind = mv % 4;
        orc_frag *F = &S->frag[sp];
        int N = F->totalNum;
        int nf = F->num_atoms;
        int k = 0;
        if (kind == 0) k = (N >= F->maxNum) ? 0 : F->confBias;
        else if (kind == 1) k = (N == 0) ? 0 : 1;
        else k = F->confBias < N ? F->confBias : N;
        if (k > ORC_MAXK) {
            free(bf);
            return ORC_E_K;
        }
        if (st) {
            st->attempts[sp][kind] += 1;
            st->evals += (kind >= 2) ? 2 * k : k;
        }
        if (S->cells && k > 0 && cells_dirty) {
            const int rc = orc_cells_build(S);
            if (rc) {
                free(bf);
                return rc;
            }
            cells_dirty = 0;
        }

        /* picks for delete / translate / rotate */
        if (kind != 0 && k > 0) orc_picks(seed, step, N, k, slots);

        /* trial evaluation (independent per trial) */
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : 1) if (nthreads > 1)
        for (int t = 0; t < k; ++t) {
            float c[3];
            float (*p)[3] = poses[t];
            if (kind == 0) {
                orc_insert_pose(S, seed, step, t, sp, c, p);
                e0[t] = orc_energy(S, c, sp, (const float (*)[3])p, -1, &ninc0[t], &ih0[t]);
                e1[t] = 0.0; ninc1[t] = 0; ih1[t] = 0;
            } else {
                int slot = slots[t];
                const orc_res *r = &S->res[F->startRes + slot];
                float cur[ORC_MAXF][3];
                for (int a = 0; a < nf; ++a)
                    for (int i = 0; i < 3; ++i) cur[a][i] = S->atom[r->atomStart + a].position[i];
                e0[t] = orc_energy(S, r->position, sp, (const float (*)[3])cur,
                                   F->startRes + slot, &ninc0[t], &ih0[t]);
                if (kind == 1) {
                    for (int i = 0; i < 3; ++i) c[i] = r->position[i];
                    memcpy(p, cur, sizeof(float) * 3 * (size_t)nf);
                    e1[t] = 0.0; ninc1[t] = 0; ih1[t] = 0;
                } else {
                    /* E_new is tested around the new centre (translate) or
                       the unchanged residue centre (rotate) */
                    if (kind == 2) orc_trn_pose(S, seed, step, t, sp, slot, c, p);
                    else orc_rot_pose(S, seed, step, t, sp, slot, c, p);
                    e1[t] = orc_energy(S, c, sp, (const float (*)[3])p,
                                       F->startRes + slot, &ninc1[t], &ih1[t]);
                }
            }
            centres[t][0] = c[0]; centres[t][1] = c[1]; centres[t][2] = c[2];
        }

        int ngroups = 0, nacc = 0;
        double B = orc_species_B(info, F);
        if (out && si < out->max_steps) {
            out->k[si] = k;
            for (int t = 0; t < k && t < out->max_k; ++t) {
                int q = T_IDX(out, si, t);
                out->centre[3 * q + 0] = centres[t][0];
                out->centre[3 * q + 1] = centres[t][1];
                out->centre[3 * q + 2] = centres[t][2];
                out->e[2 * q + 0] = e0[t];
                out->e[2 * q + 1] = e1[t];
                out->ninc[2 * q + 0] = ninc0[t];
                out->ninc[2 * q + 1] = ninc1[t];
                out->inchash[2 * q + 0] = ih0[t];
                out->inchash[2 * q + 1] = ih1[t];
                out->posehash[q] = orc_posehash(nf, (const float (*)[3])poses[t]);
                out->idx[q] = (kind == 0) ? -1 : slots[t];
            }
        }

        if (kind == 1) {
            if (k == 1) {
                uint32_t x[4];
                orc_draw(seed, (uint32_t)step, 0u, 3u, x);
                double ran = orc_u01d(x[2], x[3]);
                double E = e0[0];
                double pp = (double)N * exp(-B - ORC_BETA * (-E));
                double p = pp < 1.0 ? pp : 1.0;
                int acc = ran < p;
                int fpick = 0, facc = acc;
                if (follow) facc = follow->gacc[T_IDX(follow, si, 0)] != 0;
                if (out && si < out->max_steps) {
                    out->gpick[T_IDX(out, si, 0)] = 0;
                    out->gacc[T_IDX(out, si, 0)] = acc;
                }
                (void)fpick;
                ngroups = 1;
                if (facc) {
                    orc_delete_slot(S, sp, slots[0]);
                    cells_dirty = 1;
                    nacc = 1;
                    if (st) { st->e_tot -= E; st->accepts[sp][kind] += 1; }
                }
            }
        } else if (k > 0) {
            double *dE = e0; /* energies used for weights */
            double *dtmp = bf->dtmp;
            if (kind != 0) {
                for (int t = 0; t < k; ++t) dtmp[t] = e1[t] - e0[t];
                dE = dtmp;
            }
            for (int t = 0; t < k; ++t) unused[t] = 1;
            int cursor = 0;
            int ins_slots[ORC_MAXK];
            int n_ins = 0;
            int trial_acc[ORC_MAXK];
            for (int t = 0; t < k; ++t) trial_acc[t] = 0;
            while (1) {
                while (cursor < k && !unused[cursor]) ++cursor;
                if (cursor >= k) break;
                int it = cursor;
                unused[it] = 0;
                int m = 0;
                members[m++] = it;
                double emin = dE[it];
                for (int j = it + 1; j < k; ++j) {
                    if (unused[j] && orc_within(centres[it], centres[j], L, rc)) {
                        members[m++] = j;
                        unused[j] = 0;
                        if (dE[j] < emin) emin = dE[j];
                    }
                }
                double sumP = 0.0;
                for (int q = 0; q < m; ++q) {
                    w[q] = exp(-ORC_BETA * (dE[members[q]] - emin));
                    sumP += w[q];
                }
                uint32_t x[4];
                orc_draw(seed, (uint32_t)step, (uint32_t)ngroups, 3u, x);
                double ran_sel = orc_u01d(x[0], x[1]);
                double ran_acc = orc_u01d(x[2], x[3]);
                int pick = it, pq = 0;
                if (sumP != 0.0) {
                    double cum = 0.0;
                    pick = members[m - 1]; pq = m - 1;
                    for (int q = 0; q < m; ++q) {
                        cum += w[q] / sumP;
                        if (ran_sel < cum) { pick = members[q]; pq = q; break; }
                    }
                }
                double P = (sumP != 0.0) ? w[pq] / sumP : 1.0;
                double pp;
                if (kind == 0)
                    pp = ((double)info->cavityFactor / ((double)m * P)) / ((double)F->totalNum + 1.0)
                         * exp(B - ORC_BETA * dE[pick]);
                else
                    pp = (1.0 / ((double)m * P)) * exp(-ORC_BETA * dE[pick]);
                double p = pp < 1.0 ? pp : 1.0;
                int acc = ran_acc < p;
                int own_pick = pick, own_acc = acc;
                if (follow) {
                    pick = follow->gpick[T_IDX(follow, si, ngroups)];
                    acc = follow->gacc[T_IDX(follow, si, ngroups)] != 0;
                }
                int rec_acc = own_acc;
                if (acc) {
                    for (int j = 0; j < k; ++j)
                        if (unused[j] && orc_within(centres[pick], centres[j], L, rc)) unused[j] = 0;
                    if (kind == 0) {
                        if (F->totalNum < F->maxNum) {
                            ins_slots[n_ins] = pick;
                            slots[pick] = F->totalNum;
                            F->totalNum += 1;
                            n_ins++;
                            if (st) st->e_tot += dE[pick];
                            nacc++;
                        } else if (own_acc) {
                            rec_acc = 2;
                        }
                    } else {
                        trial_acc[pick] = 1;
                        if (st) st->e_tot += dE[pick];
                        nacc++;
                    }
                }
                if (out && si < out->max_steps && ngroups < out->max_k) {
                    out->gpick[T_IDX(out, si, ngroups)] = own_pick;
                    out->gacc[T_IDX(out, si, ngroups)] = rec_acc;
                }
                ngroups++;
            }
            /* commit (GupdateAdd / GupdateTrn / GupdateRot) */
            if (kind == 0) {
                for (int q = 0; q < n_ins; ++q) {
                    int t = ins_slots[q];
                    orc_write_pose(S, sp, slots[t], centres[t], (const float (*)[3])poses[t], 1);
                    if (out && si < out->max_steps && t < out->max_k) out->idx[T_IDX(out, si, t)] = slots[t];
                }
            } else {
                for (int t = 0; t < k; ++t)
                    if (trial_acc[t]) {
                        const float *cc = centres[t];
                        orc_write_pose(S, sp, slots[t], cc, (const float (*)[3])poses[t], 0);
                    }
            }
            if (st) st->accepts[sp][kind] += nacc;
            if (nacc) cells_dirty = 1;
        }
        if (out && si < out->max_steps) {
            out->ngroups[si] = ngroups;
            out->naccept[si] = nacc;
            out->n_after[si] = F->totalNum;
            out->etot[si] = st ? st->e_tot : 0.0;
        }
    }
    free(bf);
    return 0;
}

/* single evaluation entry for tests: energy of a pose around a centre */
double orc_energy_pose(const orc_info *info, const orc_frag *frag, const orc_res *res,
                       const orc_atom *atom, const float *ff, int species,
                       const float *centre, const float *pose, int exclude,
                       int *ninc, uint32_t *hash)
{
    orc_sys S = {info, (orc_frag *)frag, (orc_res *)res, (orc_atom *)atom, NULL, ff, NULL, NULL};
    return orc_energy(&S, centre, species, (const float (*)[3])pose, exclude, ninc, hash);
}

/* nmoves = entries of moves[]; runs steps [first, first + nsteps) of it.
   cells != 0 selects the cell-list scan (a timing variant; same inclusion
   sets, energies summed in another order). */
int orc_run_c(const orc_info *info, orc_frag *frag, orc_res *res, orc_atom *atom,
              const float *grid, const float *ff, const int *moves, int nmoves, uint32_t seed,
              int first, int nsteps, orc_trace *out, const orc_trace *follow,
              orc_stats *st, int nthreads, int cells)
{
    orc_sys S = {info, frag, res, atom, grid, ff, moves, NULL};
    int rc = orc_check(&S, nmoves, first, nsteps, follow);
    if (rc) return rc;
    orc_cells CL;
    memset(&CL, 0, sizeof(CL));
    if (cells) {
        CL.ncell = 1;
        for (int i = 0; i < 3; ++i) {
            const float L = info->cryst[i];
            int n = (int)floor((double)L / ((double)info->cutoff + 0.05));
            CL.nc[i] = n < 1 ? 1 : n;
            CL.h[i] = L / (float)CL.nc[i];
            CL.ncell *= CL.nc[i];
        }
        CL.start = (int *)malloc(sizeof(int) * (size_t)(CL.ncell + 1));
        if (!CL.start) return ORC_E_ALLOC_CELLS;
        S.cells = &CL;
    }
    rc = orc_run(&S, seed, first, nsteps, out, follow, st, nthreads);
    free(CL.start);
    free(CL.group);
    return rc;
}

/* the trial pose of trial t of move `step` (kind 0 insert, 2 translate of
   slot, 3 rotate of slot) and its test centre, for tests of the pose
   generators */
int orc_trial_pose(const orc_info *info, const orc_frag *frag, const orc_res *res, const orc_atom *atom,
                   const float *grid, uint32_t seed, int step, int t, int species, int kind, int slot,
                   float *centre, float *pose)
{
    orc_sys S = {info, (orc_frag *)frag, (orc_res *)res, (orc_atom *)atom, grid, NULL, NULL, NULL};
    if (species < 0 || species >= info->fragTypeNum) return ORC_E_MOVE;
    if (kind != 0 && (slot < 0 || slot >= frag[species].totalNum)) return ORC_E_STATE;
    float (*p)[3] = (float (*)[3])pose;
    if (kind == 0) orc_insert_pose(&S, seed, step, t, species, centre, p);
    else if (kind == 2) orc_trn_pose(&S, seed, step, t, species, slot, centre, p);
    else if (kind == 3) orc_rot_pose(&S, seed, step, t, species, slot, centre, p);
    else return ORC_E_MOVE;
    return 0;
}

uint32_t orc_group_hash_term(uint32_t g) { return orc_mix(g); }
uint32_t orc_pose_hash(int nf, const float *p) { return orc_posehash(nf, (const float (*)[3])p); }

/* ---- adjoint of a pose energy (no reference counterpart) ------------------
 * E(pose; eps, rmin, dq) with per-type parameters: type t has CHARMM
 * epsilon eps[t] (kcal/mol) and Rmin/2 rmin[t] (A), i.e. sigma_t =
 * rmin[t] 2^(5/6)/10 nm and epsilon_t = 4.184 eps[t] kJ/mol, combined as in
 * gcmc/python/simulation.py:63-72; every atom of type t carries its own
 * charge plus dq[t].  Groups are included by the FP32 group-centre test of
 * calcEnergy.  Outputs (any may be NULL): E, dE/dx of the fragment atoms
 * [nf][3], and dE/d dq[t], dE/d eps[t], dE/d rmin[t] over Y types, all FP64.
 */
static void orc_adj_group(const orc_sys *S, int g, const float c[3], int nf,
                          const float (*pose)[3], const orc_atom *tmpl,
                          const double *sig, const double *epk, const double *eps,
                          const double *dq, double *E, double *gx, double *gq,
                          double *ge, double *gr)
{
    const orc_res *r = &S->res[g];
    if (!orc_within(c, r->position, S->info->cryst, S->info->cutoff)) return;
    const float *L = S->info->cryst;
    const double cs = pow(2.0, 5.0 / 6.0) / 10.0;
    for (int a = r->atomStart; a < r->atomStart + r->atomNum; ++a) {
        const orc_atom *A = &S->atom[a];
        const int ta = A->type;
        const double qa = (double)A->charge + dq[ta];
        for (int f = 0; f < nf; ++f) {
            const int tf = tmpl[f].type;
            const double qf = (double)tmpl[f].charge + dq[tf];
            double d[3], r2 = 0.0;
            for (int i = 0; i < 3; ++i) {
                double Li = (double)L[i];
                double di = (double)A->position[i] - (double)pose[f][i];
                double q = di / Li;
                double rr = q >= 0.0 ? floor(q + 0.5) : ceil(q - 0.5);
                d[i] = di - rr * Li;
                r2 += d[i] * d[i];
            }
            const double sg = 0.5 * (sig[tf] + sig[ta]);
            const double ep = sqrt(eps[tf] * eps[ta]);
            const double s2 = sg * sg * 100.0 / r2;
            const double s6 = s2 * s2 * s2, s12 = s6 * s6;
            const double lj = 4.0 * ep * (s12 - s6);
            const double ir = 1.0 / sqrt(r2);
            const double el = ORC_KCOUL * qf * qa * ir;
            *E += lj + el;
            if (gx) {
                /* dE/d(r^2); d = x_a - x_f so dE/dx_f = -2 dE/dr2 d */
                const double dr2 = 4.0 * ep * (-6.0 * s12 + 3.0 * s6) / r2 - 0.5 * el / r2;
                for (int i = 0; i < 3; ++i) gx[3 * f + i] += -2.0 * dr2 * d[i];
            }
            if (gq) {
                gq[tf] += ORC_KCOUL * qa * ir;
                gq[ta] += ORC_KCOUL * qf * ir;
            }
            if (ge) {
                /* d eps_pair / d eps[t] = eps_pair / (2 eps[t]) per matching side */
                if (epk[tf] > 0.0) ge[tf] += lj / (2.0 * epk[tf]);
                if (epk[ta] > 0.0) ge[ta] += lj / (2.0 * epk[ta]);
            }
            if (gr) {
                /* dE/dsigma_pair * dsigma_pair/drmin[t] = ... * cs/2 per side */
                const double ds = 4.0 * ep * (12.0 * s12 - 6.0 * s6) / sg * 0.5 * cs;
                gr[tf] += ds;
                gr[ta] += ds;
            }
        }
    }
}

double orc_adjoint(const orc_info *info, const orc_frag *frag, const orc_res *res,
                   const orc_atom *atom, int species, const float *centre, const float *pose,
                   int exclude, const double *eps_kcal, const double *rmin_half,
                   const double *dq, double *gx, double *gq, double *ge, double *gr)
{
    orc_sys S = {info, (orc_frag *)frag, (orc_res *)res, (orc_atom *)atom, NULL, NULL, NULL, NULL};
    const int Y = info->ffYNum;
    double *sig = (double *)malloc(sizeof(double) * 3 * (size_t)Y);
    double *eps = sig + Y, *epk = sig + 2 * Y;
    const double cs = pow(2.0, 5.0 / 6.0) / 10.0;
    for (int t = 0; t < Y; ++t) {
        sig[t] = rmin_half[t] * cs;
        epk[t] = fabs(eps_kcal[t]);
        eps[t] = 4.184 * epk[t];
    }
    const orc_frag *F = &frag[species];
    const int nf = F->num_atoms;
    if (gx) memset(gx, 0, sizeof(double) * 3 * (size_t)nf);
    if (gq) memset(gq, 0, sizeof(double) * (size_t)Y);
    if (ge) memset(ge, 0, sizeof(double) * (size_t)Y);
    if (gr) memset(gr, 0, sizeof(double) * (size_t)Y);
    double E = 0.0;
    const float (*P)[3] = (const float (*)[3])pose;
    const int nfix = frag[0].startRes;
    for (int g = 0; g < nfix; ++g)
        orc_adj_group(&S, g, centre, nf, P, F->atoms, sig, epk, eps, dq, &E, gx, gq, ge, gr);
    for (int s = 0; s < info->fragTypeNum; ++s) {
        const int g0 = frag[s].startRes, g1 = g0 + frag[s].totalNum;
        for (int g = g0; g < g1; ++g)
            if (g != exclude)
                orc_adj_group(&S, g, centre, nf, P, F->atoms, sig, epk, eps, dq, &E, gx, gq, ge, gr);
    }
    free(sig);
    return E;
}
