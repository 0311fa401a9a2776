// engine.hip -- host side of the MI355X GCMC engine and its C ABI
// (include/pygcmc_hip.h).
//
// Replaces runGCMC_cuda (gcmc/cpp/gcmc.cu:21-118): instead of cudaMalloc +
// copies + a host loop with synchronous copies around every launch, systems
// are uploaded once into a structure-of-arrays layout with a cell list over
// group centres, and moves run with device-side decisions, either one move per
// launch pair (plain) or every system in one persistent kernel (batched).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pygcmc_hip.h"
#include "adjoint_kernels.h"

static_assert(sizeof(Atom) == 20, "Atom ABI");
static_assert(sizeof(AtomArray) == 436, "AtomArray ABI");
static_assert(sizeof(InfoStruct) == 76, "InfoStruct ABI");
static_assert(sizeof(residue) == 20, "residue ABI");

using namespace gcmc;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define HIPX(expr)                                                                         \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return fail(GCMC_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));   \
    } while (0)

uint32_t fbits(float f)
{
    uint32_t b;
    std::memcpy(&b, &f, 4);
    return b;
}
float bfloat(uint32_t b)
{
    float f;
    std::memcpy(&f, &b, 4);
    return f;
}

// smallest non-negative float d with fast_round(d / L) >= m
float shift_threshold(float L, int m)
{
    uint32_t lo = 0, hi = fbits(8.0f * L);
    while (lo < hi) {
        uint32_t mid = lo + (hi - lo) / 2;
        float d = bfloat(mid);
        if (fast_round_ref(d / L) >= (float)m) hi = mid;
        else lo = mid + 1;
    }
    return bfloat(lo);
}

// largest float x with sqrtf(x) <= rc
float r2_threshold(float rc)
{
    float x = rc * rc;
    while (x > 0.0f && sqrtf(x) > rc) x = std::nextafter(x, 0.0f);
    while (true) {
        float nx = std::nextafter(x, INFINITY);
        if (sqrtf(nx) <= rc) x = nx;
        else break;
    }
    return x;
}

struct DevBuf {
    std::vector<void *> ptrs;
    template <class T>
    hipError_t alloc(T **p, size_t n)
    {
        void *q = nullptr;
        hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
        if (e == hipSuccess) {
            ptrs.push_back(q);
            *p = (T *)q;
        }
        return e;
    }
    void release()
    {
        for (void *p : ptrs) hipFree(p);
        ptrs.clear();
    }
};

struct HostRep {
    InfoStruct info;
    std::vector<AtomArray> frag;
    std::vector<residue> res;
    std::vector<Atom> atoms;
    std::vector<float> grid, ff;
    std::vector<int> moves;
    std::vector<float4> pair;  // the pair table as uploaded (Rep::tab: replicas with equal tables)
    Rep rep{};
    DevBuf buf;
    DevBuf trbuf;
    int maxk = 1;
    int host_step = 0;
    size_t nmslot = 0;         // entries of mob_slot / mob_rec
    double est_cost = 0.0;     // estimated cost of the whole schedule (cost model of the claim priority)
    unsigned long long n_eval_now = 0;  // the State's evaluation counter as the host last saw it
    unsigned long long snap_n_eval = 0;
    double last_work = 0.0;    // measured work of the last batched run: evaluations + moves x MOVE_WORK
    // device snapshot of the state the chain changes (gcmc_engine_snapshot)
    DevBuf snap;
    float4 *s_res_pos = nullptr, *s_atom = nullptr, *s_mob_rec = nullptr;
    int *s_atype = nullptr, *s_mob_slot = nullptr, *s_mob_loc = nullptr;
    int2 *s_cell = nullptr;
    int snap_step = 0;
    float ew_beta = 0.0f;      // > 0: Ewald electrostatics with PME
    float ew_spacing = 0.0f;
    gcmc_pme *pme = nullptr;
};

}  // namespace

// PME plan internals used by the Ewald mode (pme.hip)
namespace gcmc_internal {
const float *pme_phi(const gcmc_pme *p);
const float *pme_kernel(const gcmc_pme *p);
void pme_grid(const gcmc_pme *p, int K[3]);
int pme_spread_counted(gcmc_pme *p, int cap, const int *count, const float4 *xyzq, int *reset, hipStream_t s);
}  // namespace gcmc_internal

struct gcmc_engine {
    gcmc_tuning tune{};
    std::vector<HostRep *> reps;
    bool uploaded = false;
    Rep *d_rep = nullptr;
    State *d_state = nullptr;
    Sched *d_sched = nullptr;
    LdsLayout lay{};
    bool staged = false;
    int nwg = 0;               // persistent workgroups requested
    int ncu = 0;
    int resident = 0;          // persistent workgroups that fit on the device at once
    int maxk = 1;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, evk0 = nullptr, evk1 = nullptr;
    double last_ms = 0.0;
    double kernel_ms = 0.0;   // duration of the last persistent kernel (batched mode)
    bool timed_kernel = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int graph_nsteps = -1;
    // Ewald systems, graph form: a captured chunk of ew_graph_moves moves
    hipGraph_t ew_graph = nullptr;
    hipGraphExec_t ew_gexec = nullptr;
    int ew_graph_moves = 0;
    int spin_limit = 1 << 22;  // polls without progress anywhere (~4 s)
    int nt = 256;              // threads per workgroup
    XcdMap xm{};
    int *d_cls = nullptr;      // cls_start[nclass+1] then cls_reps[R]
    unsigned long long *d_cu_inv = nullptr;  // XcdMap::cu_inv
    State *d_state_snap = nullptr;  // State of every replica at gcmc_engine_snapshot
    std::vector<int> cls_host;      // the replica -> XCD class map on the device (cls_start, cls_reps)
    std::vector<int> cls_of;        // the same, class of each replica
    // Search over XCD assignments while runs repeat one step range (tuning
    // xcd_search): the fastest assignment so far with its launch time, the
    // per-class finish times and per-replica evaluation times measured under
    // it, the exchanges already tried from it, and the assignment of the next run
    struct XcdSearch {
        std::vector<int> range;     // (first step, end) per replica of the runs compared
        std::vector<int> best, next;
        double best_ms = 0.0;
        std::vector<double> fin, evt;
        std::vector<std::pair<int, int>> tried;
        int runs = 0;
    } xs;
    double *d_adj_work = nullptr;  // column-sum partials of the adjoint backward
    size_t adj_work_n = 0;
};

namespace {

int check_system(const InfoStruct *info, const AtomArray *frag, int nmoves)
{
    if (!info || !frag) return fail(GCMC_E_ARG, "null system pointer");
    if (info->fragTypeNum < 1 || info->fragTypeNum > MAXT)
        return fail(GCMC_E_ARG, "fragTypeNum must be in [1, 32]");
    if (info->cryst[0] <= 0 || info->cryst[1] <= 0 || info->cryst[2] <= 0)
        return fail(GCMC_E_ARG, "box edges must be positive");
    if (info->totalGridNum < 1) return fail(GCMC_E_ARG, "empty cavity grid");
    if (info->totalResNum < 0 || info->totalResNum > GID_MASK)
        return fail(GCMC_E_ARG, "totalResNum must be in [0, 2^24)");
    if (nmoves < 0) return fail(GCMC_E_ARG, "negative move count");
    for (int s = 0; s < info->fragTypeNum; ++s) {
        const AtomArray &F = frag[s];
        if (F.num_atoms < 1 || F.num_atoms > MAXF)
            return fail(GCMC_E_ARG, "fragment atom count must be in [1, 20]");
        if (F.confBias < 1 || F.confBias > MAXK)
            return fail(GCMC_E_ARG, "confBias must be in [1, 2048]");
        if (F.totalNum < 0 || F.totalNum > F.maxNum)
            return fail(GCMC_E_ARG, "totalNum outside [0, maxNum]");
        if (F.startRes < 0 || F.startRes + F.maxNum > info->totalResNum)
            return fail(GCMC_E_ARG, "species slots outside residueInfo");
    }
    return 0;
}

constexpr float STENCIL_EPS = 0.05f;  // A of slack over FP32 rounding of the group test

void cell_geometry(const float cryst[3], float cell_size, int nc[3], float inv_h[3])
{
    const float h = cell_size > 0 ? cell_size : 5.0f;
    for (int i = 0; i < 3; ++i) {
        nc[i] = std::max(1, (int)std::floor(cryst[i] / h));
        nc[i] = std::min(nc[i], 255);
        inv_h[i] = (float)nc[i] / cryst[i];
    }
}

// Trial-specific pruning of the stencil (stencil_cell_near): cell edges, the
// axes whose stencil offsets are the minimum-image ones (the stencil does not
// wrap onto itself there) and the squared reach rc + slack.
void stencil_prune(const float L[3], const int nc[3], float rc, float h[3], int *prune, float *r2)
{
    *prune = 0;
    for (int i = 0; i < 3; ++i) {
        h[i] = L[i] / (float)nc[i];
        const int r = (int)std::floor((rc + STENCIL_EPS) / ((double)L[i] / nc[i])) + 1;
        if (2 * r + 1 < nc[i]) *prune |= 1 << i;
    }
    *r2 = (rc + STENCIL_EPS) * (rc + STENCIL_EPS);
}

// cell stencil: offsets whose cell box lies within rc + eps of the centre cell
std::vector<int> make_stencil(const float L[3], const int nc[3], float rc, float eps)
{
    int lo[3], hi[3];
    bool full[3];
    double h[3];
    for (int i = 0; i < 3; ++i) {
        h[i] = (double)L[i] / nc[i];
        int r = (int)std::floor((rc + eps) / h[i]) + 1;
        full[i] = (2 * r + 1 >= nc[i]);
        if (full[i]) {
            lo[i] = -((nc[i] - 1) / 2);
            hi[i] = lo[i] + nc[i] - 1;
        } else {
            lo[i] = -r;
            hi[i] = r;
        }
    }
    const double lim = (double)(rc + eps) * (rc + eps);
    std::vector<int> out;
    for (int a = lo[0]; a <= hi[0]; ++a)
        for (int b = lo[1]; b <= hi[1]; ++b)
            for (int c = lo[2]; c <= hi[2]; ++c) {
                int d[3] = {a, b, c};
                double s = 0;
                for (int i = 0; i < 3; ++i) {
                    int n = std::abs(d[i]);
                    if (full[i]) n = std::min(n, nc[i] - n);
                    double g = std::max(0, n - 1) * h[i];
                    s += g * g;
                }
                if (s <= lim) out.push_back((a + 128) | ((b + 128) << 8) | ((c + 128) << 16));
            }
    return out;
}

template <class T>
hipError_t put(T **dst, DevBuf &b, const T *src, size_t n)
{
    hipError_t e = b.alloc(dst, n);
    if (e != hipSuccess) return e;
    if (n) e = hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice);
    return e;
}

// moves of the captured chunk of the graph form for Ewald systems
constexpr int EW_GRAPH_MOVES = 32;

template <class T>
hipError_t d2d(T *dst, const T *src, size_t n, hipStream_t s)
{
    return n ? hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToDevice, s) : hipSuccess;
}

int upload_rep(gcmc_engine *E, HostRep &H)
{
    const InfoStruct &I = H.info;
    Rep &R = H.rep;
    R = Rep{};
    const int T = I.fragTypeNum;
    for (int i = 0; i < 3; ++i) {
        R.L[i] = I.cryst[i];
        R.start[i] = I.startxyz[i];
        R.invL[i] = 1.0f / I.cryst[i];
        R.ax[i].L = I.cryst[i];
        R.ax[i].t[0] = shift_threshold(I.cryst[i], 1);
        R.ax[i].t[1] = shift_threshold(I.cryst[i], 2);
        R.ax[i].t[2] = shift_threshold(I.cryst[i], 3);
    }
    R.rc = I.cutoff;
    R.r2max = r2_threshold(I.cutoff);
    R.dx = I.grid_dx;
    R.cav = I.cavityFactor;
    R.G = I.totalGridNum;
    R.T = T;
    R.X = I.ffXNum;
    R.Y = I.ffYNum;
    R.nfix = H.frag[0].startRes;
    R.nres = I.totalResNum;
    R.natom = I.totalAtomNum;
    R.nmoves = (int)H.moves.size();
    R.coul64 = E->tune.fp32_coulomb ? 0 : 1;
    R.seed = I.seed;
    R.maxk = H.maxk;

    // cell list geometry
    cell_geometry(I.cryst, E->tune.cell_size, R.nc, R.inv_h);
    R.ncell = R.nc[0] * R.nc[1] * R.nc[2];
    const float h3 = (I.cryst[0] / R.nc[0]) * (I.cryst[1] / R.nc[1]) * (I.cryst[2] / R.nc[2]);
    int cap = E->tune.bucket_cap;
    if (cap <= 0) {
        int want = (int)std::ceil(0.05 * h3 * 1.5);
        cap = 8;
        while (cap < want) cap <<= 1;
    }
    R.cap = cap;
    int nmob = 0;
    for (int s = 0; s < T; ++s) nmob += H.frag[s].maxNum;
    R.ovf_base = R.ncell * cap;

    // structure-of-arrays copies
    std::vector<float4> res_pos(I.totalResNum);
    std::vector<int2> res_rng(I.totalResNum);
    for (int g = 0; g < I.totalResNum; ++g) {
        const residue &r = H.res[g];
        res_pos[g] = make_float4(r.position[0], r.position[1], r.position[2], 0.f);
        res_rng[g] = make_int2(r.atomStart, r.atomNum);
    }
    std::vector<float4> atom(I.totalAtomNum);
    std::vector<int> atype(I.totalAtomNum);
    // every atom the energy kernel can read must name a column of the pair table
    for (int g = 0; g < I.totalResNum; ++g) {
        bool active = g < R.nfix;
        for (int s = 0; s < T && !active; ++s)
            active = g >= H.frag[s].startRes && g < H.frag[s].startRes + H.frag[s].totalNum;
        if (!active) continue;
        const residue &r = H.res[g];
        if (r.atomStart < 0 || r.atomNum < 0 || r.atomStart + r.atomNum > I.totalAtomNum)
            return fail(GCMC_E_ARG, "residue atom range outside atomInfo");
        for (int a = r.atomStart; a < r.atomStart + r.atomNum; ++a)
            if (H.atoms[a].type < 0 || H.atoms[a].type >= R.Y)
                return fail(GCMC_E_ARG, "atom type outside the pair table columns");
    }
    for (int s = 0; s < T; ++s) {
        const int nf = H.frag[s].num_atoms, base = H.res[H.frag[s].startRes].atomStart;
        for (int j = 0; j < H.frag[s].maxNum; ++j) {
            const residue &r = H.res[H.frag[s].startRes + j];
            if (r.atomNum != nf || r.atomStart != base + j * nf || r.atomStart < 0 ||
                r.atomStart + r.atomNum > I.totalAtomNum)
                return fail(GCMC_E_ARG, "fragment slots must hold contiguous atom ranges of num_atoms atoms");
        }
    }
    for (int a = 0; a < I.totalAtomNum; ++a) {
        const Atom &A = H.atoms[a];
        atom[a] = make_float4(A.position[0], A.position[1], A.position[2], A.charge);
        atype[a] = A.type;
    }
    std::vector<float4> pair((size_t)R.X * R.Y);
    for (size_t i = 0; i < pair.size(); ++i) {
        // 100 sigma^2 as an unevaluated FP32 sum hi + lo (exact to ~2^-48);
        // 4 eps is exact in FP32
        const double sg = H.ff[2 * i], ep = H.ff[2 * i + 1];
        const double s2 = sg * sg * 100.0;
        const float hi = (float)s2;
        pair[i] = make_float4(hi, (float)(4.0 * ep), (float)(s2 - (double)hi), 0.0f);
    }
    std::vector<Species> sp(T);
    for (int s = 0; s < T; ++s) {
        const AtomArray &F = H.frag[s];
        Species &S = sp[s];
        std::memset(&S, 0, sizeof(S));
        S.startRes = F.startRes;
        S.maxNum = F.maxNum;
        S.nf = F.num_atoms;
        S.conf = F.confBias;
        S.atomBase = H.res[F.startRes].atomStart;
        double nbar = (double)I.cryst[0] * (double)I.cryst[1] * (double)I.cryst[2] *
                      (double)F.conc * 0.0006023;
        S.B = BETA * (double)F.muex + std::log(nbar);
        for (int a = 0; a < F.num_atoms; ++a) {
            const Atom &A = F.atoms[a];
            S.tmpl[a] = make_float4(A.position[0], A.position[1], A.position[2], A.charge);
            S.kq[a] = (float)(KCOUL * (double)A.charge);
            S.trow[a] = A.type * R.Y;
            S.ttype[a] = A.type;
            if (A.type < 0 || A.type >= R.X)
                return fail(GCMC_E_ARG, "fragment atom type outside the pair table rows");
        }
    }
    // Rep::nowrap.  An environment atom of an included group lies within
    // rc + r_env of the trial centre and a fragment atom within r_frag of it
    // (minimum image); when rc + r_env + r_frag < L/2 on every axis, the
    // atom's one image shift around the trial centre is also the pair's.
    // A fixed group's extent is its atoms' largest distance from its centre.
    // A molecule keeps its shape under translation and rotation (about its
    // centroid, centre unchanged), so its atoms stay within rho + D of its
    // centre: rho its atoms' largest distance from their centroid, D the
    // centroid's distance from the centre -- of the template for inserted
    // molecules, of the molecules present at upload.
    {
        auto mi = [&](double d, int i) { return d - (double)R.L[i] * std::nearbyint(d / (double)R.L[i]); };
        auto extent = [&](const double *c, int n, auto pos, bool about_centroid) {
            double cen[3] = {0.0, 0.0, 0.0}, rho = 0.0, D = 0.0;
            for (int a = 0; a < n; ++a)
                for (int i = 0; i < 3; ++i) cen[i] += mi(pos(a, i) - c[i], i) / n;
            for (int a = 0; a < n; ++a) {
                double d2 = 0.0;
                for (int i = 0; i < 3; ++i) {
                    const double d = mi(pos(a, i) - c[i], i) - (about_centroid ? cen[i] : 0.0);
                    d2 += d * d;
                }
                rho = std::max(rho, std::sqrt(d2));
            }
            if (about_centroid) D = std::sqrt(cen[0] * cen[0] + cen[1] * cen[1] + cen[2] * cen[2]);
            return rho + D;
        };
        auto atoms_of = [&](int g) {
            const residue &r = H.res[g];
            return [&H, &r](int a, int i) { return (double)H.atoms[r.atomStart + a].position[i]; };
        };
        double r_env = 0.0, r_frag = 0.0;
        for (int g = 0; g < R.nfix; ++g) {
            const residue &r = H.res[g];
            const double c[3] = {r.position[0], r.position[1], r.position[2]};
            r_env = std::max(r_env, extent(c, r.atomNum, atoms_of(g), false));
        }
        for (int s = 0; s < T; ++s) {
            const AtomArray &F = H.frag[s];
            const double zero[3] = {0.0, 0.0, 0.0};
            r_frag = std::max(r_frag, extent(zero, F.num_atoms,
                                             [&F](int a, int i) { return (double)F.atoms[a].position[i]; }, true));
            for (int j = 0; j < F.totalNum; ++j) {
                const residue &r = H.res[F.startRes + j];
                const double c[3] = {r.position[0], r.position[1], r.position[2]};
                r_frag = std::max(r_frag, extent(c, r.atomNum, atoms_of(F.startRes + j), true));
            }
        }
        r_env = std::max(r_env, r_frag);
        const double half = 0.5 * std::min({(double)R.L[0], (double)R.L[1], (double)R.L[2]});
        // 0.1 A covers FP32 rounding of positions and the drift of repeated rotations
        R.nowrap = (double)R.rc + r_env + r_frag + 0.1 < half ? 1 : 0;
    }
    // fixed groups: CSR by cell
    std::vector<int> fcell(R.nfix);
    std::vector<int> fstart(R.ncell + 1, 0);
    for (int g = 0; g < R.nfix; ++g) {
        const residue &r = H.res[g];
        fcell[g] = cell_index(R.L, R.start, R.inv_h, R.nc, r.position[0], r.position[1], r.position[2]);
        fstart[fcell[g] + 1] += 1;
    }
    for (int c = 0; c < R.ncell; ++c) fstart[c + 1] += fstart[c];
    std::vector<FixRec> frec(std::max(R.nfix, 1));
    {
        std::vector<int> fill(fstart.begin(), fstart.end() - 1);
        for (int g = 0; g < R.nfix; ++g) {
            const residue &r = H.res[g];
            int bits = g;
            float w;
            std::memcpy(&w, &bits, 4);
            const int q = fill[fcell[g]]++;
            frec[q].p = make_float4(r.position[0], r.position[1], r.position[2], w);
            frec[q].rng = make_int2(r.atomStart, r.atomNum);
            frec[q].pad = make_int2(0, 0);
        }
    }
    // mobile groups: buckets + overflow
    std::vector<int> mcnt(R.ncell + 1, 0);
    std::vector<int> mslot((size_t)R.ovf_base + std::max(nmob, 1), -1);
    std::vector<int> mloc(I.totalResNum, -1);
    std::vector<float4> mrec(mslot.size(), make_float4(0.f, 0.f, 0.f, 0.f));
    for (int s = 0; s < T; ++s) {
        const AtomArray &F = H.frag[s];
        for (int j = 0; j < F.totalNum; ++j) {
            int g = F.startRes + j;
            const residue &r = H.res[g];
            int c = cell_index(R.L, R.start, R.inv_h, R.nc, r.position[0], r.position[1], r.position[2]);
            int pos;
            if (mcnt[c] < cap) pos = c * cap + mcnt[c]++;
            else pos = R.ovf_base + mcnt[R.ncell]++;
            mslot[pos] = g;
            mloc[g] = pos;
            const int bits = mob_word(g, s);
            float w;
            std::memcpy(&w, &bits, 4);
            mrec[pos] = make_float4(r.position[0], r.position[1], r.position[2], w);
        }
    }
    std::vector<int> st = make_stencil(R.L, R.nc, R.rc, STENCIL_EPS);
    R.nst = (int)st.size();
    // Per trial, a stencil cell whose box lies farther than rc + slack from
    // the trial centre holds no group that passes the group test: skipped.
    // Along an axis whose stencil offsets are the minimum-image ones (it does
    // not wrap onto itself) the box distance is measured from the centre's
    // position in its cell; other axes count no distance.
    stencil_prune(R.L, R.nc, R.rc, R.cell_h, &R.prune, &R.prune_r2);

    DevBuf &B = H.buf;
    HIPX(put((float **)&R.grid, B, H.grid.data(), H.grid.size()));
    HIPX(put((int **)&R.moves, B, H.moves.data(), H.moves.size()));
    HIPX(put(&R.res_pos, B, res_pos.data(), res_pos.size()));
    HIPX(put(&R.res_rng, B, res_rng.data(), res_rng.size()));
    HIPX(put(&R.atom, B, atom.data(), atom.size()));
    HIPX(put(&R.atype, B, atype.data(), atype.size()));
    HIPX(put((float4 **)&R.pair, B, pair.data(), pair.size()));
    HIPX(put((Species **)&R.sp, B, sp.data(), sp.size()));
    // replicas with equal pair tables share one copy held in LDS (Rep::tab):
    // the index of the first replica (in upload order) with this table
    H.pair = pair;
    R.tab = -1;
    for (size_t q = 0; q < E->reps.size(); ++q) {
        const HostRep &P = *E->reps[q];
        if (&P == &H) {
            R.tab = (int)q;
            break;
        }
        if (P.rep.X == R.X && P.rep.Y == R.Y && P.pair.size() == pair.size() &&
            std::memcmp(P.pair.data(), pair.data(), pair.size() * sizeof(float4)) == 0) {
            R.tab = P.rep.tab;
            break;
        }
    }
    {
        std::vector<int2> cellv(R.ncell + 1);
        for (int c = 0; c <= R.ncell; ++c) cellv[c] = make_int2(fstart[c], mcnt[c]);
        HIPX(put(&R.cell, B, cellv.data(), cellv.size()));
    }
    HIPX(put((FixRec **)&R.fix, B, frec.data(), frec.size()));
    HIPX(put(&R.mob_rec, B, mrec.data(), mrec.size()));
    HIPX(put(&R.mob_slot, B, mslot.data(), mslot.size()));
    H.nmslot = mslot.size();
    HIPX(put(&R.mob_loc, B, mloc.data(), mloc.size()));
    HIPX(put((int **)&R.stencil, B, st.data(), st.size()));
    HIPX(B.alloc(&R.tc, H.maxk));
    HIPX(B.alloc(&R.tslot, H.maxk));
    HIPX(B.alloc(&R.E, 2 * (size_t)H.maxk));
    HIPX(B.alloc(&R.pose, (size_t)H.maxk * MAXF));
    HIPX(B.alloc(&R.adj, (size_t)H.maxk * ADJW));
    if (H.ew_beta > 0.0f) {
        // Ewald mode: PME plan over the box, charge-delta list, and the
        // reciprocal potential of the initial state (fixed atoms and the
        // molecules present)
        int K[3];
        for (int i = 0; i < 3; ++i) K[i] = std::max(4, (int)std::lround(I.cryst[i] / H.ew_spacing));
        int rc = gcmc_pme_create(&H.pme, K, I.cryst, H.ew_beta);
        if (rc) return rc;
        R.ew = 1;
        R.ew_beta = H.ew_beta;
        for (int i = 0; i < 3; ++i) R.pk[i] = K[i];
        R.pme_phi = gcmc_internal::pme_phi(H.pme);
        R.pme_gk = gcmc_internal::pme_kernel(H.pme);
        R.pme_cap = 2 * H.maxk * MAXF;
        HIPX(B.alloc(&R.pme_delta, (size_t)R.pme_cap));
        std::vector<int> zero1(1, 0);
        HIPX(put(&R.pme_ndelta, B, zero1.data(), 1));
        std::vector<float4> q;
        const int nfix = H.frag[0].startRes;
        auto add_res = [&](int g) {
            const residue &rr = H.res[g];
            for (int a = rr.atomStart; a < rr.atomStart + rr.atomNum; ++a) {
                const Atom &A = H.atoms[a];
                q.push_back(make_float4(A.position[0], A.position[1], A.position[2], A.charge));
            }
        };
        for (int g = 0; g < nfix; ++g) add_res(g);
        for (int sp = 0; sp < T; ++sp)
            for (int j = 0; j < H.frag[sp].totalNum; ++j) add_res(H.frag[sp].startRes + j);
        float4 *dq = nullptr;
        HIPX(put(&dq, B, q.data(), q.size()));
        rc = gcmc_pme_clear(H.pme, E->stream);
        if (!rc) rc = gcmc_pme_spread(H.pme, (int)q.size(), (const float *)dq, E->stream);
        if (!rc) rc = gcmc_pme_solve(H.pme, nullptr, E->stream);
        if (rc) return rc;
        HIPX(hipStreamSynchronize(E->stream));
    }
    return 0;
}

int compute_layout(gcmc_engine *E)
{
    // the pair table held in LDS: X x Y entries of the largest table
    int nst = 1, xy = 1;
    for (HostRep *H : E->reps) {
        nst = std::max(nst, H->rep.nst);
        xy = std::max(xy, H->rep.X * H->rep.Y);
    }
    const int pair_bytes = 16 * xy;
    E->staged = E->tune.pair_lds >= 0 && pair_bytes <= 16384;
    auto up = [](int x) { return (x + 15) & ~15; };
    LdsLayout &l = E->lay;
    l.inc_cap = INC_CAP;
    int o = 0;
    l.smob = o; o = up(o + 4 * (nst + 2));
    l.sfst = o; o = up(o + 4 * (nst + 2));
    l.sfix = o; o = up(o + 4 * (nst + 2));
    l.soff = o; o = up(o + 4 * (nst + 2));
    l.incs = o; o = up(o + 4 * (l.inc_cap + 1));
    l.incn = o; o = up(o + 4 * (l.inc_cap + 1));
    l.map_cap = 4096;
    l.cand = o; o = up(o + 2 * l.map_cap);
    l.amap = o; o = up(o + 2 * l.map_cap);
    l.frag = o; o = up(o + 16 * MAXF);
    l.frow = o; o = up(o + 4 * MAXF);
    o = (o + 31) & ~31;
    l.fragd = o; o = up(o + 32 * MAXF);
    if (E->staged) {
        l.pair = o;
        o = up(o + pair_bytes);
    } else {
        l.pair = -1;
    }
    const int eval_end = o;
    // accept and prep phases reuse the evaluation region
    int a = 0;
    l.members = a; a = up(a + 4 * E->maxk);
    l.w = a; a = up(a + 8 * E->maxk);
    l.val = a; a = up(a + 8 * E->maxk);
    l.accl = a; a = up(a + 8 * E->maxk);
    l.rsel = a; a = up(a + 8 * E->maxk);
    l.racc = a; a = up(a + 8 * E->maxk);
    l.ea = a; a = up(a + 8 * E->maxk);
    // adjacency rows of a move staged in LDS for its accept phase, when they fit
    const int adj_bytes = 4 * E->maxk * ((E->maxk + 31) / 32);
    if (a + adj_bytes + 4096 <= 49152) {
        l.adjl = a;
        a = up(a + adj_bytes);
    } else {
        l.adjl = -1;
    }
    l.J = 0;  // prep runs after accept in the same workgroup
    o = std::max(eval_end, a);
    l.spt = o; o = up(o + 16 * MAXT);
    l.misc = o; o = up(o + 4 * MISC_INTS);
    l.dred = o; o = up(o + 8 * MAXWAVE);
    l.total = o;
    if (l.total > 65536)
        return fail(GCMC_E_ARG, "LDS layout exceeds 64 KiB (stencil or confBias too large; raise cell_size)");
    return 0;
}

template <int NT, bool S, bool EW>
void launch_plain_step(gcmc_engine *E, int r)
{
    const size_t lds = E->lay.total;
    hipLaunchKernelGGL((k_eval_plain<NT, S, EW>), dim3(2 * E->reps[r]->maxk), dim3(NT), lds, E->stream, E->d_rep,
                       E->d_state, r, E->lay);
    hipLaunchKernelGGL((k_accept_plain<NT, S, EW>), dim3(1), dim3(NT), lds, E->stream, E->d_rep, E->d_state, r,
                       E->lay);
}

template <int NT, bool S>
void launch_batched(gcmc_engine *E, int nsteps)
{
    const int R = (int)E->reps.size();
    const size_t lds = E->lay.total;
    hipLaunchKernelGGL(k_begin, dim3((R + 255) / 256), dim3(256), 0, E->stream, E->d_state, E->d_rep, 0, R,
                       nsteps, E->d_sched);
    hipLaunchKernelGGL((k_start<NT, S>), dim3(R), dim3(NT), lds, E->stream, E->d_rep, E->d_state, E->d_sched,
                       E->lay, 0, 1);
    if (E->resident <= 0) {
        // workgroups beyond the resident ones would only start once the
        // others have exited: launch no more than fit
        int occ = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_batched<NT, S>, NT, lds) != hipSuccess || occ < 1)
            occ = 1;
        E->resident = occ * std::max(E->ncu, 1);
    }
    const int nwg = std::min(E->nwg, E->resident);
    if (E->timed_kernel) (void)hipEventRecord(E->evk0, E->stream);
    hipLaunchKernelGGL((k_batched<NT, S>), dim3(nwg), dim3(NT), lds, E->stream, E->d_rep, E->d_state, R,
                       E->d_sched, E->lay, E->spin_limit, E->xm, std::min(std::max(E->tune.claim_affinity, 0), 254));
    if (E->timed_kernel) (void)hipEventRecord(E->evk1, E->stream);
}

template <int NT, bool S>
void launch_plain(gcmc_engine *E, int nsteps)
{
    const size_t lds = E->lay.total;
    const int R = (int)E->reps.size();
    for (int r = 0; r < R; ++r) {
        hipLaunchKernelGGL(k_begin, dim3(1), dim3(64), 0, E->stream, E->d_state, E->d_rep, r, 1, nsteps,
                           E->d_sched);
        hipLaunchKernelGGL((k_start<NT, S>), dim3(1), dim3(NT), lds, E->stream, E->d_rep, E->d_state,
                           E->d_sched, E->lay, r, 0);
        const int rem = E->reps[r]->rep.nmoves - E->reps[r]->host_step;
        const int n = std::min(nsteps, std::max(rem, 0));
        HostRep &H = *E->reps[r];
        for (int s = 0; s < n; ++s) {
            if (H.pme) launch_plain_step<NT, S, true>(E, r);
            else launch_plain_step<NT, S, false>(E, r);
            if (H.pme) {
                // charges moved by the step enter the PME grid; new potential
                gcmc_internal::pme_spread_counted(H.pme, H.rep.pme_cap, H.rep.pme_ndelta, H.rep.pme_delta,
                                                  H.rep.pme_ndelta, E->stream);
                gcmc_pme_solve(H.pme, nullptr, E->stream);
            }
        }
    }
}

// Systems with Ewald electrostatics in the batched and graph forms: move by
// move, every system's trial evaluations in one launch, its accept + prep,
// then the PME update (the accepted charge changes spread into the charge
// grid, FFT solve for the new potential) -- the reciprocal potential has to
// be current before the next move's evaluations read it.  The same kernels
// in the same order as the plain form, so the chains are identical.
template <int NT, bool S>
void launch_ewald_begin(gcmc_engine *E, int nsteps)
{
    const int R = (int)E->reps.size();
    hipLaunchKernelGGL(k_begin, dim3((R + 255) / 256), dim3(256), 0, E->stream, E->d_state, E->d_rep, 0, R, nsteps,
                       E->d_sched);
    hipLaunchKernelGGL((k_start<NT, S>), dim3(R), dim3(NT), E->lay.total, E->stream, E->d_rep, E->d_state,
                       E->d_sched, E->lay, 0, 0);
}

template <int NT, bool S>
void launch_ewald_moves(gcmc_engine *E, int nmoves)
{
    for (int m = 0; m < nmoves; ++m)
        for (int r = 0; r < (int)E->reps.size(); ++r) {
            HostRep &H = *E->reps[r];
            launch_plain_step<NT, S, true>(E, r);
            gcmc_internal::pme_spread_counted(H.pme, H.rep.pme_cap, H.rep.pme_ndelta, H.rep.pme_delta,
                                              H.rep.pme_ndelta, E->stream);
            gcmc_pme_solve(H.pme, nullptr, E->stream);
        }
}

// dispatch on (workgroup size, pair table in LDS)
template <template <int, bool> class F, class... A>
void dispatch(gcmc_engine *E, A... args)
{
    if (E->nt == 1024) E->staged ? F<1024, true>::run(E, args...) : F<1024, false>::run(E, args...);
    else if (E->nt == 512) E->staged ? F<512, true>::run(E, args...) : F<512, false>::run(E, args...);
    else E->staged ? F<256, true>::run(E, args...) : F<256, false>::run(E, args...);
}
template <int NT, bool S>
struct RunPlain {
    static void run(gcmc_engine *E, int n) { launch_plain<NT, S>(E, n); }
};
template <int NT, bool S>
struct RunBatched {
    static void run(gcmc_engine *E, int n) { launch_batched<NT, S>(E, n); }
};
template <int NT, bool S>
struct RunEwaldBegin {
    static void run(gcmc_engine *E, int n) { launch_ewald_begin<NT, S>(E, n); }
};
template <int NT, bool S>
struct RunEwaldMoves {
    static void run(gcmc_engine *E, int n) { launch_ewald_moves<NT, S>(E, n); }
};
template <int NT, bool S>
struct RunPoses {
    static void run(gcmc_engine *E, int npose, int r, int species, const float *dc, const float *dp, const int *dx,
                    double *de, int *dn, uint32_t *dh)
    {
        if (E->reps[r]->pme)
            hipLaunchKernelGGL((k_eval_poses<NT, S, true>), dim3(npose), dim3(NT), E->lay.total, E->stream, E->d_rep,
                               r, species, dc, dp, dx, de, dn, dh, E->lay);
        else
            hipLaunchKernelGGL((k_eval_poses<NT, S, false>), dim3(npose), dim3(NT), E->lay.total, E->stream,
                               E->d_rep, r, species, dc, dp, dx, de, dn, dh, E->lay);
    }
};

// The class of each replica to the device (cls_start, cls_reps).
int upload_classes(gcmc_engine *E, const std::vector<int> &cls_of)
{
    XcdMap &m = E->xm;
    const int R = (int)E->reps.size();
    std::vector<int> cls(m.nclass + 1 + R, 0);
    int o = 0;
    for (int c = 0; c < m.nclass; ++c) {
        cls[c] = o;
        for (int r = 0; r < R; ++r)
            if (cls_of[r] == c) cls[m.nclass + 1 + o++] = r;
    }
    cls[m.nclass] = o;
    HIPX(hipMemcpy(E->d_cls, cls.data(), cls.size() * sizeof(int), hipMemcpyHostToDevice));
    E->cls_host = cls;
    E->cls_of = cls_of;
    return 0;
}

// Replicas to XCD classes (each replica is served by the workgroups of one
// XCD).  Round-robin when the tuning asks for it; otherwise the heaviest
// replica first to the class with the least work so far (longest processing
// time first), work w[r] being the cost model's estimate of the schedule
// before the first run and the measured work of the last run after it: the
// run ends when its slowest XCD does, and the XCDs' loads then differ least.
// Changes which workgroups evaluate a replica's items, never the results.
std::vector<int> lpt_classes(const gcmc_engine *E, const std::vector<double> &w)
{
    const int nclass = E->xm.nclass, R = (int)E->reps.size();
    std::vector<int> cls_of(R, 0);
    if (E->tune.xcd_round_robin) {
        for (int r = 0; r < R; ++r) cls_of[r] = r % nclass;
        return cls_of;
    }
    std::vector<int> order(R);
    for (int r = 0; r < R; ++r) order[r] = r;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return w[a] > w[b]; });
    std::vector<double> load(nclass, 0.0);
    for (int r : order) {
        int c = 0;
        for (int q = 1; q < nclass; ++q)
            if (load[q] < load[c]) c = q;
        load[c] += w[r];
        cls_of[r] = c;
    }
    return cls_of;
}

int assign_classes(gcmc_engine *E, const std::vector<double> &w) { return upload_classes(E, lpt_classes(E, w)); }

// One exchange between class S (finished later) and class F under the
// assignment best (class of each replica), applied to next: a replica a of S
// for a replica b of F, the pair whose evaluation times evt (each class's
// finish time fin taken as proportional to its summed evaluation time) leave
// the later of the two classes' predicted finishes earliest; none when no
// pair outside tried would end S sooner.
bool xcd_exchange_one(const std::vector<int> &best, const std::vector<double> &fin, const std::vector<double> &evt,
                      const std::vector<std::pair<int, int>> &tried, int S, int F, std::vector<int> &next,
                      std::pair<int, int> &pair)
{
    const int R = (int)best.size();
    double eS = 0.0, eF = 0.0;
    for (int r = 0; r < R; ++r) {
        if (best[r] == S) eS += evt[r];
        if (best[r] == F) eF += evt[r];
    }
    if (eS <= 0.0 || eF <= 0.0) return false;
    const double kS = fin[S] / eS, kF = fin[F] / eF;
    double top = fin[S];
    int ba = -1, bb = -1;
    for (int a = 0; a < R; ++a) {
        if (best[a] != S) continue;
        for (int b = 0; b < R; ++b) {
            if (best[b] != F) continue;
            const double d = evt[a] - evt[b];
            if (d <= 0.0) continue;
            if (std::find(tried.begin(), tried.end(), std::make_pair(a, b)) != tried.end()) continue;
            const double pred = std::max(fin[S] - kS * d, fin[F] + kF * d);
            if (pred < top) {
                top = pred;
                ba = a;
                bb = b;
            }
        }
    }
    if (ba < 0) return false;
    next[ba] = F;
    next[bb] = S;
    pair = std::make_pair(ba, bb);
    return true;
}

// The exchanges of one try from the assignment best: the classes ordered by
// finish time, the i-th latest paired with the i-th earliest, one exchange
// per pair (xcd_exchange_one).  A try changes several classes at once; the
// run measures it as a whole and the faster assignment is kept.
bool xcd_exchanges(int nclass, const std::vector<int> &best, const std::vector<double> &fin,
                   const std::vector<double> &evt, const std::vector<std::pair<int, int>> &tried,
                   std::vector<int> &next, std::vector<std::pair<int, int>> &pairs)
{
    std::vector<int> order(nclass);
    for (int c = 0; c < nclass; ++c) order[c] = c;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return fin[a] > fin[b]; });
    next = best;
    pairs.clear();
    for (int i = 0; i < nclass / 2; ++i) {
        const int S = order[i], F = order[nclass - 1 - i];
        if (!(fin[S] > fin[F])) break;
        std::pair<int, int> pr;
        if (xcd_exchange_one(best, fin, evt, tried, S, F, next, pr)) pairs.push_back(pr);
    }
    return !pairs.empty();
}

bool propose_exchange(const gcmc_engine *E, std::vector<int> &next, std::vector<std::pair<int, int>> &pairs)
{
    const gcmc_engine::XcdSearch &X = E->xs;
    return xcd_exchanges(E->xm.nclass, X.best, X.fin, X.evt, X.tried, next, pairs);
}

// Which XCC ids the device's workgroups run on, and the assignment of
// replicas to them (assign_classes).
int build_xcd_map(gcmc_engine *E, int ncu)
{
    const int nprobe = std::max(64, 8 * ncu);
    int *d = nullptr;
    HIPX(hipMalloc(&d, nprobe * sizeof(int)));
    hipLaunchKernelGGL(k_xcc_probe, dim3(nprobe), dim3(64), 0, E->stream, d);
    std::vector<int> ids(nprobe);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(ids.data(), d, nprobe * sizeof(int), hipMemcpyDeviceToHost, E->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(E->stream);
    hipFree(d);
    if (e != hipSuccess) return fail(GCMC_E_HIP, std::string("XCC probe: ") + hipGetErrorString(e));
    XcdMap &m = E->xm;
    for (int i = 0; i < 16; ++i) m.xcc_class[i] = -1;
    m.nclass = 0;
    for (int id : ids)
        if (id >= 0 && id < 16 && m.xcc_class[id] < 0) m.xcc_class[id] = m.nclass++;
    if (m.nclass == 0) return fail(GCMC_E_DEVICE, "XCC probe returned no ids");
    const int R = (int)E->reps.size();
    HIPX(hipMalloc(&E->d_cls, (m.nclass + 1 + R) * sizeof(int)));
    HIPX(hipMalloc(&E->d_cu_inv, CU_SLOTS * sizeof(unsigned long long)));
    HIPX(hipMemset(E->d_cu_inv, 0, CU_SLOTS * sizeof(unsigned long long)));
    m.cu_inv = E->d_cu_inv;
    m.cls_start = E->d_cls;
    m.cls_reps = E->d_cls + m.nclass + 1;
    std::vector<double> w(R);
    for (int r = 0; r < R; ++r) w[r] = E->reps[r]->est_cost;
    return assign_classes(E, w);
}

int read_states(gcmc_engine *E, std::vector<State> &st)
{
    st.resize(E->reps.size());
    HIPX(hipMemcpy(st.data(), E->d_state, st.size() * sizeof(State), hipMemcpyDeviceToHost));
    return 0;
}

int sync_rep_struct(gcmc_engine *E, int r)
{
    if (!E->uploaded) return 0;
    HIPX(hipMemcpy(E->d_rep + r, &E->reps[r]->rep, sizeof(Rep), hipMemcpyHostToDevice));
    if (E->gexec) {
        hipGraphExecDestroy(E->gexec);
        hipGraphDestroy(E->graph);
        E->gexec = nullptr;
        E->graph = nullptr;
        E->graph_nsteps = -1;
    }
    if (E->ew_gexec) {
        hipGraphExecDestroy(E->ew_gexec);
        hipGraphDestroy(E->ew_graph);
        E->ew_gexec = nullptr;
        E->ew_graph = nullptr;
        E->ew_graph_moves = 0;
    }
    return 0;
}

}  // namespace

namespace gcmc_internal {
// shared by the library's other translation units (one error channel)
int set_error(int code, const std::string &msg) { return fail(code, msg); }
}  // namespace gcmc_internal

extern "C" {

const char *gcmc_last_error(void) { return g_err.c_str(); }

int gcmc_engine_create(gcmc_engine **out, const gcmc_tuning *tuning)
{
    if (!out) return fail(GCMC_E_ARG, "null engine pointer");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(GCMC_E_DEVICE, "no HIP device available");
    gcmc_engine *E = new gcmc_engine();
    if (tuning) E->tune = *tuning;
    hipError_t e = hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&E->ev0);
    if (e == hipSuccess) e = hipEventCreate(&E->ev1);
    if (e == hipSuccess) e = hipEventCreate(&E->evk0);
    if (e == hipSuccess) e = hipEventCreate(&E->evk1);
    if (e != hipSuccess) {
        delete E;
        return fail(GCMC_E_HIP, std::string("stream/event creation: ") + hipGetErrorString(e));
    }
    *out = E;
    return 0;
}

int gcmc_engine_destroy(gcmc_engine *E)
{
    if (!E) return 0;
    if (E->stream) hipStreamSynchronize(E->stream);
    if (E->gexec) hipGraphExecDestroy(E->gexec);
    if (E->graph) hipGraphDestroy(E->graph);
    if (E->ew_gexec) hipGraphExecDestroy(E->ew_gexec);
    if (E->ew_graph) hipGraphDestroy(E->ew_graph);
    for (HostRep *H : E->reps) {
        if (H->pme) gcmc_pme_destroy(H->pme);
        H->buf.release();
        H->trbuf.release();
        H->snap.release();
        delete H;
    }
    if (E->d_rep) hipFree(E->d_rep);
    if (E->d_state) hipFree(E->d_state);
    if (E->d_sched) hipFree(E->d_sched);
    if (E->d_cls) hipFree(E->d_cls);
    if (E->d_cu_inv) hipFree(E->d_cu_inv);
    if (E->d_state_snap) hipFree(E->d_state_snap);
    if (E->d_adj_work) hipFree(E->d_adj_work);
    if (E->ev0) hipEventDestroy(E->ev0);
    if (E->ev1) hipEventDestroy(E->ev1);
    if (E->evk0) hipEventDestroy(E->evk0);
    if (E->evk1) hipEventDestroy(E->evk1);
    if (E->stream) hipStreamDestroy(E->stream);
    delete E;
    return 0;
}

int gcmc_engine_add_system(gcmc_engine *E, const InfoStruct *info, const AtomArray *frag,
                           const residue *res, const Atom *atoms, const float *grid, const float *ff,
                           const int *moves, int nmoves)
{
    if (!E) return fail(GCMC_E_ARG, "null engine");
    if (E->uploaded) return fail(GCMC_E_STATE, "systems must be added before upload");
    int rc = check_system(info, frag, nmoves);
    if (rc) return rc;
    if (!res || !atoms || !grid || !ff || (nmoves > 0 && !moves))
        return fail(GCMC_E_ARG, "null array pointer");
    HostRep *H = new HostRep();
    H->info = *info;
    H->frag.assign(frag, frag + info->fragTypeNum);
    H->res.assign(res, res + info->totalResNum);
    H->atoms.assign(atoms, atoms + info->totalAtomNum);
    H->grid.assign(grid, grid + 3 * (size_t)info->totalGridNum);
    H->ff.assign(ff, ff + 2 * (size_t)info->ffXNum * info->ffYNum);
    H->moves.assign(moves, moves + nmoves);
    for (int m : H->moves) {
        if (m < 0 || m / 4 >= info->fragTypeNum) {
            delete H;
            return fail(GCMC_E_ARG, "moveArray entry names a species outside fragTypeNum");
        }
    }
    for (const AtomArray &F : H->frag) H->maxk = std::max(H->maxk, F.confBias);
    E->maxk = std::max(E->maxk, H->maxk);
    E->reps.push_back(H);
    return (int)E->reps.size() - 1;
}

int gcmc_engine_upload(gcmc_engine *E)
{
    if (!E) return fail(GCMC_E_ARG, "null engine");
    if (E->uploaded) return fail(GCMC_E_STATE, "already uploaded");
    if (E->reps.empty()) return fail(GCMC_E_STATE, "no systems added");
    for (HostRep *H : E->reps) {
        int rc = upload_rep(E, *H);
        if (rc) return rc;
    }
    // Scheduling priority of the batched form: the estimated critical-path
    // cost of a replica's moves still to run.  A move costs a fixed part
    // (publish, accept, prep, one evaluation's latency) plus one unit per
    // evaluation beyond the first that the replica's XCD runs (k trials, 2k
    // for translate / rotate with k <= the initial count of the species).
    // Prefix sums per replica, scaled so the largest total fits 24 bits.
    {
        std::vector<std::vector<unsigned long long>> cum(E->reps.size());
        unsigned long long top = 1;
        for (size_t r = 0; r < E->reps.size(); ++r) {
            const HostRep &H = *E->reps[r];
            std::vector<unsigned long long> &c = cum[r];
            c.assign(H.moves.size() + 1, 0ull);
            for (size_t s = 0; s < H.moves.size(); ++s) {
                const int kind = H.moves[s] & 3;
                const AtomArray &F = H.frag[H.moves[s] >> 2];
                long long n = kind == K_DEL ? 1 : std::min(F.confBias, H.maxk);
                if (kind == K_TRN || kind == K_ROT) n = 2 * std::min<long long>(n, std::max(F.totalNum, 1));
                c[s + 1] = c[s] + 55ull + (unsigned long long)std::max<long long>(n, 1);
            }
            top = std::max(top, c.back());
        }
        int shift = 0;
        while ((top >> shift) >= (1ull << 24)) ++shift;
        for (size_t r = 0; r < E->reps.size(); ++r) {
            HostRep &H = *E->reps[r];
            H.est_cost = (double)cum[r].back();
            std::vector<uint32_t> w(cum[r].size());
            for (size_t i = 0; i < w.size(); ++i) w[i] = (uint32_t)(cum[r][i] >> shift);
            HIPX(put((uint32_t **)&H.rep.wcum, H.buf, w.data(), w.size()));
        }
    }
    E->nt = (E->tune.wg_size == 512 || E->tune.wg_size == 1024) ? E->tune.wg_size : 256;
    int rc = compute_layout(E);
    if (rc) return rc;
    const int R = (int)E->reps.size();
    HIPX(hipMalloc(&E->d_rep, R * sizeof(Rep)));
    HIPX(hipMalloc(&E->d_state, R * sizeof(State)));
    HIPX(hipMalloc(&E->d_sched, sizeof(Sched)));
    std::vector<Rep> reps(R);
    std::vector<State> st(R);
    for (int r = 0; r < R; ++r) {
        reps[r] = E->reps[r]->rep;
        State &S = st[r];
        std::memset(&S, 0, sizeof(S));
        for (int s = 0; s < E->reps[r]->info.fragTypeNum; ++s) S.totalNum[s] = E->reps[r]->frag[s].totalNum;
    }
    HIPX(hipMemcpy(E->d_rep, reps.data(), R * sizeof(Rep), hipMemcpyHostToDevice));
    HIPX(hipMemcpy(E->d_state, st.data(), R * sizeof(State), hipMemcpyHostToDevice));
    Sched z{};
    HIPX(hipMemcpy(E->d_sched, &z, sizeof(z), hipMemcpyHostToDevice));
    int dev = 0, ncu = 0;
    HIPX(hipGetDevice(&dev));
    HIPX(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    E->ncu = ncu;
    E->resident = 0;
    E->nwg = E->tune.persistent_wgs > 0 ? E->tune.persistent_wgs : (2048 / E->nt) * std::max(ncu, 1);
    rc = build_xcd_map(E, ncu);
    if (rc) return rc;
    E->uploaded = true;
    return 0;
}

int gcmc_engine_run(gcmc_engine *E, int nsteps, int mode)
{
    if (!E || !E->uploaded) return fail(GCMC_E_STATE, "engine not uploaded");
    if (nsteps < 0) return fail(GCMC_E_ARG, "negative step count");
    if (nsteps == 0) return 0;
    int n_ew = 0;
    for (HostRep *H : E->reps) n_ew += H->pme ? 1 : 0;
    if (mode != GCMC_MODE_PLAIN && n_ew > 0 && n_ew < (int)E->reps.size())
        return fail(GCMC_E_ARG, "the batched and graph forms run either all systems with Ewald electrostatics "
                                "or none; run mixed engines in plain mode");
    if (mode != GCMC_MODE_PLAIN && mode != GCMC_MODE_BATCHED && mode != GCMC_MODE_GRAPH)
        return fail(GCMC_E_ARG, "unknown run mode");
    const bool balance = mode != GCMC_MODE_PLAIN && n_ew == 0 && !E->tune.xcd_round_robin;
    const int R = (int)E->reps.size();
    std::vector<State> st0;
    if (balance) {
        gcmc_engine::XcdSearch &X = E->xs;
        std::vector<int> range(2 * R);
        for (int r = 0; r < R; ++r) {
            const HostRep &H = *E->reps[r];
            range[2 * r] = H.host_step;
            range[2 * r + 1] = (int)std::min<long long>((long long)H.host_step + nsteps, H.rep.nmoves);
        }
        if (X.range == range && !X.next.empty()) {
            // the same steps again: the assignment chosen after the last run
            int rc = upload_classes(E, X.next);
            if (rc) return rc;
        } else {
            // other steps: balance the XCDs by the work each replica did in
            // the last run (else keep the estimate's assignment)
            X = gcmc_engine::XcdSearch{};
            X.range = range;
            std::vector<double> w(R);
            bool measured = false;
            for (int r = 0; r < R; ++r) {
                w[r] = E->reps[r]->last_work;
                measured |= w[r] > 0.0;
            }
            if (measured) {
                int rc = assign_classes(E, w);
                if (rc) return rc;
            }
        }
        int rc = read_states(E, st0);
        if (rc) return rc;
    }
    std::vector<int> step0(E->reps.size());
    for (size_t r = 0; r < step0.size(); ++r) step0[r] = E->reps[r]->host_step;
    HIPX(hipEventRecord(E->ev0, E->stream));
    if (mode != GCMC_MODE_PLAIN && n_ew > 0) {
        // moves this call runs: the most any system has left
        int n = 0;
        for (HostRep *H : E->reps) n = std::max(n, std::min(nsteps, std::max(H->rep.nmoves - H->host_step, 0)));
        dispatch<RunEwaldBegin>(E, nsteps);
        int done = 0;
        if (mode == GCMC_MODE_GRAPH && n >= EW_GRAPH_MOVES) {
            if (!E->ew_gexec) {
                HIPX(hipStreamBeginCapture(E->stream, hipStreamCaptureModeThreadLocal));
                dispatch<RunEwaldMoves>(E, EW_GRAPH_MOVES);
                HIPX(hipStreamEndCapture(E->stream, &E->ew_graph));
                HIPX(hipGraphInstantiate(&E->ew_gexec, E->ew_graph, nullptr, nullptr, 0));
                E->ew_graph_moves = EW_GRAPH_MOVES;
            }
            for (; done + E->ew_graph_moves <= n; done += E->ew_graph_moves)
                HIPX(hipGraphLaunch(E->ew_gexec, E->stream));
        }
        if (done < n) dispatch<RunEwaldMoves>(E, n - done);
    } else if (mode == GCMC_MODE_PLAIN) {
        dispatch<RunPlain>(E, nsteps);
    } else if (mode == GCMC_MODE_BATCHED) {
        E->timed_kernel = true;
        dispatch<RunBatched>(E, nsteps);
        E->timed_kernel = false;
    } else if (mode == GCMC_MODE_GRAPH) {
        if (!E->gexec || E->graph_nsteps != nsteps) {
            if (E->gexec) {
                hipGraphExecDestroy(E->gexec);
                hipGraphDestroy(E->graph);
                E->gexec = nullptr;
            }
            HIPX(hipStreamBeginCapture(E->stream, hipStreamCaptureModeThreadLocal));
            dispatch<RunBatched>(E, nsteps);
            HIPX(hipStreamEndCapture(E->stream, &E->graph));
            HIPX(hipGraphInstantiate(&E->gexec, E->graph, nullptr, nullptr, 0));
            E->graph_nsteps = nsteps;
        }
        HIPX(hipGraphLaunch(E->gexec, E->stream));
    }
    HIPX(hipGetLastError());
    HIPX(hipEventRecord(E->ev1, E->stream));
    HIPX(hipEventSynchronize(E->ev1));
    float ms = 0.f;
    HIPX(hipEventElapsedTime(&ms, E->ev0, E->ev1));
    E->last_ms = ms;
    E->kernel_ms = 0.0;
    if (mode == GCMC_MODE_BATCHED && n_ew == 0) {
        HIPX(hipEventElapsedTime(&ms, E->evk0, E->evk1));
        E->kernel_ms = ms;
    }
    std::vector<State> st;
    int rc = read_states(E, st);
    if (rc) return rc;
    Sched sch{};
    HIPX(hipMemcpy(&sch, E->d_sched, sizeof(sch), hipMemcpyDeviceToHost));
    if (sch.abort) return fail(GCMC_E_DEVICE, "persistent kernel stopped at its spin limit");
    for (size_t r = 0; r < st.size(); ++r) {
        HostRep &H = *E->reps[r];
        H.host_step = st[r].step;
        // work of this run: evaluations, plus one evaluation's worth per move
        // for its accept and prep (counts, not times: measured times would
        // follow the contention of the assignment they were measured under)
        H.last_work = (double)(st[r].n_eval - H.n_eval_now) + (double)(st[r].step - step0[r]);
        H.n_eval_now = st[r].n_eval;
        if (st[r].step != st[r].end)
            return fail(GCMC_E_DEVICE, "replica " + std::to_string(r) + " stopped at step " +
                                           std::to_string(st[r].step) + " before its end " +
                                           std::to_string(st[r].end));
    }
    if (balance) {
        // this run's time, per-replica evaluation time (workgroup-seconds) and
        // per-class finish (the latest end of a replica's chain, which runs
        // from its first publication through its evaluations, accepts and
        // preps: 100 MHz ticks)
        gcmc_engine::XcdSearch &X = E->xs;
        const int nclass = E->xm.nclass;
        const double t = mode == GCMC_MODE_BATCHED ? E->kernel_ms : E->last_ms;
        std::vector<double> evt(R), fin(nclass, 0.0);
        for (int r = 0; r < R; ++r) {
            evt[r] = (double)(st[r].t_eval - st0[r].t_eval) * 1e-8;
            const unsigned long long el = (st[r].t_step - st0[r].t_step) + (st[r].t_accept - st0[r].t_accept) +
                                          (st[r].t_prep - st0[r].t_prep);
            const int c = E->cls_of[r];
            fin[c] = std::max(fin[c], (double)el * 1e-5);
        }
        X.runs += 1;
        if (X.best.empty() || t < X.best_ms) {
            X.best = E->cls_of;
            X.best_ms = t;
            X.fin = fin;
            X.evt = evt;
            X.tried.clear();
        }
        const int tries = E->tune.xcd_search == 0 ? 3 : std::max(E->tune.xcd_search, 0);
        std::vector<std::pair<int, int>> pr;
        if (X.runs == 1) {
            // first run of these steps (on the estimate's or an earlier range's
            // assignment): next, the balance by this run's measured evaluation
            // times (a replica's evaluation time depends on the replicas it
            // shares an XCD with, so the balance is a candidate like the
            // exchanges: the fastest assignment is kept)
            X.next = lpt_classes(E, evt);
        } else if (X.runs - 1 <= tries && propose_exchange(E, X.next, pr)) {
            X.tried.insert(X.tried.end(), pr.begin(), pr.end());
        } else {
            X.next = X.best;
        }
    }
    return 0;
}

int gcmc_engine_synchronize(gcmc_engine *E)
{
    if (!E) return fail(GCMC_E_ARG, "null engine");
    HIPX(hipStreamSynchronize(E->stream));
    return 0;
}

int gcmc_engine_step(gcmc_engine *E, int r)
{
    if (!E || r < 0 || r >= (int)E->reps.size()) return fail(GCMC_E_ARG, "bad replica");
    return E->reps[r]->host_step;
}

int gcmc_engine_last_run_ms(gcmc_engine *E, double *ms)
{
    if (!E || !ms) return fail(GCMC_E_ARG, "null argument");
    *ms = E->last_ms;
    return 0;
}

int gcmc_engine_launch_info(gcmc_engine *E, int *out4)
{
    if (!E || !out4) return fail(GCMC_E_ARG, "null argument");
    if (!E->uploaded) return fail(GCMC_E_STATE, "upload() first");
    out4[0] = E->nwg;
    out4[1] = E->resident;
    out4[2] = E->nt;
    out4[3] = E->lay.total;
    return GCMC_OK;
}

int gcmc_engine_last_kernel_ms(gcmc_engine *E, double *ms)
{
    if (!E || !ms) return fail(GCMC_E_ARG, "null argument");
    *ms = E->kernel_ms;
    return 0;
}

int gcmc_engine_sub_ticks(gcmc_engine *E, int r, unsigned long long *out4)
{
    if (!E || !E->uploaded || !out4) return fail(GCMC_E_STATE, "engine not uploaded or null output");
    if (r < 0 || r >= (int)E->reps.size()) return fail(GCMC_E_ARG, "bad replica");
    HIPX(hipStreamSynchronize(E->stream));
    State S;
    HIPX(hipMemcpy(&S, E->d_state + r, sizeof(S), hipMemcpyDeviceToHost));
    for (int i = 0; i < 4; ++i) out4[i] = S.t_sub[i];
    return 0;
}

int gcmc_engine_phase_ticks(gcmc_engine *E, int r, unsigned long long *out4)
{
    if (!E || !E->uploaded || !out4) return fail(GCMC_E_STATE, "engine not uploaded or null output");
    if (r < 0 || r >= (int)E->reps.size()) return fail(GCMC_E_ARG, "bad replica");
    HIPX(hipStreamSynchronize(E->stream));
    State S;
    HIPX(hipMemcpy(&S, E->d_state + r, sizeof(S), hipMemcpyDeviceToHost));
    out4[0] = S.t_eval;
    out4[1] = S.t_accept;
    out4[2] = S.t_prep;
    out4[3] = S.t_step;
    return 0;
}

int gcmc_engine_counters(gcmc_engine *E, int r, unsigned long long *out5)
{
    if (!E || !E->uploaded || !out5) return fail(GCMC_E_STATE, "engine not uploaded or null output");
    if (r < 0 || r >= (int)E->reps.size()) return fail(GCMC_E_ARG, "bad replica");
    HIPX(hipStreamSynchronize(E->stream));
    State S;
    HIPX(hipMemcpy(&S, E->d_state + r, sizeof(S), hipMemcpyDeviceToHost));
    out5[0] = S.n_eval;
    out5[1] = S.n_pair;
    out5[2] = S.n_atom;
    out5[3] = S.n_group;
    out5[4] = S.n_fatom;
    return 0;
}

int gcmc_engine_download(gcmc_engine *E, int r, AtomArray *frag, residue *res, Atom *atoms)
{
    if (!E || !E->uploaded) return fail(GCMC_E_STATE, "engine not uploaded");
    if (r < 0 || r >= (int)E->reps.size()) return fail(GCMC_E_ARG, "bad replica");
    HostRep &H = *E->reps[r];
    const InfoStruct &I = H.info;
    HIPX(hipStreamSynchronize(E->stream));
    State S;
    HIPX(hipMemcpy(&S, E->d_state + r, sizeof(S), hipMemcpyDeviceToHost));
    std::vector<float4> res_pos(I.totalResNum), atom(I.totalAtomNum);
    std::vector<int> atype(I.totalAtomNum);
    HIPX(hipMemcpy(res_pos.data(), H.rep.res_pos, res_pos.size() * sizeof(float4), hipMemcpyDeviceToHost));
    HIPX(hipMemcpy(atom.data(), H.rep.atom, atom.size() * sizeof(float4), hipMemcpyDeviceToHost));
    HIPX(hipMemcpy(atype.data(), H.rep.atype, atype.size() * sizeof(int), hipMemcpyDeviceToHost));
    if (frag)
        for (int s = 0; s < I.fragTypeNum; ++s) frag[s].totalNum = S.totalNum[s];
    if (res)
        for (int g = 0; g < I.totalResNum; ++g) {
            res[g].position[0] = res_pos[g].x;
            res[g].position[1] = res_pos[g].y;
            res[g].position[2] = res_pos[g].z;
        }
    if (atoms)
        for (int a = 0; a < I.totalAtomNum; ++a) {
            atoms[a].position[0] = atom[a].x;
            atoms[a].position[1] = atom[a].y;
            atoms[a].position[2] = atom[a].z;
            atoms[a].charge = atom[a].w;
            atoms[a].type = atype[a];
        }
    return 0;
}

int gcmc_engine_stats(gcmc_engine *E, int r, double *energy, long long *attempts, long long *accepts)
{
    if (!E || !E->uploaded) return fail(GCMC_E_STATE, "engine not uploaded");
    if (r < 0 || r >= (int)E->reps.size()) return fail(GCMC_E_ARG, "bad replica");
    HIPX(hipStreamSynchronize(E->stream));
    State S;
    HIPX(hipMemcpy(&S, E->d_state + r, sizeof(S), hipMemcpyDeviceToHost));
    const int T = E->reps[r]->info.fragTypeNum;
    if (energy) *energy = S.energy;
    if (attempts) std::memcpy(attempts, S.att, sizeof(long long) * 4 * T);
    if (accepts) std::memcpy(accepts, S.acc, sizeof(long long) * 4 * T);
    return 0;
}

int gcmc_engine_snapshot(gcmc_engine *E)
{
    if (!E || !E->uploaded) return fail(GCMC_E_STATE, "engine not uploaded");
    for (HostRep *H : E->reps)
        if (H->pme) return fail(GCMC_E_STATE, "snapshots of systems with Ewald electrostatics are not supported");
    const int R = (int)E->reps.size();
    if (!E->d_state_snap) HIPX(hipMalloc(&E->d_state_snap, R * sizeof(State)));
    for (HostRep *Hp : E->reps) {
        HostRep &H = *Hp;
        const Rep &P = H.rep;
        if (!H.s_res_pos) {
            HIPX(H.snap.alloc(&H.s_res_pos, (size_t)P.nres));
            HIPX(H.snap.alloc(&H.s_atom, (size_t)P.natom));
            HIPX(H.snap.alloc(&H.s_atype, (size_t)P.natom));
            HIPX(H.snap.alloc(&H.s_cell, (size_t)P.ncell + 1));
            HIPX(H.snap.alloc(&H.s_mob_slot, H.nmslot));
            HIPX(H.snap.alloc(&H.s_mob_rec, H.nmslot));
            HIPX(H.snap.alloc(&H.s_mob_loc, (size_t)P.nres));
        }
        HIPX(d2d(H.s_res_pos, P.res_pos, (size_t)P.nres, E->stream));
        HIPX(d2d(H.s_atom, P.atom, (size_t)P.natom, E->stream));
        HIPX(d2d(H.s_atype, P.atype, (size_t)P.natom, E->stream));
        HIPX(d2d(H.s_cell, P.cell, (size_t)P.ncell + 1, E->stream));
        HIPX(d2d(H.s_mob_slot, P.mob_slot, H.nmslot, E->stream));
        HIPX(d2d(H.s_mob_rec, P.mob_rec, H.nmslot, E->stream));
        HIPX(d2d(H.s_mob_loc, P.mob_loc, (size_t)P.nres, E->stream));
        H.snap_step = H.host_step;
    }
    HIPX(d2d(E->d_state_snap, E->d_state, (size_t)R, E->stream));
    HIPX(hipStreamSynchronize(E->stream));
    std::vector<State> st;
    int rc = read_states(E, st);
    if (rc) return rc;
    for (int r = 0; r < R; ++r) {
        E->reps[r]->n_eval_now = E->reps[r]->snap_n_eval = st[r].n_eval;
    }
    return 0;
}

int gcmc_engine_restore(gcmc_engine *E)
{
    if (!E || !E->uploaded) return fail(GCMC_E_STATE, "engine not uploaded");
    if (!E->d_state_snap) return fail(GCMC_E_STATE, "no snapshot taken");
    for (HostRep *Hp : E->reps) {
        HostRep &H = *Hp;
        const Rep &P = H.rep;
        HIPX(d2d(P.res_pos, H.s_res_pos, (size_t)P.nres, E->stream));
        HIPX(d2d(P.atom, H.s_atom, (size_t)P.natom, E->stream));
        HIPX(d2d(P.atype, H.s_atype, (size_t)P.natom, E->stream));
        HIPX(d2d(P.cell, H.s_cell, (size_t)P.ncell + 1, E->stream));
        HIPX(d2d(P.mob_slot, H.s_mob_slot, H.nmslot, E->stream));
        HIPX(d2d(P.mob_rec, H.s_mob_rec, H.nmslot, E->stream));
        HIPX(d2d(P.mob_loc, H.s_mob_loc, (size_t)P.nres, E->stream));
        H.host_step = H.snap_step;
        H.n_eval_now = H.snap_n_eval;
    }
    HIPX(d2d(E->d_state, E->d_state_snap, E->reps.size(), E->stream));
    return 0;
}

int gcmc_engine_set_trace(gcmc_engine *E, int r, int max_steps, int max_k)
{
    if (!E || !E->uploaded) return fail(GCMC_E_STATE, "engine not uploaded");
    if (r < 0 || r >= (int)E->reps.size()) return fail(GCMC_E_ARG, "bad replica");
    if (max_steps < 0 || max_k < 0) return fail(GCMC_E_ARG, "negative trace size");
    HostRep &H = *E->reps[r];
    HIPX(hipStreamSynchronize(E->stream));
    H.trbuf.release();
    DevTrace &T = H.rep.tr;
    T = DevTrace{};
    if (max_steps > 0 && max_k > 0) {
        const size_t S = max_steps, K = max_k;
        T.max_steps = max_steps;
        T.max_k = max_k;
        T.first = H.host_step;
        DevBuf &B = H.trbuf;
        HIPX(B.alloc(&T.k, S));
        HIPX(B.alloc(&T.ngroups, S));
        HIPX(B.alloc(&T.naccept, S));
        HIPX(B.alloc(&T.n_after, S));
        HIPX(B.alloc(&T.etot, S));
        HIPX(B.alloc(&T.centre, S * K * 3));
        HIPX(B.alloc(&T.e, S * K * 2));
        HIPX(B.alloc(&T.ninc, S * K * 2));
        HIPX(B.alloc(&T.inchash, S * K * 2));
        HIPX(B.alloc(&T.posehash, S * K));
        HIPX(B.alloc(&T.idx, S * K));
        HIPX(B.alloc(&T.gpick, S * K));
        HIPX(B.alloc(&T.gacc, S * K));
        HIPX(B.alloc(&T.tick, S * 4));
        HIPX(hipMemset(T.tick, 0, S * 4 * sizeof(unsigned long long)));
        HIPX(hipMemset(T.k, 0, S * sizeof(int)));
        HIPX(hipMemset(T.ngroups, 0, S * sizeof(int)));
        HIPX(hipMemset(T.naccept, 0, S * sizeof(int)));
        HIPX(hipMemset(T.n_after, 0, S * sizeof(int)));
        HIPX(hipMemset(T.etot, 0, S * sizeof(double)));
        HIPX(hipMemset(T.centre, 0, S * K * 3 * sizeof(float)));
        HIPX(hipMemset(T.e, 0, S * K * 2 * sizeof(double)));
        HIPX(hipMemset(T.ninc, 0, S * K * 2 * sizeof(int)));
        HIPX(hipMemset(T.inchash, 0, S * K * 2 * sizeof(uint32_t)));
        HIPX(hipMemset(T.posehash, 0, S * K * sizeof(uint32_t)));
        HIPX(hipMemset(T.idx, 0xff, S * K * sizeof(int)));
        HIPX(hipMemset(T.gpick, 0xff, S * K * sizeof(int)));
        HIPX(hipMemset(T.gacc, 0, S * K * sizeof(int)));
    }
    return sync_rep_struct(E, r);
}

int gcmc_engine_get_trace(gcmc_engine *E, int r, gcmc_trace *out)
{
    if (!E || !E->uploaded || !out) return fail(GCMC_E_STATE, "engine not uploaded or null trace");
    if (r < 0 || r >= (int)E->reps.size()) return fail(GCMC_E_ARG, "bad replica");
    const DevTrace &T = E->reps[r]->rep.tr;
    if (!T.k) return fail(GCMC_E_STATE, "tracing is off for this replica");
    if (out->max_steps != T.max_steps || out->max_k != T.max_k)
        return fail(GCMC_E_ARG, "trace shape differs from gcmc_engine_set_trace");
    HIPX(hipStreamSynchronize(E->stream));
    const size_t S = T.max_steps, K = T.max_k;
    HIPX(hipMemcpy(out->k, T.k, S * sizeof(int), hipMemcpyDeviceToHost));
    HIPX(hipMemcpy(out->ngroups, T.ngroups, S * sizeof(int), hipMemcpyDeviceToHost));
    HIPX(hipMemcpy(out->naccept, T.naccept, S * sizeof(int), hipMemcpyDeviceToHost));
    HIPX(hipMemcpy(out->n_after, T.n_after, S * sizeof(int), hipMemcpyDeviceToHost));
    HIPX(hipMemcpy(out->etot, T.etot, S * sizeof(double), hipMemcpyDeviceToHost));
    HIPX(hipMemcpy(out->centre, T.centre, S * K * 3 * sizeof(float), hipMemcpyDeviceToHost));
    HIPX(hipMemcpy(out->e, T.e, S * K * 2 * sizeof(double), hipMemcpyDeviceToHost));
    HIPX(hipMemcpy(out->ninc, T.ninc, S * K * 2 * sizeof(int), hipMemcpyDeviceToHost));
    HIPX(hipMemcpy(out->inchash, T.inchash, S * K * 2 * sizeof(uint32_t), hipMemcpyDeviceToHost));
    HIPX(hipMemcpy(out->posehash, T.posehash, S * K * sizeof(uint32_t), hipMemcpyDeviceToHost));
    HIPX(hipMemcpy(out->idx, T.idx, S * K * sizeof(int), hipMemcpyDeviceToHost));
    HIPX(hipMemcpy(out->gpick, T.gpick, S * K * sizeof(int), hipMemcpyDeviceToHost));
    HIPX(hipMemcpy(out->gacc, T.gacc, S * K * sizeof(int), hipMemcpyDeviceToHost));
    if (out->tick) HIPX(hipMemcpy(out->tick, T.tick, S * 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    return 0;
}

int gcmc_engine_eval_poses(gcmc_engine *E, int r, int species, int npose, const float *centres,
                           const float *poses, const int *exclude, double *energy, int *ninc,
                           uint32_t *inchash)
{
    if (!E || !E->uploaded) return fail(GCMC_E_STATE, "engine not uploaded");
    if (r < 0 || r >= (int)E->reps.size()) return fail(GCMC_E_ARG, "bad replica");
    HostRep &H = *E->reps[r];
    if (species < 0 || species >= H.info.fragTypeNum) return fail(GCMC_E_ARG, "bad species");
    if (npose < 0) return fail(GCMC_E_ARG, "negative pose count");
    if (npose == 0) return 0;
    if (!centres || !poses || !energy) return fail(GCMC_E_ARG, "null array");
    const int nf = H.frag[species].num_atoms;
    std::vector<int> ex(npose, -1);
    if (exclude) ex.assign(exclude, exclude + npose);
    DevBuf B;
    float *dc = nullptr, *dp = nullptr;
    int *dx = nullptr, *dn = nullptr;
    double *de = nullptr;
    uint32_t *dh = nullptr;
    hipError_t e = put(&dc, B, centres, 3 * (size_t)npose);
    if (e == hipSuccess) e = put(&dp, B, poses, 3 * (size_t)npose * nf);
    if (e == hipSuccess) e = put(&dx, B, ex.data(), ex.size());
    if (e == hipSuccess) e = B.alloc(&de, npose);
    if (e == hipSuccess) e = B.alloc(&dn, npose);
    if (e == hipSuccess) e = B.alloc(&dh, npose);
    if (e == hipSuccess) {
        dispatch<RunPoses>(E, npose, r, species, (const float *)dc, (const float *)dp, (const int *)dx, de, dn, dh);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(E->stream);
    if (e == hipSuccess) e = hipMemcpy(energy, de, npose * sizeof(double), hipMemcpyDeviceToHost);
    if (e == hipSuccess && ninc) e = hipMemcpy(ninc, dn, npose * sizeof(int), hipMemcpyDeviceToHost);
    if (e == hipSuccess && inchash) e = hipMemcpy(inchash, dh, npose * sizeof(uint32_t), hipMemcpyDeviceToHost);
    B.release();
    if (e != hipSuccess) return fail(GCMC_E_HIP, std::string("eval_poses: ") + hipGetErrorString(e));
    return 0;
}

namespace {
// cell of a coordinate along one axis: truncation toward zero, non-negative
// remainder (the cavity grid's rule, gcmc/python/dataset.py:65-68), FP64
__device__ inline long long occ_cell(double x, double o, double dx, int n)
{
    const long long q = (long long)trunc((x - o) / dx);
    const long long m = q % n;
    return m < 0 ? m + n : m;
}

// one thread per (slot, atom) of species spi; slots at or beyond the
// species' current count exit
__global__ __launch_bounds__(256) void k_occupancy(const Rep *reps, const State *st, int r, int spi, double o0,
                                                   double o1, double o2, double dx, int n0, int n1, int n2,
                                                   unsigned int *counts)
{
    const Rep &C = reps[r];
    const Species &F = C.sp[spi];
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)st[r].totalNum[spi] * F.nf) return;
    const float4 p = C.atom[F.atomBase + i];
    const long long lin = (occ_cell((double)p.x, o0, dx, n0) * n1 + occ_cell((double)p.y, o1, dx, n1)) * n2 +
                          occ_cell((double)p.z, o2, dx, n2);
    atomicAdd(&counts[lin], 1u);
}
}  // namespace

int gcmc_engine_occupancy(gcmc_engine *E, int r, int species, const double *origin, double spacing,
                          const int *n, unsigned int *counts)
{
    if (!E || !E->uploaded) return fail(GCMC_E_STATE, "engine not uploaded");
    if (r < 0 || r >= (int)E->reps.size()) return fail(GCMC_E_ARG, "bad replica");
    HostRep &H = *E->reps[r];
    if (species < 0 || species >= H.info.fragTypeNum) return fail(GCMC_E_ARG, "bad species");
    if (!origin || !n || !counts) return fail(GCMC_E_ARG, "null array");
    if (!(spacing > 0.0)) return fail(GCMC_E_ARG, "spacing must be positive");
    if (n[0] < 1 || n[1] < 1 || n[2] < 1 || (long long)n[0] * n[1] * n[2] > (1ll << 31))
        return fail(GCMC_E_ARG, "grid dimensions must be 1..2^31 cells in total");
    const long long slots = (long long)H.frag[species].maxNum * H.frag[species].num_atoms;
    if (slots == 0) return 0;
    hipLaunchKernelGGL(k_occupancy, dim3((unsigned)((slots + 255) / 256)), dim3(256), 0, E->stream, E->d_rep,
                       E->d_state, r, species, origin[0], origin[1], origin[2], spacing, n[0], n[1], n[2], counts);
    HIPX(hipGetLastError());
    return 0;
}

int gcmc_engine_set_ewald(gcmc_engine *E, int r, float beta, float grid_spacing)
{
    if (!E) return fail(GCMC_E_ARG, "null engine");
    if (E->uploaded) return fail(GCMC_E_STATE, "set Ewald parameters before upload()");
    if (r < 0 || r >= (int)E->reps.size()) return fail(GCMC_E_ARG, "bad replica");
    if (!(beta > 0.0f) || !(grid_spacing > 0.0f)) return fail(GCMC_E_ARG, "beta and grid spacing must be positive");
    E->reps[r]->ew_beta = beta;
    E->reps[r]->ew_spacing = grid_spacing;
    return 0;
}

int gcmc_engine_ewald_potential(gcmc_engine *E, int r, int n, const float *xyz, float *phi)
{
    if (!E || !E->uploaded) return fail(GCMC_E_STATE, "engine not uploaded");
    if (r < 0 || r >= (int)E->reps.size()) return fail(GCMC_E_ARG, "bad replica");
    if (!E->reps[r]->pme) return fail(GCMC_E_STATE, "replica has no Ewald electrostatics");
    if (n < 0 || (n > 0 && (!xyz || !phi))) return fail(GCMC_E_ARG, "bad arguments");
    if (n == 0) return 0;
    DevBuf B;
    float *dx = nullptr, *dp = nullptr;
    hipError_t e = put(&dx, B, xyz, 3 * (size_t)n);
    if (e == hipSuccess) e = B.alloc(&dp, (size_t)n);
    if (e != hipSuccess) {
        B.release();
        return fail(GCMC_E_HIP, std::string("ewald_potential: ") + hipGetErrorString(e));
    }
    int rc = gcmc_pme_gather(E->reps[r]->pme, n, dx, dp, E->stream);
    if (!rc) {
        e = hipStreamSynchronize(E->stream);
        if (e == hipSuccess) e = hipMemcpy(phi, dp, n * sizeof(float), hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = fail(GCMC_E_HIP, std::string("ewald_potential: ") + hipGetErrorString(e));
    }
    B.release();
    return rc;
}

int gcmc_engine_adjoint(gcmc_engine *E, int r, int species, int npose, const float *centres,
                        const float *poses, const float *eps_kcal, const float *rmin_half, const float *dq,
                        double *energy, float *dEdx, float *partial, void *stream)
{
    if (!E || !E->uploaded) return fail(GCMC_E_STATE, "engine not uploaded");
    if (r < 0 || r >= (int)E->reps.size()) return fail(GCMC_E_ARG, "bad replica");
    HostRep &H = *E->reps[r];
    if (species < 0 || species >= H.info.fragTypeNum) return fail(GCMC_E_ARG, "bad species");
    if (npose < 0) return fail(GCMC_E_ARG, "negative pose count");
    if (npose == 0) return 0;
    if (!centres || !poses || !eps_kcal || !rmin_half || !energy || !dEdx || !partial)
        return fail(GCMC_E_ARG, "null array");
    if (H.pme) return fail(GCMC_E_ARG, "the adjoint covers plain Coulomb, not Ewald electrostatics");
    const int nf = H.frag[species].num_atoms;
    const int X = H.rep.X, Y = H.rep.Y;
    const int adj_off = (E->lay.total + 15) & ~15;
    const int lds = adj_off + adj_lds_bytes(X, Y);
    if (lds > 65536) return fail(GCMC_E_ARG, "adjoint LDS layout exceeds 64 KiB (too many atom types)");
    hipStream_t s = (hipStream_t)stream;
#define ADJ_LAUNCH(NFV)                                                                                         \
    hipLaunchKernelGGL((k_adjoint<256, NFV>), dim3(npose), dim3(256), lds, s, E->d_rep, r, species, centres, poses, \
                       eps_kcal, rmin_half, dq, energy, dEdx, partial, E->lay, adj_off)
    if (nf <= 4) ADJ_LAUNCH(4);
    else if (nf <= 8) ADJ_LAUNCH(8);
    else if (nf <= 12) ADJ_LAUNCH(12);
    else if (nf <= 16) ADJ_LAUNCH(16);
    else ADJ_LAUNCH(MAXF);
#undef ADJ_LAUNCH
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(GCMC_E_HIP, std::string("adjoint: ") + hipGetErrorString(e));
    return 0;
}

int gcmc_engine_adjoint_backward(gcmc_engine *E, int npose, int per_pose, int ncol, const float *dEdx,
                                 const float *partial, const float *weights, float *grad_pose, float *grad_cols,
                                 void *stream)
{
    if (!E) return fail(GCMC_E_ARG, "null engine");
    if (npose < 0 || per_pose < 0 || ncol < 0) return fail(GCMC_E_ARG, "negative size");
    if ((npose > 0 && per_pose > 0 && (!dEdx || !grad_pose)) || (ncol > 0 && (!partial || !grad_cols)))
        return fail(GCMC_E_ARG, "null array");
    hipStream_t s = (hipStream_t)stream;
    constexpr int NCH = 32;
    const size_t need = (size_t)NCH * std::max(ncol, 1);
    if (E->adj_work_n < need) {
        if (E->d_adj_work) (void)hipFree(E->d_adj_work);
        E->d_adj_work = nullptr;
        E->adj_work_n = 0;
        HIPX(hipMalloc(&E->d_adj_work, need * sizeof(double)));
        E->adj_work_n = need;
    }
    const size_t nx = (size_t)npose * per_pose;
    if (nx > 0)
        hipLaunchKernelGGL(k_scale_rows, dim3((unsigned)((nx + 255) / 256)), dim3(256), 0, s, npose, per_pose, dEdx,
                           weights, grad_pose);
    if (ncol > 0) {
        hipLaunchKernelGGL(k_adjoint_colsum, dim3((ncol + 63) / 64, NCH), dim3(1024), 0, s, npose, ncol, partial,
                           weights, E->d_adj_work);
        hipLaunchKernelGGL(k_adjoint_chunksum, dim3((ncol + 255) / 256), dim3(256), 0, s, NCH, ncol, E->d_adj_work,
                           grad_cols);
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(GCMC_E_HIP, std::string("adjoint backward: ") + hipGetErrorString(e));
    return 0;
}

}  // extern "C"

// Workspace of gcmc_adjoint_backward_noalloc: FP64 column partial sums, one
// row of at most ADJ_NOALLOC_MAX_COLS per pose chunk.  One copy per device,
// held until the process ends; every call on a device shares it (each call
// writes all the entries it then reads, so nothing needs clearing; calls
// that overlap on one device race).
namespace {
constexpr int ADJ_NOALLOC_NCH = 32;
constexpr int ADJ_NOALLOC_MAX_COLS = 4096;
constexpr int ADJ_NOALLOC_MAX_DEVICES = 64;
__device__ __align__(256) double scratch[ADJ_NOALLOC_NCH * ADJ_NOALLOC_MAX_COLS];

std::mutex g_scratch_mu;
double *g_scratch_ptr[ADJ_NOALLOC_MAX_DEVICES] = {};

// Device address of the current device's `scratch`, looked up once per device
// with that device current (the array's name in host code is only a host-side
// stand-in, not a device address).
int scratch_address(double **out)
{
    int dev = -1;
    HIPX(hipGetDevice(&dev));
    if (dev < 0 || dev >= ADJ_NOALLOC_MAX_DEVICES) return fail(GCMC_E_DEVICE, "device index beyond the workspace table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        HIPX(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(GCMC_E_HIP, "hipGetSymbolAddress returned a null workspace address");
        g_scratch_ptr[dev] = (double *)p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}
}  // namespace

extern "C" {

int gcmc_adjoint_backward_noalloc(int npose, int per_pose, int ncol, const float *dEdx, const float *partial,
                                  const float *weights, float *grad_pose, float *grad_cols, void *stream)
{
    if (npose < 0 || per_pose < 0 || ncol < 0) return fail(GCMC_E_ARG, "negative size");
    if (ncol > ADJ_NOALLOC_MAX_COLS)
        return fail(GCMC_E_ARG, "ncol " + std::to_string(ncol) + " exceeds the fixed workspace (at most " +
                                    std::to_string(ADJ_NOALLOC_MAX_COLS) + " columns)");
    if ((npose > 0 && per_pose > 0 && (!dEdx || !grad_pose)) || (ncol > 0 && (!partial || !grad_cols)))
        return fail(GCMC_E_ARG, "null array");
    double *work = nullptr;
    int rc = scratch_address(&work);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    const size_t nx = (size_t)npose * per_pose;
    if (nx > 0)
        hipLaunchKernelGGL(k_scale_rows, dim3((unsigned)((nx + 255) / 256)), dim3(256), 0, s, npose, per_pose, dEdx,
                           weights, grad_pose);
    if (ncol > 0) {
        hipLaunchKernelGGL(k_adjoint_colsum, dim3((ncol + 63) / 64, ADJ_NOALLOC_NCH), dim3(1024), 0, s, npose, ncol,
                           partial, weights, work);
        hipLaunchKernelGGL(k_adjoint_chunksum, dim3((ncol + 255) / 256), dim3(256), 0, s, ADJ_NOALLOC_NCH, ncol, work,
                           grad_cols);
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(GCMC_E_HIP, std::string("adjoint backward: ") + hipGetErrorString(e));
    return 0;
}

int gcmc_host_xcd_exchange(int nclass, int n, const int *best, const double *fin, const double *evt, int *next,
                           int *nexchange)
{
    if (nclass < 1 || n < 0 || (n > 0 && (!best || !evt || !next)) || !fin || !nexchange)
        return fail(GCMC_E_ARG, "bad arguments");
    for (int r = 0; r < n; ++r)
        if (best[r] < 0 || best[r] >= nclass) return fail(GCMC_E_ARG, "class outside [0, nclass)");
    std::vector<int> b(best, best + n), nx;
    std::vector<double> f(fin, fin + nclass), e(evt, evt + n);
    std::vector<std::pair<int, int>> pairs;
    xcd_exchanges(nclass, b, f, e, {}, nx, pairs);
    std::copy(nx.begin(), nx.end(), next);
    *nexchange = (int)pairs.size();
    return 0;
}

int gcmc_host_within(int n, const float *a, const float *b, const float *cryst, float cutoff,
                     unsigned char *out)
{
    if (n < 0 || (n > 0 && (!a || !b || !out)) || !cryst) return fail(GCMC_E_ARG, "bad arguments");
    AxisCut ax[3];
    for (int i = 0; i < 3; ++i) {
        ax[i].L = cryst[i];
        ax[i].t[0] = shift_threshold(cryst[i], 1);
        ax[i].t[1] = shift_threshold(cryst[i], 2);
        ax[i].t[2] = shift_threshold(cryst[i], 3);
    }
    const float r2max = r2_threshold(cutoff);
    for (int q = 0; q < n; ++q)
        out[q] = within(a[3 * q], a[3 * q + 1], a[3 * q + 2], b[3 * q], b[3 * q + 1], b[3 * q + 2], ax, r2max) ? 1 : 0;
    return 0;
}

int gcmc_host_stencil_covers(int n, const float *centres, const float *groups, const float *cryst,
                             const float *startxyz, float cutoff, float cell_size, unsigned char *out)
{
    if (n < 0 || (n > 0 && (!centres || !groups || !out)) || !cryst || !startxyz)
        return fail(GCMC_E_ARG, "bad arguments");
    int nc[3];
    float inv_h[3], L[3], st[3];
    for (int i = 0; i < 3; ++i) {
        L[i] = cryst[i];
        st[i] = startxyz[i];
    }
    cell_geometry(L, cell_size, nc, inv_h);
    const std::vector<int> sten = make_stencil(L, nc, cutoff, STENCIL_EPS);
    float h[3], r2;
    int prune;
    stencil_prune(L, nc, cutoff, h, &prune, &r2);
    for (int q = 0; q < n; ++q) {
        const int cc = cell_index(L, st, inv_h, nc, centres[3 * q], centres[3 * q + 1], centres[3 * q + 2]);
        const int cg = cell_index(L, st, inv_h, nc, groups[3 * q], groups[3 * q + 1], groups[3 * q + 2]);
        const int c0 = cc / (nc[1] * nc[2]), c1 = (cc / nc[2]) % nc[1], c2 = cc % nc[2];
        const int ci[3] = {c0, c1, c2};
        float u[3];
        cell_offset(L, st, h, ci, centres[3 * q], centres[3 * q + 1], centres[3 * q + 2], u);
        out[q] = 0;
        for (int code : sten) {
            const int d[3] = {(code & 255) - 128, ((code >> 8) & 255) - 128, ((code >> 16) & 255) - 128};
            if (!stencil_cell_near(u, d, h, prune, r2)) continue;     // the kernel skips this cell
            int i0 = c0 + (code & 255) - 128, i1 = c1 + ((code >> 8) & 255) - 128, i2 = c2 + ((code >> 16) & 255) - 128;
            i0 = ((i0 % nc[0]) + nc[0]) % nc[0];
            i1 = ((i1 % nc[1]) + nc[1]) % nc[1];
            i2 = ((i2 % nc[2]) + nc[2]) % nc[2];
            if ((i0 * nc[1] + i1) * nc[2] + i2 == cg) {
                out[q] = 1;
                break;
            }
        }
    }
    return 0;
}

int runGCMC_hip(const InfoStruct *info, AtomArray *fragmentInfo, residue *residueInfo, Atom *atomInfo,
                const float *grid, const float *ff, const int *moveArray)
{
    if (!info) return fail(GCMC_E_ARG, "null info");
    gcmc_engine *E = nullptr;
    gcmc_tuning t{};
    int rc = gcmc_engine_create(&E, &t);
    if (rc) return rc;
    rc = gcmc_engine_add_system(E, info, fragmentInfo, residueInfo, atomInfo, grid, ff, moveArray,
                                info->mcsteps);
    if (rc >= 0) rc = gcmc_engine_upload(E);
    if (rc >= 0) rc = gcmc_engine_run(E, info->mcsteps, GCMC_MODE_BATCHED);
    if (rc >= 0) rc = gcmc_engine_download(E, 0, fragmentInfo, residueInfo, atomInfo);
    std::string keep = g_err;
    gcmc_engine_destroy(E);
    g_err = keep;
    return rc < 0 ? rc : 0;
}

}  // extern "C"
