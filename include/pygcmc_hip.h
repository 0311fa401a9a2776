/*
 * pygcmc_hip.h -- C ABI of the MI355X GCMC hot path (libpygcmc_hip.so).
 *
 * Plain pointers and sizes only.  Every function returns 0 on success or a
 * negative status; gcmc_last_error() returns the message of the last failure
 * on the calling thread.
 */
#ifndef PYGCMC_HIP_H_
#define PYGCMC_HIP_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Struct ABI of pyGCMC, byte for byte (gcmc/cpp/gcmc.h:47-100):
 * Atom 20 B, AtomArray 436 B, InfoStruct 76 B, residue 20 B. */
typedef struct Atom {
    float position[3];
    float charge;
    int type;
} Atom;

typedef struct AtomArray {
    char name[4];
    int startRes;
    float muex;
    float conc;
    int confBias;
    float mcTime;
    int totalNum;
    int maxNum;
    int num_atoms;
    Atom atoms[20];
} AtomArray;

typedef struct InfoStruct {
    int mcsteps;
    float cutoff;
    float grid_dx;
    float startxyz[3];
    float cryst[3];
    int showInfo;
    float cavityFactor;
    int fragTypeNum;
    int totalGridNum;
    int totalResNum;
    int totalAtomNum;
    int ffXNum;
    int ffYNum;
    int PME;
    unsigned int seed;
} InfoStruct;

typedef struct residue {
    float position[3];
    int atomNum;
    int atomStart;
} residue;

#define GCMC_OK 0
#define GCMC_E_ARG (-1)
#define GCMC_E_HIP (-2)
#define GCMC_E_STATE (-3)
#define GCMC_E_DEVICE (-4)

/* ---- drop-in entry --------------------------------------------------------
 * Replaces runGCMC_cuda (gcmc/cpp/gcmc.cu:21-118, declared at
 * gcmc/cpp/gcmc.cpp:159).  Same seven arguments and the same in-place
 * contract: runs info->mcsteps moves of moveArray and writes back
 * fragmentInfo[].totalNum, residueInfo[] and atomInfo[].  The reference
 * returns void; this returns a status instead of crashing. */
int runGCMC_hip(const InfoStruct *info, AtomArray *fragmentInfo, residue *residueInfo,
                Atom *atomInfo, const float *grid, const float *ff, const int *moveArray);

const char *gcmc_last_error(void);

/* ---- engine: many independent systems ("replicas") on one GPU ------------
 * The reference runs one system per call (gcmc/cpp/gcmc.cu:21); these entry
 * points split that call into create / add / run / download so that many
 * systems of uneven size share one persistent kernel. */
typedef struct gcmc_engine gcmc_engine;

typedef struct gcmc_tuning {
    float cell_size;     /* target edge (A) of the group-centre cell list; 0 = 5 A */
    int persistent_wgs;  /* workgroups of the batched kernel; 0 = per CU by wg_size */
    int bucket_cap;      /* mobile groups per cell before overflow; 0 = auto */
    int pair_lds;        /* 0 (default) or 1 = hold the system's LJ pair table (X x Y
                            entries) in each workgroup's LDS when it fits in 16 KiB, copied
                            with global_load_lds when a workgroup turns to a system with
                            another table, faster on the 64-replica benchmark; -1 = read
                            it from global memory; both give identical results */
    int wg_size;         /* threads per workgroup: 256, 512 or 1024; 0 = 256 */
    int fp32_coulomb;    /* 1 = Coulomb pair terms in FP32 like the LJ terms; 0 (default) =
                            Coulomb in FP64 from the exactly formed FP32 displacement */
    int claim_affinity;  /* batched form: 0 = claim moves with at most 4, then at most 16
                            unclaimed items first, each class (and the larger moves) by the
                            system's estimated work left; N > 0 = the same, except that a move
                            of at least N unclaimed items of the system whose tables a workgroup
                            holds in LDS comes before other such moves.  Changes the order of
                            work, never results */
    int xcd_round_robin; /* batched form: 1 = systems to XCDs round-robin by index; 0 (default) =
                            balanced, heaviest first to the least loaded XCD, by estimated work
                            before the first run and by the work measured in the last run after
                            it.  Changes which workgroups serve a system, never results */
    int xcd_search;      /* batched form, balanced: when runs repeat one step range (the
                            snapshot restored between them), the second is balanced by the
                            evaluation time each system took in the first, and up to this many
                            of the next each try exchanges of systems between XCDs of the
                            fastest assignment so far (the i-th XCD to finish last with the
                            i-th to finish first, one exchange of two systems per pair, chosen
                            from the systems' measured evaluation times); the fastest
                            assignment is kept.  0 (default) = 3 tries, -1 = none.  Changes
                            which workgroups serve a system, never results */
} gcmc_tuning;

#define GCMC_MODE_PLAIN 0   /* one move per launch pair, systems one after another */
#define GCMC_MODE_BATCHED 1 /* all systems in one persistent kernel */
#define GCMC_MODE_GRAPH 2   /* BATCHED, launched from a captured hipGraph */

int gcmc_engine_create(gcmc_engine **out, const gcmc_tuning *tuning);
int gcmc_engine_destroy(gcmc_engine *e);
/* Adds one system (arrays as for runGCMC_hip; nmoves entries of moveArray).
 * Returns the replica index (>= 0) or a negative status. */
int gcmc_engine_add_system(gcmc_engine *e, const InfoStruct *info, const AtomArray *frag,
                           const residue *res, const Atom *atoms, const float *grid,
                           const float *ff, const int *moves, int nmoves);
/* Copies all systems to the GPU (call once after the last add). */
int gcmc_engine_upload(gcmc_engine *e);
/* Advances every replica by nsteps moves (fewer at the end of its schedule). */
int gcmc_engine_run(gcmc_engine *e, int nsteps, int mode);
int gcmc_engine_synchronize(gcmc_engine *e);
/* Current step index of a replica. */
int gcmc_engine_step(gcmc_engine *e, int replica);
/* Writes the replica's state back into caller arrays laid out as at add time. */
int gcmc_engine_download(gcmc_engine *e, int replica, AtomArray *frag, residue *res,
                         Atom *atoms);
/* Running FP64 sum of accepted energy changes and per-species move counters
 * (attempts/accepts: fragTypeNum*4 entries, index species*4 + kind). */
int gcmc_engine_stats(gcmc_engine *e, int replica, double *energy, long long *attempts,
                      long long *accepts);
/* Device copy of every replica's current state (positions, charges, types,
 * cell list, species counts, step, statistics and counters), and its return:
 * restore queues device-to-device copies on the engine's stream, so a chain
 * can be rerun from the same start with inputs resident in HBM (the
 * benchmark's passes).  The reference has no counterpart: its runGCMC_cuda
 * (gcmc/cpp/gcmc.cu:21-118) copies the host arrays in on every call.
 * Systems with Ewald electrostatics are refused. */
int gcmc_engine_snapshot(gcmc_engine *e);
int gcmc_engine_restore(gcmc_engine *e);
/* Milliseconds of the last gcmc_engine_run measured with HIP events on the
 * engine's stream: the whole run, and the persistent kernel alone (batched
 * mode; 0 otherwise). */
int gcmc_engine_last_run_ms(gcmc_engine *e, double *ms);
int gcmc_engine_last_kernel_ms(gcmc_engine *e, double *ms);

/* Persistent-kernel launch shape: out4 = {workgroups requested, workgroups
 * resident at once (known after the first batched run, else 0), threads per
 * workgroup, LDS bytes per workgroup}.  No counterpart in the reference. */
int gcmc_engine_launch_info(gcmc_engine *e, int *out4);
/* Work done by a replica since upload: [0] trial evaluations, [1] fragment-
 * environment atom pairs, [2] included environment atoms, [3] included
 * groups, [4] fragment atoms, each summed over evaluations. */
int gcmc_engine_counters(gcmc_engine *e, int replica, unsigned long long *out5);
/* Device wall-clock ticks (100 MHz) a replica spent, batched form: [0] in
 * evaluations (summed over workgroups), [1] in accept phases, [2] in prep
 * phases, [3] from publishing a move to its last evaluation finishing. */
int gcmc_engine_phase_ticks(gcmc_engine *e, int replica, unsigned long long *out4);
/* Ticks inside evaluations (summed): [0] pose and adjacency, [1] stencil
 * counts and scan, [2] candidate group tests, [3] pair passes. */
int gcmc_engine_sub_ticks(gcmc_engine *e, int replica, unsigned long long *out4);

/* ---- step trace (parity tests) --------------------------------------------
 * Per step s (relative to the step when tracing was enabled) and trial t:
 * centre, energies (insert/delete: e[0]=E; translate/rotate: E_old, E_new),
 * included-group count and set hash per evaluation, pose hash, slot, and per
 * selection group g the picked trial and accept flag. */
typedef struct gcmc_trace {
    int max_steps, max_k;
    int *k, *ngroups, *naccept, *n_after;
    double *etot;
    float *centre;      /* [S][K][3] */
    double *e;          /* [S][K][2] */
    int *ninc;          /* [S][K][2] */
    uint32_t *inchash;  /* [S][K][2] */
    uint32_t *posehash; /* [S][K]    */
    int *idx;           /* [S][K]    */
    int *gpick;         /* [S][K]    */
    int *gacc;          /* [S][K]    */
    unsigned long long *tick; /* [S][4] device ticks (100 MHz) of the batched form:
                                 move published, accept start, accept end, next
                                 move prepared; 0 in the plain form */
} gcmc_trace;

int gcmc_engine_set_trace(gcmc_engine *e, int replica, int max_steps, int max_k);
int gcmc_engine_get_trace(gcmc_engine *e, int replica, gcmc_trace *out);

/* ---- single evaluations ---------------------------------------------------
 * Energies (kJ/mol, FP32 pair terms, compensated sums, FP64 result) of npose
 * poses of one species against the replica's current state, each tested
 * around centres[i] and excluding residue exclude[i] (-1: none).  The same
 * kernel path as the MC step (calcEnergy, gcmc/cpp/gcmc_energy.h:142-204). */
int gcmc_engine_eval_poses(gcmc_engine *e, int replica, int species, int npose,
                           const float *centres, const float *poses, const int *exclude,
                           double *energy, int *ninc, uint32_t *inchash);

/* ---- occupancy maps (no counterpart in the reference: its map_generation
 * key is parsed and ignored, gcmc/scripts/mainOld.py) ----------------------
 * Adds to counts[n0][n1][n2] (a DEVICE uint32 array) one count per atom of
 * every molecule of `species` in the replica's current state, in the cell
 * int((x - origin)/spacing) mod n per axis (truncation toward zero,
 * non-negative remainder, FP64: the cavity grid's rule).  Queued on the
 * engine's stream after the runs queued before it; read counts after
 * gcmc_engine_synchronize. */
int gcmc_engine_occupancy(gcmc_engine *e, int replica, int species, const double *origin, double spacing,
                          const int *n, unsigned int *counts);

/* ---- adjoint of the pose energy (no counterpart in the reference) ----------
 * Energy of npose poses of one species (all groups of the replica's current
 * state whose centre passes the group test around centres[i]; nothing
 * excluded) with per-type parameters over the ffYNum column types:
 * eps_kcal[t] (CHARMM epsilon, kcal/mol), rmin_half[t] (CHARMM Rmin/2, A),
 * dq[t] (charge shift of every atom of type t, e; NULL = 0), combined per
 * pair by the rule of gcmc/python/simulation.py:63-72.  Writes energy[i]
 * (kJ/mol), dEdx[i][f][3] = dE_i/dx_f (kJ/mol/A) and partial[i][3][Y] =
 * dE_i/d(dq, eps_kcal, rmin_half).  All pointers are DEVICE pointers;
 * poses [npose][nf][3], centres [npose][3].  Work is queued on `stream`
 * (a hipStream_t; NULL = the default stream) and not waited for. */
int gcmc_engine_adjoint(gcmc_engine *e, int replica, int species, int npose,
                        const float *centres, const float *poses, const float *eps_kcal,
                        const float *rmin_half, const float *dq, double *energy, float *dEdx,
                        float *partial, void *stream);
/* Backward of a loss L with dL/dE_i = weights[i] (NULL = 1): grad_pose[i][k]
 * = weights[i] dEdx[i][k] (per_pose entries per pose) and grad_cols[j] =
 * sum_i weights[i] partial[i][j] over ncol columns (FP64 sums).  DEVICE
 * pointers, queued on `stream` (NULL = the default stream).  Not safe to run
 * concurrently on two streams of one engine (shared column-sum workspace). */
int gcmc_engine_adjoint_backward(gcmc_engine *e, int npose, int per_pose, int ncol,
                                 const float *dEdx, const float *partial, const float *weights,
                                 float *grad_pose, float *grad_cols, void *stream);

/* The same backward without any allocation: no counterpart in the reference
 * (pyGCMC has no adjoint); for callers that cannot allocate device memory --
 * code under HIP graph capture (a captured autograd step), tight serving
 * loops, harnesses that forbid hipMalloc.  Identical arguments and results to
 * gcmc_engine_adjoint_backward minus the engine; its column-sum workspace is
 * a fixed per-device array inside the library (1 MiB of FP64, held until
 * the process ends), so ncol must be at most 4096 (GCMC_E_ARG otherwise).
 * Every call on one device shares that workspace: calls that overlap on a
 * device (two streams, two threads) race.  DEVICE pointers, queued on
 * `stream` (NULL = the default stream). */
int gcmc_adjoint_backward_noalloc(int npose, int per_pose, int ncol, const float *dEdx,
                                  const float *partial, const float *weights, float *grad_pose,
                                  float *grad_cols, void *stream);

/* ---- Ewald electrostatics (no counterpart in the reference) ---------------
 * Before upload: give replica `replica` Ewald electrostatics with splitting
 * beta (1/A) and a PME grid of about grid_spacing A (order-4 B-splines).
 * Its trial energies then use K q q' erfc(beta r)/r in the group-cutoff pair
 * terms plus the reciprocal interaction sum_f q_f phi(x_f) with the PME
 * potential phi of the current state (the moving molecule's own share
 * removed), and the potential is updated after every move.  In the batched
 * and graph forms an engine's systems are either all Ewald systems or none:
 * Ewald systems then run move by move (each move's trials in one launch,
 * accept, PME update), the graph form replaying a captured chunk of moves;
 * the chains equal the plain form's.  The reference ignores InfoStruct.PME, and so
 * do runGCMC_hip and gcmc_engine_add_system.  eval_poses adds the same
 * reciprocal term; its exclude[i], when a molecule of the evaluated species,
 * must be one present in the state. */
int gcmc_engine_set_ewald(gcmc_engine *e, int replica, float beta, float grid_spacing);
/* Reciprocal potential (kJ/mol/e) of the replica's current state at n points
 * xyz[n][3] (host arrays). */
int gcmc_engine_ewald_potential(gcmc_engine *e, int replica, int n, const float *xyz, float *phi);

/* ---- cavity-bias grid (gcmc/python/dataset.py:21-139) --------------------
 * Occupancy of the grid_n[0] x grid_n[1] x grid_n[2] cells of edge grid_dx
 * starting at startxyz by the natom atoms xyz[natom][3] (FP64, host memory):
 * cell int((x - start)/dx) % n per axis, truncation toward zero and a
 * non-negative remainder as the reference's Python computes them.  Writes
 * the free-cell count to *npoints and, when grid is not NULL and the count is
 * at most max_points, the corners (start + i dx, FP64 rounded to FP32) of the
 * free cells to grid[3*npoints] in linear order (x slowest, z fastest).  The
 * reference lists the same cells in CPython set order.  Runs on the current
 * HIP device; synchronous. */
int gcmc_cavity_grid(int natom, const double *xyz, const double *startxyz, double grid_dx,
                     const int *grid_n, int max_points, float *grid, int *npoints);

/* ---- cavity grid in the reference's order (gcmc/python/dataset.py:66-109) -
 * The free cells of the grid_n[0] x grid_n[1] x grid_n[2] grid in the
 * iteration order of the CPython (3.8+) set the reference builds with
 * {(x, y, z) for x in range(n0) for y in range(n1) for z in range(n2)} and
 * discards the occupied cells from.  occupied[n0*n1*n2] (x slowest, z
 * fastest; nonzero = occupied) or NULL for none.  Writes the free-cell count
 * to *count and, when cells is not NULL and the count is at most max_cells,
 * the cells' (x, y, z) indices to cells[3*count].  At most 2^28 cells.  Host
 * code, no device needed. */
int gcmc_grid_set_order(const int *grid_n, const unsigned char *occupied, long long max_cells, int *cells,
                        long long *count);

/* ---- fixed-atom grouping (gcmc/python/dataset.py:175-188) ----------------
 * group[i] for the n atoms xyz[n][3] (FP64, in order): an atom opens a new
 * group when sqrt(dx**2 + dy**2 + dz**2) > cutoff from the first atom of the
 * current group (squares as Python's float power), groups numbered from 0.
 * Host code, no device needed. */
int gcmc_fixed_groups(int n, const double *xyz, double cutoff, int *group);

/* ---- smooth particle-mesh Ewald, reciprocal part ---------------------------
 * No counterpart in the reference (its PME flag, gcmc/cpp/gcmc.h:91, is never
 * read).  Orthorhombic box box[3] (A), grid[3] points (4..4096 each), Ewald
 * splitting beta (1/A), order-4 B-splines, K = 1388.431112 kJ/mol A e^-2.
 * Charge grid Q accumulates spread charges until cleared; solve computes
 * E_rec = 1/2 sum_m G(m)|FFT(Q)(m)|^2 into *energy (a DEVICE double, may be
 * NULL) and the potential grid; gather interpolates the reciprocal potential
 * per unit charge (kJ/mol/e) at points.  xyzq [n][4] = (x, y, z, q) and xyz
 * [n][3] are DEVICE float arrays; work is queued on `stream` (NULL = the
 * default stream).  One plan must not be used from two streams at once. */
typedef struct gcmc_pme gcmc_pme;
int gcmc_pme_create(gcmc_pme **out, const int *grid, const float *box, float beta);
int gcmc_pme_destroy(gcmc_pme *p);
int gcmc_pme_clear(gcmc_pme *p, void *stream);
int gcmc_pme_spread(gcmc_pme *p, int n, const float *xyzq, void *stream);
int gcmc_pme_solve(gcmc_pme *p, double *energy, void *stream);
int gcmc_pme_gather(gcmc_pme *p, int n, const float *xyz, float *phi, void *stream);

/* Counter calibration (no counterpart in the reference): `copies` plain
 * device-to-device copies of `bytes` bytes (a multiple of 16), so a profiler's
 * FETCH_SIZE / WRITE_SIZE readings can be scaled to bytes moved. */
int gcmc_copy_probe(long long bytes, int copies);

/* Host execution of the engine's FP32 group-inclusion test -- the code the
 * GPU runs, compiled for the CPU -- for n pairs of points: out[i] = 1 when
 * distanceP(a_i, b_i, cryst) <= cutoff (gcmc/cpp/gcmc_energy.h:115-122). */
int gcmc_host_within(int n, const float *a, const float *b, const float *cryst, float cutoff,
                     unsigned char *out);

/* Host check of the XCD assignment search (gcmc_tuning.xcd_search); no
 * counterpart in the reference.  From the assignment best[n] (class of each
 * system, nclass classes), the classes' finish times fin[nclass] and the
 * systems' evaluation times evt[n]: next[n] = the assignment one try of the
 * search runs (the i-th latest class paired with the i-th earliest, one
 * exchange of two systems per pair), *nexchange = exchanges made. */
int gcmc_host_xcd_exchange(int nclass, int n, const int *best, const double *fin, const double *evt, int *next,
                           int *nexchange);

/* Host check of the cell-list superset property for n (centre, group) pairs:
 * out[i] = 1 when the group's cell is one of the stencil cells the energy
 * kernel scans for the centre: a stencil cell around the centre's cell (box
 * cryst from startxyz, target cell edge cell_size) whose box is not pruned
 * by its distance from the centre itself.  Every pair that passes
 * gcmc_host_within must give 1. */
int gcmc_host_stencil_covers(int n, const float *centres, const float *groups, const float *cryst,
                             const float *startxyz, float cutoff, float cell_size,
                             unsigned char *out);

#ifdef __cplusplus
}
#endif

#endif /* PYGCMC_HIP_H_ */
