"""Benchmark: GCMC MC steps/s and trial energy evaluations/s on one GPU.

Workload: 64 independent GCMC systems whose fixed-atom counts are
log-uniform in 5,000-60,000 (box 40*(N/2000)^(1/3) A, fixed-atom density of the
2,000-atom sphere system), each with its own seed and muex and the 8 rigid
fragment types + TIP3P water (confBias 400/15/10/20/15/300/100/200/2), 10,000
moves each, FP32 pair energies with compensated sums (Coulomb terms from the
exactly formed FP32 displacement in FP64), FP64 totals and acceptance, all
systems in one persistent kernel.  Synthetic data (seeded).

One bench step is one whole pass of that workload: every system returned to
its initial state on the device (device-to-device copies from a snapshot
taken after upload), then all 10,000 moves of all 64 systems (640,000 MC
steps).  --warmup passes are untimed, --steps passes are timed; every pass
computes the same chains, and the bench checks that they end in the same
state.  The engine assigns systems to XCDs by its cost estimate for the
first pass and by the evaluation time measured in it for the second; the next three
passes each try exchanges of systems between XCDs that finished late and
early, and the fastest assignment serves every pass after them (with
--warmup 5, every timed pass).  --xcd-search -1 stops after the second
pass's balance; --xcd-round-robin 1 keeps the index order.

Prints one JSON line.  `value` is the batched form's aggregate MC steps/s over
all systems with inputs resident in HBM.  Beside it: the plain form (one move
of one system per launch pair) and the batched form on the same first moves,
the graph-launched form, the roofline of the persistent kernel, a lock-step
check of the first moves of every system against the FP64 oracle (the CPU
baseline leg, timed on the host cores), the adjoint and set-up from files.
"""

import argparse
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3   # MI355X FP32 vector peak (packed FP32)
FP32_UNPACKED_TFLOPS = 78.65  # the same without packed FP32 (this build emits none)
FP64_PEAK_TFLOPS = 78.6    # MI355X FP64 vector peak
HBM_PEAK_GBS = 8000.0
FLOP_PER_PAIR = 32         # algorithmic count per fragment-environment pair
BYTES_PER_ITEM = 20        # per included env atom / included group / fragment atom
KCAL_KJ = 4.184e-4         # 1e-4 kcal/mol in kJ/mol


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_replicas(n, moves):
    from pygcmc_amd import synth
    sizes = synth.replica_sizes(n)
    out = []
    for r, na in enumerate(sizes):
        out.append(synth.replica_system(r, int(na), mcsteps=moves))
    return out, sizes


def copies(systems):
    return [{k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in s.items()} for s in systems]


def make_engine(systems, tuning):
    from pygcmc_amd.engine import Engine
    eng = Engine(**tuning)
    for s in systems:
        eng.add_system(s)
    eng.upload()
    return eng


def counters_sum(eng, R):
    tot = np.zeros(5, dtype=np.float64)
    for r in range(R):
        c = eng.counters(r)
        tot += [c['evals'], c['pairs'], c['atoms'], c['groups'], c['frag_atoms']]
    return tot


def cpu_model():
    try:
        for line in open('/proc/cpuinfo'):
            if line.startswith('model name'):
                return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return 'unknown'


def host_threads():
    """Threads for the CPU legs: the host's cores, at most 16."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def e_tol(e):
    e = np.abs(e)
    return np.where(e <= 200.0, KCAL_KJ, np.maximum(KCAL_KJ, 1e-5 * e))


def gpu_traces(systems, moves, tuning):
    """The first `moves` moves of every system in one batched launch, traced."""
    from pygcmc_amd.engine import Engine
    eng = Engine(**tuning)
    gsys = copies(systems)
    ks = []
    for s in gsys:
        eng.add_system(s)
        ks.append(int(s['frag']['confBias'].max()))
    eng.upload()
    for r, k in enumerate(ks):
        eng.set_trace(r, moves, k)
    eng.run(moves, 'batched')
    traces = [eng.get_trace(r) for r in range(len(gsys))]
    for r in range(len(gsys)):
        eng.download(r)
    eng.close()
    return traces, gsys


def cpu_lockstep(systems, traces, gsys, moves, nthreads):
    """CPU baseline and parity sample in one: the FP64 oracle (the
    reference's algorithm, every group tested per trial) runs the first
    `moves` moves of every system on `nthreads` host threads (systems spread
    over the threads), taking the GPU's picks so the chains stay in step, and
    its trial energies, poses, inclusion sets, decisions and final states are
    compared with the GPU's."""
    from oracle import oracle as O
    R = len(systems)
    refs = copies(systems)
    otr = [None] * R
    res = [None] * R
    lock = threading.Lock()
    nxt = [0]

    def worker():
        while True:
            with lock:
                r = nxt[0]
                nxt[0] += 1
            if r >= R:
                return
            k = traces[r]['centre'].shape[1]
            ot = O.new_trace(moves, k)
            O.run(refs[r], first=0, nsteps=moves, trace=ot, follow=traces[r])
            otr[r] = ot

    t0 = time.perf_counter()
    ths = [threading.Thread(target=worker) for _ in range(nthreads)]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    dt = time.perf_counter() - t0
    worst = 0.0
    rel_hi = 0.0
    groups = 0
    mism = 0
    exact = True
    over = 0
    for r in range(R):
        gt, ot = traces[r], otr[r]
        km = np.arange(gt['centre'].shape[1])[None, :] < gt['k'][:, None]
        exact &= bool(np.array_equal(gt['k'], ot['k']) and
                      np.array_equal(gt['centre'][km], ot['centre'][km]) and
                      np.array_equal(gt['posehash'][km], ot['posehash'][km]) and
                      np.array_equal(gt['ninc'][km], ot['ninc'][km]) and
                      np.array_equal(gt['inchash'][km], ot['inchash'][km]) and
                      np.array_equal(gsys[r]['atoms'].view(np.uint8), refs[r]['atoms'].view(np.uint8)) and
                      np.array_equal(gsys[r]['frag']['totalNum'], refs[r]['frag']['totalNum']))
        eo = ot['e'][km]
        err = np.abs(gt['e'][km] - eo)
        band = np.abs(eo) <= 200.0
        if band.any():
            worst = max(worst, float(err[band].max()))
        if (~band).any():
            rel_hi = max(rel_hi, float((err[~band] / np.abs(eo[~band])).max()))
        over += int((err > e_tol(eo)).sum())
        gm = np.arange(gt['gpick'].shape[1])[None, :] < gt['ngroups'][:, None]
        groups += int(gm.sum())
        mism += int(((gt['gpick'][gm] != ot['gpick'][gm]) | (gt['gacc'][gm] != ot['gacc'][gm])).sum())
    return {'seconds': dt, 'moves': R * moves, 'max_abs_dE_kJ': worst, 'rel_err_above_200': rel_hi,
            'energies_over_bar': over, 'groups': groups, 'group_mismatches': mism,
            'bit_exact_poses_sets_states': exact}


def cpu_round_robin(systems, budget_s, nthreads, cells, chunk=5):
    """The oracle running the systems' chains from their first move, in
    chunks of `chunk` moves taken round-robin over the systems, on `nthreads`
    threads, until `budget_s` has passed: MC steps done per second.
    cells=True scans a cell list of group centres instead of every group."""
    from oracle import oracle as O
    R = len(systems)
    states = copies(systems)
    pos = [0] * R
    done = [0] * nthreads
    t_end = time.perf_counter() + budget_s

    def worker(tid):
        while time.perf_counter() < t_end:
            moved = False
            for r in range(tid, R, nthreads):
                if time.perf_counter() >= t_end:
                    break
                n = min(chunk, len(states[r]['moves']) - pos[r])
                if n <= 0:
                    continue
                O.run(states[r], first=pos[r], nsteps=n, cells=cells)
                pos[r] += n
                done[tid] += n
                moved = True
            if not moved:
                return

    t0 = time.perf_counter()
    ths = [threading.Thread(target=worker, args=(i,)) for i in range(nthreads)]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    dt = time.perf_counter() - t0
    return sum(done) / dt, sum(done), dt, max(pos)


def measured_traffic(moves, replicas, tuning):
    """HBM bytes of one k_batched launch (one pass of the workload) from the
    committed rocprofv3 --pmc record (FETCH_SIZE and WRITE_SIZE scaled by a
    copy kernel of known size), when that record was measured on this same
    workload and setting; else None."""
    path = os.path.join(ROOT, 'profiles', 'k_batched_64replicas.pmc.txt')
    if not os.path.exists(path):
        return None, None
    rec = {}
    for line in open(path):
        if '=' in line and not line.startswith('#'):
            k, v = line.strip().split('=', 1)
            rec[k] = v
    same = (int(rec.get('moves', -1)) == moves and int(rec.get('replicas', -1)) == replicas and
            float(rec.get('cell_size', -1)) == float(tuning.get('cell_size', 0)) and
            int(rec.get('wg_size', -1)) == int(tuning.get('wg_size', 0) or 256) and
            int(rec.get('pair_lds', 1)) == int(tuning.get('pair_lds', 1)) and
            int(rec.get('fp32_coulomb', 0)) == int(tuning.get('fp32_coulomb', 0)) and
            int(rec.get('xcd_round_robin', -1)) == int(tuning.get('xcd_round_robin', 0)) and
            int(rec.get('xcd_search', -2)) == int(tuning.get('xcd_search', 0)))
    if not same:
        return None, None
    return float(rec['traffic_bytes']), 'profiles/k_batched_64replicas.pmc.txt (rocprofv3 --pmc, same workload)'


def setup_bench():
    """Set-up of a 4wp7-sized system (30,520 fixed atoms, 57,077 waters,
    methanols; the size of pygcmc's logged run, test/4wp7/gcmc.75.out:60-105)
    from PDB + TOP + force-field files to the seven runGCMC arrays: file
    loading, then set-up with the reference's grid order (host) and with the
    grid built on the GPU.  pygcmc logged 14.2 s of Python set-up for its
    run (gcmc.75.out:111)."""
    import shutil
    import tempfile
    from pygcmc_amd import synth
    from pygcmc_amd.system import GCMCSystem
    d = tempfile.mkdtemp(prefix='gcmc_setup_')
    try:
        base = synth.write_large_files(d)
        out = {}
        for order in ('cpython', 'linear'):
            s = GCMCSystem(fragment_names=['MEOH', 'SOL'], conc=[0.25, 55.0], muex=[-2.04, -5.04], conf=[10, 2],
                           mctime=[0.2, 0.8], mcsteps=100, seed=1, ff_dir=os.path.join(d, 'charmm36.ff'),
                           grid_order=order)
            s.load(base + '.pdb', top=base + '.top')
            s.setup()
            out['load_s'] = round(s.timings['load_s'], 3)
            out['setup_s_' + ('host_grid' if order == 'cpython' else 'gpu_grid')] = round(s.timings['setup_s'], 3)
            info = s.arrays[0][0]
            out['pdb_atoms'] = len(s.pdb)
            out['grid_points'] = int(info['totalGridNum'])
        out['reference_python_setup_s'] = 14.2
        return out
    finally:
        shutil.rmtree(d, ignore_errors=True)


def adjoint_bench(npose, reps=5):
    """Forward + backward of the pose-energy adjoint: npose poses (equal
    shares of the 8 fragment types, uniform centres, Shoemake orientations)
    against the 20,000-atom protein of the multi-fragment system (80 A box,
    no water), timed with torch events on the current stream."""
    import torch
    from pygcmc_amd import adjoint, synth
    from pygcmc_amd.engine import Engine
    s = synth.multi_fragment(mcsteps=10)
    eng = Engine()
    eng.add_system(s)
    eng.upload()
    eps, rmin = adjoint.type_parameters(s)
    rng = np.random.default_rng(4096)
    T = 8
    per = npose // T
    batches = []
    for spi in range(T):
        F = s['frag'][spi]
        nf = int(F['num_atoms'])
        t = F['atoms'][:nf]['position'].astype(np.float64)
        info = s['info'][0]
        c = info['startxyz'].astype(np.float64) + rng.random((per, 3)) * info['cryst'].astype(np.float64)
        u1, u2, u3 = rng.random((3, per))
        q = np.stack([np.sqrt(1 - u1) * np.sin(2 * np.pi * u2), np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                      np.sqrt(u1) * np.sin(2 * np.pi * u3), np.sqrt(u1) * np.cos(2 * np.pi * u3)], 1)
        x, y, z, w = q.T
        Rm = np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w),
                       2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w),
                       2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], 1).reshape(per, 3, 3)
        p = np.einsum('nij,aj->nai', Rm, t) + c[:, None, :]
        batches.append((spi, torch.tensor(c, dtype=torch.float32, device='cuda'),
                        torch.tensor(p, dtype=torch.float32, device='cuda')))

    def once():
        e = eps.clone().requires_grad_(True)
        r = rmin.clone().requires_grad_(True)
        for spi, c, p in batches:
            pp = p.clone().requires_grad_(True)
            E = adjoint.pose_energy(eng, 0, spi, c, pp, e, r)
            E.sum().backward()
        return e.grad, r.grad

    once()
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(reps):
        once()
    ev1.record()
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / reps
    eng.close()
    nfix = int(s['frag'][0]['startRes'])
    env = int(s['res'][nfix - 1]['atomStart'] + s['res'][nfix - 1]['atomNum'])
    return {'poses': per * T, 'ms': round(ms, 3), 'poses_per_s': round(per * T / (ms / 1e3), 1),
            'env_atoms': env, 'dtype': 'fp32 (energies fp64 from compensated fp32)'}


def ewald_bench(moves, warm=32):
    """The 300,000-atom slab (150 x 150 x 120 A, 15 A pore, water +
    methylammonium / acetate, PME 1 A grid, beta 0.34 /A, 12 A cutoff) in
    Ewald mode: `moves` moves after `warm` untimed ones, in the plain form
    (launches issued move by move) and the graph form (captured 32-move
    chunks replayed); both compute the same chain."""
    from pygcmc_amd import synth
    from pygcmc_amd.engine import Engine
    s = synth.slab(mcsteps=50000)
    last = s['res'][int(s['frag'][0]['startRes']) - 1]
    out = {'fixed_atoms': int(last['atomStart'] + last['atomNum']), 'moves': moves}
    states = {}
    for mode in ('plain', 'graph'):
        g = copies([s])[0]
        eng = Engine()
        eng.add_system(g)
        eng.set_ewald(0, 0.34, 1.0)
        eng.upload()
        eng.run(warm, mode)
        ms = eng.run(moves, mode)
        eng.download(0)
        states[mode] = (g['atoms'].tobytes(), eng.stats(0)['energy'])
        out[f'{mode}_ms'] = round(ms, 2)
        out[f'{mode}_steps_per_s'] = round(moves / (ms / 1e3), 1)
        out['molecules_after'] = int(g['frag']['totalNum'].sum())
        eng.close()
    out['same_chain'] = states['plain'] == states['graph']
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=3, help='timed passes of the whole workload')
    ap.add_argument('--warmup', type=int, default=5,
                    help='untimed passes before them (5: the XCD assignment search ends within them)')
    ap.add_argument('--replicas', type=int, default=64)
    ap.add_argument('--moves', type=int, default=10000, help='moves per system in one pass')
    ap.add_argument('--plain-moves', type=int, default=200)
    ap.add_argument('--graph-passes', type=int, default=2)
    ap.add_argument('--lockstep-moves', type=int, default=200,
                    help='first moves of every system run by the FP64 oracle (CPU baseline + parity)')
    ap.add_argument('--cpu-seconds', type=float, default=10.0,
                    help='budget of each round-robin CPU leg (one core; cell-list variant on all cores)')
    # defaults: the setting pygcmc_amd.autotune picks on this workload (its
    # timings differ by under 2% across cell sizes 5-7.5 A); --autotune 1
    # searches again
    ap.add_argument('--cell-size', type=float, default=7.5)
    ap.add_argument('--wgs', type=int, default=0)
    ap.add_argument('--wg-size', type=int, default=256)
    ap.add_argument('--pair-lds', type=int, default=0)
    ap.add_argument('--fp32-coulomb', type=int, default=0)
    ap.add_argument('--xcd-round-robin', type=int, default=0)
    ap.add_argument('--xcd-search', type=int, default=0,
                    help='passes that may try an exchange of two systems between XCDs (0 = 3, -1 = none)')
    ap.add_argument('--kernel-only', action='store_true',
                    help='only the timed passes (for profiler runs): no graph, plain, lock-step, '
                         'adjoint, set-up or CPU legs')
    ap.add_argument('--adjoint-poses', type=int, default=4096)
    ap.add_argument('--ewald-moves', type=int, default=256,
                    help='moves of the 300k-atom PME slab timed in the plain and graph forms (0: skip)')
    ap.add_argument('--setup-bench', type=int, default=1,
                    help='time set-up of a 4wp7-sized system from PDB/TOP/force-field files')
    ap.add_argument('--autotune', type=int, default=0, help='pick cell/workgroup sizes by timing short runs')
    ap.add_argument('--copy-probe', type=int, default=0, metavar='MIB',
                    help='only run 3 plain copies of MIB MiB (profiler counter calibration) and exit')
    args = ap.parse_args()
    if args.copy_probe:
        import ctypes
        from pygcmc_amd import engine
        L = engine.lib()
        L.gcmc_copy_probe.argtypes = [ctypes.c_longlong, ctypes.c_int]
        engine._check(L.gcmc_copy_probe(args.copy_probe << 20, 3))
        print(json.dumps({'copy_probe_bytes': args.copy_probe << 20, 'copies': 3}), flush=True)
        return
    if args.gpus != 1:
        print(f'bench.py measures one GPU; --gpus {args.gpus} is not supported in this build '
              '(independent systems would shard one group per GPU, see DESIGN.md)', file=sys.stderr)
        sys.exit(2)
    if args.steps < 1 or args.warmup < 0 or args.moves < 1:
        print('--steps >= 1, --warmup >= 0 and --moves >= 1 are required', file=sys.stderr)
        sys.exit(2)

    M = args.moves
    t0 = time.time()
    systems, sizes = build_replicas(args.replicas, M)
    R = len(systems)
    log(f'built {R} systems ({int(sizes.sum())} fixed atoms, {M} moves each) in {time.time() - t0:.1f}s')
    tuning = dict(cell_size=args.cell_size, persistent_wgs=args.wgs, wg_size=args.wg_size, pair_lds=args.pair_lds,
                  fp32_coulomb=args.fp32_coulomb, xcd_round_robin=args.xcd_round_robin, xcd_search=args.xcd_search)
    tune_table = None
    if args.autotune:
        from pygcmc_amd.autotune import autotune
        t1 = time.time()
        tuning, tune_table = autotune(systems, log=log)
        tuning['fp32_coulomb'] = args.fp32_coulomb
        tuning['xcd_round_robin'] = args.xcd_round_robin
        tuning['xcd_search'] = args.xcd_search
        log(f'autotune: {tuning} in {time.time() - t1:.1f}s')

    eng = make_engine(copies(systems), tuning)
    eng.snapshot()

    def one_pass(mode='batched'):
        eng.restore()
        ms = eng.run(M, mode)
        for r in range(R):
            if eng.step(r) != M:
                raise RuntimeError(f'system {r} ran {eng.step(r)} of {M} moves')
        return ms

    energies_w = None
    for w in range(args.warmup):
        ms = one_pass()
        log(f'warm-up pass {w + 1}/{args.warmup}: {ms:.1f} ms')
        energies_w = [eng.stats(r)['energy'] for r in range(R)]
    run_ms, kern_ms = [], []
    eng.synchronize()
    t_wall = time.perf_counter()
    for k in range(args.steps):
        run_ms.append(one_pass())
        kern_ms.append(eng.last_kernel_ms())
    eng.synchronize()
    wall_s = time.perf_counter() - t_wall
    # every pass computes the same chains from the same start
    energies0 = [eng.stats(r)['energy'] for r in range(R)]
    if energies_w is not None and energies_w != energies0:
        raise RuntimeError('a timed pass ended in another state than the warm-up pass')
    cnt = counters_sum(eng, R)
    launch = eng.launch_info()
    steps_per_pass = R * M
    value = steps_per_pass * args.steps / wall_s
    evals, pairs, atoms, groups, fatoms = cnt
    kernel_ms = float(np.mean(kern_ms))
    log(f'timed: {args.steps} passes x {steps_per_pass} MC steps in {wall_s:.3f} s '
        f'(HIP events: run {np.mean(run_ms):.1f} ms, kernel {kernel_ms:.1f} ms per pass)')

    tflops = FLOP_PER_PAIR * pairs / (kernel_ms / 1e3) / 1e12
    alg_bytes = BYTES_PER_ITEM * (atoms + groups + fatoms)
    gbs = alg_bytes / (kernel_ms / 1e3) / 1e9
    traffic, traffic_src = measured_traffic(M, R, tuning)
    if args.kernel_only:
        eng.close()
        print(json.dumps({'value': round(value, 1), 'unit': 'MC steps/s', 'kernel_ms': kern_ms,
                          'run_ms': run_ms, 'tflops': tflops, 'algorithmic_GBps': gbs,
                          'algorithmic_bytes': alg_bytes, 'launch': launch}), flush=True)
        return

    # the same passes launched from a captured hipGraph (first launch captures)
    graph_rate = None
    if args.graph_passes > 0:
        one_pass('graph')
        gms = [one_pass('graph') for _ in range(args.graph_passes)]
        graph_rate = steps_per_pass / (np.mean(gms) / 1e3)
        same = [eng.stats(r)['energy'] for r in range(R)] == energies0
        log(f'graph: {args.graph_passes} passes, {np.mean(gms):.1f} ms each; same chains: {same}')
        if not same:
            raise RuntimeError('graph passes ended in another state than the batched passes')

    # plain vs batched on the same first moves
    P = min(args.plain_moves, M)
    eng.restore()
    ms_plain = eng.run(P, 'plain')
    plain_e = [eng.stats(r)['energy'] for r in range(R)]
    eng.restore()
    ms_b = eng.run(P, 'batched')
    same_pb = [eng.stats(r)['energy'] for r in range(R)] == plain_e
    eng.close()
    plain_rate = R * P / (ms_plain / 1e3)
    batched_same = R * P / (ms_b / 1e3)
    log(f'plain: {R * P} steps in {ms_plain:.1f} ms; batched same moves {ms_b:.1f} ms; same chains: {same_pb}')

    # CPU baseline + parity: the FP64 oracle on the first moves of every system
    nth = host_threads()
    Lm = min(args.lockstep_moves, M)
    t1 = time.time()
    traces, gsys = gpu_traces(systems, Lm, tuning)
    lk = cpu_lockstep(systems, traces, gsys, Lm, nth)
    del traces, gsys
    log(f'lock-step: {lk} (with GPU traces {time.time() - t1:.1f} s)')
    c1_rate, c1_done, c1_dt, c1_pos = cpu_round_robin(systems, args.cpu_seconds, 1, False)
    log(f'cpu 1 thread, all groups: {c1_done} steps in {c1_dt:.1f} s')
    cc_rate, cc_done, cc_dt, cc_pos = cpu_round_robin(systems, args.cpu_seconds, nth, True)
    log(f'cpu {nth} threads, cell list: {cc_done} steps in {cc_dt:.1f} s')

    adj = adjoint_bench(args.adjoint_poses) if args.adjoint_poses > 0 else None
    if adj:
        log(f"adjoint: {adj['poses']} poses fwd+bwd in {adj['ms']:.2f} ms")
    ew = ewald_bench(args.ewald_moves) if args.ewald_moves > 0 else None
    if ew:
        log(f'ewald slab: {ew}')
    setup = setup_bench() if args.setup_bench else None
    if setup:
        log(f"set-up from files: {setup}")

    cpu_rate = lk['moves'] / lk['seconds']
    model = cpu_model()
    out = {
        'metric': 'MC steps/sec + trial dE evals/sec, batched vs plain, at |dE| err <= 1e-4 kcal/mol vs FP64',
        'value': round(value, 1),
        'unit': 'MC steps/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': wall_s * 1e3 / args.steps,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'fp64 (pair terms from fp32 positions, 1/r an fp32 estimate refined by one fp64 Newton step; '
                 'fp64 sums and acceptance)' if not args.fp32_coulomb else 'fp32 pair terms, compensated sums; '
                                                                          'fp64 sums and acceptance',
        'data': 'synthetic (seeded)',
        'config': {'workload': f'{R} uneven GCMC replicas (5k-60k fixed atoms, 8 fragments + water), '
                               f'{M} moves each per step; one step = one pass from the initial state',
                   'replicas': R, 'fixed_atoms_total': int(sizes.sum()), 'moves_per_replica': M,
                   'mc_steps_per_step': steps_per_pass, 'parallelism': 'replicas in one persistent kernel'},
        'trial_evals_per_s': round(evals * args.steps / wall_s, 1),
        'pairs_per_s': pairs * args.steps / wall_s,
        'pass_ms_hip_events': [round(x, 2) for x in run_ms],
        'kernel_ms_hip_events': [round(x, 2) for x in kern_ms],
        'graph_steps_per_s': None if graph_rate is None else round(graph_rate, 1),
        'plain_steps_per_s': round(plain_rate, 1),
        'batched_steps_per_s_same_moves': round(batched_same, 1),
        'batched_over_plain': round(batched_same / plain_rate, 2),
        'plain_sample': f'first {P} moves of every system; plain and batched chains identical: {same_pb}',
        'max_abs_dE_kJ': lk['max_abs_dE_kJ'],
        'max_abs_dE_kcal': lk['max_abs_dE_kJ'] / 4.184,
        'dE_band': '|E| <= 200 kJ/mol (Boltzmann weight > e^-80); above it max relative error '
                   f"{lk['rel_err_above_200']:.1e}",
        'parity_sample': f'first {Lm} moves of all {R} systems in one batched launch, lock-step with the FP64 '
                         f"oracle: {lk['groups']} selection groups, {lk['group_mismatches']} mismatches, "
                         f"{lk['energies_over_bar']} energies over the bar, poses / inclusion sets / final "
                         f"states bit-exact: {lk['bit_exact_poses_sets_states']}",
        'decision_agreement': 1.0 - lk['group_mismatches'] / max(1, lk['groups']),
        'launch': launch,
        'tuning': tuning,
        'autotune_ms': None if tune_table is None else [[t, round(ms, 3)] for t, ms in tune_table],
        'adjoint': adj,
        'ewald_slab': ew,
        'setup_from_files': setup,
        'roofline': {'bound': 'hbm', 'kernel': 'k_batched',
                     'achieved': round(gbs, 1), 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                     'frac': round(gbs / HBM_PEAK_GBS, 4), 'traffic': traffic,
                     'traffic_source': traffic_src,
                     'algorithmic_bytes_per_launch': alg_bytes,
                     'bytes_rule': '20 B per included environment atom, included group and fragment atom, '
                                   'per evaluation',
                     'hbm_GBps_measured': None if traffic is None else round(traffic / (kernel_ms / 1e3) / 1e9, 1),
                     'kernel_ms': round(kernel_ms, 2),
                     'limiter': 'latency of L1 misses served by L2 in each CU vector-memory pipeline '
                                '(rocprofv3 counters), not HBM bandwidth or ALU throughput',
                     'flops': ({'pair_arithmetic': 'fp64', 'achieved_TFLOPs': round(tflops, 3),
                                'peak_TFLOPs': FP64_PEAK_TFLOPS, 'frac': round(tflops / FP64_PEAK_TFLOPS, 4),
                                'flop_rule': '32 FLOP per fragment-environment atom pair'}
                               if not args.fp32_coulomb else
                               {'pair_arithmetic': 'fp32', 'achieved_TFLOPs': round(tflops, 3),
                                'peak_TFLOPs': FP32_PEAK_TFLOPS, 'frac': round(tflops / FP32_PEAK_TFLOPS, 4),
                                'ceiling_without_packed_fp32_TFLOPs': FP32_UNPACKED_TFLOPS,
                                'flop_rule': '32 FLOP per fragment-environment atom pair'}),
                     'pairs': int(pairs), 'evals': int(evals)},
        'cpu_baseline': {'value': round(cpu_rate, 2), 'unit': 'MC steps/s', 'cores': nth, 'kind': 'port',
                         'cpu': model,
                         'sample': f"FP64 oracle (reference algorithm: every group tested per trial), first {Lm} "
                                   f"moves of all {R} systems = {lk['moves']} MC steps, systems spread over {nth} "
                                   f"threads, {lk['seconds']:.1f} s (the lock-step parity run)"},
        'cpu_baseline_1core': {'value': round(c1_rate, 2), 'unit': 'MC steps/s', 'cores': 1, 'kind': 'port',
                               'cpu': model,
                               'sample': f'same oracle, one thread, chunks of 5 moves round-robin over the '
                                         f'systems from move 0: {c1_done} MC steps in {c1_dt:.1f} s'},
        'cpu_baseline_cells': {'value': round(cc_rate, 2), 'unit': 'MC steps/s', 'cores': nth, 'kind': 'port',
                               'cpu': model,
                               'sample': f'oracle with a cell list of group centres (same inclusion sets), '
                                         f'{nth} threads over the systems, chunks of 5 moves round-robin from '
                                         f'move 0 (up to move {cc_pos}): {cc_done} MC steps in {cc_dt:.1f} s'},
    }
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
