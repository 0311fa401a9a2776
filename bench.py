"""Benchmark: GCMC MC steps/s and trial energy evaluations/s on one GPU.

Workload (default): 64 independent GCMC systems whose fixed-atom counts are
log-uniform in 5,000-60,000 (box 40*(N/2000)^(1/3) A, fixed-atom density of the
2,000-atom sphere system), each with its own seed and muex and the 8 rigid
fragment types + TIP3P water (confBias 400/15/10/20/15/300/100/200/2), 10,000
moves each, FP32 pair energies with compensated sums and FP64 totals and
acceptance, all systems in one persistent kernel.  Synthetic data (seeded).

Prints one JSON line.  `value` is the batched form's aggregate MC steps/s over
all systems with inputs resident in HBM; the plain form (one move of one
system per launch pair) is timed on the same moves beside it.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3   # MI355X FP32 vector (packed) peak
HBM_PEAK_GBS = 8000.0
FLOP_PER_PAIR = 32         # algorithmic count per fragment-environment pair
BYTES_PER_ITEM = 20        # per included env atom / included group / fragment atom


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_replicas(n, steps):
    from pygcmc_amd import synth
    sizes = synth.replica_sizes(n)
    out = []
    for r, na in enumerate(sizes):
        out.append(synth.replica_system(r, int(na), mcsteps=steps))
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


def cpu_baseline(states, first, budget_s, nthreads=1):
    """The FP64 oracle (reference algorithm: all groups scanned per trial;
    nthreads OpenMP threads over the trials of a move) continuing the same
    chains from step `first`."""
    from oracle import oracle as O
    done = 0
    t0 = time.perf_counter()
    for s in states:
        step = first
        while time.perf_counter() - t0 < budget_s and step < len(s['moves']):
            O.run(s, first=step, nsteps=1, nthreads=nthreads)
            step += 1
            done += 1
        if time.perf_counter() - t0 >= budget_s:
            break
    dt = time.perf_counter() - t0
    return done / dt, done, dt


def parity_sample(system, nsteps):
    """Max |dE| of the GPU against the FP64 oracle on one system (lock-step)."""
    from oracle import oracle as O
    from pygcmc_amd.engine import Engine
    g = copies([system])[0]
    k = int(g['frag']['confBias'].max())
    eng = Engine()
    eng.add_system(g)
    eng.upload()
    eng.set_trace(0, nsteps, k)
    eng.run(nsteps, 'batched')
    tr = eng.get_trace(0)
    eng.close()
    ref = copies([system])[0]
    ot = O.new_trace(nsteps, k)
    O.run(ref, first=0, nsteps=nsteps, trace=ot, follow=tr)
    m = np.arange(k)[None, :] < tr['k'][:, None]
    e_ref = ot['e'][m]
    err = np.abs(tr['e'][m] - e_ref)
    band = np.abs(e_ref) <= 200.0          # the band the 1e-4 kcal/mol bar applies to
    gm = np.arange(k)[None, :] < tr['ngroups'][:, None]
    agree = float((tr['gacc'][gm] == ot['gacc'][gm]).mean()) if gm.any() else 1.0
    rel_hi = float((err[~band] / np.abs(e_ref[~band])).max()) if (~band).any() else 0.0
    return float(err[band].max()) if band.any() else 0.0, agree, int(gm.sum()), rel_hi


def measured_traffic(steps, replicas, tuning):
    """HBM bytes of the timed k_batched launch from the committed rocprofv3
    --pmc record (FETCH_SIZE and WRITE_SIZE scaled by a copy kernel of known
    size), when that record was measured on this same workload; else None."""
    path = os.path.join(ROOT, 'profiles', 'k_batched_64replicas.pmc.txt')
    if not os.path.exists(path):
        return None, None
    rec = {}
    for line in open(path):
        if '=' in line and not line.startswith('#'):
            k, v = line.strip().split('=', 1)
            rec[k] = v
    same = (int(rec.get('steps', -1)) == steps and int(rec.get('replicas', -1)) == replicas and
            float(rec.get('cell_size', -1)) == float(tuning.get('cell_size', 0)) and
            int(rec.get('wg_size', -1)) == int(tuning.get('wg_size', 0) or 256) and
            int(rec.get('pair_lds', 1)) == int(tuning.get('pair_lds', 1)))
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=9800)
    ap.add_argument('--warmup', type=int, default=200)
    ap.add_argument('--replicas', type=int, default=64)
    ap.add_argument('--plain-steps', type=int, default=100)
    ap.add_argument('--cpu-seconds', type=float, default=15.0)
    ap.add_argument('--parity-steps', type=int, default=300)
    ap.add_argument('--mode', default='graph', choices=['batched', 'graph'])
    # defaults: the setting pygcmc_amd.autotune picks on this workload (its
    # timings differ by under 2% across cell sizes 5-7.5 A); --autotune 1
    # searches again
    ap.add_argument('--cell-size', type=float, default=7.5)
    ap.add_argument('--wgs', type=int, default=0)
    ap.add_argument('--wg-size', type=int, default=256)
    ap.add_argument('--pair-lds', type=int, default=0)
    ap.add_argument('--kernel-only', action='store_true',
                    help='only the timed batched run (for profiler runs): no graph, plain, parity, '
                         'adjoint or CPU legs')
    ap.add_argument('--adjoint-poses', type=int, default=4096)
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

    total = args.warmup + args.steps
    t0 = time.time()
    systems, sizes = build_replicas(args.replicas, total)
    R = len(systems)
    log(f'built {R} systems ({int(sizes.sum())} fixed atoms) in {time.time() - t0:.1f}s')
    tuning = dict(cell_size=args.cell_size, persistent_wgs=args.wgs, wg_size=args.wg_size, pair_lds=args.pair_lds)
    tune_table = None
    if args.autotune:
        from pygcmc_amd.autotune import autotune
        t1 = time.time()
        tuning, tune_table = autotune(systems, log=log)
        log(f'autotune: {tuning} in {time.time() - t1:.1f}s')

    # batched form over the whole schedule
    eng = make_engine(copies(systems), tuning)
    if args.warmup:
        eng.run(args.warmup, 'batched')
    states_w = [eng.download(r, copies([systems[r]])[0]) for r in range(R)]
    c0 = counters_sum(eng, R)
    ms_total = eng.run(args.steps, 'batched')
    ms_kernel = eng.last_kernel_ms()
    c1 = counters_sum(eng, R)
    launch = eng.launch_info()
    eng.close()
    steps_done = R * args.steps
    value = steps_done / (ms_total / 1e3)
    evals, pairs, atoms, groups, fatoms = c1 - c0
    log(f'batched: {steps_done} steps in {ms_total:.1f} ms (kernel {ms_kernel:.1f} ms)')

    tflops = FLOP_PER_PAIR * pairs / (ms_kernel / 1e3) / 1e12
    gbs = BYTES_PER_ITEM * (atoms + groups + fatoms) / (ms_kernel / 1e3) / 1e9
    if args.kernel_only:
        print(json.dumps({'value': round(value, 1), 'unit': 'MC steps/s', 'kernel_ms': ms_kernel,
                          'tflops': tflops, 'algorithmic_GBps': gbs, 'launch': launch}), flush=True)
        return

    # graph replay of the same chains (timing only, from the warm state)
    graph_rate = None
    if args.mode == 'graph':
        eg = make_engine(copies(states_w), tuning)
        gsteps = min(args.steps, 2000)
        ms_g = 0.0
        done = 0
        while done < gsteps:
            n = min(500, gsteps - done)
            ms_g += eg.run(n, 'graph')
            done += n
        eg.close()
        graph_rate = R * gsteps / (ms_g / 1e3)
        log(f'graph: {R * gsteps} steps in {ms_g:.1f} ms')

    # plain vs batched on the same first moves
    ps = args.plain_steps
    ep = make_engine(copies(systems), tuning)
    ms_plain = ep.run(ps, 'plain')
    ep.close()
    eb = make_engine(copies(systems), tuning)
    ms_b = eb.run(ps, 'batched')
    eb.close()
    plain_rate = R * ps / (ms_plain / 1e3)
    batched_same = R * ps / (ms_b / 1e3)
    log(f'plain: {R * ps} steps in {ms_plain:.1f} ms; batched same moves {ms_b:.1f} ms')

    # accuracy on one system against the FP64 oracle
    max_err, agree, ngroups, rel_hi = parity_sample(systems[0], args.parity_steps)
    log(f'parity: max |dE| {max_err:.2e} kJ/mol (|E| <= 200), rel {rel_hi:.1e} above; decisions agree '
        f'{agree:.6f} over {ngroups} groups')

    # CPU baseline: the oracle continuing the warm chains
    cpu_rate, cpu_done, cpu_dt = cpu_baseline(copies(states_w[:4]), args.warmup, args.cpu_seconds)
    log(f'cpu: {cpu_done} steps in {cpu_dt:.1f}s')
    ncpu = min(os.cpu_count() or 1, 16)
    cpu_rate_n, cpu_done_n, cpu_dt_n = cpu_baseline(copies(states_w[:4]), args.warmup, args.cpu_seconds / 3, ncpu)
    log(f'cpu ({ncpu} threads): {cpu_done_n} steps in {cpu_dt_n:.1f}s')

    traffic, traffic_src = measured_traffic(args.steps, R, tuning)
    adj = adjoint_bench(args.adjoint_poses) if args.adjoint_poses > 0 else None
    setup = setup_bench() if args.setup_bench else None
    if setup:
        log(f"set-up from files: {setup}")
    if adj:
        log(f"adjoint: {adj['poses']} poses fwd+bwd in {adj['ms']:.2f} ms")

    out = {
        'metric': 'MC steps/sec + trial dE evals/sec, batched vs plain, at |dE| err <= 1e-4 kcal/mol vs FP64',
        'value': round(value, 1),
        'unit': 'MC steps/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': ms_total / args.steps,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'fp32 pairs + fp64 totals',
        'data': 'synthetic (seeded)',
        'config': {'workload': f'{R} uneven GCMC replicas (5k-60k fixed atoms, 8 fragments + water, '
                               f'{total} moves each)', 'replicas': R,
                   'fixed_atoms_total': int(sizes.sum()), 'moves_per_replica': total,
                   'parallelism': 'replicas in one persistent kernel'},
        'trial_evals_per_s': round(evals / (ms_total / 1e3), 1),
        'pairs_per_s': pairs / (ms_total / 1e3),
        'graph_steps_per_s': None if graph_rate is None else round(graph_rate, 1),
        'plain_steps_per_s': round(plain_rate, 1),
        'batched_steps_per_s_same_moves': round(batched_same, 1),
        'batched_over_plain': round(batched_same / plain_rate, 2),
        'max_abs_dE_kJ': max_err,
        'max_abs_dE_kcal': max_err / 4.184,
        'dE_band': '|E| <= 200 kJ/mol (Boltzmann weight > e^-80); above it max relative error '
                   f'{rel_hi:.1e}',
        'parity_sample': f'system 0, first {args.parity_steps} moves, lock-step with the FP64 oracle',
        'decision_agreement': agree,
        'launch': launch,
        'tuning': tuning,
        'autotune_ms': None if tune_table is None else [[t, round(ms, 3)] for t, ms in tune_table],
        'adjoint': adj,
        'setup_from_files': setup,
        'roofline': {'bound': 'mfma', 'pipe': 'fp32-valu', 'kernel': 'k_batched',
                     'achieved': round(tflops, 3), 'peak': FP32_PEAK_TFLOPS, 'unit': 'TFLOP/s',
                     'frac': round(tflops / FP32_PEAK_TFLOPS, 4), 'traffic': traffic,
                     'traffic_source': traffic_src,
                     'hbm_GBps': None if traffic is None else round(traffic / (ms_kernel / 1e3) / 1e9, 1),
                     'hbm_frac': None if traffic is None else round(traffic / (ms_kernel / 1e3) / 1e9 / HBM_PEAK_GBS, 5),
                     'algorithmic_GBps': round(gbs, 1), 'kernel_ms': round(ms_kernel, 2),
                     'pairs': int(pairs), 'evals': int(evals)},
        'cpu_baseline': {'value': round(cpu_rate, 2), 'unit': 'MC steps/s', 'cores': 1, 'kind': 'port',
                         'sample': f'FP64 oracle, {cpu_done} moves of systems 0-3 continuing from '
                                   f'move {args.warmup} ({cpu_dt:.1f} s)'},
        'cpu_baseline_threads': {'value': round(cpu_rate_n, 2), 'unit': 'MC steps/s', 'cores': ncpu,
                                 'kind': 'port', 'sample': f'same, OpenMP over the trials of a move, '
                                                           f'{cpu_done_n} moves ({cpu_dt_n:.1f} s)'},
    }
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
