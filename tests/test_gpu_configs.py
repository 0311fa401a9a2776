"""Each benchmark configuration at its real workload on the GPU, against the
FP64 oracle, with the error printed per configuration.

* water sphere: 2,000 fixed atoms, 40 A box, TIP3P conf 1, all 1,000 moves,
  plain and batched forms;
* multi-fragment: 20,000 fixed atoms, 80 A box, 8 fragments + water at the
  SILCS confBias (400 / 15 / 10 / 20 / 15 / 300 / 100 / 200 / 2), the first
  300 moves;
* uneven replicas: all 64 systems (5k-60k fixed atoms) in one persistent
  launch, the first 60 moves of each, every system lock-stepped;
* adjoint: 4,096 poses (512 per fragment type, uniform centres, uniform
  random orientations) against the 20,000-atom protein: energies and
  gradients against the FP64 oracle, central differences on a sample;
* slab: 300,000 fixed atoms, 150 x 150 x 120 A, PME (1 A grid, beta 0.34,
  12 A cutoff): a 64-move window in the graph form (equal to the plain
  form's chain), then trial energies against an FP64 restatement of the same
  PME discretisation plus FP64 real-space terms.

Bars as in tests/lockstep.py (energies 4.184e-4 kJ/mol for |E| <= 200
kJ/mol, 1e-5 relative above; decisions >= 99.99 %; poses, inclusion sets,
states bit-exact) and 1e-4 relative for adjoint gradients.
"""

import threading

import numpy as np
import pytest

from oracle import ewald
from oracle import oracle as O
from pygcmc_amd import synth
from pygcmc_amd.engine import Engine

from . import lockstep as LS
from . import systems

pytestmark = pytest.mark.gpu


def _report(name, summ):
    print(f"{name}: max |dE| {summ['max_err_band']:.2e} kJ/mol (|E| <= 200), rel {summ['rel_err_above']:.1e} "
          f"above; {summ['energies']} energies, {summ['groups']} groups, {summ['mismatches']} mismatches, "
          f"{summ['accepted']} accepted", flush=True)


@pytest.mark.parametrize('mode', ['plain', 'batched'])
def test_water_sphere_config(mode):
    s = synth.water_sphere(mcsteps=1000)
    rows = LS.lockstep([s], 1000, mode)
    _report(f'water sphere, 1000 moves, {mode}', LS.summary(rows))
    assert rows[0]['accepted'] > 0


def test_multi_fragment_config():
    s = synth.multi_fragment(mcsteps=300)
    assert int(s['frag'][0]['startRes']) > 0 and len(s['atoms']) > 20000
    rows = LS.lockstep([s], 300, 'batched')
    summ = LS.summary(rows)
    _report('multi-fragment 20k / 80 A, 300 moves', summ)
    assert summ['groups'] > 1000


def test_all_64_replicas_one_launch():
    sizes = synth.replica_sizes(64)
    sy = [synth.replica_system(r, int(n), mcsteps=60) for r, n in enumerate(sizes)]
    rows = LS.lockstep(sy, 60, 'batched', nthreads=16)
    summ = LS.summary(rows)
    _report('64 replicas, 60 moves each, one launch', summ)
    assert summ['groups'] > 10000


def test_fp32_coulomb_mode_error_per_config():
    """The all-FP32 pair mode (gcmc_tuning.fp32_coulomb = 1): its trial-energy
    error against FP64 on each configuration, printed; chains follow the
    GPU's decisions so every trial is compared."""
    sizes = synth.replica_sizes(64)
    cases = [('water sphere', [synth.water_sphere(mcsteps=300)], 300),
             ('multi-fragment 20k', [synth.multi_fragment(mcsteps=100)], 100),
             ('replicas 0-7', [synth.replica_system(r, int(sizes[r]), mcsteps=40) for r in range(8)], 40)]
    for name, sy, n in cases:
        trs, gs = LS.gpu_run(sy, n, 'batched', fp32_coulomb=1)
        ots, refs = LS.oracle_follow(sy, trs, n)
        worst, rel = 0.0, 0.0
        for gt, ot in zip(trs, ots):
            km = np.arange(gt['centre'].shape[1])[None, :] < gt['k'][:, None]
            eo = ot['e'][km]
            err = np.abs(gt['e'][km] - eo)
            band = np.abs(eo) <= 200.0
            worst = max(worst, float(err[band].max()) if band.any() else 0.0)
            rel = max(rel, float((err[~band] / np.abs(eo[~band])).max()) if (~band).any() else 0.0)
        print(f'fp32 Coulomb, {name}: max |dE| {worst:.2e} kJ/mol (|E| <= 200), rel {rel:.1e} above', flush=True)
        assert worst < 0.05 and rel < 1e-4


def _threaded(fn, items, nthreads=8):
    out = [None] * len(items)
    nxt = [0]
    lock = threading.Lock()

    def worker():
        while True:
            with lock:
                i = nxt[0]
                nxt[0] += 1
            if i >= len(items):
                return
            out[i] = fn(items[i])

    ths = [threading.Thread(target=worker) for _ in range(nthreads)]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    return out


def test_adjoint_config_4096_poses():
    torch = pytest.importorskip('torch')
    from pygcmc_amd import adjoint
    s = synth.multi_fragment(mcsteps=10)
    eng = Engine()
    eng.add_system(systems.copy(s))
    eng.upload()
    eps, rmin = adjoint.type_parameters(s)
    e64, r64 = eps.double().cpu().numpy(), rmin.double().cpu().numpy()
    rng = np.random.default_rng(4096)
    Y = len(e64)
    worst_e, worst_g, worst_t, npose = 0.0, 0.0, 0.0, 0
    fd_cases = []
    try:
        for spi in range(8):
            c, p = systems.random_poses(s, spi, 512, seed=4096 + spi)
            ct = torch.tensor(c, device='cuda')
            pt = torch.tensor(p, device='cuda', requires_grad=True)
            e = eps.clone().requires_grad_(True)
            r = rmin.clone().requires_grad_(True)
            q = torch.zeros_like(eps).requires_grad_(True)
            w = rng.normal(size=len(c))
            E = adjoint.pose_energy(eng, 0, spi, ct, pt, e, r, q)
            (E * torch.tensor(w, device='cuda')).sum().backward()
            Eg = E.detach().cpu().numpy()
            gx = pt.grad.cpu().numpy() / w[:, None, None]
            refs = _threaded(lambda i: O.adjoint(s, spi, c[i], p[i], e64, r64), list(range(len(c))))
            e_ref = np.array([x[0] for x in refs])
            err = np.abs(Eg - e_ref)
            assert (err <= LS.e_tol(e_ref)).all(), (spi, float(err.max()))
            band = np.abs(e_ref) <= 200.0
            if band.any():
                worst_e = max(worst_e, float(err[band].max()))
            for i in range(len(c)):
                gref = refs[i][1]
                rel = np.linalg.norm(gx[i] - gref) / max(np.linalg.norm(gref), 1e-3)
                worst_g = max(worst_g, float(rel))
            tref = np.zeros((3, Y))
            for i in range(len(c)):
                tref += w[i] * np.stack([refs[i][2], refs[i][3], refs[i][4]])
            got = np.stack([q.grad.cpu().numpy(), e.grad.cpu().numpy(), r.grad.cpu().numpy()])
            for k in range(3):
                worst_t = max(worst_t, float(np.linalg.norm(got[k] - tref[k]) / max(np.linalg.norm(tref[k]), 1e-3)))
            npose += len(c)
            # central-difference sample: the two poses of moderate energy with the lowest index
            mod = np.nonzero(np.abs(e_ref) <= 200.0)[0][:2]
            fd_cases += [(spi, c[i], p[i], gx[i]) for i in mod]
    finally:
        eng.close()
    print(f'adjoint, {npose} poses x 20k atoms: max |dE| {worst_e:.2e} kJ/mol (|E| <= 200); pose gradients '
          f'max relative error {worst_g:.2e}; type gradients {worst_t:.2e} (vs FP64)', flush=True)
    assert npose == 4096
    assert worst_g <= 1e-4 and worst_t <= 1e-4
    worst_fd = 0.0
    for spi, c, p, g in fd_cases:
        fd = systems.fd_pose_gradient(s, spi, c, p, e64, r64)
        worst_fd = max(worst_fd, float(np.linalg.norm(g - fd) / np.linalg.norm(fd)))
    print(f'adjoint vs central differences (h = 1e-4), {len(fd_cases)} poses: max relative error {worst_fd:.2e}')
    assert len(fd_cases) >= 8 and worst_fd <= 1e-4


def test_slab_config_ewald():
    beta, spacing = 0.34, 1.0
    s = synth.slab(mcsteps=50000)
    # the plain form, then the graph form (two captured 32-move chunks) from
    # the same start: the same chain
    p = systems.copy(s)
    eng = Engine()
    eng.add_system(p)
    eng.set_ewald(0, beta, spacing)
    eng.upload()
    eng.run(64, 'plain')
    eng.download(0)
    eng.close()
    g = systems.copy(s)
    eng = Engine()
    eng.add_system(g)
    eng.set_ewald(0, beta, spacing)
    eng.upload()
    try:
        eng.run(64, 'graph')
        eng.download(0)
        assert np.array_equal(g['atoms'].view(np.uint8), p['atoms'].view(np.uint8))
        st = eng.stats(0)
        L = g['info'][0]['cryst'].astype(np.float64)
        K = [max(4, int(round(float(x) / spacing))) for x in g['info'][0]['cryst']]
        xyz, q = ewald.env_charges(g)
        env = ewald.PMEEnv(xyz, q, L, beta, K)
        worst, rel, n = 0.0, 0.0, 0
        for spi in range(int(g['info'][0]['fragTypeNum'])):
            c, p = systems.random_poses(g, spi, 8, seed=50 + spi)
            E, _, _ = eng.eval_poses(0, spi, c, p)
            ref = np.array(_threaded(lambda i: ewald.pose_energy_ewald(g, spi, c[i], p[i], beta, env),
                                     list(range(len(c)))))
            err = np.abs(E - ref)
            band = np.abs(ref) <= 200.0
            if band.any():
                worst = max(worst, float(err[band].max()))
            if (~band).any():
                rel = max(rel, float((err[~band] / np.abs(ref[~band])).max()))
            n += len(c)
    finally:
        eng.close()
    print(f'slab (300k atoms, PME 1 A): {int(g["frag"]["totalNum"].sum())} molecules after 64 moves, '
          f'{int(st["accepts"].sum())} accepted; {n} trial energies vs FP64 PME + real space: max |dE| '
          f'{worst:.2e} kJ/mol (|E| <= 200), rel {rel:.1e} above', flush=True)
    assert st['accepts'].sum() > 0
    assert worst <= 0.01 and rel <= 1e-4
