"""Ewald mode of the engine (real-space erfc pair terms + PME reciprocal
potential) against FP64 Ewald sums, the exactness of its incremental PME
updates, and its batched and graph forms against the plain form.

Bars: trial energies within the PME discretisation error of order-4
B-splines at beta = 0.34 /A -- 0.15 kJ/mol at 1 A grid spacing, 0.01 kJ/mol
at 0.5 A (measured 0.071 and 0.0028: h^4 convergence); the charge grid after a run equals a fresh spread of the final
state bit for bit (fixed-point accumulation), and runs are reproducible.
The reference has no Ewald path (its PME flag is never read): parity is
against the mathematics, not the reference.
"""

import functools

import numpy as np
import pytest

from oracle import ewald
from oracle import oracle as O
from pygcmc_amd.engine import Engine

from . import systems

pytestmark = pytest.mark.gpu

BETA = 0.34


@functools.lru_cache(maxsize=None)
def _state():
    s = systems.multi(nsteps=400, n_fixed=1500, box=32.0, seed=77, conf_scale=0.2)
    O.run(s, first=0, nsteps=250)           # populate with some molecules (plain chain)
    # runs from this state replay the schedule from move 0: another seed, or
    # an insertion trial would land exactly on the molecule the same
    # (seed, move, trial) counter inserted above (r = 0: an infinite energy)
    s['info'][0]['seed'] = np.uint32(int(s['info'][0]['seed']) ^ 0x5bd1e995)
    return s


def _kmax(L):
    return tuple(int(np.ceil(4.8 * l * BETA / np.pi)) for l in L)


def _engine(s, spacing):
    eng = Engine()
    eng.add_system(systems.copy(s))
    eng.set_ewald(0, BETA, spacing)
    eng.upload()
    return eng


@pytest.mark.parametrize('spacing,bar', [(1.0, 0.15), (0.5, 0.01)])
def test_ewald_insertion_energies(spacing, bar):
    s = _state()
    L = s['info'][0]['cryst'].astype(np.float64)
    xyz, q = ewald.env_charges(s)
    env = ewald.EwaldEnv(xyz, q, L, BETA, _kmax(L))
    eng = _engine(s, spacing)
    worst = 0.0
    for spi in (0, 6, 8):
        c, p = systems.moderate_poses(s, spi, 12, seed=5 + spi)
        E, _, _ = eng.eval_poses(0, spi, c, p)
        ref = np.array([ewald.pose_energy_ewald(s, spi, c[i], p[i], BETA, env) for i in range(len(c))])
        worst = max(worst, float(np.abs(E - ref).max()))
    eng.close()
    print(f'spacing {spacing}: max |dE| {worst:.2e} kJ/mol')
    assert worst <= bar


def test_ewald_moving_molecule_energies():
    """Old pose (the molecule's own share removed from the grid) and a
    displaced pose of present molecules, as translate/rotate/delete use."""
    s = _state()
    L = s['info'][0]['cryst'].astype(np.float64)
    eng = _engine(s, 0.5)
    rng = np.random.default_rng(9)
    worst = 0.0
    checked = 0
    for spi in range(int(s['info'][0]['fragTypeNum'])):
        F = s['frag'][spi]
        nf = int(F['num_atoms'])
        for j in range(min(2, int(F['totalNum']))):
            gid = int(F['startRes']) + j
            a0 = int(s['res'][gid]['atomStart'])
            pose = s['atoms']['position'][a0:a0 + nf].copy()
            centre = s['res'][gid]['position'].copy()
            shift = rng.uniform(0.2, 1.0, 3).astype(np.float32)
            poses = np.stack([pose, pose + shift])
            centres = np.stack([centre, centre + shift])
            E, _, _ = eng.eval_poses(0, spi, centres, poses, exclude=np.array([gid, gid]))
            xyz, q = ewald.env_charges(s, exclude=gid)
            env = ewald.EwaldEnv(xyz, q, L, BETA, _kmax(L))
            ref = [ewald.pose_energy_ewald(s, spi, centres[i], poses[i], BETA, env, exclude=gid) for i in range(2)]
            worst = max(worst, float(np.abs(E - ref).max()))
            checked += 1
    eng.close()
    print(f'moving molecules ({checked}): max |dE| {worst:.2e} kJ/mol')
    assert checked >= 3
    assert worst <= 0.01


def test_ewald_run_reproducible_and_grid_exact():
    s = _state()
    outs = []
    pts = np.random.default_rng(1).random((64, 3)).astype(np.float32) * s['info'][0]['cryst']
    for _ in range(2):
        g = systems.copy(s)
        eng = Engine()
        eng.add_system(g)
        eng.set_ewald(0, BETA, 1.0)
        eng.upload()
        eng.run(120, 'plain')
        phi = eng.ewald_potential(0, pts)
        eng.download(0)
        st = eng.stats(0)
        eng.close()
        outs.append((g, phi, st))
    (g0, phi0, st0), (g1, phi1, st1) = outs
    assert np.array_equal(g0['atoms'].view(np.uint8), g1['atoms'].view(np.uint8))
    assert np.array_equal(phi0.view(np.uint32), phi1.view(np.uint32))
    assert st0['accepts'].sum() > 0
    # the incrementally updated grid equals a fresh spread of the final state
    fresh = Engine()
    fresh.add_system(systems.copy(g0))
    fresh.set_ewald(0, BETA, 1.0)
    fresh.upload()
    phi_f = fresh.ewald_potential(0, pts)
    fresh.close()
    assert np.array_equal(phi0.view(np.uint32), phi_f.view(np.uint32))


def _ewald_run(systems_, mode, chunks, spacing=1.0, trace=0):
    eng = Engine()
    gs = [systems.copy(x) for x in systems_]
    for r, g in enumerate(gs):
        eng.add_system(g)
        eng.set_ewald(r, BETA, spacing)
    eng.upload()
    if trace:
        for r, g in enumerate(gs):
            eng.set_trace(r, trace, int(g['frag']['confBias'].max()))
    for n in chunks:
        eng.run(n, mode)
    pts = np.random.default_rng(3).random((64, 3)).astype(np.float32) * gs[0]['info'][0]['cryst']
    phi = [eng.ewald_potential(r, pts) for r in range(len(gs))]
    trs = [eng.get_trace(r) for r in range(len(gs))] if trace else None
    for r in range(len(gs)):
        eng.download(r)
    st = [eng.stats(r) for r in range(len(gs))]
    eng.close()
    return gs, phi, trs, st


def test_ewald_batched_and_graph_forms_equal_plain():
    """Ewald systems in the batched form (move by move, each move's trials in
    one launch, PME update between moves) and the graph form (a captured
    chunk of moves replayed; 100 moves = three 32-move replays + 4 eager
    moves, then 30 more eager): the same trial energies, accept flags, final
    states and reciprocal potentials as the plain form, bit for bit, two
    systems at once."""
    a = _state()
    b = systems.multi(nsteps=400, n_fixed=1200, box=30.0, seed=78, conf_scale=0.2)
    outs = {}
    for mode, chunks in (('plain', (100, 30)), ('batched', (100, 30)), ('graph', (100, 30))):
        outs[mode] = _ewald_run([a, b], mode, chunks, trace=130)
    gp, phip, trp, stp = outs['plain']
    assert sum(int(x['accepts'].sum()) for x in stp) > 0
    for mode in ('batched', 'graph'):
        g, phi, tr, st = outs[mode]
        for r in range(2):
            assert np.array_equal(tr[r]['e'].view(np.uint64), trp[r]['e'].view(np.uint64)), (mode, r)
            assert np.array_equal(tr[r]['gacc'], trp[r]['gacc']), (mode, r)
            assert np.array_equal(g[r]['atoms'].view(np.uint8), gp[r]['atoms'].view(np.uint8)), (mode, r)
            assert np.array_equal(phi[r].view(np.uint32), phip[r].view(np.uint32)), (mode, r)
            assert st[r]['energy'] == stp[r]['energy']


def test_ewald_and_plain_systems_do_not_mix_in_batched_form():
    s = _state()
    eng = Engine()
    eng.add_system(systems.copy(s))
    eng.add_system(systems.copy(s))
    eng.set_ewald(0, BETA, 1.0)
    eng.upload()
    with pytest.raises(RuntimeError, match='all systems with Ewald'):
        eng.run(5, 'batched')
    eng.run(5, 'plain')
    eng.close()
