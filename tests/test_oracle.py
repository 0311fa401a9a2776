"""Pins the CPU oracle (oracle/gcmc_oracle.c) before it is trusted as the
checker of the HIP path: published known-answer vectors, the reference's own
fixtures, closed forms and brute force."""

import json
import math
import os
import random

import numpy as np
import pytest

from oracle import oracle as O
from pygcmc_amd import dataset
from pygcmc_amd.values import BETA, COULOMB, MOLES_TO_MOLECULES

from . import systems

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')


# ---- Philox4x32-10: Random123 known-answer vectors (Salmon et al., SC'11) ----
@pytest.mark.parametrize('ctr,key,out', [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
])
def test_philox_kat(ctr, key, out):
    assert list(O.philox(ctr, key)) == out


def test_uniform_ranges():
    L = O.lib()
    assert L.orc_u01f(0) == 2.0 ** -24 and L.orc_u01f(0xffffffff) == 1.0
    assert 0.0 < L.orc_u01d(0, 0) < 1e-15 and L.orc_u01d(0xffffffff, 0xffffffff) < 1.0
    assert L.orc_below(0xffffffff, 7) == 6 and L.orc_below(0, 7) == 0


# ---- rotation: test/jupyter/test2.ipynb cells 22-23 (Rodrigues, numpy) ----
def test_rotation_notebook_known_answer():
    R = O.rotmat([29.850, 42.018, 155.751], 0.847)
    ref = np.array([[0.67341581, -0.69561841, 0.25025221],
                    [0.72709826, 0.68439014, -0.05420567],
                    [-0.13356368, 0.21846091, 0.96666208]])
    assert np.abs(R - ref).max() < 2e-7
    v = R.astype(np.float64) @ np.array([0.0, 0.0, -0.391])
    assert np.abs(v - [-0.09784862, 0.02119442, -0.37796487]).max() < 2e-7


def test_rotation_is_orthonormal_and_axis_fixed():
    rng = np.random.default_rng(3)
    for _ in range(200):
        a = rng.random(3).astype(np.float32) + 1e-3
        R = O.rotmat(a, float(rng.random() * 2 * np.pi)).astype(np.float64)
        assert np.abs(R @ R.T - np.eye(3)).max() < 1e-6
        n = a / np.linalg.norm(a)
        assert np.abs(R @ n - n).max() < 1e-6


def test_sincos_against_libm():
    xs = np.linspace(0, 2 * np.pi, 20001, dtype=np.float32)
    err = max(max(abs(O.sincosf(float(x))[0] - math.sin(float(x))),
                  abs(O.sincosf(float(x))[1] - math.cos(float(x)))) for x in xs[::7])
    assert err < 3e-7


# ---- FP32 group test: literal reference semantics -----------------------------
def test_distance_fast_round_semantics():
    L = [40.0, 40.0, 40.0]
    # half-box separations round away from zero: the image shift flips the sign
    assert O.distanceP32([20.0, 0, 0], [0, 0, 0], L) == pytest.approx(20.0)
    assert O.distanceP32([21.0, 0, 0], [0, 0, 0], L) == pytest.approx(19.0)
    assert O.distanceP32([-21.0, 0, 0], [0, 0, 0], L) == pytest.approx(19.0)
    assert O.distanceP32([3.0, 4.0, 0], [0, 0, 0], L) == 5.0


# ---- pair formula: closed forms with FP32-exact inputs --------------------------
def _two_atom_system(sigma_nm, eps_kj, q_env, q_frag, r, centre_at_atom=False):
    nb = {'A': (sigma_nm, eps_kj), 'B': (sigma_nm, eps_kj)}
    ff = dataset.ff_pair_table(['A'], ['A', 'B'], nb)
    cryst = [200.0, 200.0, 200.0]
    start, n, grid, cav = dataset.get_grid(np.array([[0.0, 0, 0]]), cryst, 16.0)
    sp = [dict(name='X', conc=1.0, muex=0.0, conf=1, mctime=1.0, template_xyz=np.zeros((1, 3)),
               charge=[q_frag], type=[0], molecules=[])]
    arrays = dataset.pack_system(
        cryst=cryst, cutoff=15.0, grid_dx=16.0, startxyz=start, grid=grid, cavity_factor=cav,
        fixed_xyz=np.array([[0.0, 0.0, 0.0]]), fixed_charge=[q_env], fixed_type=[1],
        fixed_group=np.zeros(1, np.int64), species=sp, ff=ff, ffX=1, ffY=2,
        moves=np.zeros(1, np.int32), move_counts=[1, 0, 0, 0], mcsteps=1, seed=1)
    sysd = dict(zip(('info', 'frag', 'res', 'atoms', 'grid', 'ff', 'moves'), arrays))
    pose = np.array([[r, 0.0, 0.0]], np.float32)
    centre = pose[0] if centre_at_atom else np.zeros(3, np.float32)
    return O.energy_pose(sysd, 0, centre, pose)[0]


def test_lj_closed_forms():
    # sigma 0.25 nm = 2.5 A; E(sigma) = 0, minimum -eps at 2^(1/6) sigma
    assert _two_atom_system(0.25, 1.0, 0.0, 0.0, 2.5) == 0.0
    rmin = float(np.float32(2.5 * 2 ** (1 / 6)))
    assert _two_atom_system(0.25, 1.0, 0.0, 0.0, rmin) == pytest.approx(-1.0, abs=1e-12)
    assert _two_atom_system(0.25, 1.0, 0.0, 0.0, 5.0) == pytest.approx(4 * (2.0 ** -12 - 2.0 ** -6))


def test_coulomb_closed_form():
    assert _two_atom_system(0.25, 0.0, 1.0, -1.0, 1.0) == pytest.approx(-COULOMB, abs=1e-9)
    assert _two_atom_system(0.25, 0.0, 0.5, 0.5, 2.0) == pytest.approx(COULOMB / 8, abs=1e-9)


def test_group_cutoff_is_by_centre():
    # the test is between the trial centre and the group centre: a group whose
    # centre is 15.5 A from the trial centre is skipped, but when the trial
    # centre sits on the group the pair counts at any atom distance
    assert _two_atom_system(0.25, 0.0, 1.0, 1.0, 14.5, True) == pytest.approx(COULOMB / 14.5, rel=1e-7)
    assert _two_atom_system(0.25, 0.0, 1.0, 1.0, 15.5, True) == 0.0
    assert _two_atom_system(0.25, 0.0, 1.0, 1.0, 15.5) == pytest.approx(COULOMB / 15.5, rel=1e-7)


# ---- the reference's water fixture (test/water/3water.pdb, test2.ipynb) --------
def test_water_fixture_energies():
    fx = json.load(open(os.path.join(GOLDEN, 'water3.json')))
    s, c, p, _ = systems.water3_system(True, False)
    e_coul = O.energy_pose(s, 0, c, p)[0]
    s, c, p, _ = systems.water3_system(False, True)
    e_lj = O.energy_pose(s, 0, c, p)[0]
    # inputs pass through the FP32 ABI, so agreement is at the FP32 input rounding
    assert abs(e_lj - fx['expected_lj_kj']) < 3e-4
    assert abs(e_coul - fx['expected_coulomb_kj']) < 1e-4
    # OpenMM's LJ-only difference, independent code (test2.ipynb cells 11-13)
    assert abs(e_lj - fx['openmm_lj_only_kj']) < 5e-3


# ---- move schedule, maxNum and cavity factor: test/4wp7/gcmc.75.out -------------
def test_move_schedule_4wp7():
    fx = json.load(open(os.path.join(GOLDEN, 'run4wp7.json')))
    rng = random.Random(fx['seed'])
    moves, counts = dataset.get_move(fx['names'], fx['mctime'], fx['mcsteps'], rng=rng)
    for i, n in enumerate(fx['names']):
        assert counts[4 * i:4 * i + 4] == fx['counts'][n], n


def test_maxnum_4wp7():
    fx = json.load(open(os.path.join(GOLDEN, 'run4wp7.json')))
    for name, m in fx['maxnum'].items():
        assert dataset.max_num(m['conc'], fx['volume'], m['ins'], m['n0']) == m['max'], name


def test_cavity_factor_definition():
    fx = json.load(open(os.path.join(GOLDEN, 'run4wp7.json')))
    n = fx['grid_n']
    assert fx['free_points'] / (n[0] * n[1] * n[2]) == fx['cavity_factor']


def test_grid_brute_force_and_cpython_order():
    rng = np.random.default_rng(5)
    xyz = np.round(rng.uniform(-3, 23, (300, 3)), 3)
    cryst = [20.0, 20.0, 20.0]
    start, n, grid, cav = dataset.get_grid(xyz, cryst, 1.0, order='linear')
    # brute force with Python ints, as the reference loops over atoms
    occ = set()
    for x, y, z in xyz.tolist():
        occ.add((int((x - start[0]) / 1.0) % n[0], int((y - start[1]) / 1.0) % n[1],
                 int((z - start[2]) / 1.0) % n[2]))
    free = [(a, b, c) for a in range(n[0]) for b in range(n[1]) for c in range(n[2]) if (a, b, c) not in occ]
    assert cav == len(free) / (n[0] * n[1] * n[2])
    g = grid.reshape(-1, 3)
    exp = np.array([[a + start[0], b + start[1], c + start[2]] for a, b, c in free], np.float32)
    assert np.array_equal(g, exp)
    _, _, grid2, _ = dataset.get_grid(xyz, cryst, 1.0, order='cpython')
    assert sorted(map(tuple, grid2.reshape(-1, 3).tolist())) == sorted(map(tuple, g.tolist()))


def test_set_fixcut_first_atom_rule():
    xyz = np.array([[0, 0, 0], [5, 0, 0], [6.5, 0, 0], [6.0, 0, 0], [20, 0, 0]], float)
    # 6.5 > 6 from the first atom opens group 1; 6.0 is within 6 of 6.5
    assert list(dataset.set_fixcut(xyz, 6.0)) == [0, 0, 1, 1, 2]


# ---- picks and acceptance -----------------------------------------------------------
def test_picks_are_fisher_yates():
    for N, n in ((10, 10), (57, 13), (1000, 400), (3, 1)):
        got = O.picks(123, 7, N, n)
        assert len(set(got.tolist())) == n and got.min() >= 0 and got.max() < N
        nums = list(range(N))
        for i in range(n):
            x = O.philox([7, i, 2, 0], [123, 0])[0]
            j = i + ((int(x) * (N - i)) >> 32)
            nums[i], nums[j] = nums[j], nums[i]
        assert got.tolist() == nums[:n]


def test_species_B():
    s = systems.water()
    F = s['frag'][0]
    L = s['info'][0]['cryst'].astype(np.float64)
    nbar = L[0] * L[1] * L[2] * float(F['conc']) * MOLES_TO_MOLECULES
    assert O.species_B(s, 0) == pytest.approx(BETA * float(F['muex']) + math.log(nbar), rel=1e-15)


def test_single_trial_insert_closed_form():
    """conf 1: p = min(1, cav * nbar e^{beta muex} e^{-beta E} / (N+1))."""
    s = systems.water(nsteps=300, conf=1)
    tr = O.new_trace(300, 1)
    O.run(s, trace=tr)
    info = s['info'][0]
    B = O.species_B(s, 0)
    N = 0
    checked = 0
    moves = s['moves']
    for st in range(300):
        kind = moves[st] % 4
        if kind == 0 and tr['k'][st] == 1:
            E = tr['e'][st, 0, 0]
            p = min(1.0, float(info['cavityFactor']) / (N + 1.0) * math.exp(B - BETA * E))
            x = O.philox([st, 0, 3, 0], [int(info['seed']), 0])
            ran = O.lib().orc_u01d(int(x[2]), int(x[3]))
            assert bool(tr['gacc'][st, 0]) == (ran < p)
            checked += 1
        N = tr['n_after'][st]
    assert checked > 20


def test_energy_invariances():
    s = systems.populated_water(300)
    c, p = systems.random_poses(s, 0, 8, seed=1)
    info = s['info'][0]
    L = info['cryst'].astype(np.float32)
    # coordinates on a 1/256 A lattice stay exact in FP32 after a box shift
    c = np.round(c * 256) / 256
    p = np.round(p * 256) / 256
    shift = L * np.float32([1, 0, -1])
    for i in range(8):
        e0, n0, h0 = O.energy_pose(s, 0, c[i], p[i])
        # a whole-box shift of the fragment and of its test centre changes nothing
        e1, n1, h1 = O.energy_pose(s, 0, c[i] + shift, p[i] + shift)
        assert (n0, h0) == (n1, h1) and e0 == pytest.approx(e1, rel=1e-12, abs=1e-10)


def test_brute_force_small_system():
    """Oracle energy = O(N^2) FP64 sum with the FP32 centre mask, written here
    from the group definitions (tiny system)."""
    s = systems.populated_water(200)
    c, p = systems.random_poses(s, 0, 6, seed=9)
    info = s['info'][0]
    L = info['cryst'].astype(np.float64)
    frag = s['frag'][0]
    X, Y = int(info['ffXNum']), int(info['ffYNum'])
    ff = s['ff'].reshape(X, Y, 2).astype(np.float64)
    nfix = int(frag['startRes'])
    groups = list(range(nfix)) + list(range(int(frag['startRes']), int(frag['startRes'] + frag['totalNum'])))
    for i in range(6):
        tot = 0.0
        for g in groups:
            r = s['res'][g]
            if O.distanceP32(c[i], r['position'], info['cryst']) > np.float32(info['cutoff']):
                continue
            for a in range(int(r['atomStart']), int(r['atomStart'] + r['atomNum'])):
                A = s['atoms'][a]
                for f in range(int(frag['num_atoms'])):
                    d = A['position'].astype(np.float64) - p[i][f].astype(np.float64)
                    d -= np.where(d / L >= 0, np.floor(d / L + 0.5), np.ceil(d / L - 0.5)) * L
                    r2 = float(d @ d)
                    sg, ep = ff[int(frag['atoms'][f]['type']), int(A['type'])]
                    s6 = (sg * sg * 100 / r2) ** 3
                    tot += 4 * ep * (s6 * s6 - s6) + COULOMB * float(frag['atoms'][f]['charge']) * \
                        float(A['charge']) / math.sqrt(r2)
        assert O.energy_pose(s, 0, c[i], p[i])[0] == pytest.approx(tot, rel=1e-12, abs=1e-9)


# ---- adjoint of the pose energy (FP64 oracle) ---------------------------------


def _adjoint_case():
    s = systems.multi(nsteps=10, n_fixed=3000, box=45.0)
    eps, rmin = O.type_params(s)
    return s, eps, rmin


def test_adjoint_energy_equals_table_energy():
    s, eps, rmin = _adjoint_case()
    c, p = systems.random_poses(s, 0, 20, seed=5)
    for i in range(len(c)):
        e_tab = O.energy_pose(s, 0, c[i], p[i])[0]
        e_par = O.adjoint(s, 0, c[i], p[i], eps, rmin)[0]
        # the pair table holds float32 (sigma, eps); the adjoint recombines
        assert abs(e_par - e_tab) <= 1e-6 * max(1.0, abs(e_tab))


def test_adjoint_gradients_match_finite_differences():
    s, eps, rmin = _adjoint_case()
    for spi in (0, 5, 8):
        c, p = systems.moderate_poses(s, spi, 2, seed=11 + spi)
        for i in range(len(c)):
            e, gx, gq, ge, gr = O.adjoint(s, spi, c[i], p[i], eps, rmin)
            fd = systems.fd_pose_gradient(s, spi, c[i], p[i], eps, rmin)
            assert np.linalg.norm(gx - fd) <= 1e-6 * np.linalg.norm(fd) + 1e-6
            used = np.nonzero(np.abs(gq) + np.abs(ge) + np.abs(gr))[0]
            assert len(used) > 3
            for t in used[:6]:
                for which, g in (('q', gq), ('eps', ge), ('rmin', gr)):
                    d = systems.fd_type_gradient(s, spi, c[i], p[i], eps, rmin, which, t)
                    # central differences with h = 1e-4 carry O((h/p)^2) error for
                    # parameters p ~ 0.03: 2e-5 relative
                    assert abs(g[t] - d) <= 2e-5 * max(abs(d), 1.0), (which, t, g[t], d)


def test_adjoint_epsilon_euler_identity():
    """E_LJ is homogeneous of degree 1 in the epsilons (pair epsilon =
    sqrt(eps_a eps_b)), so sum_t eps_t dE/deps_t = E - E(eps = 0)."""
    s, eps, rmin = _adjoint_case()
    for spi in (1, 6):
        c, p = systems.moderate_poses(s, spi, 3, seed=31 + spi)
        for i in range(len(c)):
            e, gx, gq, ge, gr = O.adjoint(s, spi, c[i], p[i], eps, rmin)
            e_el = O.adjoint(s, spi, c[i], p[i], np.zeros_like(eps), rmin)[0]
            assert abs(np.dot(eps, ge) - (e - e_el)) <= 1e-9 * max(1.0, abs(e - e_el))


def test_oracle_threads_do_not_change_the_chain():
    """Trials of a move are evaluated in parallel; the chain must not depend
    on the thread count."""
    s = systems.multi(nsteps=200, n_fixed=2000, box=40.0, seed=31)
    a, b = O.copy_system(s), O.copy_system(s)
    O.run(a, nthreads=1)
    O.run(b, nthreads=4)
    assert np.array_equal(a['atoms'].view(np.uint8), b['atoms'].view(np.uint8))
    assert np.array_equal(a['frag']['totalNum'], b['frag']['totalNum'])
