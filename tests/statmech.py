"""Small systems whose grand-canonical equilibrium is known in closed form,
and the statistics to compare a chain with it (shared by the oracle pins and
the GPU occupancy tests).

* ideal gas: molecules with no charge and no LJ interaction with anything.
  With k insertion trials that all fall in one selection group (cutoff above
  half the box diagonal), every weight is equal, P = 1/k, and the insertion /
  deletion acceptances of the reference (gcmc/cpp/gcmc_move_add.h:164-174,
  gcmc_move_del.h:172-176) are min(1, f lam/(N+1)) and min(1, N/lam) with
  lam = exp(B) = V c 6.023e-4 exp(beta muex) and f the cavity factor.  For
  f = 1 the number of molecules is Poisson(lam); for f < 1 the two rules are
  not each other's reverse and the stationary distribution follows from the
  ratios pi(N+1)/pi(N) = p_ins(N) / p_del(N+1) of the birth-death chain.
* LJ particles in an external field: one-atom molecules that interact with
  fixed LJ sites but not with each other.  Independent particles in the grand
  canonical ensemble: N is Poisson with mean lam <exp(-beta V(x))>, the
  average over insertion points spread uniformly over the box (a cavity grid
  of every cell), when only insertions and deletions are attempted (the
  reference's translation draws only positive displacements,
  gcmc_move_trn.h:54-66, so it is not its own reverse and would bias the
  spatial distribution).
"""

import math

import numpy as np

from pygcmc_amd import dataset, synth
from pygcmc_amd.values import BETA, MOLES_TO_MOLECULES


def _all_cells(start, n, dx):
    i, j, k = np.meshgrid(np.arange(n[0]), np.arange(n[1]), np.arange(n[2]), indexing='ij')
    idx = np.stack([i.ravel(), j.ravel(), k.ravel()], 1).astype(np.float64)
    return (np.asarray(start, np.float64)[None, :] + idx * dx).astype(np.float32), idx


def set_grid(sysd, keep=None):
    """Replace the cavity grid by every cell of the box (keep=None, cavity
    factor 1) or by the cells where keep(i, j, k) holds (cavity factor =
    their fraction)."""
    info = sysd['info'][0]
    dx = float(info['grid_dx'])
    n = [int(round(float(info['cryst'][i]) / dx)) for i in range(3)]
    g, idx = _all_cells(info['startxyz'], n, dx)
    if keep is not None:
        m = keep(idx[:, 0], idx[:, 1], idx[:, 2])
        g = g[m]
    sysd['grid'] = np.ascontiguousarray(g.reshape(-1), np.float32)
    sysd['info'][0]['totalGridNum'] = len(g)
    sysd['info'][0]['cavityFactor'] = np.float32(len(g) / (n[0] * n[1] * n[2]))
    return sysd


def lam(sysd, species=0):
    """exp(B): the mean count of the ideal gas at cavity factor 1 (FP64 from
    the FP32 fields the engines read)."""
    info = sysd['info'][0]
    F = sysd['frag'][species]
    v = float(info['cryst'][0]) * float(info['cryst'][1]) * float(info['cryst'][2])
    return v * float(F['conc']) * MOLES_TO_MOLECULES * math.exp(BETA * float(F['muex']))


def ideal_gas(seed=1, box=20.0, conc=1.0, muex=0.0, conf=1, cutoff=15.0, mcsteps=200000, name='SOL'):
    """Non-interacting TIP3P-shaped molecules (no charge, no LJ) in a box with
    four inert fixed sites (they set the box origin and take no part)."""
    xyz, el = synth.template_water()
    zero = dataset.charmm_to_gromacs(0.0, 0.0)
    spec = dict(name=name, conc=conc, muex=muex, conf=conf, mctime=1.0, template_xyz=xyz,
                charge=[0.0] * len(el), types=['HX'] * len(el), lj=[zero] * len(el), molecules=[])
    fixed = np.array([[0.0, 0.0, 0.0], [box / 2, 0.0, 0.0], [0.0, box / 2, 0.0], [0.0, 0.0, box / 2]])
    s = synth.build_system(seed=seed, fixed_xyz=fixed, fixed_charge=np.zeros(4), fixed_type_names=['XX'] * 4,
                           protein_lj={'XX': zero}, species_specs=[spec], cryst=[box] * 3, mcsteps=mcsteps,
                           cutoff=cutoff)
    return set_grid(s)


def field_particles(seed=7, box=20.0, n_sites=30, conc=1.0, muex=0.0, cutoff=12.0, mcsteps=200000,
                    site_lj=(0.25, 1.8), particle_lj=(0.2, 1.6)):
    """One-atom LJ particles (no charge) among n_sites fixed LJ sites; the
    particle-particle pair is switched off in the pair table.  The schedule
    holds insertions and deletions only (equal odds, seeded)."""
    rng = np.random.default_rng(seed)
    sites = synth.random_sphere_atoms(rng, n_sites, box * 0.6, [box] * 3, min_dist=3.0,
                                      centre=np.zeros(3))
    spec = dict(name='LJP', conc=conc, muex=muex, conf=1, mctime=1.0, template_xyz=np.zeros((1, 3)),
                charge=[0.0], types=['LJP'], lj=[dataset.charmm_to_gromacs(*particle_lj)], molecules=[])
    s = synth.build_system(seed=seed, fixed_xyz=sites, fixed_charge=np.zeros(n_sites),
                           fixed_type_names=['SITE'] * n_sites,
                           protein_lj={'SITE': dataset.charmm_to_gromacs(*site_lj)}, species_specs=[spec],
                           cryst=[box] * 3, mcsteps=mcsteps, cutoff=cutoff)
    Y = int(s['info'][0]['ffYNum'])
    t = s['types2'].index('LJP')
    s['ff'][2 * (t * Y + t)] = 0.0          # particle-particle sigma
    s['ff'][2 * (t * Y + t) + 1] = 0.0      # particle-particle epsilon
    s['moves'] = rng.integers(0, 2, mcsteps).astype(np.int32)   # species 0: insert (0) / delete (1)
    return set_grid(s)


def field_boltzmann_average(sysd, npts=400000, seed=11):
    """<exp(-beta V(x))> over x uniform in the box, V the energy of one
    particle with the fixed sites under the reference's group-centre cutoff
    (every atom of a group whose centre is within rc; gcmc_energy.h:152-191),
    by Monte Carlo quadrature in FP64.  Returns (mean, standard error)."""
    info = sysd['info'][0]
    L = info['cryst'].astype(np.float64)
    start = info['startxyz'].astype(np.float64)
    rc = float(info['cutoff'])
    nfix = int(sysd['frag'][0]['startRes'])
    res = sysd['res'][:nfix]
    Y = int(info['ffYNum'])
    tp = int(sysd['frag'][0]['atoms'][0]['type'])
    rng = np.random.default_rng(seed)
    x = start + rng.random((npts, 3)) * L
    V = np.zeros(npts)

    def mi(d):
        return d - np.round(d / L) * L

    for g in range(nfix):
        c = res[g]['position'].astype(np.float64)
        near = np.sqrt((mi(x - c) ** 2).sum(1)) <= rc
        for a in range(int(res[g]['atomStart']), int(res[g]['atomStart'] + res[g]['atomNum'])):
            at = sysd['atoms'][a]
            sig = float(sysd['ff'][2 * (tp * Y + at['type'])])
            eps = float(sysd['ff'][2 * (tp * Y + at['type']) + 1])
            r2 = (mi(x[near] - at['position'].astype(np.float64)) ** 2).sum(1)
            s6 = (100.0 * sig * sig / r2) ** 3
            V[near] += 4.0 * eps * (s6 * s6 - s6)
    w = np.exp(-BETA * V)
    return float(w.mean()), float(w.std() / math.sqrt(npts))


def birth_death_mean(lam0, cav, nmax=400):
    """Stationary mean of the count under insertion min(1, cav lam0/(N+1)) and
    deletion min(1, N/lam0) with equal odds of attempting either."""
    logp = [0.0]
    for n in range(nmax):
        pin = min(1.0, cav * lam0 / (n + 1))
        pdel = min(1.0, (n + 1) / lam0)
        logp.append(logp[-1] + math.log(pin) - math.log(pdel))
    logp = np.array(logp)
    p = np.exp(logp - logp.max())
    p /= p.sum()
    k = np.arange(len(p))
    return float((k * p).sum()), float((k * k * p).sum() - (k * p).sum() ** 2)


def block_stats(x, nblocks=40):
    """Mean of a series and its standard error from nblocks block means;
    the same for the second central moment."""
    x = np.asarray(x, np.float64)
    n = len(x) // nblocks * nblocks
    b = x[:n].reshape(nblocks, -1)
    m = b.mean(1)
    v = ((b - b.mean()) ** 2).mean(1)
    return float(m.mean()), float(m.std(ddof=1) / math.sqrt(nblocks)), float(v.mean()), \
        float(v.std(ddof=1) / math.sqrt(nblocks))


def interacting_particles(seed=13, box=18.0, conc=30.0, muex=1.0, cutoff=9.0, mcsteps=200000, n_sites=20):
    """One-atom LJ particles that interact with each other and with fixed LJ
    sites (a dense-ish LJ fluid in a field): no closed form, so its mean
    count is compared between independent chains.  Insertions, deletions
    and translations (the schedule's own move mix)."""
    s = field_particles(seed=seed, box=box, n_sites=n_sites, conc=conc, muex=muex, cutoff=cutoff,
                        mcsteps=mcsteps)
    rng = np.random.default_rng(seed + 1)
    s['moves'] = rng.choice(np.array([0, 1, 2], np.int32), mcsteps, p=[0.35, 0.35, 0.3]).astype(np.int32)
    # restore the particle-particle pair: geometric / arithmetic combination
    # of the particle's own parameters as the other pairs
    Y = int(s['info'][0]['ffYNum'])
    t = s['types2'].index('LJP')
    sig, eps = dataset.charmm_to_gromacs(0.2, 1.6)
    s['ff'][2 * (t * Y + t)] = sig
    s['ff'][2 * (t * Y + t) + 1] = eps
    return s
