"""Small seeded systems shared by the tests (built once per process)."""

import functools

import numpy as np

from oracle import oracle as O
from pygcmc_amd import synth


def copy(sysd):
    return O.copy_system(sysd)


@functools.lru_cache(maxsize=None)
def _water(nsteps, conf, n_fixed):
    return synth.water_sphere(mcsteps=nsteps, conf=conf, n_fixed=n_fixed)


def water(nsteps=1000, conf=1, n_fixed=2000):
    """Water around 2,000 fixed atoms in a 40 A box (fresh copy)."""
    return copy(_water(nsteps, conf, n_fixed))


@functools.lru_cache(maxsize=None)
def _multi(nsteps, n_fixed, box, conf_scale, seed):
    return synth.multi_fragment(seed=seed, n_fixed=n_fixed, box=box, mcsteps=nsteps,
                                conf_scale=conf_scale)


def multi(nsteps=400, n_fixed=2000, box=40.0, conf_scale=0.1, seed=2025):
    """8 fragments + water around fixed atoms (reduced sizes, fresh copy)."""
    return copy(_multi(nsteps, n_fixed, box, conf_scale, seed))


def populated_water(nsteps_pre=600, conf=4):
    """Water system advanced by the oracle so that it holds some waters."""
    s = water(nsteps=2000, conf=conf)
    O.run(s, first=0, nsteps=nsteps_pre)
    return s


def random_poses(sysd, species, n, seed=0):
    """n random rigid poses of a species: centres uniform in the box,
    uniformly random orientations (Shoemake quaternions)."""
    rng = np.random.default_rng(seed)
    info = sysd['info'][0]
    F = sysd['frag'][species]
    nf = int(F['num_atoms'])
    t = F['atoms'][:nf]['position'].astype(np.float64)
    L = info['cryst'].astype(np.float64)
    st = info['startxyz'].astype(np.float64)
    c = st + rng.random((n, 3)) * L
    u1, u2, u3 = rng.random((3, n))
    q = np.stack([np.sqrt(1 - u1) * np.sin(2 * np.pi * u2), np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                  np.sqrt(u1) * np.sin(2 * np.pi * u3), np.sqrt(u1) * np.cos(2 * np.pi * u3)], 1)
    x, y, z, w = q.T
    R = np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w),
                  2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w),
                  2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], 1).reshape(n, 3, 3)
    poses = np.einsum('nij,aj->nai', R, t) + c[:, None, :]
    return c.astype(np.float32), poses.astype(np.float32)


def water3_system(charges=True, lj=True, shift=(30.0, 40.0, 150.0)):
    """The reference's 3-water fixture as a packed system: waters 1-2 are
    active SOL slots, water 3 is the pose to insert (returned separately).
    Coordinates are translated by -shift (energy is translation invariant in
    the 211 A box) so that their FP32 storage error is ~1e-8 A."""
    import json
    import os
    from pygcmc_amd import dataset
    here = os.path.dirname(os.path.abspath(__file__))
    fx = json.load(open(os.path.join(here, 'golden', 'water3.json')))
    W = [np.array(w, dtype=np.float64) - np.array(shift) for w in fx['waters']]
    q = fx['charges'] if charges else [0.0, 0.0, 0.0]
    nb = {t: (fx['sigma_nm'][t], fx['epsilon_kj'][t] if lj else 0.0) for t in ('OT', 'HT')}
    ff = dataset.ff_pair_table(['OT', 'HT'], ['OT', 'HT'], nb)
    cryst = fx['cryst']
    allxyz = np.concatenate(W)
    start, n, grid, cav = dataset.get_grid(allxyz, cryst, 8.0)
    sp = [dict(name='SOL', conc=55.0, muex=-5.8, conf=1, mctime=1.0, template_xyz=W[0],
               charge=q, type=[0, 1, 1], molecules=[W[0], W[1]])]
    info, frag, res, atoms, grid, ff, moves = dataset.pack_system(
        cryst=cryst, cutoff=15.0, grid_dx=8.0, startxyz=start, grid=grid, cavity_factor=cav,
        fixed_xyz=np.zeros((0, 3)), fixed_charge=[], fixed_type=[], fixed_group=np.zeros(0, np.int64),
        species=sp, ff=ff, ffX=2, ffY=2, moves=np.zeros(1, np.int32), move_counts=[1, 0, 0, 0],
        mcsteps=1, seed=1)
    sysd = dict(info=info, frag=frag, res=res, atoms=atoms, grid=grid, ff=ff, moves=moves)
    pose = W[2].astype(np.float32)
    centre = pose.mean(axis=0).astype(np.float32)
    return sysd, centre, pose, fx


def moderate_poses(sysd, species, n, seed=0, emax=200.0):
    """n random poses (random_poses) whose FP64 energy lies in [-emax, emax]:
    no overlapping atoms, so gradients are well conditioned."""
    cs, ps = [], []
    s = seed
    while len(cs) < n:
        c, p = random_poses(sysd, species, 4 * n, seed=s)
        s += 1
        for i in range(len(c)):
            if abs(O.energy_pose(sysd, species, c[i], p[i])[0]) <= emax:
                cs.append(c[i])
                ps.append(p[i])
                if len(cs) == n:
                    break
    return np.array(cs, np.float32), np.array(ps, np.float32)


def fd_pose_gradient(sysd, species, centre, pose, eps, rmin, h=1e-4):
    """Central differences of the FP64 oracle energy in each pose coordinate.
    The pose is FP32, so the step is the one actually represented."""
    g = np.zeros(pose.shape, np.float64)
    for f in range(pose.shape[0]):
        for k in range(3):
            pp, pm = pose.copy(), pose.copy()
            pp[f, k] = np.float32(pose[f, k] + h)
            pm[f, k] = np.float32(pose[f, k] - h)
            ep = O.adjoint(sysd, species, centre, pp, eps, rmin)[0]
            em = O.adjoint(sysd, species, centre, pm, eps, rmin)[0]
            g[f, k] = (ep - em) / (float(pp[f, k]) - float(pm[f, k]))
    return g


def fd_type_gradient(sysd, species, centre, pose, eps, rmin, which, t, h=1e-4):
    """Central difference of the FP64 oracle energy in one per-type parameter
    (which: 'q' charge shift, 'eps', 'rmin')."""
    Y = len(eps)
    out = []
    for sgn in (1.0, -1.0):
        e, r, dq = eps.copy(), rmin.copy(), np.zeros(Y)
        {'q': dq, 'eps': e, 'rmin': r}[which][t] += sgn * h
        out.append(O.adjoint(sysd, species, centre, pose, e, r, dq)[0])
    return (out[0] - out[1]) / (2 * h)
