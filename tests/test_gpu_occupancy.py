"""Mean fragment occupancy of long GPU chains within statistical error:
against the closed forms of systems whose grand-canonical equilibrium is
known (tests/statmech.py), and against an independent FP64 oracle chain
(another seed) for an interacting system with no closed form.

Reference rules these pin on the GPU path: insertion acceptance
gcmc/cpp/gcmc_move_add.h:164-174 (cavity factor, 1/(k P_j), 1/(N+1),
exp(B - beta E)), deletion gcmc_move_del.h:166-176 (N exp(-B + beta E), no
cavity factor), translation gcmc_move_trn.h:215-222.  Errors are block
averages over 40 blocks after a burn-in; bars are 4.5 standard errors.
"""

import math

import numpy as np
import pytest

from oracle import oracle as O
from pygcmc_amd.engine import Engine

from . import statmech as SM

pytestmark = pytest.mark.gpu

NSTEPS = 200000
BURN = 2000


def _gpu_counts(sysd, nsteps=NSTEPS, seed=None):
    s = O.copy_system(sysd)
    if seed is not None:
        s['info'][0]['seed'] = np.uint32(seed)
    eng = Engine()
    eng.add_system(s)
    eng.upload()
    eng.set_trace(0, nsteps, int(s['frag']['confBias'].max()))
    eng.run(nsteps, 'batched')
    tr = eng.get_trace(0)
    eng.close()
    return tr['n_after']


@pytest.mark.parametrize('conf,cutoff,seed', [(1, 15.0, 101), (4, 18.0, 102)])
def test_gpu_ideal_gas_count_is_poisson(conf, cutoff, seed):
    s = SM.ideal_gas(conf=conf, cutoff=cutoff, seed=seed, mcsteps=NSTEPS)
    n = _gpu_counts(s)
    lam = SM.lam(s)
    m, se, v, sev = SM.block_stats(n[BURN:])
    print(f'GPU ideal gas conf {conf}: <N> {m:.4f} +- {se:.4f} (Poisson {lam:.4f}); var {v:.3f} +- {sev:.3f}')
    assert abs(m - lam) < 4.5 * se
    assert abs(v - lam) < 4.5 * sev


def test_gpu_ideal_gas_with_cavity_factor():
    s = SM.ideal_gas(conf=1, seed=103, mcsteps=NSTEPS)
    SM.set_grid(s, keep=lambda i, j, k: (i + j + k) % 2 == 0)
    f = float(s['info'][0]['cavityFactor'])
    n = _gpu_counts(s)
    mean, var = SM.birth_death_mean(SM.lam(s), f)
    m, se, v, sev = SM.block_stats(n[BURN:])
    print(f'GPU cavity factor {f}: <N> {m:.4f} +- {se:.4f} (birth-death {mean:.4f})')
    assert abs(m - mean) < 4.5 * se
    assert abs(v - var) < 4.5 * sev


def test_gpu_particles_in_a_field():
    s = SM.field_particles(mcsteps=NSTEPS, seed=104)
    w, wse = SM.field_boltzmann_average(s)
    lam = SM.lam(s)
    n = _gpu_counts(s)
    m, se, v, sev = SM.block_stats(n[BURN:])
    target = lam * w
    err = math.hypot(se, lam * wse)
    print(f'GPU field: <N> {m:.4f} +- {se:.4f}, lam <exp(-bV)> {target:.4f} +- {lam * wse:.4f}')
    assert abs(m - target) < 4.5 * err
    assert abs(target - lam) > 8 * err


def test_gpu_interacting_occupancy_matches_independent_oracle_chain():
    """LJ particles interacting with each other and with fixed sites: the
    GPU chain's mean count and variance against an FP64 oracle chain with
    another seed (independent variates, same ensemble)."""
    s = SM.interacting_particles(mcsteps=NSTEPS)
    n_gpu = _gpu_counts(s, seed=777)
    ref = O.copy_system(s)
    ref['info'][0]['seed'] = np.uint32(9191)
    tr = O.new_trace(NSTEPS, 1)
    O.run(ref, trace=tr)
    n_cpu = tr['n_after']
    burn = 20000
    m1, se1, v1, sev1 = SM.block_stats(n_gpu[burn:])
    m2, se2, v2, sev2 = SM.block_stats(n_cpu[burn:])
    lam = SM.lam(s)
    print(f'interacting: GPU <N> {m1:.3f} +- {se1:.3f} var {v1:.2f}; oracle <N> {m2:.3f} +- {se2:.3f} '
          f'var {v2:.2f} (ideal-gas lam {lam:.2f})')
    assert abs(m1 - m2) < 4.5 * math.hypot(se1, se2)
    assert abs(v1 - v2) < 4.5 * math.hypot(sev1, sev2)
    assert abs(m1 - lam) > 10 * se1            # the interactions matter
