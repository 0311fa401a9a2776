"""GPU PME reciprocal part against the FP64 Ewald sums (oracle/ewald.py).

Bars (order-4 B-splines, beta = 0.34 /A): reciprocal energy within 1e-3
relative at 1 A grid spacing and 6e-5 at 0.5 A (measured 5.7e-4 and 3.0e-5:
the h^4 convergence of cubic B-spline interpolation, checked as such);
potentials within 2e-3 of their spread; exact invariants: neutral pairs at
coincident points give zero, 1/2 sum q phi = E_rec.
"""

import numpy as np
import pytest
import torch

from oracle import ewald
from pygcmc_amd.pme import PME

pytestmark = pytest.mark.gpu


def _charges(n, L, seed):
    rng = np.random.default_rng(seed)
    xyz = rng.random((n, 3)) * L
    q = rng.normal(0, 0.5, n)
    q -= q.mean()
    return xyz, q


def _t(xyz, q=None):
    a = xyz if q is None else np.concatenate([xyz, q[:, None]], 1)
    return torch.tensor(a, dtype=torch.float32, device='cuda')


def test_pme_energy_against_ewald():
    L = [24.0, 21.0, 27.0]
    beta = 0.34
    xyz, q = _charges(200, L, 1)
    ref = ewald.recip_energy(xyz, q, L, beta, kmax=(14, 12, 16))
    rels = {}
    for spacing in (1.0, 0.5):
        grid = [int(round(l / spacing)) for l in L]
        p = PME(grid, L, beta)
        e = p.energy(_t(xyz, q))
        rels[spacing] = abs(e - ref) / abs(ref)
        print(f'spacing {spacing}: E_rec {e:.6f} ref {ref:.6f} rel {rels[spacing]:.2e}')
        p.close()
    assert rels[1.0] <= 1e-3 and rels[0.5] <= 6e-5
    # fourth-order convergence: halving the spacing cuts the error ~16x
    assert rels[1.0] / rels[0.5] >= 8.0


def test_pme_potential_against_ewald():
    L = [24.0, 21.0, 27.0]
    beta = 0.34
    xyz, q = _charges(150, L, 2)
    pts = np.random.default_rng(3).random((64, 3)) * L
    ref = ewald.recip_potential(xyz, q, L, beta, (14, 12, 16), pts)
    p = PME([24, 21, 27], L, beta)
    p.clear()
    p.spread(_t(xyz, q))
    E = p.solve()
    phi = p.gather(_t(pts)).cpu().numpy()
    err = np.abs(phi - ref).max() / (np.abs(ref).max() + 1e-9)
    print(f'potential: max err / max |phi| = {err:.2e}')
    assert err <= 2e-3
    # 1/2 sum_i q_i phi(x_i) = E_rec (same grid quantities)
    phi_s = p.gather(_t(xyz)).cpu().numpy().astype(np.float64)
    assert abs(0.5 * float((q * phi_s).sum()) - float(E.item())) <= 1e-4 * abs(float(E.item()))
    p.close()


def test_pme_neutral_coincident_pair_is_zero():
    L = [20.0, 20.0, 20.0]
    p = PME([20, 20, 20], L, 0.34)
    x = np.array([[3.3, 7.1, 11.9], [3.3, 7.1, 11.9]])
    e = p.energy(_t(x, np.array([1.0, -1.0])))
    assert abs(e) < 1e-6
    p.close()


def test_pme_slab_config_grid():
    """The 150 x 150 x 120 A box on a 1 A grid (2.7 M points), beta 0.34."""
    L = [150.0, 150.0, 120.0]
    xyz, q = _charges(100, L, 5)
    p = PME([150, 150, 120], L, 0.34)
    e = p.energy(_t(xyz, q))
    # terms beyond |k_i| = 0.4 L_i are below exp(-13.7) of the largest
    ref = ewald.recip_energy(xyz, q, L, 0.34, kmax=(60, 60, 48))
    assert abs(e - ref) <= 1e-3 * abs(ref)
    p.close()
