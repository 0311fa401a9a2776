"""The FP64 Ewald checker of the PME kernels (oracle/ewald.py), pinned by
the NaCl Madelung constant, independence of the total from the splitting
parameter, and the cardinal B-spline partition of unity."""

import math

import numpy as np

from oracle import ewald

MADELUNG_NACL = 1.747564594633182


def test_nacl_madelung_constant():
    xyz, q, L = ewald.nacl_crystal(ncell=2, a=5.64)
    beta, rc = 0.8, 5.64
    e = ewald.ewald_energy(xyz, q, L, beta, rc, kmax=(16, 16, 16))
    npair = len(q) // 2
    ref = -npair * ewald.KCOUL * MADELUNG_NACL / (5.64 / 2)
    assert abs(e - ref) <= 1e-8 * abs(ref)


def test_total_is_independent_of_beta():
    rng = np.random.default_rng(0)
    L = [20.0, 17.0, 23.0]
    xyz = rng.random((40, 3)) * L
    q = rng.normal(0, 0.5, 40)
    q -= q.mean()
    rc = 8.5
    e1 = ewald.ewald_energy(xyz, q, L, 0.55, rc, kmax=(20, 18, 22))
    e2 = ewald.ewald_energy(xyz, q, L, 0.65, rc, kmax=(24, 21, 26))
    assert abs(e1 - e2) <= 1e-8 * max(1.0, abs(e1))


def test_recip_potential_gives_recip_energy():
    """E_recip = 1/2 sum_i q_i phi_recip(x_i) (phi includes each charge's own
    periodic term)."""
    rng = np.random.default_rng(1)
    L = [15.0, 15.0, 15.0]
    xyz = rng.random((20, 3)) * L
    q = rng.normal(0, 1, 20)
    e = ewald.recip_energy(xyz, q, L, 0.4, (8, 8, 8))
    phi = ewald.recip_potential(xyz, q, L, 0.4, (8, 8, 8), xyz)
    assert abs(0.5 * float((q * phi).sum()) - e) <= 1e-10 * abs(e)


def test_bspline_partition_of_unity():
    u = np.linspace(0, 1, 101, endpoint=False)
    s = sum(ewald.bspline4(u + k) for k in range(4))
    assert np.allclose(s, 1.0, atol=1e-14)
    # first moment: sum_k (u + k) M4(u + k) = 2, the mean of M4 (symmetric about 2)
    m1 = sum((u + k) * ewald.bspline4(u + k) for k in range(4))
    assert np.allclose(m1, 2.0, atol=1e-13)
    assert math.isclose(float(ewald.bspline4(np.array([2.0]))[0]), 2.0 / 3.0)


def test_fp64_pme_converges_to_ewald_sums():
    """The FP64 SPME restatement against the Ewald sums: the reciprocal
    energy and potential converge as h^4 (order-4 B-splines) -- halving the
    grid spacing cuts the error ~16x -- and E_rec = 1/2 sum q phi."""
    rng = np.random.default_rng(5)
    L = np.array([18.0, 16.0, 20.0])
    xyz = rng.random((60, 3)) * L
    q = rng.normal(0, 0.5, 60)
    q -= q.mean()
    beta = 0.34
    kmax = tuple(int(np.ceil(4.8 * l * beta / np.pi)) for l in L)
    e_ref = ewald.recip_energy(xyz, q, L, beta, kmax)
    pts = rng.random((40, 3)) * L
    env = ewald.EwaldEnv(xyz, q, L, beta, kmax)
    phi_ref = env.phi(pts)
    errs = []
    for h in (1.0, 0.5, 0.25):
        K = [int(round(l / h)) for l in L]
        pme = ewald.PMEEnv(xyz, q, L, beta, K)
        assert abs(0.5 * float((q * pme.phi(xyz)).sum()) - pme.energy) <= 1e-9 * abs(pme.energy)
        errs.append((abs(pme.energy - e_ref), float(np.abs(pme.phi(pts) - phi_ref).max())))
    print('PME FP64 vs Ewald (energy, max phi error) at h = 1, 0.5, 0.25 A:', errs)
    for a, b in zip(errs, errs[1:]):
        assert b[1] < a[1] / 8 and b[0] < a[0] / 8     # h^4: ~16x per halving
    assert errs[-1][1] < 1e-3 and errs[-1][0] < 1e-2
