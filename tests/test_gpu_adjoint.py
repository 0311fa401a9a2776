"""The HIP adjoint (pygcmc_amd.adjoint) against the FP64 oracle gradient and
central finite differences of the FP64 energy (h = 1e-4).

Bars: energies as the MC path (4.184e-4 kJ/mol for |E| <= 200 kJ/mol);
per-pose coordinate gradients and loss gradients in the per-type parameters
within 1e-4 relative (2-norm over the pose's atoms / over the types).
"""

import numpy as np
import pytest
import torch

from oracle import oracle as O
from pygcmc_amd import adjoint
from pygcmc_amd.engine import Engine

from . import systems

pytestmark = pytest.mark.gpu

KCAL = 4.184e-4


def _setup(n_per_species=24, seed=41):
    s = systems.multi(nsteps=10, n_fixed=6000, box=60.0)
    eng = Engine()
    eng.add_system(s)
    eng.upload()
    eps, rmin = adjoint.type_parameters(s)
    poses = {}
    T = int(s['info'][0]['fragTypeNum'])
    for spi in range(T):
        poses[spi] = systems.moderate_poses(s, spi, n_per_species, seed=seed + spi)
    return s, eng, eps, rmin, poses


@pytest.fixture(scope='module')
def case():
    s, eng, eps, rmin, poses = _setup()
    yield s, eng, eps, rmin, poses
    eng.close()


def _gpu(eng, spi, c, p, eps, rmin, dq=None):
    ct = torch.tensor(c, device='cuda')
    pt = torch.tensor(p, device='cuda', requires_grad=True)
    e = eps.clone().requires_grad_(True)
    r = rmin.clone().requires_grad_(True)
    q = (torch.zeros_like(eps) if dq is None else dq.clone()).requires_grad_(True)
    E = adjoint.pose_energy(eng, 0, spi, ct, pt, e, r, q)
    return E, pt, e, r, q


def test_adjoint_energy_matches_fp64_and_mc_path(case):
    """Energies against the FP64 oracle with the same per-type parameters
    (1e-4 kcal/mol), and against the MC path's energies, whose pair table
    rounds (sigma, epsilon) once per pair instead of once per type: within
    2e-6 of the sum of |pair terms| scale, bounded here by 5e-3 kJ/mol."""
    s, eng, eps, rmin, poses = case
    e64p, r64p = eps.double().cpu().numpy(), rmin.double().cpu().numpy()
    for spi, (c, p) in poses.items():
        E, *_ = _gpu(eng, spi, c, p, eps, rmin)
        e_mc, n, _ = eng.eval_poses(0, spi, c, p)
        e64 = np.array([O.adjoint(s, spi, c[i], p[i], e64p, r64p)[0] for i in range(len(c))])
        Eg = E.detach().cpu().numpy()
        assert np.abs(Eg - e64).max() <= KCAL, (spi, np.abs(Eg - e64).max())
        assert np.abs(Eg - e_mc).max() <= 5e-3, (spi, np.abs(Eg - e_mc).max())


def test_adjoint_pose_gradients_match_fp64(case):
    s, eng, eps, rmin, poses = case
    e64, r64 = eps.double().cpu().numpy(), rmin.double().cpu().numpy()
    worst = 0.0
    for spi, (c, p) in poses.items():
        E, pt, *_ = _gpu(eng, spi, c, p, eps, rmin)
        E.sum().backward()
        g = pt.grad.cpu().numpy()
        for i in range(len(c)):
            gref = O.adjoint(s, spi, c[i], p[i], e64, r64)[1]
            rel = np.linalg.norm(g[i] - gref) / max(np.linalg.norm(gref), 1e-3)
            worst = max(worst, rel)
    print(f'pose gradient: worst relative error vs FP64 {worst:.2e}')
    assert worst <= 1e-4


def test_adjoint_type_gradients_match_fp64(case):
    s, eng, eps, rmin, poses = case
    e64, r64 = eps.double().cpu().numpy(), rmin.double().cpu().numpy()
    rng = np.random.default_rng(3)
    dq = torch.tensor(rng.normal(0, 0.02, len(e64)), dtype=torch.float32, device='cuda')
    dq64 = dq.double().cpu().numpy()
    for spi, (c, p) in poses.items():
        w = rng.normal(size=len(c))
        E, pt, e, r, q = _gpu(eng, spi, c, p, eps, rmin, dq)
        (E * torch.tensor(w, device='cuda')).sum().backward()
        ref = np.zeros((3, len(e64)))
        for i in range(len(c)):
            _, _, gq, ge, gr = O.adjoint(s, spi, c[i], p[i], e64, r64, dq64)
            ref += w[i] * np.stack([gq, ge, gr])
        got = np.stack([q.grad.cpu().numpy(), e.grad.cpu().numpy(), r.grad.cpu().numpy()])
        for k, name in enumerate(('dq', 'eps', 'rmin')):
            rel = np.linalg.norm(got[k] - ref[k]) / max(np.linalg.norm(ref[k]), 1e-3)
            print(f'species {spi} {name}: relative error {rel:.2e}')
            assert rel <= 1e-4, (spi, name, rel)


def test_adjoint_matches_finite_differences(case):
    """GPU gradients against central differences (h = 1e-4) of the FP64 energy."""
    s, eng, eps, rmin, poses = case
    e64, r64 = eps.double().cpu().numpy(), rmin.double().cpu().numpy()
    for spi in (0, 4, 8):
        c, p = poses[spi]
        c, p = c[:2], p[:2]
        E, pt, e, r, q = _gpu(eng, spi, c[:1], p[:1], eps, rmin)
        E.sum().backward()
        fd = systems.fd_pose_gradient(s, spi, c[0], p[0], e64, r64)
        g = pt.grad.cpu().numpy()[0]
        assert np.linalg.norm(g - fd) <= 1e-4 * np.linalg.norm(fd)
        for which, got in (('q', q.grad), ('eps', e.grad), ('rmin', r.grad)):
            got = got.cpu().numpy()
            for t in np.argsort(-np.abs(got))[:4]:
                d = systems.fd_type_gradient(s, spi, c[0], p[0], e64, r64, which, int(t))
                assert abs(got[t] - d) <= 1e-4 * max(abs(d), 1e-2), (spi, which, t, got[t], d)


def test_adjoint_rejects_bad_shapes(case):
    s, eng, eps, rmin, poses = case
    c, p = poses[0]
    with pytest.raises(ValueError):
        adjoint.pose_energy(eng, 0, 0, torch.tensor(c, device='cuda'),
                            torch.tensor(p[:, :-1], device='cuda'), eps, rmin)
    with pytest.raises(TypeError):
        adjoint.pose_energy(eng, 0, 0, torch.tensor(c), torch.tensor(p), eps, rmin)
