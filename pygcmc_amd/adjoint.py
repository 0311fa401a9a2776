"""Differentiable trial-pose energies (torch.autograd) over the HIP adjoint
kernel (``gcmc_engine_adjoint`` in ``include/pygcmc_hip.h``).

``pose_energy(engine, replica, species, centres, poses, eps, rmin, dq)``
returns the FP64 interaction energies (kJ/mol) of ``poses`` [P, nf, 3] of one
fragment species with the replica's current environment -- calcEnergy's
group-centre cutoff around ``centres`` [P, 3] -- as a function of the pose
coordinates and of per-type parameters over the system's type list
(``ffYNum`` entries): CHARMM epsilon ``eps`` (kcal/mol), CHARMM Rmin/2
``rmin`` (A) and a charge shift ``dq`` (e) applied to every atom of a type.
Pairs combine as sigma = (sigma_a + sigma_b)/2 and epsilon =
sqrt(epsilon_a epsilon_b) (``gcmc/python/simulation.py:63-72``).  The
reference has no gradients; this is new capability on the same energy.

Forward and backward run on the caller's current HIP stream; all tensors are
float32 on the GPU (energies are float64).  There is no CPU fallback.
"""

import ctypes

import torch

from . import engine as _engine


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p()


def _lib():
    L = _engine.lib()
    if not getattr(L, '_adjoint_types_set', False):
        P, I = ctypes.c_void_p, ctypes.c_int
        L.gcmc_engine_adjoint.argtypes = [P, I, I, I, P, P, P, P, P, P, P, P, P]
        L.gcmc_engine_adjoint_backward.argtypes = [P, I, I, I, P, P, P, P, P, P]
        L._adjoint_types_set = True
    return L


def _check_tensor(t, name, shape, dtype=torch.float32):
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError(f'{name} must be a GPU tensor')
    if t.dtype != dtype:
        raise TypeError(f'{name} must be {dtype}')
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f'{name} has shape {tuple(t.shape)}, expected {tuple(shape)}')
    return t.contiguous()


class PoseEnergy(torch.autograd.Function):
    """E[P] = energy(poses; eps, rmin, dq) with analytic gradients."""

    @staticmethod
    def forward(ctx, poses, eps, rmin, dq, eng, replica, species, centres):
        sysd = eng.systems[replica]
        Y = int(sysd['info'][0]['ffYNum'])
        nf = int(sysd['frag'][species]['num_atoms'])
        n = poses.shape[0]
        poses = _check_tensor(poses, 'poses', (n, nf, 3))
        centres = _check_tensor(centres, 'centres', (n, 3))
        eps = _check_tensor(eps, 'eps', (Y,))
        rmin = _check_tensor(rmin, 'rmin', (Y,))
        dq = _check_tensor(dq, 'dq', (Y,))
        dev = poses.device
        E = torch.empty(n, dtype=torch.float64, device=dev)
        dEdx = torch.empty((n, nf, 3), dtype=torch.float32, device=dev)
        partial = torch.empty((n, 3, Y), dtype=torch.float32, device=dev)
        stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
        _engine._check(_lib().gcmc_engine_adjoint(eng._h, replica, species, n, _ptr(centres), _ptr(poses),
                                                  _ptr(eps), _ptr(rmin), _ptr(dq), _ptr(E), _ptr(dEdx),
                                                  _ptr(partial), stream))
        ctx.save_for_backward(dEdx, partial)
        ctx.eng = eng
        ctx.Y = Y
        return E

    @staticmethod
    def backward(ctx, gE):
        dEdx, partial = ctx.saved_tensors
        n, nf, _ = dEdx.shape
        Y = ctx.Y
        w = gE.to(torch.float32).contiguous()
        gpose = torch.empty_like(dEdx)
        gcols = torch.empty(3 * Y, dtype=torch.float32, device=dEdx.device)
        stream = ctypes.c_void_p(torch.cuda.current_stream(dEdx.device).cuda_stream)
        _engine._check(_lib().gcmc_engine_adjoint_backward(ctx.eng._h, n, nf * 3, 3 * Y, _ptr(dEdx),
                                                           _ptr(partial), _ptr(w), _ptr(gpose), _ptr(gcols),
                                                           stream))
        gq, ge, gr = gcols.view(3, Y)
        return gpose, ge, gr, gq, None, None, None, None


def pose_energy(eng, replica, species, centres, poses, eps, rmin, dq=None):
    """Energies [P] (float64, kJ/mol) of poses [P, nf, 3]; differentiable in
    poses, eps (kcal/mol), rmin (Rmin/2, A) and dq (e)."""
    if dq is None:
        dq = torch.zeros_like(eps)
    return PoseEnergy.apply(poses, eps, rmin, dq, eng, replica, species, centres)


def type_parameters(sysd, device='cuda'):
    """Per-type (eps kcal/mol, Rmin/2 A) float32 tensors of a synthetic system,
    from the (sigma, epsilon) its pair table was built with."""
    from .dataset import gromacs_to_charmm
    eps, rmin = [], []
    for t in sysd['types2']:
        r, e = gromacs_to_charmm(*sysd['nb'][t])
        eps.append(e)
        rmin.append(r)
    return (torch.tensor(eps, dtype=torch.float32, device=device),
            torch.tensor(rmin, dtype=torch.float32, device=device))
