"""Smooth particle-mesh Ewald, reciprocal part, on the GPU (``gcmc_pme_*`` in
``include/pygcmc_hip.h``): order-4 B-spline spread, hipFFT, influence
function, inverse FFT and gather.  The reference declares a PME flag
(``gcmc/cpp/gcmc.h:91``) that no kernel reads; this supplies the Ewald
reciprocal energy and potential for an orthorhombic box.  Tensors are torch
GPU tensors (device memory); work runs on the current stream."""

import ctypes

import torch

from . import engine as _engine


def _lib():
    L = _engine.lib()
    if not getattr(L, '_pme_types_set', False):
        P, I = ctypes.c_void_p, ctypes.c_int
        L.gcmc_pme_create.argtypes = [ctypes.POINTER(P), P, P, ctypes.c_float]
        L.gcmc_pme_destroy.argtypes = [P]
        L.gcmc_pme_clear.argtypes = [P, P]
        L.gcmc_pme_spread.argtypes = [P, I, P, P]
        L.gcmc_pme_solve.argtypes = [P, P, P]
        L.gcmc_pme_gather.argtypes = [P, I, P, P, P]
        L._pme_types_set = True
    return L


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _dev(t, cols, name):
    if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != torch.float32:
        raise TypeError(f'{name} must be a float32 GPU tensor')
    if t.dim() != 2 or t.shape[1] != cols:
        raise ValueError(f'{name} must have shape [n, {cols}]')
    return t.contiguous()


class PME:
    """Reciprocal-space Ewald on a grid[3] mesh over box[3] (A) with
    splitting parameter beta (1/A)."""

    def __init__(self, grid, box, beta):
        import numpy as np
        self._h = ctypes.c_void_p()
        g = np.ascontiguousarray(grid, np.int32)
        b = np.ascontiguousarray(box, np.float32)
        _engine._check(_lib().gcmc_pme_create(ctypes.byref(self._h), g.ctypes.data_as(ctypes.c_void_p),
                                              b.ctypes.data_as(ctypes.c_void_p), float(beta)))

    def close(self):
        if self._h:
            _lib().gcmc_pme_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def clear(self):
        _engine._check(_lib().gcmc_pme_clear(self._h, _stream()))

    def spread(self, xyzq):
        """Adds charges (x, y, z, q) [n, 4] to the charge grid."""
        xyzq = _dev(xyzq, 4, 'xyzq')
        _engine._check(_lib().gcmc_pme_spread(self._h, xyzq.shape[0], ctypes.c_void_p(xyzq.data_ptr()),
                                              _stream()))

    def solve(self):
        """Reciprocal energy of the charge grid (kJ/mol, float64 GPU tensor);
        also prepares the potential for gather."""
        e = torch.zeros(1, dtype=torch.float64, device='cuda')
        _engine._check(_lib().gcmc_pme_solve(self._h, ctypes.c_void_p(e.data_ptr()), _stream()))
        return e

    def gather(self, xyz):
        """Reciprocal potential (kJ/mol/e) at points [n, 3]."""
        xyz = _dev(xyz, 3, 'xyz')
        phi = torch.empty(xyz.shape[0], dtype=torch.float32, device='cuda')
        _engine._check(_lib().gcmc_pme_gather(self._h, xyz.shape[0], ctypes.c_void_p(xyz.data_ptr()),
                                              ctypes.c_void_p(phi.data_ptr()), _stream()))
        return phi

    def energy(self, xyzq):
        self.clear()
        self.spread(xyzq)
        return float(self.solve().item())
