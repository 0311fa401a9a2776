"""gcmc_adjoint_backward_noalloc (fixed per-device workspace, no allocation)
against gcmc_engine_adjoint_backward (engine-allocated workspace), through
the C ABI: bit-identical outputs on two calls in a row, the second smaller."""

import ctypes

import numpy as np
import pytest

from pygcmc_amd import engine as _engine
from pygcmc_amd.engine import Engine

pytestmark = pytest.mark.gpu

P, I = ctypes.c_void_p, ctypes.c_int


def _lib():
    L = _engine.lib()
    L.gcmc_engine_adjoint_backward.argtypes = [P, I, I, I, P, P, P, P, P, P]
    L.gcmc_adjoint_backward_noalloc.argtypes = [I, I, I, P, P, P, P, P, P]
    return L


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _case(torch, L, eng, npose, per_pose, ncol, seed, weighted):
    g = torch.Generator(device='cpu').manual_seed(seed)
    dEdx = torch.randn(npose, per_pose, generator=g).cuda()
    partial = (torch.randn(npose, ncol, generator=g) * 10).cuda()
    w = torch.rand(npose, generator=g).cuda() if weighted else None
    out = []
    for noalloc in (False, True):
        gp = torch.full((npose, per_pose), float('nan'), device='cuda')
        gc = torch.full((ncol,), float('nan'), device='cuda')
        stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        if noalloc:
            rc = L.gcmc_adjoint_backward_noalloc(npose, per_pose, ncol, _ptr(dEdx), _ptr(partial), _ptr(w),
                                                 _ptr(gp), _ptr(gc), stream)
        else:
            rc = L.gcmc_engine_adjoint_backward(eng._h, npose, per_pose, ncol, _ptr(dEdx), _ptr(partial), _ptr(w),
                                                _ptr(gp), _ptr(gc), stream)
        _engine._check(rc)
        torch.cuda.synchronize()
        out.append((gp.cpu().numpy(), gc.cpu().numpy()))
    (gp0, gc0), (gp1, gc1) = out
    assert np.array_equal(gp0.view(np.uint32), gp1.view(np.uint32))
    assert np.array_equal(gc0.view(np.uint32), gc1.view(np.uint32))
    # and both are the FP64 column sums of the weighted rows
    wv = w.cpu().double().numpy() if weighted else np.ones(npose)
    ref = (wv[:, None] * partial.cpu().double().numpy()).sum(0)
    assert np.allclose(gc1, ref, rtol=1e-6, atol=1e-5)
    assert np.array_equal(gp1, (wv[:, None].astype(np.float32) * dEdx.cpu().numpy()).astype(np.float32))


def test_noalloc_backward_matches_engine_backward():
    torch = pytest.importorskip('torch')
    L = _lib()
    eng = Engine()
    try:
        print('noalloc backward: 4096 poses x 36, 3 x 1365 columns, weighted', flush=True)
        _case(torch, L, eng, 4096, 36, 4095, seed=1, weighted=True)
        print('noalloc backward: 300 poses x 9, 198 columns, unweighted', flush=True)
        _case(torch, L, eng, 300, 9, 198, seed=2, weighted=False)
    finally:
        eng.close()


def test_noalloc_backward_rejects_too_many_columns():
    pytest.importorskip('torch')
    L = _lib()
    # refused on the host before anything is launched
    rc = L.gcmc_adjoint_backward_noalloc(1, 3, 4097, None, None, None, None, None, None)
    assert rc == -1
    assert 'exceeds the fixed workspace' in L.gcmc_last_error().decode()
