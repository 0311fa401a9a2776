"""Python front end of the HIP engine (``libpygcmc_hip.so``, C ABI in
``include/pygcmc_hip.h``).

``Engine`` holds any number of independent GCMC systems ("replicas") on one
GPU.  Each system is given as the seven arrays pyGCMC passes to
``gcmc.gpu.runGCMC`` (``gcmc/python/gcmc.py:41-49``); ``run`` advances every
system by a number of moves in one of three modes:

* ``'plain'``   -- one move of one system per launch pair (the execution
  shape of the reference, ``gcmc/cpp/gcmc.cu:77-98``, without its host
  round trips),
* ``'batched'`` -- every system in one persistent kernel,
* ``'graph'``   -- the batched launch replayed from a captured hipGraph.

There is no CPU fallback: a missing library raises ``ImportError`` and a
device failure raises ``RuntimeError``.
"""

import ctypes
import os

import numpy as np

from .values import AtomArray_dtype, Atom_dtype, Info_dtype, Residue_dtype

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'libpygcmc_hip.so')

MODES = {'plain': 0, 'batched': 1, 'graph': 2}

_lib = None


class Tuning(ctypes.Structure):
    _fields_ = [('cell_size', ctypes.c_float), ('persistent_wgs', ctypes.c_int),
                ('bucket_cap', ctypes.c_int), ('pair_lds', ctypes.c_int), ('wg_size', ctypes.c_int),
                ('fp32_coulomb', ctypes.c_int), ('claim_affinity', ctypes.c_int), ('xcd_round_robin', ctypes.c_int),
                ('xcd_search', ctypes.c_int)]


class CTrace(ctypes.Structure):
    _fields_ = [('max_steps', ctypes.c_int), ('max_k', ctypes.c_int)] + \
        [(n, ctypes.c_void_p) for n in ('k', 'ngroups', 'naccept', 'n_after', 'etot', 'centre', 'e',
                                        'ninc', 'inchash', 'posehash', 'idx', 'gpick', 'gacc', 'tick')]


TRACE_FIELDS = {
    'k': (np.int32, ()), 'ngroups': (np.int32, ()), 'naccept': (np.int32, ()),
    'n_after': (np.int32, ()), 'etot': (np.float64, ()),
    'centre': (np.float32, ('K', 3)), 'e': (np.float64, ('K', 2)), 'ninc': (np.int32, ('K', 2)),
    'inchash': (np.uint32, ('K', 2)), 'posehash': (np.uint32, ('K',)), 'idx': (np.int32, ('K',)),
    'gpick': (np.int32, ('K',)), 'gacc': (np.int32, ('K',)), 'tick': (np.uint64, (4,)),
}


def lib():
    """The loaded engine library (raises ImportError when it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f'{LIB_PATH} is missing: build it with `python -m pygcmc_amd.build`')
        # torch (when installed) brings its own HIP runtime; loading it first
        # makes this library bind to that same runtime, so tensors and the
        # engine can share a process in either import order
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        P, I = ctypes.c_void_p, ctypes.c_int
        L.gcmc_last_error.restype = ctypes.c_char_p
        L.gcmc_engine_create.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(Tuning)]
        L.gcmc_engine_destroy.argtypes = [P]
        L.gcmc_engine_add_system.argtypes = [P, P, P, P, P, P, P, P, I]
        L.gcmc_engine_upload.argtypes = [P]
        L.gcmc_engine_run.argtypes = [P, I, I]
        L.gcmc_engine_synchronize.argtypes = [P]
        L.gcmc_engine_step.argtypes = [P, I]
        L.gcmc_engine_download.argtypes = [P, I, P, P, P]
        L.gcmc_engine_stats.argtypes = [P, I, P, P, P]
        L.gcmc_engine_snapshot.argtypes = [P]
        L.gcmc_engine_restore.argtypes = [P]
        L.gcmc_engine_last_run_ms.argtypes = [P, P]
        L.gcmc_engine_last_kernel_ms.argtypes = [P, P]
        L.gcmc_engine_launch_info.argtypes = [P, P]
        L.gcmc_engine_set_ewald.argtypes = [P, I, ctypes.c_float, ctypes.c_float]
        L.gcmc_engine_ewald_potential.argtypes = [P, I, I, P, P]
        L.gcmc_engine_counters.argtypes = [P, I, P]
        L.gcmc_engine_phase_ticks.argtypes = [P, I, P]
        L.gcmc_engine_sub_ticks.argtypes = [P, I, P]
        L.gcmc_engine_set_trace.argtypes = [P, I, I, I]
        L.gcmc_engine_get_trace.argtypes = [P, I, ctypes.POINTER(CTrace)]
        L.gcmc_engine_eval_poses.argtypes = [P, I, I, I, P, P, P, P, P, P]
        L.gcmc_host_within.argtypes = [I, P, P, P, ctypes.c_float, P]
        L.gcmc_host_stencil_covers.argtypes = [I, P, P, P, P, ctypes.c_float, ctypes.c_float, P]
        L.gcmc_host_xcd_exchange.argtypes = [I, I, P, P, P, P, P]
        L.runGCMC_hip.argtypes = [P, P, P, P, P, P, P]
        _lib = L
    return _lib


def _check(rc):
    if rc < 0:
        raise RuntimeError(f'pygcmc_hip error {rc}: {lib().gcmc_last_error().decode()}')
    return rc


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _arr(a, dtype, name):
    a = np.asarray(a)
    if a.dtype != dtype or not a.flags['C_CONTIGUOUS']:
        raise TypeError(f'{name} must be a C-contiguous array of dtype {dtype}')
    return a


def host_within(a, b, cryst, cutoff):
    """The engine's FP32 inclusion test run on the host (same source as the GPU)."""
    a = np.ascontiguousarray(a, np.float32).reshape(-1, 3)
    b = np.ascontiguousarray(b, np.float32).reshape(-1, 3)
    c = np.ascontiguousarray(cryst, np.float32)
    out = np.zeros(len(a), np.uint8)
    _check(lib().gcmc_host_within(len(a), _p(a), _p(b), _p(c), float(cutoff), _p(out)))
    return out.astype(bool)


def host_stencil_covers(centres, groups, cryst, startxyz, cutoff, cell_size=0.0):
    """Whether each group's cell is scanned for the paired centre (host check)."""
    a = np.ascontiguousarray(centres, np.float32).reshape(-1, 3)
    b = np.ascontiguousarray(groups, np.float32).reshape(-1, 3)
    c = np.ascontiguousarray(cryst, np.float32)
    st = np.ascontiguousarray(startxyz, np.float32)
    out = np.zeros(len(a), np.uint8)
    _check(lib().gcmc_host_stencil_covers(len(a), _p(a), _p(b), _p(c), _p(st), float(cutoff),
                                          float(cell_size), _p(out)))
    return out.astype(bool)


def host_xcd_exchange(best, fin, evt):
    """One try of the XCD assignment search (host check): from the class of
    each system, the classes' finish times and the systems' evaluation times,
    the assignment the try runs and the number of exchanges in it."""
    b = np.ascontiguousarray(best, np.int32)
    f = np.ascontiguousarray(fin, np.float64)
    e = np.ascontiguousarray(evt, np.float64)
    if len(e) != len(b):
        raise ValueError('best and evt must have one entry per system')
    nxt = np.zeros(len(b), np.int32)
    nx = np.zeros(1, np.int32)
    _check(lib().gcmc_host_xcd_exchange(len(f), len(b), _p(b), _p(f), _p(e), _p(nxt), _p(nx)))
    return nxt, int(nx[0])


class Engine:
    """Independent GCMC systems on one GPU (see module docstring)."""

    def __init__(self, cell_size=0.0, persistent_wgs=0, bucket_cap=0, pair_lds=0, wg_size=0, fp32_coulomb=0,
                 claim_affinity=0, xcd_round_robin=0, xcd_search=0):
        self._h = ctypes.c_void_p()
        t = Tuning(cell_size, persistent_wgs, bucket_cap, pair_lds, wg_size, fp32_coulomb, claim_affinity,
                   xcd_round_robin, xcd_search)
        _check(lib().gcmc_engine_create(ctypes.byref(self._h), ctypes.byref(t)))
        self.systems = []
        self._traces = {}

    def close(self):
        if self._h:
            lib().gcmc_engine_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_system(self, sysd, nmoves=None):
        """Add a system given as a dict with keys info, frag, res, atoms, grid,
        ff, moves (the runGCMC arrays).  Returns its replica index."""
        info = _arr(sysd['info'], Info_dtype, 'info')
        frag = _arr(sysd['frag'], AtomArray_dtype, 'frag')
        res = _arr(sysd['res'], Residue_dtype, 'res')
        atoms = _arr(sysd['atoms'], Atom_dtype, 'atoms')
        grid = _arr(sysd['grid'], np.float32, 'grid')
        ff = _arr(sysd['ff'], np.float32, 'ff')
        moves = _arr(sysd['moves'], np.int32, 'moves')
        n = len(moves) if nmoves is None else int(nmoves)
        r = _check(lib().gcmc_engine_add_system(self._h, _p(info), _p(frag), _p(res), _p(atoms), _p(grid),
                                                _p(ff), _p(moves), n))
        self.systems.append(sysd)
        return r

    def set_ewald(self, r, beta=0.34, grid_spacing=1.0):
        """Ewald electrostatics for replica r (before upload): real-space
        erfc pair terms and PME reciprocal potential.  The batched and graph
        forms need every system of the engine in Ewald mode (or none)."""
        _check(lib().gcmc_engine_set_ewald(self._h, int(r), float(beta), float(grid_spacing)))

    def ewald_potential(self, r, xyz):
        """Reciprocal potential (kJ/mol/e) of replica r's current state at points [n, 3]."""
        xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
        phi = np.zeros(len(xyz), np.float32)
        _check(lib().gcmc_engine_ewald_potential(self._h, int(r), len(xyz), _p(xyz), _p(phi)))
        return phi

    def upload(self):
        _check(lib().gcmc_engine_upload(self._h))

    def run(self, nsteps, mode='batched'):
        """Advance every system by nsteps moves; returns GPU milliseconds."""
        _check(lib().gcmc_engine_run(self._h, int(nsteps), MODES[mode]))
        return self.last_run_ms()

    def synchronize(self):
        """Wait for everything queued on the engine's stream."""
        _check(lib().gcmc_engine_synchronize(self._h))

    def last_run_ms(self):
        ms = ctypes.c_double()
        _check(lib().gcmc_engine_last_run_ms(self._h, ctypes.byref(ms)))
        return ms.value

    def last_kernel_ms(self):
        """Duration of the last persistent kernel (batched mode), HIP events."""
        ms = ctypes.c_double()
        _check(lib().gcmc_engine_last_kernel_ms(self._h, ctypes.byref(ms)))
        return ms.value

    def launch_info(self):
        """Persistent-kernel shape: requested and resident workgroups, threads
        per workgroup, LDS bytes per workgroup."""
        out = np.zeros(4, np.int32)
        _check(lib().gcmc_engine_launch_info(self._h, _p(out)))
        return dict(zip(('wgs', 'resident', 'threads', 'lds_bytes'), (int(x) for x in out)))

    def counters(self, r=0):
        """Work since upload: evals, pairs, env atoms, groups, fragment atoms."""
        out = np.zeros(5, np.uint64)
        _check(lib().gcmc_engine_counters(self._h, r, _p(out)))
        return dict(zip(('evals', 'pairs', 'atoms', 'groups', 'frag_atoms'), (int(x) for x in out)))

    def phase_seconds(self, r=0):
        """Device time of a replica by phase (batched form), seconds."""
        out = np.zeros(4, np.uint64)
        _check(lib().gcmc_engine_phase_ticks(self._h, r, _p(out)))
        return dict(zip(('eval', 'accept', 'prep', 'step'), (float(x) * 1e-8 for x in out)))

    def eval_sub_seconds(self, r=0):
        """Time inside evaluations by sub-phase, seconds (summed over evaluations)."""
        out = np.zeros(4, np.uint64)
        _check(lib().gcmc_engine_sub_ticks(self._h, r, _p(out)))
        return dict(zip(('pose', 'stencil', 'candidates', 'pairs'), (float(x) * 1e-8 for x in out)))

    def step(self, r=0):
        return _check(lib().gcmc_engine_step(self._h, r))

    def download(self, r=0, sysd=None):
        """Write system r's state into sysd's arrays (default: the arrays it
        was added with), as runGCMC does in place."""
        sysd = self.systems[r] if sysd is None else sysd
        _check(lib().gcmc_engine_download(self._h, r, _p(sysd['frag']), _p(sysd['res']), _p(sysd['atoms'])))
        return sysd

    def stats(self, r=0):
        T = int(self.systems[r]['info'][0]['fragTypeNum'])
        e = ctypes.c_double()
        att = np.zeros((T, 4), np.int64)
        acc = np.zeros((T, 4), np.int64)
        _check(lib().gcmc_engine_stats(self._h, r, ctypes.byref(e), _p(att), _p(acc)))
        return {'energy': e.value, 'attempts': att, 'accepts': acc}

    def snapshot(self):
        """Keep a device copy of every system's current state (see restore)."""
        _check(lib().gcmc_engine_snapshot(self._h))

    def restore(self):
        """Return every system to the last snapshot (device-to-device copies
        queued on the engine's stream): step, species counts, positions, cell
        list, statistics and work counters."""
        _check(lib().gcmc_engine_restore(self._h))

    def set_trace(self, r, max_steps, max_k):
        _check(lib().gcmc_engine_set_trace(self._h, r, int(max_steps), int(max_k)))
        self._traces[r] = (int(max_steps), int(max_k))

    def get_trace(self, r=0):
        S, K = self._traces[r]
        out = {}
        for name, (dt, shape) in TRACE_FIELDS.items():
            out[name] = np.zeros((S,) + tuple(K if s == 'K' else s for s in shape), dtype=dt)
        ct = CTrace(S, K, *[out[n].ctypes.data for n in TRACE_FIELDS])
        _check(lib().gcmc_engine_get_trace(self._h, r, ctypes.byref(ct)))
        return out

    def eval_poses(self, r, species, centres, poses, exclude=None):
        """FP32/compensated energies of poses [n, nf, 3] around centres [n, 3]."""
        c = np.ascontiguousarray(centres, np.float32)
        p = np.ascontiguousarray(poses, np.float32)
        n = len(c)
        ex = np.full(n, -1, np.int32) if exclude is None else np.ascontiguousarray(exclude, np.int32)
        E = np.zeros(n, np.float64)
        N = np.zeros(n, np.int32)
        H = np.zeros(n, np.uint32)
        _check(lib().gcmc_engine_eval_poses(self._h, r, species, n, _p(c), _p(p), _p(ex), _p(E), _p(N), _p(H)))
        return E, N, H


def run_gcmc(info, frag, res, atoms, grid, ff, moves):
    """ctypes path to ``runGCMC_hip`` (same contract as ``pygcmc_amd.gpu.runGCMC``)."""
    _check(lib().runGCMC_hip(_p(info), _p(frag), _p(res), _p(atoms), _p(grid), _p(ff), _p(moves)))


def cavity_grid(xyz, cryst, grid_dx=1.0, cavity_bias=True):
    """Cavity-bias grid built on the GPU (``gcmc_cavity_grid``): same
    definition and result as ``dataset.get_grid(order='linear')``
    (``gcmc/python/dataset.py:21-139``).  Returns ``(startxyz, grid_n,
    grid float32[3G], cavity_factor)``."""
    from .dataset import box_start
    xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
    start, end = box_start(xyz, cryst)
    if not cavity_bias:
        grid_dx = 1.0
    n = [int((end[i] - start[i]) / grid_dx) for i in range(3)]
    L = lib()
    if not getattr(L, '_grid_types_set', False):
        L.gcmc_cavity_grid.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                                       ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        L._grid_types_set = True
    st = np.array(start, np.float64)
    nn = np.array(n, np.int32)
    cnt = ctypes.c_int()
    marks = xyz if cavity_bias else xyz[:0]
    _check(L.gcmc_cavity_grid(len(marks), _p(marks), _p(st), float(grid_dx), _p(nn), 0, None,
                              ctypes.byref(cnt)))
    grid = np.empty(3 * cnt.value, np.float32)
    _check(L.gcmc_cavity_grid(len(marks), _p(marks), _p(st), float(grid_dx), _p(nn), cnt.value,
                              _p(grid), ctypes.byref(cnt)))
    total = n[0] * n[1] * n[2]
    cav = cnt.value / total if cavity_bias else 1.0
    return start, n, grid, cav
