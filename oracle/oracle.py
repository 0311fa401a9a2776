"""ctypes front end of the CPU oracle (oracle/gcmc_oracle.c).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker of the HIP path.  The product
package (pygcmc_amd) never imports this module.
"""

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'liboracle.so')

_lib = None


def build(force=False):
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(LIB_PATH) or \
            os.path.getmtime(LIB_PATH) < os.path.getmtime(os.path.join(HERE, 'gcmc_oracle.c')):
        subprocess.check_call(['make', '-s', '-C', HERE, '-B' if force else 'all'])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        L.orc_philox4x32_10.argtypes = [P, P, P]
        L.orc_u01f.argtypes = [ctypes.c_uint32]
        L.orc_u01f.restype = ctypes.c_float
        L.orc_u01d.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
        L.orc_u01d.restype = ctypes.c_double
        L.orc_below.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
        L.orc_below.restype = ctypes.c_uint32
        L.orc_sincosf.argtypes = [ctypes.c_float, P, P]
        L.orc_distanceP32.argtypes = [P, P, P]
        L.orc_distanceP32.restype = ctypes.c_float
        L.orc_rotmat.argtypes = [P, ctypes.c_float, P]
        L.orc_energy_pose.argtypes = [P, P, P, P, P, ctypes.c_int, P, P, ctypes.c_int, P, P]
        L.orc_energy_pose.restype = ctypes.c_double
        L.orc_run_c.argtypes = [P, P, P, P, P, P, P, ctypes.c_int, ctypes.c_uint32, ctypes.c_int,
                                ctypes.c_int, P, P, P, ctypes.c_int, ctypes.c_int]
        L.orc_run_c.restype = ctypes.c_int
        L.orc_picks.argtypes = [ctypes.c_uint32, ctypes.c_int, ctypes.c_int, ctypes.c_int, P]
        L.orc_species_B.argtypes = [P, P]
        L.orc_species_B.restype = ctypes.c_double
        L.orc_group_hash_term.argtypes = [ctypes.c_uint32]
        L.orc_group_hash_term.restype = ctypes.c_uint32
        L.orc_pose_hash.argtypes = [ctypes.c_int, P]
        L.orc_pose_hash.restype = ctypes.c_uint32
        L.orc_adjoint.argtypes = [P, P, P, P, ctypes.c_int, P, P, ctypes.c_int, P, P, P, P, P, P, P]
        L.orc_adjoint.restype = ctypes.c_double
        L.orc_trial_pose.argtypes = [P, P, P, P, P, ctypes.c_uint32, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_int, P, P]
        L.orc_trial_pose.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def philox(ctr, key):
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(_p(c), _p(k), _p(o))
    return o


def sincosf(x):
    s = ctypes.c_float()
    c = ctypes.c_float()
    lib().orc_sincosf(ctypes.c_float(x), ctypes.byref(s), ctypes.byref(c))
    return s.value, c.value


def distanceP32(a, b, L):
    a = np.asarray(a, np.float32); b = np.asarray(b, np.float32); L = np.asarray(L, np.float32)
    return lib().orc_distanceP32(_p(a), _p(b), _p(L))


def rotmat(axis, phi):
    a = np.asarray(axis, np.float32)
    R = np.zeros(9, np.float32)
    lib().orc_rotmat(_p(a), ctypes.c_float(phi), _p(R))
    return R.reshape(3, 3)


def picks(seed, step, N, n):
    out = np.zeros(max(n, 1), np.int32)
    lib().orc_picks(seed, step, N, n, _p(out))
    return out[:n]


def species_B(sysd, s):
    return lib().orc_species_B(_p(sysd['info']), _p(sysd['frag'][s:s + 1]))


def type_params(sysd):
    """Per-type CHARMM epsilon (kcal/mol) and Rmin/2 (A) over the all-types
    list of a synthetic system (the parameters its pair table was built from)."""
    from pygcmc_amd.dataset import gromacs_to_charmm
    eps, rmin = [], []
    for t in sysd['types2']:
        r, e = gromacs_to_charmm(*sysd['nb'][t])
        eps.append(e)
        rmin.append(r)
    return np.array(eps, np.float64), np.array(rmin, np.float64)


def adjoint(sysd, species, centre, pose, eps, rmin, dq=None, exclude=-1):
    """FP64 energy and gradients of one pose with per-type parameters:
    returns (E, dE/dx [nf,3], dE/ddq [Y], dE/deps [Y], dE/drmin [Y])."""
    Y = int(sysd['info'][0]['ffYNum'])
    nf = int(sysd['frag'][species]['num_atoms'])
    c = np.ascontiguousarray(centre, np.float32)
    p = np.ascontiguousarray(pose, np.float32)
    eps = np.ascontiguousarray(eps, np.float64)
    rmin = np.ascontiguousarray(rmin, np.float64)
    dq = np.zeros(Y) if dq is None else np.ascontiguousarray(dq, np.float64)
    gx = np.zeros((nf, 3))
    gq, ge, gr = np.zeros(Y), np.zeros(Y), np.zeros(Y)
    e = lib().orc_adjoint(_p(sysd['info']), _p(sysd['frag']), _p(sysd['res']), _p(sysd['atoms']),
                          species, _p(c), _p(p), exclude, _p(eps), _p(rmin), _p(dq),
                          _p(gx), _p(gq), _p(ge), _p(gr))
    return e, gx, gq, ge, gr


def trial_pose(sysd, step, t, species, kind, slot=-1):
    """(centre[3], pose[nf, 3]) of trial t of move `step` for the current state:
    kind 0 inserts (randomFragment), 2 translates and 3 rotates slot `slot`."""
    nf = int(sysd['frag'][species]['num_atoms'])
    c = np.zeros(3, np.float32)
    p = np.zeros((nf, 3), np.float32)
    rc = lib().orc_trial_pose(_p(sysd['info']), _p(sysd['frag']), _p(sysd['res']), _p(sysd['atoms']),
                              _p(sysd['grid']), int(sysd['info'][0]['seed']), step, t, species, kind, slot,
                              _p(c), _p(p))
    if rc != 0:
        raise ValueError(f'trial_pose refused ({rc})')
    return c, p


def energy_pose(sysd, species, centre, pose, exclude=-1):
    """FP64 energy of a fragment pose around `centre` against the current state."""
    c = np.ascontiguousarray(centre, np.float32)
    p = np.ascontiguousarray(pose, np.float32)
    n = ctypes.c_int()
    h = ctypes.c_uint32()
    e = lib().orc_energy_pose(_p(sysd['info']), _p(sysd['frag']), _p(sysd['res']),
                              _p(sysd['atoms']), _p(sysd['ff']), species, _p(c), _p(p),
                              exclude, ctypes.byref(n), ctypes.byref(h))
    return e, n.value, h.value


class Trace(ctypes.Structure):
    _fields_ = [('max_steps', ctypes.c_int), ('max_k', ctypes.c_int)] + \
        [(n, ctypes.c_void_p) for n in ('k', 'ngroups', 'naccept', 'n_after', 'etot', 'centre', 'e',
                                        'ninc', 'inchash', 'posehash', 'idx', 'gpick', 'gacc')]


class Stats(ctypes.Structure):
    _fields_ = [('e_tot', ctypes.c_double), ('attempts', (ctypes.c_longlong * 4) * 64),
                ('accepts', (ctypes.c_longlong * 4) * 64), ('evals', ctypes.c_longlong)]


TRACE_FIELDS = {
    'k': (np.int32, ()), 'ngroups': (np.int32, ()), 'naccept': (np.int32, ()),
    'n_after': (np.int32, ()), 'etot': (np.float64, ()),
    'centre': (np.float32, ('K', 3)), 'e': (np.float64, ('K', 2)), 'ninc': (np.int32, ('K', 2)),
    'inchash': (np.uint32, ('K', 2)), 'posehash': (np.uint32, ('K',)), 'idx': (np.int32, ('K',)),
    'gpick': (np.int32, ('K',)), 'gacc': (np.int32, ('K',)),
}


def new_trace(nsteps, max_k):
    """numpy arrays of a step trace (the layout the HIP engine also writes)."""
    out = {}
    for name, (dt, shape) in TRACE_FIELDS.items():
        shp = (nsteps,) + tuple(max_k if s == 'K' else s for s in shape)
        fill = -1 if name in ('gpick', 'idx') else 0
        out[name] = np.full(shp, fill, dtype=dt)
    return out


def _trace_struct(tr):
    nsteps = tr['k'].shape[0]
    max_k = tr['centre'].shape[1]
    s = Trace(nsteps, max_k)
    for name in TRACE_FIELDS:
        setattr(s, name, tr[name].ctypes.data)
    return s


def copy_system(sysd):
    return {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in sysd.items()}


RUN_ERRORS = {
    -1: 'steps [first, first + nsteps) outside the move schedule',
    -2: 'trial buffers could not be allocated',
    -3: 'a move code names a species outside fragTypeNum',
    -4: 'a trial count or fragment atom count exceeds the oracle buffers',
    -5: 'the followed trace is shorter or narrower than the run',
    -6: 'a species count outside [0, maxNum]',
    -7: 'the cell list could not be allocated',
}


def run(sysd, first=0, nsteps=None, trace=None, follow=None, nthreads=1, stats=None, cells=False):
    """Run the FP64 chain in place on sysd's arrays.  Returns (trace, stats).

    cells=True scans a cell list of group centres instead of every group (a
    CPU timing variant with the same inclusion sets; energies are summed in
    another order).  Raises ValueError, before changing any state, when the
    steps leave the move schedule, a move code or species count is invalid,
    or a trace is too small for the run."""
    moves = sysd['moves']
    if not (isinstance(moves, np.ndarray) and moves.dtype == np.int32 and moves.flags['C_CONTIGUOUS']):
        raise TypeError('moves must be a C-contiguous int32 array')
    nsteps = int(sysd['info'][0]['mcsteps']) - first if nsteps is None else int(nsteps)
    first = int(first)
    for name, tr in (('trace', trace), ('follow', follow)):
        if tr is not None and tr['k'].shape[0] < nsteps:
            raise ValueError(f'{name} holds {tr["k"].shape[0]} steps, the run needs {nsteps}')
    st = stats if stats is not None else Stats()
    ts = _trace_struct(trace) if trace is not None else None
    fs = _trace_struct(follow) if follow is not None else None
    rc = lib().orc_run_c(_p(sysd['info']), _p(sysd['frag']), _p(sysd['res']), _p(sysd['atoms']),
                         _p(sysd['grid']), _p(sysd['ff']), _p(moves), len(moves),
                         int(sysd['info'][0]['seed']), first, nsteps,
                         ctypes.byref(ts) if ts is not None else None,
                         ctypes.byref(fs) if fs is not None else None,
                         ctypes.byref(st), nthreads, 1 if cells else 0)
    if rc != 0:
        msg = RUN_ERRORS.get(rc, 'unknown error')
        raise ValueError(f'oracle run refused ({rc}): {msg} (first={first}, nsteps={nsteps}, '
                         f'schedule={len(moves)} moves)')
    return trace, st
