"""Occupancy maps of a GCMC run: per-species atom counts on a 3D grid,
accumulated on the GPU from snapshots of the engine's state, converted to
grid free energies and written as OpenDX.

pyGCMC parses ``map_generation`` / ``map_filename_prefix`` and ignores them
(``gcmc/scripts/mainOld.py``); this is the consumer those keys name.  The
grid defaults to the system box split into cells of ``spacing`` starting at
``startxyz`` -- with spacing = grid_dx, the cavity grid's cells.
"""

import ctypes

import numpy as np
import torch

from . import engine as _engine
from .values import BETA, MOLES_TO_MOLECULES


def _lib():
    L = _engine.lib()
    if not getattr(L, '_maps_types_set', False):
        P = ctypes.c_void_p
        L.gcmc_engine_occupancy.argtypes = [P, ctypes.c_int, ctypes.c_int, P, ctypes.c_double, P, P]
        L._maps_types_set = True
    return L


class OccupancyMap:
    """Counts of one species' atoms per grid cell, summed over snapshots.

    ``accumulate()`` adds the current state (queued on the engine's stream,
    after the runs before it); call it between ``eng.run`` chunks."""

    def __init__(self, eng, replica, species, spacing=1.0, origin=None, shape=None):
        info = eng.systems[replica]['info'][0]
        frag = eng.systems[replica]['frag'][species]
        if origin is None:
            origin = [float(v) for v in info['startxyz']]
        if shape is None:
            shape = [max(1, int(float(info['cryst'][i]) / spacing)) for i in range(3)]
        self.eng, self.replica, self.species = eng, replica, species
        self.origin = np.ascontiguousarray(origin, np.float64)
        self.spacing = float(spacing)
        self.shape = tuple(int(v) for v in shape)
        self.n = np.ascontiguousarray(self.shape, np.int32)
        self.nf = int(frag['num_atoms'])
        self.conc = float(frag['conc'])
        self.samples = 0
        self._counts = torch.zeros(int(np.prod(self.shape)), dtype=torch.int32, device='cuda')
        # the zero-fill runs on torch's stream, the counting kernel on the
        # engine's: wait for the fill before any accumulate() is queued
        torch.cuda.current_stream().synchronize()
        self._host = None

    def accumulate(self):
        if self._host is not None:
            raise RuntimeError('map was frozen')
        _engine._check(_lib().gcmc_engine_occupancy(
            self.eng._h, self.replica, self.species, self.origin.ctypes.data, self.spacing, self.n.ctypes.data,
            ctypes.c_void_p(self._counts.data_ptr())))
        self.samples += 1

    def counts(self):
        """int64[n0, n1, n2] counts (synchronizes the engine)."""
        if self._host is not None:
            return self._host
        self.eng.synchronize()
        return self._counts.cpu().numpy().astype(np.int64).reshape(self.shape)

    def freeze(self):
        """Keep the counts on the host and release the engine and the device
        buffer (before the engine is closed)."""
        self._host = self.counts()
        self.eng, self._counts = None, None

    def reset(self):
        self.eng.synchronize()
        self._counts.zero_()
        torch.cuda.current_stream().synchronize()
        self.samples = 0

    def free_energy(self, cap=None):
        """Grid free energy -kT ln(occupancy / bulk) in kJ/mol (kT = 1/beta of
        the engine, 300 K): bulk = samples x atoms per molecule x conc x
        6.023e-4 molecules/A^3/M x cell volume.  Empty cells get ``cap``
        (default +inf)."""
        occ = self.counts().astype(np.float64)
        bulk = self.samples * self.nf * self.conc * MOLES_TO_MOLECULES * self.spacing ** 3
        with np.errstate(divide='ignore'):
            g = -np.log(occ / bulk) / BETA
        g[occ == 0] = np.inf if cap is None else cap
        return g

    def write_dx(self, path, values=None):
        """OpenDX scalar field of values (default: the counts) at the cell
        centres, z fastest."""
        write_dx(path, self.counts() if values is None else values, self.origin + 0.5 * self.spacing, self.spacing)


def write_dx(path, values, origin, spacing):
    v = np.asarray(values, np.float64)
    n = v.shape
    finite = np.where(np.isfinite(v), v, np.finfo(np.float64).max)
    with open(path, 'w') as f:
        f.write('object 1 class gridpositions counts %d %d %d\n' % n)
        f.write('origin %.6f %.6f %.6f\n' % tuple(origin))
        f.write('delta %.6f 0 0\ndelta 0 %.6f 0\ndelta 0 0 %.6f\n' % (spacing, spacing, spacing))
        f.write('object 2 class gridconnections counts %d %d %d\n' % n)
        f.write('object 3 class array type double rank 0 items %d data follows\n' % v.size)
        flat = finite.ravel()
        for i in range(0, flat.size, 3):
            f.write(' '.join('%.6g' % x for x in flat[i:i + 3]) + '\n')
        f.write('attribute "dep" string "positions"\n')
        f.write('object "occupancy" class field\ncomponent "positions" value 1\n'
                'component "connections" value 2\ncomponent "data" value 3\n')


def read_dx(path):
    """(values[n0, n1, n2], origin[3], spacing) of a file written by write_dx."""
    with open(path) as f:
        lines = f.read().split('\n')
    n = tuple(int(v) for v in lines[0].split()[-3:])
    origin = np.array([float(v) for v in lines[1].split()[1:4]])
    spacing = float(lines[2].split()[1])
    vals = []
    for ln in lines[7:]:
        if ln.startswith('attribute'):
            break
        vals += [float(v) for v in ln.split()]
    return np.array(vals).reshape(n), origin, spacing
