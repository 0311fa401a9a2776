"""Occupancy maps: free-energy conversion and the OpenDX writer on the host
(the GPU counts are checked in test_gpu_maps.py)."""

import numpy as np

from pygcmc_amd import maps
from pygcmc_amd.values import BETA, MOLES_TO_MOLECULES


class _Fake(maps.OccupancyMap):
    def __init__(self, counts, samples, nf, conc, spacing):
        self._c = np.asarray(counts)
        self.samples, self.nf, self.conc, self.spacing = samples, nf, conc, spacing
        self.shape = self._c.shape
        self.origin = np.zeros(3)

    def counts(self):
        return self._c


def test_free_energy_closed_form():
    # bulk occupancy per cell: samples * nf * conc * 6.023e-4 * dx^3
    bulk = 10 * 3 * 55.0 * MOLES_TO_MOLECULES * 1.0
    c = np.zeros((2, 2, 2), np.int64)
    c[0, 0, 0] = round(bulk)
    c[1, 1, 1] = round(2 * bulk)
    m = _Fake(c, 10, 3, 55.0, 1.0)
    g = m.free_energy(cap=7.0)
    assert abs(g[0, 0, 0] + np.log(round(bulk) / bulk) / BETA) < 1e-12
    assert abs(g[1, 1, 1] - g[0, 0, 0] + np.log(round(2 * bulk) / round(bulk)) / BETA) < 1e-12
    assert g[0, 1, 0] == 7.0
    assert np.isinf(m.free_energy()[0, 1, 0])


def test_dx_round_trip(tmp_path):
    v = np.arange(2 * 3 * 4, dtype=np.float64).reshape(2, 3, 4) * 0.5
    maps.write_dx(str(tmp_path / 'm.dx'), v, [1.0, 2.0, 3.0], 0.5)
    back, origin, spacing = maps.read_dx(str(tmp_path / 'm.dx'))
    assert np.array_equal(back, v) and list(origin) == [1.0, 2.0, 3.0] and spacing == 0.5
    text = (tmp_path / 'm.dx').read_text().split('\n')
    assert text[0] == 'object 1 class gridpositions counts 2 3 4'
    assert text[6] == 'object 3 class array type double rank 0 items 24 data follows'
