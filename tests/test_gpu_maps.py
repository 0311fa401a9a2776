"""Occupancy maps on the GPU: counts equal a host histogram of the
downloaded state, cell for cell, and add up over snapshots."""

import numpy as np
import pytest

from . import systems

pytestmark = pytest.mark.gpu


def host_counts(sysd, species, origin, spacing, shape):
    F = sysd['frag'][species]
    nf, n = int(F['num_atoms']), int(F['totalNum'])
    r0 = int(F['startRes'])
    a0 = int(sysd['res'][r0]['atomStart'])
    xyz = sysd['atoms']['position'][a0:a0 + n * nf].astype(np.float64)
    idx = []
    for i in range(3):
        q = np.trunc((xyz[:, i] - origin[i]) / spacing).astype(np.int64)
        idx.append(np.mod(q, shape[i]))
    out = np.zeros(shape, np.int64)
    np.add.at(out, tuple(idx), 1)
    return out


def test_occupancy_matches_host_histogram():
    from pygcmc_amd.engine import Engine
    from pygcmc_amd.maps import OccupancyMap
    s = systems.multi(nsteps=600)
    with Engine() as eng:
        eng.add_system(systems.copy(s))
        eng.upload()
        T = int(s['info'][0]['fragTypeNum'])
        maps = [OccupancyMap(eng, 0, sp, spacing=1.5) for sp in range(T)]
        expect = [np.zeros(m.shape, np.int64) for m in maps]
        for chunk in range(3):
            eng.run(200, 'batched')
            for m in maps:
                m.accumulate()
            st = eng.download(0, systems.copy(s))
            for sp, m in enumerate(maps):
                expect[sp] += host_counts(st, sp, m.origin, m.spacing, m.shape)
        for sp, m in enumerate(maps):
            c = m.counts()
            assert np.array_equal(c, expect[sp]), sp
        total = sum(int(m.counts().sum()) for m in maps)
        assert total > 0
        g = maps[-1].free_energy(cap=10.0)
        assert np.isfinite(g).all()


def test_run_maps_same_chain_and_files(tmp_path):
    """GCMCSystem.run_maps (engine, chunks, snapshots) ends in the same arrays
    as the one-call run, and writes one map per species."""
    from pygcmc_amd.maps import read_dx
    from .test_io import KEYS, mini_system
    a = mini_system('top', mcsteps=300, conf=(4, 2))
    b = mini_system('top', mcsteps=300, conf=(4, 2))
    a.run()
    maps = b.run_maps(every=50, spacing=1.0, prefix=str(tmp_path / 'maps' / 'mini'))
    for k, x, y in zip(KEYS, a.arrays, b.arrays):
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), k
    assert [m.samples for m in maps] == [6, 6]
    for name in ('meoh', 'sol'):
        v, origin, spacing = read_dx(str(tmp_path / 'maps' / f'mini.{name}.gfe.dx'))
        assert v.shape == maps[0].shape and spacing == 1.0
