"""The autotuner picks a setting from its search space by timing short runs
on copies (the caller's arrays stay untouched) and caches the choice."""

import numpy as np
import pytest

from pygcmc_amd import autotune

from . import systems

pytestmark = pytest.mark.gpu


def test_autotune_picks_and_caches(tmp_path):
    base = [systems.multi(nsteps=100, n_fixed=n, box=b, seed=sd) for n, b, sd in
            ((1500, 36.0, 61), (2500, 45.0, 62))]
    before = [s['atoms'].copy() for s in base]
    space = {'cell_size': (5.0, 7.5), 'wg_size': (256,), 'persistent_wgs': (0,)}
    path = str(tmp_path / 'tune.json')
    best, table = autotune.autotune(base, space=space, steps=20, warmup=5, cache_path=path)
    assert best in [t for t, _ in table] and len(table) == 2
    assert all(ms > 0 for _, ms in table)
    for s, a in zip(base, before):
        assert np.array_equal(s['atoms'].view(np.uint8), a.view(np.uint8))
    best2, table2 = autotune.autotune(base, space=space, steps=20, warmup=5, cache_path=path)
    assert best2 == best and len(table2) == 2
