"""Native set-up (host code in libpygcmc_hip.so, no device): the cavity grid
in the reference's CPython set order and the fixed-atom grouping, each
against the literal Python construction the reference runs
(gcmc/python/dataset.py:66-109 and 175-188)."""

import numpy as np
import pytest

from pygcmc_amd import dataset as D


def _python_set_order(n, occupied=None):
    # the reference's construction: comprehension, discards, iteration
    s = {(x, y, z) for x in range(n[0]) for y in range(n[1]) for z in range(n[2])}
    if occupied is not None:
        for c in np.flatnonzero(occupied).tolist():
            s.discard((c // (n[1] * n[2]), (c // n[2]) % n[1], c % n[2]))
    return np.array(list(s), dtype=np.int64).reshape(-1, 3)


# shapes below and above the 50,000-entry growth rule, flat and ragged
@pytest.mark.parametrize('n', [(1, 1, 1), (1, 1, 9), (3, 4, 5), (17, 23, 9), (2, 150, 3), (40, 41, 42),
                               (64, 70, 55)])
def test_grid_set_order_matches_python_set(n):
    rng = np.random.default_rng(sum(n))
    occ = rng.random(n[0] * n[1] * n[2]) < 0.35
    assert np.array_equal(D.grid_set_order(n, occ), _python_set_order(n, occ))
    assert np.array_equal(D.grid_set_order(n), _python_set_order(n))


def test_grid_set_order_edges():
    n = (6, 5, 4)
    full = np.ones(120, bool)
    assert D.grid_set_order(n, full).shape == (0, 3)
    one = np.ones(120, bool)
    one[77] = False
    assert D.grid_set_order(n, one).tolist() == [[77 // 20, (77 // 4) % 5, 77 % 4]]
    with pytest.raises(ValueError):
        D.grid_set_order(n, np.zeros(119, bool))
    with pytest.raises(RuntimeError):
        D.grid_set_order((0, 5, 4))
    with pytest.raises(RuntimeError):
        D.grid_set_order((1024, 1024, 1024))


def _reference_get_grid(xyz, cryst, grid_dx):
    # gcmc/python/dataset.py:21-109, atom by atom in Python floats
    x_values, y_values, z_values = (xyz[:, i].tolist() for i in range(3))
    center = [(min(v) + max(v)) / 2.0 for v in (x_values, y_values, z_values)]
    start = [center[i] - cryst[i] / 2.0 for i in range(3)]
    end = [center[i] + cryst[i] / 2.0 for i in range(3)]
    n = [int((end[i] - start[i]) / grid_dx) for i in range(3)]
    grid = {(x, y, z) for x in range(n[0]) for y in range(n[1]) for z in range(n[2])}
    for ax, ay, az in xyz.tolist():
        grid.discard((int((ax - start[0]) / grid_dx) % n[0], int((ay - start[1]) / grid_dx) % n[1],
                      int((az - start[2]) / grid_dx) % n[2]))
    out = np.zeros(len(grid) * 3, dtype=np.float32)
    for i, (x, y, z) in enumerate(grid):
        out[i * 3] = x * grid_dx + start[0]
        out[i * 3 + 1] = y * grid_dx + start[1]
        out[i * 3 + 2] = z * grid_dx + start[2]
    return start, n, out, len(grid) / (n[0] * n[1] * n[2])


@pytest.mark.parametrize('dx', [1.0, 0.7])
def test_get_grid_cpython_order_matches_reference_construction(dx):
    rng = np.random.default_rng(5)
    cryst = [31.0, 27.5, 24.25]
    xyz = rng.random((3000, 3)) * np.array(cryst) - 3.0
    xyz[:40] += 40.0   # atoms outside the box wrap by the remainder
    start, n, grid, cav = D.get_grid(xyz, cryst, dx, True, order='cpython')
    rs, rn, rgrid, rcav = _reference_get_grid(xyz, cryst, dx)
    assert list(n) == rn and cav == rcav and np.allclose(start, rs, rtol=0, atol=0)
    assert np.array_equal(grid.view(np.uint32), rgrid.view(np.uint32))


def _reference_groups(xyz, cutoff, square):
    x0 = y0 = z0 = float('inf')
    k, out = -1, []
    for x, y, z in xyz.tolist():
        if np.sqrt(square(x - x0) + square(y - y0) + square(z - z0)) > cutoff:
            k += 1
            x0, y0, z0 = x, y, z
        out.append(k)
    return np.array(out)


def test_fixed_groups_match_python_power():
    rng = np.random.default_rng(11)
    cut = 6.0
    # a chain of 1.5 A steps: protein-like groups of several atoms
    walk = np.cumsum(rng.normal(size=(6000, 3)) * 0.87, axis=0)
    # displacements within a few ulps of the cutoff for which Python's x ** 2
    # (the C library's pow) and x * x decide the comparison differently, each
    # placed after a far spacer atom and the origin (so it is measured from
    # the origin, exactly)
    d = rng.normal(size=(300000, 3))
    d *= cut / np.sqrt((d ** 2).sum(axis=1))[:, None]
    d *= (1.0 + rng.integers(-3, 4, size=len(d)) * 2.0 ** -52)[:, None]
    flips = [v for v in d.tolist()
             if (np.sqrt(v[0] ** 2 + v[1] ** 2 + v[2] ** 2) > cut)
             != (np.sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) > cut)]
    assert len(flips) >= 5
    tail = []
    for v in flips:
        tail += [[1000.0, 1000.0, 1000.0], [0.0, 0.0, 0.0], v]
    xyz = np.concatenate([walk, np.array(tail)])
    ref = _reference_groups(xyz, cut, lambda v: v ** 2)
    mult = _reference_groups(xyz, cut, lambda v: v * v)
    assert not np.array_equal(ref, mult)
    got = D.set_fixcut(xyz, cut)
    assert np.array_equal(got, ref)
    assert got[:len(walk)].max() + 1 < len(walk) // 3
