"""Host-side logic of the engine (no GPU): the C ABI exports, the division-free
FP32 group test against the oracle's literal formula, the cell-stencil
superset property, set-up restatements and argument checking."""

import ctypes
import os
import re

import numpy as np
import pytest

from oracle import oracle as O
from pygcmc_amd import engine, synth
from pygcmc_amd.values import AtomArray_dtype, Atom_dtype, Info_dtype, Residue_dtype

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, 'include', 'pygcmc_hip.h')).read()
    names = re.findall(r'^\s*(?:int|const char \*|void)\s+\*?(\w+)\(', hdr, re.M)
    assert 'runGCMC_hip' in names and len(names) >= 15
    lib = ctypes.CDLL(engine.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n


def test_cpython_module_signature():
    from pygcmc_amd import gpu
    assert callable(gpu.runGCMC)
    with pytest.raises(TypeError):
        gpu.runGCMC(1, 2, 3)


def test_cpython_module_checks_arrays():
    from pygcmc_amd import gpu
    s = synth.water_sphere(mcsteps=10)
    with pytest.raises(TypeError):   # wrong item size
        gpu.runGCMC(s['info'], s['res'], s['res'], s['atoms'], s['grid'], s['ff'], s['moves'])
    with pytest.raises(ValueError):  # too few residues for the InfoStruct
        gpu.runGCMC(s['info'], s['frag'], s['res'][:10], s['atoms'], s['grid'], s['ff'], s['moves'])


def _boundary_points(rng, L, rc, n):
    """Pairs concentrated where the image shift or the cutoff test flips."""
    L = np.asarray(L, np.float64)
    a = rng.uniform(-0.3, 1.3, (n, 3)) * L
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    kind = rng.integers(0, 3, n)
    r = np.where(kind == 0, rc * (1 + rng.normal(0, 1e-6, n)), rng.uniform(0, 1.6 * L.max(), n))
    b = a + d * r[:, None]
    # put some axis separations right at half a box
    m = kind == 1
    b[m, 0] = a[m, 0] + L[0] * 0.5 * (1 + rng.normal(0, 1e-6, m.sum()))
    return a.astype(np.float32), b.astype(np.float32)


@pytest.mark.parametrize('L', [[40.0, 40.0, 40.0], [132.126, 132.213, 122.07], [54.31, 61.7, 80.0]])
def test_division_free_group_test_matches_literal(L):
    rng = np.random.default_rng(17)
    rc = 15.0
    a, b = _boundary_points(rng, L, rc, 20000)
    got = engine.host_within(a, b, L, rc)
    ref = np.array([O.distanceP32(a[i], b[i], L) <= np.float32(rc) for i in range(len(a))])
    assert np.array_equal(got, ref)
    assert 0 < ref.sum() < len(ref)


@pytest.mark.parametrize('cell', [0.0, 3.5, 5.0, 7.5, 16.0])
def test_stencil_covers_every_included_group(cell):
    rng = np.random.default_rng(23)
    L = [54.31, 61.7, 80.0]
    start = [-1.25, 0.5, 3.0]
    rc = 15.0
    a, b = _boundary_points(rng, L, rc, 20000)
    inc = engine.host_within(a, b, L, rc)
    cov = engine.host_stencil_covers(a, b, L, start, rc, cell)
    assert cov[inc].all()
    assert (~cov).any()  # the stencil really prunes


def test_stencil_small_box_wraps_once():
    rng = np.random.default_rng(4)
    L = [30.0, 30.0, 30.0]
    a, b = _boundary_points(rng, L, 15.0, 5000)
    cov = engine.host_stencil_covers(a, b, L, [0, 0, 0], 15.0, 5.0)
    inc = engine.host_within(a, b, L, 15.0)
    assert cov[inc].all()


def test_engine_rejects_bad_systems():
    if not engine.os.path.exists(engine.LIB_PATH):
        pytest.skip('library not built')
    s = synth.water_sphere(mcsteps=5)
    bad = O.copy_system(s)
    bad['frag'][0]['confBias'] = 5000
    try:
        eng = engine.Engine()
    except RuntimeError as e:  # no GPU in this container: creation itself reports it
        assert 'no HIP device' in str(e)
        return
    with pytest.raises(RuntimeError, match='confBias'):
        eng.add_system(bad)
    eng.close()


def test_dtypes_match_struct_sizes():
    assert (Atom_dtype.itemsize, AtomArray_dtype.itemsize, Info_dtype.itemsize,
            Residue_dtype.itemsize) == (20, 436, 76, 20)


def test_replica_sizes_and_boxes():
    sizes = synth.replica_sizes(64)
    assert len(sizes) == 64 and sizes.min() >= 5000 and sizes.max() <= 60000
    assert len(set(sizes.tolist())) == 64


def test_fragment_templates_have_config_atom_counts():
    counts = {name: len(b()[1]) for name, b, _, _ in synth.FRAGMENTS}
    assert counts == {'BENX': 12, 'PRPX': 11, 'MEOH': 6, 'FORM': 6, 'ALDX': 7, 'IMIA': 9,
                      'MAMY': 8, 'ACEY': 7}
    rng = np.random.default_rng(0)
    for name, b, net, _ in synth.FRAGMENTS:
        q = synth.fragment_charges(rng, len(b()[1]), net)
        assert abs(q.sum() - net) < 1e-12


@pytest.mark.parametrize('cell', [5.0, 6.0, 7.5])
def test_pruned_stencil_covers_cutoff_pairs_at_cell_faces(cell):
    """Centres on and next to cell faces, partners at the cutoff along axis,
    face-diagonal and body-diagonal directions: the stencil cells kept for
    the centre (pruned by their box's distance from it) hold every partner
    that passes the group test."""
    rng = np.random.default_rng(31)
    L = np.array([54.31, 80.0, 124.07])
    start = np.array([-1.25, 0.5, 3.0])
    rc = 15.0
    nc = np.maximum(1, np.floor(L / cell)).astype(int)
    h = L / nc
    fr = np.array([0.0, 1e-5, 0.25, 0.5, 1.0 - 1e-5])
    dirs = [np.array(v, np.float64) for v in
            [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1), (1, 1, 0), (-1, 1, 0),
             (1, 0, -1), (0, -1, -1), (1, 1, 1), (-1, -1, 1), (1, -1, -1), (-1, 1, -1)]]
    a, b = [], []
    for _ in range(60):
        c = rng.integers(0, nc)
        for f in rng.choice(fr, (6, 3)):
            p = start + (c + f) * h
            for d in dirs:
                d = d / np.linalg.norm(d)
                for s in (1 - 1e-6, 1.0, 1 + 1e-6):
                    a.append(p)
                    b.append(p + d * rc * s)
    a = np.array(a, np.float32)
    b = np.array(b, np.float32)
    inc = engine.host_within(a, b, L, rc)
    cov = engine.host_stencil_covers(a, b, L, start, rc, cell)
    assert inc.sum() > 1000
    assert cov[inc].all()


def _lpt(w, ncls):
    order = sorted(range(len(w)), key=lambda r: -w[r])
    load, cls = [0.0] * ncls, [0] * len(w)
    for r in order:
        c = min(range(ncls), key=lambda q: (load[q], q))
        load[c] += w[r]
        cls[r] = c
    return np.array(cls, np.int32)


@pytest.mark.parametrize('seed', [1, 2, 3])
def test_xcd_exchange_try(seed):
    """One try of the XCD assignment search (host code of the engine): the
    i-th latest class trades one system with the i-th earliest, class sizes
    stay, and each trade is the pair that, finishing times taken as
    proportional to summed evaluation times, ends the later of the two
    classes earliest -- and earlier than the slower class ended."""
    rng = np.random.default_rng(seed)
    ncls, n = 8, 64
    evt = rng.uniform(5.0, 20.0, n)
    best = _lpt(list(rng.uniform(0.5, 1.5, n) * evt), ncls)
    # finish times: proportional to the class's evaluation time with a
    # class-dependent rate (as measured: the rate varies between XCDs)
    rate = rng.uniform(9.0, 11.0, ncls)
    fin = np.array([rate[c] * evt[best == c].sum() for c in range(ncls)])
    nxt, nx = engine.host_xcd_exchange(best, fin, evt)
    assert np.array_equal(np.bincount(nxt, minlength=ncls), np.bincount(best, minlength=ncls))
    order = sorted(range(ncls), key=lambda c: -fin[c])
    partner = {}
    for i in range(ncls // 2):
        partner[order[i]] = order[ncls - 1 - i]
        partner[order[ncls - 1 - i]] = order[i]
    moved = np.nonzero(nxt != best)[0]
    assert len(moved) == 2 * nx and nx >= 1
    for r in moved:
        assert nxt[r] == partner[best[r]]
    for i in range(ncls // 2):
        S, F = order[i], order[ncls - 1 - i]
        kS, kF = fin[S] / evt[best == S].sum(), fin[F] / evt[best == F].sum()
        cand = [(max(fin[S] - kS * (evt[a] - evt[b]), fin[F] + kF * (evt[a] - evt[b])), a, b)
                for a in np.nonzero(best == S)[0] for b in np.nonzero(best == F)[0] if evt[a] > evt[b]]
        good = [c for c in cand if c[0] < fin[S]]
        a_moved = [r for r in moved if best[r] == S]
        if not good:
            assert not a_moved
            continue
        pred, a, b = min(good, key=lambda c: c[0])
        assert a_moved == [a] and nxt[a] == F and nxt[b] == S


def test_xcd_exchange_leaves_balanced_and_single_class_alone():
    evt = np.linspace(1.0, 8.0, 16)
    best = np.arange(16, dtype=np.int32) % 4
    nxt, nx = engine.host_xcd_exchange(best, np.full(4, 100.0), evt)
    assert nx == 0 and np.array_equal(nxt, best)
    nxt, nx = engine.host_xcd_exchange(np.zeros(16, np.int32), np.array([50.0]), evt)
    assert nx == 0
    with pytest.raises(RuntimeError):
        engine.host_xcd_exchange(np.full(16, 4, np.int32), np.full(4, 1.0), evt)
