"""Caller-side set-up of a GCMC system: move schedule, cavity grid, fixed-atom
grouping, force-field pair table and packing into the struct ABI.

Each function restates one piece of pyGCMC's Python set-up with the same
arithmetic (FP64 from coordinates, then FP32 in the packed arrays):

* ``get_move``      -- ``gcmc/python/dataset.py:141-171``
* ``get_grid``      -- ``gcmc/python/dataset.py:21-139`` (the reference's
  set order native)
* ``set_fixcut``    -- ``gcmc/python/dataset.py:175-188`` (native)
* ``ff_pair_table`` -- ``gcmc/python/simulation.py:37-73``
* ``pack_system``   -- ``gcmc/python/dataset.py:190-344``

The packed arrays are exactly what ``gcmc.gpu.runGCMC`` receives in the
reference (``gcmc/python/gcmc.py:41-49``).
"""

import ctypes
import random as _random

import numpy as np

from .values import (AtomArray_dtype, Atom_dtype, Info_dtype, MAX_FRAGMENT_ATOMS,
                     MOLES_TO_MOLECULES, Residue_dtype)

# gcmc/python/base.py:36-38
ATTEMPT_PROB_FRAG = [0.300, 0.300, 0.200, 0.200]
ATTEMPT_PROB_WATER = [0.250, 0.250, 0.250, 0.250]


def get_move(fragment_names, mctime, mcsteps, rng=None,
             attempt_prob_frag=ATTEMPT_PROB_FRAG, attempt_prob_water=ATTEMPT_PROB_WATER):
    """Move schedule ``moveArray[i] = species*4 + kind`` (kind 0..3 = insert,
    delete, translate, rotate) drawn from the stdlib Mersenne Twister in the
    reference's call order (``gcmc/python/dataset.py:142-150``).

    Returns ``(moveArray int32[mcsteps], counts int[4*T])``.
    """
    rng = rng if rng is not None else _random
    n_species = len(fragment_names)
    moves = np.array(rng.choices(range(n_species), weights=mctime, k=mcsteps), dtype=np.int32)
    for i, n in enumerate(moves):
        probs = attempt_prob_water if fragment_names[n] == 'SOL' else attempt_prob_frag
        moves[i] = moves[i] * 4 + rng.choices(range(len(probs)), weights=probs, k=1)[0]
    counts = [0] * (4 * n_species)
    for m in moves:
        counts[m] += 1
    return moves, counts


def box_start(xyz, cryst):
    """``startxyz`` = centre of the bounding box of all atoms minus half the box
    (``gcmc/python/dataset.py:24-46``), FP64."""
    xyz = np.asarray(xyz, dtype=np.float64)
    lo = xyz.min(axis=0)
    hi = xyz.max(axis=0)
    centre = [(float(lo[i]) + float(hi[i])) / 2.0 for i in range(3)]
    start = [centre[i] - cryst[i] / 2.0 for i in range(3)]
    end = [centre[i] + cryst[i] / 2.0 for i in range(3)]
    return start, end


def get_grid(xyz, cryst, grid_dx=1.0, cavity_bias=True, order='linear'):
    """Cavity-bias grid (``gcmc/python/dataset.py:21-139``).

    ``xyz``: all atoms of the initial structure (FP64).  A cell
    ``int((x - start)/dx) % n`` per axis is occupied when any atom falls in it;
    the grid lists the corners of the free cells.  ``order='cpython'``
    reproduces the reference's CPython ``set`` iteration order; ``'linear'``
    lists cells with x slowest, z fastest (the order the GPU builder emits).

    Returns ``(startxyz, grid_n, grid float32[3G], cavity_factor)``.
    """
    start, end = box_start(xyz, cryst)
    if not cavity_bias:
        grid_dx = 1.0
    n = [int((end[i] - start[i]) / grid_dx) for i in range(3)]
    total = n[0] * n[1] * n[2]
    if cavity_bias:
        xyz = np.asarray(xyz, dtype=np.float64)
        idx = []
        for i in range(3):
            # Python int() truncates toward zero; % with n > 0 is non-negative
            q = np.trunc((xyz[:, i] - start[i]) / grid_dx).astype(np.int64)
            idx.append(np.mod(q, n[i]))
        occ = np.zeros(total, dtype=bool)
        occ[(idx[0] * n[1] + idx[1]) * n[2] + idx[2]] = True
        free_lin = np.flatnonzero(~occ)
        cavity_factor = len(free_lin) / total
    else:
        free_lin = np.arange(total, dtype=np.int64)
        cavity_factor = 1.0
    if order == 'cpython':
        coords = grid_set_order(n, occ if cavity_bias else None).astype(np.int64)
    elif order == 'linear':
        coords = np.stack([free_lin // (n[1] * n[2]), (free_lin // n[2]) % n[1],
                           free_lin % n[2]], axis=1)
    else:
        raise ValueError(f"unknown grid order {order!r}")
    grid = np.empty(coords.shape[0] * 3, dtype=np.float32)
    for i in range(3):
        grid[i::3] = coords[:, i] * grid_dx + start[i]
    return start, n, grid, cavity_factor


def set_fixcut(xyz_fixed, fix_cutoff=6.0):
    """Consecutive grouping of fixed atoms (``gcmc/python/dataset.py:175-188``):
    an atom opens a new group when it lies more than ``fix_cutoff`` from the
    first atom of the current group (no periodic images; squares as Python's
    float ``** 2``).  Native (``gcmc_fixed_groups``).  Returns int group
    index per atom."""
    from . import engine as _engine
    xyz = np.ascontiguousarray(xyz_fixed, dtype=np.float64).reshape(-1, 3)
    out = np.empty(len(xyz), dtype=np.int32)
    L = _engine.lib()
    L.gcmc_fixed_groups.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p]
    _engine._check(L.gcmc_fixed_groups(len(xyz), xyz.ctypes.data, float(fix_cutoff), out.ctypes.data))
    return out.astype(np.int64)


def grid_set_order(n, occupied=None):
    """Free cells ``int32[G, 3]`` of an ``n[0] x n[1] x n[2]`` grid in the
    iteration order of the CPython set the reference builds and discards the
    occupied cells from (``gcmc/python/dataset.py:66-109``); ``occupied``:
    bool[n0*n1*n2], x slowest, or None.  Native (``gcmc_grid_set_order``):
    the set's slot order from CPython's tuple hash and table rules, without
    Python objects."""
    from . import engine as _engine
    L = _engine.lib()
    P = ctypes.c_void_p
    L.gcmc_grid_set_order.argtypes = [P, P, ctypes.c_longlong, P, ctypes.POINTER(ctypes.c_longlong)]
    nn = np.ascontiguousarray(n, dtype=np.int32)
    occ = None
    if occupied is not None:
        occ = np.ascontiguousarray(occupied, dtype=np.uint8).ravel()
        if occ.size != int(nn[0]) * int(nn[1]) * int(nn[2]):
            raise ValueError('occupied must hold n0*n1*n2 cells')
    cnt = ctypes.c_longlong()
    occ_p = None if occ is None else occ.ctypes.data
    _engine._check(L.gcmc_grid_set_order(nn.ctypes.data, occ_p, 0, None, ctypes.byref(cnt)))
    cells = np.empty((cnt.value, 3), dtype=np.int32)
    _engine._check(L.gcmc_grid_set_order(nn.ctypes.data, occ_p, cnt.value, cells.ctypes.data, ctypes.byref(cnt)))
    return cells


def ff_pair_table(types1, types2, nb, nbfix=None):
    """LJ pair table (``gcmc/python/simulation.py:52-73``): for each fragment
    type ``t1`` (rows) and every type ``t2`` (columns) the NBFIX pair if present,
    else Lorentz-Berthelot ``sigma=(s1+s2)/2, eps=sqrt(e1 e2)``.  ``nb[t] =
    (sigma_nm, eps_kJ)``.  Returns float32[2*X*Y]."""
    nbfix = nbfix or {}
    out = np.empty(len(types1) * len(types2) * 2, dtype=np.float32)
    k = 0
    for t1 in types1:
        for t2 in types2:
            if (t1, t2) in nbfix:
                s, e = nbfix[(t1, t2)]
            elif (t2, t1) in nbfix:
                s, e = nbfix[(t2, t1)]
            else:
                s = (nb[t1][0] + nb[t2][0]) / 2
                e = np.sqrt(nb[t1][1] * nb[t2][1])
            out[2 * k] = s
            out[2 * k + 1] = e
            k += 1
    return out


def charmm_to_gromacs(eps_kcal, rmin_half):
    """CHARMM (epsilon kcal/mol, Rmin/2 Angstrom) -> GROMACS (sigma nm, epsilon
    kJ/mol): sigma = Rmin/2 * 2^(5/6) / 10, eps = 4.184 * |epsilon|."""
    return rmin_half * 2.0 ** (5.0 / 6.0) / 10.0, 4.184 * abs(eps_kcal)


def gromacs_to_charmm(sigma_nm, eps_kj):
    """Inverse of ``charmm_to_gromacs``: (Rmin/2 in A, epsilon in kcal/mol)."""
    return sigma_nm * 10.0 / 2.0 ** (5.0 / 6.0), eps_kj / 4.184


def max_num(conc, volume, n_insert, n_initial):
    """``maxNum = max(conc*V*6.023e-4*2 + n_insert, N0 + n_insert)`` truncated to
    int (``gcmc/python/dataset.py:210-213``)."""
    m1 = conc * volume * MOLES_TO_MOLECULES * 2 + n_insert
    m2 = n_initial + n_insert
    return int(max(m1, m2))


def pack_system(*, cryst, cutoff, grid_dx, startxyz, grid, cavity_factor,
                fixed_xyz, fixed_charge, fixed_type, fixed_group,
                species, ff, ffX, ffY, moves, move_counts, mcsteps, seed,
                show_info=0, pme=0):
    """Pack a system into the seven arrays of ``gcmc.gpu.runGCMC``
    (``gcmc/python/dataset.py:190-344``).

    ``species``: list of dicts with ``name, conc, muex, conf, mctime,
    template_xyz (nf,3 FP64), charge (nf), type (nf), molecules`` (list of
    (nf,3) FP64 coordinates of the initial copies).
    """
    volume = cryst[0] * cryst[1] * cryst[2]
    T = len(species)
    frag = np.zeros(T, dtype=AtomArray_dtype)
    for i, sp in enumerate(species):
        nf = len(sp['charge'])
        if nf > MAX_FRAGMENT_ATOMS:
            raise ValueError(f"species {sp['name']} has {nf} > {MAX_FRAGMENT_ATOMS} atoms")
        frag[i]['name'] = np.array(sp['name'].ljust(4)[:4], dtype='S4')
        frag[i]['muex'] = sp['muex']
        frag[i]['conc'] = sp['conc']
        frag[i]['confBias'] = sp['conf']
        frag[i]['mcTime'] = sp['mctime']
        frag[i]['totalNum'] = len(sp['molecules'])
        frag[i]['maxNum'] = max_num(sp['conc'], volume, move_counts[i * 4], len(sp['molecules']))
        frag[i]['num_atoms'] = nf
        t = np.asarray(sp['template_xyz'], dtype=np.float64)
        centre = np.zeros(3)
        for row in t:  # sequential FP64 mean, gcmc/python/dataset.py:219-228
            centre += row
        centre /= nf
        t = t - centre
        frag[i]['atoms'][:nf]['position'] = t
        frag[i]['atoms'][:nf]['charge'] = sp['charge']
        frag[i]['atoms'][:nf]['type'] = sp['type']

    fixed_xyz = np.asarray(fixed_xyz, dtype=np.float64)
    n_fixed = len(fixed_xyz)
    n_fixed_groups = int(fixed_group[-1]) + 1 if n_fixed else 0
    total_res = n_fixed_groups + int(frag['maxNum'].sum())
    total_atoms = n_fixed + int((frag['maxNum'].astype(np.int64) * frag['num_atoms']).sum())
    res = np.zeros(total_res, dtype=Residue_dtype)
    atoms = np.zeros(total_atoms, dtype=Atom_dtype)

    if n_fixed:
        atoms[:n_fixed]['position'] = fixed_xyz
        atoms[:n_fixed]['charge'] = fixed_charge
        atoms[:n_fixed]['type'] = fixed_type
        g = np.asarray(fixed_group, dtype=np.int64)
        counts = np.bincount(g, minlength=n_fixed_groups)
        starts = np.concatenate([[0], np.cumsum(counts)[:-1]])
        # sums accumulate in float32 storage as in the reference
        # (residueInfo[...]['position'][0] += atom.x on a float32 field, the
        # add itself in FP64), atom by atom within each group
        pos = np.zeros((n_fixed_groups, 3), dtype=np.float32)
        for k in range(int(counts.max())):
            sel = np.flatnonzero(counts > k)
            pos[sel] = (pos[sel].astype(np.float64) + fixed_xyz[starts[sel] + k]).astype(np.float32)
        pos = (pos.astype(np.float64) / counts[:, None]).astype(np.float32)
        res[:n_fixed_groups]['position'] = pos
        res[:n_fixed_groups]['atomNum'] = counts
        res[:n_fixed_groups]['atomStart'] = starts

    seq = n_fixed_groups
    atom_cursor = n_fixed
    for i, sp in enumerate(species):
        frag[i]['startRes'] = seq
        nf = int(frag[i]['num_atoms'])
        m = int(frag[i]['maxNum'])
        slots = np.arange(m)
        res[seq:seq + m]['atomNum'] = nf
        res[seq:seq + m]['atomStart'] = atom_cursor + slots * nf
        n0 = int(frag[i]['totalNum'])
        if n0:
            mol = np.asarray(sp['molecules'], dtype=np.float64).reshape(n0, nf, 3)
            p = np.zeros((n0, 3), dtype=np.float32)
            for k in range(nf):  # FP64 add into FP32 storage, atom by atom
                p = (p.astype(np.float64) + mol[:, k]).astype(np.float32)
            idx = (atom_cursor + slots[:n0, None] * nf + np.arange(nf)[None, :]).ravel()
            atoms['position'][idx] = mol.reshape(-1, 3)
            atoms['charge'][idx] = np.tile(np.asarray(sp['charge'], np.float32), n0)
            atoms['type'][idx] = np.tile(np.asarray(sp['type'], np.int32), n0)
            res[seq:seq + n0]['position'] = (p.astype(np.float64) / nf).astype(np.float32)
        seq += m
        atom_cursor += m * nf

    info = np.zeros(1, dtype=Info_dtype)
    info[0]['ffXNum'] = ffX
    info[0]['ffYNum'] = ffY
    info[0]['mcsteps'] = mcsteps
    info[0]['cutoff'] = cutoff
    info[0]['grid_dx'] = grid_dx
    info[0]['startxyz'] = startxyz
    info[0]['cryst'] = cryst
    info[0]['cavityFactor'] = cavity_factor
    info[0]['fragTypeNum'] = T
    info[0]['totalGridNum'] = len(grid) // 3
    info[0]['totalResNum'] = total_res
    info[0]['totalAtomNum'] = total_atoms
    info[0]['showInfo'] = show_info
    info[0]['PME'] = pme
    info[0]['seed'] = seed
    return info, frag, res, atoms, np.ascontiguousarray(grid, dtype=np.float32), \
        np.ascontiguousarray(ff, dtype=np.float32), np.ascontiguousarray(moves, dtype=np.int32)
