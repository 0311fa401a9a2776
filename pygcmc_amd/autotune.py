"""Autotuner of the batched engine's launch and data-layout parameters.

The engine exposes tuning knobs (``gcmc_tuning`` in ``include/pygcmc_hip.h``):
cell-list edge, threads per workgroup, number of persistent workgroups,
mobile-bucket capacity and whether the pair table is staged in LDS (-1: no).  ``autotune`` times a
short batched run of the caller's systems for each candidate setting -- on
copies of the arrays, so the caller's state is untouched -- and returns the
fastest setting; ``Engine(**best)`` then runs the real workload with it.
Results are cached per system signature (box edges, atom and group counts,
species, cutoff) in a JSON file when a path is given.

Every candidate runs the same moves from the same state.  A setting changes
the speed and the order in which a trial's FP32 pair terms are summed (cell
edges reorder the included groups, the workgroup shape the lanes), never the
exact FP32 inclusion test; every setting meets the same error bar against
the FP64 oracle.
"""

import hashlib
import itertools
import json
import os

import numpy as np

from .engine import Engine

DEFAULT_SPACE = {
    'cell_size': (5.0, 6.0, 7.5, 9.0),
    'wg_size': (256, 512),
    'persistent_wgs': (0,),
    'pair_lds': (-1, 0),
}


def signature(systems):
    """Hash of what the tuning depends on: sizes and shapes, not coordinates."""
    h = hashlib.sha1()
    for s in systems:
        info = s['info'][0]
        h.update(np.asarray(info['cryst'], np.float32).tobytes())
        h.update(np.int64([int(info['totalAtomNum']), int(info['totalResNum']), int(info['fragTypeNum'])]).tobytes())
        h.update(np.float32([float(info['cutoff'])]).tobytes())
        h.update(np.asarray(s['frag']['confBias'], np.int32).tobytes())
        h.update(np.asarray(s['frag']['num_atoms'], np.int32).tobytes())
    return h.hexdigest()


def _copies(systems):
    return [{k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in s.items()} for s in systems]


def time_setting(systems, setting, steps=60, warmup=20):
    """Milliseconds of `steps` batched moves after `warmup` moves, with the
    given tuning (copies of the systems)."""
    eng = Engine(**setting)
    try:
        for s in _copies(systems):
            eng.add_system(s)
        eng.upload()
        if warmup:
            eng.run(warmup, 'batched')
        eng.run(steps, 'batched')
        return eng.last_kernel_ms()
    finally:
        eng.close()


def autotune(systems, space=None, steps=60, warmup=20, cache_path=None, log=None):
    """Fastest tuning for these systems over `space` (dict of knob -> values).
    Returns (best_setting, table) where table lists (setting, ms) tried."""
    space = dict(DEFAULT_SPACE if space is None else space)
    key = signature(systems) + f':{steps}:{warmup}:' + json.dumps(space, sort_keys=True)
    cache = {}
    if cache_path and os.path.exists(cache_path):
        with open(cache_path) as f:
            cache = json.load(f)
        if key in cache:
            return cache[key]['best'], cache[key]['table']
    names = sorted(space)
    table = []
    for values in itertools.product(*(space[n] for n in names)):
        setting = dict(zip(names, values))
        try:
            ms = time_setting(systems, setting, steps, warmup)
        except RuntimeError as e:        # e.g. a stencil that does not fit in LDS
            if log:
                log(f'autotune: {setting} rejected: {e}')
            continue
        table.append((setting, ms))
        if log:
            log(f'autotune: {setting} {ms:.2f} ms')
    if not table:
        raise RuntimeError('autotune: no candidate setting ran')
    best = min(table, key=lambda t: t[1])[0]
    if cache_path:
        cache[key] = {'best': best, 'table': table}
        os.makedirs(os.path.dirname(os.path.abspath(cache_path)), exist_ok=True)
        with open(cache_path, 'w') as f:
            json.dump(cache, f, indent=1)
    return best, table
