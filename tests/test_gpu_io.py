"""A system set up from files (PDB + TOP + force-field directory) run on the
GPU: same final arrays as the FP64 oracle on the same set-up, and the
pygcmc-style command line end to end."""

import os
import shutil

import numpy as np
import pytest

from pygcmc_amd import formats

from .test_io import KEYS, MINI, mini_system

pytestmark = pytest.mark.gpu


def test_files_to_gpu_run_matches_oracle():
    from oracle import oracle as O
    g = mini_system('top', mcsteps=400, conf=(4, 2))
    o = mini_system('top', mcsteps=400, conf=(4, 2))
    g.run()
    o.run(runner=lambda *a: O.run(dict(zip(KEYS, a))))
    for k, x, y in zip(KEYS, g.arrays, o.arrays):
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), k
    assert g.output_pdb() == o.output_pdb()
    assert g.output_top() == o.output_top()


def test_command_line(tmp_path):
    from pygcmc_amd import cli
    for f in ('mini.pdb', 'mini.top'):
        shutil.copy(os.path.join(MINI, f), tmp_path / f)
    shutil.copytree(os.path.join(MINI, 'charmm36.ff'), tmp_path / 'charmm36.ff')
    rc = cli.main(['-p', str(tmp_path / 'mini.pdb'), '-t', str(tmp_path / 'mini.top'), '-o', str(tmp_path / 'log.txt'),
                   '-n', '300', '-e', '5', '-f', '3', '--ff-dir', str(tmp_path / 'charmm36.ff'),
                   '--top-out', str(tmp_path / 'out.top'),
                   '--grid-order', 'linear', '--fragments', 'MEOH,SOL', '-u=-5.36,-5.6', '-m', '1.0,1.0',
                   '-c', '0.25,55'])
    assert rc == 0
    out = formats.read_pdb(str(tmp_path / 'mini_out.pdb'))
    log = open(tmp_path / 'log.txt').read()
    assert 'final:' in log and 'GPU simulation time' in log
    # every listed molecule is in the output structure
    atoms, _ = formats.parse_top(open(tmp_path / 'out.top').read(), str(tmp_path))
    assert len(atoms) == len(out)


def test_legacy_command_on_gpu_matches_oracle_runs(tmp_path):
    """gcmc -p params.inp (scripts/mainOld.py:26-144) on the HIP path with
    map_generation: the output PDB and TOP equal those of the same command
    with the FP64 oracle as the runner, byte for byte, and each species'
    DX map equals the map built from the oracle's chain in the same chunks
    (host histogram of each snapshot, the same free-energy formula)."""
    from oracle import oracle as O
    from pygcmc_amd import legacy
    from pygcmc_amd.maps import write_dx
    from pygcmc_amd.values import BETA, MOLES_TO_MOLECULES
    from .test_gpu_maps import host_counts
    from .test_io import FF
    for f in ('mini.pdb', 'mini.top'):
        shutil.copy(os.path.join(MINI, f), tmp_path / f)
    runs = {}
    for name, maps in (('gpu', 'yes'), ('orc', 'no')):
        inp = tmp_path / f'{name}.inp'
        inp.write_text('\n'.join([
            f'pdb:{tmp_path / "mini.pdb"}', f'top:{tmp_path / "mini.top"}',
            f'op_pdb:{tmp_path / (name + "_out.pdb")}', f'op_top:{tmp_path / (name + "_out.top")}',
            'grid_dx:   1.000', 'use_cavity_bias:yes', 'random_seed:166', f'map_generation:{maps}',
            f'map_filename_prefix:{tmp_path / name}',
            'fragname:     meoh      sol ', 'fragconc:     0.25    55.00 ', 'fragmuex:    -2.04    -5.04 ',
            'fragconf:       10        2 ', 'mctime:      0.2      0.8 ', 'mcsteps:300', '']))
        runner = None if name == 'gpu' else (lambda *a: O.run(dict(zip(KEYS, a))))
        rc = legacy.main(['-p', str(inp), '--ff-dir', FF, '--logfile', str(tmp_path / f'{name}.log')], runner=runner)
        assert rc == 0
        runs[name] = ((tmp_path / f'{name}_out.pdb').read_bytes(), (tmp_path / f'{name}_out.top').read_bytes())
    assert runs['gpu'] == runs['orc']
    # the maps: the oracle's chain in chunks of mcsteps // 100 moves
    from pygcmc_amd.system import GCMCSystem
    s = GCMCSystem(fragment_names=['MEOH', 'SOL'], conc=[0.25, 55.0], muex=[-2.04, -5.04], conf=[10, 2],
                   mctime=[0.2, 0.8], mcsteps=300, seed=166, ff_dir=FF)
    s.load(str(tmp_path / 'mini.pdb'), top=str(tmp_path / 'mini.top')).setup()
    sysd = dict(zip(KEYS, s.arrays))
    info = sysd['info'][0]
    origin = np.array([float(v) for v in info['startxyz']])
    spacing = float(info['grid_dx'])
    shape = tuple(max(1, int(float(info['cryst'][i]) / spacing)) for i in range(3))
    every = max(1, 300 // 100)
    acc = [np.zeros(shape, np.int64) for _ in range(2)]
    done = 0
    samples = 0
    while done < 300:
        n = min(every, 300 - done)
        O.run(sysd, first=done, nsteps=n)
        done += n
        samples += 1
        for sp in range(2):
            acc[sp] += host_counts(sysd, sp, origin, spacing, shape)
    for sp, name in enumerate(('meoh', 'sol')):
        F = sysd['frag'][sp]
        bulk = samples * int(F['num_atoms']) * float(F['conc']) * MOLES_TO_MOLECULES * spacing ** 3
        occ = acc[sp].astype(np.float64)
        with np.errstate(divide='ignore'):
            g = -np.log(occ / bulk) / BETA
        g[occ == 0] = 12.552
        ref = tmp_path / f'ref.{name}.gfe.dx'
        write_dx(str(ref), g, origin + 0.5 * spacing, spacing)
        assert (tmp_path / f'gpu.{name}.gfe.dx').read_text() == ref.read_text(), name
        assert acc[sp].sum() > 0 or sp == 0
