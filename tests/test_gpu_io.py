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
