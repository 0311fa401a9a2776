"""Write a small GCMC system in pyGCMC's input formats (tests/golden/io/mini):
two alanine-like residues (12 atoms) fixed in a 24 A box with 3 methanols
and 8 TIP3P waters, as PDB + GROMACS .top and PDB + CHARMM .psf, and a
minimal force-field directory (pygcmc_amd.synth.write_forcefield)."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))
from pygcmc_amd import synth  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'io', 'mini')


def main():
    synth.write_forcefield(os.path.join(OUT, 'charmm36.ff'))
    pep = [(an, ty, q, 'ALA', 1) for an, ty, q in synth.FILE_ALA] + \
          [('N', 'NH1', -0.47, 'GLY', 2), ('HN', 'H', 0.31, 'GLY', 2)]
    fixed = [p + ((6.0 + 1.2 * i, 12.0, 12.0 + 0.4 * (i % 2)),) for i, p in enumerate(pep)]
    sites = [[x, y, z] for x in (4.0, 10.0, 16.0, 21.0) for y in (4.0, 19.0) for z in (5.0, 18.0)]
    np.random.RandomState(7).shuffle(sites)
    synth.write_system_files(OUT, 'mini', (24.0, 24.0, 24.0), fixed,
                             {'MEOH': sites[:3], 'SOL': sites[3:11]})


if __name__ == '__main__':
    main()
