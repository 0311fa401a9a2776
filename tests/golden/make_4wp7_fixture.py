"""Writes tests/golden/run4wp7.json: numbers printed in the reference's only
committed run log, test/4wp7/gcmc.75.out, for its inputs test/4wp7/gcmc.75.inp.

* move schedule counts per species and move kind (gcmc.75.out:73-81) for
  random_seed 166, mcsteps 100, mctime 0.025 x 8 + 0.8;
* maxNum of BENX (gcmc.75.out:102, 287 initial, 2 inserts) and SOL (:110,
  57077 initial, 16 inserts) at volume 2132413.34447466 (:58);
* the cavity factor 1914440 free points of 132 x 132 x 122 (:88-90).
"""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))

fixture = {
    'seed': 166,
    'mcsteps': 100,
    'names': ['BENX', 'PRPX', 'DMEE', 'MEOH', 'FORM', 'IMIA', 'ACEY', 'MAMY', 'SOL'],
    'mctime': [0.025] * 8 + [0.8],
    'counts': {'BENX': [2, 1, 0, 0], 'PRPX': [0, 2, 0, 0], 'DMEE': [2, 0, 0, 0],
               'MEOH': [0, 0, 1, 0], 'FORM': [0, 0, 0, 1], 'IMIA': [0, 0, 0, 1],
               'ACEY': [2, 0, 1, 1], 'MAMY': [1, 1, 1, 0], 'SOL': [16, 32, 13, 22]},
    'volume': 2132413.34447466,
    'maxnum': {'BENX': {'conc': 0.25, 'n0': 287, 'ins': 2, 'max': 644},
               'SOL': {'conc': 55.0, 'n0': 57077, 'ins': 16, 'max': 141294}},
    'grid_n': [132, 132, 122],
    'free_points': 1914440,
    'cavity_factor': 0.9006044047027654,
}

if __name__ == '__main__':
    with open(os.path.join(HERE, 'run4wp7.json'), 'w') as f:
        json.dump(fixture, f, indent=1)
