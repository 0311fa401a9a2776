"""Writes tests/golden/water3.json: the reference's own water fixture.

Inputs are the coordinates of test/water/3water.pdb (waters 1-2 = the
environment, water 3 = the inserted fragment; CRYST1 211.426 211.426 200.447)
and the TIP3P parameters recorded in test/jupyter/test2.ipynb cell 14
(sigma nm, epsilon kJ/mol).  Expected values: the FP64 evaluation of the
reference pair formula (gcmc/cpp/gcmc_energy.h:126-138) -- LJ 180.058466 and
Coulomb 3.884917 kJ/mol -- which the notebook cross-checks against OpenMM
(LJ-only difference 180.0606 kJ/mol, cells 11-13; total 183.9285, cells 7-9).
Only numbers are stored; nothing of the reference is copied.
"""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))

WATERS = [
    [[30.871, 40.491, 153.455], [31.521, 39.751, 153.345], [31.071, 40.991, 154.295]],
    [[29.868, 41.693, 155.535], [30.450, 42.453, 155.502], [29.231, 41.908, 156.216]],
    [[32.191, 40.158, 155.174], [32.897, 40.773, 155.372], [32.576, 39.543, 154.550]],
]

fixture = {
    'cryst': [211.426, 211.426, 200.447],
    'waters': WATERS,
    'charges': [-0.834, 0.417, 0.417],
    'sigma_nm': {'OT': 0.315057422683, 'HT': 0.0400013524445},
    'epsilon_kj': {'OT': 0.63639, 'HT': 0.192464},
    'expected_lj_kj': 180.058466,
    'expected_coulomb_kj': 3.884917,
    'openmm_lj_only_kj': 180.06058589721982,
    'openmm_total_kj': 183.928515355838,
}

if __name__ == '__main__':
    with open(os.path.join(HERE, 'water3.json'), 'w') as f:
        json.dump(fixture, f, indent=1)
