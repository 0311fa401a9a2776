"""Struct ABI shared with the HIP engine, and the physical constants of the path.

The numpy dtypes are the byte layout of the C structs in ``include/pygcmc_hip.h``
and of pyGCMC's own structs (``gcmc/python/values.py:18-63`` mirrors
``gcmc/cpp/gcmc.h:47-100``), so arrays built for the reference can be handed to
``pygcmc_amd.gpu.runGCMC`` unchanged.
"""

import numpy as np

# gcmc/python/values.py:16, gcmc/cpp/gcmc.h:38
MOLES_TO_MOLECULES = 0.0006023
# gcmc/cpp/gcmc.h:28,40-41: beta = 1 / (k_B T) in mol/kJ
BOLTZMANN = 0.0083115
TEMPERATURE = 300.0
BETA = 1.0 / (BOLTZMANN * TEMPERATURE)
# gcmc/cpp/gcmc_energy.h:137, kJ/mol * Angstrom / e^2
COULOMB = 1388.431112
KCAL_TO_KJ = 4.184
# AtomArray.atoms[20] (gcmc/cpp/gcmc.h:68)
MAX_FRAGMENT_ATOMS = 20

Atom_dtype = np.dtype([
    ('position', np.float32, (3,)),
    ('charge', np.float32),
    ('type', np.int32),
])

AtomArray_dtype = np.dtype([
    ('name', 'S4'),
    ('startRes', np.int32),
    ('muex', np.float32),
    ('conc', np.float32),
    ('confBias', np.int32),
    ('mcTime', np.float32),
    ('totalNum', np.int32),
    ('maxNum', np.int32),
    ('num_atoms', np.int32),
    ('atoms', Atom_dtype, (MAX_FRAGMENT_ATOMS,)),
])

Info_dtype = np.dtype([
    ('mcsteps', np.int32),
    ('cutoff', np.float32),
    ('grid_dx', np.float32),
    ('startxyz', np.float32, (3,)),
    ('cryst', np.float32, (3,)),
    ('showInfo', np.int32),
    ('cavityFactor', np.float32),
    ('fragTypeNum', np.int32),
    ('totalGridNum', np.int32),
    ('totalResNum', np.int32),
    ('totalAtomNum', np.int32),
    ('ffXNum', np.int32),
    ('ffYNum', np.int32),
    ('PME', np.int32),
    ('seed', np.uint32),
])

Residue_dtype = np.dtype([
    ('position', np.float32, (3,)),
    ('atomNum', np.int32),
    ('atomStart', np.int32),
])

assert Atom_dtype.itemsize == 20
assert AtomArray_dtype.itemsize == 436
assert Info_dtype.itemsize == 76
assert Residue_dtype.itemsize == 20

MOVE_NAMES = ('Insert', 'Delete', 'Translate', 'Rotate')
