"""A GCMC system from files: structure + topology + force field in, the
seven ``runGCMC`` arrays, a run on the GPU, the final structure out.

``GCMCSystem`` follows pyGCMC's set-up order and arithmetic
(``gcmc/python/files.py:18-150``, ``simulation.py:16-108``,
``dataset.py:141-344``, ``dataset.py:375-441``):

1. structure (PDB) and per-atom topology (GROMACS .top or CHARMM .psf), the
   topology's residue name, type and charge replacing the PDB's;
2. fragment templates ``<ff_dir>/mol/<name>.pdb`` + ``.itp`` and LJ
   parameters from ``ff_files`` (later files override earlier ones);
3. move schedule, then cavity grid, then type lists and the pair table
   (fragment types first), fixed / mobile split by residue name, fixed-atom
   grouping, packing, and the run seed drawn last from the same stream;
4. after the run, the output PDB (and TOP) with the final molecules.

Differences from the reference are in what it does not do: files are read
relative to the topology's directory (then the working directory and
``ff_dir``) instead of a ``charmm36.ff`` link in the working directory, and
errors raise exceptions instead of exiting.
"""

import os
import random
import time

import numpy as np

from . import dataset, formats
from .values import MAX_FRAGMENT_ATOMS

# gcmc/python/base.py:18-28 (the defaults of a pygcmc run)
FRAGMENT_NAMES = ['BENX', 'PRPX', 'DMEE', 'MEOH', 'FORM', 'IMIA', 'ACEY', 'MAMY', 'SOL']
FRAGMENT_CONC = [0.25, 0.25, 0.25, 0.25, 0.25, 0.25, 0.25, 0.25, 55.00]
FRAGMENT_MUEX = [-2.79, 1.46, -1.44, -5.36, -10.92, -14.18, -97.31, -68.49, -5.6]
FRAGMENT_CONF = [1000] * 9
FRAGMENT_MCTIME = [1.0] * 8 + [10.0]
FF_FILES = ['ffnonbonded.itp', 'nbfix.itp', 'silcs.itp']


class SetupError(ValueError):
    pass


class GCMCSystem:
    """Set-up, run and output of one GCMC system.

    Species parameters are per fragment name (``fragment_names``); a single
    ``conf`` value applies to all (``gcmc/python/parameters.py:24-42``).
    ``rng``: a ``random.Random`` (default: one seeded with ``seed``, or the
    module-level stream when ``seed`` is None, as pygcmc does).
    ``grid_order``: ``'cpython'`` lists the free grid cells in the order the
    reference's Python set yields them (its arrays, bit for bit); ``'linear'``
    builds the grid on the GPU (``gcmc_cavity_grid``), x slowest."""

    def __init__(self, fragment_names=None, conc=None, muex=None, conf=None, mctime=None, mcsteps=10000,
                 cutoff=15.0, fix_cutoff=6.0, grid_dx=1.0, cavity_bias=True, seed=None, show_info=False,
                 pme=False, ff_dir='charmm36.ff', ff_files=None, grid_order='cpython', rng=None,
                 attempt_prob_frag=dataset.ATTEMPT_PROB_FRAG, attempt_prob_water=dataset.ATTEMPT_PROB_WATER):
        self.names = list(fragment_names or FRAGMENT_NAMES)
        T = len(self.names)

        def per_species(v, default, what):
            v = list(default if v is None else v)
            if len(v) != T:
                raise SetupError(f'{what}: {len(v)} values for {T} species')
            return v

        self.conc = per_species(conc, FRAGMENT_CONC if fragment_names is None else [0.25] * T, 'conc')
        self.muex = per_species(muex, FRAGMENT_MUEX if fragment_names is None else [0.0] * T, 'muex')
        conf = [1000] * T if conf is None else list(conf)
        if len(conf) == 1:
            conf = conf * T
        self.conf = per_species(conf, None, 'conf')
        if any(c <= 0 for c in self.conf):
            raise SetupError('conf values must be positive')
        self.mctime = per_species(mctime, FRAGMENT_MCTIME if fragment_names is None else [1.0] * T, 'mctime')
        self.mcsteps = int(mcsteps)
        self.cutoff = float(cutoff)
        self.fix_cutoff = float(fix_cutoff)
        self.grid_dx = float(grid_dx)
        self.cavity_bias = bool(cavity_bias) and self.grid_dx > 0
        self.show_info = bool(show_info)
        self.pme = bool(pme)
        self.ff_dir = ff_dir
        self.ff_files = list(FF_FILES if ff_files is None else ff_files)
        self.grid_order = grid_order
        self.rng = rng if rng is not None else (random.Random(seed) if seed is not None else random)
        self.attempt_prob_frag = attempt_prob_frag
        self.attempt_prob_water = attempt_prob_water
        self.pdb = None
        self.top_text = None
        self.topology = None
        self.arrays = None
        self.timings = {}

    # -- inputs ------------------------------------------------------------
    def _ff_path(self, name, base):
        for d in (base, os.getcwd()):
            p = os.path.join(d, self.ff_dir, name) if d else None
            if p and os.path.exists(p):
                return p
        p = os.path.join(self.ff_dir, name)
        return p

    def load(self, pdb, top=None, psf=None):
        """Structure and topology (``gcmc/python/files.py:18-122``)."""
        t0 = time.perf_counter()
        self.pdb_path = pdb
        self.pdb = formats.read_pdb(pdb)
        if top is None and psf is None:
            raise SetupError('a topology (.top or .psf) is required')
        self.base_dir = os.path.dirname(os.path.abspath(top or psf))
        if top is not None:
            atoms, self.top_text, _ = formats.read_top(top, search=[os.path.dirname(os.path.abspath(self.ff_dir))])
            self.topology = 'top'
        else:
            atoms = formats.read_psf(psf)
            self.topology = 'psf'
        if len(atoms) != len(self.pdb):
            raise SetupError(f'pdb atom number {len(self.pdb)} != topology atom number {len(atoms)}')
        pres = self.pdb.atoms['residue'].tolist()
        for i, a in enumerate(atoms):
            if pres[i][:3] != a[2][:3]:
                raise SetupError(f'pdb atom {i + 1} residue {pres[i]} != topology residue {a[2]}')
        self.top_atoms = atoms
        self.timings['load_s'] = time.perf_counter() - t0
        return self

    def _templates(self):
        """Fragment templates (``gcmc/python/files.py:124-150``): atoms of
        ``mol/<name>.pdb`` with type and charge from ``mol/<name>.itp``."""
        out = []
        for name in self.names:
            pdb = formats.read_pdb(self._ff_path(os.path.join('mol', name.lower() + '.pdb'), self.base_dir))
            top = formats.read_itp(self._ff_path(os.path.join('mol', name.lower() + '.itp'), self.base_dir))
            if len(top) < len(pdb):
                raise SetupError(f'template {name}: {len(pdb)} pdb atoms, {len(top)} itp atoms')
            for i in range(len(pdb)):
                if pdb.atoms['residue'][i][:3] != top[i][2][:3]:
                    raise SetupError(f'template {name} atom {i + 1}: residue {pdb.atoms["residue"][i]} != {top[i][2]}')
            if len(pdb) > MAX_FRAGMENT_ATOMS:
                raise SetupError(f'template {name} has {len(pdb)} > {MAX_FRAGMENT_ATOMS} atoms')
            out.append(dict(xyz=pdb.xyz(), names=list(pdb.atoms['name']),
                            types=[top[i][0] for i in range(len(pdb))],
                            charges=[float(top[i][4]) for i in range(len(pdb))]))
        return out

    def _forcefield(self):
        nb, nbfix = {}, {}
        for f in self.ff_files:
            a, b = formats.read_ff(self._ff_path(f, self.base_dir))
            nb.update(a)
            nbfix.update(b)
        return nb, nbfix

    # -- set-up ------------------------------------------------------------
    def setup(self):
        """The seven arrays of ``runGCMC`` (``gcmc/python/simulation.py:16-108``
        and ``dataset.py:141-344``), in ``self.arrays``."""
        if self.pdb is None:
            raise SetupError('load() first')
        t0 = time.perf_counter()
        self.templates = self._templates()
        nb, nbfix = self._forcefield()
        moves, counts = dataset.get_move(self.names, self.mctime, self.mcsteps, self.rng,
                                         self.attempt_prob_frag, self.attempt_prob_water)
        xyz = self.pdb.xyz()
        cryst = [float(c) for c in self.pdb.cryst]
        grid_dx = self.grid_dx if self.cavity_bias else 1.0
        if self.grid_order == 'linear':
            from .engine import cavity_grid
            start, n, grid, cav = cavity_grid(xyz, cryst, grid_dx, self.cavity_bias)
        else:
            start, n, grid, cav = dataset.get_grid(xyz, cryst, grid_dx, self.cavity_bias, order=self.grid_order)
        if self.cavity_bias and cav < 0.05:
            raise SetupError(f'cavity bias factor {cav:.4f} < 0.05: use a smaller grid spacing')
        # atom types: fragment types first, then the system's (simulation.py:31-45)
        types1 = []
        for t in self.templates:
            for ty in t['types']:
                if ty not in types1:
                    types1.append(ty)
        types2 = list(types1)
        seen = set(types2)
        for a in self.top_atoms:
            if a[0] not in seen:
                seen.add(a[0])
                types2.append(a[0])
        missing = [t for t in types2 if t not in nb]
        if missing:
            # a pair of two such types could still come from NBFIX; the
            # reference fails on the first Lorentz-Berthelot lookup
            need = {t for t in missing}
            for t1 in types1:
                for t2 in types2:
                    if (t1, t2) in nbfix or (t2, t1) in nbfix:
                        continue
                    if t1 in need or t2 in need:
                        raise SetupError(f'no LJ parameters for atom type {t1 if t1 in need else t2}')
        ff = dataset.ff_pair_table(types1, types2, nb, nbfix)
        index = {t: i for i, t in enumerate(types2)}
        # fixed and mobile atoms by residue name (simulation.py:84-100)
        res = [a[2] for a in self.top_atoms]
        species_of = {name: s for s, name in enumerate(self.names)}
        fixed = [i for i, r in enumerate(res) if r not in species_of]
        mobile = [[] for _ in self.names]
        for i, r in enumerate(res):
            if r in species_of:
                mobile[species_of[r]].append(i)
        species = []
        for s, name in enumerate(self.names):
            t = self.templates[s]
            nf = len(t['types'])
            idx = mobile[s]
            if len(idx) % nf:
                raise SetupError(f'{name}: {len(idx)} atoms is not a whole number of {nf}-atom molecules')
            mols = [xyz[idx[j:j + nf]] for j in range(0, len(idx), nf)]
            species.append(dict(name=name, conc=self.conc[s], muex=self.muex[s], conf=self.conf[s],
                                mctime=self.mctime[s], template_xyz=t['xyz'], charge=t['charges'],
                                type=[index[ty] for ty in t['types']], molecules=mols))
        fxyz = xyz[fixed]
        groups = dataset.set_fixcut(fxyz, self.fix_cutoff) if fixed else np.zeros(0, np.int64)
        seed = self.rng.randint(0, 2 ** 32 - 1)
        self.arrays = dataset.pack_system(
            cryst=cryst, cutoff=self.cutoff, grid_dx=grid_dx, startxyz=start, grid=grid, cavity_factor=cav,
            fixed_xyz=fxyz, fixed_charge=np.array([float(self.top_atoms[i][4]) for i in fixed]),
            fixed_type=np.array([index[self.top_atoms[i][0]] for i in fixed], np.int32),
            fixed_group=groups, species=species, ff=ff, ffX=len(types1), ffY=len(types2),
            moves=moves, move_counts=counts, mcsteps=self.mcsteps, seed=seed,
            show_info=int(self.show_info), pme=int(self.pme))
        self.types1, self.types2 = types1, types2
        self.fixed = fixed
        self.move_counts = counts
        self.timings['setup_s'] = time.perf_counter() - t0
        return self.arrays

    # -- run and output ----------------------------------------------------
    def run(self, runner=None):
        """``runGCMC`` on the arrays, in place (default: ``pygcmc_amd.gpu``)."""
        if self.arrays is None:
            self.setup()
        if runner is None:
            from .gpu import runGCMC as runner
        t0 = time.perf_counter()
        runner(*self.arrays)
        self.timings['run_s'] = time.perf_counter() - t0
        return self

    def run_maps(self, every=100, spacing=1.0, prefix=None, cap=12.552, tuning=None):
        """Run on the engine in chunks of ``every`` moves (batched form; the
        same chain as ``run``), adding one occupancy snapshot per species
        after each chunk (``pygcmc_amd.maps``).  With ``prefix``, writes
        ``<prefix>.<species>.gfe.dx``: grid free energies in kJ/mol, empty
        cells at ``cap`` (default 3 kcal/mol).  Returns the maps."""
        from .engine import Engine
        from .maps import OccupancyMap
        if self.arrays is None:
            self.setup()
        t0 = time.perf_counter()
        sysd = dict(zip(('info', 'frag', 'res', 'atoms', 'grid', 'ff', 'moves'), self.arrays))
        with Engine(**(tuning or {})) as eng:
            eng.add_system(sysd)
            eng.upload()
            maps = [OccupancyMap(eng, 0, sp, spacing) for sp in range(len(self.names))]
            done = 0
            while done < self.mcsteps:
                n = min(int(every), self.mcsteps - done)
                eng.run(n, 'batched')
                done += n
                for m in maps:
                    m.accumulate()
            for m in maps:
                m.freeze()
            eng.download(0)
        self.timings['run_s'] = time.perf_counter() - t0
        if prefix:
            d = os.path.dirname(prefix)
            if d:
                os.makedirs(d, exist_ok=True)
            for name, m in zip(self.names, maps):
                m.write_dx(f'{prefix}.{name.lower()}.gfe.dx', m.free_energy(cap=cap))
        self.maps = maps
        return maps

    def counts(self):
        frag = self.arrays[1]
        return [(self.names[s], int(frag[s]['totalNum'])) for s in range(len(self.names))]

    def output_pdb(self):
        """The output PDB of the current arrays (``gcmc/python/dataset.py:375-416``)."""
        _, frag, res, atoms = self.arrays[:4]
        lines = [self.pdb.lines[i] for i in self.fixed]
        last = self.fixed[-1] if self.fixed else None
        last_serial = self.pdb.atoms['serial'][last] if last is not None else None
        last_seq = self.pdb.atoms['sequence'][last] if last is not None else None

        def mols():
            for s, name in enumerate(self.names):
                F = frag[s]
                names = self.templates[s]['names']
                for j in range(int(F['totalNum'])):
                    r = res[int(F['startRes']) + j]
                    a0, na = int(r['atomStart']), int(r['atomNum'])
                    yield name, names, atoms['position'][a0:a0 + na]
        return formats.pdb_output(self.pdb.cryst, lines, last_serial, last_seq, mols())

    def output_top(self):
        if self.topology != 'top':
            raise SetupError('the topology was not a .top file')
        return formats.top_output(self.top_text, self.counts())

    def write(self, pdb_out=None, top_out=None):
        """Write the output PDB (default ``<input>_out.pdb``, as pygcmc) and,
        when asked, the .top with the final molecule counts."""
        if pdb_out is None:
            pdb_out = self.pdb_path.rsplit('.', 1)[0] + '_out.pdb'
        with open(pdb_out, 'w') as f:
            f.write(self.output_pdb())
        if top_out is not None:
            with open(top_out, 'w') as f:
                f.write(self.output_top())
        return pdb_out
