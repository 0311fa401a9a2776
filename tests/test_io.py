"""Format readers / writers and set-up from files (pygcmc_amd.formats,
pygcmc_amd.system), pinned by slices of pyGCMC's own test data
(tests/golden/io/ref_*, written by tests/golden/make_io_slices.py) and a
small system in the same formats (tests/golden/io/mini, make_io_system.py)."""

import os
import random
import re

import numpy as np
import pytest

from pygcmc_amd import dataset, formats
from pygcmc_amd.system import GCMCSystem

IO = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'io')
MINI = os.path.join(IO, 'mini')
FF = os.path.join(MINI, 'charmm36.ff')
KEYS = ('info', 'frag', 'res', 'atoms', 'grid', 'ff', 'moves')


def mini_system(topology='top', seed=11, mcsteps=200, conf=(2, 1)):
    s = GCMCSystem(fragment_names=['MEOH', 'SOL'], conc=[0.25, 55.0], muex=[-5.36, -5.6], conf=list(conf),
                   mctime=[1.0, 1.0], mcsteps=mcsteps, seed=seed, ff_dir=FF)
    kw = {topology: os.path.join(MINI, 'mini.' + topology)}
    s.load(os.path.join(MINI, 'mini.pdb'), **kw)
    s.setup()
    return s


# -- PDB ----------------------------------------------------------------------
def test_pdb_reference_records():
    p = formats.read_pdb(os.path.join(IO, 'ref_5uul_head.pdb'))
    assert p.cryst == [68.396, 73.809, 71.323]
    text = open(os.path.join(IO, 'ref_5uul_head.pdb')).read()
    assert len(p) == sum(1 for ln in text.split('\n') if ln.startswith('ATOM'))
    a = p.atoms[0]
    assert (a['serial'], a['name'], a['residue'], a['sequence']) == ('1', 'N', 'GLY', '0')
    assert (a['x'], a['y'], a['z']) == (23.521, 28.860, 49.148)
    assert p.lines[0].startswith('ATOM      1  N   GLY')


def test_pdb_box_from_extent_and_residue_columns():
    lines = ['ATOM      1  CB  MEOH    3      15.954   4.664   5.000  1.00  0.00',
             'HETATM    2  OG  MEOH    3      12.000   3.242   9.500  1.00  0.00',
             'ATOM      3  X   SHRT    3       1.000',            # too short: not an atom
             'ATOMX     4  OG  MEOH    3      12.000   3.242   9.500  1.00  0.00']  # 'atom ' needs a blank
    p = formats.parse_pdb('\n'.join(lines))
    assert len(p) == 2
    assert p.cryst == [15.954 - 12.0, 4.664 - 3.242, 9.5 - 5.0]
    assert list(p.atoms['residue']) == ['MEO', 'MEO']          # columns 18-20 only


def test_pdb_writer_reproduces_reference_records():
    """Records written by pyGCMC (test/5uul/*.rec_out.pdb) are reproduced by
    pdb_atom_record from their own fields, and pdb_output numbers molecules on
    from the last fixed atom."""
    # The fixture came from an earlier pyGCMC writer that right-aligned the
    # residue name and printed the box with %f; the writer in the reference
    # code (dataset.py:411, '%-4s%-4s', 'CRYST1  %.3f') is what this
    # restates.  Records of 4-letter residues read the same in both.
    lines = open(os.path.join(IO, 'ref_5uul_out_frag.pdb')).read().split('\n')
    lines = [ln for ln in lines if not (ln.startswith('ATOM') and len(ln) == 66 and ln[17] == ' ')]
    written = [ln for ln in lines if ln.startswith('ATOM') and len(ln) == 66]
    fixed = [ln for ln in lines if ln.startswith('ATOM') and len(ln) != 66]
    assert len(written) > 20 and len(fixed) > 5
    for ln in written:
        serial, name, res, seq = int(ln[6:11]), ln[13:17].strip(), ln[17:21].strip(), int(ln[22:26])
        xyz = [float(ln[30:38]), float(ln[38:46]), float(ln[46:54])]
        assert formats.pdb_atom_record(serial, name, res, seq, *xyz) == ln + '\n'
    # the slice: fixed records, then molecules of pyGCMC's writer, then END
    body = lines[:lines.index(written[0])]
    tail = written[:]
    mols, cur, last = [], None, None
    for ln in tail:
        seq = int(ln[22:26])
        if seq != last:
            cur = (ln[17:21].strip(), [], [])
            mols.append(cur)
            last = seq
        cur[1].append(ln[13:17].strip())
        cur[2].append([float(ln[30:38]), float(ln[38:46]), float(ln[46:54])])
    first = body[-1]
    out = formats.pdb_output([68.396, 73.809, 71.323], body, first[6:11].strip(), first[22:26].strip(),
                             [(r, n, x) for r, n, x in mols])
    got = out.split('\n')
    assert got[0] == 'CRYST1  68.396  73.809  71.323  90.00  90.00  90.00 P 1           1'
    assert got[1:1 + len(body)] == body
    assert got[1 + len(body):1 + len(body) + len(written)] == written
    assert got[-2] == 'END'


# -- topology -------------------------------------------------------------------
def test_top_molecule_counts_rewrite():
    """The [ molecules ] section of test/5uul/*.gc.25.top is itself pyGCMC
    writer output ('name<TAB>count'); rewriting it with its own counts gives
    it back, new counts replace, unlisted species are appended."""
    text = open(os.path.join(IO, 'ref_5uul_molecules.top')).read()
    listed = re.findall(r'^(\w+)\t(\d+)$', text, re.MULTILINE)
    assert [n for n, _ in listed] == ['BENX', 'PRPX', 'DMEE', 'MEOH', 'FORM', 'IMIA', 'ACEY', 'MAMY', 'SOL']
    same = formats.top_output(text, [(n, int(c)) for n, c in listed])
    head, body = text.split('[ molecules ]')
    assert same == head + '\n[ molecules ]\n' + body.strip() + '\n'
    new = formats.top_output(text, [('BENX', 7), ('SOL', 12345), ('NEWX', 3)])
    sect = new.split('[ molecules ]')[1]
    assert '\nBENX\t7\n' in sect and '\nSOL\t12345\n' in sect and sect.endswith('NEWX\t3\n')
    assert '\nPRPX\t41\n' in sect and 'Protein_chain_A     1' in sect


def test_top_includes_and_expansion():
    atoms, text, inc = formats.read_top(os.path.join(MINI, 'mini.top'))
    # forcefield.itp is included but defines no molecules; the #ifdef block is skipped
    assert [os.path.basename(p) for p in inc] == ['forcefield.itp', 'meoh.itp', 'tip3p.itp']
    assert len(atoms) == 12 + 3 * 6 + 8 * 3
    assert atoms[0] == ('NH1', '1', 'ALA', 'N', '-0.470')
    assert atoms[12] == ('CG331', '1', 'MEOH', 'CB', '-0.040')
    assert atoms[-1] == ('HT', '1', 'SOL', 'HW2', '0.417')
    # a missing include is an error unless skipped; a molecule with no type adds nothing
    missing, _ = formats.parse_top(open(os.path.join(IO, 'ref_5uul_molecules.top')).read(), IO, missing='skip')
    assert missing == []
    with pytest.raises(FileNotFoundError):
        formats.parse_top(open(os.path.join(IO, 'ref_5uul_molecules.top')).read(), IO)


def test_itp_blocks_and_first_atoms_section():
    text = '\n'.join(['[ moleculetype ]', 'AAA 3', '[ moleculetype ]', 'BBB 3', '[ atoms ]',
                      '1 T1 1 BBB X1 1 0.5 1.0 ; comment', '# preprocessor line', '2 T2 1 BBB X2 2 -0.5 1.0'])
    blocks = formats.itp_blocks(text)
    assert [b[0] for b in blocks] == ['title', 'moleculetype', 'moleculetype', 'atoms']
    # AAA has no [atoms] of its own: the next one is taken (reference quirk)
    assert formats._moleculetypes(blocks)[0] == ['AAA', ('T1', '1', 'BBB', 'X1', '0.5'), ('T2', '1', 'BBB', 'X2', '-0.5')]
    assert formats.parse_itp(text) == [('T1', '1', 'BBB', 'X1', '0.5'), ('T2', '1', 'BBB', 'X2', '-0.5')]


def test_psf_reference_head_matches_pdb():
    psf = formats.read_psf(os.path.join(IO, 'ref_psftest_head.psf'))
    pdb = formats.read_pdb(os.path.join(IO, 'ref_psftest_head.pdb'))
    assert len(psf) == 30 == len(pdb)
    assert psf[0] == ('NH3', '206', 'TRP', 'N', '-0.300000')
    assert psf[1] == ('HC', '206', 'TRP', 'HT1', '0.330000')
    for i in range(30):
        assert psf[i][2][:3] == pdb.atoms['residue'][i]
        assert psf[i][3] == pdb.atoms['name'][i]
    assert abs(sum(float(a[4]) for a in psf[:3]) - 0.36) < 1e-12


def test_ff_parameters_and_override_order():
    nb, nbfix = formats.read_ff(os.path.join(FF, 'ffnonbonded.itp'))
    s, e = dataset.charmm_to_gromacs(0.1521, 1.7682)
    assert abs(nb['OT'][0] - s) < 1e-12 and abs(nb['OT'][1] - e) < 5e-7
    assert nbfix == {('CG331', 'HGA3'): (0.33, 0.11)}          # [ pairtypes ] read as NBFIX
    nb2, nbfix2 = formats.read_ff(os.path.join(FF, 'nbfix.itp'))
    assert nb2 == {} and nbfix2 == {('NH1', 'OT'): (0.3, 0.7), ('HGA3', 'CG331'): (0.32, 0.12)}


# -- set-up from files -------------------------------------------------------------
def test_setup_from_top_and_psf_agree():
    a = mini_system('top').arrays
    b = mini_system('psf').arrays
    for k, x, y in zip(KEYS, a, b):
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), k


def test_setup_arrays():
    s = mini_system('top')
    info, frag, res, atoms, grid, ff, moves = s.arrays
    assert s.types1 == ['CG331', 'OG311', 'HGP1', 'HGA3', 'OT', 'HT']
    assert s.types2 == s.types1 + ['NH1', 'H', 'CT1', 'HB1', 'CT3', 'HA3', 'C', 'O']
    X, Y = len(s.types1), len(s.types2)
    assert (int(info[0]['ffXNum']), int(info[0]['ffYNum'])) == (X, Y)
    assert list(frag['totalNum']) == [3, 8] and list(frag['num_atoms']) == [6, 3]
    assert [bytes(n) for n in frag['name']] == [b'MEOH', b'SOL ']
    assert np.allclose(frag[0]['atoms']['position'][:6].mean(axis=0), 0.0, atol=1e-6)
    nb, nbfix = s._forcefield()
    row = {t: i for i, t in enumerate(s.types2)}
    ffp = ff.reshape(X, Y, 2)
    # NBFIX in either order; the [ pairtypes ] entry wins for its own key order
    assert tuple(ffp[s.types1.index('OT'), row['NH1']]) == (np.float32(0.3), np.float32(0.7))
    assert tuple(ffp[s.types1.index('CG331'), row['HGA3']]) == (np.float32(0.33), np.float32(0.11))
    assert tuple(ffp[s.types1.index('HGA3'), row['CG331']]) == (np.float32(0.32), np.float32(0.12))
    so, eo = nb['OT']
    sc, ec = nb['CT1']
    assert ffp[s.types1.index('OT'), row['CT1'], 0] == np.float32((so + sc) / 2)
    assert ffp[s.types1.index('OT'), row['CT1'], 1] == np.float32(np.sqrt(eo * ec))
    # the move schedule and the seed come from one stream, in pygcmc's order
    rng = random.Random(11)
    mv, counts = dataset.get_move(['MEOH', 'SOL'], [1.0, 1.0], 200, rng)
    assert np.array_equal(moves, mv) and int(info[0]['seed']) == rng.randint(0, 2 ** 32 - 1)
    start, n, g, cav = dataset.get_grid(s.pdb.xyz(), s.pdb.cryst, 1.0, True, order='cpython')
    assert np.array_equal(grid, g) and info[0]['cavityFactor'] == np.float32(cav)
    # fixed atoms first, grouped by the 6 A cut; molecules after
    nfix = int(frag[0]['startRes'])
    assert int(res[nfix - 1]['atomStart'] + res[nfix - 1]['atomNum']) == 12
    assert np.array_equal(dataset.set_fixcut(s.pdb.xyz()[:12], 6.0),
                          np.repeat(np.arange(nfix), res[:nfix]['atomNum']))
    mol0 = s.pdb.xyz()[12:18].astype(np.float32)
    a0 = int(res[nfix]['atomStart'])
    assert np.array_equal(atoms['position'][a0:a0 + 6], mol0)


def test_setup_errors():
    s = GCMCSystem(fragment_names=['MEOH', 'SOL'], conc=[0.25, 55.0], muex=[0, 0], mctime=[1, 1], ff_dir=FF)
    with pytest.raises(ValueError):
        s.load(os.path.join(MINI, 'mini.pdb'))                       # no topology
    with pytest.raises(ValueError):
        GCMCSystem(fragment_names=['MEOH', 'SOL'], conc=[0.25])      # one value for two species
    with pytest.raises(ValueError):
        GCMCSystem(fragment_names=['MEOH', 'SOL'], conf=[0, 1])


def test_run_with_oracle_and_outputs(tmp_path):
    """The set-up arrays through the FP64 oracle, then the output PDB/TOP:
    counts, fixed records and molecule records agree with the final arrays."""
    from oracle import oracle as O
    s = mini_system('top', mcsteps=300, conf=(4, 2))
    s.run(runner=lambda *a: O.run(dict(zip(KEYS, a))))
    info, frag, res, atoms = s.arrays[:4]
    pdb_out = s.write(str(tmp_path / 'out.pdb'), str(tmp_path / 'out.top'))
    back = formats.read_pdb(pdb_out)
    n = [int(x) for x in frag['totalNum']]
    assert len(back) == 12 + 6 * n[0] + 3 * n[1]
    assert back.lines[:12] == s.pdb.lines[:12]
    assert back.atoms['serial'][12] == '13' and back.atoms['sequence'][12] == '3'
    r0 = res[int(frag[1]['startRes'])]
    p = atoms['position'][int(r0['atomStart'])]
    k = 12 + 6 * n[0]
    assert abs(back.atoms['x'][k] - p[0]) <= 5e-4 and back.atoms['residue'][k] == 'SOL'
    top = open(tmp_path / 'out.top').read()
    sect = top.split('[ molecules ]')[1]
    assert f'\nMEOH\t{n[0]}\n' in sect and f'\nSOL\t{n[1]}\n' in sect
    re_atoms, _ = formats.parse_top(top, MINI)
    assert len(re_atoms) == len(back)


# -- legacy parameter-file command ---------------------------------------------------
def test_legacy_parameter_file_matches_reference_log():
    """test/4wp7/gcmc.75.inp parsed as pyGCMC's gcmc command did: the same
    key/value lines its log printed (gcmc.75.out)."""
    from pygcmc_amd import legacy
    cfg = legacy.load_parameters(os.path.join(IO, 'ref_4wp7_gcmc.75.inp'))
    printed = [f'{k} {v}' for k, v in cfg.items()]
    expected = open(os.path.join(IO, 'ref_4wp7_params.out')).read().rstrip('\n').split('\n')
    assert printed == expected
    assert cfg['par'] == ['charmm36.ff/ffnonbonded.itp', 'charmm36.ff/silcs.itp', 'charmm36.ff/nbfix.itp']
    assert [legacy.true_value(v) for v in ('yes', 'True', 'ON', '1', 'no', '0', 'off')] == [True] * 4 + [False] * 3


def test_legacy_command(tmp_path):
    import shutil
    from oracle import oracle as O
    from pygcmc_amd import legacy
    for f in ('mini.pdb', 'mini.top'):
        shutil.copy(os.path.join(MINI, f), tmp_path / f)
    inp = tmp_path / 'run.inp'
    inp.write_text('\n'.join([
        f'pdb:{tmp_path / "mini.pdb"}', f'top:{tmp_path / "mini.top"}',
        f'op_pdb:{tmp_path / "out.pdb"}', f'op_top:{tmp_path / "out.top"}   # written by the run',
        'grid_dx:   1.000', 'use_cavity_bias:yes', 'random_seed:166 // comment', 'map_generation:yes',
        'fragname:     meoh      sol ', 'fragconc:     0.25    55.00 ', 'fragmuex:    -2.04    -5.04 ',
        'fragconf:       10        2 ', 'mctime:      0.2      0.8 ', 'mcsteps:200', '']))
    rc = legacy.main(['-p', str(inp), '--ff-dir', FF, '--logfile', str(tmp_path / 'log')],
                     runner=lambda *a: O.run(dict(zip(KEYS, a))))
    assert rc == 0
    log = (tmp_path / 'log').read_text()
    assert 'Fragment meoh:\tmuex: -2.04\tconf: 10\tconc: 0.25\tmctime: 0.2' in log
    out = formats.read_pdb(str(tmp_path / 'out.pdb'))
    atoms, _ = formats.parse_top((tmp_path / 'out.top').read_text(), MINI)
    assert len(atoms) == len(out)
    # same run through GCMCSystem gives the same output
    s = GCMCSystem(fragment_names=['MEOH', 'SOL'], conc=[0.25, 55.0], muex=[-2.04, -5.04], conf=[10, 2],
                   mctime=[0.2, 0.8], mcsteps=200, seed=166, ff_dir=FF)
    s.load(str(tmp_path / 'mini.pdb'), top=str(tmp_path / 'mini.top')).setup()
    s.run(runner=lambda *a: O.run(dict(zip(KEYS, a))))
    assert s.output_pdb() == (tmp_path / 'out.pdb').read_text()
