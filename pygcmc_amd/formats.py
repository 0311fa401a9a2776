"""Input and output formats of a GCMC run: PDB, CHARMM PSF, GROMACS TOP/ITP
and force-field parameter files in, the output PDB and TOP out.

The readers return what pyGCMC's set-up consumes, with its field choices and
its quirks (each function cites the reference code it follows):

* ``read_pdb``   -- ``gcmc/python/protein_data.py:27-59``
* ``itp_blocks`` -- ``gcmc/python/protein_data.py:63-93``
* ``read_itp``   -- ``gcmc/python/protein_data.py:95-127``
* ``read_psf``   -- ``gcmc/python/protein_data.py:129-174``
* ``read_top``   -- ``gcmc/python/protein_data.py:179-295``
* ``read_ff``    -- ``gcmc/python/protein_data.py:316-341``
* ``pdb_output`` / ``top_output`` -- ``gcmc/python/dataset.py:375-441``

Topology atoms are tuples ``(type, resnr, residue, name, charge)`` with the
charge kept as the text of the file (the reference converts it when packing).
"""

import os
import re

import numpy as np

# ---------------------------------------------------------------------------
# PDB

PDB_ATOM_DTYPE = np.dtype([
    ('serial', 'U8'), ('name', 'U4'), ('residue', 'U4'), ('sequence', 'U6'),
    ('x', 'f8'), ('y', 'f8'), ('z', 'f8'),
])


class PDB:
    """Atoms of a PDB file: ``atoms`` (structured array, PDB_ATOM_DTYPE),
    ``lines`` (the ATOM/HETATM records as read, for re-emission), ``cryst``
    (box edges, A) and ``text`` (the whole file)."""

    def __init__(self, cryst, atoms, lines, text):
        self.cryst = cryst
        self.atoms = atoms
        self.lines = lines
        self.text = text

    def __len__(self):
        return len(self.atoms)

    def xyz(self):
        return np.stack([self.atoms['x'], self.atoms['y'], self.atoms['z']], axis=1)


def parse_pdb(text):
    """``read_pdb`` on a string (``gcmc/python/protein_data.py:27-59``).

    The box is the first ``CRYST1`` record's a, b, c; without one, the
    extent (max - min) of the atoms per axis.  An atom is an ``ATOM`` or
    ``HETATM`` record (case-insensitive, ``ATOM`` followed by a blank) longer
    than 54 characters; x, y, z from columns 31-54, serial 7-11, name 13-16,
    residue name 18-20 (three characters), residue number 23-26."""
    lines = text.strip().split('\n')
    cryst = None
    for line in lines:
        if line.lower().startswith('cryst1 '):
            cryst = [float(v) for v in line.split()[1:4]]
            break
    recs, keep = [], []
    for line in lines:
        low = line.lower()
        if (low.startswith('atom ') or low.startswith('hetatm')) and len(line) > 54:
            recs.append((line[6:11].strip(), line[12:16].strip(), line[17:20].strip(), line[22:26].strip(),
                         float(line[30:38]), float(line[38:46]), float(line[46:54])))
            keep.append(line)
    atoms = np.array(recs, dtype=PDB_ATOM_DTYPE) if recs else np.zeros(0, dtype=PDB_ATOM_DTYPE)
    if cryst is None:
        if not len(atoms):
            raise ValueError('PDB has neither a CRYST1 record nor atoms')
        cryst = [float(atoms[c].max() - atoms[c].min()) for c in ('x', 'y', 'z')]
    return PDB(cryst, atoms, keep, text)


def read_pdb(path):
    with open(path) as f:
        return parse_pdb(f.read())


# ---------------------------------------------------------------------------
# GROMACS .itp / .top and force-field files

_SECTION = re.compile(r'\[(.*)\]')
_NAME_COUNT = re.compile(r'(\w+)\s+(\d+)')
_ATOM_LINE = re.compile(r'(\d+)\s+(\S+)\s+(\d+)\s+(\S+)\s+(\S+)\s+(\d+)\s+([0-9\.\-]+)')


def itp_blocks(text):
    """Sections of a GROMACS file as ``[name, line, line, ...]`` lists, the
    first named ``title`` (``gcmc/python/protein_data.py:63-93``).  Text after
    ``;`` and after ``#`` is dropped from every line (so preprocessor lines
    vanish), blank lines are skipped, a ``[ name ]`` line opens a section."""
    blocks = [['title']]
    for line in text.split('\n'):
        line = re.sub(r';.*', '', line)
        line = re.sub(r'#.*', '', line)
        m = _SECTION.match(line)
        if m:
            blocks.append([m.group(1).strip()])
        elif line.strip():
            blocks[-1].append(line.strip())
    return blocks


def _moleculetypes(blocks):
    """``[name, atom, atom, ...]`` per [moleculetype] section: the name from the
    section's first ``name nrexcl`` line, the atoms from the first [atoms]
    section after it (``gcmc/python/protein_data.py:106-125``; that section
    may belong to a later molecule type when one has none -- kept)."""
    out = []
    for b, block in enumerate(blocks):
        if block[0] != 'moleculetype':
            continue
        for line in block[1:]:
            m = _NAME_COUNT.match(line.strip())
            if m:
                out.append([m.group(1)])
                break
        for later in blocks[b + 1:]:
            if later[0] == 'atoms':
                for line in later[1:]:
                    m = _ATOM_LINE.match(line.strip())
                    if m:
                        out[-1].append((m.group(2), m.group(3), m.group(4), m.group(5), m.group(7)))
                break
    return out


def parse_itp(text):
    """Atoms of the last molecule type in an .itp text
    (``gcmc/python/protein_data.py:95-127``)."""
    mts = _moleculetypes(itp_blocks(text))
    if not mts:
        raise ValueError('no [ moleculetype ] with a name line')
    return mts[-1][1:]


def read_itp(path):
    with open(path) as f:
        return parse_itp(f.read())


def _resolve(name, base_dir, search):
    if os.path.isabs(name):
        return name if os.path.exists(name) else None
    for d in [base_dir, os.getcwd(), *search]:
        p = os.path.join(d, name)
        if os.path.exists(p):
            return p
    return None


def parse_top(text, base_dir='.', search=(), missing='error'):
    """Per-atom topology of a GROMACS .top text (``gcmc/python/protein_data.py:
    179-295``).  Returns ``(atoms, includes)``: the atoms of every molecule
    listed in the first [molecules] section, count times each, from the first
    molecule type of that name (types defined in the .top come before those
    of the included files, in include order); molecules with no type
    contribute nothing.  Included files are the ``#include "*.itp"`` lines
    outside ``#ifdef ... #endif`` (one level, not followed further), looked up
    relative to ``base_dir``, the working directory, then ``search``.
    ``missing='error'`` raises on an include that is not found, ``'skip'``
    ignores it."""
    blocks = itp_blocks(text)
    content = re.sub(r';.*', '', text)
    content = re.sub(r'#ifdef .*?#endif', '', content, flags=re.DOTALL)
    names = [n.strip() for n in re.findall(r'^\s*#include "(.*?)"\s*$', content, re.MULTILINE)]
    includes, inc_blocks = [], []
    for n in names:
        if not n.endswith('.itp'):
            continue
        p = _resolve(n, base_dir, search)
        if p is None:
            if missing == 'error':
                raise FileNotFoundError(f'#include "{n}" not found from {base_dir!r}')
            continue
        includes.append(p)
        with open(p) as f:
            inc_blocks.append(itp_blocks(f.read()))
    molecules = []
    for block in blocks:
        if block[0] == 'molecules':
            for line in block[1:]:
                m = _NAME_COUNT.match(line.strip())
                if m:
                    molecules.append((m.group(1), int(m.group(2))))
            break
    types = _moleculetypes(blocks)
    for b in inc_blocks:
        types += _moleculetypes(b)
    atoms = []
    for name, count in molecules:
        for mt in types:
            if mt[0] == name:
                atoms += mt[1:] * count
                break
    return atoms, includes


def read_top(path, search=(), missing='error'):
    """``parse_top`` of a file; returns ``(atoms, text, includes)``."""
    with open(path) as f:
        text = f.read()
    atoms, includes = parse_top(text, os.path.dirname(os.path.abspath(path)), search, missing)
    return atoms, text, includes


def parse_psf(text):
    """Per-atom topology of a CHARMM PSF text (``gcmc/python/protein_data.py:
    129-174``): after the header line, text after ``*``, ``!`` or ``#`` is
    dropped; the first integer line is the title count (those lines are
    skipped), the next integer line the atom count N; then the first N lines
    with at least eight fields give (type, resid, resname, name, charge) from
    fields 6, 3, 4, 5, 7."""
    lines = text.strip().split('\n')[1:]
    lines = [ln.split('*', 1)[0].split('!', 1)[0].split('#', 1)[0] for ln in lines]

    def first_int(seq):
        for i, ln in enumerate(seq):
            try:
                return i, int(ln)
            except ValueError:
                continue
        raise ValueError('PSF: no count line')

    i, ntitle = first_int(lines)
    lines = lines[i + ntitle + 1:]
    i, natom = first_int(lines)
    lines = lines[i + 1:]
    pat = re.compile(r'^(\S+)\s+(\S+)\s+(\S+)\s+(\S+)\s+(\S+)\s+(\S+)\s+(\S+)\s+(\S+)')
    out = []
    for ln in lines:
        m = pat.match(ln.strip())
        if m:
            if len(out) >= natom:
                break
            out.append((m.group(6), m.group(3), m.group(4), m.group(5), m.group(7)))
    return out


def read_psf(path):
    with open(path) as f:
        return parse_psf(f.read())


_ATOMTYPE = re.compile(r'(\S+)\s+(\S+)\s+([0-9\.\-]+)\s+([0-9\.\-]+)\s+(\S+)\s+([0-9\.\-]+)+\s+([0-9\.\-]+)')
_PAIRPARAM = re.compile(r'(\S+)\s+(\S+)\s+([0-9\.\-]+)\s+([0-9\.\-]+)\s+([0-9\.\-]+)')


def parse_ff(text):
    """LJ parameters of a GROMACS force-field text (``gcmc/python/
    protein_data.py:316-341``): ``nb[type] = (sigma_nm, eps_kJ)`` from
    [atomtypes] (name, at.num, mass, charge, ptype, sigma, epsilon) and
    ``nbfix[(t1, t2)] = (sigma_nm, eps_kJ)`` from [nonbond_params] and, as the
    reference does, [pairtypes] (t1, t2, func, sigma, epsilon).  Later lines
    override earlier ones."""
    nb, nbfix = {}, {}
    for block in itp_blocks(text):
        if block[0] == 'atomtypes':
            for line in block[1:]:
                m = _ATOMTYPE.match(line.strip())
                if m:
                    nb[m.group(1)] = (float(m.group(6)), float(m.group(7)))
        if block[0] in ('nonbond_params', 'pairtypes'):
            for line in block[1:]:
                m = _PAIRPARAM.match(line.strip())
                if m:
                    nbfix[(m.group(1), m.group(2))] = (float(m.group(4)), float(m.group(5)))
    return nb, nbfix


def read_ff(path):
    with open(path) as f:
        return parse_ff(f.read())


# ---------------------------------------------------------------------------
# output

def pdb_atom_record(serial, name, residue, resseq, x, y, z):
    """One ATOM record of a molecule written by the run, as pyGCMC writes it
    (``gcmc/python/dataset.py:411``): serial and residue number wrap at 99999
    and 9999, name and residue name are cut to four characters."""
    return 'ATOM  %5d  %-4s%-4s %4d    %8.3f%8.3f%8.3f  1.00  0.00\n' % (
        (serial - 1) % 99999 + 1, name[:4], residue[:4], (resseq - 1) % 9999 + 1, x, y, z)


def pdb_output(cryst, fixed_lines, last_serial, last_sequence, molecules):
    """The output PDB (``gcmc/python/dataset.py:375-416``): a CRYST1 record,
    the fixed atoms' records as read, then every molecule of every species in
    the final state, numbered on from the last fixed atom's serial and
    residue number (1 when those are absent or not integers), and END.

    ``molecules``: iterable of ``(residue_name, atom_names, xyz[n][3])``."""
    out = ['CRYST1  %.3f  %.3f  %.3f  90.00  90.00  90.00 P 1           1\n' % (cryst[0], cryst[1], cryst[2])]
    for line in fixed_lines:
        out.append(line + '\n')
    try:
        serial = int(last_serial) + 1
    except (TypeError, ValueError):
        serial = 1
    try:
        resseq = int(last_sequence) + 1
    except (TypeError, ValueError):
        resseq = 1
    for resname, names, xyz in molecules:
        for k, p in enumerate(xyz):
            out.append(pdb_atom_record(serial, names[k], resname, resseq, p[0], p[1], p[2]))
            serial += 1
        resseq += 1
    out.append('END\n')
    return ''.join(out)


def top_output(top_text, counts):
    """The .top text with the [molecules] counts of the run's species
    (``gcmc/python/dataset.py:418-441``): everything up to the first
    ``[ molecules ]`` is kept, the section's ``name count`` entries of each
    species are rewritten to the final count (a species not listed is
    appended), and text after a second [molecules] header is dropped.

    ``counts``: list of ``(species_name, count)`` in species order."""
    parts = re.split(r'\[\s*molecules\s*\]', top_text)
    if len(parts) < 2:
        raise ValueError('topology has no [ molecules ] section')
    body = parts[1].strip() + '\n'
    listed = {n: c for n, c in re.findall(r'(\S+?)\s+(\d+)', body)}
    for name, count in counts:
        if name in listed:
            body = re.sub('%s\\s+%s' % (name, listed[name]), '%-s\\t%d' % (name, count), body)
        else:
            body += '%-s\t%d\n' % (name, count)
    return parts[0] + '\n[ molecules ]\n' + body
