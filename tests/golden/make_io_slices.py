"""Write small slices of pyGCMC's own test data files as fixtures of the
format readers and writers (run in a tree where the reference is at
/root/reference; the slices are committed, the reference is not read later).

* ref_5uul_head.pdb        -- first 40 lines of test/5uul/*.rec.pdb (header,
                              CRYST1, MODEL, atom records)
* ref_5uul_out_frag.pdb    -- lines 2440-2480 and the last 3 lines of
                              test/5uul/*.rec_out.pdb: fixed-atom records as read
                              and records written by pyGCMC's writer
* ref_5uul_molecules.top   -- test/5uul/*.gc.25.top from the '; Include silcs
                              molecules' comment to the end: includes, an
                              #ifdef block, and a [ molecules ] section that
                              pyGCMC's writer produced
* ref_psftest_head.psf     -- test/PSFtest/step1_pdbreader.psf header and its
                              first 30 atoms, the atom count line set to 30
* ref_psftest_head.pdb     -- the first 30 atom records of the matching .pdb
                              with its CRYST1 record
* ref_4wp7_gcmc.75.inp     -- test/4wp7/gcmc.75.inp, a legacy parameter file
* ref_4wp7_params.out      -- the lines of test/4wp7/gcmc.75.out where the
                              legacy command echoed the parsed parameters
"""
import os

REF = '/root/reference/test'
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'io')


def lines(path):
    with open(path) as f:
        return f.read().split('\n')


def main():
    os.makedirs(OUT, exist_ok=True)
    u = os.path.join(REF, '5uul')
    rec = lines(os.path.join(u, '5uul_proa_silcs.1.prod.25.rec.pdb'))
    with open(os.path.join(OUT, 'ref_5uul_head.pdb'), 'w') as f:
        f.write('\n'.join(rec[:40]) + '\n')
    out = lines(os.path.join(u, '5uul_proa_silcs.1.prod.25.rec_out.pdb'))
    while out and out[-1] == '':
        out.pop()
    with open(os.path.join(OUT, 'ref_5uul_out_frag.pdb'), 'w') as f:
        f.write('\n'.join(out[2439:2480] + out[-3:]) + '\n')
    top = lines(os.path.join(u, '5uul_proa_silcs.1.gc.25.top'))
    start = next(i for i, ln in enumerate(top) if ln.startswith('; Include silcs molecules'))
    with open(os.path.join(OUT, 'ref_5uul_molecules.top'), 'w') as f:
        f.write('\n'.join(top[start:]))
    p = os.path.join(REF, 'PSFtest')
    psf = lines(os.path.join(p, 'step1_pdbreader.psf'))
    nat = next(i for i, ln in enumerate(psf) if '!NATOM' in ln)
    head = psf[:nat] + ['%10d !NATOM' % 30] + psf[nat + 1:nat + 31] + ['', '%10d !NBOND: bonds' % 0, '']
    with open(os.path.join(OUT, 'ref_psftest_head.psf'), 'w') as f:
        f.write('\n'.join(head))
    pdb = lines(os.path.join(p, 'step1_pdbreader.pdb'))
    cr = [ln for ln in pdb if ln.startswith('CRYST1')][:1]
    at = [ln for ln in pdb if ln.startswith('ATOM')][:30]
    with open(os.path.join(OUT, 'ref_psftest_head.pdb'), 'w') as f:
        f.write('\n'.join(cr + at + ['END']) + '\n')
    w = os.path.join(REF, '4wp7')
    with open(os.path.join(w, 'gcmc.75.inp')) as f:
        inp = f.read()
    with open(os.path.join(OUT, 'ref_4wp7_gcmc.75.inp'), 'w') as f:
        f.write(inp)
    log = lines(os.path.join(w, 'gcmc.75.out'))
    first = next(i for i, ln in enumerate(log) if ln.startswith('par ['))
    last = next(i for i, ln in enumerate(log) if ln.startswith('Using pdb file'))
    with open(os.path.join(OUT, 'ref_4wp7_params.out'), 'w') as f:
        f.write('\n'.join(log[first:last]) + '\n')


if __name__ == '__main__':
    main()
