"""``python -m pygcmc_amd.cli``: the ``pygcmc`` command on the MI355X engine.

Same options as pyGCMC's console script (``gcmc/scripts/main.py:19-226``):
``-p`` structure, ``-t`` .top or ``-s`` .psf, ``-o`` log copy, ``-u`` muex,
``-f`` conf bias (one value for all), ``-n`` MC steps, ``-m`` mctime, ``-c``
concentrations, ``-y`` cavity-bias spacing (<= 0: none), ``-e`` seed, ``-P``
PME flag, ``-w`` show info; writes ``<pdb>_out.pdb``.  Added: ``--fragments``
(species names, default the nine of pygcmc), ``--ff-dir``
(force-field directory, default ``charmm36.ff``), ``--top-out`` (write the
.top with the final molecule counts), ``--grid-order`` (``cpython``: the
reference's grid order; ``linear``: built on the GPU).
"""

import argparse
import sys
import time

from .system import GCMCSystem


class _Tee:
    def __init__(self, *streams):
        self.streams = streams

    def write(self, s):
        for st in self.streams:
            st.write(s)

    def flush(self):
        for st in self.streams:
            st.flush()


def _floats(s):
    return [float(v) for v in s.split(',') if v]


def main(argv=None):
    ap = argparse.ArgumentParser(prog='pygcmc', description='GCMC on the MI355X engine')
    ap.add_argument('-p', '--pdb-file', required=True)
    ap.add_argument('-t', '--top-file')
    ap.add_argument('-s', '--psf-file')
    ap.add_argument('-o', '--out-file')
    ap.add_argument('-u', '--fragmuex')
    ap.add_argument('-f', '--fragconf')
    ap.add_argument('-n', '--mcsteps', type=int)
    ap.add_argument('-m', '--mctime')
    ap.add_argument('-c', '--fragconc')
    ap.add_argument('-y', '--cavitybias-dx', type=float)
    ap.add_argument('-e', '--seed', type=int)
    ap.add_argument('-P', '--PME', action='store_true')
    ap.add_argument('-w', '--show-info', action='store_true')
    ap.add_argument('--fragments', help='species names (default: the nine SILCS species of pygcmc)')
    ap.add_argument('--ff-dir', default='charmm36.ff')
    ap.add_argument('--top-out')
    ap.add_argument('--grid-order', default='cpython', choices=['cpython', 'linear'])
    a = ap.parse_args(argv)

    t0 = time.time()
    log, stdout = None, sys.stdout
    if a.out_file:
        log = open(a.out_file, 'w')
        sys.stdout = _Tee(stdout, log)
    try:
        kw = {}
        if a.fragments:
            kw['fragment_names'] = [v for v in a.fragments.split(',') if v]
        if a.fragmuex:
            kw['muex'] = _floats(a.fragmuex)
        if a.fragconf:
            kw['conf'] = [int(v) for v in a.fragconf.split(',') if v]
        if a.mcsteps is not None:
            kw['mcsteps'] = a.mcsteps
        if a.mctime:
            kw['mctime'] = _floats(a.mctime)
        if a.fragconc:
            kw['conc'] = _floats(a.fragconc)
        if a.cavitybias_dx is not None:
            kw['grid_dx'] = a.cavitybias_dx
            kw['cavity_bias'] = a.cavitybias_dx > 0
        sysm = GCMCSystem(seed=a.seed, pme=a.PME, show_info=a.show_info, ff_dir=a.ff_dir,
                          grid_order=a.grid_order, **kw)
        sysm.load(a.pdb_file, top=a.top_file, psf=a.psf_file)
        print(f'pdb atom number: {len(sysm.pdb)}; cryst: {sysm.pdb.cryst}')
        sysm.setup()
        info = sysm.arrays[0][0]
        print(f'species: {", ".join(f"{n} {c}" for n, c in sysm.counts())}')
        print(f'grid points: {int(info["totalGridNum"])}, cavity factor {float(info["cavityFactor"]):.6f}; '
              f'atom types {len(sysm.types1)} x {len(sysm.types2)}; set-up {sysm.timings["setup_s"]:.2f} s')
        sysm.run()
        print(f'GPU simulation time: {sysm.timings["run_s"]:.2f} s')
        print(f'final: {", ".join(f"{n} {c}" for n, c in sysm.counts())}')
        out = sysm.write(top_out=a.top_out)
        print(f'wrote {out}' + (f' and {a.top_out}' if a.top_out else '') + f'; total {time.time() - t0:.2f} s')
    finally:
        if log is not None:
            sys.stdout = stdout
            log.close()
    return 0


if __name__ == '__main__':
    sys.exit(main())
