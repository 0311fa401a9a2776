"""``python -m pygcmc_amd.legacy -p params.inp``: pyGCMC's ``gcmc`` command
(``gcmc/scripts/mainOld.py``), a run described by a ``key: value`` parameter
file.

Keys read (the rest -- ``par``, ``fragitp``, ``monomerdir``, ``box_size``,
cut-offs, ... -- are parsed and ignored, as pyGCMC does):
``pdb``, ``top``, ``op_pdb``, ``op_top``, ``fragname`` (upper-cased),
``fragmuex``, ``fragconf``, ``fragconc``, ``mctime`` (whitespace-separated
per species), ``mcsteps``, ``grid_dx``, ``use_cavity_bias`` (absent = no),
``random_seed``, ``PME``, and ``map_generation`` / ``map_filename_prefix``,
which pyGCMC ignores: here they run the chain in 100 chunks with an
occupancy snapshot after each and write ``<prefix>.<species>.gfe.dx``
(``pygcmc_amd.maps``).  Writes ``op_pdb`` and ``op_top``.  Added:
``--ff-dir`` (force-field directory, default ``charmm36.ff``).
"""

import argparse
import sys
import time

from .system import GCMCSystem


def load_parameters(path):
    """``key -> [values]`` of a parameter file (``gcmc/python/files.py:
    170-192``): text after ``#``, then ``//``, then ``!`` is dropped, blank
    lines are skipped, each line is ``key: value`` split at the first colon,
    both stripped; a repeated key collects its values in order."""
    out = {}
    with open(path) as f:
        for line in f:
            line = line.strip()
            if not line:
                continue
            line = line.split('#', 1)[0].split('//', 1)[0].split('!', 1)[0].strip()
            if not line:
                continue
            key, value = line.split(':', 1)
            out.setdefault(key.strip(), []).append(value.strip())
    return out


def true_value(v):
    """``bool_parameters`` (``gcmc/python/files.py:164-167``)."""
    return v.lower() in ('yes', 'true', 'on', '1')


def main(argv=None, runner=None):
    ap = argparse.ArgumentParser(prog='gcmc', description='Perform GCMC Simulation')
    ap.add_argument('-p', '--paramfile', required=True)
    ap.add_argument('-v', '--verbose', action='store_true')
    ap.add_argument('--logfile')
    ap.add_argument('--debug', action='store_true')
    ap.add_argument('--ff-dir', default='charmm36.ff')
    a = ap.parse_args(argv)
    t0 = time.time()
    log, stdout = None, sys.stdout
    if a.logfile:
        log = open(a.logfile, 'w')

        class _Tee:
            def write(self, s):
                stdout.write(s)
                log.write(s)

            def flush(self):
                stdout.flush()
                log.flush()
        sys.stdout = _Tee()
    try:
        cfg = load_parameters(a.paramfile)
        for k, v in cfg.items():
            print(k, v)
        names = [n.upper() for n in cfg['fragname'][0].split()]
        kw = dict(fragment_names=names,
                  muex=[float(v) for v in cfg['fragmuex'][0].split()],
                  conf=[int(v) for v in cfg['fragconf'][0].split()],
                  conc=[float(v) for v in cfg['fragconc'][0].split()],
                  mctime=[float(v) for v in cfg['mctime'][0].split()],
                  mcsteps=int(cfg['mcsteps'][0]), grid_dx=float(cfg['grid_dx'][0]),
                  cavity_bias=true_value(cfg.get('use_cavity_bias', ['False'])[0]),
                  pme=true_value(cfg.get('PME', ['False'])[0]), ff_dir=a.ff_dir)
        if 'random_seed' in cfg:
            kw['seed'] = int(cfg['random_seed'][0])
        s = GCMCSystem(**kw)
        for i, n in enumerate(names):
            print(f'Fragment {n.lower()}:\tmuex: {s.muex[i]}\tconf: {s.conf[i]}\tconc: {s.conc[i]}\tmctime: {s.mctime[i]}')
        s.load(cfg['pdb'][0], top=cfg['top'][0])
        s.setup()
        print(f'Python time used: {time.time() - t0} s')
        if true_value(cfg.get('map_generation', ['no'])[0]) and runner is None:
            prefix = cfg.get('map_filename_prefix', [cfg['op_pdb'][0].rsplit('.', 1)[0]])[0]
            s.run_maps(every=max(1, s.mcsteps // 100), spacing=s.grid_dx if s.cavity_bias else 1.0, prefix=prefix)
            print(f'maps: {prefix}.<species>.gfe.dx')
        else:
            s.run(runner)
        print(f'GPU simulation time: {s.timings["run_s"]:.2f} s')
        s.write(cfg['op_pdb'][0], cfg['op_top'][0])
        print(f'Time used: {time.time() - t0} s')
    finally:
        if log is not None:
            sys.stdout = stdout
            log.close()
    return 0


if __name__ == '__main__':
    sys.exit(main())
