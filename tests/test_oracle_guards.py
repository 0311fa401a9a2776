"""The CPU oracle refuses arguments that would make it read outside its
inputs (a schedule too short for the run, a bad move code, a trace too small
to follow), before any state changes."""

import numpy as np
import pytest

from oracle import oracle as O

from . import systems


def test_run_past_the_schedule_raises():
    s = systems.water(nsteps=25)
    before = O.copy_system(s)
    with pytest.raises(ValueError, match='outside the move schedule'):
        O.run(s, first=0, nsteps=300)
    with pytest.raises(ValueError, match='outside the move schedule'):
        O.run(s, first=20, nsteps=6)
    with pytest.raises(ValueError, match='outside the move schedule'):
        O.run(s, first=-1, nsteps=2)
    for k in ('frag', 'res', 'atoms'):
        assert np.array_equal(s[k].view(np.uint8), before[k].view(np.uint8))
    O.run(s, first=0, nsteps=25)     # the whole schedule is fine


def test_bad_move_code_raises():
    s = systems.water(nsteps=20)
    s['moves'][7] = 4 * int(s['info'][0]['fragTypeNum'])   # a species that does not exist
    with pytest.raises(ValueError, match='species outside fragTypeNum'):
        O.run(s, first=0, nsteps=20)
    s['moves'][7] = -1
    with pytest.raises(ValueError, match='species outside fragTypeNum'):
        O.run(s, first=0, nsteps=20)
    O.run(s, first=0, nsteps=7)      # the steps before it are valid


def test_bad_species_count_raises():
    s = systems.water(nsteps=20)
    s['frag']['totalNum'][0] = s['frag']['maxNum'][0] + 1
    with pytest.raises(ValueError, match='species count'):
        O.run(s, first=0, nsteps=5)


def test_follow_trace_too_small_raises():
    s = systems.water(nsteps=50, conf=3)
    tr = O.new_trace(10, 3)
    with pytest.raises(ValueError):
        O.run(O.copy_system(s), first=0, nsteps=50, follow=tr)
    narrow = O.new_trace(50, 2)
    with pytest.raises(ValueError, match='followed trace'):
        O.run(O.copy_system(s), first=0, nsteps=50, follow=narrow)


def test_cell_list_variant_gives_the_same_chain():
    """The cell-list scan (a CPU timing variant) includes the same groups as
    the all-groups scan of calcEnergy; only the FP64 summation order differs."""
    from pygcmc_amd import synth
    s = synth.multi_fragment(n_fixed=3000, box=48.0, mcsteps=120, conf_scale=0.05, seed=11)
    a, b = O.copy_system(s), O.copy_system(s)
    ta, tb = O.new_trace(120, 20), O.new_trace(120, 20)
    O.run(a, trace=ta)
    O.run(b, trace=tb, cells=True)
    assert np.array_equal(ta['ninc'], tb['ninc']) and np.array_equal(ta['inchash'], tb['inchash'])
    assert (np.abs(ta['e'] - tb['e']) <= 1e-9 + 1e-12 * np.abs(ta['e'])).all()
    assert np.array_equal(ta['gacc'], tb['gacc']) and np.array_equal(a['frag']['totalNum'], b['frag']['totalNum'])
