"""GPU: the reference's fatal conditions come back as host status codes.

SEREN stops with a message on these conditions (SURVEY's failure list):
  * 'Reached maximum level' when more than LEAFMAX = 8 particles share the
    deepest cell (BHgrav_build.F90:289): SERENX_ELMAX;
  * a grad-h solve with fewer particles than pp_gather cannot reach its
    neighbour number: SERENX_EFEW before any kernel runs;
  * bad arguments: SERENX_EINVAL.
Each must be an error status (SerenxError with the code) with the message,
not a crash, and the context must stay usable: the next load and evaluation
of a good cloud give the same results as a fresh context.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from helpers import CS, cloud  # noqa: E402

pytestmark = pytest.mark.gpu


def _engine():
    from seren_amd.engine import ForceEngine
    return ForceEngine(multipole="quadrupole", mac="geometric", precision="fp64",
                       thetamaxsqd=0.25, sound_speed=CS)


def _eval(eng, pos, vel, mass):
    eng.load(pos, vel, mass, np.zeros(len(mass)))
    eng.force_eval(hguess=True)
    return eng.fetch()


def test_more_than_leafmax_coincident_particles_is_elmax():
    from seren_amd.engine import ELMAX, SerenxError
    pos, vel, mass, _ = cloud(2000, "uniform", 41)
    bad = pos.copy()
    bad[100:109] = bad[100]                      # 9 particles at one point
    print("case: 9 coincident particles in a 2000-particle cloud", flush=True)
    eng = _engine()
    try:
        with pytest.raises(SerenxError) as ei:
            _eval(eng, bad, vel, mass)
        assert ei.value.code == ELMAX
        assert "maximum level" in str(ei.value)
        # exactly LEAFMAX coincident particles form one leaf: no error
        ok8 = pos.copy()
        ok8[100:108] = ok8[100]
        _eval(eng, ok8, vel, mass)
        # the context is still good: same results as a fresh one
        got = _eval(eng, pos, vel, mass)
    finally:
        eng.close()
    eng2 = _engine()
    try:
        want = _eval(eng2, pos, vel, mass)
    finally:
        eng2.close()
    for k in ("porig", "h", "rho", "a", "gpot", "pptot"):
        assert np.array_equal(got[k], want[k]), k


def test_gradh_with_fewer_particles_than_pp_gather_is_efew():
    from seren_amd.engine import EFEW, SerenxError
    pos, vel, mass, _ = cloud(40, "uniform", 42)
    eng = _engine()
    try:
        with pytest.raises(SerenxError) as ei:
            _eval(eng, pos, vel, mass)
        assert ei.value.code == EFEW
    finally:
        eng.close()


def test_bad_tuning_values_are_einval():
    from seren_amd.engine import (EINVAL, TUNE_HSOLVE_SLOTS, TUNE_HSOLVE_WAVES,
                                  TUNE_MIXED_ABSERROR, TUNE_NEAR_CAP, TUNE_NEIB_LEAVES,
                                  TUNE_PAIR_BLOCKS, TUNE_WALK_QUEUE,
                                  SerenxError)
    eng = _engine()
    try:
        for key, v in ((TUNE_HSOLVE_SLOTS, 0), (TUNE_HSOLVE_SLOTS, 97), (TUNE_HSOLVE_SLOTS, 2.5),
                       (TUNE_MIXED_ABSERROR, -1.0), (TUNE_HSOLVE_WAVES, 4),
                       (TUNE_WALK_QUEUE, -1), (TUNE_WALK_QUEUE, 65), (TUNE_NEAR_CAP, 0),
                       (TUNE_NEAR_CAP, 4097), (TUNE_NEIB_LEAVES, -1), (TUNE_NEIB_LEAVES, 513),
                       (TUNE_PAIR_BLOCKS, -1), (TUNE_PAIR_BLOCKS, 65), (99, 1.0)):
            with pytest.raises(SerenxError) as ei:
                eng.set_tuning(key, v)
            assert ei.value.code == EINVAL
        eng.set_tuning(TUNE_HSOLVE_SLOTS, 8)
        assert eng.get_tuning(TUNE_HSOLVE_SLOTS) == 8
        assert eng.get_tuning(TUNE_MIXED_ABSERROR) == 1e-4
        assert eng.get_tuning(TUNE_WALK_QUEUE) == 64   # as many blocks as fit
        assert eng.get_tuning(TUNE_NEAR_CAP) == 4096
        assert eng.get_tuning(TUNE_NEIB_LEAVES) == 512
        assert eng.get_tuning(TUNE_PAIR_BLOCKS) == 0
    finally:
        eng.close()
