"""GPU: restock force evaluations (the tree of the last build restocked with
the current positions, as tree_update.F90:24-88 does between rebuilds) run
as recorded HIP graphs give the eager restock evaluation bit for bit.

Each case replays the graphs several times, with the library's own eager
work between the replays (fetches, diagnostics, an eager restock
evaluation, a rebuild that forces a new recording), and checks every replay
against the same sequence run eagerly: h, rho, Omega, zeta, neighbour
counts and lists, accelerations and potentials, all bit-identical.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from helpers import CS, cloud  # noqa: E402

pytestmark = pytest.mark.gpu

KEYS = ("h", "rho", "omega", "zeta", "div_v", "pptot", "a", "gpot", "agravmag", "porig")


def _engine(precision):
    from seren_amd.engine import ForceEngine
    return ForceEngine(multipole="octupole", mac="gadget2", precision=precision,
                       thetamaxsqd=0.25, abserror=1e-4, sound_speed=CS)


def _sequence(eng, plan, graph):
    """plan: 'b' build evaluation, 'r' restock, 'g' restock through the
    graphs (eager when graph is False)."""
    outs = []
    for step in plan:
        if step == "b":
            eng.force_eval()
        elif step == "r" or not graph:
            eng.force_eval(restock=True)
        else:
            eng.force_eval(restock=True, graph=True)
        eng.sync()
        outs.append(eng.fetch())
        eng.diagnostics()   # eager library work between the replays
    return outs


def _compare(eager, graph, label):
    for k, (e, g) in enumerate(zip(eager, graph)):
        for key in KEYS:
            assert np.array_equal(e[key], g[key]), (label, k, key)
        ok = e["pptot"] <= e["pplist"].shape[1]
        assert np.array_equal(e["pplist"][ok], g["pplist"][ok]), (label, k)


@pytest.mark.parametrize("precision,profile,n,seed", [("fp64", "r2", 20000, 61),
                                                      ("mixed", "plummer", 12000, 62)])
def test_graph_replays_equal_eager_restock(precision, profile, n, seed):
    print(f"case graph restock {precision} {profile} n={n}", flush=True)
    pos, vel, mass, h = cloud(n, profile, seed)
    plan = "ggrgg" + "b" + "gg"
    results, info = [], None
    for graph in (False, True):
        eng = _engine(precision)
        try:
            eng.load(pos, vel, mass, h)
            eng.force_eval(hguess=True)
            results.append(_sequence(eng, plan, graph))
            if graph:
                info = eng.graph_info()
        finally:
            eng.close()
    _compare(results[0], results[1], precision)
    # the rows above pp_limit (filled between the two graphs) are exercised
    assert int(np.max(results[0][0]["pptot"])) > results[0][0]["pplist"].shape[1] \
        or profile == "plummer"
    # six replays; two recordings (A, B) before the rebuild and two after
    # (more only if a reallocation of the overflow rows forced one)
    assert info["launches"] == 6 and 4 <= info["captures"] <= 8, info


def test_graph_replays_on_a_forest():
    print("case graph restock forest of 3 clouds, mixed", flush=True)
    data = [cloud(4000, "uniform", 71), cloud(6000, "r2", 72), cloud(2500, "plummer", 73)]
    results = []
    for graph in (False, True):
        eng = _engine("mixed")
        try:
            eng.load_forest(data)
            eng.force_eval(hguess=True)
            results.append(_sequence(eng, "gggg", graph))
        finally:
            eng.close()
    _compare(results[0], results[1], "forest")


def test_restock_needs_a_built_tree():
    from seren_amd.engine import SerenxError
    eng = _engine("fp64")
    try:
        pos, vel, mass, h = cloud(3000, "uniform", 74)
        eng.load(pos, vel, mass, h)
        with pytest.raises(SerenxError):
            eng.force_eval(restock=True, graph=True)
    finally:
        eng.close()
