"""GPU: the serving loop (seren_amd.serve) with continuous batching.

Several resident clouds receive interleaved force-evaluation-plus-step
requests from concurrent worker slots, in mixed precision with FP64 guard
steps.  Each cloud's state afterwards (positions, velocities, h, rho,
accelerations, time-step levels) equals, bit for bit, the state after the
same number of requests run on that cloud alone: results do not depend on
what else is in flight.  Admission stays within the HBM budget.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from helpers import cloud  # noqa: E402

pytestmark = pytest.mark.gpu

CLOUDS = {0: (2048, "uniform", 41), 1: (4096, "r2", 42), 2: (3000, "plummer", 43),
          3: (2048, "r2", 44)}
REQUESTS = [0, 1, 2, 1, 3, 0, 2, 2, 1, 3, 0, 1, 3, 2, 0, 1]
KEYS = ("pos", "v", "h", "rho", "a", "nlevel", "nlast", "v_half")


def _server(slots):
    from seren_amd.serve import CloudServer
    return CloudServer(hbm_budget=8e9, slots=slots, precision="mixed", guard_every=2)


def test_results_same_alone_and_in_the_mix():
    print("case serve 4 clouds, 16 requests, 3 slots", flush=True)
    data = {k: cloud(*v) for k, v in CLOUDS.items()}
    mix = _server(3)
    try:
        for k, (pos, vel, mass, h) in data.items():
            mix.admit(k, pos, vel, mass, h)
        assert mix.used <= mix.budget
        mix.start()
        futs = [mix.submit(k) for k in REQUESTS]
        for f in futs:
            f.result(timeout=120)
        got = {k: mix.state(k) for k in data}
        assert sorted(mix.nsteps.values()) == sorted(REQUESTS.count(k) for k in data)
    finally:
        mix.close()
    assert mix.used == 0
    for k, (pos, vel, mass, h) in data.items():
        alone = _server(1)
        try:
            alone.admit(k, pos, vel, mass, h)
            alone.start()
            for _ in range(REQUESTS.count(k)):
                alone.submit(k).result(timeout=120)
            ref = alone.state(k)
        finally:
            alone.close()
        for key in KEYS:
            assert np.array_equal(got[k][key], ref[key]), (k, key)


def test_admission_waits_for_memory():
    from seren_amd.serve import CloudServer
    print("case serve admission budget", flush=True)
    pos, vel, mass, h = cloud(2048, "uniform", 45)
    need = CloudServer.estimate_bytes(2048)
    srv = CloudServer(hbm_budget=1.5 * need, slots=1)
    try:
        srv.admit(0, pos, vel, mass, h)
        with pytest.raises(MemoryError):
            srv.admit(1, pos, vel, mass, h, timeout=0.2)
        srv.release(0)
        srv.admit(1, pos, vel, mass, h, timeout=5)
        assert 1 in srv.engines and 0 not in srv.engines
    finally:
        srv.close()
