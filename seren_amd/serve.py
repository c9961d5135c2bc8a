"""Serving loop: force-evaluation-plus-step requests on clouds kept resident
on the device, with continuous batching.

* Every admitted cloud owns one library context (its own HIP stream and
  device state: particles, tree, neighbour lists, step state), created by
  the setup evaluation (tree h guess, h solve, forces) and kept until the
  cloud is released.  Admission charges each cloud its device bytes
  against an HBM budget (and the device's free memory); a cloud that does
  not fit waits until released clouds free memory.
* A request advances its cloud by one kick-drift-kick leapfrog step (two
  sph_integrate calls, one force evaluation of every particle), in
  mixed precision with an FP64 guard step every guard_every steps.
* Continuous batching: `slots` workers take requests as they arrive; work
  of different clouds runs concurrently on their own streams (the library
  calls release the GIL), so a request that arrives while others are on
  the GPU starts as soon as a slot is free, and its kernels overlap theirs.
  Requests on one cloud run in arrival order.

Clouds are independent, so a cloud's results do not depend on what else is
in flight: the serving test checks each cloud's state after its requests
against the same requests run alone, bit for bit.
"""
from __future__ import annotations

import dataclasses
import queue
import threading
import time
from concurrent.futures import Future
from typing import Dict, List, Optional

import numpy as np

from .engine import ForceEngine


@dataclasses.dataclass
class Request:
    cloud: int
    arrival: float
    future: Future


class CloudServer:
    def __init__(self, hbm_budget: float = 200e9, slots: int = 8, precision: str = "mixed",
                 guard_every: int = 4, engine_kw: Optional[dict] = None,
                 step_kw: Optional[dict] = None):
        self.budget = float(hbm_budget)
        self.slots = int(slots)
        self.precision = precision
        self.guard_every = int(guard_every)
        self.engine_kw = dict(multipole="octupole", mac="gadget2", abserror=1e-4,
                              sound_speed=0.2)
        self.engine_kw.update(engine_kw or {})
        self.step_kw = dict(integrator="lfkdk", nlevels=1, nbuildstep=8)
        self.step_kw.update(step_kw or {})
        self.engines: Dict[int, ForceEngine] = {}
        self.nparticles: Dict[int, int] = {}
        self.nsteps: Dict[int, int] = {}
        self.bytes: Dict[int, int] = {}
        self.locks: Dict[int, threading.Lock] = {}
        self.used = 0
        self.mem = threading.Condition()
        self.q: "queue.Queue[Optional[Request]]" = queue.Queue()
        self.workers: List[threading.Thread] = []
        self.latencies: List[float] = []
        self.particles_done = 0
        self.lat_lock = threading.Lock()

    # ---- clouds ------------------------------------------------------------
    @staticmethod
    def estimate_bytes(n: int) -> int:
        """Device bytes of a resident cloud of n particles (particle and step
        state, records, neighbour rows, h-solve lists, tree of <= 2n cells),
        rounded up; a loaded cloud is charged the larger of this and the
        bytes its context holds."""
        return int(n * 2800 + (96 << 20))

    def admit(self, cloud: int, pos, vel, mass, h=None, timeout: Optional[float] = None) -> None:
        """Reserves the cloud's device bytes against the HBM budget (waiting
        while released clouds free memory), creates its context, loads it and
        runs the setup evaluation."""
        need = self.estimate_bytes(len(mass))
        deadline = None if timeout is None else time.perf_counter() + timeout
        with self.mem:
            while self.used + need > self.budget:
                left = None if deadline is None else deadline - time.perf_counter()
                if left is not None and left <= 0:
                    raise MemoryError(f"cloud {cloud} does not fit the HBM budget")
                self.mem.wait(left)
            self.used += need
        eng = None
        try:
            # many clouds run at once: each walk stays on its cloud's one stream
            eng = ForceEngine(precision=self.precision, serial=True, **self.engine_kw)
            eng.load(pos, vel, mass, h)
            eng.force_eval(hguess=True)
            eng.step_begin(**self.step_kw)
            eng.sync()
            actual = eng.memory()["context"]
        except BaseException:
            if eng is not None:
                eng.close()
            with self.mem:
                self.used -= need
                self.mem.notify_all()
            raise
        # charge the larger of the estimate and the bytes held now: the
        # estimate covers buffers that grow later (overflow lists, sort space)
        charged = max(need, actual)
        with self.mem:
            self.used += charged - need
        self.engines[cloud] = eng
        self.nparticles[cloud] = len(mass)
        self.nsteps[cloud] = 0
        self.bytes[cloud] = charged
        self.locks[cloud] = threading.Lock()

    def release(self, cloud: int) -> None:
        with self.locks[cloud]:
            eng = self.engines.pop(cloud)
            eng.close()
        with self.mem:
            self.used -= self.bytes.pop(cloud)
            self.mem.notify_all()
        self.nparticles.pop(cloud)
        self.nsteps.pop(cloud)

    # ---- requests ----------------------------------------------------------
    def _run(self, cloud: int) -> int:
        eng = self.engines[cloud]
        with self.locks[cloud]:
            k = self.nsteps[cloud]
            guard = self.precision == "mixed" and self.guard_every > 0 and \
                (k + 1) % self.guard_every == 0
            if guard:
                eng.set_precision("fp64")
            eng.step(2)                     # one KDK leapfrog step
            if guard:
                eng.set_precision("mixed")
            self.nsteps[cloud] = k + 1
        return self.nparticles[cloud]

    def _worker(self):
        while True:
            r = self.q.get()
            if r is None:
                return
            try:
                n = self._run(r.cloud)
                done = time.perf_counter()
                with self.lat_lock:
                    self.latencies.append(done - r.arrival)
                    self.particles_done += n
                r.future.set_result(done - r.arrival)
            except BaseException as e:   # reported to the submitter
                r.future.set_exception(e)

    def start(self) -> None:
        for _ in range(self.slots):
            t = threading.Thread(target=self._worker, daemon=True)
            t.start()
            self.workers.append(t)

    def submit(self, cloud: int, arrival: Optional[float] = None) -> Future:
        f: Future = Future()
        self.q.put(Request(cloud, time.perf_counter() if arrival is None else arrival, f))
        return f

    def stop(self) -> None:
        for _ in self.workers:
            self.q.put(None)
        for t in self.workers:
            t.join()
        self.workers.clear()

    def close(self) -> None:
        self.stop()
        for c in list(self.engines):
            self.release(c)

    def state(self, cloud: int) -> Dict[str, np.ndarray]:
        eng = self.engines[cloud]
        with self.locks[cloud]:
            d = eng.fetch(neighbours=False)
            d.update(eng.fetch_step())
        return d


def poisson_run(server: CloudServer, rate: float, nrequests: int, clouds: List[int],
                probs, seed: int = 0) -> Dict[str, float]:
    """Submits nrequests at Poisson arrival times (mean rate per second) to
    clouds drawn with probabilities probs; returns latency percentiles and
    throughput over the run."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, nrequests)
    which = rng.choice(len(clouds), nrequests, p=probs)
    server.latencies.clear()
    server.particles_done = 0
    futs = []
    t0 = time.perf_counter()
    t = t0
    for g, w in zip(gaps, which):
        t += g
        now = time.perf_counter()
        if t > now:
            time.sleep(t - now)
        futs.append(server.submit(clouds[w], arrival=t))
    for f in futs:
        f.result()
    wall = time.perf_counter() - t0
    lat = np.array(server.latencies)
    return {"requests": nrequests, "rate_per_s": rate, "wall_s": wall,
            "p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "mean_ms": float(lat.mean() * 1e3),
            "particles_per_s": server.particles_done / wall,
            "offered_particles_per_s": float(
                rate * np.sum(np.array(probs) *
                              np.array([server.nparticles[c] for c in clouds])))}
