"""Serving loop: force-evaluation-plus-step requests on clouds kept resident
on the device, with continuous batching into shared device epochs.

* Clouds live in pools.  A pool is one library context holding several
  clouds as a forest (serenx_load_forest): particles, trees, neighbour
  lists and step state stay on the device between requests.  Clouds are
  packed into pools by size class, first fit up to ``pool_particles``
  particles, so the many small clouds share a pool and a 1M-particle cloud
  fills one.  Admission charges a pool's device bytes against an HBM budget
  (and the device's free memory); a pool that does not fit waits until
  released pools free memory.
* A request advances its cloud by one kick-drift-kick leapfrog step with the
  cloud's own step size (serenx_serve_step: kick, drift, one force
  evaluation, kick, next step size), in mixed precision with an FP64 guard
  step every ``guard_every`` steps of that cloud.
* Continuous batching: each pool has a worker that runs epochs.  An epoch
  takes every cloud with a pending request (the oldest request of each) and
  advances all of them with one call, i.e. one launch per kernel for the
  whole batch; requests that arrive meanwhile join the next epoch.  Pools
  run their epochs concurrently on their own streams (library calls release
  the GIL), so small-cloud epochs overlap the large clouds' steps.  Requests
  on one cloud complete in arrival order.

A cloud's particles are active only in the epochs that step it, and no walk
leaves its own cloud's tree, so its results do not depend on what else is in
the epoch or the pool: the serving test checks each cloud's state after its
requests against the same requests served with the cloud alone, bit for bit.
"""
from __future__ import annotations

import collections
import dataclasses
import threading
import time
from concurrent.futures import Future
from typing import Deque, Dict, List, Optional

import numpy as np

from .engine import ForceEngine


@dataclasses.dataclass
class Request:
    cloud: int
    arrival: float
    future: Future


class Pool:
    """Clouds resident in one forest context, stepped in epochs."""

    def __init__(self, ids: List[int], data: Dict[int, tuple], precision: str,
                 engine_kw: dict, step_kw: dict, serial: bool):
        self.ids = list(ids)
        self.index = {c: k for k, c in enumerate(self.ids)}
        self.precision = precision
        self.eng = ForceEngine(precision=precision, serial=serial, **engine_kw)
        try:
            self.off = self.eng.load_forest([data[c] for c in self.ids])
            self.eng.force_eval(hguess=True)       # setup evaluation (sph_setup)
            self.eng.serve_begin(**step_kw)
            self.eng.sync()
            self.bytes = self.eng.memory()["context"]
        except BaseException:
            self.eng.close()
            raise
        self.n = {c: int(self.off[k + 1] - self.off[k]) for c, k in self.index.items()}
        self.nsteps = {c: 0 for c in self.ids}
        self.live = set(self.ids)
        self.pending: Deque[Request] = collections.deque()
        self.cv = threading.Condition()
        self.stopping = False
        self.epochs = 0
        self.thread: Optional[threading.Thread] = None
        self.lock = threading.Lock()   # the context: one host thread at a time

    def take_epoch(self) -> List[Request]:
        """The oldest pending request of every cloud that has one."""
        batch, seen, rest = [], set(), collections.deque()
        while self.pending:
            r = self.pending.popleft()
            if r.cloud in seen:
                rest.append(r)
            else:
                seen.add(r.cloud)
                batch.append(r)
        self.pending = rest
        return batch

    def run_epoch(self, batch: List[Request], guard_every: int) -> None:
        """Steps the batch's clouds: the mixed ones in one call, the ones due
        an FP64 guard step in another."""
        groups = {"mixed": [], "fp64": []}
        for r in batch:
            k = self.nsteps[r.cloud]
            guard = self.precision == "mixed" and guard_every > 0 and (k + 1) % guard_every == 0
            groups["fp64" if guard or self.precision == "fp64" else "mixed"].append(r.cloud)
        with self.lock:
            for prec, clouds in groups.items():
                if not clouds:
                    continue
                sel = np.zeros(len(self.ids), np.uint8)
                sel[[self.index[c] for c in clouds]] = 1
                if prec != self.precision:
                    self.eng.set_precision(prec)
                self.eng.serve_step(sel)
                if prec != self.precision:
                    self.eng.set_precision(self.precision)
                for c in clouds:
                    self.nsteps[c] += 1
            self.epochs += 1

    def state(self, cloud: int) -> Dict[str, np.ndarray]:
        k = self.index[cloud]
        lo, hi = int(self.off[k]), int(self.off[k + 1])
        with self.lock:
            d = self.eng.fetch(neighbours=False)
            dt, t = self.eng.serve_info(len(self.ids))
        out = {key: v[lo:hi] for key, v in d.items() if isinstance(v, np.ndarray)}
        out["porig"] = out["porig"] - lo
        out["dt"], out["time"] = dt[k], t[k]
        return out

    def close(self) -> None:
        self.eng.close()


class CloudServer:
    def __init__(self, hbm_budget: float = 200e9, precision: str = "mixed", guard_every: int = 4,
                 pool_particles: int = 1_500_000, engine_kw: Optional[dict] = None,
                 step_kw: Optional[dict] = None, serial: bool = True, slots: int = 0):
        """slots is accepted for compatibility and ignored: concurrency is one
        epoch worker per pool, each pool on its own stream.  HIP maps a
        process's streams onto at most GPU_MAX_HW_QUEUES hardware queues (4
        unless the environment sets it before the runtime starts); pools
        that share a queue run one after the other, so a serving process
        should set it to at least the number of pools (bench.py sets 16)."""
        self.budget = float(hbm_budget)
        self.precision = precision
        self.guard_every = int(guard_every)
        self.pool_particles = int(pool_particles)
        self.serial = serial
        self.engine_kw = dict(multipole="octupole", mac="gadget2", abserror=1e-4,
                              sound_speed=0.2)
        self.engine_kw.update(engine_kw or {})
        self.step_kw = dict(accel_mult=0.5, courant_mult=0.2)
        self.step_kw.update(step_kw or {})
        self.pools: List[Pool] = []
        self.pool_of: Dict[int, Pool] = {}
        self.used = 0
        self.mem = threading.Condition()
        self.latencies: List[float] = []
        self.particles_done = 0
        self.lat_lock = threading.Lock()
        self.running = False

    # ---- clouds ------------------------------------------------------------
    @staticmethod
    def estimate_bytes(n: int) -> int:
        """Device bytes of n resident particles (particle and step state,
        records, neighbour rows, h-solve lists, tree of <= 2n cells), rounded
        up, plus a context's fixed part; a loaded pool is charged the larger
        of this and the bytes its context holds."""
        return int(n * 2800 + (96 << 20))

    def _reserve(self, need: int, timeout: Optional[float]) -> None:
        deadline = None if timeout is None else time.perf_counter() + timeout
        with self.mem:
            while self.used + need > self.budget:
                left = None if deadline is None else deadline - time.perf_counter()
                if left is not None and left <= 0:
                    raise MemoryError("the clouds do not fit the HBM budget")
                self.mem.wait(left)
            self.used += need

    def admit_many(self, clouds: Dict[int, tuple], timeout: Optional[float] = None) -> None:
        """Packs the clouds {id: (pos, vel, mass, h)} into pools (largest
        first, first fit up to pool_particles particles), reserves each
        pool's bytes against the HBM budget, loads it and runs its setup
        evaluation."""
        order = sorted(clouds, key=lambda c: -len(clouds[c][2]))
        groups: List[List[int]] = []
        sizes: List[int] = []
        for c in order:
            n = len(clouds[c][2])
            for g, s in zip(groups, sizes):
                if s + n <= self.pool_particles and len(clouds[g[0]][2]) == n:
                    g.append(c)
                    sizes[groups.index(g)] += n
                    break
            else:
                groups.append([c])
                sizes.append(n)
        for g in groups:
            g.sort()
            need = sum(self.estimate_bytes(len(clouds[c][2])) for c in g)
            self._reserve(need, timeout)
            try:
                pool = Pool(g, clouds, self.precision, self.engine_kw, self.step_kw, self.serial)
            except BaseException:
                with self.mem:
                    self.used -= need
                    self.mem.notify_all()
                raise
            charged = max(need, pool.bytes)
            with self.mem:
                self.used += charged - need
            pool.charged = charged
            self.pools.append(pool)
            for c in g:
                self.pool_of[c] = pool
            if self.running:
                self._start_pool(pool)

    def admit(self, cloud: int, pos, vel, mass, h=None, timeout: Optional[float] = None) -> None:
        self.admit_many({cloud: (pos, vel, mass, h)}, timeout)

    def release(self, cloud: int) -> None:
        """The cloud takes no more requests; its pool's memory is freed when
        every cloud of the pool is released."""
        pool = self.pool_of.pop(cloud)
        pool.live.discard(cloud)
        if pool.live:
            return
        self._stop_pool(pool)
        pool.close()
        self.pools.remove(pool)
        with self.mem:
            self.used -= pool.charged
            self.mem.notify_all()

    @property
    def nparticles(self) -> Dict[int, int]:
        return {c: p.n[c] for c, p in self.pool_of.items()}

    @property
    def nsteps(self) -> Dict[int, int]:
        return {c: p.nsteps[c] for c, p in self.pool_of.items()}

    # ---- requests ----------------------------------------------------------
    def _worker(self, pool: Pool) -> None:
        while True:
            with pool.cv:
                while not pool.pending and not pool.stopping:
                    pool.cv.wait()
                if pool.stopping and not pool.pending:
                    return
                batch = pool.take_epoch()
            try:
                pool.run_epoch(batch, self.guard_every)
                done = time.perf_counter()
                with self.lat_lock:
                    for r in batch:
                        self.latencies.append(done - r.arrival)
                        self.particles_done += pool.n[r.cloud]
                for r in batch:
                    r.future.set_result(done - r.arrival)
            except BaseException as e:   # reported to the submitters
                for r in batch:
                    r.future.set_exception(e)

    def _start_pool(self, pool: Pool) -> None:
        pool.stopping = False
        pool.thread = threading.Thread(target=self._worker, args=(pool,), daemon=True)
        pool.thread.start()

    def _stop_pool(self, pool: Pool) -> None:
        if pool.thread is None:
            return
        with pool.cv:
            pool.stopping = True
            pool.cv.notify_all()
        pool.thread.join()
        pool.thread = None

    def start(self) -> None:
        self.running = True
        for p in self.pools:
            self._start_pool(p)

    def submit(self, cloud: int, arrival: Optional[float] = None) -> Future:
        f: Future = Future()
        pool = self.pool_of[cloud]
        r = Request(cloud, time.perf_counter() if arrival is None else arrival, f)
        with pool.cv:
            pool.pending.append(r)
            pool.cv.notify()
        return f

    def stop(self) -> None:
        for p in self.pools:
            self._stop_pool(p)
        self.running = False

    def close(self) -> None:
        self.stop()
        for c in list(self.pool_of):
            self.release(c)

    def state(self, cloud: int) -> Dict[str, np.ndarray]:
        return self.pool_of[cloud].state(cloud)

    def epochs(self) -> int:
        return sum(p.epochs for p in self.pools)


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
    e0 = server.epochs()
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
    npart = server.nparticles
    return {"requests": nrequests, "rate_per_s": rate, "wall_s": wall,
            "epochs": server.epochs() - e0,
            "p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "mean_ms": float(lat.mean() * 1e3),
            "particles_per_s": server.particles_done / wall,
            "offered_particles_per_s": float(
                rate * np.sum(np.array(probs) * np.array([npart[c] for c in clouds])))}
