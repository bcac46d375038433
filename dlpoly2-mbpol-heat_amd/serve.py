"""Continuous-batching force server on one GPU.

Requests are complete force evaluations of water configurations of any size (PIMD beads of a
periodic box, isolated clusters, independent boxes).  A fixed set of worker threads takes
requests first-come first-served; each worker evaluates on a slot (engine.Slot: one library
context with its own HIP stream) of the request's size, so up to `workers` evaluations are
on the GPU at once and a request that arrives while others run starts as soon as a worker
is free: admission is per request, not per batch.

Slots are pooled by size and kind (periodic box or isolated cluster).  A worker reuses an
idle slot of the request's size and kind, creates one
when the pool has room under its memory budget, and otherwise evicts the least recently
used idle slot of another size; when the busy slots alone fill the budget it waits for one
to be released.  Every kernel sums in a fixed order and a context's results do not depend
on what else runs, so a request's result is bit-identical alone and in the mix.

Latency is measured from a request's arrival (its submit time) to the moment its result is
on the host, including any wait for a worker or a slot, slot creation, the evaluation and
the read-back of the forces.
"""
import collections
import threading
import time

import numpy as np

from . import elec_params as _ep
from . import water_params as _wp
from .engine import Slot

MB = 1 << 20


def slot_bytes(nw, cluster=False):
    """Device memory of a slot of nw waters (estimate from the buffer sizes: MB-pol cluster
    partials dominate in a periodic box, about 0.55 MB per water at liquid density, plus
    fixed tables).  An isolated cluster (imcon 0) takes every site pair into its
    electrostatics rows: (4 nw)^2 dipole tensors and partner indices, 56 bytes a pair."""
    return int(0.6 * MB * nw + 24 * MB + (56 * (4 * nw) ** 2 if cluster else 0))


def _pool_key(box):
    return (box.nw, box.imcon == 0)


class Request:
    __slots__ = ("rid", "box", "elec", "water", "tag", "t_arrival", "t_start", "t_done", "result", "forces",
                 "error", "done")

    def __init__(self, rid, box, elec, water, tag):
        self.rid, self.box, self.elec, self.water, self.tag = rid, box, elec, water, tag
        self.t_arrival = self.t_start = self.t_done = None
        self.result = self.forces = self.error = None
        self.done = threading.Event()

    @property
    def latency(self):
        return None if self.t_done is None else self.t_done - self.t_arrival


class Server:
    def __init__(self, workers=8, memory_budget=16 << 30, precision=0, solver=1, return_forces=True):
        self.precision, self.solver, self.return_forces = precision, solver, return_forces
        self.budget = memory_budget
        self._queue = collections.deque()
        self._cv = threading.Condition()
        self._idle = collections.OrderedDict()      # slot id -> slot (least recently used first)
        self._used_bytes = 0
        self._next_id = 0
        self._closing = False
        self.slots_created = 0
        self.slots_evicted = 0
        self._threads = [threading.Thread(target=self._work, daemon=True) for _ in range(workers)]
        for t in self._threads:
            t.start()

    # ---------------------------------------------------------------- client side ----
    def submit(self, box, elec=None, water=None, tag=None):
        with self._cv:
            req = Request(self._next_id, box, elec, water, tag)
            self._next_id += 1
            req.t_arrival = time.perf_counter()
            self._queue.append(req)
            self._cv.notify_all()
        return req

    def close(self):
        with self._cv:
            self._closing = True
            self._cv.notify_all()
        for t in self._threads:
            t.join()
        self._idle.clear()

    # ---------------------------------------------------------------- slot pool ----
    def _acquire(self, nw, box):
        """An idle slot for a configuration like box (nw waters, periodic or a cluster), or a
        new one within the budget (called with the lock)."""
        key = _pool_key(box)
        for sid, s in self._idle.items():
            if _pool_key(s.box) == key:
                del self._idle[sid]
                return s, False
        need = slot_bytes(nw, box.imcon == 0)
        while self._used_bytes + need > self.budget and self._idle:
            _, victim = self._idle.popitem(last=False)    # least recently used
            self._used_bytes -= slot_bytes(victim.box.nw, victim.box.imcon == 0)
            self.slots_evicted += 1
            del victim
        if self._used_bytes + need > self.budget:
            return None, False
        self._used_bytes += need
        return None, True

    def _release(self, slot):
        with self._cv:
            self._idle[id(slot)] = slot
            self._cv.notify_all()

    # ---------------------------------------------------------------- workers ----
    def _work(self):
        while True:
            with self._cv:
                while not self._queue and not self._closing:
                    self._cv.wait()
                if not self._queue:
                    return
                req = self._queue.popleft()
                need = slot_bytes(req.box.nw, req.box.imcon == 0)
                if need > self.budget:
                    req.error = MemoryError(f"a slot of {req.box.nw} waters needs {need} bytes, "
                                            f"more than the server's budget of {self.budget}")
                    req.t_done = time.perf_counter()
                    req.done.set()
                    continue
                slot, create = self._acquire(req.box.nw, req.box)
                while slot is None and not create:
                    self._cv.wait()
                    slot, create = self._acquire(req.box.nw, req.box)
            req.t_start = time.perf_counter()
            try:
                elec = req.elec or _ep.for_box(req.box.L)
                water = req.water or _wp.for_box(req.box.L)
                if create:
                    slot = Slot(req.box, precision=self.precision, solver=self.solver, elec=elec, water=water)
                    self.slots_created += 1
                else:
                    slot.set_params(elec, water)
                    slot.load(req.box)
                req.result = slot.evaluate()
                if self.return_forces:
                    req.forces = slot.forces()
            except Exception as e:        # reported on the request
                req.error = e
                if create and slot is None:
                    with self._cv:
                        self._used_bytes -= slot_bytes(req.box.nw, req.box.imcon == 0)
            req.t_done = time.perf_counter()
            if slot is not None:
                self._release(slot)
            req.done.set()


def run_poisson(server, jobs, rate, seed=0):
    """Submit jobs (list of (box, elec, water, tag)) at Poisson arrival times of the given
    rate (per second) and wait for all.  Returns (requests, wall seconds)."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, len(jobs))
    t0 = time.perf_counter()
    due = t0
    reqs = []
    for (box, elec, water, tag), g in zip(jobs, gaps):
        due += g
        delay = due - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        reqs.append(server.submit(box, elec, water, tag))
    for r in reqs:
        r.done.wait()
    return reqs, time.perf_counter() - t0


def latency_summary(reqs):
    lat = np.array([r.latency for r in reqs if r.error is None])
    errs = sorted({f"{type(r.error).__name__}: {r.error}" for r in reqs if r.error is not None})
    if lat.size == 0:
        return {"errors": len(reqs), "error_messages": errs[:3]}
    return {"requests": int(lat.size), "errors": sum(r.error is not None for r in reqs), "error_messages": errs[:3],
            "p50_ms": float(1e3 * np.percentile(lat, 50)), "p99_ms": float(1e3 * np.percentile(lat, 99)),
            "mean_ms": float(1e3 * lat.mean()), "max_ms": float(1e3 * lat.max())}
