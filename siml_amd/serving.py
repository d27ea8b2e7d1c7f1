"""Continuous-batching request loop over a resident LINGO database (no
counterpart in the reference).

A ``Server`` keeps one ``LingoDB`` in device memory and runs ``slots``
workers.  Each worker owns an executor -- by default a comparator attached
to the database, with its own HIP stream.  Requests -- a batch of query
molecules plus either ``("topk", k)`` or ``("threshold", lo)`` -- wait in an
admission queue; whenever a worker is free it takes the oldest waiting
requests of one kind (same k or same lo), up to ``max_rows`` query molecules
and within the device-memory budget, and runs them as one launch sequence.
Requests that arrive while other batches run on the device are admitted as
soon as a slot and memory free up, and the batches of different slots
execute concurrently on the device.

Memory admission.  A batch is charged an upper bound of the device memory
its call allocates: the reference set it uploads, plus the launch scratch
and result that ``siml_call_bytes`` derives from the same launch-shape code
the call runs (every candidate shape while the autotuner is on).  Neighbour
lists are data dependent; they are charged ``nbr_allowance`` bytes per query
molecule.  A request whose charge alone exceeds the budget is rejected with
``MemoryError``; otherwise the oldest request is always admitted once
nothing else holds memory, so a large request cannot starve.

Failures.  A request can be cancelled while it waits (``Request.cancel``)
and can carry a time limit for waiting (``timeout``): either ends it with
``RequestCancelled`` / ``RequestTimeout`` without running it.  An error of
a batch's call (an allocation failure, a device error) is set on every
request of that batch; the server keeps serving.  After a device error the
slot's executor is replaced; a slot whose executor cannot be rebuilt stops,
and when no slot is left the waiting requests end with the error.

Every query row is computed independently with integer intersections and
the reference's float expression, so a request's result does not depend on
what it was batched with.
"""
import threading
import time
from collections import deque

import numpy as np

from . import _lib


class RequestCancelled(Exception):
    """The request was cancelled before it ran."""


class RequestTimeout(Exception):
    """The request waited longer than its time limit."""


class Request:
    """One query batch: ragged query molecules (ids, counts, offsets, mags)
    and an operation ``("topk", k)`` or ``("threshold", lo)``.  ``timeout``
    (seconds, optional) limits how long it may wait before it runs."""

    __slots__ = ("rid", "ids", "counts", "offsets", "mags", "op", "param", "arrival", "start", "done", "result",
                 "error", "rows", "event", "timeout", "cancelled", "charge")

    def __init__(self, rid, ids, counts, offsets, mags, op, param, arrival=None, timeout=None):
        if op not in ("topk", "threshold"):
            raise ValueError(f"unknown operation {op!r}")
        self.rid = rid
        self.ids = np.ascontiguousarray(ids, np.int32)
        self.counts = np.ascontiguousarray(counts, np.int32)
        self.offsets = np.ascontiguousarray(offsets, np.int64)
        self.mags = np.ascontiguousarray(mags, np.int32)
        self.op = op
        self.param = param
        self.arrival = arrival
        self.start = None
        self.done = None
        self.result = None
        self.error = None
        self.rows = len(self.offsets) - 1
        self.event = threading.Event()
        self.timeout = timeout
        self.cancelled = False
        self.charge = None  # device bytes charged at admission, computed once

    @property
    def key(self):
        return (self.op, self.param)

    @property
    def latency(self):
        return None if self.done is None else self.done - self.arrival

    def lingos(self):
        return int(self.offsets[-1] - self.offsets[0])

    def levels(self):
        """Levels the bit-sliced tables hold for these molecules (at most
        15 per Lingo): an upper bound of their table entries."""
        c = self.counts[self.offsets[0]:self.offsets[-1]]
        return int(np.minimum(np.maximum(c, 0), 15).sum())

    def cancel(self):
        """Marks the request cancelled: if it has not started it ends with
        RequestCancelled and never runs.  Returns False when it had
        already started."""
        self.cancelled = True
        return self.start is None


def concat_requests(reqs):
    """One ragged set holding the query molecules of reqs in order."""
    ids = np.concatenate([r.ids[r.offsets[0]:r.offsets[-1]] for r in reqs]) if reqs else np.zeros(0, np.int32)
    counts = np.concatenate([r.counts[r.offsets[0]:r.offsets[-1]] for r in reqs]) if reqs else np.zeros(0, np.int32)
    lens = np.concatenate([np.diff(r.offsets) for r in reqs]) if reqs else np.zeros(0, np.int64)
    offs = np.zeros(len(lens) + 1, np.int64)
    np.cumsum(lens, out=offs[1:])
    mags = np.concatenate([r.mags for r in reqs]) if reqs else np.zeros(0, np.int32)
    return ids, counts, offs, mags


def split_results(batch, op, out):
    """Sets each request's slice of a batch result: (idx, t) rows for
    top-k, CSR (offsets from 0, columns) for threshold requests."""
    r0 = 0
    if op == "topk":
        idx, t = out
        for r in batch:
            r.result = (idx[r0:r0 + r.rows], t[r0:r0 + r.rows])
            r0 += r.rows
    else:
        roff, cols = out
        for r in batch:
            o = roff[r0:r0 + r.rows + 1]
            r.result = (o - o[0], cols[o[0]:o[-1]])
            r0 += r.rows


class GpuExecutor:
    """Runs batches on one comparator (its own HIP stream) attached to the
    database."""

    def __init__(self, db, device=0, autotune=False, group_tables=2, graphs=False):
        from .GPULingo import GPULingo
        self.g = GPULingo(device)
        self.g.attach_db(db)
        self.g.set_group_tables(group_tables)
        if autotune:
            self.g.set_autotune(True)
        if graphs:
            self.g.set_graphs(True)

    def call_bytes(self, op, rows, k):
        """siml_call_bytes of this comparator."""
        import ctypes
        b = ctypes.c_int64()
        code = _lib.OP_TOPK if op == "topk" else _lib.OP_NEIGHBORS
        self.g._h.call("siml_call_bytes", int(code), int(rows), int(k if op == "topk" else 0), ctypes.byref(b))
        return b.value

    def run(self, batch):
        ids, counts, offs, mags = concat_requests(batch)
        self.g.set_refsmiles_ragged(ids, counts, offs, mags)
        n = len(offs) - 1
        op, param = batch[0].key
        if op == "topk":
            return self.g.getTopK(0, n, int(param))
        return self.g.getNeighborsCSR(0, n, float(param))


class Server:
    """Continuous batching over ``slots`` executors sharing ``db``.

    Each default executor is a comparator with its own HIP stream, so the
    kernels of different slots run concurrently.  Two slots (the default)
    measured the lowest p99 on the 10M-molecule database up to 2000
    batches/s: one slot leaves the device idle between batches, four let
    large batches share the device with three others and stretch the tail;
    three gain only near saturation (2500 batches/s).

    ``executor_factory(slot)`` (optional) builds a slot's executor: an
    object with ``run(batch)`` returning the batch's (idx, t) or (offsets,
    columns) and ``call_bytes(op, rows, k)``; the default builds
    ``GpuExecutor``s, whose top-k batches run as HIP graphs when
    ``graphs`` is set (siml_set_graphs).
    """

    def __init__(self, db, slots=2, max_rows=2048, mem_budget=8 << 30, device=0, autotune=False, group_tables=2,
                 nbr_allowance=64 << 10, executor_factory=None, graphs=False):
        self.db = db
        self.max_rows = int(max_rows)
        self.mem_budget = int(mem_budget)
        self.nbr_allowance = int(nbr_allowance)
        self.mem_used = 0
        self.queue = deque()
        self.cv = threading.Condition()
        self.stopping = False
        self.batches = 0
        self.workers = []
        if executor_factory is None:
            def executor_factory(slot):
                return GpuExecutor(db, device, autotune, group_tables, graphs)
        self.executor_factory = executor_factory
        self.executors = [executor_factory(i) for i in range(slots)]
        self.alive = slots
        self.fatal = None
        for i in range(slots):
            t = threading.Thread(target=self._worker, args=(i,), daemon=True)
            t.start()
            self.workers.append(t)

    # -- admission ----------------------------------------------------------
    def memory_estimate(self, req):
        """Device bytes charged for `req` in a batch: its reference set
        (8 B per Lingo, 24 B per molecule, at most 16 B per level of the
        row-group tables and a group header per 32 molecules) plus the
        call's scratch and result bound; threshold requests add
        nbr_allowance bytes per molecule for their neighbour lists.
        Computed once per request."""
        if req.charge is not None:
            return req.charge
        rows = req.rows
        k = int(req.param) if req.op == "topk" else 0
        ex = next(e for e in self.executors if e is not None)
        ref = 8 * req.lingos() + 24 * rows + 16 * req.levels() + 160 * (rows // 32 + 1)
        est = ref + ex.call_bytes(req.op, rows, k)
        if req.op == "threshold":
            est += self.nbr_allowance * rows
        req.charge = est
        return est

    def submit(self, req):
        if req.arrival is None:
            req.arrival = time.perf_counter()
        with self.cv:
            if self.fatal is not None:
                self._finish(req, self.fatal)
                return req
            self.queue.append(req)
            self.cv.notify()
        return req

    def _finish(self, req, error=None, result=None):
        req.error = error
        if result is not None:
            req.result = result
        req.done = time.perf_counter()
        req.event.set()

    def _expire(self):
        """Ends the waiting requests that were cancelled, that waited past
        their time limit, or that can never fit the memory budget."""
        now = time.perf_counter()
        for r in list(self.queue):
            err = None
            if r.cancelled:
                err = RequestCancelled(f"request {r.rid} was cancelled before it ran")
            elif r.timeout is not None and now - r.arrival > r.timeout:
                err = RequestTimeout(f"request {r.rid} waited {now - r.arrival:.3f} s, more than its "
                                     f"{r.timeout} s limit")
            else:
                need = self.memory_estimate(r)
                if need > self.mem_budget:
                    err = MemoryError(f"request {r.rid} needs {need} bytes of device memory, more than the "
                                      f"{self.mem_budget}-byte budget")
            if err is not None:
                self.queue.remove(r)
                self._finish(r, err)

    def _take_batch(self):
        """Oldest requests with the key of the head of the queue that fit the
        row and memory limits (the head is always admitted once no memory is
        held, so a large request cannot starve)."""
        self._expire()
        if not self.queue:
            return None, 0
        head = self.queue[0]
        key = head.key
        batch, rows = [], 0
        need = 0
        for r in list(self.queue):
            if r.key != key:
                continue
            if batch and rows + r.rows > self.max_rows:
                break
            est = self.memory_estimate(r)
            if batch and self.mem_used + need + est > self.mem_budget:
                break
            batch.append(r)
            rows += r.rows
            need += est
        if self.mem_used > 0 and self.mem_used + need > self.mem_budget:
            return None, 0
        for r in batch:
            self.queue.remove(r)
            r.start = time.perf_counter()
        self.mem_used += need
        return batch, need

    def _blocked(self):
        head = self.queue[0]
        return self.mem_used > 0 and self.mem_used + self.memory_estimate(head) > self.mem_budget

    def _worker(self, slot):
        while True:
            with self.cv:
                while not self.stopping and (not self.queue or self._blocked()):
                    self._expire()
                    self.cv.wait(0.05)
                if self.stopping and not self.queue:
                    return
                batch, need = self._take_batch()
                if not batch:
                    continue
            ex = self.executors[slot]
            error, out = None, None
            try:
                out = ex.run(batch)
            except Exception as e:  # surfaced on the batch's requests, the loop keeps serving
                error = e
            if out is not None:
                split_results(batch, batch[0].op, out)
            replace = isinstance(error, _lib.SimlDeviceError)
            with self.cv:
                self.mem_used -= need
                self.batches += 1
                self.cv.notify_all()
            for r in batch:  # a cancel that came after the start leaves the result
                self._finish(r, error)
            if replace and not self._replace_executor(slot, error):
                return

    def _replace_executor(self, slot, error):
        """Rebuilds a slot's executor after a device error; a slot that
        cannot be rebuilt stops, and the last one to stop ends the waiting
        requests with the error."""
        try:
            self.executors[slot] = self.executor_factory(slot)
            return True
        except Exception as e:
            with self.cv:
                self.executors[slot] = None
                self.alive -= 1
                if self.alive == 0:
                    self.fatal = e
                    while self.queue:
                        self._finish(self.queue.popleft(), e)
                self.cv.notify_all()
            return False

    def close(self):
        with self.cv:
            self.stopping = True
            self.cv.notify_all()
        for t in self.workers:
            t.join()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def run_poisson(server, requests, rate, seed=0):
    """Submits `requests` with exponential inter-arrival times of mean
    1/rate seconds (their arrival times are the scheduled ones), waits for
    all, and returns (wall seconds, latencies in seconds)."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, size=len(requests))
    t0 = time.perf_counter()
    due = t0
    for req, gap in zip(requests, gaps):
        due += gap
        now = time.perf_counter()
        if due > now:
            time.sleep(due - now)
        req.arrival = due
        server.submit(req)
    for req in requests:
        req.event.wait()
    wall = time.perf_counter() - t0
    return wall, np.array([r.latency for r in requests])
