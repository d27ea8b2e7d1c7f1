"""Continuous-batching request loop over a resident LINGO database (no
counterpart in the reference).

A ``Server`` keeps one ``LingoDB`` in device memory and runs ``slots``
workers.  Each worker owns a comparator (its own HIP stream) attached to the
database.  Requests -- a batch of query molecules plus either ``("topk", k)``
or ``("threshold", lo)`` -- wait in an admission queue; whenever a worker is
free it takes the oldest waiting requests of one kind (same k or same lo),
up to ``max_rows`` query molecules and within the device-memory budget, and
runs them as one launch.  Requests that arrive while other batches run on
the device are admitted as soon as a slot and memory free up, and the
batches of different slots execute concurrently on the device.

Every query row is computed independently with integer intersections and
the reference's float expression, so a request's result does not depend on
what it was batched with.
"""
import threading
import time
from collections import deque

import numpy as np

from .GPULingo import GPULingo


class Request:
    """One query batch: ragged query molecules (ids, counts, offsets, mags)
    and an operation ``("topk", k)`` or ``("threshold", lo)``."""

    __slots__ = ("rid", "ids", "counts", "offsets", "mags", "op", "param", "arrival", "start", "done", "result",
                 "error", "rows", "event")

    def __init__(self, rid, ids, counts, offsets, mags, op, param, arrival=None):
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

    @property
    def key(self):
        return (self.op, self.param)

    @property
    def latency(self):
        return None if self.done is None else self.done - self.arrival


def concat_requests(reqs):
    """One ragged set holding the query molecules of reqs in order."""
    ids = np.concatenate([r.ids[r.offsets[0]:r.offsets[-1]] for r in reqs]) if reqs else np.zeros(0, np.int32)
    counts = np.concatenate([r.counts[r.offsets[0]:r.offsets[-1]] for r in reqs]) if reqs else np.zeros(0, np.int32)
    lens = np.concatenate([np.diff(r.offsets) for r in reqs]) if reqs else np.zeros(0, np.int64)
    offs = np.zeros(len(lens) + 1, np.int64)
    np.cumsum(lens, out=offs[1:])
    mags = np.concatenate([r.mags for r in reqs]) if reqs else np.zeros(0, np.int32)
    return ids, counts, offs, mags


class Server:
    """Continuous batching over ``slots`` comparators sharing ``db``.

    Each slot is a handle with its own HIP stream, so the kernels of
    different slots run concurrently.  Two slots (the default) measured the
    lowest p99 on the 10M-molecule database up to 2000 batches/s: one slot
    leaves the device idle between batches, four let large batches share
    the device with three others and stretch the tail; three gain only
    near saturation (2500 batches/s).
    """

    def __init__(self, db, slots=2, max_rows=2048, mem_budget=8 << 30, device=0, autotune=False, group_tables=2):
        self.db = db
        self.max_rows = int(max_rows)
        self.mem_budget = int(mem_budget)
        self.mem_used = 0
        self.queue = deque()
        self.cv = threading.Condition()
        self.stopping = False
        self.batches = 0
        self.workers = []
        self.handles = []
        for _ in range(slots):
            g = GPULingo(device)
            g.attach_db(db)
            g.set_group_tables(group_tables)
            if autotune:
                g.set_autotune(True)
            self.handles.append(g)
        for i in range(slots):
            t = threading.Thread(target=self._worker, args=(i,), daemon=True)
            t.start()
            self.workers.append(t)

    # -- admission ----------------------------------------------------------
    def memory_estimate(self, rows, op, param):
        """Device bytes a batch of `rows` query molecules may hold: its query
        set, the per-(row, wave) candidate buffers or neighbour counts, and
        the result."""
        waves = 256
        if op == "topk":
            cap = ((int(param) + 64 + 63) // 64) * 64
            return rows * (waves * cap * 8 + waves * 4 + int(param) * 8) + rows * 1024
        return rows * (waves * 12 + 64 * 1024) + rows * 1024

    def submit(self, req):
        if req.arrival is None:
            req.arrival = time.perf_counter()
        with self.cv:
            self.queue.append(req)
            self.cv.notify()
        return req

    def _take_batch(self):
        """Oldest requests with the key of the head of the queue that fit the
        row and memory limits (the head is always admitted once memory is
        free, so a large request cannot starve)."""
        head = self.queue[0]
        key = head.key
        batch, rows = [], 0
        need = 0
        for r in list(self.queue):
            if r.key != key:
                continue
            if batch and rows + r.rows > self.max_rows:
                break
            est = self.memory_estimate(r.rows, *key)
            if batch and self.mem_used + need + est > self.mem_budget:
                break
            batch.append(r)
            rows += r.rows
            need += est
        if self.mem_used + need > self.mem_budget and self.mem_used > 0:
            return None, 0
        for r in batch:
            self.queue.remove(r)
        self.mem_used += need
        return batch, need

    def _worker(self, slot):
        g = self.handles[slot]
        while True:
            with self.cv:
                while not self.stopping and (not self.queue or self._blocked()):
                    self.cv.wait(0.05)
                if self.stopping and not self.queue:
                    return
                batch, need = self._take_batch()
                if not batch:
                    continue
            t0 = time.perf_counter()
            for r in batch:
                r.start = t0
            try:
                self._run(g, batch)
            except Exception as e:  # surfaced on the request, the loop keeps serving
                for r in batch:
                    r.error = e
            t1 = time.perf_counter()
            with self.cv:
                self.mem_used -= need
                self.batches += 1
                self.cv.notify_all()
            for r in batch:
                r.done = t1
                r.event.set()

    def _blocked(self):
        head = self.queue[0]
        return self.mem_used > 0 and self.mem_used + self.memory_estimate(head.rows, *head.key) > self.mem_budget

    def _run(self, g, batch):
        ids, counts, offs, mags = concat_requests(batch)
        g.set_refsmiles_ragged(ids, counts, offs, mags)
        n = len(offs) - 1
        op, param = batch[0].key
        if op == "topk":
            idx, t = g.getTopK(0, n, int(param))
            r0 = 0
            for r in batch:
                r.result = (idx[r0:r0 + r.rows], t[r0:r0 + r.rows])
                r0 += r.rows
        else:
            roff, cols = g.getNeighborsCSR(0, n, float(param))
            r0 = 0
            for r in batch:
                o = roff[r0:r0 + r.rows + 1]
                r.result = (o - o[0], cols[o[0]:o[-1]])
                r0 += r.rows

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
