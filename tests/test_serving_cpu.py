"""The serving loop's scheduler on the CPU, with an injected executor in
place of the GPU comparators: batching rules, memory admission, rejection,
cancellation and time limits, and the propagation of allocation failures
and device errors.  The executor computes a deterministic per-row result,
so each request's slice of a batch can be checked against the request run
alone."""
import threading
import time

import numpy as np
import pytest

from siml_amd import _lib, synth
from siml_amd.serving import Request, RequestCancelled, RequestTimeout, Server


def row_key(ids, counts):
    return int(ids.astype(np.int64).sum() * 7 + counts.astype(np.int64).sum())


def fake_result(ids, counts, offs, op, param):
    """Per row: top-k rows of k columns derived from the row's contents;
    threshold rows with (row key % 5) columns."""
    n = len(offs) - 1
    keys = [row_key(ids[offs[i]:offs[i + 1]], counts[offs[i]:offs[i + 1]]) for i in range(n)]
    if op == "topk":
        k = int(param)
        idx = np.array([[(key + j) % 100003 for j in range(k)] for key in keys], np.int32).reshape(n, k)
        return idx, (idx.astype(np.float32) / 100003.0)
    lens = np.array([key % 5 for key in keys], np.int64)
    roff = np.zeros(n + 1, np.int64)
    np.cumsum(lens, out=roff[1:])
    cols = np.concatenate([np.arange(l, dtype=np.int32) + (key % 1000) for key, l in zip(keys, lens)]) \
        if n else np.zeros(0, np.int32)
    return roff, cols.astype(np.int32)


class FakeExecutor:
    """call_bytes: per_row bytes per row; run: optional gate and failure."""

    def __init__(self, shared, per_row=1000, fail=None, gate=None, delay=0.0):
        self.s = shared
        self.per_row = per_row
        self.fail = fail
        self.gate = gate
        self.delay = delay

    def call_bytes(self, op, rows, k):
        return self.per_row * rows

    def run(self, batch):
        from siml_amd.serving import concat_requests
        with self.s["lock"]:
            self.s["running"] += 1
            self.s["most"] = max(self.s["most"], self.s["running"])
            self.s["batches"].append([r.rid for r in batch])
        try:
            if self.gate is not None:
                self.gate.wait(5.0)
            if self.delay:
                time.sleep(self.delay)
            if self.fail is not None:
                err = self.fail(batch)
                if err is not None:
                    raise err
            ids, counts, offs, _ = concat_requests(batch)
            return fake_result(ids, counts, offs, *batch[0].key)
        finally:
            with self.s["lock"]:
                self.s["running"] -= 1


def shared():
    return {"lock": threading.Lock(), "running": 0, "most": 0, "batches": []}


@pytest.fixture(scope="module")
def pool():
    return synth.packed_ragged(3000, 8, 120, seed=91, dist="lognormal", median=45.0, sigma=0.35)


def requests(pool, sizes, ops=None, **kw):
    out, o = [], 0
    for i, n in enumerate(sizes):
        op, prm = (ops[i] if ops else (("topk", 5) if i % 2 == 0 else ("threshold", 0.5)))
        out.append(Request(i, *synth.ragged_slice(*pool, o, o + n), op, prm, **kw))
        o += n
    return out


def alone(r):
    offs = r.offsets - r.offsets[0]
    return fake_result(r.ids[r.offsets[0]:r.offsets[-1]], r.counts[r.offsets[0]:r.offsets[-1]], offs, r.op, r.param)


def wait_all(reqs):
    for r in reqs:
        assert r.event.wait(10.0), f"request {r.rid} did not finish"


def test_results_split_per_request_and_batching_rules(pool):
    sh = shared()
    gate = threading.Event()
    reqs = requests(pool, [3, 5, 1, 40, 7, 2, 30, 9, 11, 4, 1, 60])
    with Server(None, slots=1, max_rows=64, mem_budget=1 << 40,
                executor_factory=lambda s: FakeExecutor(sh, gate=gate)) as srv:
        first = srv.submit(requests(pool, [1], [("topk", 5)])[0])
        time.sleep(0.05)           # the first request holds the slot; the rest queue up
        for r in reqs:
            srv.submit(r)
        gate.set()
        wait_all([first] + reqs)
    for r in reqs:
        assert r.error is None
        want = alone(r)
        assert all(np.array_equal(a, b) for a, b in zip(r.result, want)), r.rid
    by_id = {r.rid: r for r in reqs}
    for b in sh["batches"][1:]:
        keys = {by_id[i].key for i in b}
        assert len(keys) == 1, "a batch mixes operations"
        assert sum(by_id[i].rows for i in b) <= 64 or len(b) == 1
    assert len(sh["batches"]) < len(reqs) + 1, "nothing was batched"


def test_request_that_cannot_fit_is_rejected(pool):
    sh = shared()
    per_row = 1 << 20
    reqs = requests(pool, [2, 40, 3], [("topk", 5)] * 3)
    with Server(None, slots=2, max_rows=1024, mem_budget=16 * per_row,
                executor_factory=lambda s: FakeExecutor(sh, per_row=per_row)) as srv:
        for r in reqs:
            srv.submit(r)
        wait_all(reqs)
    assert isinstance(reqs[1].error, MemoryError) and "budget" in str(reqs[1].error)
    assert reqs[0].error is None and reqs[2].error is None
    assert all(1 not in b for b in sh["batches"]), "the rejected request ran"


def test_admission_bounds_concurrent_memory(pool):
    """Each request is charged about 4 MiB; with a 9 MiB budget at most two
    run at once although there are four slots."""
    sh = shared()
    per_row = 1 << 20
    reqs = requests(pool, [4] * 12, [("topk", 5)] * 12)
    with Server(None, slots=4, max_rows=4, mem_budget=9 * per_row,
                executor_factory=lambda s: FakeExecutor(sh, per_row=per_row, delay=0.02)) as srv:
        for r in reqs:
            srv.submit(r)
        wait_all(reqs)
        assert srv.mem_used == 0
    assert all(r.error is None for r in reqs)
    assert sh["most"] == 2


def test_allocation_failure_reported_and_serving_continues(pool):
    sh = shared()
    boom = lambda batch: MemoryError("hipMalloc failed: out of memory") if any(r.rid == 3 for r in batch) else None
    reqs = requests(pool, [2] * 8, [("topk", 5)] * 8)
    with Server(None, slots=1, max_rows=2, executor_factory=lambda s: FakeExecutor(sh, fail=boom)) as srv:
        for r in reqs:
            srv.submit(r)
        wait_all(reqs)
    assert isinstance(reqs[3].error, MemoryError) and reqs[3].result is None
    assert all(r.error is None and r.result is not None for r in reqs if r.rid != 3)


def test_device_error_rebuilds_the_slot(pool):
    sh = shared()
    made = []
    failed = []

    def fail(batch):
        if not failed:
            failed.append(batch[0].rid)
            return _lib.SimlDeviceError("kernel launch failed")
        return None

    def factory(slot):
        made.append(slot)
        return FakeExecutor(sh, fail=fail)

    reqs = requests(pool, [2] * 6, [("threshold", 0.5)] * 6)
    with Server(None, slots=1, max_rows=2, executor_factory=factory) as srv:
        for r in reqs:
            srv.submit(r)
        wait_all(reqs)
    assert made == [0, 0], "the slot's executor was not rebuilt"
    bad = [r for r in reqs if r.error is not None]
    assert [r.rid for r in bad] == failed and isinstance(bad[0].error, _lib.SimlDeviceError)
    assert all(r.result is not None for r in reqs if r.error is None)


def test_unrecoverable_device_error_ends_waiting_requests(pool):
    sh = shared()
    made = []
    gate = threading.Event()

    def factory(slot):
        if made:
            raise _lib.SimlDeviceError("no HIP device is visible")
        made.append(slot)
        return FakeExecutor(sh, gate=gate, fail=lambda b: _lib.SimlDeviceError("device lost"))

    reqs = requests(pool, [2] * 5, [("topk", 5)] * 5)
    with Server(None, slots=1, max_rows=2, executor_factory=factory) as srv:
        for r in reqs:
            srv.submit(r)
        time.sleep(0.05)
        gate.set()
        wait_all(reqs)
        late = srv.submit(requests(pool, [1], [("topk", 5)])[0])
        assert late.event.wait(1.0) and isinstance(late.error, _lib.SimlDeviceError)
    assert all(isinstance(r.error, _lib.SimlDeviceError) for r in reqs)


def test_cancel_and_time_limit_while_waiting(pool):
    sh = shared()
    gate = threading.Event()
    first, r_cancel, r_late, r_ok = requests(pool, [1, 2, 2, 2], [("topk", 5)] * 4)
    r_late.timeout = 0.05
    with Server(None, slots=1, max_rows=1, executor_factory=lambda s: FakeExecutor(sh, gate=gate)) as srv:
        srv.submit(first)
        time.sleep(0.05)
        for r in (r_cancel, r_late, r_ok):
            srv.submit(r)
        assert r_cancel.cancel()
        time.sleep(0.2)
        gate.set()
        wait_all([first, r_cancel, r_late, r_ok])
    assert isinstance(r_cancel.error, RequestCancelled) and r_cancel.result is None
    assert isinstance(r_late.error, RequestTimeout) and r_late.result is None
    assert first.error is None and r_ok.error is None
    ran = {i for b in sh["batches"] for i in b}
    assert r_cancel.rid not in ran and r_late.rid not in ran
    assert not first.cancel(), "cancel after the start must report that it came too late"
