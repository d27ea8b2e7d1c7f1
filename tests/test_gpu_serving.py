"""Resident database, batched top-k and the continuous-batching serving
loop on the GPU: results are bit-identical to the oracle and do not depend
on what a request is batched with."""
import numpy as np
import pytest

import oracle
from siml_amd import _lib, compiler, synth
from siml_amd.GPULingo import GPULingo, LingoDB
from siml_amd.serving import Request, Server

pytestmark = pytest.mark.gpu


def bits(a):
    return np.asarray(a, np.float32).view(np.int32)


@pytest.fixture(scope="module")
def db_and_queries():
    buf, offs = synth.random_strings(20000, 8, 256, seed=9, dist="lognormal")
    db_set = compiler.pack_ragged_buffer(buf, offs)
    rng = np.random.default_rng(10)
    picks = rng.choice(20000, size=600, replace=False)
    qb, qo = synth.mutated_strings(buf, offs, picks, 3, seed=11)
    q = compiler.pack_ragged_buffer(qb, qo)
    return db_set, q


@pytest.mark.parametrize("layout", [_lib.LAYOUT_NATURAL, _lib.LAYOUT_BLOCK, _lib.LAYOUT_SORTED])
def test_resident_db_matches_oracle(db_and_queries, layout):
    db_set, q = db_and_queries
    db = LingoDB(*db_set, layout=layout)
    g = GPULingo(0)
    g.attach_db(db)
    g.set_refsmiles_ragged(*synth.ragged_slice(*q, 0, 40))
    Q = synth.padded_columns(*q, 0, 40)
    D = synth.padded_columns(*db_set, 0, len(db_set[2]) - 1)
    want = oracle.tanimoto_block(Q, D, nthreads=0)
    got = g.getMultipleRows(0, 40)
    assert np.array_equal(bits(got), bits(want))
    ti, tt = g.getTopK(0, 40, 25)
    oi, ot = oracle.topk(Q, D, 25)
    assert np.array_equal(ti, oi) and np.array_equal(bits(tt), bits(ot))
    offs, idx = g.getNeighborsCSR(0, 40, 0.5)
    lists, counts = oracle.neighbors(Q, D, 0.5)
    assert counts.sum() > 0
    assert np.array_equal(np.diff(offs), counts) and np.array_equal(idx, np.concatenate(lists))
    assert np.array_equal(g.getGlobalHistogram(0, 40), oracle.hist1000(Q, D))
    db.close()


def test_batched_equals_per_set(db_and_queries):
    """One launch over concatenated uneven query sets gives each set the
    same top-k as running the set alone."""
    db_set, q = db_and_queries
    db = LingoDB(*db_set, layout=_lib.LAYOUT_SORTED)
    g = GPULingo(0)
    g.attach_db(db)
    sizes = [1, 2, 27, 28, 29, 5, 100, 3, 57, 1]
    bounds = np.concatenate([[0], np.cumsum(sizes)])
    per = []
    for a, b in zip(bounds[:-1], bounds[1:]):
        g.set_refsmiles_ragged(*synth.ragged_slice(*q, int(a), int(b)))
        per.append(g.getTopK(0, int(b - a), 30))
    g.set_refsmiles_ragged(*synth.ragged_slice(*q, 0, int(bounds[-1])))
    bi, bt = g.getTopK(0, int(bounds[-1]), 30)
    assert np.array_equal(np.concatenate([p[0] for p in per]), bi)
    assert np.array_equal(bits(np.concatenate([p[1] for p in per])), bits(bt))
    # and the oracle over the whole database, for every row
    Q = synth.padded_columns(*q, 0, int(bounds[-1]))
    D = synth.padded_columns(*db_set, 0, len(db_set[2]) - 1)
    oi, ot = oracle.topk(Q, D, 30)
    assert np.array_equal(bi, oi) and np.array_equal(bits(bt), bits(ot))
    db.close()


def _requests(q, sizes):
    reqs, o = [], 0
    for i, n in enumerate(sizes):
        op, prm = ("topk", 20) if i % 2 == 0 else ("threshold", 0.5)
        reqs.append(Request(i, *synth.ragged_slice(*q, o, o + n), op, prm))
        o += n
    return reqs


def test_serving_alone_equals_mix(db_and_queries):
    db_set, q = db_and_queries
    db = LingoDB(*db_set, layout=_lib.LAYOUT_SORTED)
    sizes = [1, 5, 17, 40, 3, 64, 2, 100, 33, 8, 12, 70, 1, 1, 9]
    alone = _requests(q, sizes)
    with Server(db, slots=1) as srv:
        for r in alone:
            srv.submit(r).event.wait()
            assert r.error is None
    mix = _requests(q, sizes)
    with Server(db, slots=3, max_rows=128) as srv:
        for r in mix:
            srv.submit(r)
        for r in mix:
            r.event.wait()
            assert r.error is None
        batches = srv.batches
    print(f"{len(mix)} requests served in {batches} batches", flush=True)
    assert batches < len(mix)
    for a, b in zip(alone, mix):
        assert np.array_equal(a.result[0], b.result[0])
        if a.op == "topk":
            assert np.array_equal(bits(a.result[1]), bits(b.result[1]))
        else:
            assert np.array_equal(a.result[1], b.result[1])
    # and against the oracle for a few requests
    D = synth.padded_columns(*db_set, 0, len(db_set[2]) - 1)
    for r in (mix[0], mix[1], mix[7]):
        Q = synth.ragged_to_padded(r.ids, r.counts, r.offsets, r.mags)
        if r.op == "topk":
            oi, ot = oracle.topk(Q, D, 20)
            assert np.array_equal(r.result[0], oi) and np.array_equal(bits(r.result[1]), bits(ot))
        else:
            lists, counts = oracle.neighbors(Q, D, 0.5)
            assert np.array_equal(np.diff(r.result[0]), counts)
            assert np.array_equal(r.result[1], np.concatenate(lists) if lists else np.zeros(0, np.int32))
    db.close()


def test_serving_admission_budget(db_and_queries):
    """A memory budget of 1.5 times the largest request's charge serves
    every request, one at a time; a budget below a request's charge rejects
    it with MemoryError."""
    db_set, q = db_and_queries
    db = LingoDB(*db_set, layout=_lib.LAYOUT_SORTED)
    reqs = _requests(q, [30, 30, 30, 30])
    with Server(db, slots=2, max_rows=64, mem_budget=1 << 40) as probe:
        charges = [probe.memory_estimate(r) for r in reqs]
    for r in reqs:
        r.charge = None
    with Server(db, slots=2, max_rows=64, mem_budget=int(max(charges) * 1.5)) as srv:
        for r in reqs:
            srv.submit(r)
        for r in reqs:
            r.event.wait()
            assert r.error is None and r.result is not None
    small = _requests(q, [30])[0]
    with Server(db, slots=1, mem_budget=min(charges) // 2) as srv:
        srv.submit(small).event.wait()
    assert isinstance(small.error, MemoryError)
    db.close()


def _free_device_bytes():
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


@pytest.mark.parametrize("op", ["topk", "threshold"])
def test_call_bytes_bounds_device_allocation(db_and_queries, op):
    """The device memory one call allocates (free memory before and after,
    on a fresh comparator whose reference set is already uploaded) stays
    within siml_call_bytes' bound, for the default and the autotuned
    launch shapes."""
    import ctypes
    db_set, q = db_and_queries
    db = LingoDB(*db_set)
    sub = synth.ragged_slice(*q, 0, 600)
    for tune in (False, True):
        g = GPULingo(0)
        g.attach_db(db)
        g.set_autotune(tune)
        g.set_refsmiles_ragged(*sub)
        bound = ctypes.c_int64()
        code = _lib.OP_TOPK if op == "topk" else _lib.OP_NEIGHBORS
        g._h.call("siml_call_bytes", code, 600, 100 if op == "topk" else 0, ctypes.byref(bound))
        before = _free_device_bytes()
        if op == "topk":
            g.getTopK(0, 600, 100)
        else:
            offs, idx = g.getNeighborsCSR(0, 600, 0.9)
        used = before - _free_device_bytes()
        slack = 8 << 20   # allocation granularity of a handful of buffers
        nbr = 4 * len(idx) if op == "threshold" else 0
        print(f"{op} autotune={tune}: allocated {used} bytes, bound {bound.value}", flush=True)
        assert bound.value > 0 and used <= bound.value + nbr + slack
        del g
    db.close()


def test_db_file_format_and_load(db_and_queries, tmp_path):
    """The file siml_db_save writes for a NATURAL-layout database is byte
    for byte the file tests/db_file_format.py writes from the format's
    definition; a database loaded from that independent file, and a BLOCK
    layout database saved and loaded again, give the oracle's results."""
    import db_file_format as fmt
    db_set, q = db_and_queries
    ids, counts, offs, mags = db_set
    n = 3000
    sub_db = synth.ragged_slice(ids, counts, offs, mags, 0, n)
    db = LingoDB(*sub_db, layout=_lib.LAYOUT_NATURAL)
    mine = str(tmp_path / "lib.siml")
    db.save(mine)
    spec = str(tmp_path / "spec.siml")
    fmt.write(spec, *sub_db[:2], sub_db[3], sub_db[2])
    with open(mine, "rb") as f1, open(spec, "rb") as f2:
        assert f1.read() == f2.read()
    db.close()
    Q = synth.padded_columns(*q, 100, 160)
    D = synth.padded_columns(*sub_db, 0, n)
    want = oracle.tanimoto_block(Q, D, nthreads=0)
    oi, ot = oracle.topk(Q, D, 10)
    lists, counts_nb = oracle.neighbors(Q, D, 0.5)
    for layout, path in ((None, spec), (_lib.LAYOUT_BLOCK, str(tmp_path / "block.siml"))):
        if layout is not None:
            b = LingoDB(*db_set, layout=layout)
            b.save(path)
            b.close()
            D2 = synth.padded_columns(*db_set, 0, len(db_set[2]) - 1)
            want2 = oracle.tanimoto_block(Q, D2, nthreads=0)
            oi2, ot2 = oracle.topk(Q, D2, 10)
        else:
            want2, oi2, ot2 = want, oi, ot
        loaded = LingoDB.load(path)
        g = GPULingo(0)
        g.attach_db(loaded)
        g.set_refsmiles_ragged(*synth.ragged_slice(*q, 100, 160))
        assert np.array_equal(bits(g.getMultipleRows(0, 60)), bits(want2))
        ti, tt = g.getTopK(0, 60, 10)
        assert np.array_equal(ti, oi2) and np.array_equal(bits(tt), bits(ot2))
        if layout is None:
            o, i = g.getNeighborsCSR(0, 60, 0.5)
            assert np.array_equal(np.diff(o), counts_nb) and np.array_equal(i, np.concatenate(lists))
        loaded.close()


def test_autotune_changes_no_result(db_and_queries):
    """Launch shapes chosen by the autotuner give the same bits as the
    default shapes for every operation."""
    db_set, q = db_and_queries
    db = LingoDB(*db_set)
    sub = synth.ragged_slice(*q, 0, 90)
    out = []
    for tune in (False, True, True):      # the second tuned pass reuses the cached shapes
        g = GPULingo(0)
        g.attach_db(db)
        g.set_autotune(tune)
        g.set_refsmiles_ragged(*sub)
        r = [g.getMultipleRows(0, 90).view(np.int32)]
        r += list(g.getTopK(0, 90, 30))
        r += list(g.getNeighborsCSR(0, 90, 0.5))
        r.append(g.getGlobalHistogram(0, 90))
        r.append(g.getMultipleHistogrammedRows(0, 90))
        out.append(r)
        if tune:
            t, c, _ = g.launch_shape()
            assert t > 0 and c > 0
    for other in out[1:]:
        for a, b in zip(out[0], other):
            assert np.array_equal(np.asarray(a).view(np.int32) if np.asarray(a).dtype == np.float32 else a,
                                  np.asarray(b).view(np.int32) if np.asarray(b).dtype == np.float32 else b)
    db.close()


@pytest.mark.parametrize("shape", [(8192, 256, 16), (2048, 64, 16), (32768, 1024, 0), (1024, 16, 16), (4096, 256, 4)])
def test_topk_launch_shapes_and_seeding(db_and_queries, shape):
    """Top-k is the same for every launch shape, with and without seeding
    the row thresholds from a column sample, on sparse (random) and dense
    (near-duplicate) rows."""
    db_set, q = db_and_queries
    db = LingoDB(*db_set)
    D = synth.padded_columns(*db_set, 0, len(db_set[2]) - 1)
    g = GPULingo(0)
    g.attach_db(db)
    g.set_launch_shape(*shape)
    sub = synth.ragged_slice(*q, 200, 260)
    g.set_refsmiles_ragged(*sub)
    Q = synth.ragged_to_padded(*sub)
    for k in (1, 17, 100):
        ti, tt = g.getTopK(0, 60, k)
        oi, ot = oracle.topk(Q, D, k)
        assert np.array_equal(ti, oi) and np.array_equal(bits(tt), bits(ot))
    assert g.launch_shape() == (shape[0], shape[1], shape[2])
    db.close()


@pytest.mark.parametrize("tables", [1, 2, 4])
def test_group_tables_bit_identical(db_and_queries, tables):
    """Workgroups serving 1, 2 or 4 row groups give the oracle's results
    for top-k (with and without seeding), neighbour CSR (single pass and
    fill pass) and the 1000-bin histogram, also when some rows are served by
    the merge kernel in the same call (counts above 15: the merge kernel
    then indexes the same (row, wave) slots as the bit-sliced kernels)."""
    db_set, q = db_and_queries
    db = LingoDB(*db_set)
    D = synth.padded_columns(*db_set, 0, len(db_set[2]) - 1)
    chains = [b"C" * n + b"\n" for n in (22, 30, 45)]     # counts 19, 27, 42: merge-kernel rows
    Qp = synth.padded_columns(*q, 0, 300)
    Cp = compiler.cSMILEStoMatrices(chains)[:4]
    w = max(Qp[0].shape[1], Cp[0].shape[1])
    pad = lambda m: np.pad(m, ((0, 0), (0, w - m.shape[1])))
    Q = (np.vstack([pad(Qp[0]), pad(Cp[0])]), np.vstack([pad(Qp[1]), pad(Cp[1])]),
         np.vstack([Qp[2], Cp[2]]), np.vstack([Qp[3], Cp[3]]))
    n = Q[0].shape[0]
    for seed in (0, 4):
        g = GPULingo(0)
        g.attach_db(db)
        g.set_group_tables(tables)
        g.set_launch_shape(2048, 256, seed)
        g.set_refsmiles(*Q)
        ti, tt = g.getTopK(0, n, 40)
        oi, ot = oracle.topk(Q, D, 40)
        assert np.array_equal(ti, oi) and np.array_equal(bits(tt), bits(ot))
    for lo in (0.5, 0.0):   # 0.0: every pair, more than 16 per slot (fill pass)
        offs, idx = g.getNeighborsCSR(0, n, lo)
        lists, counts = oracle.neighbors(Q, D, lo)
        assert np.array_equal(np.diff(offs), counts) and np.array_equal(idx, np.concatenate(lists))
    assert np.array_equal(g.getGlobalHistogram(0, n), oracle.hist1000(Q, D))
    assert np.array_equal(g.getGlobalHistogram(17, 290), oracle.hist1000(tuple(x[17:290] for x in Q), D))
    db.close()
