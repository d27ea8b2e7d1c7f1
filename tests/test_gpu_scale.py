"""Large-database checks on the GPU: a 2M-molecule resident database (the
lognormal length law of the 10M serving database) against 512 near-duplicate
queries.  Beside oracle comparisons on sampled rows, they use properties
that hold at any size and tie independent kernels together:

* the 1000-bin histogram counts every pair exactly once;
* the neighbour search (with magnitude-band tile pruning) finds exactly the
  pairs the histogram puts in bins >= 700, since bin k holds T >= fl32(k/1000)
  and the neighbour test is T >= 0.7f = fl32(700/1000);
* the histogram does not depend on the column storage order;
* top-k lists are ordered by (T descending, column ascending) and their k-th
  value bounds the row's neighbour count.
"""
import numpy as np
import pytest

import oracle
from siml_amd import _lib, compiler, synth
from siml_amd.GPULingo import GPULingo, LingoDB

pytestmark = pytest.mark.gpu

NDB = 2_000_000
NQ = 512


@pytest.fixture(scope="module")
def big():
    buf, offs = synth.random_strings(NDB, 8, 256, seed=21, dist="lognormal", median=45.0, sigma=0.35)
    db_set = compiler.pack_ragged_buffer(buf, offs)
    rng = np.random.default_rng(22)
    picks = rng.choice(NDB, size=NQ, replace=False)
    qb, qo = synth.mutated_strings(buf, offs, picks, 2, seed=23)
    q = compiler.pack_ragged_buffer(qb, qo)
    db = LingoDB(*db_set)
    g = GPULingo(0)
    g.attach_db(db)
    g.set_refsmiles_ragged(*q)
    yield db_set, q, db, g
    db.close()


CHUNK = 1 << 18  # database columns per oracle call (bounded host memory)


def _query_rows(q, rows):
    return tuple(x[rows] for x in synth.padded_columns(*q, 0, NQ))


def _chunks(db_set):
    for c0 in range(0, NDB, CHUNK):
        c1 = min(NDB, c0 + CHUNK)
        yield c0, synth.padded_columns(*db_set, c0, c1)


def oracle_hist(q, db_set, rows):
    Q = _query_rows(q, rows)
    return sum(oracle.hist1000(Q, D) for _, D in _chunks(db_set))


def oracle_neighbours(q, db_set, rows, lo):
    Q = _query_rows(q, rows)
    parts = [[] for _ in rows]
    for c0, D in _chunks(db_set):
        lists, _ = oracle.neighbors(Q, D, lo)
        for i, l in enumerate(lists):
            parts[i].append(l.astype(np.int64) + c0)
    return [np.concatenate(p) for p in parts]


def oracle_topk(q, db_set, rows, k):
    """Chunk top-k lists merged in (T descending, column ascending) order."""
    Q = _query_rows(q, rows)
    bi = np.zeros((len(rows), 0), np.int64)
    bt = np.zeros((len(rows), 0), np.float32)
    for c0, D in _chunks(db_set):
        oi, ot = oracle.topk(Q, D, min(k, D[0].shape[0]))
        bi = np.concatenate([bi, oi.astype(np.int64) + c0], axis=1)
        bt = np.concatenate([bt, ot], axis=1)
        order = np.lexsort((bi, -bt.astype(np.float64)), axis=1)[:, :k]
        bi = np.take_along_axis(bi, order, 1)
        bt = np.take_along_axis(bt, order, 1)
    return bi, bt


def test_histogram_counts_every_pair(big):
    db_set, q, db, g = big
    h = g.getGlobalHistogram(0, NQ)
    assert h.dtype == np.int64 and int(h.sum()) == NQ * NDB
    got = g.getGlobalHistogram(0, 1) + g.getGlobalHistogram(NQ - 1, NQ)
    assert np.array_equal(got, oracle_hist(q, db_set, [0, NQ - 1]))


def test_neighbours_equal_histogram_tail(big):
    db_set, q, db, g = big
    h = g.getGlobalHistogram(0, NQ)
    offs, idx = g.getNeighborsCSR(0, NQ, 0.7)
    assert int(offs[-1]) == int(h[700:].sum())
    assert np.all(np.diff(offs) >= 0)
    for r in range(NQ):
        row = idx[offs[r]:offs[r + 1]]
        assert np.all(np.diff(row) > 0)
    rows = [1, NQ // 2]
    lists = oracle_neighbours(q, db_set, rows, 0.7)
    for i, r in enumerate(rows):
        assert np.array_equal(idx[offs[r]:offs[r + 1]], lists[i])


def test_histogram_independent_of_storage_order(big):
    db_set, q, db, g = big
    h = g.getGlobalHistogram(0, NQ)
    other = LingoDB(*db_set, layout=_lib.LAYOUT_BLOCK)
    try:
        g2 = GPULingo(0)
        g2.attach_db(other)
        g2.set_refsmiles_ragged(*q)
        assert np.array_equal(g2.getGlobalHistogram(0, NQ), h)
    finally:
        other.close()


def test_topk_order_and_bounds(big):
    db_set, q, db, g = big
    k = 50
    ti, tt = g.getTopK(0, NQ, k)
    assert ti.shape == (NQ, k) and tt.shape == (NQ, k)
    # (T descending, column ascending) within every row
    later = (tt[:, 1:] < tt[:, :-1]) | ((tt[:, 1:] == tt[:, :-1]) & (ti[:, 1:] > ti[:, :-1]))
    assert bool(np.all(later))
    # a row with c neighbours at T >= 0.7 lists min(c, k) of them first
    offs, idx = g.getNeighborsCSR(0, NQ, 0.7)
    counts = np.diff(offs)
    above = (tt >= np.float32(0.7)).sum(axis=1)
    assert np.array_equal(above, np.minimum(counts, k))
    rows = [2, NQ - 2]
    oi, ot = oracle_topk(q, db_set, rows, k)
    assert np.array_equal(ti[rows], oi)
    assert np.array_equal(tt[rows].view(np.int32), ot.view(np.int32))
