"""The HIP path at the full sizes of the benchmark configurations, checked
against the oracle and through size-independent properties:

* batched uneven query sets: a 1M-molecule resident database (lognormal
  lengths, median 45, seed 3) and 2000 query sets of geometric(1/128)
  sizes clipped to [1, 4096]; top-100 per query molecule.  One batched
  call (the sparse kernel) equals the same call through the bit-sliced
  kernels and the 2000 per-set calls on every row, and sampled rows equal
  the oracle over the whole database.
* threshold search + histogram: a 10M-molecule database (seed 4) and 1024
  near-duplicate queries; CSR lists for T >= 0.7 and the 1000-bin int64
  histogram of all 1.024e10 similarities.  The bins sum to the pair count,
  the sparse histogram equals the bit-sliced one bin for bin, the CSR is
  exactly the histogram's tail (bins 700..999: T >= fl32(0.7)), every list
  ascends, and sampled rows equal the oracle.
* serving: the same 10M database; Poisson arrivals of query batches
  (geometric(1/32) sizes clipped to [1, 512]), half top-50 and half
  threshold 0.8, continuously batched over two slots.  Every request's
  result equals the request run alone, and sampled requests equal the
  oracle.

The oracle runs over the database in column chunks, merged exactly
(CSR lists and histograms concatenate and add; top-k lists merge in
(T descending, column ascending) order).
"""
import concurrent.futures as cf
import os

import numpy as np
import pytest

import oracle
from siml_amd import _lib, compiler, synth
from siml_amd.GPULingo import GPULingo, LingoDB
from siml_amd.serving import Request, Server, run_poisson

pytestmark = pytest.mark.gpu

CHUNK = 1 << 18
THREADS = max(1, min(16, int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)))


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def say(msg):
    print(f"[configs] {msg}", flush=True)


def oracle_sweep(db_set, probes):
    """Oracle results of query rows over every database column, computed in
    one pass over column chunks (four chunks in flight).  `probes` maps a
    name to (padded query set, op, parameter) with op in "topk",
    "neighbors", "hist1000"; returns name -> result over the whole
    database: (idx, t) for top-k, a list of int64 column arrays for
    neighbours, the int64 bin counts for hist1000."""
    n = len(db_set[2]) - 1

    def one(c0):
        c1 = min(n, c0 + CHUNK)
        D = synth.padded_columns(*db_set, c0, c1)
        out = {}
        for name, (Q, op, prm) in probes.items():
            if op == "topk":
                oi, ot = oracle.topk(Q, D, min(prm, c1 - c0), nthreads=2)
                out[name] = (oi.astype(np.int64) + c0, ot)
            elif op == "neighbors":
                lists, _ = oracle.neighbors(Q, D, prm, nthreads=2)
                out[name] = [l.astype(np.int64) + c0 for l in lists]
            else:
                out[name] = oracle.hist1000(Q, D)
        return out

    acc = {}
    with cf.ThreadPoolExecutor(max_workers=max(1, THREADS // 2)) as ex:
        for part in ex.map(one, range(0, n, CHUNK)):   # ascending chunk order
            for name, r in part.items():
                op, k = probes[name][1], probes[name][2]
                if name not in acc:
                    acc[name] = r
                elif op == "topk":
                    bi = np.concatenate([acc[name][0], r[0]], axis=1)
                    bt = np.concatenate([acc[name][1], r[1]], axis=1)
                    order = np.lexsort((bi, -bt.astype(np.float64)), axis=1)[:, :k]
                    acc[name] = (np.take_along_axis(bi, order, 1), np.take_along_axis(bt, order, 1))
                elif op == "neighbors":
                    acc[name] = [np.concatenate([a, b]) for a, b in zip(acc[name], r)]
                else:
                    acc[name] = acc[name] + r
    return acc


def lognormal_set(n, seed):
    return synth.packed_ragged(n, 8, 256, seed=seed, dist="lognormal", median=45.0, sigma=0.35)


# ---------------------------------------------------------------------------
# batched uneven query sets: 1M database, 2000 sets, top-100
# ---------------------------------------------------------------------------

def test_batched_1m_database_2000_sets_top100():
    ndb, nsets, k = 1_000_000, 2000, 100
    say(f"packing the {ndb}-molecule database (seed 3) and 2000 query sets")
    db_set = lognormal_set(ndb, 3)
    sizes = synth.geometric_sizes(nsets, 1.0 / 128, 1, 4096, seed=3)
    nq = int(sizes.sum())
    q = lognormal_set(nq, 33)
    bounds = np.zeros(nsets + 1, np.int64)
    np.cumsum(sizes, out=bounds[1:])
    db = LingoDB(*db_set)
    try:
        g = GPULingo(0)
        g.attach_db(db)
        say(f"batched top-{k}: {nq} query molecules in one call")
        g.set_refsmiles_ragged(*q)
        bi, bt = g.getTopK(0, nq, k)
        assert bi.shape == (nq, k)
        assert g._h.last_kernel() == "topk_sparse"
        say(f"batched top-{k} through the bit-sliced kernels")
        gb = GPULingo(0)
        gb._h.call("siml_set_policy", _lib.POLICY_BITS)
        gb.attach_db(db)
        gb.set_refsmiles_ragged(*q)
        xi, xt = gb.getTopK(0, nq, k)
        assert gb._h.last_kernel() == "topk_bits"
        assert np.array_equal(xi, bi), f"{np.count_nonzero(xi != bi)} indices differ"
        assert np.array_equal(bits(xt), bits(bt))
        del gb, xi, xt
        say("per-set top-k: 2000 calls")
        for i in range(nsets):
            a, b = int(bounds[i]), int(bounds[i + 1])
            g.set_refsmiles_ragged(*synth.ragged_slice(*q, a, b))
            pi, pt = g.getTopK(0, b - a, k)
            assert np.array_equal(pi, bi[a:b]), f"set {i}: indices differ"
            assert np.array_equal(bits(pt), bits(bt[a:b])), f"set {i}: similarities differ"
        rng = np.random.default_rng(7)
        rows = np.sort(rng.choice(nq, size=6, replace=False))
        say(f"oracle top-{k} of rows {rows.tolist()} over the whole database")
        Q = tuple(x[rows] for x in synth.padded_columns(*q, 0, nq))
        want = oracle_sweep(db_set, {"topk": (Q, "topk", k)})["topk"]
        assert np.array_equal(bi[rows], want[0])
        assert np.array_equal(bits(bt[rows]), bits(want[1]))
    finally:
        db.close()


# ---------------------------------------------------------------------------
# 10M database: threshold CSR + 1000-bin histogram, and the serving loop
# ---------------------------------------------------------------------------

NDB = 10_000_000
NQ = 1024


@pytest.fixture(scope="module")
def db10m():
    say(f"packing the {NDB}-molecule database (seed 4) and {NQ} near-duplicate queries")
    buf, offs = synth.random_strings(NDB, 8, 256, seed=4, dist="lognormal", median=45.0, sigma=0.35)
    db_set = compiler.pack_ragged_buffer(buf, offs)
    rng = np.random.default_rng(44)
    picks = rng.choice(NDB, size=NQ, replace=False)
    q = compiler.pack_ragged_buffer(*synth.mutated_strings(buf, offs, picks, 2, 44))
    # serving pool: near duplicates of database molecules and random strings
    picks2 = rng.choice(NDB, size=3000, replace=False)
    near = compiler.pack_ragged_buffer(*synth.mutated_strings(buf, offs, picks2, 3, 45))
    del buf
    pool = synth.ragged_concat([near, lognormal_set(3000, 46)])
    say("building the resident database")
    db = LingoDB(*db_set)
    yield db_set, q, pool, db
    db.close()


def test_threshold_csr_and_histogram_10m(db10m):
    db_set, q, _, db = db10m
    g = GPULingo(0)
    g.attach_db(db)
    g.set_refsmiles_ragged(*q)
    say("1000-bin histogram of 1024 x 10M pairs")
    h = g.getGlobalHistogram(0, NQ)
    assert h.dtype == np.int64 and int(h.sum()) == NQ * NDB
    assert g._h.last_kernel() == "hist1000_sparse"
    b = GPULingo(0)
    b._h.call("siml_set_policy", _lib.POLICY_BITS)
    b.attach_db(db)
    b.set_refsmiles_ragged(*q)
    assert np.array_equal(b.getGlobalHistogram(0, NQ), h)
    assert b._h.last_kernel() == "hist1000_bits"
    del b
    say("CSR neighbour lists, T >= 0.7")
    offs, idx = g.getNeighborsCSR(0, NQ, 0.7)
    assert int(offs[-1]) == int(h[700:].sum()) > 0
    assert np.all(np.diff(offs) >= 0)
    for r in range(NQ):
        assert np.all(np.diff(idx[offs[r]:offs[r + 1]]) > 0), f"row {r} is not ascending"
    rows = [0, 1, NQ // 2, NQ - 1]
    say(f"oracle lists and histograms of rows {rows} over the whole database")
    Q = tuple(x[rows] for x in synth.padded_columns(*q, 0, NQ))
    want = oracle_sweep(db_set, {"nb": (Q, "neighbors", 0.7), "h": (Q, "hist1000", 0)})
    for i, r in enumerate(rows):
        assert np.array_equal(idx[offs[r]:offs[r + 1]], want["nb"][i]), f"row {r}: neighbours differ"
    gh = sum(g.getGlobalHistogram(r, r + 1) for r in rows)
    assert np.array_equal(gh, want["h"])
    # rows through a second, independent kernel family: the same pairs as a
    # top-k whose k-th entries reach below the threshold
    ti, tt = g.getTopK(0, NQ, 32)
    cnt = np.diff(offs)
    above = (tt >= np.float32(0.7)).sum(axis=1)
    assert np.array_equal(above, np.minimum(cnt, 32))


def _serving_requests(pool, nreq, seed):
    sizes = synth.geometric_sizes(nreq, 1.0 / 32, 1, 512, seed=seed)
    rng = np.random.default_rng(seed)
    npool = len(pool[2]) - 1
    reqs = []
    for i, s in enumerate(sizes):
        a = int(rng.integers(0, npool - int(s) + 1))
        op, prm = ("topk", 50) if i % 2 == 0 else ("threshold", 0.8)
        reqs.append(Request(i, *synth.ragged_slice(*pool, a, a + int(s)), op, prm))
    return reqs


def test_serving_10m_alone_equals_mix(db10m):
    db_set, _, pool, db = db10m
    reqs = _serving_requests(pool, 240, seed=5)
    say(f"serving {len(reqs)} requests (Poisson, 500/s) over 2 slots")
    with Server(db, slots=2, max_rows=1024) as srv:
        wall, lat = run_poisson(srv, reqs, 500.0, seed=5)
        batches = srv.batches
    errs = [r.error for r in reqs if r.error is not None]
    assert not errs, errs[:3]
    say(f"{len(reqs)} requests in {batches} batches, {wall:.2f} s, p50 {np.percentile(lat, 50) * 1e3:.1f} ms, "
        f"p99 {np.percentile(lat, 99) * 1e3:.1f} ms")
    assert batches < len(reqs)
    say("each request alone")
    g = GPULingo(0)
    g.attach_db(db)
    nonempty = 0
    for r in reqs:
        g.set_refsmiles_ragged(r.ids, r.counts, r.offsets, r.mags)
        if r.op == "topk":
            ai, at = g.getTopK(0, r.rows, 50)
            assert np.array_equal(ai, r.result[0]) and np.array_equal(bits(at), bits(r.result[1])), r.rid
        else:
            ao, ac = g.getNeighborsCSR(0, r.rows, 0.8)
            assert np.array_equal(ao, r.result[0]) and np.array_equal(ac, r.result[1]), r.rid
            nonempty += int(ao[-1] > 0)
    assert nonempty > 0, "no threshold request found a neighbour"
    small = [r for r in reqs if r.rows <= 3]
    sample = [next(r for r in small if r.op == "topk"), next(r for r in small if r.op == "threshold")]
    say(f"oracle over the whole database for requests {[r.rid for r in sample]}")
    probes = {}
    for r in sample:
        Q = synth.ragged_to_padded(r.ids, r.counts, r.offsets, r.mags)
        probes[r.rid] = (Q, "topk", 50) if r.op == "topk" else (Q, "neighbors", 0.8)
    want = oracle_sweep(db_set, probes)
    for r in sample:
        if r.op == "topk":
            assert np.array_equal(r.result[0], want[r.rid][0])
            assert np.array_equal(bits(r.result[1]), bits(want[r.rid][1]))
        else:
            lists = want[r.rid]
            assert np.array_equal(np.diff(r.result[0]), [len(l) for l in lists])
            assert np.array_equal(r.result[1], np.concatenate(lists))
