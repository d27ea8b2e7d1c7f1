"""Launch modes on the GPU, each giving the oracle's results and the
ordinary launches' bits.

Graph mode (siml_set_graphs): top-k and 1000-bin histogram calls recorded by
stream capture and launched as one HIP graph, over repeated replays with
eager work of the same and of another handle in between, for calls whose
kernel sequence repeats (the cached graph is updated in place) and calls
whose sequence differs (seeded and unseeded top-k, merge-kernel rows, 1 to
300 rows).

Persistent grid (siml_set_persistent): the bit-sliced similarity pass run by
resident workgroups taking row-group slots from per-XCD work queues (the
handles keep the bit-sliced kernels: policy BITS).
"""
import numpy as np
import pytest

import oracle
from siml_amd import _lib, compiler, synth
from siml_amd.GPULingo import GPULingo, LingoDB

pytestmark = pytest.mark.gpu

# torch (caller-owned device arrays for the allocation-free entry point) is
# imported before the library first touches the device, so that the process
# loads one HIP runtime: torch's
try:
    import torch
except ImportError:  # pragma: no cover
    torch = None


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


@pytest.fixture(scope="module")
def db_and_queries():
    buf, offs = synth.random_strings(20000, 8, 256, seed=9, dist="lognormal")
    db_set = compiler.pack_ragged_buffer(buf, offs)
    rng = np.random.default_rng(12)
    picks = rng.choice(20000, size=400, replace=False)
    qb, qo = synth.mutated_strings(buf, offs, picks, 3, seed=13)
    q = compiler.pack_ragged_buffer(qb, qo)
    return db_set, q


def with_merge_rows(Qp):
    """Query rows plus three long carbon chains (Lingo counts above 15: rows
    the merge kernel serves in the same call)."""
    chains = [b"C" * n + b"\n" for n in (22, 30, 45)]
    Cp = compiler.cSMILEStoMatrices(chains)[:4]
    w = max(Qp[0].shape[1], Cp[0].shape[1])
    pad = lambda m: np.pad(m, ((0, 0), (0, w - m.shape[1])))
    return (np.ascontiguousarray(np.vstack([pad(Qp[0]), pad(Cp[0])])),
            np.ascontiguousarray(np.vstack([pad(Qp[1]), pad(Cp[1])])),
            np.ascontiguousarray(np.vstack([Qp[2], Cp[2]])),
            np.ascontiguousarray(np.vstack([Qp[3], Cp[3]])))


def test_graph_topk_and_hist_match_eager_and_oracle(db_and_queries):
    db_set, q = db_and_queries
    db = LingoDB(*db_set)
    D = synth.padded_columns(*db_set, 0, len(db_set[2]) - 1)
    Q = with_merge_rows(synth.padded_columns(*q, 0, 399))
    n = Q[0].shape[0]
    eager, graph, other = GPULingo(0), GPULingo(0), GPULingo(0)
    for g in (eager, graph, other):
        g.attach_db(db)
        g.set_refsmiles(*Q)
    graph.set_graphs(True)
    # (rows, k, seed stride): repeated sequences replay the cached graph with
    # new arguments; the others instantiate new ones
    calls = [((0, 60), 17, 1), ((0, 60), 17, 1), ((60, 120), 17, 1), ((5, 6), 1, 1), ((0, 300), 100, 1),
             ((0, 300), 100, 0), ((n - 40, n), 40, 1), ((n - 40, n), 40, 0), ((0, 60), 17, 1), ((200, 260), 50, 4)]
    for i, ((rb, rl), k, seed) in enumerate(calls):
        print(f"graph top-k call {i}: rows [{rb}, {rl}) k={k} seed={seed}", flush=True)
        for g in (eager, graph):
            g.set_launch_shape(2048, 256, seed)
        gi, gt = graph.getTopK(rb, rl, k)
        ei, et = eager.getTopK(rb, rl, k)
        assert np.array_equal(gi, ei) and np.array_equal(bits(gt), bits(et)), i
        oi, ot = oracle.topk(tuple(x[rb:rl] for x in Q), D, k)
        assert np.array_equal(gi, oi) and np.array_equal(bits(gt), bits(ot)), i
        # eager work between replays, on this handle and on another
        offs, _ = graph.getNeighborsCSR(rb, rl, 0.6)
        assert offs[-1] >= 0
        other.getMultipleRows(rb, min(rl, rb + 8))
    for rb, rl in ((0, n), (17, 290), (0, n), (5, 6)):
        print(f"graph hist1000: rows [{rb}, {rl})", flush=True)
        h = graph.getGlobalHistogram(rb, rl)
        assert np.array_equal(h, oracle.hist1000(tuple(x[rb:rl] for x in Q), D))
        assert int(h.sum()) == (rl - rb) * len(db)
        other.getTopK(rb, min(rl, rb + 4), 3)
    launches, builds = graph.graph_stats()
    print(f"graph launches {launches}, instantiations {builds}", flush=True)
    assert launches == len(calls) + 4
    assert 1 <= builds < launches
    assert eager.graph_stats() == (0, 0)
    db.close()


def test_graph_topk_noalloc_and_autotune(db_and_queries):
    """The allocation-free entry point and the autotuner's candidate runs go
    through graph mode too, with results equal to the eager handle's."""
    if torch is None:
        pytest.skip("torch is not importable")
    db_set, q = db_and_queries
    db = LingoDB(*db_set)
    Q = synth.padded_columns(*q, 0, 199)
    eager, graph = GPULingo(0), GPULingo(0)
    for g in (eager, graph):
        g.attach_db(db)
        g.set_refsmiles(*Q)
    graph.set_graphs(True)
    graph.set_autotune(True)
    for rb, rl, k in ((0, 150, 30), (10, 40, 8), (0, 150, 30)):
        print(f"graph noalloc/autotune: rows [{rb}, {rl}) k={k}", flush=True)
        ti, tt = graph.getTopK(rb, rl, k)
        ei, et = eager.getTopK(rb, rl, k)
        assert np.array_equal(ti, ei) and np.array_equal(bits(tt), bits(et))
        idx = torch.full((rl - rb, k), -7, dtype=torch.int32, device="cuda:0")
        t = torch.full((rl - rb, k), 7.0, dtype=torch.float32, device="cuda:0")
        torch.cuda.synchronize()
        graph._h.call("siml_topk_noalloc", rb, rl, k, idx.data_ptr(), t.data_ptr())
        graph._h.call("siml_synchronize")
        assert np.array_equal(idx.cpu().numpy(), ei) and np.array_equal(bits(t.cpu().numpy()), bits(et))
    launches, builds = graph.graph_stats()
    assert launches > 3 and builds >= 1
    db.close()


def test_persistent_grid_matches_oracle(db_and_queries):
    """Top-k and the 1000-bin histogram on the persistent grid (per-XCD work
    queues) equal the ordinary grid's results and the oracle's, for 1 to
    403 rows (with merge-kernel rows), seeded and unseeded top-k, several
    column-range counts, the allocation-free entry point, and graph mode."""
    db_set, q = db_and_queries
    db = LingoDB(*db_set)
    D = synth.padded_columns(*db_set, 0, len(db_set[2]) - 1)
    Q = with_merge_rows(synth.padded_columns(*q, 0, 399))
    n = Q[0].shape[0]
    ordinary, pers = GPULingo(0), GPULingo(0)
    for g in (ordinary, pers):
        g._h.call("siml_set_policy", _lib.POLICY_BITS)
        g.attach_db(db)
        g.set_refsmiles(*Q)
    for wg, cap, seed in ((8192, 1024, 1), (2048, 256, 0), (65536, 1024, 16), (512, 64, 4)):
        for g in (ordinary, pers):
            g.set_launch_shape(wg, cap, seed)
        pers.set_persistent(True)
        for (rb, rl), k in (((0, 60), 17), ((5, 6), 1), ((0, n), 100), ((n - 40, n), 40), ((100, 365), 9)):
            print(f"persistent top-k: shape {(wg, cap, seed)} rows [{rb}, {rl}) k={k}", flush=True)
            pi, pt = pers.getTopK(rb, rl, k)
            oi, ot = ordinary.getTopK(rb, rl, k)
            assert np.array_equal(pi, oi) and np.array_equal(bits(pt), bits(ot))
            wi, wt = oracle.topk(tuple(x[rb:rl] for x in Q), D, k)
            assert np.array_equal(pi, wi) and np.array_equal(bits(pt), bits(wt))
        for rb, rl in ((0, n), (17, 290), (5, 6)):
            print(f"persistent hist1000: shape {(wg, cap, seed)} rows [{rb}, {rl})", flush=True)
            h = pers.getGlobalHistogram(rb, rl)
            assert np.array_equal(h, oracle.hist1000(tuple(x[rb:rl] for x in Q), D))
    pers.set_graphs(True)
    for rb, rl, k in ((0, 200, 30), (0, 200, 30), (7, 9, 5)):
        print(f"persistent + graph mode: rows [{rb}, {rl}) k={k}", flush=True)
        pi, pt = pers.getTopK(rb, rl, k)
        wi, wt = oracle.topk(tuple(x[rb:rl] for x in Q), D, k)
        assert np.array_equal(pi, wi) and np.array_equal(bits(pt), bits(wt))
    assert np.array_equal(pers.getGlobalHistogram(0, n), oracle.hist1000(Q, D))
    if torch is not None:
        idx = torch.full((150, 30), -7, dtype=torch.int32, device="cuda:0")
        t = torch.full((150, 30), 7.0, dtype=torch.float32, device="cuda:0")
        torch.cuda.synchronize()
        pers._h.call("siml_topk_noalloc", 0, 150, 30, idx.data_ptr(), t.data_ptr())
        pers._h.call("siml_synchronize")
        wi, wt = oracle.topk(tuple(x[0:150] for x in Q), D, 30)
        assert np.array_equal(idx.cpu().numpy(), wi) and np.array_equal(bits(t.cpu().numpy()), bits(wt))
    db.close()
