"""Pins the oracle's Tanimoto (restating _CPULingo.c:53-71, :239-271,
:427-462 and the single-row kernel of kernels.py:122-159,234-270) with an
exact-rational brute force, closed forms and invariants."""
import numpy as np
from hypothesis import given, settings, strategies as st

import oracle
from oracle import brute


def _set_from_bags(bags, mags=None):
    n = len(bags)
    w = max([len(b[0]) for b in bags] + [1])
    L = np.zeros((n, w), np.int32)
    C = np.zeros((n, w), np.int32)
    lens = np.zeros((n, 1), np.int32)
    for i, (ids, cnt) in enumerate(bags):
        L[i, :len(ids)] = ids
        C[i, :len(ids)] = cnt
        lens[i] = len(ids)
    if mags is None:
        mags = C.sum(1).astype(np.int32)
    return L, C, lens, np.asarray(mags, np.int32).reshape(-1, 1)


bag_strategy = st.dictionaries(st.integers(-6, 12), st.integers(1, 5), max_size=10).map(
    lambda d: (sorted(d), [d[k] for k in sorted(d)]))


@settings(max_examples=300, deadline=None)
@given(st.lists(bag_strategy, min_size=1, max_size=6), st.lists(bag_strategy, min_size=1, max_size=6))
def test_block_equals_fraction_bruteforce(rbags, qbags):
    ref = _set_from_bags(rbags)
    qry = _set_from_bags(qbags)
    got = oracle.tanimoto_block(ref, qry)
    for i, (ia, ca) in enumerate(rbags):
        for j, (ib, cb) in enumerate(qbags):
            want = brute.tanimoto(ia, ca, int(ref[3][i, 0]), ib, cb, int(qry[3][j, 0]))
            assert got[i, j].view(np.uint32) == want.view(np.uint32)


@settings(max_examples=200, deadline=None)
@given(st.lists(bag_strategy, min_size=1, max_size=4), st.lists(bag_strategy, min_size=1, max_size=4),
       st.integers(0, 40), st.integers(0, 40))
def test_block_with_explicit_mags(rbags, qbags, extra_r, extra_q):
    # magnitudes larger than the count sums (still exact small integers)
    rm = [sum(c) + extra_r for _, c in rbags]
    qm = [sum(c) + extra_q for _, c in qbags]
    ref = _set_from_bags(rbags, rm)
    qry = _set_from_bags(qbags, qm)
    got = oracle.tanimoto_block(ref, qry)
    for i, (ia, ca) in enumerate(rbags):
        for j, (ib, cb) in enumerate(qbags):
            want = brute.tanimoto(ia, ca, rm[i], ib, cb, qm[j])
            assert got[i, j].view(np.uint32) == want.view(np.uint32)


def test_closed_forms(demo_lines):
    ref = oracle.pack(demo_lines[:300])
    M = oracle.tanimoto_block(ref, ref)
    mags = ref[3].ravel()
    d = np.diag(M)
    assert np.all(d[mags > 0] == np.float32(1.0))            # T(A,A) = 1
    assert np.array_equal(M.view(np.uint32), M.T.view(np.uint32))  # bit-exact symmetry
    assert np.all((M >= 0) & (M <= 1))
    mm = mags.astype(np.float32)
    bound = np.minimum.outer(mm, mm) / np.maximum.outer(mm, mm)
    assert np.all(M <= np.where(np.isnan(bound), 0, bound))  # T <= min/max


def test_zero_mag_and_disjoint():
    ref = _set_from_bags([([1, 2], [1, 1]), ([], [])])
    qry = _set_from_bags([([3], [1]), ([1], [2]), ([], [])])
    M = oracle.tanimoto_block(ref, qry)
    assert M[0, 0] == 0.0                   # disjoint
    assert M[0, 1] == np.float32(1 / 3)     # I=1, U=2+2-1
    assert np.all(M[1] == 0.0) and M[0, 2] == 0.0   # zero magnitude


def test_serial_equals_parallel(demo_lines):
    """examples/benchmark.py:200: CPU == CPU-parallel, exactly."""
    ref = oracle.pack(demo_lines[:256])
    qry = oracle.pack(demo_lines)
    a = oracle.tanimoto_block(ref, qry, nthreads=1)
    b = oracle.tanimoto_block(ref, qry, nthreads=4)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_intersection_counts(demo_lines):
    ref = oracle.pack(demo_lines[:64])
    qry = oracle.pack(demo_lines[:200])
    inter = oracle.intersection_block(ref, qry)
    for i in range(0, 64, 9):
        for j in range(0, 200, 13):
            ia = ref[0][i, :ref[2][i, 0]].tolist()
            ca = ref[1][i, :ref[2][i, 0]].tolist()
            ib = qry[0][j, :qry[2][j, 0]].tolist()
            cb = qry[1][j, :qry[2][j, 0]].tolist()
            want = sum(min(a, b) for x, a in zip(ia, ca) for y, b in zip(ib, cb) if x == y)
            assert inter[i, j] == want


def test_single_row_recomputes_magnitudes(demo_lines):
    ref = list(oracle.pack(demo_lines[:40]))
    qry = list(oracle.pack(demo_lines[100:300]))
    true_ref, true_qry = tuple(ref), tuple(qry)
    # corrupt the given magnitudes: the single-row path must ignore them
    ref[3] = ref[3] + 7
    qry[3] = qry[3] * 2 + 1
    block_true = oracle.tanimoto_block(true_ref, true_qry)
    for r in (0, 5, 39):
        got = oracle.single_row(tuple(ref), r, tuple(qry))
        assert np.array_equal(got.view(np.uint32), block_true[r].view(np.uint32))


def test_single_row_zero_test_on_lengths():
    # lengths 0 but non-zero counts sum is impossible; use a zero-length
    # column with a non-zero given magnitude: the block gives 0/(ma+mb)=0,
    # the single row gives 0 by its length rule -- both exactly +0.0
    ref = _set_from_bags([([1, 2], [1, 1])])
    qry = _set_from_bags([([], [])], mags=[5])
    assert oracle.single_row(ref, 0, qry)[0].view(np.uint32) == 0
    assert oracle.tanimoto_block(ref, qry)[0, 0].view(np.uint32) == 0
