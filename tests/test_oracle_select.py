"""Pins the oracle's histograms (_cpuutil.cpp:52-72, kernels.py:606-667 and
the 1000-bin extension), neighbour search (kernels.py:393-475) and top-k
against numpy float32 arithmetic and filters over the full matrix."""
import numpy as np
import pytest

import oracle
from siml_amd.GPULingo import reference_neighbor_counts


def _np_bin101_f32(t):
    p = np.asarray(t, np.float32) * np.float32(100)
    b = np.trunc(p)
    return np.clip(np.where(np.isnan(b) | (b < 0), 0, b), 0, 100).astype(np.int64)


EDGES = np.array([np.float32(k / 1000.0) for k in range(1001)], dtype=np.float32)


def _np_bin1000(t):
    t = np.asarray(t, np.float32)
    k = np.searchsorted(EDGES, t, side="right") - 1
    return np.clip(k, 0, 999)


def test_division_and_bins_exhaustive():
    """Every (I, U) with 0 <= I <= U <= 600: the oracle's Tanimoto equals
    numpy's IEEE float32 I/U, and its bins equal numpy's float32 binning."""
    width = 600
    for I in range(0, width + 1, 1):
        U = np.arange(max(I, 1), width + 1)
        ids = np.arange(width, dtype=np.int32)
        ref = (ids.reshape(1, -1), np.ones((1, width), np.int32), np.array([[width]], np.int32),
               np.array([[1]], np.int32))
        q_l = np.zeros((len(U), max(I, 1)), np.int32)
        q_c = np.zeros_like(q_l)
        if I:
            q_l[:] = ids[:I]
            q_c[:] = 1
        q_len = np.full((len(U), 1), I, np.int32)
        q_mag = (U + I - 1).astype(np.int32).reshape(-1, 1)   # U = 1 + mb - I
        T = oracle.tanimoto_block(ref, (q_l, q_c, q_len, q_mag))[0]
        want = np.float32(I) / U.astype(np.float32)
        assert np.array_equal(T.view(np.uint32), want.view(np.uint32)), I
        if I % 20 == 0:
            assert [oracle.bin101_f32(x) for x in T] == _np_bin101_f32(T).tolist()
            assert [oracle.bin1000(x) for x in T] == _np_bin1000(T).tolist()


def test_bin_edge_values():
    # bin 100 only for T == 1; values just below 1 go to 99
    assert oracle.bin101_f32(1.0) == 100
    assert oracle.bin101_f32(np.nextafter(np.float32(1), np.float32(0))) == 99
    assert oracle.bin1000(1.0) == 999 and oracle.bin1000(0.0) == 0
    for k in (1, 7, 100, 333, 700, 999):
        e = EDGES[k]
        assert oracle.bin1000(e) == k
        assert oracle.bin1000(np.nextafter(e, np.float32(0))) == k - 1


def test_row_hist101_float32_vs_float64():
    """_cpuutil.cpp:92,123: float64 input is multiplied in double; the
    float32 product can round up to the next integer."""
    x32 = np.array([[k / 100.0 for k in range(101)]], dtype=np.float32)
    h32 = oracle.row_hist101(x32)
    h64 = oracle.row_hist101(x32.astype(np.float64))
    b32 = _np_bin101_f32(x32[0])
    b64 = np.clip(np.trunc(x32[0].astype(np.float64) * 100.0), 0, 100).astype(np.int64)
    assert np.array_equal(h32[0], np.bincount(b32, minlength=101))
    assert np.array_equal(h64[0], np.bincount(b64, minlength=101))
    assert not np.array_equal(h32, h64)   # 0.57f is one value that differs


def test_hist_rows_and_hist1000(demo_lines):
    ref = oracle.pack(demo_lines[:64])
    qry = oracle.pack(demo_lines[:1500])
    M = oracle.tanimoto_block(ref, qry)
    H = oracle.hist_rows101(ref, qry)
    assert np.all(H.sum(1) == 1500)
    for r in range(64):
        assert np.array_equal(H[r], np.bincount(_np_bin101_f32(M[r]), minlength=101))
    G = oracle.hist1000(ref, qry)
    assert G.sum() == 64 * 1500
    assert np.array_equal(G, np.bincount(_np_bin1000(M.ravel()), minlength=1000))


@pytest.mark.parametrize("lo,hi", [(0.0, 1.1), (0.3, 1.1), (0.5, 0.9), (0.7, 1.1), (1.0, 1.1)])
def test_neighbors_equal_matrix_filter(demo_lines, lo, hi):
    ref = oracle.pack(demo_lines[:80])
    qry = oracle.pack(demo_lines[:2000])
    M = oracle.tanimoto_block(ref, qry)
    lists, counts = oracle.neighbors(ref, qry, lo, hi)
    for r in range(80):
        want = np.nonzero((M[r] >= np.float32(lo)) & (M[r] < np.float32(hi)))[0]
        assert np.array_equal(lists[r], want)
        assert counts[r] == len(want)


def test_neighbors_zero_magnitudes():
    # both magnitudes 0 -> bound 0/0 = NaN -> excluded even for lo <= 0;
    # one magnitude 0 -> bound 0 -> included for lo <= 0 (kernels.py:440-448)
    L = np.array([[1], [0]], np.int32)
    C = np.array([[1], [0]], np.int32)
    lens = np.array([[1], [0]], np.int32)
    mags = np.array([[1], [0]], np.int32)
    s = (L, C, lens, mags)
    lists, counts = oracle.neighbors(s, s, 0.0)
    assert lists[0].tolist() == [0, 1]
    assert lists[1].tolist() == [0]


def _simulated_reference_counts(lists, nq, pitch):
    """Literal walk of gpuGetNeighbors_multirow's wave loop (kernels.py:433-469)."""
    out = []
    for lst in lists:
        s = set(lst.tolist())
        cnt = 0
        for base in range(0, nq, 128):
            cnt += sum(1 for j in range(base, min(base + 128, nq)) if j in s)
            if cnt > pitch:
                break
        out.append(cnt)
    return out


@pytest.mark.parametrize("pitch", [0, 16, 32, 64, 4096])
def test_reference_overflow_counts(demo_lines, pitch):
    ref = oracle.pack(demo_lines[:40])
    qry = oracle.pack(demo_lines[:1200])
    lists, counts = oracle.neighbors(ref, qry, 0.15)
    offs = np.zeros(41, np.int64)
    offs[1:] = np.cumsum(counts)
    idx = np.concatenate(lists).astype(np.int32)
    got = reference_neighbor_counts(offs, idx, pitch)
    assert got.tolist() == _simulated_reference_counts(lists, 1200, pitch)


@pytest.mark.parametrize("k", [1, 5, 100, 1300])
def test_topk_equals_lexsort(demo_lines, k):
    ref = oracle.pack(demo_lines[:30])
    qry = oracle.pack(demo_lines[:1200])
    M = oracle.tanimoto_block(ref, qry)
    oi, ot = oracle.topk(ref, qry, k)
    cols = np.arange(1200)
    for r in range(30):
        order = np.lexsort((cols, -M[r]))[:k]
        n = min(k, 1200)
        assert np.array_equal(oi[r, :n], order)
        assert np.array_equal(ot[r, :n].view(np.uint32), M[r, order].view(np.uint32))
        assert np.all(oi[r, n:] == -1) and np.all(np.isneginf(ot[r, n:]))
