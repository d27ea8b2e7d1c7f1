"""Bit-exact parity of the HIP path with the CPU oracle.

All comparisons are on the float32 bit patterns (view as uint32) or on
integer arrays: the bar is identical intersection counts, identical float32
similarities, identical neighbour sets, histogram counts and top-k lists.
"""
import numpy as np
import pytest

import oracle
from siml_amd import _lib, synth
from siml_amd.CPULingo import CPULingo
from siml_amd.GPULingo import GPULingo, reference_neighbor_counts
from siml_amd.utilities import rowHistogram101

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def assert_same_f32(got, want):
    assert got.shape == want.shape
    bad = np.count_nonzero(bits(got) != bits(want))
    assert bad == 0, f"{bad} of {want.size} float32 values differ"


@pytest.fixture(scope="module")
def gpu():
    return GPULingo(0)


@pytest.fixture(scope="module")
def demo_set(demo_lines):
    return oracle.pack(demo_lines)


def _set(g, ref, qry):
    g.set_refsmiles(*ref)
    g.set_qsmiles(*qry)


def test_multirow_demo4096_full(gpu, demo_set):
    """examples/benchmark.py:123-192: N^2 matrix in 256-row blocks, GPU == CPU."""
    _set(gpu, demo_set, demo_set)
    want = oracle.tanimoto_block(demo_set, demo_set, nthreads=0)
    got = np.empty_like(want)
    for base in range(0, 4096, 256):
        got[base:base + 256] = gpu.getMultipleRows(base, min(base + 256, 4096))
    assert_same_f32(got, want)


def test_multirow_all_pairs_4096(gpu):
    s = synth.packed_padded(4096, 20, 80, seed=1)
    _set(gpu, s, s)
    got = gpu.getMultipleRows(0, 4096)
    assert_same_f32(got, oracle.tanimoto_block(s, s, nthreads=0))


def test_multirow_all_pairs_32768_sampled(gpu):
    """32768 x 32768 full matrix on the device; sampled rows checked exactly,
    diagonal == 1 and symmetry on the sampled block."""
    s = synth.packed_padded(32768, 20, 120, seed=2)
    _set(gpu, s, s)
    gpu.getMultipleRows_async(0, 32768)
    rng = np.random.default_rng(0)
    rows = np.unique(np.concatenate([[0, 32767], rng.choice(32768, 64, replace=False)]))
    got = np.empty((len(rows), 32768), np.float32)
    for i, r in enumerate(rows):
        gpu._h.call("siml_peek_rows_f32", int(r), int(r) + 1, _lib.f32p(got[i:i + 1]), 32768)
    sub = tuple(x[rows] for x in s)
    assert_same_f32(got, oracle.tanimoto_block(sub, s, nthreads=0))
    assert np.all(got[np.arange(len(rows)), rows] == np.float32(1.0))
    blk = got[:, rows]
    assert np.array_equal(bits(blk), bits(blk.T))
    gpu._retrieve = None  # leave the 4 GiB result on the device


def test_multirow_all_pairs_32768_full_both_kernels():
    """The whole 32768 x 32768 matrix (1.07e9 pairs of the all-pairs
    workload) equal to the oracle on every pair, through the sparse kernel
    (AUTO's choice here) and through the bit-sliced kernel: the
    reference's own acceptance check, exact equality over the full N^2
    matrix (examples/benchmark.py:189-206), at this size.  Compared in
    blocks of 4096 rows."""
    s = synth.packed_padded(32768, 20, 120, seed=2)
    hs = []
    for policy, name in ((_lib.POLICY_AUTO, "rows_sparse"), (_lib.POLICY_BITS, "rows_bits")):
        g = GPULingo(0)
        g._h.call("siml_set_policy", policy)
        _set(g, s, s)
        g.getMultipleRows_async(0, 32768)
        assert g._h.last_kernel() == name
        hs.append(g)
    got = np.empty((4096, 32768), np.float32)
    for b0 in range(0, 32768, 4096):
        want = oracle.tanimoto_block(tuple(x[b0:b0 + 4096] for x in s), s, nthreads=0)
        for g in hs:
            g._h.call("siml_peek_rows_f32", b0, b0 + 4096, _lib.f32p(got), 32768)
            assert_same_f32(got, want)
    for g in hs:
        g._retrieve = None  # leave the 4 GiB results on the device


def test_single_row_semantics(gpu, demo_lines):
    ref = list(oracle.pack(demo_lines[:50]))
    qry = list(oracle.pack(demo_lines[1000:3000]))
    ref_bad = (ref[0], ref[1], ref[2], ref[3] + 3)
    qry_bad = (qry[0], qry[1], qry[2], qry[3] * 2)
    _set(gpu, ref_bad, qry_bad)
    for r in (0, 17, 49):
        assert_same_f32(gpu.getTanimotoRow(r)[0], oracle.single_row(ref_bad, r, qry_bad))


def test_hist_rows(gpu, demo_set, demo_lines):
    ref = oracle.pack(demo_lines[:300])
    _set(gpu, ref, demo_set)
    got = gpu.getMultipleHistogrammedRows(0, 300)
    assert got.shape == (300, 101) and got.dtype == np.int32
    assert np.array_equal(got, oracle.hist_rows101(ref, demo_set))
    assert np.all(got.sum(1) == 4096)


@pytest.mark.parametrize("lo,hi", [(0.0, 1.1), (0.25, 1.1), (0.5, 0.8), (0.7, 1.1), (1.0, 1.1)])
def test_neighbors_csr(gpu, demo_set, demo_lines, lo, hi):
    ref = oracle.pack(demo_lines[:200])
    _set(gpu, ref, demo_set)
    offs, idx = gpu.getNeighborsCSR(0, 200, lo, hi)
    lists, counts = oracle.neighbors(ref, demo_set, lo, hi)
    assert np.array_equal(np.diff(offs), counts)
    want = np.concatenate(lists) if len(lists) else np.zeros(0, np.int32)
    assert np.array_equal(idx, want)


@pytest.mark.parametrize("maxn", [None, 8, 40])
def test_get_neighbors_reference_shape(gpu, demo_set, demo_lines, maxn):
    ref = oracle.pack(demo_lines[:64])
    _set(gpu, ref, demo_set)
    mat, cnt = gpu.getNeighbors(0, 64, 0.2, maxneighbors=maxn)
    lists, full = oracle.neighbors(ref, demo_set, 0.2)
    pitch = 16 * (((4096 if maxn is None else maxn) + 15) // 16)
    offs = np.zeros(65, np.int64)
    offs[1:] = np.cumsum(full)
    want_cnt = reference_neighbor_counts(offs, np.concatenate(lists).astype(np.int32), pitch)
    assert cnt.dtype == np.int32 and np.array_equal(cnt, want_cnt)
    for r in range(64):
        k = min(int(cnt[r]), mat.shape[1])
        assert np.array_equal(mat[r, :k], lists[r][:k])


def test_row_histogram101(gpu):
    rng = np.random.default_rng(3)
    x = rng.random((37, 5000), dtype=np.float32)
    x[0, :101] = np.arange(101, dtype=np.float32) / np.float32(100)
    x[1, :4] = [-0.5, np.nan, 2.0, 1.0]
    assert np.array_equal(rowHistogram101(x), oracle.row_hist101(x))
    x64 = x.astype(np.float64)
    assert np.array_equal(rowHistogram101(x64), oracle.row_hist101(x64))


def test_division_exhaustive(gpu):
    """All (I, U) with 0 <= I <= U <= 600: device T == numpy float32 I/U.
    The row is one Lingo with count 600 and magnitude 1; column (I, U) is
    that Lingo with count I (or another Lingo when I = 0) and magnitude
    U + I - 1, so isct = I and the union is exactly U."""
    Is, Us = [], []
    for I in range(601):
        U = np.arange(max(I, 1), 601)
        Is.append(np.full(len(U), I))
        Us.append(U)
    I = np.concatenate(Is)
    U = np.concatenate(Us)
    n = len(I)
    ref = (np.zeros((1, 1), np.int32), np.array([[600]], np.int32), np.ones((1, 1), np.int32),
           np.ones((1, 1), np.int32))
    q_l = np.where(I == 0, 1, 0).astype(np.int32).reshape(-1, 1)
    q_c = np.where(I == 0, 1, I).astype(np.int32).reshape(-1, 1)
    qry = (q_l, q_c, np.ones((n, 1), np.int32), (U + I - 1).astype(np.int32).reshape(-1, 1))
    _set(gpu, ref, qry)
    got = gpu.getMultipleRows(0, 1)[0]
    assert_same_f32(got, I.astype(np.float32) / U.astype(np.float32))
    assert_same_f32(got, oracle.tanimoto_block(ref, qry)[0])


def test_cpulingo_dropin(demo_set, demo_lines):
    c = CPULingo()
    ref = oracle.pack(demo_lines[:500])
    c.set_refsmiles(*ref)
    c.set_qsmiles(*demo_set)
    want = oracle.tanimoto_block(ref, demo_set, nthreads=0)
    assert_same_f32(c.getMultipleRows(0, 500, nprocs=4), want)
    assert_same_f32(c.getMultipleRows(-10, 1000), want[-10:])      # Python slicing
    assert c.getMultipleRows(400, 100).shape == (0, 4096)
    assert_same_f32(c.getTanimotoRow(7), want[7:8])
    assert c.retrieveAsyncResult() is c.retrieveAsyncResult()
    # default magnitudes = lengths.sum(1) (CPULingo.py:82,98)
    c2 = CPULingo()
    c2.set_refsmiles(ref[0], ref[1], ref[2])
    c2.set_qsmiles(demo_set[0], demo_set[1], demo_set[2])
    r2 = (ref[0], ref[1], ref[2], ref[2].sum(1).astype(np.int32))
    q2 = (demo_set[0], demo_set[1], demo_set[2], demo_set[2].sum(1).astype(np.int32))
    assert_same_f32(c2.getMultipleRows(0, 100), oracle.tanimoto_block(tuple(x[:100] for x in r2), q2, 0))


def test_stateless_block(demo_lines):
    import ctypes
    from siml_amd import _lib
    ref = oracle.pack(demo_lines[:100])
    qry = oracle.pack(demo_lines[100:900])
    out = np.empty((100, 800), np.float32)
    L = _lib.load()
    p = _lib.i32p
    _lib.check(L.siml_tanimoto_block(0, p(ref[0]), p(ref[1]), p(ref[3]), p(ref[2]), 100, ref[0].shape[1],
                                     ref[0].shape[1], p(qry[0]), p(qry[1]), p(qry[3]), p(qry[2]), 800,
                                     qry[0].shape[1], qry[0].shape[1],
                                     out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), 800))
    assert_same_f32(out, oracle.tanimoto_block(ref, qry))


def test_async_protocol(gpu, demo_set):
    _set(gpu, demo_set, demo_set)
    gpu.getMultipleRows_async(0, 64)
    while not gpu.asyncOperationsDone():
        pass
    a = gpu.retrieveAsyncResult()
    assert a.shape == (64, 4096)
    assert a.base is not None and a.base.shape[1] % 16 == 0  # padded pitch view
    with pytest.raises(RuntimeError):
        gpu.retrieveAsyncResult()
    with pytest.raises(ValueError):
        gpu.getMultipleRows_async(10, 5000)


def test_unsorted_and_duplicate_lists(gpu):
    """The j-outer merge equals the three-way merge state for state, so even
    unsorted or duplicated Lingo lists give the oracle's values."""
    rng = np.random.default_rng(7)
    n, w = 300, 24
    L = rng.integers(-20, 20, size=(n, w)).astype(np.int32)
    C = rng.integers(1, 4, size=(n, w)).astype(np.int32)
    lens = rng.integers(0, w + 1, size=(n, 1)).astype(np.int32)
    mags = rng.integers(0, 60, size=(n, 1)).astype(np.int32)
    s = (L, C, lens, mags)
    _set(gpu, s, s)
    assert_same_f32(gpu.getMultipleRows(0, n), oracle.tanimoto_block(s, s))


def test_edge_sizes(gpu, demo_set, demo_lines):
    ref = oracle.pack(demo_lines[:3])
    empty = tuple(x[:0] for x in demo_set)
    _set(gpu, ref, empty)
    assert gpu.getMultipleRows(0, 3).shape == (3, 0)
    _set(gpu, ref, demo_set)
    assert gpu.getMultipleRows(1, 1).shape == (0, 4096)
    # a ragged column count (not a multiple of 64) and zero-length molecules
    qry = oracle.pack([b"", b"C", b"CCCC"] + demo_lines[:61])
    _set(gpu, ref, qry)
    assert_same_f32(gpu.getMultipleRows(0, 3), oracle.tanimoto_block(ref, qry))


def test_row_length_limit(gpu):
    from siml_amd import _lib
    w = _lib.MAX_ROW_LINGOS
    big = (np.arange(w, dtype=np.int32).reshape(1, -1), np.ones((1, w), np.int32),
           np.array([[w]], np.int32), np.array([[w]], np.int32))
    qry = (np.arange(0, 2 * w, 2, dtype=np.int32).reshape(1, -1), np.ones((1, w), np.int32),
           np.array([[w]], np.int32), np.array([[w]], np.int32))
    _set(gpu, big, qry)
    assert_same_f32(gpu.getMultipleRows(0, 1), oracle.tanimoto_block(big, qry))
    over = (np.arange(w + 1, dtype=np.int32).reshape(1, -1), np.ones((1, w + 1), np.int32),
            np.array([[w + 1]], np.int32), np.array([[w + 1]], np.int32))
    with pytest.raises(ValueError, match="limit"):
        gpu.set_refsmiles(*over)


def test_ragged_sets_equal_padded(gpu):
    ids, cnt, offs, mags = synth.packed_ragged(2000, 8, 256, seed=21, dist="lognormal")
    pad = synth.ragged_to_padded(ids, cnt, offs, mags)
    gpu.set_refsmiles_ragged(ids[:offs[300]], cnt[:offs[300]], offs[:301], mags[:300])
    gpu.set_qsmiles_ragged(ids, cnt, offs, mags)
    got = gpu.getMultipleRows(0, 300)
    assert_same_f32(got, oracle.tanimoto_block(tuple(x[:300] for x in pad), pad, 0))


KERNEL_NAMES = {_lib.POLICY_AUTO: "rows_bits", _lib.POLICY_MERGE: "rows_merge"}
POLICIES = [_lib.POLICY_AUTO, _lib.POLICY_MERGE]


@pytest.mark.parametrize("policy", POLICIES)
def test_policies_bit_identical(demo_lines, policy):
    g = GPULingo(0)
    g._h.call("siml_set_policy", policy)
    ref = oracle.pack(demo_lines[:777])
    qry = oracle.pack(demo_lines[500:4096])
    _set(g, ref, qry)
    want = oracle.tanimoto_block(ref, qry, nthreads=0)
    assert_same_f32(g.getMultipleRows(0, 777), want)
    assert_same_f32(g.getMultipleRows(5, 13), want[5:13])     # range inside one group
    assert_same_f32(g.getMultipleRows(250, 251), want[250:251])
    assert g._h.last_kernel() == KERNEL_NAMES[policy]


def test_mixed_grouped_and_merge_rows(demo_lines):
    """Rows that cannot use the multi-row kernels (count > 65535, unsorted
    ids, too many ids) are served by the merge kernel in the same call."""
    g = GPULingo(0)
    ref = [x.copy() for x in oracle.pack(demo_lines[:100])]
    L, C, n, m = ref
    C[3, 0] = 70000                      # count beyond u16
    m[3, 0] += 70000 - 1
    if n[10, 0] >= 2:                    # unsorted row
        L[10, [0, 1]] = L[10, [1, 0]]
    big = 1000                           # more ids than one table holds
    wide = np.zeros((100, max(L.shape[1], big)), np.int32)
    wide[:, :L.shape[1]] = L
    cwide = np.zeros_like(wide)
    cwide[:, :C.shape[1]] = C
    wide[20, :big] = np.arange(big)
    cwide[20, :big] = 1
    n[20, 0] = big
    m[20, 0] = big
    ref = (wide, cwide, n, m)
    qry = oracle.pack(demo_lines[:2000])
    _set(g, ref, qry)
    got = g.getMultipleRows(0, 100)
    assert g._h.last_kernel() == "rows_bits+merge"
    assert_same_f32(got, oracle.tanimoto_block(ref, qry, nthreads=0))


@pytest.mark.parametrize("policy", POLICIES)
@pytest.mark.parametrize("k", [1, 8, 100, 256])
def test_topk(demo_lines, policy, k):
    g = GPULingo(0)
    g._h.call("siml_set_policy", policy)
    ref = oracle.pack(demo_lines[:70])
    qry = oracle.pack(demo_lines)
    _set(g, ref, qry)
    idx, t = g.getTopK(0, 70, k)
    oi, ot = oracle.topk(ref, qry, k)
    assert np.array_equal(idx, oi)
    assert_same_f32(t, ot)


def test_topk_more_than_columns_and_ties(demo_lines):
    g = GPULingo(0)
    ref = oracle.pack(demo_lines[:20])
    qry = oracle.pack(demo_lines[:37] + [b"CCCC"] * 30)   # exact ties
    _set(g, ref, qry)
    idx, t = g.getTopK(3, 20, 100)
    oi, ot = oracle.topk(tuple(x[3:20] for x in ref), qry, 100)
    assert np.array_equal(idx, oi)
    assert_same_f32(t, ot)
    assert np.all(idx[:, 67:] == -1)


@pytest.mark.parametrize("policy", POLICIES)
def test_global_histogram(demo_lines, demo_set, policy):
    g = GPULingo(0)
    g._h.call("siml_set_policy", policy)
    ref = oracle.pack(demo_lines[:333])
    _set(g, ref, demo_set)
    h = g.getGlobalHistogram(0, 333)
    assert h.dtype == np.int64 and h.sum() == 333 * 4096
    assert np.array_equal(h, oracle.hist1000(ref, demo_set))
    h2 = g.getGlobalHistogram(100, 101)
    assert np.array_equal(h2, oracle.hist1000(tuple(x[100:101] for x in ref), demo_set))


@pytest.mark.parametrize("policy", POLICIES)
def test_global_histogram_caller_magnitudes(demo_lines, policy):
    """Magnitudes given by the caller below the count sums (the CPULingo
    default of lengths, CPULingo.py:82): unions of 0 or less give T = +inf
    or negative values; the bins still match the oracle, including the
    pairs the kernels count by (I, U)."""
    g = GPULingo(0)
    g._h.call("siml_set_policy", policy)
    ref = oracle.pack(demo_lines[:200] + [b"CCCCCCCC\n", b"CCCCC\n", b"CCCCCCCCCCCC\n"])
    qry = oracle.pack(demo_lines[1000:2500] + [b"CCCCCC\n", b"CCCCCCCCCC\n"])
    ref = ref[:3] + (np.ascontiguousarray(ref[2].ravel().astype(np.int32)),)
    qry = qry[:3] + (np.ascontiguousarray(qry[2].ravel().astype(np.int32)),)
    _set(g, ref, qry)
    n = ref[0].shape[0]
    h = g.getGlobalHistogram(0, n)
    assert h.sum() == n * qry[0].shape[0]
    assert np.array_equal(h, oracle.hist1000(ref, qry))


@pytest.mark.parametrize("policy", POLICIES)
def test_hist_and_neighbors_both_policies(demo_lines, demo_set, policy):
    g = GPULingo(0)
    g._h.call("siml_set_policy", policy)
    ref = oracle.pack(demo_lines[1000:1150])
    _set(g, ref, demo_set)
    assert np.array_equal(g.getMultipleHistogrammedRows(3, 150), oracle.hist_rows101(
        tuple(x[3:150] for x in ref), demo_set))
    offs, idx = g.getNeighborsCSR(0, 150, 0.4)
    lists, counts = oracle.neighbors(ref, demo_set, 0.4)
    assert np.array_equal(np.diff(offs), counts) and np.array_equal(idx, np.concatenate(lists))


@pytest.mark.parametrize("layout", [_lib.LAYOUT_NATURAL, _lib.LAYOUT_SORTED])
def test_neighbor_lists_longer_than_sort_tiles(layout):
    """lo = 0 makes all 20000 columns neighbours of every row: rows of five
    4096-entry sort tiles and more, merged over three passes; the higher
    thresholds give ragged rows that straddle tile boundaries, and lo > 1
    gives no neighbour at all.  Slot and row offsets come from the device
    scan; lists are complete and ascending in both storage orders."""
    g = GPULingo(0)
    g._h.call("siml_set_layout", layout)
    qry = synth.packed_padded(20000, 10, 90, seed=17)
    ref = synth.packed_padded(75, 10, 90, seed=18)
    _set(g, ref, qry)
    for lo, base, limit in ((0.0, 0, 75), (0.3, 0, 75), (0.45, 5, 70), (1.05, 0, 75)):
        offs, idx = g.getNeighborsCSR(base, limit, lo)
        lists, counts = oracle.neighbors(tuple(x[base:limit] for x in ref), qry, lo)
        assert offs[0] == 0 and np.array_equal(np.diff(offs), counts), lo
        want = np.concatenate(lists) if len(lists) else np.zeros(0, np.int32)
        assert np.array_equal(idx, want), lo
    assert offs[-1] == 0


@pytest.mark.parametrize("layout", [_lib.LAYOUT_NATURAL, _lib.LAYOUT_BLOCK, _lib.LAYOUT_SORTED])
def test_layouts_bit_identical(demo_lines, demo_set, layout):
    """Column storage order (natural or length-sorted blocks) changes no bit
    of any result, including neighbour order and top-k ties."""
    g = GPULingo(0)
    g._h.call("siml_set_layout", layout)
    ref = oracle.pack(demo_lines[2000:2300])
    qry = oracle.pack(demo_lines[:3000] + [b"CCCC"] * 40)
    _set(g, ref, qry)
    assert_same_f32(g.getMultipleRows(0, 300), oracle.tanimoto_block(ref, qry, nthreads=0))
    assert_same_f32(g.getTanimotoRow(5)[0], oracle.single_row(ref, 5, qry))
    offs, idx = g.getNeighborsCSR(0, 300, 0.2)
    lists, counts = oracle.neighbors(ref, qry, 0.2)
    assert np.array_equal(np.diff(offs), counts) and np.array_equal(idx, np.concatenate(lists))
    ti, tt = g.getTopK(0, 300, 50)
    oi, ot = oracle.topk(ref, qry, 50)
    assert np.array_equal(ti, oi) and np.array_equal(bits(tt), bits(ot))
    assert np.array_equal(g.getGlobalHistogram(0, 300), oracle.hist1000(ref, qry))
    g._h.call("siml_set_policy", _lib.POLICY_MERGE)
    assert_same_f32(g.getMultipleRows(10, 40), oracle.tanimoto_block(tuple(x[10:40] for x in ref), qry, 0))


@pytest.mark.parametrize("policy", POLICIES)
def test_repeated_lingos(demo_lines, policy):
    """Multiplicities above 15 (long carbon chains): such rows take the
    merge kernel, such columns have their levels above 15 dropped from the
    level stream; every output stays bit-exact."""
    g = GPULingo(0)
    g._h.call("siml_set_policy", policy)
    chains = ["C" * n for n in (4, 5, 18, 19, 40, 80)] + ["CC(C)" + "C" * n + "O" for n in (10, 17, 30)]
    ref = oracle.pack(chains + demo_lines[:60])
    qry = oracle.pack(demo_lines[100:700] + chains)
    _set(g, ref, qry)
    n = ref[0].shape[0]
    assert_same_f32(g.getMultipleRows(0, n), oracle.tanimoto_block(ref, qry, nthreads=0))
    assert np.array_equal(g.getMultipleHistogrammedRows(0, n), oracle.hist_rows101(ref, qry))
    ti, tt = g.getTopK(0, n, 20)
    oi, ot = oracle.topk(ref, qry, 20)
    assert np.array_equal(ti, oi) and np.array_equal(bits(tt), bits(ot))
    offs, idx = g.getNeighborsCSR(0, n, 0.3)
    lists, counts = oracle.neighbors(ref, qry, 0.3)
    assert np.array_equal(np.diff(offs), counts) and np.array_equal(idx, np.concatenate(lists))


@pytest.mark.gpu
def test_short_division_exhaustive():
    """The short division the bit-sliced kernels use when magnitudes are
    small and consistent (method 1) equals IEEE float32 division on every
    integer numerator in [0, 65536) and denominator in [1, 131072): the
    whole domain of (I, U) those magnitudes allow."""
    import ctypes
    lib = _lib.load()
    bad, fn, fd = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    _lib.check(lib.siml_selftest_division(0, 1, ctypes.byref(bad), ctypes.byref(fn), ctypes.byref(fd)))
    print(f"method 1: {bad.value} mismatches, first n={fn.value} d={fd.value}", flush=True)
    assert bad.value == 0 and fn.value == -1 and fd.value == -1


@pytest.mark.parametrize("lo", [0.0, 0.05, 0.6, 0.95])
def test_neighbors_single_and_two_pass(demo_lines, demo_set, lo):
    """Slots with at most 16 neighbours take the single pass; any larger
    slot (lo = 0 makes every pair a neighbour) takes the fill pass.  Both
    give the complete, ascending lists."""
    g = GPULingo(0)
    ref = oracle.pack(demo_lines[3000:3060])
    _set(g, ref, demo_set)
    offs, idx = g.getNeighborsCSR(0, 60, lo)
    lists, counts = oracle.neighbors(ref, demo_set, lo)
    assert np.array_equal(np.diff(offs), counts)
    assert np.array_equal(idx, np.concatenate(lists))


@pytest.mark.parametrize("stride", [16, 4, 2])
@pytest.mark.parametrize("k", [1, 50, 256])
def test_topk_seeded_real_smiles(demo_lines, k, stride):
    """Real SMILES (dense similarities): top-k with the row thresholds seeded
    from every stride-th query tile equals the oracle."""
    g = GPULingo(0)
    g.set_launch_shape(2048, 256, stride)
    ref = oracle.pack(demo_lines[100:220])
    qry = oracle.pack(demo_lines)
    _set(g, ref, qry)
    idx, t = g.getTopK(0, 120, k)
    oi, ot = oracle.topk(ref, qry, k)
    assert np.array_equal(idx, oi)
    assert_same_f32(t, ot)
    assert g.launch_shape()[2] == stride


@pytest.mark.parametrize("stride", [16, 4])
@pytest.mark.parametrize("k", [1, 50])
def test_topk_seeded_admission_with_ties(demo_lines, k, stride):
    """Seeded rows skip pairs that cannot reach their seed's T before the
    division (an I, ma, mb test with a margin).  Eight copies of every query
    molecule make every T value, the seed's included, shared by at least
    eight columns: ties at and next to the admission boundary, where a lost
    pair would change the (T desc, column asc) lists."""
    g = GPULingo(0)
    g.set_launch_shape(8192, 1024, stride)
    ref = oracle.pack(demo_lines[300:420])
    qry = oracle.pack(demo_lines[:1024] * 8)
    _set(g, ref, qry)
    idx, t = g.getTopK(0, 120, k)
    oi, ot = oracle.topk(ref, qry, k)
    assert np.array_equal(idx, oi)
    assert_same_f32(t, ot)


def test_valu_peak_microbenchmark():
    """The measured int32 VALU rate of the device (independent add and min
    chains) is positive and not above the vector peak: 256 CUs x 128 lanes
    per clock (4 SIMDs, a wave64 instruction every 2 cycles) x 2.4 GHz."""
    import ctypes
    lib = _lib.load()
    for op in (0, 1):
        rate = ctypes.c_double()
        _lib.check(lib.siml_selftest_valu(0, op, ctypes.byref(rate)))
        print(f"op {op}: {rate.value:.4g} lane-ops/s", flush=True)
        assert 0 < rate.value <= 256 * 128 * 2.4e9 * 1.02


@pytest.mark.parametrize("policy", POLICIES)
def test_evaluated_work_counter(demo_lines, demo_set, policy):
    """The device counter of evaluated neighbour pairs.  With lo = 0 nothing
    is pruned: every pair is evaluated once per pass, and the merge work is
    exactly nq * sum(La) + nr * sum(Lb) per pass.  With lo = 0.7 the pairs
    whose magnitude bound fl(min/max) reaches lo must all be evaluated
    (pruning is lossless); the merge kernel evaluates exactly those."""
    g = GPULingo(0)
    g._h.call("siml_set_policy", policy)
    ref = oracle.pack(demo_lines[500:700])
    _set(g, ref, demo_set)
    h = g._h
    h.call("siml_set_timing", 1)
    nr, nq = ref[0].shape[0], demo_set[0].shape[0]
    la = ref[2].ravel().astype(np.int64)
    lb = demo_set[2].ravel().astype(np.int64)
    h.work_stats(reset=True)
    g.getNeighborsCSR(0, nr, 0.0)
    pairs, lsum = h.work_stats(reset=True)
    passes = pairs // (nr * nq)
    assert passes in (1, 2) and pairs == passes * nr * nq, (pairs, nr * nq)
    assert lsum == passes * (nq * int(la.sum()) + nr * int(lb.sum()))
    ma = ref[3].ravel().astype(np.float32)[:, None]
    mb = demo_set[3].ravel().astype(np.float32)[None, :]
    bound = np.minimum(ma, mb) / np.maximum(ma, mb)
    need = int(np.count_nonzero(bound >= np.float32(0.7)))
    offs, idx = g.getNeighborsCSR(0, nr, 0.7)
    pairs, lsum = h.work_stats(reset=True)
    lists, counts = oracle.neighbors(ref, demo_set, 0.7)
    assert np.array_equal(np.diff(offs), counts) and np.array_equal(idx, np.concatenate(lists))
    assert need <= pairs <= 2 * nr * nq
    if policy == _lib.POLICY_MERGE:
        assert pairs == 2 * need
    h.call("siml_set_timing", 0)
    g.getNeighborsCSR(0, nr, 0.7)
    assert h.work_stats(reset=True) == (0, 0)


def test_strided_result_copies(demo_lines):
    """Results whose host row pitch differs from the device pitch (CPULingo's
    contiguous arrays for a column count that is not a multiple of 16, and
    explicit pitches) and large results of equal pitch go through the
    staged copies (pinned slabs of 64 MiB above 32 MiB, two slabs here);
    every element lands in its row."""
    ref = synth.packed_padded(17000, 20, 120, seed=31)
    qry = synth.packed_padded(1001, 20, 120, seed=32)
    c = CPULingo()
    c.set_refsmiles(*ref)
    c.set_qsmiles(*qry)
    got = c.getMultipleRows(0, 17000)
    assert got.shape == (17000, 1001) and got.flags["C_CONTIGUOUS"]
    want = oracle.tanimoto_block(ref, qry, nthreads=0)
    assert_same_f32(got, want)
    g = GPULingo(0)
    _set(g, ref, qry)
    # GPULingo's padded host array has the device pitch: the contiguous path
    assert_same_f32(g.getMultipleRows(0, 17000), want)
    h = g._h
    h.call("siml_multirow_async", 100, 140)
    peek = np.full((40, 1003), -1.0, np.float32)
    h.call("siml_peek_rows_f32", 5, 25, _lib.f32p(peek[5:]), 1003)
    assert_same_f32(peek[5:25, :1001], want[105:125])
    assert np.all(peek[5:25, 1001:] == -1.0) and np.all(peek[:5] == -1.0) and np.all(peek[25:] == -1.0)
    out = np.full((40, 1020), -1.0, np.float32)
    h.retrieve_f32(out, 1020)
    assert_same_f32(out[:, :1001], want[100:140])
    assert np.all(out[:, 1001:] == -1.0)


def _key_slots(keys):
    """The two candidate table slots of level keys (the row-group hash of
    siml_amd/csrc/bitgroups.h), vectorised over a uint32 array."""
    h = keys.astype(np.uint64)
    m = np.uint64(0xFFFFFFFF)
    h = (h * np.uint64(0x9E3779B1)) & m
    h ^= h >> np.uint64(15)
    h = (h * np.uint64(0x85EBCA77)) & m
    h ^= h >> np.uint64(13)
    h = (h * np.uint64(0xC2B2AE3D)) & m
    h ^= h >> np.uint64(16)
    h1 = h >> np.uint64(20)
    h2 = (h >> np.uint64(4)) & np.uint64(4095)
    h2 = np.where(h2 == h1, h2 ^ np.uint64(1), h2)
    return h1, h2


def _colliding_ids():
    """Three ASCII Lingo ids whose level-1 keys (id | 0x80) share both slots."""
    b = np.arange(0x21, 0x7F, dtype=np.int64)
    x = (0x43 | (b[:, None, None] << 8) | (b[None, :, None] << 16) | (b[None, None, :] << 24)).reshape(-1)
    h1, h2 = _key_slots((x | 0x80).astype(np.uint32))
    pair = np.minimum(h1, h2) * np.uint64(4096) + np.maximum(h1, h2)
    order = np.argsort(pair, kind="stable")
    ps = pair[order]
    first = np.nonzero(ps[2:] == ps[:-2])[0][0]
    return [int(v) for v in sorted(x[order[first:first + 3]])]


def test_colliding_level_keys(demo_lines):
    """Rows whose level keys share both table slots: a group whose insertion
    runs out of displacements gives rows back, and a row that cannot be
    placed at all goes to the merge kernel; every pair stays bit-exact."""
    a, b, c = _colliding_ids()
    ids, counts, offs, mags = synth.packed_ragged(300, 20, 80, seed=41)
    extra_rows = [[a, b, 0x01010101], [c, 0x02010101], [a, b, c]]
    extra_cols = [[a, b, c], [c, 0x02010101], [a]]

    def with_rows(rows):
        I, C, O = [ids], [counts], [offs]
        end = int(offs[-1])
        for r in rows:
            r = sorted(r)
            I.append(np.array(r, np.int32))
            C.append(np.ones(len(r), np.int32))
            end += len(r)
            O.append(np.array([end], np.int64))
        i, cnt, o = np.concatenate(I), np.concatenate(C), np.concatenate(O)
        m = np.add.reduceat(cnt, o[:-1]).astype(np.int32)
        return synth.ragged_to_padded(i, cnt, o, m)

    ref, qry = with_rows(extra_rows), with_rows(extra_cols)
    g = GPULingo(0)
    _set(g, ref, qry)
    got = g.getMultipleRows(0, ref[0].shape[0])
    want = oracle.tanimoto_block(ref, qry, nthreads=0)
    assert_same_f32(got, want)
    # the planted columns meet the planted rows
    assert want[-1, -3] == np.float32(1.0) and want[-2, -2] == np.float32(1.0)


def test_named_retrieve_methods(gpu, demo_set, demo_lines):
    """GPULingo.py:538-628: the retrieve methods the reference binds per
    operation, called by name."""
    ref = oracle.pack(demo_lines[:32])
    _set(gpu, ref, demo_set)
    gpu.getMultipleRows_async(0, 32)
    with pytest.raises(ValueError):
        gpu.retrieveAsyncMatrixVectorResult()
    m = gpu.retrieveAsyncMatrixResult()
    assert_same_f32(m, oracle.tanimoto_block(ref, demo_set, nthreads=0))
    gpu.getMultipleHistogrammedRows_async(0, 32)
    h = gpu.retrieveAsyncMatrixResult()
    assert h.shape == (32, 101) and np.all(h.sum(axis=1) == 4096)
    lists, _ = oracle.neighbors(ref, demo_set, 0.3)
    gpu.getNeighbors_async(0, 32, 0.3, maxneighbors=40)
    full, cnt = gpu.retrieveAsyncMatrixVectorResult()
    assert full.shape == (32, 40)
    gpu.getNeighbors_async(0, 32, 0.3, maxneighbors=40)
    trimmed, cnt2 = gpu.retrieveAsyncMatrixVectorResult_trimmed()
    assert np.array_equal(cnt, cnt2) and trimmed.shape[1] == min(int(cnt.max()), 48)
    for r in range(32):
        k = min(int(cnt[r]), 40)
        assert np.array_equal(full[r, :k], lists[r][:k])
        assert np.array_equal(trimmed[r, :min(k, trimmed.shape[1])], lists[r][:min(k, trimmed.shape[1])])
    with pytest.raises(RuntimeError):
        gpu.retrieveAsyncResultBase()
    with pytest.raises(RuntimeError):
        gpu.retrieveAsyncMatrixResult()
