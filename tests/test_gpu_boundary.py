"""The C-ABI boundary on the GPU: the stateless entry points under
concurrent callers, the query-set reuse of siml_tanimoto_block, CPULingo
re-reading the caller's arrays, and the short float32 division of the
bit-sliced kernels over every (I, U) a 2000-level row can produce."""
import ctypes
import threading

import numpy as np
import pytest

import oracle
from siml_amd import _lib, synth
from siml_amd.CPULingo import CPULingo
from siml_amd.GPULingo import GPULingo
from siml_amd.utilities import rowHistogram101

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def tanimoto_block(ref, qry):
    """siml_tanimoto_block through ctypes, the call a _CPULingo binding
    makes (_CPULingo.c:83-274): (lingos, counts, MAGS, LENGTHS) per set."""
    lib = _lib.load()
    rl, rc, rn, rm = (np.ascontiguousarray(x, np.int32) for x in ref)
    ql, qc, qn, qm = (np.ascontiguousarray(x, np.int32) for x in qry)
    out = np.empty((rl.shape[0], ql.shape[0]), np.float32)
    _lib.check(lib.siml_tanimoto_block(0, _lib.i32p(rl), _lib.i32p(rc), _lib.i32p(rm), _lib.i32p(rn),
                                       rl.shape[0], rl.shape[1], rl.shape[1],
                                       _lib.i32p(ql), _lib.i32p(qc), _lib.i32p(qm), _lib.i32p(qn),
                                       ql.shape[0], ql.shape[1], ql.shape[1],
                                       out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), out.shape[1]))
    return out


def test_stateless_entry_points_from_four_threads():
    """Four host threads call siml_tanimoto_block and rowHistogram101 at the
    same time with sets of different sizes (so the shared handle's buffers
    are resized between calls); every result equals the oracle.  ctypes
    releases the GIL, which the reference held around these calls
    (_CPULingo.c:83-108)."""
    cases = []
    for t in range(4):
        ref = synth.packed_padded(100 + 150 * t, 20, 120, seed=40 + t)
        qry = synth.packed_padded(3000 + 2000 * t, 20, 120, seed=50 + t)
        mat = np.random.default_rng(60 + t).random((50 + 40 * t, 4000 + 3000 * t), dtype=np.float32)
        cases.append((ref, qry, mat, oracle.tanimoto_block(ref, qry, nthreads=0), oracle.row_hist101(mat)))
    errors = []
    barrier = threading.Barrier(4)

    def work(t):
        ref, qry, mat, want, want_h = cases[t]
        try:
            barrier.wait()
            for it in range(6):
                if (it + t) % 2 == 0:
                    got = tanimoto_block(ref, qry)
                    if not np.array_equal(bits(got), bits(want)):
                        errors.append(f"thread {t} iteration {it}: tanimoto block differs")
                else:
                    if not np.array_equal(rowHistogram101(mat), want_h):
                        errors.append(f"thread {t} iteration {it}: row histogram differs")
        except Exception as e:  # reported below
            errors.append(f"thread {t}: {e!r}")

    threads = [threading.Thread(target=work, args=(t,)) for t in range(4)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    assert not errors, errors


def test_tanimoto_block_reuses_and_rereads_query_set():
    """The same query set with successive row blocks (the reference's
    256-row pattern, examples/benchmark.py:123-167), then the query arrays
    edited in place: every block equals the oracle on the arrays as they
    are at the call."""
    ref = synth.packed_padded(600, 20, 80, seed=71)
    qry = [x.copy() for x in synth.packed_padded(2500, 20, 80, seed=72)]
    for base in range(0, 600, 256):
        sub = tuple(x[base:base + 256] for x in ref)
        assert np.array_equal(bits(tanimoto_block(sub, qry)), bits(oracle.tanimoto_block(sub, qry, nthreads=0)))
    qry[1][7, 0] += 2          # one more copy of molecule 7's first Lingo
    qry[3][7, 0] += 2
    qry[3][11, 0] += 5         # another magnitude
    sub = tuple(x[:256] for x in ref)
    got = tanimoto_block(sub, qry)
    want = oracle.tanimoto_block(sub, qry, nthreads=0)
    assert np.array_equal(bits(got), bits(want))
    assert not np.array_equal(bits(got[:, 7]), bits(oracle.tanimoto_block(sub, synth.packed_padded(
        2500, 20, 80, seed=72), nthreads=0)[:, 7]))


def test_cpulingo_sees_in_place_edits():
    """CPULingo.getMultipleRows reads the arrays at each call, as the
    reference does (CPULingo.py:126-138)."""
    ref = [x.copy() for x in synth.packed_padded(300, 20, 80, seed=73)]
    qry = [x.copy() for x in synth.packed_padded(900, 20, 80, seed=74)]
    c = CPULingo()
    c.set_refsmiles(ref[0], ref[1], ref[2])
    c.set_qsmiles(qry[0], qry[1], qry[2])

    def want():
        r = (ref[0], ref[1], ref[2], np.int32(ref[2].sum(1)))
        q = (qry[0], qry[1], qry[2], np.int32(qry[2].sum(1)))
        return oracle.tanimoto_block(r, q, nthreads=0)

    assert np.array_equal(bits(c.getMultipleRows(0, 300)), bits(want()))
    ref[1][5, 0] += 1          # counts only: the default magnitudes stay
    qry[1][9, 0] += 3
    assert np.array_equal(bits(c.getMultipleRows(0, 300)), bits(want()))


def prefix_sets(sizes):
    """Molecules whose levels are the first `a` of a fixed sequence: the
    x-th Lingo (ASCII bytes, ascending in x) holds levels 15x+1 .. 15x+15,
    so molecule a has the first ceil(a/15) Lingos with count 15 (the last
    a - 15 (ceil(a/15) - 1)).  Two such molecules a, b intersect in exactly
    min(a, b) and their count sums are a and b."""
    w = max((a + 14) // 15 for a in sizes) or 1
    L = np.zeros((len(sizes), w), np.int32)
    C = np.zeros((len(sizes), w), np.int32)
    n = np.zeros((len(sizes), 1), np.int32)
    for i, a in enumerate(sizes):
        k = (a + 14) // 15
        x = np.arange(k)
        L[i, :k] = 0x30303030 + (x % 64) + ((x // 64) << 8)
        C[i, :k] = 15
        if k:
            C[i, k - 1] = a - 15 * (k - 1)
        n[i, 0] = k
    return L, C, n, np.ascontiguousarray(C.sum(1, dtype=np.int32).reshape(-1, 1))


def test_fast_division_every_union_up_to_2000():
    """Rows a and columns b = 0..2000 levels, magnitudes = count sums (so
    the bit-sliced kernels take the short division): T(a, b) is
    min(a, b) / max(a, b), i.e. every (I, U) with 1 <= I <= U <= 2000, and
    must equal numpy's float32 division bit for bit; a row or column of
    magnitude 0 gives 0."""
    sizes = list(range(0, 2001))
    s = prefix_sets(sizes)
    g = GPULingo(0)
    g._h.call("siml_set_policy", _lib.POLICY_BITS)
    g.set_refsmiles(*s)
    g.set_qsmiles(*s)
    got = g.getMultipleRows(0, len(sizes))
    # rows whose ~2000 keys the two-choice table cannot place are left to the
    # merge kernel (IEEE division); the bit-sliced kernels serve the rest
    groups, merged = ctypes.c_int64(), ctypes.c_int64()
    _lib.check(_lib.load().siml_row_groups(g._h.h, ctypes.byref(groups), ctypes.byref(merged)))
    print(f"{groups.value} row groups, {merged.value} rows on the merge kernel", flush=True)
    assert g._h.last_kernel().startswith("rows_bits") and merged.value <= len(sizes) // 10
    a = np.array(sizes, np.int64)
    lo = np.minimum(a[:, None], a[None, :]).astype(np.float32)
    hi = np.maximum(a[:, None], a[None, :]).astype(np.float32)
    with np.errstate(invalid="ignore", divide="ignore"):
        want = np.where((a[:, None] == 0) | (a[None, :] == 0), np.float32(0), lo / hi).astype(np.float32)
    assert np.array_equal(bits(got), bits(want))
    # the same through the 1000-bin histogram (bins of the short quotient)
    h = g.getGlobalHistogram(0, len(sizes))
    assert np.array_equal(h, oracle.hist1000(s, s))


def test_fast_division_sparse_kernel():
    """The same prefix molecules of 0..255 levels (every row within the
    sparse kernel's byte counts) on the sparse path: every (I, U) with
    1 <= I <= U <= 255 against numpy's float32 division."""
    sizes = list(range(0, 256))
    s = prefix_sets(sizes)
    g = GPULingo(0)
    g._h.call("siml_set_policy", _lib.POLICY_SPARSE)
    g.set_refsmiles(*s)
    g.set_qsmiles(*s)
    got = g.getMultipleRows(0, len(sizes))
    assert g._h.last_kernel() == "rows_sparse"
    a = np.array(sizes, np.int64)
    lo = np.minimum(a[:, None], a[None, :]).astype(np.float32)
    hi = np.maximum(a[:, None], a[None, :]).astype(np.float32)
    with np.errstate(invalid="ignore", divide="ignore"):
        want = np.where((a[:, None] == 0) | (a[None, :] == 0), np.float32(0), lo / hi).astype(np.float32)
    assert np.array_equal(bits(got), bits(want))
