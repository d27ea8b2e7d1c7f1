"""The sparse top-k (one workgroup per row walking the query set's inverted
index chunk by chunk, a running threshold in LDS): identical indices and
float32 values to the oracle's top-k and to the bit-sliced path, over
several column chunks (uint16 and int32 postings), every column layout,
column ranges split over several workgroups per row (merged by the final
selection), rows with fewer than k positive pairs (the T = 0 fill), exact
ties, fewer columns than k, and the fall-back of rows the sparse path does
not take."""
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


def same(got, want):
    gi, gt = got
    wi, wt = want
    assert gi.shape == wi.shape
    bad = np.count_nonzero(gi != wi)
    assert bad == 0, f"{bad} of {wi.size} indices differ"
    bad = np.count_nonzero(np.ascontiguousarray(gt).view(np.uint32) != np.ascontiguousarray(wt).view(np.uint32))
    assert bad == 0, f"{bad} of {wt.size} values differ"


def handle(policy, ref, qry, layout=None):
    g = GPULingo(0)
    g._h.call("siml_set_policy", policy)
    if layout is not None:
        g._h.call("siml_set_layout", layout)
    g.set_refsmiles(*ref)
    g.set_qsmiles(*qry)
    return g


_CHUNKS = {}


def chunk_sets():
    if not _CHUNKS:
        _CHUNKS["ref"] = synth.packed_padded(3000, 8, 120, seed=101, dist="lognormal")
        _CHUNKS["qry"] = synth.packed_padded(100000, 8, 200, seed=102, dist="lognormal")
    return _CHUNKS["ref"], _CHUNKS["qry"]


@pytest.mark.parametrize("k", [1, 37, 100, 256])
@pytest.mark.parametrize("layout", [_lib.LAYOUT_NATURAL, _lib.LAYOUT_BLOCK, _lib.LAYOUT_SORTED])
def test_chunks_and_layouts(layout, k):
    """3000 rows against 100000 lognormal molecules (four column chunks,
    int32 postings, one workgroup per row)."""
    ref, qry = chunk_sets()
    g = handle(_lib.POLICY_SPARSE, ref, qry, layout)
    got = g.getTopK(0, 3000, k)
    assert g._h.last_kernel() == "topk_sparse"
    rows = np.arange(0, 3000, 7)
    if ("want", k) not in _CHUNKS:
        _CHUNKS["want", k] = oracle.topk(tuple(x[rows] for x in ref), qry, k, nthreads=0)
    same(tuple(x[rows] for x in got), _CHUNKS["want", k])


def test_split_rows_and_auto():
    """Few rows (column ranges split over several workgroups per row, then
    merged) against 60000 molecules (uint16 postings, two chunks); AUTO
    takes the sparse path on these strings and BITS gives the same
    results."""
    ref = synth.packed_padded(40, 10, 150, seed=111)
    qry = synth.packed_padded(60000, 10, 150, seed=112)
    want = oracle.topk(ref, qry, 100, nthreads=0)
    g = handle(_lib.POLICY_AUTO, ref, qry)
    same(g.getTopK(0, 40, 100), want)
    assert g._h.last_kernel() == "topk_sparse"
    same(g.getTopK(5, 6, 100), tuple(x[5:6] for x in want))
    b = handle(_lib.POLICY_BITS, ref, qry)
    same(b.getTopK(0, 40, 100), want)
    assert b._h.last_kernel() == "topk_bits"


def test_few_positive_pairs_zero_fill():
    """Rows sharing few or no Lingos with the columns: fewer than k pairs
    with T > 0, the rest of the row is T = 0 columns in ascending order
    (skipping the positive ones), in both launch shapes."""
    rng = np.random.default_rng(121)
    rare = np.frombuffer(b"FP#@", np.uint8)
    strings = [bytes(rng.choice(rare, 12)) + b"\n" for _ in range(30)]
    strings += [b"CCCC\n", b"FFFFCCO\n", b"\n", b"c1ccccc1\n"]
    ref = oracle.pack(strings)
    qry = synth.packed_padded(70000, 10, 120, seed=122)
    want = oracle.topk(ref, qry, 256, nthreads=0)
    g = handle(_lib.POLICY_SPARSE, ref, qry)
    same(g.getTopK(0, len(strings), 256), want)       # split launch (few rows)
    assert g._h.last_kernel() == "topk_sparse"
    many = oracle.pack(strings * 80)                   # one workgroup per row
    g2 = handle(_lib.POLICY_SPARSE, many, qry)
    got = g2.getTopK(0, len(strings) * 80, 256)
    assert g2._h.last_kernel() == "topk_sparse"
    for r in range(len(strings)):
        same(tuple(x[[r, r + len(strings) * 79]] for x in got), tuple(x[[r, r]] for x in want))


def test_ties_and_fewer_columns():
    """Every column repeated (exact ties broken by the smaller column) over
    two chunks; then fewer columns than k (-1 / -inf padding)."""
    base = synth.random_smiles(500, 10, 60, seed=131)
    cols = [s.encode() + b"\n" for s in base] * 80                  # 40000 columns
    qry = oracle.pack(cols)
    ref = oracle.pack([s.encode() + b"\n" for s in base[:64]])
    want = oracle.topk(ref, qry, 200, nthreads=0)
    g = handle(_lib.POLICY_SPARSE, ref, qry)
    same(g.getTopK(0, 64, 200), want)
    assert g._h.last_kernel() == "topk_sparse"
    small = oracle.pack(cols[:50])
    want2 = oracle.topk(ref, small, 100, nthreads=0)
    g2 = handle(_lib.POLICY_SPARSE, ref, small)
    got2 = g2.getTopK(0, 64, 100)
    same(got2, want2)
    assert np.all(got2[0][:, 50:] == -1)


def test_rows_the_sparse_path_does_not_take():
    """A row with a count above 15 has no row group: the call keeps the
    bit-sliced and merge kernels, with the same results."""
    strings = [s.encode() + b"\n" for s in synth.random_smiles(50, 20, 80, seed=141)]
    strings[9] = b"C" * 40 + b"\n"
    ref = oracle.pack(strings)
    qry = synth.packed_padded(40000, 10, 120, seed=142)
    g = handle(_lib.POLICY_SPARSE, ref, qry)
    same(g.getTopK(0, 50, 64), oracle.topk(ref, qry, 64, nthreads=0))
    assert g._h.last_kernel() == "topk_bits"
    same(g.getTopK(10, 50, 64), oracle.topk(tuple(x[10:50] for x in ref), qry, 64, nthreads=0))
    assert g._h.last_kernel() == "topk_sparse"


def test_resident_database_ragged_rows():
    """Ragged query rows against a resident ragged database of 150000
    molecules (its index built once and shared by two handles)."""
    buf, offs = synth.random_strings(150000, 8, 256, seed=151, dist="lognormal")
    db_set = compiler.pack_ragged_buffer(buf, offs)
    q = synth.packed_ragged(500, 8, 256, seed=152, dist="lognormal")
    db = LingoDB(*db_set)
    rows = np.arange(0, 500, 11)
    D = synth.padded_columns(*db_set, 0, len(db_set[2]) - 1)
    Q = tuple(x[rows] for x in synth.ragged_to_padded(*q))
    want = oracle.topk(Q, D, 50, nthreads=0)
    for policy in (_lib.POLICY_SPARSE, _lib.POLICY_AUTO):
        g = GPULingo(0)
        g._h.call("siml_set_policy", policy)
        g.attach_db(db)
        g.set_refsmiles_ragged(*q)
        got = g.getTopK(0, 500, 50)
        assert g._h.last_kernel() == "topk_sparse"
        same(tuple(x[rows] for x in got), want)
    db.close()


def test_long_columns_two_byte_magnitudes():
    """Columns of up to 700 characters (magnitudes above 255: the two-byte
    magnitude copy) against short rows, 40000 columns in two chunks."""
    ref = synth.packed_padded(2500, 10, 100, seed=161)
    qry = synth.packed_padded(40000, 150, 700, seed=162)
    assert int(qry[3].max()) > 255
    g = handle(_lib.POLICY_SPARSE, ref, qry)
    got = g.getTopK(0, 2500, 100)
    assert g._h.last_kernel() == "topk_sparse"
    rows = np.arange(0, 2500, 9)
    same(tuple(x[rows] for x in got), oracle.topk(tuple(x[rows] for x in ref), qry, 100, nthreads=0))


@pytest.mark.parametrize("nq", [32768, 65536, 65537])
def test_chunk_and_posting_width_boundaries(nq):
    """One full chunk; two full chunks with uint16 postings (the widest set
    they hold); int32 postings with a last chunk of one column -- top-k and
    the histogram, in one launch shape and split over column ranges."""
    qry = synth.packed_padded(nq, 10, 120, seed=171)
    for nr in (12, 2100):
        ref = synth.packed_padded(nr, 10, 120, seed=172 + nr)
        g = handle(_lib.POLICY_SPARSE, ref, qry)
        rows = np.arange(0, nr, max(1, nr // 12))
        got = g.getTopK(0, nr, 64)
        assert g._h.last_kernel() == "topk_sparse"
        same(tuple(x[rows] for x in got), oracle.topk(tuple(x[rows] for x in ref), qry, 64, nthreads=0))
        h = g.getGlobalHistogram(0, 12)
        assert g._h.last_kernel() == "hist1000_sparse"
        assert np.array_equal(h, oracle.hist1000(tuple(x[:12] for x in ref), qry))


def test_row_groups_built_on_first_use():
    """A reference set whose calls take the sparse path builds no row
    groups; the first call that needs them (the bit-sliced path under BITS)
    builds them, with the same results; the allocation-free top-k before
    that uses the merge kernel."""
    if torch is None:
        pytest.skip("torch is not importable")
    ref = synth.packed_padded(200, 10, 120, seed=181)
    qry = synth.packed_padded(40000, 10, 120, seed=182)
    want = oracle.topk(ref, qry, 30, nthreads=0)
    g = handle(_lib.POLICY_SPARSE, ref, qry)
    same(g.getTopK(0, 200, 30), want)
    assert g._h.last_kernel() == "topk_sparse"
    assert g._h.row_groups() == (0, 0)
    idx = torch.full((200, 30), -7, dtype=torch.int32, device="cuda:0")
    t = torch.full((200, 30), 7.0, dtype=torch.float32, device="cuda:0")
    torch.cuda.synchronize()
    g._h.call("siml_topk_noalloc", 0, 200, 30, idx.data_ptr(), t.data_ptr())
    g._h.call("siml_synchronize")
    same((idx.cpu().numpy(), t.cpu().numpy()), want)
    assert g._h.row_groups() == (0, 0)
    g._h.call("siml_set_policy", _lib.POLICY_BITS)
    same(g.getTopK(0, 200, 30), want)
    assert g._h.last_kernel() == "topk_bits"
    groups, merged = g._h.row_groups()
    assert groups > 0 and merged == 0
