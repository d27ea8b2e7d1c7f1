"""The sparse path of the float32 matrix (an inverted index of the query
set's levels, per-pair counts in LDS): bit-identical to the oracle and to
the bit-sliced kernel under every policy, on real SMILES (where AUTO keeps
the bit-sliced kernel: common Lingos make long postings) and on synthetic
strings (where AUTO takes the sparse path), with partial column chunks and
float4 groups, odd row ranges, rows the sparse kernel leaves to the merge
kernel, magnitudes other than the count sums,
every column layout, ragged sets and a resident database shared by two
handles, query sets too wide for uint16 postings, and query sets of one to
five molecules."""
import numpy as np
import pytest

import oracle
from siml_amd import _lib, compiler, synth
from siml_amd.GPULingo import GPULingo, LingoDB

pytestmark = pytest.mark.gpu

NAMES = {_lib.POLICY_AUTO: None, _lib.POLICY_BITS: "rows_bits", _lib.POLICY_SPARSE: "rows_sparse",
         _lib.POLICY_MERGE: "rows_merge"}


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def assert_same_f32(got, want):
    assert got.shape == want.shape
    bad = np.count_nonzero(bits(got) != bits(want))
    assert bad == 0, f"{bad} of {want.size} float32 values differ"


def handle(policy, ref, qry, layout=None):
    g = GPULingo(0)
    g._h.call("siml_set_policy", policy)
    if layout is not None:
        g._h.call("siml_set_layout", layout)
    g.set_refsmiles(*ref)
    g.set_qsmiles(*qry)
    return g


@pytest.mark.parametrize("policy", [_lib.POLICY_AUTO, _lib.POLICY_BITS, _lib.POLICY_SPARSE])
def test_real_smiles_every_policy(demo_lines, policy):
    """ZINC molecules: AUTO keeps the bit-sliced kernel (common Lingos give
    long postings); the sparse path gives the same bits."""
    ref = oracle.pack(demo_lines[:777])
    qry = oracle.pack(demo_lines[500:4096])
    g = handle(policy, ref, qry)
    want = oracle.tanimoto_block(ref, qry, nthreads=0)
    assert_same_f32(g.getMultipleRows(0, 777), want)
    assert g._h.last_kernel() == (NAMES[policy] or "rows_bits")
    assert_same_f32(g.getMultipleRows(5, 13), want[5:13])
    assert_same_f32(g.getMultipleRows(250, 251), want[250:251])


def test_synthetic_auto_takes_sparse_path():
    """8192 seeded U[20,120] strings against themselves (the all-pairs
    workload's law at a quarter of its size): AUTO takes the sparse path;
    300 sampled rows bit-identical to the oracle and to the bit-sliced
    kernel, the diagonal 1."""
    s = synth.packed_padded(8192, 20, 120, seed=2)
    rows = np.sort(np.random.default_rng(3).choice(8192, 300, replace=False))
    want = oracle.tanimoto_block(tuple(x[rows] for x in s), s, nthreads=0)
    g = handle(_lib.POLICY_AUTO, s, s)
    got = g.getMultipleRows(0, 8192)
    assert g._h.last_kernel() == "rows_sparse"
    assert_same_f32(got[rows], want)
    assert np.all(got[np.arange(8192), np.arange(8192)] == np.float32(1.0))
    b = handle(_lib.POLICY_BITS, s, s)
    assert_same_f32(b.getMultipleRows(0, 8192)[rows], want)
    assert b._h.last_kernel() == "rows_bits"


@pytest.mark.parametrize("layout", [_lib.LAYOUT_NATURAL, _lib.LAYOUT_BLOCK, _lib.LAYOUT_SORTED])
def test_partial_chunks_and_row_ranges(layout):
    """32774 query molecules: two column chunks of the sparse kernel, the
    second of 6 columns (a partial float4); row ranges of one to 300 rows
    that start at odd and even rows."""
    ref = synth.packed_padded(300, 10, 120, seed=31)
    qry = synth.packed_padded(32774, 10, 120, seed=32)
    g = handle(_lib.POLICY_SPARSE, ref, qry, layout)
    want = oracle.tanimoto_block(ref, qry, nthreads=0)
    for rb, rl in ((0, 300), (5, 6), (3, 10), (297, 300), (1, 2)):
        assert_same_f32(g.getMultipleRows(rb, rl), want[rb:rl])
        assert g._h.last_kernel() == "rows_sparse"


def test_rows_left_to_the_merge_kernel():
    """Rows the sparse kernel does not take -- a count above 15 (no group),
    unsorted ids (no group), more than 255 levels (a byte cannot hold the
    count) -- go to the merge kernel in the same call."""
    strings = [s.encode() + b"\n" for s in synth.random_smiles(60, 20, 80, seed=41)]
    rng = np.random.default_rng(42)
    letters = np.frombuffer(b"CNOSPFcnos()=#[]@+-", np.uint8)
    strings[7] = b"C" * 40 + b"\n"                                  # "CCCC" 37 times
    strings[20] = bytes(rng.choice(letters, 300)) + b"\n"           # ~297 distinct levels
    ref = [x.copy() for x in oracle.pack(strings)]
    L, C, n, m = ref
    if n[33, 0] >= 2:                                               # unsorted row
        L[33, [0, 1]] = L[33, [1, 0]]
    qry = synth.packed_padded(5000, 10, 200, seed=43)
    want = oracle.tanimoto_block(tuple(ref), qry, nthreads=0)
    g = handle(_lib.POLICY_SPARSE, tuple(ref), qry)
    assert_same_f32(g.getMultipleRows(0, 60), want)
    assert g._h.last_kernel() == "rows_sparse+merge"
    assert_same_f32(g.getMultipleRows(20, 21), want[20:21])
    assert g._h.last_kernel() == "rows_merge"


def test_magnitudes_other_than_count_sums():
    """Magnitudes from the reference's default (the unique-Lingo counts, not
    the count sums) take the IEEE division, not the short one."""
    ref = list(synth.packed_padded(200, 20, 120, seed=51))
    qry = list(synth.packed_padded(3000, 20, 120, seed=52))
    ref[3] = ref[2].copy()
    qry[3] = qry[2].copy()
    ref[3][:5] = 0                                                 # zero magnitudes: the zero rule
    want = oracle.tanimoto_block(tuple(ref), tuple(qry), nthreads=0)
    g = handle(_lib.POLICY_SPARSE, tuple(ref), tuple(qry))
    assert_same_f32(g.getMultipleRows(0, 200), want)
    assert g._h.last_kernel() == "rows_sparse"


def test_ragged_sets_and_shared_database():
    """Ragged reference rows against a resident database (its index built
    once, by whichever handle asks first, and shared)."""
    buf, offs = synth.random_strings(30000, 8, 256, seed=61, dist="lognormal")
    db_set = compiler.pack_ragged_buffer(buf, offs)
    q = synth.packed_ragged(150, 8, 256, seed=62, dist="lognormal")
    db = LingoDB(*db_set)
    D = synth.padded_columns(*db_set, 0, len(db_set[2]) - 1)
    Q = synth.ragged_to_padded(*q)
    want = oracle.tanimoto_block(Q, D, nthreads=0)
    hs = []
    for policy in (_lib.POLICY_SPARSE, _lib.POLICY_AUTO):
        g = GPULingo(0)
        g._h.call("siml_set_policy", policy)
        g.attach_db(db)
        g.set_refsmiles_ragged(*q)
        hs.append(g)
    assert_same_f32(hs[0].getMultipleRows(0, 150), want)
    assert hs[0]._h.last_kernel() == "rows_sparse"
    assert_same_f32(hs[1].getMultipleRows(17, 150), want[17:])
    assert hs[1]._h.last_kernel() in ("rows_sparse", "rows_bits")
    db.close()


def test_wide_query_set_int32_postings():
    """70000 query molecules (postings kept as int32, three column chunks)
    with IEEE-division magnitudes on some rows and the short division on
    the rest of the calls."""
    ref = list(synth.packed_padded(40, 10, 150, seed=71))
    qry = synth.packed_padded(70000, 10, 150, seed=72)
    want = oracle.tanimoto_block(tuple(ref), qry, nthreads=0)
    g = handle(_lib.POLICY_SPARSE, tuple(ref), qry)
    assert_same_f32(g.getMultipleRows(0, 40), want)
    assert g._h.last_kernel() == "rows_sparse"
    ref[3] = ref[2].copy()                                          # magnitudes != count sums
    want2 = oracle.tanimoto_block(tuple(ref), qry, nthreads=0)
    g2 = handle(_lib.POLICY_SPARSE, tuple(ref), qry)
    assert_same_f32(g2.getMultipleRows(7, 40), want2[7:])


@pytest.mark.parametrize("nq", [1, 3, 5])
def test_tiny_query_sets(nq):
    """One to five query molecules: a single partial float4 per row."""
    ref = synth.packed_padded(9, 10, 60, seed=81)
    qry = synth.packed_padded(nq, 10, 60, seed=82)
    want = oracle.tanimoto_block(ref, qry, nthreads=0)
    g = handle(_lib.POLICY_SPARSE, ref, qry)
    assert_same_f32(g.getMultipleRows(0, 9), want)
    assert_same_f32(g.getMultipleRows(4, 5), want[4:5])
