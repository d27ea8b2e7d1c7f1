"""The sparse 1000-bin histogram (one workgroup per row and column range
walking the query set's inverted index chunk by chunk): identical int64 bin
counts to the oracle's and to the bit-sliced path's, summing to the pair
count, over several column chunks (uint16 and int32 postings), every column
layout, column ranges split over several workgroups per row, magnitudes of
one and two bytes, a resident database, and the fall-back of rows the
sparse path does not take."""
import numpy as np
import pytest

import oracle
from siml_amd import _lib, compiler, synth
from siml_amd.GPULingo import GPULingo, LingoDB

pytestmark = pytest.mark.gpu


def handle(policy, ref, qry, layout=None):
    g = GPULingo(0)
    g._h.call("siml_set_policy", policy)
    if layout is not None:
        g._h.call("siml_set_layout", layout)
    g.set_refsmiles(*ref)
    g.set_qsmiles(*qry)
    return g


_SETS = {}


def sets():
    if not _SETS:
        _SETS["ref"] = synth.packed_padded(600, 8, 120, seed=201, dist="lognormal")
        _SETS["qry"] = synth.packed_padded(100000, 8, 200, seed=202, dist="lognormal")
        _SETS["want"] = oracle.hist1000(_SETS["ref"], _SETS["qry"])
    return _SETS["ref"], _SETS["qry"], _SETS["want"]


@pytest.mark.parametrize("layout", [_lib.LAYOUT_NATURAL, _lib.LAYOUT_BLOCK, _lib.LAYOUT_SORTED])
def test_chunks_layouts_and_splits(layout):
    """600 rows against 100000 lognormal molecules: four column chunks,
    int32 postings, four column ranges per row; then 7 rows (every chunk
    its own range) and one row."""
    ref, qry, want = sets()
    g = handle(_lib.POLICY_SPARSE, ref, qry, layout)
    h = g.getGlobalHistogram(0, 600)
    assert g._h.last_kernel() == "hist1000_sparse"
    assert np.array_equal(h, want)
    assert int(h.sum()) == 600 * 100000
    for rb, rl in ((100, 107), (599, 600)):
        h = g.getGlobalHistogram(rb, rl)
        assert g._h.last_kernel() == "hist1000_sparse"
        assert np.array_equal(h, oracle.hist1000(tuple(x[rb:rl] for x in ref), qry))


def test_auto_and_bits_agree():
    """AUTO takes the sparse path on these strings (60000 columns, uint16
    postings, two chunks); BITS gives the same bins."""
    ref = synth.packed_padded(300, 10, 150, seed=211)
    qry = synth.packed_padded(60000, 10, 150, seed=212)
    want = oracle.hist1000(ref, qry)
    g = handle(_lib.POLICY_AUTO, ref, qry)
    assert np.array_equal(g.getGlobalHistogram(0, 300), want)
    assert g._h.last_kernel() == "hist1000_sparse"
    b = handle(_lib.POLICY_BITS, ref, qry)
    assert np.array_equal(b.getGlobalHistogram(0, 300), want)
    assert b._h.last_kernel() == "hist1000_bits"


def test_near_duplicates_long_columns_and_fallback():
    """Rows that are mutated copies of columns (high similarities, every
    region of the histogram), columns of up to 700 characters (two-byte
    magnitudes), and a row with a count above 15 (no row group: the call
    keeps the bit-sliced and merge kernels)."""
    buf, offs = synth.random_strings(40000, 150, 700, seed=221)
    qry = compiler.pack_ragged_buffer(buf, offs)
    Q = synth.padded_columns(*qry, 0, 39999)
    assert int(Q[3].max()) > 255
    picks = np.random.default_rng(222).choice(40000, size=200, replace=False)
    rb_, ro_ = synth.mutated_strings(buf, offs, picks, 3, seed=223)
    rows = [rb_[ro_[i]:ro_[i + 1]][:100] + b"\n" for i in range(200)]
    ref = oracle.pack(rows)
    g = handle(_lib.POLICY_SPARSE, ref, Q)
    h = g.getGlobalHistogram(0, 200)
    assert g._h.last_kernel() == "hist1000_sparse"
    assert np.array_equal(h, oracle.hist1000(ref, Q))
    rows[17] = b"C" * 40 + b"\n"
    ref2 = oracle.pack(rows)
    g2 = handle(_lib.POLICY_SPARSE, ref2, Q)
    assert np.array_equal(g2.getGlobalHistogram(0, 200), oracle.hist1000(ref2, Q))
    assert g2._h.last_kernel() == "hist1000_bits"


def test_resident_database():
    """Ragged rows against a resident ragged database of 150000 molecules
    (its index shared by two handles)."""
    buf, offs = synth.random_strings(150000, 8, 256, seed=231, dist="lognormal")
    db_set = compiler.pack_ragged_buffer(buf, offs)
    q = synth.packed_ragged(60, 8, 256, seed=232, dist="lognormal")
    db = LingoDB(*db_set)
    want = oracle.hist1000(synth.ragged_to_padded(*q), synth.padded_columns(*db_set, 0, len(db_set[2]) - 1))
    for policy in (_lib.POLICY_SPARSE, _lib.POLICY_AUTO):
        g = GPULingo(0)
        g._h.call("siml_set_policy", policy)
        g.attach_db(db)
        g.set_refsmiles_ragged(*q)
        assert np.array_equal(g.getGlobalHistogram(0, 60), want)
        assert g._h.last_kernel() == "hist1000_sparse"
    db.close()
