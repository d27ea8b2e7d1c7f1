"""Pins the oracle's packer (oracle_normalize / oracle_lingo_bag /
oracle_pack, restating _ccompiler.cpp:47-381) with hand-derived known
answers -- one per transition of the digit state machine -- and with an
independent Counter-based Lingo bag."""
import numpy as np
import pytest
from hypothesis import given, settings, strategies as st

import oracle
from oracle import brute

# (input, normalised) -- derived by hand from _ccompiler.cpp:178-264
KAT_NORMALISE = [
    # NORMAL state (:229-256)
    (b"C1", b"C0"),                 # digit -> '0', go to REPLACE
    (b"CCO", b"CCO"),               # plain copy
    (b"C%12C", b"C%0C"),            # '%' copied, CHOMP drops '1', '2' -> '0'
    (b"[13C]", b"[13C]"),           # '[' -> SPECIALNUMS keeps digits verbatim
    (b"C-2", b"C-2"),
    (b"C+2", b"C+2"),
    (b"NH2", b"NH2"),               # 'H' is special too
    (b"CC O", b"CC"),               # blank ends the SMILES
    (b"CCO\nC", b"CCO"),            # newline ends it
    (b"CCO\rC", b"CCO"),            # carriage return ends it
    (b"CC\0CC", b"CC"),             # NUL ends it
    # REPLACE state (:198-221)
    (b"C12", b"C00"),               # digit -> '0', stay
    (b"C1 C", b"C0"),               # blank ends
    (b"C1\nC", b"C0"),              # newline ends
    (b"C1%23", b"C0%0"),            # '%' -> CHOMP
    (b"C1+2", b"C0+0"),             # '+' is a plain copy here (no SPECIALNUMS)
    (b"C1[2", b"C0[0"),
    (b"1[2", b"0[0"),
    (b"C1C", b"C0C"),               # other -> copy, back to NORMAL
    # CHOMP state (:194-197)
    (b"%+2", b"%0"),                # drops even a special character
    (b"% C", b"%C"),                # and a blank
    (b"%12C", b"%0C"),
    # SPECIALNUMS state (:222-228)
    (b"H12", b"H12"),               # digits copied verbatim
    (b"H2O1", b"H2O0"),             # non-digit copied, back to NORMAL
    (b"C+ N", b"C+ N"),             # blank copied verbatim (name swallowed)
    (b"C-\nCC", b"C-\nCC"),         # newline copied verbatim
    (b"[+2", b"[+0"),               # the terminating '+' is not re-examined
    (b"+12", b"+12"),
    (b"", b""),
]


@pytest.mark.parametrize("inp,want", KAT_NORMALISE)
def test_normalise_kat(inp, want):
    assert oracle.normalize(inp) == want


def test_bag_kats():
    l, c, n, m = oracle.pack([b"CCO", b"CCCC", b"CCCCC", b"c1ccccc1"])
    assert n.ravel().tolist() == [0, 1, 1, 4]
    assert m.ravel().tolist() == [0, 1, 2, 5]
    assert l.shape == (4, 4) and l.dtype == np.int32
    assert l[1, 0] == 0x43434343 and c[1, 0] == 1
    assert l[2, 0] == 0x43434343 and c[2, 0] == 2
    # "c0ccccc0": windows c0cc, 0ccc, cccc, cccc, ccc0
    assert l[3].tolist() == [0x30636363, 0x63633063, 0x63636330, 0x63636363]
    assert c[3].tolist() == [1, 1, 1, 2]
    assert not l[0].any() and not c[0].any()


def test_bag_signed_order():
    # "ABC\x80" is a negative int32 and must sort before positive ids
    l, c, n, m = oracle.pack([b"ABC\x80ABCD"])
    ids = l[0, :n[0, 0]]
    assert ids[0] < 0
    assert np.all(np.diff(ids.astype(np.int64)) > 0)
    assert m[0, 0] == 5 and c[0].sum() == 5


def _check_against_counter(strings):
    l, c, n, m = oracle.pack(strings)
    for r, s in enumerate(strings):
        norm = oracle.normalize(s)
        ids, cnt, mag = brute.bag(norm)
        L = n[r, 0]
        assert L == len(ids)
        assert l[r, :L].tolist() == ids
        assert c[r, :L].tolist() == cnt
        assert m[r, 0] == mag
        assert not l[r, L:].any() and not c[r, L:].any()


def test_bag_matches_counter_on_demo(demo_lines):
    _check_against_counter(demo_lines[:1024])


@settings(max_examples=200, deadline=None)
@given(st.lists(st.binary(min_size=0, max_size=40), min_size=1, max_size=8))
def test_bag_matches_counter_random_bytes(strings):
    _check_against_counter(strings)


@settings(max_examples=100, deadline=None)
@given(st.lists(st.text(alphabet="CNOc()=#123%+-H[] \n", max_size=30), min_size=1, max_size=6))
def test_bag_matches_counter_smiles_like(strings):
    _check_against_counter([s.encode() for s in strings])


def test_demo_invariants(demo_lines):
    l, c, n, m = oracle.pack(demo_lines)
    assert l.shape == (4096, int(n.max()))
    norms = [oracle.normalize(s) for s in demo_lines]
    assert m.ravel().tolist() == [max(len(x) - 3, 0) for x in norms]
    assert np.array_equal(c.sum(1), m.ravel())
    for r in range(0, 4096, 7):
        ids = l[r, :n[r, 0]].astype(np.int64)
        assert np.all(np.diff(ids) > 0)
    # every demo line is 'SMILES NAME\n': the name never reaches the Lingos
    assert all(b" " not in x for x in norms)


def test_plain_strings_agree_with_digit_only_variant():
    # with no '%', '+', '-', 'H', '[' the C machine reduces to "1-9 -> 0"
    rng = np.random.default_rng(5)
    alpha = np.frombuffer(b"CNOScno()=#123456", np.uint8)
    strings = [alpha[rng.integers(0, alpha.size, rng.integers(0, 50))].tobytes() for _ in range(300)]
    for s in strings:
        assert oracle.normalize(s) == brute.digits_only(s)
