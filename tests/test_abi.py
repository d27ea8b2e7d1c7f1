"""C-ABI library checks that need no GPU: every symbol declared in
include/siml_amd.h is exported, the Python binding declares them, and the
host-side packer (product code) equals the oracle's packer."""
import os
import re
import subprocess

import numpy as np
import pytest
from hypothesis import given, settings, strategies as st

import oracle
from siml_amd import _lib, compiler

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "siml_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(siml_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 30
    lib = _lib.load()
    for n in names:
        assert hasattr(lib, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (siml_[a-z0-9_]+)", out))
    assert set(names) <= exported
    assert set(names) == set(_lib.SIGNATURES)


def test_error_reporting_without_gpu():
    lib = _lib.load()
    assert lib.siml_version() == 1
    with pytest.raises(ValueError):
        _lib.check(lib.siml_pack_measure(None, None, -1, None, None, 0))
    assert "negative" in lib.siml_last_error().decode()


def _same_as_oracle(strings):
    l, c, n, m, names = compiler.cSMILEStoMatrices(strings)
    ol, oc, on, om = oracle.pack(strings)
    assert l.shape == ol.shape and l.dtype == np.int32
    assert np.array_equal(l, ol) and np.array_equal(c, oc)
    assert np.array_equal(n, on) and np.array_equal(m, om)
    assert n.shape == (len(strings), 1) and m.shape == (len(strings), 1)
    return names


def test_packer_demo(demo_lines):
    names = _same_as_oracle(demo_lines)
    assert names[0] == "ZINC00000023" and len(names) == 4096


def test_packer_threads_agree(demo_lines):
    a = compiler.cSMILEStoMatrices(demo_lines * 3, nthreads=1)
    b = compiler.cSMILEStoMatrices(demo_lines * 3, nthreads=8)
    for x, y in zip(a[:4], b[:4]):
        assert np.array_equal(x, y)


@settings(max_examples=200, deadline=None)
@given(st.lists(st.binary(max_size=48), min_size=1, max_size=10))
def test_packer_random_bytes(strings):
    _same_as_oracle(strings)


@settings(max_examples=100, deadline=None)
@given(st.lists(st.text(alphabet="CNOSc()=#0123456789%+-H[]@ \n\r", max_size=40), min_size=1, max_size=8))
def test_packer_smiles_like(strings):
    _same_as_oracle(strings)


def test_packer_ragged_matches_padded():
    from siml_amd import synth
    ids, cnt, offs, mags = synth.packed_ragged(3000, 8, 256, seed=9, dist="lognormal")
    L, C, n, m = synth.packed_padded(3000, 8, 256, seed=9, dist="lognormal")
    L2, C2, n2, m2 = synth.ragged_to_padded(ids, cnt, offs, mags)
    assert np.array_equal(L, L2) and np.array_equal(C, C2)
    assert np.array_equal(n, n2) and np.array_equal(m, m2)


def test_packer_empty_and_names():
    l, c, n, m, names = compiler.cSMILEStoMatrices([])
    assert l.shape == (0, 0) and n.shape == (0, 1)
    l, c, n, m, names = compiler.cSMILEStoMatrices(["CC", "CCO name1\n", b"CCCC  x y"])
    assert l.shape == (3, 1) and n.ravel().tolist() == [0, 0, 1]
    assert names == ["", "name1", "x"]


def test_python_variant_packer():
    """compiler.py:151-194: digits 1-9 -> 0 on the first field only; mags is
    the 1-D row sum of counts."""
    l, c, n, m, names = compiler.SMILEStoMatrices(["c1ccccc1 benzene", "C%12 x", "CC"])
    assert m.shape == (3,) and m.tolist() == [5, 1, 0]
    assert l[0, :4].tolist() == [0x30636363, 0x63633063, 0x63636330, 0x63636363]
    assert names == ["benzene", "x", ""]
    lc, cc = compiler.SMILEStoMultiset("CCCCC")
    assert lc.shape == (1, 1) and lc[0, 0] == 0x43434343 and cc[0, 0] == 2
    assert compiler.preprocessNumbers("C1CC2") == "C0CC0"
