"""Golden fixtures (tests/golden/golden_small.npz, written by
tests/golden/make_golden.py): the oracle and the host packer reproduce them
on the CPU; the HIP path reproduces them on the GPU."""
import os

import numpy as np
import pytest

import oracle
from siml_amd import compiler

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden_small.npz")


@pytest.fixture(scope="module")
def gold():
    z = np.load(GOLDEN, allow_pickle=False)
    return {k: z[k] for k in z.files}


def _packed(g):
    return g["lingos"], g["counts"], g["lengths"], g["mags"]


def test_oracle_reproduces_golden(gold):
    strings = [bytes(s) for s in gold["strings"]]
    p = oracle.pack(strings)
    for a, b in zip(p, _packed(gold)):
        assert np.array_equal(a, b)
    M = oracle.tanimoto_block(p, p)
    assert np.array_equal(M.view(np.uint32), gold["tanimoto"].view(np.uint32))
    assert np.array_equal(oracle.hist_rows101(p, p), gold["hist101"])
    assert np.array_equal(oracle.hist1000(p, p), gold["hist1000"])
    lists, counts = oracle.neighbors(p, p, 0.3)
    assert np.array_equal(np.concatenate(lists), gold["nbr_indices"])
    ti, tt = oracle.topk(p, p, 8)
    assert np.array_equal(ti, gold["topk_idx"]) and np.array_equal(tt.view(np.uint32), gold["topk_t"].view(np.uint32))
    assert np.array_equal(oracle.single_row(p, 7, p).view(np.uint32), gold["single_row_7"].view(np.uint32))


def test_host_packer_reproduces_golden(gold):
    strings = [bytes(s) for s in gold["strings"]]
    l, c, n, m, _ = compiler.cSMILEStoMatrices(strings)
    for a, b in zip((l, c, n, m), _packed(gold)):
        assert np.array_equal(a, b)


@pytest.mark.gpu
def test_gpu_reproduces_golden(gold):
    from siml_amd.GPULingo import GPULingo
    g = GPULingo(0)
    p = _packed(gold)
    g.set_refsmiles(*p)
    g.set_qsmiles(*p)
    n = p[0].shape[0]
    M = g.getMultipleRows(0, n)
    assert np.array_equal(M.view(np.uint32), gold["tanimoto"].view(np.uint32))
    assert np.array_equal(g.getMultipleHistogrammedRows(0, n), gold["hist101"])
    offs, idx = g.getNeighborsCSR(0, n, 0.3)
    assert np.array_equal(offs, gold["nbr_offsets"]) and np.array_equal(idx, gold["nbr_indices"])
    assert np.array_equal(g.getTanimotoRow(7)[0].view(np.uint32), gold["single_row_7"].view(np.uint32))
