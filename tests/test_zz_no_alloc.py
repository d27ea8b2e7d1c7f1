"""siml_topk_noalloc, the allocation-free top-k entry point, through the
C-ABI: the same indices and float32 bit patterns as siml_topk_async (the
existing entry point) on calls in a row over one handle, the second with a
smaller input so that the shared static scratch array holds the first
call's contents; launches narrowed to fit that array; rejection of a call
that cannot fit; a pending result of the handle left untouched.

The result arrays are device memory from torch (the caller's memory); the
entry point itself allocates nothing.
"""
import numpy as np
import pytest

import oracle
from siml_amd import _lib, synth
from siml_amd.GPULingo import GPULingo

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def noalloc_topk(g, base, limit, k):
    n = limit - base
    idx = torch.full((max(n, 1), k), -7, dtype=torch.int32, device="cuda:0")
    t = torch.full((max(n, 1), k), 7.0, dtype=torch.float32, device="cuda:0")
    torch.cuda.synchronize()
    g._h.call("siml_topk_noalloc", base, limit, k, idx.data_ptr(), t.data_ptr())
    g._h.call("siml_synchronize")
    return idx[:n].cpu().numpy(), t[:n].cpu().numpy()


def async_topk(g, base, limit, k):
    g._h.call("siml_topk_async", base, limit, k)
    return g._h.retrieve_topk()


@pytest.fixture(scope="module")
def sets():
    ref = synth.packed_padded(400, 10, 90, seed=71)
    qry = synth.packed_padded(40000, 10, 90, seed=72)
    return ref, qry


def test_two_calls_match_existing_entry_point(sets):
    ref, qry = sets
    g = GPULingo(0)
    g.set_refsmiles(*ref)
    g.set_qsmiles(*qry)
    for base, limit, k in ((0, 300, 50), (10, 47, 20)):
        ni, nt = noalloc_topk(g, base, limit, k)
        ai, at = async_topk(g, base, limit, k)
        print(f"rows [{base}, {limit}) k={k}", flush=True)
        assert np.array_equal(ni, ai), (base, limit, k)
        assert np.array_equal(bits(nt), bits(at)), (base, limit, k)
    oi, ot = oracle.topk(tuple(x[10:47] for x in ref), qry, 20)
    assert np.array_equal(ni, oi) and np.array_equal(bits(nt), bits(ot))


def test_narrowed_launch_and_pending_result(sets):
    """400 rows with k = 256 need candidate buffers beyond the scratch array
    at the handle's launch shape: the call runs fewer column ranges and gives
    the same result.  A pending result of the handle (a histogram queued
    before) is still retrievable afterwards."""
    ref, qry = sets
    g = GPULingo(0)
    g.set_refsmiles(*ref)
    g.set_qsmiles(*qry)
    need = np.zeros(1, np.int64)
    g._h.call("siml_call_bytes", _lib.OP_TOPK, 400, 256, _lib.i64p(need))
    print(f"400 rows k=256: {int(need[0])} bytes at the handle's launch shape", flush=True)
    assert need[0] > _lib.NOALLOC_SCRATCH_BYTES
    ai, at = async_topk(g, 0, 400, 256)
    g._h.call("siml_hist1000_async", 0, 400)
    ni, nt = noalloc_topk(g, 0, 400, 256)
    assert np.array_equal(ni, ai) and np.array_equal(bits(nt), bits(at))
    assert np.array_equal(g._h.retrieve_i64(), oracle.hist1000(ref, qry))


def test_rejects_call_beyond_scratch():
    """Candidate buffers of one column range for 12000 rows with k = 256
    exceed the scratch array: a clear error, and the handle still works."""
    ref = synth.packed_padded(12000, 10, 40, seed=73)
    qry = synth.packed_padded(3000, 10, 40, seed=74)
    g = GPULingo(0)
    g.set_refsmiles(*ref)
    g.set_qsmiles(*qry)
    with pytest.raises(ValueError, match="scratch"):
        noalloc_topk(g, 0, 12000, 256)
    ni, nt = noalloc_topk(g, 100, 164, 8)
    oi, ot = oracle.topk(tuple(x[100:164] for x in ref), qry, 8)
    assert np.array_equal(ni, oi) and np.array_equal(bits(nt), bits(ot))
