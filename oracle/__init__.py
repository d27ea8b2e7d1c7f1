"""CPU oracle for the LINGO similarity path (TEST INFRASTRUCTURE ONLY).

This package is the checker for the HIP product in ``siml_amd``: only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it.  It wraps ``siml_oracle.c`` (a plain-C
restatement of the reference CPU path, each function citing the reference
file:line it follows) through ctypes, and ``brute.py`` (independent
pure-Python computations on tiny inputs, used to pin the C restatement).

Pinning of every function is listed in the header of ``siml_oracle.c``;
no function here is "parity unpinned".
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)
_i64 = ctypes.c_int64
_int = ctypes.c_int


def build():
    """Compile liboracle.so with the host C compiler (gcc)."""
    subprocess.check_call(["make", "-s", "-B", "-C", _HERE, "liboracle.so"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.oracle_normalize.argtypes = [ctypes.c_char_p, _i64, ctypes.c_char_p]
        L.oracle_normalize.restype = _i64
        L.oracle_pack.argtypes = [ctypes.c_char_p, _i64p, _i64, _i32p, _i32p, _i32p, _i32p, _i64]
        L.oracle_pack.restype = _int
        L.oracle_tanimoto_block.argtypes = [_i32p, _i32p, _i32p, _i32p, _i64, _i64,
                                            _i32p, _i32p, _i32p, _i32p, _i64, _i64, _f32p, _i64, _int]
        L.oracle_tanimoto_block.restype = None
        L.oracle_intersection_block.argtypes = [_i32p, _i32p, _i32p, _i64, _i64,
                                                _i32p, _i32p, _i32p, _i64, _i64, _i32p, _int]
        L.oracle_intersection_block.restype = None
        L.oracle_single_row.argtypes = [_i32p, _i32p, ctypes.c_int32, _i32p, _i32p, _i32p, _i64, _i64, _f32p]
        L.oracle_single_row.restype = None
        L.oracle_row_hist101_f32.argtypes = [_f32p, _i64, _i64, _i64, _u32p]
        L.oracle_row_hist101_f64.argtypes = [_f64p, _i64, _i64, _i64, _u32p]
        L.oracle_hist_rows101.argtypes = [_i32p, _i32p, _i32p, _i32p, _i64, _i64,
                                          _i32p, _i32p, _i32p, _i32p, _i64, _i64, _i32p, _int]
        L.oracle_hist1000.argtypes = [_i32p, _i32p, _i32p, _i32p, _i64, _i64,
                                      _i32p, _i32p, _i32p, _i32p, _i64, _i64, _i64p]
        L.oracle_neighbors.argtypes = [_i32p, _i32p, _i32p, _i32p, _i64, _i64,
                                       _i32p, _i32p, _i32p, _i32p, _i64, _i64,
                                       ctypes.c_float, ctypes.c_float, _i32p, _i64p, _int]
        L.oracle_topk.argtypes = [_i32p, _i32p, _i32p, _i32p, _i64, _i64,
                                  _i32p, _i32p, _i32p, _i32p, _i64, _i64, ctypes.c_int32,
                                  _i32p, _f32p, _int]
        L.oracle_bin1000.argtypes = [ctypes.c_float]
        L.oracle_bin1000.restype = ctypes.c_int32
        L.oracle_bin101_f32.argtypes = [ctypes.c_float]
        L.oracle_bin101_f32.restype = ctypes.c_uint32
        L.oracle_bin101_f64.argtypes = [ctypes.c_double]
        L.oracle_bin101_f64.restype = ctypes.c_uint32
        L.oracle_max_threads.restype = _int
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


def _c32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _vec(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32).reshape(-1))


def normalize(s):
    """Digit-normalised form of one SMILES string (bytes)."""
    b = s.encode() if isinstance(s, str) else bytes(s)
    out = ctypes.create_string_buffer(len(b) + 1)
    n = lib().oracle_normalize(b, len(b), out)
    return out.raw[:n]


def pack(strings):
    """(lingos, counts, lengths, mags) with the cSMILEStoMatrices layout:
    (N, W) int32 matrices and (N, 1) int32 vectors."""
    bs = [s.encode() if isinstance(s, str) else bytes(s) for s in strings]
    offs = np.zeros(len(bs) + 1, dtype=np.int64)
    offs[1:] = np.cumsum([len(b) for b in bs]) if bs else []
    buf = b"".join(bs)
    n = len(bs)
    lengths = np.zeros((n, 1), dtype=np.int32)
    mags = np.zeros((n, 1), dtype=np.int32)
    L = lib()
    L.oracle_pack(buf, _p(offs, _i64p), n, _p(lengths, _i32p), _p(mags, _i32p), None, None, 0)
    w = int(lengths.max()) if n else 0
    lingos = np.zeros((n, w), dtype=np.int32)
    counts = np.zeros((n, w), dtype=np.int32)
    if w > 0:
        rc = L.oracle_pack(buf, _p(offs, _i64p), n, _p(lengths, _i32p), _p(mags, _i32p),
                           _p(lingos, _i32p), _p(counts, _i32p), w)
        assert rc == 0
    return lingos, counts, lengths, mags


class _Set:
    def __init__(self, lingos, counts, lengths, mags):
        self.l = _c32(lingos)
        self.c = _c32(counts)
        self.len = _vec(lengths)
        self.mag = _vec(mags)
        self.n = self.l.shape[0]
        self.pitch = self.l.shape[1] if self.l.ndim == 2 else 0
        if self.pitch == 0:  # keep a valid pointer for zero-width sets
            self.l = np.zeros((self.n, 1), np.int32)
            self.c = np.zeros((self.n, 1), np.int32)
            self.pitch = 1

    def args(self):
        return (_p(self.l, _i32p), _p(self.c, _i32p), _p(self.mag, _i32p), _p(self.len, _i32p),
                self.n, self.pitch)


def tanimoto_block(ref, query, nthreads=1):
    """float32 (nr, nq) block, _CPULingo getTanimotoBlock[Parallel].
    ref/query are (lingos, counts, lengths, mags) tuples."""
    R, Q = _Set(*ref), _Set(*query)
    out = np.empty((R.n, Q.n), dtype=np.float32)
    if R.n and Q.n:
        lib().oracle_tanimoto_block(*R.args(), *Q.args(), _p(out, _f32p), Q.n, nthreads)
    return out


def intersection_block(ref, query, nthreads=0):
    R, Q = _Set(*ref), _Set(*query)
    out = np.empty((R.n, Q.n), dtype=np.int32)
    if R.n and Q.n:
        lib().oracle_intersection_block(_p(R.l, _i32p), _p(R.c, _i32p), _p(R.len, _i32p), R.n, R.pitch,
                                        _p(Q.l, _i32p), _p(Q.c, _i32p), _p(Q.len, _i32p), Q.n, Q.pitch,
                                        _p(out, _i32p), nthreads)
    return out


def single_row(ref, row, query):
    """GPULingo.getTanimotoRow semantics (recomputed magnitudes)."""
    R, Q = _Set(*ref), _Set(*query)
    out = np.empty(Q.n, dtype=np.float32)
    a = np.ascontiguousarray(R.l[row])
    ac = np.ascontiguousarray(R.c[row])
    if Q.n:
        lib().oracle_single_row(_p(a, _i32p), _p(ac, _i32p), int(R.len[row]), _p(Q.l, _i32p), _p(Q.c, _i32p),
                                _p(Q.len, _i32p), Q.n, Q.pitch, _p(out, _f32p))
    return out


def row_hist101(mat):
    """_cpuutil.rowHistogram101 on a float32 or float64 (R, C) matrix."""
    mat = np.asarray(mat)
    if mat.ndim != 2:
        raise ValueError("Data given to rowHistogram101 was not of dimension 2")
    hist = np.zeros((mat.shape[0], 101), dtype=np.uint32)
    if mat.dtype == np.float32:
        m = np.ascontiguousarray(mat)
        lib().oracle_row_hist101_f32(_p(m, _f32p), m.shape[0], m.shape[1], m.shape[1], _p(hist, _u32p))
    elif mat.dtype == np.float64:
        m = np.ascontiguousarray(mat)
        lib().oracle_row_hist101_f64(_p(m, _f64p), m.shape[0], m.shape[1], m.shape[1], _p(hist, _u32p))
    else:
        raise ValueError("Data given to rowHistogram101 was not of float (float32) or double (float64) type")
    return hist


def hist_rows101(ref, query, nthreads=0):
    R, Q = _Set(*ref), _Set(*query)
    hist = np.zeros((R.n, 101), dtype=np.int32)
    if R.n:
        lib().oracle_hist_rows101(*R.args(), *Q.args(), _p(hist, _i32p), nthreads)
    return hist


def hist1000(ref, query):
    R, Q = _Set(*ref), _Set(*query)
    hist = np.zeros(1000, dtype=np.int64)
    if R.n and Q.n:
        lib().oracle_hist1000(*R.args(), *Q.args(), _p(hist, _i64p))
    return hist


def neighbors(ref, query, lo, hi=1.1, nthreads=0):
    """Complete neighbour lists: (list of int32 arrays, int64 counts)."""
    R, Q = _Set(*ref), _Set(*query)
    idx = np.empty((R.n, max(Q.n, 1)), dtype=np.int32)
    counts = np.zeros(R.n, dtype=np.int64)
    if R.n:
        lib().oracle_neighbors(*R.args(), *Q.args(), ctypes.c_float(np.float32(lo)),
                               ctypes.c_float(np.float32(hi)), _p(idx, _i32p), _p(counts, _i64p), nthreads)
    return [idx[r, :counts[r]].copy() for r in range(R.n)], counts


def topk(ref, query, k, nthreads=0):
    R, Q = _Set(*ref), _Set(*query)
    oi = np.empty((R.n, k), dtype=np.int32)
    ot = np.empty((R.n, k), dtype=np.float32)
    if R.n and k > 0:
        lib().oracle_topk(*R.args(), *Q.args(), k, _p(oi, _i32p), _p(ot, _f32p), nthreads)
    return oi, ot


def bin1000(t):
    return int(lib().oracle_bin1000(ctypes.c_float(np.float32(t))))


def bin101_f32(t):
    return int(lib().oracle_bin101_f32(ctypes.c_float(np.float32(t))))


def max_threads():
    return int(lib().oracle_max_threads())
