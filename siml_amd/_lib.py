"""ctypes binding of libsiml_amd.so (declarations in include/siml_amd.h).

The library is built in-tree by ``siml_amd._build``; importing this module
never builds it and never falls back to another implementation: a missing
library raises ImportError.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsiml_amd.so")

SIML_OK = 0
SIML_EINVAL = -1
SIML_ENOMEM = -2
SIML_EDEVICE = -3
SIML_ESTATE = -4
SIML_ELIMIT = -5

RESULT_NONE, RESULT_F32, RESULT_I32, RESULT_CSR, RESULT_TOPK, RESULT_I64 = range(6)
MAX_ROW_LINGOS = 8192
MAX_TOPK = 256
POLICY_AUTO, POLICY_MERGE, POLICY_BITS, POLICY_SPARSE = 0, 1, 2, 3
LAYOUT_NATURAL, LAYOUT_BLOCK, LAYOUT_SORTED = 0, 1, 2
OP_TOPK, OP_NEIGHBORS, OP_HIST1000 = 1, 2, 3
NOALLOC_SCRATCH_BYTES = 128 << 20

_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)
_i64 = ctypes.c_int64
_int = ctypes.c_int
_vp = ctypes.c_void_p

# name -> (restype, argtypes); the exported surface of include/siml_amd.h
SIGNATURES = {
    "siml_last_error": (ctypes.c_char_p, []),
    "siml_version": (_int, []),
    "siml_device_count": (_int, [ctypes.POINTER(_int)]),
    "siml_pack_measure": (_int, [ctypes.c_char_p, _i64p, _i64, _i32p, _i32p, _int]),
    "siml_pack_fill": (_int, [ctypes.c_char_p, _i64p, _i64, _i32p, _i32p, _i64, _int]),
    "siml_pack_fill_ragged": (_int, [ctypes.c_char_p, _i64p, _i64, _i64p, _i32p, _i32p, _int]),
    "siml_tanimoto_block": (_int, [_int, _i32p, _i32p, _i32p, _i32p, _i64, _i64, _i64,
                                   _i32p, _i32p, _i32p, _i32p, _i64, _i64, _i64, _f32p, _i64]),
    "siml_row_hist101_f32": (_int, [_int, _f32p, _i64, _i64, _i64, _u32p]),
    "siml_row_hist101_f64": (_int, [_int, _f64p, _i64, _i64, _i64, _u32p]),
    "siml_create": (_int, [_int, ctypes.POINTER(_vp)]),
    "siml_destroy": (_int, [_vp]),
    "siml_set_ref": (_int, [_vp, _i32p, _i32p, _i32p, _i32p, _i64, _i64, _i64]),
    "siml_set_query": (_int, [_vp, _i32p, _i32p, _i32p, _i32p, _i64, _i64, _i64]),
    "siml_set_ref_ragged": (_int, [_vp, _i32p, _i32p, _i32p, _i64p, _i64]),
    "siml_set_query_ragged": (_int, [_vp, _i32p, _i32p, _i32p, _i64p, _i64]),
    "siml_set_sizes": (_int, [_vp, _i64p, _i64p]),
    "siml_db_create": (_int, [_int, _i32p, _i32p, _i32p, _i64p, _i64, _int, ctypes.POINTER(_vp)]),
    "siml_db_destroy": (_int, [_vp]),
    "siml_db_size": (_int, [_vp, _i64p]),
    "siml_attach_db": (_int, [_vp, _vp]),
    "siml_db_save": (_int, [_vp, ctypes.c_char_p]),
    "siml_db_load": (_int, [_int, ctypes.c_char_p, ctypes.POINTER(_vp)]),
    "siml_multirow_async": (_int, [_vp, _i64, _i64]),
    "siml_single_row_async": (_int, [_vp, _i64]),
    "siml_hist_rows_async": (_int, [_vp, _i64, _i64]),
    "siml_neighbors_async": (_int, [_vp, _i64, _i64, ctypes.c_float, ctypes.c_float]),
    "siml_topk_async": (_int, [_vp, _i64, _i64, ctypes.c_int32]),
    "siml_hist1000_async": (_int, [_vp, _i64, _i64]),
    "siml_topk_noalloc": (_int, [_vp, _i64, _i64, ctypes.c_int32, _vp, _vp]),
    "siml_query_done": (_int, [_vp, ctypes.POINTER(_int)]),
    "siml_synchronize": (_int, [_vp]),
    "siml_result_info": (_int, [_vp, ctypes.POINTER(_int), _i64p, _i64p, _i64p]),
    "siml_retrieve_f32": (_int, [_vp, _f32p, _i64]),
    "siml_retrieve_i32": (_int, [_vp, _i32p, _i64]),
    "siml_retrieve_csr": (_int, [_vp, _i64p, _i32p]),
    "siml_retrieve_topk": (_int, [_vp, _i32p, _f32p]),
    "siml_retrieve_i64": (_int, [_vp, _i64p]),
    "siml_peek_rows_f32": (_int, [_vp, _i64, _i64, _f32p, _i64]),
    "siml_set_timing": (_int, [_vp, _int]),
    "siml_set_policy": (_int, [_vp, _int]),
    "siml_set_layout": (_int, [_vp, _int]),
    "siml_kernel_stats": (_int, [_vp, ctypes.POINTER(ctypes.c_double), _i64p, _i64p, _int]),
    "siml_last_kernel": (ctypes.c_char_p, [_vp]),
    "siml_work_stats": (_int, [_vp, _i64p, _i64p, _int]),
    "siml_selftest_division": (_int, [_int, _int, _i64p, _i64p, _i64p]),
    "siml_set_autotune": (_int, [_vp, _int]),
    "siml_set_graphs": (_int, [_vp, _int]),
    "siml_set_persistent": (_int, [_vp, _int]),
    "siml_graph_stats": (_int, [_vp, _i64p, _i64p]),
    "siml_selftest_valu": (_int, [_int, _int, ctypes.POINTER(ctypes.c_double)]),
    "siml_selftest_lds": (_int, [_int, _int, ctypes.POINTER(ctypes.c_double)]),
    "siml_row_groups": (_int, [_vp, _i64p, _i64p]),
    "siml_kernel_split": (_int, [_vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
    "siml_launch_shape": (_int, [_vp, _i64p, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]),
    "siml_set_launch_shape": (_int, [_vp, _i64, ctypes.c_int32, ctypes.c_int32]),
    "siml_set_group_tables": (_int, [_vp, _int]),
    "siml_call_bytes": (_int, [_vp, _int, _i64, ctypes.c_int32, _i64p]),
}

_lib = None


def load():
    """Load libsiml_amd.so and declare its signatures.  Raises ImportError if
    the library has not been built (there is no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run `python -m siml_amd._build` "
                              "(or __graft_entry__.build()) first")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class SimlDeviceError(RuntimeError):
    """A HIP runtime or kernel error reported by libsiml_amd.so."""


def check(rc):
    """Map a SIML_E* return code to the Python exception the reference
    raises for that class of failure (ValueError for argument errors)."""
    if rc == SIML_OK:
        return
    msg = load().siml_last_error().decode(errors="replace")
    if rc in (SIML_EINVAL, SIML_ELIMIT):
        raise ValueError(msg)
    if rc == SIML_ENOMEM:
        raise MemoryError(msg)
    if rc == SIML_ESTATE:
        raise RuntimeError(msg)
    raise SimlDeviceError(msg)


def ptr(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def i32p(a):
    return ptr(a, _i32p)


def i64p(a):
    return ptr(a, _i64p)


def f32p(a):
    return ptr(a, _f32p)


def device_count():
    n = _int(0)
    check(load().siml_device_count(ctypes.byref(n)))
    return n.value


class Handle:
    """Owns one siml_ctx (one HIP stream on one device)."""

    def __init__(self, device=0):
        self._lib = load()
        h = _vp()
        check(self._lib.siml_create(int(device), ctypes.byref(h)))
        self.h = h
        self.device = int(device)

    def close(self):
        if getattr(self, "h", None):
            self._lib.siml_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def call(self, name, *args):
        check(getattr(self._lib, name)(self.h, *args))

    def result_info(self):
        kind, rows, cols, nnz = _int(0), _i64(0), _i64(0), _i64(0)
        self.call("siml_result_info", ctypes.byref(kind), ctypes.byref(rows), ctypes.byref(cols),
                  ctypes.byref(nnz))
        return kind.value, rows.value, cols.value, nnz.value

    def done(self):
        d = _int(0)
        self.call("siml_query_done", ctypes.byref(d))
        return bool(d.value)

    def kernel_stats(self, reset=False):
        ms, n, pairs = ctypes.c_double(0), _i64(0), _i64(0)
        self.call("siml_kernel_stats", ctypes.byref(ms), ctypes.byref(n), ctypes.byref(pairs), int(reset))
        return ms.value, n.value, pairs.value

    def kernel_split(self):
        """(similarity ms, top-k selection ms) of kernel_stats' time since
        its last reset."""
        a, b = ctypes.c_double(0), ctypes.c_double(0)
        self.call("siml_kernel_split", ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def row_groups(self):
        """(bit-sliced row groups, rows on the merge kernel) of the
        reference set."""
        g, m = _i64(0), _i64(0)
        self.call("siml_row_groups", ctypes.byref(g), ctypes.byref(m))
        return g.value, m.value

    def work_stats(self, reset=False):
        """(pairs evaluated, sum of La + Lb over them) by the neighbour
        kernels while timing was on, since the last reset."""
        pairs, lsum = _i64(0), _i64(0)
        self.call("siml_work_stats", ctypes.byref(pairs), ctypes.byref(lsum), int(reset))
        return pairs.value, lsum.value

    def last_kernel(self):
        return self._lib.siml_last_kernel(self.h).decode()

    def sizes(self):
        a, b = _i64(0), _i64(0)
        self.call("siml_set_sizes", ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    # -- sets -------------------------------------------------------------
    def set_padded(self, which, lingos, counts, lengths, mags):
        n, w = lingos.shape
        fn = "siml_set_ref" if which == "ref" else "siml_set_query"
        lingos = lingos if w > 0 else np.zeros((n, 1), np.int32)
        counts = counts if w > 0 else np.zeros((n, 1), np.int32)
        pitch = lingos.shape[1]
        self.call(fn, i32p(lingos), i32p(counts), i32p(mags), i32p(lengths), n, w, pitch)

    def set_ragged(self, which, ids, counts, offsets, mags):
        fn = "siml_set_ref_ragged" if which == "ref" else "siml_set_query_ragged"
        ids = ids if ids.size else np.zeros(1, np.int32)
        counts = counts if counts.size else np.zeros(1, np.int32)
        self.call(fn, i32p(ids), i32p(counts), i32p(mags), i64p(offsets), len(offsets) - 1)

    # -- retrieval --------------------------------------------------------
    def retrieve_f32(self, out, pitch):
        self.call("siml_retrieve_f32", f32p(out), int(pitch))

    def retrieve_i32(self, out, pitch):
        self.call("siml_retrieve_i32", i32p(out), int(pitch))

    def retrieve_csr(self):
        kind, rows, cols, nnz = self.result_info()
        if kind != RESULT_CSR:
            raise RuntimeError("no pending neighbour result")
        offs = np.zeros(rows + 1, dtype=np.int64)
        idx = np.empty(max(nnz, 1), dtype=np.int32)
        self.call("siml_retrieve_csr", i64p(offs), i32p(idx))
        return offs, idx[:nnz]

    def retrieve_topk(self):
        kind, rows, cols, nnz = self.result_info()
        if kind != RESULT_TOPK:
            raise RuntimeError("no pending top-k result")
        idx = np.empty((rows, cols), dtype=np.int32)
        t = np.empty((rows, cols), dtype=np.float32)
        self.call("siml_retrieve_topk", i32p(idx), f32p(t))
        return idx, t

    def retrieve_i64(self):
        kind, rows, cols, nnz = self.result_info()
        if kind != RESULT_I64:
            raise RuntimeError("no pending int64 result")
        out = np.zeros(cols, dtype=np.int64)
        self.call("siml_retrieve_i64", i64p(out))
        return out
