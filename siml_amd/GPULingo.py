"""GPULingo: LINGO similarities on an MI355X (drop-in for pysiml.GPULingo).

Method names, argument meaning, return shapes and dtypes follow
GPULingo.py:158-648 of the reference; computation runs in the HIP kernels of
libsiml_amd.so.  Rows are the reference set (set_refsmiles), columns the
query set (set_qsmiles).

Deliberate differences from the reference, all on paths where it fails or is
undefined:
  * argument errors raise ValueError (the reference raises a bare ``raise``
    or the undefined ``NotImplementedException``);
  * there are no 32K/64K set-size limits (GPULingo.py:206-207, 269-270);
  * neighbour lists are computed completely; getNeighbors reports the
    reference's overflow count (kernels.py:466-469) but its matrix never
    exceeds the padded pitch, and unused entries are -1 rather than
    uninitialised memory.
Additive methods: getTopK, getGlobalHistogram, getNeighborsCSR.
"""
import numpy as np

from . import _lib


def _check_int32_matrix(a, name):
    a = np.asarray(a)
    if a.dtype != np.int32:
        raise ValueError(f"{name} was not of type int32")
    if a.ndim != 2:
        raise ValueError(f"{name} did not have dimension 2")
    return np.ascontiguousarray(a)


def _check_vector(v, n, name):
    v = np.asarray(v)
    if v.dtype != np.int32:
        raise ValueError(f"{name} was not of type int32")
    if not (v.ndim == 1 or (v.ndim == 2 and v.shape[1] == 1)):
        raise ValueError(f"{name} was not 1-D")
    if v.shape[0] != n:
        raise ValueError(f"{name} length did not equal number of rows of the Lingo matrix")
    return np.ascontiguousarray(v.reshape(-1))


def normalise_set(lingos, counts, lengths, mags, label):
    """Validate a SMILES set the way _CPULingo.c:139-215 does and return
    contiguous int32 arrays (mags may be None)."""
    lingos = _check_int32_matrix(lingos, f"{label} Lingo matrix")
    counts = _check_int32_matrix(counts, f"{label} Lingo count matrix")
    if lingos.shape != counts.shape:
        raise ValueError(f"{label} Lingo and Lingo count matrix did not have identical shapes")
    n = lingos.shape[0]
    lengths = _check_vector(lengths, n, f"{label} length vector")
    if mags is not None:
        mags = _check_vector(mags, n, f"{label} magnitude vector")
    return lingos, counts, lengths, mags


def reference_neighbor_counts(offsets, indices, pitch):
    """Per-row counts as gpuGetNeighbors_multirow reports them
    (kernels.py:433-474): it scans the columns in waves of 128, adds each
    wave's neighbours and stops after the first wave that takes the count
    past `pitch`, so an overflowing row reports the number of neighbours
    with column < 128 * (wave of neighbour `pitch` + 1)."""
    rows = len(offsets) - 1
    out = np.empty(rows, dtype=np.int32)
    for r in range(rows):
        lst = indices[offsets[r]:offsets[r + 1]]
        c = len(lst)
        if c <= pitch:
            out[r] = c
        else:
            limit = 128 * (int(lst[pitch]) // 128 + 1)
            out[r] = int(np.searchsorted(lst, limit, side="left"))
    return out


class GPULingo:
    """Object to handle computation of LINGO similarities on the GPU with
    HIP device ID *deviceID* (GPULingo.py:158-174)."""

    def __init__(self, deviceID=0):
        self.deviceID = deviceID
        self.nref = 0
        self.nquery = 0
        self.rlengths = None
        self.qshape = None
        self._h = _lib.Handle(deviceID)
        self._retrieve = None
        self._pending_kind = None

    # -- set-up ------------------------------------------------------------
    def set_refsmiles(self, refsmilesmat, refcountsmat, reflengths, refmags=None):
        """GPULingo.py:186-243.  Without *refmags* the magnitudes are the
        row sums of the count matrix (accumulateRefMagnitudes,
        kernels.py:45-61)."""
        l, c, n, m = normalise_set(refsmilesmat, refcountsmat, reflengths, refmags, "Reference")
        self._h.set_padded("ref", l, c, n, m)
        self.rlengths = n
        self.rshape = l.shape
        self.nref = l.shape[0]

    def set_qsmiles(self, qsmilesmat, qcountsmat, querylengths, querymags=None):
        """GPULingo.py:245-314 (magnitudes: accumulateQueryMagnitudes,
        kernels.py:63-79)."""
        l, c, n, m = normalise_set(qsmilesmat, qcountsmat, querylengths, querymags, "Query")
        self._h.set_padded("query", l, c, n, m)
        self.qshape = l.shape
        self.nquery = l.shape[0]

    def set_refsmiles_ragged(self, ids, counts, offsets, mags=None):
        """Ragged reference set (no counterpart in the reference)."""
        ids, counts, offsets, mags = _ragged(ids, counts, offsets, mags)
        self._h.set_ragged("ref", ids, counts, offsets, mags)
        self.nref = len(offsets) - 1
        self.rlengths = np.diff(offsets).astype(np.int32)

    def set_qsmiles_ragged(self, ids, counts, offsets, mags=None):
        """Ragged query set (no counterpart in the reference)."""
        ids, counts, offsets, mags = _ragged(ids, counts, offsets, mags)
        self._h.set_ragged("query", ids, counts, offsets, mags)
        self.nquery = len(offsets) - 1
        self.qshape = (self.nquery, 0)

    def attach_db(self, db):
        """Use the resident database *db* (a LingoDB on the same device) as
        the query set; several comparators may share one database."""
        self._h.call("siml_attach_db", db.h)
        self._db = db  # keep the database object alive while attached
        self.nquery = len(db)
        self.qshape = (self.nquery, 0)


    def set_autotune(self, on=True):
        """Launch shapes chosen by measurement on the first call of each
        shape bucket (results are unchanged; see siml_set_autotune)."""
        self._h.call("siml_set_autotune", 1 if on else 0)

    def set_graphs(self, on=True):
        """Top-k and global-histogram calls recorded by stream capture and
        launched as one HIP graph each (results are unchanged; see
        siml_set_graphs)."""
        self._h.call("siml_set_graphs", 1 if on else 0)

    def set_persistent(self, on=True):
        """Top-k and global-histogram similarity passes on a persistent grid
        with per-XCD work queues (results are unchanged; turns the autotuner
        off; see siml_set_persistent)."""
        self._h.call("siml_set_persistent", 1 if on else 0)

    def graph_stats(self):
        """(graph launches, graph instantiations) of this comparator."""
        import ctypes
        n, b = ctypes.c_int64(), ctypes.c_int64()
        self._h.call("siml_graph_stats", ctypes.byref(n), ctypes.byref(b))
        return n.value, b.value

    def launch_shape(self):
        """(workgroups aimed at per launch, most waves per top-k row,
        top-k seed stride: 0 = no seeding) of the last call."""
        import ctypes
        t, c, sd = ctypes.c_int64(), ctypes.c_int32(), ctypes.c_int32()
        self._h.call("siml_launch_shape", ctypes.byref(t), ctypes.byref(c), ctypes.byref(sd))
        return t.value, c.value, sd.value

    def set_group_tables(self, max_tables=2):
        """Most row groups (1, 2 or 4) one workgroup of the bit-sliced
        kernels serves; results do not depend on it."""
        self._h.call("siml_set_group_tables", int(max_tables))

    def set_launch_shape(self, target_wg=8192, topk_wave_cap=1024, topk_seed=True):
        """Fixes the launch shape (turns the autotuner off).  topk_seed: the
        top-k seed stride (0 or False: none; True or 1: a stride from the
        query set's size, tiles / 512 clamped to [16, 64]; an int in
        [2, 64]: every topk_seed-th tile)."""
        seed = 1 if topk_seed is True else int(topk_seed)
        self._h.call("siml_set_launch_shape", int(target_wg), int(topk_wave_cap), seed)

    # -- helpers -----------------------------------------------------------
    def _rows(self, rowbase, rowlimit):
        rowbase, rowlimit = int(rowbase), int(rowlimit)
        if rowbase < 0 or rowlimit > self.nref or rowlimit < rowbase:
            raise ValueError(f"rows [{rowbase}, {rowlimit}) outside [0, {self.nref})")
        return rowbase, rowlimit

    def _pend(self, fn, kind="matrix"):
        self._retrieve = fn
        self._pending_kind = kind

    # -- similarity rows -----------------------------------------------------
    def getTanimotoRow(self, rowidx):
        """Single Tanimoto row, synchronous (GPULingo.py:316-325)."""
        self.getTanimotoRow_async(rowidx)
        return self.retrieveAsyncResult()

    def getTanimotoRow_async(self, rowidx):
        """GPULingo.py:327-352 / gpuLingoSim3_srgf_T_tex2ns: magnitudes are
        recomputed from the counts and the zero test is on lengths."""
        rowidx = int(rowidx)
        if rowidx < 0 or rowidx >= self.nref:
            raise ValueError(f"row {rowidx} outside [0, {self.nref})")
        self._h.call("siml_single_row_async", rowidx)
        self._pend(self._retrieve_matrix)

    def getMultipleRows(self, rowbase, rowlimit):
        """GPULingo.py:354-363."""
        self.getMultipleRows_async(rowbase, rowlimit)
        return self.retrieveAsyncResult()

    def getMultipleRows_async(self, rowbase, rowlimit):
        """GPULingo.py:365-405 / gpuLingoSim3_multirow."""
        rowbase, rowlimit = self._rows(rowbase, rowlimit)
        self._h.call("siml_multirow_async", rowbase, rowlimit)
        self._pend(self._retrieve_matrix)

    def getMultipleHistogrammedRows(self, rowbase, rowlimit):
        """GPULingo.py:407-417."""
        self.getMultipleHistogrammedRows_async(rowbase, rowlimit)
        return self.retrieveAsyncResult()

    def getMultipleHistogrammedRows_async(self, rowbase, rowlimit):
        """GPULingo.py:419-464 / gpuHistogrammedLingo_multirow: int32
        (rows, 101), bin (unsigned)(T*100)."""
        rowbase, rowlimit = self._rows(rowbase, rowlimit)
        self._h.call("siml_hist_rows_async", rowbase, rowlimit)
        self._pend(self._retrieve_hist, "histogram")

    def getNeighbors(self, rowbase, rowlimit, lowerbound, upperbound=1.1, maxneighbors=None):
        """GPULingo.py:516-535."""
        self.getNeighbors_async(rowbase, rowlimit, lowerbound, upperbound, maxneighbors)
        return self.retrieveAsyncResult()

    def getNeighbors_async(self, rowbase, rowlimit, lowerbound, upperbound=1.1, maxneighbors=None):
        """GPULingo.py:466-514 / gpuGetNeighbors_multirow: (matrix, counts)
        with ascending query indices, lowerbound <= T < upperbound (float32
        bounds)."""
        rowbase, rowlimit = self._rows(rowbase, rowlimit)
        if maxneighbors is None:
            maxneighbors = self.nquery
        pitch = 16 * ((int(maxneighbors) + 15) // 16)
        self._h.call("siml_neighbors_async", rowbase, rowlimit, float(np.float32(lowerbound)),
                     float(np.float32(upperbound)))
        self._pend(lambda trimmed=True: self._retrieve_neighbors(pitch, maxneighbors, trimmed), "neighbors")

    # -- additive -------------------------------------------------------------
    def getNeighborsCSR(self, rowbase, rowlimit, lowerbound, upperbound=1.1):
        """Complete neighbour lists as CSR: (offsets int64 (rows+1,),
        indices int32 (nnz,)), ascending within each row."""
        rowbase, rowlimit = self._rows(rowbase, rowlimit)
        self._h.call("siml_neighbors_async", rowbase, rowlimit, float(np.float32(lowerbound)),
                     float(np.float32(upperbound)))
        self._retrieve = None
        return self._h.retrieve_csr()

    def getTopK(self, rowbase, rowlimit, k):
        """Top-k query columns per row: (indices int32 (rows,k), T float32
        (rows,k)); larger T first, ties by smaller index."""
        rowbase, rowlimit = self._rows(rowbase, rowlimit)
        self._h.call("siml_topk_async", rowbase, rowlimit, int(k))
        self._retrieve = None
        return self._h.retrieve_topk()

    def getGlobalHistogram(self, rowbase, rowlimit):
        """int64 counts of all T of the block in 1000 bins with float32 edges
        fl32(k/1000)."""
        rowbase, rowlimit = self._rows(rowbase, rowlimit)
        self._h.call("siml_hist1000_async", rowbase, rowlimit)
        self._retrieve = None
        return self._h.retrieve_i64()

    # -- async protocol -------------------------------------------------------
    def _retrieve_matrix(self):
        kind, rows, cols, _ = self._h.result_info()
        pitch = 16 * ((cols + 15) // 16)
        out = np.empty((rows, pitch), dtype=np.float32)
        self._h.retrieve_f32(out, pitch)
        return out[:, :cols]

    def _retrieve_hist(self):
        kind, rows, cols, _ = self._h.result_info()
        out = np.empty((rows, 112), dtype=np.int32)
        self._h.retrieve_i32(out, 112)
        return out[:, :101]

    def _retrieve_neighbors(self, pitch, maxneighbors, trimmed):
        offs, idx = self._h.retrieve_csr()
        counts = reference_neighbor_counts(offs, idx, pitch)
        rows = len(offs) - 1
        if trimmed:
            width = int(min(counts.max(), pitch)) if rows else 0
        else:
            width = int(maxneighbors)
        mat = np.full((rows, width), -1, dtype=np.int32)
        for r in range(rows):
            lst = idx[offs[r]:offs[r + 1]][:width]
            mat[r, :len(lst)] = lst
        return mat, counts

    def retrieveAsyncResult(self):
        """Result of the last asynchronous call (GPULingo.py:537-641); blocks
        until it is complete.  Valid once per asynchronous call."""
        fn, self._retrieve = self._retrieve, None
        if fn is None:
            raise RuntimeError("Error: tried to call retrieveAsyncResult without a pending asynchronous result")
        return fn()

    def _retrieve_kind(self, kinds, what):
        if self._retrieve is not None and self._pending_kind not in kinds:
            raise ValueError(f"{what}: the pending result is a {self._pending_kind} result")

    def retrieveAsyncMatrixResult(self):
        """GPULingo.py:538-556: the pending single-row, multi-row or
        histogrammed-row matrix."""
        self._retrieve_kind(("matrix", "histogram"), "retrieveAsyncMatrixResult")
        return self.retrieveAsyncResult()

    def retrieveAsyncMatrixVectorResult(self):
        """GPULingo.py:559-584: the pending neighbour search as (matrix with
        maxneighbors columns, counts); entries past a row's count are -1."""
        self._retrieve_kind(("neighbors",), "retrieveAsyncMatrixVectorResult")
        fn, self._retrieve = self._retrieve, None
        if fn is None:
            raise RuntimeError("Error: tried to call retrieveAsyncResult without a pending asynchronous result")
        return fn(trimmed=False)

    def retrieveAsyncMatrixVectorResult_trimmed(self):
        """GPULingo.py:586-623: the pending neighbour search as (matrix with
        max(counts) columns, counts) -- what retrieveAsyncResult returns after
        getNeighbors_async (GPULingo.py:513)."""
        self._retrieve_kind(("neighbors",), "retrieveAsyncMatrixVectorResult_trimmed")
        return self.retrieveAsyncResult()

    def retrieveAsyncResultBase(self):
        """GPULingo.py:625-628: the retrieve of an object with no pending
        result."""
        raise RuntimeError("Error: tried to call retrieveAsyncResult without a pending asynchronous result")

    def asyncOperationsDone(self):
        """True once all work queued on this object has completed
        (GPULingo.py:643-647)."""
        return self._h.done()


def _ragged(ids, counts, offsets, mags):
    ids = np.ascontiguousarray(ids, dtype=np.int32).reshape(-1)
    counts = np.ascontiguousarray(counts, dtype=np.int32).reshape(-1)
    offsets = np.ascontiguousarray(offsets, dtype=np.int64).reshape(-1)
    if ids.shape != counts.shape:
        raise ValueError("ids and counts differ in length")
    if offsets.size == 0 or offsets[-1] > ids.size:
        raise ValueError("offsets do not fit the id array")
    if mags is not None:
        mags = np.ascontiguousarray(mags, dtype=np.int32).reshape(-1)
        if mags.size != offsets.size - 1:
            raise ValueError("magnitude vector length does not match offsets")
    return ids, counts, offsets, mags


class LingoDB:
    """A query (column) set built once in device memory and shared by any
    number of comparators on the same device (no counterpart in the
    reference; the reference re-uploads its query set per comparator,
    GPULingo.py:245-314).  Input is ragged: molecule r owns
    ids[offsets[r]:offsets[r+1]]."""

    def __init__(self, ids, counts, offsets, mags=None, device=0, layout=_lib.LAYOUT_SORTED):
        import ctypes
        ids, counts, offsets, mags = _ragged(ids, counts, offsets, mags)
        self._lib = _lib.load()
        h = ctypes.c_void_p()
        ids_ = ids if ids.size else np.zeros(1, np.int32)
        counts_ = counts if counts.size else np.zeros(1, np.int32)
        _lib.check(self._lib.siml_db_create(int(device), _lib.i32p(ids_), _lib.i32p(counts_), _lib.i32p(mags),
                                            _lib.i64p(offsets), len(offsets) - 1, int(layout), ctypes.byref(h)))
        self.h = h
        self.device = int(device)
        self.n = len(offsets) - 1

    @classmethod
    def load(cls, path, device=0):
        """A database saved with save(), uploaded without rebuilding."""
        import ctypes
        import os
        self = cls.__new__(cls)
        self._lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(self._lib.siml_db_load(int(device), os.fsencode(path), ctypes.byref(h)))
        self.h = h
        self.device = int(device)
        n = ctypes.c_int64()
        _lib.check(self._lib.siml_db_size(h, ctypes.byref(n)))
        self.n = n.value
        return self

    def save(self, path):
        """Writes the built database (device arrays, format-versioned)."""
        import os
        _lib.check(self._lib.siml_db_save(self.h, os.fsencode(path)))

    def __len__(self):
        return self.n

    def close(self):
        if getattr(self, "h", None):
            self._lib.siml_db_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
