"""CPULingo interface served by the MI355X kernels (drop-in for
pysiml.CPULingo).

The reference's CPULingo (CPULingo.py:43-197) wraps the OpenMP C extension
_CPULingo.  This class keeps its interface and semantics exactly -- Python
slicing of the row range, the default magnitudes ``int32(lengths.sum(1))``
(CPULingo.py:82,98), validation errors raised at compute time, a single
synchronous "async" slot -- and computes the block with the HIP kernels of
libsiml_amd.so (the same kernels as GPULingo.getMultipleRows, whose float
expression is the one of _CPULingo.c:70).  ``nprocs`` is accepted and has no
effect; ``supportsParallel()`` is True.  The HIP device is
``$SIML_DEVICE`` (default 0).
"""
import os

import numpy as np

from . import _lib
from .GPULingo import normalise_set


class CPULingo:
    """Object to handle computation of LINGO similarities (CPULingo.py:43)."""

    def __init__(self):
        self.rlingos = None
        self.rcounts = None
        self.rlengths = None
        self.rmags = None
        self.qlingos = None
        self.qcounts = None
        self.qlengths = None
        self.qmags = None
        self.lastasyncresult = None
        self._h = None
        self._snap = {}   # set name -> contents last uploaded

    def supportsParallel(self):
        """CPULingo.py:61-68; the device path is always parallel."""
        return True

    def set_refsmiles(self, refsmilesmat, refcountsmat, reflengths, refmags=None):
        """CPULingo.py:70-84: stores the set; without *refmags* the
        magnitudes are int32(reflengths.sum(1)), i.e. the number of unique
        Lingos, exactly as the reference does."""
        self.rlingos = refsmilesmat
        self.rcounts = refcountsmat
        self.rlengths = reflengths
        self.rmags = refmags if refmags is not None else np.int32(reflengths.sum(1))

    def set_qsmiles(self, qsmilesmat, qcountsmat, querylengths, querymags=None):
        """CPULingo.py:86-101."""
        self.qlingos = qsmilesmat
        self.qcounts = qcountsmat
        self.qlengths = querylengths
        self.qmags = querymags if querymags is not None else np.int32(querylengths.sum(1))

    def _checkInitialized(self):
        """CPULingo.py:103-108."""
        if self.rlingos is None or self.rcounts is None or self.rlengths is None or self.rmags is None or \
           self.qlingos is None or self.qcounts is None or self.qlengths is None or self.qmags is None:
            raise ValueError("CPULingo object not fully initialized")

    def _handle(self):
        if self._h is None:
            self._h = _lib.Handle(int(os.environ.get("SIML_DEVICE", "0")))
        return self._h

    def _sync(self, which):
        """Uploads set `which` ("ref" or "query") unless the device already
        holds exactly its current contents.  The reference reads the arrays
        on every call (CPULingo.py:126-138), so an in-place edit after
        set_refsmiles/set_qsmiles is seen by the next call here too."""
        if which == "ref":
            arrays, label = (self.rlingos, self.rcounts, self.rlengths, self.rmags), "Reference"
        else:
            arrays, label = (self.qlingos, self.qcounts, self.qlengths, self.qmags), "Query"
        cur = normalise_set(*arrays, label)
        snap = self._snap.get(which)
        if snap is not None and all(a.shape == b.shape and np.array_equal(a, b) for a, b in zip(snap, cur)):
            return
        self._handle().set_padded(which, *cur)
        self._snap[which] = tuple(np.array(a, copy=True) for a in cur)

    def getMultipleRows_async(self, rowbase, rowlimit, nprocs=1):
        """CPULingo.py:111-141: rows self.rlingos[rowbase:rowlimit] (Python
        slice semantics) against the whole query set; synchronous."""
        self._checkInitialized()
        h = self._handle()
        nref = np.asarray(self.rlingos).shape[0]
        start, stop, _ = slice(rowbase, rowlimit).indices(nref)
        stop = max(stop, start)
        self._sync("ref")
        self._sync("query")
        h.call("siml_multirow_async", start, stop)
        kind, rows, cols, _ = h.result_info()
        out = np.empty((rows, cols), dtype=np.float32)
        h.retrieve_f32(out, max(cols, 1))
        self.lastasyncresult = out

    def getTanimotoRow_async(self, row):
        """CPULingo.py:143-155."""
        self.getMultipleRows_async(row, row + 1)

    def getTanimotoRow(self, row):
        """CPULingo.py:157-163."""
        self.getTanimotoRow_async(row)
        return self.retrieveAsyncResult()

    def getMultipleRows(self, rowbase, rowlimit, nprocs=1):
        """CPULingo.py:165-179."""
        self.getMultipleRows_async(rowbase, rowlimit, nprocs)
        return self.retrieveAsyncResult()

    def retrieveAsyncResult(self):
        """CPULingo.py:181-190: the last result (valid until the next call)."""
        return self.lastasyncresult

    def asyncOperationsDone(self):
        """CPULingo.py:192-197."""
        return True
