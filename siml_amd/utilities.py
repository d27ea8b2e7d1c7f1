"""Utilities (drop-in for pysiml.utilities)."""
import os

import numpy as np

from . import _lib


def rowHistogram101(datamatrix):
    """utilities.py:44-55 / _cpuutil.rowHistogram101 (_cpuutil.cpp:74-126):
    uint32 (rows, 101) row histograms with bin min((unsigned)(x*100), 100),
    computed on the GPU.  float64 input is multiplied in double
    (_cpuutil.cpp:92,123).  Values whose product is negative or NaN fall in
    bin 0 and values at or above 1.0 in bin 100 (the documented bins,
    utilities.py:48-53)."""
    a = np.asarray(datamatrix)
    if a.dtype == np.float32:
        fn, ct = "siml_row_hist101_f32", _lib._f32p
    elif a.dtype == np.float64:
        fn, ct = "siml_row_hist101_f64", _lib._f64p
    else:
        raise ValueError("Data given to rowHistogram101 was not of float (float32) or double (float64) type")
    if a.ndim != 2:
        raise ValueError("Data given to rowHistogram101 was not of dimension 2")
    a = np.ascontiguousarray(a)
    hist = np.zeros((a.shape[0], 101), dtype=np.uint32)
    if a.shape[0]:
        dev = int(os.environ.get("SIML_DEVICE", "0"))
        data = a if a.shape[1] else np.zeros((a.shape[0], 1), a.dtype)
        _lib.check(getattr(_lib.load(), fn)(dev, _lib.ptr(data, ct), a.shape[0], a.shape[1],
                                             max(a.shape[1], 1), _lib.ptr(hist, _lib._u32p)))
    return hist
