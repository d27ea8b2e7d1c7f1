"""SMILES -> SIML set conversion (drop-in for pysiml.compiler).

``cSMILEStoMatrices`` packs with the native packer of libsiml_amd.so
(siml_pack_measure / siml_pack_fill), following _ccompiler.SMILEStoMatrices
(_ccompiler.cpp:276-381) and compiler.cSMILEStoMatrices (compiler.py:53-82).
``SMILEStoMatrices``, ``SMILEStoMultiset`` and ``preprocessNumbers`` follow the
pure-Python variants (compiler.py:98-194).

Strings may be ``str`` (encoded as UTF-8, which is what a Python-2 ``str``
read from a UTF-8 file holds) or ``bytes``.
"""
import numpy as np

from . import _lib

__all__ = ["cSMILEStoMatrices", "SMILEStoMatrices", "SMILEStoMultiset", "preprocessNumbers",
           "pack_ragged"]


def _as_bytes(s):
    if isinstance(s, (bytes, bytearray, memoryview)):
        return bytes(s)
    if isinstance(s, str):
        return s.encode("utf-8")
    raise TypeError("Non-string element in sequence")


def _buffer(strings):
    bs = [_as_bytes(s) for s in strings]
    offs = np.zeros(len(bs) + 1, dtype=np.int64)
    if bs:
        np.cumsum([len(b) for b in bs], out=offs[1:])
    return b"".join(bs) or b"\0", offs, len(bs)


def _getSMILESname(smi):
    """compiler.py:47-52: second whitespace-separated field, or ""."""
    fields = smi.split()
    if len(fields) > 1:
        f = fields[1]
        return f.decode("utf-8", "replace") if isinstance(f, bytes) else f
    return ""


def _materialise(smileslist):
    try:
        return list(smileslist)
    except TypeError:
        raise TypeError("Was not given an iterable object")


def cSMILEStoMatrices(smileslist, nthreads=0):
    """Convert the sequence of SMILES strings *smileslist* into a SIML SMILES
    set and list of molecule names (compiler.py:53-82).

    Returns (lingos (N,W) int32, counts (N,W) int32, lengths (N,1) int32,
    mags (N,1) int32, names).  The digit handling is that of the C compiler
    (_ccompiler.cpp:178-264), not of :func:`SMILEStoMatrices`.
    """
    items = _materialise(smileslist)
    buf, offs, n = _buffer(items)
    lib = _lib.load()
    lengths = np.zeros((n, 1), dtype=np.int32)
    mags = np.zeros((n, 1), dtype=np.int32)
    _lib.check(lib.siml_pack_measure(buf, _lib.i64p(offs), n, _lib.i32p(lengths), _lib.i32p(mags), nthreads))
    w = int(lengths.max()) if n else 0
    lingos = np.zeros((n, w), dtype=np.int32)
    counts = np.zeros((n, w), dtype=np.int32)
    if n and w:
        _lib.check(lib.siml_pack_fill(buf, _lib.i64p(offs), n, _lib.i32p(lingos), _lib.i32p(counts), w, nthreads))
    names = [_getSMILESname(s) for s in items]
    return lingos, counts, lengths, mags, names


def pack_ragged(smileslist, nthreads=0):
    """Ragged form of cSMILEStoMatrices for large sets (no counterpart in
    the reference): (ids int32, counts int32, offsets int64 (N+1,),
    mags int32 (N,)).  Molecule r owns ids[offsets[r]:offsets[r+1]]."""
    items = smileslist if isinstance(smileslist, list) else _materialise(smileslist)
    buf, offs, n = _buffer(items)
    return pack_ragged_buffer(buf, offs, nthreads)


def pack_ragged_buffer(buf, offs, nthreads=0):
    """pack_ragged over one byte buffer with N+1 string offsets."""
    n = len(offs) - 1
    lib = _lib.load()
    offs = np.ascontiguousarray(offs, dtype=np.int64)
    lengths = np.zeros(n, dtype=np.int32)
    mags = np.zeros(n, dtype=np.int32)
    _lib.check(lib.siml_pack_measure(buf, _lib.i64p(offs), n, _lib.i32p(lengths), _lib.i32p(mags), nthreads))
    row_off = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(lengths, out=row_off[1:])
    total = int(row_off[-1])
    ids = np.zeros(max(total, 1), dtype=np.int32)
    counts = np.zeros(max(total, 1), dtype=np.int32)
    _lib.check(lib.siml_pack_fill_ragged(buf, _lib.i64p(offs), n, _lib.i64p(row_off), _lib.i32p(ids),
                                         _lib.i32p(counts), nthreads))
    return ids[:total], counts[:total], row_off, mags


_XTABLE = bytes.maketrans(b"123456789", b"000000000")


def preprocessNumbers(smi, xtable=None):
    """compiler.py:135-149: map every digit 1-9 to 0."""
    if isinstance(smi, str):
        return smi.translate(str.maketrans("123456789", "000000000"))
    return bytes(smi).translate(_XTABLE if xtable is None else xtable)


def _windows_multiset(b):
    """Sorted (signed int32) unique 4-byte windows of b with counts."""
    if len(b) < 4:
        return np.zeros(0, np.int32), np.zeros(0, np.int32)
    u = np.frombuffer(b, dtype=np.uint8).astype(np.uint32)
    w = (u[:-3] | (u[1:-2] << 8) | (u[2:-1] << 16) | (u[3:] << 24)).view(np.int32)
    ids, cnt = np.unique(w, return_counts=True)
    return ids.astype(np.int32), cnt.astype(np.int32)


def SMILEStoMultiset(smiles):
    """compiler.py:98-132: (lingos (L,1), counts (L,1)) of the raw string,
    with no preprocessing."""
    ids, cnt = _windows_multiset(_as_bytes(smiles))
    return ids.reshape(-1, 1), cnt.reshape(-1, 1)


def SMILEStoMatrices(smileslist):
    """Pure-Python-variant packer (compiler.py:151-194): first field of each
    line, digits 1-9 -> 0, all windows.  Returns (lingos (N,W), counts (N,W),
    lengths (N,1), mags (N,) = row sums of counts, names)."""
    bags, names = [], []
    for line in smileslist:
        fields = _as_bytes(line).split()
        smiles = preprocessNumbers(fields[0])
        bags.append(_windows_multiset(smiles))
        names.append(fields[1].decode("utf-8", "replace") if len(fields) > 1 else "")
    maxw = max((b[0].shape[0] for b in bags), default=0)
    lingomatrix = np.zeros((len(bags), maxw), dtype=np.int32)
    countmatrix = np.zeros_like(lingomatrix)
    lingolengths = np.empty((len(bags), 1), dtype=np.int32)
    for i, (ids, cnt) in enumerate(bags):
        lingolengths[i] = ids.shape[0]
        lingomatrix[i, :ids.shape[0]] = ids
        countmatrix[i, :ids.shape[0]] = cnt
    lingomags = np.int32(countmatrix.sum(1))
    return lingomatrix, countmatrix, lingolengths, lingomags, names
