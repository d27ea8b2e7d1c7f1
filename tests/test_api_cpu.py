"""Host-side behaviour of the drop-in API that needs no GPU: argument
validation and error classes (_CPULingo.c:73-80,139-215; CPULingo.py:103-108;
_cpuutil.cpp:90-112)."""
import numpy as np
import pytest

from siml_amd.CPULingo import CPULingo
from siml_amd.GPULingo import normalise_set
from siml_amd.utilities import rowHistogram101


def _set(n=3, w=4):
    return (np.zeros((n, w), np.int32), np.zeros((n, w), np.int32), np.zeros((n, 1), np.int32),
            np.zeros((n, 1), np.int32))


def test_cpulingo_not_initialised():
    c = CPULingo()
    assert c.supportsParallel() and c.asyncOperationsDone()
    with pytest.raises(ValueError, match="not fully initialized"):
        c.getMultipleRows(0, 1)


def test_cpulingo_default_mags_are_lengths_sum():
    """CPULingo.py:82,98: without mags the magnitude is int32(lengths.sum(1))."""
    c = CPULingo()
    L, C, n, _ = _set()
    n[:] = [[2], [0], [5]]
    c.set_refsmiles(L, C, n)
    c.set_qsmiles(L, C, n)
    assert c.rmags.tolist() == [2, 0, 5] and c.qmags.tolist() == [2, 0, 5]


@pytest.mark.parametrize("bad,msg", [
    (lambda s: (s[0].astype(np.int64),) + s[1:], "was not of type int32"),
    (lambda s: (s[0][:, 0],) + s[1:], "did not have dimension 2"),
    (lambda s: (s[0], s[1][:, :3]) + s[2:], "identical shapes"),
    (lambda s: s[:2] + (s[2][:2],) + s[3:], "length vector length"),
    (lambda s: s[:2] + (np.zeros((3, 2), np.int32),) + s[3:], "was not 1-D"),
    (lambda s: s[:3] + (s[3].astype(np.float32),), "was not of type int32"),
])
def test_set_validation(bad, msg):
    with pytest.raises(ValueError, match=msg):
        normalise_set(*bad(_set()), "Reference")


def test_row_histogram_argument_errors():
    with pytest.raises(ValueError, match="float32"):
        rowHistogram101(np.zeros((2, 2), np.int32))
    with pytest.raises(ValueError, match="dimension 2"):
        rowHistogram101(np.zeros(4, np.float32))


def test_db_load_rejects_bad_files(tmp_path):
    """siml_db_load validates the file before touching a device."""
    from siml_amd.GPULingo import LingoDB
    missing = tmp_path / "missing.db"
    with pytest.raises(ValueError, match="cannot open"):
        LingoDB.load(str(missing))
    bad = tmp_path / "bad.db"
    bad.write_bytes(b"not a database file at all" * 4)
    with pytest.raises(ValueError, match="not a database file"):
        LingoDB.load(str(bad))
    old = tmp_path / "old.db"
    hdr = bytearray(b"SIMLDB\x00\x01" + b"\x00" * 64)
    hdr[8:12] = (1).to_bytes(4, "little")  # an older format version
    old.write_bytes(bytes(hdr))
    with pytest.raises(ValueError, match="format version"):
        LingoDB.load(str(old))
