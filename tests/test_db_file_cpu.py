"""The database file loader rejects inconsistent files on the host, before
anything reaches a device (no GPU needed): every offset, length, index and
flag the kernels rely on is checked against the arrays.  The files are
written from the format's definition (tests/db_file_format.py) and then
corrupted one field at a time."""
import ctypes
import struct

import numpy as np
import pytest

from siml_amd import _lib, compiler, synth

import db_file_format as fmt


@pytest.fixture(scope="module")
def small_set():
    buf, offs = synth.random_strings(150, 8, 120, seed=81, dist="lognormal", median=45.0, sigma=0.35)
    return compiler.pack_ragged_buffer(buf, offs)   # ids, counts, offsets, mags


def load_rc(path):
    lib = _lib.load()
    db = ctypes.c_void_p()
    rc = lib.siml_db_load(0, str(path).encode(), ctypes.byref(db))
    msg = lib.siml_last_error().decode()
    if rc == _lib.SIML_OK:
        lib.siml_db_destroy(db)
    return rc, msg


def devices():
    n = ctypes.c_int()
    _lib.load().siml_device_count(ctypes.byref(n))
    return n.value


def test_valid_file_passes_validation(small_set, tmp_path):
    ids, counts, offs, mags = small_set
    path = tmp_path / "ok.siml"
    hdr, _ = fmt.write(path, ids, counts, mags, offs)
    assert hdr["levels_ok"] and hdr["mags_small"] and hdr["n"] == 150
    rc, msg = load_rc(path)
    # without a device the load stops after validation, at the device check
    assert rc == (_lib.SIML_OK if devices() else _lib.SIML_EDEVICE), msg


def _corrupt(small_set, tmp_path, edit, name):
    ids, counts, offs, mags = small_set
    hdr, arrays = fmt.build_arrays(ids, counts, mags, offs)
    arrays = [np.array(a, copy=True) for a in arrays]
    kw = edit(hdr, arrays)
    kw = kw if isinstance(kw, dict) else {}
    path = tmp_path / f"{name}.siml"
    path.write_bytes(fmt.file_bytes(hdr, arrays, **kw))
    return load_rc(path)


CASES = {
    "magic": (lambda h, a: {"magic": b"SIMLDX\x00\x01"}, "not a database file"),
    "version": (lambda h, a: {"version": 2}, "format version"),
    "ntiles": (lambda h, a: h.update(ntiles=h["ntiles"] + 1), "inconsistent header"),
    "flag_value": (lambda h, a: h.update(levels_ok=2), "inconsistent header"),
    "levels_without_sorted": (lambda h, a: h.update(sorted_unique=0), "inconsistent header"),
    "stream_size": (lambda h, a: h.update(tiles_total=h["tiles_total"] - 64), "offsets"),
    "stream_past_end": (lambda h, a: h.update(tiles_total=h["tiles_total"] + 64), "truncated"),
    "tile_offset": (lambda h, a: a[1].__setitem__(1, a[1][1] + 64), "tile 0 offsets"),
    "length_past_tile": (lambda h, a: a[3].__setitem__(5, a[2][0] + 1), "longer than its tile"),
    "maxlen": (lambda h, a: h.update(maxlen=h["maxlen"] + 1), "maximum length"),
    "perm_duplicate": (lambda h, a: (h.update(identity=0), a[6].__setitem__(3, 4)), "not a permutation"),
    "perm_range": (lambda h, a: a[6].__setitem__(3, 150), "not a permutation"),
    "perm_identity_flag": (lambda h, a: (a[6].__setitem__(3, 4), a[6].__setitem__(4, 3)), "layout flags"),
    "mags_small_flag": (lambda h, a: (a[4].__setitem__(7, a[5][7] - 1)), "small-magnitude flag"),
    "tile_mag_range": (lambda h, a: a[7].__setitem__((0, 1), a[7][0, 1] + 1), "magnitude range"),
    "csum": (lambda h, a: a[5].__setitem__(9, a[5][9] - 1), "count sum"),
    "unsorted": (lambda h, a: a[0].__setitem__((int(a[1][0]) + 64 + 2, 0), 0), "flags or count sum"),
    "negative_count": (lambda h, a: a[0].__setitem__((int(a[1][0]) + 2, 1), -1), "flags or count sum"),
    "level_length": (lambda h, a: a[10].__setitem__(0, a[10][0] + 1), "level offsets"),
    "level_stream_size": (lambda h, a: h.update(btiles_total=h["btiles_total"] - 64), "level offsets"),
    "slot_word": (lambda h, a: a[8].__setitem__((0, 1), a[8][0, 1] ^ 0x8000), "slot word"),
    "empty_level_word": (lambda h, a: a[8].__setitem__((len(a[8]) - 1, 1), 8), "slot word"),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_corrupt_file_rejected(small_set, tmp_path, name):
    edit, words = CASES[name]
    rc, msg = _corrupt(small_set, tmp_path, edit, name)
    assert rc == _lib.SIML_EINVAL and words in msg, (rc, msg)


def test_truncated_file_rejected(small_set, tmp_path):
    ids, counts, offs, mags = small_set
    hdr, arrays = fmt.build_arrays(ids, counts, mags, offs)
    data = fmt.file_bytes(hdr, arrays)
    for cut in (10, 64, len(data) // 2, len(data) - 1):
        path = tmp_path / f"cut{cut}.siml"
        path.write_bytes(data[:cut])
        rc, msg = load_rc(path)
        assert rc == _lib.SIML_EINVAL, (cut, msg)


def test_huge_header_counts_rejected_without_allocating(tmp_path):
    """Counts in the header far beyond the file's size are refused before
    any host array is sized by them."""
    hdr = dict(layout=0, n=2 ** 31 - 1, ntiles=(2 ** 31 - 1 + 63) // 64, tiles_total=2 ** 40, btiles_total=0,
               maxlen=5, identity=1, blocked=1, levels_ok=0, sorted_unique=1, counts_nonneg=1, mags_small=0)
    path = tmp_path / "huge.siml"
    path.write_bytes(fmt.header_bytes(hdr) + b"\x00" * 4096)
    rc, msg = load_rc(path)
    assert rc == _lib.SIML_EINVAL and "truncated" in msg


def test_format_writer_levels_match_probe_rule():
    """The writer's level keys and slot words follow the probe rule
    (pinned independently of the library in tests/test_bitgroups_cpu.py):
    distinct slots, 8-byte aligned offsets inside the 32 KiB table."""
    keys = fmt.level_keys(np.array([0x43434343, 0x2843434F]), np.array([1, 15]))
    assert keys[0] == 0x43434343 | 0x80 and keys[1] == 0x2843434F | 0x80 | 0x8000 | 0x800000 | 0x80000000
    w = fmt.slot_words(np.arange(1, 100000, dtype=np.uint32))
    lo, hi = w & 0xFFFF, w >> 16
    assert np.all(lo % 8 == 0) and np.all(hi % 8 == 0) and np.all(lo != hi)
    assert np.all(lo < 4096 * 8) and np.all(hi < 4096 * 8)
    assert struct.calcsize("<iiqqqqi6B2x") + 8 == 60
