"""An independent writer of the on-disk database format (test
infrastructure): the file siml_db_save writes for a set in the NATURAL
layout, produced here from the format's definition in numpy.

Format (all little-endian): a 64-byte header, then 11 arrays (8 when the
level stream is absent), each starting at the next multiple of 64 bytes:

  header  magic "SIMLDB\\0\\1", int32 version 3, int32 layout, int64 n,
          int64 ntiles, int64 tiles_total, int64 btiles_total, int32 maxlen,
          uint8 identity, blocked, levels_ok, sorted_unique, counts_nonneg,
          mags_small, 2 pad bytes (zero), 4 bytes of struct padding (zero)
  tiles   int2 (id, count) merge stream: step j of lane l of tile t at
          toff[t] + 64 j + l, zero padded to the tile's longest molecule
  toff    int64 (ntiles + 1) exclusive scan of 64 * tlen
  tlen    int32 per tile, its longest molecule (unique Lingos)
  len, mag, csum, perm   int32 per storage position
  tmag    int2 per tile (min, max) magnitude
  btiles  int2 (key, slot word) level stream: level j of lane l of tile t at
          btoff[t] + (j // 2) * 128 + 2 l + (j % 2); levels are (x, t) for
          ASCII ids x in id order, t = 1 .. min(count, 15); step counts
          rounded up to 8; one spare block of 8 x 64 zero levels at the end
  btoff   int64 (ntiles + 1), btlen int32 per tile
"""
import struct

import numpy as np

WAVE = 64
MAGIC = b"SIMLDB\x00\x01"
VERSION = 3
MAX_LEVEL = 15
SLOTS = 4096


def level_keys(ids, t):
    """key(x, t) = x | (t&1) << 7 | (t&2) << 14 | (t&4) << 21 | (t&8) << 28."""
    x = np.asarray(ids, np.uint32)
    t = np.asarray(t, np.uint32)
    return (x | ((t & 1) << 7) | ((t & 2) << 14) | ((t & 4) << 21) | ((t & 8) << 28)).astype(np.uint32)


def slot_words(keys):
    """Byte offsets of a key's two candidate slots of the 4096-slot table,
    h1 in the low half and h2 in the high half (a 32-bit mix of the key)."""
    m = np.uint64(0xFFFFFFFF)
    h = (np.asarray(keys, np.uint64) * np.uint64(0x9E3779B1)) & m
    h ^= h >> np.uint64(15)
    h = (h * np.uint64(0x85EBCA77)) & m
    h ^= h >> np.uint64(13)
    h = (h * np.uint64(0xC2B2AE3D)) & m
    h ^= h >> np.uint64(16)
    h1 = h >> np.uint64(20)
    h2 = (h >> np.uint64(4)) & np.uint64(SLOTS - 1)
    h2 = np.where(h2 == h1, h2 ^ np.uint64(1), h2)
    return ((h1 * np.uint64(8)) | ((h2 * np.uint64(8)) << np.uint64(16))).astype(np.uint32)


def ascii_id(x):
    return (np.asarray(x, np.int64) & 0x80808080) == 0


def build_arrays(ids, counts, mags, offsets):
    """Arrays and header fields of a NATURAL-layout database."""
    ids = np.asarray(ids, np.int32)
    counts = np.asarray(counts, np.int32)
    offsets = np.asarray(offsets, np.int64)
    n = len(offsets) - 1
    lens = np.diff(offsets).astype(np.int32)
    csum = np.array([int(counts[offsets[r]:offsets[r + 1]].astype(np.uint32).sum()) & 0xFFFFFFFF
                     for r in range(n)], np.uint32).view(np.int32)
    mags = csum.copy() if mags is None else np.asarray(mags, np.int32).reshape(-1)
    sorted_unique = all(np.all(np.diff(ids[offsets[r]:offsets[r + 1]]) > 0) for r in range(n))
    counts_nonneg = bool(np.all(counts[:offsets[-1]] >= 0))
    mags_small = counts_nonneg and bool(np.all((mags >= 0) & (mags < 65536) & (mags.astype(np.int64) >=
                                                                                csum.astype(np.int64))))
    levels_ok = sorted_unique and counts_nonneg
    ntiles = (n + WAVE - 1) // WAVE
    tlen = np.array([lens[t * WAVE:(t + 1) * WAVE].max(initial=0) for t in range(ntiles)], np.int32)
    toff = np.zeros(ntiles + 1, np.int64)
    np.cumsum(tlen.astype(np.int64) * WAVE, out=toff[1:])
    tiles = np.zeros((int(toff[-1]), 2), np.int32)
    for p in range(n):
        t, lane = divmod(p, WAVE)
        at = toff[t] + lane + WAVE * np.arange(lens[p])
        tiles[at, 0] = ids[offsets[p]:offsets[p + 1]]
        tiles[at, 1] = counts[offsets[p]:offsets[p + 1]]
    tmag = np.array([[mags[t * WAVE:(t + 1) * WAVE].min(), mags[t * WAVE:(t + 1) * WAVE].max()]
                     for t in range(ntiles)], np.int32).reshape(-1, 2)
    arrays = [tiles, toff, tlen, lens, mags, csum, np.arange(n, dtype=np.int32), tmag]
    btiles_total = 0
    if levels_ok:
        per = []
        for p in range(n):
            x = ids[offsets[p]:offsets[p + 1]]
            c = counts[offsets[p]:offsets[p + 1]]
            keep = ascii_id(x) & (c > 0)
            xs = np.repeat(x[keep], np.minimum(c[keep], MAX_LEVEL))
            ts = np.concatenate([np.arange(1, min(k, MAX_LEVEL) + 1) for k in c[keep]]) if keep.any() \
                else np.zeros(0, np.int64)
            per.append(level_keys(xs, ts))
        btlen = np.array([(max((len(per[p]) for p in range(t * WAVE, min(n, (t + 1) * WAVE))), default=0) + 7)
                          // 8 * 8 for t in range(ntiles)], np.int32)
        btoff = np.zeros(ntiles + 1, np.int64)
        np.cumsum(btlen.astype(np.int64) * WAVE, out=btoff[1:])
        bt = np.zeros((int(btoff[-1]) + 8 * WAVE, 2), np.uint32)
        for p in range(n):
            t, lane = divmod(p, WAVE)
            j = np.arange(len(per[p]))
            at = btoff[t] + (j // 2) * 2 * WAVE + 2 * lane + (j % 2)
            bt[at, 0] = per[p]
            bt[at, 1] = slot_words(per[p])
        arrays += [bt.view(np.int32), btoff, btlen]
        btiles_total = len(bt)
    hdr = dict(layout=0, n=n, ntiles=ntiles, tiles_total=int(toff[-1]), btiles_total=btiles_total,
               maxlen=int(lens.max(initial=0)), identity=1, blocked=1, levels_ok=int(levels_ok),
               sorted_unique=int(sorted_unique), counts_nonneg=int(counts_nonneg), mags_small=int(mags_small))
    return hdr, arrays


def header_bytes(h, version=VERSION, magic=MAGIC):
    b = magic + struct.pack("<iiqqqqi6B2x", version, h["layout"], h["n"], h["ntiles"], h["tiles_total"],
                            h["btiles_total"], h["maxlen"], h["identity"], h["blocked"], h["levels_ok"],
                            h["sorted_unique"], h["counts_nonneg"], h["mags_small"])
    return b + b"\x00" * (64 - len(b))


def file_bytes(hdr, arrays, **kw):
    out = bytearray(header_bytes(hdr, **kw))
    for a in arrays:
        out += b"\x00" * (-len(out) % 64)
        out += np.ascontiguousarray(a).tobytes()
    return bytes(out)


def write(path, ids, counts, mags, offsets):
    hdr, arrays = build_arrays(ids, counts, mags, offsets)
    with open(path, "wb") as f:
        f.write(file_bytes(hdr, arrays))
    return hdr, arrays
