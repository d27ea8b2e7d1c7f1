"""Seeded synthetic SMILES-like workloads.

Strings are drawn i.i.d. over the alphabet 'CNOSPFcnos()=#[]@+-123456' with
either uniform lengths U[lo, hi] or lognormal lengths (median, sigma) clipped
to [lo, hi].  They are packed with the native packer, so ring digits are
normalised exactly as cSMILEStoMatrices does.

Workloads (sizes from the benchmark definitions of this project):
  all_pairs_4096   4096 strings, U[20, 80], seed 1
  all_pairs_32768  32768 strings, U[20, 120], seed 2 (rows == columns)
  db_lognormal     lognormal median 45, sigma 0.35, clipped [8, 256]
"""
import numpy as np

ALPHABET = np.frombuffer(b"CNOSPFcnos()=#[]@+-123456", dtype=np.uint8)


def random_strings(n, lo, hi, seed, dist="uniform", median=45.0, sigma=0.35):
    """(buffer bytes, offsets int64 (n+1,)) of n random strings."""
    rng = np.random.default_rng(seed)
    if dist == "uniform":
        lens = rng.integers(lo, hi + 1, size=n, dtype=np.int64)
    elif dist == "lognormal":
        lens = np.rint(median * np.exp(sigma * rng.standard_normal(n))).astype(np.int64)
        lens = np.clip(lens, lo, hi)
    else:
        raise ValueError(f"unknown length distribution {dist!r}")
    offs = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(lens, out=offs[1:])
    chars = ALPHABET[rng.integers(0, ALPHABET.size, size=int(offs[-1]))]
    return chars.tobytes(), offs


def random_smiles(n, lo, hi, seed, dist="uniform", **kw):
    """The same strings as a list of str."""
    buf, offs = random_strings(n, lo, hi, seed, dist, **kw)
    return [buf[offs[i]:offs[i + 1]].decode("ascii") for i in range(n)]


def packed_padded(n, lo, hi, seed, dist="uniform", **kw):
    """cSMILEStoMatrices-layout set (lingos, counts, lengths (n,1), mags (n,1))."""
    from .compiler import cSMILEStoMatrices
    lingos, counts, lengths, mags, _ = cSMILEStoMatrices(random_smiles(n, lo, hi, seed, dist, **kw))
    return lingos, counts, lengths, mags


def packed_ragged(n, lo, hi, seed, dist="uniform", nthreads=0, **kw):
    """Ragged set (ids, counts, offsets (n+1,), mags (n,))."""
    from .compiler import pack_ragged_buffer
    buf, offs = random_strings(n, lo, hi, seed, dist, **kw)
    return pack_ragged_buffer(buf, offs, nthreads)


def ragged_to_padded(ids, counts, offsets, mags):
    """Padded (lingos, counts, lengths (n,1), mags (n,1)) from a ragged set."""
    n = len(offsets) - 1
    lens = np.diff(offsets).astype(np.int32)
    w = int(lens.max()) if n else 0
    L = np.zeros((n, w), np.int32)
    C = np.zeros((n, w), np.int32)
    if n and w:
        row = np.repeat(np.arange(n), lens)
        col = np.arange(int(offsets[-1])) - np.repeat(offsets[:-1], lens)
        L[row, col] = ids
        C[row, col] = counts
    return L, C, lens.reshape(-1, 1), np.asarray(mags, np.int32).reshape(-1, 1)


def ragged_slice(ids, counts, offsets, mags, a, b):
    """Molecules [a, b) of a ragged set as a new ragged set."""
    o0, o1 = int(offsets[a]), int(offsets[b])
    return ids[o0:o1], counts[o0:o1], offsets[a:b + 1] - o0, mags[a:b]


def ragged_concat(parts):
    """One ragged set holding the molecules of each (ids, counts, offsets,
    mags) part in order."""
    ids = np.concatenate([p[0] for p in parts]).astype(np.int32, copy=False)
    counts = np.concatenate([p[1] for p in parts]).astype(np.int32, copy=False)
    lens = np.concatenate([np.diff(p[2]) for p in parts])
    offs = np.zeros(len(lens) + 1, np.int64)
    np.cumsum(lens, out=offs[1:])
    mags = np.concatenate([p[3] for p in parts]).astype(np.int32, copy=False)
    return ids, counts, offs, mags


def geometric_sizes(n, p, lo, hi, seed):
    """n set sizes drawn from a geometric law with success probability p,
    clipped to [lo, hi]."""
    rng = np.random.default_rng(seed)
    return np.clip(rng.geometric(p, size=n), lo, hi).astype(np.int64)


def mutated_strings(buf, offs, picks, edits, seed):
    """(buffer, offsets) of copies of strings `picks` of (buf, offs), each
    with `edits` random characters replaced over the alphabet: near
    duplicates of database molecules."""
    rng = np.random.default_rng(seed)
    src = np.frombuffer(buf, dtype=np.uint8)
    out, lens = [], []
    for i in picks:
        s = src[offs[i]:offs[i + 1]].copy()
        if s.size:
            pos = rng.integers(0, s.size, size=edits)
            s[pos] = ALPHABET[rng.integers(0, ALPHABET.size, size=edits)]
        out.append(s)
        lens.append(s.size)
    o = np.zeros(len(lens) + 1, np.int64)
    np.cumsum(lens, out=o[1:])
    return (np.concatenate(out).tobytes() if out else b""), o


def padded_columns(ids, counts, offsets, mags, a, b):
    """cSMILEStoMatrices-layout (lingos, counts, lengths, mags) of molecules
    [a, b) of a ragged set."""
    return ragged_to_padded(*ragged_slice(ids, counts, offsets, mags, a, b))
