"""Independent pure-Python computations used to pin the C oracle on small
inputs (TEST INFRASTRUCTURE ONLY).

Nothing here restates the C merge: the Lingo bag is a collections.Counter
over all 4-byte windows, the intersection is Counter '&', and the Tanimoto is
an exact rational rounded once to float32.
"""
from collections import Counter
from fractions import Fraction

import numpy as np


def windows(norm: bytes):
    """Every overlapping 4-byte window as a little-endian signed int32."""
    return [int(np.frombuffer(norm[k:k + 4], dtype="<i4")[0]) for k in range(len(norm) - 3)]


def bag(norm: bytes):
    """(sorted ids, counts, magnitude) of a normalised string via Counter."""
    c = Counter(windows(norm))
    ids = sorted(c)
    return ids, [c[i] for i in ids], max(len(norm) - 3, 0)


def f32_of_fraction(fr: Fraction) -> np.float32:
    """Round an exact rational to float32 (one correct rounding).

    float(Fraction) is correctly rounded to binary64; the second rounding to
    binary32 is innocuous for a quotient of integers below 2**24 because
    53 >= 2*24 + 2."""
    return np.float32(float(fr))


def tanimoto(ids_a, cnt_a, mag_a, ids_b, cnt_b, mag_b) -> np.float32:
    """I/(ma+mb-I) with I = sum of min counts, from Counters."""
    if mag_a == 0 or mag_b == 0:
        return np.float32(0.0)
    ca = Counter(dict(zip(ids_a, cnt_a)))
    cb = Counter(dict(zip(ids_b, cnt_b)))
    inter = sum((ca & cb).values())
    return f32_of_fraction(Fraction(inter, mag_a + mag_b - inter))


def digits_only(s: bytes) -> bytes:
    """The pure-Python packer's preprocessing (compiler.py:148,164,171-173):
    first whitespace-separated field with every 1-9 mapped to 0."""
    f = s.split()
    first = f[0] if f else b""
    return first.translate(bytes.maketrans(b"123456789", b"000000000"))
