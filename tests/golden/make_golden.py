"""Writes the golden fixtures of tests/golden/ (run from the repo root:
``python tests/golden/make_golden.py``).

The reference ships no expected outputs, so the vectors are produced by the
CPU oracle and, before being written, cross-checked entry by entry against
the independent brute force of oracle/brute.py (Counter bags, exact
rationals).  Inputs: the first 96 lines of examples/demo4096.smi (kept here
as demo4096.smi) and 160 seeded synthetic strings.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from oracle import brute  # noqa: E402

ALPHABET = np.frombuffer(b"CNOSPFcnos()=#[]@+-123456", dtype=np.uint8)


def inputs():
    with open(os.path.join(HERE, "demo4096.smi"), "rb") as f:
        demo = f.readlines()[:96]
    rng = np.random.default_rng(20240601)
    synth = [ALPHABET[rng.integers(0, ALPHABET.size, rng.integers(4, 60))].tobytes() for _ in range(160)]
    return demo + synth


def cross_check(strings, packed, M):
    l, c, n, m = packed
    for r, s in enumerate(strings):
        ids, cnt, mag = brute.bag(oracle.normalize(s))
        assert l[r, :n[r, 0]].tolist() == ids and c[r, :n[r, 0]].tolist() == cnt and m[r, 0] == mag
    bags = [brute.bag(oracle.normalize(s)) for s in strings]
    for i in range(len(strings)):
        for j in range(len(strings)):
            t = brute.tanimoto(bags[i][0], bags[i][1], bags[i][2], bags[j][0], bags[j][1], bags[j][2])
            assert M[i, j].view(np.uint32) == t.view(np.uint32), (i, j)


def main():
    strings = inputs()
    packed = oracle.pack(strings)
    M = oracle.tanimoto_block(packed, packed)
    cross_check(strings, packed, M)
    lists, counts = oracle.neighbors(packed, packed, 0.3)
    offs = np.zeros(len(strings) + 1, np.int64)
    offs[1:] = np.cumsum(counts)
    ti, tt = oracle.topk(packed, packed, 8)
    np.savez_compressed(
        os.path.join(HERE, "golden_small.npz"),
        strings=np.array(strings, dtype=object).astype("S"),
        lingos=packed[0], counts=packed[1], lengths=packed[2], mags=packed[3],
        tanimoto=M,
        hist101=oracle.hist_rows101(packed, packed),
        hist1000=oracle.hist1000(packed, packed),
        nbr_offsets=offs, nbr_indices=np.concatenate(lists).astype(np.int32),
        topk_idx=ti, topk_t=tt,
        single_row_7=oracle.single_row(packed, 7, packed),
    )
    print("wrote golden_small.npz:", len(strings), "strings")


if __name__ == "__main__":
    main()
