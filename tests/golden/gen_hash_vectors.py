"""Writes tests/golden/hash_vectors.json: known answers for the integer hashes of
the VHS passes, computed with Python integers masked to 32 bits (independent of
the C/HIP uint32 code), from the definitions in
src/main/resources/shaders/post/p07_ccd_noise.frag:50-54 (pcgHash),
p12_dropout.frag:32-41 (one-round uHash, fHash) and p13_headswitch.frag:40-49
(two-round uHash, fHash). fHash's float is numpy float32(u) / float32(4294967295.0)
(the literal rounds to 2^32).

Run: python tests/golden/gen_hash_vectors.py
"""
import json
import os

import numpy as np

M = 0xFFFFFFFF


def pcg_hash(v):
    v = (v * 747796405 + 2891336453) & M
    w = (((v >> ((v >> 28) + 4)) ^ v) * 277803737) & M
    return ((w >> 22) ^ w) & M


def uhash1(x):
    x ^= x >> 16
    x = (x * 0x45D9F3B) & M
    x ^= x >> 16
    return x


def uhash2(x):
    x = (((x >> 16) ^ x) * 0x45D9F3B) & M
    x = (((x >> 16) ^ x) * 0x45D9F3B) & M
    return ((x >> 16) ^ x) & M


def fbits(u):
    f = np.float32(u) / np.float32(4294967295.0)
    return int(np.array([f], np.float32).view(np.uint32)[0])


def main():
    rng = np.random.default_rng(11)
    vals = [0, 1, 2, 0xDEADBEEF, 0xCAFEBABE, 0xBEEF, 0xFFFFFFFF, 123456789] + [int(v) for v in rng.integers(0, 2**32, 24)]
    pairs = [(vals[i], vals[(i * 7 + 3) % len(vals)]) for i in range(len(vals))]
    out = {
        "pcg": [[a, pcg_hash(a)] for a in vals],
        "uhash1": [[a, uhash1(a)] for a in vals],
        "uhash2": [[a, uhash2(a)] for a in vals],
        "fhash1": [[a, b, fbits(uhash1(a ^ ((uhash1(b) * 2654435761) & M)))] for a, b in pairs],
        "fhash2": [[a, b, fbits(uhash2(a ^ ((uhash2(b) * 2654435761) & M)))] for a, b in pairs],
    }
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hash_vectors.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=0)
    print(path)


if __name__ == "__main__":
    main()
