"""Writes tests/golden/pcg_vectors.json: known-answer vectors for the PCG RNG.

The values are computed here with Python's arbitrary-precision integers masked
to 32 bits, an arithmetic route independent of the C/HIP uint32 code, from the
definitions in src/main/resources/shaders/pathtrace.comp:80-98 (pcg, rand,
initSeed) and photon_trace.comp:68-72 (photon seeds; added with the SPPM path).
rand() = float32(u) / 2^32 is computed with numpy float32 rounding (RNE).

Run: python tests/golden/gen_pcg_vectors.py
"""
import json
import os

import numpy as np

M32 = 0xFFFFFFFF


def pcg(state):
    state = (state * 747796405 + 2891336453) & M32
    w = (((state >> ((state >> 28) + 4)) ^ state) * 277803737) & M32
    return state, ((w >> 22) ^ w) & M32


def init_seed(x, y, frame):
    s = (x + y * 7919 + frame * 6271) & M32
    s, _ = pcg(s)
    s, _ = pcg(s)
    return s


def main():
    cases = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (319, 179, 15), (853, 479, 127), (5, 7, 99991),
             (853, 479, 4095 * 3 + 2), (12, 34, 0x7FFFFFFF // 6271), (100, 200, 3)]
    out = []
    for (x, y, f) in cases:
        s = init_seed(x, y, f)
        seq = []
        st = s
        for _ in range(12):
            st, v = pcg(st)
            seq.append(v)
        rnd = [float(np.float32(v) / np.float32(4294967296.0)) for v in seq]
        out.append({"x": x, "y": y, "frame": f, "seed": s, "pcg": seq, "rand": rnd})
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pcg_vectors.json")
    with open(path, "w") as fh:
        json.dump({"source": "pathtrace.comp:80-98", "cases": out}, fh, indent=1)
    print(path)


if __name__ == "__main__":
    main()
