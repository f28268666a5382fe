"""Writes tests/golden/murmur64a.txt: MurmurHash64A vectors computed by the reference's own
MurmurHash2_64.cu, compiled from the reference checkout by oracle/build_ref.sh.

    sh oracle/build_ref.sh && python tests/golden/make_murmur_vectors.py
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
lib = C.CDLL(os.path.join(ROOT, "oracle", "_ref", "libmurmur_ref.so"))
f = getattr(lib, "_Z13MurmurHash64APKvij")
f.restype = C.c_uint64
f.argtypes = [C.c_char_p, C.c_int, C.c_uint32]

r = np.random.default_rng(20240601)
lines = ["# hexkey seed MurmurHash64A (reference MurmurHash2_64.cu:63-104)"]
keys = [b"", b"a", b"BUILDING\0\0", b"AUTOMOBILE", bytes(8), bytes(range(16))]
keys += [r.integers(0, 256, n, dtype=np.uint8).tobytes() for n in range(1, 41)]
keys += [np.array([k, d, 0], np.int64).tobytes() for k, d in ((1, 19950101), (33, 19921231), (-5, 7))]
for i, k in enumerate(keys):
    for seed in (0, 100, int(r.integers(0, 1 << 32))) if i % 3 == 0 else (100,):
        lines.append(f"{k.hex() or '-'} {seed} {f(k, len(k), seed):016x}")
with open(os.path.join(HERE, "murmur64a.txt"), "w") as out:
    out.write("\n".join(lines) + "\n")
print(len(lines) - 1, "vectors")
