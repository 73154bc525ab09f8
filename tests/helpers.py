"""Shared test helpers: exact-integer views of entry matrices and fixtures."""
import json
import os
import random

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def parse_hex(s):
    return -int(s[1:], 16) if s.startswith("-") else int(s, 16)


def load_json(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def limbs_needed(v):
    return (abs(v).bit_length() + 63) // 64


def to_entries(vals, cap):
    rows, cols = len(vals), len(vals[0]) if vals else 0
    m = np.zeros((rows, cols, cap + 1), dtype=np.uint64)
    for y in range(rows):
        for x in range(cols):
            v = vals[y][x]
            mag, n = abs(v), 0
            while mag:
                m[y, x, n] = mag & 0xFFFFFFFFFFFFFFFF
                mag >>= 64
                n += 1
            m[y, x, cap] = np.uint64((-n) & 0xFFFFFFFFFFFFFFFF) if v < 0 else np.uint64(n)
    return m


def entry_int(e):
    cap = e.shape[-1] - 1
    s = int(np.int64(e[cap]))
    v = 0
    for i in range(abs(s) - 1, -1, -1):
        v = (v << 64) | int(e[i])
    return -v if s < 0 else v


def to_ints(m):
    return [[entry_int(m[y, x]) for x in range(m.shape[1])] for y in range(m.shape[0])]


def is_normalized(m):
    cap = m.shape[-1] - 1
    flat = m.reshape(-1, cap + 1)
    sz = np.abs(flat[:, cap].view(np.int64))
    for i in range(flat.shape[0]):
        if sz[i] > cap or (sz[i] and flat[i, sz[i] - 1] == 0):
            return False
    return True


def rand_ints(rows, cols, bits, seed, mode="rand"):
    rng = random.Random(seed)
    out = []
    for _ in range(rows):
        r = []
        for _ in range(cols):
            if mode == "ones":
                v = (1 << bits) - 1
            else:
                v = rng.getrandbits(bits) if bits else 0
            r.append(-v if rng.getrandbits(1) else v)
        out.append(r)
    return out


def matmul(A, B):
    K, M, N = len(A), len(B), len(B[0]) if B else 0
    return [[sum(A[y][k] * B[k][x] for k in range(M)) for x in range(N)] for y in range(K)]
