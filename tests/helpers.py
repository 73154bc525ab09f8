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


def mod_q_entries(X, q, chunk=65536):
    """Every entry of an entry matrix mod q (signed entries mapped to [0, q)),
    for q < 2^26 and at most 2048 limbs per entry: limb i contributes
    (limb mod q) (2^(64 i) mod q) < 2^52, so a row of them sums below 2^63."""
    cap = X.shape[-1] - 1
    flat = X.reshape(-1, cap + 1)
    sz = flat[:, cap].view(np.int64)
    n = np.abs(sz)
    L = int(n.max()) if n.size else 0
    out = np.zeros(flat.shape[0], dtype=np.uint64)
    if L:
        assert L <= 2048 and q < (1 << 26)
        pw = np.array([pow(2, 64 * i, q) for i in range(L)], dtype=np.uint64)
        qq = np.uint64(q)
        for c0 in range(0, flat.shape[0], chunk):
            lim = flat[c0:c0 + chunk, :L] % qq
            lim[np.arange(L)[None, :] >= n[c0:c0 + chunk, None]] = 0
            out[c0:c0 + chunk] = (lim * pw[None, :]).sum(axis=1) % qq
    out = np.where(sz < 0, (np.uint64(q) - out) % np.uint64(q), out)
    return out.astype(np.int64).reshape(X.shape[:-1])


def matmul_mod(a, b, q):
    """(a @ b) mod q for int64 matrices with entries in [0, q), q < 2^26."""
    out = np.zeros((a.shape[0], b.shape[1]), dtype=np.int64)
    for k0 in range(0, a.shape[1], 256):   # partial sums stay below 2^63
        out = (out + a[:, k0:k0 + 256] @ b[k0:k0 + 256, :]) % q
    return out
