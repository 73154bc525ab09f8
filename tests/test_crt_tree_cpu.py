"""The kept big-modulus CRT API (mat-gpu/gpu_crt.h:6-14, crt.cpp:16-164) on
the library's own arithmetic, checked against Python integers and the
invariants crt.cpp documents.  Host-only functions: no GPU."""
import ctypes
import math
import random

import numpy as np
import pytest

from helpers import entry_int, to_entries

CAP = 40
W = CAP + 1


def ent(v):
    return to_entries([[v]], CAP)[0, 0].copy()


@pytest.fixture(scope="module")
def L():
    import mgx
    lib = mgx.lib()
    vp, i64 = ctypes.c_void_p, ctypes.c_int64
    lib.mgx_crt_split.argtypes = [vp, vp, vp, vp, vp, vp, i64, i64]
    lib.mgx_crt_merge.argtypes = [vp] * 7 + [i64]
    lib.mgx_crt_make_split_tree.argtypes = [ctypes.POINTER(vp)] * 4 + [vp, ctypes.c_uint64, i64]
    lib.mgx_crt_split_list.argtypes = [vp, vp, vp, vp, ctypes.c_int, i64]
    lib.mgx_crt_merge_list.argtypes = [vp, vp, vp, vp, vp, vp, ctypes.c_int, i64]
    return lib


def gmp_gcdext_cofactors(a, b):
    """s, t with g = s a + t b and |s| < b/(2g), |t| < a/(2g) (mpz_gcdext's normalisation)."""
    g = math.gcd(a, b)
    s = pow(a // g, -1, b // g) if b // g > 1 else 0
    if 2 * s > b // g:
        s -= b // g
    t = (g - s * a) // b
    return g, s, t


def ref_split(x, lst):
    """crt.cpp:16-63 restated on Python integers."""
    m = math.isqrt(x)
    while any(math.gcd(m, l) != 1 for l in lst):
        m += 1
    n = m + 3
    while True:
        g, mf, nf = gmp_gcdext_cofactors(m, n)
        if g == 1 and all(math.gcd(n, l) == 1 for l in lst):
            return m, n, mf, nf
        n += 2 if m % 2 == 0 else 1


def test_split_and_merge(L):
    rng = random.Random(4)
    for bits in [8, 64, 65, 200, 1000]:
        for _ in range(6):
            x = rng.getrandbits(bits) | 1
            lst = [rng.getrandbits(20) | 1 for _ in range(3)]
            le = [ent(v) for v in lst]
            ptrs = (ctypes.c_void_p * len(le))(*[e.ctypes.data for e in le])
            m, n, mf, nf = ent(0), ent(0), ent(0), ent(0)
            assert L.mgx_crt_split(m.ctypes.data, n.ctypes.data, mf.ctypes.data, nf.ctypes.data, ent(x).ctypes.data,
                                   ptrs, len(lst), CAP) == 0
            got = tuple(entry_int(e) for e in (m, n, mf, nf))
            assert got == ref_split(x, lst)
            M, N, MF, NF = got
            assert M * N > x and math.gcd(M, N) == 1 and MF * M + NF * N == 1 and M < N
            # merge(split(y)) == y for y < m n
            y = rng.randrange(M * N)
            out = ent(0)
            assert L.mgx_crt_merge(out.ctypes.data, ent(y % M).ctypes.data, ent(y % N).ctypes.data, m.ctypes.data,
                                   n.ctypes.data, mf.ctypes.data, nf.ctypes.data, CAP) == 0
            assert entry_int(out) == y


def test_split_tree_roundtrip(L):
    rng = random.Random(9)
    for bits, limit in [(300, 1 << 40), (1000, 1 << 62), (64, 1000)]:
        x = rng.getrandbits(bits) | (1 << (bits - 1))
        ms, ns, mfs, nfs = (ctypes.c_void_p() for _ in range(4))
        length = L.mgx_crt_make_split_tree(ctypes.byref(ms), ctypes.byref(ns), ctypes.byref(mfs), ctypes.byref(nfs),
                                           ent(x).ctypes.data, limit, CAP)
        assert length > 0
        arr = lambda p: np.ctypeslib.as_array((ctypes.c_uint64 * (length * W)).from_address(p.value)).reshape(length, W)
        M, N, MF, NF = (arr(p) for p in (ms, ns, mfs, nfs))
        # restated tree: split, then continue with x mod n (crt.cpp:108-117)
        xr, lst = x, []
        for i in range(length):
            assert xr >= limit
            want = ref_split(xr, lst)
            assert (entry_int(M[i]), entry_int(N[i]), entry_int(MF[i]), entry_int(NF[i])) == want
            lst += [want[0], want[1]]
            xr %= want[1]
        assert xr < limit
        # split_list / merge_list against crt.cpp:126-164 restated; the tree of
        # x itself merges back to x only level by level, so check the restatement
        mv = [(entry_int(M[i]), entry_int(N[i]), entry_int(MF[i]), entry_int(NF[i])) for i in range(length)]

        def split_list(v, lvl):
            if lvl == length:
                return [v]
            return split_list(v % mv[lvl][0], lvl + 1) + split_list(v % mv[lvl][1], lvl + 1)

        def merge_list(sp, lvl):
            if lvl == length:
                return sp[0]
            h = len(sp) // 2
            a, b = merge_list(sp[:h], lvl + 1), merge_list(sp[h:], lvl + 1)
            m, n, mf, nf = mv[lvl]
            return (a * nf * n + b * mf * m) % (m * n)

        for y in [x, rng.randrange(mv[0][0] * mv[0][1]), rng.randrange(limit)]:
            leaves = np.zeros((1 << length, W), dtype=np.uint64)
            assert L.mgx_crt_split_list(leaves.ctypes.data, ent(y).ctypes.data, ms.value, ns.value, length, CAP) == 0
            assert [entry_int(leaves[i]) for i in range(1 << length)] == split_list(y, 0)
            back = ent(0)
            assert L.mgx_crt_merge_list(back.ctypes.data, leaves.ctypes.data, ms.value, ns.value, mfs.value,
                                        nfs.value, length, CAP) == 0
            assert entry_int(back) == merge_list(split_list(y, 0), 0)
            if y < limit:
                assert entry_int(back) == y   # every level reduces to the identity
        libc = ctypes.CDLL(None)
        for p in (ms, ns, mfs, nfs):
            libc.free(p)
