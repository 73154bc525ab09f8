"""The kept split-tree API on the device (gpu_crt_split_list / gpu_crt_merge_list,
mat-gpu/gpu_crt.h:10-12, crt.cpp:126-164) for many entries at once, against the
library's host functions (pinned in test_crt_tree_cpu.py against a restatement of
crt.cpp) and the restated recursion itself."""
import ctypes
import random

import numpy as np
import pytest

from helpers import entry_int

pytestmark = pytest.mark.gpu


def ent(mgx, v, cap):
    e = np.zeros(cap + 1, dtype=np.uint64)
    mgx.int_to_entry(v, cap, e)
    return e


@pytest.mark.parametrize("bits,limit,cap", [(300, 1 << 40, 8), (1000, 1 << 62, 20), (2000, 1 << 30, 40),
                                            (64, 1000, 3), (100, 1 << 62, 3)])
def test_split_merge_list_device(gpu, bits, limit, cap):
    mgx = gpu
    L = mgx.lib()
    W = cap + 1
    rng = random.Random(bits)
    x0 = rng.getrandbits(bits) | (1 << (bits - 1))
    ms, ns, mfs, nfs = (ctypes.c_void_p() for _ in range(4))
    length = L.mgx_crt_make_split_tree(ctypes.byref(ms), ctypes.byref(ns), ctypes.byref(mfs), ctypes.byref(nfs),
                                       ent(mgx, x0, cap).ctypes.data, limit, cap)
    assert length >= 0
    arr = lambda p: np.ctypeslib.as_array((ctypes.c_uint64 * max(length * W, 1)).from_address(p.value)).copy() \
        if length else np.zeros(W, dtype=np.uint64)
    tree = [arr(p) for p in (ms, ns, mfs, nfs)]
    mv = [tuple(entry_int(t[i * W:(i + 1) * W]) for t in tree) for i in range(length)]
    top = mv[0][0] * mv[0][1] if length else limit
    vals = [x0, 0, 1, -1, -x0, rng.randrange(top), -rng.randrange(top)]
    vals += [rng.randrange(-(1 << bits) + 1, 1 << bits) for _ in range(300)]
    X = np.stack([ent(mgx, v, cap) for v in vals])
    n, nl = len(vals), 1 << length
    dX = mgx.DeviceBuffer.from_array(X)
    dT = [mgx.DeviceBuffer.from_array(t) for t in tree]
    dR = mgx.DeviceBuffer(n * nl * W * 8)
    dB = mgx.DeviceBuffer(n * W * 8)
    mgx.check(L.mgx_crt_split_list_device(dR.ptr, dX.ptr, n, dT[0].ptr, dT[1].ptr, length, cap, None))
    mgx.check(L.mgx_crt_merge_list_device(dB.ptr, dR.ptr, n, dT[0].ptr, dT[1].ptr, dT[2].ptr, dT[3].ptr, length, cap,
                                          None))
    R = dR.to_array((n, nl, W))
    B = dB.to_array((n, W))

    def split_list(v, lvl):   # crt.cpp:126-142 restated
        if lvl == length:
            return [v]
        return split_list(v % mv[lvl][0], lvl + 1) + split_list(v % mv[lvl][1], lvl + 1)

    def merge_list(sp, lvl):   # crt.cpp:145-164 with crt_merge, crt.cpp:67-80
        if lvl == length:
            return sp[0]
        h = len(sp) // 2
        a, b = merge_list(sp[:h], lvl + 1), merge_list(sp[h:], lvl + 1)
        m, nn, mf, nf = mv[lvl]
        return (a * nf * nn + b * mf * m) % (m * nn)

    for i, v in enumerate(vals):
        want = split_list(v, 0) if length else [v]
        assert [entry_int(R[i, j]) for j in range(nl)] == want, (i, v)
        assert entry_int(B[i]) == merge_list(want, 0), (i, v)
        if length and i < 40:   # the host functions agree as well
            leaves = np.zeros((nl, W), dtype=np.uint64)
            assert L.mgx_crt_split_list(leaves.ctypes.data, X[i].ctypes.data, ms.value, ns.value, length, cap) == 0
            assert (leaves == R[i]).all()
        if 0 <= v < limit:
            assert entry_int(B[i]) == v   # every level reduces to the identity
    for p in (ms, ns, mfs, nfs):
        if p.value:
            ctypes.CDLL(None).free(p)


def test_split_merge_device_arguments(gpu):
    """Bad arguments are refused before any launch; no entries is a no-op."""
    mgx = gpu
    L = mgx.lib()
    d = mgx.DeviceBuffer(64)
    assert L.mgx_crt_split_list_device(None, d.ptr, 1, d.ptr, d.ptr, 1, 4, None) != 0
    assert L.mgx_crt_split_list_device(d.ptr, d.ptr, -1, d.ptr, d.ptr, 1, 4, None) != 0
    assert L.mgx_crt_split_list_device(d.ptr, d.ptr, 1, None, d.ptr, 1, 4, None) != 0
    assert L.mgx_crt_merge_list_device(d.ptr, d.ptr, 1, d.ptr, d.ptr, d.ptr, None, 1, 4, None) != 0
    assert L.mgx_crt_merge_list_device(d.ptr, d.ptr, 1, d.ptr, d.ptr, d.ptr, d.ptr, 21, 4, None) != 0
    assert L.mgx_crt_split_list_device(d.ptr, d.ptr, 0, d.ptr, d.ptr, 1, 4, None) == 0
    assert L.mgx_crt_merge_list_device(d.ptr, d.ptr, 0, d.ptr, d.ptr, d.ptr, d.ptr, 1, 4, None) == 0
