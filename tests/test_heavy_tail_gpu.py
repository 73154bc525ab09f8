"""Heavy-tailed entry mixes through the learned dispatch (MGX_ROUTE_AUTO with
the committed cost model): 90% of the entries 256-bit and 10% 32768-bit.
Products of 32768-bit entries exceed the plain kernel's widest class, so the
only route that can take them is the multi-modular one; the dispatch must
pick it and the result must be exact.

Small shapes are checked entry by entry against Python integers; the full
n=768 product by an independent modulus over a block of rows and by exact
sampled entries from the oracle.
"""
import ctypes
import os
import random

import numpy as np
import pytest

import oracle
from helpers import entry_int, is_normalized, matmul, matmul_mod, mod_q_entries, to_ints

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MODEL = os.path.join(ROOT, "bachelor-thesis_amd", "costmodel_params.txt")
LO, HI, FRAC = 256, 32768, 0.10


@pytest.fixture(scope="module")
def model(gpu):
    L = gpu.bind_sched()
    cm = ctypes.c_void_p()
    gpu.check(L.mgx_costmodel_load(ctypes.byref(cm), MODEL.encode()))
    gpu.check(L.mgx_dispatch_set_model(cm))
    yield cm
    L.mgx_dispatch_set_model(None)
    L.mgx_costmodel_destroy(cm)


def mixed(mgx, rows, cols, cap, seed):
    L = mgx.bind_sched()
    d = mgx.DeviceBuffer(rows * cols * (cap + 1) * 8)
    mgx.check(L.mgx_random_matrix_device(d.ptr, cols, rows, cols, cap, LO, HI, FRAC, seed, None))
    return d


def cap_for(n):
    return (2 * HI + n.bit_length() + 2 + 63) // 64 + 1


@pytest.mark.parametrize("K,M,N", [(8, 8, 8), (5, 17, 6)])
def test_heavy_tail_small_exact(gpu, model, K, M, N):
    mgx = gpu
    L = mgx.bind_sched()
    cap = cap_for(M)
    dA, dB = mixed(mgx, K, M, cap, 11 + K), mixed(mgx, M, N, cap, 12 + K)
    dC = mgx.DeviceBuffer(K * N * (cap + 1) * 8)
    r = ctypes.c_int()
    mgx.check(L.mgx_dispatch_choose(K, M, N, dA.ptr, M, dB.ptr, N, cap, ctypes.byref(r)))
    A, B = to_ints(dA.to_array((K, M, cap + 1))), to_ints(dB.to_array((M, N, cap + 1)))
    if max(abs(v).bit_length() for row in A + B for v in row) > 16384:
        assert r.value == 0   # only the multi-modular route takes such operands
    mgx.check(L.mgx_matmul_route(-1, K, M, N, dC.ptr, N, dA.ptr, M, dB.ptr, N, cap, None))
    C = dC.to_array((K, N, cap + 1))
    assert to_ints(C) == matmul(A, B)
    assert is_normalized(C)


def test_heavy_tail_full_size(gpu, model):
    """n=768, 90% 256-bit / 10% 32768-bit entries, learned dispatch."""
    mgx = gpu
    L = mgx.bind_sched()
    n = 768
    cap = cap_for(n)
    W = cap + 1
    dA, dB = mixed(mgx, n, n, cap, 21), mixed(mgx, n, n, cap, 22)
    dC = mgx.DeviceBuffer(n * n * W * 8)
    r = ctypes.c_int()
    mgx.check(L.mgx_dispatch_choose(n, n, n, dA.ptr, n, dB.ptr, n, cap, ctypes.byref(r)))
    assert r.value == 0
    mgx.check(L.mgx_matmul_route(-1, n, n, n, dC.ptr, n, dA.ptr, n, dB.ptr, n, cap, None))
    A, B, C = dA.to_array((n, n, W)), dB.to_array((n, n, W)), dC.to_array((n, n, W))
    dA.free(), dB.free(), dC.free()
    big = (np.abs(A[..., cap].view(np.int64)) > LO // 64).mean()
    assert 0.08 < big < 0.12
    q = 67108859   # a prime below 2^26, not a CRT prime (those lie in [16384, 65280) or above 2^30)
    rows = np.arange(0, n, 16)
    am = mod_q_entries(np.ascontiguousarray(A[rows]), q)
    bm = mod_q_entries(B, q)
    assert (mod_q_entries(np.ascontiguousarray(C[rows]), q) == matmul_mod(am, bm, q)).all()
    rng = random.Random(3)
    rs, cs = rng.sample(range(n), 2), rng.sample(range(n), 2)
    sub = oracle.mat_mul(np.ascontiguousarray(A[rs]), np.ascontiguousarray(B[:, cs]), cap)
    assert to_ints(sub) == [[entry_int(C[y, x]) for x in cs] for y in rs]
    assert is_normalized(np.ascontiguousarray(C[rows]))
