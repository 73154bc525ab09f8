"""Out-of-core products through mgx_matmul_host with a transfer limit
(MGX_PARAM_TRANSFER, the reference's TRANSFER_SIZE, mat-gpu/mat-params.h:6):
square products recurse over host-memory blocks with the method's own 2 x 2
sequence until a block fits (mat-gpu/mat-pregpu.cu:10-96, the *_pregpu
drivers), other shapes are tiled; every result is compared with the oracle."""
import numpy as np
import pytest

import oracle
from helpers import is_normalized, to_ints

pytestmark = pytest.mark.gpu

TRANSFER = 4


@pytest.fixture
def transfer(gpu):
    def set_(t):
        gpu.call("mgx_set_param", TRANSFER, t)
    yield set_
    gpu.call("mgx_set_param", TRANSFER, 0)


def host_product(mgx, method, A, B, cap):
    K, M = A.shape[:2]
    N = B.shape[1]
    C = np.zeros((K, N, cap + 1), dtype=np.uint64)
    secs = mgx.ctypes.c_double()
    mgx.call("mgx_matmul_host", method, K, M, N, C.ctypes.data, np.ascontiguousarray(A).ctypes.data,
             np.ascontiguousarray(B).ctypes.data, cap, mgx.ctypes.byref(secs))
    return C


@pytest.mark.parametrize("method", [3, 35, 4])
def test_host_recursion_two_levels(gpu, transfer, method):
    mgx = gpu
    n, bits, cap = 128, 300, 13
    A, B = mgx.random_matrix(n, n, cap, bits, 70 + method), mgx.random_matrix(n, n, cap, bits, 80 + method)
    transfer(32)   # 128 -> 64 -> 32: two host levels
    C = host_product(mgx, method, A, B, cap)
    assert to_ints(C) == to_ints(oracle.mat_mul(A, B, cap))
    assert is_normalized(C)


@pytest.mark.parametrize("method", [37, 357, 47])
def test_host_recursion_interleaved_leaves(gpu, transfer, method):
    mgx = gpu
    n, bits, cap = 512, 200, 9
    A, B = mgx.random_matrix(n, n, cap, bits, 90 + method), mgx.random_matrix(n, n, cap, bits, 91 + method)
    transfer(256)   # one host level; each 256 x 256 block multiplies interleaved on the device
    C = host_product(mgx, method, A, B, cap)
    rows = [0, 200, 256, 511]
    assert to_ints(np.ascontiguousarray(C[rows])) == to_ints(oracle.mat_mul(np.ascontiguousarray(A[rows]), B, cap))
    assert is_normalized(C)


@pytest.mark.parametrize("method,K,M,N", [(35, 70, 45, 90), (8, 96, 96, 96), (2, 33, 70, 20), (0, 40, 64, 31)])
def test_host_tiled(gpu, transfer, method, K, M, N):
    mgx = gpu
    bits = 400
    cap = (2 * bits + 8 + 63) // 64 + 2
    A, B = mgx.random_matrix(K, M, cap, bits, K + method), mgx.random_matrix(M, N, cap, bits, N + method)
    transfer(32)
    C = host_product(mgx, method, A, B, cap)
    assert to_ints(C) == to_ints(oracle.mat_mul(A, B, cap))
    assert is_normalized(C)


def test_transfer_limit_off_is_unchanged(gpu, transfer):
    mgx = gpu
    n, bits, cap = 48, 500, 18
    A, B = mgx.random_matrix(n, n, cap, bits, 5), mgx.random_matrix(n, n, cap, bits, 6)
    transfer(0)
    C0 = host_product(mgx, 35, A, B, cap)
    transfer(16)
    C1 = host_product(mgx, 35, A, B, cap)
    assert to_ints(C0) == to_ints(C1) == to_ints(oracle.mat_mul(A, B, cap))
