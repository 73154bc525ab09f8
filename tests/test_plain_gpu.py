"""Parity of the plain route (per-entry mpn arithmetic, recursion) with the
oracle and exact integers (bit-exact)."""
import numpy as np
import pytest

import oracle
from helpers import is_normalized, matmul, rand_ints, to_entries, to_ints

pytestmark = pytest.mark.gpu


def dev(mgx, a):
    return mgx.DeviceBuffer.from_array(a)


def run_mul(mgx, fn, A, B, cap, C0=None, rs=None):
    K, M = A.shape[:2]
    N = B.shape[1]
    dA, dB = dev(mgx, A), dev(mgx, B)
    dC = dev(mgx, C0) if C0 is not None else mgx.DeviceBuffer(K * N * (cap + 1) * 8)
    if C0 is None:
        dC.zero()
    mgx.call(fn, K, M, N, dC.ptr, N, dA.ptr, M, dB.ptr, N, cap, None)
    mgx.call("mgx_check_error", None)
    return dC.to_array((K, N, cap + 1))


@pytest.mark.parametrize("bits", [1, 31, 32, 63, 64, 65, 1000, 2048, 2049, 4096, 5000])
@pytest.mark.parametrize("K,M,N", [(1, 1, 1), (3, 4, 5), (9, 7, 2), (16, 16, 16)])
def test_launcher_mul(gpu, bits, K, M, N):
    A, B = rand_ints(K, M, bits, bits + K), rand_ints(M, N, bits, bits + N + 1)
    cap = (2 * bits + 8 + 63) // 64 + 1
    C = run_mul(gpu, "mgx_launcher_mul", to_entries(A, cap), to_entries(B, cap), cap)
    assert to_ints(C) == matmul(A, B)
    assert is_normalized(C)


@pytest.mark.parametrize("toom", [0, 2])
def test_launcher_mul_karatsuba(gpu, toom):
    mgx = gpu
    old = mgx.lib().mgx_get_param(1)
    mgx.call("mgx_set_param", 1, toom)
    try:
        for bits in [4096, 4000, 8192]:
            K, M, N = 5, 6, 4
            A = rand_ints(K, M, bits, 3)
            B = rand_ints(M, N, bits, 4, mode="ones" if bits == 8192 else "rand")
            cap = (2 * bits + 8 + 63) // 64 + 1
            C = run_mul(mgx, "mgx_launcher_mul", to_entries(A, cap), to_entries(B, cap), cap)
            assert to_ints(C) == matmul(A, B), (toom, bits)
    finally:
        mgx.call("mgx_set_param", 1, old)


def test_launcher_addmul(gpu):
    K, M, N, bits = 6, 5, 7, 700
    cap = (2 * bits + 8 + 63) // 64 + 2
    A, B = rand_ints(K, M, bits, 1), rand_ints(M, N, bits, 2)
    C0 = rand_ints(K, N, 1500, 3)
    C0[0][0] = -sum(A[0][k] * B[k][0] for k in range(M))     # cancels to zero
    C = run_mul(gpu, "mgx_launcher_addmul", to_entries(A, cap), to_entries(B, cap), cap, C0=to_entries(C0, cap))
    AB = matmul(A, B)
    assert to_ints(C) == [[C0[y][x] + AB[y][x] for x in range(N)] for y in range(K)]
    assert to_ints(C)[0][0] == 0
    assert is_normalized(C)


@pytest.mark.parametrize("sub", [False, True])
def test_launcher_add_sub(gpu, sub):
    mgx = gpu
    H, W, cap = 9, 11, 10
    A = rand_ints(H, W, 500, 5)
    B = rand_ints(H, W, 300, 6)
    A[0][0], B[0][0] = (1 << 512) - 1, -1 if sub else 1          # carry through every limb
    A[1][1], B[1][1] = 12345, 12345 if sub else -12345           # exact zero
    A[2][2], B[2][2] = 0, 0
    A[3][3], B[3][3] = -(1 << 256), (1 << 256) if sub else -(1 << 256)
    dA, dB = dev(mgx, to_entries(A, cap)), dev(mgx, to_entries(B, cap))
    dC = mgx.DeviceBuffer(H * W * (cap + 1) * 8)
    mgx.call("mgx_launcher_sub" if sub else "mgx_launcher_add", W, H, dC.ptr, W, dA.ptr, W, dB.ptr, W, cap, None)
    mgx.call("mgx_check_error", None)
    C = dC.to_array((H, W, cap + 1))
    want = [[A[y][x] - B[y][x] if sub else A[y][x] + B[y][x] for x in range(W)] for y in range(H)]
    assert to_ints(C) == want
    assert is_normalized(C)
    # in place: C = C + B (the recursion's ADD(c11, C, c11, C, T, T))
    mgx.call("mgx_add_matrices", W, H, dA.ptr, W, dA.ptr, W, dB.ptr, W, cap, None)
    assert to_ints(dA.to_array((H, W, cap + 1))) == [[A[y][x] + B[y][x] for x in range(W)] for y in range(H)]


def test_interleave_roundtrip(gpu):
    mgx = gpu
    n, cap = 512, 4
    X = mgx.random_matrix(2, 256, cap, 200, 9)
    d = dev(mgx, X)
    mgx.call("mgx_interleave", d.ptr, n, cap, None)
    Y = d.to_array(X.size).copy()
    ref = X.reshape(-1).copy()
    oracle.lib().or_transform_interleave(ref.ctypes.data, 2, 256, (cap + 1) * 8, 8)
    assert (Y == ref).all()
    mgx.call("mgx_deinterleave", d.ptr, n, cap, None)
    assert (d.to_array(X.shape) == X).all()


@pytest.mark.parametrize("fn", ["mgx_launcher_mul_strided", "mgx_launcher_addmul_strided"])
def test_strided_launchers(gpu, fn):
    mgx = gpu
    K, M, N, bits, cap = 2, 256, 256, 300, 12
    A, B = mgx.random_matrix(K, M, cap, bits, 1), mgx.random_matrix(M, N, cap, bits, 2)
    C0 = mgx.random_matrix(K, N, cap, 400, 3)
    dA, dB, dC = dev(mgx, A), dev(mgx, B), dev(mgx, C0)
    for d, cnt in ((dA, K * M), (dB, M * N), (dC, K * N)):
        mgx.call("mgx_interleave", d.ptr, cnt, cap, None)
    mgx.call(fn, K, M, N, dC.ptr, N, dA.ptr, M, dB.ptr, N, cap, None)
    mgx.call("mgx_check_error", None)
    mgx.call("mgx_deinterleave", dC.ptr, K * N, cap, None)
    got = dC.to_array((K, N, cap + 1))
    want = oracle.mat_mul(A, B, cap, addmul_into=C0 if "addmul" in fn else None)
    assert to_ints(got) == to_ints(want)


@pytest.mark.parametrize("method", [0, 1, 2, 3])
@pytest.mark.parametrize("n,leaf", [(32, 8), (48, 12), (64, 64)])
def test_recursion_methods(gpu, method, n, leaf):
    mgx = gpu
    old = mgx.lib().mgx_get_param(0)
    mgx.call("mgx_set_param", 0, leaf)
    try:
        bits, cap = 300, 13
        A, B = rand_ints(n, n, bits, method + n), rand_ints(n, n, bits, method + n + 50)
        C0 = rand_ints(n, n, 600, 7) if method == 3 else [[0] * n for _ in range(n)]
        dA, dB, dC = dev(mgx, to_entries(A, cap)), dev(mgx, to_entries(B, cap)), dev(mgx, to_entries(C0, cap))
        mgx.call("mgx_recursive", method, 0, n, dC.ptr, n, dA.ptr, n, dB.ptr, n, cap, None)
        AB = matmul(A, B)
        assert to_ints(dC.to_array((n, n, cap + 1))) == [[C0[y][x] + AB[y][x] for x in range(n)] for y in range(n)]
    finally:
        mgx.call("mgx_set_param", 0, old)


@pytest.mark.parametrize("K,M,N", [(5, 9, 7), (40, 24, 33), (20, 20, 20), (17, 64, 3)])
def test_squarify(gpu, K, M, N):
    mgx = gpu
    old = mgx.lib().mgx_get_param(0)
    mgx.call("mgx_set_param", 0, 4)
    try:
        bits, cap = 128, 6
        A, B = rand_ints(K, M, bits, K), rand_ints(M, N, bits, N)
        C = run_mul(mgx, None or "mgx_naive_smallrect", to_entries(A, cap), to_entries(B, cap), cap)
        assert to_ints(C) == matmul(A, B)
        dA, dB = dev(mgx, to_entries(A, cap)), dev(mgx, to_entries(B, cap))
        dC = mgx.DeviceBuffer(K * N * (cap + 1) * 8)
        for method in (0, 1, 2):
            dC.zero()
            mgx.call("mgx_squarify_mul", method, K, M, N, dC.ptr, N, dA.ptr, M, dB.ptr, N, cap, None)
            assert to_ints(dC.to_array((K, N, cap + 1))) == matmul(A, B), method
    finally:
        mgx.call("mgx_set_param", 0, old)


@pytest.mark.parametrize("method", [1, 2, 3, 35, 37, 357, 4, 47, 8])
def test_matmul_host_methods(gpu, method):
    mgx = gpu
    n, bits, cap = 256, 256, 10
    A, B = mgx.random_matrix(n, n, cap, bits, 11), mgx.random_matrix(n, n, cap, bits, 12)
    C = np.zeros((n, n, cap + 1), dtype=np.uint64)
    secs = mgx.ctypes.c_double()
    mgx.call("mgx_matmul_host", method, n, n, n, C.ctypes.data, A.ctypes.data, B.ctypes.data, cap,
             mgx.ctypes.byref(secs))
    rows = [0, 17, 255]
    sub = oracle.mat_mul(np.ascontiguousarray(A[rows]), B, cap)
    assert to_ints(np.ascontiguousarray(C[rows])) == to_ints(sub)
    assert is_normalized(C)


def test_plain_capacity_error(gpu):
    mgx = gpu
    A = [[(1 << 300) - 1]]
    cap = 6
    with pytest.raises(mgx.MgxError) as ei:
        run_mul(mgx, "mgx_launcher_mul", to_entries(A, cap), to_entries(A, cap), cap)
    assert ei.value.status == 1


def test_plain_equals_crt_config3_sized_block(gpu):
    """Both routes on the same 2048-bit inputs give identical bytes."""
    mgx = gpu
    n, bits, cap = 256, 2048, 66
    A, B = mgx.random_matrix(n, n, cap, bits, 21), mgx.random_matrix(n, n, cap, bits, 22)
    P = run_mul(mgx, "mgx_launcher_mul", A, B, cap)
    from test_crt_gpu import run_crt
    Q = run_crt(mgx, A, B, cap)
    sz = P[..., cap].view(np.int64)
    assert (sz == Q[..., cap].view(np.int64)).all()
    L = np.abs(sz)
    mask = np.arange(cap)[None, None, :] < L[..., None]
    assert (np.where(mask, P[..., :cap], 0) == np.where(mask, Q[..., :cap], 0)).all()


def test_scratch_reuse_across_streams(gpu):
    # Recursion temporaries come from the library's block cache: blocks released
    # on one stream are taken by work queued on another without a host
    # synchronization in between; every product must still be exact, also
    # after the idle blocks are unmapped.
    mgx = gpu
    old = mgx.lib().mgx_get_param(0)
    mgx.call("mgx_set_param", 0, 8)
    streams = [mgx.c_vp(), mgx.c_vp()]
    for s in streams:
        mgx.call("mgx_stream_create", mgx.ctypes.byref(s))
    try:
        bits, cap = 300, 13
        jobs = []
        for i, (n, si) in enumerate([(64, 0), (32, 1), (64, 1), (32, 0), (64, 0)]):
            A, B = rand_ints(n, n, bits, 900 + i), rand_ints(n, n, bits, 950 + i)
            dA, dB = dev(mgx, to_entries(A, cap)), dev(mgx, to_entries(B, cap))
            dC = mgx.DeviceBuffer(n * n * (cap + 1) * 8)
            dC.zero()
            jobs.append((n, A, B, dA, dB, dC, streams[si].value))
        mgx.call("mgx_device_sync")
        for n, A, B, dA, dB, dC, s in jobs:
            mgx.call("mgx_recursive", 2, 0, n, dC.ptr, n, dA.ptr, n, dB.ptr, n, cap, s)
        for s in streams:
            mgx.call("mgx_stream_sync", s.value)
        for n, A, B, dA, dB, dC, s in jobs:
            assert to_ints(dC.to_array((n, n, cap + 1))) == matmul(A, B)
        mgx.call("mgx_scratch_release")
        n, A, B, dA, dB, dC, s = jobs[0]
        dC.zero()
        mgx.call("mgx_device_sync")
        mgx.call("mgx_recursive", 1, 0, n, dC.ptr, n, dA.ptr, n, dB.ptr, n, cap, s)
        mgx.call("mgx_stream_sync", s)
        assert to_ints(dC.to_array((n, n, cap + 1))) == matmul(A, B)
    finally:
        for s in streams:
            mgx.call("mgx_stream_destroy", s.value)
        mgx.call("mgx_set_param", 0, old)
