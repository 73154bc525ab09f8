"""Parity of the INT8-MFMA multi-modular route with the oracle (bit-exact)."""
import random

import numpy as np
import pytest

import oracle
from helpers import entry_int, is_normalized, matmul, rand_ints, to_entries, to_ints

pytestmark = pytest.mark.gpu


def run_crt(mgx, A, B, cap, plan=None):
    K, M = A.shape[:2]
    N = B.shape[1]
    dA, dB = mgx.DeviceBuffer.from_array(A), mgx.DeviceBuffer.from_array(B)
    dC = mgx.DeviceBuffer(max(K * N, 1) * (cap + 1) * 8)
    dC.zero()
    mgx.call("mgx_crt_mfma_matmul", plan, K, M, N, dC.ptr, N, dA.ptr, M, dB.ptr, N, cap, None)
    return dC.to_array((K, N, cap + 1))


@pytest.fixture(params=[2, 4], ids=["S2", "S4"])
def slices(request, gpu):
    """Digits per residue for the plans a test makes: 2 (primes below 65280) or 4 (below 2^31)."""
    gpu.call("mgx_set_param", 3, request.param)
    yield request.param
    gpu.call("mgx_set_param", 3, 0)


def product_plan_M(primes):
    M = 1
    for p in primes:
        M *= p
    return M


@pytest.fixture(params=[0, 1], ids=["fast", "tiles"])
def reduce_kernel(request, gpu):
    """MGX_PARAM_REDUCE: two-digit plans take the constant-resident reduce (entries of at
    most 256 bytes) and the tiled MFMA GEMM (over 512 bytes) with 0, the default, and the
    entry-resident and 64-row tiles with 1."""
    gpu.call("mgx_set_param", 6, request.param)
    yield request.param
    gpu.call("mgx_set_param", 6, 0)


@pytest.mark.parametrize("transpose", [0, 1])
# 64-, 128- and 256-byte entries (constant-resident or entry-resident), 384 bytes
# (entry-resident, two K halves), generic (> 512 bytes: 5 and 9 table tiles of 128 primes)
@pytest.mark.parametrize("bits,cap", [(200, 5), (700, 14), (2000, 34), (3000, 50), (5000, 82), (9000, 142)])
def test_reduce_kernel(gpu, slices, reduce_kernel, transpose, bits, cap):
    mgx = gpu
    if reduce_kernel and (slices == 4 or 2048 < bits <= 4096):
        pytest.skip("the reduce kernel choice applies to two-digit plans outside 257-512-byte entries")
    rows, K = 70, 45
    vals = rand_ints(rows, K, bits, 3 + transpose)
    vals[0][0], vals[1][1], vals[2][2] = 0, (1 << bits) - 1, -((1 << bits) - 1)
    X = to_entries(vals if not transpose else [list(r) for r in zip(*vals)], cap)
    plan = mgx.crt_plan(K, bits, bits)
    try:
        ps = mgx.crt_plan_primes(plan)
        S = mgx.lib().mgx_crt_plan_slices(plan)
        assert S == slices
        limit = 65279 if S == 2 else (1 << 31) - 1
        assert ps == oracle.primes(len(ps), limit=limit)
        assert len(ps) == oracle.lib().or_crt_prime_count_below(limit, 2 * bits + 6 + 2)
        rp, kp = mgx.c_i64(), mgx.c_i64()
        mgx.call("mgx_crt_residue_dims", plan, rows, K, mgx.ctypes.byref(rp), mgx.ctypes.byref(kp))
        rpad, kpad = rp.value, kp.value
        dX = mgx.DeviceBuffer.from_array(X)
        dR = mgx.DeviceBuffer(S * len(ps) * rpad * kpad)
        xrs = X.shape[1]
        mgx.call("mgx_crt_reduce", plan, transpose, rows, K, dX.ptr, xrs, cap, 0, dR.ptr, None)
        mgx.call("mgx_check_error", None)
        R = dR.to_array((len(ps), S, rpad, kpad), dtype=np.int8)
        if S == 2:   # planes in 16-row x 64-column blocks, a band's blocks in column order
            R = R.reshape(len(ps), S, rpad // 16, kpad // 64, 16, 64).transpose(0, 1, 2, 4, 3, 5).reshape(
                len(ps), S, rpad, kpad)
        for qi in [0, 1, len(ps) // 2, len(ps) - 1]:
            p = ps[qi]
            for r in range(rows):
                for k in range(K):
                    res = vals[r][k] % p
                    cen = res - p if res > p // 2 else res
                    d = oracle.slices(cen)   # four digits; the top S..3 vanish when |cen| < 2^15
                    assert d[S:] == [0] * (4 - S)
                    assert [int(R[qi, s, r, k]) for s in range(S)] == d[:S], (qi, r, k)
            assert not R[qi, :, rows:, :].any() and not R[qi, :, :, K:].any()   # zero padding
    finally:
        mgx.call("mgx_crt_plan_destroy", plan)


def test_residue_gemm_kernel(gpu, slices):
    mgx = gpu
    K, M, N, bits, cap = 37, 50, 29, 260, 10
    A, B = rand_ints(K, M, bits, 21), rand_ints(M, N, bits, 22)
    plan = mgx.crt_plan(M, bits, bits)
    try:
        ps = mgx.crt_plan_primes(plan)
        S = mgx.lib().mgx_crt_plan_slices(plan)
        Mprod = product_plan_M(ps)
        rp, kp, rpb = mgx.c_i64(), mgx.c_i64(), mgx.c_i64()
        mgx.call("mgx_crt_residue_dims", plan, K, M, mgx.ctypes.byref(rp), mgx.ctypes.byref(kp))
        mgx.call("mgx_crt_residue_dims", plan, N, M, mgx.ctypes.byref(rpb), None)
        a, b = to_entries(A, cap), to_entries(B, cap)
        dA, dB = mgx.DeviceBuffer.from_array(a), mgx.DeviceBuffer.from_array(b)
        dRA = mgx.DeviceBuffer(S * len(ps) * rp.value * kp.value)
        dRB = mgx.DeviceBuffer(S * len(ps) * rpb.value * kp.value)
        ystride = mgx.lib().mgx_crt_y_stride(K, N)
        assert ystride == (K * N + 255) // 256 * 256
        nplanes = (S * len(ps) + 3) // 4
        dY = mgx.DeviceBuffer(nplanes * ystride * 4)
        mgx.call("mgx_crt_reduce", plan, 0, K, M, dA.ptr, M, cap, 0, dRA.ptr, None)
        mgx.call("mgx_crt_reduce", plan, 1, N, M, dB.ptr, N, cap, 0, dRB.ptr, None)
        mgx.call("mgx_crt_residue_gemm", plan, K, M, N, dRA.ptr, dRB.ptr, dY.ptr, None)
        mgx.call("mgx_check_error", None)
        Y = dY.to_array((nplanes, ystride), dtype=np.uint32)
        C = matmul(A, B)
        for qi in [0, 3, len(ps) - 1]:
            p = ps[qi]
            inv = pow((Mprod // p) % p, p - 2, p)
            for i in range(K):
                for j in range(N):
                    if S == 4:
                        w = int(Y[qi, i * N + j])
                    else:   # plane q/2, low half for even q, high half for odd q
                        w = (int(Y[qi // 2, i * N + j]) >> (16 * (qi % 2))) & 0xffff
                    d = [((w >> (8 * s)) & 255) - (256 if (w >> (8 * s)) & 128 else 0) for s in range(S)]
                    y = sum(t * 256 ** s for s, t in enumerate(d)) % p
                    assert y == (C[i][j] % p) * inv % p, (qi, i, j)
    finally:
        mgx.call("mgx_crt_plan_destroy", plan)


SHAPES = [(1, 1, 1), (1, 5, 1), (3, 4, 5), (7, 1, 9), (33, 17, 65), (64, 64, 64), (130, 70, 129), (5, 300, 6)]


@pytest.mark.parametrize("K,M,N", SHAPES)
@pytest.mark.parametrize("bits", [1, 64, 200, 1024])
def test_crt_matmul_small(gpu, slices, K, M, N, bits):
    if bits == 1024 and K * M * N > 64 * 64 * 64:
        pytest.skip("oracle time")
    A, B = rand_ints(K, M, bits, K * 7 + bits), rand_ints(M, N, bits, N * 11 + bits)
    cap = (2 * bits + 12 + 63) // 64 + 1
    C = run_crt(gpu, to_entries(A, cap), to_entries(B, cap), cap)
    gpu.call("mgx_check_error", None)
    assert to_ints(C) == matmul(A, B)
    assert is_normalized(C)


def test_crt_edge_values(gpu, slices):
    # zeros, +-1, all-ones magnitudes, exact cancellation to zero, one-sided signs
    bits, cap = 2048, 66
    top = (1 << bits) - 1
    A = [[top, -top, 0, 1], [-1, 0, top, top], [0, 0, 0, 0]]
    B = [[top, 1, 0], [top, -1, 5], [7, 0, -top], [-7, 0, top]]
    C = run_crt(gpu, to_entries(A, cap), to_entries(B, cap), cap)
    assert to_ints(C) == matmul(A, B)
    assert is_normalized(C)


@pytest.mark.parametrize("M", [8300, 33000])
def test_crt_long_inner_dimension(gpu, slices, M):
    # K beyond one int32 accumulation chunk (8192 for S = 4, 32768 for S = 2) takes the folding variant
    K, N, bits, cap = 3, 2, 100, 6
    A, B = rand_ints(K, M, bits, 1), rand_ints(M, N, bits, 2)
    C = run_crt(gpu, to_entries(A, cap), to_entries(B, cap), cap)
    assert to_ints(C) == matmul(A, B)


def test_crt_capacity_error(gpu, slices):
    mgx = gpu
    A = [[(1 << 300) - 1]]
    cap = 6   # the product needs 10 limbs
    with pytest.raises(mgx.MgxError) as ei:
        run_crt(mgx, to_entries(A, cap), to_entries(A, cap), cap)
    assert ei.value.status == 1


def test_crt_explicit_plan_reused(gpu, slices):
    mgx = gpu
    plan = mgx.crt_plan(40, 500, 300)
    try:
        for seed in range(3):
            A, B = rand_ints(9, 40, 500, seed), rand_ints(40, 11, 300, seed + 100)
            C = run_crt(mgx, to_entries(A, 14), to_entries(B, 14), 14, plan=plan)
            assert to_ints(C) == matmul(A, B)
    finally:
        mgx.call("mgx_crt_plan_destroy", plan)


def test_crt_full_size_properties(gpu, slices):
    """n=1024, 2048-bit entries: an independent modulus, sampled exact entries."""
    mgx = gpu
    n, bits, cap = 1024, 2048, 66
    A, B = mgx.random_matrix(n, n, cap, bits, 1), mgx.random_matrix(n, n, cap, bits, 2)
    C = run_crt(mgx, A, B, cap)
    assert is_normalized(C[:8])
    q = 67108859   # largest prime below 2^26, not a CRT prime (those are below 2^16 or above 2^30)

    def mod_q(X):
        r = np.zeros(X.shape[:2], dtype=np.uint64)
        sz = X[..., cap].view(np.int64)
        t64 = np.uint64(pow(2, 64, q))
        for li in range(cap - 1, -1, -1):
            limb = np.where(np.abs(sz) > li, X[..., li], np.uint64(0)) % np.uint64(q)
            r = (r * t64 + limb) % np.uint64(q)
        return np.where(sz < 0, (np.uint64(q) - r) % np.uint64(q), r).astype(np.int64)

    am, bm = mod_q(A), mod_q(B)
    want = np.zeros((n, n), dtype=np.int64)
    for k0 in range(0, n, 256):   # keep partial sums below 2^63
        want = (want + am[:, k0:k0 + 256] @ bm[k0:k0 + 256, :]) % q
    assert (mod_q(C) == want).all()
    rng = random.Random(5)
    rows, cols = rng.sample(range(n), 3), rng.sample(range(n), 3)
    sub = oracle.mat_mul(np.ascontiguousarray(A[rows]), np.ascontiguousarray(B[:, cols]), cap)
    assert to_ints(sub) == [[entry_int(C[r, c]) for c in cols] for r in rows]


def test_crt_stages_captured_in_graph(gpu, slices):
    # The stage entry points only enqueue work; captured into a HIP graph
    # (reconstruct then takes graph-owned scratch) and replayed, with eager
    # work between the replays, every replay gives the eager product.  The
    # graph calls go to the HIP runtime the library itself uses.
    import ctypes
    mgx = gpu
    hip = ctypes.CDLL("libamdhip64.so")
    vp = ctypes.c_void_p
    K, M, N, bits, cap = 40, 33, 36, 400, 16
    A, B = rand_ints(K, M, bits, 71), rand_ints(M, N, bits, 72)
    plan = mgx.crt_plan(M, bits, bits)
    stream = vp()
    mgx.call("mgx_stream_create", ctypes.byref(stream))
    graph, exe = vp(), vp()
    try:
        P = mgx.lib().mgx_crt_plan_primes(plan)
        S = mgx.lib().mgx_crt_plan_slices(plan)
        rp, kp, rpb = mgx.c_i64(), mgx.c_i64(), mgx.c_i64()
        mgx.call("mgx_crt_residue_dims", plan, K, M, mgx.ctypes.byref(rp), mgx.ctypes.byref(kp))
        mgx.call("mgx_crt_residue_dims", plan, N, M, mgx.ctypes.byref(rpb), None)
        dA, dB = mgx.DeviceBuffer.from_array(to_entries(A, cap)), mgx.DeviceBuffer.from_array(to_entries(B, cap))
        dRA = mgx.DeviceBuffer(S * P * rp.value * kp.value)
        dRB = mgx.DeviceBuffer(S * P * rpb.value * kp.value)
        dY = mgx.DeviceBuffer((S * P + 3) // 4 * mgx.lib().mgx_crt_y_stride(K, N) * 4)
        dC = mgx.DeviceBuffer(K * N * (cap + 1) * 8)
        want = matmul(A, B)

        def stages(s):
            mgx.call("mgx_crt_reduce", plan, 0, K, M, dA.ptr, M, cap, 0, dRA.ptr, s)
            mgx.call("mgx_crt_reduce", plan, 1, N, M, dB.ptr, N, cap, 0, dRB.ptr, s)
            mgx.call("mgx_crt_residue_gemm", plan, K, M, N, dRA.ptr, dRB.ptr, dY.ptr, s)
            mgx.call("mgx_crt_reconstruct", plan, K, N, dY.ptr, dC.ptr, N, cap, s)

        stages(None)
        mgx.call("mgx_check_error", None)
        assert to_ints(dC.to_array((K, N, cap + 1))) == want
        assert hip.hipStreamBeginCapture(stream, 1) == 0   # thread-local capture
        stages(stream.value)
        assert hip.hipStreamEndCapture(stream, ctypes.byref(graph)) == 0
        assert hip.hipGraphInstantiate(ctypes.byref(exe), graph, None, None, ctypes.c_size_t(0)) == 0
        for rep in range(3):
            dC.zero()
            mgx.call("mgx_device_sync")
            assert hip.hipGraphLaunch(exe, stream) == 0
            mgx.call("mgx_stream_sync", stream.value)
            mgx.call("mgx_check_error", None)
            assert to_ints(dC.to_array((K, N, cap + 1))) == want, rep
            stages(None)   # eager work between replays
            mgx.call("mgx_device_sync")
    finally:
        if exe.value:
            hip.hipGraphExecDestroy(exe)
        if graph.value:
            hip.hipGraphDestroy(graph)
        mgx.call("mgx_stream_destroy", stream.value)
        mgx.call("mgx_crt_plan_destroy", plan)
