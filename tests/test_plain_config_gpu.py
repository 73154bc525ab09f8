"""Plain route at the plain-route workload's entry size (4096-bit operands,
products of ~8200 bits) and the interleaved ("strided") paths, against exact
Python integers and the oracle.

The recursion is the reference's Strassen / Winograd / Winograd-addmul
(mat-gpu/mat-gpu.cu:166-365, AoS) and its interleaved form
(mat-gpu/mat-gpu.cu:367-475, launchers mat-gpu-launchers.cu:80-147).  Leaves
are forced small so that every test splits at least twice; 4096-bit operands
grow to 4098 and 4100 bits after two levels of operand sums, which moves the
leaf products from two 64-word blocks to three.
"""
import functools
import random

import numpy as np
import pytest

import oracle
from helpers import entry_int, is_normalized, matmul, rand_ints, to_entries, to_ints

pytestmark = pytest.mark.gpu

REC = {"naive": 0, "strassen": 1, "winograd": 2, "winograd_addmul": 3}


@functools.lru_cache(maxsize=None)
def case(n, bits, seed):
    A, B = rand_ints(n, n, bits, seed), rand_ints(n, n, bits, seed + 1)
    C0 = rand_ints(n, n, 2 * bits, seed + 2)
    return A, B, C0, matmul(A, B)


class Params:
    """Set library tunables for a test and restore them afterwards."""

    def __init__(self, mgx, **kv):
        self.mgx, self.kv, self.old = mgx, kv, {}

    def __enter__(self):
        ids = {"leaf": 0, "toom": 1, "depth": 2, "slices": 3}
        for k, v in self.kv.items():
            self.old[ids[k]] = self.mgx.lib().mgx_get_param(ids[k])
            self.mgx.call("mgx_set_param", ids[k], v)

    def __exit__(self, *a):
        for i, v in self.old.items():
            self.mgx.call("mgx_set_param", i, v)


@pytest.mark.parametrize("toom", [0, 2])
@pytest.mark.parametrize("method", ["strassen", "winograd", "winograd_addmul"])
@pytest.mark.parametrize("n", [32, 64])
def test_recursion_4096_bit(gpu, n, method, toom):
    mgx = gpu
    bits = 4096
    cap = (2 * bits + 16 + 63) // 64 + 2   # products plus the sum over k and the recursion's growth
    A, B, C0, AB = case(n, bits, 7 * n)
    addmul = method == "winograd_addmul"
    with Params(mgx, leaf=8, toom=toom):
        dA, dB = mgx.DeviceBuffer.from_array(to_entries(A, cap)), mgx.DeviceBuffer.from_array(to_entries(B, cap))
        dC = mgx.DeviceBuffer.from_array(to_entries(C0 if addmul else [[0] * n for _ in range(n)], cap))
        mgx.call("mgx_" + method, n, dC.ptr, n, dA.ptr, n, dB.ptr, n, cap, None)
        got = dC.to_array((n, n, cap + 1))
    want = [[C0[y][x] + AB[y][x] for x in range(n)] for y in range(n)] if addmul else AB
    assert to_ints(got) == want
    assert is_normalized(got)


def mod_q(X, q):
    """Each entry mod q (signed entries mapped to [0, q)), limb by limb."""
    cap = X.shape[-1] - 1
    r = np.zeros(X.shape[:2], dtype=np.uint64)
    sz = X[..., cap].view(np.int64)
    t64 = np.uint64(pow(2, 64, q))
    for li in range(int(np.abs(sz).max()) - 1, -1, -1):
        limb = np.where(np.abs(sz) > li, X[..., li], np.uint64(0)) % np.uint64(q)
        r = (r * t64 + limb) % np.uint64(q)
    return np.where(sz < 0, (np.uint64(q) - r) % np.uint64(q), r).astype(np.int64)


def matmul_mod(a, b, q):
    out = np.zeros((a.shape[0], b.shape[1]), dtype=np.int64)
    for k0 in range(0, a.shape[1], 256):   # partial sums stay below 2^63 for q < 2^26
        out = (out + a[:, k0:k0 + 256] @ b[k0:k0 + 256, :]) % q
    return out


def test_winograd_full_size_4096_bit(gpu):
    """n=256, 4096-bit entries, Winograd two levels deep (leaf 64): an
    independent modulus over every entry plus sampled exact entries."""
    mgx = gpu
    n, bits = 256, 4096
    cap = (2 * bits + 16 + 63) // 64 + 2
    A, B = mgx.random_matrix(n, n, cap, bits, 31), mgx.random_matrix(n, n, cap, bits, 32)
    with Params(mgx, leaf=64):
        dA, dB = mgx.DeviceBuffer.from_array(A), mgx.DeviceBuffer.from_array(B)
        dC = mgx.DeviceBuffer(n * n * (cap + 1) * 8)
        dC.zero()
        mgx.call("mgx_winograd", n, dC.ptr, n, dA.ptr, n, dB.ptr, n, cap, None)
        C = dC.to_array((n, n, cap + 1))
    q = 67108859   # a prime below 2^26
    assert (mod_q(C, q) == matmul_mod(mod_q(A, q), mod_q(B, q), q)).all()
    rng = random.Random(9)
    rows, cols = rng.sample(range(n), 3), rng.sample(range(n), 3)
    sub = oracle.mat_mul(np.ascontiguousarray(A[rows]), np.ascontiguousarray(B[:, cols]), cap)
    assert to_ints(sub) == [[entry_int(C[r, c]) for c in cols] for r in rows]
    assert is_normalized(C)


# ---------------------------------------------------------------- interleaved layout

def interleaved(mgx, X):
    d = mgx.DeviceBuffer.from_array(X)
    mgx.call("mgx_interleave", d.ptr, X.shape[0] * X.shape[1], X.shape[2] - 1, None)
    return d


@pytest.mark.parametrize("sub", [False, True])
def test_strided_add_sub(gpu, sub):
    mgx = gpu
    H, W, cap = 3, 512, 9
    A = rand_ints(H, W, 500, 61)
    B = rand_ints(H, W, 400, 62)
    A[0][0], B[0][0] = (1 << 512) - 1, -1 if sub else 1        # carry through every limb
    A[1][300], B[1][300] = 777, 777 if sub else -777           # exact zero
    A[2][511], B[2][511] = 0, -5
    a, b = to_entries(A, cap), to_entries(B, cap)
    dA, dB = interleaved(mgx, a), interleaved(mgx, b)
    dC = mgx.DeviceBuffer(H * W * (cap + 1) * 8)
    fn = "mgx_launcher_sub_strided" if sub else "mgx_launcher_add_strided"
    mgx.call(fn, W, H, dC.ptr, W, dA.ptr, W, dB.ptr, W, cap, None)
    mgx.call("mgx_check_error", None)
    mgx.call("mgx_deinterleave", dC.ptr, H * W, cap, None)
    got = dC.to_array((H, W, cap + 1))
    want = oracle.mat_aors(a, b, cap, sub=sub)
    assert to_ints(got) == to_ints(want)
    assert to_ints(got) == [[A[y][x] - B[y][x] if sub else A[y][x] + B[y][x] for x in range(W)] for y in range(H)]
    assert is_normalized(got)
    # the matrix-level form, in place (the recursion's C = C + T)
    fn2 = "mgx_sub_matrices_strided" if sub else "mgx_add_matrices_strided"
    mgx.call(fn2, W, H, dA.ptr, W, dA.ptr, W, dB.ptr, W, cap, None)
    mgx.call("mgx_deinterleave", dA.ptr, H * W, cap, None)
    assert to_ints(dA.to_array((H, W, cap + 1))) == to_ints(want)


@pytest.mark.parametrize("method", ["naive_recursive", "strassen", "winograd", "winograd_addmul"])
def test_strided_recursion_splits(gpu, method):
    """N = 512 interleaved: one real split into 256 x 256 quadrants (the
    interleaved leaf is at least one interleave block), checked on sampled
    rows against the oracle."""
    mgx = gpu
    n, bits, cap = 512, 256, 11
    A, B = mgx.random_matrix(n, n, cap, bits, 41), mgx.random_matrix(n, n, cap, bits, 42)
    C0 = mgx.random_matrix(n, n, cap, 500, 43) if method == "winograd_addmul" else np.zeros((n, n, cap + 1), np.uint64)
    dA, dB, dC = interleaved(mgx, A), interleaved(mgx, B), interleaved(mgx, C0)
    mgx.call("mgx_%s_strided" % method, n, dC.ptr, n, dA.ptr, n, dB.ptr, n, cap, None)
    mgx.call("mgx_deinterleave", dC.ptr, n * n, cap, None)
    C = dC.to_array((n, n, cap + 1))
    rows = [0, 255, 256, 511]
    want = oracle.mat_mul(np.ascontiguousarray(A[rows]), B, cap,
                          addmul_into=np.ascontiguousarray(C0[rows]) if method == "winograd_addmul" else None)
    assert to_ints(np.ascontiguousarray(C[rows])) == to_ints(want)
    assert is_normalized(C)


@pytest.mark.parametrize("method", [37, 357, 47])
def test_host_interleaved_methods_split(gpu, method):
    """mat-gpu's -m 37 / 357 / 47 on host matrices at N = 512 (one split)."""
    mgx = gpu
    n, bits, cap = 512, 200, 9
    A, B = mgx.random_matrix(n, n, cap, bits, 51), mgx.random_matrix(n, n, cap, bits, 52)
    C = np.zeros((n, n, cap + 1), dtype=np.uint64)
    secs = mgx.ctypes.c_double()
    mgx.call("mgx_matmul_host", method, n, n, n, C.ctypes.data, A.ctypes.data, B.ctypes.data, cap,
             mgx.ctypes.byref(secs))
    rows = [3, 300, 511]
    assert to_ints(np.ascontiguousarray(C[rows])) == to_ints(oracle.mat_mul(np.ascontiguousarray(A[rows]), B, cap))


def test_concurrent_routes_on_two_streams(gpu):
    """Winograd-route products with different leaf/depth choices queued on two
    streams from two host threads: the routes pass their recursion shape per
    call, so neither disturbs the other, and the global tunables are unchanged."""
    import threading
    mgx = gpu
    L = mgx.bind_sched()
    before = [mgx.lib().mgx_get_param(i) for i in range(4)]
    jobs = []
    for i, (n, route) in enumerate([(40, 2), (48, 3), (36, 2), (44, 3)]):
        bits, cap = 300, 13
        A, B = rand_ints(n, n, bits, 300 + i), rand_ints(n, n, bits, 400 + i)
        jobs.append((n, route, A, B, mgx.DeviceBuffer.from_array(to_entries(A, cap)),
                     mgx.DeviceBuffer.from_array(to_entries(B, cap)), mgx.DeviceBuffer(n * n * (cap + 1) * 8), cap))
    errors = []

    def worker(sel):
        s = mgx.c_vp()
        mgx.call("mgx_stream_create", mgx.ctypes.byref(s))
        try:
            for rep in range(3):
                for n, route, A, B, dA, dB, dC, cap in [jobs[j] for j in sel]:
                    st = L.mgx_matmul_route(route, n, n, n, dC.ptr, n, dA.ptr, n, dB.ptr, n, cap, s.value)
                    if st:
                        errors.append(mgx.lib().mgx_last_error())
        finally:
            mgx.call("mgx_stream_destroy", s.value)

    ts = [threading.Thread(target=worker, args=([0, 1],)), threading.Thread(target=worker, args=([2, 3],))]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errors
    for n, route, A, B, dA, dB, dC, cap in jobs:
        assert to_ints(dC.to_array((n, n, cap + 1))) == matmul(A, B)
    assert [mgx.lib().mgx_get_param(i) for i in range(4)] == before
