"""The batched executor (device work queues feeding persistent kernels,
captured in one hipGraph) on batches that reach every kernel class, each
product checked against exact Python integers; the plan cache's pinning of a
live batch's plans; and the device check of the plans' per-operand bit
bounds.

Classes reached: entry-resident reduce (entries up to 256 bytes and up to
512 bytes), generic reduce (entries over 512 bytes), residue GEMM with two
and with four digits per residue, both with K beyond one int32 accumulation
chunk, reconstruct for both digit counts, the carry pass, rectangular
K x 2K x K products, empty products, the plain per-entry kernel at
several block classes (up to 8 blocks of 64 words), and plain products with
one or two Winograd levels (block-addition stages around the plain kernel).
"""
import ctypes

import numpy as np
import pytest

from helpers import is_normalized, matmul, rand_ints, to_entries, to_ints

pytestmark = pytest.mark.gpu

vp, i64 = ctypes.c_void_p, ctypes.c_int64
CRT, PLAIN, AUTO = 0, 1, -1


def cap_for(bits, M):
    return (2 * bits + max(M, 2).bit_length() + 2 + 63) // 64 + 1


def build(mgx, specs, seed0):
    """Device buffers and mgx_product records for (K, M, N, bits, route) specs."""
    bufs, prods, wants = [], [], []
    for i, spec in enumerate(specs):
        K, M, N, bits, route = spec[:5]
        cap = spec[5] if len(spec) > 5 else cap_for(bits, M)
        A, B = rand_ints(K, M, bits, seed0 + 2 * i), rand_ints(M, N, bits, seed0 + 2 * i + 1)
        if K and M:
            A[0][0] = (1 << bits) - 1            # all-ones magnitude
        if M and N:
            B[0][0] = -((1 << bits) - 1)
        dA = mgx.DeviceBuffer.from_array(to_entries(A, cap)) if K * M else mgx.DeviceBuffer(8)
        dB = mgx.DeviceBuffer.from_array(to_entries(B, cap))
        dC = mgx.DeviceBuffer(max(K * N, 1) * (cap + 1) * 8)
        bufs.append((dA, dB, dC, K, N, cap))
        prods.append(mgx.Product(K, M, N, dC.ptr, N, dA.ptr, M, dB.ptr, N, cap, 0, 0, route))
        wants.append(matmul(A, B) if K else [])
    return bufs, prods, wants


def run_and_check(mgx, L, bufs, prods, wants, reps=2):
    arr = (mgx.Product * len(prods))(*prods)
    b = vp()
    mgx.check(L.mgx_batch_create(ctypes.byref(b), ctypes.cast(arr, vp), len(prods), None))
    try:
        routes = [L.mgx_batch_route(b, i) for i in range(len(prods))]
        for rep in range(reps):
            for (_, _, dC, K, N, cap) in bufs:
                dC.zero()
            mgx.check(L.mgx_batch_run(b, None))
            mgx.call("mgx_check_error", None)
            for i, ((_, _, dC, K, N, cap), want) in enumerate(zip(bufs, wants)):
                if K == 0 or N == 0:
                    continue
                got = dC.to_array((K, N, cap + 1))
                assert to_ints(got) == want, (rep, i, prods[i].K, prods[i].M, prods[i].N)
                assert is_normalized(got)
            check_queues(mgx, L, b)
        return routes
    finally:
        L.mgx_batch_destroy(b)


def check_queues(mgx, L, b):
    """Every persistent launch drained its queue: each item claimed once and
    each workgroup left after one failing claim (counter = items + grid)."""
    nq = L.mgx_batch_queue_count(b)
    assert nq > 0
    for q in range(nq):
        kind, items, grid, claims = ctypes.c_int(), i64(), ctypes.c_int(), i64()
        mgx.check(L.mgx_batch_queue_info(b, q, ctypes.byref(kind), ctypes.byref(items), ctypes.byref(grid),
                                         ctypes.byref(claims)))
        assert 0 < grid.value <= max(items.value, 1)
        # eight counters per queue, one failing claim per workgroup on each
        assert claims.value == items.value + 8 * grid.value, (q, kind.value, items.value, grid.value, claims.value)


def test_batch_every_class_two_digits(gpu):
    mgx = gpu
    L = mgx.bind_sched()
    specs = [
        (24, 24, 24, 5000, CRT),     # entries over 512 bytes: generic reduce
        (20, 30, 20, 3000, CRT),     # 257-512 bytes: entry-resident reduce, 8 K steps
        (16, 16, 16, 1500, CRT),     # up to 256 bytes
        (12, 24, 12, 800, CRT),      # rectangular K x 2K x K
        (2, 33000, 3, 100, CRT),     # K beyond one 32768-deep accumulation chunk
        (1, 1, 1, 64, CRT),
        (0, 4, 4, 100, CRT),         # empty: nothing to do
        (5, 6, 4, 7000, PLAIN),      # plain kernel, 4 blocks of 64 words
        (3, 4, 3, 15000, PLAIN),     # plain kernel, 8 blocks
        (6, 6, 6, 300, PLAIN),       # plain kernel, 1 block
        (9, 18, 9, 2000, PLAIN),     # plain, rectangular
    ]
    bufs, prods, wants = build(mgx, specs, 500)
    routes = run_and_check(mgx, L, bufs, prods, wants)
    assert routes[:7] == [CRT] * 7 and routes[7:] == [PLAIN] * 4


def test_batch_every_class_four_digits(gpu):
    mgx = gpu
    L = mgx.bind_sched()
    old = mgx.lib().mgx_get_param(3)
    mgx.call("mgx_set_param", 3, 4)   # plans made while this is set use primes below 2^31, 4 digits
    try:
        specs = [
            (20, 20, 20, 700, CRT),
            (10, 8300, 3, 100, CRT),  # beyond one 8192-deep chunk
            (8, 8, 8, 5000, CRT),     # generic reduce
            (8, 10, 8, 3000, CRT),    # entry-resident, 8 K steps
            (7, 14, 7, 1200, CRT),    # rectangular
        ]
        bufs, prods, wants = build(mgx, specs, 700)
        routes = run_and_check(mgx, L, bufs, prods, wants)
        assert routes == [CRT] * 5
    finally:
        mgx.call("mgx_set_param", 3, old)


def test_batch_winograd_levels(gpu):
    """Plain products with one or two Winograd levels run as batch stages:
    block additions in their own queues around the plain kernel's queues.
    Odd K, M and N at either level (peeled rows and columns, the rank-one
    fix-up), shapes too small to split, and the fall-backs: a leaf past the
    plain kernel's widest class, and a capacity with no room for the
    temporaries, run with fewer levels."""
    mgx = gpu
    L = mgx.bind_sched()
    W1, W2 = 2, 3
    specs = [
        (16, 16, 16, 300, W1),
        (17, 9, 13, 700, W1),        # odd K, M and N
        (20, 12, 28, 2000, W2),      # even at level 0, odd at level 1
        (9, 18, 9, 1500, W2),        # rectangular, odd outside
        (12, 11, 10, 4000, W2),      # odd inner dimension at level 0
        (2, 2, 2, 64, W2),           # level 1 is 1 x 1 x 1
        (1, 5, 7, 100, W1),          # one row: no split
        (6, 6, 6, 15000, W1),        # leaves of 8 blocks
        (4, 6, 4, 16382, W2),        # second-level leaves would pass 512 words: one level
        (5, 6, 4, 16383, W2),        # first-level leaves would too: flat plain kernel
        (8, 8, 8, 1020, W1, 32),     # 2044-bit products in 32 limbs: no room for the temporaries
        (24, 24, 24, 1500, CRT),     # a multi-modular product in the same batch
    ]
    bufs, prods, wants = build(mgx, specs, 900)
    routes = run_and_check(mgx, L, bufs, prods, wants)
    assert routes == [W1, W1, W2, W2, W2, W2, W1, W1, W1, PLAIN, PLAIN, CRT]


@pytest.mark.parametrize("mode", [1, 2])
def test_batch_reconstruct_modes(gpu, mode):
    """Reconstruct mode 1 gives every two-digit product the word-buffer GEMM and
    the carry pass, mode 2 the one-pass reconstruct, row blocks past 9
    included (the default splits at 9); both exact."""
    mgx = gpu
    L = mgx.bind_sched()
    old = mgx.lib().mgx_get_param(7)
    mgx.call("mgx_set_param", 7, mode)
    try:
        specs = [
            (16, 16, 16, 3000, CRT),     # 6 row blocks
            (8, 9, 8, 8000, CRT),        # 17 row blocks
            (4, 5, 6, 16000, CRT),       # 33 row blocks
            (33, 20, 17, 700, CRT),
        ]
        bufs, prods, wants = build(mgx, specs, 1100 + mode)
        routes = run_and_check(mgx, L, bufs, prods, wants)
        assert routes == [CRT] * 4
    finally:
        mgx.call("mgx_set_param", 7, old)


def test_batch_plans_survive_cache_turnover(gpu):
    """A live batch keeps its plans pinned: after more than 256 other shapes
    have passed through the plan cache, replays are still exact."""
    mgx = gpu
    L = mgx.bind_sched()
    specs = [(10, 12, 9, 900, CRT), (7, 7, 7, 4000, CRT), (6, 9, 5, 6000, CRT)]
    bufs, prods, wants = build(mgx, specs, 800)
    arr = (mgx.Product * len(prods))(*prods)
    b = vp()
    mgx.check(L.mgx_batch_create(ctypes.byref(b), ctypes.cast(arr, vp), len(prods), None))
    try:
        tot, pin = i64(), i64()
        mgx.call("mgx_crt_plan_cache_stats", ctypes.byref(tot), ctypes.byref(pin))
        assert pin.value >= 3
        # 300 new shapes (distinct bit lengths) through the one-product route
        cap = 6
        one = mgx.DeviceBuffer.from_array(to_entries([[1]], cap))
        out = mgx.DeviceBuffer((cap + 1) * 8)
        for bits in range(2, 302):
            x = mgx.DeviceBuffer.from_array(to_entries([[(1 << (bits - 1)) | 1]], cap))
            mgx.call("mgx_crt_mfma_matmul", None, 1, 1, 1, out.ptr, 1, x.ptr, 1, one.ptr, 1, cap, None)
            assert to_ints(out.to_array((1, 1, cap + 1))) == [[(1 << (bits - 1)) | 1]]
        mgx.call("mgx_crt_plan_cache_stats", ctypes.byref(tot), ctypes.byref(pin))
        assert pin.value >= 3 and tot.value <= 256 + pin.value
        for rep in range(2):
            for (_, _, dC, K, N, cap_) in bufs:
                dC.zero()
            mgx.check(L.mgx_batch_run(b, None))
            mgx.call("mgx_check_error", None)
            for (_, _, dC, K, N, cap_), want in zip(bufs, wants):
                assert to_ints(dC.to_array((K, N, cap_ + 1))) == want, rep
    finally:
        L.mgx_batch_destroy(b)
    mgx.call("mgx_crt_plan_cache_stats", ctypes.byref(tot), ctypes.byref(pin))
    assert pin.value == 0


@pytest.mark.parametrize("plan_bits,entry_bits", [(100, 101), (100, 120), (100, 500), (4000, 4001), (5000, 5050)])
def test_entry_above_plan_bound_is_rejected(gpu, plan_bits, entry_bits):
    """Entries between the plan's bound and its byte-rounded width must not
    pass silently: the device check reports MGX_ERR_SHAPE (both reduce tiles)."""
    mgx = gpu
    K, M, N = 3, 5, 4
    cap = cap_for(entry_bits, M)
    A = rand_ints(K, M, plan_bits, 1)
    B = rand_ints(M, N, plan_bits, 2)
    A[1][2] = (1 << (entry_bits - 1)) + 12345          # one entry just over the bound
    plan = mgx.crt_plan(M, plan_bits, plan_bits)
    try:
        dA, dB = mgx.DeviceBuffer.from_array(to_entries(A, cap)), mgx.DeviceBuffer.from_array(to_entries(B, cap))
        dC = mgx.DeviceBuffer(K * N * (cap + 1) * 8)
        with pytest.raises(mgx.MgxError) as ei:
            mgx.call("mgx_crt_mfma_matmul", plan, K, M, N, dC.ptr, N, dA.ptr, M, dB.ptr, N, cap, None)
        assert ei.value.status == 2
        # the same operand on the B side: C^T = B^T A^T
        dBt = mgx.DeviceBuffer.from_array(to_entries([list(r) for r in zip(*B)], cap))
        dAt = mgx.DeviceBuffer.from_array(to_entries([list(r) for r in zip(*A)], cap))
        with pytest.raises(mgx.MgxError) as ei:
            mgx.call("mgx_crt_mfma_matmul", plan, N, M, K, dC.ptr, K, dBt.ptr, M, dAt.ptr, K, cap, None)
        assert ei.value.status == 2
        # at the bound itself it is exact
        A[1][2] = -((1 << plan_bits) - 1)
        dA2 = mgx.DeviceBuffer.from_array(to_entries(A, cap))
        mgx.call("mgx_crt_mfma_matmul", plan, K, M, N, dC.ptr, N, dA2.ptr, M, dB.ptr, N, cap, None)
        assert to_ints(dC.to_array((K, N, cap + 1))) == matmul(A, B)
    finally:
        mgx.call("mgx_crt_plan_destroy", plan)


def test_batch_understated_bits_rejected(gpu):
    mgx = gpu
    L = mgx.bind_sched()
    K, M, N, bits = 4, 6, 5, 700
    cap = cap_for(bits, M)
    A, B = rand_ints(K, M, bits, 3), rand_ints(M, N, bits, 4)
    dA, dB = mgx.DeviceBuffer.from_array(to_entries(A, cap)), mgx.DeviceBuffer.from_array(to_entries(B, cap))
    dC = mgx.DeviceBuffer(K * N * (cap + 1) * 8)
    arr = (mgx.Product * 1)(mgx.Product(K, M, N, dC.ptr, N, dA.ptr, M, dB.ptr, N, cap, 600, bits, CRT))
    b = vp()
    mgx.check(L.mgx_batch_create(ctypes.byref(b), ctypes.cast(arr, vp), 1, None))
    try:
        mgx.check(L.mgx_batch_run(b, None))
        with pytest.raises(mgx.MgxError) as ei:
            mgx.call("mgx_check_error", None)
        assert ei.value.status == 2
    finally:
        L.mgx_batch_destroy(b)
