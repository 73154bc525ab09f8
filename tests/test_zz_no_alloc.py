"""The CRT product's no-allocation entry point against the allocating one:
bit-identical results on two calls in a row (the second smaller, so it runs
over what the first left in the shared scratch array), the empty inner
dimension, and the size limit."""
import ctypes

import numpy as np
import pytest

from helpers import is_normalized, matmul, rand_ints, to_entries, to_ints

pytestmark = pytest.mark.gpu


def same_entries(X, Y, cap):
    sx, sy = X[..., cap].view(np.int64), Y[..., cap].view(np.int64)
    if not (sx == sy).all():
        return False
    mask = np.arange(cap)[None, None, :] < np.abs(sx)[..., None]
    return bool((np.where(mask, X[..., :cap], 0) == np.where(mask, Y[..., :cap], 0)).all())


@pytest.mark.parametrize("slices", [2, 4])
def test_noalloc_matches_allocating_entry(gpu, slices):
    mgx = gpu
    old = mgx.lib().mgx_get_param(3)
    mgx.call("mgx_set_param", 3, slices)
    stream = ctypes.c_void_p()
    mgx.call("mgx_stream_create", ctypes.byref(stream))
    s = stream.value
    try:
        bits = 1000
        plan = mgx.crt_plan(200, bits, bits)
        try:
            assert mgx.lib().mgx_crt_plan_slices(plan) == slices
            for step, (K, M, N, seed) in enumerate([(160, 200, 150, 1), (40, 60, 30, 2)]):
                cap = (2 * bits + 8 + 2 + 63) // 64 + 1
                need = mgx.lib().mgx_crt_noalloc_scratch_bytes(plan, K, M, N)
                assert 0 < need <= 128 << 20
                A, B = rand_ints(K, M, bits, seed), rand_ints(M, N, bits, seed + 10)
                dA, dB = mgx.DeviceBuffer.from_array(to_entries(A, cap)), mgx.DeviceBuffer.from_array(to_entries(B, cap))
                dC, dR = (mgx.DeviceBuffer(K * N * (cap + 1) * 8) for _ in range(2))
                mgx.call("mgx_crt_mfma_matmul", plan, K, M, N, dR.ptr, N, dA.ptr, M, dB.ptr, N, cap, s)
                mgx.call("mgx_crt_mfma_matmul_noalloc", plan, K, M, N, dC.ptr, N, dA.ptr, M, dB.ptr, N, cap, s)
                mgx.call("mgx_check_error", s)
                got, ref = dC.to_array((K, N, cap + 1)), dR.to_array((K, N, cap + 1))
                assert same_entries(got, ref, cap), step
                assert is_normalized(got)
                if step == 1:
                    assert to_ints(got) == matmul(A, B)
        finally:
            mgx.call("mgx_crt_plan_destroy", plan)
    finally:
        mgx.call("mgx_stream_destroy", s)
        mgx.call("mgx_set_param", 3, old)


def test_noalloc_empty_inner_dimension_and_limit(gpu):
    mgx = gpu
    cap = 5
    plan = mgx.crt_plan(4096, 2048, 2048)
    try:
        K, N = 3, 4
        dC = mgx.DeviceBuffer.from_array(to_entries(rand_ints(K, N, 200, 3), cap))
        dummy = mgx.DeviceBuffer(8)
        mgx.call("mgx_crt_mfma_matmul_noalloc", plan, K, 0, N, dC.ptr, N, dummy.ptr, 0, dummy.ptr, N, cap, None)
        mgx.call("mgx_check_error", None)
        assert to_ints(dC.to_array((K, N, cap + 1))) == [[0] * N for _ in range(K)]
        # n = 1024 with 2048-bit entries needs about 1.6 GB of scratch: refused before any launch
        n = 1024
        assert mgx.lib().mgx_crt_noalloc_scratch_bytes(plan, n, n, n) > 128 << 20
        with pytest.raises(mgx.MgxError) as ei:
            mgx.call("mgx_crt_mfma_matmul_noalloc", plan, n, n, n, dummy.ptr, n, dummy.ptr, n, dummy.ptr, n, 66, None)
        assert ei.value.status == 2
        with pytest.raises(mgx.MgxError) as ei:
            mgx.call("mgx_crt_mfma_matmul_noalloc", None, 1, 1, 1, dummy.ptr, 1, dummy.ptr, 1, dummy.ptr, 1, 1, None)
        assert ei.value.status == 5
    finally:
        mgx.call("mgx_crt_plan_destroy", plan)
