"""The one-pass reconstruct (MGX_PARAM_RECON = 2 forces it; the default 0 takes
it for two-digit products of a workgroup per CU and at most 9 row blocks) against
the two-pass one (word buffer and carry pass, MGX_PARAM_RECON = 1) and against
Python integers.  Cases: products near their bound (the top row block fixes every
sign), products far below an explicit plan's bound and zero ones (the workgroup
sweeps further blocks for the sign), row-block counts from 1 to over 30, entry
counts that leave a partial block of 256, and the capacity check."""
import pytest

from helpers import is_normalized, matmul, rand_ints, to_entries, to_ints

pytestmark = pytest.mark.gpu

RECON = 7


def run(mgx, A, B, cap, mode, plan=None):
    K, M = A.shape[:2]
    N = B.shape[1]
    dA, dB = mgx.DeviceBuffer.from_array(A), mgx.DeviceBuffer.from_array(B)
    dC = mgx.DeviceBuffer(max(K * N, 1) * (cap + 1) * 8)
    dC.zero()
    mgx.call("mgx_set_param", RECON, mode)
    try:
        mgx.call("mgx_crt_mfma_matmul", plan, K, M, N, dC.ptr, N, dA.ptr, M, dB.ptr, N, cap, None)
        mgx.call("mgx_check_error", None)
    finally:
        mgx.call("mgx_set_param", RECON, 0)
    return dC.to_array((K, N, cap + 1))


@pytest.mark.parametrize("K,M,N,bits", [(1, 1, 1, 64), (5, 3, 7, 200), (17, 33, 19, 2048), (64, 20, 70, 700),
                                        (3, 9, 300, 4000), (9, 4, 11, 9000)])
def test_fused_matches_two_pass_and_python(gpu, K, M, N, bits):
    A, B = rand_ints(K, M, bits, 3 * K + bits), rand_ints(M, N, bits, 5 * N + bits)
    cap = (2 * bits + 12 + 63) // 64 + 1
    eA, eB = to_entries(A, cap), to_entries(B, cap)
    C0 = run(gpu, eA, eB, cap, 2)
    C1 = run(gpu, eA, eB, cap, 1)
    assert (C0 == C1).all()
    assert to_ints(C0) == matmul(A, B)
    assert is_normalized(C0)


def test_fused_gamma_words_either_side_of_a_block_boundary(gpu):
    """The one-pass reconstruct takes gamma from two extra result words when they
    fit in the last row block (plans whose word count is not a multiple of 32)
    and sums it on the vector ALUs otherwise: bit lengths around 1000 bits put
    the word count on both sides of 64."""
    for bits in range(980, 1044, 4):
        A, B = rand_ints(4, 9, bits, bits), rand_ints(9, 5, bits, bits + 1)
        cap = (2 * bits + 12 + 63) // 64 + 1
        eA, eB = to_entries(A, cap), to_entries(B, cap)
        C0 = run(gpu, eA, eB, cap, 2)
        assert to_ints(C0) == matmul(A, B), bits
        assert is_normalized(C0)


@pytest.mark.parametrize("small_bits", [0, 1, 40, 900])
def test_fused_far_below_bound(gpu, small_bits):
    # a plan for 3000-bit operands, entries of small_bits bits: |C| is far below
    # the top row block, whose words leave the sign open
    mgx = gpu
    K, M, N, cap = 20, 13, 300, 100
    plan = mgx.crt_plan(M, 3000, 3000)
    try:
        A = rand_ints(K, M, small_bits, 7) if small_bits else [[0] * M for _ in range(K)]
        B = rand_ints(M, N, small_bits, 8) if small_bits else [[0] * N for _ in range(M)]
        if small_bits:
            A[0][0], B[0][0] = -A[0][0] or -1, B[0][0] or 1   # at least one negative product term
        eA, eB = to_entries(A, cap), to_entries(B, cap)
        C0 = run(mgx, eA, eB, cap, 2, plan)
        C1 = run(mgx, eA, eB, cap, 1, plan)
        assert (C0 == C1).all()
        assert to_ints(C0) == matmul(A, B)
        assert is_normalized(C0)
    finally:
        mgx.call("mgx_crt_plan_destroy", plan)


def test_fused_mixed_magnitudes(gpu):
    # rows of full-size, tiny and zero entries side by side in one workgroup's 256 entries
    bits, cap = 1500, 50
    K, M, N = 6, 5, 90
    A = rand_ints(K, M, bits, 21)
    A[1] = [x % 1000 for x in A[1]]
    A[2] = [0] * M
    A[3] = [-x for x in A[3]]
    B = rand_ints(M, N, bits, 22)
    eA, eB = to_entries(A, cap), to_entries(B, cap)
    C0 = run(gpu, eA, eB, cap, 2)
    assert (C0 == run(gpu, eA, eB, cap, 1)).all()
    assert to_ints(C0) == matmul(A, B)


def test_fused_capacity_error(gpu):
    mgx = gpu
    A = [[(1 << 300) - 1]]
    cap = 6   # the product needs 10 limbs
    for mode in (2, 1):
        with pytest.raises(mgx.MgxError) as ei:
            run(mgx, to_entries(A, cap), to_entries(A, cap), cap, mode)
        assert ei.value.status == 1
