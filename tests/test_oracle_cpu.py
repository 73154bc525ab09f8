"""Pins the oracle (the CPU restatement) against exact mathematics: Python
integers, the golden fixtures, GMP itself and closed forms.  No GPU."""
import ctypes
import random

import numpy as np
import pytest

import oracle
from helpers import (entry_int, is_normalized, load_json, matmul, parse_hex, rand_ints, to_entries, to_ints)


def P(a):
    return a.ctypes.data


def test_golden_matmul(orc):
    for case in load_json("kat_matmul.json"):
        A = [[parse_hex(v) for v in r] for r in case["A"]]
        B = [[parse_hex(v) for v in r] for r in case["B"]]
        C = [[parse_hex(v) for v in r] for r in case["C"]]
        # the reference's addmul reserves max(|w|, |u|+|v|) + 1 limbs (mpz-aorsmul.cu:98)
        cap = (2 * case["bits"] + 8 + 63) // 64 + 2
        got = oracle.mat_mul(to_entries(A, cap), to_entries(B, cap), cap, nthreads=2)
        assert to_ints(got) == C, (case["K"], case["M"], case["N"], case["bits"], case["mode"])
        assert is_normalized(got)


@pytest.mark.parametrize("seed", range(6))
def test_mpz_ops_vs_python(orc, seed):
    rng = random.Random(seed)
    cap = 12
    for _ in range(300):
        bu, bv = rng.choice([0, 1, 63, 64, 65, 128, 300]), rng.choice([0, 1, 64, 200, 320])
        u = rng.getrandbits(bu) * rng.choice([1, -1]) if bu else 0
        v = rng.getrandbits(bv) * rng.choice([1, -1]) if bv else 0
        if rng.random() < 0.1:
            v = -u
        w0 = rng.getrandbits(300) * rng.choice([1, -1])
        U, V, W = to_entries([[u]], cap)[0, 0], to_entries([[v]], cap)[0, 0], to_entries([[w0]], cap)[0, 0]
        out = np.zeros(cap + 1, dtype=np.uint64)
        assert orc.or_mpz_add(P(out), P(U), P(V), cap) == 0 and entry_int(out) == u + v
        assert orc.or_mpz_sub(P(out), P(U), P(V), cap) == 0 and entry_int(out) == u - v
        assert orc.or_mpz_mul(P(out), P(U), P(V), cap) == 0 and entry_int(out) == u * v
        acc = W.copy()
        assert orc.or_mpz_addmul(P(acc), P(U), P(V), cap) == 0 and entry_int(acc) == w0 + u * v
        acc = W.copy()
        assert orc.or_mpz_submul(P(acc), P(U), P(V), cap) == 0 and entry_int(acc) == w0 - u * v
        assert is_normalized(acc[None, None, :])


def test_mpz_capacity_is_reported(orc):
    cap = 2
    U = to_entries([[(1 << 127) + 5]], cap)[0, 0]
    out = np.zeros(cap + 1, dtype=np.uint64)
    # mpz_add reallocs abs_usize + 1 limbs (mat-gpu/mpz-add.cu:73-74): 3 > cap
    assert orc.or_mpz_add(P(out), P(U), P(U), cap) == 1
    assert orc.or_mpz_mul(P(out), P(U), P(U), cap) == 1


@pytest.mark.parametrize("thr", [4, 8, 24])
def test_toom22_vs_schoolbook(orc, thr):
    rng = random.Random(thr)
    orc.or_set_toom22_threshold(thr)
    try:
        for an, bn in [(thr, thr), (thr + 1, thr), (2 * thr + 1, 2 * thr + 1), (3 * thr, 2 * thr + 1), (40, 33),
                       (64, 64), (65, 64), (129, 128)]:
            for pattern in ["rand", "ones", "sparse"]:
                if pattern == "ones":
                    a, b = (1 << (64 * an)) - 1, (1 << (64 * bn)) - 1
                elif pattern == "sparse":
                    a, b = (1 << (64 * an - 1)) | 1, (1 << (64 * bn - 1)) | (1 << 64)
                else:
                    a = rng.getrandbits(64 * an) | (1 << (64 * an - 1))
                    b = rng.getrandbits(64 * bn) | (1 << (64 * bn - 1))
                A = np.array([(a >> (64 * i)) & (2**64 - 1) for i in range(an)], dtype=np.uint64)
                B = np.array([(b >> (64 * i)) & (2**64 - 1) for i in range(bn)], dtype=np.uint64)
                R = np.zeros(an + bn, dtype=np.uint64)
                orc.or_mpn_mul(P(R), P(A), an, P(B), bn)
                got = sum(int(R[i]) << (64 * i) for i in range(an + bn))
                assert got == a * b, (an, bn, pattern)
    finally:
        orc.or_set_toom22_threshold(24)


def test_mpn_carry_primitives(orc):
    n = 5
    ones = np.full(n, 2**64 - 1, dtype=np.uint64)
    one = np.zeros(n, dtype=np.uint64)
    one[0] = 1
    r = np.zeros(n, dtype=np.uint64)
    assert orc.or_mpn_add_n(P(r), P(ones), P(one), n) == 1 and not r.any()   # 2^320 - 1 + 1
    assert orc.or_mpn_sub_n(P(r), P(one), P(ones), n) == 1                     # borrow out
    assert int(r[0]) == 2 and all(int(x) == 0 for x in r[1:])
    # addmul_1 / submul_1 carry-out limbs (GMP-documented return values)
    u = np.full(3, 2**64 - 1, dtype=np.uint64)
    acc = np.full(3, 2**64 - 1, dtype=np.uint64)
    cy = orc.or_mpn_addmul_1(P(acc), P(u), 3, 2**64 - 1)
    full = (2**192 - 1) + (2**192 - 1) * (2**64 - 1)
    assert int(cy) == full >> 192 and sum(int(acc[i]) << (64 * i) for i in range(3)) == full % 2**192
    acc = np.zeros(3, dtype=np.uint64)
    bw = orc.or_mpn_submul_1(P(acc), P(u), 3, 3)
    val = 0 - (2**192 - 1) * 3
    assert int(bw) == (-val + 2**192 - 1) // 2**192
    assert sum(int(acc[i]) << (64 * i) for i in range(3)) == val % 2**192


@pytest.mark.parametrize("method", [0, 1, 2])
@pytest.mark.parametrize("n,base", [(8, 2), (16, 4), (12, 3), (32, 8)])
def test_recursive_sequences(orc, method, n, base):
    cap = 6
    A, B = rand_ints(n, n, 100, 1 + n + method), rand_ints(n, n, 100, 2 + n + method)
    C = np.zeros((n, n, cap + 1), dtype=np.uint64)
    a, b = to_entries(A, cap), to_entries(B, cap)
    assert orc.or_mat_recursive(method, n, P(C), n, P(a), n, P(b), n, cap, base) == 0
    assert to_ints(C) == matmul(A, B)
    # Winograd addmul: C += A*B
    W0 = rand_ints(n, n, 150, 9)
    C = to_entries(W0, cap)
    assert orc.or_mat_winograd_addmul(n, P(C), n, P(a), n, P(b), n, cap, base) == 0
    AB = matmul(A, B)
    assert to_ints(C) == [[W0[y][x] + AB[y][x] for x in range(n)] for y in range(n)]


@pytest.mark.parametrize("K,M,N", [(5, 9, 7), (9, 9, 9), (12, 12, 12), (3, 20, 17), (24, 10, 30)])
def test_squarify(orc, K, M, N):
    cap = 5
    A, B = rand_ints(K, M, 90, K), rand_ints(M, N, 90, N)
    C = np.zeros((K, N, cap + 1), dtype=np.uint64)
    a, b = to_entries(A, cap), to_entries(B, cap)
    assert orc.or_mat_squarify(2, K, M, N, P(C), N, P(a), M, P(b), N, cap, 2) == 0
    assert to_ints(C) == matmul(A, B)


def test_interleave_closed_form(orc):
    # transform_interleave (mat-gpu/main.cu:137-156): word w of entry j -> w*npart + j
    ngroup, npart, W = 3, 256, 5
    x = np.arange(ngroup * npart * W, dtype=np.uint64)
    y = x.copy()
    orc.or_transform_interleave(P(y), ngroup, npart, W * 8, 8)
    for g in range(ngroup):
        for j in range(0, npart, 37):
            for w in range(W):
                assert y[g * npart * W + w * npart + j] == x[g * npart * W + j * W + w]
    # the inverse (npart = W words, partsz = npart*8 bytes) restores the input
    orc.or_transform_interleave(P(y), ngroup, W, npart * 8, 8)
    assert (y == x).all()


def test_text_format(orc):
    cases = [(0, "0"), (1, "1"), (-1, "-1"), (255, "ff"), (-(1 << 64), "-10000000000000000"), ((1 << 130) - 1, "3" + "f" * 32)]
    cap = 4
    for v, s in cases:
        e = to_entries([[v]], cap)[0, 0]
        buf = ctypes.create_string_buffer(200)
        n = orc.or_mpz_out_str(buf, 200, P(e), cap)
        assert buf.value.decode() == s and n == len(s)
        back = np.zeros(cap + 1, dtype=np.uint64)
        assert orc.or_mpz_inp_str(P(back), cap, s.encode()) == 0
        assert entry_int(back) == v
    bad = np.zeros(cap + 1, dtype=np.uint64)
    assert orc.or_mpz_inp_str(P(bad), cap, b"12g4") == 2


def test_crt_primes_and_counts(orc):
    kat = load_json("kat_crt.json")
    ps = kat["primes"]
    assert oracle.primes(len(ps)) == ps
    for bits, n in kat["prime_counts"].items():
        assert orc.or_crt_prime_count(int(bits)) == n
    # two-digit residues: primes below 65280
    ps16 = kat["primes16"]
    assert oracle.primes(len(ps16), limit=65279) == ps16
    for bits, n in kat["prime_counts16"].items():
        assert orc.or_crt_prime_count_below(65279, int(bits)) == n
    assert orc.or_crt_prime_count_below(65279, 200000) == -1   # not enough 16-bit primes
    # Miller-Rabin against trial division on small odd numbers
    for n in range(3, 4000, 2):
        assert bool(orc.or_is_prime_u32(n)) == all(n % d for d in range(2, int(n ** 0.5) + 1))


def test_crt_residues_and_slices(orc):
    kat = load_json("kat_crt.json")
    ps = kat["primes"][:8]
    cap = 40
    for item in kat["residues"]:
        x = parse_hex(item["x"])
        e = to_entries([[x]], cap)
        for p, r in zip(ps, item["r"]):
            assert orc.or_mod_entry(P(e[0, 0]), cap, p) == r
    for item in kat["slices"]:
        assert oracle.slices(item["r"]) == item["d"]


def test_crt_reconstruct_vs_python(orc):
    ps = oracle.primes(70)
    rng = random.Random(11)
    cap = 40
    M = 1
    for p in ps:
        M *= p
    for _ in range(50):
        x = rng.randrange(-(M // 2) + 1, M // 2)
        res = np.array([x % p for p in ps], dtype=np.uint32)
        e = np.zeros(cap + 1, dtype=np.uint64)
        pr = np.array(ps, dtype=np.uint32)
        assert orc.or_crt_reconstruct(P(e), cap, P(res), P(pr), len(ps)) == 0
        assert entry_int(e) == x
        assert is_normalized(e[None, None, :])


def test_crt_matmul_vs_naive(orc):
    K, M, N, bits, cap = 6, 9, 5, 300, 12
    A, B = rand_ints(K, M, bits, 5), rand_ints(M, N, bits, 6)
    a, b = to_entries(A, cap), to_entries(B, cap)
    need = 2 * bits + 4 + 2
    P_ = orc.or_crt_prime_count(need)
    pr = np.array(oracle.primes(P_), dtype=np.uint32)
    C = np.zeros((K, N, cap + 1), dtype=np.uint64)
    assert orc.or_mat_mul_crt(K, M, N, P(C), N, P(a), M, P(b), N, cap, P(pr), P_, 2) == 0
    assert to_ints(C) == matmul(A, B)


def test_oracle_vs_gmp(orc):
    """GMP (libgmp.so.10, the arithmetic mat-cpu links) agrees with the oracle."""
    ver = orc.or_matcpu_gmp_version()
    if ver is None:
        pytest.skip("libgmp.so.10 not loadable")
    n, bits, cap = 16, 700, 25
    A, B = rand_ints(n, n, bits, 31), rand_ints(n, n, bits, 32)
    a, b = to_entries(A, cap), to_entries(B, cap)
    C = np.zeros((n, n, cap + 1), dtype=np.uint64)
    secs = ctypes.c_double()
    # mat-cpu method 35 restated: Winograd DAG, -B 2 -J 8, 4 threads
    assert orc.or_matcpu_winograd_timed(n, P(C), P(a), P(b), cap, 2, 8, 4, ctypes.byref(secs)) == 0
    assert to_ints(C) == matmul(A, B)
    assert to_ints(oracle.mat_mul(a, b, cap)) == to_ints(C)
    yx = np.array([[3, 4], [0, 0], [15, 15]], dtype=np.int64)
    out = np.zeros((3, cap + 1), dtype=np.uint64)
    assert orc.or_matcpu_sample(n, P(a), P(b), cap, P(yx), 3, P(out), 2, ctypes.byref(secs)) == 0
    AB = matmul(A, B)
    assert [entry_int(out[i]) for i in range(3)] == [AB[3][4], AB[0][0], AB[15][15]]


def test_mlp_gradient_finite_differences(orc):
    D, H, O, B = 5, 7, 3, 4
    npar = orc.or_mlp_param_count(D, H, O)
    rng = np.random.default_rng(0)
    params = rng.normal(0, 0.5, npar)
    X, T = rng.normal(size=(B, D)), rng.normal(size=(B, O))
    mask = (rng.random((B, O)) < 0.7).astype(np.float64)
    grad, Y = np.zeros(npar), np.zeros((B, O))
    loss = orc.or_mlp_loss_grad(D, H, O, B, P(params), P(X), P(T), P(mask), P(grad), P(Y))
    # forward restated with numpy
    W1 = params[:H * D].reshape(H, D); b1 = params[H * D:H * D + H]
    o = H * D + H
    W2 = params[o:o + H * H].reshape(H, H); b2 = params[o + H * H:o + H * H + H]
    o += H * H + H
    W3 = params[o:o + O * H].reshape(O, H); b3 = params[o + O * H:]
    h1 = np.maximum(X @ W1.T + b1, 0); h2 = np.maximum(h1 @ W2.T + b2, 0); y = h2 @ W3.T + b3
    assert np.allclose(Y, y)
    assert np.isclose(loss, (((y - T) * mask) ** 2).sum() / mask.sum())
    eps = 1e-6
    for i in range(0, npar, 7):
        pp, pm = params.copy(), params.copy()
        pp[i] += eps; pm[i] -= eps
        lp = orc.or_mlp_loss_grad(D, H, O, B, P(pp), P(X), P(T), P(mask), None, P(Y))
        lm = orc.or_mlp_loss_grad(D, H, O, B, P(pm), P(X), P(T), P(mask), None, P(Y))
        assert abs((lp - lm) / (2 * eps) - grad[i]) < 1e-5 * max(1, abs(grad[i]))
