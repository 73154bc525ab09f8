"""ctypes loader for liboracle.so -- test infrastructure only.

Used by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg, and
only as the checker.  Matrices are numpy uint64 arrays (rows, cols, cap+1)
in the same entry layout as the product library.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "liboracle.so")
_lib = None
c_i64, c_vp = ctypes.c_int64, ctypes.c_void_p


def build():
    subprocess.run(["make", "-s", "-C", HERE, "-j8"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(SO):
            build()
        L = ctypes.CDLL(SO)
        mm = [c_i64, c_i64, c_i64, c_vp, c_i64, c_vp, c_i64, c_vp, c_i64, c_i64]
        sigs = {
            "or_mat_mul": (mm + [ctypes.c_int], ctypes.c_int),
            "or_mat_addmul": (mm + [ctypes.c_int], ctypes.c_int),
            "or_mat_add": ([c_i64, c_i64, c_vp, c_i64, c_vp, c_i64, c_vp, c_i64, c_i64], ctypes.c_int),
            "or_mat_sub": ([c_i64, c_i64, c_vp, c_i64, c_vp, c_i64, c_vp, c_i64, c_i64], ctypes.c_int),
            "or_mat_recursive": ([ctypes.c_int, c_i64, c_vp, c_i64, c_vp, c_i64, c_vp, c_i64, c_i64, c_i64], ctypes.c_int),
            "or_mat_winograd_addmul": ([c_i64, c_vp, c_i64, c_vp, c_i64, c_vp, c_i64, c_i64, c_i64], ctypes.c_int),
            "or_mat_squarify": ([ctypes.c_int, c_i64, c_i64, c_i64, c_vp, c_i64, c_vp, c_i64, c_vp, c_i64, c_i64, c_i64],
                                ctypes.c_int),
            "or_transform_interleave": ([c_vp, c_i64, c_i64, c_i64, c_i64], None),
            "or_mpz_out_str": ([ctypes.c_char_p, c_i64, c_vp, c_i64], c_i64),
            "or_mpz_inp_str": ([c_vp, c_i64, ctypes.c_char_p], ctypes.c_int),
            "or_crt_primes": ([c_i64, c_vp], ctypes.c_int),
            "or_crt_prime_count": ([c_i64], c_i64),
            "or_crt_primes_below": ([ctypes.c_uint64, c_i64, c_vp], ctypes.c_int),
            "or_crt_prime_count_below": ([ctypes.c_uint64, c_i64], c_i64),
            "or_mod_entry": ([c_vp, c_i64, ctypes.c_uint32], ctypes.c_uint32),
            "or_balanced_slices": ([ctypes.c_int32, c_vp], None),
            "or_crt_reconstruct": ([c_vp, c_i64, c_vp, c_vp, c_i64], ctypes.c_int),
            "or_mat_mul_crt": (mm + [c_vp, c_i64, ctypes.c_int], ctypes.c_int),
            "or_is_prime_u32": ([ctypes.c_uint64], ctypes.c_int),
            "or_matcpu_gmp_version": ([], ctypes.c_char_p),
            "or_matcpu_winograd_timed": ([c_i64, c_vp, c_vp, c_vp, c_i64, c_i64, c_i64, ctypes.c_int,
                                          ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
            "or_matcpu_sample": ([c_i64, c_vp, c_vp, c_i64, c_vp, c_i64, c_vp, ctypes.c_int,
                                  ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
            "or_mlp_param_count": ([ctypes.c_int, ctypes.c_int, ctypes.c_int], c_i64),
            "or_mlp_loss_grad": ([ctypes.c_int] * 4 + [c_vp] * 6, ctypes.c_double),
            "or_mpz_add": ([c_vp, c_vp, c_vp, c_i64], ctypes.c_int),
            "or_mpz_sub": ([c_vp, c_vp, c_vp, c_i64], ctypes.c_int),
            "or_mpz_mul": ([c_vp, c_vp, c_vp, c_i64], ctypes.c_int),
            "or_mpz_addmul": ([c_vp, c_vp, c_vp, c_i64], ctypes.c_int),
            "or_mpz_submul": ([c_vp, c_vp, c_vp, c_i64], ctypes.c_int),
            "or_mpn_mul": ([c_vp, c_vp, c_i64, c_vp, c_i64], ctypes.c_uint64),
            "or_mpn_toom22_mul": ([c_vp, c_vp, c_i64, c_vp, c_i64, c_vp], None),
            "or_set_toom22_threshold": ([c_i64], None),
            "or_mpn_add_n": ([c_vp, c_vp, c_vp, c_i64], ctypes.c_uint64),
            "or_mpn_sub_n": ([c_vp, c_vp, c_vp, c_i64], ctypes.c_uint64),
            "or_mpn_addmul_1": ([c_vp, c_vp, c_i64, ctypes.c_uint64], ctypes.c_uint64),
            "or_mpn_submul_1": ([c_vp, c_vp, c_i64, ctypes.c_uint64], ctypes.c_uint64),
            "or_mpn_mul_1": ([c_vp, c_vp, c_i64, ctypes.c_uint64], ctypes.c_uint64),
        }
        for n, (a, r) in sigs.items():
            f = getattr(L, n)
            f.argtypes = a
            f.restype = r
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data


def mat_mul(A, B, cap, nthreads=8, addmul_into=None):
    K, M = A.shape[:2]
    N = B.shape[1]
    C = np.zeros((K, N, cap + 1), dtype=np.uint64) if addmul_into is None else addmul_into.copy()
    f = lib().or_mat_addmul if addmul_into is not None else lib().or_mat_mul
    st = f(K, M, N, _p(C), N, _p(A), M, _p(B), N, cap, nthreads)
    if st:
        raise OverflowError("oracle: capacity exceeded")
    return C


def mat_aors(A, B, cap, sub=False):
    H, W = A.shape[:2]
    C = np.zeros_like(A)
    f = lib().or_mat_sub if sub else lib().or_mat_add
    if f(W, H, _p(C), W, _p(A), W, _p(B), W, cap):
        raise OverflowError("oracle: capacity exceeded")
    return C


def primes(count, limit=(1 << 31) - 1):
    out = np.zeros(count, dtype=np.uint32)
    assert lib().or_crt_primes_below(limit, count, _p(out)) == 0
    return [int(x) for x in out]


def mod_matrix(X, p):
    cap = X.shape[-1] - 1
    flat = np.ascontiguousarray(X.reshape(-1, cap + 1))
    return np.array([lib().or_mod_entry(_p(flat[i]), cap, p) for i in range(flat.shape[0])],
                    dtype=np.uint64).reshape(X.shape[:2])


def slices(r):
    d = np.zeros(4, dtype=np.int8)
    lib().or_balanced_slices(int(r), _p(d))
    return [int(x) for x in d]
