"""ctypes binding of libmgx.so (the C ABI in include/mgx.h).

Host matrices are numpy uint64 arrays of shape (rows, cols, cap + 1): the
last word of each entry is the signed size (int64), the others are the
limbs, little endian -- the layout of mat-gpu's struct gpu___mpz_t in its
64-bit-limb configuration (mat-gpu/gmp-es-mpz.h:22-27, gmp-es.h:27-31).

The binding refuses to work when the library is missing: there is no
fallback arithmetic anywhere in this package.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libmgx.so")

_lib = None

c_i64 = ctypes.c_int64
c_vp = ctypes.c_void_p


class MgxError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("mgx status %d: %s" % (status, msg))
        self.status = status


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libmgx.so is not built (run __graft_entry__.build())")
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def _declare(L):
    mul_args = [c_i64, c_i64, c_i64, c_vp, c_i64, c_vp, c_i64, c_vp, c_i64, c_i64, c_vp]
    aors_args = [c_i64, c_i64, c_vp, c_i64, c_vp, c_i64, c_vp, c_i64, c_i64, c_vp]
    rec_args = [c_i64, c_vp, c_i64, c_vp, c_i64, c_vp, c_i64, c_i64, c_vp]
    sigs = {
        "mgx_init": [ctypes.c_int],
        "mgx_check_error": [c_vp],
        "mgx_malloc": [ctypes.POINTER(c_vp), ctypes.c_size_t],
        "mgx_free": [c_vp],
        "mgx_scratch_release": [],
        "mgx_memcpy_htod": [c_vp, c_vp, ctypes.c_size_t, c_vp],
        "mgx_memcpy_dtoh": [c_vp, c_vp, ctypes.c_size_t, c_vp],
        "mgx_memset": [c_vp, ctypes.c_int, ctypes.c_size_t, c_vp],
        "mgx_stream_create": [ctypes.POINTER(c_vp)],
        "mgx_stream_destroy": [c_vp],
        "mgx_stream_sync": [c_vp],
        "mgx_device_sync": [],
        "mgx_event_create": [ctypes.POINTER(c_vp)],
        "mgx_event_destroy": [c_vp],
        "mgx_event_record": [c_vp, c_vp],
        "mgx_event_elapsed": [ctypes.POINTER(ctypes.c_float), c_vp, c_vp],
        "mgx_random_matrix": [c_vp, c_i64, c_i64, c_i64, c_i64, c_i64, ctypes.c_uint64],
        "mgx_entry_to_hex": [ctypes.c_char_p, c_i64, c_vp, c_i64],
        "mgx_entry_from_hex": [c_vp, c_i64, ctypes.c_char_p],
        "mgx_launcher_mul": mul_args, "mgx_launcher_addmul": mul_args,
        "mgx_launcher_mul_strided": mul_args, "mgx_launcher_addmul_strided": mul_args,
        "mgx_launcher_add": aors_args, "mgx_launcher_sub": aors_args,
        "mgx_launcher_add_strided": aors_args, "mgx_launcher_sub_strided": aors_args,
        "mgx_add_matrices": aors_args, "mgx_sub_matrices": aors_args,
        "mgx_add_matrices_strided": aors_args, "mgx_sub_matrices_strided": aors_args,
        "mgx_interleave": [c_vp, c_i64, c_i64, c_vp],
        "mgx_deinterleave": [c_vp, c_i64, c_i64, c_vp],
        "mgx_naive_smallrect": mul_args,
        "mgx_recursive": [ctypes.c_int, ctypes.c_int] + rec_args,
        "mgx_naive_recursive": rec_args, "mgx_naive_recursive_strided": rec_args,
        "mgx_strassen": rec_args, "mgx_strassen_strided": rec_args,
        "mgx_winograd": rec_args, "mgx_winograd_strided": rec_args,
        "mgx_winograd_addmul": rec_args, "mgx_winograd_addmul_strided": rec_args,
        "mgx_squarify_mul": [ctypes.c_int] + mul_args,
        "mgx_matmul_host": [ctypes.c_int, c_i64, c_i64, c_i64, c_vp, c_vp, c_vp, c_i64,
                            ctypes.POINTER(ctypes.c_double)],
        "mgx_crt_plan_create": [ctypes.POINTER(c_vp), c_i64, c_i64, c_i64],
        "mgx_crt_plan_destroy": [c_vp],
        "mgx_crt_plan_prime_list": [c_vp, c_vp],
        "mgx_crt_plan_cache_stats": [ctypes.POINTER(c_i64), ctypes.POINTER(c_i64)],
        "mgx_crt_scan_bits": [c_i64, c_i64, c_vp, c_i64, c_i64, ctypes.POINTER(c_i64), c_vp],
        "mgx_crt_mfma_matmul": [c_vp] + mul_args,
        "mgx_crt_mfma_matmul_noalloc": [c_vp] + mul_args,
        "mgx_crt_residue_dims": [c_vp, c_i64, c_i64, ctypes.POINTER(c_i64), ctypes.POINTER(c_i64)],
        "mgx_crt_reduce": [c_vp, ctypes.c_int, c_i64, c_i64, c_vp, c_i64, c_i64, c_i64, c_vp, c_vp],
        "mgx_crt_residue_gemm": [c_vp, c_i64, c_i64, c_i64, c_vp, c_vp, c_vp, c_vp],
        "mgx_crt_reconstruct": [c_vp, c_i64, c_i64, c_vp, c_vp, c_i64, c_i64, c_vp],
        "mgx_set_param": [ctypes.c_int, c_i64],
        "mgx_get_param": [ctypes.c_int],
        # kept reference CRT API (host)
        "mgx_crt_split": [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, c_i64],
        "mgx_crt_merge": [c_vp] * 7 + [c_i64],
        "mgx_crt_make_split_tree": [ctypes.POINTER(c_vp)] * 4 + [c_vp, ctypes.c_uint64, c_i64],
        "mgx_crt_split_list": [c_vp, c_vp, c_vp, c_vp, ctypes.c_int, c_i64],
        "mgx_crt_merge_list": [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, ctypes.c_int, c_i64],
        # mgx_sched.h
        "mgx_matmul_route": [ctypes.c_int] + mul_args,
        "mgx_matmul_schedule": [ctypes.c_int] + mul_args,
        "mgx_schedule_get": [ctypes.c_int, c_vp],
        "mgx_tune_time_schedules": [c_i64, c_i64, c_i64, c_vp, c_vp, c_vp, c_i64, ctypes.c_int, c_vp,
                                    ctypes.POINTER(ctypes.c_uint)],
        "mgx_dispatch_choose_schedule": [c_i64, c_i64, c_i64, c_vp, c_i64, c_vp, c_i64, c_i64,
                                         ctypes.POINTER(ctypes.c_int)],
        "mgx_random_matrix_device": [c_vp, c_i64, c_i64, c_i64, c_i64, c_i64, c_i64, ctypes.c_double,
                                     ctypes.c_uint64, c_vp],
        "mgx_tune_features": [c_i64, c_i64, c_i64, c_vp, ctypes.c_double, c_i64, ctypes.c_double, c_vp],
        "mgx_tune_features_device": [c_i64, c_i64, c_i64, c_vp, c_i64, c_vp, c_i64, c_i64, c_vp],
        "mgx_tune_time_routes": [c_i64, c_i64, c_i64, c_vp, c_vp, c_vp, c_i64, ctypes.c_int, c_vp,
                                 ctypes.POINTER(ctypes.c_uint)],
        "mgx_costmodel_create": [ctypes.POINTER(c_vp), ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64],
        "mgx_costmodel_destroy": [c_vp],
        "mgx_costmodel_get_params": [c_vp, c_vp],
        "mgx_costmodel_set_params": [c_vp, c_vp],
        "mgx_costmodel_set_normalisation": [c_vp, c_vp, c_vp],
        "mgx_costmodel_loss_grad": [c_vp, c_vp, c_vp, c_vp, ctypes.c_int, c_vp, ctypes.POINTER(ctypes.c_float)],
        "mgx_costmodel_train": [c_vp, c_vp, c_vp, c_vp, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                ctypes.POINTER(ctypes.c_float)],
        "mgx_costmodel_predict": [c_vp, c_vp, ctypes.c_int, c_vp],
        "mgx_costmodel_load": [ctypes.POINTER(c_vp), ctypes.c_char_p],
        "mgx_dispatch_set_model": [c_vp],
        "mgx_dispatch_choose": [c_i64, c_i64, c_i64, c_vp, c_i64, c_vp, c_i64, c_i64, ctypes.POINTER(ctypes.c_int)],
        "mgx_measure_peaks": [ctypes.POINTER(ctypes.c_double)] * 3,
        "mgx_batch_create": [ctypes.POINTER(c_vp), c_vp, c_i64, c_vp],
        "mgx_batch_run": [c_vp, c_vp],
        "mgx_batch_route": [c_vp, c_i64],
        "mgx_batch_destroy": [c_vp],
        "mgx_batch_queue_info": [c_vp, c_i64, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(c_i64),
                                 ctypes.POINTER(ctypes.c_int), ctypes.POINTER(c_i64)],
        "mgx_queue_map": [c_vp, ctypes.c_int, c_i64, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(c_i64)],
        "mgx_queue_claim": [c_vp, c_i64, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)],
        "mgx_crt_split_list_device": [c_vp, c_vp, c_i64, c_vp, c_vp, ctypes.c_int, c_i64, c_vp],
        "mgx_crt_merge_list_device": [c_vp, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, ctypes.c_int, c_i64, c_vp],
    }
    for name, args in sigs.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    L.mgx_last_error.argtypes = []
    L.mgx_last_error.restype = ctypes.c_char_p
    L.mgx_version.argtypes = []
    L.mgx_version.restype = ctypes.c_char_p
    L.mgx_costmodel_param_count.argtypes = [c_vp]
    L.mgx_costmodel_param_count.restype = c_i64
    L.mgx_dispatch_last.argtypes = []
    L.mgx_costmodel_outputs.argtypes = [c_vp]
    L.mgx_costmodel_outputs.restype = c_i64
    L.mgx_batch_scratch_bytes.argtypes = [c_vp]
    L.mgx_batch_scratch_bytes.restype = c_i64
    L.mgx_batch_queue_count.argtypes = [c_vp]
    L.mgx_batch_queue_count.restype = c_i64
    L.mgx_entry_to_hex.restype = c_i64
    L.mgx_get_param.restype = c_i64
    L.mgx_crt_plan_primes.argtypes = [c_vp]
    L.mgx_crt_plan_primes.restype = c_i64
    L.mgx_crt_y_stride.argtypes = [c_i64, c_i64]
    L.mgx_crt_y_stride.restype = c_i64
    L.mgx_crt_noalloc_scratch_bytes.argtypes = [c_vp, c_i64, c_i64, c_i64]
    L.mgx_crt_noalloc_scratch_bytes.restype = c_i64
    L.mgx_crt_plan_slices.argtypes = [c_vp]
    L.mgx_crt_plan_slices.restype = c_i64
    L.mgx_queue_claim.restype = c_i64


class Product(ctypes.Structure):
    """mgx_product (include/mgx_sched.h): one product of a batch."""
    _fields_ = [("K", c_i64), ("M", c_i64), ("N", c_i64), ("C", c_vp), ("Crs", c_i64), ("A", c_vp), ("Ars", c_i64),
                ("B", c_vp), ("Brs", c_i64), ("cap", c_i64), ("bitsA", c_i64), ("bitsB", c_i64),
                ("route", ctypes.c_int)]


def bind_sched():
    """The library with every entry point of mgx_sched.h declared (all of
    them are declared when the library is first loaded)."""
    return lib()


def check(status):
    if status != 0:
        raise MgxError(status, lib().mgx_last_error().decode())
    return status


def call(name, *args):
    return check(getattr(lib(), name)(*args))


# ---------------------------------------------------------------- device memory
class DeviceBuffer:
    """A block of device memory holding a numpy array's bytes."""

    def __init__(self, nbytes):
        p = c_vp()
        call("mgx_malloc", ctypes.byref(p), max(int(nbytes), 1))
        self.ptr = p.value
        self.nbytes = int(nbytes)

    @classmethod
    def from_array(cls, a, stream=None):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        if a.nbytes:
            call("mgx_memcpy_htod", b.ptr, a.ctypes.data, a.nbytes, stream)
        return b

    def to_array(self, shape, dtype=np.uint64, stream=None):
        out = np.empty(shape, dtype=dtype)
        if out.nbytes:
            call("mgx_memcpy_dtoh", out.ctypes.data, self.ptr, out.nbytes, stream)
        return out

    def upload(self, a, stream=None):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        call("mgx_memcpy_htod", self.ptr, a.ctypes.data, a.nbytes, stream)

    def zero(self, stream=None):
        call("mgx_memset", self.ptr, 0, self.nbytes, stream)

    def free(self):
        if self.ptr:
            lib().mgx_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def init(device=0):
    call("mgx_init", device)


def sync():
    call("mgx_device_sync")


# ---------------------------------------------------------------- host entries
def random_matrix(rows, cols, cap, nbits, seed):
    a = np.zeros((rows, cols, cap + 1), dtype=np.uint64)
    call("mgx_random_matrix", a.ctypes.data, cols, rows, cols, cap, nbits, seed)
    return a


def entry_to_int(e):
    cap = e.shape[-1] - 1
    s = int(np.int64(e[cap]))
    n = abs(s)
    v = 0
    for i in range(n - 1, -1, -1):
        v = (v << 64) | int(e[i])
    return -v if s < 0 else v


def matrix_to_ints(m):
    return [[entry_to_int(m[y, x]) for x in range(m.shape[1])] for y in range(m.shape[0])]


def int_to_entry(v, cap, out=None):
    e = np.zeros(cap + 1, dtype=np.uint64) if out is None else out
    mag = abs(v)
    n = 0
    while mag:
        if n >= cap:
            raise ValueError("value needs more than %d limbs" % cap)
        e[n] = mag & 0xFFFFFFFFFFFFFFFF
        mag >>= 64
        n += 1
    for i in range(n, cap):
        e[i] = 0
    e[cap] = np.uint64((-n) & 0xFFFFFFFFFFFFFFFF) if v < 0 else np.uint64(n)
    return e


def ints_to_matrix(vals, cap):
    rows, cols = len(vals), len(vals[0]) if vals else 0
    m = np.zeros((rows, cols, cap + 1), dtype=np.uint64)
    for y in range(rows):
        for x in range(cols):
            int_to_entry(vals[y][x], cap, m[y, x])
    return m


def normalized(m):
    """True when every entry's top limb is non-zero (or the entry is zero)."""
    cap = m.shape[-1] - 1
    sz = m[..., cap].view(np.int64)
    n = np.abs(sz)
    flat = m.reshape(-1, cap + 1)
    for i, k in enumerate(n.reshape(-1)):
        if k and flat[i, k - 1] == 0:
            return False
    return True


def crt_plan(K, bitsA, bitsB):
    p = c_vp()
    call("mgx_crt_plan_create", ctypes.byref(p), K, bitsA, bitsB)
    return p.value


def crt_plan_primes(plan):
    n = lib().mgx_crt_plan_primes(plan)
    out = np.zeros(n, dtype=np.uint32)
    call("mgx_crt_plan_prime_list", plan, out.ctypes.data)
    return [int(x) for x in out]
