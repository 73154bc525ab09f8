// common.h -- internal helpers shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>
#include "../../include/mgx.h"

namespace mgx {

// Thread-local error text behind mgx_last_error().
void set_error(const char *fmt, ...);
int fail(int status, const char *fmt, ...);

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

#define MGX_HIP(call)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            return ::mgx::fail(MGX_ERR_HIP, "%s: %s (%s:%d)", #call, hipGetErrorString(e_), \
                               __FILE__, __LINE__);                                        \
    } while (0)

#define MGX_TRY(expr)                     \
    do {                                  \
        int s_ = (expr);                  \
        if (s_ != MGX_OK) return s_;      \
    } while (0)

// Device-side error word: kernels set bits, the host checks after a sync.
enum : unsigned { DEV_ERR_CAPACITY = 1u, DEV_ERR_INPUT = 2u, DEV_ERR_WIDTH = 4u };
unsigned *device_error_word();          // per-device, zero-initialised at init
int check_device_error(hipStream_t s);  // synchronizes s, reads and clears the word

// Stream-ordered scratch: blocks cached per device and reused behind an
// event recorded at release (graph-owned pool memory while a stream captures).
int scratch_alloc(void **p, size_t bytes, hipStream_t s);
int scratch_free(void *p, hipStream_t s);
int scratch_release_device(int dev);   // unmaps the device's idle cached blocks

int ensure_init();

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }
inline int64_t round_up(int64_t a, int64_t b) { return ceil_div(a, b) * b; }

// Tunables (mgx_set_param).
int64_t param(int which);

// Plain route (plain.hip, rec.cpp).  toom < 0 takes MGX_PARAM_TOOM22.
int plain_mul(bool strided, bool addmul, int64_t K, int64_t M, int64_t N, uint64_t *C, int64_t Crs, const uint64_t *A,
              int64_t Ars, const uint64_t *B, int64_t Brs, int64_t cap, int64_t opwords, hipStream_t s, int toom = -1);
int plain_aors(bool strided, bool sub, int64_t Wd, int64_t H, uint64_t *C, int64_t Crs, const uint64_t *A, int64_t Ars,
               const uint64_t *B, int64_t Brs, int64_t cap, hipStream_t s);
// squarify + recursion with the leaf size, depth limit and Toom cutoff given
// per call (the global tunables are left alone)
int squarify_with(int method, int64_t K, int64_t M, int64_t N, uint64_t *C, int64_t Crs, const uint64_t *A,
                  int64_t Ars, const uint64_t *B, int64_t Brs, int64_t cap, int64_t opbits, int64_t leaf,
                  int64_t maxdepth, hipStream_t s, int toom = -1);

// A plan from the per-device cache, pinned: it is not evicted (freed) until
// every pin is released with release_plan.  The batched executor holds its
// plans pinned for its lifetime, a one-product call for the call.
// slices: the digit count the plan must use (< 0: MGX_PARAM_CRT_SLICES).
int cached_plan(int64_t K, int64_t bitsA, int64_t bitsB, const mgx_crt_plan **out, int slices = -1);
int plan_create_slices(mgx_crt_plan **out, int64_t K, int64_t bitsA, int64_t bitsB, int64_t want);
void release_plan(const mgx_crt_plan *pl);
void plan_cache_stats(int64_t *total, int64_t *pinned);
// C = A*B through a plan: stage launches on s, then one synchronization for
// the device error word.
// gemm_rows: residue-GEMM tile height for two-digit plans (128, or 64; 0 = 128).
int crt_matmul_with(const mgx_crt_plan *pl, int64_t K, int64_t M, int64_t N, mgx_limb *C, int64_t Crs,
                    const mgx_limb *A, int64_t Ars, const mgx_limb *B, int64_t Brs, int64_t cap, hipStream_t s,
                    int gemm_rows = 0);
int residue_gemm_launch(const mgx_crt_plan *pl, int64_t K, int64_t M, int64_t N, const int8_t *RA, const int8_t *RB,
                        uint32_t *Y, hipStream_t stream, int rows);


}  // namespace mgx
