/*
 * mgx.h -- C ABI of the MI355X exact big-integer matrix multiplier.
 *
 * Drop-in for the GPU path of mat-gpu (tomsmeding/bachelor-thesis).  Every
 * entry point takes plain pointers and sizes; device pointers are HIP device
 * memory, streams are hipStream_t passed as void* (NULL = the null stream).
 *
 * ENTRY TYPE.  An entry is the reference's fixed-capacity signed integer
 * struct gpu___mpz_t {limb ptr[cap]; mp_size_t size;}
 * (mat-gpu/gmp-es-mpz.h:22-27) in its 64-bit-limb configuration
 * (mat-gpu/gmp-es.h:27-31): cap little-endian uint64_t limbs followed by one
 * int64_t size word, |size| = number of limbs in use, sign of size = sign of
 * the value, zero <=> size 0.  cap is a run-time argument here (the
 * reference fixes it at compile time, gmp-es-mpz.h:18); one call uses one cap
 * for all its matrices.  Outputs are always normalized (top limb non-zero);
 * inputs may carry zero top limbs.
 *
 * MATRICES.  Row-major, rowstride counted in entries (mat-gpu/main.cu:26).
 * C is K x N, A is K x M, B is M x N (mat-gpu/main.cu:242-245).
 *
 * INTERLEAVED ("strided") LAYOUT.  Within each block of MGX_INTERLEAVE_BLOCK
 * consecutive entries of a row, word w of entry j is stored at word position
 * w * MGX_INTERLEAVE_BLOCK + j (mat-gpu/main.cu:137-156,
 * mat-gpu/mat-params.h:4); rowstrides must be multiples of the block.
 *
 * ERRORS.  Every function returns an mgx_status; mgx_last_error() gives a
 * message.  The reference aborts instead (cusuc.h:7-15, mpz-realloc.cu:5-11).
 *
 * THREADING.  Calls are stream-ordered and asynchronous unless stated; a
 * call that must return a host value synchronizes its stream.  The library
 * is thread-compatible: concurrent calls from several host threads are safe
 * when they use different streams and different matrices.
 */
#ifndef MGX_H
#define MGX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MGX_INTERLEAVE_BLOCK 256

typedef uint64_t mgx_limb;

typedef enum {
    MGX_OK = 0,
    MGX_ERR_CAPACITY = 1,    /* a value needs more than cap limbs */
    MGX_ERR_SHAPE = 2,       /* sizes, strides or alignment not supported */
    MGX_ERR_HIP = 3,         /* a HIP runtime call failed */
    MGX_ERR_OOM = 4,         /* device allocation failed */
    MGX_ERR_ARG = 5,         /* invalid argument (null pointer, bad enum) */
    MGX_ERR_UNSUPPORTED = 6, /* not built for this device or configuration */
    MGX_ERR_PARSE = 7        /* malformed text input */
} mgx_status;

/* ---------------- library ---------------- */
int mgx_init(int device);
const char *mgx_last_error(void);
const char *mgx_version(void);
/* Synchronizes the stream and reports (then clears) errors that kernels
 * flagged on the device since the last check: MGX_ERR_CAPACITY when an
 * output needed more than cap limbs.  The launchers are asynchronous and
 * report through this; matrix-level calls check it themselves. */
int mgx_check_error(void *stream);

/* Device memory and stream helpers (so that C, ctypes or cgo callers need no
 * other HIP binding). */
int mgx_malloc(void **p, size_t bytes);
int mgx_free(void *p);
/* The library keeps the scratch blocks of its products (residues, Y planes,
 * recursion temporaries) mapped between calls; this unmaps the current
 * device's idle ones.  Housekeeping with no counterpart in the reference,
 * which allocates per call (mat-gpu/mat-gpu.cu:98). */
int mgx_scratch_release(void);
int mgx_memcpy_htod(void *dst, const void *src, size_t bytes, void *stream);
int mgx_memcpy_dtoh(void *dst, const void *src, size_t bytes, void *stream);
int mgx_memset(void *dst, int value, size_t bytes, void *stream);
int mgx_stream_create(void **stream);
int mgx_stream_destroy(void *stream);
int mgx_stream_sync(void *stream);
int mgx_device_sync(void);
/* Elapsed milliseconds between two events recorded on a stream. */
int mgx_event_create(void **ev);
int mgx_event_destroy(void *ev);
int mgx_event_record(void *ev, void *stream);
int mgx_event_elapsed(float *ms, void *start, void *stop);

/* ---------------- host utilities (no device) ---------------- */
/* Seeded entries: magnitude of nbits uniform bits (mpz_urandomb semantics,
 * mat-cpu/mat.c:40) and a uniform sign (mat-cpu/mat.c:41), from a
 * SplitMix64-seeded xoshiro256** stream shared by every path. */
int mgx_random_matrix(mgx_limb *host, int64_t W, int64_t H, int64_t rowstride, int64_t cap,
                      int64_t nbits, uint64_t seed);
/* Hex text in mat-gpu's format (mat-gpu/mpz.cu:85-204). */
int64_t mgx_entry_to_hex(char *buf, int64_t buflen, const mgx_limb *entry, int64_t cap);
int mgx_entry_from_hex(mgx_limb *entry, int64_t cap, const char *text);

/* ---------------- launchers (mat-gpu/mat-gpu-launchers.h:7-61) ---------------- */
/* C (=|+=) A*B per output entry.  Replaces launcher_mul / launcher_addmul
 * (mat-gpu-launchers.cu:5-63). */
int mgx_launcher_mul(int64_t K, int64_t M, int64_t N, mgx_limb *C, int64_t Crs, const mgx_limb *A, int64_t Ars,
                     const mgx_limb *B, int64_t Brs, int64_t cap, void *stream);
int mgx_launcher_addmul(int64_t K, int64_t M, int64_t N, mgx_limb *C, int64_t Crs, const mgx_limb *A,
                        int64_t Ars, const mgx_limb *B, int64_t Brs, int64_t cap, void *stream);
int mgx_launcher_mul_strided(int64_t K, int64_t M, int64_t N, mgx_limb *C, int64_t Crs, const mgx_limb *A,
                             int64_t Ars, const mgx_limb *B, int64_t Brs, int64_t cap, void *stream);
int mgx_launcher_addmul_strided(int64_t K, int64_t M, int64_t N, mgx_limb *C, int64_t Crs, const mgx_limb *A,
                                int64_t Ars, const mgx_limb *B, int64_t Brs, int64_t cap, void *stream);
/* C = A +- B elementwise, W x H.  Replaces launcher_add / launcher_sub
 * (mat-gpu-launchers.cu:65-147). */
int mgx_launcher_add(int64_t W, int64_t H, mgx_limb *C, int64_t Crs, const mgx_limb *A, int64_t Ars,
                     const mgx_limb *B, int64_t Brs, int64_t cap, void *stream);
int mgx_launcher_sub(int64_t W, int64_t H, mgx_limb *C, int64_t Crs, const mgx_limb *A, int64_t Ars,
                     const mgx_limb *B, int64_t Brs, int64_t cap, void *stream);
int mgx_launcher_add_strided(int64_t W, int64_t H, mgx_limb *C, int64_t Crs, const mgx_limb *A, int64_t Ars,
                             const mgx_limb *B, int64_t Brs, int64_t cap, void *stream);
int mgx_launcher_sub_strided(int64_t W, int64_t H, mgx_limb *C, int64_t Crs, const mgx_limb *A, int64_t Ars,
                             const mgx_limb *B, int64_t Brs, int64_t cap, void *stream);
/* Interleave / de-interleave whole blocks of MGX_INTERLEAVE_BLOCK entries in
 * place on the device (mat-gpu/main.cu:137-156 with atomsz = 8). */
int mgx_interleave(mgx_limb *M, int64_t nentries, int64_t cap, void *stream);
int mgx_deinterleave(mgx_limb *M, int64_t nentries, int64_t cap, void *stream);

/* ---------------- matrix-level plain route (mat-gpu/mat-gpu.h:6-84) ---------------- */
int mgx_add_matrices(int64_t W, int64_t H, mgx_limb *C, int64_t Crs, const mgx_limb *A, int64_t Ars,
                     const mgx_limb *B, int64_t Brs, int64_t cap, void *stream);
int mgx_sub_matrices(int64_t W, int64_t H, mgx_limb *C, int64_t Crs, const mgx_limb *A, int64_t Ars,
                     const mgx_limb *B, int64_t Brs, int64_t cap, void *stream);
int mgx_add_matrices_strided(int64_t W, int64_t H, mgx_limb *C, int64_t Crs, const mgx_limb *A, int64_t Ars,
                             const mgx_limb *B, int64_t Brs, int64_t cap, void *stream);
int mgx_sub_matrices_strided(int64_t W, int64_t H, mgx_limb *C, int64_t Crs, const mgx_limb *A, int64_t Ars,
                             const mgx_limb *B, int64_t Brs, int64_t cap, void *stream);
int mgx_naive_smallrect(int64_t K, int64_t M, int64_t N, mgx_limb *C, int64_t Crs, const mgx_limb *A,
                        int64_t Ars, const mgx_limb *B, int64_t Brs, int64_t cap, void *stream);

/* Square N x N recursion over 2x2 blocks down to the leaf size (the
 * reference's BASE_SIZE = 256, mat-gpu/mat-params.h:5; here a tunable,
 * mgx_set_param(MGX_PARAM_LEAF, ...)).  Temporaries come from the
 * library's stream-ordered scratch cache; no host synchronization inside. */
typedef enum {
    MGX_REC_NAIVE = 0,          /* naive_recursive_gpu (mat-gpu.cu:79-164) */
    MGX_REC_STRASSEN = 1,       /* strassen_gpu (mat-gpu.cu:166-265) */
    MGX_REC_WINOGRAD = 2,       /* winograd_gpu (mat-gpu.cu:267-315,367-420) */
    MGX_REC_WINOGRAD_ADDMUL = 3 /* winograd_addmul_gpu (mat-gpu.cu:317-365,422-475) */
} mgx_rec_method;
int mgx_recursive(int method, int strided, int64_t N, mgx_limb *C, int64_t Crs, const mgx_limb *A, int64_t Ars,
                  const mgx_limb *B, int64_t Brs, int64_t cap, void *stream);
int mgx_naive_recursive(int64_t N, mgx_limb *C, int64_t Crs, const mgx_limb *A, int64_t Ars, const mgx_limb *B,
                        int64_t Brs, int64_t cap, void *stream);
int mgx_naive_recursive_strided(int64_t N, mgx_limb *C, int64_t Crs, const mgx_limb *A, int64_t Ars,
                                const mgx_limb *B, int64_t Brs, int64_t cap, void *stream);
int mgx_strassen(int64_t N, mgx_limb *C, int64_t Crs, const mgx_limb *A, int64_t Ars, const mgx_limb *B,
                 int64_t Brs, int64_t cap, void *stream);
int mgx_strassen_strided(int64_t N, mgx_limb *C, int64_t Crs, const mgx_limb *A, int64_t Ars, const mgx_limb *B,
                         int64_t Brs, int64_t cap, void *stream);
int mgx_winograd(int64_t N, mgx_limb *C, int64_t Crs, const mgx_limb *A, int64_t Ars, const mgx_limb *B,
                 int64_t Brs, int64_t cap, void *stream);
int mgx_winograd_strided(int64_t N, mgx_limb *C, int64_t Crs, const mgx_limb *A, int64_t Ars, const mgx_limb *B,
                         int64_t Brs, int64_t cap, void *stream);
int mgx_winograd_addmul(int64_t N, mgx_limb *C, int64_t Crs, const mgx_limb *A, int64_t Ars, const mgx_limb *B,
                        int64_t Brs, int64_t cap, void *stream);
int mgx_winograd_addmul_strided(int64_t N, mgx_limb *C, int64_t Crs, const mgx_limb *A, int64_t Ars,
                                const mgx_limb *B, int64_t Brs, int64_t cap, void *stream);
/* Any K x M x N: square power-of-two-times-small blocks go to the recursion
 * `method`, the rest is split as squarify_mul_pregpu does
 * (mat-gpu/mat-pregpu.cu:469-576). */
int mgx_squarify_mul(int method, int64_t K, int64_t M, int64_t N, mgx_limb *C, int64_t Crs, const mgx_limb *A,
                     int64_t Ars, const mgx_limb *B, int64_t Brs, int64_t cap, void *stream);

/* Host-memory entry: copies A and B to the device, multiplies with the
 * mat-gpu method code (1, 2, 3, 35, 37, 357, 4, 47 as in
 * mat-gpu/main.cu:358-502; 8 = CRT-MFMA; 0 = cost-model choice), copies C
 * back.  Replaces the *_pregpu layer (mat-gpu/mat-pregpu.h:6-63).
 * Out of core (MGX_PARAM_TRANSFER = T > 0 and a side above T): square
 * products of methods 3/35/4/37/357/47 recurse over 2 x 2 blocks in host
 * memory with the method's own sequence until a block side is at most T
 * (as transfer_recurse, mat-gpu/mat-pregpu.cu:10-96), block additions run
 * on the device over staged copies; other shapes and methods are tiled in
 * T x T blocks whose products accumulate on the device. */
int mgx_matmul_host(int method, int64_t K, int64_t M, int64_t N, mgx_limb *C, const mgx_limb *A,
                    const mgx_limb *B, int64_t cap, double *seconds);

/* ---------------- multi-modular CRT route on INT8 MFMA (new; no reference counterpart) ---------------- */
typedef struct mgx_crt_plan mgx_crt_plan;
/* Plan for products with inner dimension K and |A| < 2^bitsA, |B| < 2^bitsB:
 * the smallest prefix of the primes below 65280 (S = 2 digits per residue;
 * the default while the primes above 16384 suffice) or below 2^31 (S = 4),
 * descending, whose product exceeds 2^(bitsA + bitsB + ceil(log2 K) + 2),
 * with all per-prime constants resident on the device (MGX_PARAM_CRT_SLICES
 * forces S).  Synchronous. */
int mgx_crt_plan_create(mgx_crt_plan **plan, int64_t K, int64_t bitsA, int64_t bitsB);
/* Largest bit length of |x| over a rows x cols matrix (synchronizes). */
int mgx_crt_scan_bits(int64_t rows, int64_t cols, const mgx_limb *X, int64_t Xrs, int64_t cap, int64_t *bits,
                      void *stream);
int mgx_crt_plan_destroy(mgx_crt_plan *plan);
int64_t mgx_crt_plan_primes(const mgx_crt_plan *plan);
int64_t mgx_crt_plan_slices(const mgx_crt_plan *plan);   /* S: 2 (primes < 65280) or 4 (primes < 2^31) */
int mgx_crt_plan_prime_list(const mgx_crt_plan *plan, uint32_t *out);
/* C = A*B through the plan (plan NULL: scan the inputs' limb counts on the
 * device, which synchronizes the stream once, and plan internally for
 * |A| < 2^(64 max|size_A|), |B| < 2^(64 max|size_B|)).  With a
 * caller's plan, M must not exceed its K and every |A| < 2^bitsA, |B| < 2^bitsB
 * (checked on the device: MGX_ERR_SHAPE otherwise, C undefined). */
int mgx_crt_mfma_matmul(const mgx_crt_plan *plan, int64_t K, int64_t M, int64_t N, mgx_limb *C, int64_t Crs,
                        const mgx_limb *A, int64_t Ars, const mgx_limb *B, int64_t Brs, int64_t cap,
                        void *stream);
/* The same product with no allocation of any kind, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  Its scratch is one
 * fixed device array per GPU of MGX_NOALLOC_SCRATCH_MB MiB, held from
 * library load until the process ends; products needing more scratch than
 * that (mgx_crt_noalloc_scratch_bytes) fail with MGX_ERR_SHAPE.  plan must be
 * a caller's plan (mgx_crt_plan_create) and mgx_init must have run on the
 * device.  Fully asynchronous: it only enqueues kernels on the stream and
 * does not synchronize; read errors with mgx_check_error.  Every call on one
 * GPU uses that GPU's one scratch array: calls that overlap on a GPU (other
 * streams, other threads, an earlier call still running) race and corrupt
 * each other's results; order them on one stream. */
#define MGX_NOALLOC_SCRATCH_MB 128
int mgx_crt_mfma_matmul_noalloc(const mgx_crt_plan *plan, int64_t K, int64_t M, int64_t N, mgx_limb *C, int64_t Crs,
                                const mgx_limb *A, int64_t Ars, const mgx_limb *B, int64_t Brs, int64_t cap,
                                void *stream);
/* Scratch bytes a product needs on the CRT route (-1 on bad arguments). */
int64_t mgx_crt_noalloc_scratch_bytes(const mgx_crt_plan *plan, int64_t K, int64_t M, int64_t N);
/* Stage kernels, exposed for testing and profiling.  Residues are planar
 * int8 slices R[q][s] (s < S) of rpad x kpad bytes (mgx_crt_residue_dims):
 * row-major for S = 4; for S = 2 in blocks of 16 rows x 64 columns (1 KB,
 * row-major inside), the blocks of each 16-row band in column order, i.e.
 * (row, k) at (row/16)*16*kpad + (k/64)*1024 + (row%16)*64 + k%64.  Y holds y_q = (C_ij mod p_q) (M/p_q)^-1 mod
 * p_q, centred and written as S balanced int8 digits, at entry i*N+j of
 * uint32 word planes of ystride = mgx_crt_y_stride(K, N) words (K*N rounded
 * up to 256): plane q for S = 4 (byte s = digit s); for S = 2 plane q/2,
 * low half for even q, high half for odd q (ceil(P/2) planes).
 * mgx_crt_reconstruct takes its own stream-ordered scratch (about
 * 8 * ystride bytes per 32-bit word of M). */
int mgx_crt_residue_dims(const mgx_crt_plan *plan, int64_t rows, int64_t K, int64_t *rpad, int64_t *kpad);
/* mgx_crt_reduce: transpose 0 reduces the K-major operand A (rows x K, bound
 * bitsA), transpose 1 the operand B stored K x rows (bound bitsB); an entry at
 * or above its operand's bound sets the device error word (MGX_ERR_SHAPE at
 * the next check).  nbytes is reserved and must be 0. */
int mgx_crt_reduce(const mgx_crt_plan *plan, int transpose, int64_t rows, int64_t K, const mgx_limb *X,
                   int64_t Xrs, int64_t cap, int64_t nbytes, int8_t *R, void *stream);
int mgx_crt_residue_gemm(const mgx_crt_plan *plan, int64_t K, int64_t M, int64_t N, const int8_t *RA,
                         const int8_t *RB, uint32_t *Y, void *stream);
int mgx_crt_reconstruct(const mgx_crt_plan *plan, int64_t K, int64_t N, const uint32_t *Y, mgx_limb *C,
                        int64_t Crs, int64_t cap, void *stream);
int64_t mgx_crt_y_stride(int64_t K, int64_t N);
/* The library's per-device plan cache (plans made for mgx_crt_mfma_matmul
 * with plan NULL and for batches): number of cached plans and how many are
 * pinned by a call or a live batch.  Pinned plans are never evicted; at most
 * 256 unpinned ones are kept.  Housekeeping with no reference counterpart. */
int mgx_crt_plan_cache_stats(int64_t *total, int64_t *pinned);

/* ---------------- kept reference CRT API (mat-gpu/gpu_crt.h:6-14; host, and a device batch form) ---------------- */
/* Same arguments and meaning as gpu_crt_* on entries of capacity cap, using
 * this library's own arithmetic (no GMP).  make_split_tree allocates the
 * four arrays (cap+1 words per entry) with malloc; the caller frees them
 * (mat-gpu/gpu_crt.cu:90-120).  Returns the tree length, or -status. */
int mgx_crt_split(mgx_limb *m, mgx_limb *n, mgx_limb *mf, mgx_limb *nf, const mgx_limb *x,
                  const mgx_limb *const *list, int64_t listlen, int64_t cap);
int mgx_crt_merge(mgx_limb *x, const mgx_limb *a, const mgx_limb *b, const mgx_limb *m, const mgx_limb *n,
                  const mgx_limb *mf, const mgx_limb *nf, int64_t cap);
int mgx_crt_make_split_tree(mgx_limb **ms, mgx_limb **ns, mgx_limb **mfs, mgx_limb **nfs, const mgx_limb *x,
                            uint64_t limit, int64_t cap);
int mgx_crt_split_list(mgx_limb *res, const mgx_limb *x, const mgx_limb *ms, const mgx_limb *ns, int length,
                       int64_t cap);
int mgx_crt_merge_list(mgx_limb *x, const mgx_limb *split, const mgx_limb *ms, const mgx_limb *ns,
                       const mgx_limb *mfs, const mgx_limb *nfs, int length, int64_t cap);
/* gpu_crt_split_list / gpu_crt_merge_list (mat-gpu/gpu_crt.h:10-12, crt.cpp:126-164) on the device, for
 * count entries at once through one shared tree: X holds count entries, res (split) count * 2^length
 * leaves (entry e's leaves at e * 2^length .. , in the reference's order); the tree arrays (length
 * entries each, as from mgx_crt_make_split_tree) are device copies.  All cap+1 words per entry; device
 * pointers; synchronous (MGX_ERR_CAPACITY when a value exceeds cap). */
int mgx_crt_split_list_device(mgx_limb *res, const mgx_limb *X, int64_t count, const mgx_limb *ms,
                              const mgx_limb *ns, int length, int64_t cap, void *stream);
int mgx_crt_merge_list_device(mgx_limb *X, const mgx_limb *split, int64_t count, const mgx_limb *ms,
                              const mgx_limb *ns, const mgx_limb *mfs, const mgx_limb *nfs, int length, int64_t cap,
                              void *stream);

/* ---------------- tunables ---------------- */
typedef enum {
    MGX_PARAM_LEAF = 0,          /* recursion leaf size (reference BASE_SIZE = 256) */
    MGX_PARAM_TOOM22 = 1,        /* Toom-22 block cutoff in 64-limb blocks (0 = never) */
    MGX_PARAM_WINOGRAD_DEPTH = 2, /* max recursion levels for the plain route (-1 = to the leaf) */
    MGX_PARAM_CRT_SLICES = 3,     /* CRT digits per residue for new plans: 0 = 2 when primes below 65280
                                     suffice, else 4 (default); 2 or 4 to force */
    MGX_PARAM_TRANSFER = 4,       /* mgx_matmul_host: largest block side moved to the device at once
                                     (reference TRANSFER_SIZE = 512, mat-gpu/mat-params.h:6); 0 = the whole
                                     product (default) */
    MGX_PARAM_GEMM_ROWS = 5,      /* residue-GEMM tile height of two-digit plans: 0 or 128 (default), or 64
                                     (twice the workgroups, for small products) */
    MGX_PARAM_REDUCE = 6,         /* reduce kernels of two-digit plans: 0 (default) = entries of at most
                                     256 bytes with the whole constant table in LDS (one workgroup per CU)
                                     whenever it fits, entries over 512 bytes as a tiled MFMA GEMM;
                                     1 = entry-resident tiles of 128 entries, and 64-row tiles past 512 bytes */
    MGX_PARAM_RECON = 7           /* reconstruct of two-digit plans: 0 (default) = one pass writing C's limbs
                                     when the product has at most 9 row blocks of 128 result bytes (products up
                                     to ~4600 bits) and, in single-product calls, a block of 256 entries per CU,
                                     else two; 1 = always a word-buffer GEMM pass and a carry pass; 2 = always
                                     one pass (batches read it when they are created) */
} mgx_param;
int mgx_set_param(int param, int64_t value);
int64_t mgx_get_param(int param);

#ifdef __cplusplus
}
#endif
#endif
