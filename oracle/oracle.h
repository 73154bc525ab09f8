/*
 * oracle.h -- CPU restatement of the reference's exact big-integer matrix
 * product, used ONLY as test infrastructure (the checker).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load liboracle.so.  The product library never links or calls it.
 *
 * Limbs are 64-bit: the reference's non-USE_32 configuration
 * (mat-gpu/gmp-es.h:27-31).  An "entry" is the reference's fixed-capacity
 * signed integer struct gpu___mpz_t (mat-gpu/gmp-es-mpz.h:22-27) with
 * cap limbs followed by a signed size word: cap+1 words of 8 bytes,
 * normalized sign-magnitude (top limb nonzero, zero <=> size 0).
 * Matrices are row-major with rowstride counted in entries
 * (mat-gpu/main.cu:26).
 *
 * Every function cites the reference file:line it restates.  The multi-modular
 * (CRT) functions have no reference counterpart: they restate the mathematics
 * (Chinese remainder theorem, balanced digit slicing, Barrett reduction).
 */
#ifndef MGX_ORACLE_H
#define MGX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef uint64_t olimb;

/* ---- mpn layer (mat-gpu/gmp-es.h:667-720 prototypes) ---- */
olimb or_mpn_add_n(olimb *rp, const olimb *up, const olimb *vp, int64_t n);
olimb or_mpn_sub_n(olimb *rp, const olimb *up, const olimb *vp, int64_t n);
olimb or_mpn_add_1(olimb *rp, const olimb *up, int64_t n, olimb v);
olimb or_mpn_sub_1(olimb *rp, const olimb *up, int64_t n, olimb v);
olimb or_mpn_add(olimb *rp, const olimb *up, int64_t un, const olimb *vp, int64_t vn);
olimb or_mpn_sub(olimb *rp, const olimb *up, int64_t un, const olimb *vp, int64_t vn);
int or_mpn_cmp(const olimb *up, const olimb *vp, int64_t n);
olimb or_mpn_mul_1(olimb *rp, const olimb *up, int64_t n, olimb v);
olimb or_mpn_addmul_1(olimb *rp, const olimb *up, int64_t n, olimb v);
olimb or_mpn_submul_1(olimb *rp, const olimb *up, int64_t n, olimb v);
void or_mpn_mul_basecase(olimb *rp, const olimb *up, int64_t un, const olimb *vp, int64_t vn);
void or_mpn_toom22_mul(olimb *pp, const olimb *ap, int64_t an, const olimb *bp, int64_t bn, olimb *scratch);
int64_t or_mpn_toom22_itch(int64_t an);
olimb or_mpn_mul(olimb *rp, const olimb *up, int64_t un, const olimb *vp, int64_t vn);
void or_set_toom22_threshold(int64_t limbs);
int64_t or_get_toom22_threshold(void);

/* ---- entry (mpz) layer; AoS entries of capacity cap ---- */
/* Status: 0 ok, 1 capacity overflow (the reference asserts: mpz-realloc.cu:5-11). */
int or_mpz_add(olimb *w, const olimb *u, const olimb *v, int64_t cap);
int or_mpz_sub(olimb *w, const olimb *u, const olimb *v, int64_t cap);
int or_mpz_mul(olimb *w, const olimb *u, const olimb *v, int64_t cap);
int or_mpz_addmul(olimb *w, const olimb *u, const olimb *v, int64_t cap);
int or_mpz_submul(olimb *w, const olimb *u, const olimb *v, int64_t cap);
void or_mpz_neg(olimb *w, const olimb *u, int64_t cap);

/* ---- matrix layer (C is K x N, A is K x M, B is M x N; mat-gpu/main.cu:242) ---- */
int or_mat_mul(int64_t K, int64_t M, int64_t N, olimb *C, int64_t Crs, const olimb *A, int64_t Ars,
               const olimb *B, int64_t Brs, int64_t cap, int nthreads);
int or_mat_addmul(int64_t K, int64_t M, int64_t N, olimb *C, int64_t Crs, const olimb *A, int64_t Ars,
                  const olimb *B, int64_t Brs, int64_t cap, int nthreads);
int or_mat_add(int64_t W, int64_t H, olimb *C, int64_t Crs, const olimb *A, int64_t Ars,
               const olimb *B, int64_t Brs, int64_t cap);
int or_mat_sub(int64_t W, int64_t H, olimb *C, int64_t Crs, const olimb *A, int64_t Ars,
               const olimb *B, int64_t Brs, int64_t cap);
/* method: 0 naive recursive, 1 Strassen (3 temporaries), 2 Winograd; base: leaf size */
int or_mat_recursive(int method, int64_t N, olimb *C, int64_t Crs, const olimb *A, int64_t Ars,
                     const olimb *B, int64_t Brs, int64_t cap, int64_t base);
int or_mat_winograd_addmul(int64_t N, olimb *C, int64_t Crs, const olimb *A, int64_t Ars,
                           const olimb *B, int64_t Brs, int64_t cap, int64_t base);
int or_mat_squarify(int method, int64_t K, int64_t M, int64_t N, olimb *C, int64_t Crs, const olimb *A,
                    int64_t Ars, const olimb *B, int64_t Brs, int64_t cap, int64_t base);

/* interleave (mat-gpu/main.cu:137-156): ngroup groups of npart parts of partsz bytes */
void or_transform_interleave(void *p, int64_t ngroup, int64_t npart, int64_t partsz, int64_t atomsz);

/* ---- text I/O (mat-gpu/mpz.cu:85-204) ---- */
int64_t or_mpz_out_str(char *buf, int64_t buflen, const olimb *e, int64_t cap);
int or_mpz_inp_str(olimb *e, int64_t cap, const char *str);

/* ---- multi-modular route (mathematics; no reference counterpart) ---- */
int or_crt_primes(int64_t count, uint32_t *out);             /* largest primes below 2^31, descending */
int or_crt_primes_below(uint64_t limit, int64_t count, uint32_t *out);  /* largest primes <= limit */
int64_t or_crt_prime_count_below(uint64_t limit, int64_t bits_needed);   /* -1: not enough primes */
int64_t or_crt_prime_count(int64_t bits_needed);              /* primes needed to exceed 2^bits_needed */
uint32_t or_mod_entry(const olimb *e, int64_t cap, uint32_t p); /* signed entry mod p, in [0,p) */
void or_balanced_slices(int32_t r, int8_t d[4]);              /* r = sum d_i 2^(8i), |r| < 2^30 */
int or_crt_reconstruct(olimb *e, int64_t cap, const uint32_t *res, const uint32_t *primes, int64_t P);
int or_mat_mul_crt(int64_t K, int64_t M, int64_t N, olimb *C, int64_t Crs, const olimb *A, int64_t Ars,
                   const olimb *B, int64_t Brs, int64_t cap, const uint32_t *primes, int64_t P, int nthreads);

int or_is_prime_u32(uint64_t n);
int or_mpz_cmp(const olimb *u, const olimb *v, int64_t cap);

/* ---- CPU baseline: mat-cpu method 35 (Winograd job DAG on a pthread pool) ---- */
/* Status 3: libgmp.so.10 could not be loaded. */
const char *or_matcpu_gmp_version(void);
int or_matcpu_winograd(int64_t N, olimb *C, const olimb *A, const olimb *B, int64_t cap,
                       int64_t minblock, int64_t jobsize, int nthreads);
int or_matcpu_winograd_timed(int64_t N, olimb *C, const olimb *A, const olimb *B, int64_t cap,
                             int64_t minblock, int64_t jobsize, int nthreads, double *seconds);
int or_matcpu_sample(int64_t N, const olimb *A, const olimb *B, int64_t cap, const int64_t *yx, int64_t cnt,
                     olimb *out, int nthreads, double *seconds);

/* ---- cost-model MLP in float64 (gradient reference for the HIP kernels) ---- */
/* Layers: h1 = relu(W1 x + b1), h2 = relu(W2 h1 + b2), y = W3 h2 + b3.
 * Parameters packed as W1[H][D] b1[H] W2[H][H] b2[H] W3[O][H] b3[O].
 * loss = sum over (b, o) with mask[b][o] != 0 of (y[b][o] - t[b][o])^2 / count. */
int64_t or_mlp_param_count(int D, int H, int O);
double or_mlp_loss_grad(int D, int H, int O, int B, const double *params, const double *X, const double *T,
                        const double *mask, double *grad, double *Y);

#ifdef __cplusplus
}
#endif
#endif
