/*
 * crt.c -- oracle (test infrastructure only): the multi-modular route.
 *
 * The reference has no word-size-prime CRT (its crt.cpp splits one big modulus
 * into a tree and is never called), so these functions restate mathematics:
 * deterministic Miller-Rabin, residues of signed entries, balanced base-256
 * digits, and Garner's mixed-radix reconstruction (Knuth, TAOCP vol. 2,
 * 4.3.2).  Garner is deliberately a different algorithm from the explicit-CRT
 * sum the GPU reconstruct kernel uses.
 */
#include "oracle.h"
#include <stdlib.h>
#include <string.h>
#include <pthread.h>

typedef unsigned __int128 u128;

static uint64_t mulmod(uint64_t a, uint64_t b, uint64_t m) { return (uint64_t)((u128)a * b % m); }
static uint64_t powmod(uint64_t a, uint64_t e, uint64_t m)
{
    uint64_t r = 1 % m;
    a %= m;
    while (e) { if (e & 1) r = mulmod(r, a, m); a = mulmod(a, a, m); e >>= 1; }
    return r;
}

/* Miller-Rabin with bases 2, 7, 61: exact for n < 4,759,123,141. */
int or_is_prime_u32(uint64_t n)
{
    if (n < 2) return 0;
    static const uint64_t small[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61};
    for (unsigned i = 0; i < sizeof small / sizeof small[0]; i++) {
        if (n == small[i]) return 1;
        if (n % small[i] == 0) return 0;
    }
    uint64_t d = n - 1; int s = 0;
    while (!(d & 1)) { d >>= 1; s++; }
    static const uint64_t bases[] = {2, 7, 61};
    for (int i = 0; i < 3; i++) {
        uint64_t x = powmod(bases[i], d, n);
        if (x == 1 || x == n - 1) continue;
        int comp = 1;
        for (int r = 1; r < s; r++) { x = mulmod(x, x, n); if (x == n - 1) { comp = 0; break; } }
        if (comp) return 0;
    }
    return 1;
}

/* The count largest primes <= limit, descending (crt.hip primes_for_bits:
 * limit 2^31 - 1 for four digits per residue, 65279 for two). */
int or_crt_primes_below(uint64_t limit, int64_t count, uint32_t *out)
{
    uint64_t c = limit;
    for (int64_t i = 0; i < count; c--) {
        if (c < 3) return 1;
        if (or_is_prime_u32(c)) out[i++] = (uint32_t)c;
    }
    return 0;
}

int or_crt_primes(int64_t count, uint32_t *out) { return or_crt_primes_below((1ull << 31) - 1, count, out); }

/* Smallest P with prod(first P primes <= limit) > 2^bits_needed, by exact
 * big-integer comparison (a limb vector product); -1 if the primes above 2
 * run out first. */
int64_t or_crt_prime_count_below(uint64_t limit, int64_t bits_needed)
{
    int64_t cap = bits_needed / 64 + 4;
    olimb *m = (olimb *)calloc((size_t)cap + 64, sizeof(olimb));
    int64_t n = 1; m[0] = 1;
    int64_t P = 0;
    uint64_t c = limit;
    for (;;) {
        /* m > 2^bits_needed <=> bit length of m > bits_needed, or == and m is not a power of two */
        int64_t top = n - 1;
        while (top > 0 && m[top] == 0) top--;
        int64_t bl = top * 64 + (64 - __builtin_clzll(m[top]));
        if (bl - 1 > bits_needed) break;
        if (bl - 1 == bits_needed) {
            int pow2 = 1;
            for (int64_t i = 0; i < top; i++) if (m[i]) pow2 = 0;
            if (m[top] & (m[top] - 1)) pow2 = 0;
            if (!pow2) break;
        }
        while (c >= 3 && !or_is_prime_u32(c)) c--;
        if (c < 3) { P = -1; break; }
        m[n] = or_mpn_mul_1(m, m, n, c);
        n++;
        while (n > 1 && m[n - 1] == 0) n--;
        c--;
        P++;
    }
    free(m);
    return P;
}

int64_t or_crt_prime_count(int64_t bits_needed) { return or_crt_prime_count_below((1ull << 31) - 1, bits_needed); }

/* Signed entry modulo p, in [0, p): Horner over the limbs, high to low. */
uint32_t or_mod_entry(const olimb *e, int64_t cap, uint32_t p)
{
    int64_t s = ((const int64_t *)e)[cap];
    int64_t n = s < 0 ? -s : s;
    uint64_t r = 0;
    for (int64_t i = n - 1; i >= 0; i--) r = (uint64_t)((((u128)r) << 64 | e[i]) % p);
    if (s < 0 && r) r = p - r;
    return (uint32_t)r;
}

/* r = sum_i d_i 256^i with d_i in [-128, 127]; exact for -2155905152 <= r <= 2139062143. */
void or_balanced_slices(int32_t r, int8_t d[4])
{
    int64_t v = r;
    for (int i = 0; i < 4; i++) {
        int64_t t = ((v + 128) & 255) - 128;
        d[i] = (int8_t)t;
        v = (v - t) / 256;
    }
}

static uint64_t inv_mod(uint64_t a, uint64_t p) { return powmod(a, p - 2, p); }

/* Garner: x = v_0 + v_1 p_0 + v_2 p_0 p_1 + ..., with 0 <= v_i < p_i, so
 * 0 <= x < M; then x - M if x > M/2.  Status 1 when the result needs more
 * than cap limbs. */
int or_crt_reconstruct(olimb *e, int64_t cap, const uint32_t *res, const uint32_t *primes, int64_t P)
{
    int64_t L = P + 4;
    olimb *x = (olimb *)calloc((size_t)L, sizeof(olimb));
    olimb *m = (olimb *)calloc((size_t)L, sizeof(olimb));
    olimb *t = (olimb *)calloc((size_t)L, sizeof(olimb));
    int64_t xn = 1, mn = 1;
    m[0] = 1; x[0] = 0;
    for (int64_t i = 0; i < P; i++) {
        uint64_t p = primes[i];
        /* x mod p and m mod p */
        uint64_t xm = 0, mm = 0;
        for (int64_t j = xn - 1; j >= 0; j--) xm = (uint64_t)((((u128)xm) << 64 | x[j]) % p);
        for (int64_t j = mn - 1; j >= 0; j--) mm = (uint64_t)((((u128)mm) << 64 | m[j]) % p);
        uint64_t v = mulmod((res[i] % p + p - xm) % p, inv_mod(mm, p), p);
        /* x += v * m */
        memset(t, 0, (size_t)L * sizeof(olimb));
        t[mn] = or_mpn_mul_1(t, m, mn, v);
        int64_t tn = mn + 1;
        if (tn > xn) { memset(x + xn, 0, (size_t)(tn - xn) * sizeof(olimb)); xn = tn; }
        olimb cy = or_mpn_add(x, x, xn, t, tn);
        if (cy) x[xn++] = cy;
        while (xn > 1 && x[xn - 1] == 0) xn--;
        /* m *= p */
        m[mn] = or_mpn_mul_1(m, m, mn, p);
        mn++;
        while (mn > 1 && m[mn - 1] == 0) mn--;
    }
    while (xn > 0 && x[xn - 1] == 0) xn--;
    while (mn > 0 && m[mn - 1] == 0) mn--;
    /* compare 2x with m */
    olimb *x2 = t;
    memset(x2, 0, (size_t)L * sizeof(olimb));
    memcpy(x2, x, (size_t)xn * sizeof(olimb));
    int64_t x2n = xn + 1;
    x2[xn] = or_mpn_add_n(x2, x2, x2, xn);
    int64_t cmpn = x2n > mn ? x2n : mn;
    for (int64_t j = mn; j < cmpn; j++) m[j] = 0;
    int neg = or_mpn_cmp(x2, m, cmpn) > 0;
    int64_t sign = 1;
    if (neg) {
        /* x = m - x */
        for (int64_t j = xn; j < mn; j++) x[j] = 0;
        or_mpn_sub_n(x, m, x, mn);
        xn = mn;
        while (xn > 0 && x[xn - 1] == 0) xn--;
        sign = -1;
    }
    int st = 0;
    if (xn > cap) st = 1;
    else {
        memcpy(e, x, (size_t)xn * sizeof(olimb));
        ((int64_t *)e)[cap] = sign * xn;
    }
    free(x); free(m); free(t);
    return st;
}

typedef struct {
    int64_t K, M, N, Crs, Ars, Brs, cap, P;
    olimb *C;
    const olimb *A, *B;
    const uint32_t *primes;
    uint32_t *ra, *rb;
    int nthreads, tid, status;
} crt_job;

static void *crt_worker(void *arg)
{
    crt_job *j = (crt_job *)arg;
    uint32_t *res = (uint32_t *)malloc((size_t)j->P * sizeof(uint32_t));
    for (int64_t y = j->tid; y < j->K; y += j->nthreads)
        for (int64_t x = 0; x < j->N; x++) {
            for (int64_t q = 0; q < j->P; q++) {
                uint64_t p = j->primes[q];
                u128 acc = 0;
                for (int64_t k = 0; k < j->M; k++)
                    acc += (u128)j->ra[(q * j->K + y) * j->M + k] * j->rb[(q * j->N + x) * j->M + k];
                res[q] = (uint32_t)(acc % p);
            }
            if (or_crt_reconstruct(j->C + (y * j->Crs + x) * (j->cap + 1), j->cap, res, j->primes, j->P))
                j->status = 1;
        }
    free(res);
    return NULL;
}

/* C = A*B through residues: C mod p_q = (A mod p_q)(B mod p_q) mod p_q for
 * every prime, then Garner per entry. */
int or_mat_mul_crt(int64_t K, int64_t M, int64_t N, olimb *C, int64_t Crs, const olimb *A, int64_t Ars,
                   const olimb *B, int64_t Brs, int64_t cap, const uint32_t *primes, int64_t P, int nthreads)
{
    int64_t W = cap + 1;
    uint32_t *ra = (uint32_t *)malloc((size_t)(P * K * M) * sizeof(uint32_t));
    uint32_t *rb = (uint32_t *)malloc((size_t)(P * N * M) * sizeof(uint32_t));
    for (int64_t q = 0; q < P; q++) {
        for (int64_t y = 0; y < K; y++)
            for (int64_t k = 0; k < M; k++) ra[(q * K + y) * M + k] = or_mod_entry(A + (y * Ars + k) * W, cap, primes[q]);
        for (int64_t k = 0; k < M; k++)
            for (int64_t x = 0; x < N; x++) rb[(q * N + x) * M + k] = or_mod_entry(B + (k * Brs + x) * W, cap, primes[q]);
    }
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 64) nthreads = 64;
    pthread_t th[64];
    crt_job jobs[64];
    for (int t = 0; t < nthreads; t++) {
        crt_job j = {K, M, N, Crs, Ars, Brs, cap, P, C, A, B, primes, ra, rb, nthreads, t, 0};
        jobs[t] = j;
        pthread_create(&th[t], NULL, crt_worker, &jobs[t]);
    }
    int st = 0;
    for (int t = 0; t < nthreads; t++) { pthread_join(th[t], NULL); st |= jobs[t].status; }
    free(ra); free(rb);
    return st;
}
