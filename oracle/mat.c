/*
 * mat.c -- oracle (test infrastructure only): matrix layer on entry arrays.
 *
 * C is K x N, A is K x M, B is M x N (mat-gpu/main.cu:242-245); rowstrides
 * count entries (mat-gpu/main.cu:26).  An entry is cap+1 words.
 */
#include "oracle.h"
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define ENT(P, rs, y, x, cap) ((P) + ((int64_t)(y) * (rs) + (x)) * ((cap) + 1))

typedef struct {
    int64_t K, M, N, Crs, Ars, Brs, cap;
    olimb *C;
    const olimb *A, *B;
    int addmul, nthreads, tid, status;
} mm_job;

/* mat-gpu/mat-core.h:7-53 (mat_mul_1 / mat_addmul_1): the k-loop of mpz
 * addmul per output entry; mat-cpu/mat.c:72-93 is the same on GMP. */
static void *mm_worker(void *arg)
{
    mm_job *j = (mm_job *)arg;
    for (int64_t y = j->tid; y < j->K; y += j->nthreads) {
        for (int64_t x = 0; x < j->N; x++) {
            olimb *c = ENT(j->C, j->Crs, y, x, j->cap);
            int64_t k0 = 0;
            if (!j->addmul) {
                if (j->M == 0) { c[j->cap] = 0; continue; }
                if (or_mpz_mul(c, ENT(j->A, j->Ars, y, 0, j->cap), ENT(j->B, j->Brs, 0, x, j->cap), j->cap)) {
                    j->status = 1; return NULL;
                }
                k0 = 1;
            }
            for (int64_t k = k0; k < j->M; k++) {
                if (or_mpz_addmul(c, ENT(j->A, j->Ars, y, k, j->cap), ENT(j->B, j->Brs, k, x, j->cap), j->cap)) {
                    j->status = 1; return NULL;
                }
            }
        }
    }
    return NULL;
}

static int mm(int64_t K, int64_t M, int64_t N, olimb *C, int64_t Crs, const olimb *A, int64_t Ars,
              const olimb *B, int64_t Brs, int64_t cap, int nthreads, int addmul)
{
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    mm_job jobs[256];
    for (int t = 0; t < nthreads; t++) {
        mm_job j = {K, M, N, Crs, Ars, Brs, cap, C, A, B, addmul, nthreads, t, 0};
        jobs[t] = j;
    }
    if (nthreads == 1) { mm_worker(&jobs[0]); return jobs[0].status; }
    for (int t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, mm_worker, &jobs[t]);
    int st = 0;
    for (int t = 0; t < nthreads; t++) { pthread_join(th[t], NULL); st |= jobs[t].status; }
    return st;
}

int or_mat_mul(int64_t K, int64_t M, int64_t N, olimb *C, int64_t Crs, const olimb *A, int64_t Ars,
               const olimb *B, int64_t Brs, int64_t cap, int nthreads)
{
    return mm(K, M, N, C, Crs, A, Ars, B, Brs, cap, nthreads, 0);
}

int or_mat_addmul(int64_t K, int64_t M, int64_t N, olimb *C, int64_t Crs, const olimb *A, int64_t Ars,
                  const olimb *B, int64_t Brs, int64_t cap, int nthreads)
{
    return mm(K, M, N, C, Crs, A, Ars, B, Brs, cap, nthreads, 1);
}

/* mat-gpu/mat-gpu-launchers.cu:65-78 / :107-120 (launcher_add / launcher_sub) */
static int aors_mat(int64_t W, int64_t H, olimb *C, int64_t Crs, const olimb *A, int64_t Ars,
                    const olimb *B, int64_t Brs, int64_t cap, int sub)
{
    for (int64_t y = 0; y < H; y++)
        for (int64_t x = 0; x < W; x++) {
            int st = sub ? or_mpz_sub(ENT(C, Crs, y, x, cap), ENT(A, Ars, y, x, cap), ENT(B, Brs, y, x, cap), cap)
                         : or_mpz_add(ENT(C, Crs, y, x, cap), ENT(A, Ars, y, x, cap), ENT(B, Brs, y, x, cap), cap);
            if (st) return st;
        }
    return 0;
}

int or_mat_add(int64_t W, int64_t H, olimb *C, int64_t Crs, const olimb *A, int64_t Ars,
               const olimb *B, int64_t Brs, int64_t cap)
{
    return aors_mat(W, H, C, Crs, A, Ars, B, Brs, cap, 0);
}

int or_mat_sub(int64_t W, int64_t H, olimb *C, int64_t Crs, const olimb *A, int64_t Ars,
               const olimb *B, int64_t Brs, int64_t cap)
{
    return aors_mat(W, H, C, Crs, A, Ars, B, Brs, cap, 1);
}

/* ---- 2x2 block recursion (mat-gpu/mat-defines.h sequences) ---- */
enum { A11, A12, A21, A22, B11, B12, B21, B22, C11, C12, C21, C22, P1, P2, P3, P4 };
enum { OADD, OSUB, OMUL, OAML };
typedef struct { unsigned char op, d, a, b; } step;

/* mat-gpu/mat-defines.h:21-36 (NAIVE_RECURSIVE_SEQUENCE; T is p1) */
static const step seq_naive[] = {
    {OMUL, C11, A11, B11}, {OMUL, P1, A12, B21}, {OADD, C11, C11, P1},
    {OMUL, C12, A11, B12}, {OMUL, P1, A12, B22}, {OADD, C12, C12, P1},
    {OMUL, C21, A21, B11}, {OMUL, P1, A22, B21}, {OADD, C21, C21, P1},
    {OMUL, C22, A21, B12}, {OMUL, P1, A22, B22}, {OADD, C22, C22, P1},
};
/* mat-gpu/mat-defines.h:77-117 (STRASSEN_SEQUENCE, three temporaries) */
static const step seq_strassen[] = {
    {OADD, P1, A11, A22}, {OADD, P2, B11, B22}, {OMUL, C11, P1, P2},
    {OSUB, P1, B12, B22}, {OMUL, C12, A11, P1}, {OADD, C22, C11, C12},
    {OADD, P1, A21, A22}, {OMUL, C21, P1, B11}, {OSUB, C22, C22, C21},
    {OSUB, P2, A12, A22}, {OADD, P3, B21, B22}, {OMUL, P1, P2, P3}, {OADD, C11, C11, P1},
    {OSUB, P2, A21, A11}, {OADD, P3, B11, B12}, {OMUL, P1, P2, P3}, {OADD, C22, C22, P1},
    {OSUB, P2, B21, B11}, {OMUL, P1, A22, P2}, {OADD, C11, C11, P1}, {OADD, C21, C21, P1},
    {OADD, P2, A11, A12}, {OMUL, P1, P2, B22}, {OSUB, C11, C11, P1}, {OADD, C12, C12, P1},
};
/* mat-gpu/mat-defines.h:144-164 (WINOGRAD_SEQUENCE) */
static const step seq_winograd[] = {
    {OADD, P1, A21, A22}, {OSUB, P2, B12, B11}, {OMUL, C22, P1, P2},
    {OSUB, P1, P1, A11}, {OSUB, P2, B22, P2}, {OMUL, P3, A11, B11},
    {OMUL, C11, A12, B21}, {OADD, C11, C11, P3}, {OAML, P3, P1, P2},
    {OSUB, P1, A12, P1}, {OSUB, P2, B21, P2}, {OMUL, C12, P1, B22},
    {OADD, C12, C12, C22}, {OADD, C12, C12, P3}, {OMUL, C21, A22, P2},
    {OSUB, P1, A11, A21}, {OSUB, P2, B22, B12}, {OAML, P3, P1, P2},
    {OADD, C21, C21, P3}, {OADD, C22, C22, P3},
};
/* mat-gpu/mat-defines.h:166-187 (WINOGRAD_ADDMUL_SEQUENCE: C += A*B) */
static const step seq_winograd_addmul[] = {
    {OADD, P1, A21, A22}, {OSUB, P2, B12, B11}, {OMUL, P3, P1, P2},
    {OADD, C12, C12, P3}, {OADD, C22, C22, P3}, {OSUB, P1, P1, A11},
    {OSUB, P2, B22, P2}, {OMUL, P3, A11, B11}, {OADD, C11, C11, P3},
    {OAML, P3, P1, P2}, {OAML, C11, A12, B21}, {OSUB, P1, A12, P1},
    {OSUB, P2, B21, P2}, {OAML, C12, P1, B22}, {OADD, C12, C12, P3},
    {OAML, C21, A22, P2}, {OSUB, P1, A11, A21}, {OSUB, P2, B22, B12},
    {OAML, P3, P1, P2}, {OADD, C21, C21, P3}, {OADD, C22, C22, P3},
};

static int rec(int method, int addmul, int64_t N, olimb *C, int64_t Crs, const olimb *A, int64_t Ars,
               const olimb *B, int64_t Brs, int64_t cap, int64_t base);

static int run_seq(const step *s, int ns, int method, int64_t N, olimb *C, int64_t Crs, const olimb *A,
                   int64_t Ars, const olimb *B, int64_t Brs, int64_t cap, int64_t base)
{
    int64_t h = N / 2, W = cap + 1;
    olimb *T = (olimb *)calloc((size_t)(4 * h * h * W), sizeof(olimb));
    olimb *ptr[16];
    int64_t rs[16];
    for (int q = 0; q < 4; q++) {
        int64_t oy = (q >> 1) * h, ox = (q & 1) * h;
        ptr[A11 + q] = (olimb *)ENT(A, Ars, oy, ox, cap); rs[A11 + q] = Ars;
        ptr[B11 + q] = (olimb *)ENT(B, Brs, oy, ox, cap); rs[B11 + q] = Brs;
        ptr[C11 + q] = ENT(C, Crs, oy, ox, cap); rs[C11 + q] = Crs;
        ptr[P1 + q] = T + q * h * h * W; rs[P1 + q] = h;
    }
    int st = 0;
    for (int i = 0; i < ns && !st; i++) {
        const step *p = &s[i];
        switch (p->op) {
        case OADD: st = or_mat_add(h, h, ptr[p->d], rs[p->d], ptr[p->a], rs[p->a], ptr[p->b], rs[p->b], cap); break;
        case OSUB: st = or_mat_sub(h, h, ptr[p->d], rs[p->d], ptr[p->a], rs[p->a], ptr[p->b], rs[p->b], cap); break;
        case OMUL: st = rec(method, 0, h, ptr[p->d], rs[p->d], ptr[p->a], rs[p->a], ptr[p->b], rs[p->b], cap, base); break;
        case OAML: st = rec(method, 1, h, ptr[p->d], rs[p->d], ptr[p->a], rs[p->a], ptr[p->b], rs[p->b], cap, base); break;
        }
    }
    free(T);
    return st;
}

/* mat-gpu/mat-gpu.cu:79-475: recurse while N > base (N even), leaf naive. */
static int rec(int method, int addmul, int64_t N, olimb *C, int64_t Crs, const olimb *A, int64_t Ars,
               const olimb *B, int64_t Brs, int64_t cap, int64_t base)
{
    if (N <= base || (N & 1)) {
        return addmul ? or_mat_addmul(N, N, N, C, Crs, A, Ars, B, Brs, cap, 1)
                      : or_mat_mul(N, N, N, C, Crs, A, Ars, B, Brs, cap, 1);
    }
    if (addmul) {
        if (method == 2)
            return run_seq(seq_winograd_addmul, sizeof seq_winograd_addmul / sizeof(step), method, N, C, Crs, A, Ars, B, Brs, cap, base);
        /* naive/Strassen addmul: T = A*B; C += T */
        olimb *T = (olimb *)calloc((size_t)(N * N * (cap + 1)), sizeof(olimb));
        int st = rec(method, 0, N, T, N, A, Ars, B, Brs, cap, base);
        if (!st) st = or_mat_add(N, N, C, Crs, C, Crs, T, N, cap);
        free(T);
        return st;
    }
    switch (method) {
    case 0: return run_seq(seq_naive, sizeof seq_naive / sizeof(step), method, N, C, Crs, A, Ars, B, Brs, cap, base);
    case 1: return run_seq(seq_strassen, sizeof seq_strassen / sizeof(step), method, N, C, Crs, A, Ars, B, Brs, cap, base);
    default: return run_seq(seq_winograd, sizeof seq_winograd / sizeof(step), method, N, C, Crs, A, Ars, B, Brs, cap, base);
    }
}

int or_mat_recursive(int method, int64_t N, olimb *C, int64_t Crs, const olimb *A, int64_t Ars,
                     const olimb *B, int64_t Brs, int64_t cap, int64_t base)
{
    return rec(method, 0, N, C, Crs, A, Ars, B, Brs, cap, base);
}

int or_mat_winograd_addmul(int64_t N, olimb *C, int64_t Crs, const olimb *A, int64_t Ars,
                           const olimb *B, int64_t Brs, int64_t cap, int64_t base)
{
    return rec(2, 1, N, C, Crs, A, Ars, B, Brs, cap, base);
}

static int64_t odd_part(int64_t n) { while (n > 0 && !(n & 1)) n >>= 1; return n; }

/* mat-gpu/mat-pregpu.cu:469-576 (naive_smallrect_recurse + squarify_mul_pregpu):
 * small or rectangular products are split at the largest power of two not
 * above the smallest dimension; squares whose odd part is <= base recurse. */
int or_mat_squarify(int method, int64_t K, int64_t M, int64_t N, olimb *C, int64_t Crs, const olimb *A,
                    int64_t Ars, const olimb *B, int64_t Brs, int64_t cap, int64_t base)
{
    if (K == 0 || M == 0 || N == 0) return 0;
    if (K <= base || M <= base || N <= base) return or_mat_mul(K, M, N, C, Crs, A, Ars, B, Brs, cap, 1);
    if (K == M && M == N && odd_part(K) <= base) return or_mat_recursive(method, K, C, Crs, A, Ars, B, Brs, cap, base);
    int64_t mind = K < M ? K : M;
    if (N < mind) mind = N;
    int64_t lt = 1;
    while (lt * 2 <= mind) lt *= 2;
    int64_t W = cap + 1;
    olimb *T = (olimb *)calloc((size_t)(K * N * W), sizeof(olimb));
    int tK = lt < K, tM = lt < M, tN = lt < N, st = 0;
#define SUBM(P, rs, iy, ix) ((P) + ((iy) * lt * (rs) + (ix) * lt) * W)
#define MUL(k, m, n, c, crs, a, ars, b, brs) if (!st) st = or_mat_squarify(method, k, m, n, c, crs, a, ars, b, brs, cap, base)
#define ADD(h, w, c, crs, a, ars, b, brs) if (!st) st = or_mat_add(w, h, c, crs, a, ars, b, brs, cap)
    MUL(lt, lt, lt, SUBM(C, Crs, 0, 0), Crs, SUBM(A, Ars, 0, 0), Ars, SUBM(B, Brs, 0, 0), Brs);
    if (tM) { MUL(lt, M - lt, lt, SUBM(T, N, 0, 0), N, SUBM(A, Ars, 0, 1), Ars, SUBM(B, Brs, 1, 0), Brs);
              ADD(lt, lt, SUBM(C, Crs, 0, 0), Crs, SUBM(C, Crs, 0, 0), Crs, SUBM(T, N, 0, 0), N); }
    if (tN) {
        MUL(lt, lt, N - lt, SUBM(C, Crs, 0, 1), Crs, SUBM(A, Ars, 0, 0), Ars, SUBM(B, Brs, 0, 1), Brs);
        if (tM) { MUL(lt, M - lt, N - lt, SUBM(T, N, 0, 1), N, SUBM(A, Ars, 0, 1), Ars, SUBM(B, Brs, 1, 1), Brs);
                  ADD(lt, N - lt, SUBM(C, Crs, 0, 1), Crs, SUBM(C, Crs, 0, 1), Crs, SUBM(T, N, 0, 1), N); }
    }
    if (tK) {
        MUL(K - lt, lt, lt, SUBM(C, Crs, 1, 0), Crs, SUBM(A, Ars, 1, 0), Ars, SUBM(B, Brs, 0, 0), Brs);
        if (tM) { MUL(K - lt, M - lt, lt, SUBM(T, N, 1, 0), N, SUBM(A, Ars, 1, 1), Ars, SUBM(B, Brs, 1, 0), Brs);
                  ADD(K - lt, lt, SUBM(C, Crs, 1, 0), Crs, SUBM(C, Crs, 1, 0), Crs, SUBM(T, N, 1, 0), N); }
        if (tN) {
            MUL(K - lt, lt, N - lt, SUBM(C, Crs, 1, 1), Crs, SUBM(A, Ars, 1, 0), Ars, SUBM(B, Brs, 0, 1), Brs);
            if (tM) { MUL(K - lt, M - lt, N - lt, SUBM(T, N, 1, 1), N, SUBM(A, Ars, 1, 1), Ars, SUBM(B, Brs, 1, 1), Brs);
                      ADD(K - lt, N - lt, SUBM(C, Crs, 1, 1), Crs, SUBM(C, Crs, 1, 1), Crs, SUBM(T, N, 1, 1), N); }
        }
    }
#undef SUBM
#undef MUL
#undef ADD
    free(T);
    return st;
}

/* mat-gpu/main.cu:137-156 (transform_interleave): within each group of npart
 * parts of partsz bytes, atom a of part q moves to slot (npart*a + q). */
void or_transform_interleave(void *p_, int64_t ngroup, int64_t npart, int64_t partsz, int64_t atomsz)
{
    int64_t natom = partsz / atomsz;
    unsigned char *tmp = (unsigned char *)malloc((size_t)(npart * partsz));
    unsigned char *p = (unsigned char *)p_;
    for (int64_t g = 0; g < ngroup; g++) {
        for (int64_t q = 0; q < npart; q++)
            for (int64_t a = 0; a < natom; a++)
                memcpy(tmp + atomsz * (npart * a + q), p + atomsz * (natom * q + a), (size_t)atomsz);
        memcpy(p, tmp, (size_t)(npart * partsz));
        p += npart * partsz;
    }
    free(tmp);
}
