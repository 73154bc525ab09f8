/*
 * matcpu.c -- the CPU baseline (test/bench infrastructure only): a restatement
 * of mat-cpu's method 35, i.e. GMP mpz arithmetic (mat-cpu/mat.c:54-334) under
 * a Winograd job DAG (mat-cpu/matj.c:345-560, squarify :680-780) run by a
 * pthread control/worker scheduler that hands out ready ("source") jobs
 * depth-first, newest first (mat-cpu/job.c:448-650).
 *
 * GMP is loaded at run time from libgmp.so.10 (no header needed); it is the
 * arithmetic the reference links (mat-cpu/README.txt:1-7).  Entries cross the
 * boundary as the fixed-capacity limb arrays used everywhere else.
 *
 * Dependencies between jobs of one recursion level are derived from the
 * blocks each job reads and writes (read-after-write, write-after-read,
 * write-after-write), which is what the reference's hand-written edges and
 * "write conflict" edges encode (mat-cpu/matj.c:401-420).
 */
#include "oracle.h"
#include <dlfcn.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include <time.h>

typedef struct { int alloc; int size; uint64_t *d; } gmpz;   /* GMP's __mpz_struct ABI */

static struct {
    void *h;
    void (*init)(gmpz *);
    void (*clear)(gmpz *);
    void (*add)(gmpz *, const gmpz *, const gmpz *);
    void (*sub)(gmpz *, const gmpz *, const gmpz *);
    void (*mul)(gmpz *, const gmpz *, const gmpz *);
    void (*addmul)(gmpz *, const gmpz *, const gmpz *);
    void (*import)(gmpz *, size_t, int, size_t, int, size_t, const void *);
    void *(*export_)(void *, size_t *, int, size_t, int, size_t, const gmpz *);
    const char *version;
} G;

static int gmp_load(void)
{
    if (G.h) return 0;
    const char *names[] = {"libgmp.so.10", "libgmp.so", NULL};
    for (int i = 0; names[i] && !G.h; i++) G.h = dlopen(names[i], RTLD_NOW | RTLD_LOCAL);
    if (!G.h) return 1;
    G.init = (void (*)(gmpz *))dlsym(G.h, "__gmpz_init");
    G.clear = (void (*)(gmpz *))dlsym(G.h, "__gmpz_clear");
    G.add = (void (*)(gmpz *, const gmpz *, const gmpz *))dlsym(G.h, "__gmpz_add");
    G.sub = (void (*)(gmpz *, const gmpz *, const gmpz *))dlsym(G.h, "__gmpz_sub");
    G.mul = (void (*)(gmpz *, const gmpz *, const gmpz *))dlsym(G.h, "__gmpz_mul");
    G.addmul = (void (*)(gmpz *, const gmpz *, const gmpz *))dlsym(G.h, "__gmpz_addmul");
    G.import = (void (*)(gmpz *, size_t, int, size_t, int, size_t, const void *))dlsym(G.h, "__gmpz_import");
    G.export_ = (void *(*)(void *, size_t *, int, size_t, int, size_t, const gmpz *))dlsym(G.h, "__gmpz_export");
    const char **v = (const char **)dlsym(G.h, "__gmp_version");
    G.version = v ? *v : "?";
    if (!G.init || !G.clear || !G.add || !G.sub || !G.mul || !G.addmul || !G.import || !G.export_) {
        dlclose(G.h); G.h = NULL; return 1;
    }
    return 0;
}

const char *or_matcpu_gmp_version(void) { return gmp_load() ? NULL : G.version; }

/* ---- sequential GMP matrix ops (mat-cpu/mat.c:54-334) ---- */
typedef gmpz *mat;
#define EL(P, rs, y, x) ((P) + (int64_t)(y) * (rs) + (x))

static void m_add(int64_t n, mat C, int64_t cr, const gmpz *A, int64_t ar, const gmpz *B, int64_t br, int sub)
{
    for (int64_t y = 0; y < n; y++)
        for (int64_t x = 0; x < n; x++)
            (sub ? G.sub : G.add)(EL(C, cr, y, x), EL(A, ar, y, x), EL(B, br, y, x));
}

static void m_mul(int64_t n, mat C, int64_t cr, const gmpz *A, int64_t ar, const gmpz *B, int64_t br, int addmul)
{
    for (int64_t y = 0; y < n; y++)
        for (int64_t x = 0; x < n; x++) {
            gmpz *c = EL(C, cr, y, x);
            int64_t k0 = 0;
            if (!addmul) { G.mul(c, EL(A, ar, y, 0), EL(B, br, 0, x)); k0 = 1; }
            for (int64_t k = k0; k < n; k++) G.addmul(c, EL(A, ar, y, k), EL(B, br, k, x));
        }
}

static mat m_init(int64_t n)
{
    mat T = (mat)malloc((size_t)(n * n) * sizeof(gmpz));
    for (int64_t i = 0; i < n * n; i++) G.init(&T[i]);
    return T;
}
static void m_free(int64_t n, mat T)
{
    for (int64_t i = 0; i < n * n; i++) G.clear(&T[i]);
    free(T);
}

enum { A11, A12, A21, A22, B11, B12, B21, B22, C11, C12, C21, C22, P1, P2, P3 };
enum { OADD, OSUB, OMUL, OAML };
typedef struct { unsigned char op, d, a, b; } step;
/* mat-cpu/mat.c:193-262 and matj.c:401-420 (Winograd, 7 products) */
static const step seq_w[] = {
    {OADD, P1, A21, A22}, {OSUB, P2, B12, B11}, {OMUL, C22, P1, P2},
    {OSUB, P1, P1, A11}, {OSUB, P2, B22, P2}, {OMUL, P3, A11, B11},
    {OMUL, C11, A12, B21}, {OADD, C11, C11, P3}, {OAML, P3, P1, P2},
    {OSUB, P1, A12, P1}, {OSUB, P2, B21, P2}, {OMUL, C12, P1, B22},
    {OADD, C12, C12, C22}, {OADD, C12, C12, P3}, {OMUL, C21, A22, P2},
    {OSUB, P1, A11, A21}, {OSUB, P2, B22, B12}, {OAML, P3, P1, P2},
    {OADD, C21, C21, P3}, {OADD, C22, C22, P3},
};
/* mat-cpu/mat.c:264-334 and matj.c:457-560 (C += A*B) */
static const step seq_wa[] = {
    {OADD, P1, A21, A22}, {OSUB, P2, B12, B11}, {OMUL, P3, P1, P2},
    {OADD, C12, C12, P3}, {OADD, C22, C22, P3}, {OSUB, P1, P1, A11},
    {OSUB, P2, B22, P2}, {OMUL, P3, A11, B11}, {OADD, C11, C11, P3},
    {OAML, P3, P1, P2}, {OAML, C11, A12, B21}, {OSUB, P1, A12, P1},
    {OSUB, P2, B21, P2}, {OAML, C12, P1, B22}, {OADD, C12, C12, P3},
    {OAML, C21, A22, P2}, {OSUB, P1, A11, A21}, {OSUB, P2, B22, B12},
    {OAML, P3, P1, P2}, {OADD, C21, C21, P3}, {OADD, C22, C22, P3},
};

static void quad_ptrs(int64_t n, mat C, int64_t cr, const gmpz *A, int64_t ar, const gmpz *B, int64_t br,
                      mat T, gmpz **ptr, int64_t *rs)
{
    int64_t h = n / 2;
    for (int q = 0; q < 4; q++) {
        int64_t oy = (q >> 1) * h, ox = (q & 1) * h;
        ptr[A11 + q] = (gmpz *)EL(A, ar, oy, ox); rs[A11 + q] = ar;
        ptr[B11 + q] = (gmpz *)EL(B, br, oy, ox); rs[B11 + q] = br;
        ptr[C11 + q] = EL(C, cr, oy, ox); rs[C11 + q] = cr;
    }
    for (int q = 0; q < 3; q++) { ptr[P1 + q] = EL(T, n, (q >> 1) * h, (q & 1) * h); rs[P1 + q] = n; }
}

/* mat-cpu/mat.c:193-334 with nlevels levels of recursion, naive leaves */
static void m_winograd(int64_t n, mat C, int64_t cr, const gmpz *A, int64_t ar, const gmpz *B, int64_t br,
                       int nlevels, int addmul)
{
    if (nlevels <= 0 || (n & 1)) { m_mul(n, C, cr, A, ar, B, br, addmul); return; }
    mat T = m_init(n);
    gmpz *ptr[15];
    int64_t rs[15];
    quad_ptrs(n, C, cr, A, ar, B, br, T, ptr, rs);
    const step *s = addmul ? seq_wa : seq_w;
    int ns = addmul ? (int)(sizeof seq_wa / sizeof(step)) : (int)(sizeof seq_w / sizeof(step));
    int64_t h = n / 2;
    for (int i = 0; i < ns; i++) {
        const step *p = &s[i];
        if (p->op == OADD || p->op == OSUB)
            m_add(h, ptr[p->d], rs[p->d], ptr[p->a], rs[p->a], ptr[p->b], rs[p->b], p->op == OSUB);
        else
            m_winograd(h, ptr[p->d], rs[p->d], ptr[p->a], rs[p->a], ptr[p->b], rs[p->b], nlevels - 1, p->op == OAML);
    }
    m_free(n, T);
}

/* ---- job DAG + scheduler (mat-cpu/job.c) ---- */
typedef struct job {
    int kind;              /* 0 add, 1 sub, 2 winograd leaf, 3 addmul leaf, 4 free T, 5 noop */
    int64_t n, cr, ar, br;
    gmpz *C; const gmpz *A, *B;
    int nlevels;
    int ndeps;             /* unmet dependencies */
    int *rev; int nrev, caprev;
} job;

typedef struct {
    job *jobs; int njobs, cap;
    int64_t minblock, jobsize;
} dag;

static int new_job(dag *g, int kind)
{
    if (g->njobs == g->cap) { g->cap = g->cap ? 2 * g->cap : 256; g->jobs = (job *)realloc(g->jobs, (size_t)g->cap * sizeof(job)); }
    job *j = &g->jobs[g->njobs];
    memset(j, 0, sizeof *j);
    j->kind = kind;
    return g->njobs++;
}

static void add_dep(dag *g, int from, int to)
{
    if (from < 0) return;
    job *f = &g->jobs[from];
    for (int i = 0; i < f->nrev; i++) if (f->rev[i] == to) return;
    if (f->nrev == f->caprev) { f->caprev = f->caprev ? 2 * f->caprev : 4; f->rev = (int *)realloc(f->rev, (size_t)f->caprev * sizeof(int)); }
    f->rev[f->nrev++] = to;
    g->jobs[to].ndeps++;
}

typedef struct { int *v; int n; } deps;

static int floor_log2(int64_t x) { int l = 0; while (x > 1) { x >>= 1; l++; } return l; }

/* Builds the DAG for C (=|+=) A*B of size n after the jobs in gd; returns a
 * job that completes after everything inside (mat-cpu/matj.c:345-446). */
static int build(dag *g, int64_t n, gmpz *C, int64_t cr, const gmpz *A, int64_t ar, const gmpz *B, int64_t br,
                 int addmul, deps gd)
{
    if (n <= g->jobsize || (n & 1)) {
        int j = new_job(g, addmul ? 3 : 2);
        job *J = &g->jobs[j];
        J->n = n; J->C = C; J->cr = cr; J->A = A; J->ar = ar; J->B = B; J->br = br;
        J->nlevels = n <= g->minblock ? 0 : floor_log2(n / g->minblock);
        for (int i = 0; i < gd.n; i++) add_dep(g, gd.v[i], j);
        return j;
    }
    mat T = m_init(n);
    gmpz *ptr[15];
    int64_t rs[15];
    quad_ptrs(n, C, cr, A, ar, B, br, T, ptr, rs);
    const step *s = addmul ? seq_wa : seq_w;
    int ns = addmul ? (int)(sizeof seq_wa / sizeof(step)) : (int)(sizeof seq_w / sizeof(step));
    int last_writer[15], readers[15][24], nreaders[15];
    for (int b = 0; b < 15; b++) { last_writer[b] = -1; nreaders[b] = 0; }
    int all[24], nall = 0;
    int64_t h = n / 2;
    for (int i = 0; i < ns; i++) {
        const step *p = &s[i];
        int d[64], nd = 0;
        int srcs[2] = {p->a, p->b};
        for (int q = 0; q < 2; q++) if (last_writer[srcs[q]] >= 0) d[nd++] = last_writer[srcs[q]];
        if (last_writer[p->d] >= 0) d[nd++] = last_writer[p->d];
        for (int r = 0; r < nreaders[p->d] && nd < 64; r++) d[nd++] = readers[p->d][r];
        /* a job with a same-level predecessor already waits for gd through it */
        if (nd == 0) for (int r = 0; r < gd.n && nd < 64; r++) d[nd++] = gd.v[r];
        deps dd = {d, nd};
        int j;
        if (p->op == OADD || p->op == OSUB) {
            j = new_job(g, p->op == OSUB);
            job *J = &g->jobs[j];
            J->n = h; J->C = ptr[p->d]; J->cr = rs[p->d]; J->A = ptr[p->a]; J->ar = rs[p->a]; J->B = ptr[p->b]; J->br = rs[p->b];
            for (int r = 0; r < nd; r++) add_dep(g, d[r], j);
        } else {
            j = build(g, h, ptr[p->d], rs[p->d], ptr[p->a], rs[p->a], ptr[p->b], rs[p->b], p->op == OAML, dd);
        }
        for (int q = 0; q < 2; q++) readers[srcs[q]][nreaders[srcs[q]]++] = j;
        last_writer[p->d] = j;
        nreaders[p->d] = 0;
        all[nall++] = j;
    }
    int cl = new_job(g, 4);
    g->jobs[cl].n = n; g->jobs[cl].C = T;
    for (int i = 0; i < nall; i++) add_dep(g, all[i], cl);
    return cl;
}

typedef struct {
    dag *g;
    pthread_mutex_t mu;
    pthread_cond_t cv;
    int *ready; int nready;     /* source jobs, newest first */
    int running, done;
} sched;

static void run_job(job *J)
{
    switch (J->kind) {
    case 0: m_add(J->n, J->C, J->cr, J->A, J->ar, J->B, J->br, 0); break;
    case 1: m_add(J->n, J->C, J->cr, J->A, J->ar, J->B, J->br, 1); break;
    case 2: m_winograd(J->n, J->C, J->cr, J->A, J->ar, J->B, J->br, J->nlevels, 0); break;
    case 3: m_winograd(J->n, J->C, J->cr, J->A, J->ar, J->B, J->br, J->nlevels, 1); break;
    case 4: m_free(J->n, J->C); break;
    default: break;
    }
}

static void *worker(void *arg)
{
    sched *S = (sched *)arg;
    pthread_mutex_lock(&S->mu);
    for (;;) {
        while (S->nready == 0 && S->done < S->g->njobs) pthread_cond_wait(&S->cv, &S->mu);
        if (S->done == S->g->njobs) break;
        int j = S->ready[0];                      /* index 0: depth-first (job.c:514-515) */
        memmove(S->ready, S->ready + 1, (size_t)(S->nready - 1) * sizeof(int));
        S->nready--;
        pthread_mutex_unlock(&S->mu);
        run_job(&S->g->jobs[j]);
        pthread_mutex_lock(&S->mu);
        job *J = &S->g->jobs[j];
        for (int r = 0; r < J->nrev; r++) {
            int t = J->rev[r];
            if (--S->g->jobs[t].ndeps == 0) {     /* new sources are prepended (job.c:573-577) */
                memmove(S->ready + 1, S->ready, (size_t)S->nready * sizeof(int));
                S->ready[0] = t;
                S->nready++;
            }
        }
        S->done++;
        pthread_cond_broadcast(&S->cv);
    }
    pthread_mutex_unlock(&S->mu);
    return NULL;
}

/* Squarify (mat-cpu/matj.c:680-780) is not needed for the square power-of-two
 * shapes this baseline is timed on; other shapes run one sequential job. */
int or_matcpu_winograd(int64_t N, olimb *Cout, const olimb *Ain, const olimb *Bin, int64_t cap,
                       int64_t minblock, int64_t jobsize, int nthreads)
{
    return or_matcpu_winograd_timed(N, Cout, Ain, Bin, cap, minblock, jobsize, nthreads, NULL);
}

static void to_gmp(gmpz *z, const olimb *e, int64_t cap)
{
    int64_t s = ((const int64_t *)e)[cap];
    int64_t n = s < 0 ? -s : s;
    G.import(z, (size_t)n, -1, 8, 0, 0, e);
    if (s < 0) z->size = -z->size;
}

static int from_gmp(olimb *e, int64_t cap, const gmpz *z)
{
    int64_t n = z->size < 0 ? -z->size : z->size;
    if (n > cap) return 1;
    size_t cnt = 0;
    if (n) G.export_(e, &cnt, -1, 8, 0, 0, z);
    ((int64_t *)e)[cap] = z->size < 0 ? -(int64_t)cnt : (int64_t)cnt;
    return 0;
}

int or_matcpu_winograd_timed(int64_t N, olimb *Cout, const olimb *Ain, const olimb *Bin, int64_t cap,
                             int64_t minblock, int64_t jobsize, int nthreads, double *seconds)
{
    if (gmp_load()) return 3;
    if (nthreads < 1) nthreads = 1;
    int64_t W = cap + 1;
    mat A = m_init(N), B = m_init(N), C = m_init(N);
    for (int64_t i = 0; i < N * N; i++) { to_gmp(&A[i], Ain + i * W, cap); to_gmp(&B[i], Bin + i * W, cap); }

    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    dag g = {NULL, 0, 0, minblock, jobsize};
    deps none = {NULL, 0};
    build(&g, N, C, N, A, N, B, N, 0, none);
    sched S;
    S.g = &g;
    pthread_mutex_init(&S.mu, NULL);
    pthread_cond_init(&S.cv, NULL);
    S.ready = (int *)malloc((size_t)g.njobs * sizeof(int));
    S.nready = 0; S.running = 0; S.done = 0;
    for (int i = g.njobs - 1; i >= 0; i--) if (g.jobs[i].ndeps == 0) S.ready[S.nready++] = i;
    /* ready list in submission order, first submitted first */
    for (int i = 0; i < S.nready / 2; i++) { int t = S.ready[i]; S.ready[i] = S.ready[S.nready - 1 - i]; S.ready[S.nready - 1 - i] = t; }
    pthread_t *th = (pthread_t *)malloc((size_t)nthreads * sizeof(pthread_t));
    for (int t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, worker, &S);
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    if (seconds) *seconds = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);

    int st = 0;
    for (int64_t i = 0; i < N * N && !st; i++) st = from_gmp(Cout + i * W, cap, &C[i]);
    for (int i = 0; i < g.njobs; i++) free(g.jobs[i].rev);
    free(g.jobs); free(S.ready); free(th);
    m_free(N, A); m_free(N, B); m_free(N, C);
    return st;
}

/* A bounded sample of the same product: the exact entries C[y][x] for the
 * listed (y, x) pairs, by GMP dot products (mat-cpu/mat.c:72-82, method 1),
 * spread over nthreads threads.  Returns the wall time in *seconds. */
typedef struct { int64_t N, cap, cnt, tid, nt; const olimb *A, *B; const int64_t *yx; olimb *out; int st; } samp;

static void *samp_worker(void *arg)
{
    samp *s = (samp *)arg;
    int64_t W = s->cap + 1;
    gmpz acc, a, b;
    G.init(&acc); G.init(&a); G.init(&b);
    for (int64_t i = s->tid; i < s->cnt; i += s->nt) {
        int64_t y = s->yx[2 * i], x = s->yx[2 * i + 1];
        acc.size = 0;
        for (int64_t k = 0; k < s->N; k++) {
            to_gmp(&a, s->A + (y * s->N + k) * W, s->cap);
            to_gmp(&b, s->B + (k * s->N + x) * W, s->cap);
            G.addmul(&acc, &a, &b);
        }
        if (from_gmp(s->out + i * W, s->cap, &acc)) s->st = 1;
    }
    G.clear(&acc); G.clear(&a); G.clear(&b);
    return NULL;
}

int or_matcpu_sample(int64_t N, const olimb *A, const olimb *B, int64_t cap, const int64_t *yx, int64_t cnt,
                     olimb *out, int nthreads, double *seconds)
{
    if (gmp_load()) return 3;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    samp ss[256];
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int t = 0; t < nthreads; t++) {
        samp s = {N, cap, cnt, t, nthreads, A, B, yx, out, 0};
        ss[t] = s;
        pthread_create(&th[t], NULL, samp_worker, &ss[t]);
    }
    int st = 0;
    for (int t = 0; t < nthreads; t++) { pthread_join(th[t], NULL); st |= ss[t].st; }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    if (seconds) *seconds = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
    return st;
}
