// rec.cpp -- matrix-level plain route: 2x2 block recursion (naive, Strassen,
// Winograd, Winograd-addmul), squarify for rectangular shapes, and the
// host-memory entry with mat-gpu's method codes.
//
// Everything is stream-ordered: temporaries come from the library's scratch
// cache and go back to it behind an event, no host synchronization happens
// inside the recursion (the reference allocates and synchronizes after every launch,
// mat-gpu/mat-gpu.cu:20,389).  The public matrix-level calls synchronize once
// at the end and report the device error word as a status.
#include "common.h"
#include "../../include/mgx_sched.h"
#include <chrono>
#include <algorithm>

namespace mgx {
int plain_mul(bool strided, bool addmul, int64_t K, int64_t M, int64_t N, uint64_t *C, int64_t Crs, const uint64_t *A,
              int64_t Ars, const uint64_t *B, int64_t Brs, int64_t cap, int64_t opwords, hipStream_t s);
int plain_aors(bool strided, bool sub, int64_t Wd, int64_t H, uint64_t *C, int64_t Crs, const uint64_t *A, int64_t Ars,
               const uint64_t *B, int64_t Brs, int64_t cap, hipStream_t s);
}

extern "C" int mgx_crt_scan_bits(int64_t rows, int64_t cols, const mgx_limb *X, int64_t Xrs, int64_t cap, int64_t *bits,
                                 void *stream);

using namespace mgx;

namespace {

enum { A11, A12, A21, A22, B11, B12, B21, B22, C11, C12, C21, C22, P1, P2, P3, P4 };
enum { OADD, OSUB, OMUL, OAML };
struct Step { unsigned char op, d, a, b; };

// mat-gpu/mat-defines.h:21-36
const Step seq_naive[] = {
    {OMUL, C11, A11, B11}, {OMUL, P1, A12, B21}, {OADD, C11, C11, P1},
    {OMUL, C12, A11, B12}, {OMUL, P1, A12, B22}, {OADD, C12, C12, P1},
    {OMUL, C21, A21, B11}, {OMUL, P1, A22, B21}, {OADD, C21, C21, P1},
    {OMUL, C22, A21, B12}, {OMUL, P1, A22, B22}, {OADD, C22, C22, P1},
};
// mat-gpu/mat-defines.h:77-117 (three temporaries)
const Step seq_strassen[] = {
    {OADD, P1, A11, A22}, {OADD, P2, B11, B22}, {OMUL, C11, P1, P2},
    {OSUB, P1, B12, B22}, {OMUL, C12, A11, P1}, {OADD, C22, C11, C12},
    {OADD, P1, A21, A22}, {OMUL, C21, P1, B11}, {OSUB, C22, C22, C21},
    {OSUB, P2, A12, A22}, {OADD, P3, B21, B22}, {OMUL, P1, P2, P3}, {OADD, C11, C11, P1},
    {OSUB, P2, A21, A11}, {OADD, P3, B11, B12}, {OMUL, P1, P2, P3}, {OADD, C22, C22, P1},
    {OSUB, P2, B21, B11}, {OMUL, P1, A22, P2}, {OADD, C11, C11, P1}, {OADD, C21, C21, P1},
    {OADD, P2, A11, A12}, {OMUL, P1, P2, B22}, {OSUB, C11, C11, P1}, {OADD, C12, C12, P1},
};
// mat-gpu/mat-defines.h:144-164
const Step seq_winograd[] = {
    {OADD, P1, A21, A22}, {OSUB, P2, B12, B11}, {OMUL, C22, P1, P2},
    {OSUB, P1, P1, A11}, {OSUB, P2, B22, P2}, {OMUL, P3, A11, B11},
    {OMUL, C11, A12, B21}, {OADD, C11, C11, P3}, {OAML, P3, P1, P2},
    {OSUB, P1, A12, P1}, {OSUB, P2, B21, P2}, {OMUL, C12, P1, B22},
    {OADD, C12, C12, C22}, {OADD, C12, C12, P3}, {OMUL, C21, A22, P2},
    {OSUB, P1, A11, A21}, {OSUB, P2, B22, B12}, {OAML, P3, P1, P2},
    {OADD, C21, C21, P3}, {OADD, C22, C22, P3},
};
// mat-gpu/mat-defines.h:166-187
const Step seq_winograd_addmul[] = {
    {OADD, P1, A21, A22}, {OSUB, P2, B12, B11}, {OMUL, P3, P1, P2},
    {OADD, C12, C12, P3}, {OADD, C22, C22, P3}, {OSUB, P1, P1, A11},
    {OSUB, P2, B22, P2}, {OMUL, P3, A11, B11}, {OADD, C11, C11, P3},
    {OAML, P3, P1, P2}, {OAML, C11, A12, B21}, {OSUB, P1, A12, P1},
    {OSUB, P2, B21, P2}, {OAML, C12, P1, B22}, {OADD, C12, C12, P3},
    {OAML, C21, A22, P2}, {OSUB, P1, A11, A21}, {OSUB, P2, B22, B12},
    {OAML, P3, P1, P2}, {OADD, C21, C21, P3}, {OADD, C22, C22, P3},
};

struct Ctx {
    hipStream_t s;
    int64_t cap;
    bool strided;
    int64_t leaf;
    int64_t maxdepth;
};

bool can_split(const Ctx &c, int64_t N, int depth)
{
    if (N <= c.leaf || (N & 1)) return false;
    if (c.maxdepth >= 0 && depth >= c.maxdepth) return false;
    if (c.strided && (N / 2) % MGX_INTERLEAVE_BLOCK) return false;
    return true;
}

// C (=|+=) A*B, N x N; opbits bounds the operands' bit lengths at this level
int rec(const Ctx &c, int method, bool addmul, int64_t N, uint64_t *C, int64_t Crs, const uint64_t *A, int64_t Ars,
        const uint64_t *B, int64_t Brs, int64_t opbits, int depth)
{
    if (!can_split(c, N, depth))
        return plain_mul(c.strided, addmul, N, N, N, C, Crs, A, Ars, B, Brs, c.cap, (opbits + 31) / 32, c.s);
    const int64_t h = N / 2, W = c.cap + 1;
    const Step *seq;
    size_t ns;
    int sub_method = method;
    if (addmul && method != MGX_REC_WINOGRAD && method != MGX_REC_WINOGRAD_ADDMUL) {
        // naive / Strassen addmul: T = A*B, then C += T
        uint64_t *T = nullptr;
        MGX_TRY(scratch_alloc((void **)&T, (size_t)(N * N * W * 8), c.s));
        int st = rec(c, method, false, N, T, N, A, Ars, B, Brs, opbits, depth);
        if (st == MGX_OK) st = plain_aors(c.strided, false, N, N, C, Crs, C, Crs, T, N, c.cap, c.s);
        scratch_free(T, c.s);
        return st;
    }
    switch (method) {
    case MGX_REC_NAIVE: seq = seq_naive; ns = sizeof seq_naive / sizeof(Step); break;
    case MGX_REC_STRASSEN: seq = seq_strassen; ns = sizeof seq_strassen / sizeof(Step); break;
    default:
        if (addmul) { seq = seq_winograd_addmul; ns = sizeof seq_winograd_addmul / sizeof(Step); }
        else { seq = seq_winograd; ns = sizeof seq_winograd / sizeof(Step); }
        sub_method = MGX_REC_WINOGRAD;
        break;
    }
    uint64_t *T = nullptr;
    MGX_TRY(scratch_alloc((void **)&T, (size_t)(N * N * W * 8), c.s));
    uint64_t *ptr[16];
    int64_t rs[16];
    for (int q = 0; q < 4; q++) {
        int64_t oy = (q >> 1) * h, ox = (q & 1) * h;
        ptr[A11 + q] = (uint64_t *)A + (oy * Ars + ox) * W; rs[A11 + q] = Ars;
        ptr[B11 + q] = (uint64_t *)B + (oy * Brs + ox) * W; rs[B11 + q] = Brs;
        ptr[C11 + q] = C + (oy * Crs + ox) * W; rs[C11 + q] = Crs;
        ptr[P1 + q] = T + (oy * N + ox) * W; rs[P1 + q] = N;
    }
    const int64_t subbits = opbits + 2;   // operand sums grow by at most two bits per level
    int st = MGX_OK;
    for (size_t i = 0; i < ns && st == MGX_OK; i++) {
        const Step &p = seq[i];
        switch (p.op) {
        case OADD:
        case OSUB:
            st = plain_aors(c.strided, p.op == OSUB, h, h, ptr[p.d], rs[p.d], ptr[p.a], rs[p.a], ptr[p.b], rs[p.b], c.cap, c.s);
            break;
        case OMUL:
            st = rec(c, sub_method, false, h, ptr[p.d], rs[p.d], ptr[p.a], rs[p.a], ptr[p.b], rs[p.b], subbits, depth + 1);
            break;
        case OAML:
            st = rec(c, sub_method, true, h, ptr[p.d], rs[p.d], ptr[p.a], rs[p.a], ptr[p.b], rs[p.b], subbits, depth + 1);
            break;
        }
    }
    scratch_free(T, c.s);
    return st;
}

int64_t odd_part(int64_t n)
{
    while (n > 0 && !(n & 1)) n >>= 1;
    return n;
}

int squarify(const Ctx &c, int method, int64_t K, int64_t M, int64_t N, uint64_t *C, int64_t Crs, const uint64_t *A,
             int64_t Ars, const uint64_t *B, int64_t Brs, int64_t opbits)
{
    if (K == 0 || M == 0 || N == 0) return MGX_OK;
    if (K <= c.leaf || M <= c.leaf || N <= c.leaf)
        return plain_mul(false, false, K, M, N, C, Crs, A, Ars, B, Brs, c.cap, (opbits + 31) / 32, c.s);
    if (K == M && M == N && odd_part(K) <= c.leaf) return rec(c, method, false, K, C, Crs, A, Ars, B, Brs, opbits, 0);
    int64_t mind = std::min(std::min(K, M), N), lt = 1;
    while (lt * 2 <= mind) lt *= 2;
    const int64_t W = c.cap + 1;
    uint64_t *T = nullptr;
    MGX_TRY(scratch_alloc((void **)&T, (size_t)(K * N * W * 8), c.s));
    auto sub = [&](const uint64_t *P, int64_t rs, int64_t iy, int64_t ix) { return (uint64_t *)P + (iy * lt * rs + ix * lt) * W; };
    const bool tK = lt < K, tM = lt < M, tN = lt < N;
    int st = MGX_OK;
    // mat-gpu/mat-pregpu.cu:505-523
    auto MUL = [&](int64_t k, int64_t m, int64_t n, uint64_t *cc, int64_t crs, const uint64_t *aa, int64_t ars,
                   const uint64_t *bb, int64_t brs) {
        if (st == MGX_OK) st = squarify(c, method, k, m, n, cc, crs, aa, ars, bb, brs, opbits);
    };
    auto ADD = [&](int64_t hh, int64_t ww, uint64_t *cc, int64_t crs, const uint64_t *aa, int64_t ars, const uint64_t *bb,
                   int64_t brs) {
        if (st == MGX_OK) st = plain_aors(false, false, ww, hh, cc, crs, aa, ars, bb, brs, c.cap, c.s);
    };
    MUL(lt, lt, lt, sub(C, Crs, 0, 0), Crs, sub(A, Ars, 0, 0), Ars, sub(B, Brs, 0, 0), Brs);
    if (tM) { MUL(lt, M - lt, lt, sub(T, N, 0, 0), N, sub(A, Ars, 0, 1), Ars, sub(B, Brs, 1, 0), Brs);
              ADD(lt, lt, sub(C, Crs, 0, 0), Crs, sub(C, Crs, 0, 0), Crs, sub(T, N, 0, 0), N); }
    if (tN) {
        MUL(lt, lt, N - lt, sub(C, Crs, 0, 1), Crs, sub(A, Ars, 0, 0), Ars, sub(B, Brs, 0, 1), Brs);
        if (tM) { MUL(lt, M - lt, N - lt, sub(T, N, 0, 1), N, sub(A, Ars, 0, 1), Ars, sub(B, Brs, 1, 1), Brs);
                  ADD(lt, N - lt, sub(C, Crs, 0, 1), Crs, sub(C, Crs, 0, 1), Crs, sub(T, N, 0, 1), N); }
    }
    if (tK) {
        MUL(K - lt, lt, lt, sub(C, Crs, 1, 0), Crs, sub(A, Ars, 1, 0), Ars, sub(B, Brs, 0, 0), Brs);
        if (tM) { MUL(K - lt, M - lt, lt, sub(T, N, 1, 0), N, sub(A, Ars, 1, 1), Ars, sub(B, Brs, 1, 0), Brs);
                  ADD(K - lt, lt, sub(C, Crs, 1, 0), Crs, sub(C, Crs, 1, 0), Crs, sub(T, N, 1, 0), N); }
        if (tN) {
            MUL(K - lt, lt, N - lt, sub(C, Crs, 1, 1), Crs, sub(A, Ars, 1, 0), Ars, sub(B, Brs, 0, 1), Brs);
            if (tM) { MUL(K - lt, M - lt, N - lt, sub(T, N, 1, 1), N, sub(A, Ars, 1, 1), Ars, sub(B, Brs, 1, 1), Brs);
                      ADD(K - lt, N - lt, sub(C, Crs, 1, 1), Crs, sub(C, Crs, 1, 1), Crs, sub(T, N, 1, 1), N); }
        }
    }
    scratch_free(T, c.s);
    return st;
}

int input_bits(int64_t K, int64_t M, int64_t N, const uint64_t *A, int64_t Ars, const uint64_t *B, int64_t Brs,
               int64_t cap, hipStream_t s, int64_t *bits)
{
    int64_t ba = 0, bb = 0;
    MGX_TRY(mgx_crt_scan_bits(K, M, (const mgx_limb *)A, Ars, cap, &ba, s));
    MGX_TRY(mgx_crt_scan_bits(M, N, (const mgx_limb *)B, Brs, cap, &bb, s));
    *bits = std::max<int64_t>(std::max(ba, bb), 1);
    return MGX_OK;
}

Ctx make_ctx(hipStream_t s, int64_t cap, bool strided)
{
    Ctx c;
    c.s = s;
    c.cap = cap;
    c.strided = strided;
    c.leaf = std::max<int64_t>(param(MGX_PARAM_LEAF), 1);
    if (strided) c.leaf = std::max<int64_t>(c.leaf, MGX_INTERLEAVE_BLOCK);
    c.maxdepth = param(MGX_PARAM_WINOGRAD_DEPTH);
    return c;
}

}  // namespace

namespace mgx {
// operand bits known: no scan, no sync (used by the batch path and the tuner)
int recursive_nosync(int method, bool strided, int64_t N, uint64_t *C, int64_t Crs, const uint64_t *A, int64_t Ars,
                     const uint64_t *B, int64_t Brs, int64_t cap, int64_t opbits, hipStream_t s)
{
    Ctx c = make_ctx(s, cap, strided);
    bool addmul = method == MGX_REC_WINOGRAD_ADDMUL;
    return rec(c, method, addmul, N, C, Crs, A, Ars, B, Brs, opbits, 0);
}
int squarify_nosync(int method, int64_t K, int64_t M, int64_t N, uint64_t *C, int64_t Crs, const uint64_t *A,
                    int64_t Ars, const uint64_t *B, int64_t Brs, int64_t cap, int64_t opbits, hipStream_t s)
{
    Ctx c = make_ctx(s, cap, false);
    return squarify(c, method, K, M, N, C, Crs, A, Ars, B, Brs, opbits);
}
// the same with the leaf size and depth given per call (the routes and the
// tuner; the global tunables are left alone, so concurrent calls do not race)
int squarify_with(int method, int64_t K, int64_t M, int64_t N, uint64_t *C, int64_t Crs, const uint64_t *A,
                  int64_t Ars, const uint64_t *B, int64_t Brs, int64_t cap, int64_t opbits, int64_t leaf,
                  int64_t maxdepth, hipStream_t s)
{
    Ctx c = make_ctx(s, cap, false);
    c.leaf = std::max<int64_t>(leaf, 1);
    c.maxdepth = maxdepth;
    return squarify(c, method, K, M, N, C, Crs, A, Ars, B, Brs, opbits);
}
}  // namespace mgx

extern "C" {

int mgx_check_error(void *stream);

#define MGX_AORS_ARGS int64_t Wd, int64_t H, mgx_limb *C, int64_t Crs, const mgx_limb *A, int64_t Ars, \
                      const mgx_limb *B, int64_t Brs, int64_t cap, void *stream
int mgx_add_matrices(MGX_AORS_ARGS)
{
    MGX_TRY(plain_aors(false, false, Wd, H, C, Crs, A, Ars, B, Brs, cap, as_stream(stream)));
    return check_device_error(as_stream(stream));
}
int mgx_sub_matrices(MGX_AORS_ARGS)
{
    MGX_TRY(plain_aors(false, true, Wd, H, C, Crs, A, Ars, B, Brs, cap, as_stream(stream)));
    return check_device_error(as_stream(stream));
}
int mgx_add_matrices_strided(MGX_AORS_ARGS)
{
    MGX_TRY(plain_aors(true, false, Wd, H, C, Crs, A, Ars, B, Brs, cap, as_stream(stream)));
    return check_device_error(as_stream(stream));
}
int mgx_sub_matrices_strided(MGX_AORS_ARGS)
{
    MGX_TRY(plain_aors(true, true, Wd, H, C, Crs, A, Ars, B, Brs, cap, as_stream(stream)));
    return check_device_error(as_stream(stream));
}

int mgx_naive_smallrect(int64_t K, int64_t M, int64_t N, mgx_limb *C, int64_t Crs, const mgx_limb *A, int64_t Ars,
                        const mgx_limb *B, int64_t Brs, int64_t cap, void *stream)
{
    hipStream_t s = as_stream(stream);
    int64_t bits;
    MGX_TRY(input_bits(K, M, N, A, Ars, B, Brs, cap, s, &bits));
    MGX_TRY(plain_mul(false, false, K, M, N, C, Crs, A, Ars, B, Brs, cap, (bits + 31) / 32, s));
    return check_device_error(s);
}

int mgx_recursive(int method, int strided, int64_t N, mgx_limb *C, int64_t Crs, const mgx_limb *A, int64_t Ars,
                  const mgx_limb *B, int64_t Brs, int64_t cap, void *stream)
{
    MGX_TRY(ensure_init());
    if (method < 0 || method > MGX_REC_WINOGRAD_ADDMUL) return fail(MGX_ERR_ARG, "unknown method %d", method);
    if (N < 0 || cap < 1) return fail(MGX_ERR_SHAPE, "bad shape");
    if (N == 0) return MGX_OK;
    if (strided && (N % MGX_INTERLEAVE_BLOCK || Crs % MGX_INTERLEAVE_BLOCK || Ars % MGX_INTERLEAVE_BLOCK ||
                    Brs % MGX_INTERLEAVE_BLOCK))
        return fail(MGX_ERR_SHAPE, "interleaved recursion needs N and rowstrides multiples of %d", MGX_INTERLEAVE_BLOCK);
    hipStream_t s = as_stream(stream);
    int64_t bits;
    if (strided) bits = 32 * 2 * cap;   // interleaved inputs: bound by capacity
    else MGX_TRY(input_bits(N, N, N, A, Ars, B, Brs, cap, s, &bits));
    MGX_TRY(recursive_nosync(method, strided != 0, N, C, Crs, A, Ars, B, Brs, cap, bits, s));
    return check_device_error(s);
}

#define MGX_REC_ARGS int64_t N, mgx_limb *C, int64_t Crs, const mgx_limb *A, int64_t Ars, const mgx_limb *B, \
                     int64_t Brs, int64_t cap, void *stream
int mgx_naive_recursive(MGX_REC_ARGS) { return mgx_recursive(MGX_REC_NAIVE, 0, N, C, Crs, A, Ars, B, Brs, cap, stream); }
int mgx_naive_recursive_strided(MGX_REC_ARGS) { return mgx_recursive(MGX_REC_NAIVE, 1, N, C, Crs, A, Ars, B, Brs, cap, stream); }
int mgx_strassen(MGX_REC_ARGS) { return mgx_recursive(MGX_REC_STRASSEN, 0, N, C, Crs, A, Ars, B, Brs, cap, stream); }
int mgx_strassen_strided(MGX_REC_ARGS) { return mgx_recursive(MGX_REC_STRASSEN, 1, N, C, Crs, A, Ars, B, Brs, cap, stream); }
int mgx_winograd(MGX_REC_ARGS) { return mgx_recursive(MGX_REC_WINOGRAD, 0, N, C, Crs, A, Ars, B, Brs, cap, stream); }
int mgx_winograd_strided(MGX_REC_ARGS) { return mgx_recursive(MGX_REC_WINOGRAD, 1, N, C, Crs, A, Ars, B, Brs, cap, stream); }
int mgx_winograd_addmul(MGX_REC_ARGS) { return mgx_recursive(MGX_REC_WINOGRAD_ADDMUL, 0, N, C, Crs, A, Ars, B, Brs, cap, stream); }
int mgx_winograd_addmul_strided(MGX_REC_ARGS) { return mgx_recursive(MGX_REC_WINOGRAD_ADDMUL, 1, N, C, Crs, A, Ars, B, Brs, cap, stream); }

int mgx_squarify_mul(int method, int64_t K, int64_t M, int64_t N, mgx_limb *C, int64_t Crs, const mgx_limb *A,
                     int64_t Ars, const mgx_limb *B, int64_t Brs, int64_t cap, void *stream)
{
    MGX_TRY(ensure_init());
    if (method < 0 || method > MGX_REC_WINOGRAD) return fail(MGX_ERR_ARG, "unknown method %d", method);
    if (K < 0 || M < 0 || N < 0 || cap < 1) return fail(MGX_ERR_SHAPE, "bad shape");
    hipStream_t s = as_stream(stream);
    if (K == 0 || N == 0) return MGX_OK;
    if (M == 0) {
        MGX_HIP(hipMemset2DAsync(C, (size_t)(Crs * (cap + 1) * 8), 0, (size_t)(N * (cap + 1) * 8), (size_t)K, s));
        return check_device_error(s);
    }
    int64_t bits;
    MGX_TRY(input_bits(K, M, N, A, Ars, B, Brs, cap, s, &bits));
    MGX_TRY(squarify_nosync(method, K, M, N, C, Crs, A, Ars, B, Brs, cap, bits, s));
    return check_device_error(s);
}

int mgx_crt_mfma_matmul(const mgx_crt_plan *plan, int64_t K, int64_t M, int64_t N, mgx_limb *C, int64_t Crs,
                        const mgx_limb *A, int64_t Ars, const mgx_limb *B, int64_t Brs, int64_t cap, void *stream);

// mat-gpu/main.cu:358-502 method codes on host matrices (rowstrides = widths)
int mgx_matmul_host(int method, int64_t K, int64_t M, int64_t N, mgx_limb *C, const mgx_limb *A, const mgx_limb *B,
                    int64_t cap, double *seconds)
{
    MGX_TRY(ensure_init());
    auto t0 = std::chrono::steady_clock::now();
    hipStream_t s = nullptr;
    MGX_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    const int64_t W = cap + 1;
    uint64_t *dA = nullptr, *dB = nullptr, *dC = nullptr;
    int st = MGX_OK;
    auto al = [&](uint64_t **p, int64_t n) {
        if (st == MGX_OK && hipMalloc(p, (size_t)std::max<int64_t>(n, 1) * W * 8) != hipSuccess)
            st = fail(MGX_ERR_OOM, "device allocation of %lld entries", (long long)n);
    };
    al(&dA, K * M); al(&dB, M * N); al(&dC, K * N);
    if (st == MGX_OK && K * M) st = hipMemcpyAsync(dA, A, (size_t)(K * M * W * 8), hipMemcpyHostToDevice, s) == hipSuccess ? MGX_OK : MGX_ERR_HIP;
    if (st == MGX_OK && M * N) st = hipMemcpyAsync(dB, B, (size_t)(M * N * W * 8), hipMemcpyHostToDevice, s) == hipSuccess ? MGX_OK : MGX_ERR_HIP;
    if (st == MGX_OK && K * N) st = hipMemsetAsync(dC, 0, (size_t)(K * N * W * 8), s) == hipSuccess ? MGX_OK : MGX_ERR_HIP;
    const bool square = K == M && M == N;
    if (st == MGX_OK) {
        switch (method) {
        case 1:   // host arithmetic in the reference; here the device naive kernel
        case 2:
            st = mgx_naive_smallrect(K, M, N, dC, N, dA, M, dB, N, cap, s);
            break;
        case 3: st = mgx_squarify_mul(MGX_REC_STRASSEN, K, M, N, dC, N, dA, M, dB, N, cap, s); break;
        case 35: st = mgx_squarify_mul(MGX_REC_WINOGRAD, K, M, N, dC, N, dA, M, dB, N, cap, s); break;
        case 4: st = mgx_squarify_mul(MGX_REC_NAIVE, K, M, N, dC, N, dA, M, dB, N, cap, s); break;
        case 37:
        case 357:
        case 47: {
            if (!square || N % MGX_INTERLEAVE_BLOCK) { st = fail(MGX_ERR_SHAPE, "method %d needs square N %% %d == 0", method, MGX_INTERLEAVE_BLOCK); break; }
            int rm = method == 37 ? MGX_REC_STRASSEN : method == 357 ? MGX_REC_WINOGRAD : MGX_REC_NAIVE;
            int64_t bits = 0;
            st = input_bits(N, N, N, dA, N, dB, N, cap, s, &bits);
            if (st == MGX_OK) st = mgx_interleave(dA, N * N, cap, s);
            if (st == MGX_OK) st = mgx_interleave(dB, N * N, cap, s);
            if (st == MGX_OK) st = recursive_nosync(rm, true, N, dC, N, dA, N, dB, N, cap, bits, s);
            if (st == MGX_OK) st = mgx_deinterleave(dC, N * N, cap, s);
            if (st == MGX_OK) st = check_device_error(s);
            break;
        }
        case 8: st = mgx_crt_mfma_matmul(nullptr, K, M, N, dC, N, dA, M, dB, N, cap, s); break;
        case 0: st = mgx_matmul_route(MGX_ROUTE_AUTO, K, M, N, dC, N, dA, M, dB, N, cap, s); break;
        default: st = fail(MGX_ERR_ARG, "invalid method %d", method);
        }
    }
    if (st == MGX_OK && K * N)
        st = hipMemcpyAsync(C, dC, (size_t)(K * N * W * 8), hipMemcpyDeviceToHost, s) == hipSuccess ? MGX_OK : MGX_ERR_HIP;
    hipStreamSynchronize(s);
    hipFree(dA); hipFree(dB); hipFree(dC);
    hipStreamDestroy(s);
    if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return st;
}

}  // extern "C"
