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
    int toom;   // Toom-22 block cutoff for the leaf products (< 0: MGX_PARAM_TOOM22)
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
        return plain_mul(c.strided, addmul, N, N, N, C, Crs, A, Ars, B, Brs, c.cap, (opbits + 31) / 32, c.s, c.toom);
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
        return plain_mul(false, false, K, M, N, C, Crs, A, Ars, B, Brs, c.cap, (opbits + 31) / 32, c.s, c.toom);
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
    c.toom = -1;
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
                  int64_t maxdepth, hipStream_t s, int toom)
{
    Ctx c = make_ctx(s, cap, false);
    c.leaf = std::max<int64_t>(leaf, 1);
    c.maxdepth = maxdepth;
    c.toom = toom;
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

}  // extern "C"

namespace {

// C = A*B on device matrices (rowstrides = widths) with a mat-gpu method code
// (mat-gpu/main.cu:358-502); C must start zeroed for the methods that
// accumulate.  Interleaved methods need a square side that is a multiple of
// the interleave block; other shapes take the same recursion uninterleaved.
int device_method(int method, int64_t K, int64_t M, int64_t N, uint64_t *dC, uint64_t *dA, uint64_t *dB, int64_t cap,
                  hipStream_t s)
{
    const bool square = K == M && M == N;
    switch (method) {
    case 1:   // host arithmetic in the reference; here the device naive kernel
    case 2: return mgx_naive_smallrect(K, M, N, dC, N, dA, M, dB, N, cap, s);
    case 3: return mgx_squarify_mul(MGX_REC_STRASSEN, K, M, N, dC, N, dA, M, dB, N, cap, s);
    case 35: return mgx_squarify_mul(MGX_REC_WINOGRAD, K, M, N, dC, N, dA, M, dB, N, cap, s);
    case 4: return mgx_squarify_mul(MGX_REC_NAIVE, K, M, N, dC, N, dA, M, dB, N, cap, s);
    case 37:
    case 357:
    case 47: {
        const int rm = method == 37 ? MGX_REC_STRASSEN : method == 357 ? MGX_REC_WINOGRAD : MGX_REC_NAIVE;
        if (!square || N % MGX_INTERLEAVE_BLOCK) return mgx_squarify_mul(rm, K, M, N, dC, N, dA, M, dB, N, cap, s);
        int64_t bits = 0;
        MGX_TRY(input_bits(N, N, N, dA, N, dB, N, cap, s, &bits));
        MGX_TRY(mgx_interleave(dA, N * N, cap, s));
        MGX_TRY(mgx_interleave(dB, N * N, cap, s));
        MGX_TRY(recursive_nosync(rm, true, N, dC, N, dA, N, dB, N, cap, bits, s));
        MGX_TRY(mgx_deinterleave(dC, N * N, cap, s));
        // leave the inputs as they came
        MGX_TRY(mgx_deinterleave(dA, N * N, cap, s));
        MGX_TRY(mgx_deinterleave(dB, N * N, cap, s));
        return check_device_error(s);
    }
    case 8: return mgx_crt_mfma_matmul(nullptr, K, M, N, dC, N, dA, M, dB, N, cap, s);
    case 0: return mgx_matmul_route(MGX_ROUTE_AUTO, K, M, N, dC, N, dA, M, dB, N, cap, s);
    default: return fail(MGX_ERR_ARG, "invalid method %d", method);
    }
}

// ---- out of core: matrices in host memory, blocks staged through device scratch ----
int put_block(uint64_t *d, const uint64_t *h, int64_t rows, int64_t cols, int64_t hrs, int64_t W, hipStream_t s)
{
    if (rows * cols)
        MGX_HIP(hipMemcpy2DAsync(d, (size_t)(cols * W * 8), h, (size_t)(hrs * W * 8), (size_t)(cols * W * 8),
                                 (size_t)rows, hipMemcpyHostToDevice, s));
    return MGX_OK;
}
int get_block(uint64_t *h, int64_t hrs, const uint64_t *d, int64_t rows, int64_t cols, int64_t W, hipStream_t s)
{
    if (rows * cols)
        MGX_HIP(hipMemcpy2DAsync(h, (size_t)(hrs * W * 8), d, (size_t)(cols * W * 8), (size_t)(cols * W * 8),
                                 (size_t)rows, hipMemcpyDeviceToHost, s));
    return MGX_OK;
}

struct HostCtx {
    int method;      // mat-gpu method code run on each transferred block
    int64_t T, cap;
    hipStream_t s;
};

// Device scratch for one staged step, released at scope end.
struct DevBlocks {
    hipStream_t s;
    uint64_t *p[4] = {nullptr, nullptr, nullptr, nullptr};
    explicit DevBlocks(hipStream_t s_) : s(s_) {}
    int take(int i, int64_t entries, int64_t W) { return scratch_alloc((void **)&p[i], (size_t)std::max<int64_t>(entries, 1) * W * 8, s); }
    ~DevBlocks()
    {
        for (auto q : p) scratch_free(q, s);
    }
};

// dst = a +- b on h x w host blocks (dst may be a or b)
int host_aors(const HostCtx &c, bool sub, int64_t h, int64_t w, uint64_t *dst, int64_t drs, const uint64_t *a,
              int64_t ars, const uint64_t *b, int64_t brs)
{
    const int64_t W = c.cap + 1;
    DevBlocks d(c.s);
    MGX_TRY(d.take(0, h * w, W));
    MGX_TRY(d.take(1, h * w, W));
    MGX_TRY(put_block(d.p[0], a, h, w, ars, W, c.s));
    MGX_TRY(put_block(d.p[1], b, h, w, brs, W, c.s));
    MGX_TRY(plain_aors(false, sub, w, h, d.p[0], w, d.p[0], w, d.p[1], w, c.cap, c.s));
    MGX_TRY(get_block(dst, drs, d.p[0], h, w, W, c.s));
    return check_device_error(c.s);
}

// C (=|+=) A*B for K x M and M x N host blocks small enough to transfer whole
int device_block(const HostCtx &c, bool addmul, int64_t K, int64_t M, int64_t N, uint64_t *C, int64_t Crs,
                 const uint64_t *A, int64_t Ars, const uint64_t *B, int64_t Brs)
{
    const int64_t W = c.cap + 1;
    DevBlocks d(c.s);
    MGX_TRY(d.take(0, K * M, W));
    MGX_TRY(d.take(1, M * N, W));
    MGX_TRY(d.take(2, K * N, W));
    MGX_TRY(put_block(d.p[0], A, K, M, Ars, W, c.s));
    MGX_TRY(put_block(d.p[1], B, M, N, Brs, W, c.s));
    MGX_HIP(hipMemsetAsync(d.p[2], 0, (size_t)(K * N * W * 8), c.s));
    MGX_TRY(device_method(c.method, K, M, N, d.p[2], d.p[0], d.p[1], c.cap, c.s));
    if (addmul) {   // C += the block product
        MGX_TRY(d.take(3, K * N, W));
        MGX_TRY(put_block(d.p[3], C, K, N, Crs, W, c.s));
        MGX_TRY(plain_aors(false, false, N, K, d.p[2], N, d.p[2], N, d.p[3], N, c.cap, c.s));
    }
    MGX_TRY(get_block(C, Crs, d.p[2], K, N, W, c.s));
    return check_device_error(c.s);
}

// Pinned host temporaries of the host-level recursion.
struct HostTemp {
    uint64_t *p = nullptr;
    ~HostTemp()
    {
        if (p) (void)hipHostFree(p);
    }
};

// C (=|+=) A*B, N x N in host memory: the method's 2 x 2 sequence
// (mat-gpu/mat-defines.h) over host blocks until a side is at most T
// (mat-gpu/mat-pregpu.cu:10-96 transfer_recurse and the *_pregpu drivers).
int host_rec(const HostCtx &c, int method, bool addmul, int64_t N, uint64_t *C, int64_t Crs, const uint64_t *A,
             int64_t Ars, const uint64_t *B, int64_t Brs)
{
    if (N <= c.T || (N & 1)) return device_block(c, addmul, N, N, N, C, Crs, A, Ars, B, Brs);
    const int64_t h = N / 2, W = c.cap + 1;
    HostTemp T;
    if (hipHostMalloc((void **)&T.p, (size_t)(N * N * W * 8), hipHostMallocDefault) != hipSuccess)
        return fail(MGX_ERR_OOM, "pinned host temporary of %lld entries", (long long)(N * N));
    if (addmul && method != MGX_REC_WINOGRAD) {   // naive / Strassen addmul: T = A*B, then C += T
        MGX_TRY(host_rec(c, method, false, N, T.p, N, A, Ars, B, Brs));
        return host_aors(c, false, N, N, C, Crs, C, Crs, T.p, N);
    }
    const Step *seq;
    size_t ns;
    switch (method) {
    case MGX_REC_NAIVE: seq = seq_naive; ns = sizeof seq_naive / sizeof(Step); break;
    case MGX_REC_STRASSEN: seq = seq_strassen; ns = sizeof seq_strassen / sizeof(Step); break;
    default:
        if (addmul) { seq = seq_winograd_addmul; ns = sizeof seq_winograd_addmul / sizeof(Step); }
        else { seq = seq_winograd; ns = sizeof seq_winograd / sizeof(Step); }
        break;
    }
    uint64_t *ptr[16];
    int64_t rs[16];
    for (int q = 0; q < 4; q++) {
        const int64_t oy = (q >> 1) * h, ox = (q & 1) * h;
        ptr[A11 + q] = (uint64_t *)A + (oy * Ars + ox) * W; rs[A11 + q] = Ars;
        ptr[B11 + q] = (uint64_t *)B + (oy * Brs + ox) * W; rs[B11 + q] = Brs;
        ptr[C11 + q] = C + (oy * Crs + ox) * W; rs[C11 + q] = Crs;
        ptr[P1 + q] = T.p + (oy * N + ox) * W; rs[P1 + q] = N;
    }
    for (size_t i = 0; i < ns; i++) {
        const Step &p = seq[i];
        if (p.op == OADD || p.op == OSUB)
            MGX_TRY(host_aors(c, p.op == OSUB, h, h, ptr[p.d], rs[p.d], ptr[p.a], rs[p.a], ptr[p.b], rs[p.b]));
        else
            MGX_TRY(host_rec(c, method, p.op == OAML, h, ptr[p.d], rs[p.d], ptr[p.a], rs[p.a], ptr[p.b], rs[p.b]));
    }
    return MGX_OK;
}

// Any shape: T x T tiles of C, each accumulated on the device over the
// T-wide slices of the inner dimension.
int host_tiled(const HostCtx &c, int64_t K, int64_t M, int64_t N, uint64_t *C, const uint64_t *A, const uint64_t *B)
{
    const int64_t W = c.cap + 1, T = c.T;
    DevBlocks d(c.s);
    MGX_TRY(d.take(0, T * T, W));
    MGX_TRY(d.take(1, T * T, W));
    MGX_TRY(d.take(2, T * T, W));
    MGX_TRY(d.take(3, T * T, W));
    for (int64_t i0 = 0; i0 < K; i0 += T)
        for (int64_t j0 = 0; j0 < N; j0 += T) {
            const int64_t ki = std::min(T, K - i0), nj = std::min(T, N - j0);
            MGX_HIP(hipMemsetAsync(d.p[2], 0, (size_t)(ki * nj * W * 8), c.s));
            for (int64_t k0 = 0; k0 < M; k0 += T) {
                const int64_t mk = std::min(T, M - k0);
                MGX_TRY(put_block(d.p[0], A + (i0 * M + k0) * W, ki, mk, M, W, c.s));
                MGX_TRY(put_block(d.p[1], B + (k0 * N + j0) * W, mk, nj, N, W, c.s));
                MGX_HIP(hipMemsetAsync(d.p[3], 0, (size_t)(ki * nj * W * 8), c.s));
                MGX_TRY(device_method(c.method, ki, mk, nj, d.p[3], d.p[0], d.p[1], c.cap, c.s));
                MGX_TRY(plain_aors(false, false, nj, ki, d.p[2], nj, d.p[2], nj, d.p[3], nj, c.cap, c.s));
            }
            MGX_TRY(get_block(C + (i0 * N + j0) * W, N, d.p[2], ki, nj, W, c.s));
            MGX_TRY(check_device_error(c.s));
        }
    return MGX_OK;
}

}  // namespace

extern "C" {

// mat-gpu/main.cu:358-502 method codes on host matrices (rowstrides = widths)
int mgx_matmul_host(int method, int64_t K, int64_t M, int64_t N, mgx_limb *C, const mgx_limb *A, const mgx_limb *B,
                    int64_t cap, double *seconds)
{
    MGX_TRY(ensure_init());
    if (K < 0 || M < 0 || N < 0 || cap < 1) return fail(MGX_ERR_SHAPE, "bad shape");
    switch (method) {
    case 0: case 1: case 2: case 3: case 35: case 4: case 37: case 357: case 47: case 8: break;
    default: return fail(MGX_ERR_ARG, "invalid method %d", method);
    }
    if ((method == 37 || method == 357 || method == 47) && (K != M || M != N || N % MGX_INTERLEAVE_BLOCK))
        return fail(MGX_ERR_SHAPE, "method %d needs square N %% %d == 0", method, MGX_INTERLEAVE_BLOCK);
    auto t0 = std::chrono::steady_clock::now();
    hipStream_t s = nullptr;
    MGX_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    const int64_t W = cap + 1, T = param(MGX_PARAM_TRANSFER);
    int st = MGX_OK;
    if (T > 0 && std::max(std::max(K, M), N) > T) {
        // out of core: the host keeps the matrices, blocks of side <= T move
        HostCtx c{method, T, cap, s};
        const bool square = K == M && M == N;
        const int rm = method == 3 || method == 37 ? MGX_REC_STRASSEN
                     : method == 35 || method == 357 ? MGX_REC_WINOGRAD
                     : method == 4 || method == 47 ? MGX_REC_NAIVE : -1;
        if (square && rm >= 0) st = host_rec(c, rm, false, N, C, N, A, N, B, N);
        else st = host_tiled(c, K, M, N, C, A, B);
    } else {
        uint64_t *dA = nullptr, *dB = nullptr, *dC = nullptr;
        auto al = [&](uint64_t **p, int64_t n) {
            if (st == MGX_OK && hipMalloc(p, (size_t)std::max<int64_t>(n, 1) * W * 8) != hipSuccess)
                st = fail(MGX_ERR_OOM, "device allocation of %lld entries", (long long)n);
        };
        al(&dA, K * M); al(&dB, M * N); al(&dC, K * N);
        if (st == MGX_OK && K * M) st = hipMemcpyAsync(dA, A, (size_t)(K * M * W * 8), hipMemcpyHostToDevice, s) == hipSuccess ? MGX_OK : MGX_ERR_HIP;
        if (st == MGX_OK && M * N) st = hipMemcpyAsync(dB, B, (size_t)(M * N * W * 8), hipMemcpyHostToDevice, s) == hipSuccess ? MGX_OK : MGX_ERR_HIP;
        if (st == MGX_OK && K * N) st = hipMemsetAsync(dC, 0, (size_t)(K * N * W * 8), s) == hipSuccess ? MGX_OK : MGX_ERR_HIP;
        if (st == MGX_OK && K * N) st = device_method(method, K, M, N, dC, dA, dB, cap, s);
        if (st == MGX_OK && K * N)
            st = hipMemcpyAsync(C, dC, (size_t)(K * N * W * 8), hipMemcpyDeviceToHost, s) == hipSuccess ? MGX_OK : MGX_ERR_HIP;
        hipStreamSynchronize(s);
        hipFree(dA); hipFree(dB); hipFree(dC);
    }
    hipStreamSynchronize(s);
    hipStreamDestroy(s);
    if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return st;
}

}  // extern "C"
