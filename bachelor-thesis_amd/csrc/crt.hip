// crt.hip -- the multi-modular route on gfx950 INT8 MFMA.
//
// C = A*B for signed big-integer matrices, exactly, through P word-size
// primes p_q < 2^31 with M = prod p_q > 4 max|C|:
//
//   1. reduce:      every entry of A and B mod p_q, centred to |r| <= (p-1)/2
//                   and split into four balanced int8 digits (planar slices).
//                   Itself an int8 GEMM: entry bytes (offset by -128) times the
//                   balanced digits of 2^(8b) mod p_q.
//   2. residue GEMM: per prime, the 16 slice products A^(s) B^(t) on
//                   v_mfma_i32_32x32x32_i8, accumulated per diagonal s+t in
//                   int32 over K-chunks of at most 8192, folded mod p_q in
//                   64-bit integers (Barrett), then y_q = c_q (M/p_q)^-1 mod p_q
//                   is written as four balanced int8 digits.
//   3. reconstruct: explicit CRT, C = sum_q y_q (M/p_q) - beta M, again an int8
//                   GEMM (the digits of (M/p_q) 256^s against the digits of
//                   y_q) with the carry propagation fused into the epilogue;
//                   beta comes from a 32-bit fixed-point sum of y_q/p_q.
//
// No floating point anywhere.  The reference has no counterpart (its
// crt.cpp splits one modulus into a tree and is never called); the maths is
// the Chinese remainder theorem, Barrett reduction and balanced radix-256
// digits.
#include "common.h"
#include "bigint_host.h"
#include <vector>
#include <cstring>
#include <algorithm>

namespace mgx {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

struct PrimeConst {
    uint32_t p;
    uint32_t inv;    // (M/p)^-1 mod p
    uint64_t mu;     // floor(2^64 / p)
    uint64_t off;    // a multiple of p that is >= 2^62
    uint64_t fx;     // floor(2^63 / p)
    int32_t w[7];    // 2^(8d) mod p, centred
    uint32_t kofs;   // 128 * sum_{b<kb} 2^(8b) mod p
};

__device__ __forceinline__ uint32_t mod_u64(uint64_t u, uint32_t p, uint64_t mu)
{
    uint64_t q = __umul64hi(u, mu);          // q in {floor(u/p) - 1, floor(u/p)}
    uint64_t r = u - q * (uint64_t)p;        // r < 2p
    return (uint32_t)(r >= p ? r - p : r);
}
__device__ __forceinline__ uint32_t mod_s64(int64_t x, const PrimeConst &c)
{
    return mod_u64((uint64_t)(x + (int64_t)c.off), c.p, c.mu);   // |x| < 2^62
}
// four balanced base-256 digits of |r| < 2^30, byte s = digit s
__device__ __forceinline__ uint32_t slice4(int32_t r)
{
    uint32_t w = 0;
    int32_t v = r;
#pragma unroll
    for (int s = 0; s < 4; s++) {
        int32_t t = ((v + 128) & 255) - 128;
        w |= (uint32_t)(t & 255) << (8 * s);
        v = (v - t) >> 8;
    }
    return w;
}
__device__ __forceinline__ int32_t unslice4(uint32_t w)
{
    return (int32_t)(int8_t)(w & 255) + ((int32_t)(int8_t)((w >> 8) & 255) << 8) +
           ((int32_t)(int8_t)((w >> 16) & 255) << 16) + ((int32_t)(int8_t)(w >> 24) << 24);
}
__device__ __forceinline__ int32_t centre(uint32_t r, uint32_t p) { return r > (p >> 1) ? (int32_t)r - (int32_t)p : (int32_t)r; }

__device__ __forceinline__ v16i mfma32(v4i a, v4i b, v16i c)
{
    return __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c, 0, 0, 0);
}

constexpr int LDS_PITCH = 80;   // 64 data bytes + 16 padding per row

// ------------------------------------------------------------------------
// bit-length scan: out[0] = max over entries of bit length of |x|
// ------------------------------------------------------------------------
__global__ void scan_bits_kernel(const uint64_t *X, int64_t rs, int64_t W, int64_t rows, int64_t cols, int *out)
{
    int64_t n = rows * cols;
    int best = 0;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        int64_t r = e / cols, c = e % cols;
        const uint64_t *p = X + (r * rs + c) * W;
        int64_t s = (int64_t)p[W - 1];
        int64_t a = s < 0 ? -s : s;
        while (a > 0 && p[a - 1] == 0) a--;
        if (a > 0) {
            int b = (int)((a - 1) * 64 + 64 - __clzll(p[a - 1]));
            best = max(best, b);
        }
    }
    for (int o = 32; o > 0; o >>= 1) best = max(best, __shfl_xor(best, o));
    if ((threadIdx.x & 63) == 0) atomicMax(out, best);
}

// ------------------------------------------------------------------------
// 1. reduce.  Output rows (q,s) x entries; tile 64 rows x 128 entries.
// ------------------------------------------------------------------------
__global__ __launch_bounds__(256) void crt_reduce_kernel(
    const PrimeConst *__restrict__ pc, const int8_t *__restrict__ cst, int P, int nrb, int kb,
    const uint64_t *__restrict__ X, int64_t Xrs, int64_t W, int transpose, int64_t rows, int64_t K,
    int64_t kpad, int64_t nent, int8_t *__restrict__ R, int64_t plane, unsigned *__restrict__ err)
{
    __shared__ __attribute__((aligned(16))) int8_t LA[64 * LDS_PITCH];
    __shared__ __attribute__((aligned(16))) int8_t LB[128 * LDS_PITCH];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int rb = blockIdx.x % nrb;
    const int64_t e0 = (int64_t)(blockIdx.x / nrb) * 128;

    // the two entries this thread stages
    const uint64_t *ep[2];
    int64_t nlimb[2];
    for (int h = 0; h < 2; h++) {
        int64_t e = e0 + (tid >> 2) + 64 * h;
        int64_t r = e / kpad, k = e % kpad;
        ep[h] = nullptr; nlimb[h] = 0;
        if (e < nent && r < rows && k < K) {
            ep[h] = transpose ? X + (k * Xrs + r) * W : X + (r * Xrs + k) * W;
            int64_t s = (int64_t)ep[h][W - 1];
            nlimb[h] = s < 0 ? -s : s;
            if (rb == 0)   // limbs beyond the plan's bit bound must be zero
                for (int64_t li = kb / 8; li < nlimb[h]; li++)
                    if (ep[h][li]) { atomicOr(err, DEV_ERR_INPUT); break; }
        }
    }
    const int part = tid & 3;
    v16i acc[2];
    for (int i = 0; i < 2; i++) acc[i] = (v16i){0};

    for (int k0 = 0; k0 < kb; k0 += 64) {
        v4i a = *(const v4i *)(cst + (int64_t)(rb * 64 + (tid >> 2)) * kb + k0 + 16 * part);
        uint64_t b[2][2];
        for (int h = 0; h < 2; h++)
            for (int u = 0; u < 2; u++) {
                int64_t li = (k0 + 16 * part) / 8 + u;
                uint64_t v = (ep[h] && li < nlimb[h]) ? ep[h][li] : 0;
                b[h][u] = v ^ 0x8080808080808080ull;   // byte - 128
            }
        __syncthreads();
        *(v4i *)(LA + (tid >> 2) * LDS_PITCH + 16 * part) = a;
        for (int h = 0; h < 2; h++) {
            uint64_t *d = (uint64_t *)(LB + ((tid >> 2) + 64 * h) * LDS_PITCH + 16 * part);
            d[0] = b[h][0]; d[1] = b[h][1];
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 2; kk++) {
            v4i bf = *(const v4i *)(LB + (wave * 32 + (lane & 31)) * LDS_PITCH + kk * 32 + 16 * (lane >> 5));
#pragma unroll
            for (int rt = 0; rt < 2; rt++) {
                v4i af = *(const v4i *)(LA + (rt * 32 + (lane & 31)) * LDS_PITCH + kk * 32 + 16 * (lane >> 5));
                acc[rt] = mfma32(af, bf, acc[rt]);
            }
        }
    }

    // epilogue: lane owns entry column e, 4 primes x 4 slices per row tile
    const int64_t e = e0 + wave * 32 + (lane & 31);
    if (e >= nent) return;
    int64_t r = e / kpad, k = e % kpad;
    bool real = r < rows && k < K;
    bool neg = false;
    if (real) {
        const uint64_t *p = transpose ? X + (k * Xrs + r) * W : X + (r * Xrs + k) * W;
        neg = (int64_t)p[W - 1] < 0;
    }
#pragma unroll
    for (int rt = 0; rt < 2; rt++)
#pragma unroll
        for (int g = 0; g < 4; g++) {
            int q = rb * 16 + rt * 8 + 2 * g + (lane >> 5);
            if (q >= P) continue;
            uint32_t word = 0;
            if (real) {
                const PrimeConst c = pc[q];
                int64_t v = (int64_t)acc[rt][4 * g] + ((int64_t)acc[rt][4 * g + 1] << 8) +
                            ((int64_t)acc[rt][4 * g + 2] << 16) + ((int64_t)acc[rt][4 * g + 3] << 24) + c.kofs;
                uint32_t res = mod_s64(v, c);
                if (neg && res) res = c.p - res;
                word = slice4(centre(res, c.p));
            }
            int8_t *o = R + (int64_t)q * 4 * plane + e;
#pragma unroll
            for (int s = 0; s < 4; s++) o[s * plane] = (int8_t)(word >> (8 * s));
        }
}

// ------------------------------------------------------------------------
// 2. residue GEMM.  Tile 128 (i) x 64 (j) per prime; 4 waves of 64 x 32.
// ------------------------------------------------------------------------
constexpr int KCHUNK = 8192;   // 4 pairs * 8192 * 2^14 = 2^29 < 2^31

template <bool CHUNKED>
__global__ __launch_bounds__(256, 1) void crt_gemm_kernel(
    const PrimeConst *__restrict__ pc, const int8_t *__restrict__ RA, const int8_t *__restrict__ RB,
    int64_t planeA, int64_t planeB, int64_t kpad, int64_t rowsC, int64_t colsC, int ntm, int ntn,
    int64_t total, uint32_t *__restrict__ Y, int64_t ystride)
{
    __shared__ __attribute__((aligned(16))) int8_t LA[4 * 128 * LDS_PITCH];
    __shared__ __attribute__((aligned(16))) int8_t LB[4 * 64 * LDS_PITCH];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wi = wave >> 1, wj = wave & 1;

    // XCD-aware order: consecutive tiles (same prime) land on one XCD
    int64_t per = (total + 7) / 8;
    int64_t v = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
    if (v >= total) return;
    const int64_t tiles = (int64_t)ntm * ntn;
    const int q = (int)(v / tiles);
    const int64_t t = v % tiles;
    const int64_t i0 = (t / ntn) * 128, j0 = (t % ntn) * 64;
    const PrimeConst c = pc[q];

    const int8_t *gA = RA + (int64_t)q * 4 * planeA;
    const int8_t *gB = RB + (int64_t)q * 4 * planeB;

    v16i acc[7][2];
#pragma unroll
    for (int d = 0; d < 7; d++)
#pragma unroll
        for (int rt = 0; rt < 2; rt++) acc[d][rt] = (v16i){0};

    v4i pa[8], pb[4];
#define GLOAD(k0_)                                                                              \
    do {                                                                                        \
        _Pragma("unroll") for (int u = 0; u < 8; u++) {                                         \
            int ch = tid + 256 * u, s = ch >> 9, row = (ch >> 2) & 127, prt = ch & 3;           \
            pa[u] = *(const v4i *)(gA + s * planeA + (i0 + row) * kpad + (k0_) + 16 * prt);     \
        }                                                                                       \
        _Pragma("unroll") for (int u = 0; u < 4; u++) {                                         \
            int ch = tid + 256 * u, s = ch >> 8, row = (ch >> 2) & 63, prt = ch & 3;            \
            pb[u] = *(const v4i *)(gB + s * planeB + (j0 + row) * kpad + (k0_) + 16 * prt);     \
        }                                                                                       \
    } while (0)
    // Fold the diagonal sums into the residue mod p.  A product whose K is
    // longer than one chunk keeps its running residue in its own Y element
    // between chunks (each thread owns its elements), so no register holds
    // it across the MFMA loop.  The last chunk writes y = c (M/p)^-1 mod p
    // as four balanced digits.
#define FOLD(first_, last_)                                                                     \
    do {                                                                                        \
        _Pragma("unroll") for (int rt = 0; rt < 2; rt++)                                        \
            _Pragma("unroll") for (int r = 0; r < 16; r++) {                                    \
                int64_t i = i0 + wi * 64 + rt * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);  \
                int64_t j = j0 + wj * 32 + (lane & 31);                                         \
                if (i < rowsC && j < colsC) {                                                   \
                    uint32_t *yp = Y + (int64_t)q * ystride + i * colsC + j;                    \
                    int64_t x = (first_) ? 0 : (int64_t)*yp;                                    \
                    _Pragma("unroll") for (int d = 0; d < 7; d++) x += (int64_t)acc[d][rt][r] * c.w[d]; \
                    uint32_t res = mod_s64(x, c);                                               \
                    if (last_) {                                                                \
                        uint32_t y = mod_u64((uint64_t)res * c.inv, c.p, c.mu);                 \
                        *yp = slice4(centre(y, c.p));                                           \
                    } else {                                                                    \
                        *yp = res;                                                              \
                    }                                                                           \
                }                                                                               \
            }                                                                                   \
        _Pragma("unroll") for (int d = 0; d < 7; d++)                                           \
            _Pragma("unroll") for (int rt = 0; rt < 2; rt++) acc[d][rt] = (v16i){0};            \
    } while (0)

    GLOAD(0);
    int64_t since = 0;
    bool first = true;
    for (int64_t k0 = 0; k0 < kpad; k0 += 64) {
        __syncthreads();
#pragma unroll
        for (int u = 0; u < 8; u++) {
            int ch = tid + 256 * u, s = ch >> 9, row = (ch >> 2) & 127, prt = ch & 3;
            *(v4i *)(LA + (s * 128 + row) * LDS_PITCH + 16 * prt) = pa[u];
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            int ch = tid + 256 * u, s = ch >> 8, row = (ch >> 2) & 63, prt = ch & 3;
            *(v4i *)(LB + (s * 64 + row) * LDS_PITCH + 16 * prt) = pb[u];
        }
        __syncthreads();
        if (k0 + 64 < kpad) GLOAD(k0 + 64);
#pragma unroll
        for (int kk = 0; kk < 2; kk++) {
            v4i bf[4], af[4][2];
            const int ko = kk * 32 + 16 * (lane >> 5);
#pragma unroll
            for (int s = 0; s < 4; s++) {
                bf[s] = *(const v4i *)(LB + (s * 64 + wj * 32 + (lane & 31)) * LDS_PITCH + ko);
#pragma unroll
                for (int rt = 0; rt < 2; rt++)
                    af[s][rt] = *(const v4i *)(LA + (s * 128 + wi * 64 + rt * 32 + (lane & 31)) * LDS_PITCH + ko);
            }
#pragma unroll
            for (int s = 0; s < 4; s++)
#pragma unroll
                for (int u = 0; u < 4; u++)
#pragma unroll
                    for (int rt = 0; rt < 2; rt++) acc[s + u][rt] = mfma32(af[s][rt], bf[u], acc[s + u][rt]);
        }
        since += 64;
        if (CHUNKED && since == KCHUNK && k0 + 64 < kpad) {
            FOLD(first, false);
            first = false;
            since = 0;
        }
    }
    if (first) FOLD(true, true);
    else FOLD(false, true);
#undef GLOAD
#undef FOLD
}

// ------------------------------------------------------------------------
// 3. reconstruct.  64 entries per block; row blocks of 128 result bytes.
// ------------------------------------------------------------------------
__global__ __launch_bounds__(256) void crt_reconstruct_kernel(
    const PrimeConst *__restrict__ pc, int P, const int8_t *__restrict__ Mc, int Kr, int nrb,
    const uint32_t *__restrict__ M32, const uint32_t *__restrict__ Y, int64_t ystride, int64_t nent,
    int64_t colsC, uint64_t *__restrict__ C, int64_t Crs, int64_t W, int64_t cap, int64_t out_limbs,
    unsigned *__restrict__ err)
{
    __shared__ __attribute__((aligned(16))) int8_t LA[128 * LDS_PITCH];
    __shared__ __attribute__((aligned(16))) int8_t LB[64 * LDS_PITCH];
    __shared__ int32_t LS[64][132];
    __shared__ int64_t s_fix[4][64];
    __shared__ int32_t s_cnt[4][64];
    __shared__ int64_t s_carry[64];
    __shared__ int32_t s_gamma[64];
    __shared__ int64_t s_seg[64][4][2];    // local carry-out, first digit
    __shared__ int32_t s_segf[64][4];      // bit0: digits 1..7 all ones, bit1: all zero
    __shared__ int64_t s_cin[64][4];
    __shared__ int32_t s_top[64];          // highest non-zero limb (two's complement words)
    __shared__ int32_t s_topn[64];         // highest non-zero limb after negation
    __shared__ int32_t s_flags[64];        // bit0: some limb >= cap non-zero, bit1: limbs >= cap not all ones, bit2: some limb < cap non-zero

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t e0 = (int64_t)blockIdx.x * 64;
    const int64_t lim = min(cap, out_limbs);

    // phase 0: beta (rounded fixed-point sum of y_q/p_q) and the count of negative digits
    {
        int el = tid & 63, ph = tid >> 6;
        int64_t fix = 0;
        int32_t cnt = 0;
        if (e0 + el < nent)
            for (int q = ph; q < P; q += 4) {
                int32_t yc = unslice4(Y[(int64_t)q * ystride + e0 + el]);
                uint32_t p = pc[q].p;
                uint32_t y = yc < 0 ? (uint32_t)(yc + (int32_t)p) : (uint32_t)yc;
                fix += (int64_t)(((uint64_t)y * pc[q].fx) >> 31);
                cnt += yc < 0;
            }
        s_fix[ph][el] = fix;
        s_cnt[ph][el] = cnt;
        __syncthreads();
        if (tid < 64) {
            int64_t f = s_fix[0][tid] + s_fix[1][tid] + s_fix[2][tid] + s_fix[3][tid];
            int32_t n = s_cnt[0][tid] + s_cnt[1][tid] + s_cnt[2][tid] + s_cnt[3][tid];
            s_gamma[tid] = (int32_t)((f + (1ll << 31)) >> 32) - n;
            s_carry[tid] = 0;
            s_top[tid] = -1;
            s_topn[tid] = -1;
            s_flags[tid] = 0;
        }
    }

    const int el_c = tid >> 2, seg = tid & 3;
    const int64_t e_c = e0 + el_c;
    uint64_t *cp = nullptr;
    if (e_c < nent) cp = C + ((e_c / colsC) * Crs + (e_c % colsC)) * W;

    for (int rb = 0; rb < nrb; rb++) {
        v16i acc[2] = {(v16i){0}, (v16i){0}};
        for (int k0 = 0; k0 < Kr; k0 += 64) {
            __syncthreads();
            for (int u = 0; u < 2; u++) {
                int ch = tid + 256 * u, row = ch >> 2, prt = ch & 3;
                *(v4i *)(LA + row * LDS_PITCH + 16 * prt) =
                    *(const v4i *)(Mc + (int64_t)(rb * 128 + row) * Kr + k0 + 16 * prt);
            }
            {
                int qq = tid >> 4, ec = tid & 15, q = k0 / 4 + qq;
                uint32_t w4[4] = {0, 0, 0, 0};
                if (q < P) {
                    const uint32_t *src = Y + (int64_t)q * ystride + e0 + 4 * ec;
                    v4i t4 = *(const v4i *)src;
                    w4[0] = t4[0]; w4[1] = t4[1]; w4[2] = t4[2]; w4[3] = t4[3];
                }
                for (int u = 0; u < 4; u++) *(uint32_t *)(LB + (4 * ec + u) * LDS_PITCH + 4 * qq) = w4[u];
            }
            __syncthreads();
#pragma unroll
            for (int kk = 0; kk < 2; kk++) {
                const int ko = kk * 32 + 16 * (lane >> 5);
                v4i af = *(const v4i *)(LA + (wave * 32 + (lane & 31)) * LDS_PITCH + ko);
#pragma unroll
                for (int ct = 0; ct < 2; ct++) {
                    v4i bf = *(const v4i *)(LB + (ct * 32 + (lane & 31)) * LDS_PITCH + ko);
                    acc[ct] = mfma32(af, bf, acc[ct]);
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int ct = 0; ct < 2; ct++)
#pragma unroll
            for (int r = 0; r < 16; r++)
                LS[ct * 32 + (lane & 31)][wave * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3)] = acc[ct][r];
        __syncthreads();

        // carry phase: thread (entry el_c, segment seg) owns words seg*8 .. seg*8+7 of the block
        uint32_t dg[8];
        int64_t carry = 0;
        const int64_t gam = s_gamma[el_c];
#pragma unroll
        for (int vv = 0; vv < 8; vv++) {
            int wb = seg * 8 + vv;
            int64_t x = (int64_t)LS[el_c][4 * wb] + ((int64_t)LS[el_c][4 * wb + 1] << 8) +
                        ((int64_t)LS[el_c][4 * wb + 2] << 16) + ((int64_t)LS[el_c][4 * wb + 3] << 24);
            x -= gam * (int64_t)M32[rb * 32 + wb];
            x += carry;
            dg[vv] = (uint32_t)x;
            carry = x >> 32;
        }
        bool all1 = true, all0 = true;
#pragma unroll
        for (int vv = 1; vv < 8; vv++) { all1 &= dg[vv] == 0xffffffffu; all0 &= dg[vv] == 0; }
        s_seg[el_c][seg][0] = carry;
        s_seg[el_c][seg][1] = dg[0];
        s_segf[el_c][seg] = (all1 ? 1 : 0) | (all0 ? 2 : 0);
        __syncthreads();
        if (seg == 0) {
            int64_t cin = s_carry[el_c];
            for (int g = 0; g < 4; g++) {
                s_cin[el_c][g] = cin;
                int64_t t0 = s_seg[el_c][g][1] + cin;
                int64_t ov = 0;
                if (t0 >= (1ll << 32) && (s_segf[el_c][g] & 1)) ov = 1;
                if (t0 < 0 && (s_segf[el_c][g] & 2)) ov = -1;
                cin = s_seg[el_c][g][0] + ov;
            }
            s_carry[el_c] = cin;
        }
        __syncthreads();
        int64_t cin = s_cin[el_c][seg];
#pragma unroll
        for (int vv = 0; vv < 8; vv++) {
            int64_t x = (int64_t)dg[vv] + cin;
            dg[vv] = (uint32_t)x;
            cin = x >> 32;
        }
        if (cp) {
            int topl = -1, fl = 0;
#pragma unroll
            for (int m = 0; m < 4; m++) {
                int64_t L = (int64_t)rb * 16 + seg * 4 + m;
                uint64_t limb = (uint64_t)dg[2 * m] | ((uint64_t)dg[2 * m + 1] << 32);
                if (L < lim) {
                    cp[L] = limb;
                    if (limb) { topl = (int)L; fl |= 4; }
                } else if (L < out_limbs) {
                    if (limb) fl |= 1;
                    if (limb != ~0ull) fl |= 2;
                }
            }
            if (topl >= 0) atomicMax(&s_top[el_c], topl);
            if (fl) atomicOr(&s_flags[el_c], fl);
        }
    }
    __threadfence();
    __syncthreads();

    // negative results: |C| = ~w + 1 over the stored limbs; the limbs are
    // split in four contiguous ranges and the +1 passes a range iff all of
    // its limbs are zero.  No early exits: every thread reaches each barrier.
    const bool neg = cp && s_carry[el_c] < 0;
    const int64_t per = (lim + 3) / 4, lo = seg * per, hi = min(lim, lo + per);
    bool allz = true;
    if (neg)
        for (int64_t L = lo; L < hi; L++) allz &= cp[L] == 0;
    s_segf[el_c][seg] = allz;
    __syncthreads();
    if (neg) {
        int cin = 1;
        for (int g = 0; g < seg; g++) cin &= s_segf[el_c][g];
        int topl = -1;
        for (int64_t L = lo; L < hi; L++) {
            uint64_t m = ~cp[L] + (uint64_t)cin;
            cin = cin && m == 0;
            cp[L] = m;
            if (m) topl = (int)L;
        }
        if (topl >= 0) atomicMax(&s_topn[el_c], topl);
    }
    __syncthreads();
    if (cp && seg == 0) {
        const int fl = s_flags[el_c];
        bool bad = false;
        if (out_limbs > cap) bad = neg ? ((fl & 2) || !(fl & 4)) : (fl & 1);
        if (bad) atomicOr(err, DEV_ERR_CAPACITY);
        int64_t sz = (neg ? s_topn[el_c] : s_top[el_c]) + 1;
        cp[W - 1] = (uint64_t)(neg ? -sz : sz);
    }
}

}  // namespace mgx

// ========================================================================
// host side: plan and launches
// ========================================================================
using namespace mgx;

struct mgx_crt_plan {
    int64_t K, bitsA, bitsB, bound;   // |C| < 2^bound
    int64_t P, kb, Kr, rows4pad, Lbp, out_limbs;
    std::vector<uint32_t> primes;
    PrimeConst *d_pc = nullptr;
    int8_t *d_cst = nullptr;
    int8_t *d_mc = nullptr;
    uint32_t *d_m32 = nullptr;
};

namespace {

bool is_prime(uint64_t n)
{
    auto mulmod = [](uint64_t a, uint64_t b, uint64_t m) { return (uint64_t)((unsigned __int128)a * b % m); };
    auto powmod = [&](uint64_t a, uint64_t e, uint64_t m) {
        uint64_t r = 1;
        a %= m;
        while (e) { if (e & 1) r = mulmod(r, a, m); a = mulmod(a, a, m); e >>= 1; }
        return r;
    };
    if (n < 2) return false;
    for (uint64_t sp : {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61}) {
        if (n == sp) return true;
        if (n % sp == 0) return false;
    }
    uint64_t d = n - 1;
    int s = 0;
    while (!(d & 1)) { d >>= 1; s++; }
    for (uint64_t a : {2, 7, 61}) {   // exact below 4,759,123,141
        uint64_t x = powmod(a, d, n);
        if (x == 1 || x == n - 1) continue;
        bool comp = true;
        for (int r = 1; r < s; r++) { x = mulmod(x, x, n); if (x == n - 1) { comp = false; break; } }
        if (comp) return false;
    }
    return true;
}

std::vector<uint32_t> &prime_cache()
{
    static std::vector<uint32_t> v;
    return v;
}

// primes below 2^31, descending, until their product exceeds 2^bits
std::vector<uint32_t> primes_for_bits(int64_t bits, hb::Nat &M)
{
    auto &cache = prime_cache();
    std::vector<uint32_t> out;
    M = hb::from_u64(1);
    uint64_t c = cache.empty() ? (1ull << 31) - 1 : cache.back() - 1;
    size_t i = 0;
    for (;;) {
        // M > 2^bits ?
        int64_t bl = hb::bitlen(M);
        if (bl - 1 > bits) break;
        if (bl - 1 == bits) {
            bool pow2 = true;
            for (size_t w = 0; w + 1 < M.size(); w++) pow2 &= M[w] == 0;
            pow2 &= (M.back() & (M.back() - 1)) == 0;
            if (!pow2) break;
        }
        uint32_t p;
        if (i < cache.size()) p = cache[i];
        else {
            while (!is_prime(c)) c--;
            p = (uint32_t)c--;
            cache.push_back(p);
        }
        i++;
        out.push_back(p);
        M = hb::mul_small(M, p);
    }
    return out;
}

uint64_t powmod64(uint64_t a, uint64_t e, uint64_t m)
{
    uint64_t r = 1 % m;
    a %= m;
    while (e) {
        if (e & 1) r = (uint64_t)((unsigned __int128)r * a % m);
        a = (uint64_t)((unsigned __int128)a * a % m);
        e >>= 1;
    }
    return r;
}

int64_t clog2(int64_t k)
{
    int64_t l = 0;
    while ((1ll << l) < k) l++;
    return l;
}

// balanced base-256 digits of a natural number, ndig digits
void balanced_digits(const hb::Nat &v, int64_t ndig, int8_t *out, int64_t stride)
{
    std::vector<uint8_t> bytes(v.size() * 4);
    for (size_t i = 0; i < v.size(); i++)
        for (int b = 0; b < 4; b++) bytes[4 * i + b] = (uint8_t)(v[i] >> (8 * b));
    int carry = 0;
    for (int64_t d = 0; d < ndig; d++) {
        int x = (d < (int64_t)bytes.size() ? bytes[d] : 0) + carry;
        int t = ((x + 128) & 255) - 128;
        out[d * stride] = (int8_t)t;
        carry = (x - t) >> 8;
    }
}

}  // namespace

extern "C" {

int mgx_crt_plan_create(mgx_crt_plan **out, int64_t K, int64_t bitsA, int64_t bitsB)
{
    if (!out || K < 1 || bitsA < 0 || bitsB < 0) return fail(MGX_ERR_ARG, "bad plan arguments");
    MGX_TRY(ensure_init());
    if (bitsA == 0) bitsA = 1;
    if (bitsB == 0) bitsB = 1;
    auto *pl = new mgx_crt_plan();
    pl->K = K;
    pl->bitsA = bitsA;
    pl->bitsB = bitsB;
    pl->bound = bitsA + bitsB + clog2(K);
    hb::Nat M;
    pl->primes = primes_for_bits(pl->bound + 2, M);
    const int64_t P = (int64_t)pl->primes.size();
    pl->P = P;
    pl->kb = round_up(ceil_div(std::max(bitsA, bitsB), 8), 64);
    pl->rows4pad = round_up(4 * P, 64);
    pl->Kr = round_up(4 * P, 64);
    int64_t mbytes = ceil_div(hb::bitlen(M), 8);
    pl->Lbp = round_up(mbytes + 1, 128);
    pl->out_limbs = ceil_div(pl->bound, 64);

    // per-prime constants
    std::vector<PrimeConst> pc(P);
    std::vector<int8_t> cst((size_t)pl->rows4pad * pl->kb, 0);
    std::vector<int8_t> mc((size_t)pl->Lbp * pl->Kr, 0);
    for (int64_t q = 0; q < P; q++) {
        uint32_t p = pl->primes[q];
        PrimeConst &c = pc[q];
        c.p = p;
        hb::Nat Mq = hb::divmod_small(M, p, nullptr);
        uint32_t mqmod = hb::mod_small(Mq, p);
        c.inv = (uint32_t)powmod64(mqmod, p - 2, p);
        c.mu = ~0ull / p;
        c.off = ((1ull << 62) / p + 1) * (uint64_t)p;
        c.fx = (1ull << 63) / p;
        for (int d = 0; d < 7; d++) {
            uint64_t w = powmod64(2, 8 * d, p);
            c.w[d] = w > p / 2 ? (int32_t)w - (int32_t)p : (int32_t)w;
        }
        uint64_t ks = 0, pw = 1;
        for (int64_t b = 0; b < pl->kb; b++) {
            // reduce-GEMM constants: digits of centred 2^(8b) mod p
            int32_t cen = pw > p / 2 ? (int32_t)pw - (int32_t)p : (int32_t)pw;
            int32_t vv = cen;
            for (int s = 0; s < 4; s++) {
                int32_t t = ((vv + 128) & 255) - 128;
                cst[(size_t)(4 * q + s) * pl->kb + b] = (int8_t)t;
                vv = (vv - t) >> 8;
            }
            ks = (ks + pw) % p;
            pw = (pw * 256) % p;
        }
        c.kofs = (uint32_t)(ks * 128 % p);
        // reconstruct constants: digits of (M/p) 256^s
        for (int s = 0; s < 4; s++) {
            hb::Nat v = hb::shl(Mq, 8 * s);
            balanced_digits(v, pl->Lbp, mc.data() + 4 * q + s, pl->Kr);
        }
    }
    std::vector<uint32_t> m32(pl->Lbp / 4, 0);
    for (size_t i = 0; i < M.size() && i < m32.size(); i++) m32[i] = M[i];

    hipError_t e;
    if ((e = hipMalloc(&pl->d_pc, sizeof(PrimeConst) * P)) != hipSuccess ||
        (e = hipMalloc(&pl->d_cst, cst.size())) != hipSuccess || (e = hipMalloc(&pl->d_mc, mc.size())) != hipSuccess ||
        (e = hipMalloc(&pl->d_m32, m32.size() * 4)) != hipSuccess) {
        mgx_crt_plan_destroy(pl);
        return fail(MGX_ERR_OOM, "plan allocation: %s", hipGetErrorString(e));
    }
    MGX_HIP(hipMemcpy(pl->d_pc, pc.data(), sizeof(PrimeConst) * P, hipMemcpyHostToDevice));
    MGX_HIP(hipMemcpy(pl->d_cst, cst.data(), cst.size(), hipMemcpyHostToDevice));
    MGX_HIP(hipMemcpy(pl->d_mc, mc.data(), mc.size(), hipMemcpyHostToDevice));
    MGX_HIP(hipMemcpy(pl->d_m32, m32.data(), m32.size() * 4, hipMemcpyHostToDevice));
    *out = pl;
    return MGX_OK;
}

int mgx_crt_plan_destroy(mgx_crt_plan *pl)
{
    if (!pl) return MGX_OK;
    hipFree(pl->d_pc);
    hipFree(pl->d_cst);
    hipFree(pl->d_mc);
    hipFree(pl->d_m32);
    delete pl;
    return MGX_OK;
}

int64_t mgx_crt_plan_primes(const mgx_crt_plan *pl) { return pl ? pl->P : -1; }

int mgx_crt_plan_prime_list(const mgx_crt_plan *pl, uint32_t *out)
{
    if (!pl || !out) return fail(MGX_ERR_ARG, "null argument");
    memcpy(out, pl->primes.data(), pl->primes.size() * 4);
    return MGX_OK;
}

int mgx_crt_residue_dims(const mgx_crt_plan *pl, int64_t rows, int64_t K, int64_t *rpad, int64_t *kpad)
{
    if (!pl) return fail(MGX_ERR_ARG, "null plan");
    if (rpad) *rpad = round_up(std::max<int64_t>(rows, 1), 128);
    if (kpad) *kpad = round_up(std::max<int64_t>(K, 1), 64);
    return MGX_OK;
}

int mgx_crt_reduce(const mgx_crt_plan *pl, int transpose, int64_t rows, int64_t K, const mgx_limb *X, int64_t Xrs,
                   int64_t cap, int64_t nbytes, int8_t *R, void *stream)
{
    (void)nbytes;
    if (!pl || !X || !R) return fail(MGX_ERR_ARG, "null argument");
    if ((int64_t)8 * cap < pl->kb && false) return fail(MGX_ERR_SHAPE, "cap");
    int64_t rpad, kpad;
    mgx_crt_residue_dims(pl, rows, K, &rpad, &kpad);
    int64_t nent = rpad * kpad, plane = rpad * kpad;
    int nrb = (int)(pl->rows4pad / 64);
    int64_t nblk = ceil_div(nent, 128) * nrb;
    if (nblk > 0x7fffffffll) return fail(MGX_ERR_SHAPE, "matrix too large for one reduce launch");
    hipLaunchKernelGGL(crt_reduce_kernel, dim3((unsigned)nblk), dim3(256), 0, as_stream(stream), pl->d_pc,
                       pl->d_cst, (int)pl->P, nrb, (int)pl->kb, (const uint64_t *)X, Xrs, cap + 1, transpose, rows,
                       K, kpad, nent, R, plane, device_error_word());
    MGX_HIP(hipGetLastError());
    return MGX_OK;
}

int mgx_crt_residue_gemm(const mgx_crt_plan *pl, int64_t K, int64_t M, int64_t N, const int8_t *RA,
                         const int8_t *RB, uint32_t *Y, void *stream)
{
    if (!pl || !RA || !RB || !Y) return fail(MGX_ERR_ARG, "null argument");
    int64_t rpadA, rpadB, kpad;
    mgx_crt_residue_dims(pl, K, M, &rpadA, &kpad);
    mgx_crt_residue_dims(pl, N, M, &rpadB, nullptr);
    int ntm = (int)(rpadA / 128), ntn = (int)(rpadB / 64);
    int64_t total = (int64_t)ntm * ntn * pl->P;
    int64_t grid = round_up(total, 8);
    int64_t ystride = round_up(K * N, 64);
    // products with K beyond one chunk take the variant that folds between chunks
    if (kpad <= KCHUNK)
        hipLaunchKernelGGL(crt_gemm_kernel<false>, dim3((unsigned)grid), dim3(256), 0, as_stream(stream), pl->d_pc,
                           RA, RB, rpadA * kpad, rpadB * kpad, kpad, K, N, ntm, ntn, total, Y, ystride);
    else
        hipLaunchKernelGGL(crt_gemm_kernel<true>, dim3((unsigned)grid), dim3(256), 0, as_stream(stream), pl->d_pc,
                           RA, RB, rpadA * kpad, rpadB * kpad, kpad, K, N, ntm, ntn, total, Y, ystride);
    MGX_HIP(hipGetLastError());
    return MGX_OK;
}

int mgx_crt_reconstruct(const mgx_crt_plan *pl, int64_t K, int64_t N, const uint32_t *Y, mgx_limb *C, int64_t Crs,
                        int64_t cap, void *stream)
{
    if (!pl || !Y || !C) return fail(MGX_ERR_ARG, "null argument");
    int64_t nent = K * N, ystride = round_up(K * N, 64);
    int64_t grid = ceil_div(nent, 64);
    unsigned *err = device_error_word();
    hipLaunchKernelGGL(crt_reconstruct_kernel, dim3((unsigned)grid), dim3(256), 0, as_stream(stream), pl->d_pc,
                       (int)pl->P, pl->d_mc, (int)pl->Kr, (int)(pl->Lbp / 128), pl->d_m32, Y, ystride, nent, N,
                       (uint64_t *)C, Crs, cap + 1, cap, pl->out_limbs, err);
    MGX_HIP(hipGetLastError());
    return MGX_OK;
}

int mgx_crt_scan_bits(int64_t rows, int64_t cols, const mgx_limb *X, int64_t Xrs, int64_t cap, int64_t *bits,
                      void *stream)
{
    MGX_TRY(ensure_init());
    hipStream_t s = as_stream(stream);
    int *d = nullptr;
    MGX_TRY(scratch_alloc((void **)&d, sizeof(int), s));
    MGX_HIP(hipMemsetAsync(d, 0, sizeof(int), s));
    int64_t n = rows * cols;
    int grid = (int)std::min<int64_t>(std::max<int64_t>(ceil_div(n, 256), 1), 4096);
    if (n > 0)
        hipLaunchKernelGGL(scan_bits_kernel, dim3(grid), dim3(256), 0, s, (const uint64_t *)X, Xrs, cap + 1, rows,
                           cols, d);
    int h = 0;
    MGX_HIP(hipMemcpyAsync(&h, d, sizeof(int), hipMemcpyDeviceToHost, s));
    MGX_TRY(scratch_free(d, s));
    MGX_HIP(hipStreamSynchronize(s));
    *bits = h;
    return MGX_OK;
}

int mgx_crt_mfma_matmul(const mgx_crt_plan *plan, int64_t K, int64_t M, int64_t N, mgx_limb *C, int64_t Crs,
                        const mgx_limb *A, int64_t Ars, const mgx_limb *B, int64_t Brs, int64_t cap, void *stream)
{
    MGX_TRY(ensure_init());
    if (K < 0 || M < 0 || N < 0 || cap < 1) return fail(MGX_ERR_SHAPE, "bad shape");
    if (K == 0 || N == 0) return MGX_OK;
    hipStream_t s = as_stream(stream);
    if (M == 0) {   // empty inner dimension: C = 0
        MGX_HIP(hipMemset2DAsync(C, (size_t)(Crs * (cap + 1) * 8), 0, (size_t)(N * (cap + 1) * 8), (size_t)K, s));
        MGX_HIP(hipStreamSynchronize(s));
        return MGX_OK;
    }
    mgx_crt_plan *own = nullptr;
    const mgx_crt_plan *pl = plan;
    if (!pl) {
        int64_t ba, bb;
        MGX_TRY(mgx_crt_scan_bits(K, M, A, Ars, cap, &ba, stream));
        MGX_TRY(mgx_crt_scan_bits(M, N, B, Brs, cap, &bb, stream));
        MGX_TRY(mgx_crt_plan_create(&own, M, ba, bb));
        pl = own;
    } else if (plan->K < M) {
        return fail(MGX_ERR_SHAPE, "plan made for inner dimension %lld < %lld", (long long)plan->K, (long long)M);
    }
    int64_t rpadA, rpadB, kpad;
    mgx_crt_residue_dims(pl, K, M, &rpadA, &kpad);
    mgx_crt_residue_dims(pl, N, M, &rpadB, nullptr);
    int8_t *RA = nullptr, *RB = nullptr;
    uint32_t *Y = nullptr;
    int st = MGX_OK;
    size_t bytesA = (size_t)(4 * pl->P * rpadA * kpad), bytesB = (size_t)(4 * pl->P * rpadB * kpad);
    size_t bytesY = (size_t)pl->P * round_up(K * N, 64) * 4;
    if ((st = scratch_alloc((void **)&RA, bytesA, s)) == MGX_OK && (st = scratch_alloc((void **)&RB, bytesB, s)) == MGX_OK &&
        (st = scratch_alloc((void **)&Y, bytesY, s)) == MGX_OK) {
        if (st == MGX_OK) st = mgx_crt_reduce(pl, 0, K, M, A, Ars, cap, 0, RA, stream);
        if (st == MGX_OK) st = mgx_crt_reduce(pl, 1, N, M, B, Brs, cap, 0, RB, stream);
        if (st == MGX_OK) st = mgx_crt_residue_gemm(pl, K, M, N, RA, RB, Y, stream);
        if (st == MGX_OK) st = mgx_crt_reconstruct(pl, K, N, Y, C, Crs, cap, stream);
    }
    scratch_free(RA, s);
    scratch_free(RB, s);
    scratch_free(Y, s);
    if (own) {
        hipStreamSynchronize(s);
        mgx_crt_plan_destroy(own);
    }
    if (st == MGX_OK) st = check_device_error(s);
    return st;
}

}  // extern "C"
