// crt_kernels.h -- device side of the multi-modular route: per-prime
// constants, modular helpers and the three tile functions (reduce, residue
// GEMM, reconstruct) shared by the one-product kernels and the batched
// persistent kernels.  See crt.hip for the method.
#pragma once
#include "common.h"

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
// 1. reduce.  Output rows (q,s) x entries; tile 64 rows x 128 entries.
// ------------------------------------------------------------------------
__device__ __forceinline__ void reduce_tile(int64_t blk,
    const PrimeConst *__restrict__ pc, const int8_t *__restrict__ cst, int P, int nrb, int kb,
    const uint64_t *__restrict__ X, int64_t Xrs, int64_t W, int transpose, int64_t rows, int64_t K,
    int64_t kpad, int64_t nent, int8_t *__restrict__ R, int64_t plane, unsigned *__restrict__ err)
{
    __shared__ __attribute__((aligned(16))) int8_t LA[64 * LDS_PITCH];
    __shared__ __attribute__((aligned(16))) int8_t LB[128 * LDS_PITCH];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int rb = (int)(blk % nrb);
    const int64_t e0 = (blk / nrb) * 128;

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
__device__ __forceinline__ void gemm_tile(int64_t v,
    const PrimeConst *__restrict__ pc, const int8_t *__restrict__ RA, const int8_t *__restrict__ RB,
    int64_t planeA, int64_t planeB, int64_t kpad, int64_t rowsC, int64_t colsC, int ntm, int ntn,
    int64_t total, uint32_t *__restrict__ Y, int64_t ystride)
{
    __shared__ __attribute__((aligned(16))) int8_t LA[4 * 128 * LDS_PITCH];
    __shared__ __attribute__((aligned(16))) int8_t LB[4 * 64 * LDS_PITCH];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wi = wave >> 1, wj = wave & 1;

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
__device__ __forceinline__ void reconstruct_tile(int64_t blk,
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
    const int64_t e0 = blk * 64;
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
