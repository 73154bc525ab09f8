// crt_kernels.h -- device side of the multi-modular route: per-prime
// constants, modular helpers and the tile functions (reduce, residue GEMM,
// reconstruct) shared by the one-product kernels and the batched persistent
// kernels.  See crt.hip for the method.
//
// A plan slices every residue into S balanced int8 digits: S = 4 for primes
// below 2^31, S = 2 for primes below 65280 (|centred residue| <= 32639, the
// range of two balanced digits).  Rows of the reduce output and columns of
// the reconstruction constants are indexed S*q + s; the residue GEMM's
// output Y is packed as 32-bit words of four digits, so word plane w holds
// prime w (S = 4) or primes 2w and 2w+1 in its low and high halves (S = 2).
#pragma once
#include "common.h"
#include <cstring>
#include <vector>

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
    uint32_t off32;  // a multiple of p in (2^31, 2^31 + p] (p < 2^16)
    int32_t f45;     // floor(2^45 / p) (p < 2^16)
    uint32_t kneg;   // -127 * sum_{b<kb} 2^(8b) mod p
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
// u mod p for 2^14 <= p < 2^16 and any u < 2^32, given negp = -p and mu32 =
// floor(2^32 / p): the quotient q = floor(u mu32 / 2^32) is short of
// floor(u / p) by at most 1 (u mu32 / 2^32 > u / p - u / 2^32 > u / p - 1),
// so u - q p < 2p; q < 2^18 and p < 2^16 take the signed 24-bit
// multiply-add (only the low 32 bits of the result matter)
__device__ __forceinline__ uint32_t bred32_lazy(uint32_t u, int32_t negp, uint32_t mu32)
{
    return u + (uint32_t)__mul24((int32_t)__umulhi(u, mu32), negp);
}
__device__ __forceinline__ uint32_t bred32(uint32_t u, int32_t negp, uint32_t mu32)
{
    const uint32_t r = bred32_lazy(u, negp, mu32);
    return min(r, r + (uint32_t)negp);   // r - p wraps above r when r < p
}
// a mod p for signed |a| < 2^31 - 2^17: off32 is a multiple of p in (2^31, 2^31 + p]
__device__ __forceinline__ uint32_t smod32(int32_t a, int32_t negp, uint32_t mu32, uint32_t off32)
{
    return bred32((uint32_t)a + off32, negp, mu32);
}
// the same without the correction: a value congruent to a mod p in [0, 2p)
__device__ __forceinline__ uint32_t smod32_lazy(int32_t a, int32_t negp, uint32_t mu32, uint32_t off32)
{
    return bred32_lazy((uint32_t)a + off32, negp, mu32);
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
// two balanced digits of -32896 <= r <= 32639 in the low 16 bits: with
// u = r + 128, the low digit's byte is (u & 255) ^ 128 and the high digit is
// u >> 8 (arithmetic), i.e. bits 8..15 of u
__device__ __forceinline__ uint32_t slice2(int32_t r)
{
    return ((uint32_t)(r + 128) ^ 0x80u) & 0xffffu;
}
__device__ __forceinline__ int32_t unslice2(uint32_t h)
{
    return (int32_t)(int8_t)(h & 255) + ((int32_t)(int8_t)((h >> 8) & 255) << 8);
}
template <int S>
__device__ __forceinline__ uint32_t slice_s(int32_t r) { return S == 4 ? slice4(r) : slice2(r); }

__device__ __forceinline__ v16i mfma32(v4i a, v4i b, v16i c)
{
    return __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c, 0, 0, 0);
}
// 16x16x64: lane l gives row (col) l%16, K bytes 16(l/16)..+15; output
// register r of lane l is row 4(l/16)+r, column l%16
__device__ __forceinline__ v4i mfma16(v4i a, v4i b, v4i c)
{
    return __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, c, 0, 0, 0);
}

constexpr int LDS_PITCH = 80;   // 64 data bytes + 16 padding per row

// True when the entry's magnitude (limbs 0 .. a-1) has a bit at or above
// position xbits, i.e. |x| >= 2^xbits: the plan's modulus covers only
// products of entries below its per-operand bounds.  Zero top limbs are
// allowed; the loop usually inspects one limb (the top one).
__device__ __forceinline__ bool over_bound(const uint64_t *__restrict__ e, int64_t a, int xbits)
{
    for (int64_t li = xbits >> 6; li < a; li++)
        if (e[li] & (li == (xbits >> 6) ? ~0ull << (xbits & 63) : ~0ull)) return true;
    return false;
}


// Byte offset of residue (row r, column k) in an S = 2 residue plane: blocks
// of 16 rows x 64 columns (1 KB, row-major inside), the blocks of a 16-row band
// in column order, so that the residue GEMM's 16-row x 64-byte loads are
// whole contiguous kilobytes.  S = 4 planes are plain row-major.
__host__ __device__ __forceinline__ int64_t res_off(int S, int64_t r, int64_t k, int64_t kpad)
{
    return S == 2 ? (r >> 4) * (kpad << 4) + ((k >> 6) << 10) + ((r & 15) << 6) + (k & 63) : r * kpad + k;
}

// word of entry el in an LO4 row: the 16-byte group of entries el/4 moves to
// group (el/4) ^ 1 in the upper half of the row, so that the transpose's
// 16-byte reads of groups 2c and 2c + 1 (c = 0..15) are free of bank conflicts
__device__ __forceinline__ int lo4_swz(int el) { return el ^ (((el >> 6) & 1) << 2); }
// The residues of a 64-row x 128-entry tile (row block rb, entries e0 ..
// e0 + 127, all inside the plane) from LDS to their planes: LO4 holds, for
// row group qg (rows 4 qg .. 4 qg + 3) and entry el, the four rows' bytes as
// one dword at LO4[qg * 128 + lo4_swz(el)].  Thread (qg, ch) takes entries
// 8 ch .. 8 ch + 7; a 4 x 8 byte transpose gives each row an 8-byte segment,
// and the 16 threads of one qg write 128 consecutive bytes per row.
template <int S>
__device__ __forceinline__ void store_res_tile(const uint32_t *LO4, int tid, int64_t e0, int64_t kpad, int rb, int P,
                                               int8_t *__restrict__ R, int64_t plane)
{
    const int qg = tid >> 4, ch = tid & 15;
    const int64_t eo = res_off(S, (e0 + 8 * ch) / kpad, (e0 + 8 * ch) % kpad, kpad);   // 8 columns, one row
    const uint4 a = *(const uint4 *)(LO4 + qg * 128 + lo4_swz(8 * ch));
    const uint4 b = *(const uint4 *)(LO4 + qg * 128 + lo4_swz(8 * ch + 4));
    const uint32_t d[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int64_t grow = (int64_t)rb * 64 + 4 * qg + k;   // = S*q + s
        uint32_t lo = 0, hi = 0;
#pragma unroll
        for (int x = 0; x < 4; x++) {
            lo |= ((d[x] >> (8 * k)) & 255u) << (8 * x);
            hi |= ((d[x + 4] >> (8 * k)) & 255u) << (8 * x);
        }
        if (grow < (int64_t)S * P) *(uint2 *)(R + grow * plane + eo) = make_uint2(lo, hi);
    }
}

// ------------------------------------------------------------------------
// 1. reduce.  Output rows S*q+s x entries; tile 64 rows x 128 entries.
// ------------------------------------------------------------------------
template <int S>
__device__ __forceinline__ void reduce_tile(int64_t blk,
    const PrimeConst *__restrict__ pc, const int8_t *__restrict__ cst, int P, int nrb, int kb, int xbits,
    const uint64_t *__restrict__ X, int64_t Xrs, int64_t W, int transpose, int64_t rows, int64_t K,
    int64_t kpad, int64_t nent, int8_t *__restrict__ R, int64_t plane, unsigned *__restrict__ err)
{
    __shared__ __attribute__((aligned(16))) int8_t LA[64 * LDS_PITCH];
    __shared__ __attribute__((aligned(16))) int8_t LB[128 * LDS_PITCH];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int rb = (int)(blk % nrb);
    const int64_t e0 = (blk / nrb) * 128;

    // the two entries this thread stages: limb base (padding entries read
    // entry 0 and mask everything) and limb count
    int64_t base[2];
    int nl[2];
    const int lim = (int)min<int64_t>(kb / 8, W - 1);   // limb loads clamped below both bounds
    for (int h = 0; h < 2; h++) {
        const int64_t e = e0 + (tid >> 2) + 64 * h, r = e / kpad, k = e % kpad;
        const bool real = e < nent && r < rows && k < K;
        base[h] = real ? (transpose ? (k * Xrs + r) * W : (r * Xrs + k) * W) : 0;
        const int64_t sz = (int64_t)X[base[h] + W - 1];
        const int64_t a = real ? min(sz < 0 ? -sz : sz, W - 1) : 0;
        nl[h] = (int)min<int64_t>(a, lim);
        if (rb == 0 && over_bound(X + base[h], a, xbits)) atomicOr(err, DEV_ERR_INPUT);
    }
    const int part = tid & 3;
    v16i acc[2];
    for (int i = 0; i < 2; i++) acc[i] = (v16i){0};

    // the next step's constants and limbs are loaded (unconditionally, limb
    // index clamped) before this step's MFMAs, and masked at |size| and
    // biased to byte - 128 when they are stored to LDS
    v4i an;
    uint64_t bn[2][2];
#define RLOAD(k0_)                                                                                    \
    {                                                                                                 \
        an = *(const v4i *)(cst + (int64_t)(rb * 64 + (tid >> 2)) * kb + (k0_) + 16 * part);        \
        _Pragma("unroll") for (int h = 0; h < 2; h++)                                                 \
        _Pragma("unroll") for (int u = 0; u < 2; u++)                                                 \
            bn[h][u] = X[base[h] + min(((k0_) + 16 * part) / 8 + u, lim - 1)];                       \
    }
    RLOAD(0);
    for (int k0 = 0; k0 < kb; k0 += 64) {
        __syncthreads();
        *(v4i *)(LA + (tid >> 2) * LDS_PITCH + 16 * part) = an;
#pragma unroll
        for (int h = 0; h < 2; h++) {
            uint64_t *d = (uint64_t *)(LB + ((tid >> 2) + 64 * h) * LDS_PITCH + 16 * part);
#pragma unroll
            for (int u = 0; u < 2; u++) {
                const int li = (k0 + 16 * part) / 8 + u;
                d[u] = (li < nl[h] ? bn[h][u] : 0) ^ 0x8080808080808080ull;
            }
        }
        __syncthreads();
        if (k0 + 64 < kb) RLOAD(k0 + 64);
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
#undef RLOAD

    // epilogue: lane owns entry el = wave*32 + lane%32; register 4g+u of row
    // tile rt is row rt*32 + 8g + 4(lane/32) + u = S*q + s of this row block.
    // The four rows' digits of each (rt, g) form one dword in LDS (in LB,
    // free once every wave is past its last MFMA), stored as row segments.
    // The plane holds whole tiles (nent is a multiple of 128), so every entry
    // of the tile exists; padding entries have no limbs and are stored as 0.
    const int el = wave * 32 + (lane & 31);
    const int64_t e = e0 + el, r = e / kpad, k = e % kpad;
    const bool real = e < nent && r < rows && k < K;
    bool neg = false;
    if (real) {
        const uint64_t *p = transpose ? X + (k * Xrs + r) * W : X + (r * Xrs + k) * W;
        neg = (int64_t)p[W - 1] < 0;
    }
    uint32_t *const LO4 = (uint32_t *)LB;
    __syncthreads();
#pragma unroll
    for (int rt = 0; rt < 2; rt++)
#pragma unroll
        for (int g = 0; g < 4; g++) {
            uint32_t dw = 0;
#pragma unroll
            for (int w = 0; w < 4 / S; w++) {
                const int q = rb * (64 / S) + (rt * 32 + 8 * g + 4 * (lane >> 5)) / S + w;
                uint32_t word = 0;
                if (real && q < P && S == 2) {
                    // 32-bit steps: |acc| <= kb 2^14 < 2^28 (S = 2 plans bound
                    // products below 2^17 bits: kb < 2^14), so acc_1 is reduced
                    // first (below 2p) and acc_0 + 256 r1 + kofs + off32 + h lies in
                    // [0, 2^32); with h = floor(p/2), the residue minus h is the
                    // centred residue
                    const PrimeConst &c = pc[q];
                    const uint32_t mu32 = (uint32_t)(c.mu >> 32), hp = c.p >> 1;
                    const int32_t negp = -(int32_t)c.p;
                    const uint32_t r1 = smod32_lazy(acc[rt][4 * g + 2 * w + 1], negp, mu32, c.off32);
                    const uint32_t u = (uint32_t)(acc[rt][4 * g + 2 * w] + 256 * (int32_t)r1) + c.kofs + c.off32 + hp;
                    const int32_t xr = (int32_t)bred32(u, negp, mu32) - (int32_t)hp;
                    word = slice2(neg ? -xr : xr);
                } else if (real && q < P) {
                    const PrimeConst c = pc[q];
                    int64_t v = c.kofs;
#pragma unroll
                    for (int s = 0; s < S; s++) v += (int64_t)acc[rt][4 * g + S * w + s] << (8 * s);
                    uint32_t res = mod_s64(v, c);
                    if (neg && res) res = c.p - res;
                    word = slice_s<S>(centre(res, c.p));
                }
                dw |= word << (16 * w);
            }
            LO4[(rt * 8 + 2 * g + (lane >> 5)) * 128 + lo4_swz(el)] = dw;
        }
    __syncthreads();
    store_res_tile<S>(LO4, tid, e0, kpad, rb, P, R, plane);
}


// Entry-resident reduce for entries of at most 256 bytes: every lane keeps
// half of one entry's bytes in registers for the whole sweep (they are the
// MFMA B operand of every prime block), and the workgroup sweeps every prime
// block over its 128 entries; the residues of a prime block leave through
// LDS as 8-byte row segments.  KS: 64-byte K steps (kb <= 64 KS; the
// constant columns kb .. 64 KS are zero).
// Limbs 4 k32 + 2 hh and 4 k32 + 2 hh + 1 (bytes 32 k32 + 16 hh .. + 15) of
// the entry at X + base, k32 < K32, for the entry-resident MFMA operands.
// When every such limb lies below the size word, two limbs per 16-byte load
// (entries are 8-byte aligned; limbs past the plan's bytes are masked by the
// caller): half the load instructions of the limb-wise path, whose 64 lanes
// each touch a cache line of their own.  Else limb-wise, indices clamped at
// lim - 1 (lim <= W - 1: never the size word's neighbour past the entry).
template <int K32>
__device__ __forceinline__ void entry_limbs(const uint64_t *__restrict__ X, int64_t base, int64_t W, int lim, int hh,
                                            uint64_t (&v)[K32][2])
{
    if (4 * K32 <= W - 1) {
        typedef uint32_t v4u8 __attribute__((ext_vector_type(4), aligned(8)));
#pragma unroll
        for (int k32 = 0; k32 < K32; k32++) {
            const v4u8 t = *(const v4u8 *)(X + base + 4 * k32 + 2 * hh);
            v[k32][0] = t.x | ((uint64_t)t.y << 32);
            v[k32][1] = t.z | ((uint64_t)t.w << 32);
        }
    } else {
#pragma unroll
        for (int k32 = 0; k32 < K32; k32++)
#pragma unroll
            for (int u = 0; u < 2; u++) v[k32][u] = X[base + min(4 * k32 + 2 * hh + u, lim - 1)];
    }
}

constexpr int RES_KB = 512;   // widest entry (bytes) of the entry-resident tile
constexpr int RES_KG = 2;   // K steps whose A fragments are read ahead of their MFMAs

template <int S, int KS>
__device__ __forceinline__ void reduce_tile_res(int64_t blk,
    const PrimeConst *__restrict__ pc, const int8_t *__restrict__ cst, int P, int nrb, int kb, int xbits,
    const uint64_t *__restrict__ X, int64_t Xrs, int64_t W, int transpose, int64_t rows, int64_t K,
    int64_t kpad, int8_t *__restrict__ R, int64_t plane, unsigned *__restrict__ err)
{
    static_assert(KS >= 1 && 64 * KS <= RES_KB, "K steps");
    constexpr int PITCH = 64 * KS + 16;   // LDS row of the constants: 64 KS bytes + 16 padding
    constexpr int K32 = 2 * KS;   // 32-byte MFMA steps
    __shared__ __attribute__((aligned(16))) int8_t LC[64 * PITCH];
    __shared__ __attribute__((aligned(16))) int8_t LO[64 * 128];
    // S = 2, per sign of the entry and prime of the row block: -p, floor(2^32/p),
    // the offset K and 128 - floor(p/2) (see the epilogue); the negative
    // entries' table starts at row 33, so that the two tables' rows of one
    // prime sit in different banks
    __shared__ uint4 s_pc[65];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t e0 = blk * 128;
    const int cpe = kb / 16;   // 16-byte chunks per entry
    const int el = wave * 32 + (lane & 31), hh = lane >> 5;

    // This lane's entry, bytes 32 k32 + 16 hh .. + 15 for every k32 < kb/32
    // (limbs 4 k32 + 2 hh, + 1): loaded with no condition, all in flight at
    // once (padding reads entry 0, limb indices clamped below the cap), then
    // masked at |size| and biased to byte - 128, or (S = 2, negative entry) to
    // 127 - byte, which makes the products sum to -|x| plus a constant.
    bool neg;
    v4i bf[K32];
    {
        const int64_t e = e0 + el, r = e / kpad, k = e % kpad;
        const bool real = r < rows && k < K;
        const int64_t base = real ? (transpose ? (k * Xrs + r) * W : (r * Xrs + k) * W) : 0;
        const int lim = (int)min<int64_t>(kb / 8, W - 1);
        const int64_t sz = (int64_t)X[base + W - 1];
        uint64_t v[K32][2];
        entry_limbs<K32>(X, base, W, lim, hh, v);
        const int64_t a = real ? min(sz < 0 ? -sz : sz, W - 1) : 0;
        const int nl = (int)min<int64_t>(a, lim);
        neg = real && sz < 0;
        const uint64_t flip = (S == 2 && neg) ? 0x7f7f7f7f7f7f7f7full : 0x8080808080808080ull;
#pragma unroll
        for (int k32 = 0; k32 < K32; k32++) {
            const uint64_t d0 = (4 * k32 + 2 * hh < nl ? v[k32][0] : 0) ^ flip;
            const uint64_t d1 = (4 * k32 + 2 * hh + 1 < nl ? v[k32][1] : 0) ^ flip;
            bf[k32] = (v4i){(int)(uint32_t)d0, (int)(uint32_t)(d0 >> 32), (int)(uint32_t)d1, (int)(uint32_t)(d1 >> 32)};
        }
        if (hh == 0 && over_bound(X + base, a, xbits)) atomicOr(err, DEV_ERR_INPUT);
    }
    // the constants of row block rb: 64 rows x kb bytes, at most KS 16-byte
    // chunks per thread, one block ahead in registers
    v4i cpre[KS];
    int cst_off[KS], lc_off[KS];
#pragma unroll
    for (int u = 0; u < KS; u++) {
        const int c = tid + 256 * u, row = c / cpe, part = c - row * cpe;
        cst_off[u] = c < 64 * cpe ? row * kb + 16 * part : -1;
        lc_off[u] = row * PITCH + 16 * part;
    }
    // columns kb .. 64 KS: zero once (the row blocks' copies never write them)
    for (int c = tid; c < 64 * (4 * KS - cpe); c += 256) {
        const int row = c / (4 * KS - cpe), part = cpe + c % (4 * KS - cpe);
        *(v4i *)(LC + row * PITCH + 16 * part) = (v4i){0, 0, 0, 0};
    }
#define CLOAD(rb_)                                                                                    \
    _Pragma("unroll") for (int u = 0; u < KS; u++)                                                     \
        if (cst_off[u] >= 0) cpre[u] = *(const v4i *)(cst + (int64_t)(rb_) * 64 * kb + cst_off[u]);
    // S = 2: the row block's 32 primes' epilogue constants (thread 32 sg + j:
    // sign sg, prime j), one block ahead in registers
    uint4 pcv = make_uint4(0, 0, 0, 0);
#define PCLOAD(rb_)                                                                                   \
    if (S == 2 && tid < 64) {                                                                         \
        const PrimeConst &c_ = pc[min((rb_) * 32 + (tid & 31), P - 1)];                               \
        const uint32_t h_ = c_.p >> 1;                                                                \
        pcv = make_uint4((uint32_t)-(int32_t)c_.p, (uint32_t)(c_.mu >> 32),                          \
                         (tid >> 5 ? c_.kneg : c_.kofs) + c_.off32 + h_, 128u - h_);                  \
    }
    CLOAD(0);
    PCLOAD(0);
    for (int rb = 0; rb < nrb; rb++) {
        __syncthreads();   // LC, LO and s_pc free again
#pragma unroll
        for (int u = 0; u < KS; u++)
            if (cst_off[u] >= 0) *(v4i *)(LC + lc_off[u]) = cpre[u];
        if (S == 2 && tid < 64) s_pc[(tid >> 5) * 33 + (tid & 31)] = pcv;
        __syncthreads();
        if (rb + 1 < nrb) {
            CLOAD(rb + 1);
            PCLOAD(rb + 1);
        }
        // A fragments in groups of RES_KG K steps, all read before the group's
        // MFMAs, so that the MFMAs do not each wait on an LDS read
        v16i acc[2];
#pragma unroll
        for (int k0 = 0; k0 < K32; k0 += RES_KG) {
            v4i af[RES_KG][2];
#pragma unroll
            for (int k32 = k0; k32 < k0 + RES_KG && k32 < K32; k32++)
#pragma unroll
                for (int rt = 0; rt < 2; rt++)
                    af[k32 - k0][rt] = *(const v4i *)(LC + (rt * 32 + (lane & 31)) * PITCH + k32 * 32 + 16 * hh);
#pragma unroll
            for (int k32 = k0; k32 < k0 + RES_KG && k32 < K32; k32++)
#pragma unroll
                for (int rt = 0; rt < 2; rt++)
                    acc[rt] = mfma32(af[k32 - k0][rt], bf[k32], k32 == 0 ? (v16i){0} : acc[rt]);
        }
        // the residue digits of rows 4 qg .. 4 qg + 3 of this row block
        // (qg = 8 rt + 2 g + lane / 32; S = 2: primes 2 qg, 2 qg + 1 with two
        // digits each) for entry el form one dword of LO4[qg][el]
        uint32_t *const LO4 = (uint32_t *)LO;
#pragma unroll
        for (int rt = 0; rt < 2; rt++)
#pragma unroll
            for (int g = 0; g < 4; g++) {
                uint32_t dw = 0;
                if (S == 2) {
                    // |acc| <= kb 2^14 and v = acc_0 + 256 acc_1 is congruent to
                    // x - 128 sum 2^(8b) (to -x + 127 sum 2^(8b) for a negative
                    // entry).  For kb <= 256, |v| < 2^31 - 2^22; wider entries first
                    // reduce acc_1 below 2p (plus p floor(floor(2^32/p)/2), a
                    // multiple of p in (2^31 - 2p, 2^31]).  u = v + K with K =
                    // (128 resp. -127) sum 2^(8b) mod p + off32 + h then lies in
                    // [0, 2^32) and is congruent to x + h; with h = floor(p/2),
                    // (u mod p) - h is the centred residue and (u mod p) + 128 - h
                    // its two digits' bytes plus 128.  Padding entries have no limbs
                    // and come out 0.
                    const uint4 *tab = s_pc + (neg ? 33 : 0) + 16 * rt + 4 * g + 2 * hh;
                    uint32_t xw[2];
#pragma unroll
                    for (int w = 0; w < 2; w++) {
                        const uint4 k4 = tab[w];   // -p, mu32, K, 128 - h
                        int32_t hi = acc[rt][4 * g + 2 * w + 1];
                        if (KS > 4)
                            hi = (int32_t)bred32_lazy((uint32_t)hi + __umul24(k4.y >> 1, (uint32_t)-(int32_t)k4.x),
                                                      (int32_t)k4.x, k4.y);
                        const uint32_t u = (uint32_t)(acc[rt][4 * g + 2 * w] + hi * 256) + k4.z;
                        xw[w] = bred32(u, (int32_t)k4.x, k4.y) + k4.w;
                    }
                    dw = ((xw[0] & 0xffffu) | (xw[1] << 16)) ^ 0x00800080u;
                } else {
#pragma unroll
                    for (int w = 0; w < 4 / S; w++) {
                        const int ql = (rt * 32 + 8 * g + 4 * (lane >> 5)) / S + w, q = rb * (64 / S) + ql;
                        // (rows past the last prime take its constants and are not stored;
                        // padding entries have no limbs and come out 0)
                        const PrimeConst c = pc[min(q, P - 1)];
                        int64_t v = c.kofs;
#pragma unroll
                        for (int s = 0; s < S; s++) v += (int64_t)acc[rt][4 * g + S * w + s] << (8 * s);
                        uint32_t res = mod_s64(v, c);
                        if (neg && res) res = c.p - res;
                        dw |= slice_s<S>(centre(res, c.p)) << (16 * w);
                    }
                }
                LO4[(rt * 8 + 2 * g + (lane >> 5)) * 128 + lo4_swz(el)] = dw;
            }
        __syncthreads();
        store_res_tile<S>(LO4, tid, e0, kpad, rb, P, R, plane);
    }
#undef CLOAD
#undef PCLOAD
}

// ------------------------------------------------------------------------
// 1b. constant-resident reduce (S = 2, entries of at most 256 bytes).  One
//     workgroup of 8 waves per CU keeps the whole constant table (2P rows x
//     kb bytes) and the per-prime epilogue constants in LDS, loaded once.
//     The workgroup takes a 4-row x 64-column patch of a residue plane (256
//     entries) at a time; wave w holds 32 of them (patch row w/2, columns
//     32 (w%2) .. + 31), the lane's half of an entry's bytes in registers as
//     the MFMA A operand (rows = entries).  The constants are the B operand:
//     row block rb of the table holds digit 0 of primes 32 rb .. 32 rb + 31,
//     then digit 1 of the same primes, so that a lane's output column is one
//     prime (both digits in the same register of the two 32-column tiles)
//     and its epilogue constants are read once per row block.  The residues
//     of a row block go through a plane-major LDS tile (64 planes x 256
//     entries) and leave as 256 contiguous bytes per plane (a 4 x 64 patch
//     of the 16 x 64 blocked plane), two barriers per row block.
//     Table layout (rows of 64 KS bytes, zero past kb): row block rb < nrb - 1
//     at rows 64 rb + 32 d + q % 32; the
//     last block's nlast primes at 64 (nrb - 1) + nlast d + q % 32 (2P rows in
//     all); 16-byte chunk c of row r at slot c ^ cres_swz(r), so that the rows
//     one fragment read touches per LDS cycle hit distinct banks for kb of 64,
//     128 or 256 bytes.
// ------------------------------------------------------------------------
// one operand of a reduce launch that takes A and B together
struct ResSide {
    const uint64_t *X;
    int64_t Xrs, rows, plane;
    int transpose;
    int bits;   // |entries| < 2^bits (the plan's bound for this operand)
    int8_t *R;
};

constexpr int CRES_THREADS = 512;              // eight waves, two per SIMD
constexpr int CRES_LO_PITCH = 260;             // bytes per plane of the output tile (bank shift 1)
constexpr int CRES_MAX_LDS = 160 * 1024;
// LDS row of the table: kb rounded up to 64, 128 or 256 bytes (the columns
// past kb are zero: the entries' bytes there enter as -128)
inline int64_t cres_row_bytes(int64_t kb) { return kb <= 64 ? 64 : kb <= 128 ? 128 : 256; }
inline int64_t cres_lds_bytes(int64_t S, int64_t P, int64_t kb)
{
    return S * P * cres_row_bytes(kb) + round_up(P, 32) * 16 + 64 * CRES_LO_PITCH;
}
__device__ __forceinline__ int cres_swz(int r, int C) { return ((r * C) >> 4) & (C - 1); }
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

template <int KS>
__device__ __forceinline__ void reduce_cres(const PrimeConst *__restrict__ pc, const int8_t *__restrict__ cst, int P,
    int kb, int64_t W, int64_t K, int64_t kpad, const ResSide &s0, const ResSide &s1, int64_t nu0, int64_t nu,
    unsigned *__restrict__ err)
{
    static_assert(KS >= 1 && KS <= 4, "entries of at most 256 bytes");
    constexpr int K32 = 2 * KS, C = 4 * KS;   // 32-byte MFMA steps; 16-byte chunks per row
    extern __shared__ __attribute__((aligned(16))) int8_t LDYN[];
    const int SP = 2 * P, nrb = (P + 31) / 32, lastq = 32 * (nrb - 1), lastrow = 2 * lastq, nlast = P - lastq;
    int8_t *const LC = LDYN;
    uint4 *const s_pc = (uint4 *)(LDYN + (int64_t)SP * 64 * KS);
    uint8_t *const LO = (uint8_t *)(s_pc + 32 * nrb);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int c32 = lane & 31, hh = lane >> 5;
    {
        const int nch = SP * C;
        for (int c0 = 0; c0 < nch; c0 += 8 * CRES_THREADS) {
            v4i t[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int c = min(c0 + tid + CRES_THREADS * u, nch - 1), lr = c / C, ch = c - lr * C;
                int q, d;
                if (lr < lastrow) { q = (lr >> 6) * 32 + (lr & 31); d = (lr >> 5) & 1; }
                else { const int t_ = lr - lastrow; d = t_ / nlast; q = lastq + t_ - d * nlast; }
                t[u] = 16 * ch < kb ? *(const v4i *)(cst + (int64_t)(2 * q + d) * kb + 16 * ch) : (v4i){0, 0, 0, 0};
            }
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int c = c0 + tid + CRES_THREADS * u, lr = c / C, ch = c - lr * C;
                if (c < nch) *(v4i *)(LC + lr * 64 * KS + 16 * (ch ^ cres_swz(lr, C))) = t[u];
            }
        }
        // per prime: -p, floor(2^32/p), and the offsets K + off32 + h of a
        // positive and of a negative entry (h = floor(p/2)); rows past the last
        // prime repeat it
        for (int i = tid; i < 32 * nrb; i += CRES_THREADS) {
            const PrimeConst &c_ = pc[min(i, P - 1)];
            const uint32_t h_ = c_.p >> 1;
            s_pc[i] = make_uint4((uint32_t)-(int32_t)c_.p, (uint32_t)(c_.mu >> 32), c_.kofs + c_.off32 + h_,
                                 c_.kneg + c_.off32 + h_);
        }
        __syncthreads();
    }
    // unit uu: a 4-row x 64-column patch of side 0 (uu < nu0) or side 1;
    // this lane's entry: limb base (padding reads entry 0), size word, limbs
    uint64_t raw[K32][2];
    int64_t rsz = 0;
    bool rreal = false;
    auto load = [&](int64_t uu) {
        const bool s = uu >= nu0;
        const int64_t ul = uu - (s ? nu0 : 0), cbs = kpad >> 6;
        const int64_t r = 4 * (ul / cbs) + (wave >> 1), k = 64 * (ul % cbs) + 32 * (wave & 1) + c32;
        const int64_t Xrs = s ? s1.Xrs : s0.Xrs;
        rreal = r < (s ? s1.rows : s0.rows) && k < K;
        const int64_t base = rreal ? ((s ? s1.transpose : s0.transpose) ? (k * Xrs + r) * W : (r * Xrs + k) * W) : 0;
        const uint64_t *X = s ? s1.X : s0.X;
        const int lim = (int)min<int64_t>(kb / 8, W - 1);
        rsz = (int64_t)X[base + W - 1];
        entry_limbs<K32>(X, base, W, lim, hh, raw);
        if (hh == 0 && rreal && over_bound(X + base, min(rsz < 0 ? -rsz : rsz, W - 1), s ? s1.bits : s0.bits))
            atomicOr(err, DEV_ERR_INPUT);
    };
    // bytes - 128, or 127 - byte for a negative entry (the products then sum
    // to -|x| plus a constant), masked at |size|
    auto operand = [&](v4i (&a)[K32], bool &ng) {
        ng = rreal && rsz < 0;
        const int nl = rreal ? (int)min<int64_t>(min(rsz < 0 ? -rsz : rsz, W - 1), (int64_t)(kb / 8)) : 0;
        const uint64_t flip = ng ? 0x7f7f7f7f7f7f7f7full : 0x8080808080808080ull;
#pragma unroll
        for (int k32 = 0; k32 < K32; k32++) {
            const uint64_t d0 = (4 * k32 + 2 * hh < nl ? raw[k32][0] : 0) ^ flip;
            const uint64_t d1 = (4 * k32 + 2 * hh + 1 < nl ? raw[k32][1] : 0) ^ flip;
            a[k32] = (v4i){(int)(uint32_t)d0, (int)(uint32_t)(d0 >> 32), (int)(uint32_t)d1, (int)(uint32_t)(d1 >> 32)};
        }
    };
    int64_t uu = blockIdx.x;
    v4i af[K32];
    bool neg = false;
    if (uu < nu) {
        load(uu);
        operand(af, neg);
    }
    // the next unit's limbs are loaded before a unit's row blocks and turned
    // into its operand after the MFMAs of row block cb, before that block's
    // stores: vmcnt counts stores and loads alike and in order, so waiting
    // for limbs loaded behind a row block's stores would wait for the stores
    const int cb = min(1, nrb - 1);
    for (; uu < nu; uu += gridDim.x) {
        // output register 4g + i of this lane belongs to entry 8g + 4hh + i of
        // the wave; bit 4g + i of m16: that entry is negative
        uint32_t m16;
        {
            const uint32_t nm = (uint32_t)__ballot(neg) >> (4 * hh);   // bit e: entry e (lanes e and e + 32 agree)
            m16 = (nm & 0xfu) | ((nm >> 4) & 0xf0u) | ((nm >> 8) & 0xf00u) | ((nm >> 12) & 0xf000u);
        }
        const bool s = uu >= nu0;
        const int64_t ul = uu - (s ? nu0 : 0), cbs = kpad >> 6;
        int8_t *const R = s ? s1.R : s0.R;
        const int64_t plane = s ? s1.plane : s0.plane;
        const bool more = uu + gridDim.x < nu;
        if (more) load(uu + gridDim.x);
        v4i afn[K32];
        bool negn = false;
        // store phase: thread t takes plane t / 8 of the row block and 32
        // entries of the patch: row (t % 8) / 2, columns 32 (t % 2) .. + 31
        const int sp = tid >> 3, sj = tid & 7;
        int8_t *const Rb = R + res_off(2, 4 * (ul / cbs) + (sj >> 1), 64 * (ul % cbs) + 32 * (sj & 1), kpad);
        for (int rb = 0; rb < nrb; rb++) {
            v16i acc[2];   // acc[d]: the digit-d column of this lane's prime
#pragma unroll
            for (int k32 = 0; k32 < K32; k32++)
#pragma unroll
                for (int d = 0; d < 2; d++) {
                    const int r = rb < nrb - 1 ? rb * 64 + d * 32 + c32 : lastrow + d * nlast + min(c32, nlast - 1);
                    const v4i bfr = *(const v4i *)(LC + r * 64 * KS + 16 * ((2 * k32 + hh) ^ cres_swz(r, C)));
                    acc[d] = mfma32(af[k32], bfr, k32 == 0 ? (v16i){0} : acc[d]);
                }
            if (more && rb == cb) operand(afn, negn);
            // |acc| <= kb 2^14 <= 2^22, so v = acc_0 + 256 acc_1 is below 2^31 -
            // 2^22 in magnitude and congruent to x - 128 sum 2^(8b) (to -x + 127
            // sum 2^(8b) for a negative entry); u = v + K with K = (128 resp.
            // -127) sum 2^(8b) mod p + off32 + h lies in [0, 2^32) and is
            // congruent to x + h, so (u mod p) - h is the centred residue and
            // (u mod p) + 128 - h carries its two balanced digits: byte 0 ^ 0x80
            // and byte 1.  Padding entries have no limbs and come out 0.
            const uint4 k4 = s_pc[rb * 32 + c32];   // -p, mu32, K (positive), K (negative)
            const uint32_t c128h = 128u - ((uint32_t)-(int32_t)k4.x >> 1), dk = k4.w - k4.z;
#pragma unroll
            for (int g = 0; g < 4; g++) {
                uint32_t v[4];
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const uint32_t sel = (uint32_t)__builtin_amdgcn_sbfe((int)m16, 4 * g + i, 1);   // 0 or ~0
                    const uint32_t u = (uint32_t)(acc[0][4 * g + i] + (acc[1][4 * g + i] << 8)) + k4.z + (dk & sel);
                    v[i] = bred32(u, (int32_t)k4.x, k4.y) + c128h;
                }
                // digit d of entries 8g + 4hh .. + 3: byte d of v[0..3]
                const uint32_t t01 = __builtin_amdgcn_perm(v[1], v[0], 0x05010400u);
                const uint32_t t23 = __builtin_amdgcn_perm(v[3], v[2], 0x05010400u);
                uint8_t *const lo = LO + 2 * c32 * CRES_LO_PITCH + 32 * wave + 8 * g + 4 * hh;
                *(uint32_t *)lo = __builtin_amdgcn_perm(t23, t01, 0x05040100u) ^ 0x80808080u;
                *(uint32_t *)(lo + CRES_LO_PITCH) = __builtin_amdgcn_perm(t23, t01, 0x07060302u);
            }
            lds_barrier();
            {
                const uint32_t *src = (const uint32_t *)(LO + sp * CRES_LO_PITCH + 32 * sj);
                uint32_t w8[8];
#pragma unroll
                for (int x = 0; x < 8; x++) w8[x] = src[x];
                const int grow = rb * 64 + sp;   // plane 2 q + d
                if (grow < SP) {
                    *(uint4 *)(Rb + grow * plane) = make_uint4(w8[0], w8[1], w8[2], w8[3]);
                    *(uint4 *)(Rb + grow * plane + 16) = make_uint4(w8[4], w8[5], w8[6], w8[7]);
                }
            }
            lds_barrier();
        }
#pragma unroll
        for (int k32 = 0; k32 < K32; k32++) af[k32] = afn[k32];
        neg = negn;
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
// 2b. residue GEMM for two slices (primes below 65280).  One workgroup
//     takes a (64 WM) x 128 tile of a pair of primes 2w, 2w+1 in turn; 2 WM waves
//     of 64 x 64 on v_mfma_i32_32x32x32_i8, 3 diagonal accumulators
//     (s + t = 0, 1, 2).  K runs in steps of 64 bytes through a ring of LDS
//     stages filled by global_load_lds_dwordx4 (global memory straight into
//     LDS, no staging registers) STAGES - 1 steps ahead.  The launches use
//     WM = 2 with two stages, two workgroups per CU: one workgroup's fold
//     (VALU) then runs beside the other's MFMAs, which measured faster than
//     deeper rings or 256-row tiles with one workgroup per CU.  The pair's digits leave as
//     whole 32-bit words (prime 2w low, 2w+1 high) through LDS.
//
//     Stage layout: [A | B][slice][row 0..127][4 chunks of 16 bytes], the
//     chunk of K bytes 16c..16c+15 of row r at slot c ^ g(r) with
//     g(r) = (r >> 2) & 3: rows 64 bytes apart share banks every fourth row
//     (64 four-byte banks), and g gives rows r, r + 4, r + 8, r + 12 distinct
//     slots, which made the fragment reads free of bank conflicts
//     (SQ_LDS_BANK_CONFLICT 0; XOR patterns on row bits 0-1 cost 2-4x).  One DMA
//     instruction of a wave writes 1 KB (16 rows) at base + 16*lane; lane l
//     takes row 16*rb16 + l/4, slot l%4, i.e. chunk (l & 3) ^ g(row).
// ------------------------------------------------------------------------
constexpr int KCHUNK2 = 32768;   // 2 pairs * 32768 * 2^14 = 2^30 < 2^31
// Tile shapes: WM wave rows of 64 (WG tile 64*WM x 128, 2*WM waves) and a
// ring of STAGES LDS stages; each stage holds 2 slices x 64*WM rows of A and
// 2 slices x 128 rows of B, 64 bytes of K each.
template <int WM> struct G2Shape {
    static constexpr int M = 64 * WM, THREADS = 128 * WM, WAVES = 2 * WM;
    static constexpr int A_BYTES = 2 * M * 64, STAGE = A_BYTES + 2 * 128 * 64;
    static constexpr int A_DMA = 2 * M / 16, DMA = A_DMA + 16, DMA_PER_WAVE = DMA / WAVES;
};

typedef __attribute__((address_space(3))) void lds_void;
// word 3 of a raw buffer resource on gfx9-family chips (32-bit data format)
constexpr int BUF_DWORD3 = 0x00020000;

__device__ __forceinline__ uint32_t lds_offset(const void *p) { return (uint32_t)(uintptr_t)(const lds_void *)p; }
__device__ __forceinline__ int g2_swz(int r) { return (r >> 2) & 3; }

// 16 bytes from sbase + voff (sbase wave-uniform, voff per lane) to LDS at
// lds_base + 16 * lane (lds_base wave-uniform).  Issued as inline assembly so
// that the compiler does not wait on it: the caller orders it with
// wait_vmcnt and a barrier.
__device__ __forceinline__ void dma16(const int8_t *sbase, uint32_t voff, uint32_t lds_base)
{
    asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %1, %2" ::"s"(lds_base), "v"(voff),
                 "s"(sbase)
                 : "memory", "m0");
}
// a pointer that is the same in every lane, as the compiler can see it
__device__ __forceinline__ const int8_t *uniform_ptr(const int8_t *p)
{
    const uint64_t v = (uint64_t)(uintptr_t)p;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return (const int8_t *)(uintptr_t)(((uint64_t)hi << 32) | lo);
}
// wait until at most N vector memory operations of this wave are in flight
template <int N>
__device__ __forceinline__ void wait_vmcnt()
{
    __builtin_amdgcn_s_waitcnt((N & 15) | (((N >> 4) & 3) << 14) | 0x0F70);
}
// wait until at most n * PER of this wave's DMAs are in flight (n < 4)
template <int PER>
__device__ __forceinline__ void wait_stages(int n)
{
    if (n >= 3) wait_vmcnt<3 * PER>();
    else if (n == 2) wait_vmcnt<2 * PER>();
    else if (n == 1) wait_vmcnt<PER>();
    else wait_vmcnt<0>();
}

template <bool CHUNKED, int WM, int STAGES>
__device__ __forceinline__ void gemm_tile2(int64_t v,
    const PrimeConst *__restrict__ pc, int P, const int8_t *__restrict__ RA, const int8_t *__restrict__ RB,
    int64_t planeA, int64_t planeB, int64_t kpad, int64_t rowsC, int64_t colsC, int ntm, int ntn,
    uint32_t *__restrict__ Y, int64_t ystride)
{
    using Sh = G2Shape<WM>;
    static_assert(Sh::DMA % Sh::WAVES == 0, "DMA instructions must split evenly over the waves");
    static_assert(STAGES * Sh::STAGE >= Sh::M * 128 * 4, "the output words reuse the ring");
    __shared__ __attribute__((aligned(1024))) int8_t LS[STAGES * Sh::STAGE];
    uint32_t *LW = (uint32_t *)LS;   // after the K loop: the M x 128 output words
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wave_u = __builtin_amdgcn_readfirstlane(wave);
    const int wi = wave >> 1, wj = wave & 1, l32 = lane & 31, h = lane >> 5;

    const int64_t tiles = (int64_t)ntm * ntn;
    const int pair = (int)(v / tiles);
    const int64_t t = v % tiles;
    const int64_t i0 = (t / ntn) * Sh::M, j0 = (t % ntn) * 128;
    uint32_t *Yw = Y + (int64_t)pair * ystride;
    const int nsteps = (int)(kpad / 64);
    const uint32_t ls0 = __builtin_amdgcn_readfirstlane(lds_offset(LS));

    uint32_t hold[2][2][8];   // digits of prime 2w, two 16-bit values per register

    for (int half = 0; half < 2; half++) {
        const int q = 2 * pair + half;
        if (q >= P) break;
        const PrimeConst c = pc[q];
        const uint32_t mu32 = (uint32_t)(c.mu >> 32), off32 = c.off32, hp = c.p >> 1, c128h = 128u - hp;
        const int32_t negp = -(int32_t)c.p;
        const int8_t *gA = RA + (int64_t)q * 2 * planeA;
        const int8_t *gB = RB + (int64_t)q * 2 * planeB;
        // this wave's DMA sources per stage; DMA instruction idx fills 16 rows
        // (1 KB): idx < A_DMA: A slice idx / (M/16), rows 16 (idx % (M/16));
        // else B slice (idx - A_DMA) / 8, rows 16 ((idx - A_DMA) % 8).  The
        // 16-row block's start is a wave-uniform base, the lane's row and
        // chunk a 32-bit offset.
        // g2_swz(16 rb16 + lane / 4) = (lane >> 4) & 3 for every rb16, so the
        // lane's chunk, hence its offset, is the same for every DMA
        const int8_t *sbase[Sh::DMA_PER_WAVE];
        uint32_t dst[Sh::DMA_PER_WAVE];
        const uint32_t c0 = (lane & 3) ^ ((lane >> 4) & 3);
        // (residue planes in 16-row x 64-byte blocks: row l/4 of a block is 64 bytes in)
        const uint32_t voff = (uint32_t)((lane >> 2) * 64) + 16 * c0;
#pragma unroll
        for (int u = 0; u < Sh::DMA_PER_WAVE; u++) {
            const int idx = wave_u * Sh::DMA_PER_WAVE + u;
            const bool isA = idx < Sh::A_DMA;
            const int j = isA ? idx : idx - Sh::A_DMA, rbn = isA ? Sh::M / 16 : 8;
            const int s = j / rbn, rb16 = j % rbn, rows = isA ? Sh::M : 128;
            sbase[u] = uniform_ptr(isA ? gA + s * planeA + (i0 + rb16 * 16) * kpad
                                       : gB + s * planeB + (j0 + rb16 * 16) * kpad);
            dst[u] = (isA ? 0 : Sh::A_BYTES) + (s * rows + rb16 * 16) * 64;
        }
#define ISSUE(k_)                                                                                     \
        _Pragma("unroll") for (int u = 0; u < Sh::DMA_PER_WAVE; u++)                                  \
            dma16(sbase[u] + (int64_t)(k_) * 1024, voff,                                                \
                  __builtin_amdgcn_readfirstlane(ls0 + ((k_) % STAGES) * Sh::STAGE + dst[u]));

        v16i acc[3][2][2];
#pragma unroll
        for (int d = 0; d < 3; d++)
#pragma unroll
            for (int rt = 0; rt < 2; rt++)
#pragma unroll
                for (int ct = 0; ct < 2; ct++) acc[d][rt][ct] = (v16i){0};

        // x = acc0 + 256 (acc1 + 256 acc2) mod p plus the running residue of
        // earlier chunks (kept in [0, p) in Y), in signed 32-bit Barrett steps
        // (|acc_d| <= 2^30 within a chunk; the two inner steps leave values
        // below 2p < 2^17 uncorrected, so every sum stays below 2^31 - 2^17);
        // the last fold turns it into y = x (M/p)^-1 mod p, centred, as two
        // balanced digits: with h = floor(p/2), (x (M/p)^-1 + h) mod p - h is
        // the centred residue, and adding 128 gives the digits' bytes
#define FOLD2(first_, last_, h_)                                                                     \
        _Pragma("unroll") for (int rt = 0; rt < 2; rt++)                                             \
        _Pragma("unroll") for (int ct = 0; ct < 2; ct++)                                             \
        _Pragma("unroll") for (int r = 0; r < 16; r++) {                                             \
            const int64_t i = i0 + wi * 64 + rt * 32 + 8 * (r >> 2) + 4 * h + (r & 3);               \
            const int64_t j = j0 + wj * 64 + ct * 32 + l32;                                          \
            uint16_t *yh = (uint16_t *)(Yw + i * colsC + j) + (h_);                                  \
            const bool in = i < rowsC && j < colsC;                                                  \
            const int32_t prev = (first_ || !in) ? 0 : (int32_t)*yh;                                 \
            const uint32_t r2 = smod32_lazy(acc[2][rt][ct][r], negp, mu32, off32);                   \
            const uint32_t z = smod32_lazy(acc[1][rt][ct][r] + 256 * (int32_t)r2, negp, mu32, off32); \
            uint32_t res = smod32(acc[0][rt][ct][r] + 256 * (int32_t)z + prev, negp, mu32, off32);  \
            if (last_)   /* res inv + h < p^2 + p < 2^32 */                                          \
                res = ((bred32(__umul24(res, c.inv) + hp, negp, mu32) + c128h) ^ 0x80u) & 0xffffu;    \
            if (!CHUNKED && last_) {                                                                 \
                if ((h_) == 0) {                                                                     \
                    if (r & 1) hold[rt][ct][r >> 1] |= res << 16;                                    \
                    else hold[rt][ct][r >> 1] = res;                                                 \
                } else {                                                                             \
                    LW[(wi * 64 + rt * 32 + 8 * (r >> 2) + 4 * h + (r & 3)) * 128 + wj * 64 + ct * 32 + l32] = \
                        ((hold[rt][ct][r >> 1] >> (16 * (r & 1))) & 0xffff) | (res << 16);           \
                }                                                                                    \
            } else if (in) {                                                                         \
                *yh = (uint16_t)res;                                                                 \
            }                                                                                        \
        }                                                                                            \
        _Pragma("unroll") for (int d = 0; d < 3; d++)                                                \
        _Pragma("unroll") for (int rt = 0; rt < 2; rt++)                                             \
        _Pragma("unroll") for (int ct = 0; ct < 2; ct++) acc[d][rt][ct] = (v16i){0};

        __syncthreads();   // every wave is done with the previous prime's stages
#pragma unroll
        for (int k = 0; k < STAGES - 1; k++)
            if (k < nsteps) { ISSUE(k); }
        int64_t since = 0;
        bool first = true;
        for (int k = 0; k < nsteps; k++) {
            // stage k has landed once only the later stages' DMAs (up to STAGES - 2) are in flight
            wait_stages<Sh::DMA_PER_WAVE>(min(nsteps - 1, k + STAGES - 2) - k);
            __syncthreads();
            // stage (k + STAGES - 1) % STAGES was last read in step k-1, which every wave has finished
            if (k + STAGES - 1 < nsteps) { ISSUE(k + STAGES - 1); }
            const int8_t *sa = LS + (k % STAGES) * Sh::STAGE, *sb = sa + Sh::A_BYTES;
#pragma unroll
            for (int kk = 0; kk < 2; kk++) {
                const int ch = 2 * kk + h;
                v4i bf[2][2];
#pragma unroll
                for (int u = 0; u < 2; u++)
#pragma unroll
                    for (int x = 0; x < 2; x++) {
                        const int rb = wj * 64 + x * 32 + l32;
                        bf[u][x] = *(const v4i *)(sb + (u * 128 + rb) * 64 + 16 * (ch ^ g2_swz(rb)));
                    }
#pragma unroll
                for (int s = 0; s < 2; s++) {   // one slice of A at a time: fewer live fragments
                    v4i af[2];
#pragma unroll
                    for (int x = 0; x < 2; x++) {
                        const int ra = wi * 64 + x * 32 + l32;
                        af[x] = *(const v4i *)(sa + (s * Sh::M + ra) * 64 + 16 * (ch ^ g2_swz(ra)));
                    }
#pragma unroll
                    for (int u = 0; u < 2; u++)
#pragma unroll
                        for (int rt = 0; rt < 2; rt++)
#pragma unroll
                            for (int ct = 0; ct < 2; ct++)
                                acc[s + u][rt][ct] = mfma32(af[rt], bf[u][ct], acc[s + u][rt][ct]);
                }
            }
            since += 64;
            if (CHUNKED && since == KCHUNK2 && k + 1 < nsteps) {
                FOLD2(first, false, half);
                first = false;
                since = 0;
            }
        }
        __syncthreads();   // the output words reuse the stages
        if (CHUNKED) {
            if (first) { FOLD2(true, true, half); }
            else { FOLD2(false, true, half); }
        } else if (half == 0) {   // one K chunk: this is the only fold
            FOLD2(true, true, 0);
        } else {
            FOLD2(true, true, 1);
        }
    }
#undef ISSUE
#undef FOLD2
    if (!CHUNKED) {
        if (2 * pair + 1 >= P) {   // odd prime count: the pair's high half is zero
#pragma unroll
            for (int rt = 0; rt < 2; rt++)
#pragma unroll
                for (int ct = 0; ct < 2; ct++)
#pragma unroll
                    for (int r = 0; r < 16; r++)
                        LW[(wi * 64 + rt * 32 + 8 * (r >> 2) + 4 * h + (r & 3)) * 128 + wj * 64 + ct * 32 + l32] =
                            (hold[rt][ct][r >> 1] >> (16 * (r & 1))) & 0xffff;
        }
        __syncthreads();
        // THREADS/32 rows of 128 words per pass, four words per thread: 512
        // contiguous bytes per row, 16-byte stores when the row's words are
        // 16-byte aligned (colsC a multiple of 4) and inside C
        const int64_t jq = j0 + 4 * (tid & 31);
        const bool vec = (colsC & 3) == 0 && jq + 3 < colsC;
        for (int pass = 0; pass < Sh::M * 32 / Sh::THREADS; pass++) {
            const int row = (Sh::THREADS / 32) * pass + (tid >> 5);
            const int64_t i = i0 + row;
            if (i >= rowsC) continue;
            const uint4 w = *(const uint4 *)(LW + row * 128 + 4 * (tid & 31));
            if (vec) {
                *(uint4 *)(Yw + i * colsC + jq) = w;
            } else {
                if (jq < colsC) Yw[i * colsC + jq] = w.x;
                if (jq + 1 < colsC) Yw[i * colsC + jq + 1] = w.y;
                if (jq + 2 < colsC) Yw[i * colsC + jq + 2] = w.z;
                if (jq + 3 < colsC) Yw[i * colsC + jq + 3] = w.w;
            }
        }
    }
}

// ------------------------------------------------------------------------
// 1c. reduce as an MFMA GEMM (S = 2, entries wider than the entry-resident
//     tiles take).  The residues of an entry are sum_b byte_b (digits of
//     2^(8b) mod p), a product [entries x kb] x [kb x 2P]; for wide entries
//     this is the largest MFMA work of a product (2 S P kb ops per entry), so
//     it runs as a tiled GEMM fed by LDS DMA like the residue GEMM.
//     Workgroup tile: 128 entries (a 2-row x 64-column patch of the residue
//     plane) x 256 table rows (128 primes in groups of 32: the digit-0 rows,
//     then the digit-1 rows of the same primes); 4 waves of 64 entries x 128
//     table rows.  The entries are the MFMA A operand, so a lane's output
//     column is one prime's digit and both digits of a prime meet in one
//     register of two column tiles.  K runs in 64-byte steps through a
//     RG_STAGES-deep LDS ring; the entries' 16-byte chunks come straight from
//     their limbs (per-lane 64-bit addresses, kept inside the entry) and are
//     masked at |size| and biased (byte - 128, or 127 - byte for a negative
//     entry) when a wave reads its fragments.  The table is read from a
//     copy in 1 KB blocks (rg_table_blocked), so that each table DMA
//     instruction reads one contiguous kilobyte: rows kb bytes apart (a power
//     of two) put a DMA's sixteen rows on few memory channels, which measured
//     1.5x slower for the whole kernel.  Stage layout: [entries 128 x
//     64 B][table 256 x 64 B], chunk c of row r at slot c ^ g2_swz(r).  The
//     epilogue is the constant-resident reduce's with the digit-1 sums
//     reduced first (|acc| <= kb 2^14 < 2^27 for the widths two-digit plans
//     allow); the residues leave through LDS as 128 contiguous bytes per
//     plane.  Tile order: groups of RG_GE entry tiles; inside a group the
//     entry tile varies fastest, so the workgroups an XCD runs at once share
//     a few entry tiles and a few table tiles in its L2.
// ------------------------------------------------------------------------
constexpr int RG_STAGES = 3;
constexpr int RG_STAGE = 128 * 64 + 256 * 64;   // 24 KB
constexpr int RG_OPITCH = 144;                  // bytes per plane of the output tile
constexpr int RG_GE = 8;                        // entry tiles per group
static_assert(RG_STAGES * RG_STAGE >= 256 * RG_OPITCH, "the output tile reuses the ring");
inline int64_t rg_tiles(int64_t nent, int64_t P) { return ceil_div(nent, 128) * ceil_div(P, 128); }
// The reduce table (rows S q + s of kb bytes, S = 2) in 1 KB blocks: block
// (qb, d, c) at ((2 qb + d) (kb / 64) + c) KB holds bytes 64 c .. 64 c + 63 of
// table row 2 q + d for the primes q = 16 qb .. 16 qb + 15 (one 64-byte row
// each, zero past the last prime); 16 qb runs to P rounded up to 128, the
// primes a table tile spans.
inline std::vector<int8_t> rg_table_blocked(const int8_t *cst, int64_t P, int64_t kb)
{
    const int64_t nqb = ceil_div(P, 128) * 8, nc = kb / 64;
    std::vector<int8_t> out((size_t)(nqb * 2 * nc * 1024), 0);
    for (int64_t qb = 0; qb < nqb; qb++)
        for (int64_t d = 0; d < 2; d++)
            for (int64_t c = 0; c < nc; c++)
                for (int64_t i = 0; i < 16; i++) {
                    const int64_t q = 16 * qb + i;
                    if (q < P)
                        memcpy(out.data() + ((2 * qb + d) * nc + c) * 1024 + 64 * i, cst + (2 * q + d) * kb + 64 * c, 64);
                }
    return out;
}

// 16 bytes from the lane's own address to LDS at lds_base + 16 * lane
__device__ __forceinline__ void dma16v(const void *addr, uint32_t lds_base)
{
    asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %1, off" ::"s"(lds_base), "v"(addr) : "memory", "m0");
}

__device__ __forceinline__ void reduce_gemm_tile(int64_t v, const PrimeConst *__restrict__ pc,
    const int8_t *__restrict__ cstb, int P, int kb, int xbits, const uint64_t *__restrict__ X, int64_t Xrs, int64_t W,
    int transpose, int64_t rows, int64_t K, int64_t kpad, int64_t nent, int8_t *__restrict__ R, int64_t plane,
    unsigned *__restrict__ err)
{
    constexpr int PER = 6;   // DMA instructions per wave and stage: 2 of entries, 4 of the table
    __shared__ __attribute__((aligned(1024))) int8_t LS[RG_STAGES * RG_STAGE];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wave_u = __builtin_amdgcn_readfirstlane(wave);
    const int wi = wave >> 1, wj = wave & 1, c32 = lane & 31, h = lane >> 5;
    // tile v -> (entry tile et, table tile ct)
    const int64_t nct = (P + 127) / 128, net = (nent + 127) / 128, gsz = RG_GE * nct;
    const int64_t grp = v / gsz, inner = v % gsz, ge = min<int64_t>(RG_GE, net - grp * RG_GE);
    const int64_t ct = inner / ge, et = grp * RG_GE + inner % ge, cbs = kpad >> 6;
    // the patch: rows r0, r0 + 1, columns col0 ..  A transposed operand's
    // entry (r, k) sits at row k of the matrix, so its patches run down the
    // rows first: the tiles of a group then read the same 64 matrix rows (a
    // few pages), where column-first patches would spread one group over
    // every matrix row (measured 1.6x slower)
    const int64_t npr = nent / kpad / 2;
    const int64_t r0 = transpose ? 2 * (et % npr) : 2 * (et / cbs);
    const int64_t col0 = transpose ? 64 * (et / npr) : 64 * (et % cbs);
    const int nsteps = kb / 64;
    const int lim = (int)min<int64_t>(kb / 8, W - 1);   // limbs an entry can contribute
    const uint32_t ls0 = __builtin_amdgcn_readfirstlane(lds_offset(LS));

    // entry el = 64 row + column of the patch: limb base (padding: entry 0),
    // limbs below |size| that count, sign
    auto entry = [&](int el, int64_t &base, int &nl, bool &neg) {
        const int64_t r = r0 + (el >> 6), k = col0 + (el & 63);
        const bool real = r < rows && k < K;
        base = real ? (transpose ? (k * Xrs + r) * W : (r * Xrs + k) * W) : 0;
        const int64_t sz = (int64_t)X[base + W - 1];
        const int64_t a = real ? min(sz < 0 ? -sz : sz, W - 1) : 0;
        nl = (int)min<int64_t>(a, lim);
        neg = real && sz < 0;
    };
    // DMA sources: this wave's entry instructions fill entry rows 32 wave + 16 u
    // .. + 15, its table instructions table rows 64 wave + 16 u ..; lane l takes
    // row + l / 4, slot l % 4, i.e. chunk (l % 4) ^ g2_swz(row) = (l % 4) ^ ((l >> 4) & 3)
    const int chunk = (lane & 3) ^ ((lane >> 4) & 3);
    const int8_t *eaddr[2];
    const int8_t *tbase[4];   // table rows 64 wave + 16 u ..: block (qb, d) = (8 ct + 2 wave + u % 2, u / 2)
#pragma unroll
    for (int u = 0; u < 2; u++) {
        int64_t b;
        int nl_;
        bool ng;
        entry(32 * wave + 16 * u + (lane >> 2), b, nl_, ng);
        eaddr[u] = (const int8_t *)(X + b) + 16 * chunk;
    }
#pragma unroll
    for (int u = 0; u < 4; u++)
        tbase[u] = uniform_ptr(cstb + ((2 * (8 * ct + 2 * wave_u + (u & 1)) + (u >> 1)) * (int64_t)(kb / 64)) * 1024);
    const uint32_t tvoff = (uint32_t)(64 * (lane >> 2) + 16 * chunk);
    // bytes of an entry at offsets >= emax come from its last 16 bytes (masked)
    const int64_t emax = 8 * W - 16;
#define ISSUE(k_)                                                                                       \
    {                                                                                                   \
        const uint32_t sb_ = ls0 + ((k_) % RG_STAGES) * RG_STAGE;                                       \
        const int64_t ko_ = (int64_t)(k_) * 64;                                                         \
        _Pragma("unroll") for (int u = 0; u < 2; u++)                                                   \
            dma16v(eaddr[u] + min<int64_t>(ko_, emax - 16 * chunk),                                     \
                   __builtin_amdgcn_readfirstlane(sb_ + (32 * wave_u + 16 * u) * 64));                  \
        _Pragma("unroll") for (int u = 0; u < 4; u++)                                                   \
            dma16(tbase[u] + (k_) * 1024, tvoff,                                                        \
                  __builtin_amdgcn_readfirstlane(sb_ + 128 * 64 + (64 * wave_u + 16 * u) * 64));         \
    }

    // the fragments' entries (rows 64 wi + 32 rt + c32): limbs below |size| and sign
    int fnl[2];
    bool fneg[2];
#pragma unroll
    for (int rt = 0; rt < 2; rt++) {
        int64_t b;
        entry(64 * wi + 32 * rt + c32, b, fnl[rt], fneg[rt]);
        if (ct == 0 && h == 0) {
            const int64_t r = r0 + ((64 * wi + 32 * rt + c32) >> 6), k = col0 + ((64 * wi + 32 * rt + c32) & 63);
            if (r < rows && k < K) {
                const int64_t sz = (int64_t)X[b + W - 1];
                if (over_bound(X + b, min(sz < 0 ? -sz : sz, W - 1), xbits)) atomicOr(err, DEV_ERR_INPUT);
            }
        }
    }

    v16i acc[2][4];
#pragma unroll
    for (int rt = 0; rt < 2; rt++)
#pragma unroll
        for (int c = 0; c < 4; c++) acc[rt][c] = (v16i){0};
#pragma unroll
    for (int k = 0; k < RG_STAGES - 1; k++)
        if (k < nsteps) { ISSUE(k); }
    for (int k = 0; k < nsteps; k++) {
        wait_stages<PER>(min(nsteps - 1, k + RG_STAGES - 2) - k);
        __syncthreads();
        if (k + RG_STAGES - 1 < nsteps) { ISSUE(k + RG_STAGES - 1); }
        const int8_t *se = LS + (k % RG_STAGES) * RG_STAGE, *st = se + 128 * 64;
        // every fragment of the step is read before the first MFMA, so that
        // the second half's LDS reads run under the first half's MFMAs
        v4i raw[2][2], bfr[2][4];
#pragma unroll
        for (int kk = 0; kk < 2; kk++) {
            const int ch = 2 * kk + h;
#pragma unroll
            for (int rt = 0; rt < 2; rt++) {
                const int row = 64 * wi + 32 * rt + c32;
                raw[kk][rt] = *(const v4i *)(se + row * 64 + 16 * (ch ^ g2_swz(row)));
            }
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const int row = 128 * wj + 32 * c + c32;
                bfr[kk][c] = *(const v4i *)(st + row * 64 + 16 * (ch ^ g2_swz(row)));
            }
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int kk = 0; kk < 2; kk++) {
            v4i af[2];
#pragma unroll
            for (int rt = 0; rt < 2; rt++) {
                // limbs 8k + 4kk + 2h and + 1 of the entry; masked at |size|, then biased
                const int l0 = 8 * k + 4 * kk + 2 * h;
                const uint32_t flip = fneg[rt] ? 0x7f7f7f7fu : 0x80808080u;
                const v4i &r_ = raw[kk][rt];
                af[rt] = (v4i){(int)((l0 < fnl[rt] ? (uint32_t)r_[0] : 0u) ^ flip),
                               (int)((l0 < fnl[rt] ? (uint32_t)r_[1] : 0u) ^ flip),
                               (int)((l0 + 1 < fnl[rt] ? (uint32_t)r_[2] : 0u) ^ flip),
                               (int)((l0 + 1 < fnl[rt] ? (uint32_t)r_[3] : 0u) ^ flip)};
            }
#pragma unroll
            for (int rt = 0; rt < 2; rt++)
#pragma unroll
                for (int c = 0; c < 4; c++) acc[rt][c] = mfma32(af[rt], bfr[kk][c], acc[rt][c]);
        }
    }
#undef ISSUE
    __syncthreads();   // the output tile reuses the ring

    // epilogue: lane column c32 of column tiles 2 pg and 2 pg + 1 is prime
    // q = 128 ct + 64 wj + 32 pg + c32 (digits 0 and 1); register 4g + i of
    // row tile rt is entry 64 wi + 32 rt + 8g + 4h + i of the patch
    uint8_t *const LO = (uint8_t *)LS;
#pragma unroll
    for (int pg = 0; pg < 2; pg++) {
        const int64_t q = min<int64_t>(128 * ct + 64 * wj + 32 * pg + c32, P - 1);
        const PrimeConst &c_ = pc[q];
        const uint32_t hp = c_.p >> 1, mu32 = (uint32_t)(c_.mu >> 32), c128h = 128u - hp;
        const int32_t negp = -(int32_t)c_.p;
        const uint32_t kpos = c_.kofs + c_.off32 + hp, dk = c_.kneg + c_.off32 + hp - kpos;
#pragma unroll
        for (int rt = 0; rt < 2; rt++) {
            uint32_t m16;
            {
                const uint32_t nm = (uint32_t)__ballot(fneg[rt]) >> (4 * h);   // bit e: entry 32 rt + e (of the wave's 64)
                m16 = (nm & 0xfu) | ((nm >> 4) & 0xf0u) | ((nm >> 8) & 0xf00u) | ((nm >> 12) & 0xf000u);
            }
#pragma unroll
            for (int g = 0; g < 4; g++) {
                uint32_t vv[4];
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const uint32_t sel = (uint32_t)__builtin_amdgcn_sbfe((int)m16, 4 * g + i, 1);
                    // |acc| < 2^27: the digit-1 sum reduced below 2p first, then
                    // u = acc_0 + 256 r1 + K lies in [0, 2^32)
                    const uint32_t r1 = smod32_lazy(acc[rt][2 * pg + 1][4 * g + i], negp, mu32, c_.off32);
                    const uint32_t u = (uint32_t)(acc[rt][2 * pg][4 * g + i] + (int32_t)(r1 << 8)) + kpos + (dk & sel);
                    vv[i] = bred32(u, negp, mu32) + c128h;
                }
                const uint32_t t01 = __builtin_amdgcn_perm(vv[1], vv[0], 0x05010400u);
                const uint32_t t23 = __builtin_amdgcn_perm(vv[3], vv[2], 0x05010400u);
                const int pl = 2 * (64 * wj + 32 * pg + c32);   // plane of digit 0 in the tile
                uint8_t *const lo = LO + pl * RG_OPITCH + 64 * wi + 32 * rt + 8 * g + 4 * h;
                *(uint32_t *)lo = __builtin_amdgcn_perm(t23, t01, 0x05040100u) ^ 0x80808080u;
                *(uint32_t *)(lo + RG_OPITCH) = __builtin_amdgcn_perm(t23, t01, 0x07060302u);
            }
        }
    }
    __syncthreads();
    // 256 planes x 128 bytes: eight lanes per plane, 16 bytes each (entries
    // 16 j .. 16 j + 15: patch row j / 4, columns 16 (j % 4) ..)
    const int j = lane & 7;
    int8_t *const Rb = R + res_off(2, r0 + (j >> 2), col0 + 16 * (j & 3), kpad);
#pragma unroll
    for (int pass = 0; pass < 8; pass++) {
        const int pl = 32 * pass + 8 * wave + (lane >> 3);
        const int64_t gpl = 256 * ct + pl;   // global plane 2 q + d
        const uint4 w4 = *(const uint4 *)(LO + pl * RG_OPITCH + 16 * j);
        if (gpl < 2 * (int64_t)P) *(uint4 *)(Rb + gpl * plane) = w4;
    }
}

// ------------------------------------------------------------------------
// 3. reconstruct, in two passes.
//    (a) recon_gemm_tile: X' = sum_q yc_q (M/p_q), with yc_q the centred
//        y_q, as an int8 GEMM of the digits of (M/p_q) 256^s (rows = result
//        byte positions) against the digits of yc_q (columns = entries).  The
//        epilogue folds each group of four byte rows into one signed 32-bit
//        word value (|value| < 2^49); the row block's 32 words of an entry
//        then run a carry chain, leaving 32-bit words Wv[t][e] and the
//        block's carry-out (|c| < 2^18) in Cb[rb][e], which belongs at word
//        32 (rb + 1), or at word nw for the top block.  The
//        workgroups also form gamma = round(sum_q yc_q / p_q) = round(X'/M),
//        so that C = X' - gamma M (|C| <= M/4), in fixed point: row block rb
//        takes the K steps ks = rb mod nrb and writes its partial sum to
//        G[rb][e] (int64; S = 2: sum of yc floor(2^45/p), each term short by
//        less than 2^-30; S = 4: sum of floor(y floor(2^63/p) / 2^31) - 2^32
//        [yc < 0] with y = yc mod p, each short by less than 2^-31), far
//        inside the 1/4 margin for any prime count the plans allow.
//    (b) recon_carry_block: one thread per entry finds the sign of
//        X' - gamma M from its top words down (the top block's carry first),
//        then runs the carry chain of
//        |X' - gamma M| low to high; limbs leave through LDS so that each
//        entry's limbs are stored as whole 128-byte lines.
// ------------------------------------------------------------------------
constexpr int RECON_ROWS = 128;   // result bytes per workgroup
constexpr int RECON_ENT = 256;    // entries per workgroup pass (64 per wave)
// RESIDENT: the workgroup's whole row block of Mc (128 x Kr bytes) and (S = 2)
// floor(2^45/p) of every prime sit in dynamic LDS, loaded once, and the K loop runs without
// barriers; the workgroup then sweeps epw consecutive entry blocks with
// them (recon_epw: more for larger products, so that fewer workgroups load
// the block while every CU still gets several).  Used while that fits beside
// a second workgroup.
constexpr int RECON_RES_MAX_LDS = 144 * 1024;
inline int recon_epw(int64_t nent, int nrb)
{
    const int64_t blocks = (nent + RECON_ENT - 1) / RECON_ENT;
    return blocks * nrb >= 8 * 2048 ? 8 : blocks * nrb >= 4 * 1024 ? 4 : 2;
}
// Wv, Cb and G are stored in blocks of RECON_ENT entries: index t of entry e
// (t < nt) at (e / RECON_ENT) nt RECON_ENT + t RECON_ENT + e % RECON_ENT, so
// that a reconstruct workgroup's words and the carry pass's reads are long
// contiguous runs.
__device__ __forceinline__ int64_t recon_idx(int64_t t, int64_t e, int64_t nt)
{
    return (e / RECON_ENT) * nt * RECON_ENT + t * RECON_ENT + e % RECON_ENT;
}
static_assert(RECON_ROWS / 4 == 32, "the carry pass takes one row block of words per 16-limb chunk");
// (S = 2: floor(2^45/p) for every digit row q < Kr/S, zero past the last prime)
inline int64_t recon_res_lds(int64_t Kr, int64_t S) { return (int64_t)RECON_ROWS * (Kr + 16) + 4 * (Kr / S); }

// The 32 words of a reconstruct tile's row block for the lane's entry
// (128 result bytes x 256 entries; before the exchange, lane c32 + 32 h of
// wave w holds rows 8g + 4h + u of row tile rt for entries e0 + 64 w + 32 ct
// + c32).  Row 8g + 4h + u of row tile rt is byte u of word j = rt*8 + 2g + h
// of the block.  Lane c32 + 32h ends up with all 32 words of entry e0 + 64
// wave + 32 h + c32 (|word| < 2^49): v_permlane32_swap exchanges the upper
// half-wave of the ct = 0 values with the lower half-wave of the ct = 1
// values, after which the first operand holds word 2g' and the second word
// 2g' + 1 of the lane's own entry.
__device__ __forceinline__ void recon_words(const v16i (&acc)[4][2], int64_t (&wv)[32])
{
#pragma unroll
    for (int rt = 0; rt < 4; rt++)
#pragma unroll
        for (int g = 0; g < 4; g++) {
            uint64_t v[2];
#pragma unroll
            for (int ct = 0; ct < 2; ct++)
                v[ct] = (uint64_t)((int64_t)acc[rt][ct][4 * g] + ((int64_t)acc[rt][ct][4 * g + 1] << 8) +
                                   ((int64_t)acc[rt][ct][4 * g + 2] << 16) +
                                   ((int64_t)acc[rt][ct][4 * g + 3] << 24));
            const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)v[0], (unsigned)v[1], false, false);
            const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(v[0] >> 32), (unsigned)(v[1] >> 32),
                                                             false, false);
            wv[rt * 8 + 2 * g] = (int64_t)(((uint64_t)hi[0] << 32) | lo[0]);
            wv[rt * 8 + 2 * g + 1] = (int64_t)(((uint64_t)hi[1] << 32) | lo[1]);
        }
}

// Epilogue of a reconstruct tile: the row block's 32 words of each entry,
// carried, and its partial gamma sums.
__device__ __forceinline__ void recon_store(const v16i (&acc)[4][2], const int64_t (&gsum)[2], int64_t e0, int wave,
    int h, int c32, int rb, int nw, int nrb, int64_t nent, uint32_t *__restrict__ Wv, int32_t *__restrict__ Cb,
    int64_t *__restrict__ G)
{
    {
        int64_t wv[32];
        recon_words(acc, wv);
        const int64_t e = e0 + 64 * wave + 32 * h + c32;
        const int nv = min(RECON_ROWS / 4, nw - rb * (RECON_ROWS / 4));   // words of this block below nw
        int64_t carry = 0;
        if (nv == RECON_ROWS / 4 && e0 + RECON_ENT <= nent) {
            // the whole block for every entry (e0 is a multiple of RECON_ENT):
            // word j of entry e at a wave-uniform base + 4 (e % RECON_ENT) +
            // j RECON_ENT 4, stored with no address arithmetic and no branch;
            // cache policy 2 (non-temporal): the carry pass reads each word once
            const auto rs = __builtin_amdgcn_make_buffer_rsrc(
                (void *)(Wv + recon_idx((int64_t)rb * (RECON_ROWS / 4), e0, nw)), 0, -1, BUF_DWORD3);
            const int voff = (int)(e % RECON_ENT) * 4;
#pragma unroll
            for (int j = 0; j < RECON_ROWS / 4; j++) {
                const int64_t x = wv[j] + carry;   // |wv| < 2^49
                carry = x >> 32;
                __builtin_amdgcn_raw_buffer_store_b32((uint32_t)x, rs, voff, j * RECON_ENT * 4, 2);
            }
        } else {
#pragma unroll
            for (int j = 0; j < RECON_ROWS / 4; j++) {
                if (j < nv) {
                    const int64_t x = wv[j] + carry;
                    carry = x >> 32;
                    if (e < nent) Wv[recon_idx(rb * (RECON_ROWS / 4) + j, e, nw)] = (uint32_t)x;
                }
            }
        }
        if (e < nent) Cb[recon_idx(rb, e, nrb)] = (int32_t)carry;
    }
#pragma unroll
    for (int ct = 0; ct < 2; ct++) {   // lanes l and l + 32 hold the two halves of an entry's sum
        const uint64_t b = (uint64_t)gsum[ct];
        const uint32_t lo = __shfl_xor((unsigned)b, 32), hi = __shfl_xor((unsigned)(b >> 32), 32);
        const int64_t f = gsum[ct] + (int64_t)(((uint64_t)hi << 32) | lo);
        const int64_t e = e0 + 64 * wave + 32 * ct + c32;
        if (h == 0 && e < nent) G[recon_idx(rb, e, nrb)] = f;
    }
}

// gamma term of digit row q (S = 2: word wd's half sub; S = 4: the whole word)
// in the fixed point of the partial sums; zero for rows past the last prime
template <int S>
__device__ __forceinline__ int64_t gamma_term(uint32_t wd, int sub, int q, int P, const PrimeConst *__restrict__ pc,
                                              int32_t f45)
{
    if (S == 2) return (int64_t)unslice2(wd >> (16 * sub)) * f45;
    const PrimeConst &c_ = pc[min(q, P - 1)];
    const int32_t yc = unslice4(wd);
    const uint32_t y = yc < 0 ? (uint32_t)(yc + (int32_t)c_.p) : (uint32_t)yc;
    return q < P ? (int64_t)(((uint64_t)y * c_.fx) >> 31) - (yc < 0 ? (1ll << 32) : 0) : 0;
}

// Streaming reconstruct tile (S = 2, row blocks of Mc too large for LDS): the
// same 128-byte x 256-entry tile and epilogue as recon_gemm_tile, with both
// operands of every 64-byte K step brought into a ring of RS_STAGES LDS
// stages by global_load_lds_dwordx4, two steps ahead, one barrier per step.
// Stage layout: Mc rows 128 x 64 bytes (chunk c of row r at slot c ^
// g2_swz(r)), then the step's 16 word planes of Y, 256 entries each: plane p
// at rs_yplane(p) = 1024 p + 128 (p / 4), so that the four planes of a lane's
// fragment (8 kk + 4 h + i, i < 4) lie a multiple of 256 bytes apart (two
// reads each take one ds_read2st64_b32) and the two half-waves' groups start
// 32 banks apart.  Per wave and stage: two Mc DMAs (rows 16 (2 wave + u) ..)
// and four Y DMAs (planes 4 wave + u); the per-prime gamma constants come
// through the scalar cache, so that the only vector memory operations in the
// loop are the DMAs the waits count.
constexpr int RS_STAGES = 3;
__host__ __device__ constexpr int rs_yplane(int p) { return 1024 * p + 128 * (p / 4); }
constexpr int RS_STAGE = 128 * 64 + rs_yplane(16);
// One K sweep of a streaming tile: acc (128 result bytes from Mc row mrow0 on
// x the 256 entries from e0 on) and the gamma terms of the steps gmode picks
// (0: none, 1: k = rb mod nrb, 2: every step) added to gsum.  LS: the ring.
__device__ __forceinline__ void recon_sweep(int8_t *LS, const PrimeConst *__restrict__ pc, int P,
    const int8_t *__restrict__ Mc, int Kr, int64_t mrow0, const uint32_t *__restrict__ Y, int64_t ystride, int64_t e0,
    int gmode, int rb, int nrb, v16i (&acc)[4][2], int64_t (&gsum)[2])
{
    constexpr int PER = 6;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, h = lane >> 5, c32 = lane & 31;
    const int wave_u = __builtin_amdgcn_readfirstlane(wave);
    const int nsteps = Kr / 64, nyp = (2 * P + 3) / 4;
    const uint32_t ls0 = __builtin_amdgcn_readfirstlane(lds_offset(LS));
    // Mc: lane l takes row 16 (2 wave + u) + l / 4, slot l % 4, i.e. chunk (l % 4) ^ ((l >> 4) & 3)
    const uint32_t mvoff = (uint32_t)((lane >> 2) * Kr + 16 * ((lane & 3) ^ ((lane >> 4) & 3)));
    const int8_t *mbase[2];
#pragma unroll
    for (int u = 0; u < 2; u++) mbase[u] = uniform_ptr(Mc + (mrow0 + 16 * (2 * wave_u + u)) * Kr);
    const int8_t *ybase = uniform_ptr((const int8_t *)(Y + e0));
#define ISSUE(k_)                                                                                      \
    {                                                                                                  \
        const uint32_t sb_ = ls0 + ((k_) % RS_STAGES) * RS_STAGE;                                      \
        _Pragma("unroll") for (int u = 0; u < 2; u++)                                                  \
            dma16(mbase[u] + (k_) * 64, mvoff, __builtin_amdgcn_readfirstlane(sb_ + 16 * (2 * wave_u + u) * 64)); \
        _Pragma("unroll") for (int u = 0; u < 4; u++) {                                                \
            const int gp_ = min(16 * (k_) + 4 * wave_u + u, nyp - 1);                                  \
            dma16(ybase + (int64_t)gp_ * ystride * 4, (uint32_t)(16 * lane),                           \
                  __builtin_amdgcn_readfirstlane(sb_ + 128 * 64 + rs_yplane(4 * wave_u + u)));        \
        }                                                                                              \
    }

#pragma unroll
    for (int rt = 0; rt < 4; rt++)
#pragma unroll
        for (int ct = 0; ct < 2; ct++) acc[rt][ct] = (v16i){0};
#pragma unroll
    for (int k = 0; k < RS_STAGES - 1; k++)
        if (k < nsteps) { ISSUE(k); }
    for (int k = 0; k < nsteps; k++) {
        wait_stages<PER>(min(nsteps - 1, k + RS_STAGES - 2) - k);
        __syncthreads();
        if (k + RS_STAGES - 1 < nsteps) { ISSUE(k + RS_STAGES - 1); }
        const int8_t *sa = LS + (k % RS_STAGES) * RS_STAGE;
        const uint32_t *sy = (const uint32_t *)(sa + 128 * 64);
        v4i af[2][4], yf[2][2];
#pragma unroll
        for (int kk = 0; kk < 2; kk++) {
#pragma unroll
            for (int ct = 0; ct < 2; ct++)
#pragma unroll
                for (int i = 0; i < 4; i++)
                    yf[ct][kk][i] = (int)sy[rs_yplane(8 * kk + 4 * h + i) / 4 + 64 * wave + 32 * ct + c32];
#pragma unroll
            for (int rt = 0; rt < 4; rt++) {
                const int row = rt * 32 + c32;
                af[kk][rt] = *(const v4i *)(sa + row * 64 + 16 * ((2 * kk + h) ^ g2_swz(row)));
            }
        }
        if (gmode == 2 || (gmode == 1 && k % nrb == rb)) {   // gamma terms of this step
#pragma unroll
            for (int kk = 0; kk < 2; kk++)
#pragma unroll
                for (int i = 0; i < 4; i++)
#pragma unroll
                    for (int sub = 0; sub < 2; sub++) {
                        // both half-waves' primes from wave-uniform addresses; none past the last
                        const int q0 = 2 * (16 * k + 8 * kk + i) + sub, q1 = q0 + 8;
                        const int32_t f0 = q0 < P ? pc[min(q0, P - 1)].f45 : 0;
                        const int32_t f1 = q1 < P ? pc[min(q1, P - 1)].f45 : 0;
                        const int32_t f = h ? f1 : f0;
#pragma unroll
                        for (int ct = 0; ct < 2; ct++) gsum[ct] += gamma_term<2>((uint32_t)yf[ct][kk][i], sub, 0, P, pc, f);
                    }
        }
#pragma unroll
        for (int kk = 0; kk < 2; kk++)
#pragma unroll
            for (int rt = 0; rt < 4; rt++)
#pragma unroll
                for (int ct = 0; ct < 2; ct++) acc[rt][ct] = mfma32(af[kk][rt], yf[ct][kk], acc[rt][ct]);
    }
#undef ISSUE
}

__device__ __forceinline__ void recon_stream_tile(int64_t blk, const PrimeConst *__restrict__ pc, int P,
    const int8_t *__restrict__ Mc, int Kr, int nrb, int nw, const uint32_t *__restrict__ Y, int64_t ystride,
    int64_t nent, uint32_t *__restrict__ Wv, int32_t *__restrict__ Cb, int64_t *__restrict__ G)
{
    __shared__ __attribute__((aligned(1024))) int8_t LS[RS_STAGES * RS_STAGE];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, h = lane >> 5, c32 = lane & 31;
    const int rb = (int)(blk % nrb);
    const int64_t e0 = (blk / nrb) * RECON_ENT;
    v16i acc[4][2];
    int64_t gsum[2] = {0, 0};
    recon_sweep(LS, pc, P, Mc, Kr, (int64_t)rb * RECON_ROWS, Y, ystride, e0, 1, rb, nrb, acc, gsum);
    recon_store(acc, gsum, e0, wave, h, c32, rb, nw, nrb, nent, Wv, Cb, G);
}

template <int S, bool RESIDENT>
__device__ __forceinline__ void recon_gemm_tile_regs(int64_t blk, const PrimeConst *__restrict__ pc, int P,
    const int8_t *__restrict__ Mc, int Kr, int nrb, int nw, const uint32_t *__restrict__ Y, int64_t ystride,
    int64_t nent, uint32_t *__restrict__ Wv, int32_t *__restrict__ Cb, int64_t *__restrict__ G, int epw)
{
    __shared__ __attribute__((aligned(16))) int8_t LAs[RESIDENT ? 16 : RECON_ROWS * LDS_PITCH];
    extern __shared__ __attribute__((aligned(16))) int8_t LDYN[];
    int8_t *const LA = RESIDENT ? LDYN : LAs;
    const int apitch = RESIDENT ? Kr + 16 : LDS_PITCH;
    int32_t *const s_f45 = (int32_t *)(LDYN + RECON_ROWS * (Kr + 16));   // RESIDENT only
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, h = lane >> 5, c32 = lane & 31;
    const int rb = (int)(blk % nrb);
    const int64_t eb0 = (blk / nrb) * (RESIDENT ? epw : 1);

    if (RESIDENT) {
        // 128 rows x Kr bytes in 16-byte chunks, eight loads in flight per thread
        const int cpr = Kr / 16, nch = RECON_ROWS * cpr;
        for (int c0 = 0; c0 < nch; c0 += 8 * 256) {
            v4i t[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int c = min(c0 + tid + 256 * u, nch - 1), row = c / cpr, part = c - row * cpr;
                t[u] = *(const v4i *)(Mc + (int64_t)(rb * RECON_ROWS + row) * Kr + 16 * part);
            }
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int c = c0 + tid + 256 * u, row = c / cpr, part = c - row * cpr;
                if (c < nch) *(v4i *)(LA + row * apitch + 16 * part) = t[u];
            }
        }
        if (S == 2)
            for (int q = tid; q < Kr / S; q += 256) s_f45[q] = q < P ? pc[q].f45 : 0;
        __syncthreads();
    }

    for (int64_t eb = eb0; eb < eb0 + (RESIDENT ? epw : 1); eb++) {
    const int64_t e0 = eb * RECON_ENT;
    if (e0 >= nent) break;
    const uint32_t *yp = Y + e0 + 64 * wave + c32;

    v16i acc[4][2];
#pragma unroll
    for (int rt = 0; rt < 4; rt++)
#pragma unroll
        for (int ct = 0; ct < 2; ct++) acc[rt][ct] = (v16i){0};
    // Mc (L2-resident) is loaded one step ahead; Y two steps ahead into three
    // register sets that rotate by unrolling the loop by three, so that no
    // copy waits on a load just issued.
    v4i pa[2], y0[2][2], y1[2][2], y2[2][2];
    int64_t gsum[2] = {0, 0};
    // A: 128 rows x 64 bytes of Mc, two 16-byte chunks per thread
#define ALOAD(dst, k0_)                                                                                \
    _Pragma("unroll") for (int u = 0; u < 2; u++) {                                                    \
        const int ch = tid + 256 * u;                                                                  \
        dst[u] = *(const v4i *)(Mc + (int64_t)(rb * RECON_ROWS + (ch >> 2)) * Kr + (k0_) + 16 * (ch & 3)); \
    }
    // B fragments straight from Y: K index 4w+b of an entry is byte b of word
    // plane w.  While the planes span less than 4 GB they are read through one
    // buffer resource based at the workgroup's first entry (which the compiler
    // can see is wave-uniform): a load is the lane's fixed offset (its wave's
    // entries, entry column, half h: planes 4h .. 4h+3 of each group of 8)
    // plus a wave-uniform plane offset, with no per-load address arithmetic,
    // and planes past the last lie beyond the resource's range, where buffer
    // loads return 0 (the Mc columns k >= S*P are zero anyway).  Larger Y
    // regions are read with global loads, planes clamped to the last.
    const int nyp = (S * P + 3) / 4;
    const uint32_t ysb = (uint32_t)ystride * 4;
    const bool ybuf = (int64_t)nyp * ystride * 4 < (1ll << 32);
    const uint32_t yvoff = (uint32_t)(64 * wave + c32) * 4 + 4u * (uint32_t)h * ysb;
    // range: the bytes from entry e0 of plane 0 to the end of the last plane
    // (below 2^32 whenever ybuf, and e0 < ystride), in 32-bit scalar arithmetic
    const auto yrs = __builtin_amdgcn_make_buffer_rsrc(
        (void *)(Y + e0), 0, (int)((uint32_t)nyp * ysb - (uint32_t)e0 * 4u), BUF_DWORD3);
#define YLOAD(dst, k0_)                                                                  \
    if (ybuf) {                                                                          \
        const uint32_t so_ = (uint32_t)((k0_) / 4) * ysb;                                \
        _Pragma("unroll") for (int ct = 0; ct < 2; ct++)                                 \
        _Pragma("unroll") for (int kk = 0; kk < 2; kk++)                                 \
        _Pragma("unroll") for (int i = 0; i < 4; i++)                                    \
            dst[ct][kk][i] = (int)__builtin_amdgcn_raw_buffer_load_b32(                  \
                yrs, yvoff + 128 * ct, so_ + (8 * kk + i) * ysb, 0);                      \
    } else {                                                                             \
        _Pragma("unroll") for (int ct = 0; ct < 2; ct++)                                 \
        _Pragma("unroll") for (int kk = 0; kk < 2; kk++)                                 \
        _Pragma("unroll") for (int i = 0; i < 4; i++) {                                  \
            const int w = min((k0_) / 4 + 8 * kk + 4 * h + i, nyp - 1);                  \
            dst[ct][kk][i] = (int)yp[(int64_t)w * ystride + 32 * ct];                    \
        }                                                                                \
    }
#define RSTEP(ycur, ynext2)                                                                          \
    {                                                                                                \
        if (!RESIDENT) {                                                                             \
            __syncthreads();                                                                         \
            _Pragma("unroll") for (int u = 0; u < 2; u++) {                                          \
                const int ch = tid + 256 * u;                                                        \
                *(v4i *)(LA + (ch >> 2) * LDS_PITCH + 16 * (ch & 3)) = pa[u];                        \
            }                                                                                        \
            __syncthreads();                                                                         \
            if (k0 + 64 < Kr) { ALOAD(pa, k0 + 64); }                                                \
        }                                                                                            \
        if (k0 + 128 < Kr) { YLOAD(ynext2, k0 + 128); }                                              \
        if ((k0 / 64) % nrb == rb) {   /* this row block's share of the gamma sums */               \
            _Pragma("unroll") for (int kk = 0; kk < 2; kk++)                                         \
            _Pragma("unroll") for (int i = 0; i < 4; i++)                                            \
            _Pragma("unroll") for (int sub = 0; sub < 4 / S; sub++) {                                \
                const int q = (4 / S) * (k0 / 4 + 8 * kk + 4 * h + i) + sub;                         \
                /* no branch: rows past the last prime add nothing */                                \
                const int32_t f = S != 2 ? 0 : RESIDENT ? s_f45[q] : (q < P ? pc[min(q, P - 1)].f45 : 0); \
                _Pragma("unroll") for (int ct = 0; ct < 2; ct++)                                     \
                    gsum[ct] += gamma_term<S>((uint32_t)ycur[ct][kk][i], sub, q, P, pc, f);          \
            }                                                                                        \
        }                                                                                            \
        /* all eight A fragments of the step first, so that no MFMA waits on its own LDS read */     \
        v4i af[2][4];                                                                                \
        _Pragma("unroll") for (int kk = 0; kk < 2; kk++)                                             \
        _Pragma("unroll") for (int rt = 0; rt < 4; rt++)                                             \
            af[kk][rt] = *(const v4i *)(LA + (rt * 32 + c32) * apitch + (RESIDENT ? k0 : 0) + kk * 32 + 16 * h); \
        _Pragma("unroll") for (int kk = 0; kk < 2; kk++)                                             \
        _Pragma("unroll") for (int rt = 0; rt < 4; rt++)                                             \
        _Pragma("unroll") for (int ct = 0; ct < 2; ct++) acc[rt][ct] = mfma32(af[kk][rt], ycur[ct][kk], acc[rt][ct]); \
        k0 += 64;                                                                                    \
        if (k0 >= Kr) break;                                                                         \
    }
    if (!RESIDENT) { ALOAD(pa, 0); }
    YLOAD(y0, 0);
    if (64 < Kr) { YLOAD(y1, 64); }
    for (int k0 = 0;;) {
        RSTEP(y0, y2);
        RSTEP(y1, y0);
        RSTEP(y2, y1);
    }
#undef RSTEP
#undef ALOAD
#undef YLOAD
    recon_store(acc, gsum, e0, wave, h, c32, rb, nw, nrb, nent, Wv, Cb, G);
    }   // entry blocks
}

// One reconstruct tile: the streaming tile for two-digit plans whose row
// blocks do not fit LDS, else the register-fed tile above (RESIDENT: the row
// block sits in LDS for epw entry blocks).
template <int S, bool RESIDENT>
__device__ __forceinline__ void recon_gemm_tile(int64_t blk, const PrimeConst *__restrict__ pc, int P,
    const int8_t *__restrict__ Mc, int Kr, int nrb, int nw, const uint32_t *__restrict__ Y, int64_t ystride,
    int64_t nent, uint32_t *__restrict__ Wv, int32_t *__restrict__ Cb, int64_t *__restrict__ G, int epw = 1)
{
    if constexpr (S == 2 && !RESIDENT)
        recon_stream_tile(blk, pc, P, Mc, Kr, nrb, nw, Y, ystride, nent, Wv, Cb, G);
    else
        recon_gemm_tile_regs<S, RESIDENT>(blk, pc, P, Mc, Kr, nrb, nw, Y, ystride, nent, Wv, Cb, G, epw);
}

constexpr int CARRY_ENT = 256;   // entries per workgroup
constexpr int CARRY_PITCH = 17;  // limbs per LDS row (16 + 1 padding)

__device__ __forceinline__ void recon_carry_block(int64_t blk, const uint32_t *__restrict__ Wv,
    const int32_t *__restrict__ Cb, int64_t ystride, int nw, int nrb, const int64_t *__restrict__ G, int gshift,
    const uint32_t *__restrict__ M32, int64_t nent, int64_t colsC,
    uint64_t *__restrict__ C, int64_t Crs, int64_t W, int64_t cap, unsigned *__restrict__ err)
{
    __shared__ uint64_t LS[CARRY_ENT * CARRY_PITCH];
    __shared__ int64_t s_base[CARRY_ENT];
    const int tid = threadIdx.x;
    const int64_t e = blk * CARRY_ENT + tid;
    const bool valid = e < nent;
    s_base[tid] = valid ? ((e / colsC) * Crs + e % colsC) * W : -1;
    int64_t gam = 0;
    if (valid) {   // gamma: the rounded sum of the row blocks' fixed-point partial sums
        int64_t f = 0;
        for (int rb = 0; rb < nrb; rb++) f += G[recon_idx(rb, e, nrb)];
        gam = (f + (1ll << (gshift - 1))) >> gshift;
    }
    // this entry's words and block carries: stride 256 within its entry block
    const uint32_t *wp = Wv + (valid ? recon_idx(0, e, nw) : 0);
    const int32_t *cp = Cb + (valid ? recon_idx(0, e, nrb) : 0);
    constexpr int BW = RECON_ROWS / 4;   // words per row block

    // Sign: X' - gamma M = sum_t x_t 2^(32t) + c_top 2^(32 nw) with
    // x_t = Wv[t] - gam M32[t] (+ the carry of the block below at a block's
    // first word), |x_t| < 2^47, so the words below t add less than 2^16
    // units of 2^(32t); once the top part H_t reaches 2^28 in magnitude its
    // sign is the sign of the whole, and below that bound H_t * 2^32 + x_t
    // cannot overflow.
    bool neg = false;
    if (valid) {
        int64_t H = cp[(int64_t)(nrb - 1) * RECON_ENT];
        for (int t = nw - 1; t >= 0; t--) {
            int64_t x = (int64_t)wp[(int64_t)t * RECON_ENT] - gam * (int64_t)M32[t];
            if (t % BW == 0 && t > 0) x += cp[(int64_t)(t / BW - 1) * RECON_ENT];
            H = H * 4294967296ll + x;
            if (H >= (1ll << 28) || H <= -(1ll << 28)) break;
        }
        neg = H < 0;
    }

    const int nlimb = nw / 2, lim = (int)min<int64_t>(cap, nlimb);
    int64_t carry = 0;
    int64_t top = -1;
    bool hi_nz = false;
    for (int c0 = 0; c0 < nlimb; c0 += 16) {
        // unconditional loads (word index clamped, padding threads read entry 0)
        // so that all 32 are in flight at once; masked afterwards.  Word
        // indices are 32-bit (nw < 2^31), so their arithmetic stays in
        // scalar integer registers.
        int64_t xw[32];
#pragma unroll
        for (int j = 0; j < 32; j++) {
            const int t = min(2 * c0 + j, nw - 1);
            xw[j] = (int64_t)__builtin_nontemporal_load(wp + (int64_t)t * RECON_ENT) - gam * (int64_t)M32[t];
        }
        // 2 c0 is a multiple of 32 = BW: word 2 c0 takes the carry of block 2 c0 / BW - 1
        const int64_t cin = c0 > 0 ? (int64_t)cp[(int64_t)(2 * c0 / BW - 1) * RECON_ENT] : 0;
        xw[0] += cin;
#pragma unroll
        for (int j = 0; j < 32; j++)
            if (!valid || 2 * c0 + j >= nw) xw[j] = 0;
#pragma unroll
        for (int j = 0; j < 16; j++) {
            int64_t x0 = neg ? -xw[2 * j] : xw[2 * j];
            x0 += carry;
            carry = x0 >> 32;
            int64_t x1 = neg ? -xw[2 * j + 1] : xw[2 * j + 1];
            x1 += carry;
            carry = x1 >> 32;
            const uint64_t limb = (uint64_t)(uint32_t)x0 | ((uint64_t)(uint32_t)x1 << 32);
            const int64_t L = c0 + j;
            if (L < lim) {
                LS[tid * CARRY_PITCH + j] = limb;
                if (limb) top = L;
            } else if (L < nlimb) {
                hi_nz |= limb != 0;
            }
        }
        __syncthreads();
        // two limbs per thread and 16-byte store (entries are 8-byte aligned)
        typedef uint64_t v2u64 __attribute__((ext_vector_type(2), aligned(8)));
        const int64_t n = min<int64_t>(16, lim - c0);
        if (n > 0)
            for (int f = tid; f < CARRY_ENT * 8; f += CARRY_ENT) {
                const int ent = f >> 3, j = 2 * (f & 7);
                const int64_t base = s_base[ent];
                const uint64_t *src = LS + ent * CARRY_PITCH + j;
                if (j + 1 < n && base >= 0) *(v2u64 *)(C + base + c0 + j) = (v2u64){src[0], src[1]};
                else if (j < n && base >= 0) C[base + c0 + j] = src[0];
            }
        __syncthreads();
    }
    if (valid) {
        if (hi_nz) atomicOr(err, DEV_ERR_CAPACITY);
        C[s_base[tid] + W - 1] = (uint64_t)(neg ? -(top + 1) : top + 1);
    }
}

// ------------------------------------------------------------------------
// 3c. one-pass reconstruct (S = 2).  A workgroup takes 256 entries (e0 ..)
//     through every row block with recon_sweep and writes their limbs of
//     C = X' - gamma M itself: no word buffer between two passes.  The row
//     blocks are aligned to the top word: block nrb - 1 holds words nw - 32 ..
//     nw - 1, block b >= 1 words nw - 32 (nrb - b) .., block 0 the rest from
//     word 0 (nw - 32 (nrb - 1) of them; nw is even, so every block starts
//     on a limb).  The top block is swept first, with every K step's gamma
//     terms: it gives gamma and, from its words down, the sign of C as the
//     carry pass finds it (once the top part H reaches 2^28 in magnitude the
//     words below, |word| < 2^50, cannot change its sign); its words stay in
//     registers, carried within the block.  While an entry's sign is still
//     open (|C| below 2^(32 (nw - 32) + 28): products far below their
//     bound, zero ones) the workgroup sweeps further blocks downwards, which
//     only extend H (exact: raw words, no carries).  Then the blocks below
//     the top are swept from block 0 up and turned into limbs of |C| with one
//     running carry, each block's 16 limbs per entry leaving through LDS as
//     128-byte runs, and the top block's words close the chain; its carry-out
//     lies beyond the limbs (|C| < 2^(32 nw)), as in the carry pass.
// ------------------------------------------------------------------------
__device__ __forceinline__ int recon_block_start(int b, int nw, int nrb) { return b == 0 ? 0 : nw - 32 * (nrb - b); }

__device__ __forceinline__ void recon_fused_block(int64_t eb, const PrimeConst *__restrict__ pc, int P,
    const int8_t *__restrict__ Mc, int Kr, int nrb, int nw, int gw, const uint32_t *__restrict__ Y, int64_t ystride,
    int64_t nent, const uint32_t *__restrict__ M32, int64_t colsC, uint64_t *__restrict__ C, int64_t Crs, int64_t W,
    int64_t cap, unsigned *__restrict__ err)
{
    __shared__ __attribute__((aligned(1024))) int8_t LS[RS_STAGES * RS_STAGE];   // the K ring, then limb staging
    __shared__ int64_t s_base[RECON_ENT];
    static_assert(RS_STAGES * RS_STAGE >= RECON_ENT * CARRY_PITCH * 8, "the limb staging reuses the ring");
    const int tid = threadIdx.x, h = (tid & 63) >> 5;
    const int64_t e0 = eb * RECON_ENT;
    const int64_t e = e0 + tid;   // after recon_words, thread 64 wave + 32 h + c32 holds entry e0 + tid
    const bool valid = e < nent;
    s_base[tid] = valid ? ((e / colsC) * Crs + e % colsC) * W : -1;   // first read after the sweeps' barriers
    const int T = nrb - 1;
    v16i acc[4][2];
    int64_t gsum[2] = {0, 0};
    int64_t wv[32];

    // the top block, with gamma: from its words nw and nw + 1 when Mc carries the gamma rows (gw = 2),
    // else from every step's gamma terms
    const int nwg = nw + gw;
    const int wst = recon_block_start(T, nwg, nrb), nvt = nw - wst;
    recon_sweep(LS, pc, P, Mc, Kr, 4 * (int64_t)wst, Y, ystride, e0, gw ? 0 : 2, 0, nrb, acc, gsum);
    recon_words(acc, wv);
    int64_t gam;
    if (gw) {
        int64_t F = 0;
#pragma unroll
        for (int j = 0; j < 32; j++)
            F += j == nvt ? wv[j] : j == nvt + 1 ? (int64_t)((uint64_t)wv[j] << 32) : 0;
        gam = (F + (1ll << 44)) >> 45;
    } else {
        // lanes l and l + 32 hold the two K halves of the same entries' sums
        int64_t F[2];
#pragma unroll
        for (int ct = 0; ct < 2; ct++) {
            const uint64_t b = (uint64_t)gsum[ct];
            const uint32_t lo = __shfl_xor((unsigned)b, 32), hi = __shfl_xor((unsigned)(b >> 32), 32);
            F[ct] = gsum[ct] + (int64_t)(((uint64_t)hi << 32) | lo);
        }
        gam = ((h ? F[1] : F[0]) + (1ll << 44)) >> 45;
    }
    uint32_t topw[32];   // the top block's words, carried within it
    {
        int64_t c = 0;
#pragma unroll
        for (int j = 0; j < 32; j++) {
            const int64_t x = (j < nvt ? wv[j] : 0) + c;
            c = x >> 32;
            topw[j] = (uint32_t)x;
        }
    }
    int64_t H = 0;
    bool det = !valid;
#pragma unroll
    for (int j = 31; j >= 0; j--)
        if (j < nvt && !det) {
            H = H * 4294967296ll + (wv[j] - gam * (int64_t)M32[wst + j]);
            det = H >= (1ll << 28) || H <= -(1ll << 28);
        }
    for (int b = T - 1; b >= 0 && __syncthreads_or(!det); b--) {
        const int ws = recon_block_start(b, nwg, nrb), nv = recon_block_start(b + 1, nwg, nrb) - ws;
        recon_sweep(LS, pc, P, Mc, Kr, 4 * (int64_t)ws, Y, ystride, e0, 0, 0, nrb, acc, gsum);
        recon_words(acc, wv);
#pragma unroll
        for (int j = 31; j >= 0; j--)
            if (j < nv && !det) {
                H = H * 4294967296ll + (wv[j] - gam * (int64_t)M32[ws + j]);
                det = H >= (1ll << 28) || H <= -(1ll << 28);
            }
    }
    const bool neg = H < 0;   // H is C itself when no word fixed the sign
    __syncthreads();   // the last sweep's ring reads are done before the next sweep's DMA

    // limbs of |C|, block 0 up
    const int nlimb = nw / 2, lim = (int)min<int64_t>(cap, nlimb);
    const int64_t sgn = neg ? -1 : 1;
    uint64_t *const LST = (uint64_t *)LS;
    int64_t carry = 0, top = -1;
    bool hi_nz = false;
    uint64_t lb[16];
    // one block's limbs (limbs L0 .. L0 + nl - 1) through LDS to C
    auto emit = [&](int L0, int nl) {
        __syncthreads();   // every wave is done with the ring's last stage
#pragma unroll
        for (int j = 0; j < 16; j++) LST[tid * CARRY_PITCH + j] = lb[j];
        __syncthreads();
        // two limbs per thread and 16-byte store (entries are 8-byte aligned)
        typedef uint64_t v2u64 __attribute__((ext_vector_type(2), aligned(8)));
        const int n = max(0, min(nl, lim - L0));
        for (int f = tid; f < RECON_ENT * 8; f += RECON_ENT) {
            const int ent = f >> 3, j = 2 * (f & 7);
            const int64_t base = s_base[ent];
            const uint64_t *src = LST + ent * CARRY_PITCH + j;
            if (j + 1 < n && base >= 0) *(v2u64 *)(C + base + L0 + j) = (v2u64){src[0], src[1]};
            else if (j < n && base >= 0) C[base + L0 + j] = src[0];
        }
        __syncthreads();   // the staging is read before the next sweep refills the ring
    };
    // limbs of block (ws, nv) from its raw words, or from the top block's carried ones
#define LIMBS(wsrc, ws, nv)                                                                            \
    _Pragma("unroll") for (int j = 0; j < 16; j++) {                                                    \
        uint64_t limb = 0;                                                                              \
        if (2 * j < (nv)) {                                                                             \
            const int64_t x0 = sgn * ((int64_t)wsrc[2 * j] - gam * (int64_t)M32[(ws) + 2 * j]) + carry; \
            carry = x0 >> 32;                                                                           \
            const int64_t x1 = sgn * ((int64_t)wsrc[2 * j + 1] - gam * (int64_t)M32[(ws) + 2 * j + 1]) + carry; \
            carry = x1 >> 32;                                                                           \
            limb = (uint64_t)(uint32_t)x0 | ((uint64_t)(uint32_t)x1 << 32);                             \
            const int L = (ws) / 2 + j;                                                                 \
            if (L < lim) {                                                                              \
                if (limb) top = L;                                                                      \
            } else if (L < nlimb) {                                                                     \
                hi_nz |= limb != 0;                                                                     \
            }                                                                                           \
        }                                                                                               \
        lb[j] = limb;                                                                                   \
    }
    // a whole block below the capacity: the same limbs with no per-limb
    // conditions, the sign applied as (v ^ s) - s, gamma M32 as an unsigned
    // 32 x 32-bit product corrected for a negative gamma
    const int64_t smask = neg ? -1 : 0;
    const uint32_t gamu = (uint32_t)gam, gneg = gam < 0 ? ~0u : 0u;
    auto limbs_full = [&](int ws) {
#pragma unroll
        for (int j = 0; j < 16; j++) {
            int64_t x[2];
#pragma unroll
            for (int u = 0; u < 2; u++) {
                const uint32_t m = M32[ws + 2 * j + u];
                const uint64_t gm = (uint64_t)gamu * m - ((uint64_t)(gneg & m) << 32);   // gamma M32 mod 2^64
                const int64_t v = wv[2 * j + u] - (int64_t)gm;
                x[u] = ((v ^ smask) - smask) + carry;
                carry = x[u] >> 32;
            }
            const uint64_t limb = (uint64_t)(uint32_t)x[0] | ((uint64_t)(uint32_t)x[1] << 32);
            top = limb ? ws / 2 + j : top;
            lb[j] = limb;
        }
    };
    for (int b = 0; b < T; b++) {
        const int ws = recon_block_start(b, nwg, nrb), nv = recon_block_start(b + 1, nwg, nrb) - ws;
        recon_sweep(LS, pc, P, Mc, Kr, 4 * (int64_t)ws, Y, ystride, e0, 0, 0, nrb, acc, gsum);
        recon_words(acc, wv);
        if (nv == 32 && ws / 2 + 16 <= lim) limbs_full(ws);
        else LIMBS(wv, ws, nv);
        emit(ws / 2, nv / 2);
    }
    LIMBS(topw, wst, nvt);
    emit(wst / 2, nvt / 2);
#undef LIMBS
    if (valid) {
        if (hi_nz) atomicOr(err, DEV_ERR_CAPACITY);
        C[s_base[tid] + W - 1] = (uint64_t)(neg ? -(top + 1) : top + 1);
    }
}

// ------------------------------------------------------------------------
// host side shared by the one-product launches and the batch
// ------------------------------------------------------------------------
inline int64_t crt_ystride(int64_t nent) { return round_up(nent < 1 ? 1 : nent, 256); }

// Device pointers and sizes of a plan (mgx_crt_plan is private to crt.hip).
struct PlanView {
    const PrimeConst *pc;
    const int8_t *cst, *mc;
    const int8_t *cstb;   // the table in 1 KB blocks for the tiled reduce (two-digit plans, kb > RES_KB), else null
    const int8_t *mcg;    // mc with the gamma rows (gw = 2 words above nw) for the one-pass reconstruct, else null
    const uint32_t *m32;
    int64_t S, P, kb, Kr, rows4pad, Lbp, out_limbs, nw, gw;
    int64_t bitsA, bitsB;   // the operands' bit bounds (|A| < 2^bitsA, |B| < 2^bitsB)
};
int plan_view(const mgx_crt_plan *pl, PlanView *v);
}  // namespace mgx
