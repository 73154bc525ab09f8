// plain_kernels.h -- device side of the plain route shared by the
// one-product kernels and the batched persistent kernels.  See plain.hip.
#pragma once
#include "common.h"

namespace mgx {

// ---------------------------------------------------------------- helpers
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint32_t rot1(uint32_t v)   // lane l <- lane l-1
{
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x13C, 0xF, 0xF, false);
}
__device__ __forceinline__ uint32_t rdlane(uint32_t v, int i)
{
    return (uint32_t)__builtin_amdgcn_readlane((int)v, i);
}
__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

struct Acc96 {
    uint64_t lo;
    uint32_t hi;
};
__device__ __forceinline__ void add96(Acc96 &a, uint64_t p)
{
    unsigned long long c;
    a.lo = __builtin_addcll(a.lo, p, 0ull, &c);
    a.hi += (uint32_t)c;
}
__device__ __forceinline__ void add96(Acc96 &a, const Acc96 &b)
{
    unsigned long long c;
    a.lo = __builtin_addcll(a.lo, b.lo, 0ull, &c);
    a.hi += b.hi + (uint32_t)c;
}
__device__ __forceinline__ uint32_t shfl_up1(uint32_t v, uint32_t from_prev_group_lane63)
{
    uint32_t r = (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x138, 0xF, 0xF, true);   // wave_shr:1
    return lane_id() == 0 ? from_prev_group_lane63 : r;
}

// entry word addressing (32-bit words of the limb array)
template <bool STRIDED>
struct Ent {
    const uint32_t *base;   // word 0 of the entry (AoS) or of its interleave block
    int64_t j;              // position within the interleave block (STRIDED)
    int64_t W;              // words of 64 bits per entry (cap + 1)
    __device__ __forceinline__ uint32_t word(int64_t w) const
    {
        if (STRIDED) return base[2 * ((w >> 1) * MGX_INTERLEAVE_BLOCK + j) + (w & 1)];
        return base[w];
    }
    __device__ __forceinline__ int64_t size() const
    {
        const int64_t *p = (const int64_t *)base;
        if (STRIDED) return p[(W - 1) * MGX_INTERLEAVE_BLOCK + j];
        return p[W - 1];
    }
};
template <bool STRIDED>
__device__ __forceinline__ Ent<STRIDED> ent(const uint64_t *M, int64_t rs, int64_t y, int64_t x, int64_t W)
{
    int64_t e = y * rs + x;
    Ent<STRIDED> r;
    r.W = W;
    if (STRIDED) {
        r.base = (const uint32_t *)(M + (e / MGX_INTERLEAVE_BLOCK) * MGX_INTERLEAVE_BLOCK * W);
        r.j = e % MGX_INTERLEAVE_BLOCK;
    } else {
        r.base = (const uint32_t *)(M + e * W);
        r.j = 0;
    }
    return r;
}
template <bool STRIDED>
__device__ __forceinline__ void put_word(uint64_t *M, int64_t rs, int64_t y, int64_t x, int64_t W, int64_t w, uint32_t v)
{
    int64_t e = y * rs + x;
    if (STRIDED) {
        uint32_t *b = (uint32_t *)(M + (e / MGX_INTERLEAVE_BLOCK) * MGX_INTERLEAVE_BLOCK * W);
        b[2 * ((w >> 1) * MGX_INTERLEAVE_BLOCK + e % MGX_INTERLEAVE_BLOCK) + (w & 1)] = v;
    } else {
        ((uint32_t *)(M + e * W))[w] = v;
    }
}
template <bool STRIDED>
__device__ __forceinline__ void put_size(uint64_t *M, int64_t rs, int64_t y, int64_t x, int64_t W, int64_t s)
{
    int64_t e = y * rs + x;
    if (STRIDED) M[(e / MGX_INTERLEAVE_BLOCK) * MGX_INTERLEAVE_BLOCK * W + (W - 1) * MGX_INTERLEAVE_BLOCK + e % MGX_INTERLEAVE_BLOCK] = (uint64_t)s;
    else M[e * W + W - 1] = (uint64_t)s;
}

// Carry-lookahead over one 64-lane group: lane value v (33-bit: d + x, where
// x in {0,1}); returns the final 32-bit digit and updates the group carry.
__device__ __forceinline__ uint32_t resolve_add(uint32_t d, uint32_t x, uint32_t &gcarry)
{
    uint64_t v = (uint64_t)d + x;
    uint64_t G = ballot(v >> 32), P = ballot((uint32_t)v == 0xffffffffu && !(v >> 32));
    // carries into each lane = carry bits of the integer sum (G|P) + G + cin
    uint64_t a = G | P;
    unsigned __int128 t = (unsigned __int128)a + G + gcarry;
    uint64_t cin = (uint64_t)t ^ a ^ G;
    gcarry = (uint32_t)(t >> 64);
    return (uint32_t)v + (uint32_t)((cin >> lane_id()) & 1);
}
// Borrow-lookahead for big - small over one group.
__device__ __forceinline__ uint32_t resolve_sub(uint32_t big, uint32_t small, uint32_t &gborrow)
{
    uint64_t G = ballot(big < small), P = ballot(big == small);
    uint64_t a = G | P;
    unsigned __int128 t = (unsigned __int128)a + G + gborrow;
    uint64_t bin = (uint64_t)t ^ a ^ G;
    gborrow = (t >> 64) != 0;
    return big - small - (uint32_t)((bin >> lane_id()) & 1);
}

// Column sums (96-bit, lane l of group g = column 64g+l) to normalized
// 32-bit digits in place (d[g] per lane).
template <int G>
__device__ __forceinline__ void columns_to_digits(const Acc96 (&acc)[G], uint32_t (&d)[G])
{
    uint32_t c1prev = 0, midlast = 0, hilast = 0, hilast2 = 0;
    uint32_t c1[G];
    // word_c = lo(col c) + mid(col c-1) + hi(col c-2)
#pragma unroll
    for (int g = 0; g < G; g++) {
        uint32_t lo = (uint32_t)acc[g].lo, mid = (uint32_t)(acc[g].lo >> 32), hi = acc[g].hi;
        uint32_t m1 = shfl_up1(mid, midlast);
        uint32_t h1 = shfl_up1(hi, hilast);
        uint32_t h2 = shfl_up1(h1, hilast2);
        uint64_t s = (uint64_t)lo + m1 + h2;
        d[g] = (uint32_t)s;
        c1[g] = (uint32_t)(s >> 32);
        midlast = rdlane(mid, 63);
        hilast2 = rdlane(h1, 63);
        hilast = rdlane(hi, 63);
    }
    // add the small carries (0..2) one column up, then resolve the 0/1 carries
    uint32_t gcarry = 0;
#pragma unroll
    for (int g = 0; g < G; g++) {
        uint32_t cprev = shfl_up1(c1[g], c1prev);
        c1prev = rdlane(c1[g], 63);
        uint64_t t = (uint64_t)d[g] + cprev;
        uint32_t dd = (uint32_t)t, c2 = (uint32_t)(t >> 32);
        d[g] = dd;
        c1[g] = c2;
    }
    uint32_t c2prev = 0;
#pragma unroll
    for (int g = 0; g < G; g++) {
        uint32_t x = shfl_up1(c1[g], c2prev);
        c2prev = rdlane(c1[g], 63);
        d[g] = resolve_add(d[g], x, gcarry);
    }
}

// --------------------------------------------------------- mul / addmul
// Block product a (this lane's word a_l) x b: column l (lo) and l+64 (hi).
// Only the first cnt words of a can be non-zero (cnt wave-uniform, 1..64),
// so the rotation stops after cnt steps.
struct ColPair {
    Acc96 lo, hi;
};
// Every step's product goes into one running sum per lane (no branch on
// the lane); column l's sum is complete after step l, so lane l's copy of the
// running sum then is its lo, and hi is the rest.  Lanes l >= cnt take no
// copy: every step belongs to their lo.
// lane l <- lane l + j of its row of 16 (j = 1 .. 7; 0 past the row's end)
__device__ __forceinline__ uint32_t row_shl(uint32_t v, int j)
{
    switch (j) {
    case 1: return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x101, 0xF, 0xF, true);
    case 2: return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x102, 0xF, 0xF, true);
    case 3: return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x103, 0xF, 0xF, true);
    case 4: return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x104, 0xF, 0xF, true);
    case 5: return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x105, 0xF, 0xF, true);
    case 6: return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x106, 0xF, 0xF, true);
    default: return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x107, 0xF, 0xF, true);
    }
}
// Whole blocks: the copy is taken once per group of 8 steps, after the
// group's last step s, by the group's 8 lanes; lane l's copy then also holds
// the steps l < i <= s, i.e. a_{l+j} b_{64-j} for j = 1 .. 7 - (l mod 8),
// which are subtracted afterwards (a_{l+j} by a DPP row shift, b_{64-j} by
// v_readlane).
// b enters as its two 16-bit halves, each rotated: a product of a 32-bit word
// and a 16-bit half is below 2^48, so 64 of them sum below 2^54 and every step
// is a multiply-add into a 64-bit sum with no carry to detect.
__device__ __forceinline__ Acc96 join16(uint64_t lo16sum, uint64_t hi16sum)
{
    Acc96 r;
    unsigned long long c;
    r.lo = __builtin_addcll(lo16sum, hi16sum << 16, 0ull, &c);
    r.hi = (uint32_t)(hi16sum >> 48) + (uint32_t)c;
    return r;
}
__device__ __forceinline__ void block_steps64_grouped(uint32_t a, uint32_t b, uint32_t l, Acc96 &acc, Acc96 &snap)
{
    uint32_t bl = b & 0xffffu, bh = b >> 16;
    uint64_t sl = 0, sh = 0, cl = 0, ch = 0;
#pragma unroll
    for (int i = 0; i < 64; i++) {
        const uint32_t ai = rdlane(a, i);
        sl = (uint64_t)ai * bl + sl;
        sh = (uint64_t)ai * bh + sh;
        if ((i & 7) == 7 && (l >> 3) == (uint32_t)(i >> 3)) { cl = sl; ch = sh; }
        bl = rot1(bl);
        bh = rot1(bh);
    }
    acc = join16(sl, sh);
    snap = join16(cl, ch);
    Acc96 e = {0, 0};
#pragma unroll
    for (int j = 1; j < 8; j++) {
        // row_shl:j: lane l reads lane l + j of its row of 16 (groups of 8 lie inside a row)
        uint32_t aj = row_shl(a, j);
        aj = (int)(l & 7) + j <= 7 ? aj : 0u;
        add96(e, (uint64_t)aj * rdlane(b, 64 - j));
    }
    unsigned long long bw;
    snap.lo = __builtin_subcll(snap.lo, e.lo, 0ull, &bw);
    snap.hi = snap.hi - e.hi - (uint32_t)bw;
}
// The same for a block of cnt < 64 words of a: steps stop at cnt, the last
// group's copy is taken after step cnt - 1, and lanes l >= cnt take the
// whole sum (every step is below their column).
__device__ __forceinline__ void block_steps_grouped(uint32_t a, uint32_t b, uint32_t l, int cnt, Acc96 &acc,
                                                    Acc96 &snap)
{
    uint32_t bl = b & 0xffffu, bh = b >> 16;
    uint64_t sl = 0, sh = 0, cl = 0, ch = 0;
#pragma unroll
    for (int g = 0; g < 8; g++) {
        if (8 * g >= cnt) break;
#pragma unroll
        for (int u = 0; u < 8; u++) {
            if (8 * g + u >= cnt) break;
            const uint32_t ai = rdlane(a, 8 * g + u);
            sl = (uint64_t)ai * bl + sl;
            sh = (uint64_t)ai * bh + sh;
            bl = rot1(bl);
            bh = rot1(bh);
        }
        if ((l >> 3) == (uint32_t)g) { cl = sl; ch = sh; }
    }
    acc = join16(sl, sh);
    snap = join16(cl, ch);
    if (l >= (uint32_t)cnt) { snap = acc; return; }
    // lane l's copy was taken after step min(8 (l / 8) + 7, cnt - 1)
    const int r = min((int)(l | 7), cnt - 1) - (int)l;
    Acc96 e = {0, 0};
#pragma unroll
    for (int j = 1; j < 8; j++) {
        uint32_t aj = row_shl(a, j);
        aj = j <= r ? aj : 0u;
        add96(e, (uint64_t)aj * rdlane(b, 64 - j));
    }
    unsigned long long bw;
    snap.lo = __builtin_subcll(snap.lo, e.lo, 0ull, &bw);
    snap.hi = snap.hi - e.hi - (uint32_t)bw;
}
// A few words of a (cnt < 16): the copy at every step, no corrections.
__device__ __forceinline__ void block_steps_short(uint32_t a, uint32_t b, uint32_t l, int cnt, Acc96 &acc,
                                                  Acc96 &snap)
{
    uint32_t bl = b & 0xffffu, bh = b >> 16;
    uint64_t sl = 0, sh = 0, cl = 0, ch = 0;
#pragma unroll
    for (int i = 0; i < 15; i++) {
        if (i >= cnt) break;
        const uint32_t ai = rdlane(a, i);
        sl = (uint64_t)ai * bl + sl;
        sh = (uint64_t)ai * bh + sh;
        if (l == (uint32_t)i) { cl = sl; ch = sh; }
        bl = rot1(bl);
        bh = rot1(bh);
    }
    acc = join16(sl, sh);
    snap = join16(cl, ch);
}
__device__ __noinline__ ColPair block_product_fn(uint32_t a, uint32_t b, int cnt)
{
    const uint32_t l = lane_id();
    Acc96 acc = {0, 0}, snap = {0, 0};
    if (cnt == 64) block_steps64_grouped(a, b, l, acc, snap);
    else if (cnt < 16) block_steps_short(a, b, l, cnt, acc, snap);
    else block_steps_grouped(a, b, l, cnt, acc, snap);
    if (l >= (uint32_t)cnt) snap = acc;
    ColPair r;
    r.lo = snap;
    unsigned long long bw;
    r.hi.lo = __builtin_subcll(acc.lo, snap.lo, 0ull, &bw);
    r.hi.hi = acc.hi - snap.hi - (uint32_t)bw;
    return r;
}
// The product is symmetric, so the block with fewer valid words is the one
// iterated: a 65-word Winograd sum times a 64-word entry costs 64 + 1 steps.
__device__ __forceinline__ void block_product(uint32_t a, uint32_t b, int cnta, int cntb, Acc96 &lo, Acc96 &hi)
{
    ColPair r = cnta <= cntb ? block_product_fn(a, b, cnta) : block_product_fn(b, a, cntb);
    add96(lo, r.lo);
    add96(hi, r.hi);
}
__device__ __forceinline__ int block_words(int64_t n, int blk)
{
    const int64_t c = n - 64 * (int64_t)blk;
    return c <= 0 ? 0 : c >= 64 ? 64 : (int)c;
}

template <int G>
__device__ __forceinline__ void acc_add(Acc96 (&acc)[G], int g, const Acc96 &v)
{
#pragma unroll
    for (int h = 0; h < G; h++)
        if (h == g) add96(acc[h], v);
}

// register-array element with a run-time (wave-uniform) index
template <int NB>
__device__ __forceinline__ uint32_t sel(const uint32_t (&x)[NB], int i)
{
    uint32_t r = 0;
#pragma unroll
    for (int g = 0; g < NB; g++)
        if (g == i) r = x[g];
    return r;
}

// one-level Karatsuba on the block halves needs |x0 - x1| per operand;
// returns true when x0 < x1
template <int NB>
__device__ __forceinline__ bool abs_diff_halves(const uint32_t (&x)[NB], int h, uint32_t (&out)[NB])
{
    int sign = 0;
#pragma unroll
    for (int g = NB - 1; g >= 0; g--) {
        if (g >= h || sign) continue;
        uint32_t xh = sel<NB>(x, g + h);
        uint64_t ne = ballot(x[g] != xh);
        if (ne) {
            int top = 63 - __clzll(ne);
            uint32_t u = rdlane(x[g], top), v = rdlane(xh, top);
            sign = u > v ? 1 : -1;
        }
    }
    uint32_t br = 0;
#pragma unroll
    for (int g = 0; g < NB; g++) {
        if (g >= h) { out[g] = 0; continue; }
        uint32_t xh = sel<NB>(x, g + h);
        uint32_t u = sign >= 0 ? x[g] : xh, v = sign >= 0 ? xh : x[g];
        out[g] = resolve_sub(u, v, br);
    }
    return sign < 0;
}

// One output entry per wavefront; blk is the group of four entries of this
// workgroup (a block index, or a work-queue item in the batched executor).
template <int NB, bool STRIDED, bool ADDMUL>
__device__ __forceinline__ void plain_entries(int64_t blk, int64_t K, int64_t M, int64_t N, uint64_t *__restrict__ C,
                                              int64_t Crs, const uint64_t *__restrict__ A, int64_t Ars,
                                              const uint64_t *__restrict__ B, int64_t Brs, int64_t W, int64_t cap,
                                              int toom, unsigned *__restrict__ err)
{
    constexpr int G = 2 * NB + 1;
    const int64_t idx = blk * 4 + (threadIdx.x >> 6);
    if (idx >= K * N) return;
    const int64_t y = idx / N, x = idx % N;
    const uint32_t l = lane_id();

    Acc96 accP[G], accN[G];
#pragma unroll
    for (int g = 0; g < G; g++) { accP[g] = {0, 0}; accN[g] = {0, 0}; }

    if (ADDMUL) {
        Ent<STRIDED> c = ent<STRIDED>(C, Crs, y, x, W);
        int64_t sc = c.size(), nw = 2 * (sc < 0 ? -sc : sc);
#pragma unroll
        for (int g = 0; g < G; g++) {
            int64_t w = 64 * g + l;
            uint32_t v = w < nw ? c.word(w) : 0;
            if (sc < 0) accN[g].lo = v;
            else accP[g].lo = v;
        }
    }

    for (int64_t k = 0; k < M; k++) {
        Ent<STRIDED> ea = ent<STRIDED>(A, Ars, y, k, W), eb = ent<STRIDED>(B, Brs, k, x, W);
        int64_t sa = ea.size(), sb = eb.size();
        if (sa == 0 || sb == 0) continue;
        int64_t na = 2 * (sa < 0 ? -sa : sa), nb = 2 * (sb < 0 ? -sb : sb);
        int nba = (int)((na + 63) >> 6), nbb = (int)((nb + 63) >> 6);
        if (nba > NB || nbb > NB) { if (l == 0) atomicOr(err, DEV_ERR_WIDTH); continue; }
        uint32_t a[NB], b[NB];
#pragma unroll
        for (int g = 0; g < NB; g++) {
            int64_t w = 64 * g + l;
            a[g] = w < na ? ea.word(w) : 0;
            b[g] = w < nb ? eb.word(w) : 0;
        }
        const bool neg = (sa < 0) != (sb < 0);
        const int nbm = nba > nbb ? nba : nbb;
        // Karatsuba pays only when both top halves are nearly full blocks
        if (toom && nbm >= 2 && (nbm & 1) == 0 && nbm >= toom && nba == nbb && na >= 64 * nbm - 16 &&
            nb >= 64 * nbm - 16) {
            // Karatsuba: a b = a0 b0 (1 + X) + a1 b1 (X + X^2) - (a0 - a1)(b0 - b1) X, X = 2^(32*64*h)
            const int h = nbm / 2;
            uint32_t da[NB], db[NB];
            bool sda = abs_diff_halves<NB>(a, h, da), sdb = abs_diff_halves<NB>(b, h, db);
            const bool mneg = neg != (sda != sdb);   // sign of -(a0-a1)(b0-b1) relative to +
            // block indices are compile-time constants (guarded), so every
            // register array is addressed statically
#pragma unroll
            for (int ib = 0; ib < NB / 2; ib++)
#pragma unroll
                for (int jb = 0; jb < NB / 2; jb++) {
                    if (ib >= h || jb >= h) continue;
                    const uint32_t a0 = a[ib], b0 = b[jb], am = da[ib], bm = db[jb];
                    const uint32_t a1 = sel<NB>(a, ib + h), b1 = sel<NB>(b, jb + h);
                    Acc96 z0l = {0, 0}, z0h = {0, 0}, z2l = {0, 0}, z2h = {0, 0}, zml = {0, 0}, zmh = {0, 0};
                    block_product(a0, b0, 64, 64, z0l, z0h);
                    block_product(a1, b1, block_words(na, ib + h), block_words(nb, jb + h), z2l, z2h);
                    block_product(am, bm, 64, 64, zml, zmh);
                    const int g0 = ib + jb;
                    Acc96 sl = z0l, sh = z0h;   // z0 + z2 lands at offset h
                    add96(sl, z2l);
                    add96(sh, z2h);
#define PUT(toN, g, v)                           \
    do {                                         \
        if (toN) acc_add<G>(accN, (g), (v));     \
        else acc_add<G>(accP, (g), (v));         \
    } while (0)
                    PUT(neg, g0, z0l); PUT(neg, g0 + 1, z0h);
                    PUT(neg, g0 + h, sl); PUT(neg, g0 + h + 1, sh);
                    PUT(neg, g0 + 2 * h, z2l); PUT(neg, g0 + 2 * h + 1, z2h);
                    // -(a0-a1)(b0-b1) X with its sign
                    PUT(!mneg, g0 + h, zml); PUT(!mneg, g0 + h + 1, zmh);
#undef PUT
                }
        } else {
#pragma unroll
            for (int ib = 0; ib < NB; ib++)
#pragma unroll
                for (int jb = 0; jb < NB; jb++) {
                    if (ib >= nba || jb >= nbb) continue;
                    Acc96 zl = {0, 0}, zh = {0, 0};
                    block_product(a[ib], b[jb], block_words(na, ib), block_words(nb, jb), zl, zh);
                    if (neg) { add96(accN[ib + jb], zl); add96(accN[ib + jb + 1], zh); }
                    else { add96(accP[ib + jb], zl); add96(accP[ib + jb + 1], zh); }
                }
        }
    }

    // normalize both magnitudes, subtract the smaller from the larger
    uint32_t dP[G], dN[G];
    columns_to_digits<G>(accP, dP);
    columns_to_digits<G>(accN, dN);
    int cmp = 0;
#pragma unroll
    for (int g = G - 1; g >= 0; g--) {
        if (cmp) continue;
        uint64_t ne = ballot(dP[g] != dN[g]);
        if (ne) {
            int top = 63 - __clzll(ne);
            cmp = rdlane(dP[g], top) > rdlane(dN[g], top) ? 1 : -1;
        }
    }
    uint32_t r[G];
    uint32_t bw = 0;
    int topw = -1;
#pragma unroll
    for (int g = 0; g < G; g++) {
        r[g] = cmp >= 0 ? resolve_sub(dP[g], dN[g], bw) : resolve_sub(dN[g], dP[g], bw);
        uint64_t nz = ballot(r[g] != 0);
        if (nz) topw = 64 * g + 63 - __clzll(nz);
    }
    const int64_t limbs = (topw + 2) / 2;
    if (limbs > cap) {
        if (l == 0) atomicOr(err, DEV_ERR_CAPACITY);
        return;
    }
#pragma unroll
    for (int g = 0; g < G; g++) {
        int64_t w = 64 * g + l;
        if (w < 2 * limbs) put_word<STRIDED>(C, Crs, y, x, W, w, r[g]);
    }
    if (l == 0) put_size<STRIDED>(C, Crs, y, x, W, cmp >= 0 ? limbs : -limbs);
}

// Entry idx (row-major over an H x Wd block) of C = A + (sub ? -B : B), one
// wavefront per entry, mpz_add semantics (mat-gpu/mpz-add.cu:48-123).
template <bool STRIDED>
__device__ __forceinline__ void aors_entry(int64_t idx, int64_t Wd, int64_t H, uint64_t *__restrict__ C, int64_t Crs,
                                           const uint64_t *__restrict__ A, int64_t Ars,
                                           const uint64_t *__restrict__ B, int64_t Brs, int64_t W, int64_t cap,
                                           int sub, unsigned *__restrict__ err)
{
    if (idx >= Wd * H) return;
    const int64_t y = idx / Wd, x = idx % Wd;
    const uint32_t l = lane_id();
    Ent<STRIDED> ea = ent<STRIDED>(A, Ars, y, x, W), eb = ent<STRIDED>(B, Brs, y, x, W);
    int64_t sa = ea.size(), sb = eb.size();
    if (sub) sb = -sb;
    const int64_t na = 2 * (sa < 0 ? -sa : sa), nb = 2 * (sb < 0 ? -sb : sb);
    const int ng = (int)((std::max(na, nb) + 64) >> 6);   // one extra word for the carry
    const bool same = (sa >= 0) == (sb >= 0) || sa == 0 || sb == 0;
    // pass 1 (different signs): compare magnitudes from the top group down
    int cmp = 0;
    if (!same)
        for (int g = ng - 1; g >= 0 && !cmp; g--) {
            int64_t w = 64 * g + l;
            uint32_t u = w < na ? ea.word(w) : 0, v = w < nb ? eb.word(w) : 0;
            uint64_t ne = ballot(u != v);
            if (ne) {
                int top = 63 - __clzll(ne);
                cmp = rdlane(u, top) > rdlane(v, top) ? 1 : -1;
            }
        }
    // pass 2: low to high; group g is written after it is read, and C may
    // alias A or B only entry-for-entry, so later groups are still intact
    int topw = -1;
    uint32_t cy = 0, lastc = 0;
    for (int g = 0; g < ng; g++) {
        int64_t w = 64 * g + l;
        uint32_t u = w < na ? ea.word(w) : 0, v = w < nb ? eb.word(w) : 0;
        uint32_t r;
        if (same) {
            uint64_t s = (uint64_t)u + v;
            uint32_t c = (uint32_t)(s >> 32);
            uint32_t xin = shfl_up1(c, lastc);   // the raw carry of lane l enters lane l+1
            lastc = rdlane(c, 63);
            r = resolve_add((uint32_t)s, xin, cy);
        } else {
            r = cmp >= 0 ? resolve_sub(u, v, cy) : resolve_sub(v, u, cy);
        }
        uint64_t nz = ballot(r != 0);
        if (nz) topw = 64 * g + 63 - __clzll(nz);
        if (w < 2 * cap) put_word<STRIDED>(C, Crs, y, x, W, w, r);
    }
    const int64_t limbs = (topw + 2) / 2;
    if (limbs > cap) { if (l == 0) atomicOr(err, DEV_ERR_CAPACITY); return; }
    bool negr = same ? (sa < 0 || (sa == 0 && sb < 0)) : (cmp >= 0 ? sa < 0 : sb < 0);
    if (l == 0) put_size<STRIDED>(C, Crs, y, x, W, negr ? -limbs : limbs);
}

}  // namespace mgx
