/*
 * mpn.c -- oracle (test infrastructure only): natural-number limb arithmetic
 * on 64-bit limbs, restating the GMP 6.1.2 generic mpn code that the
 * reference carries in mat-gpu/.  Each function cites the reference
 * definition it follows.  unsigned __int128 plays the role of umul_ppmm
 * (mat-gpu/gmp-es.h:594-617).
 */
#include "oracle.h"
#include <string.h>
#include <stdlib.h>

typedef unsigned __int128 u128;

/* mat-gpu/add_n.cu:35-59 */
olimb or_mpn_add_n(olimb *rp, const olimb *up, const olimb *vp, int64_t n)
{
    olimb cy = 0;
    for (int64_t i = 0; i < n; i++) {
        olimb u = up[i], v = vp[i];
        olimb s = u + v;
        olimb c1 = s < u;
        olimb r = s + cy;
        olimb c2 = r < s;
        rp[i] = r;
        cy = c1 | c2;
    }
    return cy;
}

/* mat-gpu/sub_n.cu:35-59 */
olimb or_mpn_sub_n(olimb *rp, const olimb *up, const olimb *vp, int64_t n)
{
    olimb bw = 0;
    for (int64_t i = 0; i < n; i++) {
        olimb u = up[i], v = vp[i];
        olimb d = u - v;
        olimb b1 = u < v;
        olimb r = d - bw;
        olimb b2 = d < bw;
        rp[i] = r;
        bw = b1 | b2;
    }
    return bw;
}

/* mat-gpu/addsub.cu:19-58 (mpn_add_1 / MPN_INCR) */
olimb or_mpn_add_1(olimb *rp, const olimb *up, int64_t n, olimb v)
{
    olimb cy = v;
    for (int64_t i = 0; i < n; i++) {
        olimb r = up[i] + cy;
        cy = r < cy;
        rp[i] = r;
    }
    return cy;
}

/* mat-gpu/addsub.cu:60-100 (mpn_sub_1) */
olimb or_mpn_sub_1(olimb *rp, const olimb *up, int64_t n, olimb v)
{
    olimb bw = v;
    for (int64_t i = 0; i < n; i++) {
        olimb u = up[i];
        rp[i] = u - bw;
        bw = u < bw;
    }
    return bw;
}

/* mat-gpu/addsub.cu:186-230 (mpn_add: un >= vn) */
olimb or_mpn_add(olimb *rp, const olimb *up, int64_t un, const olimb *vp, int64_t vn)
{
    olimb cy = or_mpn_add_n(rp, up, vp, vn);
    return or_mpn_add_1(rp + vn, up + vn, un - vn, cy);
}

/* mat-gpu/addsub.cu:232-276 (mpn_sub: un >= vn) */
olimb or_mpn_sub(olimb *rp, const olimb *up, int64_t un, const olimb *vp, int64_t vn)
{
    olimb bw = or_mpn_sub_n(rp, up, vp, vn);
    return or_mpn_sub_1(rp + vn, up + vn, un - vn, bw);
}

/* mat-gpu/addsub.cu:278-300 (mpn_cmp) */
int or_mpn_cmp(const olimb *up, const olimb *vp, int64_t n)
{
    for (int64_t i = n - 1; i >= 0; i--) {
        if (up[i] != vp[i]) return up[i] > vp[i] ? 1 : -1;
    }
    return 0;
}

/* mat-gpu/mul_1.cu:36-59 */
olimb or_mpn_mul_1(olimb *rp, const olimb *up, int64_t n, olimb v)
{
    olimb cl = 0;
    for (int64_t i = 0; i < n; i++) {
        u128 p = (u128)up[i] * v + cl;
        rp[i] = (olimb)p;
        cl = (olimb)(p >> 64);
    }
    return cl;
}

/* mat-gpu/addmul_1.cu:38-63 */
olimb or_mpn_addmul_1(olimb *rp, const olimb *up, int64_t n, olimb v)
{
    olimb cl = 0;
    for (int64_t i = 0; i < n; i++) {
        /* u*v + r + cl < 2^128: (2^64-1)^2 + 2(2^64-1) = 2^128 - 1 */
        u128 p = (u128)up[i] * v + rp[i] + cl;
        rp[i] = (olimb)p;
        cl = (olimb)(p >> 64);
    }
    return cl;
}

/* mat-gpu/submul_1.cu:38-67 */
olimb or_mpn_submul_1(olimb *rp, const olimb *up, int64_t n, olimb v)
{
    olimb cl = 0;
    for (int64_t i = 0; i < n; i++) {
        u128 p = (u128)up[i] * v + cl;
        olimb lo = (olimb)p;
        cl = (olimb)(p >> 64);
        olimb r = rp[i];
        olimb d = r - lo;
        cl += d > r;
        rp[i] = d;
    }
    return cl;
}

/* mat-gpu/mul_basecase.cu:53-76: row 0 by mul_1, rows 1.. by addmul_1 */
void or_mpn_mul_basecase(olimb *rp, const olimb *up, int64_t un, const olimb *vp, int64_t vn)
{
    rp[un] = or_mpn_mul_1(rp, up, un, vp[0]);
    for (int64_t j = 1; j < vn; j++) rp[un + j] = or_mpn_addmul_1(rp + j, up, un, vp[j]);
}

/* Toom-22 cutoff in limbs.  The reference pins MUL_TOOM22_THRESHOLD to the
 * toom33 threshold (mat-gpu/gmp-es.h:393-395), which makes toom22 dead code;
 * here it is a runtime parameter so that the algorithm is reachable. */
static int64_t g_toom22_threshold = 24;
void or_set_toom22_threshold(int64_t limbs) { g_toom22_threshold = limbs < 4 ? 4 : limbs; }
int64_t or_get_toom22_threshold(void) { return g_toom22_threshold; }

/* mat-gpu/gmp-es.h:459-462 (mpn_toom22_mul_itch) */
int64_t or_mpn_toom22_itch(int64_t an) { return 2 * (an + 64); }

static void mul_rec(olimb *p, const olimb *a, int64_t an, const olimb *b, int64_t bn, olimb *ws);

static int zero_p(const olimb *p, int64_t n)
{
    for (int64_t i = 0; i < n; i++) if (p[i]) return 0;
    return 1;
}

/* Increment/decrement of a limb vector by a small value (MPN_INCR_U / MPN_DECR_U,
 * mat-gpu/gmp-es.h:246-352). */
static void incr_u(olimb *p, int64_t n, olimb v)
{
    for (int64_t i = 0; i < n && v; i++) { olimb r = p[i] + v; v = r < v; p[i] = r; }
}
static void decr_u(olimb *p, int64_t n, olimb v)
{
    for (int64_t i = 0; i < n && v; i++) { olimb u = p[i]; p[i] = u - v; v = u < v; }
}

/* mat-gpu/toom22_mul.cu:108-235: Karatsuba with evaluation points 0, -1, inf.
 * a has an limbs, b has bn limbs, an >= bn, split at n = ceil(an/2). */
void or_mpn_toom22_mul(olimb *pp, const olimb *ap, int64_t an, const olimb *bp, int64_t bn, olimb *scratch)
{
    int64_t s = an >> 1, n = an - s, t = bn - n;
    const olimb *a0 = ap, *a1 = ap + n, *b0 = bp, *b1 = bp + n;
    olimb *asm1 = pp, *bsm1 = pp + n;
    int vm1_neg = 0;

    /* |a0 - a1| (toom22_mul.cu:140-164) */
    if (s == n) {
        if (or_mpn_cmp(a0, a1, n) < 0) { or_mpn_sub_n(asm1, a1, a0, n); vm1_neg = 1; }
        else or_mpn_sub_n(asm1, a0, a1, n);
    } else {
        if (a0[s] == 0 && or_mpn_cmp(a0, a1, s) < 0) {
            or_mpn_sub_n(asm1, a1, a0, s); asm1[s] = 0; vm1_neg = 1;
        } else {
            asm1[s] = a0[s] - or_mpn_sub_n(asm1, a0, a1, s);
        }
    }
    /* |b0 - b1| (toom22_mul.cu:167-191) */
    if (t == n) {
        if (or_mpn_cmp(b0, b1, n) < 0) { or_mpn_sub_n(bsm1, b1, b0, n); vm1_neg ^= 1; }
        else or_mpn_sub_n(bsm1, b0, b1, n);
    } else {
        if (zero_p(b0 + t, n - t) && or_mpn_cmp(b0, b1, t) < 0) {
            or_mpn_sub_n(bsm1, b1, b0, t);
            memset(bsm1 + t, 0, (size_t)(n - t) * sizeof(olimb));
            vm1_neg ^= 1;
        } else {
            or_mpn_sub(bsm1, b0, n, b1, t);
        }
    }

    olimb *v0 = pp, *vinf = pp + 2 * n, *vm1 = scratch, *ws = scratch + 2 * n;
    mul_rec(vm1, asm1, n, bsm1, n, ws);                 /* vm1: 2n limbs */
    if (s > t) mul_rec(vinf, a1, s, b1, t, ws);         /* vinf: s+t limbs */
    else mul_rec(vinf, a1, s, b1, s, ws);
    mul_rec(v0, ap, n, bp, n, ws);                      /* v0: 2n limbs */

    /* interpolation (toom22_mul.cu:208-229) */
    olimb cy = or_mpn_add_n(pp + 2 * n, v0 + n, vinf, n);
    olimb cy2 = cy + or_mpn_add_n(pp + n, pp + 2 * n, v0, n);
    cy += or_mpn_add(pp + 2 * n, pp + 2 * n, n, vinf + n, s + t - n);
    if (vm1_neg) cy += or_mpn_add_n(pp + n, pp + n, vm1, 2 * n);
    else cy -= or_mpn_sub_n(pp + n, pp + n, vm1, 2 * n);
    incr_u(pp + 2 * n, s + t, cy2);
    if (cy <= 2) incr_u(pp + 3 * n, s + t - n, cy);
    else decr_u(pp + 3 * n, s + t - n, 1);
}

/* TOOM22_MUL_REC (mat-gpu/toom22_mul.cu:60-106): basecase below the cutoff,
 * toom22 when balanced enough, otherwise split the longer operand. */
static void mul_rec(olimb *p, const olimb *a, int64_t an, const olimb *b, int64_t bn, olimb *ws)
{
    if (an < bn) { const olimb *t = a; a = b; b = t; int64_t tn = an; an = bn; bn = tn; }
    if (bn < g_toom22_threshold) { or_mpn_mul_basecase(p, a, an, b, bn); return; }
    if (4 * an < 5 * bn) { or_mpn_toom22_mul(p, a, an, b, bn, ws); return; }
    /* unbalanced: chunk a into bn-limb pieces (mat-gpu/mul.cu:93-149 chunking) */
    olimb *tmp = ws;
    olimb *ws2 = ws + 2 * bn;
    mul_rec(p, a, bn, b, bn, ws2);
    int64_t done = bn;
    while (done < an) {
        int64_t c = an - done < bn ? an - done : bn;
        mul_rec(tmp, a + done, c, b, bn, ws2);
        /* p[done .. done+bn) already holds the high part; add tmp (c+bn limbs) */
        olimb cy = or_mpn_add_n(p + done, p + done, tmp, bn);
        memcpy(p + done + bn, tmp + bn, (size_t)c * sizeof(olimb));
        incr_u(p + done + bn, c, cy);
        done += c;
    }
}

/* mat-gpu/mul.cu:60-149 (mpn_mul: un >= vn >= 1); returns the top limb. */
olimb or_mpn_mul(olimb *rp, const olimb *up, int64_t un, const olimb *vp, int64_t vn)
{
    if (un < vn) { const olimb *t = up; up = vp; vp = t; int64_t tn = un; un = vn; vn = tn; }
    if (vn < g_toom22_threshold) {
        or_mpn_mul_basecase(rp, up, un, vp, vn);
    } else {
        int64_t itch = 8 * (un + vn) + 512;
        olimb stackbuf[4096];
        olimb *ws = itch <= 4096 ? stackbuf : (olimb *)malloc((size_t)itch * sizeof(olimb));
        mul_rec(rp, up, un, vp, vn, ws);
        if (ws != stackbuf) free(ws);
    }
    return rp[un + vn - 1];
}
