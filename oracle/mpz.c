/*
 * mpz.c -- oracle (test infrastructure only): signed fixed-capacity entries.
 *
 * An entry is cap 64-bit limbs followed by one signed size word, the layout of
 * the reference's struct gpu___mpz_t (mat-gpu/gmp-es-mpz.h:22-27) in its
 * 64-bit-limb configuration (mat-gpu/gmp-es.h:27-31).  The reference asserts
 * when an operation needs more than cap limbs (mat-gpu/mpz-realloc.cu:5-11);
 * here that is status 1 and the destination is left unspecified.
 */
#include "oracle.h"
#include <string.h>
#include <stdlib.h>
#include <ctype.h>

#define SIZ(e, cap) (((int64_t *)(e))[cap])
static int64_t iabs64(int64_t x) { return x < 0 ? -x : x; }

static int64_t normalize(const olimb *p, int64_t n)
{
    while (n > 0 && p[n - 1] == 0) n--;
    return n;
}

/* mat-gpu/mpz-add.cu:48-123 (mpz_add; mpz_sub is the same with v negated,
 * mat-gpu/mpz-sub.cu:1-2).  The realloc to abs_usize+1 limbs is the
 * capacity condition (mpz-add.cu:73-74). */
static int aors(olimb *w, const olimb *u, const olimb *v, int64_t cap, int negv)
{
    int64_t usize = SIZ(u, cap), vsize = negv ? -SIZ(v, cap) : SIZ(v, cap);
    int64_t au = iabs64(usize), av = iabs64(vsize);
    if (au < av) {
        const olimb *t = u; u = v; v = t;
        int64_t s = usize; usize = vsize; vsize = s;
        s = au; au = av; av = s;
    }
    if (au + 1 > cap) return 1;
    olimb tmp[cap + 1];
    int64_t wsize;
    if ((usize ^ vsize) < 0) {
        if (au != av) {
            or_mpn_sub(tmp, u, au, v, av);
            wsize = normalize(tmp, au);
            if (usize < 0) wsize = -wsize;
        } else if (or_mpn_cmp(u, v, au) < 0) {
            or_mpn_sub_n(tmp, v, u, au);
            wsize = normalize(tmp, au);
            if (usize >= 0) wsize = -wsize;
        } else {
            or_mpn_sub_n(tmp, u, v, au);
            wsize = normalize(tmp, au);
            if (usize < 0) wsize = -wsize;
        }
    } else {
        olimb cy = or_mpn_add(tmp, u, au, v, av);
        tmp[au] = cy;
        wsize = au + (cy != 0);
        if (usize < 0) wsize = -wsize;
    }
    memcpy(w, tmp, (size_t)iabs64(wsize) * sizeof(olimb));
    SIZ(w, cap) = wsize;
    return 0;
}

int or_mpz_add(olimb *w, const olimb *u, const olimb *v, int64_t cap) { return aors(w, u, v, cap, 0); }
int or_mpz_sub(olimb *w, const olimb *u, const olimb *v, int64_t cap) { return aors(w, u, v, cap, 1); }

/* mat-gpu/mpz-mul.cu:50-110 (mpz_mul). */
int or_mpz_mul(olimb *w, const olimb *u, const olimb *v, int64_t cap)
{
    int64_t us = SIZ(u, cap), vs = SIZ(v, cap);
    int64_t sign = us ^ vs;
    int64_t au = iabs64(us), av = iabs64(vs);
    if (au < av) { const olimb *t = u; u = v; v = t; int64_t s = au; au = av; av = s; }
    if (av == 0) { SIZ(w, cap) = 0; return 0; }
    int64_t wsize = au + av;
    if (av == 1) {
        if (au + 1 > cap) return 1;
    } else if (wsize > cap) {
        return 1;
    }
    olimb tmp[wsize];
    or_mpn_mul(tmp, u, au, v, av);
    wsize -= tmp[wsize - 1] == 0;
    memcpy(w, tmp, (size_t)wsize * sizeof(olimb));
    SIZ(w, cap) = sign < 0 ? -wsize : wsize;
    return 0;
}

/* mat-gpu/mpz-aorsmul.cu:69-161 (aorsmul): w += sub? -x*y : x*y.
 * The realloc to max(wsize,tsize)+1 limbs is the capacity condition
 * (mpz-aorsmul.cu:98). */
static int aorsmul(olimb *w, const olimb *x, const olimb *y, int64_t cap, int64_t sub)
{
    int64_t xs = SIZ(x, cap), ys = SIZ(y, cap);
    if (xs == 0 || ys == 0) return 0;
    if (iabs64(ys) > iabs64(xs)) { const olimb *t = x; x = y; y = t; int64_t s = xs; xs = ys; ys = s; }
    sub ^= ys; ys = iabs64(ys);
    sub ^= xs; xs = iabs64(xs);
    int64_t wss = SIZ(w, cap);
    sub ^= wss;
    int64_t ws = iabs64(wss);
    int64_t ts = xs + ys;
    if ((ws > ts ? ws : ts) + 1 > cap) return 1;
    olimb tp[ts];
    or_mpn_mul(tp, x, xs, y, ys);
    ts -= tp[ts - 1] == 0;
    if (wss == 0) {
        memcpy(w, tp, (size_t)ts * sizeof(olimb));
        SIZ(w, cap) = sub >= 0 ? ts : -ts;
        return 0;
    }
    olimb res[cap + 1];
    int64_t rs;
    if (sub >= 0) {
        if (ws >= ts) { olimb c = or_mpn_add(res, w, ws, tp, ts); res[ws] = c; rs = ws + (c != 0); }
        else { olimb c = or_mpn_add(res, tp, ts, w, ws); res[ts] = c; rs = ts + (c != 0); }
    } else {
        /* mpn_cmp_twosizes_lt: |w| < |t| flips the sign (mpz-aorsmul.cu:141-152) */
        int lt = ws != ts ? ws < ts : or_mpn_cmp(w, tp, ws) < 0;
        if (lt) { or_mpn_sub(res, tp, ts, w, ws); rs = normalize(res, ts); wss = -wss; }
        else { or_mpn_sub(res, w, ws, tp, ts); rs = normalize(res, ws); }
    }
    memcpy(w, res, (size_t)rs * sizeof(olimb));
    SIZ(w, cap) = wss >= 0 ? rs : -rs;
    return 0;
}

/* mat-gpu/mpz-aorsmul.cu:268-290 */
int or_mpz_addmul(olimb *w, const olimb *u, const olimb *v, int64_t cap) { return aorsmul(w, u, v, cap, 0); }
int or_mpz_submul(olimb *w, const olimb *u, const olimb *v, int64_t cap) { return aorsmul(w, u, v, cap, -1); }

/* mat-gpu/mpz-neg.cu:35-57 (in-place use is the only one in the reference) */
void or_mpz_neg(olimb *w, const olimb *u, int64_t cap)
{
    int64_t s = SIZ(u, cap);
    if (w != u) memcpy(w, u, (size_t)iabs64(s) * sizeof(olimb));
    SIZ(w, cap) = -s;
}

int or_mpz_cmp(const olimb *u, const olimb *v, int64_t cap)
{
    int64_t us = SIZ(u, cap), vs = SIZ(v, cap);
    int64_t au = normalize(u, iabs64(us)), av = normalize(v, iabs64(vs));
    int64_t su = us < 0 ? -au : au, sv = vs < 0 ? -av : av;
    if (su != sv) return su < sv ? -1 : 1;
    int c = or_mpn_cmp(u, v, au);
    return su < 0 ? -c : c;
}

/* mat-gpu/mpz.cu:85-104 (mpz_out_str_16) and :182-204 (out_file): lower-case
 * hex, '-' prefix, "0" for zero.  Returns the string length, or -1 if buflen is
 * too small.  The text is NUL-terminated. */
int64_t or_mpz_out_str(char *buf, int64_t buflen, const olimb *e, int64_t cap)
{
    int64_t size = SIZ(e, cap);
    int64_t n = 0;
    if (size == 0) {
        if (buflen < 2) return -1;
        buf[0] = '0'; buf[1] = 0; return 1;
    }
    if (size < -1 || (size == -1 && e[0] != 0)) {
        if (n + 1 >= buflen) return -1;
        buf[n++] = '-';
        size = -size;
    }
    if (size < 0) size = -size;
    int print = 0;
    for (int64_t i = size - 1; i >= 0; i--) {
        for (int k = 15; k >= 0; k--) {
            char c = "0123456789abcdef"[(e[i] >> (4 * k)) & 0xf];
            if (c != '0') print = 1;
            if (print) { if (n + 1 >= buflen) return -1; buf[n++] = c; }
        }
    }
    if (!print) { if (n + 1 >= buflen) return -1; buf[n++] = '0'; }
    buf[n] = 0;
    return n;
}

/* mat-gpu/mpz.cu:106-146 (inp_buffer_hex, mpz_inp_str): optional leading
 * space, '-', "0x"; hex digits only.  Status: 0 ok, 1 capacity, 2 bad char. */
int or_mpz_inp_str(olimb *e, int64_t cap, const char *str)
{
    while (isspace((unsigned char)*str)) str++;
    int neg = 0;
    if (*str == '-') { neg = 1; str++; }
    if (str[0] == '0' && str[1] == 'x') str += 2;
    int64_t nch = 0;
    while (str[nch] && !isspace((unsigned char)str[nch])) nch++;
    if (nch > cap * 16) return 1;
    memset(e, 0, (size_t)cap * sizeof(olimb));
    for (int64_t i = 0; i < nch; i++) {
        int c = tolower((unsigned char)str[nch - 1 - i]);
        int d;
        if (c >= '0' && c <= '9') d = c - '0';
        else if (c >= 'a' && c <= 'f') d = c - 'a' + 10;
        else return 2;
        e[i / 16] |= (olimb)d << (4 * (i % 16));
    }
    /* inp_buffer_hex returns ceil(nch/16) limbs, unnormalized (mpz.cu:106-118) */
    int64_t sz = (nch + 15) / 16;
    SIZ(e, cap) = neg ? -sz : sz;
    return 0;
}
