// crt_tree.cpp -- the reference's big-modulus CRT API (mat-gpu/gpu_crt.h:6-14,
// mat-gpu/crt.cpp:16-164) on the library's entries, with this library's own
// host arithmetic instead of GMP.  Host functions; no device work.
//
// x is split into a tree of coprime pairs (m, n) with m ~ sqrt(x), n > m and
// m n > x; mf m + nf n = 1 comes from the extended Euclidean algorithm with
// GMP's normalisation of the cofactors (|mf| < n/2, |nf| < m/2), so the
// results are the ones mpz_gcdext gives (crt.cpp:44).
#include "common.h"
#include "bigint_host.h"
#include <cstdlib>
#include <cstring>
#include <vector>

using namespace mgx;
using hb::Nat;

namespace {

struct Int {
    bool neg = false;
    Nat mag;
};

Int norm(Int a)
{
    hb::trim(a.mag);
    if (a.mag.empty()) a.neg = false;
    return a;
}
Int from_nat(const Nat &n) { return norm(Int{false, n}); }
Int neg(Int a) { a.neg = !a.neg; return norm(a); }
Int add(const Int &a, const Int &b)
{
    if (a.neg == b.neg) return norm(Int{a.neg, hb::add(a.mag, b.mag)});
    int c = hb::cmp(a.mag, b.mag);
    if (c == 0) return Int{};
    if (c > 0) return norm(Int{a.neg, hb::sub(a.mag, b.mag)});
    return norm(Int{b.neg, hb::sub(b.mag, a.mag)});
}
Int sub(const Int &a, const Int &b) { return add(a, neg(b)); }
Int mul(const Int &a, const Int &b) { return norm(Int{a.neg != b.neg, hb::mul(a.mag, b.mag)}); }
// floor division with non-negative remainder for positive d (mpz_mod semantics)
Nat mod_pos(const Int &a, const Nat &d)
{
    Nat q, r;
    hb::divmod(a.mag, d, q, r);
    if (a.neg && !r.empty()) r = hb::sub(d, r);
    return r;
}
Nat isqrt(const Nat &x)
{
    if (x.empty()) return Nat();
    Nat y = hb::shl(hb::from_u64(1), (hb::bitlen(x) + 1) / 2);
    for (;;) {
        Nat q, r;
        hb::divmod(x, y, q, r);
        Nat z = hb::shr(hb::add(q, y), 1);
        if (hb::cmp(z, y) >= 0) return y;
        y = z;
    }
}
Nat gcd(Nat a, Nat b)
{
    while (!b.empty()) {
        Nat q, r;
        hb::divmod(a, b, q, r);
        a = b;
        b = r;
    }
    return a;
}
bool is_one(const Nat &a) { return a.size() == 1 && a[0] == 1; }

// g = s a + t b for a, b > 0, normalised as mpz_gcdext does
void gcdext(const Nat &a, const Nat &b, Nat &g, Int &s, Int &t)
{
    Int r0 = from_nat(a), r1 = from_nat(b), s0 = from_nat(hb::from_u64(1)), s1;
    while (!r1.mag.empty()) {
        Nat q, r;
        hb::divmod(r0.mag, r1.mag, q, r);
        Int nr = from_nat(r), ns = sub(s0, mul(from_nat(q), s1));
        r0 = r1; r1 = nr; s0 = s1; s1 = ns;
    }
    g = r0.mag;
    // s is determined mod b/g; pick the representative with |s| < b/(2g)
    Nat bg, rr;
    hb::divmod(b, g, bg, rr);
    Nat sm = mod_pos(s0, bg);
    Int sv = from_nat(sm);
    if (hb::cmp(hb::add(sm, sm), bg) > 0) sv = sub(sv, from_nat(bg));
    s = sv;
    // t = (g - s a) / b
    Int num = sub(from_nat(g), mul(s, from_nat(a)));
    Nat q, r;
    hb::divmod(num.mag, b, q, r);
    t = norm(Int{num.neg, q});
}

Int from_entry(const mgx_limb *e, int64_t cap)
{
    int64_t s = ((const int64_t *)e)[cap];
    int64_t n = s < 0 ? -s : s;
    Int r;
    r.neg = s < 0;
    for (int64_t i = 0; i < n; i++) {
        r.mag.push_back((uint32_t)e[i]);
        r.mag.push_back((uint32_t)(e[i] >> 32));
    }
    return norm(r);
}
int to_entry(mgx_limb *e, int64_t cap, const Int &v)
{
    int64_t limbs = ((int64_t)v.mag.size() + 1) / 2;
    if (limbs > cap) return fail(MGX_ERR_CAPACITY, "CRT value needs %lld limbs > cap %lld", (long long)limbs, (long long)cap);
    for (int64_t i = 0; i < cap; i++) {
        uint64_t lo = 2 * i < (int64_t)v.mag.size() ? v.mag[2 * i] : 0;
        uint64_t hi = 2 * i + 1 < (int64_t)v.mag.size() ? v.mag[2 * i + 1] : 0;
        e[i] = lo | (hi << 32);
    }
    ((int64_t *)e)[cap] = v.neg ? -limbs : limbs;
    return MGX_OK;
}

// mat-gpu/crt.cpp:16-63
void split(const Nat &x, const std::vector<Nat> &list, Nat &m, Nat &n, Int &mf, Int &nf)
{
    m = isqrt(x);
    const Nat one = hb::from_u64(1);
    for (;;) {
        bool ok = true;
        for (const Nat &l : list)
            if (!is_one(gcd(m, l))) { ok = false; break; }
        if (ok) break;
        m = hb::add(m, one);
    }
    n = hb::add(m, hb::from_u64(3));
    const bool m_even = m.empty() || !(m[0] & 1);
    for (;;) {
        Nat g;
        gcdext(m, n, g, mf, nf);
        if (is_one(g)) {
            bool ok = true;
            for (const Nat &l : list)
                if (!is_one(gcd(n, l))) { ok = false; break; }
            if (ok) break;
        }
        n = hb::add(n, hb::from_u64(m_even ? 2 : 1));
    }
}

// mat-gpu/crt.cpp:66-78: x = (a nf n + b mf m) mod (m n)
Int merge(const Int &a, const Int &b, const Nat &m, const Nat &n, const Int &mf, const Int &nf)
{
    Int x = add(mul(mul(a, nf), from_nat(n)), mul(mul(b, mf), from_nat(m)));
    return from_nat(mod_pos(x, hb::mul(m, n)));
}

void split_list(std::vector<Int> &res, size_t at, const Int &x, const std::vector<Nat> &ms, const std::vector<Nat> &ns,
                size_t lvl, int length)
{
    if (length <= 0) { res[at] = x; return; }
    split_list(res, at, from_nat(mod_pos(x, ms[lvl])), ms, ns, lvl + 1, length - 1);
    split_list(res, at + ((size_t)1 << (length - 1)), from_nat(mod_pos(x, ns[lvl])), ms, ns, lvl + 1, length - 1);
}

Int merge_list(const std::vector<Int> &split, size_t at, const std::vector<Nat> &ms, const std::vector<Nat> &ns,
               const std::vector<Int> &mfs, const std::vector<Int> &nfs, size_t lvl, int length)
{
    if (length <= 0) return split[at];
    Int a = merge_list(split, at, ms, ns, mfs, nfs, lvl + 1, length - 1);
    Int b = merge_list(split, at + ((size_t)1 << (length - 1)), ms, ns, mfs, nfs, lvl + 1, length - 1);
    return merge(a, b, ms[lvl], ns[lvl], mfs[lvl], nfs[lvl]);
}

}  // namespace

extern "C" {

int mgx_crt_split(mgx_limb *m, mgx_limb *n, mgx_limb *mf, mgx_limb *nf, const mgx_limb *x, const mgx_limb *const *list,
                  int64_t listlen, int64_t cap)
{
    if (!m || !n || !mf || !nf || !x || (listlen && !list)) return fail(MGX_ERR_ARG, "null argument");
    Int xv = from_entry(x, cap);
    if (xv.neg || xv.mag.empty()) return fail(MGX_ERR_ARG, "crt_split needs x >= 1");
    std::vector<Nat> l;
    for (int64_t i = 0; i < listlen; i++) l.push_back(from_entry(list[i], cap).mag);
    Nat mm, nn;
    Int mfv, nfv;
    split(xv.mag, l, mm, nn, mfv, nfv);
    MGX_TRY(to_entry(m, cap, from_nat(mm)));
    MGX_TRY(to_entry(n, cap, from_nat(nn)));
    MGX_TRY(to_entry(mf, cap, mfv));
    return to_entry(nf, cap, nfv);
}

int mgx_crt_merge(mgx_limb *x, const mgx_limb *a, const mgx_limb *b, const mgx_limb *m, const mgx_limb *n,
                  const mgx_limb *mf, const mgx_limb *nf, int64_t cap)
{
    if (!x || !a || !b || !m || !n || !mf || !nf) return fail(MGX_ERR_ARG, "null argument");
    Int r = merge(from_entry(a, cap), from_entry(b, cap), from_entry(m, cap).mag, from_entry(n, cap).mag,
                  from_entry(mf, cap), from_entry(nf, cap));
    return to_entry(x, cap, r);
}

// mat-gpu/crt.cpp:91-120 and gpu_crt.cu:90-120: returns the length (>= 0) or -status
int mgx_crt_make_split_tree(mgx_limb **ms, mgx_limb **ns, mgx_limb **mfs, mgx_limb **nfs, const mgx_limb *x,
                            uint64_t limit, int64_t cap)
{
    if (!ms || !ns || !mfs || !nfs || !x || limit < 2) return -fail(MGX_ERR_ARG, "bad argument");
    Int xv = from_entry(x, cap);
    if (xv.neg) return -fail(MGX_ERR_ARG, "crt tree needs x >= 0");
    std::vector<Nat> vm, vn, lst;
    std::vector<Int> vmf, vnf;
    Nat xr = xv.mag;
    const Nat lim = hb::from_u64(limit);
    while (hb::cmp(xr, lim) >= 0) {
        Nat m, n;
        Int mf, nf;
        split(xr, lst, m, n, mf, nf);
        lst.push_back(m);
        lst.push_back(n);
        vm.push_back(m); vn.push_back(n); vmf.push_back(mf); vnf.push_back(nf);
        xr = mod_pos(from_nat(xr), n);
    }
    const size_t len = vm.size(), W = (size_t)cap + 1;
    mgx_limb *a[4];
    for (int i = 0; i < 4; i++) a[i] = (mgx_limb *)calloc(len ? len * W : 1, sizeof(mgx_limb));
    for (size_t i = 0; i < len; i++) {
        int st;
        if ((st = to_entry(a[0] + i * W, cap, from_nat(vm[i]))) || (st = to_entry(a[1] + i * W, cap, from_nat(vn[i]))) ||
            (st = to_entry(a[2] + i * W, cap, vmf[i])) || (st = to_entry(a[3] + i * W, cap, vnf[i]))) {
            for (int k = 0; k < 4; k++) free(a[k]);
            return -st;
        }
    }
    *ms = a[0]; *ns = a[1]; *mfs = a[2]; *nfs = a[3];
    return (int)len;
}

// mat-gpu/crt.cpp:126-142: res holds 2^length entries
int mgx_crt_split_list(mgx_limb *res, const mgx_limb *x, const mgx_limb *ms, const mgx_limb *ns, int length, int64_t cap)
{
    if (!res || !x || length < 0 || (length && (!ms || !ns))) return fail(MGX_ERR_ARG, "bad argument");
    const size_t W = (size_t)cap + 1;
    std::vector<Nat> vm, vn;
    for (int i = 0; i < length; i++) {
        vm.push_back(from_entry(ms + i * W, cap).mag);
        vn.push_back(from_entry(ns + i * W, cap).mag);
    }
    std::vector<Int> out((size_t)1 << length);
    split_list(out, 0, from_entry(x, cap), vm, vn, 0, length);
    for (size_t i = 0; i < out.size(); i++) MGX_TRY(to_entry(res + i * W, cap, out[i]));
    return MGX_OK;
}

// mat-gpu/crt.cpp:145-164
int mgx_crt_merge_list(mgx_limb *x, const mgx_limb *split_, const mgx_limb *ms, const mgx_limb *ns, const mgx_limb *mfs,
                       const mgx_limb *nfs, int length, int64_t cap)
{
    if (!x || !split_ || length < 0 || (length && (!ms || !ns || !mfs || !nfs))) return fail(MGX_ERR_ARG, "bad argument");
    const size_t W = (size_t)cap + 1;
    std::vector<Nat> vm, vn;
    std::vector<Int> vmf, vnf, sp((size_t)1 << length);
    for (int i = 0; i < length; i++) {
        vm.push_back(from_entry(ms + i * W, cap).mag);
        vn.push_back(from_entry(ns + i * W, cap).mag);
        vmf.push_back(from_entry(mfs + i * W, cap));
        vnf.push_back(from_entry(nfs + i * W, cap));
    }
    for (size_t i = 0; i < sp.size(); i++) sp[i] = from_entry(split_ + i * W, cap);
    return to_entry(x, cap, merge_list(sp, 0, vm, vn, vmf, vnf, 0, length));
}

}  // extern "C"
