// bigint_host.h -- small host-side natural-number helpers on 32-bit words
// (little endian), used to build per-prime constants and by the kept
// gpu_crt_* API.  Values are std::vector<uint32_t> without leading zeros.
#pragma once
#include <cstdint>
#include <vector>
#include <algorithm>

namespace mgx {
namespace hb {

using Nat = std::vector<uint32_t>;

inline void trim(Nat &a) { while (!a.empty() && a.back() == 0) a.pop_back(); }
inline Nat from_u64(uint64_t v)
{
    Nat r;
    while (v) { r.push_back((uint32_t)v); v >>= 32; }
    return r;
}
inline int cmp(const Nat &a, const Nat &b)
{
    if (a.size() != b.size()) return a.size() < b.size() ? -1 : 1;
    for (size_t i = a.size(); i-- > 0;)
        if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
    return 0;
}
inline Nat add(const Nat &a, const Nat &b)
{
    Nat r(std::max(a.size(), b.size()) + 1, 0);
    uint64_t c = 0;
    for (size_t i = 0; i < r.size(); i++) {
        uint64_t s = c + (i < a.size() ? a[i] : 0) + (i < b.size() ? b[i] : 0);
        r[i] = (uint32_t)s;
        c = s >> 32;
    }
    trim(r);
    return r;
}
// a - b, requires a >= b
inline Nat sub(const Nat &a, const Nat &b)
{
    Nat r(a.size(), 0);
    int64_t br = 0;
    for (size_t i = 0; i < a.size(); i++) {
        int64_t d = (int64_t)a[i] - (i < b.size() ? b[i] : 0) - br;
        br = d < 0;
        r[i] = (uint32_t)(d + (br << 32));
    }
    trim(r);
    return r;
}
inline Nat mul_small(const Nat &a, uint32_t m)
{
    Nat r(a.size() + 1, 0);
    uint64_t c = 0;
    for (size_t i = 0; i < a.size(); i++) {
        uint64_t t = (uint64_t)a[i] * m + c;
        r[i] = (uint32_t)t;
        c = t >> 32;
    }
    r[a.size()] = (uint32_t)c;
    trim(r);
    return r;
}
inline Nat mul(const Nat &a, const Nat &b)
{
    if (a.empty() || b.empty()) return Nat();
    Nat r(a.size() + b.size(), 0);
    for (size_t i = 0; i < a.size(); i++) {
        uint64_t c = 0;
        for (size_t j = 0; j < b.size(); j++) {
            uint64_t t = (uint64_t)a[i] * b[j] + r[i + j] + c;
            r[i + j] = (uint32_t)t;
            c = t >> 32;
        }
        r[i + b.size()] = (uint32_t)c;
    }
    trim(r);
    return r;
}
inline Nat divmod_small(const Nat &a, uint32_t d, uint32_t *rem)
{
    Nat q(a.size(), 0);
    uint64_t r = 0;
    for (size_t i = a.size(); i-- > 0;) {
        uint64_t cur = (r << 32) | a[i];
        q[i] = (uint32_t)(cur / d);
        r = cur % d;
    }
    if (rem) *rem = (uint32_t)r;
    trim(q);
    return q;
}
inline uint32_t mod_small(const Nat &a, uint32_t d)
{
    uint64_t r = 0;
    for (size_t i = a.size(); i-- > 0;) r = ((r << 32) | a[i]) % d;
    return (uint32_t)r;
}
inline int64_t bitlen(const Nat &a)
{
    if (a.empty()) return 0;
    return (int64_t)(a.size() - 1) * 32 + (32 - __builtin_clz(a.back()));
}
inline Nat shl(const Nat &a, int64_t bits)
{
    if (a.empty()) return Nat();
    int64_t w = bits / 32, b = bits % 32;
    Nat r(a.size() + w + 1, 0);
    for (size_t i = 0; i < a.size(); i++) {
        uint64_t v = (uint64_t)a[i] << b;
        r[i + w] |= (uint32_t)v;
        r[i + w + 1] |= (uint32_t)(v >> 32);
    }
    trim(r);
    return r;
}
inline Nat shr(const Nat &a, int64_t bits)
{
    int64_t w = bits / 32, b = bits % 32;
    if ((int64_t)a.size() <= w) return Nat();
    Nat r(a.size() - w, 0);
    for (size_t i = 0; i < r.size(); i++) {
        uint64_t v = a[i + w];
        if (i + w + 1 < a.size()) v |= (uint64_t)a[i + w + 1] << 32;
        r[i] = (uint32_t)(v >> b);
    }
    trim(r);
    return r;
}
// Knuth algorithm D (TAOCP vol. 2, 4.3.1): q = a / b, r = a % b, b != 0.
inline void divmod(const Nat &a, const Nat &b, Nat &q, Nat &r)
{
    if (cmp(a, b) < 0) { q.clear(); r = a; return; }
    if (b.size() == 1) { uint32_t rr; q = divmod_small(a, b[0], &rr); r = from_u64(rr); return; }
    int s = __builtin_clz(b.back());
    Nat v = shl(b, s), u = shl(a, s);
    size_t n = v.size();
    if (u.size() == a.size()) u.push_back(0);
    if (u.size() < a.size() + 1) u.resize(a.size() + 1, 0);
    size_t m = u.size() - n;
    q.assign(m, 0);
    for (size_t j = m; j-- > 0;) {
        uint64_t num = ((uint64_t)u[j + n] << 32) | u[j + n - 1];
        uint64_t qh = num / v[n - 1], rh = num % v[n - 1];
        while (qh >= (1ull << 32) || qh * v[n - 2] > ((rh << 32) | u[j + n - 2])) {
            qh--; rh += v[n - 1];
            if (rh >= (1ull << 32)) break;
        }
        int64_t br = 0;
        uint64_t c = 0;
        for (size_t i = 0; i < n; i++) {
            uint64_t p = qh * v[i] + c;
            c = p >> 32;
            int64_t t = (int64_t)u[i + j] - (int64_t)(uint32_t)p - br;
            br = t < 0;
            u[i + j] = (uint32_t)(t + (br << 32));
        }
        int64_t t = (int64_t)u[j + n] - (int64_t)c - br;
        br = t < 0;
        u[j + n] = (uint32_t)(t + (br << 32));
        if (br) {
            qh--;
            uint64_t cc = 0;
            for (size_t i = 0; i < n; i++) {
                uint64_t s2 = (uint64_t)u[i + j] + v[i] + cc;
                u[i + j] = (uint32_t)s2;
                cc = s2 >> 32;
            }
            u[j + n] += (uint32_t)cc;
        }
        q[j] = (uint32_t)qh;
    }
    trim(q);
    u.resize(n);
    trim(u);
    r = shr(u, s);
}

}  // namespace hb
}  // namespace mgx
