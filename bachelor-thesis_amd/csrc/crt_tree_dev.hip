// crt_tree_dev.hip -- the reference's split-tree API (mat-gpu/gpu_crt.h:10-12:
// gpu_crt_split_list, gpu_crt_merge_list) on the device, over many entries
// at once: every entry of a device array is reduced through one shared tree
// of coprime pairs (ms[l], ns[l]) made by mgx_crt_make_split_tree, and
// merged back.  In the reference these functions convert to GMP on the host
// and recurse there (mat-gpu/gpu_crt.cu:123-194, crt.cpp:126-164); here one
// thread takes one entry and walks the tree level by level with its own
// multi-precision arithmetic (32-bit words, Knuth's division), in a
// thread-private work area.  The tree is small and read by every thread.
//
// Layout: the leaves of entry e are res[e * 2^length + i], i < 2^length,
// in the reference's order: leaf i is the residue along the path whose bit
// for level l (most significant first) is 0 for ms[l] and 1 for ns[l]
// (crt.cpp:131-141).  Level l of the walk keeps node p (l bits) at slot
// p 2^(length - l), the first leaf of its subtree, so the leaves array is
// the work area of the split (the output) and of the merge (a copy of the
// input).
#include "common.h"
#include <vector>

using namespace mgx;

namespace {

// ---- magnitudes as little-endian 32-bit words; n = used words ----
__host__ __device__ int wnorm(const uint32_t *a, int n)
{
    while (n > 0 && a[n - 1] == 0) n--;
    return n;
}
__host__ __device__ int wcmp(const uint32_t *a, int an, const uint32_t *b, int bn)
{
    if (an != bn) return an < bn ? -1 : 1;
    for (int i = an - 1; i >= 0; i--)
        if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
    return 0;
}
// r = a + b (r may be a)
__host__ __device__ int wadd(uint32_t *r, const uint32_t *a, int an, const uint32_t *b, int bn)
{
    if (an < bn) { const uint32_t *t = a; a = b; b = t; int tn = an; an = bn; bn = tn; }
    uint64_t c = 0;
    for (int i = 0; i < an; i++) {
        c += (uint64_t)a[i] + (i < bn ? b[i] : 0u);
        r[i] = (uint32_t)c;
        c >>= 32;
    }
    r[an] = (uint32_t)c;
    return wnorm(r, an + 1);
}
// r = a - b for a >= b (r may be a)
__host__ __device__ int wsub(uint32_t *r, const uint32_t *a, int an, const uint32_t *b, int bn)
{
    int64_t c = 0;
    for (int i = 0; i < an; i++) {
        c += (int64_t)a[i] - (int64_t)(i < bn ? b[i] : 0u);
        r[i] = (uint32_t)c;
        c >>= 32;   // arithmetic: 0 or -1
    }
    return wnorm(r, an);
}
// r = a b (r distinct from a and b)
__host__ __device__ int wmul(uint32_t *r, const uint32_t *a, int an, const uint32_t *b, int bn)
{
    if (an == 0 || bn == 0) return 0;
    for (int i = 0; i < an + bn; i++) r[i] = 0;
    for (int i = 0; i < an; i++) {
        uint64_t c = 0;
        const uint64_t ai = a[i];
        for (int j = 0; j < bn; j++) {
            c += ai * b[j] + r[i + j];
            r[i + j] = (uint32_t)c;
            c >>= 32;
        }
        r[i + bn] = (uint32_t)c;
    }
    return wnorm(r, an + bn);
}
// r = a mod m (m > 0; r may be a); u: an + 1 words of work space, v: mn words
// (Knuth, TAOCP 4.3.1, algorithm D: normalise so the top word of m has its top
// bit set; then each quotient digit estimate from the top two words of the
// remainder is at most 2 too large and is corrected before the subtraction)
__host__ __device__ int wmod(uint32_t *r, const uint32_t *a, int an, const uint32_t *m, int mn, uint32_t *u, uint32_t *v)
{
    if (wcmp(a, an, m, mn) < 0) {
        for (int i = 0; i < an; i++) r[i] = a[i];
        return an;
    }
    if (mn == 1) {
        uint64_t rem = 0;
        for (int i = an - 1; i >= 0; i--) rem = ((rem << 32) | a[i]) % m[0];
        r[0] = (uint32_t)rem;
        return wnorm(r, 1);
    }
    const int s = __builtin_clz(m[mn - 1]);
    for (int i = mn - 1; i > 0; i--) v[i] = s ? (m[i] << s) | (m[i - 1] >> (32 - s)) : m[i];
    v[0] = m[0] << s;
    u[an] = s ? a[an - 1] >> (32 - s) : 0;
    for (int i = an - 1; i > 0; i--) u[i] = s ? (a[i] << s) | (a[i - 1] >> (32 - s)) : a[i];
    u[0] = a[0] << s;
    const uint64_t B = 1ull << 32, vt = v[mn - 1], vs = v[mn - 2];
    for (int j = an - mn; j >= 0; j--) {
        const uint64_t num = ((uint64_t)u[j + mn] << 32) | u[j + mn - 1];
        uint64_t qh = num / vt, rh = num % vt;
        while (qh >= B || qh * vs > ((rh << 32) | u[j + mn - 2])) {
            qh--;
            rh += vt;
            if (rh >= B) break;
        }
        // u[j .. j + mn] -= qh v
        int64_t borrow = 0;
        uint64_t carry = 0;
        for (int i = 0; i < mn; i++) {
            const uint64_t p = qh * v[i] + carry;
            carry = p >> 32;
            const int64_t t = (int64_t)u[i + j] - (int64_t)(uint32_t)p + borrow;
            u[i + j] = (uint32_t)t;
            borrow = t >> 32;
        }
        const int64_t t = (int64_t)u[j + mn] - (int64_t)carry + borrow;
        u[j + mn] = (uint32_t)t;
        if (t < 0) {   // qh was one too large: add v back
            uint64_t c = 0;
            for (int i = 0; i < mn; i++) {
                c += (uint64_t)u[i + j] + v[i];
                u[i + j] = (uint32_t)c;
                c >>= 32;
            }
            u[j + mn] += (uint32_t)c;
        }
    }
    for (int i = 0; i < mn; i++) r[i] = s ? (u[i] >> s) | (u[i + 1] << (32 - s)) : u[i];
    return wnorm(r, mn);
}

// ---- entries (cap 64-bit limbs + signed size) <-> words ----
__host__ __device__ int from_entry(uint32_t *w, const uint64_t *e, int64_t cap, bool *neg)
{
    const int64_t sz = (int64_t)e[cap];
    int64_t n = sz < 0 ? -sz : sz;
    if (n > cap) n = cap;
    if (neg) *neg = sz < 0;
    for (int64_t i = 0; i < n; i++) {
        w[2 * i] = (uint32_t)e[i];
        w[2 * i + 1] = (uint32_t)(e[i] >> 32);
    }
    return wnorm(w, (int)(2 * n));
}
// false when the value needs more than cap limbs
__host__ __device__ bool to_entry(uint64_t *e, const uint32_t *w, int n, int64_t cap, bool neg)
{
    const int64_t limbs = (n + 1) / 2;
    if (limbs > cap) return false;
    for (int64_t i = 0; i < cap; i++) {
        const uint32_t lo = 2 * i < n ? w[2 * i] : 0u, hi = 2 * i + 1 < n ? w[2 * i + 1] : 0u;
        e[i] = (uint64_t)lo | ((uint64_t)hi << 32);
    }
    e[cap] = (uint64_t)(neg && limbs ? -limbs : limbs);
    return true;
}

// Per-thread work areas in 32-bit words (c = cap, so an entry has at most 2c
// words): the split needs the node value, the residue, m, n and the division's
// two buffers; the merge the two children, m, n, mf, nf, the products (a nf:
// 4c + 1 words, times n: 6c + 2), their sum, m n (4c + 1) and the division.
__host__ __device__ inline int64_t split_words(int64_t c) { return 6 * (2 * c + 4); }
__host__ __device__ inline int64_t merge_words_per_thread(int64_t c) { return 6 * (2 * c + 4) + 5 * (6 * c + 8) + 3 * (4 * c + 4); }

// entry e through the tree; false when a value exceeds cap
__host__ __device__ bool split_one(int64_t e, uint64_t *res, const uint64_t *X, const uint64_t *ms,
    const uint64_t *ns, int length, int64_t cap, uint32_t *scratch)
{
    const int64_t W = cap + 1, nleaf = 1ll << length, e2 = 2 * cap + 4;
    uint32_t *Xv = scratch + e * split_words(cap), *Rv = Xv + e2, *Mv = Rv + e2, *Nv = Mv + e2, *U = Nv + e2,
             *V = U + e2;
    uint64_t *leaves = res + e * nleaf * W;
    bool neg;
    const int xn = from_entry(Xv, X + e * W, cap, &neg);
    bool ok = true;
    if (length == 0) ok = to_entry(leaves, Xv, xn, cap, neg);
    for (int l = 0; l < length && ok; l++) {
        const int64_t step = nleaf >> l, half = step >> 1;
        const int mn = from_entry(Mv, ms + l * W, cap, nullptr), nn = from_entry(Nv, ns + l * W, cap, nullptr);
        if (mn == 0 || nn == 0) { ok = false; break; }
        for (int64_t i = 0; i < nleaf && ok; i += step) {
            // the node: x at level 0, else the slot's (non-negative) residue
            const int an = l == 0 ? xn : from_entry(Xv, leaves + i * W, cap, nullptr);
            // the n child first (its slot is still free), then the m child in
            // place; mpz_mod semantics: for a negative x the residue is
            // d - (|x| mod d) when |x| mod d is not 0 (crt.cpp:135,139)
            for (int child = 1; child >= 0 && ok; child--) {
                const uint32_t *d = child ? Nv : Mv;
                const int dn = child ? nn : mn;
                int rn = wmod(Rv, Xv, an, d, dn, U, V);
                if (l == 0 && neg && rn > 0) rn = wsub(Rv, d, dn, Rv, rn);
                ok = to_entry(leaves + (i + (child ? half : 0)) * W, Rv, rn, cap, false);
            }
        }
    }
    return ok;
}
__global__ __launch_bounds__(256) void split_list_kernel(uint64_t *res, const uint64_t *X, int64_t count,
    const uint64_t *ms, const uint64_t *ns, int length, int64_t cap, uint32_t *scratch, unsigned *err)
{
    const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e < count && !split_one(e, res, X, ms, ns, length, cap, scratch)) atomicOr(err, DEV_ERR_CAPACITY);
}

// entry e's leaves (overwritten) merged into X[e]; false when a value exceeds cap
__host__ __device__ bool merge_one(int64_t e, uint64_t *X, uint64_t *leaves_all, const uint64_t *ms,
    const uint64_t *ns, const uint64_t *mfs, const uint64_t *nfs, int length, int64_t cap, uint32_t *scratch)
{
    const int64_t W = cap + 1, nleaf = 1ll << length, e2 = 2 * cap + 4, e6 = 6 * cap + 8, e4 = 4 * cap + 4;
    uint32_t *Av = scratch + e * merge_words_per_thread(cap), *Bv = Av + e2, *Mv = Bv + e2, *Nv = Mv + e2,
             *MF = Nv + e2, *NF = MF + e2;
    uint32_t *T = NF + e2, *T1 = T + e6, *T2 = T1 + e6, *S = T2 + e6, *U = S + e6;
    uint32_t *MNv = U + e6, *V = MNv + e4, *O = V + e4;
    uint64_t *leaves = leaves_all + e * nleaf * W;
    bool ok = true;
    // from the leaves up: level l merges slots i and i + half into slot i
    // (crt.cpp:147-164 with crt_merge, crt.cpp:67-80)
    for (int l = length - 1; l >= 0 && ok; l--) {
        const int64_t step = nleaf >> l, half = step >> 1;
        bool mfneg, nfneg;
        const int mn = from_entry(Mv, ms + l * W, cap, nullptr), nn = from_entry(Nv, ns + l * W, cap, nullptr);
        const int mfn = from_entry(MF, mfs + l * W, cap, &mfneg), nfn = from_entry(NF, nfs + l * W, cap, &nfneg);
        if (mn == 0 || nn == 0) { ok = false; break; }
        const int mnn = wmul(MNv, Mv, mn, Nv, nn);
        for (int64_t i = 0; i < nleaf && ok; i += step) {
            const int an = from_entry(Av, leaves + i * W, cap, nullptr);
            const int bn = from_entry(Bv, leaves + (i + half) * W, cap, nullptr);
            // t1 = a nf n (sign of nf), t2 = b mf m (sign of mf), s = t1 + t2
            int tn = wmul(T, Av, an, NF, nfn);
            const int t1n = wmul(T1, T, tn, Nv, nn);
            tn = wmul(T, Bv, bn, MF, mfn);
            const int t2n = wmul(T2, T, tn, Mv, mn);
            bool sneg;
            int sn;
            if (nfneg == mfneg) {
                sn = wadd(S, T1, t1n, T2, t2n);
                sneg = nfneg;
            } else if (wcmp(T1, t1n, T2, t2n) >= 0) {
                sn = wsub(S, T1, t1n, T2, t2n);
                sneg = nfneg;
            } else {
                sn = wsub(S, T2, t2n, T1, t1n);
                sneg = mfneg;
            }
            // x = s mod (m n), non-negative (mpz_mod)
            int rn = wmod(O, S, sn, MNv, mnn, U, V);
            if (sneg && rn > 0) rn = wsub(O, MNv, mnn, O, rn);
            ok = to_entry(leaves + i * W, O, rn, cap, false);
        }
    }
    if (ok) {
        const int xn = from_entry(Av, leaves, cap, nullptr);
        ok = to_entry(X + e * W, Av, xn, cap, false);
    }
    return ok;
}
__global__ __launch_bounds__(256) void merge_list_kernel(uint64_t *X, uint64_t *leaves_all, int64_t count,
    const uint64_t *ms, const uint64_t *ns, const uint64_t *mfs, const uint64_t *nfs, int length, int64_t cap,
    uint32_t *scratch, unsigned *err)
{
    const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e < count && !merge_one(e, X, leaves_all, ms, ns, mfs, nfs, length, cap, scratch))
        atomicOr(err, DEV_ERR_CAPACITY);
}

}  // namespace

extern "C" {

int mgx_crt_split_list_device(mgx_limb *res, const mgx_limb *X, int64_t count, const mgx_limb *ms,
                              const mgx_limb *ns, int length, int64_t cap, void *stream)
{
    MGX_TRY(ensure_init());
    if (!res || !X || (length > 0 && (!ms || !ns))) return fail(MGX_ERR_ARG, "null argument");
    if (count < 0 || cap < 1 || length < 0 || length > 20) return fail(MGX_ERR_SHAPE, "bad shape");
    if (count == 0) return MGX_OK;
    hipStream_t s = as_stream(stream);
    uint32_t *scr = nullptr;
    MGX_TRY(scratch_alloc((void **)&scr, (size_t)(count * split_words(cap) * 4), s));
    hipLaunchKernelGGL(split_list_kernel, dim3((unsigned)ceil_div(count, 256)), dim3(256), 0, s, (uint64_t *)res,
                       (const uint64_t *)X, count, (const uint64_t *)ms, (const uint64_t *)ns, length, cap, scr,
                       device_error_word());
    const hipError_t e = hipGetLastError();
    MGX_TRY(scratch_free(scr, s));
    if (e != hipSuccess) return fail(MGX_ERR_HIP, "split_list_kernel: %s", hipGetErrorString(e));
    return check_device_error(s);
}

int mgx_crt_merge_list_device(mgx_limb *X, const mgx_limb *split, int64_t count, const mgx_limb *ms,
                              const mgx_limb *ns, const mgx_limb *mfs, const mgx_limb *nfs, int length, int64_t cap,
                              void *stream)
{
    MGX_TRY(ensure_init());
    if (!X || !split || (length > 0 && (!ms || !ns || !mfs || !nfs))) return fail(MGX_ERR_ARG, "null argument");
    if (count < 0 || cap < 1 || length < 0 || length > 20) return fail(MGX_ERR_SHAPE, "bad shape");
    if (count == 0) return MGX_OK;
    hipStream_t s = as_stream(stream);
    // the leaves are copied to a work array that the merge overwrites level by level
    const size_t leaf_bytes = (size_t)(count << length) * (size_t)(cap + 1) * 8;
    char *scr = nullptr;
    MGX_TRY(scratch_alloc((void **)&scr, leaf_bytes + (size_t)(count * merge_words_per_thread(cap) * 4), s));
    MGX_HIP(hipMemcpyAsync(scr, split, leaf_bytes, hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(merge_list_kernel, dim3((unsigned)ceil_div(count, 256)), dim3(256), 0, s, (uint64_t *)X,
                       (uint64_t *)scr, count, (const uint64_t *)ms, (const uint64_t *)ns, (const uint64_t *)mfs,
                       (const uint64_t *)nfs, length, cap, (uint32_t *)(scr + leaf_bytes), device_error_word());
    const hipError_t e = hipGetLastError();
    MGX_TRY(scratch_free(scr, s));
    if (e != hipSuccess) return fail(MGX_ERR_HIP, "merge_list_kernel: %s", hipGetErrorString(e));
    return check_device_error(s);
}

}  // extern "C"
