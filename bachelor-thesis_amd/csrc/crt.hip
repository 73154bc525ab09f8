// crt.hip -- the multi-modular route on gfx950 INT8 MFMA.
//
// C = A*B for signed big-integer matrices, exactly, through P primes with
// M = prod p_q > 4 max|C|: S = 2 digits per residue with primes below 65280
// (the default while enough of them exist), else S = 4 with primes below 2^31.
// Two digits need 4 slice products per prime, four need 16, and the 16-bit
// primes are only twice as many, so S = 2 halves the residue GEMM:
//
//   1. reduce:      every entry of A and B mod p_q, centred to |r| <= (p-1)/2
//                   and split into S balanced int8 digits (planar slices).
//                   Itself an int8 GEMM: entry bytes (offset by -128) times the
//                   balanced digits of 2^(8b) mod p_q.
//   2. residue GEMM: per prime, the S^2 slice products A^(s) B^(t) on INT8
//                   MFMA, accumulated per diagonal s+t in int32 over K-chunks
//                   (8192 for S = 4, 32768 for S = 2), folded mod p_q in
//                   64-bit integers (Barrett), then y_q = c_q (M/p_q)^-1 mod p_q
//                   is written as S balanced int8 digits.
//   3. reconstruct: explicit CRT, C = sum_q y_q (M/p_q) - beta M, again an int8
//                   GEMM (the digits of (M/p_q) 256^s against the digits of
//                   y_q) whose epilogue leaves one signed 32-bit-word column
//                   sum per word; beta comes from a 32-bit fixed-point sum of
//                   y_q/p_q.  A carry pass per entry then finds the sign and
//                   writes sign and magnitude.
//
// No floating point anywhere.  The reference has no counterpart (its
// crt.cpp splits one modulus into a tree and is never called); the maths is
// the Chinese remainder theorem, Barrett reduction and balanced radix-256
// digits.
#include "crt_kernels.h"
#include "bigint_host.h"
#include <vector>
#include <cstring>
#include <algorithm>
#include <array>
#include <atomic>
#include <mutex>

namespace mgx {

// ------------------------------------------------------------------------
// bit-length scan: out[0] = max over entries of bit length of |x|; one
// atomic per workgroup.  Only the size word and the top limb of an entry are
// read; each thread has the loads of four entries in flight at once (the scan
// is latency-bound: two dependent 8-byte loads per 536-byte entry).
// ------------------------------------------------------------------------
// Up to two matrices in one launch (workgroups [0, nblk0) scan the first).
struct ScanArgs {
    const uint64_t *X[2];
    int64_t rs[2], rows[2], cols[2];
    int nblk0;
};

// EXACT: bit length of the top non-zero limb (two cache lines per entry:
// size word and top limb); otherwise the limb-granular bound 64 |size| from
// the size words alone (one cache line per entry), which is what plans need.
template <bool EXACT>
__global__ __launch_bounds__(256) void scan_bits_kernel(ScanArgs sa, int64_t W, int *out_pair)
{
    __shared__ int s_best[4];
    const int m = (int)blockIdx.x < sa.nblk0 ? 0 : 1;
    const int64_t blk = m ? blockIdx.x - sa.nblk0 : blockIdx.x, nblk = m ? gridDim.x - sa.nblk0 : sa.nblk0;
    const uint64_t *X = sa.X[m];
    const int64_t rs = sa.rs[m], cols = sa.cols[m];
    int *out = out_pair + m;
    const int64_t n = sa.rows[m] * cols, stride = nblk * blockDim.x;
    int best = 0;
    for (int64_t e0 = blk * (int64_t)blockDim.x + threadIdx.x; e0 < n; e0 += 4 * stride) {
        const uint64_t *p[4];
        int64_t a[4];
        uint64_t top[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            int64_t e = e0 + j * stride;
            p[j] = X + ((e / cols) * rs + e % cols) * W;
            int64_t s = e < n ? (int64_t)p[j][W - 1] : 0;
            a[j] = min(s < 0 ? -s : s, W - 1);
        }
        if (EXACT) {
#pragma unroll
            for (int j = 0; j < 4; j++) top[j] = a[j] > 0 ? p[j][a[j] - 1] : 0;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                while (a[j] > 0 && top[j] == 0) {   // zero top limbs: not normalized
                    a[j]--;
                    top[j] = a[j] > 0 ? p[j][a[j] - 1] : 0;
                }
                if (a[j] > 0) best = max(best, (int)((a[j] - 1) * 64 + 64 - __clzll(top[j])));
            }
        } else {
#pragma unroll
            for (int j = 0; j < 4; j++) best = max(best, (int)(64 * a[j]));
        }
    }
    for (int o = 32; o > 0; o >>= 1) best = max(best, __shfl_xor(best, o));
    if ((threadIdx.x & 63) == 0) s_best[threadIdx.x >> 6] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
        best = max(max(s_best[0], s_best[1]), max(s_best[2], s_best[3]));
        if (best > 0) atomicMax(out, best);
    }
}

// ------------------------------------------------------------------------
// one-product launches: one workgroup per tile
// ------------------------------------------------------------------------
template <int S>
__global__ __launch_bounds__(256) void crt_reduce_kernel(
    const PrimeConst *__restrict__ pc, const int8_t *__restrict__ cst, int P, int nrb, int kb, int xbits,
    const uint64_t *__restrict__ X, int64_t Xrs, int64_t W, int transpose, int64_t rows, int64_t K,
    int64_t kpad, int64_t nent, int8_t *__restrict__ R, int64_t plane, unsigned *__restrict__ err)
{
    reduce_tile<S>(blockIdx.x, pc, cst, P, nrb, kb, xbits, X, Xrs, W, transpose, rows, K, kpad, nent, R, plane, err);
}

// Up to two operands (A and B of one product share the plan, K and kpad) in
// one launch: workgroups [0, nblk0) reduce side 0, the rest side 1.
template <int S, int KS>
__global__ __launch_bounds__(256) void crt_reduce_res_kernel(
    const PrimeConst *__restrict__ pc, const int8_t *__restrict__ cst, int P, int nrb, int kb, int64_t W, int64_t K,
    int64_t kpad, ResSide s0, ResSide s1, int64_t nblk0, unsigned *__restrict__ err)
{
    const bool second = (int64_t)blockIdx.x >= nblk0;
    const ResSide &sd = second ? s1 : s0;
    reduce_tile_res<S, KS>(second ? blockIdx.x - nblk0 : blockIdx.x, pc, cst, P, nrb, kb, sd.bits, sd.X, sd.Xrs, W,
                           sd.transpose, sd.rows, K, kpad, sd.R, sd.plane, err);
}
// The reduce of wide entries as a tiled MFMA GEMM (two-digit plans), tiles
// in XCD-aware order: the tiles an XCD runs together share entry and table
// tiles in its L2.
__global__ __launch_bounds__(256, 2) void crt_reduce_gemm_kernel(
    const PrimeConst *__restrict__ pc, const int8_t *__restrict__ cstb, int P, int kb, int xbits,
    const uint64_t *__restrict__ X, int64_t Xrs, int64_t W, int transpose, int64_t rows, int64_t K, int64_t kpad,
    int64_t nent, int8_t *__restrict__ R, int64_t plane, int64_t total, unsigned *__restrict__ err)
{
    const int64_t per = (total + 7) / 8;
    const int64_t v = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
    if (v >= total) return;
    reduce_gemm_tile(v, pc, cstb, P, kb, xbits, X, Xrs, W, transpose, rows, K, kpad, nent, R, plane, err);
}
// The constant-resident reduce: one workgroup per CU, units [0, nu0) of 256
// entries (4 x 64 patches) from side 0, the rest from side 1.
template <int KS>
__global__ __launch_bounds__(CRES_THREADS, 1) void crt_reduce_cres_kernel(
    const PrimeConst *__restrict__ pc, const int8_t *__restrict__ cst, int P, int kb, int64_t W, int64_t K,
    int64_t kpad, ResSide s0, ResSide s1, int64_t nu0, int64_t nu, unsigned *__restrict__ err)
{
    reduce_cres<KS>(pc, cst, P, kb, W, K, kpad, s0, s1, nu0, nu, err);
}
// the entry-resident reduce for a plan's byte bound kb (<= RES_KB, a multiple of 64)
template <int S>
static auto reduce_res_for(int64_t kb) -> decltype(&crt_reduce_res_kernel<S, 4>)
{
    return kb <= 64 ? crt_reduce_res_kernel<S, 1> : kb <= 128 ? crt_reduce_res_kernel<S, 2>
         : kb <= 256 ? crt_reduce_res_kernel<S, 4> : crt_reduce_res_kernel<S, 8>;
}

template <bool CHUNKED>
__global__ __launch_bounds__(256, 1) void crt_gemm_kernel(
    const PrimeConst *__restrict__ pc, const int8_t *__restrict__ RA, const int8_t *__restrict__ RB,
    int64_t planeA, int64_t planeB, int64_t kpad, int64_t rowsC, int64_t colsC, int ntm, int ntn,
    int64_t total, uint32_t *__restrict__ Y, int64_t ystride)
{
    // XCD-aware order: consecutive tiles (same prime) land on one XCD
    int64_t per = (total + 7) / 8;
    int64_t v = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
    if (v >= total) return;
    gemm_tile<CHUNKED>(v, pc, RA, RB, planeA, planeB, kpad, rowsC, colsC, ntm, ntn, total, Y, ystride);
}

// Two workgroups of 4 waves share a CU (two LDS stages of 32 KB each), so
// one workgroup's fold overlaps the other's MFMAs.
template <bool CHUNKED, int WM, int STAGES>
__global__ __launch_bounds__(128 * WM, 2) void crt_gemm2_kernel(
    const PrimeConst *__restrict__ pc, int P, const int8_t *__restrict__ RA, const int8_t *__restrict__ RB,
    int64_t planeA, int64_t planeB, int64_t kpad, int64_t rowsC, int64_t colsC, int ntm, int ntn, int64_t total,
    uint32_t *__restrict__ Y, int64_t ystride)
{
    // XCD-aware order: consecutive tiles (same prime pair) land on one XCD
    int64_t per = (total + 7) / 8;
    int64_t v = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
    if (v >= total) return;
    gemm_tile2<CHUNKED, WM, STAGES>(v, pc, P, RA, RB, planeA, planeB, kpad, rowsC, colsC, ntm, ntn, Y, ystride);
}

template <int S, bool RESIDENT>
__global__ __launch_bounds__(256, 2) void crt_recon_gemm_kernel(
    const PrimeConst *__restrict__ pc, int P, const int8_t *__restrict__ Mc, int Kr, int nrb, int nw,
    const uint32_t *__restrict__ Y, int64_t ystride, int64_t nent, uint32_t *__restrict__ Wv,
    int32_t *__restrict__ Cb, int64_t *__restrict__ G, int64_t total, int epw)
{
    // XCD-aware order: the row blocks of one entry block share an L2
    int64_t per = (total + 7) / 8;
    int64_t v = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
    if (v >= total) return;
    recon_gemm_tile<S, RESIDENT>(v, pc, P, Mc, Kr, nrb, nw, Y, ystride, nent, Wv, Cb, G, epw);
}

// One-pass reconstruct of two-digit plans: one workgroup per 256 entries, in
// XCD-aware order (neighbouring entry blocks share Mc and Y lines in an L2)
__global__ __launch_bounds__(256, 2) void crt_recon_fused_kernel(
    const PrimeConst *__restrict__ pc, int P, const int8_t *__restrict__ Mc, int Kr, int nrb, int nw,
    const uint32_t *__restrict__ Y, int64_t ystride, int64_t nent, const uint32_t *__restrict__ M32, int64_t colsC,
    uint64_t *__restrict__ C, int64_t Crs, int64_t W, int64_t cap, unsigned *__restrict__ err, int64_t nblk, int gw)
{
    const int64_t per = (nblk + 7) / 8;
    const int64_t v = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
    if (v >= nblk) return;
    recon_fused_block(v, pc, P, Mc, Kr, nrb, nw, gw, Y, ystride, nent, M32, colsC, C, Crs, W, cap, err);
}

__global__ __launch_bounds__(256) void crt_recon_carry_kernel(
    const uint32_t *__restrict__ Wv, const int32_t *__restrict__ Cb, int64_t ystride, int nw, int nrb,
    const int64_t *__restrict__ G, int gshift,
    const uint32_t *__restrict__ M32, int64_t nent, int64_t colsC, uint64_t *__restrict__ C, int64_t Crs, int64_t W,
    int64_t cap, unsigned *__restrict__ err)
{
    recon_carry_block(blockIdx.x, Wv, Cb, ystride, nw, nrb, G, gshift, M32, nent, colsC, C, Crs, W, cap, err);
}

}  // namespace mgx

// ========================================================================
// host side: plan and launches
// ========================================================================
using namespace mgx;

struct mgx_crt_plan {
    int64_t K, bitsA, bitsB, bound;   // |C| < 2^bound
    int64_t S;    // balanced int8 digits per residue (2 or 4)
    int64_t P, kb, Kr, rows4pad, Lbp, out_limbs;   // rows4pad: S*P rounded up to 64
    int64_t nw;   // 32-bit words of the reconstruction's column sums kept (even)
    std::vector<uint32_t> primes;
    PrimeConst *d_pc = nullptr;
    int8_t *d_cst = nullptr;
    int8_t *d_cstb = nullptr;   // the reduce table in 1 KB blocks (rg_table_blocked), when the tiled reduce takes it
    int8_t *d_mc = nullptr;
    // the one-pass reconstruct's copy of d_mc with the gamma rows: rows 4 nw + t hold the digits of the
    // per-prime gamma constants (gw = 2 extra words), when they add no row block; else null and gw = 0
    int8_t *d_mcg = nullptr;
    int64_t gw = 0, Lg = 0;
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

// Largest admissible prime for S digits: the centred residue must be a
// sum of S balanced digits, |r| <= 127 * (256^S - 1) / 255.
uint64_t prime_limit(int S) { return S == 2 ? 65279 : (1ull << 31) - 1; }
constexpr uint64_t PRIME_FLOOR_S2 = 16384;   // below this, 16-bit primes add too few bits

std::vector<uint32_t> &prime_cache(int S)
{
    static std::vector<uint32_t> v2, v4;
    return S == 2 ? v2 : v4;
}
// guards both prime caches: plans may be created from several host threads
std::mutex g_prime_mu;

// primes below the limit for S, descending, until their product exceeds
// 2^bits; empty when the primes above the floor do not suffice (S = 2)
std::vector<uint32_t> primes_for_bits(int64_t bits, hb::Nat &M, int S)
{
    std::lock_guard<std::mutex> lk(g_prime_mu);
    auto &cache = prime_cache(S);
    std::vector<uint32_t> out;
    M = hb::from_u64(1);
    uint64_t c = cache.empty() ? prime_limit(S) : cache.back() - 1;
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
            while (c >= 3 && !is_prime(c)) c--;
            if (S == 2 && c < PRIME_FLOOR_S2) { out.clear(); return out; }
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
    return plan_create_slices(out, K, bitsA, bitsB, param(MGX_PARAM_CRT_SLICES));
}

}  // extern "C"

namespace mgx {
// A plan with the digit count given (0: 2 when the primes below 65280
// suffice, else 4; 2 or 4 to force).
int plan_create_slices(mgx_crt_plan **out, int64_t K, int64_t bitsA, int64_t bitsB, int64_t want)
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
    if (want != 0 && want != 2 && want != 4) {
        delete pl;
        return fail(MGX_ERR_ARG, "slice count %lld is not 0, 2 or 4", (long long)want);
    }
    pl->S = 4;
    if (want != 4) {
        pl->primes = primes_for_bits(pl->bound + 2, M, 2);
        if (!pl->primes.empty()) pl->S = 2;
        else if (want == 2) {
            delete pl;
            return fail(MGX_ERR_UNSUPPORTED, "a %lld-bit bound needs more primes below 65280 than exist",
                        (long long)pl->bound);
        }
    }
    if (pl->S == 4) pl->primes = primes_for_bits(pl->bound + 2, M, 4);
    const int64_t P = (int64_t)pl->primes.size(), S = pl->S;
    pl->P = P;
    pl->kb = round_up(ceil_div(std::max(bitsA, bitsB), 8), 64);
    pl->rows4pad = round_up(S * P, 64);
    pl->Kr = round_up(S * P, 64);
    if (pl->Kr > 131072) {   // int32 column sums of 4P digit products must stay below 2^31
        delete pl;
        return fail(MGX_ERR_UNSUPPORTED, "%lld primes exceed the 32768-prime limit", (long long)P);
    }
    int64_t mbytes = ceil_div(hb::bitlen(M), 8);
    pl->Lbp = round_up(mbytes + 1, 128);
    pl->out_limbs = ceil_div(pl->bound, 64);
    // sum_q yc_q (M/p_q) has at most mbytes + 1 bytes; |C| at most out_limbs limbs
    pl->nw = round_up(std::max(ceil_div(mbytes + 1, 4), 2 * pl->out_limbs), 2);

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
        c.off32 = p < 65536 ? (uint32_t)(((1ull << 31) / p + 1) * p) : 0;
        c.f45 = p < 65536 ? (int32_t)((1ull << 45) / p) : 0;
        uint64_t ks = 0, pw = 1;
        for (int64_t b = 0; b < pl->kb; b++) {
            // reduce-GEMM constants: digits of centred 2^(8b) mod p
            int32_t cen = pw > p / 2 ? (int32_t)pw - (int32_t)p : (int32_t)pw;
            int32_t vv = cen;
            for (int s = 0; s < S; s++) {
                int32_t t = ((vv + 128) & 255) - 128;
                cst[(size_t)(S * q + s) * pl->kb + b] = (int8_t)t;
                vv = (vv - t) >> 8;
            }
            ks = (ks + pw) % p;
            pw = (pw * 256) % p;
        }
        c.kofs = (uint32_t)(ks * 128 % p);
        c.kneg = (uint32_t)((p - ks * 127 % p) % p);
        // reconstruct constants: digits of (M/p) 256^s
        for (int s = 0; s < S; s++) {
            hb::Nat v = hb::shl(Mq, 8 * s);
            balanced_digits(v, pl->Lbp, mc.data() + S * q + s, pl->Kr);
        }
    }
    std::vector<uint32_t> m32(pl->Lbp / 4, 0);
    for (size_t i = 0; i < M.size() && i < m32.size(); i++) m32[i] = M[i];
    std::vector<int8_t> cstb;
    if (S == 2 && pl->kb > RES_KB) cstb = rg_table_blocked(cst.data(), P, pl->kb);
    // Gamma rows for the one-pass reconstruct: with f_q = floor(2^45 / p_q) in balanced digits f_qj
    // (j = 0..3), row 4 nw + t, column 2q + s holds f_q,(t - s), so the GEMM's words nw and nw + 1 of an
    // entry carry sum_q y_q f_q (y_q = y_q0 + 256 y_q1, its two digits): the fixed-point gamma the
    // reconstruct would otherwise sum on the vector ALUs.  Only when the two extra words add no row block.
    std::vector<int8_t> mcg;
    if (S == 2 && ceil_div(4 * pl->nw + 8, RECON_ROWS) == ceil_div(4 * pl->nw, RECON_ROWS)) {
        pl->Lg = round_up(4 * pl->nw + 8, RECON_ROWS);
        mcg.assign((size_t)pl->Lg * pl->Kr, 0);
        std::copy(mc.begin(), mc.begin() + (size_t)std::min(pl->Lbp, 4 * pl->nw) * pl->Kr, mcg.begin());
        bool ok = true;
        for (int64_t q = 0; q < P && ok; q++) {
            int32_t vv = pc[q].f45, fd[4];
            for (int j = 0; j < 4; j++) {
                fd[j] = ((vv + 128) & 255) - 128;
                vv = (vv - fd[j]) >> 8;
            }
            ok = vv == 0 && pc[q].f45 > 0;
            for (int t = 0; t < 5 && ok; t++)
                for (int sd = 0; sd < 2; sd++)
                    if (t - sd >= 0 && t - sd < 4)
                        mcg[(size_t)(4 * pl->nw + t) * pl->Kr + 2 * q + sd] = (int8_t)fd[t - sd];
        }
        if (ok) pl->gw = 2;
        else mcg.clear();
    }

    hipError_t e;
    if ((e = hipMalloc(&pl->d_pc, sizeof(PrimeConst) * P)) != hipSuccess ||
        (e = hipMalloc(&pl->d_cst, cst.size())) != hipSuccess || (e = hipMalloc(&pl->d_mc, mc.size())) != hipSuccess ||
        (e = hipMalloc(&pl->d_m32, m32.size() * 4)) != hipSuccess ||
        (!cstb.empty() && (e = hipMalloc(&pl->d_cstb, cstb.size())) != hipSuccess) ||
        (!mcg.empty() && (e = hipMalloc(&pl->d_mcg, mcg.size())) != hipSuccess)) {
        mgx_crt_plan_destroy(pl);
        return fail(MGX_ERR_OOM, "plan allocation: %s", hipGetErrorString(e));
    }
    MGX_HIP(hipMemcpy(pl->d_pc, pc.data(), sizeof(PrimeConst) * P, hipMemcpyHostToDevice));
    MGX_HIP(hipMemcpy(pl->d_cst, cst.data(), cst.size(), hipMemcpyHostToDevice));
    MGX_HIP(hipMemcpy(pl->d_mc, mc.data(), mc.size(), hipMemcpyHostToDevice));
    MGX_HIP(hipMemcpy(pl->d_m32, m32.data(), m32.size() * 4, hipMemcpyHostToDevice));
    if (!cstb.empty()) MGX_HIP(hipMemcpy(pl->d_cstb, cstb.data(), cstb.size(), hipMemcpyHostToDevice));
    if (!mcg.empty()) MGX_HIP(hipMemcpy(pl->d_mcg, mcg.data(), mcg.size(), hipMemcpyHostToDevice));
    *out = pl;
    return MGX_OK;
}
}  // namespace mgx

namespace mgx {
namespace {
int device_cus()
{
    static std::atomic<int> cus[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    int n = cus[dev].load(std::memory_order_relaxed);
    if (!n) {
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n < 1) n = 256;
        cus[dev].store(n, std::memory_order_relaxed);
    }
    return n;
}
}  // namespace

// The reduce of one or two operands of a plan whose entries fit the
// entry-resident tiles (kb <= RES_KB): nA (nB) padded entries of side a (b;
// nB = 0: one side).  Two-digit plans whose constant table fits LDS take the
// constant-resident kernel (one workgroup per CU), unless MGX_PARAM_REDUCE
// asks for the tiles.
int reduce_sides_launch(const mgx_crt_plan *pl, const ResSide &sa, const ResSide &sb, int64_t nA, int64_t nB,
                        int64_t K, int64_t kpad, int64_t cap, hipStream_t s)
{
    const int64_t lds = cres_lds_bytes(pl->S, pl->P, pl->kb);
    if (pl->S == 2 && pl->kb <= 256 && lds <= CRES_MAX_LDS && param(MGX_PARAM_REDUCE) == 0) {
        auto kern = pl->kb <= 64 ? crt_reduce_cres_kernel<1> : pl->kb <= 128 ? crt_reduce_cres_kernel<2>
                                                                              : crt_reduce_cres_kernel<4>;
        MGX_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        // (nA and nB are multiples of 128 x 64: whole 4 x 64 patches)
        const int64_t nu0 = nA / 256, nu = nu0 + nB / 256;
        const int64_t grid = std::min<int64_t>(device_cus(), nu);
        if (grid > 0)
            hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(CRES_THREADS), (unsigned)lds, s, pl->d_pc, pl->d_cst,
                               (int)pl->P, (int)pl->kb, cap + 1, K, kpad, sa, sb, nu0, nu, device_error_word());
        MGX_HIP(hipGetLastError());
        return MGX_OK;
    }
    auto res = pl->S == 2 ? reduce_res_for<2>(pl->kb) : reduce_res_for<4>(pl->kb);
    const int64_t nblk0 = ceil_div(nA, 128), nblk = nblk0 + ceil_div(nB, 128);
    if (nblk > 0x7fffffffll) return fail(MGX_ERR_SHAPE, "matrices too large for one reduce launch");
    hipLaunchKernelGGL(res, dim3((unsigned)nblk), dim3(256), 0, s, pl->d_pc, pl->d_cst, (int)pl->P,
                       (int)(pl->rows4pad / 64), (int)pl->kb, cap + 1, K, kpad, sa, sb, nblk0, device_error_word());
    MGX_HIP(hipGetLastError());
    return MGX_OK;
}
}  // namespace mgx

extern "C" {

int mgx_crt_plan_destroy(mgx_crt_plan *pl)
{
    if (!pl) return MGX_OK;
    (void)hipFree(pl->d_pc);
    (void)hipFree(pl->d_cst);
    (void)hipFree(pl->d_cstb);
    (void)hipFree(pl->d_mcg);
    (void)hipFree(pl->d_mc);
    (void)hipFree(pl->d_m32);
    delete pl;
    return MGX_OK;
}

int64_t mgx_crt_plan_primes(const mgx_crt_plan *pl) { return pl ? pl->P : -1; }
int64_t mgx_crt_plan_slices(const mgx_crt_plan *pl) { return pl ? pl->S : -1; }

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
    if (!pl || !X || !R) return fail(MGX_ERR_ARG, "null argument");
    if (nbytes != 0) return fail(MGX_ERR_ARG, "nbytes is reserved and must be 0");
    if (rows < 0 || K < 0 || cap < 1) return fail(MGX_ERR_SHAPE, "bad shape");
    const int xbits = (int)(transpose ? pl->bitsB : pl->bitsA);
    int64_t rpad, kpad;
    mgx_crt_residue_dims(pl, rows, K, &rpad, &kpad);
    int64_t nent = rpad * kpad, plane = rpad * kpad;
    int nrb = (int)(pl->rows4pad / 64);
    const bool resident = pl->kb <= RES_KB;   // the entries' bytes fit registers: read them once
    int64_t nblk = ceil_div(nent, 128) * (resident ? 1 : nrb);
    if (nblk > 0x7fffffffll) return fail(MGX_ERR_SHAPE, "matrix too large for one reduce launch");
    auto res = pl->S == 2 ? reduce_res_for<2>(pl->kb) : reduce_res_for<4>(pl->kb);
    auto gen = pl->S == 2 ? crt_reduce_kernel<2> : crt_reduce_kernel<4>;
    if (resident) {
        const ResSide sd{(const uint64_t *)X, Xrs, rows, plane, transpose, xbits, R};
        return reduce_sides_launch(pl, sd, sd, nent, 0, K, kpad, cap, as_stream(stream));
    } else if (pl->S == 2 && param(MGX_PARAM_REDUCE) == 0) {   // wide entries: the reduce as a tiled MFMA GEMM
        const int64_t total = rg_tiles(nent, pl->P), grid = round_up(total, 8);
        if (grid > 0x7fffffffll) return fail(MGX_ERR_SHAPE, "matrix too large for one reduce launch");
        hipLaunchKernelGGL(crt_reduce_gemm_kernel, dim3((unsigned)grid), dim3(256), 0, as_stream(stream), pl->d_pc,
                           pl->d_cstb, (int)pl->P, (int)pl->kb, xbits, (const uint64_t *)X, Xrs, cap + 1, transpose, rows,
                           K, kpad, nent, R, plane, total, device_error_word());
    } else
        hipLaunchKernelGGL(gen, dim3((unsigned)nblk), dim3(256), 0, as_stream(stream), pl->d_pc, pl->d_cst,
                           (int)pl->P, nrb, (int)pl->kb, xbits, (const uint64_t *)X, Xrs, cap + 1, transpose, rows, K, kpad,
                           nent, R, plane, device_error_word());
    MGX_HIP(hipGetLastError());
    return MGX_OK;
}

int mgx_crt_residue_gemm(const mgx_crt_plan *pl, int64_t K, int64_t M, int64_t N, const int8_t *RA,
                         const int8_t *RB, uint32_t *Y, void *stream)
{
    return residue_gemm_launch(pl, K, M, N, RA, RB, Y, as_stream(stream), (int)param(MGX_PARAM_GEMM_ROWS));
}

}  // extern "C"

namespace mgx {
// rows: the S = 2 tile's height, 128 (also for 0) or 64 (more workgroups for
// small products); S = 4 products have one tile shape
int residue_gemm_launch(const mgx_crt_plan *pl, int64_t K, int64_t M, int64_t N, const int8_t *RA, const int8_t *RB,
                        uint32_t *Y, hipStream_t stream, int rows)
{
    if (!pl || !RA || !RB || !Y) return fail(MGX_ERR_ARG, "null argument");
    int64_t rpadA, rpadB, kpad;
    mgx_crt_residue_dims(pl, K, M, &rpadA, &kpad);
    mgx_crt_residue_dims(pl, N, M, &rpadB, nullptr);
    int64_t ystride = crt_ystride(K * N);
    if (pl->S == 2) {   // rows x 128 tiles of prime pairs (rpadA is a multiple of 128)
        const int wm = rows == 64 ? 1 : 2;
        int ntm = (int)(rpadA / (64 * wm)), ntn = (int)(rpadB / 128);
        int64_t total = (int64_t)ntm * ntn * ceil_div(pl->P, 2);
        int64_t grid = round_up(total, 8);
        if (grid > 0x7fffffffll) return fail(MGX_ERR_SHAPE, "product too large for one GEMM launch");
        auto kern = wm == 1 ? (kpad > KCHUNK2 ? crt_gemm2_kernel<true, 1, 2> : crt_gemm2_kernel<false, 1, 2>)
                            : (kpad > KCHUNK2 ? crt_gemm2_kernel<true, 2, 2> : crt_gemm2_kernel<false, 2, 2>);
        hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(128 * wm), 0, stream, pl->d_pc, (int)pl->P, RA,
                           RB, rpadA * kpad, rpadB * kpad, kpad, K, N, ntm, ntn, total, Y, ystride);
        MGX_HIP(hipGetLastError());
        return MGX_OK;
    }
    int ntm = (int)(rpadA / 128), ntn = (int)(rpadB / 64);
    int64_t total = (int64_t)ntm * ntn * pl->P;
    int64_t grid = round_up(total, 8);
    if (grid > 0x7fffffffll) return fail(MGX_ERR_SHAPE, "product too large for one GEMM launch");
    // products with K beyond one chunk take the variant that folds between chunks
    if (kpad <= KCHUNK)
        hipLaunchKernelGGL(crt_gemm_kernel<false>, dim3((unsigned)grid), dim3(256), 0, stream, pl->d_pc,
                           RA, RB, rpadA * kpad, rpadB * kpad, kpad, K, N, ntm, ntn, total, Y, ystride);
    else
        hipLaunchKernelGGL(crt_gemm_kernel<true>, dim3((unsigned)grid), dim3(256), 0, stream, pl->d_pc,
                           RA, RB, rpadA * kpad, rpadB * kpad, kpad, K, N, ntm, ntn, total, Y, ystride);
    MGX_HIP(hipGetLastError());
    return MGX_OK;
}
}  // namespace mgx

namespace {
int recon_row_blocks(const mgx_crt_plan *pl) { return (int)ceil_div(4 * pl->nw, RECON_ROWS); }

// Both reconstruct passes with caller-provided scratch: Wv holds nw x ystride
// 32-bit words, then Cb, the row blocks' carries (nrb x ystride int32), and
// G, their fixed-point partial gamma sums (nrb x ystride int64).
int reconstruct_with(const mgx_crt_plan *pl, int64_t K, int64_t N, const uint32_t *Y, mgx_limb *C, int64_t Crs,
                     int64_t cap, char *scr, hipStream_t s)
{
    const int64_t nent = K * N, ystride = crt_ystride(nent);
    const int nrb = recon_row_blocks(pl);
    // one pass, no scratch, when the entry blocks alone (one workgroup each, every row block in turn) give
    // every CU one and the row blocks are few (measured: ahead from 256 blocks of 256 entries at 5 and 9 row
    // blocks, behind at 17 and more, where a workgroup's sweeps run long and the last ones run alone);
    // otherwise the two passes, which spread the row blocks over workgroups
    const int64_t nblk = ceil_div(nent, RECON_ENT);
    const int64_t rmode = param(MGX_PARAM_RECON);
    if (pl->S == 2 && 4 * pl->nw <= pl->Lbp &&
        (rmode == 2 || (rmode == 0 && nrb <= 9 && nblk >= device_cus()))) {
        if (round_up(nblk, 8) > 0x7fffffffll) return fail(MGX_ERR_SHAPE, "product too large for one reconstruct launch");
        if (nblk > 0)
            hipLaunchKernelGGL(crt_recon_fused_kernel, dim3((unsigned)round_up(nblk, 8)), dim3(256), 0, s, pl->d_pc,
                               (int)pl->P, pl->gw ? pl->d_mcg : pl->d_mc, (int)pl->Kr, nrb, (int)pl->nw, Y, ystride,
                               nent, pl->d_m32, N, (uint64_t *)C, Crs, cap + 1, cap, device_error_word(), nblk,
                               (int)pl->gw);
        MGX_HIP(hipGetLastError());
        return MGX_OK;
    }
    uint32_t *Wv = (uint32_t *)scr;
    int32_t *Cb = (int32_t *)(scr + pl->nw * ystride * 4);
    int64_t *G = (int64_t *)(scr + (pl->nw + nrb) * ystride * 4);
    const int64_t lds = recon_res_lds(pl->Kr, pl->S);
    const bool resident = lds <= RECON_RES_MAX_LDS;
    // resident workgroups sweep epw entry blocks each
    const int epw = resident ? recon_epw(nent, nrb) : 1;
    const int64_t total = ceil_div(nent, RECON_ENT * epw) * nrb;
    if (round_up(total, 8) > 0x7fffffffll) return fail(MGX_ERR_SHAPE, "product too large for one reconstruct launch");
    auto kern = pl->S == 2 ? (resident ? crt_recon_gemm_kernel<2, true> : crt_recon_gemm_kernel<2, false>)
                           : (resident ? crt_recon_gemm_kernel<4, true> : crt_recon_gemm_kernel<4, false>);
    if (resident) MGX_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(kern, dim3((unsigned)round_up(total, 8)), dim3(256), resident ? (unsigned)lds : 0u, s, pl->d_pc,
                       (int)pl->P, pl->d_mc, (int)pl->Kr, nrb, (int)pl->nw, Y, ystride, nent, Wv, Cb, G, total, epw);
    MGX_HIP(hipGetLastError());
    hipLaunchKernelGGL(crt_recon_carry_kernel, dim3((unsigned)ceil_div(nent, CARRY_ENT)), dim3(256), 0, s, Wv, Cb,
                       ystride, (int)pl->nw, nrb, G, pl->S == 2 ? 45 : 32, pl->d_m32, nent, N, (uint64_t *)C, Crs,
                       cap + 1, cap,
                       device_error_word());
    MGX_HIP(hipGetLastError());
    return MGX_OK;
}

size_t recon_scratch_bytes(const mgx_crt_plan *pl, int64_t nent)
{
    return (size_t)((pl->nw * 4 + recon_row_blocks(pl) * 12) * crt_ystride(nent));
}
}  // namespace

extern "C" {

int mgx_crt_reconstruct(const mgx_crt_plan *pl, int64_t K, int64_t N, const uint32_t *Y, mgx_limb *C, int64_t Crs,
                        int64_t cap, void *stream)
{
    if (!pl || !Y || !C) return fail(MGX_ERR_ARG, "null argument");
    if (K == 0 || N == 0) return MGX_OK;
    hipStream_t s = as_stream(stream);
    char *scr = nullptr;
    MGX_TRY(scratch_alloc((void **)&scr, recon_scratch_bytes(pl, K * N), s));
    int st = reconstruct_with(pl, K, N, Y, C, Crs, cap, scr, s);
    scratch_free(scr, s);
    return st;
}

int64_t mgx_crt_y_stride(int64_t K, int64_t N) { return crt_ystride(K * N); }

int mgx_crt_plan_cache_stats(int64_t *total, int64_t *pinned)
{
    if (!total || !pinned) return fail(MGX_ERR_ARG, "null argument");
    plan_cache_stats(total, pinned);
    return MGX_OK;
}

}  // extern "C"

namespace {
// Bit lengths of up to two matrices with one stream synchronization.
int scan_bits_pair(int nmat, const int64_t *rows, const int64_t *cols, const mgx_limb *const *X, const int64_t *Xrs,
                   int64_t cap, int64_t *bits, hipStream_t s, bool exact)
{
    int *d = nullptr;
    MGX_TRY(scratch_alloc((void **)&d, 2 * sizeof(int), s));
    MGX_HIP(hipMemsetAsync(d, 0, 2 * sizeof(int), s));
    ScanArgs sa{};
    int grid[2] = {0, 0};
    for (int m = 0; m < nmat; m++) {
        const int64_t n = rows[m] * cols[m];
        sa.X[m] = (const uint64_t *)X[m];
        sa.rs[m] = Xrs[m];
        sa.rows[m] = rows[m];
        sa.cols[m] = cols[m];
        grid[m] = n > 0 ? (int)std::min<int64_t>(ceil_div(n, 1024), 2048) : 0;
    }
    sa.nblk0 = grid[0];
    if (grid[0] + grid[1] > 0)
        hipLaunchKernelGGL(exact ? scan_bits_kernel<true> : scan_bits_kernel<false>, dim3(grid[0] + grid[1]), dim3(256),
                           0, s, sa, cap + 1, d);
    int h[2] = {0, 0};
    MGX_HIP(hipMemcpyAsync(h, d, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
    MGX_TRY(scratch_free(d, s));
    MGX_HIP(hipStreamSynchronize(s));
    for (int m = 0; m < nmat; m++) bits[m] = h[m];
    return MGX_OK;
}
}  // namespace

extern "C" {

int mgx_crt_scan_bits(int64_t rows, int64_t cols, const mgx_limb *X, int64_t Xrs, int64_t cap, int64_t *bits,
                      void *stream)
{
    MGX_TRY(ensure_init());
    return scan_bits_pair(1, &rows, &cols, &X, &Xrs, cap, bits, as_stream(stream), true);
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
    const mgx_crt_plan *pl = plan;
    if (!pl) {
        const int64_t rows[2] = {K, M}, cols[2] = {M, N}, rs[2] = {Ars, Brs};
        const mgx_limb *X[2] = {A, B};
        int64_t bits[2];
        // limb-granular bounds (64 |size|): one cache line per entry, and the
        // same plan as exact bit lengths whenever the top limbs are full
        MGX_TRY(scan_bits_pair(2, rows, cols, X, rs, cap, bits, s, false));
        MGX_TRY(cached_plan(M, bits[0], bits[1], &pl));   // pinned until this call has synchronized
    } else if (plan->K < M) {
        return fail(MGX_ERR_SHAPE, "plan made for inner dimension %lld < %lld", (long long)plan->K, (long long)M);
    }
    int st = crt_matmul_with(pl, K, M, N, C, Crs, A, Ars, B, Brs, cap, s, (int)param(MGX_PARAM_GEMM_ROWS));
    if (!plan) release_plan(pl);
    return st;
}

}  // extern "C"

namespace mgx {
namespace {
// One scratch block per product: residues of A and B, Y, then the
// reconstruction's column sums and gamma (which reuse the residue bytes when
// those are larger).  Every byte a stage reads was written by an earlier
// stage of the same product (the residue planes including their zero
// padding), so the block needs no clearing.
struct CrtLayout {
    int64_t rpadA, rpadB, kpad;
    size_t bytesA, offY, offR, total;
};
CrtLayout crt_layout(const mgx_crt_plan *pl, int64_t K, int64_t M, int64_t N)
{
    CrtLayout L;
    mgx_crt_residue_dims(pl, K, M, &L.rpadA, &L.kpad);
    mgx_crt_residue_dims(pl, N, M, &L.rpadB, nullptr);
    L.bytesA = (size_t)(pl->S * pl->P * L.rpadA * L.kpad);
    const size_t bytesB = (size_t)(pl->S * pl->P * L.rpadB * L.kpad);
    const size_t bytesY = (size_t)ceil_div(pl->S * pl->P, 4) * crt_ystride(K * N) * 4;
    const size_t bytesR = recon_scratch_bytes(pl, K * N);
    L.offY = (size_t)round_up((int64_t)(L.bytesA + bytesB), 256);
    L.offR = bytesR <= L.offY ? 0 : (size_t)round_up((int64_t)(L.offY + bytesY), 256);
    L.total = std::max(L.offY + bytesY, L.offR + bytesR);
    return L;
}

// The stage launches of one product on s (no allocation, no synchronization).
int crt_launch(const mgx_crt_plan *pl, int64_t K, int64_t M, int64_t N, mgx_limb *C, int64_t Crs, const mgx_limb *A,
               int64_t Ars, const mgx_limb *B, int64_t Brs, int64_t cap, char *scr, const CrtLayout &L, hipStream_t s,
               int gemm_rows)
{
    void *stream = s;
    int8_t *RA = (int8_t *)scr, *RB = (int8_t *)scr + L.bytesA;
    uint32_t *Y = (uint32_t *)(scr + L.offY);
    if (pl->kb <= RES_KB) {   // both operands in one launch: one tail instead of two
        const int64_t nA = L.rpadA * L.kpad, nB = L.rpadB * L.kpad;
        const ResSide sa{(const uint64_t *)A, Ars, K, nA, 0, (int)pl->bitsA, RA},
            sb{(const uint64_t *)B, Brs, N, nB, 1, (int)pl->bitsB, RB};
        MGX_TRY(reduce_sides_launch(pl, sa, sb, nA, nB, M, L.kpad, cap, s));
    } else {
        MGX_TRY(mgx_crt_reduce(pl, 0, K, M, A, Ars, cap, 0, RA, stream));
        MGX_TRY(mgx_crt_reduce(pl, 1, N, M, B, Brs, cap, 0, RB, stream));
    }
    MGX_TRY(residue_gemm_launch(pl, K, M, N, RA, RB, Y, s, gemm_rows));
    return reconstruct_with(pl, K, N, Y, C, Crs, cap, scr + L.offR, s);
}
}  // namespace

// The product through a plan (stream-ordered), then one synchronization for
// the device error word.
int crt_matmul_with(const mgx_crt_plan *pl, int64_t K, int64_t M, int64_t N, mgx_limb *C, int64_t Crs,
                    const mgx_limb *A, int64_t Ars, const mgx_limb *B, int64_t Brs, int64_t cap, hipStream_t s,
                    int gemm_rows)
{
    const CrtLayout L = crt_layout(pl, K, M, N);
    char *scr = nullptr;
    int st = scratch_alloc((void **)&scr, L.total, s);
    if (st == MGX_OK) {
        st = crt_launch(pl, K, M, N, C, Crs, A, Ars, B, Brs, cap, scr, L, s, gemm_rows);
        scratch_free(scr, s);
    }
    if (st == MGX_OK) st = check_device_error(s);
    return st;
}

// ------------------------------------------------------------------------
// The entry point that allocates nothing: its scratch block is one
// file-scope device array per GPU, of NOALLOC_SCRATCH_BYTES, held until the
// process ends.  In host code `scratch` names a host-side stand-in, never
// the device memory: its device address is looked up once per device (at
// library init, with that device current) and only that address is passed
// to kernels.
// ------------------------------------------------------------------------
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)MGX_NOALLOC_SCRATCH_MB << 20;
static __device__ __align__(256) char scratch[NOALLOC_SCRATCH_BYTES];
namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_noalloc_mu;
char *g_noalloc_ptr[NOALLOC_MAX_DEVICES] = {};
}  // namespace

int noalloc_scratch_init(int dev)
{
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail(MGX_ERR_UNSUPPORTED, "device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (g_noalloc_ptr[dev]) return MGX_OK;
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(MGX_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_noalloc_ptr[dev] = (char *)p;
    return MGX_OK;
}

__global__ void zero_entries_kernel(uint64_t *C, int64_t Crs, int64_t K, int64_t N, int64_t W)
{
    const int64_t total = K * N * W;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = t / W;
        C[((e / N) * Crs + e % N) * W + t % W] = 0;
    }
}

int plan_view(const mgx_crt_plan *pl, PlanView *v)
{
    v->pc = pl->d_pc; v->cst = pl->d_cst; v->cstb = pl->d_cstb; v->mc = pl->d_mc; v->m32 = pl->d_m32;
    v->mcg = pl->d_mcg; v->gw = pl->gw;
    v->S = pl->S; v->P = pl->P; v->kb = pl->kb; v->Kr = pl->Kr; v->rows4pad = pl->rows4pad; v->Lbp = pl->Lbp;
    v->out_limbs = pl->out_limbs;
    v->nw = pl->nw;
    v->bitsA = pl->bitsA;
    v->bitsB = pl->bitsB;
    return MGX_OK;
}

// Plans made for (device, ceil(log2 K), bitsA, bitsB, slice setting) are kept
// for the life of the process, so that repeated products of one shape pay for
// the constant tables once.  Every taker pins its plan and releases it when no
// queued or captured work of its own reads the plan any more.  At most
// PLAN_CACHE_MAX unpinned plans are kept: a new shape evicts the least
// recently used unpinned plan (pinned ones are never freed, so a batch with
// more distinct shapes than that simply keeps them all).
namespace {
constexpr size_t PLAN_CACHE_MAX = 256;
struct CachedPlan {
    std::array<int64_t, 5> key;
    mgx_crt_plan *plan;
    int pins;
    uint64_t used;   // logical clock of the last take
};
std::mutex g_plan_mu;
std::vector<CachedPlan> g_plans;
uint64_t g_plan_clock = 0;
}  // namespace

int cached_plan(int64_t K, int64_t bitsA, int64_t bitsB, const mgx_crt_plan **out, int slices)
{
    int dev = 0;
    MGX_HIP(hipGetDevice(&dev));
    const int64_t want = slices >= 0 ? slices : param(MGX_PARAM_CRT_SLICES);
    std::array<int64_t, 5> key = {dev, clog2(K), bitsA, bitsB, want};
    std::lock_guard<std::mutex> lk(g_plan_mu);
    for (auto &c : g_plans)
        if (c.key == key) {
            c.pins++;
            c.used = ++g_plan_clock;
            *out = c.plan;
            return MGX_OK;
        }
    mgx_crt_plan *pl = nullptr;
    MGX_TRY(plan_create_slices(&pl, (int64_t)1 << clog2(K), bitsA, bitsB, want));
    size_t unpinned = 0;
    for (auto &c : g_plans) unpinned += c.pins == 0;
    if (unpinned >= PLAN_CACHE_MAX) {
        size_t victim = g_plans.size();
        for (size_t i = 0; i < g_plans.size(); i++)
            if (g_plans[i].pins == 0 && (victim == g_plans.size() || g_plans[i].used < g_plans[victim].used))
                victim = i;
        // an unpinned plan has no queued work of a pinning call left (each
        // releases after its own synchronization); the device-wide wait
        // covers work on other streams that a caller-given plan shares
        (void)hipDeviceSynchronize();
        mgx_crt_plan_destroy(g_plans[victim].plan);
        g_plans.erase(g_plans.begin() + (ptrdiff_t)victim);
    }
    g_plans.push_back({key, pl, 1, ++g_plan_clock});
    *out = pl;
    return MGX_OK;
}

void release_plan(const mgx_crt_plan *pl)
{
    std::lock_guard<std::mutex> lk(g_plan_mu);
    for (auto &c : g_plans)
        if (c.plan == pl && c.pins > 0) {
            c.pins--;
            return;
        }
}

// Cache occupancy for tests: total plans and pinned plans.
void plan_cache_stats(int64_t *total, int64_t *pinned)
{
    std::lock_guard<std::mutex> lk(g_plan_mu);
    int64_t p = 0;
    for (auto &c : g_plans) p += c.pins > 0;
    *total = (int64_t)g_plans.size();
    *pinned = p;
}
}  // namespace mgx

extern "C" {

int mgx_crt_mfma_matmul_noalloc(const mgx_crt_plan *plan, int64_t K, int64_t M, int64_t N, mgx_limb *C, int64_t Crs,
                                const mgx_limb *A, int64_t Ars, const mgx_limb *B, int64_t Brs, int64_t cap,
                                void *stream)
{
    if (!plan) return fail(MGX_ERR_ARG, "the no-allocation product needs a plan made with mgx_crt_plan_create");
    if (K < 0 || M < 0 || N < 0 || cap < 1) return fail(MGX_ERR_SHAPE, "bad shape");
    if (plan->K < M)
        return fail(MGX_ERR_SHAPE, "plan made for inner dimension %lld < %lld", (long long)plan->K, (long long)M);
    int dev = 0;
    MGX_HIP(hipGetDevice(&dev));
    char *scr = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_noalloc_mu);
        if (dev >= 0 && dev < NOALLOC_MAX_DEVICES) scr = g_noalloc_ptr[dev];
    }
    if (!scr || !device_error_word())
        return fail(MGX_ERR_ARG, "call mgx_init on device %d before the no-allocation product", dev);
    if (K == 0 || N == 0) return MGX_OK;
    hipStream_t s = as_stream(stream);
    if (M == 0) {   // empty inner dimension: C = 0 (a fill kernel, not a memset)
        const int64_t total = K * N * (cap + 1);
        hipLaunchKernelGGL(zero_entries_kernel, dim3((unsigned)std::min<int64_t>(ceil_div(total, 256), 4096)),
                           dim3(256), 0, s, (uint64_t *)C, Crs, K, N, cap + 1);
        MGX_HIP(hipGetLastError());
        return MGX_OK;
    }
    const CrtLayout L = crt_layout(plan, K, M, N);
    if (L.total > NOALLOC_SCRATCH_BYTES)
        return fail(MGX_ERR_SHAPE, "this product needs %zu bytes of scratch; the no-allocation entry point has %zu",
                    L.total, NOALLOC_SCRATCH_BYTES);
    return crt_launch(plan, K, M, N, C, Crs, A, Ars, B, Brs, cap, scr, L, s, (int)param(MGX_PARAM_GEMM_ROWS));
}

int64_t mgx_crt_noalloc_scratch_bytes(const mgx_crt_plan *plan, int64_t K, int64_t M, int64_t N)
{
    if (!plan || K < 0 || M < 1 || N < 0) return -1;
    return (int64_t)crt_layout(plan, K, M, N).total;
}

}  // extern "C"
