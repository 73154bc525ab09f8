// tune.hip -- routes, seeded device inputs, product features, the exhaustive
// autotuner and the cost-model dispatch.
#include "common.h"
#include "../../include/mgx_sched.h"
#include <algorithm>
#include <chrono>
#include <cmath>
#include <map>
#include <mutex>
#include <vector>

namespace mgx {
int plain_mul(bool strided, bool addmul, int64_t K, int64_t M, int64_t N, uint64_t *C, int64_t Crs, const uint64_t *A,
              int64_t Ars, const uint64_t *B, int64_t Brs, int64_t cap, int64_t opwords, hipStream_t s);
int squarify_with(int method, int64_t K, int64_t M, int64_t N, uint64_t *C, int64_t Crs, const uint64_t *A,
                  int64_t Ars, const uint64_t *B, int64_t Brs, int64_t cap, int64_t opbits, int64_t leaf,
                  int64_t maxdepth, hipStream_t s);

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ uint64_t hash3(uint64_t seed, uint64_t e, uint64_t l)
{
    return mix64(mix64(seed ^ mix64(e)) + l);
}

// one thread per (entry, word); the thread of word `cap` writes the size
__global__ void random_device_kernel(uint64_t *M, int64_t Wd, int64_t H, int64_t rs, int64_t cap, int64_t blo,
                                     int64_t bhi, uint64_t thr, uint64_t seed)
{
    const int64_t Wn = cap + 1, total = Wd * H * Wn;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        int64_t e = t / Wn, l = t % Wn;
        int64_t y = e / Wd, x = e % Wd;
        uint64_t *ent = M + (y * rs + x) * Wn;
        int64_t bits = hash3(seed, e, 0xffffffffull) < thr ? bhi : blo;
        int64_t nl = (bits + 63) / 64;
        auto limb = [&](int64_t i) {
            uint64_t v = hash3(seed, e, (uint64_t)i);
            if (i == nl - 1 && (bits & 63)) v &= (1ull << (bits & 63)) - 1;
            return v;
        };
        if (l < cap) {
            ent[l] = l < nl ? limb(l) : 0;
        } else {
            int64_t sz = nl;
            while (sz > 0 && limb(sz - 1) == 0) sz--;
            bool neg = hash3(seed, e, 0xfffffffeull) & 1;
            ent[cap] = (uint64_t)(neg ? -sz : sz);
        }
    }
}

// per-matrix statistics: hist[6], sum of bits, max bits, sum of 2048-bit blocks
__global__ void feature_scan_kernel(const uint64_t *X, int64_t rs, int64_t Wn, int64_t rows, int64_t cols,
                                    unsigned long long *acc)
{
    const int64_t n = rows * cols;
    unsigned long long h[6] = {0, 0, 0, 0, 0, 0}, sb = 0, mb = 0, sk = 0;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t *p = X + ((e / cols) * rs + e % cols) * Wn;
        int64_t s = (int64_t)p[Wn - 1], a = s < 0 ? -s : s;
        while (a > 0 && p[a - 1] == 0) a--;
        int64_t b = a ? (a - 1) * 64 + 64 - __clzll(p[a - 1]) : 0;
        int bin = b <= 128 ? 0 : b <= 512 ? 1 : b <= 2048 ? 2 : b <= 8192 ? 3 : b <= 32768 ? 4 : 5;
#pragma unroll
        for (int i = 0; i < 6; i++) h[i] += bin == i;
        sb += b;
        mb = b > (int64_t)mb ? b : mb;
        sk += (b + 2047) / 2048;
    }
    for (int i = 0; i < 6; i++) atomicAdd(&acc[i], h[i]);
    atomicAdd(&acc[6], sb);
    atomicMax(&acc[7], mb);
    atomicAdd(&acc[8], sk);
}

}  // namespace mgx

using namespace mgx;

namespace {

int plain_route(int route, int64_t K, int64_t M, int64_t N, mgx_limb *C, int64_t Crs, const mgx_limb *A, int64_t Ars,
                const mgx_limb *B, int64_t Brs, int64_t cap, int64_t opbits, hipStream_t s)
{
    if (route == MGX_ROUTE_PLAIN) return plain_mul(false, false, K, M, N, C, Crs, A, Ars, B, Brs, cap, (opbits + 31) / 32, s);
    return squarify_with(MGX_REC_WINOGRAD, K, M, N, C, Crs, A, Ars, B, Brs, cap, opbits, 2,
                         route == MGX_ROUTE_WINOGRAD1 ? 1 : 2, s);
}

struct Stats {
    unsigned long long v[9];
};

int scan_stats(int64_t rows, int64_t cols, const mgx_limb *X, int64_t Xrs, int64_t cap, Stats &st, hipStream_t s)
{
    unsigned long long *d = nullptr;
    MGX_TRY(scratch_alloc((void **)&d, sizeof(Stats), s));
    MGX_HIP(hipMemsetAsync(d, 0, sizeof(Stats), s));
    int64_t n = rows * cols;
    if (n > 0) {
        int grid = (int)std::min<int64_t>(std::max<int64_t>(ceil_div(n, 256), 1), 2048);
        hipLaunchKernelGGL(feature_scan_kernel, dim3(grid), dim3(256), 0, s, (const uint64_t *)X, Xrs, cap + 1, rows,
                           cols, d);
    }
    MGX_HIP(hipMemcpyAsync(&st, d, sizeof(Stats), hipMemcpyDeviceToHost, s));
    MGX_TRY(scratch_free(d, s));
    MGX_HIP(hipStreamSynchronize(s));
    return MGX_OK;
}

mgx_costmodel *g_model = nullptr;
std::mutex g_mu;
std::map<std::vector<int>, int> g_tuned;   // exhaustive choices by feature bucket

double now() { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); }

}  // namespace

extern "C" {

int mgx_random_matrix_device(mgx_limb *Mx, int64_t Wd, int64_t H, int64_t rs, int64_t cap, int64_t blo, int64_t bhi,
                             double frac_hi, uint64_t seed, void *stream)
{
    MGX_TRY(ensure_init());
    if ((blo + 63) / 64 > cap || (bhi + 63) / 64 > cap) return fail(MGX_ERR_CAPACITY, "bits exceed cap");
    uint64_t thr = frac_hi <= 0 ? 0 : frac_hi >= 1 ? ~0ull : (uint64_t)(frac_hi * 18446744073709551616.0);
    int64_t total = Wd * H * (cap + 1);
    if (!total) return MGX_OK;
    unsigned grid = (unsigned)std::min<int64_t>(ceil_div(total, 256), 65536);
    hipLaunchKernelGGL(random_device_kernel, dim3(grid), dim3(256), 0, as_stream(stream), (uint64_t *)Mx, Wd, H, rs,
                       cap, blo, bhi, thr, seed);
    MGX_HIP(hipGetLastError());
    return MGX_OK;
}

int mgx_tune_features(int64_t K, int64_t M, int64_t N, const int64_t *hist6, double mean_bits, int64_t max_bits,
                      double mean_blocks, float *f)
{
    double tot = 0;
    for (int i = 0; i < 6; i++) tot += (double)hist6[i];
    if (tot <= 0) tot = 1;
    f[0] = (float)std::log2((double)std::max<int64_t>(K, 1));
    f[1] = (float)std::log2((double)std::max<int64_t>(M, 1));
    f[2] = (float)std::log2((double)std::max<int64_t>(N, 1));
    f[3] = f[0] + f[1] + f[2];
    f[4] = (float)std::log2(std::max(mean_bits, 1.0));
    f[5] = (float)std::log2((double)std::max<int64_t>(max_bits, 1));
    f[6] = (float)std::log2(std::max(mean_blocks, 1.0));
    for (int i = 0; i < 6; i++) f[7 + i] = (float)(hist6[i] / tot);
    return MGX_OK;
}

int mgx_tune_features_device(int64_t K, int64_t M, int64_t N, const mgx_limb *A, int64_t Ars, const mgx_limb *B,
                             int64_t Brs, int64_t cap, float *f)
{
    MGX_TRY(ensure_init());
    Stats a, b;
    MGX_TRY(scan_stats(K, M, A, Ars, cap, a, nullptr));
    MGX_TRY(scan_stats(M, N, B, Brs, cap, b, nullptr));
    int64_t hist[6];
    for (int i = 0; i < 6; i++) hist[i] = (int64_t)(a.v[i] + b.v[i]);
    double cnt = (double)(K * M + M * N);
    if (cnt <= 0) cnt = 1;
    return mgx_tune_features(K, M, N, hist, (double)(a.v[6] + b.v[6]) / cnt, (int64_t)std::max(a.v[7], b.v[7]),
                             (double)(a.v[8] + b.v[8]) / cnt, f);
}

int mgx_matmul_route(int route, int64_t K, int64_t M, int64_t N, mgx_limb *C, int64_t Crs, const mgx_limb *A,
                     int64_t Ars, const mgx_limb *B, int64_t Brs, int64_t cap, void *stream)
{
    MGX_TRY(ensure_init());
    hipStream_t s = as_stream(stream);
    if (route == MGX_ROUTE_AUTO) MGX_TRY(mgx_dispatch_choose(K, M, N, A, Ars, B, Brs, cap, &route));
    if (route == MGX_ROUTE_CRT) return mgx_crt_mfma_matmul(nullptr, K, M, N, C, Crs, A, Ars, B, Brs, cap, stream);
    if (route < 0 || route >= MGX_ROUTES) return fail(MGX_ERR_ARG, "unknown route %d", route);
    if (K == 0 || N == 0) return MGX_OK;
    int64_t ba = 1, bb = 1;
    MGX_TRY(mgx_crt_scan_bits(K, M, A, Ars, cap, &ba, stream));
    MGX_TRY(mgx_crt_scan_bits(M, N, B, Brs, cap, &bb, stream));
    if (M == 0) {
        MGX_HIP(hipMemset2DAsync(C, (size_t)(Crs * (cap + 1) * 8), 0, (size_t)(N * (cap + 1) * 8), (size_t)K, s));
    } else {
        MGX_TRY(plain_route(route, K, M, N, C, Crs, A, Ars, B, Brs, cap, std::max<int64_t>(std::max(ba, bb), 1), s));
    }
    return check_device_error(s);
}

int mgx_tune_time_routes(int64_t K, int64_t M, int64_t N, mgx_limb *C, const mgx_limb *A, const mgx_limb *B,
                         int64_t cap, int reps, double *times, unsigned *estimated)
{
    MGX_TRY(ensure_init());
    if (reps < 1) reps = 1;
    unsigned est = 0;
    auto run = [&](int r, int64_t rows, double *out) -> int {
        int st0 = mgx_matmul_route(r, rows, M, N, C, N, A, M, B, N, cap, nullptr);   // warm-up (plans, pools)
        if (st0 == MGX_ERR_UNSUPPORTED || st0 == MGX_ERR_CAPACITY) {   // route cannot take this product
            *out = 1e30;
            return MGX_OK;
        }
        MGX_TRY(st0);
        std::vector<double> t;
        for (int i = 0; i < reps; i++) {
            double t0 = now();
            MGX_TRY(mgx_matmul_route(r, rows, M, N, C, N, A, M, B, N, cap, nullptr));
            t.push_back(now() - t0);
        }
        std::sort(t.begin(), t.end());
        *out = t[t.size() / 2];
        return MGX_OK;
    };
    MGX_TRY(run(MGX_ROUTE_CRT, K, &times[MGX_ROUTE_CRT]));
    // one-row-block estimate of the plain route; plain routes that cannot
    // come within 2x of the CRT time even with two Winograd levels are not run
    int64_t rows = std::min<int64_t>(K, 8);
    double t_sub;
    MGX_TRY(run(MGX_ROUTE_PLAIN, rows, &t_sub));
    double naive_est = t_sub * (double)K / (double)rows;
    bool hopeless = naive_est * 0.7 > 2.0 * times[MGX_ROUTE_CRT] && K > rows;
    for (int r = MGX_ROUTE_PLAIN; r < MGX_ROUTES; r++) {
        if (hopeless) {
            double f = r == MGX_ROUTE_PLAIN ? 1.0 : r == MGX_ROUTE_WINOGRAD1 ? 0.875 : 0.765625;
            times[r] = naive_est * f;
            est |= 1u << r;
        } else if (r == MGX_ROUTE_PLAIN && rows == K) {
            times[r] = t_sub;
        } else {
            MGX_TRY(run(r, K, &times[r]));
        }
    }
    if (estimated) *estimated = est;
    return MGX_OK;
}

int mgx_dispatch_set_model(mgx_costmodel *cm)
{
    std::lock_guard<std::mutex> lk(g_mu);
    g_model = cm;
    return MGX_OK;
}

int mgx_dispatch_choose(int64_t K, int64_t M, int64_t N, const mgx_limb *A, int64_t Ars, const mgx_limb *B,
                        int64_t Brs, int64_t cap, int *route)
{
    float f[MGX_FEATURES];
    MGX_TRY(mgx_tune_features_device(K, M, N, A, Ars, B, Brs, cap, f));
    mgx_costmodel *cm;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        cm = g_model;
    }
    if (cm) {
        float y[MGX_ROUTES];
        MGX_TRY(mgx_costmodel_predict(cm, f, 1, y));
        // the plain kernel takes operands of at most 512 32-bit words, and
        // each Winograd level adds up to two bits to them: routes that
        // cannot take the product are never chosen, whatever the prediction
        const double maxbits = std::exp2((double)f[5]);
        int best = MGX_ROUTE_CRT;
        for (int r = 1; r < MGX_ROUTES; r++) {
            const int levels = r == MGX_ROUTE_WINOGRAD1 ? 1 : r == MGX_ROUTE_WINOGRAD2 ? 2 : 0;
            if (maxbits + 2 * levels <= 512 * 32 && y[r] < y[best]) best = r;
        }
        *route = best;
        return MGX_OK;
    }
    // fallback: the exhaustive autotuner's choice for this feature bucket,
    // timed once on a same-shaped seeded problem and remembered
    std::vector<int> key;
    for (int i = 0; i < MGX_FEATURES; i++) key.push_back((int)std::lround(f[i] * (i >= 7 ? 8 : 2)));
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_tuned.find(key);
        if (it != g_tuned.end()) { *route = it->second; return MGX_OK; }
    }
    int64_t maxbits = (int64_t)std::exp2(f[5]) + 1;
    int64_t tcap = (2 * maxbits + 64 + 63) / 64 + 2;
    mgx_limb *tA = nullptr, *tB = nullptr, *tC = nullptr;
    hipStream_t s = nullptr;
    int st = scratch_alloc((void **)&tA, (size_t)(K * M * (tcap + 1) * 8), s);
    if (st == MGX_OK) st = scratch_alloc((void **)&tB, (size_t)(M * N * (tcap + 1) * 8), s);
    if (st == MGX_OK) st = scratch_alloc((void **)&tC, (size_t)(K * N * (tcap + 1) * 8), s);
    double times[MGX_ROUTES];
    if (st == MGX_OK) st = mgx_random_matrix_device(tA, M, K, M, tcap, maxbits, maxbits, 0, 1, s);
    if (st == MGX_OK) st = mgx_random_matrix_device(tB, N, M, N, tcap, maxbits, maxbits, 0, 2, s);
    if (st == MGX_OK) st = mgx_tune_time_routes(K, M, N, tC, tA, tB, tcap, 1, times, nullptr);
    scratch_free(tA, s);
    scratch_free(tB, s);
    scratch_free(tC, s);
    MGX_TRY(st);
    int best = 0;
    for (int r = 1; r < MGX_ROUTES; r++)
        if (times[r] < times[best]) best = r;
    std::lock_guard<std::mutex> lk(g_mu);
    g_tuned[key] = best;
    *route = best;
    return MGX_OK;
}

}  // extern "C"
