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
int costmodel_predict1(mgx_costmodel *cm, const float *x, float *y, hipStream_t s);
int costmodel_view(const mgx_costmodel *cm, const float **params, const float **norm, int *D, int *H, int *O);

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

// Statistics of up to two matrices in one launch (workgroups [0, nblk0)
// scan the first), with the loads of four entries in flight per thread: the
// dispatch's one pass over the inputs, whose maximum bit lengths then also
// size the multi-modular plan.  Bit lengths are limb-granular (64 |size|,
// from the size words alone: one cache line per entry).
constexpr int STAT_PITCH = 16;   // 64-bit counters 128 bytes apart
struct StatArgs {
    const uint64_t *X[2];
    int64_t rs[2], rows[2], cols[2];
    int nblk0;
};
__global__ __launch_bounds__(256) void feature_scan2_kernel(StatArgs sa, int64_t Wn, unsigned long long *acc_pair)
{
    const int m = (int)blockIdx.x < sa.nblk0 ? 0 : 1;
    const int64_t blk = m ? blockIdx.x - sa.nblk0 : blockIdx.x, nblk = m ? gridDim.x - sa.nblk0 : sa.nblk0;
    const uint64_t *X = sa.X[m];
    const int64_t rs = sa.rs[m], cols = sa.cols[m], n = sa.rows[m] * cols, stride = nblk * blockDim.x;
    // counter c of matrix m sits alone in a 128-byte line: atomics to one
    // line serialize in one L2 channel
    unsigned long long *acc = acc_pair + STAT_PITCH * 9 * m;
    __shared__ unsigned long long s_part[4][9];
    unsigned long long h[6] = {0, 0, 0, 0, 0, 0}, sb = 0, mb = 0, sk = 0;
    for (int64_t e0 = blk * (int64_t)blockDim.x + threadIdx.x; e0 < n; e0 += 4 * stride) {
        const uint64_t *p[4];
        int64_t a[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int64_t e = e0 + j * stride;
            p[j] = X + ((e / cols) * rs + e % cols) * Wn;
            const int64_t sz = e < n ? (int64_t)p[j][Wn - 1] : 0;
            a[j] = min(sz < 0 ? -sz : sz, Wn - 1);
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
            if (e0 + j * stride >= n) continue;
            const int64_t b = 64 * a[j];   // limb-granular bit length, from the size word alone
            const int bin = b <= 128 ? 0 : b <= 512 ? 1 : b <= 2048 ? 2 : b <= 8192 ? 3 : b <= 32768 ? 4 : 5;
#pragma unroll
            for (int i = 0; i < 6; i++) h[i] += bin == i;
            sb += b;
            mb = b > (int64_t)mb ? b : mb;
            sk += (b + 2047) / 2048;
        }
    }
    // wave sums first: one atomic per wave and counter
    for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
        for (int i = 0; i < 6; i++) h[i] += __shfl_xor(h[i], o);
        sb += __shfl_xor(sb, o);
        sk += __shfl_xor(sk, o);
        mb = max(mb, (unsigned long long)__shfl_xor(mb, o));
    }
    if ((threadIdx.x & 63) == 0) {
        const int w = threadIdx.x >> 6;
        for (int i = 0; i < 6; i++) s_part[w][i] = h[i];
        s_part[w][6] = sb;
        s_part[w][7] = mb;
        s_part[w][8] = sk;
    }
    __syncthreads();
    if (threadIdx.x < 9) {   // one atomic per counter per workgroup
        const int c = threadIdx.x;
        unsigned long long v = s_part[0][c];
        for (int w = 1; w < 4; w++) v = c == 7 ? max(v, s_part[w][c]) : v + s_part[w][c];
        if (v) {
            if (c == 7) atomicMax(&acc[STAT_PITCH * c], v);
            else atomicAdd(&acc[STAT_PITCH * c], v);
        }
    }
}

// The dispatch's features and inference on the device, queued right after
// the statistics scan so that one copy and one synchronization bring back
// both the bounds and the chosen schedule.  One workgroup: thread 0 forms the
// features exactly as mgx_tune_features does (in double, then float),
// threads j < H evaluate hidden unit j of each layer, threads o < O the
// outputs, and thread 0 picks the fastest predicted schedule that can take
// the operands.
struct DispatchTab {
    int plain[MGX_SCHEDULES];   // schedule is a plain route (bounded by the 512-word kernel)
    int depth[MGX_SCHEDULES];
};
__global__ __launch_bounds__(256) void dispatch_kernel(unsigned long long *stats, int64_t K, int64_t M, int64_t N,
                                                       const float *__restrict__ params, const float *__restrict__ norm,
                                                       int D, int H, int O, DispatchTab tab)
{
    __shared__ float x[32], h1[256], h2[256], y[64];
    const unsigned long long *sa = stats, *sb = stats + STAT_PITCH * 9;
    const int j = threadIdx.x;
    if (j == 0) {
        double hist[6], tot = 0;
        for (int i = 0; i < 6; i++) {
            hist[i] = (double)(sa[STAT_PITCH * i] + sb[STAT_PITCH * i]);
            tot += hist[i];
        }
        if (tot <= 0) tot = 1;
        double cnt = (double)(K * M + M * N);
        if (cnt <= 0) cnt = 1;
        const double mean_bits = (double)(sa[STAT_PITCH * 6] + sb[STAT_PITCH * 6]) / cnt;
        const unsigned long long mb = max(sa[STAT_PITCH * 7], sb[STAT_PITCH * 7]);
        const double mean_blocks = (double)(sa[STAT_PITCH * 8] + sb[STAT_PITCH * 8]) / cnt;
        float f[MGX_FEATURES];
        f[0] = (float)log2((double)max<int64_t>(K, 1));
        f[1] = (float)log2((double)max<int64_t>(M, 1));
        f[2] = (float)log2((double)max<int64_t>(N, 1));
        f[3] = f[0] + f[1] + f[2];
        f[4] = (float)log2(fmax(mean_bits, 1.0));
        f[5] = (float)log2((double)max(mb, 1ull));
        f[6] = (float)log2(fmax(mean_blocks, 1.0));
        for (int i = 0; i < 6; i++) f[7 + i] = (float)(hist[i] / tot);
        for (int i = 0; i < D; i++) x[i] = (f[i] - norm[i]) / norm[D + i];
    }
    __syncthreads();
    const float *W1 = params, *b1 = W1 + H * D, *W2 = b1 + H, *b2 = W2 + H * H, *W3 = b2 + H, *b3 = W3 + O * H;
    if (j < H) {
        float a = b1[j];
        for (int i = 0; i < D; i++) a = fmaf(W1[j * D + i], x[i], a);
        h1[j] = a > 0.f ? a : 0.f;
    }
    __syncthreads();
    if (j < H) {
        float a = b2[j];
        for (int i = 0; i < H; i++) a = fmaf(W2[j * H + i], h1[i], a);
        h2[j] = a > 0.f ? a : 0.f;
    }
    __syncthreads();
    if (j < O) {
        float a = b3[j];
        for (int i = 0; i < H; i++) a = fmaf(W3[j * H + i], h2[i], a);
        y[j] = a;
    }
    __syncthreads();
    if (j == 0) {
        const unsigned long long maxbits = max(sa[STAT_PITCH * 7], sb[STAT_PITCH * 7]);
        int best = 0;
        float by = 3.0e38f;
        for (int o = 0; o < O; o++)
            if ((!tab.plain[o] || maxbits + 2 * (unsigned long long)tab.depth[o] <= 512 * 32) && y[o] < by) {
                by = y[o];
                best = o;
            }
        stats[STAT_PITCH * 18] = (unsigned long long)best;   // one more line after the counters
    }
}

}  // namespace mgx

using namespace mgx;

namespace {

// The tuner's schedule space: route, digit count and residue-GEMM tile of
// the multi-modular route, Toom-22 cutoff and Winograd depth of the plain
// routes.  Values -1 (slices, toom) and 0 (gemm_rows) take the global
// tunables.
const mgx_schedule SCHEDULES[MGX_SCHEDULES] = {
    {MGX_ROUTE_CRT, 0, 128, 0, 0},       // two digits while the 16-bit primes suffice, 128-row tiles
    {MGX_ROUTE_CRT, 0, 64, 0, 0},        // the same with 64-row tiles (twice the workgroups)
    {MGX_ROUTE_CRT, 4, 128, 0, 0},       // primes below 2^31, four digits
    {MGX_ROUTE_PLAIN, 0, 0, 2, 0},       // per-entry kernel, Karatsuba from two 64-word blocks
    {MGX_ROUTE_PLAIN, 0, 0, 0, 0},       // per-entry kernel, schoolbook only
    {MGX_ROUTE_WINOGRAD1, 0, 0, 2, 1},   // one Winograd level over the per-entry kernel
    {MGX_ROUTE_WINOGRAD2, 0, 0, 2, 2},   // two levels
    {MGX_ROUTE_WINOGRAD1, 0, 0, 0, 1},
    {MGX_ROUTE_WINOGRAD2, 0, 0, 0, 2},
};

// the schedule a fixed route runs (global tunables for its parameters)
mgx_schedule route_schedule(int route)
{
    mgx_schedule sc{route, -1, 0, -1, route == MGX_ROUTE_WINOGRAD1 ? 1 : route == MGX_ROUTE_WINOGRAD2 ? 2 : 0};
    return sc;
}

// operands of maxbits bits fit the plain kernel's widest class (512 words)
// after the schedule's Winograd levels (two bits of growth each)
bool schedule_feasible(const mgx_schedule &sc, int64_t maxbits)
{
    return sc.route == MGX_ROUTE_CRT || maxbits + 2 * sc.depth <= 512 * 32;
}

// C = A*B through a schedule, operand bit bounds known; synchronizes once
// (device error word)
int run_schedule(const mgx_schedule &sc, int64_t K, int64_t M, int64_t N, mgx_limb *C, int64_t Crs, const mgx_limb *A,
                 int64_t Ars, const mgx_limb *B, int64_t Brs, int64_t cap, int64_t ba, int64_t bb, hipStream_t s)
{
    if (sc.route == MGX_ROUTE_CRT) {
        const mgx_crt_plan *pl = nullptr;
        MGX_TRY(cached_plan(M, ba, bb, &pl, sc.slices));
        const int rows = sc.gemm_rows > 0 ? sc.gemm_rows : (int)param(MGX_PARAM_GEMM_ROWS);
        int st = crt_matmul_with(pl, K, M, N, C, Crs, A, Ars, B, Brs, cap, s, rows);
        release_plan(pl);
        return st;
    }
    const int64_t opbits = std::max(ba, bb);
    if (sc.route == MGX_ROUTE_PLAIN)
        MGX_TRY(plain_mul(false, false, K, M, N, C, Crs, A, Ars, B, Brs, cap, (opbits + 31) / 32, s, sc.toom));
    else
        MGX_TRY(squarify_with(MGX_REC_WINOGRAD, K, M, N, C, Crs, A, Ars, B, Brs, cap, opbits, 2, sc.depth, s, sc.toom));
    return check_device_error(s);
}

// per matrix: bit-length histogram (6 bins), sum of bits, max bits, sum of 2048-bit blocks
struct Stats {
    unsigned long long v[9];
};

int costmodel_view_locked(const float **params, const float **norm, int *D, int *H, int *O);

// Statistics of A (K x M) and B (M x N) with one launch and one
// synchronization; with `schedule`, also the installed cost model's choice
// (-1 when no model with one output per schedule is installed).
int scan_stats_pair(int64_t K, int64_t M, int64_t N, const mgx_limb *A, int64_t Ars, const mgx_limb *B, int64_t Brs,
                    int64_t cap, Stats &sa_out, Stats &sb_out, hipStream_t s, int *schedule = nullptr)
{
    unsigned long long *d = nullptr;
    const size_t bytes = (2 * 9 + 1) * STAT_PITCH * 8;
    MGX_TRY(scratch_alloc((void **)&d, bytes, s));
    MGX_HIP(hipMemsetAsync(d, 0, bytes, s));
    StatArgs sa{};
    const int64_t rows[2] = {K, M}, cols[2] = {M, N}, rs[2] = {Ars, Brs};
    const mgx_limb *X[2] = {A, B};
    int grid[2] = {0, 0};
    for (int m = 0; m < 2; m++) {
        const int64_t n = rows[m] * cols[m];
        sa.X[m] = (const uint64_t *)X[m];
        sa.rs[m] = rs[m];
        sa.rows[m] = rows[m];
        sa.cols[m] = cols[m];
        grid[m] = n > 0 ? (int)std::min<int64_t>(ceil_div(n, 1024), 2048) : 0;
    }
    sa.nblk0 = grid[0];
    if (grid[0] + grid[1] > 0)
        hipLaunchKernelGGL(feature_scan2_kernel, dim3(grid[0] + grid[1]), dim3(256), 0, s, sa, cap + 1, d);
    MGX_HIP(hipGetLastError());
    bool inferred = false;
    if (schedule) {
        const float *params, *norm;
        int D, H, O;
        if (costmodel_view_locked(&params, &norm, &D, &H, &O) == 0 && O == MGX_SCHEDULES && D == MGX_FEATURES) {
            DispatchTab tab;
            for (int i = 0; i < MGX_SCHEDULES; i++) {
                tab.plain[i] = SCHEDULES[i].route != MGX_ROUTE_CRT;
                tab.depth[i] = SCHEDULES[i].depth;
            }
            hipLaunchKernelGGL(dispatch_kernel, dim3(1), dim3(256), 0, s, d, K, M, N, params, norm, D, H, O, tab);
            MGX_HIP(hipGetLastError());
            inferred = true;
        }
    }
    unsigned long long h[(2 * 9 + 1) * STAT_PITCH];
    MGX_HIP(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, s));
    MGX_TRY(scratch_free(d, s));
    MGX_HIP(hipStreamSynchronize(s));
    for (int c = 0; c < 9; c++) {
        sa_out.v[c] = h[STAT_PITCH * c];
        sb_out.v[c] = h[STAT_PITCH * (9 + c)];
    }
    if (schedule) *schedule = inferred ? (int)h[STAT_PITCH * 18] : -1;
    return MGX_OK;
}

int features_of(int64_t K, int64_t M, int64_t N, const Stats &a, const Stats &b, float *f)
{
    int64_t hist[6];
    for (int i = 0; i < 6; i++) hist[i] = (int64_t)(a.v[i] + b.v[i]);
    double cnt = (double)(K * M + M * N);
    if (cnt <= 0) cnt = 1;
    return mgx_tune_features(K, M, N, hist, (double)(a.v[6] + b.v[6]) / cnt, (int64_t)std::max(a.v[7], b.v[7]),
                             (double)(a.v[8] + b.v[8]) / cnt, f);
}

int choose_schedule(int64_t K, int64_t M, int64_t N, const float *f, int *schedule, hipStream_t s);

mgx_costmodel *g_model = nullptr;
std::mutex g_mu;
std::map<std::vector<int>, int> g_tuned;   // exhaustive schedule choices by feature bucket

// the installed model's device view, under the dispatch lock
int costmodel_view_locked(const float **params, const float **norm, int *D, int *H, int *O)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_model) return 1;
    return costmodel_view(g_model, params, norm, D, H, O);
}

thread_local int g_last_schedule = -1;   // the schedule this thread's last MGX_ROUTE_AUTO call ran

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
    MGX_TRY(scan_stats_pair(K, M, N, A, Ars, B, Brs, cap, a, b, nullptr));
    return features_of(K, M, N, a, b, f);
}

int mgx_schedule_get(int index, mgx_schedule *out)
{
    if (index < 0 || index >= MGX_SCHEDULES || !out) return fail(MGX_ERR_ARG, "no schedule %d", index);
    *out = SCHEDULES[index];
    return MGX_OK;
}

// One pass over the inputs gives both the dispatch's features and the bit
// bounds whichever schedule runs needs.
static int matmul_sched(int index, const mgx_schedule *given, int64_t K, int64_t M, int64_t N, mgx_limb *C,
                        int64_t Crs, const mgx_limb *A, int64_t Ars, const mgx_limb *B, int64_t Brs, int64_t cap,
                        void *stream)
{
    MGX_TRY(ensure_init());
    if (K < 0 || M < 0 || N < 0 || cap < 1) return fail(MGX_ERR_SHAPE, "bad shape");
    hipStream_t s = as_stream(stream);
    if (K == 0 || N == 0) return MGX_OK;
    if (M == 0) {
        MGX_HIP(hipMemset2DAsync(C, (size_t)(Crs * (cap + 1) * 8), 0, (size_t)(N * (cap + 1) * 8), (size_t)K, s));
        return check_device_error(s);
    }
    Stats sa, sb;
    int inferred = -1;   // the model's choice, when it ran on the device beside the scan
    MGX_TRY(scan_stats_pair(K, M, N, A, Ars, B, Brs, cap, sa, sb, s,
                            !given && index == MGX_ROUTE_AUTO ? &inferred : nullptr));
    const int64_t ba = std::max<int64_t>((int64_t)sa.v[7], 1), bb = std::max<int64_t>((int64_t)sb.v[7], 1);
    mgx_schedule sc;
    if (given) {
        sc = *given;
    } else {
        if (index == MGX_ROUTE_AUTO && inferred >= 0) {
            index = inferred;
        } else if (index == MGX_ROUTE_AUTO) {
            float f[MGX_FEATURES];
            MGX_TRY(features_of(K, M, N, sa, sb, f));
            MGX_TRY(choose_schedule(K, M, N, f, &index, s));
        }
        if (index < 0 || index >= MGX_SCHEDULES) return fail(MGX_ERR_ARG, "no schedule %d", index);
        sc = SCHEDULES[index];
    }
    if (!schedule_feasible(sc, std::max(ba, bb)))
        return fail(MGX_ERR_UNSUPPORTED, "operands of %lld bits exceed the plain kernel's 512-word limit",
                    (long long)std::max(ba, bb));
    if (!given) g_last_schedule = index;
    return run_schedule(sc, K, M, N, C, Crs, A, Ars, B, Brs, cap, ba, bb, s);
}

int mgx_matmul_schedule(int index, int64_t K, int64_t M, int64_t N, mgx_limb *C, int64_t Crs, const mgx_limb *A,
                        int64_t Ars, const mgx_limb *B, int64_t Brs, int64_t cap, void *stream)
{
    return matmul_sched(index, nullptr, K, M, N, C, Crs, A, Ars, B, Brs, cap, stream);
}

int mgx_matmul_route(int route, int64_t K, int64_t M, int64_t N, mgx_limb *C, int64_t Crs, const mgx_limb *A,
                     int64_t Ars, const mgx_limb *B, int64_t Brs, int64_t cap, void *stream)
{
    if (route == MGX_ROUTE_AUTO) return matmul_sched(MGX_ROUTE_AUTO, nullptr, K, M, N, C, Crs, A, Ars, B, Brs, cap, stream);
    if (route < 0 || route >= MGX_ROUTES) return fail(MGX_ERR_ARG, "unknown route %d", route);
    const mgx_schedule sc = route_schedule(route);
    return matmul_sched(0, &sc, K, M, N, C, Crs, A, Ars, B, Brs, cap, stream);
}

int mgx_tune_time_schedules(int64_t K, int64_t M, int64_t N, mgx_limb *C, const mgx_limb *A, const mgx_limb *B,
                            int64_t cap, int reps, double *times, unsigned *estimated)
{
    MGX_TRY(ensure_init());
    if (reps < 1) reps = 1;
    unsigned est = 0;
    // median wall time of reps calls after one warm-up (plans, scratch);
    // 1e30 when the schedule cannot take the product
    auto run = [&](int i, int64_t rows, double *out) -> int {
        int st0 = mgx_matmul_schedule(i, rows, M, N, C, N, A, M, B, N, cap, nullptr);
        if (st0 == MGX_ERR_UNSUPPORTED || st0 == MGX_ERR_CAPACITY) {
            *out = 1e30;
            return MGX_OK;
        }
        MGX_TRY(st0);
        std::vector<double> t;
        for (int r = 0; r < reps; r++) {
            const double t0 = now();
            MGX_TRY(mgx_matmul_schedule(i, rows, M, N, C, N, A, M, B, N, cap, nullptr));
            t.push_back(now() - t0);
        }
        std::sort(t.begin(), t.end());
        *out = t[t.size() / 2];
        return MGX_OK;
    };
    double best_crt = 1e30;
    for (int i = 0; i < MGX_SCHEDULES; i++)
        if (SCHEDULES[i].route == MGX_ROUTE_CRT) {
            MGX_TRY(run(i, K, &times[i]));
            best_crt = std::min(best_crt, times[i]);
        }
    // The flat kernel's time on one block of rows estimates the plain
    // schedules; when even two Winograd levels could not come within 2x of
    // the best multi-modular schedule they are not run but extrapolated
    // (flagged in *estimated): the flat time scaled by the rows, times 7/8
    // per Winograd level.
    const int64_t rows = std::min<int64_t>(K, 8);
    double flat[2] = {1e30, 1e30};   // toom on, off
    int flat_idx[2] = {-1, -1};
    for (int i = 0; i < MGX_SCHEDULES; i++)
        if (SCHEDULES[i].route == MGX_ROUTE_PLAIN) {
            const int k = SCHEDULES[i].toom ? 0 : 1;
            flat_idx[k] = i;
            MGX_TRY(run(i, rows, &flat[k]));
        }
    const double scale = (double)K / (double)rows;
    const bool hopeless = K > rows && std::min(flat[0], flat[1]) * scale * 0.7 > 2.0 * best_crt;
    for (int i = 0; i < MGX_SCHEDULES; i++) {
        const mgx_schedule &sc = SCHEDULES[i];
        if (sc.route == MGX_ROUTE_CRT) continue;
        const double base = flat[sc.toom ? 0 : 1];
        if (base >= 1e29) {
            times[i] = 1e30;   // the plain kernel cannot take these operands
        } else if (hopeless) {
            times[i] = base * scale * std::pow(0.875, sc.depth);
            est |= 1u << i;
        } else if (sc.route == MGX_ROUTE_PLAIN && rows == K) {
            times[i] = base;
        } else {
            MGX_TRY(run(i, K, &times[i]));
        }
    }
    if (estimated) *estimated = est;
    return MGX_OK;
}

int mgx_tune_time_routes(int64_t K, int64_t M, int64_t N, mgx_limb *C, const mgx_limb *A, const mgx_limb *B,
                         int64_t cap, int reps, double *times, unsigned *estimated)
{
    double t[MGX_SCHEDULES];
    unsigned est = 0;
    MGX_TRY(mgx_tune_time_schedules(K, M, N, C, A, B, cap, reps, t, &est));
    // each route at its default schedule: CRT two digits 128-row tiles,
    // plain and Winograd with the Karatsuba cutoff of two blocks
    const int pick[MGX_ROUTES] = {0, 3, 5, 6};
    unsigned e = 0;
    for (int r = 0; r < MGX_ROUTES; r++) {
        times[r] = t[pick[r]];
        if (est & (1u << pick[r])) e |= 1u << r;
    }
    if (estimated) *estimated = e;
    return MGX_OK;
}

int mgx_dispatch_last(void) { return g_last_schedule; }

int mgx_dispatch_set_model(mgx_costmodel *cm)
{
    std::lock_guard<std::mutex> lk(g_mu);
    g_model = cm;
    return MGX_OK;
}

int mgx_dispatch_choose_schedule(int64_t K, int64_t M, int64_t N, const mgx_limb *A, int64_t Ars, const mgx_limb *B,
                                 int64_t Brs, int64_t cap, int *schedule)
{
    float f[MGX_FEATURES];
    MGX_TRY(mgx_tune_features_device(K, M, N, A, Ars, B, Brs, cap, f));
    return choose_schedule(K, M, N, f, schedule, nullptr);
}

int mgx_dispatch_choose(int64_t K, int64_t M, int64_t N, const mgx_limb *A, int64_t Ars, const mgx_limb *B,
                        int64_t Brs, int64_t cap, int *route)
{
    int sc = 0;
    MGX_TRY(mgx_dispatch_choose_schedule(K, M, N, A, Ars, B, Brs, cap, &sc));
    *route = SCHEDULES[sc].route;
    return MGX_OK;
}

}  // extern "C"

namespace {
// The schedule for a product's features: the cost model's prediction when a
// model is installed (one output per schedule; a model with one output per
// route picks among the routes' default schedules), else the exhaustive
// autotuner's remembered choice for the feature bucket.  Schedules that
// cannot take the operands are never chosen.
int choose_schedule(int64_t K, int64_t M, int64_t N, const float *f, int *schedule, hipStream_t s_pred)
{
    mgx_costmodel *cm;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        cm = g_model;
    }
    const int64_t maxbits = (int64_t)std::llround(std::exp2((double)f[5]));
    if (cm) {
        const int O = (int)mgx_costmodel_outputs(cm);
        float y[64];
        if (O != MGX_SCHEDULES && O != MGX_ROUTES) return fail(MGX_ERR_ARG, "cost model has %d outputs", O);
        MGX_TRY(costmodel_predict1(cm, f, y, s_pred));
        const int route_pick[MGX_ROUTES] = {0, 3, 5, 6};
        int best = 0;
        float by = 1e30f;
        for (int o = 0; o < O; o++) {
            const int sc = O == MGX_SCHEDULES ? o : route_pick[o];
            if (schedule_feasible(SCHEDULES[sc], maxbits) && y[o] < by) {
                by = y[o];
                best = sc;
            }
        }
        *schedule = best;
        return MGX_OK;
    }
    // fallback: the exhaustive autotuner's choice for this feature bucket,
    // timed once on a same-shaped seeded problem and remembered
    std::vector<int> key;
    for (int i = 0; i < MGX_FEATURES; i++) key.push_back((int)std::lround(f[i] * (i >= 7 ? 8 : 2)));
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_tuned.find(key);
        if (it != g_tuned.end()) { *schedule = it->second; return MGX_OK; }
    }
    const int64_t tbits = maxbits + 1;
    const int64_t tcap = (2 * tbits + 64 + 63) / 64 + 2;
    mgx_limb *tA = nullptr, *tB = nullptr, *tC = nullptr;
    hipStream_t s = nullptr;
    int st = scratch_alloc((void **)&tA, (size_t)(K * M * (tcap + 1) * 8), s);
    if (st == MGX_OK) st = scratch_alloc((void **)&tB, (size_t)(M * N * (tcap + 1) * 8), s);
    if (st == MGX_OK) st = scratch_alloc((void **)&tC, (size_t)(K * N * (tcap + 1) * 8), s);
    double times[MGX_SCHEDULES];
    if (st == MGX_OK) st = mgx_random_matrix_device(tA, M, K, M, tcap, tbits, tbits, 0, 1, s);
    if (st == MGX_OK) st = mgx_random_matrix_device(tB, N, M, N, tcap, tbits, tbits, 0, 2, s);
    if (st == MGX_OK) st = mgx_tune_time_schedules(K, M, N, tC, tA, tB, tcap, 1, times, nullptr);
    scratch_free(tA, s);
    scratch_free(tB, s);
    scratch_free(tC, s);
    MGX_TRY(st);
    int best = 0;
    for (int i = 1; i < MGX_SCHEDULES; i++)
        if (times[i] < times[best]) best = i;
    std::lock_guard<std::mutex> lk(g_mu);
    g_tuned[key] = best;
    *schedule = best;
    return MGX_OK;
}
}  // namespace
