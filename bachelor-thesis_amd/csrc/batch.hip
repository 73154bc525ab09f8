// batch.hip -- many independent exact products as one captured hipGraph.
//
// Replaces the reference's host job scheduler (mat-gpu/job.cu:32-336: a host
// DAG on up to four streams with a polled chunk pool, never called): every
// product of the batch is planned once on the host, then each stage of each
// route is ONE persistent launch whose workgroups pull tiles from a device
// work queue (an atomic counter over the concatenated tile ranges of all
// products of that stage).  Stages are ordered by the launches themselves;
// the whole sequence, counter reset included (a fill kernel, no memset
// nodes), is captured once and replayed with hipGraphLaunch.
#include "crt_kernels.h"
#include "plain_kernels.h"
#include "queue.h"
#include "../../include/mgx_sched.h"
#include <algorithm>
#include <vector>

namespace mgx {

struct ReduceJob {
    const PrimeConst *pc;
    const int8_t *cst, *cstb;   // cstb: the table in 1 KB blocks (tiled reduce)
    int P, nrb, kb, xbits, transpose, resident;
    const uint64_t *X;
    int64_t Xrs, W, rows, K, kpad, nent, plane;
    int8_t *R;
};
struct GemmJob {
    const PrimeConst *pc;
    const int8_t *RA, *RB;
    int64_t planeA, planeB, kpad, rowsC, colsC, total, ystride;
    int ntm, ntn, P;
    uint32_t *Y;
};
struct ReconJob {
    const PrimeConst *pc;
    const int8_t *Mc;
    const uint32_t *M32, *Y;
    int P, Kr, nrb, nw;
    int64_t ystride, nent, colsC, Crs, W, cap;
    uint64_t *C;
    uint32_t *Wv;  // carried 32-bit words, nw x ystride
    int32_t *Cb;   // row-block carries, nrb x ystride
    int64_t *G;    // fixed-point gamma partial sums, nrb x ystride
    int gshift;    // their fraction bits
    int gw;        // one-pass reconstruct: words of gamma rows above nw in Mc (0 or 2)
};
struct PlainJob {
    int64_t K, M, N, Crs, Ars, Brs, W, cap;
    uint64_t *C;
    const uint64_t *A, *B;
    int toom;   // Toom-22 cutoff of this product's schedule
};
// C = A +- B over an H x Wd block (one Winograd block addition)
struct AorsJob {
    int64_t Wd, H, Crs, Ars, Brs, W, cap;
    uint64_t *C;
    const uint64_t *A, *B;
    int sub;
};

struct Queue {
    unsigned long long *counter;   // QRANGES counters, QPITCH apart
    const int64_t *prefix;         // njobs + 1 tile offsets
    int njobs;
};

// The workgroup's claiming state (used by thread 0): the range it claims from
// and how often it has moved on.
struct Claim {
    int cur, moved;
    __device__ Claim() : cur((int)(blockIdx.x % QRANGES)), moved(0) {}
};

__device__ __forceinline__ bool next_item(const Queue &q, Claim &cl, int &job, int64_t &tile)
{
    __shared__ int64_t s_item;
    if (threadIdx.x == 0) {
        const int64_t total = q.prefix[q.njobs];
        int64_t it;
        do {
            it = queue_claim_once([&](int x) { return atomicAdd(q.counter + QPITCH * x, 1ull); }, total, cl.cur,
                                  cl.moved);
        } while (it == -1);
        s_item = it;
    }
    __syncthreads();
    const int64_t it = s_item;
    __syncthreads();
    return it >= 0 && queue_map(q.prefix, q.njobs, it, &job, &tile);
}

constexpr int NSLOTS = 40;   // queues of a batch (counter slots)
__global__ void reset_counters_kernel(unsigned long long *c, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) c[i] = 0;
}

// Entry-resident and generic reduce jobs run in kernels of their own, so
// that each gets its own occupancy (the generic tile needs few registers and
// hides its memory latency with many waves).  Entry-resident jobs run in
// two kernels, entries up to 256 bytes and up to 512 bytes, each serving
// every byte bound below its own (the constant columns past kb are zero):
// each inlined tile function adds its own LDS arrays and registers.
template <int S, int KS>
__global__ __launch_bounds__(256) void batch_reduce_res_kernel(Queue q, const ReduceJob *jobs, unsigned *err)
{
    int j;
    int64_t t;
    Claim cl;
    while (next_item(q, cl, j, t)) {
        const ReduceJob &b = jobs[j];
        reduce_tile_res<S, KS>(t, b.pc, b.cst, b.P, b.nrb, b.kb, b.xbits, b.X, b.Xrs, b.W, b.transpose, b.rows, b.K,
                               b.kpad,
                               b.R, b.plane, err);
        __syncthreads();
    }
}

// reduce tile kind for a byte bound: 0 entry-resident up to 256 bytes, 1 generic, 2 entry-resident up to 512
inline int reduce_kind(int64_t kb) { return kb <= 256 ? 0 : kb <= RES_KB ? 2 : 1; }

// two-digit plans with entries over 512 bytes: the reduce as a tiled MFMA GEMM
__global__ __launch_bounds__(256, 2) void batch_reduce_gemm_kernel(Queue q, const ReduceJob *jobs, unsigned *err)
{
    int j;
    int64_t t;
    Claim cl;
    while (next_item(q, cl, j, t)) {
        const ReduceJob &b = jobs[j];
        reduce_gemm_tile(t, b.pc, b.cstb, b.P, b.kb, b.xbits, b.X, b.Xrs, b.W, b.transpose, b.rows, b.K, b.kpad, b.nent,
                         b.R, b.plane, err);
    }
}

template <int S>
__global__ __launch_bounds__(256) void batch_reduce_gen_kernel(Queue q, const ReduceJob *jobs, unsigned *err)
{
    int j;
    int64_t t;
    Claim cl;
    while (next_item(q, cl, j, t)) {
        const ReduceJob &b = jobs[j];
        reduce_tile<S>(t, b.pc, b.cst, b.P, b.nrb, b.kb, b.xbits, b.X, b.Xrs, b.W, b.transpose, b.rows, b.K, b.kpad,
                       b.nent,
                       b.R, b.plane, err);
        __syncthreads();
    }
}

template <bool CHUNKED>
__global__ __launch_bounds__(256, 1) void batch_gemm_kernel(Queue q, const GemmJob *jobs)
{
    int j;
    int64_t t;
    Claim cl;
    while (next_item(q, cl, j, t)) {
        const GemmJob &g = jobs[j];
        gemm_tile<CHUNKED>(t, g.pc, g.RA, g.RB, g.planeA, g.planeB, g.kpad, g.rowsC, g.colsC, g.ntm, g.ntn, g.total,
                           g.Y, g.ystride);
        __syncthreads();
    }
}

template <bool CHUNKED>
__global__ __launch_bounds__(256, 2) void batch_gemm2_kernel(Queue q, const GemmJob *jobs)
{
    int j;
    int64_t t;
    Claim cl;
    while (next_item(q, cl, j, t)) {
        const GemmJob &g = jobs[j];
        gemm_tile2<CHUNKED, 2, 2>(t, g.pc, g.P, g.RA, g.RB, g.planeA, g.planeB, g.kpad, g.rowsC, g.colsC, g.ntm,
                                  g.ntn, g.Y, g.ystride);
        __syncthreads();
    }
}

template <int S>
__global__ __launch_bounds__(256, 2) void batch_recon_gemm_kernel(Queue q, const ReconJob *jobs)
{
    int j;
    int64_t t;
    Claim cl;
    while (next_item(q, cl, j, t)) {
        const ReconJob &r = jobs[j];
        recon_gemm_tile<S, false>(t, r.pc, r.P, r.Mc, r.Kr, r.nrb, r.nw, r.Y, r.ystride, r.nent, r.Wv, r.Cb, r.G);
        __syncthreads();
    }
}

// two-digit products of at most 9 row blocks: the one-pass reconstruct, one item per 256 entries
__global__ __launch_bounds__(256, 2) void batch_recon_fused_kernel(Queue q, const ReconJob *jobs, unsigned *err)
{
    int j;
    int64_t t;
    Claim cl;
    while (next_item(q, cl, j, t)) {
        const ReconJob &r = jobs[j];
        recon_fused_block(t, r.pc, r.P, r.Mc, r.Kr, r.nrb, r.nw, r.gw, r.Y, r.ystride, r.nent, r.M32, r.colsC, r.C,
                          r.Crs, r.W, r.cap, err);
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void batch_recon_carry_kernel(Queue q, const ReconJob *jobs, unsigned *err)
{
    int j;
    int64_t t;
    Claim cl;
    while (next_item(q, cl, j, t)) {
        const ReconJob &r = jobs[j];
        recon_carry_block(t, r.Wv, r.Cb, r.ystride, r.nw, r.nrb, r.G, r.gshift, r.M32, r.nent, r.colsC, r.C, r.Crs, r.W,
                          r.cap, err);
        __syncthreads();
    }
}

template <int NB>
__global__ __launch_bounds__(256) void batch_plain_kernel(Queue q, const PlainJob *jobs, unsigned *err)
{
    int j;
    int64_t t;
    Claim cl;
    while (next_item(q, cl, j, t)) {
        const PlainJob &p = jobs[j];
        plain_entries<NB, false, false>(t, p.K, p.M, p.N, p.C, p.Crs, p.A, p.Ars, p.B, p.Brs, p.W, p.cap, p.toom, err);
    }
}

// tile = four entries, one per wavefront
__global__ __launch_bounds__(256) void batch_aors_kernel(Queue q, const AorsJob *jobs, unsigned *err)
{
    int j;
    int64_t t;
    Claim cl;
    while (next_item(q, cl, j, t)) {
        const AorsJob &a = jobs[j];
        aors_entry<false>(4 * t + (threadIdx.x >> 6), a.Wd, a.H, a.C, a.Crs, a.A, a.Ars, a.B, a.Brs, a.W, a.cap,
                          a.sub, err);
    }
}

}  // namespace mgx

using namespace mgx;

struct BatchQueue {
    int kind;        // mgx_batch_queue_info kinds
    int slot;        // counter index
    int64_t items;   // tiles over all jobs of the queue
    int grid;        // workgroups of its persistent launch
};
struct mgx_batch {
    std::vector<int> routes;
    std::vector<BatchQueue> queues;   // in launch order
    unsigned long long *counters = nullptr;
    std::vector<const mgx_crt_plan *> plans;   // pinned in the plan cache while the batch lives
    void *scratch = nullptr;       // residues, Y, column sums, job tables, prefixes, counters
    int64_t scratch_bytes = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

const int NB_CLASSES[] = {1, 2, 3, 4, 6, 8};
const int NCLASS = 6;

struct Layout {
    int64_t off = 0;
    int64_t take(int64_t bytes)
    {
        int64_t o = off;
        off += round_up(std::max<int64_t>(bytes, 1), 256);
        return o;
    }
};

template <int NB>
void launch_plain(int grid, hipStream_t s, Queue q, const PlainJob *jobs, unsigned *err)
{
    hipLaunchKernelGGL(batch_plain_kernel<NB>, dim3(grid), dim3(256), 0, s, q, jobs, err);
}

// plain kernel class (blocks of 64 words) for operands of opbits bits
int plain_class(int64_t opbits)
{
    const int64_t words = 2 * ((opbits + 63) / 64);
    int k = 0;
    while (k < NCLASS - 1 && 64 * NB_CLASSES[k] < words) k++;
    return k;
}

// Winograd levels of a plain-routed product as batch jobs: the seven
// products and the block additions of mat-gpu/mat-defines.h:144-164 (the
// recursion rec.cpp runs one launch at a time) regrouped into stages whose
// jobs are independent, so that each stage is one queue for the whole batch:
//   pre-adds   S1 = A21 + A22, S3 = A11 - A21, T1 = B12 - B11, T3 = B22 - B12
//              S2 = S1 - A11, T2 = B22 - T1
//              S4 = A12 - S2, T4 = T2 - B21
//   products   P1 = A11 B11, P2 = A12 B21, P3 = S4 B22, P4 = A22 T4,
//              P5 = S1 T1, P6 = S2 T2, P7 = S3 T3
//   post-adds  C11 = P1 + P2, U2 = P1 + P6 (into P6)
//              U3 = U2 + P7 (into P7), U4 = U2 + P5 (into P1)
//              C12 = U4 + P3, C21 = U3 - P4, C22 = U3 + P5
// A level splits the even part of K x M x N; an odd last row of C and an odd
// last column are plain products of their own, and an odd inner dimension
// leaves the rank-one term A[:, M-1] B[M-1, :], a plain product into a
// temporary added to C after the level's post-adds (the fix-up).  Stage
// numbers: pre-adds 3L + {0, 1, 2} for level L = 0, 1; then, inner level
// first, the post-adds and the fix-up at 6 + 4 (1 - L) + {0, 1, 2, 3}.  The
// products of the last level run in the plain kernel queues between.
constexpr int WSTAGES = 14;
struct WinoGen {
    int depth, toom;
    int64_t W, cap, bits;   // entry words, limbs, operand bit bound at level 0
    uintptr_t tmp;          // the product's temporaries (0 while sizing)
    int64_t used = 0;
    std::vector<AorsJob> *stage;                        // WSTAGES lists
    std::vector<std::pair<PlainJob, int64_t>> leaves;   // products and their operand bits

    uint64_t *take(int64_t entries)
    {
        uint64_t *p = (uint64_t *)(tmp + used);
        used += round_up(std::max<int64_t>(entries, 1) * W * 8, 256);
        return p;
    }
    template <class T>
    T *at(T *X, int64_t rs, int64_t y, int64_t x) const { return X + (y * rs + x) * W; }
    void mul(int64_t K, int64_t M, int64_t N, uint64_t *C, int64_t Crs, const uint64_t *A, int64_t Ars,
             const uint64_t *B, int64_t Brs, int64_t opbits)
    {
        PlainJob q;
        q.K = K; q.M = M; q.N = N; q.Crs = Crs; q.Ars = Ars; q.Brs = Brs; q.W = W; q.cap = cap;
        q.C = C; q.A = A; q.B = B; q.toom = toom;
        leaves.push_back({q, opbits});
    }
    void add(int st, int64_t H, int64_t Wd, uint64_t *C, int64_t Crs, const uint64_t *A, int64_t Ars,
             const uint64_t *B, int64_t Brs, bool sub)
    {
        AorsJob a;
        a.Wd = Wd; a.H = H; a.Crs = Crs; a.Ars = Ars; a.Brs = Brs; a.W = W; a.cap = cap;
        a.C = C; a.A = A; a.B = B; a.sub = sub;
        stage[st].push_back(a);
    }
    void gen(int level, int64_t K, int64_t M, int64_t N, uint64_t *C, int64_t Crs, const uint64_t *A, int64_t Ars,
             const uint64_t *B, int64_t Brs)
    {
        const int64_t ob = bits + 2 * level;   // operand sums grow by two bits a level
        if (level == depth || K < 2 || M < 2 || N < 2) {
            mul(K, M, N, C, Crs, A, Ars, B, Brs, ob);
            return;
        }
        const int64_t K2 = K & ~1ll, M2 = M & ~1ll, N2 = N & ~1ll, h = K2 / 2, m = M2 / 2, w = N2 / 2;
        const int pre = 3 * level, post = 6 + 4 * (1 - level);
        if (K & 1) mul(1, M, N, at(C, Crs, K2, 0), Crs, at(A, Ars, K2, 0), Ars, B, Brs, ob);
        if (N & 1) mul(K2, M, 1, at(C, Crs, 0, N2), Crs, A, Ars, at(B, Brs, 0, N2), Brs, ob);
        if (M & 1) {
            uint64_t *R = take(K2 * N2);
            mul(K2, 1, N2, R, N2, at(A, Ars, 0, M2), Ars, at(B, Brs, M2, 0), Brs, ob);
            add(post + 3, K2, N2, C, Crs, C, Crs, R, N2, false);
        }
        const uint64_t *A11 = A, *A12 = at(A, Ars, 0, m), *A21 = at(A, Ars, h, 0), *A22 = at(A, Ars, h, m);
        const uint64_t *B11 = B, *B12 = at(B, Brs, 0, w), *B21 = at(B, Brs, m, 0), *B22 = at(B, Brs, m, w);
        uint64_t *S[4], *T[4], *P[7];
        for (auto &x : S) x = take(h * m);
        for (auto &x : T) x = take(m * w);
        for (auto &x : P) x = take(h * w);
        add(pre, h, m, S[0], m, A21, Ars, A22, Ars, false);
        add(pre, h, m, S[2], m, A11, Ars, A21, Ars, true);
        add(pre, m, w, T[0], w, B12, Brs, B11, Brs, true);
        add(pre, m, w, T[2], w, B22, Brs, B12, Brs, true);
        add(pre + 1, h, m, S[1], m, S[0], m, A11, Ars, true);
        add(pre + 1, m, w, T[1], w, B22, Brs, T[0], w, true);
        add(pre + 2, h, m, S[3], m, A12, Ars, S[1], m, true);
        add(pre + 2, m, w, T[3], w, T[1], w, B21, Brs, true);
        gen(level + 1, h, m, w, P[0], w, A11, Ars, B11, Brs);
        gen(level + 1, h, m, w, P[1], w, A12, Ars, B21, Brs);
        gen(level + 1, h, m, w, P[2], w, S[3], m, B22, Brs);
        gen(level + 1, h, m, w, P[3], w, A22, Ars, T[3], w);
        gen(level + 1, h, m, w, P[4], w, S[0], m, T[0], w);
        gen(level + 1, h, m, w, P[5], w, S[1], m, T[1], w);
        gen(level + 1, h, m, w, P[6], w, S[2], m, T[2], w);
        uint64_t *C11 = C, *C12 = at(C, Crs, 0, w), *C21 = at(C, Crs, h, 0), *C22 = at(C, Crs, h, w);
        add(post, h, w, C11, Crs, P[0], w, P[1], w, false);
        add(post, h, w, P[5], w, P[0], w, P[5], w, false);
        add(post + 1, h, w, P[6], w, P[5], w, P[6], w, false);
        add(post + 1, h, w, P[0], w, P[5], w, P[4], w, false);
        add(post + 2, h, w, C12, Crs, P[0], w, P[2], w, false);
        add(post + 2, h, w, C21, Crs, P[6], w, P[3], w, true);
        add(post + 2, h, w, C22, Crs, P[6], w, P[4], w, false);
    }
};

}  // namespace

extern "C" {

int mgx_batch_create(mgx_batch **out, const mgx_product *prods, int64_t count, void *stream)
{
    MGX_TRY(ensure_init());
    if (!out || (count && !prods)) return fail(MGX_ERR_ARG, "null argument");
    hipStream_t s = as_stream(stream);
    auto *b = new mgx_batch();
    b->routes.assign(count, MGX_ROUTE_CRT);
    int st = MGX_OK;
    std::vector<int64_t> bitsA(count), bitsB(count);
    std::vector<int> slices(count, -1), tooms(count, (int)param(MGX_PARAM_TOOM22)), wdepth(count, 0);
    std::vector<PlanView> views(count);
    for (int64_t i = 0; i < count && st == MGX_OK; i++) {
        const mgx_product &p = prods[i];
        if (p.K < 0 || p.M < 1 || p.N < 0 || p.cap < 1 || !p.A || !p.B || !p.C) {
            st = fail(MGX_ERR_SHAPE, "product %lld: bad shape or pointer", (long long)i);
            break;
        }
        bitsA[i] = p.bitsA;
        bitsB[i] = p.bitsB;
        if (bitsA[i] <= 0) st = mgx_crt_scan_bits(p.K, p.M, p.A, p.Ars, p.cap, &bitsA[i], stream);
        if (st == MGX_OK && bitsB[i] <= 0) st = mgx_crt_scan_bits(p.M, p.N, p.B, p.Brs, p.cap, &bitsB[i], stream);
        bitsA[i] = std::max<int64_t>(bitsA[i], 1);
        bitsB[i] = std::max<int64_t>(bitsB[i], 1);
        int r = p.route;
        if (st == MGX_OK && r == MGX_ROUTE_AUTO) {
            int sci = 0;
            st = mgx_dispatch_choose_schedule(p.K, p.M, p.N, p.A, p.Ars, p.B, p.Brs, p.cap, &sci);
            mgx_schedule sc{};
            if (st == MGX_OK) st = mgx_schedule_get(sci, &sc);
            r = sc.route;
            slices[i] = sc.slices;   // the batch's residue GEMM has one tile shape
            tooms[i] = sc.toom;
        }
        if (r == MGX_ROUTE_WINOGRAD1 || r == MGX_ROUTE_WINOGRAD2) {
            // fewer levels when the leaf operands (two bits more a level)
            // outgrow the plain kernel or the temporaries (at most
            // M 2^(bitsA + bitsB + 3 levels + 2)) outgrow the capacity
            const int64_t mb = std::max(bitsA[i], bitsB[i]), lgM = 64 - __builtin_clzll((uint64_t)std::max<int64_t>(p.M, 1));
            int d = r == MGX_ROUTE_WINOGRAD1 ? 1 : 2;
            while (d > 0 && ((mb + 2 * d + 63) / 64 * 2 > 64 * 8 || 64 * p.cap < bitsA[i] + bitsB[i] + 3 * d + 3 + lgM)) d--;
            r = d == 0 ? MGX_ROUTE_PLAIN : d == 1 ? MGX_ROUTE_WINOGRAD1 : MGX_ROUTE_WINOGRAD2;
            wdepth[i] = d;
        }
        if (r == MGX_ROUTE_PLAIN && (std::max(bitsA[i], bitsB[i]) + 63) / 64 * 2 > 64 * 8)
            r = MGX_ROUTE_CRT;   // beyond the plain kernel's widest class
        b->routes[i] = r;
        if (st == MGX_OK && r == MGX_ROUTE_CRT) {
            const mgx_crt_plan *pl = nullptr;
            st = cached_plan(p.M, bitsA[i], bitsB[i], &pl, slices[i]);
            if (st == MGX_OK) {
                b->plans.push_back(pl);   // the graph reads its tables on every replay
                st = plan_view(pl, &views[i]);
            }
        }
    }
    if (st != MGX_OK) { mgx_batch_destroy(b); return st; }

    // ---- job tables (sizes first, then one scratch block, then real pointers).
    // Classes: reduce by slice count (sc = 1 for S = 2) and tile (2 kind + sc;
    // kind 0: entries up to 256 bytes, 1: generic, 2: up to 512 bytes),
    // reconstruct-GEMM by slice count, residue GEMM by slice count and
    // chunking (2 sc + chunked); the carry pass takes every CRT product in one
    // queue; two-digit products of at most 9 row blocks take the one-pass
    // reconstruct instead of the reconstruct GEMM and the carry pass.
    std::vector<ReduceJob> rj[6];
    // two-digit generic reduce jobs (class 3) as tiled MFMA GEMMs (fixed at creation: the graph is captured here)
    const bool rgemm = param(MGX_PARAM_REDUCE) == 0;
    std::vector<GemmJob> gj[4];
    std::vector<ReconJob> cj[2], kj, fj;
    std::vector<PlainJob> pj[NCLASS];
    std::vector<int64_t> rp[6], gp[4], cp[2], kp{0}, fp{0}, pp[NCLASS];
    for (auto &v : rp) v.push_back(0);
    for (auto &v : gp) v.push_back(0);
    for (auto &v : cp) v.push_back(0);
    for (auto &v : pp) v.push_back(0);
    struct Dims { int64_t rpadA, rpadB, kpad, ystride; };
    std::vector<Dims> dims(count);
    std::vector<int> pclass(count, -1), crt(count, 0);
    size_t nred[6] = {0, 0, 0, 0, 0, 0}, ngem[4] = {0, 0, 0, 0}, nrec[2] = {0, 0}, nrec_all = 0, nfus = 0;
    std::vector<int> fused(count, 0);
    const int64_t rmode = param(MGX_PARAM_RECON);
    size_t npl[NCLASS] = {0, 0, 0, 0, 0, 0};
    std::vector<AorsJob> aj[WSTAGES];
    std::vector<int64_t> wbytes(count, 0);
    auto wino = [&](int64_t i, uintptr_t tmp) {
        const mgx_product &p = prods[i];
        WinoGen g;
        g.depth = wdepth[i]; g.toom = tooms[i]; g.W = p.cap + 1; g.cap = p.cap;
        g.bits = std::max(bitsA[i], bitsB[i]);
        g.tmp = tmp;
        g.stage = aj;
        g.gen(0, p.K, p.M, p.N, (uint64_t *)p.C, p.Crs, (const uint64_t *)p.A, p.Ars, (const uint64_t *)p.B, p.Brs);
        return g;
    };
    for (int64_t i = 0; i < count; i++) {
        const mgx_product &p = prods[i];
        if (p.K == 0 || p.N == 0) continue;
        if (b->routes[i] == MGX_ROUTE_CRT) {
            Dims &d = dims[i];
            const int sc = views[i].S == 2;
            d.rpadA = round_up(p.K, 128); d.rpadB = round_up(p.N, 128);
            d.kpad = round_up(p.M, 64); d.ystride = crt_ystride(p.K * p.N);
            crt[i] = 1;
            nred[2 * reduce_kind(views[i].kb) + sc] += 2;
            ngem[2 * sc + (d.kpad > (sc ? KCHUNK2 : KCHUNK))]++;
            const PlanView &v = views[i];
            fused[i] = rmode != 1 && sc && (rmode == 2 || ceil_div(4 * v.nw, RECON_ROWS) <= 9) && 4 * v.nw <= v.Lbp;
            if (fused[i]) {
                nfus++;
            } else {
                nrec[sc]++;
                nrec_all++;
            }
        } else if (wdepth[i] > 0) {
            const WinoGen g = wino(i, 0);   // sizing: temporaries from offset 0, jobs counted
            wbytes[i] = g.used;
            for (auto &lf : g.leaves) npl[plain_class(lf.second)]++;
            pclass[i] = 0;
        } else {
            pclass[i] = plain_class(std::max(bitsA[i], bitsB[i]));
            npl[pclass[i]]++;
        }
    }
    size_t nas[WSTAGES];
    for (int st2 = 0; st2 < WSTAGES; st2++) {
        nas[st2] = aj[st2].size();
        aj[st2].clear();
    }
    Layout L;
    std::vector<int64_t> o_ra(count), o_rb(count), o_y(count), o_wv(count), o_ga(count);
    for (int64_t i = 0; i < count; i++) {
        if (!crt[i]) continue;
        const Dims &d = dims[i];
        const PlanView &v = views[i];
        o_ra[i] = L.take(v.S * v.P * d.rpadA * d.kpad);
        o_rb[i] = L.take(v.S * v.P * d.rpadB * d.kpad);
        o_y[i] = L.take(ceil_div(v.S * v.P, 4) * d.ystride * 4);
        if (!fused[i]) {   // the two passes' words, then G (int64) and Cb (int32)
            o_wv[i] = L.take(v.nw * d.ystride * 4);
            o_ga[i] = L.take(ceil_div(4 * v.nw, RECON_ROWS) * d.ystride * 12);
        }
    }
    std::vector<int64_t> o_wt(count, 0);
    for (int64_t i = 0; i < count; i++)
        if (wbytes[i]) o_wt[i] = L.take(wbytes[i]);
    const int64_t o_ctr = L.take(NSLOTS * QRANGES * QPITCH * 8);
    int64_t o_aj[WSTAGES], o_ap[WSTAGES];
    for (int c = 0; c < WSTAGES; c++) { o_aj[c] = L.take((int64_t)(nas[c] * sizeof(AorsJob))); o_ap[c] = L.take((int64_t)(nas[c] + 1) * 8); }
    int64_t o_rj[6], o_rp[6], o_gj[4], o_gp[4], o_cj[2], o_cp[2], o_pj[NCLASS], o_pp[NCLASS];
    for (int c = 0; c < 6; c++) { o_rj[c] = L.take((int64_t)(nred[c] * sizeof(ReduceJob))); o_rp[c] = L.take((int64_t)(nred[c] + 1) * 8); }
    for (int c = 0; c < 4; c++) { o_gj[c] = L.take((int64_t)(ngem[c] * sizeof(GemmJob))); o_gp[c] = L.take((int64_t)(ngem[c] + 1) * 8); }
    for (int c = 0; c < 2; c++) { o_cj[c] = L.take((int64_t)(nrec[c] * sizeof(ReconJob))); o_cp[c] = L.take((int64_t)(nrec[c] + 1) * 8); }
    const int64_t o_kj = L.take((int64_t)(nrec_all * sizeof(ReconJob))), o_kp = L.take((int64_t)(nrec_all + 1) * 8);
    const int64_t o_fj = L.take((int64_t)(nfus * sizeof(ReconJob))), o_fp = L.take((int64_t)(nfus + 1) * 8);
    for (int c = 0; c < NCLASS; c++) { o_pj[c] = L.take((int64_t)(npl[c] * sizeof(PlainJob))); o_pp[c] = L.take((int64_t)(npl[c] + 1) * 8); }
    b->scratch_bytes = L.off;
    if (hipMalloc(&b->scratch, (size_t)b->scratch_bytes) != hipSuccess) {
        b->scratch = nullptr;
        mgx_batch_destroy(b);
        return fail(MGX_ERR_OOM, "batch scratch of %lld bytes", (long long)L.off);
    }
    char *base = (char *)b->scratch;
    for (int64_t i = 0; i < count; i++) {
        const mgx_product &p = prods[i];
        if (crt[i]) {
            const PlanView &v = views[i];
            const Dims &d = dims[i];
            const int sc = v.S == 2;
            const int nrb = (int)(v.rows4pad / 64);
            const int kind = reduce_kind(v.kb);
            const bool resident = kind != 1;
            for (int t = 0; t < 2; t++) {
                ReduceJob r;
                r.pc = v.pc; r.cst = v.cst; r.cstb = v.cstb; r.P = (int)v.P; r.nrb = nrb; r.kb = (int)v.kb; r.transpose = t;
                r.xbits = (int)(t ? v.bitsB : v.bitsA);
                r.resident = resident;
                r.X = (const uint64_t *)(t ? p.B : p.A); r.Xrs = t ? p.Brs : p.Ars; r.W = p.cap + 1;
                r.rows = t ? p.N : p.K; r.K = p.M; r.kpad = d.kpad;
                r.nent = (t ? d.rpadB : d.rpadA) * d.kpad; r.plane = r.nent;
                r.R = (int8_t *)(base + (t ? o_rb[i] : o_ra[i]));
                const int rc = 2 * kind + sc;
                rj[rc].push_back(r);
                rp[rc].push_back(rp[rc].back() + (rc == 3 && rgemm ? rg_tiles(r.nent, r.P)
                                                                    : ceil_div(r.nent, 128) * (resident ? 1 : nrb)));
            }
            GemmJob g;
            g.pc = v.pc; g.RA = (const int8_t *)(base + o_ra[i]); g.RB = (const int8_t *)(base + o_rb[i]);
            g.planeA = d.rpadA * d.kpad; g.planeB = d.rpadB * d.kpad; g.kpad = d.kpad; g.rowsC = p.K; g.colsC = p.N;
            g.P = (int)v.P;
            g.ntm = (int)(d.rpadA / 128);
            g.ntn = (int)(d.rpadB / (sc ? 128 : 64));
            g.total = (int64_t)g.ntm * g.ntn * (sc ? ceil_div(v.P, 2) : v.P);
            g.ystride = d.ystride; g.Y = (uint32_t *)(base + o_y[i]);
            const int gc = 2 * sc + (d.kpad > (sc ? KCHUNK2 : KCHUNK));
            gj[gc].push_back(g);
            gp[gc].push_back(gp[gc].back() + g.total);
            ReconJob c;
            c.pc = v.pc; c.Mc = v.mc; c.M32 = v.m32; c.Y = g.Y; c.P = (int)v.P; c.Kr = (int)v.Kr;
            c.nrb = (int)ceil_div(4 * v.nw, RECON_ROWS); c.nw = (int)v.nw; c.ystride = d.ystride;
            c.nent = p.K * p.N; c.colsC = p.N; c.Crs = p.Crs; c.W = p.cap + 1; c.cap = p.cap; c.C = (uint64_t *)p.C;
            c.gshift = v.S == 2 ? 45 : 32;
            c.gw = 0;
            if (fused[i]) {
                c.Wv = nullptr; c.G = nullptr; c.Cb = nullptr;
                if (v.gw) { c.Mc = v.mcg; c.gw = (int)v.gw; }
                fj.push_back(c);
                fp.push_back(fp.back() + ceil_div(c.nent, RECON_ENT));
            } else {
                c.Wv = (uint32_t *)(base + o_wv[i]); c.G = (int64_t *)(base + o_ga[i]);
                c.Cb = (int32_t *)(base + o_ga[i] + c.nrb * d.ystride * 8);
                cj[sc].push_back(c);
                cp[sc].push_back(cp[sc].back() + ceil_div(c.nent, RECON_ENT) * c.nrb);
                kj.push_back(c);
                kp.push_back(kp.back() + ceil_div(c.nent, CARRY_ENT));
            }
        } else if (pclass[i] >= 0 && wdepth[i] > 0) {
            const WinoGen g = wino(i, (uintptr_t)(base + o_wt[i]));
            for (auto &lf : g.leaves) {
                const int k = plain_class(lf.second);
                pj[k].push_back(lf.first);
                pp[k].push_back(pp[k].back() + ceil_div(lf.first.K * lf.first.N, 4));
            }
        } else if (pclass[i] >= 0) {
            PlainJob q;
            q.K = p.K; q.M = p.M; q.N = p.N; q.Crs = p.Crs; q.Ars = p.Ars; q.Brs = p.Brs; q.W = p.cap + 1; q.cap = p.cap;
            q.C = (uint64_t *)p.C; q.A = (const uint64_t *)p.A; q.B = (const uint64_t *)p.B;
            q.toom = tooms[i];
            pj[pclass[i]].push_back(q);
            pp[pclass[i]].push_back(pp[pclass[i]].back() + ceil_div(p.K * p.N, 4));
        }
    }
    std::vector<int64_t> ap[WSTAGES];
    for (int c = 0; c < WSTAGES; c++) {
        if (aj[c].size() != nas[c]) st = fail(MGX_ERR_UNSUPPORTED, "batch: Winograd stage %d has %zu jobs, sized for %zu", c, aj[c].size(), nas[c]);
        ap[c].push_back(0);
        for (const AorsJob &a : aj[c]) ap[c].push_back(ap[c].back() + ceil_div(a.Wd * a.H, 4));
    }
    auto up = [&](int64_t o, const void *src, size_t n) -> int {
        if (n) MGX_HIP(hipMemcpy(base + o, src, n, hipMemcpyHostToDevice));
        return MGX_OK;
    };
    for (int c = 0; c < 6 && st == MGX_OK; c++) {
        st = up(o_rj[c], rj[c].data(), rj[c].size() * sizeof(ReduceJob));
        if (st == MGX_OK) st = up(o_rp[c], rp[c].data(), rp[c].size() * 8);
    }
    for (int c = 0; c < 2 && st == MGX_OK; c++) {
        st = up(o_cj[c], cj[c].data(), cj[c].size() * sizeof(ReconJob));
        if (st == MGX_OK) st = up(o_cp[c], cp[c].data(), cp[c].size() * 8);
    }
    for (int c = 0; c < 4 && st == MGX_OK; c++) {
        st = up(o_gj[c], gj[c].data(), gj[c].size() * sizeof(GemmJob));
        if (st == MGX_OK) st = up(o_gp[c], gp[c].data(), gp[c].size() * 8);
    }
    if (st == MGX_OK) st = up(o_kj, kj.data(), kj.size() * sizeof(ReconJob));
    if (st == MGX_OK) st = up(o_kp, kp.data(), kp.size() * 8);
    if (st == MGX_OK) st = up(o_fj, fj.data(), fj.size() * sizeof(ReconJob));
    if (st == MGX_OK) st = up(o_fp, fp.data(), fp.size() * 8);
    for (int c = 0; c < NCLASS && st == MGX_OK; c++) {
        st = up(o_pj[c], pj[c].data(), pj[c].size() * sizeof(PlainJob));
        if (st == MGX_OK) st = up(o_pp[c], pp[c].data(), pp[c].size() * 8);
    }
    for (int c = 0; c < WSTAGES && st == MGX_OK; c++) {
        st = up(o_aj[c], aj[c].data(), aj[c].size() * sizeof(AorsJob));
        if (st == MGX_OK) st = up(o_ap[c], ap[c].data(), ap[c].size() * 8);
    }
    if (st != MGX_OK) { mgx_batch_destroy(b); return st; }

    // ---- capture: reset counters, plain classes, reduce, residue GEMM, reconstruct (GEMM, carry).
    // Counter slots: 0-1 entry-resident reduce (256 bytes), 2-5 residue GEMM, 6-7 reconstruct GEMM, 8 carry,
    // 9-14 plain classes, 15-16 generic reduce, 17-18 entry-resident reduce (512 bytes), 19 one-pass
    // reconstruct, 20-33 the Winograd addition stages (pre-adds before the plain classes, post-adds after).
    unsigned long long *ctr = (unsigned long long *)(base + o_ctr);
    b->counters = ctr;
    unsigned *err = device_error_word();
    auto queue = [&](int slot, int64_t o_prefix, size_t njobs) {
        Queue q;
        q.counter = ctr + (int64_t)slot * QRANGES * QPITCH;
        q.prefix = (const int64_t *)(base + o_prefix);
        q.njobs = (int)njobs;
        return q;
    };
    hipStream_t cs = nullptr;
    if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) { mgx_batch_destroy(b); return fail(MGX_ERR_HIP, "stream"); }
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(reset_counters_kernel, dim3(NSLOTS * QRANGES * QPITCH / 256), dim3(256), 0, cs, ctr,
                           NSLOTS * QRANGES * QPITCH);
        auto adds = [&](int c) {
            if (aj[c].empty()) return;
            int grid = (int)std::min<int64_t>(ap[c].back(), 1024);
            b->queues.push_back({MGX_QUEUE_ADDSUB, 20 + c, ap[c].back(), grid});
            hipLaunchKernelGGL(batch_aors_kernel, dim3(grid), dim3(256), 0, cs, queue(20 + c, o_ap[c], aj[c].size()),
                               (const AorsJob *)(base + o_aj[c]), err);
        };
        for (int c = 0; c < 6; c++) adds(c);
        for (int c = 0; c < NCLASS; c++) {
            if (pj[c].empty()) continue;
            Queue q = queue(9 + c, o_pp[c], pj[c].size());
            int grid = (int)std::min<int64_t>(pp[c].back(), 2048);
            b->queues.push_back({MGX_QUEUE_PLAIN, 9 + c, pp[c].back(), grid});
            const PlainJob *jobs = (const PlainJob *)(base + o_pj[c]);
            switch (NB_CLASSES[c]) {
            case 1: launch_plain<1>(grid, cs, q, jobs, err); break;
            case 2: launch_plain<2>(grid, cs, q, jobs, err); break;
            case 3: launch_plain<3>(grid, cs, q, jobs, err); break;
            case 4: launch_plain<4>(grid, cs, q, jobs, err); break;
            case 6: launch_plain<6>(grid, cs, q, jobs, err); break;
            default: launch_plain<8>(grid, cs, q, jobs, err); break;
            }
        }
        for (int c = 6; c < WSTAGES; c++) adds(c);
        for (int c = 0; c < 6; c++) {
            if (rj[c].empty()) continue;
            const int kind = c >> 1, two = c & 1;
            int grid = (int)std::min<int64_t>(rp[c].back(), kind == 1 ? 2048 : 1024);
            if (c == 3 && rgemm) grid = (int)std::min<int64_t>(rp[c].back(), 512);
            auto kern = c == 3 && rgemm ? batch_reduce_gemm_kernel
                      : kind == 1 ? (two ? batch_reduce_gen_kernel<2> : batch_reduce_gen_kernel<4>)
                      : kind == 2 ? (two ? batch_reduce_res_kernel<2, 8> : batch_reduce_res_kernel<4, 8>)
                                  : (two ? batch_reduce_res_kernel<2, 4> : batch_reduce_res_kernel<4, 4>);
            const int slot = kind == 0 ? two : kind == 1 ? 15 + two : 17 + two;
            b->queues.push_back({MGX_QUEUE_REDUCE, slot, rp[c].back(), grid});
            hipLaunchKernelGGL(kern, dim3(grid), dim3(256), 0, cs, queue(slot, o_rp[c], rj[c].size()),
                               (const ReduceJob *)(base + o_rj[c]), err);
        }
        for (int c = 0; c < 4; c++) {
            if (gj[c].empty()) continue;
            int grid = (int)std::min<int64_t>(gp[c].back(), 256);
            b->queues.push_back({MGX_QUEUE_RESIDUE_GEMM, 2 + c, gp[c].back(), grid});
            auto kern = c == 0 ? batch_gemm_kernel<false> : c == 1 ? batch_gemm_kernel<true>
                      : c == 2 ? batch_gemm2_kernel<false> : batch_gemm2_kernel<true>;
            hipLaunchKernelGGL(kern, dim3(grid), dim3(256), 0, cs, queue(2 + c, o_gp[c], gj[c].size()),
                               (const GemmJob *)(base + o_gj[c]));
        }
        for (int c = 0; c < 2; c++) {
            if (cj[c].empty()) continue;
            int grid = (int)std::min<int64_t>(cp[c].back(), 1024);
            b->queues.push_back({MGX_QUEUE_RECONSTRUCT, 6 + c, cp[c].back(), grid});
            hipLaunchKernelGGL(c ? batch_recon_gemm_kernel<2> : batch_recon_gemm_kernel<4>, dim3(grid), dim3(256), 0,
                               cs, queue(6 + c, o_cp[c], cj[c].size()), (const ReconJob *)(base + o_cj[c]));
        }
        if (!fj.empty()) {
            int grid = (int)std::min<int64_t>(fp.back(), 512);
            b->queues.push_back({MGX_QUEUE_RECONSTRUCT, 19, fp.back(), grid});
            hipLaunchKernelGGL(batch_recon_fused_kernel, dim3(grid), dim3(256), 0, cs, queue(19, o_fp, fj.size()),
                               (const ReconJob *)(base + o_fj), err);
        }
        if (!kj.empty()) {
            int grid = (int)std::min<int64_t>(kp.back(), 1024);
            b->queues.push_back({MGX_QUEUE_CARRY, 8, kp.back(), grid});
            hipLaunchKernelGGL(batch_recon_carry_kernel, dim3(grid), dim3(256), 0, cs, queue(8, o_kp, kj.size()),
                               (const ReconJob *)(base + o_kj), err);
        }
        e = hipStreamEndCapture(cs, &b->graph);
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
    hipStreamDestroy(cs);
    if (e != hipSuccess) {
        mgx_batch_destroy(b);
        return fail(MGX_ERR_HIP, "batch graph capture: %s", hipGetErrorString(e));
    }
    (void)s;
    *out = b;
    return MGX_OK;
}

int mgx_batch_run(mgx_batch *b, void *stream)
{
    if (!b) return fail(MGX_ERR_ARG, "null batch");
    MGX_HIP(hipGraphLaunch(b->exec, as_stream(stream)));
    return MGX_OK;
}

int mgx_batch_route(const mgx_batch *b, int64_t i)
{
    if (!b || i < 0 || i >= (int64_t)b->routes.size()) return -1;
    return b->routes[i];
}

int64_t mgx_batch_scratch_bytes(const mgx_batch *b) { return b ? b->scratch_bytes : -1; }

int64_t mgx_batch_queue_count(const mgx_batch *b) { return b ? (int64_t)b->queues.size() : -1; }

int mgx_batch_queue_info(const mgx_batch *b, int64_t i, int *kind, int64_t *items, int *grid, int64_t *claims)
{
    if (!b || i < 0 || i >= (int64_t)b->queues.size()) return fail(MGX_ERR_ARG, "no queue %lld", (long long)i);
    const BatchQueue &q = b->queues[i];
    if (kind) *kind = q.kind;
    if (items) *items = q.items;
    if (grid) *grid = q.grid;
    if (claims) {
        unsigned long long c[QRANGES * QPITCH];
        MGX_HIP(hipDeviceSynchronize());
        MGX_HIP(hipMemcpy(c, b->counters + (int64_t)q.slot * QRANGES * QPITCH, sizeof c, hipMemcpyDeviceToHost));
        int64_t sum = 0;
        for (int x = 0; x < QRANGES; x++) sum += (int64_t)c[QPITCH * x];
        *claims = sum;
    }
    return MGX_OK;
}

int mgx_queue_map(const int64_t *prefix, int njobs, int64_t item, int *job, int64_t *tile)
{
    if (!prefix || !job || !tile) return fail(MGX_ERR_ARG, "null argument");
    return queue_map(prefix, njobs, item, job, tile) ? 0 : 1;
}

int64_t mgx_queue_claim(uint64_t *counters, int64_t total, int *cur, int *moved)
{
    if (!counters || !cur || !moved || *cur < 0 || *cur >= QRANGES) return -3;
    return queue_claim_once([&](int x) { return counters[x]++; }, total, *cur, *moved);
}

int mgx_batch_destroy(mgx_batch *b)
{
    if (!b) return MGX_OK;
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    if (b->scratch || !b->plans.empty()) (void)hipDeviceSynchronize();   // replays still queued
    if (b->scratch) hipFree(b->scratch);
    for (const mgx_crt_plan *pl : b->plans) release_plan(pl);
    delete b;
    return MGX_OK;
}

}  // extern "C"
