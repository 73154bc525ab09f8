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
};
struct PlainJob {
    int64_t K, M, N, Crs, Ars, Brs, W, cap;
    uint64_t *C;
    const uint64_t *A, *B;
    int toom;   // Toom-22 cutoff of this product's schedule
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

constexpr int NSLOTS = 32;   // queues of a batch (counter slots)
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
        recon_fused_block(t, r.pc, r.P, r.Mc, r.Kr, r.nrb, r.nw, r.Y, r.ystride, r.nent, r.M32, r.colsC, r.C, r.Crs,
                          r.W, r.cap, err);
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
    std::vector<int> slices(count, -1), tooms(count, (int)param(MGX_PARAM_TOOM22));
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
        // the batch runs plain products as one flat per-entry launch
        if (r == MGX_ROUTE_WINOGRAD1 || r == MGX_ROUTE_WINOGRAD2) r = MGX_ROUTE_PLAIN;
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
    const bool fuse_ok = param(MGX_PARAM_RECON) != 1;
    size_t npl[NCLASS] = {0, 0, 0, 0, 0, 0};
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
            fused[i] = fuse_ok && sc && ceil_div(4 * v.nw, RECON_ROWS) <= 9 && 4 * v.nw <= v.Lbp;
            if (fused[i]) {
                nfus++;
            } else {
                nrec[sc]++;
                nrec_all++;
            }
        } else {
            int64_t words = 2 * ((std::max(bitsA[i], bitsB[i]) + 63) / 64);
            int k = 0;
            while (k < NCLASS - 1 && 64 * NB_CLASSES[k] < words) k++;
            pclass[i] = k;
            npl[k]++;
        }
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
    const int64_t o_ctr = L.take(NSLOTS * QRANGES * QPITCH * 8);
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
            if (fused[i]) {
                c.Wv = nullptr; c.G = nullptr; c.Cb = nullptr;
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
        } else if (pclass[i] >= 0) {
            PlainJob q;
            q.K = p.K; q.M = p.M; q.N = p.N; q.Crs = p.Crs; q.Ars = p.Ars; q.Brs = p.Brs; q.W = p.cap + 1; q.cap = p.cap;
            q.C = (uint64_t *)p.C; q.A = (const uint64_t *)p.A; q.B = (const uint64_t *)p.B;
            q.toom = tooms[i];
            pj[pclass[i]].push_back(q);
            pp[pclass[i]].push_back(pp[pclass[i]].back() + ceil_div(p.K * p.N, 4));
        }
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
    if (st != MGX_OK) { mgx_batch_destroy(b); return st; }

    // ---- capture: reset counters, plain classes, reduce, residue GEMM, reconstruct (GEMM, carry).
    // Counter slots: 0-1 entry-resident reduce (256 bytes), 2-5 residue GEMM, 6-7 reconstruct GEMM, 8 carry,
    // 9-14 plain classes, 15-16 generic reduce, 17-18 entry-resident reduce (512 bytes), 19 one-pass
    // reconstruct.
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
