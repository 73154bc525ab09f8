/*
 * mgx_sched.h -- scheduling side of the library: seeded device inputs, the
 * route autotuner, the learned cost model and the batched executor.
 *
 * None of these has a counterpart in the reference, whose job scheduler
 * (mat-gpu/job.cu:32-336) is never called; the batched executor replaces it
 * with a device work queue feeding persistent kernels.  The scheduling
 * choices never change the arithmetic: every route is exact.
 */
#ifndef MGX_SCHED_H
#define MGX_SCHED_H

#include "mgx.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- routes ---------------- */
typedef enum {
    MGX_ROUTE_CRT = 0,        /* INT8-MFMA multi-modular route */
    MGX_ROUTE_PLAIN = 1,      /* per-entry mpn kernel, no recursion */
    MGX_ROUTE_WINOGRAD1 = 2,  /* plain route, one Winograd level (squarify) */
    MGX_ROUTE_WINOGRAD2 = 3,  /* plain route, two Winograd levels (squarify) */
    MGX_ROUTES = 4,
    MGX_ROUTE_AUTO = -1       /* cost model, falling back to the autotuner */
} mgx_route;

/* C = A*B on the device through one route (its parameters from the global
 * tunables), or MGX_ROUTE_AUTO: the schedule the cost model picks.
 * Synchronous with status. */
int mgx_matmul_route(int route, int64_t K, int64_t M, int64_t N, mgx_limb *C, int64_t Crs, const mgx_limb *A,
                     int64_t Ars, const mgx_limb *B, int64_t Brs, int64_t cap, void *stream);

/* ---------------- schedules: the autotuner's search space ---------------- */
/* A schedule fixes what the reference fixes at compile time and sweeps by
 * rebuilding (mat-gpu/mat-params.h:4-14, Makefile:19-22): the route, the
 * multi-modular route's digit count (hence prime count: two digits use
 * about twice as many primes below 65280 as four digits use below 2^31) and
 * residue-GEMM tile, and the plain routes' Toom-22 cutoff and Winograd
 * depth.  Every schedule is exact; only the time differs. */
#define MGX_SCHEDULES 9
typedef struct {
    int route;       /* mgx_route */
    int slices;      /* CRT: digits per residue (0: two while the primes below 65280 suffice, else four) */
    int gemm_rows;   /* CRT with two digits: residue-GEMM tile height, 128 or 64 */
    int toom;        /* plain routes: Toom-22 cutoff in 64-word blocks (0: schoolbook only) */
    int depth;       /* Winograd levels (0 for the flat plain kernel and the CRT route) */
} mgx_schedule;
int mgx_schedule_get(int index, mgx_schedule *out);
/* C = A*B through schedule index (or MGX_ROUTE_AUTO).  Synchronous. */
int mgx_matmul_schedule(int index, int64_t K, int64_t M, int64_t N, mgx_limb *C, int64_t Crs, const mgx_limb *A,
                        int64_t Ars, const mgx_limb *B, int64_t Brs, int64_t cap, void *stream);

/* Device-side seeded entries (a counter-based SplitMix64 hash of seed, entry
 * and limb): each entry has bits_hi bits with probability frac_hi, else
 * bits_lo bits, magnitude uniform below 2^bits, sign uniform. */
int mgx_random_matrix_device(mgx_limb *M, int64_t W, int64_t H, int64_t rowstride, int64_t cap, int64_t bits_lo,
                             int64_t bits_hi, double frac_hi, uint64_t seed, void *stream);

/* ---------------- features and the autotuner ---------------- */
#define MGX_FEATURES 13
/* Features of a product from its shape and the bit lengths of its entries
 * (host-side histograms: counts of entries of A and B in the bit-length bins
 * <=128, <=512, <=2048, <=8192, <=32768, larger). */
int mgx_tune_features(int64_t K, int64_t M, int64_t N, const int64_t *hist6, double mean_bits, int64_t max_bits,
                      double mean_blocks, float *out);
/* Features of a device-resident product (scans the entries' size words, so
 * bit lengths are limb-granular, 64 |size|; synchronizes). */
int mgx_tune_features_device(int64_t K, int64_t M, int64_t N, const mgx_limb *A, int64_t Ars, const mgx_limb *B,
                             int64_t Brs, int64_t cap, float *out);
/* The exhaustive autotuner: times every schedule on the given device
 * matrices (C is scratch, rowstrides = widths), median of reps calls after
 * one warm-up; schedules that cannot take the product get 1e30, and plain
 * schedules that cannot come within 2x of the best multi-modular one even
 * with two Winograd levels are extrapolated from one block of 8 rows
 * instead of run (bit i of *estimated).  Times in seconds. */
int mgx_tune_time_schedules(int64_t K, int64_t M, int64_t N, mgx_limb *C, const mgx_limb *A, const mgx_limb *B,
                            int64_t cap, int reps, double *times, unsigned *estimated);
/* The same for the four routes at their default schedules (0, 3, 5, 6). */
int mgx_tune_time_routes(int64_t K, int64_t M, int64_t N, mgx_limb *C, const mgx_limb *A, const mgx_limb *B,
                         int64_t cap, int reps, double *times, unsigned *estimated);

/* ---------------- learned cost model (FP32 MLP, hand-written HIP kernels) ---------------- */
typedef struct mgx_costmodel mgx_costmodel;
int mgx_costmodel_create(mgx_costmodel **cm, int D, int H, int O, uint64_t seed);
int mgx_costmodel_destroy(mgx_costmodel *cm);
int64_t mgx_costmodel_param_count(const mgx_costmodel *cm);
int64_t mgx_costmodel_outputs(const mgx_costmodel *cm);   /* O: MGX_SCHEDULES (or MGX_ROUTES for a route model) */
int mgx_costmodel_get_params(const mgx_costmodel *cm, float *out);
int mgx_costmodel_set_params(mgx_costmodel *cm, const float *in);
int mgx_costmodel_set_normalisation(mgx_costmodel *cm, const float *mean, const float *scale);
int mgx_costmodel_loss_grad(mgx_costmodel *cm, const float *X, const float *T, const float *mask, int B,
                            float *grad, float *loss);
int mgx_costmodel_train(mgx_costmodel *cm, const float *X, const float *T, const float *mask, int B, int epochs,
                        float lr, float *loss);
int mgx_costmodel_predict(mgx_costmodel *cm, const float *X, int B, float *Y);
/* Load a trained model from the text file costmodel.py writes
 * ("mgx-costmodel 1 D H O", means, scales, packed parameters). */
int mgx_costmodel_load(mgx_costmodel **cm, const char *path);
/* Installs the model used by MGX_ROUTE_AUTO (NULL: none; the autotuner's
 * exhaustive timing decides instead).  The library does not take ownership. */
int mgx_dispatch_set_model(mgx_costmodel *cm);
/* The schedule (and its route) MGX_ROUTE_AUTO would take for a
 * device-resident product. */
int mgx_dispatch_choose_schedule(int64_t K, int64_t M, int64_t N, const mgx_limb *A, int64_t Ars, const mgx_limb *B,
                                 int64_t Brs, int64_t cap, int *schedule);
int mgx_dispatch_choose(int64_t K, int64_t M, int64_t N, const mgx_limb *A, int64_t Ars, const mgx_limb *B,
                        int64_t Brs, int64_t cap, int *route);
/* The schedule the calling thread's last MGX_ROUTE_AUTO product ran (-1:
 * none yet).  With a model installed, that product's features and inference
 * ran on the device right after its statistics scan (one launch, no extra
 * synchronization). */
int mgx_dispatch_last(void);

/* ---------------- measured ceilings (no reference counterpart) ---------------- */
/* Measures, on the current device, the INT8 MFMA issue rate
 * (v_mfma_i32_32x32x32_i8, operations/s), the 32x32+64-bit integer
 * multiply-add rate (v_mad_u64_u32, the plain route's limb product, per
 * second) and the device copy bandwidth (bytes read + written per second).
 * Any output pointer may be NULL.  Synchronous; takes 2 GiB of scratch. */
int mgx_measure_peaks(double *int8_ops_per_s, double *mad_per_s, double *copy_bytes_per_s);

/* ---------------- batched executor ---------------- */
typedef struct {
    int64_t K, M, N;            /* C is K x N, A is K x M, B is M x N */
    mgx_limb *C;
    int64_t Crs;
    const mgx_limb *A;
    int64_t Ars;
    const mgx_limb *B;
    int64_t Brs;
    int64_t cap;
    int64_t bitsA, bitsB;       /* bounds on |entries| (0: scanned at build) */
    int route;                  /* mgx_route; MGX_ROUTE_AUTO lets the cost model choose */
} mgx_product;
typedef struct mgx_batch mgx_batch;
/* Plans every product, allocates all scratch, builds the per-stage work
 * queues and captures the whole batch (one persistent launch per stage per
 * route class) as a hipGraph.  Synchronous. */
int mgx_batch_create(mgx_batch **b, const mgx_product *products, int64_t count, void *stream);
/* Replays the captured graph on the stream (asynchronous). */
int mgx_batch_run(mgx_batch *b, void *stream);
/* Route chosen for product i, and the number of graph nodes. */
int mgx_batch_route(const mgx_batch *b, int64_t i);
int64_t mgx_batch_scratch_bytes(const mgx_batch *b);
int mgx_batch_destroy(mgx_batch *b);

/* Work queues of a batch, in launch order (one per persistent launch).  A
 * queue is the concatenation of its jobs' tile ranges, split into 8
 * contiguous item ranges with a counter each (one per XCD); a workgroup
 * claims from its XCD's range first and moves on to the next range when a
 * claim there fails, so after a replay the counters of a queue sum to
 * items + 8 grid (claims).  Reading the claims synchronizes the device.
 * Diagnostics with no reference counterpart (the reference's host
 * scheduler, mat-gpu/job.cu:183-336, has none). */
typedef enum {
    MGX_QUEUE_PLAIN = 0, MGX_QUEUE_REDUCE = 1, MGX_QUEUE_RESIDUE_GEMM = 2, MGX_QUEUE_RECONSTRUCT = 3,
    MGX_QUEUE_CARRY = 4,
    MGX_QUEUE_ADDSUB = 5   /* block additions of the Winograd levels of plain products */
} mgx_queue_kind;
int64_t mgx_batch_queue_count(const mgx_batch *b);
int mgx_batch_queue_info(const mgx_batch *b, int64_t i, int *kind, int64_t *items, int *grid, int64_t *claims);
/* The queues' item mapping (host copy of the device code): item -> (job,
 * tile) for prefix[0] = 0 <= ... <= prefix[njobs] = items.  Returns 0, or 1
 * when item is past the end (the claim that ends a workgroup's loop). */
int mgx_queue_map(const int64_t *prefix, int njobs, int64_t item, int *job, int64_t *tile);
/* One claim attempt (host copy of the device code, counters[0..7] the
 * ranges' counters): the claimed item; -1 when range *cur is drained (*cur
 * and *moved advance); -2 when all 8 ranges have been found drained; -3 on
 * bad arguments.  A workgroup starts with *cur = its XCD, *moved = 0. */
int64_t mgx_queue_claim(uint64_t *counters, int64_t total, int *cur, int *moved);

#ifdef __cplusplus
}
#endif
#endif
