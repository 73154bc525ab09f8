// peaks.hip -- on-device measurements of the ceilings the kernels are measured
// against: the INT8 MFMA issue rate (v_mfma_i32_32x32x32_i8, the residue
// GEMM's instruction), the 32x32->64-bit integer multiply-add rate
// (v_mad_u64_u32, the plain route's limb product) and the device-to-device
// copy bandwidth (HBM).  Each probe is a fixed, bounded loop over
// independent operations whose result is stored once per thread, so the
// compiler cannot drop the work.
#include "common.h"
#include "../../include/mgx_sched.h"
#include <algorithm>

namespace mgx {

typedef int v4i_p __attribute__((ext_vector_type(4)));
typedef int v16i_p __attribute__((ext_vector_type(16)));

// every wave: 4 independent accumulator chains of `iters` MFMAs each
__global__ __launch_bounds__(256) void probe_mfma_kernel(int iters, int seed, int *out)
{
    const v4i_p a = {seed + (int)threadIdx.x, seed ^ 0x5a5a, seed * 3, 7};
    const v4i_p b = {seed - 1, (int)threadIdx.x, seed ^ 0x3c3c, 11};
    v16i_p c0 = {0}, c1 = {0}, c2 = {0}, c3 = {0};
    for (int i = 0; i < iters; i++) {
        c0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c0, 0, 0, 0);
        c1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c1, 0, 0, 0);
        c2 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c2, 0, 0, 0);
        c3 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c3, 0, 0, 0);
    }
    int r = 0;
#pragma unroll
    for (int k = 0; k < 16; k++) r += c0[k] ^ c1[k] ^ c2[k] ^ c3[k];
    out[blockIdx.x * blockDim.x + threadIdx.x] = r;
}

// every lane: 8 independent chains of `iters` 32x32+64 -> 64-bit multiply-adds
__global__ __launch_bounds__(256) void probe_mad_kernel(int iters, uint32_t seed, uint64_t *out)
{
    const uint32_t t = threadIdx.x + seed;
    uint32_t m[8];
    uint64_t acc[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
        m[k] = t * 2654435761u + 97u * (uint32_t)k + 1u;
        acc[k] = (uint64_t)k * 0x9e3779b97f4a7c15ull + t;
    }
    for (int i = 0; i < iters; i++) {
#pragma unroll
        for (int k = 0; k < 8; k++) acc[k] = (uint64_t)m[k] * (uint32_t)acc[k] + acc[k];
    }
    uint64_t r = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) r ^= acc[k];
    out[blockIdx.x * blockDim.x + threadIdx.x] = r;
}

__global__ __launch_bounds__(256) void probe_copy_kernel(const uint4 *__restrict__ in, uint4 *__restrict__ out, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = in[i];
}

}  // namespace mgx

using namespace mgx;

namespace {
// median of `reps` event-timed launches, in milliseconds
template <typename F>
int time_ms(F launch, int reps, hipStream_t s, float *ms_out)
{
    hipEvent_t e0, e1;
    MGX_HIP(hipEventCreate(&e0));
    MGX_HIP(hipEventCreate(&e1));
    launch();   // warm-up
    std::vector<float> t;
    for (int r = 0; r < reps; r++) {
        MGX_HIP(hipEventRecord(e0, s));
        launch();
        MGX_HIP(hipEventRecord(e1, s));
        MGX_HIP(hipEventSynchronize(e1));
        float ms = 0;
        MGX_HIP(hipEventElapsedTime(&ms, e0, e1));
        t.push_back(ms);
    }
    MGX_HIP(hipGetLastError());
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    std::sort(t.begin(), t.end());
    *ms_out = t[t.size() / 2];
    return MGX_OK;
}
}  // namespace

extern "C" {

int mgx_measure_peaks(double *int8_ops_per_s, double *mad_per_s, double *copy_bytes_per_s)
{
    MGX_TRY(ensure_init());
    hipStream_t s = nullptr;
    int dev = 0, ncu = 0;
    MGX_HIP(hipGetDevice(&dev));
    MGX_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const int grid = 8 * ncu;   // two 256-thread workgroups per SIMD's worth of waves
    void *buf = nullptr;
    const size_t copy_bytes = (size_t)1 << 30;
    MGX_TRY(scratch_alloc(&buf, 2 * copy_bytes, s));
    int st = MGX_OK;
    float ms = 0;
    if (int8_ops_per_s) {
        const int iters = 2048;
        st = time_ms([&] { hipLaunchKernelGGL(probe_mfma_kernel, dim3(grid), dim3(256), 0, s, iters, 1, (int *)buf); },
                     5, s, &ms);
        // per wave per iteration: 4 MFMAs of 2 * 32 * 32 * 32 operations
        if (st == MGX_OK) *int8_ops_per_s = (double)grid * 4 * iters * 4 * 65536.0 / (ms * 1e-3);
    }
    if (st == MGX_OK && mad_per_s) {
        const int iters = 4096;
        st = time_ms([&] { hipLaunchKernelGGL(probe_mad_kernel, dim3(grid), dim3(256), 0, s, iters, 1u, (uint64_t *)buf); },
                     5, s, &ms);
        if (st == MGX_OK) *mad_per_s = (double)grid * 256 * iters * 8 / (ms * 1e-3);
    }
    if (st == MGX_OK && copy_bytes_per_s) {
        const int64_t n = (int64_t)(copy_bytes / sizeof(uint4));
        const uint4 *in = (const uint4 *)buf;
        uint4 *out = (uint4 *)((char *)buf + copy_bytes);
        MGX_HIP(hipMemsetAsync(buf, 1, copy_bytes, s));
        st = time_ms([&] { hipLaunchKernelGGL(probe_copy_kernel, dim3(grid * 4), dim3(256), 0, s, in, out, n); }, 5, s,
                     &ms);
        if (st == MGX_OK) *copy_bytes_per_s = 2.0 * (double)copy_bytes / (ms * 1e-3);
    }
    scratch_free(buf, s);
    MGX_HIP(hipStreamSynchronize(s));
    return st;
}

}  // extern "C"
