// plain.hip -- the plain route: per-entry mpn arithmetic on gfx950 VALU.
//
// One wavefront per output entry.  Operands are read as 32-bit words (the
// 64-bit limbs of the entry type, little endian), one word per lane per
// 64-word block.  A product of two blocks is computed column-wise (GMP's
// mul_basecase result, mat-gpu/mul_basecase.cu:53-96, in product-scanning
// order): at step i lane l multiplies a_i (read from lane i into a scalar
// register) by b_{(l-i) mod 64} (b rotated one lane per step with DPP
// wave_ror) and adds the 64-bit product into the 96-bit sum of column l or
// l+64.  Column sums of all k products of C[y][x] = sum_k A[y][k] B[k][x]
// accumulate in two unsigned sets, one per sign of the product
// (mat-gpu/mpz-aorsmul.cu:69-161 decides the same sign per step); the
// entry is normalized once at the end: columns to digits, carries resolved
// across the wave with ballot masks (carry-lookahead by one 64-bit integer
// add of the generate/propagate masks), the two magnitudes compared and
// subtracted the same way, and the sign-magnitude result written.
//
// Operands of two or more blocks may use one level of Karatsuba
// (mat-gpu/toom22_mul.cu:108-235: evaluation at 0, -1 and infinity) on block
// halves: |a0 - a1| |b0 - b1| replaces one of the four half products.
#include "plain_kernels.h"
#include <algorithm>

namespace mgx {

template <int NB, bool STRIDED, bool ADDMUL>
__global__ __launch_bounds__(256) void plain_mul_kernel(int64_t K, int64_t M, int64_t N, uint64_t *__restrict__ C,
                                                        int64_t Crs, const uint64_t *__restrict__ A, int64_t Ars,
                                                        const uint64_t *__restrict__ B, int64_t Brs, int64_t W,
                                                        int64_t cap, int toom, unsigned *__restrict__ err)
{
    plain_entries<NB, STRIDED, ADDMUL>(blockIdx.x, K, M, N, C, Crs, A, Ars, B, Brs, W, cap, toom, err);
}

// ---------------------------------------------------------------- add / sub
template <bool STRIDED>
__global__ __launch_bounds__(256) void plain_aors_kernel(int64_t Wd, int64_t H, uint64_t *__restrict__ C,
                                                         int64_t Crs, const uint64_t *__restrict__ A, int64_t Ars,
                                                         const uint64_t *__restrict__ B, int64_t Brs, int64_t W,
                                                         int64_t cap, int sub, unsigned *__restrict__ err)
{
    aors_entry<STRIDED>((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6), Wd, H, C, Crs, A, Ars, B, Brs, W, cap, sub, err);
}

// ---------------------------------------------------------- interleave
// out[block][w][j] = in[block][j][w]  (dir 0)  and the inverse (dir 1)
__global__ void interleave_kernel(const uint64_t *__restrict__ in, uint64_t *__restrict__ out, int64_t nblocks,
                                  int64_t W, int dir)
{
    const int64_t per = MGX_INTERLEAVE_BLOCK * W;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nblocks * per;
         t += (int64_t)gridDim.x * blockDim.x) {
        int64_t b = t / per, r = t % per;
        int64_t w = r / MGX_INTERLEAVE_BLOCK, j = r % MGX_INTERLEAVE_BLOCK;   // strided position
        if (dir == 0) out[b * per + r] = in[b * per + j * W + w];
        else out[b * per + j * W + w] = in[b * per + r];
    }
}

}  // namespace mgx

// ========================================================================
// host launchers
// ========================================================================
using namespace mgx;

namespace {

template <int NB, bool S, bool AM>
void launch_mul_nb(dim3 g, hipStream_t s, int64_t K, int64_t M, int64_t N, uint64_t *C, int64_t Crs, const uint64_t *A,
                   int64_t Ars, const uint64_t *B, int64_t Brs, int64_t W, int64_t cap, int toom, unsigned *err)
{
    hipLaunchKernelGGL((plain_mul_kernel<NB, S, AM>), g, dim3(256), 0, s, K, M, N, C, Crs, A, Ars, B, Brs, W, cap,
                       toom, err);
}

template <bool S, bool AM>
int launch_mul(int64_t K, int64_t M, int64_t N, uint64_t *C, int64_t Crs, const uint64_t *A, int64_t Ars,
               const uint64_t *B, int64_t Brs, int64_t cap, int64_t opwords, hipStream_t s, int toom)
{
    MGX_TRY(ensure_init());
    if (K < 0 || M < 0 || N < 0 || cap < 1) return fail(MGX_ERR_SHAPE, "bad shape");
    if (S && (Crs % MGX_INTERLEAVE_BLOCK || Ars % MGX_INTERLEAVE_BLOCK || Brs % MGX_INTERLEAVE_BLOCK))
        return fail(MGX_ERR_SHAPE, "interleaved rowstrides must be multiples of %d", MGX_INTERLEAVE_BLOCK);
    if (K * N == 0) return MGX_OK;
    // operands are at most cap limbs; wider than the widest kernel class (512
    // words) the kernel itself reports the operands it cannot take
    int64_t nbneed = std::min<int64_t>((opwords + 63) / 64, 8);
    dim3 g((unsigned)ceil_div(K * N, 4));
    unsigned *err = device_error_word();
    if (toom < 0) toom = (int)param(MGX_PARAM_TOOM22);
    const int64_t W = cap + 1;
#define MGX_CASE(nb) \
    if (nbneed <= nb) { launch_mul_nb<nb, S, AM>(g, s, K, M, N, C, Crs, A, Ars, B, Brs, W, cap, toom, err); MGX_HIP(hipGetLastError()); return MGX_OK; }
    MGX_CASE(1) MGX_CASE(2) MGX_CASE(3) MGX_CASE(4) MGX_CASE(6) MGX_CASE(8)
#undef MGX_CASE
    return fail(MGX_ERR_UNSUPPORTED, "operands of %lld words exceed the plain route's 512-word limit", (long long)opwords);
}

template <bool S>
int launch_aors(int64_t Wd, int64_t H, uint64_t *C, int64_t Crs, const uint64_t *A, int64_t Ars, const uint64_t *B,
                int64_t Brs, int64_t cap, int sub, hipStream_t s)
{
    MGX_TRY(ensure_init());
    if (Wd < 0 || H < 0 || cap < 1) return fail(MGX_ERR_SHAPE, "bad shape");
    if (S && (Crs % MGX_INTERLEAVE_BLOCK || Ars % MGX_INTERLEAVE_BLOCK || Brs % MGX_INTERLEAVE_BLOCK))
        return fail(MGX_ERR_SHAPE, "interleaved rowstrides must be multiples of %d", MGX_INTERLEAVE_BLOCK);
    if (Wd * H == 0) return MGX_OK;
    hipLaunchKernelGGL((plain_aors_kernel<S>), dim3((unsigned)ceil_div(Wd * H, 4)), dim3(256), 0, s, Wd, H, C, Crs, A,
                       Ars, B, Brs, cap + 1, cap, sub, device_error_word());
    MGX_HIP(hipGetLastError());
    return MGX_OK;
}

}  // namespace

namespace mgx {
// operand-size-aware entry used by the recursion drivers
int plain_mul(bool strided, bool addmul, int64_t K, int64_t M, int64_t N, uint64_t *C, int64_t Crs, const uint64_t *A,
              int64_t Ars, const uint64_t *B, int64_t Brs, int64_t cap, int64_t opwords, hipStream_t s, int toom)
{
    if (strided) return addmul ? launch_mul<true, true>(K, M, N, C, Crs, A, Ars, B, Brs, cap, opwords, s, toom)
                               : launch_mul<true, false>(K, M, N, C, Crs, A, Ars, B, Brs, cap, opwords, s, toom);
    return addmul ? launch_mul<false, true>(K, M, N, C, Crs, A, Ars, B, Brs, cap, opwords, s, toom)
                  : launch_mul<false, false>(K, M, N, C, Crs, A, Ars, B, Brs, cap, opwords, s, toom);
}
int plain_aors(bool strided, bool sub, int64_t Wd, int64_t H, uint64_t *C, int64_t Crs, const uint64_t *A, int64_t Ars,
               const uint64_t *B, int64_t Brs, int64_t cap, hipStream_t s)
{
    return strided ? launch_aors<true>(Wd, H, C, Crs, A, Ars, B, Brs, cap, sub, s)
                   : launch_aors<false>(Wd, H, C, Crs, A, Ars, B, Brs, cap, sub, s);
}
}  // namespace mgx

extern "C" {

#define MGX_MUL_ARGS int64_t K, int64_t M, int64_t N, mgx_limb *C, int64_t Crs, const mgx_limb *A, int64_t Ars, \
                     const mgx_limb *B, int64_t Brs, int64_t cap, void *stream
#define MGX_AORS_ARGS int64_t Wd, int64_t H, mgx_limb *C, int64_t Crs, const mgx_limb *A, int64_t Ars, \
                      const mgx_limb *B, int64_t Brs, int64_t cap, void *stream

int mgx_launcher_mul(MGX_MUL_ARGS) { return plain_mul(false, false, K, M, N, C, Crs, A, Ars, B, Brs, cap, 2 * cap, as_stream(stream)); }
int mgx_launcher_addmul(MGX_MUL_ARGS) { return plain_mul(false, true, K, M, N, C, Crs, A, Ars, B, Brs, cap, 2 * cap, as_stream(stream)); }
int mgx_launcher_mul_strided(MGX_MUL_ARGS) { return plain_mul(true, false, K, M, N, C, Crs, A, Ars, B, Brs, cap, 2 * cap, as_stream(stream)); }
int mgx_launcher_addmul_strided(MGX_MUL_ARGS) { return plain_mul(true, true, K, M, N, C, Crs, A, Ars, B, Brs, cap, 2 * cap, as_stream(stream)); }
int mgx_launcher_add(MGX_AORS_ARGS) { return plain_aors(false, false, Wd, H, C, Crs, A, Ars, B, Brs, cap, as_stream(stream)); }
int mgx_launcher_sub(MGX_AORS_ARGS) { return plain_aors(false, true, Wd, H, C, Crs, A, Ars, B, Brs, cap, as_stream(stream)); }
int mgx_launcher_add_strided(MGX_AORS_ARGS) { return plain_aors(true, false, Wd, H, C, Crs, A, Ars, B, Brs, cap, as_stream(stream)); }
int mgx_launcher_sub_strided(MGX_AORS_ARGS) { return plain_aors(true, true, Wd, H, C, Crs, A, Ars, B, Brs, cap, as_stream(stream)); }

static int interleave_dir(mgx_limb *Mx, int64_t nentries, int64_t cap, int dir, void *stream)
{
    MGX_TRY(ensure_init());
    if (nentries % MGX_INTERLEAVE_BLOCK) return fail(MGX_ERR_SHAPE, "entry count must be a multiple of %d", MGX_INTERLEAVE_BLOCK);
    hipStream_t s = as_stream(stream);
    size_t bytes = (size_t)nentries * (cap + 1) * 8;
    if (!bytes) return MGX_OK;
    void *tmp = nullptr;
    MGX_TRY(scratch_alloc(&tmp, bytes, s));
    int64_t nblocks = nentries / MGX_INTERLEAVE_BLOCK;
    int64_t total = nentries * (cap + 1);
    unsigned grid = (unsigned)std::min<int64_t>(ceil_div(total, 256), 65536);
    hipLaunchKernelGGL(interleave_kernel, dim3(grid), dim3(256), 0, s, (const uint64_t *)Mx, (uint64_t *)tmp, nblocks,
                       cap + 1, dir);
    MGX_HIP(hipGetLastError());
    MGX_HIP(hipMemcpyAsync(Mx, tmp, bytes, hipMemcpyDeviceToDevice, s));
    return scratch_free(tmp, s);
}
int mgx_interleave(mgx_limb *Mx, int64_t n, int64_t cap, void *stream) { return interleave_dir(Mx, n, cap, 0, stream); }
int mgx_deinterleave(mgx_limb *Mx, int64_t n, int64_t cap, void *stream) { return interleave_dir(Mx, n, cap, 1, stream); }

int mgx_check_error(void *stream) { return check_device_error(as_stream(stream)); }

}  // extern "C"
