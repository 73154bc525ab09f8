// mlp.hip -- the learned cost model: a three-layer ReLU perceptron in FP32
// with hand-written forward, backward and Adam kernels.
//
//   h1 = relu(W1 x + b1), h2 = relu(W2 h1 + b2), y = W3 h2 + b3
//   loss = sum_{b,o: mask} (y - t)^2 / count(mask)
//
// Parameters are packed W1[H][D] b1[H] W2[H][H] b2[H] W3[O][H] b3[O] (the
// same packing the float64 oracle uses).  It only ever chooses a schedule;
// no result of the arithmetic routes passes through it.
#include "common.h"
#include <vector>
#include <cmath>
#include <cstring>

namespace mgx {

// C[m][n] = act( sum_k A(m,k) B(k,n) + bias[n] ), A(m,k) = A[m*as0 + k*as1],
// B(k,n) = B[k*bs0 + n*bs1]; act 0 = identity, 1 = relu.  accumulate: C += .
__global__ void mlp_gemm_kernel(int Mr, int Nc, int Kd, const float *__restrict__ A, int as0, int as1,
                                const float *__restrict__ B, int bs0, int bs1, const float *__restrict__ bias,
                                float *__restrict__ C, int act, int accumulate)
{
    __shared__ float As[16][17], Bs[16][17];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int row = blockIdx.y * 16 + ty, col = blockIdx.x * 16 + tx;
    float acc = 0.f;
    for (int k0 = 0; k0 < Kd; k0 += 16) {
        int ka = k0 + tx, kb = k0 + ty;
        As[ty][tx] = (row < Mr && ka < Kd) ? A[(int64_t)row * as0 + (int64_t)ka * as1] : 0.f;
        Bs[ty][tx] = (kb < Kd && col < Nc) ? B[(int64_t)kb * bs0 + (int64_t)col * bs1] : 0.f;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 16; k++) acc = fmaf(As[ty][k], Bs[k][tx], acc);
        __syncthreads();
    }
    if (row < Mr && col < Nc) {
        if (bias) acc += bias[col];
        if (act == 1) acc = acc > 0.f ? acc : 0.f;
        float *c = C + (int64_t)row * Nc + col;
        *c = accumulate ? *c + acc : acc;
    }
}

// dY = 2 (Y - T) mask / count ; also writes the per-element squared error
__global__ void mlp_loss_kernel(int n, const float *Y, const float *T, const float *mask, float inv_count, float *dY,
                                float *err)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float m = mask[i] != 0.f ? 1.f : 0.f;
    float e = (Y[i] - T[i]) * m;
    dY[i] = 2.f * e * inv_count;
    err[i] = e * e * inv_count;
}

// dZ = dH where H > 0
__global__ void mlp_relu_back_kernel(int n, const float *H, float *dH)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && !(H[i] > 0.f)) dH[i] = 0.f;
}

// column sums of a rows x cols matrix into out (bias gradients)
__global__ void mlp_colsum_kernel(int rows, int cols, const float *X, float *out)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    float s = 0.f;
    for (int r = 0; r < rows; r++) s += X[(int64_t)r * cols + c];
    out[c] = s;
}

__global__ void mlp_sum_kernel(int n, const float *x, float *out)
{
    __shared__ float sh[256];
    float s = 0.f;
    for (int i = threadIdx.x; i < n; i += 256) s += x[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0];
}

// Adam (Kingma & Ba) with bias correction folded into the step size
__global__ void mlp_adam_kernel(int n, float *p, const float *g, float *m, float *v, float lr_t, float b1, float b2,
                                float eps)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float gi = g[i];
    float mi = b1 * m[i] + (1.f - b1) * gi;
    float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr_t * mi / (sqrtf(vi) + eps);
}

// One sample through all three layers in one workgroup (the dispatch's
// inference: one feature vector per product).  The normalised input comes by
// value; thread j computes hidden unit j of each layer (H <= 256).
constexpr int PREDICT1_MAX_D = 32, PREDICT1_MAX_H = 256, PREDICT1_MAX_O = 64;
struct Predict1In {
    float x[PREDICT1_MAX_D];
};
__global__ __launch_bounds__(256) void mlp_predict1_kernel(const float *__restrict__ params, int D, int H, int O,
                                                           Predict1In in, float *__restrict__ out)
{
    __shared__ float h1[PREDICT1_MAX_H], h2[PREDICT1_MAX_H];
    const float *W1 = params, *b1 = W1 + H * D, *W2 = b1 + H, *b2 = W2 + H * H, *W3 = b2 + H, *b3 = W3 + O * H;
    const int j = threadIdx.x;
    if (j < H) {
        float a = b1[j];
        for (int i = 0; i < D; i++) a = fmaf(W1[j * D + i], in.x[i], a);
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
        out[j] = a;
    }
}

}  // namespace mgx

using namespace mgx;

struct mgx_costmodel {
    int D, H, O;
    int64_t np;
    float *params = nullptr, *grad = nullptr, *m = nullptr, *v = nullptr;
    float *d_out = nullptr, *h_out = nullptr;   // one-sample prediction: device result, pinned host copy
    float *d_norm = nullptr;   // mean[D] then scale[D] on the device (the dispatch kernel normalises there)
    int64_t step = 0;
    // feature normalisation (host): x' = (x - mean) / scale
    std::vector<float> mean, scale;
};

namespace {

struct Work {
    float *X, *T, *Mk, *H1, *H2, *Y, *dY, *dH2, *dH1, *err, *loss;
};

int gemm(hipStream_t s, int Mr, int Nc, int Kd, const float *A, int as0, int as1, const float *B, int bs0, int bs1,
         const float *bias, float *C, int act, int acc)
{
    dim3 g((unsigned)ceil_div(Nc, 16), (unsigned)ceil_div(Mr, 16));
    hipLaunchKernelGGL(mlp_gemm_kernel, g, dim3(256), 0, s, Mr, Nc, Kd, A, as0, as1, B, bs0, bs1, bias, C, act, acc);
    MGX_HIP(hipGetLastError());
    return MGX_OK;
}

// forward for a batch already on the device; fills H1, H2, Y
int forward(mgx_costmodel *cm, const Work &w, int Bn, hipStream_t s)
{
    const int D = cm->D, H = cm->H, O = cm->O;
    const float *W1 = cm->params, *b1 = W1 + H * D, *W2 = b1 + H, *b2 = W2 + H * H, *W3 = b2 + H, *b3 = W3 + O * H;
    MGX_TRY(gemm(s, Bn, H, D, w.X, D, 1, W1, 1, D, b1, w.H1, 1, 0));       // X W1^T
    MGX_TRY(gemm(s, Bn, H, H, w.H1, H, 1, W2, 1, H, b2, w.H2, 1, 0));      // H1 W2^T
    MGX_TRY(gemm(s, Bn, O, H, w.H2, H, 1, W3, 1, H, b3, w.Y, 0, 0));       // H2 W3^T
    return MGX_OK;
}

// gradient of the masked MSE into cm->grad, loss into w.loss[0]
int backward(mgx_costmodel *cm, const Work &w, int Bn, float inv_count, hipStream_t s)
{
    const int D = cm->D, H = cm->H, O = cm->O;
    const float *W2 = cm->params + H * D + H, *W3 = W2 + H * H + H;
    float *gW1 = cm->grad, *gb1 = gW1 + H * D, *gW2 = gb1 + H, *gb2 = gW2 + H * H, *gW3 = gb2 + H, *gb3 = gW3 + O * H;
    int n = Bn * O;
    hipLaunchKernelGGL(mlp_loss_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s, n, w.Y, w.T, w.Mk, inv_count,
                       w.dY, w.err);
    hipLaunchKernelGGL(mlp_sum_kernel, dim3(1), dim3(256), 0, s, n, w.err, w.loss);
    MGX_TRY(gemm(s, O, H, Bn, w.dY, 1, O, w.H2, H, 1, nullptr, gW3, 0, 0));   // dY^T H2
    hipLaunchKernelGGL(mlp_colsum_kernel, dim3((unsigned)ceil_div(O, 64)), dim3(64), 0, s, Bn, O, w.dY, gb3);
    MGX_TRY(gemm(s, Bn, H, O, w.dY, O, 1, W3, H, 1, nullptr, w.dH2, 0, 0));   // dY W3
    hipLaunchKernelGGL(mlp_relu_back_kernel, dim3((unsigned)ceil_div(Bn * H, 256)), dim3(256), 0, s, Bn * H, w.H2, w.dH2);
    MGX_TRY(gemm(s, H, H, Bn, w.dH2, 1, H, w.H1, H, 1, nullptr, gW2, 0, 0));  // dH2^T H1
    hipLaunchKernelGGL(mlp_colsum_kernel, dim3((unsigned)ceil_div(H, 64)), dim3(64), 0, s, Bn, H, w.dH2, gb2);
    MGX_TRY(gemm(s, Bn, H, H, w.dH2, H, 1, W2, H, 1, nullptr, w.dH1, 0, 0));  // dH2 W2
    hipLaunchKernelGGL(mlp_relu_back_kernel, dim3((unsigned)ceil_div(Bn * H, 256)), dim3(256), 0, s, Bn * H, w.H1, w.dH1);
    MGX_TRY(gemm(s, H, D, Bn, w.dH1, 1, H, w.X, D, 1, nullptr, gW1, 0, 0));   // dH1^T X
    hipLaunchKernelGGL(mlp_colsum_kernel, dim3((unsigned)ceil_div(H, 64)), dim3(64), 0, s, Bn, H, w.dH1, gb1);
    MGX_HIP(hipGetLastError());
    return MGX_OK;
}

int alloc_work(mgx_costmodel *cm, int Bn, Work &w, hipStream_t s)
{
    const int D = cm->D, H = cm->H, O = cm->O;
    size_t floats = (size_t)Bn * (D + 2 * O + O + 2 * H + O + 2 * H + O) + 16;
    float *base = nullptr;
    MGX_TRY(scratch_alloc((void **)&base, floats * 4, s));
    float *p = base;
    w.X = p; p += (size_t)Bn * D;
    w.T = p; p += (size_t)Bn * O;
    w.Mk = p; p += (size_t)Bn * O;
    w.Y = p; p += (size_t)Bn * O;
    w.H1 = p; p += (size_t)Bn * H;
    w.H2 = p; p += (size_t)Bn * H;
    w.dY = p; p += (size_t)Bn * O;
    w.dH2 = p; p += (size_t)Bn * H;
    w.dH1 = p; p += (size_t)Bn * H;
    w.err = p; p += (size_t)Bn * O;
    w.loss = p;
    return MGX_OK;
}

void normalise(const mgx_costmodel *cm, const float *X, int Bn, std::vector<float> &out)
{
    out.resize((size_t)Bn * cm->D);
    for (int b = 0; b < Bn; b++)
        for (int d = 0; d < cm->D; d++)
            out[(size_t)b * cm->D + d] = (X[(size_t)b * cm->D + d] - cm->mean[d]) / cm->scale[d];
}

}  // namespace

namespace mgx { int costmodel_predict1(mgx_costmodel *cm, const float *x, float *y, hipStream_t s); }

extern "C" {

int mgx_costmodel_destroy(mgx_costmodel *cm);

int mgx_costmodel_create(mgx_costmodel **out, int D, int H, int O, uint64_t seed)
{
    MGX_TRY(ensure_init());
    if (!out || D < 1 || H < 1 || O < 1) return fail(MGX_ERR_ARG, "bad cost-model shape");
    auto *cm = new mgx_costmodel();
    cm->D = D; cm->H = H; cm->O = O;
    cm->np = (int64_t)H * D + H + (int64_t)H * H + H + (int64_t)O * H + O;
    cm->mean.assign(D, 0.f);
    cm->scale.assign(D, 1.f);
    std::vector<float> p(cm->np, 0.f);
    // He-uniform initialisation from a seeded SplitMix64 stream
    uint64_t x = seed;
    auto rnd = [&]() {
        uint64_t z = (x += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        z ^= z >> 31;
        return (float)((double)(z >> 11) / 9007199254740992.0) * 2.f - 1.f;
    };
    int64_t o = 0;
    float a1 = std::sqrt(6.f / D), a2 = std::sqrt(6.f / H);
    for (int64_t i = 0; i < (int64_t)H * D; i++) p[o++] = a1 * rnd();
    o += H;
    for (int64_t i = 0; i < (int64_t)H * H; i++) p[o++] = a2 * rnd();
    o += H;
    for (int64_t i = 0; i < (int64_t)O * H; i++) p[o++] = a2 * rnd();
    size_t bytes = (size_t)cm->np * 4;
    if (hipMalloc(&cm->params, bytes) != hipSuccess || hipMalloc(&cm->grad, bytes) != hipSuccess ||
        hipMalloc(&cm->m, bytes) != hipSuccess || hipMalloc(&cm->v, bytes) != hipSuccess ||
        hipMalloc(&cm->d_out, 4 * PREDICT1_MAX_O) != hipSuccess || hipMalloc(&cm->d_norm, 8 * (size_t)D) != hipSuccess ||
        hipHostMalloc(&cm->h_out, 4 * PREDICT1_MAX_O, hipHostMallocDefault) != hipSuccess) {
        mgx_costmodel_destroy(cm);
        return fail(MGX_ERR_OOM, "cost-model allocation");
    }
    MGX_HIP(hipMemcpy(cm->params, p.data(), bytes, hipMemcpyHostToDevice));
    MGX_HIP(hipMemset(cm->m, 0, bytes));
    MGX_HIP(hipMemset(cm->v, 0, bytes));
    MGX_HIP(hipMemset(cm->grad, 0, bytes));
    MGX_HIP(hipMemcpy(cm->d_norm, cm->mean.data(), 4 * (size_t)D, hipMemcpyHostToDevice));
    MGX_HIP(hipMemcpy(cm->d_norm + D, cm->scale.data(), 4 * (size_t)D, hipMemcpyHostToDevice));
    *out = cm;
    return MGX_OK;
}

int mgx_costmodel_destroy(mgx_costmodel *cm)
{
    if (!cm) return MGX_OK;
    hipFree(cm->params); hipFree(cm->grad); hipFree(cm->m); hipFree(cm->v); hipFree(cm->d_out); hipFree(cm->d_norm);
    if (cm->h_out) hipHostFree(cm->h_out);
    delete cm;
    return MGX_OK;
}

int64_t mgx_costmodel_param_count(const mgx_costmodel *cm) { return cm ? cm->np : -1; }
int64_t mgx_costmodel_outputs(const mgx_costmodel *cm) { return cm ? cm->O : -1; }

int mgx_costmodel_get_params(const mgx_costmodel *cm, float *out)
{
    MGX_HIP(hipMemcpy(out, cm->params, (size_t)cm->np * 4, hipMemcpyDeviceToHost));
    return MGX_OK;
}

int mgx_costmodel_set_params(mgx_costmodel *cm, const float *in)
{
    MGX_HIP(hipMemcpy(cm->params, in, (size_t)cm->np * 4, hipMemcpyHostToDevice));
    return MGX_OK;
}

// Feature normalisation applied by train/predict (mean and scale per input).
int mgx_costmodel_set_normalisation(mgx_costmodel *cm, const float *mean, const float *scale)
{
    for (int d = 0; d < cm->D; d++) {
        cm->mean[d] = mean[d];
        cm->scale[d] = scale[d] > 0 ? scale[d] : 1.f;
    }
    MGX_HIP(hipMemcpy(cm->d_norm, cm->mean.data(), 4 * (size_t)cm->D, hipMemcpyHostToDevice));
    MGX_HIP(hipMemcpy(cm->d_norm + cm->D, cm->scale.data(), 4 * (size_t)cm->D, hipMemcpyHostToDevice));
    return MGX_OK;
}

// Text form written by costmodel.py: "mgx-costmodel 1", then D H O, then D
// means, D scales and the packed parameters, whitespace separated.
int mgx_costmodel_load(mgx_costmodel **out, const char *path)
{
    if (!out || !path) return fail(MGX_ERR_ARG, "null argument");
    FILE *f = fopen(path, "r");
    if (!f) return fail(MGX_ERR_ARG, "cannot open cost model %s", path);
    char tag[32] = {0};
    int version = 0, D = 0, H = 0, O = 0;
    int st = MGX_OK;
    if (fscanf(f, "%31s %d %d %d %d", tag, &version, &D, &H, &O) != 5 || strcmp(tag, "mgx-costmodel") != 0 ||
        version != 1 || D < 1 || H < 1 || O < 1 || D > 4096 || H > 4096 || O > 4096)
        st = fail(MGX_ERR_PARSE, "%s is not a cost model file", path);
    mgx_costmodel *cm = nullptr;
    if (st == MGX_OK) st = mgx_costmodel_create(&cm, D, H, O, 0);
    if (st == MGX_OK) {
        std::vector<float> mean(D), scale(D), p(cm->np);
        bool ok = true;
        for (auto *v : {&mean, &scale, &p})
            for (float &x : *v) ok = ok && fscanf(f, "%f", &x) == 1;
        if (!ok) st = fail(MGX_ERR_PARSE, "%s: truncated cost model", path);
        if (st == MGX_OK) st = mgx_costmodel_set_params(cm, p.data());
        if (st == MGX_OK) st = mgx_costmodel_set_normalisation(cm, mean.data(), scale.data());
        if (st != MGX_OK) {
            mgx_costmodel_destroy(cm);
            cm = nullptr;
        }
    }
    fclose(f);
    if (st == MGX_OK) *out = cm;
    return st;
}

// Loss and gradient (into grad_out, host) for a batch; X is raw (normalised here).
int mgx_costmodel_loss_grad(mgx_costmodel *cm, const float *X, const float *T, const float *mask, int Bn,
                            float *grad_out, float *loss_out)
{
    hipStream_t s = nullptr;
    Work w;
    MGX_TRY(alloc_work(cm, Bn, w, s));
    std::vector<float> xn;
    normalise(cm, X, Bn, xn);
    double cnt = 0;
    for (int64_t i = 0; i < (int64_t)Bn * cm->O; i++) cnt += mask[i] != 0.f;
    MGX_HIP(hipMemcpyAsync(w.X, xn.data(), xn.size() * 4, hipMemcpyHostToDevice, s));
    MGX_HIP(hipMemcpyAsync(w.T, T, (size_t)Bn * cm->O * 4, hipMemcpyHostToDevice, s));
    MGX_HIP(hipMemcpyAsync(w.Mk, mask, (size_t)Bn * cm->O * 4, hipMemcpyHostToDevice, s));
    MGX_TRY(forward(cm, w, Bn, s));
    MGX_TRY(backward(cm, w, Bn, (float)(1.0 / (cnt > 0 ? cnt : 1)), s));
    if (grad_out) MGX_HIP(hipMemcpyAsync(grad_out, cm->grad, (size_t)cm->np * 4, hipMemcpyDeviceToHost, s));
    float l = 0;
    MGX_HIP(hipMemcpyAsync(&l, w.loss, 4, hipMemcpyDeviceToHost, s));
    MGX_TRY(scratch_free(w.X, s));
    MGX_HIP(hipStreamSynchronize(s));
    if (loss_out) *loss_out = l;
    return MGX_OK;
}

// Full-batch Adam for `epochs` steps; returns the final loss.
int mgx_costmodel_train(mgx_costmodel *cm, const float *X, const float *T, const float *mask, int Bn, int epochs,
                        float lr, float *loss_out)
{
    hipStream_t s = nullptr;
    Work w;
    MGX_TRY(alloc_work(cm, Bn, w, s));
    std::vector<float> xn;
    normalise(cm, X, Bn, xn);
    double cnt = 0;
    for (int64_t i = 0; i < (int64_t)Bn * cm->O; i++) cnt += mask[i] != 0.f;
    MGX_HIP(hipMemcpyAsync(w.X, xn.data(), xn.size() * 4, hipMemcpyHostToDevice, s));
    MGX_HIP(hipMemcpyAsync(w.T, T, (size_t)Bn * cm->O * 4, hipMemcpyHostToDevice, s));
    MGX_HIP(hipMemcpyAsync(w.Mk, mask, (size_t)Bn * cm->O * 4, hipMemcpyHostToDevice, s));
    const float b1 = 0.9f, b2 = 0.999f;
    for (int e = 0; e < epochs; e++) {
        MGX_TRY(forward(cm, w, Bn, s));
        MGX_TRY(backward(cm, w, Bn, (float)(1.0 / (cnt > 0 ? cnt : 1)), s));
        cm->step++;
        float lr_t = lr * std::sqrt(1.f - std::pow(b2, (float)cm->step)) / (1.f - std::pow(b1, (float)cm->step));
        hipLaunchKernelGGL(mlp_adam_kernel, dim3((unsigned)ceil_div(cm->np, 256)), dim3(256), 0, s, (int)cm->np,
                           cm->params, cm->grad, cm->m, cm->v, lr_t, b1, b2, 1e-8f);
    }
    MGX_TRY(forward(cm, w, Bn, s));
    MGX_TRY(backward(cm, w, Bn, (float)(1.0 / (cnt > 0 ? cnt : 1)), s));
    float l = 0;
    MGX_HIP(hipMemcpyAsync(&l, w.loss, 4, hipMemcpyDeviceToHost, s));
    MGX_TRY(scratch_free(w.X, s));
    MGX_HIP(hipStreamSynchronize(s));
    if (loss_out) *loss_out = l;
    return MGX_OK;
}

int mgx_costmodel_predict(mgx_costmodel *cm, const float *X, int Bn, float *Y)
{
    if (Bn == 1 && cm->D <= PREDICT1_MAX_D && cm->H <= PREDICT1_MAX_H && cm->O <= PREDICT1_MAX_O)
        return costmodel_predict1(cm, X, Y, nullptr);   // one sample: the fused kernel
    hipStream_t s = nullptr;
    Work w;
    MGX_TRY(alloc_work(cm, Bn, w, s));
    std::vector<float> xn;
    normalise(cm, X, Bn, xn);
    MGX_HIP(hipMemcpyAsync(w.X, xn.data(), xn.size() * 4, hipMemcpyHostToDevice, s));
    MGX_TRY(forward(cm, w, Bn, s));
    MGX_HIP(hipMemcpyAsync(Y, w.Y, (size_t)Bn * cm->O * 4, hipMemcpyDeviceToHost, s));
    MGX_TRY(scratch_free(w.X, s));
    MGX_HIP(hipStreamSynchronize(s));
    return MGX_OK;
}

}  // extern "C"

namespace mgx {
// Device pointers and shape of a model, for kernels that run its inference
// inside a larger launch sequence.
int costmodel_view(const mgx_costmodel *cm, const float **params, const float **norm, int *D, int *H, int *O)
{
    if (cm->H > PREDICT1_MAX_H || cm->D > PREDICT1_MAX_D || cm->O > PREDICT1_MAX_O) return 1;
    *params = cm->params;
    *norm = cm->d_norm;
    *D = cm->D;
    *H = cm->H;
    *O = cm->O;
    return 0;
}

// The dispatch's inference for one raw feature vector x (normalised here),
// on the caller's stream: one launch, one 4 O-byte copy, one synchronization.
int costmodel_predict1(mgx_costmodel *cm, const float *x, float *y, hipStream_t s)
{
    if (cm->D > PREDICT1_MAX_D || cm->H > PREDICT1_MAX_H || cm->O > PREDICT1_MAX_O)
        return mgx_costmodel_predict(cm, x, 1, y);
    Predict1In in{};
    for (int d = 0; d < cm->D; d++) in.x[d] = (x[d] - cm->mean[d]) / cm->scale[d];
    hipLaunchKernelGGL(mlp_predict1_kernel, dim3(1), dim3(256), 0, s, cm->params, cm->D, cm->H, cm->O, in, cm->d_out);
    MGX_HIP(hipGetLastError());
    MGX_HIP(hipMemcpyAsync(cm->h_out, cm->d_out, (size_t)cm->O * 4, hipMemcpyDeviceToHost, s));
    MGX_HIP(hipStreamSynchronize(s));
    memcpy(y, cm->h_out, (size_t)cm->O * 4);
    return MGX_OK;
}
}  // namespace mgx
