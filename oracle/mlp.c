/*
 * mlp.c -- oracle (test infrastructure only): the cost model's three-layer
 * ReLU perceptron in float64, forward pass and reverse-mode gradient of the
 * masked mean-squared error.  The reference has no learned component; this
 * is textbook backpropagation, used to check the HIP training kernels.
 */
#include "oracle.h"
#include <stdlib.h>
#include <string.h>

int64_t or_mlp_param_count(int D, int H, int O)
{
    return (int64_t)H * D + H + (int64_t)H * H + H + (int64_t)O * H + O;
}

double or_mlp_loss_grad(int D, int H, int O, int B, const double *params, const double *X, const double *T,
                        const double *mask, double *grad, double *Y)
{
    const double *W1 = params, *b1 = W1 + (int64_t)H * D, *W2 = b1 + H, *b2 = W2 + (int64_t)H * H,
                 *W3 = b2 + H, *b3 = W3 + (int64_t)O * H;
    double *gW1 = grad, *gb1 = gW1 + (int64_t)H * D, *gW2 = gb1 + H, *gb2 = gW2 + (int64_t)H * H,
           *gW3 = gb2 + H, *gb3 = gW3 + (int64_t)O * H;
    if (grad) memset(grad, 0, (size_t)or_mlp_param_count(D, H, O) * sizeof(double));
    double *h1 = (double *)malloc(sizeof(double) * (size_t)H), *h2 = (double *)malloc(sizeof(double) * (size_t)H);
    double *d1 = (double *)malloc(sizeof(double) * (size_t)H), *d2 = (double *)malloc(sizeof(double) * (size_t)H);
    double *dy = (double *)malloc(sizeof(double) * (size_t)O);
    double cnt = 0;
    for (int64_t i = 0; i < (int64_t)B * O; i++) cnt += mask[i] != 0;
    if (cnt == 0) cnt = 1;
    double loss = 0;
    for (int b = 0; b < B; b++) {
        const double *x = X + (int64_t)b * D;
        for (int h = 0; h < H; h++) {
            double s = b1[h];
            for (int d = 0; d < D; d++) s += W1[(int64_t)h * D + d] * x[d];
            h1[h] = s > 0 ? s : 0;
        }
        for (int h = 0; h < H; h++) {
            double s = b2[h];
            for (int g = 0; g < H; g++) s += W2[(int64_t)h * H + g] * h1[g];
            h2[h] = s > 0 ? s : 0;
        }
        for (int o = 0; o < O; o++) {
            double s = b3[o];
            for (int h = 0; h < H; h++) s += W3[(int64_t)o * H + h] * h2[h];
            if (Y) Y[(int64_t)b * O + o] = s;
            double m = mask[(int64_t)b * O + o] != 0;
            double e = (s - T[(int64_t)b * O + o]) * m;
            loss += e * e / cnt;
            dy[o] = 2 * e / cnt;
        }
        if (!grad) continue;
        for (int o = 0; o < O; o++) {
            gb3[o] += dy[o];
            for (int h = 0; h < H; h++) gW3[(int64_t)o * H + h] += dy[o] * h2[h];
        }
        for (int h = 0; h < H; h++) {
            double s = 0;
            for (int o = 0; o < O; o++) s += W3[(int64_t)o * H + h] * dy[o];
            d2[h] = h2[h] > 0 ? s : 0;
        }
        for (int h = 0; h < H; h++) {
            gb2[h] += d2[h];
            for (int g = 0; g < H; g++) gW2[(int64_t)h * H + g] += d2[h] * h1[g];
        }
        for (int g = 0; g < H; g++) {
            double s = 0;
            for (int h = 0; h < H; h++) s += W2[(int64_t)h * H + g] * d2[h];
            d1[g] = h1[g] > 0 ? s : 0;
        }
        for (int h = 0; h < H; h++) {
            gb1[h] += d1[h];
            for (int d = 0; d < D; d++) gW1[(int64_t)h * D + d] += d1[h] * x[d];
        }
    }
    free(h1); free(h2); free(d1); free(d2); free(dy);
    return loss;
}
