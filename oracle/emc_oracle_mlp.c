/*
 * emc_oracle_mlp.c -- FP64 restatement of the orientation-proposal MLP -- TEST INFRASTRUCTURE ONLY
 * (the checker for the HIP forward/backward/Adam kernels; loaded only by tests/ and bench.py's
 * baseline leg).
 *
 * No reference counterpart: the reference always searches every rotation.  The network follows
 * the proposal design: x (log1p of the 2x2-pooled pattern, standardised) -> 512 -> 512 -> C,
 * GELU (erf form) after the two hidden layers, softmax over the C coarse rotation cells,
 * cross-entropy against a target distribution over the cells, Adam (Kingma & Ba 2015).
 *
 * Layouts: W1 [H][D], W2 [H][H], W3 [C][H] (row = output unit), biases [H], [H], [C];
 * activations [B][*] row-major.
 * Pinned by: central finite differences of the loss (test_oracle_mlp.py) and the Adam update's
 * closed form for a constant gradient.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_API __attribute__((visibility("default")))

static double gelu(double x) { return 0.5 * x * (1.0 + erf(x / sqrt(2.0))); }
static double gelu_d(double x)
{
    return 0.5 * (1.0 + erf(x / sqrt(2.0))) + x * exp(-0.5 * x * x) / sqrt(2.0 * M_PI);
}

/* out[b][o] = sum_k in[b][k] W[o][k] + bias[o] */
static void dense(const double *in, int B, int K, const double *W, const double *bias, int O, double *out)
{
#pragma omp parallel for schedule(static)
    for (int b = 0; b < B; b++)
        for (int o = 0; o < O; o++) {
            double acc = bias[o];
            const double *w = W + (size_t)o * K, *x = in + (size_t)b * K;
            for (int k = 0; k < K; k++) acc += x[k] * w[k];
            out[(size_t)b * O + o] = acc;
        }
}

/* Forward.  pre1/pre2 [B][H] pre-activations, h1/h2 [B][H], z [B][C] logits, p [B][C] softmax. */
ORC_API void orc_mlp_forward(const double *x, int B, int D, int H, int C, const double *W1, const double *b1,
                             const double *W2, const double *b2, const double *W3, const double *b3, double *pre1,
                             double *h1, double *pre2, double *h2, double *z, double *p)
{
    dense(x, B, D, W1, b1, H, pre1);
    for (size_t k = 0; k < (size_t)B * H; k++) h1[k] = gelu(pre1[k]);
    dense(h1, B, H, W2, b2, H, pre2);
    for (size_t k = 0; k < (size_t)B * H; k++) h2[k] = gelu(pre2[k]);
    dense(h2, B, H, W3, b3, C, z);
    for (int b = 0; b < B; b++) {
        const double *zb = z + (size_t)b * C;
        double m = zb[0];
        for (int c = 1; c < C; c++) m = zb[c] > m ? zb[c] : m;
        double s = 0.0;
        for (int c = 0; c < C; c++) s += exp(zb[c] - m);
        for (int c = 0; c < C; c++) p[(size_t)b * C + c] = exp(zb[c] - m) / s;
    }
}

/* Mean cross-entropy -(1/B) sum_b sum_c t log p. */
ORC_API double orc_mlp_loss(const double *p, const double *t, int B, int C)
{
    double L = 0.0;
    for (size_t k = 0; k < (size_t)B * C; k++)
        if (t[k] > 0.0) L -= t[k] * log(p[k] > 1e-300 ? p[k] : 1e-300);
    return L / B;
}

/* Backward of the mean cross-entropy: gradients of all parameters (same layouts). */
ORC_API void orc_mlp_backward(const double *x, const double *t, int B, int D, int H, int C, const double *W2,
                              const double *W3, const double *pre1, const double *h1, const double *pre2,
                              const double *h2, const double *p, double *gW1, double *gb1, double *gW2, double *gb2,
                              double *gW3, double *gb3)
{
    double *dz = (double *)malloc(sizeof(double) * (size_t)B * C);
    double *d2 = (double *)malloc(sizeof(double) * (size_t)B * H);
    double *d1 = (double *)malloc(sizeof(double) * (size_t)B * H);
    for (size_t k = 0; k < (size_t)B * C; k++) dz[k] = (p[k] - t[k]) / B;
    /* layer 3 */
#pragma omp parallel for schedule(static)
    for (int o = 0; o < C; o++) {
        double gb = 0.0;
        for (int k = 0; k < H; k++) {
            double acc = 0.0;
            for (int b = 0; b < B; b++) acc += dz[(size_t)b * C + o] * h2[(size_t)b * H + k];
            gW3[(size_t)o * H + k] = acc;
        }
        for (int b = 0; b < B; b++) gb += dz[(size_t)b * C + o];
        gb3[o] = gb;
    }
#pragma omp parallel for schedule(static)
    for (int b = 0; b < B; b++)
        for (int k = 0; k < H; k++) {
            double acc = 0.0;
            for (int o = 0; o < C; o++) acc += dz[(size_t)b * C + o] * W3[(size_t)o * H + k];
            d2[(size_t)b * H + k] = acc * gelu_d(pre2[(size_t)b * H + k]);
        }
    /* layer 2 */
#pragma omp parallel for schedule(static)
    for (int o = 0; o < H; o++) {
        double gb = 0.0;
        for (int k = 0; k < H; k++) {
            double acc = 0.0;
            for (int b = 0; b < B; b++) acc += d2[(size_t)b * H + o] * h1[(size_t)b * H + k];
            gW2[(size_t)o * H + k] = acc;
        }
        for (int b = 0; b < B; b++) gb += d2[(size_t)b * H + o];
        gb2[o] = gb;
    }
#pragma omp parallel for schedule(static)
    for (int b = 0; b < B; b++)
        for (int k = 0; k < H; k++) {
            double acc = 0.0;
            for (int o = 0; o < H; o++) acc += d2[(size_t)b * H + o] * W2[(size_t)o * H + k];
            d1[(size_t)b * H + k] = acc * gelu_d(pre1[(size_t)b * H + k]);
        }
    /* layer 1 */
#pragma omp parallel for schedule(static)
    for (int o = 0; o < H; o++) {
        double gb = 0.0;
        for (int k = 0; k < D; k++) {
            double acc = 0.0;
            for (int b = 0; b < B; b++) acc += d1[(size_t)b * H + o] * x[(size_t)b * D + k];
            gW1[(size_t)o * D + k] = acc;
        }
        for (int b = 0; b < B; b++) gb += d1[(size_t)b * H + o];
        gb1[o] = gb;
    }
    free(dz);
    free(d2);
    free(d1);
}

/* Adam step t (1-based) on n parameters in place. */
ORC_API void orc_adam(double *w, const double *g, double *m, double *v, long n, int t, double lr, double b1,
                      double b2, double eps)
{
    const double c1 = 1.0 - pow(b1, t), c2 = 1.0 - pow(b2, t);
    for (long k = 0; k < n; k++) {
        m[k] = b1 * m[k] + (1.0 - b1) * g[k];
        v[k] = b2 * v[k] + (1.0 - b2) * g[k] * g[k];
        w[k] -= lr * (m[k] / c1) / (sqrt(v[k] / c2) + eps);
    }
}
