/*
 * og_learn.c -- FP64 oracle of the learned tile smoother: a 3-layer 3D conv
 * net on one halo'd 10^3 tile (1 -> 16 -> 16 -> 1 channels, 3^3 kernels,
 * zero "same" padding, ReLU after layers 1 and 2, no biases), its loss, the
 * gradient of the loss with respect to the weights, and the Adam update.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  No reference counterpart.
 *
 * Layout.  A tile volume is 1000 values, index (i * 10 + j) * 10 + k.  The
 * weights are one flat array of OG_NET_W = 7776 values:
 *   W1[16][27] at 0, W2[16 out][16 in][27] at 432, W3[16 in][27] at 7344,
 * tap index (di + 1) * 9 + (dj + 1) * 3 + (dk + 1), offsets di, dj, dk in -1..1.
 * Loss of one tile, d = y - e on the 10^3 volume:
 *   L = sum over the interior 8^3 cells p of (6 d[p] - sum of the 6 neighbours of p)^2
 * and the batch loss is the mean of L over the tiles.
 */
#include "og_internal.h"

#include <stdlib.h>
#include <string.h>

#define NV 1000
#define NC 16

static inline int vid(int i, int j, int k) { return (i * 10 + j) * 10 + k; }

/* out[co][p] = sum_ci sum_tap w[co][ci][tap] in[ci][p + off(tap)], zero outside */
static void conv(const double* w, int cin, int cout, const double* in, double* out)
{
    for (int co = 0; co < cout; co++)
        for (int i = 0; i < 10; i++)
            for (int j = 0; j < 10; j++)
                for (int k = 0; k < 10; k++) {
                    double s = 0;
                    for (int ci = 0; ci < cin; ci++)
                        for (int t = 0; t < 27; t++) {
                            int a = i + t / 9 - 1, b = j + (t / 3) % 3 - 1, c = k + t % 3 - 1;
                            if (a < 0 || a > 9 || b < 0 || b > 9 || c < 0 || c > 9) continue;
                            s += w[((size_t)co * cin + ci) * 27 + t] * in[(size_t)ci * NV + vid(a, b, c)];
                        }
                    out[(size_t)co * NV + vid(i, j, k)] = s;
                }
}

/* gin[ci][q] += sum_co sum_tap w[co][ci][tap] gout[co][q - off(tap)];
 * gw[co][ci][tap] += sum_p gout[co][p] in[ci][p + off(tap)] */
static void conv_back(const double* w, int cin, int cout, const double* in, const double* gout, double* gin, double* gw)
{
    for (int co = 0; co < cout; co++)
        for (int i = 0; i < 10; i++)
            for (int j = 0; j < 10; j++)
                for (int k = 0; k < 10; k++) {
                    double g = gout[(size_t)co * NV + vid(i, j, k)];
                    if (g == 0) continue;
                    for (int ci = 0; ci < cin; ci++)
                        for (int t = 0; t < 27; t++) {
                            int a = i + t / 9 - 1, b = j + (t / 3) % 3 - 1, c = k + t % 3 - 1;
                            if (a < 0 || a > 9 || b < 0 || b > 9 || c < 0 || c > 9) continue;
                            size_t wi = ((size_t)co * cin + ci) * 27 + t, xi = (size_t)ci * NV + vid(a, b, c);
                            if (gw) gw[wi] += g * in[xi];
                            if (gin) gin[xi] += g * w[wi];
                        }
                }
}

void og_net_forward(const double* w, const double* x, double* y)
{
    double* a1 = (double*)malloc(sizeof(double) * NC * NV);
    double* a2 = (double*)malloc(sizeof(double) * NC * NV);
    conv(w, 1, NC, x, a1);
    for (int i = 0; i < NC * NV; i++) a1[i] = a1[i] > 0 ? a1[i] : 0;
    conv(w + 432, NC, NC, a1, a2);
    for (int i = 0; i < NC * NV; i++) a2[i] = a2[i] > 0 ? a2[i] : 0;
    /* layer 3: W3[in][tap] as a [1 out][16 in][27] kernel */
    conv(w + 7344, NC, 1, a2, y);
    free(a1);
    free(a2);
}

/* s = 7-point Laplacian of d on the interior 8^3, 0 elsewhere */
static void lap_interior(const double* d, double* s)
{
    memset(s, 0, sizeof(double) * NV);
    for (int i = 1; i <= 8; i++)
        for (int j = 1; j <= 8; j++)
            for (int k = 1; k <= 8; k++)
                s[vid(i, j, k)] = 6 * d[vid(i, j, k)] - d[vid(i - 1, j, k)] - d[vid(i + 1, j, k)] - d[vid(i, j - 1, k)] -
                                  d[vid(i, j + 1, k)] - d[vid(i, j, k - 1)] - d[vid(i, j, k + 1)];
}

/* mean loss over n tiles (x, e: n x 1000 floats) and its gradient (OG_NET_W values) */
double og_net_loss_grad(const double* w, int n, const float* x, const float* e, double* grad)
{
    double* xin = (double*)malloc(sizeof(double) * NV);
    double* a1 = (double*)malloc(sizeof(double) * NC * NV);
    double* a2 = (double*)malloc(sizeof(double) * NC * NV);
    double* y = (double*)malloc(sizeof(double) * NV);
    double* d = (double*)malloc(sizeof(double) * NV);
    double* s = (double*)malloc(sizeof(double) * NV);
    double* gy = (double*)malloc(sizeof(double) * NV);
    double* g2 = (double*)malloc(sizeof(double) * NC * NV);
    double* g1 = (double*)malloc(sizeof(double) * NC * NV);
    if (grad) memset(grad, 0, sizeof(double) * OG_NET_W);
    double loss = 0;
    for (int b = 0; b < n; b++) {
        for (int i = 0; i < NV; i++) xin[i] = x[(size_t)b * NV + i];
        conv(w, 1, NC, xin, a1);
        for (int i = 0; i < NC * NV; i++) a1[i] = a1[i] > 0 ? a1[i] : 0;
        conv(w + 432, NC, NC, a1, a2);
        for (int i = 0; i < NC * NV; i++) a2[i] = a2[i] > 0 ? a2[i] : 0;
        conv(w + 7344, NC, 1, a2, y);
        for (int i = 0; i < NV; i++) d[i] = y[i] - e[(size_t)b * NV + i];
        lap_interior(d, s);
        for (int i = 0; i < NV; i++) loss += s[i] * s[i];
        if (!grad) continue;
        /* dL/dy = 2 Lap^T s (s is zero off the interior) */
        for (int i = 0; i < 10; i++)
            for (int j = 0; j < 10; j++)
                for (int k = 0; k < 10; k++) {
                    double v = 6 * s[vid(i, j, k)];
                    if (i > 0) v -= s[vid(i - 1, j, k)];
                    if (i < 9) v -= s[vid(i + 1, j, k)];
                    if (j > 0) v -= s[vid(i, j - 1, k)];
                    if (j < 9) v -= s[vid(i, j + 1, k)];
                    if (k > 0) v -= s[vid(i, j, k - 1)];
                    if (k < 9) v -= s[vid(i, j, k + 1)];
                    gy[vid(i, j, k)] = 2 * v / n;
                }
        memset(g2, 0, sizeof(double) * NC * NV);
        conv_back(w + 7344, NC, 1, a2, gy, g2, grad + 7344);
        for (int i = 0; i < NC * NV; i++) if (a2[i] <= 0) g2[i] = 0;
        memset(g1, 0, sizeof(double) * NC * NV);
        conv_back(w + 432, NC, NC, a1, g2, g1, grad + 432);
        for (int i = 0; i < NC * NV; i++) if (a1[i] <= 0) g1[i] = 0;
        conv_back(w, 1, NC, xin, g1, NULL, grad);
    }
    free(xin); free(a1); free(a2); free(y); free(d); free(s); free(gy); free(g2); free(g1);
    return loss / n;
}

/* Adam (Kingma and Ba, 2015) at step t >= 1 */
void og_adam_step(double* w, double* m, double* v, const double* g, int n, int t, double lr, double b1, double b2, double eps)
{
    double c1 = 1 - pow(b1, t), c2 = 1 - pow(b2, t);
    for (int i = 0; i < n; i++) {
        m[i] = b1 * m[i] + (1 - b1) * g[i];
        v[i] = b2 * v[i] + (1 - b2) * g[i] * g[i];
        w[i] -= lr * (m[i] / c1) / (sqrt(v[i] / c2) + eps);
    }
}
