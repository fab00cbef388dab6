/*
 * Double-precision CPU reference of the learned policy's cost model.
 * TEST INFRASTRUCTURE ONLY (same rules as oracle.c).
 *
 * Network: 8 -> 32 -> 32 -> 3, ReLU after the two hidden layers, linear out.
 * Parameters packed as float32[1443]:
 *   W1[32][8], b1[32], W2[32][32], b2[32], W3[3][32], b3[3]   (row-major, out x in)
 * Loss over a batch of B samples with targets t[B][3] and 0/1 mask[B][3]:
 *   loss = 1/(2B) * sum_b sum_j mask[b][j] * (y[b][j] - t[b][j])^2
 * No reference counterpart (the reference uses a fixed heuristic).
 * Pinned by central finite differences of oracle_mlp_loss (tests), which
 * checks every analytic gradient entry independently of how it is derived.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define NI 8
#define NH 32
#define NO 3
#define OFF_W1 0
#define OFF_B1 (OFF_W1 + NH * NI)
#define OFF_W2 (OFF_B1 + NH)
#define OFF_B2 (OFF_W2 + NH * NH)
#define OFF_W3 (OFF_B2 + NH)
#define OFF_B3 (OFF_W3 + NO * NH)
#define NPARAM (OFF_B3 + NO)

int32_t oracle_mlp_nparam(void) { return NPARAM; }

static void fwd1(const double *p, const float *x, double *h1, double *h2, double *y)
{
    for (int o = 0; o < NH; o++) {
        double s = p[OFF_B1 + o];
        for (int i = 0; i < NI; i++) s += p[OFF_W1 + o * NI + i] * (double)x[i];
        h1[o] = s > 0 ? s : 0;
    }
    for (int o = 0; o < NH; o++) {
        double s = p[OFF_B2 + o];
        for (int i = 0; i < NH; i++) s += p[OFF_W2 + o * NH + i] * h1[i];
        h2[o] = s > 0 ? s : 0;
    }
    for (int o = 0; o < NO; o++) {
        double s = p[OFF_B3 + o];
        for (int i = 0; i < NH; i++) s += p[OFF_W3 + o * NH + i] * h2[i];
        y[o] = s;
    }
}

static double *widen(const float *params)
{
    double *p = (double *)malloc(sizeof(double) * NPARAM);
    for (int i = 0; i < NPARAM; i++) p[i] = (double)params[i];
    return p;
}

void oracle_mlp_forward(const float *params, const float *x, int32_t B, double *out)
{
    double *p = widen(params);
    double h1[NH], h2[NH];
    for (int b = 0; b < B; b++) fwd1(p, x + (size_t)b * NI, h1, h2, out + (size_t)b * NO);
    free(p);
}

double oracle_mlp_loss(const double *p, const float *x, const float *t, const float *mask,
                       int32_t B)
{
    double h1[NH], h2[NH], y[NO], loss = 0;
    for (int b = 0; b < B; b++) {
        fwd1(p, x + (size_t)b * NI, h1, h2, y);
        for (int j = 0; j < NO; j++) {
            double d = y[j] - (double)t[(size_t)b * NO + j];
            loss += (double)mask[(size_t)b * NO + j] * d * d;
        }
    }
    return loss / (2.0 * B);
}

/* Loss and its analytic gradient w.r.t. all 1443 parameters (grad[NPARAM]). */
double oracle_mlp_loss_grad(const float *params, const float *x, const float *t,
                            const float *mask, int32_t B, double *grad)
{
    double *p = widen(params);
    memset(grad, 0, sizeof(double) * NPARAM);
    double h1[NH], h2[NH], y[NO], dy[NO], dh2[NH], dh1[NH], loss = 0;
    for (int b = 0; b < B; b++) {
        const float *xb = x + (size_t)b * NI;
        fwd1(p, xb, h1, h2, y);
        for (int j = 0; j < NO; j++) {
            double d = y[j] - (double)t[(size_t)b * NO + j];
            double mk = (double)mask[(size_t)b * NO + j];
            loss += mk * d * d;
            dy[j] = mk * d / B;
        }
        for (int j = 0; j < NO; j++) {
            grad[OFF_B3 + j] += dy[j];
            for (int i = 0; i < NH; i++) grad[OFF_W3 + j * NH + i] += dy[j] * h2[i];
        }
        for (int i = 0; i < NH; i++) {
            double s = 0;
            for (int j = 0; j < NO; j++) s += p[OFF_W3 + j * NH + i] * dy[j];
            dh2[i] = h2[i] > 0 ? s : 0;
        }
        for (int o = 0; o < NH; o++) {
            grad[OFF_B2 + o] += dh2[o];
            for (int i = 0; i < NH; i++) grad[OFF_W2 + o * NH + i] += dh2[o] * h1[i];
        }
        for (int i = 0; i < NH; i++) {
            double s = 0;
            for (int o = 0; o < NH; o++) s += p[OFF_W2 + o * NH + i] * dh2[o];
            dh1[i] = h1[i] > 0 ? s : 0;
        }
        for (int o = 0; o < NH; o++) {
            grad[OFF_B1 + o] += dh1[o];
            for (int i = 0; i < NI; i++) grad[OFF_W1 + o * NI + i] += dh1[o] * (double)xb[i];
        }
    }
    free(p);
    return loss / (2.0 * B);
}
