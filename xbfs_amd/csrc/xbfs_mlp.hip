// Learned direction/queue cost model: hand-written HIP kernels for the FP32
// MLP 8 -> 32 -> 32 -> 3 (ReLU) that predicts the log2 time of each mode at
// a level (no reference counterpart; it replaces the decision block
// bfs_adaptiveFQ_wb_async.cuh:1047-1078 when loaded).
//
// Parameters are packed as W1[32][8], b1[32], W2[32][32], b2[32], W3[3][32],
// b3[3] (1443 floats).  Loss over n samples with targets y[n][3] and 0/1
// mask[n][3]:  L = 1/(2n) sum_i sum_j mask_ij (out_ij - y_ij)^2.
// Gradients are reduced deterministically: each block writes a partial
// gradient, a second kernel sums the partials in block order.  The network is
// far too small for matrix cores; everything runs on the VALU.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "xbfs_policy.h"

#define XB_MLP_BLOCK 256
#define XB_MLP_CHUNK 64 /* samples staged in LDS per block iteration */
#define XB_OFF_W1 0
#define XB_OFF_B1 (XB_OFF_W1 + XB_MLP_NH * XB_MLP_NI)
#define XB_OFF_W2 (XB_OFF_B1 + XB_MLP_NH)
#define XB_OFF_B2 (XB_OFF_W2 + XB_MLP_NH * XB_MLP_NH)
#define XB_OFF_W3 (XB_OFF_B2 + XB_MLP_NH)
#define XB_OFF_B3 (XB_OFF_W3 + XB_MLP_NO * XB_MLP_NH)

// forward for many samples: one thread per sample
__global__ void __launch_bounds__(XB_MLP_BLOCK) k_mlp_forward(const float *__restrict__ p,
                                                              const float *__restrict__ X,
                                                              int32_t n, float *__restrict__ out)
{
    __shared__ float sp[XB_MLP_NPARAM];
    for (int i = threadIdx.x; i < XB_MLP_NPARAM; i += blockDim.x) sp[i] = p[i];
    __syncthreads();
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < n;
         s += (int64_t)gridDim.x * blockDim.x) {
        float x[XB_MLP_NI], y[XB_MLP_NO];
        for (int i = 0; i < XB_MLP_NI; i++) x[i] = X[s * XB_MLP_NI + i];
        xbfs_mlp_forward1(sp, x, y);
        for (int j = 0; j < XB_MLP_NO; j++) out[s * XB_MLP_NO + j] = y[j];
    }
}

// Partial loss and gradient per block.  Each block walks chunks of 64
// samples: thread t < 64 runs sample t's forward and backward pass and stages
// its activations (x, h1, h2) and deltas (d1, d2, dy) in LDS; then all 256
// threads accumulate their share of the 1443 gradient entries over the chunk.
__global__ void __launch_bounds__(XB_MLP_BLOCK) k_mlp_grad(const float *__restrict__ p,
                                                           const float *__restrict__ X,
                                                           const float *__restrict__ Y,
                                                           const float *__restrict__ M, int32_t n,
                                                           float *__restrict__ part_grad,
                                                           float *__restrict__ part_loss)
{
    __shared__ float sp[XB_MLP_NPARAM];
    __shared__ float sx[XB_MLP_CHUNK][XB_MLP_NI];
    __shared__ float sh1[XB_MLP_CHUNK][XB_MLP_NH];
    __shared__ float sh2[XB_MLP_CHUNK][XB_MLP_NH];
    __shared__ float sd1[XB_MLP_CHUNK][XB_MLP_NH];
    __shared__ float sd2[XB_MLP_CHUNK][XB_MLP_NH];
    __shared__ float sdy[XB_MLP_CHUNK][XB_MLP_NO];
    __shared__ float sloss[XB_MLP_CHUNK];
    for (int i = threadIdx.x; i < XB_MLP_NPARAM; i += blockDim.x) sp[i] = p[i];
    const float *W1 = sp + XB_OFF_W1, *b1 = sp + XB_OFF_B1, *W2 = sp + XB_OFF_W2;
    const float *b2 = sp + XB_OFF_B2, *W3 = sp + XB_OFF_W3, *b3 = sp + XB_OFF_B3;
    const float inv_n = 1.0f / (float)n;
    // each thread owns gradient entries e = tid + k*256 (k < 6)
    float acc[(XB_MLP_NPARAM + XB_MLP_BLOCK - 1) / XB_MLP_BLOCK];
    for (int k = 0; k < (int)(sizeof(acc) / sizeof(acc[0])); k++) acc[k] = 0.0f;
    float loss_acc = 0.0f;
    __syncthreads();
    const int64_t nchunks = ((int64_t)n + XB_MLP_CHUNK - 1) / XB_MLP_CHUNK;
    for (int64_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
        const int t = threadIdx.x;
        if (t < XB_MLP_CHUNK) {
            const int64_t s = c * XB_MLP_CHUNK + t;
            float x[XB_MLP_NI], h1[XB_MLP_NH], h2[XB_MLP_NH], y[XB_MLP_NO], dy[XB_MLP_NO];
            const bool valid = s < n;
            for (int i = 0; i < XB_MLP_NI; i++) x[i] = valid ? X[s * XB_MLP_NI + i] : 0.0f;
            for (int o = 0; o < XB_MLP_NH; o++) {
                float a = b1[o];
                for (int i = 0; i < XB_MLP_NI; i++) a += W1[o * XB_MLP_NI + i] * x[i];
                h1[o] = a > 0.0f ? a : 0.0f;
            }
            for (int o = 0; o < XB_MLP_NH; o++) {
                float a = b2[o];
                for (int i = 0; i < XB_MLP_NH; i++) a += W2[o * XB_MLP_NH + i] * h1[i];
                h2[o] = a > 0.0f ? a : 0.0f;
            }
            float l = 0.0f;
            for (int j = 0; j < XB_MLP_NO; j++) {
                float a = b3[j];
                for (int i = 0; i < XB_MLP_NH; i++) a += W3[j * XB_MLP_NH + i] * h2[i];
                y[j] = a;
                const float mk = valid ? M[s * XB_MLP_NO + j] : 0.0f;
                const float d = valid ? (a - Y[s * XB_MLP_NO + j]) : 0.0f;
                l += mk * d * d;
                dy[j] = mk * d * inv_n;
            }
            sloss[t] = l;
            for (int i = 0; i < XB_MLP_NH; i++) {
                float g = 0.0f;
                for (int j = 0; j < XB_MLP_NO; j++) g += W3[j * XB_MLP_NH + i] * dy[j];
                sd2[t][i] = h2[i] > 0.0f ? g : 0.0f;
            }
            for (int i = 0; i < XB_MLP_NH; i++) {
                float g = 0.0f;
                for (int o = 0; o < XB_MLP_NH; o++) g += W2[o * XB_MLP_NH + i] * sd2[t][o];
                sd1[t][i] = h1[i] > 0.0f ? g : 0.0f;
            }
            for (int i = 0; i < XB_MLP_NI; i++) sx[t][i] = x[i];
            for (int i = 0; i < XB_MLP_NH; i++) {
                sh1[t][i] = h1[i];
                sh2[t][i] = h2[i];
            }
            for (int j = 0; j < XB_MLP_NO; j++) sdy[t][j] = dy[j];
        }
        __syncthreads();
        if (threadIdx.x == 0)
            for (int k = 0; k < XB_MLP_CHUNK; k++) loss_acc += sloss[k];
        for (int k = 0; k < (int)(sizeof(acc) / sizeof(acc[0])); k++) {
            const int e = threadIdx.x + k * XB_MLP_BLOCK;
            if (e >= XB_MLP_NPARAM) break;
            float g = 0.0f;
            if (e < XB_OFF_B1) {
                const int o = e / XB_MLP_NI, i = e % XB_MLP_NI;
                for (int s = 0; s < XB_MLP_CHUNK; s++) g += sd1[s][o] * sx[s][i];
            } else if (e < XB_OFF_W2) {
                const int o = e - XB_OFF_B1;
                for (int s = 0; s < XB_MLP_CHUNK; s++) g += sd1[s][o];
            } else if (e < XB_OFF_B2) {
                const int r = e - XB_OFF_W2, o = r / XB_MLP_NH, i = r % XB_MLP_NH;
                for (int s = 0; s < XB_MLP_CHUNK; s++) g += sd2[s][o] * sh1[s][i];
            } else if (e < XB_OFF_W3) {
                const int o = e - XB_OFF_B2;
                for (int s = 0; s < XB_MLP_CHUNK; s++) g += sd2[s][o];
            } else if (e < XB_OFF_B3) {
                const int r = e - XB_OFF_W3, j = r / XB_MLP_NH, i = r % XB_MLP_NH;
                for (int s = 0; s < XB_MLP_CHUNK; s++) g += sdy[s][j] * sh2[s][i];
            } else {
                const int j = e - XB_OFF_B3;
                for (int s = 0; s < XB_MLP_CHUNK; s++) g += sdy[s][j];
            }
            acc[k] += g;
        }
        __syncthreads();
    }
    for (int k = 0; k < (int)(sizeof(acc) / sizeof(acc[0])); k++) {
        const int e = threadIdx.x + k * XB_MLP_BLOCK;
        if (e < XB_MLP_NPARAM) part_grad[(int64_t)blockIdx.x * XB_MLP_NPARAM + e] = acc[k];
    }
    if (threadIdx.x == 0) part_loss[blockIdx.x] = loss_acc * 0.5f * inv_n;
}

// grad[e] = sum over blocks b (in order) of part_grad[b][e]; loss likewise
__global__ void k_mlp_reduce(const float *__restrict__ part_grad, const float *__restrict__ part_loss,
                             int32_t nblocks, float *__restrict__ grad, float *__restrict__ loss)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < XB_MLP_NPARAM) {
        float s = 0.0f;
        for (int b = 0; b < nblocks; b++) s += part_grad[(int64_t)b * XB_MLP_NPARAM + e];
        grad[e] = s;
    }
    if (e == 0) {
        float s = 0.0f;
        for (int b = 0; b < nblocks; b++) s += part_loss[b];
        *loss = s;
    }
}

// Adam step t (1-based), bias-corrected; state m, v in device memory.
__global__ void k_mlp_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                           float *__restrict__ v, int32_t t, float lr, float b1, float b2, float eps,
                           const float *__restrict__ loss, float *__restrict__ loss_curve)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < XB_MLP_NPARAM) {
        const float gi = g[e];
        const float mi = b1 * m[e] + (1.0f - b1) * gi;
        const float vi = b2 * v[e] + (1.0f - b2) * gi * gi;
        m[e] = mi;
        v[e] = vi;
        float c1 = 1.0f, c2 = 1.0f, pb1 = b1, pb2 = b2;
        for (int k = 0; k < t; k++) { c1 *= pb1; c2 *= pb2; }
        const float mh = mi / (1.0f - c1), vh = vi / (1.0f - c2);
        p[e] -= lr * mh / (sqrtf(vh) + eps);
    }
    if (e == 0 && loss_curve) loss_curve[t - 1] = *loss;
}
