// Learned voice-tail predictor for gfx950: a 12 -> 64 -> 64 -> 1 ReLU MLP in
// FP32 (5057 parameters), trained and evaluated by these kernels. No reference
// counterpart: it predicts the exact voice tails of the voice-silence policy
// (render.hip judges them), from features known when a job is planned.
//
// One 256-thread block takes a tile of 32 samples with the parameters and the
// activations in LDS; layers are small dense products computed as 8 outputs
// per thread. The backward pass writes one partial gradient per block, and a
// second kernel sums the partials in block order (deterministic) and applies
// Adam, so a training run is bit-reproducible.
#include <hip/hip_runtime.h>

#include "../../../include/clay_hip.h"

namespace {

constexpr int kIn = CLAY_MLP_INPUTS, kHid = CLAY_MLP_HIDDEN, kParams = CLAY_MLP_PARAMS;
constexpr int oW1 = 0, oB1 = oW1 + kIn * kHid, oW2 = oB1 + kHid, oB2 = oW2 + kHid * kHid, oW3 = oB2 + kHid,
              oB3 = oW3 + kHid;
static_assert(oB3 + 1 == kParams, "parameter layout");
constexpr int kTile = 32;     // samples per block (LDS: parameters + 3 activation tiles, < 64 KB)
constexpr int kThreads = 256;
constexpr int kPad = kHid + 1;  // LDS row pitch (bank spread)

} // namespace

// Forward (and with Backward, the gradient of sum_s scale * (yhat_s - y_s)^2)
// over samples [first + 32 * block, ...) of `index` (or the identity), n total.
// partial: (kParams + 1) floats per block, the last one the block's loss sum.
template<bool Backward>
__global__ void __launch_bounds__(kThreads) clayMlpPass(const float* __restrict__ params,
                                                        const float* __restrict__ x, const float* __restrict__ y,
                                                        const int* __restrict__ index, int n, int first, float scale,
                                                        float* __restrict__ partial, float* __restrict__ out)
{
    __shared__ float P[kParams];
    __shared__ float X[kTile][kIn + 1];
    __shared__ float H1[kTile][kPad], H2[kTile][kPad];  // H2's storage holds D1 in the backward pass
    __shared__ float D2[kTile][kPad];
    __shared__ float dY[kTile], loss[kTile];
    const int t = threadIdx.x;
    const int base = first + int(blockIdx.x) * kTile;
    for (int p = t; p < kParams; p += kThreads)
        P[p] = params[p];
    for (int e = t; e < kTile * kIn; e += kThreads) {
        const int s = e / kIn, i = e % kIn, k = base + s;
        float v = 0.f;
        if (k < n)
            v = x[size_t(index ? index[k] : k) * kIn + size_t(i)];
        X[s][i] = v;
    }
    __syncthreads();
    const int j = t & (kHid - 1), s0 = t >> 6;  // hidden unit; samples s0, s0 + 4, ...
    for (int s = s0; s < kTile; s += 4) {
        float h = P[oB1 + j];
        for (int i = 0; i < kIn; ++i)
            h = fmaf(X[s][i], P[oW1 + i * kHid + j], h);
        H1[s][j] = fmaxf(h, 0.f);
    }
    __syncthreads();
    for (int s = s0; s < kTile; s += 4) {
        float h = P[oB2 + j];
        for (int i = 0; i < kHid; ++i)
            h = fmaf(H1[s][i], P[oW2 + i * kHid + j], h);
        H2[s][j] = fmaxf(h, 0.f);
    }
    __syncthreads();
    if (t < kTile) {
        float yhat = P[oB3];
        for (int i = 0; i < kHid; ++i)
            yhat = fmaf(H2[t][i], P[oW3 + i], yhat);
        const int k = base + t;
        const bool valid = k < n;
        if (!Backward) {
            if (valid)
                out[index ? index[k] : k] = yhat;
        } else {
            const float e = valid ? yhat - y[index ? index[k] : k] : 0.f;
            dY[t] = 2.f * scale * e;
            loss[t] = scale * e * e;
        }
    }
    if (!Backward)
        return;
    __syncthreads();
    float* g = partial + size_t(blockIdx.x) * (kParams + 1);
    // Output layer and D2 = dY w3 (H2 > 0).
    if (t < kHid) {
        float w = 0.f;
        for (int s = 0; s < kTile; ++s)
            w = fmaf(H2[s][t], dY[s], w);
        g[oW3 + t] = w;
    }
    if (t == kHid) {
        float b = 0.f, l = 0.f;
        for (int s = 0; s < kTile; ++s) {
            b += dY[s];
            l += loss[s];
        }
        g[oB3] = b;
        g[kParams] = l;
    }
    for (int s = s0; s < kTile; s += 4)
        D2[s][j] = H2[s][j] > 0.f ? dY[s] * P[oW3 + j] : 0.f;
    __syncthreads();
    float (*D1)[kPad] = H2;  // H2 is dead from here on
    // dW2[i][j] = sum_s H1[s][i] D2[s][j], 16 entries per thread.
    for (int e = t; e < kHid * kHid; e += kThreads) {
        const int i = e / kHid, jj = e % kHid;
        float w = 0.f;
        for (int s = 0; s < kTile; ++s)
            w = fmaf(H1[s][i], D2[s][jj], w);
        g[oW2 + e] = w;
    }
    if (t < kHid) {
        float b = 0.f;
        for (int s = 0; s < kTile; ++s)
            b += D2[s][t];
        g[oB2 + t] = b;
    }
    // D1[s][i] = (sum_j D2[s][j] W2[i][j]) (H1 > 0).
    for (int s = s0; s < kTile; s += 4) {
        float d = 0.f;
        for (int jj = 0; jj < kHid; ++jj)
            d = fmaf(D2[s][jj], P[oW2 + j * kHid + jj], d);
        D1[s][j] = H1[s][j] > 0.f ? d : 0.f;
    }
    __syncthreads();
    for (int e = t; e < kIn * kHid; e += kThreads) {
        const int i = e / kHid, jj = e % kHid;
        float w = 0.f;
        for (int s = 0; s < kTile; ++s)
            w = fmaf(X[s][i], D1[s][jj], w);
        g[oW1 + e] = w;
    }
    if (t < kHid) {
        float b = 0.f;
        for (int s = 0; s < kTile; ++s)
            b += D1[s][t];
        g[oB1 + t] = b;
    }
}

// Sums the block partials in block order. With adam != 0 applies one Adam step
// (bias-corrected, step >= 1) to params; otherwise writes the summed gradient
// to grad. The loss sum goes to *loss.
__global__ void __launch_bounds__(kThreads) clayMlpReduce(const float* __restrict__ partial, int blocks,
                                                         float* __restrict__ params, float* __restrict__ m,
                                                         float* __restrict__ v, float* __restrict__ grad,
                                                         float* __restrict__ loss, int adam, float lr, float beta1,
                                                         float beta2, float eps, int step)
{
    const int p = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (p > kParams)
        return;
    float sum = 0.f;
    for (int b = 0; b < blocks; ++b)
        sum += partial[size_t(b) * (kParams + 1) + size_t(p)];
    if (p == kParams) {
        *loss = sum;
        return;
    }
    if (!adam) {
        grad[p] = sum;
        return;
    }
    const float mp = fmaf(beta1, m[p], (1.f - beta1) * sum);
    const float vp = fmaf(beta2, v[p], (1.f - beta2) * sum * sum);
    m[p] = mp;
    v[p] = vp;
    const float mh = mp / (1.f - powf(beta1, float(step)));
    const float vh = vp / (1.f - powf(beta2, float(step)));
    params[p] -= lr * mh / (sqrtf(vh) + eps);
}

template __global__ void clayMlpPass<true>(const float*, const float*, const float*, const int*, int, int, float,
                                           float*, float*);
template __global__ void clayMlpPass<false>(const float*, const float*, const float*, const int*, int, int, float,
                                            float*, float*);
