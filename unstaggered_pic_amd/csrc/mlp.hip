// mlp.hip -- training and standalone inference of the learned Picard guess
// (network layout in mlp.hpp).  Training is FP32 on the vector ALUs: one
// thread per sample computes the forward and backward pass, weight gradients
// accumulate in LDS and then in a global FP32 buffer; Adam updates the
// parameters; a moments kernel sets the feature normalisation.  None of this
// has a reference counterpart: the reference's inner solve starts from v^n
// (P3:5318).
#include "kernels.hpp"
#include "mlp.hpp"

namespace upic {

namespace {

inline unsigned nb(int64_t n, int t = 256) { return (unsigned)((n + t - 1) / t); }

// Feature normalisation and target scale from the training set:
// in_mean, in_scale = 1/std (1 for a constant feature), out_scale = rms.
__global__ void k_mlp_moments(const float* __restrict__ X, int ld, const float* __restrict__ T, int64_t n,
                              double* __restrict__ acc) {
    double s[12] = {}, s2[12] = {}, t2[3] = {};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
#pragma unroll
        for (int j = 0; j < 12; ++j) {
            const double v = X[i * ld + j];
            s[j] += v;
            s2[j] += v * v;
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double v = T[i * ld + c];
            t2[c] += v * v;
        }
    }
#pragma unroll
    for (int j = 0; j < 12; ++j) {
        atomicAdd(&acc[j], s[j]);
        atomicAdd(&acc[12 + j], s2[j]);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) atomicAdd(&acc[24 + c], t2[c]);
}

__global__ void k_mlp_set_norm(const double* __restrict__ acc, int64_t n, float* __restrict__ P) {
    const int j = threadIdx.x;
    const double inv = 1.0 / (double)(n > 0 ? n : 1);
    if (j < 12) {
        const double mu = acc[j] * inv;
        const double var = fmax(acc[12 + j] * inv - mu * mu, 0.0);
        P[kMlpInMean + j] = (float)mu;
        P[kMlpInScale + j] = var > 1e-30 ? (float)(1.0 / sqrt(var)) : 1.0f;
    } else if (j < 15) {
        const double ms = acc[24 + (j - 12)] * inv;
        P[kMlpOutScale + (j - 12)] = ms > 1e-60 ? (float)sqrt(ms) : 1.0f;
    }
}

__device__ __forceinline__ float sigm(float z) { return 1.0f / (1.0f + __expf(-z)); }

// Loss and gradient of one minibatch, in two passes.  Pass 1, one thread per
// sample: forward and backward through the network, writing the sample's
// activations and back-propagated errors to a scratch row
//   [x 12 | h1 64 | dz1 64 | h2 64 | dz2 64 | dy 3]  (kActRow floats);
// loss = sum_b sum_c (y_bc - t_bc)^2 / (3 B) with t = T / out_scale.  Pass 2,
// one thread per trainable entry: the weight gradient is the sum over the
// batch of an outer-product entry (dz . x, dz . h, dy . h) and a bias
// gradient the sum of an error.  Samples are rows idx[b] (idx NULL: b) of X
// and T (row stride ld).
constexpr int kActX = 0, kActH1 = 12, kActD1 = 76, kActH2 = 140, kActD2 = 204, kActDy = 268, kActRow = 272;

__global__ __launch_bounds__(256) void k_mlp_fwdbwd(const float* __restrict__ P, const float* __restrict__ X,
                                                    const float* __restrict__ T, int ld, const int* __restrict__ idx,
                                                    int B, float* __restrict__ act, double* __restrict__ loss) {
    __shared__ float w[kMlpNParam];
    for (int i = threadIdx.x; i < kMlpNParam; i += blockDim.x) w[i] = P[i];
    __syncthreads();
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    double lsum = 0.0;
    if (b < B) {
        const int64_t s = idx ? idx[b] : b;
        float* row = act + (int64_t)b * kActRow;
        float x[12];
#pragma unroll
        for (int k = 0; k < 12; ++k) {
            x[k] = (X[s * ld + k] - w[kMlpInMean + k]) * w[kMlpInScale + k];
            row[kActX + k] = x[k];
        }
        // x, z1, h1 live in registers (fully unrolled); the rest of the
        // sample's vectors go through its scratch row (rolled loops keep the
        // register footprint small)
        float z1[64], h1[64];
#pragma unroll
        for (int i = 0; i < 64; ++i) {
            float a = w[kMlpB1 + i];
#pragma unroll
            for (int k = 0; k < 12; ++k) a += w[kMlpW1 + i * 12 + k] * x[k];
            z1[i] = a;
            h1[i] = silu_f(a);
            row[kActH1 + i] = h1[i];
        }
#pragma unroll 1
        for (int i = 0; i < 64; ++i) {
            float a = w[kMlpB2 + i];
#pragma unroll
            for (int k = 0; k < 64; ++k) a += w[kMlpW2 + i * 64 + k] * h1[k];
            row[kActD2 + i] = a;  // z2 for now
            row[kActH2 + i] = silu_f(a);
        }
        float dy[3];
        const float sc = 2.0f / (3.0f * (float)B);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            float a = w[kMlpB3 + c];
#pragma unroll 8
            for (int k = 0; k < 64; ++k) a += w[kMlpW3 + c * 64 + k] * row[kActH2 + k];
            const float e = a - T[s * ld + c] / w[kMlpOutScale + c];
            lsum += (double)e * (double)e;
            dy[c] = sc * e;
            row[kActDy + c] = dy[c];
        }
#pragma unroll 1
        for (int k = 0; k < 64; ++k) {
            const float z = row[kActD2 + k];
            const float sg = sigm(z);
            float dh = 0.0f;
#pragma unroll
            for (int c = 0; c < 3; ++c) dh += w[kMlpW3 + c * 64 + k] * dy[c];
            row[kActD2 + k] = dh * sg * (1.0f + z * (1.0f - sg));
        }
        float dh1[64];
#pragma unroll
        for (int k = 0; k < 64; ++k) dh1[k] = 0.0f;
#pragma unroll 1
        for (int i = 0; i < 64; ++i) {
            const float d = row[kActD2 + i];
#pragma unroll
            for (int k = 0; k < 64; ++k) dh1[k] += w[kMlpW2 + i * 64 + k] * d;
        }
#pragma unroll
        for (int k = 0; k < 64; ++k) {
            const float sg = sigm(z1[k]);
            row[kActD1 + k] = dh1[k] * sg * (1.0f + z1[k] * (1.0f - sg));
        }
    }
    for (int o = 32; o > 0; o >>= 1) lsum += __shfl_down(lsum, o, 64);
    if ((threadIdx.x & 63) == 0 && loss) atomicAdd(loss, lsum / (3.0 * (double)B));
}

__global__ __launch_bounds__(256) void k_mlp_wgrad(const float* __restrict__ act, int B, float* __restrict__ G) {
    const int e = kMlpTrain0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= kMlpNParam) return;
    int ia, ib;  // error column, input column (-1: bias)
    if (e < kMlpB1) { const int r = e - kMlpW1; ia = kActD1 + r / 12; ib = kActX + r % 12; }
    else if (e < kMlpW2) { ia = kActD1 + (e - kMlpB1); ib = -1; }
    else if (e < kMlpB2) { const int r = e - kMlpW2; ia = kActD2 + r / 64; ib = kActH1 + r % 64; }
    else if (e < kMlpW3) { ia = kActD2 + (e - kMlpB2); ib = -1; }
    else if (e < kMlpB3) { const int r = e - kMlpW3; ia = kActDy + r / 64; ib = kActH2 + r % 64; }
    else { ia = kActDy + (e - kMlpB3); ib = -1; }
    float acc = 0.0f;
    for (int b = 0; b < B; ++b) {
        const float* row = act + (int64_t)b * kActRow;
        acc += row[ia] * (ib >= 0 ? row[ib] : 1.0f);
    }
    G[e] += acc;
}

// Adam (bias-corrected) on the trainable entries; step t >= 1.
__global__ void k_mlp_adam(float* __restrict__ P, const float* __restrict__ G, float* __restrict__ M,
                           float* __restrict__ V, float lr, float b1c, float b2c) {
    const int i = kMlpTrain0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kMlpNParam) return;
    const float g = G[i];
    const float m = 0.9f * M[i] + 0.1f * g;
    const float v = 0.999f * V[i] + 0.001f * g * g;
    M[i] = m;
    V[i] = v;
    P[i] -= lr * (m / b1c) / (sqrtf(v / b2c) + 1.0e-8f);
}

// Inference blob from the parameters (biases, normalisation, BF16 fragments).
__global__ void k_mlp_pack(const float* __restrict__ P, float* __restrict__ blob) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < kBlobF1) {
        float v = 0.0f;
        if (e < 12) v = P[kMlpInMean + e];
        else if (e < 24) v = P[kMlpInScale + e - 12];
        else if (e < 27) v = P[kMlpOutScale + e - 24];
        else if (e >= kBlobB1 && e < kBlobB1 + 64) v = P[kMlpB1 + e - kBlobB1];
        else if (e >= kBlobB2 && e < kBlobB2 + 64) v = P[kMlpB2 + e - kBlobB2];
        else if (e >= kBlobB3 && e < kBlobB3 + 3) v = P[kMlpB3 + e - kBlobB3];
        blob[e] = v;
    }
    // fragments: element j of lane l of fragment (ks, cb) holds W[n][k] with
    // k = 32 ks + 8 (l / 16) + j, n = 16 cb + l % 16 (zero outside the layer)
    __bf16* fb = reinterpret_cast<__bf16*>(blob + kBlobF1);
    const int nfrag_el = 14 * 512;
    if (e < nfrag_el) {
        const int frag = e / 512, lane = (e / 8) % 64, j = e % 8;
        int layer, ks, cb;
        if (frag < 4) { layer = 0; ks = 0; cb = frag; }
        else if (frag < 12) { layer = 1; ks = (frag - 4) / 4; cb = (frag - 4) % 4; }
        else { layer = 2; ks = frag - 12; cb = 0; }
        const int k = 32 * ks + 8 * (lane / 16) + j, n = 16 * cb + lane % 16;
        float v = 0.0f;
        if (layer == 0 && k < 12 && n < 64) v = P[kMlpW1 + n * 12 + k];
        if (layer == 1 && k < 64 && n < 64) v = P[kMlpW2 + n * 64 + k];
        if (layer == 2 && k < 64 && n < 3) v = P[kMlpW3 + n * 64 + k];
        fb[e] = (__bf16)v;
    }
}

// Standalone inference (the same device function the P3 push calls).
__global__ __launch_bounds__(256) void k_mlp_forward(const float* __restrict__ blob, const float* __restrict__ X,
                                                     int64_t n, float* __restrict__ Y) {
    extern __shared__ unsigned char lds[];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    float f[12], y[3];
#pragma unroll
    for (int k = 0; k < 12; ++k) f[k] = i < n ? X[i * 12 + k] : 0.0f;
    mlp_wave_infer(blob, f, y, lds + (threadIdx.x / 64) * kMlpLdsWave);
    if (i < n) {
        Y[i * 3 + 0] = y[0];
        Y[i * 3 + 1] = y[1];
        Y[i * 3 + 2] = y[2];
    }
}

// Minibatch indices: a splitmix64 hash of (seed, step, slot) modulo n.
__global__ void k_mlp_batch_idx(uint64_t seed, uint64_t step, int B, int64_t n, int* __restrict__ idx) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    uint64_t z = seed + 0x9E3779B97F4A7C15ull * (step * 1000003ull + (uint64_t)b + 1ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    idx[b] = (int)(z % (uint64_t)n);
}

}  // namespace

void launch_mlp_moments(hipStream_t s, const float* rows, int64_t n, double* acc, float* P) {
    hipLaunchKernelGGL(k_mlp_moments, dim3(64), dim3(256), 0, s, rows, 15, rows + 12, n, acc);
    hipLaunchKernelGGL(k_mlp_set_norm, dim3(1), dim3(64), 0, s, acc, n, P);
}
void launch_mlp_grad(hipStream_t s, const float* P, const float* rows, const int* idx, int B, float* G, double* loss,
                     float* act) {
    if (B <= 0) return;
    hipLaunchKernelGGL(k_mlp_fwdbwd, dim3(nb(B)), dim3(256), 0, s, P, rows, rows + 12, 15, idx, B, act, loss);
    hipLaunchKernelGGL(k_mlp_wgrad, dim3(nb(kMlpNParam - kMlpTrain0)), dim3(256), 0, s, act, B, G);
}
size_t mlp_grad_scratch_floats(int B) { return (size_t)B * kActRow; }
void launch_mlp_adam(hipStream_t s, float* P, const float* G, float* M, float* V, float lr, int t) {
    const float b1c = 1.0f - powf(0.9f, (float)t), b2c = 1.0f - powf(0.999f, (float)t);
    hipLaunchKernelGGL(k_mlp_adam, dim3(nb(kMlpNParam - kMlpTrain0)), dim3(256), 0, s, P, G, M, V, lr, b1c, b2c);
}
void launch_mlp_pack(hipStream_t s, const float* P, float* blob) {
    hipLaunchKernelGGL(k_mlp_pack, dim3(nb(14 * 512)), dim3(256), 0, s, P, blob);
}
void launch_mlp_forward(hipStream_t s, const float* blob, const float* X, int64_t n, float* Y) {
    if (n > 0) hipLaunchKernelGGL(k_mlp_forward, dim3(nb(n)), dim3(256), 4 * kMlpLdsWave, s, blob, X, n, Y);
}
void launch_mlp_batch_idx(hipStream_t s, uint64_t seed, uint64_t step, int B, int64_t n, int* idx) {
    if (B > 0) hipLaunchKernelGGL(k_mlp_batch_idx, dim3(nb(B)), dim3(256), 0, s, seed, step, B, n, idx);
}

}  // namespace upic
