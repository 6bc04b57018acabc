// mlp.hpp -- the learned initial guess of the per-particle Picard solve:
// a 12 -> 64 -> 64 -> 3 MLP with SiLU activations.
//
// The network corrects a physics predictor of the orbit-averaged velocity
// (see mixed.hip): guess = v^n + delta_phys + out_scale * MLP(normalised
// features).  It only supplies the starting iterate of the damped Picard
// iteration of solve_local_particle_substep (P3:4930-4983, whose plain start
// is v^n, P3:5318); the iteration to the unchanged tolerance still decides
// convergence, so the converged state does not depend on the network.
//
// Parameter vector (FP32, kMlpNParam floats):
//   [in_mean 12 | in_scale 12 | out_scale 3 | W1 64x12 | b1 64 | W2 64x64 |
//    b2 64 | W3 3x64 | b3 3]      (W row-major: out x in)
// Inference blob (device, built from the parameters by launch_mlp_pack):
//   floats [0, 176): in_mean, in_scale, out_scale, pad, b1, b2, b3 (16)
//   then BF16 MFMA B-operand fragments of the three layers.
//
// Inference runs per wavefront on MFMA (v_mfma_f32_16x16x32_bf16): the 64
// particles of a wave are the M dimension (four 16-row blocks), BF16
// operands, FP32 accumulation; SiLU and biases in FP32.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace upic {

constexpr int kMlpIn = 12, kMlpHid = 64, kMlpOut = 3;
constexpr int kMlpInMean = 0, kMlpInScale = 12, kMlpOutScale = 24;
constexpr int kMlpW1 = 27, kMlpB1 = kMlpW1 + kMlpHid * kMlpIn;       // 795
constexpr int kMlpW2 = kMlpB1 + kMlpHid, kMlpB2 = kMlpW2 + kMlpHid * kMlpHid;  // 859, 4955
constexpr int kMlpW3 = kMlpB2 + kMlpHid, kMlpB3 = kMlpW3 + kMlpOut * kMlpHid;  // 5019, 5211
constexpr int kMlpNParam = kMlpB3 + kMlpOut;                          // 5214
constexpr int kMlpTrain0 = kMlpW1;                                    // first trainable entry

// blob offsets (floats)
constexpr int kBlobMean = 0, kBlobScale = 12, kBlobOut = 24, kBlobB1 = 32, kBlobB2 = 96, kBlobB3 = 160;
constexpr int kBlobF1 = 176;                 // 4 fragments x 64 lanes x 8 bf16 = 1024 floats
constexpr int kBlobF2 = kBlobF1 + 1024;      // 8 fragments
constexpr int kBlobF3 = kBlobF2 + 2048;      // 2 fragments
constexpr int kBlobN = kBlobF3 + 512;        // 3760 floats

// per-wave LDS scratch of the inference (bytes): X 64x40 bf16, H1/H2 16x72
// bf16, outputs 64x4 float
constexpr int kMlpXS = 40, kMlpHS = 72;
constexpr int kMlpLdsWave = 64 * kMlpXS * 2 + 2 * 16 * kMlpHS * 2 + 64 * 4 * 4;  // 13,824

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

// SiLU with the hardware exponential and reciprocal (v_exp_f32, v_rcp_f32):
// about 1 ulp, far below the BF16 resolution of the network, and a few
// instructions instead of a correctly rounded division per activation.
__device__ __forceinline__ float silu_f(float z) { return z * __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }

__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Every lane of the wavefront must call this (no divergence): lane l brings
// the 12 raw features of its particle and receives the 3 scaled outputs.
// `lds` is this wave's kMlpLdsWave bytes of LDS.
__device__ __forceinline__ void mlp_wave_infer(const float* __restrict__ blob, const float f[12], float y[3],
                                               unsigned char* lds) {
    const int lane = threadIdx.x & 63;
    const int r16 = lane & 15, g4 = lane >> 4;
    __bf16* X = reinterpret_cast<__bf16*>(lds);
    __bf16* H1 = X + 64 * kMlpXS;
    __bf16* H2 = H1 + 16 * kMlpHS;
    float* O = reinterpret_cast<float*>(H2 + 16 * kMlpHS);
#pragma unroll
    for (int j = 0; j < 32; ++j) {
        const float v = j < kMlpIn ? (f[j] - blob[kBlobMean + j]) * blob[kBlobScale + j] : 0.0f;
        X[lane * kMlpXS + j] = (__bf16)v;
    }
    wave_sync_lds();
    const bf16x8* F1 = reinterpret_cast<const bf16x8*>(blob + kBlobF1);
    const bf16x8* F2 = reinterpret_cast<const bf16x8*>(blob + kBlobF2);
    const bf16x8* F3 = reinterpret_cast<const bf16x8*>(blob + kBlobF3);
    for (int rb = 0; rb < 4; ++rb) {
        // layer 1: [16 x 32] x [32 x 64]
        const bf16x8 a = *reinterpret_cast<const bf16x8*>(X + (16 * rb + r16) * kMlpXS + 8 * g4);
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) {
            floatx4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, F1[cb * 64 + lane], acc, 0, 0, 0);
            const float b = blob[kBlobB1 + 16 * cb + r16];
#pragma unroll
            for (int r = 0; r < 4; ++r) H1[(4 * g4 + r) * kMlpHS + 16 * cb + r16] = (__bf16)silu_f(acc[r] + b);
        }
        wave_sync_lds();
        // layer 2: [16 x 64] x [64 x 64]
        const bf16x8 a0 = *reinterpret_cast<const bf16x8*>(H1 + r16 * kMlpHS + 8 * g4);
        const bf16x8 a1 = *reinterpret_cast<const bf16x8*>(H1 + r16 * kMlpHS + 32 + 8 * g4);
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) {
            floatx4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, F2[cb * 64 + lane], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, F2[(4 + cb) * 64 + lane], acc, 0, 0, 0);
            const float b = blob[kBlobB2 + 16 * cb + r16];
#pragma unroll
            for (int r = 0; r < 4; ++r) H2[(4 * g4 + r) * kMlpHS + 16 * cb + r16] = (__bf16)silu_f(acc[r] + b);
        }
        wave_sync_lds();
        // layer 3: [16 x 64] x [64 x 16] (3 live columns)
        const bf16x8 c0 = *reinterpret_cast<const bf16x8*>(H2 + r16 * kMlpHS + 8 * g4);
        const bf16x8 c1 = *reinterpret_cast<const bf16x8*>(H2 + r16 * kMlpHS + 32 + 8 * g4);
        floatx4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(c0, F3[lane], acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(c1, F3[64 + lane], acc, 0, 0, 0);
        if (r16 < kMlpOut) {
            const float b = blob[kBlobB3 + r16], s = blob[kBlobOut + r16];
#pragma unroll
            for (int r = 0; r < 4; ++r) O[(16 * rb + 4 * g4 + r) * 4 + r16] = (acc[r] + b) * s;
        }
        wave_sync_lds();
    }
    y[0] = O[lane * 4 + 0];
    y[1] = O[lane * 4 + 1];
    y[2] = O[lane * 4 + 2];
    wave_sync_lds();
}

}  // namespace upic
