// Krylov-step controller network: 8 -> 64 -> 64 -> 2, tanh hidden layers,
// sigmoid outputs.  Weights, inputs and activations are bfloat16; every
// sum is FP32.  The parameter vector (FP32 master copy and its bfloat16
// image share the layout) is
//   W1[64][8] b1[64] W2[64][64] b2[64] W3[2][64] b3[2].
#pragma once
#include <cstdint>

#include <hip/hip_runtime.h>

namespace espec {

constexpr int kMlpIn = 8, kMlpHid = 64, kMlpOut = 2;
constexpr int kMlpW1 = 0;
constexpr int kMlpB1 = kMlpW1 + kMlpHid * kMlpIn;
constexpr int kMlpW2 = kMlpB1 + kMlpHid;
constexpr int kMlpB2 = kMlpW2 + kMlpHid * kMlpHid;
constexpr int kMlpW3 = kMlpB2 + kMlpHid;
constexpr int kMlpB3 = kMlpW3 + kMlpOut * kMlpHid;
constexpr int kMlpNPar = kMlpB3 + kMlpOut;  // 4866
constexpr int kMlpBatch = 64;               // samples per training minibatch

// Round to the nearest bfloat16, ties to even (NaN stays NaN).
__host__ __device__ inline uint16_t f2bf(float x)
{
    uint32_t u;
    __builtin_memcpy(&u, &x, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t h)
{
    const uint32_t u = (uint32_t)h << 16;
    float x;
    __builtin_memcpy(&x, &u, 4);
    return x;
}
__host__ __device__ inline float bfr(float x) { return bf2f(f2bf(x)); }

// Inference for n feature vectors (one wavefront each) with bfloat16
// parameters wb; out[s][0..1] are the sigmoid outputs.
hipError_t launch_mlp_infer(const uint16_t *wb, const float *feat, float *out, int n, hipStream_t s);

}  // namespace espec
