// Kernels of the Krylov-step controller network (layout in mlp.hpp):
// inference, the gradient of one minibatch, and device-resident training
// with Adam.  bfloat16 operands, FP32 sums; the FP32 master parameters are
// updated and their bfloat16 image refreshed after every minibatch.
//
// No reference counterpart: the controller picks the Krylov order and the
// time-step scale of the short-iterative Lanczos propagator (sil.hip),
// which falls back to its fixed-order step when the Krylov residual
// estimate of a chosen step exceeds the tolerance.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

#include "api_common.hpp"
#include "mlp.hpp"

namespace espec {

namespace {

__device__ __forceinline__ float wave_sum_f(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ float sigm(float y) { return 1.0f / (1.0f + expf(-y)); }

__device__ __forceinline__ float w(const uint16_t *wb, int i) { return bf2f(wb[i]); }

}  // namespace

// One wavefront per sample: lane j is hidden unit j of both layers.
__global__ __launch_bounds__(64) void mlp_infer(const uint16_t *wb, const float *feat, float *out,
                                                int n)
{
    __shared__ float h1s[kMlpHid];
    const int s = blockIdx.x;
    if (s >= n) return;
    const int j = threadIdx.x;
    float z = 0;
    for (int k = 0; k < kMlpIn; ++k) z += w(wb, kMlpW1 + j * kMlpIn + k) * bfr(feat[s * kMlpIn + k]);
    h1s[j] = bfr(tanhf(z + w(wb, kMlpB1 + j)));
    __syncthreads();
    float z2 = 0;
    for (int k = 0; k < kMlpHid; ++k) z2 += w(wb, kMlpW2 + j * kMlpHid + k) * h1s[k];
    const float h2 = bfr(tanhf(z2 + w(wb, kMlpB2 + j)));
    for (int o = 0; o < kMlpOut; ++o) {
        const float y = wave_sum_f(w(wb, kMlpW3 + o * kMlpHid + j) * h2);
        if (j == 0) out[s * kMlpOut + o] = sigm(y + w(wb, kMlpB3 + o));
    }
}

namespace {

constexpr int kTrainNT = 256;
constexpr int B = kMlpBatch;

struct MlpLds {
    float x[B][kMlpIn];
    float h1[B][kMlpHid];
    float h2[B][kMlpHid];   // after the output layer's backward: dz1
    float d2[B][kMlpHid];   // dz2
    float dy[B][kMlpOut];
    float red[kTrainNT / 64];
};

// Forward and backward of one minibatch (rows idx[0..B) of x, t) for the
// loss mean((sigmoid(y) - t)^2) over the B x 2 outputs.  Every operand of
// a product is a bfloat16 value (weights, inputs, activations and the
// back-propagated dz), every sum FP32.  grad (global) receives dL/dparam;
// returns the loss in thread 0.
__device__ float minibatch(const uint16_t *wb, const float *X, const float *T, const int *idx,
                           float *grad, MlpLds &L)
{
    const int t = threadIdx.x;
    for (int e = t; e < B * kMlpIn; e += kTrainNT) {
        const int b = e / kMlpIn, k = e % kMlpIn;
        L.x[b][k] = bfr(X[(long long)idx[b] * kMlpIn + k]);
    }
    __syncthreads();
    for (int e = t; e < B * kMlpHid; e += kTrainNT) {
        const int b = e / kMlpHid, j = e % kMlpHid;
        float z = 0;
        for (int k = 0; k < kMlpIn; ++k) z += w(wb, kMlpW1 + j * kMlpIn + k) * L.x[b][k];
        L.h1[b][j] = bfr(tanhf(z + w(wb, kMlpB1 + j)));
    }
    __syncthreads();
    for (int e = t; e < B * kMlpHid; e += kTrainNT) {
        const int b = e / kMlpHid, j = e % kMlpHid;
        float z = 0;
        for (int k = 0; k < kMlpHid; ++k) z += w(wb, kMlpW2 + j * kMlpHid + k) * L.h1[b][k];
        L.h2[b][j] = bfr(tanhf(z + w(wb, kMlpB2 + j)));
    }
    __syncthreads();
    float lacc = 0;
    for (int e = t; e < B * kMlpOut; e += kTrainNT) {
        const int b = e / kMlpOut, o = e % kMlpOut;
        float y = 0;
        for (int j = 0; j < kMlpHid; ++j) y += w(wb, kMlpW3 + o * kMlpHid + j) * L.h2[b][j];
        const float p = sigm(y + w(wb, kMlpB3 + o));
        const float r = p - T[(long long)idx[b] * kMlpOut + o];
        lacc += r * r;
        L.dy[b][o] = bfr(2.0f * r / (float)(B * kMlpOut) * p * (1.0f - p));
    }
    lacc = wave_sum_f(lacc);
    if ((t & 63) == 0) L.red[t >> 6] = lacc;
    __syncthreads();
    // output layer
    for (int e = t; e < kMlpOut * kMlpHid; e += kTrainNT) {
        const int o = e / kMlpHid, j = e % kMlpHid;
        float g = 0;
        for (int b = 0; b < B; ++b) g += L.dy[b][o] * L.h2[b][j];
        grad[kMlpW3 + e] = g;
    }
    if (t < kMlpOut) {
        float g = 0;
        for (int b = 0; b < B; ++b) g += L.dy[b][t];
        grad[kMlpB3 + t] = g;
    }
    for (int e = t; e < B * kMlpHid; e += kTrainNT) {
        const int b = e / kMlpHid, j = e % kMlpHid;
        float d = 0;
        for (int o = 0; o < kMlpOut; ++o) d += w(wb, kMlpW3 + o * kMlpHid + j) * L.dy[b][o];
        L.d2[b][j] = bfr(d * (1.0f - L.h2[b][j] * L.h2[b][j]));
    }
    __syncthreads();
    // second hidden layer
    for (int e = t; e < kMlpHid * kMlpHid; e += kTrainNT) {
        const int j = e / kMlpHid, k = e % kMlpHid;
        float g = 0;
        for (int b = 0; b < B; ++b) g += L.d2[b][j] * L.h1[b][k];
        grad[kMlpW2 + e] = g;
    }
    for (int j = t; j < kMlpHid; j += kTrainNT) {
        float g = 0;
        for (int b = 0; b < B; ++b) g += L.d2[b][j];
        grad[kMlpB2 + j] = g;
    }
    for (int e = t; e < B * kMlpHid; e += kTrainNT) {  // dz1 overwrites h2
        const int b = e / kMlpHid, k = e % kMlpHid;
        float d = 0;
        for (int j = 0; j < kMlpHid; ++j) d += w(wb, kMlpW2 + j * kMlpHid + k) * L.d2[b][j];
        L.h2[b][k] = bfr(d * (1.0f - L.h1[b][k] * L.h1[b][k]));
    }
    __syncthreads();
    // first hidden layer
    for (int e = t; e < kMlpHid * kMlpIn; e += kTrainNT) {
        const int k = e / kMlpIn, i = e % kMlpIn;
        float g = 0;
        for (int b = 0; b < B; ++b) g += L.h2[b][k] * L.x[b][i];
        grad[kMlpW1 + e] = g;
    }
    for (int k = t; k < kMlpHid; k += kTrainNT) {
        float g = 0;
        for (int b = 0; b < B; ++b) g += L.h2[b][k];
        grad[kMlpB1 + k] = g;
    }
    __syncthreads();  // grad complete and visible to the block
    float loss = 0;
    if (t == 0)
        for (int q = 0; q < kTrainNT / 64; ++q) loss += L.red[q];
    return loss / (float)(B * kMlpOut);
}

}  // namespace

__global__ __launch_bounds__(kTrainNT) void mlp_grad_kernel(const uint16_t *wb, const float *X,
                                                            const float *T, const int *idx,
                                                            float *grad, float *loss)
{
    __shared__ MlpLds L;
    const float l = minibatch(wb, X, T, idx, grad, L);
    if (threadIdx.x == 0) *loss = l;
}

struct AdamArgs {
    float *master, *m, *v, *grad;
    uint16_t *wb;
    float lr, b1, b2, eps;
    int step0;  // Adam steps taken before this launch
};

// nb minibatches in sequence, each followed by an Adam update of all
// parameters; loss[q] is minibatch q's loss.  One workgroup.
__global__ __launch_bounds__(kTrainNT) void mlp_train_kernel(AdamArgs a, const float *X,
                                                             const float *T, const int *idx,
                                                             int nb, float *loss)
{
    __shared__ MlpLds L;
    for (int q = 0; q < nb; ++q) {
        const float l = minibatch(a.wb, X, T, idx + (long long)q * B, a.grad, L);
        if (threadIdx.x == 0) loss[q] = l;
        const int step = a.step0 + q + 1;
        const float c1 = 1.0f - powf(a.b1, (float)step), c2 = 1.0f - powf(a.b2, (float)step);
        for (int p = threadIdx.x; p < kMlpNPar; p += kTrainNT) {
            const float g = a.grad[p];
            const float m = a.b1 * a.m[p] + (1.0f - a.b1) * g;
            const float v = a.b2 * a.v[p] + (1.0f - a.b2) * g * g;
            a.m[p] = m;
            a.v[p] = v;
            const float x = a.master[p] - a.lr * (m / c1) / (sqrtf(v / c2) + a.eps);
            a.master[p] = x;
            a.wb[p] = f2bf(x);
        }
        __syncthreads();  // the next minibatch reads the new bfloat16 weights
    }
}

hipError_t launch_mlp_infer(const uint16_t *wb, const float *feat, float *out, int n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(mlp_infer, dim3(n), dim3(64), 0, s, wb, feat, out, n);
    return hipGetLastError();
}

}  // namespace espec

using namespace espec;

namespace {

std::vector<uint16_t> to_bf16(const float *p, int n)
{
    std::vector<uint16_t> h(n);
    for (int i = 0; i < n; ++i) h[i] = f2bf(p[i]);
    return h;
}

}  // namespace

extern "C" {

int espec_mlp_infer(const float *params, const float *feat, int n, float *out)
{
    if (!params || !feat || !out || n < 1) return ESPEC_ERR_ARG;
    const std::vector<uint16_t> hb = to_bf16(params, kMlpNPar);
    DevBuf wb, f, o;
    ESPEC_HIP_TRY(wb.upload(hb.data(), sizeof(uint16_t) * hb.size()));
    ESPEC_HIP_TRY(f.upload(feat, sizeof(float) * kMlpIn * (size_t)n));
    ESPEC_HIP_TRY(o.alloc(sizeof(float) * kMlpOut * (size_t)n));
    ESPEC_HIP_TRY(launch_mlp_infer(wb.as<uint16_t>(), f.as<float>(), o.as<float>(), n, nullptr));
    ESPEC_HIP_TRY(hipMemcpy(out, o.p, sizeof(float) * kMlpOut * (size_t)n, hipMemcpyDeviceToHost));
    return ESPEC_OK;
}

int espec_mlp_grad(const float *params, const float *x, const float *t, float *grad, float *loss)
{
    if (!params || !x || !t || !grad || !loss) return ESPEC_ERR_ARG;
    const std::vector<uint16_t> hb = to_bf16(params, kMlpNPar);
    std::vector<int> idx(kMlpBatch);
    std::iota(idx.begin(), idx.end(), 0);
    DevBuf wb, dx, dt, di, dg, dl;
    ESPEC_HIP_TRY(wb.upload(hb.data(), sizeof(uint16_t) * hb.size()));
    ESPEC_HIP_TRY(dx.upload(x, sizeof(float) * kMlpIn * kMlpBatch));
    ESPEC_HIP_TRY(dt.upload(t, sizeof(float) * kMlpOut * kMlpBatch));
    ESPEC_HIP_TRY(di.upload(idx.data(), sizeof(int) * idx.size()));
    ESPEC_HIP_TRY(dg.alloc(sizeof(float) * kMlpNPar));
    ESPEC_HIP_TRY(dl.alloc(sizeof(float)));
    hipLaunchKernelGGL(mlp_grad_kernel, dim3(1), dim3(kTrainNT), 0, nullptr, wb.as<uint16_t>(),
                       dx.as<float>(), dt.as<float>(), di.as<int>(), dg.as<float>(), dl.as<float>());
    ESPEC_HIP_TRY(hipGetLastError());
    ESPEC_HIP_TRY(hipMemcpy(grad, dg.p, sizeof(float) * kMlpNPar, hipMemcpyDeviceToHost));
    ESPEC_HIP_TRY(hipMemcpy(loss, dl.p, sizeof(float), hipMemcpyDeviceToHost));
    return ESPEC_OK;
}

int espec_mlp_train(float *params, const float *x, const float *t, int n, int epochs, float lr,
                    unsigned seed, float *loss)
{
    if (!params || !x || !t || n < kMlpBatch || epochs < 1 || !(lr > 0)) return ESPEC_ERR_ARG;
    const int nb = n / kMlpBatch;
    const std::vector<uint16_t> hb = to_bf16(params, kMlpNPar);
    std::vector<float> zero(kMlpNPar, 0.0f);
    DevBuf master, mm, vv, gg, wb, dx, dt, di, dl;
    ESPEC_HIP_TRY(master.upload(params, sizeof(float) * kMlpNPar));
    ESPEC_HIP_TRY(mm.upload(zero.data(), sizeof(float) * kMlpNPar));
    ESPEC_HIP_TRY(vv.upload(zero.data(), sizeof(float) * kMlpNPar));
    ESPEC_HIP_TRY(gg.alloc(sizeof(float) * kMlpNPar));
    ESPEC_HIP_TRY(wb.upload(hb.data(), sizeof(uint16_t) * hb.size()));
    ESPEC_HIP_TRY(dx.upload(x, sizeof(float) * kMlpIn * (size_t)n));
    ESPEC_HIP_TRY(dt.upload(t, sizeof(float) * kMlpOut * (size_t)n));
    ESPEC_HIP_TRY(di.alloc(sizeof(int) * (size_t)nb * kMlpBatch));
    ESPEC_HIP_TRY(dl.alloc(sizeof(float) * (size_t)nb));
    AdamArgs a{master.as<float>(), mm.as<float>(), vv.as<float>(), gg.as<float>(),
               wb.as<uint16_t>(), lr, 0.9f, 0.999f, 1e-8f, 0};
    std::vector<int> perm(n);
    std::iota(perm.begin(), perm.end(), 0);
    std::mt19937 rng(seed);
    std::vector<float> lh(nb);
    for (int ep = 0; ep < epochs; ++ep) {
        std::shuffle(perm.begin(), perm.end(), rng);
        ESPEC_HIP_TRY(hipMemcpy(di.p, perm.data(), sizeof(int) * (size_t)nb * kMlpBatch,
                                hipMemcpyHostToDevice));
        hipLaunchKernelGGL(mlp_train_kernel, dim3(1), dim3(kTrainNT), 0, nullptr, a,
                           dx.as<float>(), dt.as<float>(), di.as<int>(), nb, dl.as<float>());
        ESPEC_HIP_TRY(hipGetLastError());
        ESPEC_HIP_TRY(hipMemcpy(lh.data(), dl.p, sizeof(float) * nb, hipMemcpyDeviceToHost));
        if (loss) {
            double s = 0;
            for (float v : lh) s += v;
            loss[ep] = (float)(s / nb);
        }
        a.step0 += nb;
    }
    ESPEC_HIP_TRY(hipMemcpy(params, master.p, sizeof(float) * kMlpNPar, hipMemcpyDeviceToHost));
    return ESPEC_OK;
}

}  // extern "C"
