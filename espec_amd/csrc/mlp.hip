// Kernels of the Krylov-step controller network (layout in mlp.hpp):
// inference, the gradient of one minibatch, and device-resident training
// with Adam.  Every layer product and every weight gradient is a chain of
// 16x16x16 bfloat16 matrix-core products with FP32 accumulation
// (v_mfma_f32_16x16x16_bf16); the FP32 master parameters are updated and
// their bfloat16 image refreshed after every minibatch.
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

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// d + A B for one 16x16x16 block on the matrix cores.  a(i, k) and b(k, j)
// return bfloat16-representable values; lane l supplies A's row i = l % 16
// and B's column j = l % 16 at k = 4 (l / 16) + v, v < 4.  On return,
// element v of lane l holds D's row 4 (l / 16) + v, column l % 16.
template <class FA, class FB>
__device__ __forceinline__ f32x4 mma(f32x4 d, FA a, FB b)
{
    const int l = threadIdx.x & 63, r = l & 15, q = (l >> 4) * 4;
    bf16x4 x, y;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        x[v] = (short)f2bf(a(r, q + v));
        y[v] = (short)f2bf(b(q + v, r));
    }
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(x, y, d, 0, 0, 0);
}

}  // namespace

// One wavefront per 16 samples: the three layers as matrix-core products
// over [16 samples x 64 units] tiles (inputs padded from 8 to 16, outputs
// from 2 to 16 with zeros), activations exchanged through LDS.
__global__ __launch_bounds__(64) void mlp_infer(const uint16_t *wb, const float *feat, float *out,
                                                int n)
{
    __shared__ float hs[16][kMlpHid + 1];
    const int s0 = blockIdx.x * 16, l = threadIdx.x, c = l & 15, r0 = (l >> 4) * 4;
#pragma unroll
    for (int nt = 0; nt < kMlpHid / 16; ++nt) {
        f32x4 d = {0, 0, 0, 0};
        d = mma(d,
                [&](int i, int k) {
                    return (k < kMlpIn && s0 + i < n) ? bfr(feat[(long long)(s0 + i) * kMlpIn + k])
                                                      : 0.0f;
                },
                [&](int k, int j) { return k < kMlpIn ? w(wb, kMlpW1 + (nt * 16 + j) * kMlpIn + k) : 0.0f; });
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int j = nt * 16 + c;
            hs[r0 + v][j] = bfr(tanhf(d[v] + w(wb, kMlpB1 + j)));
        }
    }
    __syncthreads();
    f32x4 d2[kMlpHid / 16] = {};
#pragma unroll
    for (int kt = 0; kt < kMlpHid / 16; ++kt)
#pragma unroll
        for (int nt = 0; nt < kMlpHid / 16; ++nt)
            d2[nt] = mma(d2[nt], [&](int i, int k) { return hs[i][kt * 16 + k]; },
                         [&](int k, int j) { return w(wb, kMlpW2 + (nt * 16 + j) * kMlpHid + kt * 16 + k); });
    __syncthreads();  // every read of the first layer's activations is done
#pragma unroll
    for (int nt = 0; nt < kMlpHid / 16; ++nt)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int j = nt * 16 + c;
            hs[r0 + v][j] = bfr(tanhf(d2[nt][v] + w(wb, kMlpB2 + j)));
        }
    __syncthreads();
    f32x4 y = {0, 0, 0, 0};
#pragma unroll
    for (int kt = 0; kt < kMlpHid / 16; ++kt)
        y = mma(y, [&](int i, int k) { return hs[i][kt * 16 + k]; },
                [&](int k, int j) { return j < kMlpOut ? w(wb, kMlpW3 + j * kMlpHid + kt * 16 + k) : 0.0f; });
    if (c < kMlpOut)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int sm = s0 + r0 + v;
            if (sm < n) out[(long long)sm * kMlpOut + c] = sigm(y[v] + w(wb, kMlpB3 + c));
        }
}

namespace {

constexpr int kTrainNT = 256;  // four wavefronts
constexpr int B = kMlpBatch;
constexpr int NT16 = kMlpHid / 16;

struct MlpLds {
    float x[B][kMlpIn];
    float h1[B][kMlpHid + 1];
    float h2[B][kMlpHid + 1];  // after the output layer's backward: dz1
    float d2[B][kMlpHid + 1];  // dz2
    float dy[B][kMlpOut];
    float red[kTrainNT / 64];
};

// Forward and backward of one minibatch (rows idx[0..B) of x, t) for the
// loss mean((sigmoid(y) - t)^2) over the B x 2 outputs.  Every operand of
// a product is a bfloat16 value (weights, inputs, activations and the
// back-propagated dz), every sum FP32.  Wavefront q takes samples
// 16q..16q+15 in the forward pass and the dz1 product, and the q-th
// 16-column (or 16-row) block of each weight gradient.  grad (global)
// receives dL/dparam; returns the loss in thread 0.
__device__ float minibatch(const uint16_t *wb, const float *X, const float *T, const int *idx,
                           float *grad, MlpLds &L)
{
    const int t = threadIdx.x, q = t >> 6, l = t & 63, c = l & 15, r0 = (l >> 4) * 4;
    const int R = 16 * q;
    for (int e = t; e < B * kMlpIn; e += kTrainNT) {
        const int b = e / kMlpIn, k = e % kMlpIn;
        L.x[b][k] = bfr(X[(long long)idx[b] * kMlpIn + k]);
    }
    __syncthreads();
    // forward, samples R..R+15
#pragma unroll
    for (int nt = 0; nt < NT16; ++nt) {
        f32x4 d = {0, 0, 0, 0};
        d = mma(d, [&](int i, int k) { return k < kMlpIn ? L.x[R + i][k] : 0.0f; },
                [&](int k, int j) { return k < kMlpIn ? w(wb, kMlpW1 + (nt * 16 + j) * kMlpIn + k) : 0.0f; });
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int j = nt * 16 + c;
            L.h1[R + r0 + v][j] = bfr(tanhf(d[v] + w(wb, kMlpB1 + j)));
        }
    }
    __syncthreads();
    {
        f32x4 d2[NT16] = {};
#pragma unroll
        for (int kt = 0; kt < NT16; ++kt)
#pragma unroll
            for (int nt = 0; nt < NT16; ++nt)
                d2[nt] = mma(d2[nt], [&](int i, int k) { return L.h1[R + i][kt * 16 + k]; },
                             [&](int k, int j) { return w(wb, kMlpW2 + (nt * 16 + j) * kMlpHid + kt * 16 + k); });
#pragma unroll
        for (int nt = 0; nt < NT16; ++nt)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int j = nt * 16 + c;
                L.h2[R + r0 + v][j] = bfr(tanhf(d2[nt][v] + w(wb, kMlpB2 + j)));
            }
    }
    __syncthreads();
    float lacc = 0;
    {
        f32x4 y = {0, 0, 0, 0};
#pragma unroll
        for (int kt = 0; kt < NT16; ++kt)
            y = mma(y, [&](int i, int k) { return L.h2[R + i][kt * 16 + k]; },
                    [&](int k, int j) { return j < kMlpOut ? w(wb, kMlpW3 + j * kMlpHid + kt * 16 + k) : 0.0f; });
        if (c < kMlpOut)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int b = R + r0 + v;
                const float p = sigm(y[v] + w(wb, kMlpB3 + c));
                const float r = p - T[(long long)idx[b] * kMlpOut + c];
                lacc += r * r;
                L.dy[b][c] = bfr(2.0f * r / (float)(B * kMlpOut) * p * (1.0f - p));
            }
    }
    lacc = wave_sum_f(lacc);
    if (l == 0) L.red[q] = lacc;
    __syncthreads();
    // output layer: dW3 = dy^T h2 (rows o < 2 of a 16-row block), block q of its columns
    {
        f32x4 d = {0, 0, 0, 0};
#pragma unroll
        for (int kt = 0; kt < B / 16; ++kt)
            d = mma(d, [&](int i, int k) { return i < kMlpOut ? L.dy[kt * 16 + k][i] : 0.0f; },
                    [&](int k, int j) { return L.h2[kt * 16 + k][R + j]; });
        if (r0 == 0)
#pragma unroll
            for (int v = 0; v < kMlpOut; ++v) grad[kMlpW3 + v * kMlpHid + R + c] = d[v];
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
    // second hidden layer: dW2 = dz2^T h1 (row block q), dz1 = dz2 W2 (sample block q)
    {
        f32x4 g[NT16] = {}, d1[NT16] = {};
#pragma unroll
        for (int kt = 0; kt < B / 16; ++kt)
#pragma unroll
            for (int nt = 0; nt < NT16; ++nt)
                g[nt] = mma(g[nt], [&](int i, int k) { return L.d2[kt * 16 + k][R + i]; },
                            [&](int k, int j) { return L.h1[kt * 16 + k][nt * 16 + j]; });
#pragma unroll
        for (int kt = 0; kt < NT16; ++kt)
#pragma unroll
            for (int nt = 0; nt < NT16; ++nt)
                d1[nt] = mma(d1[nt], [&](int i, int k) { return L.d2[R + i][kt * 16 + k]; },
                             [&](int k, int j) { return w(wb, kMlpW2 + (kt * 16 + k) * kMlpHid + nt * 16 + j); });
#pragma unroll
        for (int nt = 0; nt < NT16; ++nt)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = R + r0 + v, k = nt * 16 + c;
                grad[kMlpW2 + i * kMlpHid + k] = g[nt][v];
                // dz1 overwrites h2 (rows R..R+15 are this wavefront's alone here)
                L.h2[i][k] = bfr(d1[nt][v] * (1.0f - L.h1[i][k] * L.h1[i][k]));
            }
    }
    for (int j = t; j < kMlpHid; j += kTrainNT) {
        float g = 0;
        for (int b = 0; b < B; ++b) g += L.d2[b][j];
        grad[kMlpB2 + j] = g;
    }
    __syncthreads();
    // first hidden layer: dW1 = dz1^T x (row block q, inputs padded to 16)
    {
        f32x4 d = {0, 0, 0, 0};
#pragma unroll
        for (int kt = 0; kt < B / 16; ++kt)
            d = mma(d, [&](int i, int k) { return L.h2[kt * 16 + k][R + i]; },
                    [&](int k, int j) { return j < kMlpIn ? L.x[kt * 16 + k][j] : 0.0f; });
        if (c < kMlpIn)
#pragma unroll
            for (int v = 0; v < 4; ++v) grad[kMlpW1 + (R + r0 + v) * kMlpIn + c] = d[v];
    }
    for (int k = t; k < kMlpHid; k += kTrainNT) {
        float g = 0;
        for (int b = 0; b < B; ++b) g += L.h2[b][k];
        grad[kMlpB1 + k] = g;
    }
    __syncthreads();  // grad complete and visible to the block
    float loss = 0;
    if (t == 0)
        for (int k = 0; k < kTrainNT / 64; ++k) loss += L.red[k];
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
    hipLaunchKernelGGL(mlp_infer, dim3((n + 15) / 16), dim3(64), 0, s, wb, feat, out, n);
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
