// Streaming copy kernel used to calibrate HBM bandwidth and the rocprof
// FETCH_SIZE / WRITE_SIZE counters against a transfer of known size.
#include "api_common.hpp"

namespace {

// One 16-byte element per thread over a grid that covers the buffer: the
// workgroups in flight then sweep a contiguous window of HBM, which on
// MI355X streams faster (about 6.2 TB/s at 1 GiB) than a grid-stride loop
// over a fixed grid (about 4.7 TB/s).
__global__ __launch_bounds__(256) void copy_probe(const float4 *__restrict__ x, float4 *__restrict__ y,
                                                  long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = x[i];
}

}  // namespace

extern "C" int espec_bw_probe(size_t nbytes, int iters, double *ms, double *bytes)
{
    if (nbytes < 16 || iters < 1 || !ms || !bytes) return ESPEC_ERR_ARG;
    const long long n = (long long)(nbytes / 16);
    espec::DevBuf a, b;
    ESPEC_HIP_TRY(a.alloc(n * 16));
    ESPEC_HIP_TRY(b.alloc(n * 16));
    ESPEC_HIP_TRY(hipMemset(a.p, 0, n * 16));
    hipEvent_t e0, e1;
    ESPEC_HIP_TRY(hipEventCreate(&e0));
    ESPEC_HIP_TRY(hipEventCreate(&e1));
    const long long grid = (n + 255) / 256;
    if (grid > 0x7fffffffLL) return ESPEC_ERR_ARG;
    hipLaunchKernelGGL(copy_probe, dim3((unsigned)grid), dim3(256), 0, 0, a.as<float4>(), b.as<float4>(), n);
    ESPEC_HIP_TRY(hipGetLastError());
    ESPEC_HIP_TRY(hipEventRecord(e0, 0));
    for (int k = 0; k < iters; ++k)
        hipLaunchKernelGGL(copy_probe, dim3((unsigned)grid), dim3(256), 0, 0, a.as<float4>(), b.as<float4>(), n);
    ESPEC_HIP_TRY(hipGetLastError());
    ESPEC_HIP_TRY(hipEventRecord(e1, 0));
    ESPEC_HIP_TRY(hipEventSynchronize(e1));
    float t = 0;
    ESPEC_HIP_TRY(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = t / iters;
    *bytes = 2.0 * n * 16;
    return ESPEC_OK;
}
