// Instruction-rate probe for the roofline report: the arithmetic rate one
// instruction form reaches on this device with no memory traffic in the
// timed loop (8 independent accumulators per wave, 8 waves per SIMD), to set
// beside the published peak.  Not on the solver's path.
#include "bse_common.hpp"
#include "../../include/bse_mi355x.h"
#include <algorithm>

typedef float pr_f4 __attribute__((ext_vector_type(4)));
typedef double pr_d4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_probe_f32(int iters, float* out)
{
    pr_f4 acc[8];
    for (int i = 0; i < 8; ++i) acc[i] = pr_f4{0, 0, 0, 0};
    const float a = 1.0f + threadIdx.x * 1e-7f, b = 1.0f - threadIdx.x * 1e-7f;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int i = 0; i < 8; ++i) acc[i] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[i], 0, 0, 0);
    }
    float s = 0.0f;
    for (int i = 0; i < 8; ++i) s += acc[i][0] + acc[i][1] + acc[i][2] + acc[i][3];
    out[blockIdx.x * 256 + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_probe_f64(int iters, double* out)
{
    pr_d4 acc[8];
    for (int i = 0; i < 8; ++i) acc[i] = pr_d4{0, 0, 0, 0};
    const double a = 1.0 + threadIdx.x * 1e-9, b = 1.0 - threadIdx.x * 1e-9;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int i = 0; i < 8; ++i) acc[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[i], 0, 0, 0);
    }
    double s = 0.0;
    for (int i = 0; i < 8; ++i) s += acc[i][0] + acc[i][1] + acc[i][2] + acc[i][3];
    out[blockIdx.x * 256 + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_probe_fma64(int iters, double* out)
{
    double acc[16];
    for (int i = 0; i < 16; ++i) acc[i] = threadIdx.x * 1e-9 + i;
    const double a = 1.0 - 1e-12, b = 1e-13;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int i = 0; i < 16; ++i) acc[i] = fma(acc[i], a, b);
    }
    double s = 0.0;
    for (int i = 0; i < 16; ++i) s += acc[i];
    out[blockIdx.x * 256 + threadIdx.x] = s;
}

extern "C" int bse_probe_rate(int which, double* tflops)
{
    if (!tflops || which < 0 || which > 2) return BSE_ERR_ARG;
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return BSE_ERR_HIP;
    const int nblk = std::max(1, ncu) * 8, iters = which == 2 ? 4000 : (which == 1 ? 2000 : 4000);
    // flops per launch: 4 waves x 8 accumulators x 2 m n k per MFMA, or 256
    // threads x 16 chains x 2 per FMA
    const double flops = which == 2 ? (double)nblk * 256 * iters * 16 * 2.0
                                    : (double)nblk * 4 * iters * 8 * 2048.0;
    void* out = nullptr;
    if (hipMalloc(&out, sizeof(double) * nblk * 256) != hipSuccess) return BSE_ERR_ALLOC;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = BSE_OK;
    double best = 0.0;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = BSE_ERR_HIP;
    for (int rep = 0; rep < 3 && rc == BSE_OK; ++rep) {
        (void)hipEventRecord(e0, 0);
        if (which == 0)
            hipLaunchKernelGGL(k_probe_f32, dim3(nblk), dim3(256), 0, 0, iters, (float*)out);
        else if (which == 1)
            hipLaunchKernelGGL(k_probe_f64, dim3(nblk), dim3(256), 0, 0, iters, (double*)out);
        else
            hipLaunchKernelGGL(k_probe_fma64, dim3(nblk), dim3(256), 0, 0, iters, (double*)out);
        (void)hipEventRecord(e1, 0);
        float ms = 0.0f;
        if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
            rc = BSE_ERR_HIP;
            break;
        }
        if (ms > 0.0f) best = std::max(best, flops / (ms * 1e-3) / 1e12);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(out);
    if (rc == BSE_OK) *tflops = best;
    return rc;
}
