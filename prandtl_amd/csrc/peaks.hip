// Machine peaks measured on the device the library runs on: an FMA-chain
// kernel (FP64 or unpacked FP32 vector arithmetic) and a streaming copy.
// The roofline of the RHS kernels is quoted against the datasheet peaks;
// these give the achievable ones beside them.
#include <hip/hip_runtime.h>

#include "../../include/prandtl_hip.h"

namespace {

constexpr int PEAK_BLOCK = 256;
constexpr int PEAK_CHAINS = 8;   // independent FMA chains per thread

// Each thread runs PEAK_CHAINS independent chains v <- v a + b, so the FMA
// pipeline always has an instruction ready; the sums are written only when
// they equal a value they never reach (kept live, never stored).
template <typename T>
__global__ void __launch_bounds__(PEAK_BLOCK) k_peak_fma(T *sink, int iters, T a, T b)
{
    T v[PEAK_CHAINS];
#pragma unroll
    for (int c = 0; c < PEAK_CHAINS; c++) v[c] = T(threadIdx.x + c) * T(1e-3);
    for (int i = 0; i < iters; i++) {
#pragma unroll
        for (int c = 0; c < PEAK_CHAINS; c++) v[c] = fma(v[c], a, b);
    }
    T s = v[0];
#pragma unroll
    for (int c = 1; c < PEAK_CHAINS; c++) s += v[c];
    if (s == T(-12345.0)) sink[0] = s;
}

// grid-stride copy of n 16-byte words
__global__ void __launch_bounds__(PEAK_BLOCK) k_peak_copy(const double2 *__restrict__ src, double2 *__restrict__ dst,
                                                         size_t n)
{
    const size_t stride = (size_t)gridDim.x * PEAK_BLOCK;
    for (size_t i = (size_t)blockIdx.x * PEAK_BLOCK + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

int peak_err(hipError_t e) { return e == hipSuccess ? PRANDTL_OK : PRANDTL_ERR_HIP; }

} // namespace

extern "C" {

int prandtl_peak_fma(int bits, void *sink, double *tflops, void *stream)
{
    if ((bits != 64 && bits != 32) || !sink || !tflops) return PRANDTL_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    int dev = 0, ncu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return peak_err(e);
    // 8 workgroups of 4 waves per CU: 8 waves per SIMD
    const int blocks = 8 * ncu, iters = 4096;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    float ms = 0.0f;
    for (int rep = 0; rep < 2 && e == hipSuccess; rep++) {   // the first launch warms up
        e = hipEventRecord(e0, s);
        if (bits == 64)
            hipLaunchKernelGGL(k_peak_fma<double>, dim3(blocks), dim3(PEAK_BLOCK), 0, s, static_cast<double *>(sink),
                               iters, 0.999999, 1e-7);
        else
            hipLaunchKernelGGL(k_peak_fma<float>, dim3(blocks), dim3(PEAK_BLOCK), 0, s, static_cast<float *>(sink),
                               iters, 0.999999f, 1e-7f);
        if (e == hipSuccess) e = hipGetLastError();
        if (e == hipSuccess) e = hipEventRecord(e1, s);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (e != hipSuccess) return peak_err(e);
    const double flops = 2.0 * PEAK_CHAINS * (double)iters * blocks * PEAK_BLOCK;
    *tflops = flops / (ms * 1e-3) * 1e-12;
    return PRANDTL_OK;
}

int prandtl_peak_copy(const void *src, void *dst, size_t bytes, int reps, double *us, void *stream)
{
    if (!src || !dst || !us || reps < 1 || bytes < 16 || bytes % 16) return PRANDTL_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    int dev = 0, ncu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return peak_err(e);
    const size_t n = bytes / 16;
    const size_t want = (n + PEAK_BLOCK - 1) / PEAK_BLOCK;
    const int blocks = (int)(want < (size_t)(16 * ncu) ? want : (size_t)(16 * ncu));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_peak_copy, dim3(blocks), dim3(PEAK_BLOCK), 0, s, static_cast<const double2 *>(src),
                           static_cast<double2 *>(dst), n);
        e = hipEventRecord(e0, s);
    }
    for (int r = 0; r < reps && e == hipSuccess; r++) {
        hipLaunchKernelGGL(k_peak_copy, dim3(blocks), dim3(PEAK_BLOCK), 0, s, static_cast<const double2 *>(src),
                           static_cast<double2 *>(dst), n);
        e = hipGetLastError();
    }
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventRecord(e1, s);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (e != hipSuccess) return peak_err(e);
    *us = 1e3 * ms / reps;
    return PRANDTL_OK;
}

} // extern "C"
