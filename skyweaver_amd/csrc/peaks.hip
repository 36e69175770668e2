// peaks.hip -- microbenchmarks that confirm the roofline peaks on the device
// the library runs on: dense int8 MFMA rate, FP32 FMA rate and HBM copy
// bandwidth.  The copy kernel also calibrates hardware byte counters (a copy
// of N bytes must fetch N and write N bytes).  No reference counterpart.
#include "sw_common.hpp"

namespace sw {
namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

// 4 independent v_mfma_i32_32x32x32_i8 accumulation chains per wave
__global__ void __launch_bounds__(256) mfma_i8_peak_kernel(int* out, int iters)
{
    v4i a = {(int)threadIdx.x, 0x01020304, 0x05060708, 0x090a0b0c};
    v4i b = {0x0d0e0f10, (int)blockIdx.x, 0x11121314, 0x15161718};
    v16i c0 = {0}, c1 = {0}, c2 = {0}, c3 = {0};
    for(int i = 0; i < iters; ++i) {
        c0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c0, 0, 0, 0);
        c1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c1, 0, 0, 0);
        c2 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c2, 0, 0, 0);
        c3 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c3, 0, 0, 0);
    }
    int s = 0;
#pragma unroll
    for(int r = 0; r < 16; ++r) s += c0[r] ^ c1[r] ^ c2[r] ^ c3[r];
    out[blockIdx.x * 256 + threadIdx.x] = s;
}

// 8 independent FMA chains per lane
__global__ void __launch_bounds__(256) fma_f32_peak_kernel(float* out, int iters)
{
    float x[8];
#pragma unroll
    for(int k = 0; k < 8; ++k) x[k] = (float)(threadIdx.x + k);
    const float m = 0.999f, a = 0.001f;
    for(int i = 0; i < iters; ++i)
#pragma unroll
        for(int k = 0; k < 8; ++k) x[k] = fmaf(x[k], m, a);
    float s = 0;
#pragma unroll
    for(int k = 0; k < 8; ++k) s += x[k];
    out[blockIdx.x * 256 + threadIdx.x] = s;
}

__global__ void __launch_bounds__(256) copy_kernel(const uint4* __restrict__ in, uint4* __restrict__ out,
                                                   size_t n)
{
    for(size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) out[i] = in[i];
}

} // namespace
} // namespace sw

extern "C" int sw_microbench(int which, double* result, double* seconds, void* stream)
{
    SW_CHECK_ARG(result && seconds, "null argument");
    hipStream_t st = sw::as_stream(stream);
    hipEvent_t e0, e1;
    SW_HIP_TRY(hipEventCreate(&e0));
    SW_HIP_TRY(hipEventCreate(&e1));
    int dev = 0, ncu = 0;
    SW_HIP_TRY(hipGetDevice(&dev));
    SW_HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    float ms = 0;
    int rc = SW_OK;
    if(which == 0 || which == 2) {
        const int blocks = ncu * 8, iters = which == 0 ? 4096 : 65536;
        void* out = nullptr;
        SW_HIP_TRY(hipMalloc(&out, (size_t)blocks * 256 * 4));
        for(int rep = 0; rep < 2; ++rep) {  // the first launch warms up clocks
            SW_HIP_TRY(hipEventRecord(e0, st));
            if(which == 0)
                hipLaunchKernelGGL(sw::mfma_i8_peak_kernel, dim3(blocks), dim3(256), 0, st,
                                   static_cast<int*>(out), iters);
            else
                hipLaunchKernelGGL(sw::fma_f32_peak_kernel, dim3(blocks), dim3(256), 0, st,
                                   static_cast<float*>(out), iters);
            SW_HIP_TRY(hipEventRecord(e1, st));
            SW_HIP_TRY(hipEventSynchronize(e1));
        }
        SW_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        const double waves = (double)blocks * 4;
        // MFMA: 4 chains x 32*32*32 MACs x 2 ops; FMA: 8 chains x 64 lanes x 2 flops
        const double ops = which == 0 ? waves * iters * 4 * 65536.0 : waves * iters * 8 * 64 * 2.0;
        *result = ops / (ms * 1e-3) / 1e12;  // T(FL)OP/s
        (void)hipFree(out);
    } else if(which == 1) {
        const size_t bytes = (size_t)1 << 30;
        void *a = nullptr, *b = nullptr;
        SW_HIP_TRY(hipMalloc(&a, bytes));
        SW_HIP_TRY(hipMalloc(&b, bytes));
        SW_HIP_TRY(hipMemsetAsync(a, 1, bytes, st));
        for(int rep = 0; rep < 3; ++rep) {
            SW_HIP_TRY(hipEventRecord(e0, st));
            hipLaunchKernelGGL(sw::copy_kernel, dim3(ncu * 16), dim3(256), 0, st,
                               static_cast<const uint4*>(a), static_cast<uint4*>(b), bytes / 16);
            SW_HIP_TRY(hipEventRecord(e1, st));
            SW_HIP_TRY(hipEventSynchronize(e1));
        }
        SW_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        *result = 2.0 * bytes / (ms * 1e-3) / 1e9;  // GB/s (read + write)
        (void)hipFree(a);
        (void)hipFree(b);
    } else {
        sw::set_error("sw_microbench: which must be 0 (int8 MFMA), 1 (HBM copy) or 2 (FP32 FMA)");
        rc = SW_ERR_INVALID;
    }
    *seconds = ms * 1e-3;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}
