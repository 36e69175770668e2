// sw_common.hpp -- shared host/device helpers for the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "skyweaver_amd.h"

// Epilogues must reproduce the CPU oracle's float operation order exactly:
// no multiply-add contraction anywhere in this library.  Kernels that want
// an FMA (the FFT butterflies) call fmaf() explicitly.
#pragma clang fp contract(off)

namespace sw {

// ---------------------------------------------------------------- errors
void set_error(const std::string& msg);
const std::string& get_error();

struct Error {
    int code;
    std::string msg;
};

#define SW_CHECK_ARG(cond, msg)                                  \
    do {                                                         \
        if(!(cond)) {                                            \
            ::sw::set_error(std::string(__func__) + ": " + (msg)); \
            return SW_ERR_INVALID;                               \
        }                                                        \
    } while(0)

#define SW_HIP_TRY(expr)                                                       \
    do {                                                                       \
        hipError_t _e = (expr);                                                \
        if(_e != hipSuccess) {                                                 \
            ::sw::set_error(std::string(__func__) + ": " #expr " failed: " +   \
                            hipGetErrorString(_e));                            \
            return SW_ERR_HIP;                                                 \
        }                                                                      \
    } while(0)

// Check the launch that was just enqueued.
#define SW_LAUNCH_CHECK() SW_HIP_TRY(hipGetLastError())

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Library-owned scratch buffer (grown on demand, freed at exit), used when a
// caller passes a NULL workspace.  Not thread-safe across concurrent calls.
void* scratch(size_t bytes, int slot);

// The overlap a coherent-dedispersion plan gives a DM when none is
// configured (csrc/codedisp.hip).
int codedisp_overlap(double low_freq, double chan_bw, double tsamp, float dm);

// ---------------------------------------------------------- device helpers
// A raw buffer resource (gfx9 format) over [base, base + bytes): a buffer
// access takes one 32-bit VGPR offset plus a scalar offset, so a thread's
// run of accesses at uniform strides costs no per-access address
// arithmetic.  Callers keep every access inside the range they give.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_rsrc(const void* base, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}

struct cf32 {
    float x, y;
};

// |x|^2 = x.x*x.x + x.y*x.y (cpp/skyweaver/beamformer_utils.cuh:103-106)
__host__ __device__ __forceinline__ float mag2(float re, float im)
{
    return re * re + im * im;
}

// Stokes parameters of two polarisations (cpp/skyweaver/beamformer_utils.cuh:125-140).
// U and V are 2*Re / 2*Im of p0 * conj(p1) with the textbook complex product.
template <int S>
__host__ __device__ __forceinline__ float stokes(float r0, float i0, float r1,
                                                 float i1)
{
    if constexpr(S == SW_STOKES_I) {
        return mag2(r0, i0) + mag2(r1, i1);
    } else if constexpr(S == SW_STOKES_Q) {
        return mag2(r0, i0) - mag2(r1, i1);
    } else if constexpr(S == SW_STOKES_U) {
        // (r0*r1) - (i0*(-i1)) == (r0*r1) + (i0*i1) exactly in IEEE
        return 2 * ((r0 * r1) + (i0 * i1));
    } else {
        // (r0*(-i1)) + (i0*r1) == (i0*r1) - (r0*i1) exactly in IEEE
        return 2 * ((i0 * r1) - (r0 * i1));
    }
}

// clamp<int8_t>(float), cpp/skyweaver/beamformer_utils.cuh:186-195
__host__ __device__ __forceinline__ int8_t clamp_i8(float v)
{
    return static_cast<int8_t>(fmaxf(-128.0f, fminf(127.0f, v)));
}

// rescale, cpp/skyweaver/beamformer_utils.cuh:174-184
template <int S>
__host__ __device__ __forceinline__ float rescale(float power, float offset,
                                                  float scale)
{
    if constexpr(S == SW_STOKES_I)
        return rintf((power - offset) / scale);
    else
        return rintf(power / scale);
}

// ib_subtract, cpp/skyweaver/beamformer_utils.cuh:165-172 (multiplier 127^2 = 16129)
__host__ __device__ __forceinline__ float ib_subtract(float power, float ib,
                                                      float scale)
{
    return rintf((power - ib * 16129.0f) / scale);
}

__host__ __device__ constexpr int nstokes(int mode) { return mode == SW_STOKES_IQUV ? 4 : 1; }

} // namespace sw
