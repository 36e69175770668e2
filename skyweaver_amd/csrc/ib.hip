// ib.hip -- incoherent beamformer: weighted Stokes power summed over
// antennas and time-scrunched, for every beamset in one launch.
//
// Reference: IncoherentBeamformer::beamform / icbf_ftpa_general_k,
// sw/src/IncoherentBeamformer.cu:16-183; oracle
// sw/test/src/IncoherentBeamformerTester.cu:41-103:
//   P[s][t'][f] = sum_{t in window} sum_a Stokes(v[f][t][0][a], v[f][t][1][a]) * w[s][a]
//   raw = P, out = clamp(rescale(P, offset[s][f], scale[s][f]))
// Unlike the reference kernel (which applies antenna_weights[a] of beamset 0
// to every beamset and caps the antenna count at 64), every beamset uses its
// own weight row and any antenna count works.
//
// Each lane owns one output sample (t', f) and accumulates in exactly the
// oracle's order (time ascending, antenna ascending, one rounding for the
// product and one for the sum), so raw powers are bit-identical to the
// oracle for any weights, not only when the sums happen to be exact.  The
// arithmetic is tiny next to the HBM read of the FTPA block.
#include "sw_common.hpp"

namespace sw {
namespace {

constexpr int kThreads = 256;

template <int S, int SMAX>
__global__ void __launch_bounds__(kThreads)
ib_kernel(const uint8_t* __restrict__ ftpa, float* __restrict__ raw, int8_t* __restrict__ out,
          const float* __restrict__ scale, const float* __restrict__ offset,
          const float* __restrict__ antw, int nchans, int nsamples, int nant, int tscrunch,
          int nbeamsets)
{
    constexpr int NV = nstokes(S);
    const int f   = blockIdx.y;
    const int tp  = blockIdx.x * kThreads + threadIdx.x;   // output sample
    const int nsamps_out = nsamples / tscrunch;
    if(tp >= nsamps_out) return;
    float power[SMAX][NV];
#pragma unroll
    for(int s = 0; s < SMAX; ++s)
#pragma unroll
        for(int k = 0; k < NV; ++k) power[s][k] = 0.0f;

    const size_t row_bytes = (size_t)4 * nant;       // P * A * 2
    const uint8_t* base = ftpa + ((size_t)f * nsamples + (size_t)tp * tscrunch) * row_bytes;
    for(int tt = 0; tt < tscrunch; ++tt) {
        const uint8_t* row = base + (size_t)tt * row_bytes;
        for(int a0 = 0; a0 < nant; a0 += 4) {
            // 4 antennas of each polarisation: 8 bytes each
            const uint2 q0 = *reinterpret_cast<const uint2*>(row + (size_t)a0 * 2);
            const uint2 q1 = *reinterpret_cast<const uint2*>(row + (size_t)(nant + a0) * 2);
            const uint32_t w0[2] = {q0.x, q0.y};
            const uint32_t w1[2] = {q1.x, q1.y};
#pragma unroll
            for(int j = 0; j < 4; ++j) {
                const uint32_t d0 = w0[j >> 1] >> (16 * (j & 1));
                const uint32_t d1 = w1[j >> 1] >> (16 * (j & 1));
                const float r0 = (float)(int8_t)(d0), i0 = (float)(int8_t)(d0 >> 8);
                const float r1 = (float)(int8_t)(d1), i1 = (float)(int8_t)(d1 >> 8);
                float st[NV];
                if constexpr(NV == 1) {
                    st[0] = stokes<S>(r0, i0, r1, i1);
                } else {
                    st[0] = stokes<SW_STOKES_I>(r0, i0, r1, i1);
                    st[1] = stokes<SW_STOKES_Q>(r0, i0, r1, i1);
                    st[2] = stokes<SW_STOKES_U>(r0, i0, r1, i1);
                    st[3] = stokes<SW_STOKES_V>(r0, i0, r1, i1);
                }
#pragma unroll
                for(int s = 0; s < SMAX; ++s) {
                    if(s < nbeamsets) {
                        const float w = antw[s * nant + a0 + j];
#pragma unroll
                        for(int k = 0; k < NV; ++k) power[s][k] += st[k] * w;
                    }
                }
            }
        }
    }
#pragma unroll
    for(int s = 0; s < SMAX; ++s) {
        if(s < nbeamsets) {
            const size_t o  = ((size_t)s * nsamps_out + tp) * nchans + f;
            const float sc  = scale[s * nchans + f];
            const float off = offset[s * nchans + f];
            if constexpr(NV == 1) {
                raw[o] = power[s][0];
                out[o] = clamp_i8(rescale<S>(power[s][0], off, sc));
            } else {
                reinterpret_cast<float4*>(raw)[o] =
                    make_float4(power[s][0], power[s][1], power[s][2], power[s][3]);
                char4 q;
                q.x = clamp_i8(rescale<SW_STOKES_I>(power[s][0], off, sc));
                q.y = clamp_i8(rescale<SW_STOKES_Q>(power[s][1], off, sc));
                q.z = clamp_i8(rescale<SW_STOKES_U>(power[s][2], off, sc));
                q.w = clamp_i8(rescale<SW_STOKES_V>(power[s][3], off, sc));
                reinterpret_cast<char4*>(out)[o] = q;
            }
        }
    }
}

template <int S>
int launch_ib(const uint8_t* ftpa, float* raw, int8_t* out, const float* scale,
              const float* offset, const float* antw, const sw_bf_dims& d, hipStream_t st)
{
    const int nout = d.nsamples / d.tscrunch;
    dim3 grid((nout + kThreads - 1) / kThreads, d.nchans);
#define SW_IB_LAUNCH(SM)                                                                     \
    hipLaunchKernelGGL((ib_kernel<S, SM>), grid, dim3(kThreads), 0, st, ftpa, raw, out,       \
                       scale, offset, antw, d.nchans, d.nsamples, d.nantennas, d.tscrunch,   \
                       d.nbeamsets)
    if(d.nbeamsets <= 1)
        SW_IB_LAUNCH(1);
    else if(d.nbeamsets <= 4)
        SW_IB_LAUNCH(4);
    else if(d.nbeamsets <= 8)
        SW_IB_LAUNCH(8);
    else
        SW_IB_LAUNCH(16);
#undef SW_IB_LAUNCH
    SW_LAUNCH_CHECK();
    return SW_OK;
}

} // namespace
} // namespace sw

extern "C" int sw_ib_beamform(const void* d_ftpa, float* d_raw, void* d_out,
                              const float* d_scale, const float* d_offset,
                              const float* d_ant_weights, const sw_bf_dims* dims,
                              void* stream)
{
    SW_CHECK_ARG(dims, "null dims");
    const sw_bf_dims& d = *dims;
    SW_CHECK_ARG(d_ftpa && d_raw && d_out && d_scale && d_offset && d_ant_weights, "null buffer");
    SW_CHECK_ARG(d.nbeamsets > 0, "Number of beamsets must be greater than zero");
    SW_CHECK_ARG(d.nbeamsets <= 16, "at most 16 beamsets are supported");
    SW_CHECK_ARG(d.nchans > 0 && d.nchans <= 65535, "bad channel count");
    SW_CHECK_ARG(d.nantennas > 0 && d.nantennas % 4 == 0, "antenna count must be a multiple of 4");
    SW_CHECK_ARG(d.tscrunch > 0 && d.nsamples > 0 && d.nsamples % d.tscrunch == 0,
                 "nsamples must be a positive multiple of tscrunch");
    SW_CHECK_ARG(d.stokes >= SW_STOKES_I && d.stokes <= SW_STOKES_IQUV, "bad stokes mode");
    SW_CHECK_ARG((reinterpret_cast<uintptr_t>(d_ftpa) & 7) == 0, "input must be 8-byte aligned");
    auto* in  = static_cast<const uint8_t*>(d_ftpa);
    auto* out = static_cast<int8_t*>(d_out);
    hipStream_t st = sw::as_stream(stream);
    switch(d.stokes) {
    case SW_STOKES_I: return sw::launch_ib<SW_STOKES_I>(in, d_raw, out, d_scale, d_offset, d_ant_weights, d, st);
    case SW_STOKES_Q: return sw::launch_ib<SW_STOKES_Q>(in, d_raw, out, d_scale, d_offset, d_ant_weights, d, st);
    case SW_STOKES_U: return sw::launch_ib<SW_STOKES_U>(in, d_raw, out, d_scale, d_offset, d_ant_weights, d, st);
    case SW_STOKES_V: return sw::launch_ib<SW_STOKES_V>(in, d_raw, out, d_scale, d_offset, d_ant_weights, d, st);
    default: return sw::launch_ib<SW_STOKES_IQUV>(in, d_raw, out, d_scale, d_offset, d_ant_weights, d, st);
    }
}
