// ib.hip -- incoherent beamformer: weighted Stokes power summed over
// antennas and time-scrunched, for every beamset in one launch.
//
// Reference: IncoherentBeamformer::beamform / icbf_ftpa_general_k,
// cpp/skyweaver/src/IncoherentBeamformer.cu:16-183; oracle
// cpp/skyweaver/test/src/IncoherentBeamformerTester.cu:41-103:
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
// rows are staged through LDS so that the HBM reads are coalesced.
//
// Exact fast path (Stokes I, antenna weights all 0 or 1): every term
// w_a (re^2 + im^2) is a non-negative integer, so while the total stays
// below 2^24 every partial sum of the oracle's float chain is exact and the
// float result equals the integer sum.  The kernel then sums with
// v_dot4_i32_i8 (two antennas of one polarisation per word, masked by the
// beamset's 0/1 weights) in int32 and converts once; a wave whose total
// reaches 2^24 on any lane recomputes its outputs with the float chain.
#include "sw_common.hpp"

namespace sw {
namespace {

constexpr int kOut = 64;        // output samples per workgroup (one wave)
constexpr int kMaskWords = 64;  // 0/1 weight masks for up to 128 antennas

// One wave per (64 consecutive output samples, channel).  For each step tt of
// the scrunch window the 64 rows t'*ts + tt (one per output) are staged in
// LDS with coalesced 16-byte loads; each lane then accumulates its own row,
// antennas ascending, for every beamset.
template <int S, int SMAX>
__global__ void __launch_bounds__(kOut)
ib_kernel(const uint8_t* __restrict__ ftpa, float* __restrict__ raw, int8_t* __restrict__ out,
          const float* __restrict__ scale, const float* __restrict__ offset,
          const float* __restrict__ antw, int nchans, int nsamples, int nant, int tscrunch,
          int nbeamsets)
{
    constexpr int NV = nstokes(S);
    extern __shared__ uint4 rows[];           // [kOut][rowq + 1]
    const int f    = blockIdx.y;
    const int lane = threadIdx.x;
    const int nsamps_out = nsamples / tscrunch;
    const int tp0  = blockIdx.x * kOut;
    const int tp   = tp0 + lane;
    const int rowq = nant / 4;                // 16-byte chunks per (t) row: 4*nant bytes
    const int ldsq = rowq + 1;
    const int nrows = min(kOut, nsamps_out - tp0);
    float power[SMAX][NV];
#pragma unroll
    for(int s = 0; s < SMAX; ++s)
#pragma unroll
        for(int k = 0; k < NV; ++k) power[s][k] = 0.0f;
    const uint4* chan = reinterpret_cast<const uint4*>(ftpa + (size_t)f * nsamples * 4 * nant);

    // ---- exact integer path
    __shared__ uint32_t wmask[SMAX][kMaskWords];  // per word: 0xFFFF per antenna of weight 1
    bool exact = false;
    if constexpr(S == SW_STOKES_I) {
        if(nant <= 2 * kMaskWords) {
            bool binary = true;
            for(int i = lane; i < nbeamsets * nant / 2; i += kOut) {
                const int sb = i / (nant / 2), j = i % (nant / 2);
                const float w0 = antw[sb * nant + 2 * j], w1 = antw[sb * nant + 2 * j + 1];
                binary = binary && (w0 == 0.0f || w0 == 1.0f) && (w1 == 0.0f || w1 == 1.0f);
                wmask[sb][j] = (w0 != 0.0f ? 0x0000FFFFu : 0u) | (w1 != 0.0f ? 0xFFFF0000u : 0u);
            }
            exact = __syncthreads_and(binary);
        }
    }
    if(exact) {
        int isum[SMAX];
#pragma unroll
        for(int s = 0; s < SMAX; ++s) isum[s] = 0;
        const int hw = nant / 2;  // words per polarisation
        for(int tt = 0; tt < tscrunch; ++tt) {
            __syncthreads();
            for(int i = lane; i < nrows * rowq; i += kOut) {
                const int r = i / rowq, q = i % rowq;
                rows[r * ldsq + q] = chan[((size_t)(tp0 + r) * tscrunch + tt) * rowq + q];
            }
            __syncthreads();
            if(lane < nrows) {
                const uint32_t* rw = reinterpret_cast<const uint32_t*>(rows + lane * ldsq);
#pragma unroll
                for(int s = 0; s < SMAX; ++s) {
                    if(s < nbeamsets) {
                        int acc = 0;
                        for(int j = 0; j < hw; ++j) {
                            const int m = (int)wmask[s][j];
                            const int x0 = (int)rw[j], x1 = (int)rw[hw + j];
                            acc = __builtin_amdgcn_sdot4(x0 & m, x0, acc, false);
                            acc = __builtin_amdgcn_sdot4(x1 & m, x1, acc, false);
                        }
                        isum[s] += acc;
                    }
                }
            }
        }
        bool big = false;
#pragma unroll
        for(int s = 0; s < SMAX; ++s) {
            big = big || (isum[s] >= (1 << 24));
            power[s][0] = (float)isum[s];
        }
        exact = !__any(big);  // wave-uniform
    }
    if(!exact) {
#pragma unroll
    for(int s = 0; s < SMAX; ++s)
#pragma unroll
        for(int k = 0; k < NV; ++k) power[s][k] = 0.0f;
    for(int tt = 0; tt < tscrunch; ++tt) {
        __syncthreads();
        for(int i = lane; i < nrows * rowq; i += kOut) {
            const int r = i / rowq, q = i % rowq;
            rows[r * ldsq + q] = chan[((size_t)(tp0 + r) * tscrunch + tt) * rowq + q];
        }
        __syncthreads();
        if(lane < nrows) {
            const uint8_t* rb = reinterpret_cast<const uint8_t*>(rows + lane * ldsq);
            for(int a4 = 0; a4 < nant; a4 += 4) {
                // 4 antennas of pol 0 and of pol 1 (8 bytes each, 8-byte aligned)
                const uint2 q0 = *reinterpret_cast<const uint2*>(rb + a4 * 2);
                const uint2 q1 = *reinterpret_cast<const uint2*>(rb + (nant + a4) * 2);
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
                            const float w = antw[s * nant + a4 + j];
#pragma unroll
                            for(int k = 0; k < NV; ++k) power[s][k] += st[k] * w;
                        }
                    }
                }
            }
        }
    }
    }
    if(tp >= nsamps_out) return;
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
    dim3 grid((nout + kOut - 1) / kOut, d.nchans);
    const size_t lds = (size_t)kOut * (d.nantennas / 4 + 1) * 16;
#define SW_IB_LAUNCH(SM)                                                                     \
    hipLaunchKernelGGL((ib_kernel<S, SM>), grid, dim3(kOut), lds, st, ftpa, raw, out,         \
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
    SW_CHECK_ARG((reinterpret_cast<uintptr_t>(d_ftpa) & 15) == 0, "input must be 16-byte aligned");
    SW_CHECK_ARG(d.nantennas <= 512, "at most 512 antennas");
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
