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

// Time-major input (the coherent dedisperser's ring): sample t of column
// c = p A + a of channel f at ring + ((f 2A + c) ld + w0 + t) 2.  Stokes I
// with 0/1 antenna weights only.  One lane per output sample t' of channel
// f; a lane reads its ts consecutive samples of every column (64 lanes read
// 64 ts 2 contiguous bytes), sums re^2 + im^2 of the column with
// v_dot4_i32_i8 and adds the column sum to the beamsets that weight the
// antenna 1.  As in ib_kernel the integer total equals the oracle's float
// chain while below 2^24; a lane at or above it recomputes its outputs with
// the float chain in the oracle's order (time, then antenna ascending).
constexpr int kTmThreads = 256;
constexpr int kTmMaxCols = 256;  // 128 antennas

template <int SMAX, bool V16>
__global__ void __launch_bounds__(kTmThreads)
ib_tm_kernel(const uint8_t* __restrict__ ring, long ld, long w0, float* __restrict__ raw,
             int8_t* __restrict__ out, const float* __restrict__ scale, const float* __restrict__ offset,
             const float* __restrict__ antw, int nchans, int nsamples, int nant, int ts, int nbeamsets)
{
    __shared__ uint8_t colw[SMAX][kTmMaxCols];  // weight (0/1) of column c in beamset s
    const int PA = 2 * nant;
    for(int i = threadIdx.x; i < nbeamsets * PA; i += kTmThreads) {
        const int sb = i / PA, c = i % PA;
        colw[sb][c] = antw[sb * nant + c % nant] != 0.0f ? 1 : 0;
    }
    __syncthreads();
    const int f = blockIdx.y;
    const int nsamps_out = nsamples / ts;
    const int tp = blockIdx.x * kTmThreads + threadIdx.x;
    if(tp >= nsamps_out) return;  // after the only barrier
    const uint8_t* base = ring + ((size_t)f * PA * ld + w0 + (size_t)tp * ts) * 2;
    int isum[SMAX];
#pragma unroll
    for(int s = 0; s < SMAX; ++s) isum[s] = 0;
#pragma unroll 4
    for(int c = 0; c < PA; ++c) {
        const uint8_t* x = base + (size_t)c * ld * 2;
        int acc = 0;  // sum over the window of re^2 + im^2 of column c
        if constexpr(V16) {  // ts % 8 == 0: 16-byte aligned runs
            for(int k = 0; k < ts * 2; k += 16) {
                const uint4 v = *reinterpret_cast<const uint4*>(x + k);
                acc = __builtin_amdgcn_sdot4((int)v.x, (int)v.x, acc, false);
                acc = __builtin_amdgcn_sdot4((int)v.y, (int)v.y, acc, false);
                acc = __builtin_amdgcn_sdot4((int)v.z, (int)v.z, acc, false);
                acc = __builtin_amdgcn_sdot4((int)v.w, (int)v.w, acc, false);
            }
        } else {
            for(int k = 0; k < ts; ++k) {
                const int v = (int)*reinterpret_cast<const uint16_t*>(x + 2 * k);
                acc = __builtin_amdgcn_sdot4(v, v, acc, false);
            }
        }
#pragma unroll
        for(int s = 0; s < SMAX; ++s)
            if(s < nbeamsets && colw[s][c]) isum[s] += acc;
    }
    float power[SMAX];
    bool big = false;
#pragma unroll
    for(int s = 0; s < SMAX; ++s) {
        big = big || isum[s] >= (1 << 24);
        power[s] = (float)isum[s];
    }
    if(big) {  // the oracle's float chain
#pragma unroll
        for(int s = 0; s < SMAX; ++s) power[s] = 0.0f;
        for(int tt = 0; tt < ts; ++tt)
            for(int a = 0; a < nant; ++a) {
                const uint16_t d0 = *reinterpret_cast<const uint16_t*>(base + ((size_t)a * ld + tt) * 2);
                const uint16_t d1 = *reinterpret_cast<const uint16_t*>(base + ((size_t)(nant + a) * ld + tt) * 2);
                const float st = stokes<SW_STOKES_I>((float)(int8_t)d0, (float)(int8_t)(d0 >> 8),
                                                     (float)(int8_t)d1, (float)(int8_t)(d1 >> 8));
#pragma unroll
                for(int s = 0; s < SMAX; ++s)
                    if(s < nbeamsets) power[s] += st * antw[s * nant + a];
            }
    }
#pragma unroll
    for(int s = 0; s < SMAX; ++s) {
        if(s < nbeamsets) {
            const size_t o = ((size_t)s * nsamps_out + tp) * nchans + f;
            raw[o] = power[s];
            out[o] = clamp_i8(rescale<SW_STOKES_I>(power[s], offset[s * nchans + f], scale[s * nchans + f]));
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

// Incoherent beam from the dedisperser's time-major ring (see ib_tm_kernel):
// Stokes I, 0/1 antenna weights, d.nantennas <= 128.
int ib_beamform_tm(const void* d_ring, long ld, long w0, float* d_raw, void* d_out, const float* d_scale,
                   const float* d_offset, const float* d_ant_weights, const sw_bf_dims& d, hipStream_t st)
{
    SW_CHECK_ARG(d_ring && d_raw && d_out && d_scale && d_offset && d_ant_weights, "null buffer");
    SW_CHECK_ARG(d.stokes == SW_STOKES_I && 2 * d.nantennas <= kTmMaxCols, "time-major IB: Stokes I, <= 128 antennas");
    SW_CHECK_ARG(d.nbeamsets > 0 && d.nbeamsets <= 16 && d.nchans > 0 && d.nchans <= 65535, "bad dimensions");
    SW_CHECK_ARG(d.tscrunch > 0 && d.nsamples % d.tscrunch == 0, "nsamples must be a multiple of tscrunch");
    SW_CHECK_ARG((reinterpret_cast<uintptr_t>(d_ring) & 15) == 0 && ld % 8 == 0 && w0 % 8 == 0,
                 "ring must be 16-byte aligned with 8-sample aligned rows");
    const int nout = d.nsamples / d.tscrunch;
    const dim3 grid((nout + kTmThreads - 1) / kTmThreads, d.nchans);
    const bool v16 = d.tscrunch % 8 == 0;
    auto* in = static_cast<const uint8_t*>(d_ring);
    auto* out = static_cast<int8_t*>(d_out);
#define SW_IBT(SM)                                                                                         \
    if(v16)                                                                                                \
        hipLaunchKernelGGL((ib_tm_kernel<SM, true>), grid, dim3(kTmThreads), 0, st, in, ld, w0, d_raw, out, \
                           d_scale, d_offset, d_ant_weights, d.nchans, d.nsamples, d.nantennas, d.tscrunch, \
                           d.nbeamsets);                                                                   \
    else                                                                                                   \
        hipLaunchKernelGGL((ib_tm_kernel<SM, false>), grid, dim3(kTmThreads), 0, st, in, ld, w0, d_raw, out,\
                           d_scale, d_offset, d_ant_weights, d.nchans, d.nsamples, d.nantennas, d.tscrunch, \
                           d.nbeamsets);
    if(d.nbeamsets <= 1) {
        SW_IBT(1)
    } else if(d.nbeamsets <= 4) {
        SW_IBT(4)
    } else {
        SW_IBT(16)
    }
#undef SW_IBT
    SW_LAUNCH_CHECK();
    return SW_OK;
}

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
