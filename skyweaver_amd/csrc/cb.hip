// cb.hip -- coherent beamformer: complex int8 beam sums over antennas, Stokes
// detection, time scrunch, incoherent-beam subtraction and int8 requantisation.
//
// Reference: CoherentBeamformer::beamform / bf_ftpa_general_k,
// cpp/skyweaver/src/CoherentBeamformer.cu:13-269 (dp4a form); oracle
// cpp/skyweaver/test/src/CoherentBeamformerTester.cu:43-152.  Per (f, b, t, p):
//   V = sum_a W[f][b][a] * v[f][t][p][a]           (exact integers)
//   power(t') = sum_{t in window, ascending} Stokes(V_p0, V_p1)   (float)
//   out[t'][f][b] = clamp(rint((power - ib[s][t'][f]*127^2) / scale[s][f]))
//   (or rescale with the offset when IB subtraction is off)
//
// Two forms of the antenna sum share the float epilogue, whose operation
// order is the oracle's, so both are bit-identical to it:
//  * plain: one lane per beam, four v_dot4_i32_i8 accumulators
//    (re*re, im*im, re*im, im*re) per 4 antennas -- the reference's dp4a
//    arithmetic on VALU, one DM per launch;
//  * mfma: v_mfma_i32_32x32x32_i8 with time samples as rows and beams as
//    columns.  K is [re | im] of the antennas (K' = 2A):
//        Re = sum_a vr*wr + vi*(~wi) + sum_a vi      (-x == ~x + 1, no overflow)
//        Im = sum_a vr*wi + vi*wr
//    The per-row correction sum_a vi is itself an MFMA against a ones
//    operand and seeds the Re accumulator.  Weights are repacked once per
//    weight update into [F][B][2][2A] rows ([wr | ~wi], [wi | wr]).
#include "sw_common.hpp"

#include <type_traits>

namespace sw {
namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

__device__ __forceinline__ uint32_t re_bytes(uint32_t d0, uint32_t d1)
{
    return __builtin_amdgcn_perm(d1, d0, 0x06040200u); // d0.b0 d0.b2 d1.b0 d1.b2
}
__device__ __forceinline__ uint32_t im_bytes(uint32_t d0, uint32_t d1)
{
    return __builtin_amdgcn_perm(d1, d0, 0x07050301u); // d0.b1 d0.b3 d1.b1 d1.b3
}

// ------------------------------------------------------------ epilogue
// The FP32 division x / d as the compiler expands it (two v_div_scale, v_rcp,
// five FMAs, v_div_fmas, v_div_fixup) reduces, when its scale steps are
// identities, to five FMA-type steps on the divisor's refined reciprocal
// r1 = fma(fma(-d, r0, 1), r0, r0), r0 = rcp(d): bit for bit the same
// quotient.  The scale steps are identities for d in [2^-40, 2^60]
// (kFastDivLo, kFastDivHi) and quotients within [2^-100, 2^90]: every
// numerator here is an integer-valued power below 2^50, its difference with
// a float offset, or 0 (+-0 gives +0 where the expansion gives +-0: the same
// int8 after rint and the clamp).  The caller checks the divisors.
constexpr float kFastDivLo = 0x1p-40f, kFastDivHi = 0x1p60f;
__device__ __forceinline__ float div_recip(float d)
{
    const float r0 = __builtin_amdgcn_rcpf(d);
    return __fmaf_rn(__fmaf_rn(-d, r0, 1.0f), r0, r0);
}
__device__ __forceinline__ float div_by(float x, float d, float r1)
{
    const float q0 = x * r1;
    const float e0 = __fmaf_rn(-d, q0, x);
    const float q1 = __fmaf_rn(e0, r1, q0);
    const float e1 = __fmaf_rn(-d, q1, x);
    return __fmaf_rn(e1, r1, q1);
}

template <int S>
struct Epi {
    static constexpr int NV = nstokes(S);
    // Stokes of one sample from the two polarisations' complex beam sums.
    __device__ __forceinline__ static void detect(float r0, float i0, float r1, float i1,
                                                  float* st)
    {
        if constexpr(NV == 1) {
            st[0] = stokes<S>(r0, i0, r1, i1);
        } else {
            st[0] = stokes<SW_STOKES_I>(r0, i0, r1, i1);
            st[1] = stokes<SW_STOKES_Q>(r0, i0, r1, i1);
            st[2] = stokes<SW_STOKES_U>(r0, i0, r1, i1);
            st[3] = stokes<SW_STOKES_V>(r0, i0, r1, i1);
        }
    }
    // Requantise and store one output (t', f, b).
    __device__ __forceinline__ static void store(const float* power, int8_t* out,
                                                 const float* ib_raw, const float* scale,
                                                 const float* offset, const int* bmap,
                                                 int tout, int f, int b, int nchans,
                                                 int nbeams, int nsamps_out, int ib_sub)
    {
        const int s     = bmap[b];
        const int so    = s * nchans + f;
        const float sc  = scale[so];
        const size_t oi = ((size_t)tout * nchans + f) * nbeams + b;
        const size_t ii = ((size_t)s * nsamps_out + tout) * nchans + f;
        if constexpr(NV == 1) {
            float r;
            if(ib_sub)
                r = ib_subtract(power[0], ib_raw[ii], sc);
            else
                r = rescale<S>(power[0], offset[so], sc);
            out[oi] = clamp_i8(r);
        } else {
            const float4 ib = reinterpret_cast<const float4*>(ib_raw)[ii];
            float r[4];
            if(ib_sub) {
                r[0] = ib_subtract(power[0], ib.x, sc);
                r[1] = ib_subtract(power[1], ib.y, sc);
                r[2] = ib_subtract(power[2], ib.z, sc);
                r[3] = ib_subtract(power[3], ib.w, sc);
            } else {
                const float off = offset[so];
                r[0] = rescale<SW_STOKES_I>(power[0], off, sc);
                r[1] = rescale<SW_STOKES_Q>(power[1], off, sc);
                r[2] = rescale<SW_STOKES_U>(power[2], off, sc);
                r[3] = rescale<SW_STOKES_V>(power[3], off, sc);
            }
            char4 q;
            q.x = clamp_i8(r[0]);
            q.y = clamp_i8(r[1]);
            q.z = clamp_i8(r[2]);
            q.w = clamp_i8(r[3]);
            reinterpret_cast<char4*>(out)[oi] = q;
        }
    }
    // store_pre of a single-Stokes output through a buffer resource, at
    // byte offset vo + so; FD: the division by sc is div_by with r1
    template <bool FD>
    __device__ __forceinline__ static void store_pre_buf(const float* power, __amdgpu_buffer_rsrc_t rs, int vo,
                                                         int so, const float* ib, float sc, float r1, float off,
                                                         int ib_sub)
    {
        static_assert(NV == 1, "single Stokes outputs");
        float r;
        if constexpr(FD) {
            // the numerators of ib_subtract and rescale, in their operation order
            const float x = ib_sub ? power[0] - ib[0] * 16129.0f : (S == SW_STOKES_I ? power[0] - off : power[0]);
            r = rintf(div_by(x, sc, r1));
        } else {
            r = ib_sub ? ib_subtract(power[0], ib[0], sc) : rescale<S>(power[0], off, sc);
        }
        __builtin_amdgcn_raw_buffer_store_b8((uint8_t)clamp_i8(r), rs, vo, so, 0);
    }
    // As store(), with the beamset's scale / offset and the incoherent-beam
    // powers already loaded.
    __device__ __forceinline__ static void store_pre(const float* power, int8_t* out, const float* ib,
                                                     float sc, float off, int tout, int f, int b,
                                                     int nchans, int nbeams, int ib_sub)
    {
        const size_t oi = ((size_t)tout * nchans + f) * nbeams + b;
        if constexpr(NV == 1) {
            const float r = ib_sub ? ib_subtract(power[0], ib[0], sc) : rescale<S>(power[0], off, sc);
            out[oi] = clamp_i8(r);
        } else {
            float r[4];
            if(ib_sub) {
#pragma unroll
                for(int k = 0; k < 4; ++k) r[k] = ib_subtract(power[k], ib[k], sc);
            } else {
                r[0] = rescale<SW_STOKES_I>(power[0], off, sc);
                r[1] = rescale<SW_STOKES_Q>(power[1], off, sc);
                r[2] = rescale<SW_STOKES_U>(power[2], off, sc);
                r[3] = rescale<SW_STOKES_V>(power[3], off, sc);
            }
            char4 q;
            q.x = clamp_i8(r[0]);
            q.y = clamp_i8(r[1]);
            q.z = clamp_i8(r[2]);
            q.w = clamp_i8(r[3]);
            reinterpret_cast<char4*>(out)[oi] = q;
        }
    }
};

// ======================================================== plain (v_dot4) form
constexpr int kPlainThreads = 256;  // 4 waves x 64 beams
constexpr int kBeamLds = 1024;      // beams whose epilogue parameters the MFMA form stages in LDS
constexpr int kPlainTB      = 64;   // samples staged per LDS pass

template <int S, int A4>
__global__ void __launch_bounds__(kPlainThreads)
cb_plain_kernel(const uint8_t* __restrict__ ftpa, const uint8_t* __restrict__ weights,
                int8_t* __restrict__ out, const float* __restrict__ scale,
                const float* __restrict__ offset, const int* __restrict__ bmap,
                const float* __restrict__ ib_raw, int nchans, int nsamples, int nbeams,
                int tscrunch, int tb_per_block, int ib_sub)
{
    constexpr int NV = nstokes(S);
    constexpr int A  = A4 * 4;
    // LDS: [kPlainTB][2][A4] of (re x4, im x4)
    __shared__ uint2 vt[kPlainTB * 2 * A4];
    const int f   = blockIdx.y;
    const int tid = threadIdx.x;
    const int b   = blockIdx.z * kPlainThreads + tid;
    const bool active = b < nbeams;
    // beam weights in registers, transposed to (re x4, im x4) per 4 antennas
    uint32_t wr[A4], wi[A4];
    {
        const uint2* wrow =
            reinterpret_cast<const uint2*>(weights + ((size_t)f * nbeams + (active ? b : 0)) * A * 2);
#pragma unroll
        for(int g = 0; g < A4; ++g) {
            const uint2 q = wrow[g];
            wr[g] = active ? re_bytes(q.x, q.y) : 0u;
            wi[g] = active ? im_bytes(q.x, q.y) : 0u;
        }
    }
    const int nsamps_out = nsamples / tscrunch;
    const int t_begin    = blockIdx.x * tb_per_block;
    const int t_end      = min(nsamples, t_begin + tb_per_block);
    float power[NV];
#pragma unroll
    for(int k = 0; k < NV; ++k) power[k] = 0.0f;
    for(int tc = t_begin; tc < t_end; tc += kPlainTB) {
        const int nt = min(kPlainTB, t_end - tc);
        __syncthreads();
        // stage nt samples x 2 pols x A4 groups: 8 bytes each
        const uint2* src = reinterpret_cast<const uint2*>(ftpa + ((size_t)f * nsamples + tc) * 4 * A);
        for(int i = tid; i < nt * 2 * A4; i += kPlainThreads) {
            const uint2 q = src[i];
            vt[i] = make_uint2(re_bytes(q.x, q.y), im_bytes(q.x, q.y));
        }
        __syncthreads();
        for(int t = 0; t < nt; ++t) {
            float re[2], im[2];
#pragma unroll
            for(int p = 0; p < 2; ++p) {
                int xx = 0, yy = 0, xy = 0, yx = 0;
                const uint2* v = vt + (t * 2 + p) * A4;
#pragma unroll
                for(int g = 0; g < A4; ++g) {
                    const uint2 q = v[g];   // broadcast read
                    xx = __builtin_amdgcn_sdot4((int)wr[g], (int)q.x, xx, false);
                    yy = __builtin_amdgcn_sdot4((int)wi[g], (int)q.y, yy, false);
                    xy = __builtin_amdgcn_sdot4((int)wr[g], (int)q.y, xy, false);
                    yx = __builtin_amdgcn_sdot4((int)wi[g], (int)q.x, yx, false);
                }
                re[p] = (float)xx - (float)yy;
                im[p] = (float)xy + (float)yx;
            }
            float st[NV];
            Epi<S>::detect(re[0], im[0], re[1], im[1], st);
#pragma unroll
            for(int k = 0; k < NV; ++k) power[k] += st[k];
            const int tg = tc + t;
            if((tg + 1) % tscrunch == 0) {
                if(active)
                    Epi<S>::store(power, out, ib_raw, scale, offset, bmap, tg / tscrunch, f, b,
                                  nchans, nbeams, nsamps_out, ib_sub);
#pragma unroll
                for(int k = 0; k < NV; ++k) power[k] = 0.0f;
            }
        }
    }
}

// ========================================================== MFMA form
constexpr int kMfmaWaves = 4;  // default waves per workgroup, 32 time rows each (WV)

// Packed weights: for each (f, b) two rows of 2A bytes:
//   row 0 (real part):      [ wr[0..A) | ~wi[0..A) ]
//   row 1 (imaginary part): [ wi[0..A) |  wr[0..A) ]
__global__ void cb_pack_kernel(const uint8_t* __restrict__ w, uint8_t* __restrict__ packed,
                               int nant, size_t nrows)
{
    // one thread per 4 antennas of one (f, b) row
    const size_t i   = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int a4     = nant / 4;
    if(i >= nrows * a4) return;
    const size_t row = i / a4;
    const int g      = (int)(i % a4);
    const uint2 q    = reinterpret_cast<const uint2*>(w + row * nant * 2)[g];
    const uint32_t r = re_bytes(q.x, q.y), im = im_bytes(q.x, q.y);
    uint32_t* o = reinterpret_cast<uint32_t*>(packed + row * nant * 4);
    const int qa = nant / 4;  // dwords per A bytes
    o[g]          = r;
    o[qa + g]     = ~im;
    o[2 * qa + g] = im;
    o[3 * qa + g] = r;
}

// Per-output epilogue parameters, loaded before the beam tile's MFMAs so
// that their latency hides behind them: beamset scale / offset of the lane's
// beam and (for few outputs per lane) the incoherent-beam powers.
template <int S, int NOUT>
struct EpiParams {
    static constexpr int NV = nstokes(S);
    static constexpr bool kPreIb = NOUT * NV <= 4;
    float sc = 0, off = 0, r1 = 0;
    float ib[kPreIb ? NOUT : 1][NV];
    int s = 0;
};

// The incoherent beam computed by the coherent beamformer itself from the
// time rows it has staged (time-major input, Stokes I, 0/1 antenna weights,
// tscrunch 4 ... 32): the integer sums of ib_tm_kernel, with its float-chain
// fallback at 2^24, so the results are identical and the ring is read once.
struct FusedIb {
    float* raw = nullptr;  // [S][T/ts][F]; null: not fused (ib_raw is an input)
    int8_t* out = nullptr;
    const float* scale = nullptr;   // [S][F]
    const float* offset = nullptr;  // [S][F]
    const float* antw = nullptr;    // [S][A], 0 or 1
};

// TMI: the voltages come time-major from the coherent dedisperser's ring
// (sample t of column c = p A + a of channel f at ftpa + ((f 2A + c) ld + w0
// + t) 2) and are corner-turned through LDS into the FTPA rows the operands
// are read from; else ftpa is FTPA [F][T][2][A].
//
// Stokes I at tscrunch 8 with 64 antennas and 4 waves takes 176 VGPRs
// unconstrained, two waves per SIMD; asked for three (168 VGPRs, two values
// spilled) it runs three workgroups per CU, which its LDS allows.  Other
// shapes keep the compiler's allocation (three waves would spill them far
// more).
template <int S, int TS, int NSTEPS, int WV>
constexpr int cb_min_waves()
{
    return S == SW_STOKES_I && TS == 8 && NSTEPS == 4 && WV == 4 ? 3 : 1;
}
//
// FD: every scale lies in [kFastDivLo, kFastDivHi] (checked by the caller),
// so the requantisation divides with div_by.
template <int S, int TS, int NSTEPS, bool TMI, int WV, bool FD = false>
__global__ void __launch_bounds__(WV * 64, (cb_min_waves<S, TS, NSTEPS, WV>()))
cb_mfma_kernel(const uint8_t* __restrict__ ftpa, const uint8_t* __restrict__ packed,
               int8_t* __restrict__ out, const float* __restrict__ scale,
               const float* __restrict__ offset, const int* __restrict__ bmap,
               const float* __restrict__ ib_raw, int nchans, int nsamples, int nbeams,
               int ntchunks, int ib_sub, int nbeamsets, long ld, long w0, FusedIb fib)
{
    constexpr int NV = nstokes(S);
    constexpr int A  = NSTEPS * 16;  // antennas (K' = 2A = 32 * NSTEPS)
    constexpr int KP = 2 * A;
    // weight tile: 32 beams x 2 rows x KP bytes, rows padded by 16 bytes so
    // that the 32 lanes of a half-wave read distinct banks
    constexpr int WPITCH = 2 * KP + 16;          // bytes per beam in LDS
    constexpr int WTILE  = 32 * 2 * KP / 16;     // uint4 per tile in global memory
    constexpr int WPER   = (WTILE + WV * 64 - 1) / (WV * 64);
    constexpr int NOUT   = TS <= 16 ? 16 / TS : 1;  // outputs per lane per tile
    // LDS: the weight tiles (double-buffered), or first, with TMI, the
    // workgroup's 32 WV time rows of both polarisations (rows of
    // 4A bytes padded by 16)
    constexpr int WL16   = 32 * WPITCH / 16;
    constexpr int SPITCH = 4 * A + 16;
    constexpr int ST16   = TMI ? 32 * WV * SPITCH / 16 : 0;
    __shared__ uint4 smem[2 * WL16 > ST16 ? 2 * WL16 : ST16];
    uint4 (*wlds)[WL16] = reinterpret_cast<uint4 (*)[WL16]>(smem);
    // epilogue parameters of the channel's beams and the workgroup's
    // incoherent-beam powers, staged once (up to kBeamLds beams)
    constexpr bool kPre = EpiParams<S, NOUT>::kPreIb;
    constexpr int NO_WG = 32 * WV / (TS <= 32 ? TS : 32);  // output samples of the workgroup
    __shared__ int tb_s[kPre ? kBeamLds : 1];
    __shared__ float tb_sc[kPre ? kBeamLds : 1], tb_off[kPre ? kBeamLds : 1];
    __shared__ float tb_r1[kPre && FD ? kBeamLds : 1];  // div_recip of the scales
    __shared__ float tb_ib[kPre ? 16 * NO_WG * NV : 1];
    // XCD-aware mapping: consecutive workgroups go round-robin over the 8
    // XCDs; keep all time chunks of a channel on one XCD so its packed
    // weights stay resident in that XCD's L2.
    int f, tc;
    {
        const int bid = blockIdx.x;
        if((nchans & 7) == 0) {
            const int xcd = bid & 7, local = bid >> 3;
            f  = xcd + 8 * (local / ntchunks);
            tc = local % ntchunks;
        } else {
            f  = bid / ntchunks;
            tc = bid % ntchunks;
        }
    }
    const int tid  = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int t0   = (tc * WV + wave) * 32;  // first time row of this wave
    // waves past the end of the gulp still take part in the weight staging
    // and its barriers, but compute on row 0 and store nothing
    const bool active = t0 < nsamples;
    const int tl   = active ? t0 : 0;
    const int row = lane & 31;  // A-operand row / output column index
    const int grp = lane >> 5;  // k group (16 bytes)
    // A-operand row i = 8a + 4g + b carries time sample 16g + 4a + b, so that
    // accumulator register r of a lane in half g holds time 16g + r: each
    // lane sees 16 consecutive samples of its beam and the time scrunch needs
    // no exchange between lanes.
    const int trow = 16 * ((row >> 2) & 1) + 4 * (row >> 3) + (row & 3);

    const int nsamps_out = nsamples / TS;
    const bool tabs = kPre && nbeams <= kBeamLds && nbeamsets <= 16;  // uniform
    // TMI: the workgroup's time rows, 16-byte runs of 8 samples of each
    // column, issued first so that their latency overlaps the table loads
    constexpr int NST = TMI ? 2 * A * 4 * WV / (WV * 64) : 1;  // runs per thread
    uint4 run[NST];
    if constexpr(TMI) {
        const int t0w = tc * WV * 32;
#pragma unroll
        for(int k = 0; k < NST; ++k) {
            const int i = tid + k * WV * 64;
            const int c = i % (2 * A), q = i / (2 * A);
            run[k] = make_uint4(0, 0, 0, 0);
            if(t0w + 8 * q < nsamples)
                run[k] = *reinterpret_cast<const uint4*>(ftpa + (((size_t)f * 2 * A + c) * ld + w0 + t0w + 8 * q) * 2);
        }
    }
    if constexpr(kPre) {
        if(tabs) {
            // all of a thread's loads in flight together (beam map first,
            // then the scales and offsets it selects)
            constexpr int NB = kBeamLds / (WV * 64);
            int sbv[NB];
#pragma unroll
            for(int k = 0; k < NB; ++k) {
                const int b = tid + k * WV * 64;
                sbv[k] = b < nbeams ? bmap[b] : 0;
            }
            float scv[NB], ofv[NB];
#pragma unroll
            for(int k = 0; k < NB; ++k) {
                const int b = tid + k * WV * 64;
                scv[k] = b < nbeams ? scale[sbv[k] * nchans + f] : 0.0f;
                ofv[k] = (b < nbeams && !ib_sub) ? offset[sbv[k] * nchans + f] : 0.0f;
            }
#pragma unroll
            for(int k = 0; k < NB; ++k) {
                const int b = tid + k * WV * 64;
                if(b < nbeams) {
                    tb_s[b]   = sbv[k];
                    tb_sc[b]  = scv[k];
                    tb_off[b] = ofv[k];
                    if constexpr(FD) tb_r1[b] = div_recip(scv[k]);
                }
            }
            const int o0 = tc * WV * 32 / TS;
            for(int i = tid; !(TMI && fib.raw) && i < nbeamsets * NO_WG * NV; i += WV * 64) {
                const int k = i % NV, o = (i / NV) % NO_WG, sb = i / (NV * NO_WG);
                float v = 0.0f;
                if(ib_sub && o0 + o < nsamps_out) v = ib_raw[(((size_t)sb * nsamps_out + o0 + o) * nchans + f) * NV + k];
                tb_ib[i] = v;
            }
        }
    }
    // fused incoherent beam: lane l reads 16-byte runs (words 4 g .. 4 g + 3,
    // g = l mod LPR, word w = columns 2w, 2w + 1: one pol, antennas 2w mod A
    // and the next) of rows l / LPR + RPI i; its antenna weights (0/1) as bit
    // masks, bit 2 sb (+1) of ibm[j] = word 4 g + j's first (second) antenna
    // in beamset sb, loaded while the rows are staged
    constexpr int LPR = A / 4 > 0 ? A / 4 : 1;  // lanes per row
    constexpr int RPI = 64 / LPR;               // rows per read
    unsigned ibm[4] = {0, 0, 0, 0};
    if constexpr(TMI && S == SW_STOKES_I && kPre && 256 % A == 0 && TS >= RPI) {
        if(fib.raw) {
#pragma unroll
            for(int j = 0; j < 4; ++j) {
                const int a = (2 * (4 * (lane % LPR) + j)) % A;
                for(int sb = 0; sb < nbeamsets; ++sb)
                    ibm[j] |= (fib.antw[sb * A + a] != 0.0f ? 1u : 0u) << (2 * sb) |
                              (fib.antw[sb * A + a + 1] != 0.0f ? 2u : 0u) << (2 * sb);
            }
        }
    }
    // ---- A operands: 32 time rows x K' for both polarisations
    if constexpr(TMI) {
        // corner turn of the runs loaded above into rows
        uint8_t* stg = reinterpret_cast<uint8_t*>(smem);
#pragma unroll
        for(int k = 0; k < NST; ++k) {
            const int i = tid + k * WV * 64;
            const int c = i % (2 * A), q = i / (2 * A);
            const uint32_t w[4] = {run[k].x, run[k].y, run[k].z, run[k].w};
#pragma unroll
            for(int j = 0; j < 8; ++j)
                *reinterpret_cast<uint16_t*>(stg + (8 * q + j) * SPITCH + c * 2) = (uint16_t)(w[j >> 1] >> (16 * (j & 1)));
        }
        __syncthreads();
        if constexpr(S == SW_STOKES_I && kPre && 256 % A == 0 && TS >= RPI) {
            if(fib.raw) {  // uniform: the incoherent beam from the staged rows
                const int o0 = tc * NO_WG;
                const int r0 = lane / LPR, g = lane % LPR;
                // one beamset over every antenna (uniform): a word's two
                // antennas are summed by one dot product, no masks
                const bool all1 =
                    nbeamsets == 1 && __all(((ibm[0] & ibm[1] & ibm[2] & ibm[3]) & 3u) == 3u);
#pragma unroll
                for(int oi = 0; oi < NO_WG / WV; ++oi) {
                    const int o = wave * (NO_WG / WV) + oi;  // this wave's own rows
                    int lo[4] = {0, 0, 0, 0}, hi[4] = {0, 0, 0, 0}, tot = 0;
#pragma unroll
                    for(int i = 0; i < TS / RPI; ++i) {
                        const uint4 v = *reinterpret_cast<const uint4*>(stg + (o * TS + RPI * i + r0) * SPITCH + 16 * g);
                        const int w[4] = {(int)v.x, (int)v.y, (int)v.z, (int)v.w};
                        if(all1) {
#pragma unroll
                            for(int j = 0; j < 4; ++j) tot = __builtin_amdgcn_sdot4(w[j], w[j], tot, false);
                        } else {
#pragma unroll
                            for(int j = 0; j < 4; ++j) {
                                lo[j] = __builtin_amdgcn_sdot4(w[j], w[j] & 0xFFFF, lo[j], false);
                                hi[j] = __builtin_amdgcn_sdot4(w[j], (int)((unsigned)w[j] & 0xFFFF0000u), hi[j], false);
                            }
                        }
                    }
                    for(int sb = 0; sb < nbeamsets; ++sb) {
                        int part = tot;
                        if(!all1) {
#pragma unroll
                            for(int j = 0; j < 4; ++j) {
                                if(ibm[j] >> (2 * sb) & 1u) part += lo[j];
                                if(ibm[j] >> (2 * sb) & 2u) part += hi[j];
                            }
                        }
#pragma unroll
                        for(int m = 32; m > 0; m >>= 1) part += __shfl_xor(part, m, 64);
                        float power = (float)part;  // the float chain's value while below 2^24
                        if(part >= (1 << 24)) {     // the oracle's float chain: time, then antenna ascending
                            power = 0.0f;
                            for(int tt = 0; tt < TS; ++tt) {
                                const uint8_t* r = stg + (o * TS + tt) * SPITCH;
                                for(int a = 0; a < A; ++a) {
                                    const uint16_t d0 = *reinterpret_cast<const uint16_t*>(r + a * 2);
                                    const uint16_t d1 = *reinterpret_cast<const uint16_t*>(r + (A + a) * 2);
                                    const float st =
                                        stokes<SW_STOKES_I>((float)(int8_t)d0, (float)(int8_t)(d0 >> 8),
                                                            (float)(int8_t)d1, (float)(int8_t)(d1 >> 8));
                                    power += st * fib.antw[sb * A + a];
                                }
                            }
                        }
                        if(lane == 0) {
                            if(o0 + o < nsamps_out) {
                                const size_t oi2 = ((size_t)sb * nsamps_out + o0 + o) * nchans + f;
                                fib.raw[oi2] = power;
                                fib.out[oi2] = clamp_i8(rescale<SW_STOKES_I>(power, fib.offset[sb * nchans + f],
                                                                             fib.scale[sb * nchans + f]));
                            }
                            if(tabs) tb_ib[sb * NO_WG + o] = ib_sub ? power : 0.0f;
                        }
                    }
                }
            }
        }
    }
    v4i va[2][NSTEPS];
#pragma unroll
    for(int p = 0; p < 2; ++p) {
        const uint8_t* vrow = TMI ? reinterpret_cast<const uint8_t*>(smem) + (wave * 32 + trow) * SPITCH + p * A * 2
                                  : ftpa + (((size_t)f * nsamples + tl + trow) * 2 + p) * A * 2;
#pragma unroll
        for(int s = 0; s < NSTEPS; ++s) {
            const int k  = 32 * s + 16 * grp;
            const bool imag = k >= A;
            const int a0 = imag ? k - A : k;
            const uint4 lo = *reinterpret_cast<const uint4*>(vrow + a0 * 2);
            const uint4 hi = *reinterpret_cast<const uint4*>(vrow + a0 * 2 + 16);
            v4i x;
            if(!imag) {
                x[0] = re_bytes(lo.x, lo.y); x[1] = re_bytes(lo.z, lo.w);
                x[2] = re_bytes(hi.x, hi.y); x[3] = re_bytes(hi.z, hi.w);
            } else {
                x[0] = im_bytes(lo.x, lo.y); x[1] = im_bytes(lo.z, lo.w);
                x[2] = im_bytes(hi.x, hi.y); x[3] = im_bytes(hi.z, hi.w);
            }
            va[p][s] = x;
        }
    }
    // ---- per-row correction sum_a vi (seed of the real accumulators)
    v16i corr[2];
#pragma unroll
    for(int p = 0; p < 2; ++p) {
        v16i c = {0};
#pragma unroll
        for(int s = 0; s < NSTEPS; ++s) {
            const int k = 32 * s + 16 * grp;
            const int one = (k >= A) ? 0x01010101 : 0;
            const v4i ones = {one, one, one, one};
            c = __builtin_amdgcn_mfma_i32_32x32x32_i8(va[p][s], ones, c, 0, 0, 0);
        }
        corr[p] = c;
    }

    const int nbt = (nbeams + 31) / 32;
    const uint4* wsrc = reinterpret_cast<const uint4*>(packed + (size_t)f * nbeams * 2 * KP);
    const size_t wtotal = (size_t)nbeams * 2 * KP / 16;  // uint4 of this channel
    const int nfull = (int)(wtotal / WTILE);             // whole tiles
    // the channel's weights as a buffer: a whole tile's loads take the
    // lane's fixed offset and the tile's scalar offset bt WTILE 16, which is
    // at most (nfull - 1) WTILE 16, so every load lies in [0, 16 wtotal)
    const auto wrs = buffer_rsrc(wsrc, (uint32_t)(wtotal * 16));
    // global -> registers for tile bt (beams past the end read as zero)
    auto fetch = [&](int bt, uint4* r) {
        if(bt < nfull) {  // a whole tile (uniform)
#pragma unroll
            for(int i = 0; i < WPER; ++i) {
                const int idx = tid + i * WV * 64;
                if(WTILE % (WV * 64) == 0 || idx < WTILE) {
                    const auto v = __builtin_amdgcn_raw_buffer_load_b128(wrs, idx * 16, bt * WTILE * 16, 0);
                    r[i] = make_uint4(v[0], v[1], v[2], v[3]);
                } else {
                    r[i] = make_uint4(0, 0, 0, 0);
                }
            }
        } else {
#pragma unroll
            for(int i = 0; i < WPER; ++i) {
                const int idx = tid + i * WV * 64;
                const size_t g = (size_t)bt * WTILE + idx;
                r[i] = (idx < WTILE && g < wtotal) ? wsrc[g] : make_uint4(0, 0, 0, 0);
            }
        }
    };
    auto stash = [&](int buf, const uint4* r) {
#pragma unroll
        for(int i = 0; i < WPER; ++i) {
            const int idx = tid + i * WV * 64;
            if(idx < WTILE) {
                const int beam = idx / (2 * KP / 16), c16 = idx % (2 * KP / 16);
                wlds[buf][beam * (WPITCH / 16) + c16] = r[i];
            }
        }
    };
    // single-Stokes outputs (TS <= 16, parameters preloaded) as a buffer
    // over the workgroup's output rows [tw0, tw0 + NO_WG) of channel f:
    // output (tout, f, b) at byte (tout - tw0) F B + row (the lane's fixed
    // offset) + 32 bt (the tile's scalar offset); b < nbeams keeps every
    // store below (NO_WG - 1) F B + nbeams, the range given
    constexpr bool kBufOut = NV == 1 && TS <= 16 && EpiParams<S, NOUT>::kPreIb;
    const int tw0 = tc * NO_WG;
    const auto ors = buffer_rsrc(out + ((size_t)tw0 * nchans + f) * nbeams,
                                 (uint32_t)((size_t)(NO_WG - 1) * nchans * nbeams + nbeams));
    int ovo[kBufOut ? NOUT : 1];
#pragma unroll
    for(int o = 0; o < (kBufOut ? NOUT : 1); ++o)
        ovo[o] = ((t0 + 16 * grp + o * TS) / TS - tw0) * nchans * nbeams + row;
    uint4 wreg[WPER];
    fetch(0, wreg);
    if constexpr(TMI) __syncthreads();  // every wave has its operands: the rows' LDS becomes the weight tiles
    stash(0, wreg);
    __syncthreads();

    for(int bt = 0; bt < nbt; ++bt) {
        const int buf = bt & 1;
        if(bt + 1 < nbt) fetch(bt + 1, wreg);  // in flight during this tile
        const int b = bt * 32 + row;
        const bool bvalid = b < nbeams;
        // epilogue parameters of this lane's beam
        EpiParams<S, NOUT> ep;
        if(tabs) {
            if constexpr(kPre) {
                if(active && bvalid) {
                    ep.s = tb_s[b];
                    ep.sc = tb_sc[b];
                    if constexpr(FD) ep.r1 = tb_r1[b];
                    ep.off = tb_off[b];
#pragma unroll
                    for(int o = 0; o < NOUT; ++o) {
                        const int ol = TS <= 16 ? (32 * wave + 16 * grp + o * TS) / TS : wave;
#pragma unroll
                        for(int k = 0; k < NV; ++k) ep.ib[o][k] = tb_ib[(ep.s * NO_WG + ol) * NV + k];
                    }
                }
            }
        } else if(active && bvalid) {
            ep.s = bmap[b];
            const int so = ep.s * nchans + f;
            ep.sc = scale[so];
            if constexpr(FD) ep.r1 = div_recip(ep.sc);
            if(!ib_sub) ep.off = offset[so];
            if constexpr(EpiParams<S, NOUT>::kPreIb) {
                if(ib_sub) {
#pragma unroll
                    for(int o = 0; o < NOUT; ++o) {
                        const int tout = TS <= 16 ? (t0 + 16 * grp + o * TS) / TS : t0 / 32;
                        const size_t ii = ((size_t)ep.s * nsamps_out + tout) * nchans + f;
                        if constexpr(NV == 1) {
                            ep.ib[o][0] = ib_raw[ii];
                        } else {
                            const float4 v = reinterpret_cast<const float4*>(ib_raw)[ii];
                            ep.ib[o][0] = v.x;
                            ep.ib[o][1] = v.y;
                            ep.ib[o][2] = v.z;
                            ep.ib[o][3] = v.w;
                        }
                    }
                }
            }
        }
        const uint8_t* wl = reinterpret_cast<const uint8_t*>(wlds[buf]) + row * WPITCH + 16 * grp;
        v16i acc_re[2] = {corr[0], corr[1]};
        v16i acc_im[2] = {v16i{0}, v16i{0}};
#pragma unroll
        for(int s = 0; s < NSTEPS; ++s) {
            const v4i wre = *reinterpret_cast<const v4i*>(wl + 32 * s);
            const v4i wim = *reinterpret_cast<const v4i*>(wl + KP + 32 * s);
#pragma unroll
            for(int p = 0; p < 2; ++p) {
                acc_re[p] = __builtin_amdgcn_mfma_i32_32x32x32_i8(va[p][s], wre, acc_re[p], 0, 0, 0);
                acc_im[p] = __builtin_amdgcn_mfma_i32_32x32x32_i8(va[p][s], wim, acc_im[p], 0, 0, 0);
            }
        }
        // ---- detect: register r of this lane holds time sample 16*grp + r
        //      of beam `b`
        float st[16][NV];
#pragma unroll
        for(int r = 0; r < 16; ++r)
            Epi<S>::detect((float)acc_re[0][r], (float)acc_im[0][r], (float)acc_re[1][r],
                           (float)acc_im[1][r], st[r]);
        if constexpr(TS <= 16) {
#pragma unroll
            for(int o = 0; o < 16 / TS; ++o) {
                float power[NV];
#pragma unroll
                for(int k = 0; k < NV; ++k) power[k] = 0.0f;
#pragma unroll
                for(int q = 0; q < TS; ++q)
#pragma unroll
                    for(int k = 0; k < NV; ++k) power[k] += st[o * TS + q][k];
                if(active && bvalid) {
                    const int tout = (t0 + 16 * grp + o * TS) / TS;
                    if constexpr(kBufOut)
                        Epi<S>::template store_pre_buf<FD>(power, ors, ovo[o], 32 * bt, ep.ib[o], ep.sc, ep.r1,
                                                           ep.off, ib_sub);
                    else if constexpr(EpiParams<S, NOUT>::kPreIb)
                        Epi<S>::store_pre(power, out, ep.ib[o], ep.sc, ep.off, tout, f, b, nchans, nbeams,
                                          ib_sub);
                    else
                        Epi<S>::store(power, out, ib_raw, scale, offset, bmap, tout, f, b, nchans, nbeams,
                                      nsamps_out, ib_sub);
                }
            }
        } else {  // TS == 32: the upper half continues the lower half's sum
            float power[NV];
#pragma unroll
            for(int k = 0; k < NV; ++k) power[k] = 0.0f;
#pragma unroll
            for(int q = 0; q < 16; ++q)
#pragma unroll
                for(int k = 0; k < NV; ++k) power[k] += st[q][k];
            float carried[NV];
#pragma unroll
            for(int k = 0; k < NV; ++k) carried[k] = __shfl_xor(power[k], 32);
            if(grp) {
#pragma unroll
                for(int k = 0; k < NV; ++k) power[k] = carried[k];
#pragma unroll
                for(int q = 0; q < 16; ++q)
#pragma unroll
                    for(int k = 0; k < NV; ++k) power[k] += st[q][k];
                if(active && bvalid)
                    Epi<S>::store_pre(power, out, ep.ib[0], ep.sc, ep.off, t0 / 32, f, b, nchans, nbeams,
                                      ib_sub);
            }
        }
        if(bt + 1 < nbt) stash(buf ^ 1, wreg);
        __syncthreads();
    }
}

// ------------------------------------------------------------ dispatch
template <int S>
int launch_plain(const uint8_t* ftpa, const uint8_t* w, int8_t* out, const float* sc,
                 const float* off, const int* bmap, const float* ib, const sw_bf_dims& d,
                 hipStream_t st)
{
    const int tb_per_block = max(1024, d.tscrunch * ((1024 + d.tscrunch - 1) / d.tscrunch));
    dim3 grid((d.nsamples + tb_per_block - 1) / tb_per_block, d.nchans,
              (d.nbeams + kPlainThreads - 1) / kPlainThreads);
#define SW_PLAIN(A4V)                                                                      \
    case A4V:                                                                              \
        hipLaunchKernelGGL((cb_plain_kernel<S, A4V>), grid, dim3(kPlainThreads), 0, st,    \
                           ftpa, w, out, sc, off, bmap, ib, d.nchans, d.nsamples,          \
                           d.nbeams, d.tscrunch, tb_per_block, d.ib_subtract);             \
        break;
    switch(d.nantennas / 4) {
        SW_PLAIN(1) SW_PLAIN(2) SW_PLAIN(3) SW_PLAIN(4) SW_PLAIN(5) SW_PLAIN(6) SW_PLAIN(7)
        SW_PLAIN(8) SW_PLAIN(10) SW_PLAIN(12) SW_PLAIN(14) SW_PLAIN(16) SW_PLAIN(20)
        SW_PLAIN(24) SW_PLAIN(28) SW_PLAIN(32)
    default:
        set_error("sw_cb_beamform: plain form supports A in {4..32, 40, 48, 56, 64, 80, 96, 112, 128}");
        return SW_ERR_UNSUPPORTED;
    }
#undef SW_PLAIN
    SW_LAUNCH_CHECK();
    return SW_OK;
}

// LDS bytes of cb_mfma_kernel<S, TS, NSTEPS, TMI, WV> (its __shared__ arrays)
template <int S, int TS, int NSTEPS, bool TMI, int WV>
constexpr size_t cb_mfma_lds_bytes()
{
    constexpr int A = NSTEPS * 16, KP = 2 * A, NV = nstokes(S);
    constexpr int WL16 = 32 * (2 * KP + 16) / 16;
    constexpr int ST16 = TMI ? 32 * WV * (4 * A + 16) / 16 : 0;
    constexpr int NOUT = TS <= 16 ? 16 / TS : 1;
    constexpr bool kPre = NOUT * NV <= 4;
    constexpr int NO_WG = 32 * WV / (TS <= 32 ? TS : 32);
    return (size_t)16 * (2 * WL16 > ST16 ? 2 * WL16 : ST16) + (kPre ? (size_t)kBeamLds * 16 : 16) +
           (kPre ? (size_t)16 * NO_WG * NV * 4 : 4);
}

// workgroup shapes (waves of 32 time rows) instantiated per form: the
// pipeline's forms (Stokes I, on the ring or on FTPA, and full Stokes on
// FTPA) take 2, 4 or 8 waves, the others 4
template <int S, bool TMI>
constexpr bool cb_wave_variants()
{
    return S == SW_STOKES_I || (S == SW_STOKES_IQUV && !TMI);
}

template <int S, int TS, int NS, bool TMI, int WV, bool FD = false>
int launch_mfma_wv(const uint8_t* ftpa, const uint8_t* packed, int8_t* out, const float* sc, const float* off,
                   const int* bmap, const float* ib, const sw_bf_dims& d, hipStream_t st, long ld, long w0,
                   const FusedIb& fib)
{
    if constexpr(cb_mfma_lds_bytes<S, TS, NS, TMI, WV>() <= 163840) {
        const int ntchunks = (d.nsamples + 32 * WV - 1) / (32 * WV);
        hipLaunchKernelGGL((cb_mfma_kernel<S, TS, NS, TMI, WV, FD>), dim3(ntchunks * d.nchans), dim3(WV * 64), 0, st,
                           ftpa, packed, out, sc, off, bmap, ib, d.nchans, d.nsamples, d.nbeams, ntchunks,
                           d.ib_subtract, d.nbeamsets, ld, w0, fib);
        SW_LAUNCH_CHECK();
        return SW_OK;
    } else {
        set_error("sw_cb_beamform: this workgroup shape does not fit the LDS at this antenna count");
        return SW_ERR_UNSUPPORTED;
    }
}

template <int S, int TS, bool TMI>
int launch_mfma_ts(const uint8_t* ftpa, const uint8_t* packed, int8_t* out, const float* sc,
                   const float* off, const int* bmap, const float* ib, const sw_bf_dims& d,
                   hipStream_t st, long ld, long w0, const FusedIb& fib, int waves)
{
#define SW_MFMA_WV(NS, WV) return launch_mfma_wv<S, TS, NS, TMI, WV>(ftpa, packed, out, sc, off, bmap, ib, d, st, ld, w0, fib)
#define SW_MFMA(NS)                                                   \
    case NS:                                                          \
        if constexpr(cb_wave_variants<S, TMI>()) {                    \
            if(waves == 2) SW_MFMA_WV(NS, 2);                         \
            if(waves == 8) SW_MFMA_WV(NS, 8);                         \
        }                                                             \
        SW_MFMA_WV(NS, kMfmaWaves);
    if(waves != 2 && waves != 4 && waves != 8) {
        set_error("sw_cb_beamform: waves per workgroup must be 2, 4 or 8");
        return SW_ERR_INVALID;
    }
    if(waves != kMfmaWaves && !cb_wave_variants<S, TMI>()) {
        set_error("sw_cb_beamform: this Stokes mode runs with 4 waves per workgroup only");
        return SW_ERR_UNSUPPORTED;
    }
    switch(d.nantennas / 16) {
        SW_MFMA(1) SW_MFMA(2) SW_MFMA(3) SW_MFMA(4) SW_MFMA(5) SW_MFMA(6) SW_MFMA(7) SW_MFMA(8)
    default:
        set_error("sw_cb_beamform: MFMA form supports A in {16, 32, ..., 128}");
        return SW_ERR_UNSUPPORTED;
    }
#undef SW_MFMA
#undef SW_MFMA_WV
}

template <int S, bool TMI>
int launch_mfma(const uint8_t* ftpa, const uint8_t* packed, int8_t* out, const float* sc,
                const float* off, const int* bmap, const float* ib, const sw_bf_dims& d,
                hipStream_t st, long ld = 0, long w0 = 0, const FusedIb& fib = FusedIb{}, int waves = kMfmaWaves)
{
    switch(d.tscrunch) {
    case 1: return launch_mfma_ts<S, 1, TMI>(ftpa, packed, out, sc, off, bmap, ib, d, st, ld, w0, fib, waves);
    case 2: return launch_mfma_ts<S, 2, TMI>(ftpa, packed, out, sc, off, bmap, ib, d, st, ld, w0, fib, waves);
    case 4: return launch_mfma_ts<S, 4, TMI>(ftpa, packed, out, sc, off, bmap, ib, d, st, ld, w0, fib, waves);
    case 8: return launch_mfma_ts<S, 8, TMI>(ftpa, packed, out, sc, off, bmap, ib, d, st, ld, w0, fib, waves);
    case 16: return launch_mfma_ts<S, 16, TMI>(ftpa, packed, out, sc, off, bmap, ib, d, st, ld, w0, fib, waves);
    case 32: return launch_mfma_ts<S, 32, TMI>(ftpa, packed, out, sc, off, bmap, ib, d, st, ld, w0, fib, waves);
    default:
        set_error("sw_cb_beamform: MFMA form supports tscrunch in {1, 2, 4, 8, 16, 32}");
        return SW_ERR_UNSUPPORTED;
    }
}

int check_dims(const sw_bf_dims& d)
{
    SW_CHECK_ARG(d.nchans > 0 && d.nbeams > 0 && d.nantennas > 0, "empty dimensions");
    SW_CHECK_ARG(d.nantennas % 4 == 0, "Number of antennas must be a multiple of 4");
    SW_CHECK_ARG(d.nbeamsets > 0, "Number of beamsets must be greater than zero");
    SW_CHECK_ARG(d.tscrunch > 0 && d.nsamples > 0 && d.nsamples % d.tscrunch == 0,
                 "nsamples must be a positive multiple of tscrunch");
    SW_CHECK_ARG(d.stokes >= SW_STOKES_I && d.stokes <= SW_STOKES_IQUV, "bad stokes mode");
    return SW_OK;
}

} // namespace

// exposed to the pipeline
int cb_mfma_supported(const sw_bf_dims& d)
{
    return d.nantennas % 16 == 0 && d.nantennas <= 128 && d.nsamples % 32 == 0 &&
           (d.tscrunch == 1 || d.tscrunch == 2 || d.tscrunch == 4 || d.tscrunch == 8 ||
            d.tscrunch == 16 || d.tscrunch == 32);
}

// the incoherent beam can be fused into the time-major MFMA form: Stokes I,
// epilogue tables in LDS (kBeamLds beams, 16 beamsets), tscrunch 4 ... 32,
// 16, 32, 64 or 128 antennas
int cb_fused_ib_supported(const sw_bf_dims& d)
{
    // 16-byte row reads: A divides 256 and a read's 256 / A rows fit an output
    return cb_mfma_supported(d) && d.stokes == SW_STOKES_I && d.nbeams <= kBeamLds && d.nbeamsets <= 16 &&
           d.tscrunch >= 4 && 256 % d.nantennas == 0 && d.tscrunch * d.nantennas >= 256;
}

} // namespace sw

extern "C" size_t sw_cb_workspace_size(const sw_bf_dims* dims)
{
    if(!dims) return 0;
    return (size_t)dims->nchans * dims->nbeams * dims->nantennas * 4;
}

extern "C" int sw_cb_pack_weights(const void* d_weights, void* d_packed, int nchans,
                                  int nbeams, int nantennas, void* stream)
{
    SW_CHECK_ARG(d_weights && d_packed, "null buffer");
    SW_CHECK_ARG(nchans > 0 && nbeams > 0 && nantennas > 0 && nantennas % 4 == 0,
                 "bad weight dimensions");
    SW_CHECK_ARG((reinterpret_cast<uintptr_t>(d_weights) & 7) == 0 &&
                     (reinterpret_cast<uintptr_t>(d_packed) & 15) == 0,
                 "weights must be 8-byte and packed weights 16-byte aligned");
    const size_t nrows = (size_t)nchans * nbeams;
    const size_t n     = nrows * (nantennas / 4);
    hipLaunchKernelGGL(sw::cb_pack_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       sw::as_stream(stream), static_cast<const uint8_t*>(d_weights),
                       static_cast<uint8_t*>(d_packed), nantennas, nrows);
    SW_LAUNCH_CHECK();
    return SW_OK;
}

namespace sw {
// The MFMA form on the coherent dedisperser's time-major ring (Stokes I):
// sample t of column (f, p, a) at d_ring + ((f 2A + p A + a) ld + w0 + t) 2.
// With d_ib_out_raw non-null the incoherent beam is computed here too (see
// FusedIb; requires cb_fused_ib_supported) and written to d_ib_out_raw /
// d_ib_out, and d_ib_raw is not read.
int cb_beamform_packed_tm(const void* d_ring, long ld, long w0, const void* d_packed, const float* d_scale,
                          const float* d_offset, const int* d_beamset_map, const float* d_ib_raw, void* d_out,
                          const sw_bf_dims& d, hipStream_t st, float* d_ib_out_raw, void* d_ib_out,
                          const float* d_ib_scale, const float* d_ib_offset, const float* d_ant_weights, int waves,
                          bool fastdiv)
{
    int rc = check_dims(d);
    if(rc) return rc;
    SW_CHECK_ARG(d_ring && d_packed && d_scale && d_offset && d_beamset_map && d_ib_raw && d_out, "null buffer");
    SW_CHECK_ARG((reinterpret_cast<uintptr_t>(d_ring) & 15) == 0 && (reinterpret_cast<uintptr_t>(d_packed) & 15) == 0
                     && ld % 8 == 0 && w0 % 8 == 0 && d.nsamples % 8 == 0,
                 "ring and packed weights must be 16-byte aligned, rows 8-sample aligned");
    SW_CHECK_ARG(d.stokes == SW_STOKES_I, "time-major beamformer input: Stokes I only");
    if(!cb_mfma_supported(d)) {
        set_error("cb_beamform_packed_tm: shape not supported by the MFMA form");
        return SW_ERR_UNSUPPORTED;
    }
    FusedIb fib;
    if(d_ib_out_raw) {
        SW_CHECK_ARG(cb_fused_ib_supported(d), "fused incoherent beam: unsupported shape");
        SW_CHECK_ARG(d_ib_out && d_ib_scale && d_ib_offset && d_ant_weights, "null buffer");
        fib = FusedIb{d_ib_out_raw, static_cast<int8_t*>(d_ib_out), d_ib_scale, d_ib_offset, d_ant_weights};
    }
    const auto* ring = static_cast<const uint8_t*>(d_ring);
    const auto* pk = static_cast<const uint8_t*>(d_packed);
    auto* out = static_cast<int8_t*>(d_out);
    // every scale checked by the caller: the division by div_by, for the
    // 64-antenna, 4-wave shapes at tscrunch 16 and 8
    if(fastdiv && waves == kMfmaWaves && d.nantennas == 64 && d.tscrunch == 16)
        return launch_mfma_wv<SW_STOKES_I, 16, 4, true, kMfmaWaves, true>(ring, pk, out, d_scale, d_offset,
                                                                          d_beamset_map, d_ib_raw, d, st, ld, w0, fib);
    if(fastdiv && waves == kMfmaWaves && d.nantennas == 64 && d.tscrunch == 8)
        return launch_mfma_wv<SW_STOKES_I, 8, 4, true, kMfmaWaves, true>(ring, pk, out, d_scale, d_offset,
                                                                         d_beamset_map, d_ib_raw, d, st, ld, w0, fib);
    return launch_mfma<SW_STOKES_I, true>(ring, pk, out, d_scale, d_offset, d_beamset_map, d_ib_raw, d, st, ld, w0,
                                          fib, waves);
}

// sw_cb_beamform_packed with `waves` waves (32 time rows each) per workgroup
int cb_beamform_packed_waves(const void* d_ftpa, const void* d_packed, const float* d_scale, const float* d_offset,
                             const int* d_beamset_map, const float* d_ib_raw, void* d_out, const sw_bf_dims& d,
                             hipStream_t st, int waves)
{
    int rc = check_dims(d);
    if(rc) return rc;
    SW_CHECK_ARG(d_ftpa && d_packed && d_scale && d_offset && d_beamset_map && d_ib_raw && d_out,
                 "null buffer");
    SW_CHECK_ARG((reinterpret_cast<uintptr_t>(d_ftpa) & 15) == 0 &&
                     (reinterpret_cast<uintptr_t>(d_packed) & 15) == 0,
                 "voltages and packed weights must be 16-byte aligned");
    if(!cb_mfma_supported(d)) {
        set_error("sw_cb_beamform_packed: shape not supported by the MFMA form "
                  "(needs A % 16 == 0, A <= 128, T % 32 == 0, tscrunch in {1..32, pow2})");
        return SW_ERR_UNSUPPORTED;
    }
    auto* in  = static_cast<const uint8_t*>(d_ftpa);
    auto* pk  = static_cast<const uint8_t*>(d_packed);
    auto* out = static_cast<int8_t*>(d_out);
    const FusedIb none{};
    switch(d.stokes) {
    case SW_STOKES_I: return launch_mfma<SW_STOKES_I, false>(in, pk, out, d_scale, d_offset, d_beamset_map, d_ib_raw, d, st, 0, 0, none, waves);
    case SW_STOKES_Q: return launch_mfma<SW_STOKES_Q, false>(in, pk, out, d_scale, d_offset, d_beamset_map, d_ib_raw, d, st, 0, 0, none, waves);
    case SW_STOKES_U: return launch_mfma<SW_STOKES_U, false>(in, pk, out, d_scale, d_offset, d_beamset_map, d_ib_raw, d, st, 0, 0, none, waves);
    case SW_STOKES_V: return launch_mfma<SW_STOKES_V, false>(in, pk, out, d_scale, d_offset, d_beamset_map, d_ib_raw, d, st, 0, 0, none, waves);
    default: return launch_mfma<SW_STOKES_IQUV, false>(in, pk, out, d_scale, d_offset, d_beamset_map, d_ib_raw, d, st, 0, 0, none, waves);
    }
}
} // namespace sw

extern "C" int sw_cb_beamform_packed(const void* d_ftpa, const void* d_packed,
                                     const float* d_scale, const float* d_offset,
                                     const int* d_beamset_map, const float* d_ib_raw,
                                     void* d_out, const sw_bf_dims* dims, void* stream)
{
    SW_CHECK_ARG(dims, "null dims");
    return sw::cb_beamform_packed_waves(d_ftpa, d_packed, d_scale, d_offset, d_beamset_map, d_ib_raw, d_out, *dims,
                                        sw::as_stream(stream), sw::kMfmaWaves);
}

extern "C" int sw_cb_beamform_packed_tile(const void* d_ftpa, const void* d_packed, const float* d_scale,
                                          const float* d_offset, const int* d_beamset_map, const float* d_ib_raw,
                                          void* d_out, const sw_bf_dims* dims, int waves, void* stream)
{
    SW_CHECK_ARG(dims, "null dims");
    return sw::cb_beamform_packed_waves(d_ftpa, d_packed, d_scale, d_offset, d_beamset_map, d_ib_raw, d_out, *dims,
                                        sw::as_stream(stream), waves);
}

extern "C" int sw_cb_beamform(const void* d_ftpa, const void* d_weights, const float* d_scale,
                              const float* d_offset, const int* d_beamset_map,
                              const float* d_ib_raw, void* d_out, const sw_bf_dims* dims,
                              int form, void* d_workspace, void* stream)
{
    SW_CHECK_ARG(dims, "null dims");
    int rc = sw::check_dims(*dims);
    if(rc) return rc;
    const sw_bf_dims& d = *dims;
    SW_CHECK_ARG(d_ftpa && d_weights && d_scale && d_offset && d_beamset_map && d_ib_raw && d_out,
                 "null buffer");
    if(form == SW_CB_MFMA) {
        void* ws = d_workspace ? d_workspace : sw::scratch(sw_cb_workspace_size(dims), 1);
        if(!ws) return SW_ERR_ALLOC;
        rc = sw_cb_pack_weights(d_weights, ws, d.nchans, d.nbeams, d.nantennas, stream);
        if(rc) return rc;
        return sw_cb_beamform_packed(d_ftpa, ws, d_scale, d_offset, d_beamset_map, d_ib_raw,
                                     d_out, dims, stream);
    }
    SW_CHECK_ARG(form == SW_CB_PLAIN, "unknown beamformer form");
    SW_CHECK_ARG((reinterpret_cast<uintptr_t>(d_ftpa) & 7) == 0 &&
                     (reinterpret_cast<uintptr_t>(d_weights) & 7) == 0,
                 "voltages and weights must be 8-byte aligned");
    auto* in  = static_cast<const uint8_t*>(d_ftpa);
    auto* w   = static_cast<const uint8_t*>(d_weights);
    auto* out = static_cast<int8_t*>(d_out);
    hipStream_t st = sw::as_stream(stream);
    switch(d.stokes) {
    case SW_STOKES_I: return sw::launch_plain<SW_STOKES_I>(in, w, out, d_scale, d_offset, d_beamset_map, d_ib_raw, d, st);
    case SW_STOKES_Q: return sw::launch_plain<SW_STOKES_Q>(in, w, out, d_scale, d_offset, d_beamset_map, d_ib_raw, d, st);
    case SW_STOKES_U: return sw::launch_plain<SW_STOKES_U>(in, w, out, d_scale, d_offset, d_beamset_map, d_ib_raw, d, st);
    case SW_STOKES_V: return sw::launch_plain<SW_STOKES_V>(in, w, out, d_scale, d_offset, d_beamset_map, d_ib_raw, d, st);
    default: return sw::launch_plain<SW_STOKES_IQUV>(in, w, out, d_scale, d_offset, d_beamset_map, d_ib_raw, d, st);
    }
}
