// idisp.hip -- incoherent (post-detection) dedispersion of TFB powers on the
// GPU, and the DedispersionPlan block parser.
//
// Reference: IncoherentDedisperser::prepare / dedisperse,
// cpp/skyweaver/src/IncoherentDedisperser.cu:36-125, with DMSampleDelay
// (cpp/skyweaver/dedispersion_utils.cuh:19-40):
//   delta[d][f] = (int)(K dm_d (1/f_0^2 - 1/f^2) / tsamp + 0.5), f_0 = lowest channel,
//   tsamp = cb_tscrunch / (bandwidth / nchans)
//   out[t'][d][b] = (int8)clamp(sum_{s<ts} sum_f in[D + t' ts + s - delta[d][f]][f][b]
//                               / (float)sqrt(F ts))
// with D = max delta and t' < (T - D) / ts.  The sums are exact integers, so
// the summation order is free; the division and the truncating clamp are the
// reference's.
//
// Kernel design: a workgroup owns 64 words (256 bytes) of a (t, f) row,
// TT output samples and a tile of up to 4 DPW consecutive DMs (one wave per
// DM residue).  For every channel it stages the union of the rows its DMs
// read into LDS, biased by XOR 0x80 so every byte is an unsigned x + 128; each
// lane then adds its word into two registers of 16-bit lanes (even and odd
// bytes), i.e. four beam sums per two integer adds.  The 16-bit lanes hold
// at most 256 biased bytes; longer sums (nchans * ts > 256) are flushed into
// 32-bit accumulators.
//
// Matrix-core form (idisp_mfma_kernel, used whenever nchans % 4 == 0 and the
// rows of one DM fit the stage): one DM's channel sum is a chain of
// v_mfma_i32_16x16x64_i8 products.  The B operand carries data: lane
// (n, g) = (l % 16, l / 16) holds the 16 beam bytes [b0, b0 + 16) of channel
// f0 + g at the input row of output sample n.  The A operand is a constant
// selector, row m holding a single 1 at byte m of every k-group, so
// C[m][n] = sum over g of beam b0 + m of channel f0 + g at sample n: four
// channels x 16 beams x 16 samples per instruction, summed exactly in int32.
// For tscrunch <= 2 with 16-byte rows the same products run in
// idisp_dma_kernel: 16 waves of one DM each, 32 output samples per
// workgroup, each quad of channels loaded straight into one of two LDS stage
// buffers while the other quad's products run.
#include "sw_common.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <sstream>
#include <string>
#include <type_traits>
#include <vector>

struct sw_idisp {
    int nchans = 0, ndms = 0, tscrunch = 1;
    int max_delay = 0;
    float scale = 1.0f;
    std::vector<float> dms;
    std::vector<int> delays;          // [D][F]
    // device tiling
    int dpw = 2, tt = 16, rows_cap = 0;
    bool wide = false;
    bool mfma = false;                // matrix-core kernel (else the 16-bit-lane one)
    std::vector<int2> tiles;          // (first DM, count)
    int* d_delays = nullptr;
    int2* d_span = nullptr;           // [ntiles][F] (delta_min, delta_max)
    int2* d_tiles = nullptr;
};

namespace sw {
namespace {

constexpr double kDmConstHzS = 4.148806423e15;  // Hz^2 pc^-1 cm^3 s

// one step of 16-bit-lane accumulation of a biased word
__device__ __forceinline__ void id_add(uint32_t w, uint32_t& e, uint32_t& o)
{
    e += w & 0x00FF00FFu;
    o += (w >> 8) & 0x00FF00FFu;
}

__device__ __forceinline__ int8_t id_quant(int32_t sum, float scale)
{
    // reference: clamp<int8_t>(float_sum / scale_factor), truncating cast
    const float v = (float)sum / scale;
    return (int8_t)fmaxf(-128.0f, fminf(127.0f, v));
}

constexpr int kPfRows = 64;  // stage rows up to which the next channel is prefetched in registers

// Input word (t, f, w) lives at t * stride_t + f * stride_f + w: TFB
// (stride_t = F * nwords, stride_f = nwords) or channel-major FTB
// (stride_t = nwords, stride_f = T * nwords), which keeps a channel's rows
// together.  PF: the next channel's stage (<= kPfRows rows) is loaded into
// registers while the current one is summed.
template <int DPW, int TT, bool WIDE, bool PF>
__global__ void __launch_bounds__(256)
idisp_kernel(const uint32_t* __restrict__ tfb, uint32_t* __restrict__ tdb,
             const int* __restrict__ delays, const int2* __restrict__ span,
             const int2* __restrict__ tiles, int nchans, int nwords, int ndms, int ts,
             int max_delay, long nsamples, int nout, int tb0, float scale, long stride_t, long stride_f)
{
    extern __shared__ uint4 stage[];  // [rows][16] uint4 = [rows][64] words
    const uint32_t* rows = reinterpret_cast<const uint32_t*>(stage);
    const int tile  = blockIdx.x;
    const int chunk = blockIdx.y;
    const int tb    = tb0 + blockIdx.z;
    const int wave  = threadIdx.x >> 6;
    const int lane  = threadIdx.x & 63;
    const int2 td   = tiles[tile];
    const int t0    = tb * TT;
    const long in0  = (long)max_delay + (long)t0 * ts;

    uint32_t e[DPW][TT], o[DPW][TT];
    uint32_t w32[WIDE ? DPW : 1][WIDE ? TT : 1][4];
#pragma unroll
    for(int j = 0; j < DPW; ++j)
#pragma unroll
        for(int i = 0; i < TT; ++i) {
            e[j][i] = 0;
            o[j][i] = 0;
            if constexpr(WIDE) {
#pragma unroll
                for(int k = 0; k < 4; ++k) w32[j][i][k] = 0;
            }
        }
    int pending = 0;  // biased bytes in the 16-bit lanes (WIDE only)

    // 16-byte pieces of channel f's stage (rows 16-byte aligned: nwords % 4 == 0)
    auto piece = [&](int f, long lo, int i) {
        const int r = i >> 4, q = i & 15;
        const int w4 = chunk * 64 + q * 4;
        uint4 v = make_uint4(0x80808080u, 0x80808080u, 0x80808080u, 0x80808080u);
        if(w4 < nwords && lo + r < nsamples) {
            v = *reinterpret_cast<const uint4*>(tfb + (lo + r) * stride_t + f * stride_f + w4);
            v.x ^= 0x80808080u;
            v.y ^= 0x80808080u;
            v.z ^= 0x80808080u;
            v.w ^= 0x80808080u;
        }
        return v;
    };
    constexpr int NPF = PF ? kPfRows * 16 / 256 : 1;
    uint4 pre[NPF];
    if constexpr(PF) {
        const int2 sp = span[(size_t)tile * nchans];
        const int nr = TT * ts + sp.y - sp.x;
#pragma unroll
        for(int k = 0; k < NPF; ++k) {
            const int i = threadIdx.x + 256 * k;
            if(i < nr * 16) pre[k] = piece(0, in0 - sp.y, i);
        }
    }

    for(int f = 0; f < nchans; ++f) {
        const int2 sp   = span[(size_t)tile * nchans + f];
        const long lo   = in0 - sp.y;
        const int nrows = TT * ts + sp.y - sp.x;
        __syncthreads();
        if constexpr(PF) {
#pragma unroll
            for(int k = 0; k < NPF; ++k) {
                const int i = threadIdx.x + 256 * k;
                if(i < nrows * 16) stage[i] = pre[k];
            }
            if(f + 1 < nchans) {  // next channel's rows, in flight during the sums
                const int2 sn = span[(size_t)tile * nchans + f + 1];
                const int nr = TT * ts + sn.y - sn.x;
#pragma unroll
                for(int k = 0; k < NPF; ++k) {
                    const int i = threadIdx.x + 256 * k;
                    if(i < nr * 16) pre[k] = piece(f + 1, in0 - sn.y, i);
                }
            }
        } else if((nwords & 3) == 0) {
            for(int i = threadIdx.x; i < nrows * 16; i += 256) stage[i] = piece(f, lo, i);
        } else {
            uint32_t* st32 = reinterpret_cast<uint32_t*>(stage);
            for(int i = threadIdx.x; i < nrows * 64; i += 256) {
                const int r = i >> 6, q = i & 63;
                const int w = chunk * 64 + q;
                uint32_t v = 0x80808080u;
                if(w < nwords && lo + r < nsamples) v = tfb[(lo + r) * stride_t + f * stride_f + w] ^ 0x80808080u;
                st32[i] = v;
            }
        }
        __syncthreads();
#pragma unroll
        for(int j = 0; j < DPW; ++j) {
            const int dl = wave + 4 * j;
            if(dl < td.y) {
                const int delta = delays[(size_t)(td.x + dl) * nchans + f];
                const uint32_t* src = rows + (sp.y - delta) * 64 + lane;
#pragma unroll
                for(int i = 0; i < TT; ++i)
                    for(int s = 0; s < ts; ++s) id_add(src[(i * ts + s) * 64], e[j][i], o[j][i]);
            }
        }
        if constexpr(WIDE) {
            pending += ts;
            if(pending + ts > 256 || f == nchans - 1) {
#pragma unroll
                for(int j = 0; j < DPW; ++j)
#pragma unroll
                    for(int i = 0; i < TT; ++i) {
                        w32[j][i][0] += e[j][i] & 0xFFFFu;
                        w32[j][i][1] += o[j][i] & 0xFFFFu;
                        w32[j][i][2] += e[j][i] >> 16;
                        w32[j][i][3] += o[j][i] >> 16;
                        e[j][i] = 0;
                        o[j][i] = 0;
                    }
                pending = 0;
            }
        }
    }

    const int word = chunk * 64 + lane;
    if(word >= nwords) return;
    const int32_t bias = 128 * nchans * ts;
#pragma unroll
    for(int j = 0; j < DPW; ++j) {
        const int dl = wave + 4 * j;
        if(dl >= td.y) continue;
#pragma unroll
        for(int i = 0; i < TT; ++i) {
            if(t0 + i >= nout) continue;
            int32_t s[4];
            if constexpr(WIDE) {
#pragma unroll
                for(int k = 0; k < 4; ++k) s[k] = (int32_t)w32[j][i][k] - bias;
            } else {
                s[0] = (int32_t)(e[j][i] & 0xFFFFu) - bias;
                s[1] = (int32_t)(o[j][i] & 0xFFFFu) - bias;
                s[2] = (int32_t)(e[j][i] >> 16) - bias;
                s[3] = (int32_t)(o[j][i] >> 16) - bias;
            }
            const uint32_t packed = (uint32_t)(uint8_t)id_quant(s[0], scale) |
                                    ((uint32_t)(uint8_t)id_quant(s[1], scale) << 8) |
                                    ((uint32_t)(uint8_t)id_quant(s[2], scale) << 16) |
                                    ((uint32_t)(uint8_t)id_quant(s[3], scale) << 24);
            tdb[((size_t)(t0 + i) * ndms + td.x + dl) * nwords + word] = packed;
        }
    }
}

constexpr int kRowsCap = 256;  // staged rows per channel: 64 KB of LDS

// ---------------------------------------------------------------- MFMA form

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int kMfWaves = 8;    // waves per workgroup
constexpr int kMfDpw   = 2;    // DMs per wave
constexpr int kDmaWaves = 16;  // waves (one DM each) per workgroup of the double-buffered form
constexpr int kMfRows  = 128;  // staged rows per channel
constexpr int kMfPitch = 144;  // bytes between staged rows: 128 B of beams + 16;
                               // 16 consecutive rows start on distinct banks

template <typename P>
__device__ __forceinline__ P zero_piece()
{
    if constexpr(std::is_same<P, uint4>::value)
        return make_uint4(0, 0, 0, 0);
    else
        return 0u;
}

// A workgroup of NW waves owns TT = 16 NTT output samples, 128 bytes of beams
// (a chunk) and a tile of up to NW * DPW consecutive DMs (wave w takes the
// tile's DMs w, w + NW, ...).  Per quad of channels it stages into LDS the
// union of the rows its DMs read, [in0 - max delay, in0 + TT ts - min delay)
// per channel, while the rows of the next two quads are loaded into
// registers.  Input byte (t, f, b) lives at t * stride_t + f * stride_f + b
// (TFB or FTB).
//
// Work items are dealt to the 8 XCDs in contiguous runs, and a run is walked
// in groups of kIdTbGroup time blocks x kIdTileGroup DM tiles of one chunk:
// the 32 workgroups that one XCD runs together read overlapping rows (next
// time blocks; DM tiles of neighbouring delays), which its L2 then serves.
constexpr int kIdTbGroup = 8, kIdTileGroup = 4;

template <int NW, int DPW, int NTT, int RCAP, bool VEC16, bool PF>
__global__ void __launch_bounds__(NW * 64, PF ? 1 : 2)
idisp_mfma_kernel(const int8_t* __restrict__ in, int8_t* __restrict__ tdb, const int* __restrict__ delays,
                  const int2* __restrict__ span, const int2* __restrict__ tiles, int ntiles, int nchunks,
                  int ntb, long w0, long nwork, int nchans, int W, int ndms, int ts, int max_delay,
                  long nsamples, int nout, float scale, long stride_t, long stride_f)
{
    constexpr int NT    = NW * 64;
    constexpr int TT    = 16 * NTT;
    constexpr int PIECE = VEC16 ? 16 : 4;  // bytes per staging load
    constexpr int PPR   = 128 / PIECE;     // loads per staged row
    static_assert((RCAP * PPR) % NT == 0, "a thread's loads must stay in one channel each");
    constexpr int NPF = 4 * RCAP * PPR / NT;
    using P = typename std::conditional<VEC16, uint4, uint32_t>::type;
    __shared__ __attribute__((aligned(16))) int8_t stage[4 * RCAP * kMfPitch];

    const long per = (nwork + 7) / 8;
    const long wl  = (long)(blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if(wl >= nwork) return;  // whole workgroup, before any barrier
    const long w = w0 + wl;
    constexpr int GRP = kIdTbGroup * kIdTileGroup;
    const int gi = (int)(w % GRP);
    const long rest  = w / GRP;
    const int chunk  = (int)(rest % nchunks);
    const long blk   = rest / nchunks;
    const int ntbb   = (ntb + kIdTbGroup - 1) / kIdTbGroup;
    const int tb     = (int)(blk % ntbb) * kIdTbGroup + gi % kIdTbGroup;
    const int tile   = (int)(blk / ntbb) * kIdTileGroup + gi / kIdTbGroup;
    if(tb >= ntb || tile >= ntiles) return;  // whole workgroup, before any barrier
    const int2 td    = tiles[tile];
    const int t0     = tb * TT;
    const long in0   = (long)max_delay + (long)t0 * ts;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int n = lane & 15, g = lane >> 4;
    const int b0 = chunk * 128;
    const int nq = nchans / 4;
    const int2* tspan = span + (size_t)tile * nchans;

    P preA[NPF], preB[NPF];
    auto fetch = [&](int q, P (&pre)[NPF]) {
#pragma unroll
        for(int k = 0; k < NPF; ++k) {
            const int s = threadIdx.x + NT * k;
            const int j = (NT * k) / (RCAP * PPR);  // compile-time: NT divides RCAP * PPR
            const int r = (s / PPR) % RCAP, c = s % PPR;
            const int2 spj = tspan[4 * q + j];
            const long row = in0 - spj.y + r;
            const int b = b0 + c * PIECE;
            P v = zero_piece<P>();
            if(r < TT * ts + spj.y - spj.x && b < W && row < nsamples)
                v = *reinterpret_cast<const P*>(in + row * stride_t + (long)(4 * q + j) * stride_f + b);
            pre[k] = v;
        }
    };
    // only the rows the quad's DMs read are written: a channel of quad q
    // needs TT ts + (its span of delays) rows, often far fewer than RCAP
    auto stash = [&](int q, const P (&pre)[NPF]) {
#pragma unroll
        for(int k = 0; k < NPF; ++k) {
            const int s = threadIdx.x + NT * k;
            const int j = (NT * k) / (RCAP * PPR);
            const int r = (s / PPR) % RCAP, c = s % PPR;
            const int2 spj = tspan[4 * q + j];
            if(r < TT * ts + spj.y - spj.x)
                *reinterpret_cast<P*>(stage + (j * RCAP + r) * kMfPitch + c * PIECE) = pre[k];
        }
    };
    // this lane's channel of quad q: its span and the delays of the wave's DMs
    auto lane_delays = [&](int q, int2& spg, int (&dly)[DPW]) {
        spg = tspan[4 * q + g];
#pragma unroll
        for(int jd = 0; jd < DPW; ++jd) {
            const int dl = wave + NW * jd;
            dly[jd] = dl < td.y ? delays[(size_t)(td.x + dl) * nchans + 4 * q + g] : spg.y;
        }
    };

    v4i sel = {0, 0, 0, 0};
    sel[n >> 2] = 1 << (8 * (n & 3));
    v4i acc[DPW][NTT][8];
#pragma unroll
    for(int jd = 0; jd < DPW; ++jd)
#pragma unroll
        for(int tt = 0; tt < NTT; ++tt)
#pragma unroll
            for(int bg = 0; bg < 8; ++bg) acc[jd][tt][bg] = v4i{0, 0, 0, 0};

    auto products = [&](int2 spc, const int (&dc)[DPW]) {
#pragma unroll
        for(int jd = 0; jd < DPW; ++jd) {
            if(wave + NW * jd >= td.y) continue;  // wave-uniform
            const int8_t* base = stage + (g * RCAP + spc.y - dc[jd] + n * ts) * kMfPitch;
#pragma unroll
            for(int tt = 0; tt < NTT; ++tt)
                for(int s = 0; s < ts; ++s) {
                    const int8_t* rp = base + (tt * 16 * ts + s) * kMfPitch;
                    if constexpr(!PF) {
                        // the row's eight reads in flight together (the
                        // scheduler would otherwise issue them one by one)
                        v4i bv[8];
#pragma unroll
                        for(int bg = 0; bg < 8; ++bg) bv[bg] = *reinterpret_cast<const v4i*>(rp + bg * 16);
                        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                        for(int bg = 0; bg < 8; ++bg)
                            acc[jd][tt][bg] =
                                __builtin_amdgcn_mfma_i32_16x16x64_i8(sel, bv[bg], acc[jd][tt][bg], 0, 0, 0);
                    } else {
#pragma unroll
                        for(int bg = 0; bg < 8; ++bg) {
                            const v4i b = *reinterpret_cast<const v4i*>(rp + bg * 16);
                            acc[jd][tt][bg] = __builtin_amdgcn_mfma_i32_16x16x64_i8(sel, b, acc[jd][tt][bg], 0, 0, 0);
                        }
                    }
                }
        }
    };

    if constexpr(!PF) {
        // no register prefetch (two workgroups share a CU and hide each
        // other's loads): stage quad q, then its products
        for(int q = 0; q < nq; ++q) {
            int2 spc;
            int dc[DPW];
            lane_delays(q, spc, dc);  // issued with the row loads, before any wait
            fetch(q, preA);
            __syncthreads();  // the previous quad's reads are done
            stash(q, preA);
            __syncthreads();
            products(spc, dc);
        }
    } else {
    // quads alternate between two register sets: A holds even quads, B odd
    int2 spA, spB;
    int dA[DPW], dB[DPW];
    fetch(0, preA);
    lane_delays(0, spA, dA);
    if(nq > 1) {
        fetch(1, preB);
        lane_delays(1, spB, dB);
    }
    for(int q = 0; q < nq; q += 2) {
        {
            __syncthreads();  // the previous quad's reads are done
            stash(q, preA);
            __syncthreads();
            const int2 spc = spA;
            int dc[DPW];
#pragma unroll
            for(int jd = 0; jd < DPW; ++jd) dc[jd] = dA[jd];
            if(q + 2 < nq) {  // in flight during two quads' products
                fetch(q + 2, preA);
                lane_delays(q + 2, spA, dA);
            }
            products(spc, dc);
        }
        if(q + 1 < nq) {
            __syncthreads();
            stash(q + 1, preB);
            __syncthreads();
            const int2 spc = spB;
            int dc[DPW];
#pragma unroll
            for(int jd = 0; jd < DPW; ++jd) dc[jd] = dB[jd];
            if(q + 3 < nq) {
                fetch(q + 3, preB);
                lane_delays(q + 3, spB, dB);
            }
            products(spc, dc);
        }
    }
    }

    // C lane (n, g), register r: beam b0 + 16 bg + 4 g + r at sample t0 + 16 tt + n
#pragma unroll
    for(int jd = 0; jd < DPW; ++jd) {
        const int dl = wave + NW * jd;
        if(dl >= td.y) continue;
#pragma unroll
        for(int tt = 0; tt < NTT; ++tt) {
            const int t = t0 + 16 * tt + n;
            if(t >= nout) continue;
            int8_t* o = tdb + ((size_t)t * ndms + td.x + dl) * W;
#pragma unroll
            for(int bg = 0; bg < 8; ++bg) {
                const int b = b0 + 16 * bg + 4 * g;
                if(b >= W) continue;
                const v4i a = acc[jd][tt][bg];
                const uint32_t packed = (uint32_t)(uint8_t)id_quant(a[0], scale) |
                                        ((uint32_t)(uint8_t)id_quant(a[1], scale) << 8) |
                                        ((uint32_t)(uint8_t)id_quant(a[2], scale) << 16) |
                                        ((uint32_t)(uint8_t)id_quant(a[3], scale) << 24);
                *reinterpret_cast<uint32_t*>(o + b) = packed;
            }
        }
    }
}

// Double-buffered form: a workgroup of NW = 16 waves (one DM each) owns TT =
// 16 NTT output samples, a 128-byte chunk of beams and a tile of up to 16
// consecutive DMs.  Each quad's rows go from global memory straight into one
// of two LDS stage buffers (global_load_lds_dwordx4: lane L of a wave's load
// writes the 16-byte slot L of a 1 KB block) while the previous quad's
// products read the other buffer, one barrier per quad and no staging
// registers.  The stage keeps idisp_mfma_kernel's layout (row pitch
// kMfPitch: nine 16-byte slots, the ninth padding that no lane loads), and
// the products read it exactly as there.  The two buffers take 144 KB: one
// workgroup of 16 waves per CU.  Slots that no lane loads (padding, rows
// past the tile's span or the input, beams past W) keep stale bytes: no
// stored output reads them (rows past the span are read by no DM of the
// tile, rows past the input feed only output samples >= nout, beams past W
// only beams >= W, and the one-hot products sum each beam from its own
// bytes).
template <int NW, int NTT, int RCAP>
__global__ void __launch_bounds__(NW * 64, 1)
idisp_dma_kernel(const int8_t* __restrict__ in, int8_t* __restrict__ tdb, const int* __restrict__ delays,
                 const int2* __restrict__ span, const int2* __restrict__ tiles, int ntiles, int nchunks,
                 int ntb, long w0, long nwork, int nchans, int W, int ndms, int ts, int max_delay,
                 long nsamples, int nout, float scale, long stride_t, long stride_f)
{
    constexpr int TT  = 16 * NTT;
    constexpr int SPR = kMfPitch / 16;  // 16-byte slots per staged row
    constexpr int SPC = RCAP * SPR;     // slots per channel
    static_assert(kMfPitch % 16 == 0 && SPC % 64 == 0, "a wave's 64-slot block stays in one channel");
    constexpr int NBLK = 4 * SPC / 64;          // 1 KB blocks per quad
    constexpr int NB   = (NBLK + NW - 1) / NW;  // blocks per wave per quad (the last round partial)
    constexpr int QB   = 4 * RCAP * kMfPitch;
    // two arrays, not one: the compiler then knows that a buffer's reads do
    // not touch the loads in flight into the other (with one array it waits
    // for every row load before any read)
    __shared__ __attribute__((aligned(16))) int8_t stage0[QB];
    __shared__ __attribute__((aligned(16))) int8_t stage1[QB];

    const long per = (nwork + 7) / 8;
    const long wl  = (long)(blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if(wl >= nwork) return;  // whole workgroup, before any barrier
    const long w = w0 + wl;
    constexpr int GRP = kIdTbGroup * kIdTileGroup;
    const int gi = (int)(w % GRP);
    const long rest  = w / GRP;
    const int chunk  = (int)(rest % nchunks);
    const long blk   = rest / nchunks;
    const int ntbb   = (ntb + kIdTbGroup - 1) / kIdTbGroup;
    const int tb     = (int)(blk % ntbb) * kIdTbGroup + gi % kIdTbGroup;
    const int tile   = (int)(blk / ntbb) * kIdTileGroup + gi / kIdTbGroup;
    if(tb >= ntb || tile >= ntiles) return;  // whole workgroup, before any barrier
    const int2 td    = tiles[tile];
    const int t0     = tb * TT;
    const long in0   = (long)max_delay + (long)t0 * ts;
    const int lane = threadIdx.x & 63;
    // the wave index made uniform: the spans below are scalar loads whose
    // waits do not hold back the row loads
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int n = lane & 15, g = lane >> 4;
    const int b0 = chunk * 128;
    const int nq = nchans / 4;
    const int2* tspan = span + (size_t)tile * nchans;
    const bool active = wv < td.y;  // this wave's DM is in the tile

    // quad q's rows into stg: block k NW + wave of the quad, slot lane
    auto issue = [&](int q, int8_t* stg) {
        // the quad's four spans: scalar loads in flight together
        const int2 sq0 = tspan[4 * q], sq1 = tspan[4 * q + 1], sq2 = tspan[4 * q + 2], sq3 = tspan[4 * q + 3];
        int ln = lane;
        asm volatile("" : "+v"(ln));  // the slot arithmetic is redone per quad, not held in registers
#pragma unroll
        for(int k = 0; k < NB; ++k) {
            const int bk = k * NW + wv;  // wave-uniform
            if(NBLK % NW != 0 && bk >= NBLK) break;
            const int j  = bk * 64 / SPC;  // channel of the quad, wave-uniform
            const int s  = bk * 64 - j * SPC + ln;
            const int r = s / SPR, c = s % SPR;
            const int2 spj = j == 0 ? sq0 : j == 1 ? sq1 : j == 2 ? sq2 : sq3;
            const long row = in0 - spj.y + r;
            const int b = b0 + c * 16;
            if(c < 8 && r < TT * ts + spj.y - spj.x && b < W && row < nsamples)
                __builtin_amdgcn_global_load_lds(in + row * stride_t + (long)(4 * q + j) * stride_f + b,
                                                 (__attribute__((address_space(3))) void*)(stg + bk * 1024), 16,
                                                 0, 0);
        }
    };
    // this lane's channel of quad q: its span and the delay of the wave's DM
    // (a wave past the tile's DMs loads the last DM's and does not use it, so
    // neither load waits on the other)
    const int dmw = td.x + min(wv, td.y - 1);
    auto lane_delay = [&](int q, int2& spg, int& dly) {
        spg = tspan[4 * q + g];
        dly = delays[(size_t)dmw * nchans + 4 * q + g];
    };

    v4i sel = {0, 0, 0, 0};
    sel[n >> 2] = 1 << (8 * (n & 3));
    v4i acc[NTT][8];
#pragma unroll
    for(int tt = 0; tt < NTT; ++tt)
#pragma unroll
        for(int bg = 0; bg < 8; ++bg) acc[tt][bg] = v4i{0, 0, 0, 0};

    auto products = [&](const int8_t* stg, int2 spc, int dc) {
        if(!active) return;  // wave-uniform
        const int8_t* base = stg + (g * RCAP + spc.y - dc + n * ts) * kMfPitch;
#pragma unroll
        for(int tt = 0; tt < NTT; ++tt)
            for(int s = 0; s < ts; ++s) {
                const int8_t* rp = base + (tt * 16 * ts + s) * kMfPitch;
                v4i bv[8];  // the row's eight reads in flight together
#pragma unroll
                for(int bg = 0; bg < 8; ++bg) bv[bg] = *reinterpret_cast<const v4i*>(rp + bg * 16);
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for(int bg = 0; bg < 8; ++bg)
                    acc[tt][bg] = __builtin_amdgcn_mfma_i32_16x16x64_i8(sel, bv[bg], acc[tt][bg], 0, 0, 0);
            }
    };

    // quads alternate between the buffers: even quads in stage0, odd in stage1
    int2 spA, spB;
    int dA, dB;
    lane_delay(0, spA, dA);
    issue(0, stage0);
    for(int q = 0; q < nq; q += 2) {
        __builtin_amdgcn_s_waitcnt(0);  // this wave's loads of quad q
        __syncthreads();                // every wave's; quad q - 1's products are done, stage1 is free
        if(q + 1 < nq) {
            lane_delay(q + 1, spB, dB);
            issue(q + 1, stage1);
        }
        products(stage0, spA, dA);
        if(q + 1 < nq) {
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();  // quad q + 1 staged; quad q's products are done, stage0 is free
            if(q + 2 < nq) {
                lane_delay(q + 2, spA, dA);
                issue(q + 2, stage0);
            }
            products(stage1, spB, dB);
        }
    }

    // C lane (n, g), register r: beam b0 + 16 bg + 4 g + r at sample t0 + 16 tt + n
    if(!active) return;
#pragma unroll
    for(int tt = 0; tt < NTT; ++tt) {
        const int t = t0 + 16 * tt + n;
        if(t >= nout) continue;
        int8_t* o = tdb + ((size_t)t * ndms + td.x + wv) * W;
#pragma unroll
        for(int bg = 0; bg < 8; ++bg) {
            const int b = b0 + 16 * bg + 4 * g;
            if(b >= W) continue;
            const v4i a = acc[tt][bg];
            const uint32_t packed = (uint32_t)(uint8_t)id_quant(a[0], scale) |
                                    ((uint32_t)(uint8_t)id_quant(a[1], scale) << 8) |
                                    ((uint32_t)(uint8_t)id_quant(a[2], scale) << 16) |
                                    ((uint32_t)(uint8_t)id_quant(a[3], scale) << 24);
            *reinterpret_cast<uint32_t*>(o + b) = packed;
        }
    }
}

void free_plan_device(sw_idisp* p)
{
    (void)hipFree(p->d_delays);
    (void)hipFree(p->d_span);
    (void)hipFree(p->d_tiles);
    p->d_delays = nullptr;
    p->d_span = nullptr;
    p->d_tiles = nullptr;
}

} // namespace
} // namespace sw

extern "C" int sw_id_delays(const float* dms, int ndms, const double* chan_freqs, int nchans,
                            double tsamp, int* h_delays, int* max_delay)
{
    SW_CHECK_ARG(dms && chan_freqs && h_delays && ndms > 0 && nchans > 0 && tsamp > 0, "bad argument");
    // DMSampleDelay(dm, f_ref = channel_frequencies().front(), tsamp)(f)
    const double f_ref = chan_freqs[0];
    int mx = 0;
    for(int d = 0; d < ndms; ++d)
        for(int c = 0; c < nchans; ++c) {
            const double f = chan_freqs[c];
            const double delay = sw::kDmConstHzS * (1 / (f_ref * f_ref) - 1 / (f * f)) * (double)dms[d];
            const int v = (int)(delay / tsamp + 0.5);
            h_delays[(size_t)d * nchans + c] = v;
            mx = std::max(mx, v);
        }
    if(max_delay) *max_delay = mx;
    return SW_OK;
}

extern "C" int sw_channel_frequencies(double centre_frequency, double bandwidth, int nchans,
                                      double* h_freqs)
{
    // PipelineConfig::calculate_channel_frequencies,
    // cpp/skyweaver/src/PipelineConfig.cpp:188-203
    SW_CHECK_ARG(h_freqs && nchans > 0, "bad argument");
    const double chbw = bandwidth / nchans;
    const double fbottom = centre_frequency - bandwidth / 2.0;
    for(int c = 0; c < nchans; ++c) h_freqs[c] = fbottom + chbw / 2.0 + (chbw * c);
    return SW_OK;
}

extern "C" int sw_ddplan_parse_block(const char* spec, float* coherent_dm, int* tscrunch,
                                     float* h_dms, int max_dms, int* ndms)
{
    // DedispersionPlan::add_block(std::string), cpp/skyweaver/src/DedispersionPlan.cpp:17-78:
    // "cdm", "cdm:tscrunch" or "cdm:begin:end:step:tscrunch"; the incoherent
    // DMs of the last form are accumulated in float (dm += step while dm < end),
    // and an empty range yields {begin}.
    SW_CHECK_ARG(spec && coherent_dm && tscrunch && ndms, "null argument");
    std::vector<std::string> parts;
    {
        std::stringstream ss(spec);
        std::string item;
        while(std::getline(ss, item, ':')) parts.push_back(item);
    }
    std::vector<float> dms;
    try {
        if(parts.size() == 1 || parts.size() == 2) {
            *coherent_dm = std::stof(parts[0]);
            *tscrunch = parts.size() == 2 ? std::stoi(parts[1]) : 1;
            dms.push_back(*coherent_dm);
        } else if(parts.size() == 5) {
            *coherent_dm = std::stof(parts[0]);
            const float b = std::stof(parts[1]), e = std::stof(parts[2]), st = std::stof(parts[3]);
            *tscrunch = std::stoi(parts[4]);
            for(float dm = b; dm < e; dm += st) {
                dms.push_back(dm);
                if(dms.size() > (size_t)1 << 24) break;
            }
            if(dms.empty()) dms.push_back(b);
        } else {
            sw::set_error(std::string("Block defition string of unexpected size: ") + spec);
            return SW_ERR_INVALID;
        }
    } catch(std::invalid_argument& e) {
        sw::set_error(std::string("Error parsing block definition string: ") + e.what());
        return SW_ERR_INVALID;
    } catch(std::out_of_range&) {
        sw::set_error("Block definition value out of range");
        return SW_ERR_INVALID;
    }
    *ndms = (int)dms.size();
    if(h_dms) {
        SW_CHECK_ARG(max_dms >= (int)dms.size(), "DM buffer too small");
        std::memcpy(h_dms, dms.data(), sizeof(float) * dms.size());
    }
    return SW_OK;
}

extern "C" int sw_idisp_create(const sw_idisp_config* cfg, sw_idisp** out)
{
    SW_CHECK_ARG(cfg && out, "null argument");
    *out = nullptr;
    SW_CHECK_ARG(cfg->nchans > 0 && cfg->ndms > 0 && cfg->dms, "empty plan");
    SW_CHECK_ARG(cfg->tscrunch > 0 && cfg->cb_tscrunch > 0, "tscrunch must be positive");
    SW_CHECK_ARG(cfg->bandwidth != 0.0, "bandwidth must be non-zero");
    auto p = std::make_unique<sw_idisp>();
    p->nchans = cfg->nchans;
    p->ndms = cfg->ndms;
    p->tscrunch = cfg->tscrunch;
    p->dms.assign(cfg->dms, cfg->dms + cfg->ndms);
    std::vector<double> freqs(cfg->nchans);
    if(cfg->chan_freqs)
        freqs.assign(cfg->chan_freqs, cfg->chan_freqs + cfg->nchans);
    else
        sw_channel_frequencies(cfg->centre_frequency, cfg->bandwidth, cfg->nchans, freqs.data());
    const double chbw = cfg->bandwidth / cfg->nchans;
    const double tsamp = cfg->cb_tscrunch / chbw;
    p->delays.resize((size_t)cfg->ndms * cfg->nchans);
    sw_id_delays(p->dms.data(), p->ndms, freqs.data(), p->nchans, tsamp, p->delays.data(), &p->max_delay);
    for(int v: p->delays)
        SW_CHECK_ARG(v >= 0, "channel frequencies must not lie below the first channel's");
    p->scale = (float)std::sqrt((double)((size_t)p->nchans * p->tscrunch));

    // tiling: TT output samples per workgroup, tiles of consecutive DMs whose
    // union of read rows fits the LDS stage
    int cap = sw::kRowsCap, dmax_tile = 8;
    p->mfma = cfg->kernel == 0 && p->nchans % 4 == 0 && 16 * p->tscrunch <= sw::kMfRows;
    if(p->mfma) {
        // 32 output samples per workgroup of 16 waves (one DM each), the
        // quads' rows double-buffered in LDS; 16 samples per workgroup of 8
        // waves (two DMs each) where 32 samples' rows would not fit the stage
        p->tt = 32 * p->tscrunch <= sw::kMfRows / 2 ? 32 : 16;
        cap = sw::kMfRows;
        dmax_tile = p->tt == 32 ? sw::kDmaWaves : sw::kMfWaves * sw::kMfDpw;
    } else {
        p->wide = (long)p->nchans * p->tscrunch > 256;
        p->tt = p->wide ? 8 : 16;
        if(p->tt * p->tscrunch > sw::kRowsCap / 2) p->tt = 4;
        if(p->tt * p->tscrunch > sw::kRowsCap) {
            sw::set_error("incoherent dedispersion tscrunch too large (at most 64)");
            return SW_ERR_UNSUPPORTED;
        }
        p->dpw = 2;
    }
    const int F = p->nchans;
    std::vector<int2> span;
    for(int d = 0; d < p->ndms;) {
        int n = 1;
        while(n < dmax_tile && d + n < p->ndms) {
            int worst = 0;
            for(int f = 0; f < F; ++f) {
                int lo = INT32_MAX, hi = INT32_MIN;
                for(int k = 0; k <= n; ++k) {
                    lo = std::min(lo, p->delays[(size_t)(d + k) * F + f]);
                    hi = std::max(hi, p->delays[(size_t)(d + k) * F + f]);
                }
                worst = std::max(worst, hi - lo);
            }
            if(worst + p->tt * p->tscrunch > cap) break;
            ++n;
        }
        p->tiles.push_back(make_int2(d, n));
        int rows = 0;
        for(int f = 0; f < F; ++f) {
            int lo = INT32_MAX, hi = INT32_MIN;
            for(int k = 0; k < n; ++k) {
                lo = std::min(lo, p->delays[(size_t)(d + k) * F + f]);
                hi = std::max(hi, p->delays[(size_t)(d + k) * F + f]);
            }
            span.push_back(make_int2(lo, hi));
            rows = std::max(rows, hi - lo + p->tt * p->tscrunch);
        }
        p->rows_cap = std::max(p->rows_cap, rows);
        d += n;
    }
    SW_HIP_TRY(hipMalloc(&p->d_delays, sizeof(int) * p->delays.size()));
    SW_HIP_TRY(hipMalloc(&p->d_span, sizeof(int2) * span.size()));
    SW_HIP_TRY(hipMalloc(&p->d_tiles, sizeof(int2) * p->tiles.size()));
    SW_HIP_TRY(hipMemcpy(p->d_delays, p->delays.data(), sizeof(int) * p->delays.size(), hipMemcpyHostToDevice));
    SW_HIP_TRY(hipMemcpy(p->d_span, span.data(), sizeof(int2) * span.size(), hipMemcpyHostToDevice));
    SW_HIP_TRY(hipMemcpy(p->d_tiles, p->tiles.data(), sizeof(int2) * p->tiles.size(), hipMemcpyHostToDevice));
    *out = p.release();
    return SW_OK;
}

extern "C" int sw_idisp_destroy(sw_idisp* p)
{
    if(p) {
        sw::free_plan_device(p);
        delete p;
    }
    return SW_OK;
}

extern "C" const char* sw_idisp_kernel(const sw_idisp* p)
{
    if(!p) return "";
    if(!p->mfma) return "idisp_kernel";
    return p->tt == 32 ? "idisp_dma_kernel" : "idisp_mfma_kernel";
}

extern "C" int sw_idisp_info(const sw_idisp* p, int* max_delay, int* h_delays, float* scale)
{
    SW_CHECK_ARG(p, "null plan");
    if(max_delay) *max_delay = p->max_delay;
    if(h_delays) std::memcpy(h_delays, p->delays.data(), sizeof(int) * p->delays.size());
    if(scale) *scale = p->scale;
    return SW_OK;
}

namespace sw {
int idisp_launch(sw_idisp* p, const void* d_in, size_t nsamples, int nbeams, int nvec, void* d_tdb, bool ftb,
                 hipStream_t st);
}

extern "C" int sw_idisp_dedisperse(sw_idisp* p, const void* d_tfb, size_t nsamples, int nbeams, int nvec,
                                   void* d_tdb, void* stream)
{
    return sw::idisp_launch(p, d_tfb, nsamples, nbeams, nvec, d_tdb, false, sw::as_stream(stream));
}

extern "C" int sw_idisp_dedisperse_ftb(sw_idisp* p, const void* d_ftb, size_t nsamples, int nbeams, int nvec,
                                       void* d_tdb, void* stream)
{
    return sw::idisp_launch(p, d_ftb, nsamples, nbeams, nvec, d_tdb, true, sw::as_stream(stream));
}

int sw::idisp_launch(sw_idisp* p, const void* d_tfb, size_t nsamples, int nbeams, int nvec, void* d_tdb, bool ftb,
                     hipStream_t st)
{
    SW_CHECK_ARG(p && d_tfb && d_tdb, "null argument");
    SW_CHECK_ARG(nbeams > 0 && (nvec == 1 || nvec == 4), "nvec must be 1 (int8) or 4 (char4)");
    if(nsamples <= (size_t)p->max_delay) {
        sw::set_error("Fewer than max_sample_delay samples passed to dedisperse method");
        return SW_ERR_INVALID;
    }
    if((nsamples - p->max_delay) % p->tscrunch != 0) {
        sw::set_error("(nsamples - max_sample_delay) must be a multiple of tscrunch;");
        return SW_ERR_INVALID;
    }
    const long nbv = (long)nbeams * nvec;
    SW_CHECK_ARG(nbv % 4 == 0, "nbeams * nvec must be a multiple of 4");
    SW_CHECK_ARG(((uintptr_t)d_tfb & 15) == 0 && ((uintptr_t)d_tdb & 3) == 0, "misaligned buffer");
    const int nwords = (int)(nbv / 4);
    const int nout = (int)((nsamples - p->max_delay) / p->tscrunch);
    const int nchunks = (nwords + 63) / 64;
    const int ntb = (nout + p->tt - 1) / p->tt;
    const size_t lds = (size_t)p->rows_cap * 256;
    const long stride_t = ftb ? nwords : (long)p->nchans * nwords;
    const long stride_f = ftb ? (long)nsamples * nwords : nwords;
    if(p->mfma) {
        const long nchunks128 = (nbv + 127) / 128;
        const long ntb128 = (nout + p->tt - 1) / p->tt;
        const long ntiles = (long)p->tiles.size();
        const long nwork = (ntiles + sw::kIdTileGroup - 1) / sw::kIdTileGroup * sw::kIdTileGroup * nchunks128 *
                           ((ntb128 + sw::kIdTbGroup - 1) / sw::kIdTbGroup * sw::kIdTbGroup);
        const long max_launch = 8L << 20;  // work items per launch (grid x threads < 2^32)
        const bool v16 = nbv % 16 == 0 && stride_t % 4 == 0 && stride_f % 4 == 0;  // strides in words
        for(long w0 = 0; w0 < nwork; w0 += max_launch) {
            const long nw = std::min(max_launch, nwork - w0);
            const unsigned nblk = (unsigned)(8 * ((nw + 7) / 8));
#define SW_IDM_LAUNCH(NTT, V16)                                                                              \
    hipLaunchKernelGGL((sw::idisp_mfma_kernel<sw::kMfWaves, sw::kMfDpw, NTT, sw::kMfRows, V16, NTT == 2>), dim3(nblk), \
                       dim3(sw::kMfWaves * 64), 0, st, static_cast<const int8_t*>(d_tfb),                   \
                       static_cast<int8_t*>(d_tdb), p->d_delays, p->d_span, p->d_tiles, (int)ntiles,          \
                       (int)nchunks128, (int)ntb128, w0, nw, p->nchans, (int)nbv, p->ndms, p->tscrunch, p->max_delay,      \
                       (long)nsamples, nout, p->scale, stride_t * 4, stride_f * 4)
            if(p->tt == 32) {
                if(v16)
                    hipLaunchKernelGGL((sw::idisp_dma_kernel<sw::kDmaWaves, 2, sw::kMfRows>), dim3(nblk),
                                       dim3(sw::kDmaWaves * 64), 0, st, static_cast<const int8_t*>(d_tfb),
                                       static_cast<int8_t*>(d_tdb), p->d_delays, p->d_span, p->d_tiles,
                                       (int)ntiles, (int)nchunks128, (int)ntb128, w0, nw, p->nchans, (int)nbv,
                                       p->ndms, p->tscrunch, p->max_delay, (long)nsamples, nout, p->scale,
                                       stride_t * 4, stride_f * 4);
                else
                    SW_IDM_LAUNCH(2, false);
            } else {
                if(v16)
                    SW_IDM_LAUNCH(1, true);
                else
                    SW_IDM_LAUNCH(1, false);
            }
#undef SW_IDM_LAUNCH
            SW_LAUNCH_CHECK();
        }
        return SW_OK;
    }
    const bool pf = p->rows_cap <= sw::kPfRows && (nwords & 3) == 0;
    for(int tb0 = 0; tb0 < ntb; tb0 += 65535) {
        const dim3 grid((unsigned)p->tiles.size(), (unsigned)nchunks, (unsigned)std::min(65535, ntb - tb0));
#define SW_ID_LAUNCH(TT, WIDE, PF)                                                                         \
    hipLaunchKernelGGL((sw::idisp_kernel<2, TT, WIDE, PF>), grid, dim3(256), lds, st,                     \
                       static_cast<const uint32_t*>(d_tfb), static_cast<uint32_t*>(d_tdb), p->d_delays,   \
                       p->d_span, p->d_tiles, p->nchans, nwords, p->ndms, p->tscrunch, p->max_delay,      \
                       (long)nsamples, nout, tb0, p->scale, stride_t, stride_f)
        if(p->wide) {
            if(p->tt == 8)
                SW_ID_LAUNCH(8, true, false);
            else
                SW_ID_LAUNCH(4, true, false);
        } else if(p->tt == 16) {
            if(pf)
                SW_ID_LAUNCH(16, false, true);
            else
                SW_ID_LAUNCH(16, false, false);
        } else {
            SW_ID_LAUNCH(4, false, false);
        }
#undef SW_ID_LAUNCH
        SW_LAUNCH_CHECK();
    }
    return SW_OK;
}
