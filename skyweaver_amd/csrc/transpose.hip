// transpose.hip -- TAFTP -> FTPA split transpose with antenna zero padding.
//
// Reference: Transposer::transpose / transpose_k, cpp/skyweaver/src/Transposer.cu:13-184,
// oracle cpp/skyweaver/test/src/TransposerTester.cu:30-72:
//   out[f][h*256 + s][p][a] = in[h][a][f][s][p]  (a < nant_in), 0 otherwise.
//
// One workgroup per (heap h, channel f).  For each heap the input holds, per
// antenna, one contiguous 1 KiB run (256 samples x 2 pols x complex int8);
// the output for (f, h) is one contiguous run of 256 x 2 x A_out x 2 bytes.
// Antennas are staged through LDS in groups of 64 rows x 256 dwords (a dword
// = one sample's two polarisations).  The dword column is XOR-swizzled with
// 8*(row/8) so that both the 16-byte row writes and the column gathers of
// the output phase are bank-conflict free.  HBM-bound: 2*T*F*P*(A_in+A_out)
// bytes per call.
#include "sw_common.hpp"

namespace sw {
namespace {

constexpr int kHeap     = 256; // SKYWEAVER_NSAMPLES_PER_HEAP
constexpr int kGroup    = 64;  // antennas staged per LDS pass
constexpr int kThreads  = 256;

__device__ __forceinline__ int swz(int row) { return (row & ~7); }

template <int APT> // antennas per output thread: 8 -> 16-byte stores, 4 -> 8-byte
__global__ void __launch_bounds__(kThreads)
transpose_taftp_ftpa(const uint32_t* __restrict__ in, uint8_t* __restrict__ out,
                     int nant_in, int nant_out, int nchans, int nheaps)
{
    __shared__ uint32_t tile[kGroup * kHeap];
    const int h   = blockIdx.x;
    const int f   = blockIdx.y;
    const int tid = threadIdx.x;
    // input dword index of (h, a, f, s): ((h*A_in + a)*F + f)*256 + s
    const size_t in_heap = (size_t)h * nant_in * nchans;
    // output byte index of (f, t, p, a): ((f*T + t)*2 + p)*A_out*2 + a*2
    const size_t T = (size_t)nheaps * kHeap;
    uint8_t* out_fh = out + ((size_t)f * T + (size_t)h * kHeap) * 2 * nant_out * 2;

    for(int g0 = 0; g0 < nant_out; g0 += kGroup) {
        const int ga_out = min(kGroup, nant_out - g0);          // rows written
        const int ga_in  = max(0, min(kGroup, nant_in - g0));    // rows loaded
        // ---- load ga_in rows of 256 dwords, 16 bytes per thread-iteration
        for(int idx = tid; idx < ga_in * (kHeap / 4); idx += kThreads) {
            const int r = idx >> 6;          // row within group
            const int c = (idx & 63) << 2;   // first dword of the chunk
            const uint4 v = *reinterpret_cast<const uint4*>(
                in + ((in_heap + (size_t)(g0 + r) * nchans + f) * kHeap + c));
            *reinterpret_cast<uint4*>(&tile[r * kHeap + (c ^ swz(r))]) = v;
        }
        __syncthreads();
        // ---- write: thread -> (sample t, antenna group of APT)
        const int nag = ga_out / APT;
        for(int idx = tid; idx < kHeap * nag; idx += kThreads) {
            const int t  = idx / nag;
            const int g  = idx - t * nag;
            const int r0 = g * APT;
            uint32_t p0[APT / 2], p1[APT / 2];
#pragma unroll
            for(int j = 0; j < APT; j += 2) {
                const int ra = r0 + j, rb = r0 + j + 1;
                const uint32_t da = (ra < ga_in) ? tile[ra * kHeap + (t ^ swz(ra))] : 0u;
                const uint32_t db = (rb < ga_in) ? tile[rb * kHeap + (t ^ swz(rb))] : 0u;
                // da = [p0re p0im p1re p1im] of antenna ra (little-endian bytes)
                p0[j / 2] = __builtin_amdgcn_perm(db, da, 0x05040100u); // lo16(da) | lo16(db)<<16
                p1[j / 2] = __builtin_amdgcn_perm(db, da, 0x07060302u); // hi16(da) | hi16(db)<<16
            }
            uint8_t* o0 = out_fh + ((size_t)t * 2 * nant_out + g0 + r0) * 2;
            uint8_t* o1 = o0 + (size_t)nant_out * 2;
            if constexpr(APT == 8) {
                *reinterpret_cast<uint4*>(o0) = make_uint4(p0[0], p0[1], p0[2], p0[3]);
                *reinterpret_cast<uint4*>(o1) = make_uint4(p1[0], p1[1], p1[2], p1[3]);
            } else {
                *reinterpret_cast<uint2*>(o0) = make_uint2(p0[0], p0[1]);
                *reinterpret_cast<uint2*>(o1) = make_uint2(p1[0], p1[1]);
            }
        }
        __syncthreads();
    }
}

} // namespace
} // namespace sw

extern "C" int sw_transpose(const void* d_taftp, void* d_ftpa, int nheap_groups,
                            int nant_in, int nant_out, int nchans, void* stream)
{
    SW_CHECK_ARG(d_taftp && d_ftpa, "null buffer");
    SW_CHECK_ARG(nheap_groups > 0 && nchans > 0, "empty input");
    SW_CHECK_ARG(nant_in > 0 && nant_in <= nant_out,
                 "Input number of antennas must be <= to the maximum nantennas");
    SW_CHECK_ARG(nant_out % 4 == 0, "output antenna count must be a multiple of 4");
    SW_CHECK_ARG(nheap_groups <= 65535 * 1024 && nchans <= 65535, "grid too large");
    SW_CHECK_ARG((reinterpret_cast<uintptr_t>(d_taftp) & 15) == 0 &&
                     (reinterpret_cast<uintptr_t>(d_ftpa) & 15) == 0,
                 "buffers must be 16-byte aligned");
    dim3 grid(nheap_groups, nchans);
    auto* in  = static_cast<const uint32_t*>(d_taftp);
    auto* out = static_cast<uint8_t*>(d_ftpa);
    // 16-byte stores need every 64-antenna group (and the remainder) to be a
    // multiple of 8 antennas.
    if(nant_out % 8 == 0)
        hipLaunchKernelGGL(sw::transpose_taftp_ftpa<8>, grid, dim3(sw::kThreads), 0,
                           sw::as_stream(stream), in, out, nant_in, nant_out, nchans,
                           nheap_groups);
    else
        hipLaunchKernelGGL(sw::transpose_taftp_ftpa<4>, grid, dim3(sw::kThreads), 0,
                           sw::as_stream(stream), in, out, nant_in, nant_out, nchans,
                           nheap_groups);
    SW_LAUNCH_CHECK();
    return SW_OK;
}
