// layout.hip -- time-major voltage layout used inside the pipeline.
//
// The coherent dedisperser transforms along time, so inside the pipeline
// the voltages of a gulp are kept time-major, "FPAT": [F][P][A][T] complex
// int8 (column c = p*A + a holds T contiguous samples).  The beamformers
// consume the reference's FTPA layout ([F][T][P][A]); a tiled corner turn
// converts between the two.
//   * transpose_taftp_fpat: TAFTP [T/256][A_in][F][256][P] -> FPAT, antennas
//     >= A_in zero.  Each (heap, antenna, channel) run of 256 samples x 2
//     pols is split into two 512-byte runs: no LDS needed.
//   * corner_turn: [F][R][C] -> [F][C][R] of 16-bit elements through 64x64
//     LDS tiles (FTPA <-> FPAT).
//   * statistics on FPAT: contiguous per-column reductions, same exact
//     integer sums and FP64 finalisation as statistics.hip.
#include "sw_common.hpp"

#include <mutex>
#include <string>

namespace sw {
namespace {

__global__ void __launch_bounds__(256)
taftp_to_fpat_kernel(const uint32_t* __restrict__ in, uint8_t* __restrict__ out, int nant_in,
                     int nant_out, int nchans, int nheaps)
{
    // one thread per 4 samples (16 input bytes) of one (heap, antenna, channel)
    const size_t idx   = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t total = (size_t)nheaps * nant_out * nchans * 64;
    if(idx >= total) return;
    const int q = (int)(idx & 63);                 // 4-sample group within the heap
    size_t r    = idx >> 6;
    const int f = (int)(r % nchans);
    r /= nchans;
    const int a = (int)(r % nant_out);
    const int h = (int)(r / nant_out);
    const size_t T = (size_t)nheaps * 256;
    uint32_t p0lo = 0, p0hi = 0, p1lo = 0, p1hi = 0;
    if(a < nant_in) {
        const uint4 v = *reinterpret_cast<const uint4*>(in + (((size_t)h * nant_in + a) * nchans + f) * 256 + q * 4);
        // v.x = [p0re p0im p1re p1im] of sample 4q, etc.
        p0lo = __builtin_amdgcn_perm(v.y, v.x, 0x05040100u);
        p0hi = __builtin_amdgcn_perm(v.w, v.z, 0x05040100u);
        p1lo = __builtin_amdgcn_perm(v.y, v.x, 0x07060302u);
        p1hi = __builtin_amdgcn_perm(v.w, v.z, 0x07060302u);
    }
    const size_t t0 = (size_t)h * 256 + q * 4;
    uint8_t* o0 = out + (((size_t)f * 2 + 0) * nant_out + a) * T * 2 + t0 * 2;
    uint8_t* o1 = out + (((size_t)f * 2 + 1) * nant_out + a) * T * 2 + t0 * 2;
    *reinterpret_cast<uint2*>(o0) = make_uint2(p0lo, p0hi);
    *reinterpret_cast<uint2*>(o1) = make_uint2(p1lo, p1hi);
}

// [F][R][C] -> [F][C][R] of E (16-bit complex int8 or 8-byte complex
// float) elements, 64 x 64 tiles; source rows are ld >= C elements apart,
// destination channels R * dcc >= R * C elements apart
template <typename E>
__global__ void __launch_bounds__(256)
corner_turn_kernel(const E* __restrict__ src, E* __restrict__ dst, int R, int C, int ld, int dcc,
                   int tiles_r, int tiles_c)
{
    __shared__ E tile[64][65];
    const int bid = blockIdx.x;
    const int f   = bid / (tiles_r * tiles_c);
    const int rem = bid % (tiles_r * tiles_c);
    const int tr = rem / tiles_c, tc = rem % tiles_c;
    const int r0 = tr * 64, c0 = tc * 64;
    const E* s = src + (size_t)f * R * ld;
    E* d       = dst + (size_t)f * R * dcc;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for(int i = ty; i < 64; i += 4) {
        const int r = r0 + i, c = c0 + tx;
        if(r < R && c < C) tile[i][tx] = s[(size_t)r * ld + c];
    }
    __syncthreads();
    for(int i = ty; i < 64; i += 4) {
        const int c = c0 + i, r = r0 + tx;
        if(r < R && c < C) d[(size_t)c * R + r] = tile[tx][i];
    }
}

// As corner_turn_kernel with 8-byte global accesses (R % 4 == 0 and
// C % 4 == 0): each thread loads 4 consecutive columns of a row and stores 4
// consecutive rows of a column.
__global__ void __launch_bounds__(256)
corner_turn_x4_kernel(const uint16_t* __restrict__ src, uint16_t* __restrict__ dst, int R, int C, int ld,
                      int dcc, int tiles_r, int tiles_c)
{
    __shared__ uint16_t tile[64][68];  // 136-byte rows: 8-byte aligned, banks staggered
    const int bid = blockIdx.x;
    const int f   = bid / (tiles_r * tiles_c);
    const int rem = bid % (tiles_r * tiles_c);
    const int tr = rem / tiles_c, tc = rem % tiles_c;
    const int r0 = tr * 64, c0 = tc * 64;
    const uint16_t* s = src + (size_t)f * R * ld;
    uint16_t* d       = dst + (size_t)f * R * dcc;
    const int g = threadIdx.x & 15, q = threadIdx.x >> 4;
#pragma unroll
    for(int k = 0; k < 4; ++k) {
        const int i = q + 16 * k, r = r0 + i, c = c0 + 4 * g;
        uint2 v = make_uint2(0, 0);
        if(r < R && c < C) v = *reinterpret_cast<const uint2*>(s + (size_t)r * ld + c);
        *reinterpret_cast<uint2*>(&tile[i][4 * g]) = v;
    }
    __syncthreads();
#pragma unroll
    for(int k = 0; k < 4; ++k) {
        const int j = q + 16 * k, c = c0 + j, r = r0 + 4 * g;
        if(r < R && c < C) {
            const uint32_t lo = (uint32_t)tile[4 * g][j] | ((uint32_t)tile[4 * g + 1][j] << 16);
            const uint32_t hi = (uint32_t)tile[4 * g + 2][j] | ((uint32_t)tile[4 * g + 3][j] << 16);
            *reinterpret_cast<uint2*>(d + (size_t)c * R + r) = make_uint2(lo, hi);
        }
    }
}

// exact int64 power sums of one column chunk; part [chunk][col][4]
__global__ void __launch_bounds__(256)
fpat_stats_partial(const uint32_t* __restrict__ tm, long long* __restrict__ part, int ncols, int T,
                   int chunk_len)
{
    const int col   = blockIdx.x;
    const int chunk = blockIdx.y;
    const int t0 = chunk * chunk_len, t1 = min(T, t0 + chunk_len);
    // each dword = two complex samples
    const uint32_t* base = tm + (size_t)col * (T / 2);
    long long S1 = 0, S2 = 0, S3 = 0, S4 = 0;
    int s1 = 0, s2 = 0, s3 = 0;
    int iter = 0;
    for(int w = t0 / 2 + threadIdx.x; w < t1 / 2; w += 256) {
        const uint32_t v = base[w];
#pragma unroll
        for(int b = 0; b < 4; ++b) {
            const int x = (int)(int8_t)(v >> (8 * b));
            const int x2 = x * x;
            s1 += x;
            s2 += x2;
            s3 += x * x2;
            S4 += (long long)(x2 * x2);
        }
        if(++iter == 128) {  // 512 values: |sum x^3| <= 2^30
            S1 += s1; S2 += s2; S3 += s3;
            s1 = s2 = s3 = 0;
            iter = 0;
        }
    }
    S1 += s1; S2 += s2; S3 += s3;
    __shared__ long long red[4][256];
    red[0][threadIdx.x] = S1;
    red[1][threadIdx.x] = S2;
    red[2][threadIdx.x] = S3;
    red[3][threadIdx.x] = S4;
    __syncthreads();
    for(int st = 128; st > 0; st >>= 1) {
        if(threadIdx.x < st)
#pragma unroll
            for(int m = 0; m < 4; ++m) red[m][threadIdx.x] += red[m][threadIdx.x + st];
        __syncthreads();
    }
    if(threadIdx.x < 4) part[((size_t)chunk * ncols + col) * 4 + threadIdx.x] = red[threadIdx.x][0];
}

// TAFTP -> FPAT as taftp_to_fpat_kernel, fused with the statistics' exact
// power sums, so that the gulp is read from HBM once for both.  One wave per
// (chunk of kFusedHeaps heaps, antenna, channel): per heap it loads the
// channel's 256 samples x 2 pols (16 bytes a lane), writes the two pols'
// 512-byte runs and adds the sums of x, x^2, x^3, x^4 over re and im with
// dot-product instructions: sdot4 for x and x^2; for x^3 and x^4 the values
// are sign-extended to 16-bit pairs (v_perm), squared with a packed 16-bit
// multiply (x^2 <= 2^14 fits) and summed by sdot2 / udot2.  Bounds per lane
// and pol over one chunk: |sum x^3| <= 16 heaps x 8 values x 2^21 = 2^28 and
// a heap's x^4 sum <= 8 x 2^28 = 2^31 (unsigned), so the 32-bit partials are
// exact; the wave reduction and the x^4 total run in 64 bits.  Partials go
// to part[chunk][col][4] (col = (f*2 + p)*A + a) for fpat_stats_finalize.
// A chunk may also span up to kFusedHeapsMax heaps (|sum x^3| <= 2^30 per
// lane and pol), which the no-allocation entry point uses to fit its fixed
// scratch array.
constexpr int kFusedHeaps = 16;
constexpr int kFusedHeapsMax = 64;

typedef short sw_short2 __attribute__((ext_vector_type(2)));
typedef unsigned short sw_ushort2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ void power_sums(uint32_t x, int& s1, int& s2, int& s3, uint32_t& s4)
{
    s1 = __builtin_amdgcn_sdot4((int)x, 0x01010101, s1, false);
    s2 = __builtin_amdgcn_sdot4((int)x, (int)x, s2, false);
    // [x0, x2] and [x1, x3] as sign-extended 16-bit pairs (selector 8 / 9
    // replicate the sign bits of bytes 1 / 3)
    const sw_short2 e0 = __builtin_bit_cast(sw_short2, __builtin_amdgcn_perm(x << 8, x << 8, 0x09030801u));
    const sw_short2 e1 = __builtin_bit_cast(sw_short2, __builtin_amdgcn_perm(x, x, 0x09030801u));
    const sw_short2 q0 = e0 * e0, q1 = e1 * e1;
    s3 = __builtin_amdgcn_sdot2(q0, e0, s3, false);
    s3 = __builtin_amdgcn_sdot2(q1, e1, s3, false);
    s4 = __builtin_amdgcn_udot2(__builtin_bit_cast(sw_ushort2, q0), __builtin_bit_cast(sw_ushort2, q0), s4, false);
    s4 = __builtin_amdgcn_udot2(__builtin_bit_cast(sw_ushort2, q1), __builtin_bit_cast(sw_ushort2, q1), s4, false);
}

__device__ __forceinline__ long long wave_sum(long long v)
{
#pragma unroll
    for(int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

__global__ void __launch_bounds__(256)
taftp_to_fpat_stats_kernel(const uint4* __restrict__ in, uint2* __restrict__ out, long long* __restrict__ part,
                           int nant_in, int nant_out, int nchans, int nheaps, int nchunks, int chunk_heaps)
{
    const int lane    = threadIdx.x & 63;
    const size_t wave = (size_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if(wave >= (size_t)nchunks * nant_out * nchans) return;  // whole waves only
    const int f     = (int)(wave % nchans);
    const size_t r  = wave / nchans;
    const int a     = (int)(r % nant_out);
    const int chunk = (int)(r / nant_out);
    const int h0 = chunk * chunk_heaps, h1 = min(nheaps, h0 + chunk_heaps);
    const size_t col_words = (size_t)nheaps * 64;  // one column: T samples = T/4 uint2
    uint2* o0 = out + (((size_t)f * 2 + 0) * nant_out + a) * col_words + lane;
    uint2* o1 = out + (((size_t)f * 2 + 1) * nant_out + a) * col_words + lane;
    const uint4* src = in + ((size_t)a * nchans + f) * 64 + lane;
    const size_t heap_stride = (size_t)nant_in * nchans * 64;
    int s1[2] = {0, 0}, s2[2] = {0, 0}, s3[2] = {0, 0};
    unsigned long long s4[2] = {0, 0};
    auto heap = [&](int h, const uint4& v) {
        // v.x = [p0re p0im p1re p1im] of sample 4*lane, etc.
        const uint32_t p0lo = __builtin_amdgcn_perm(v.y, v.x, 0x05040100u);
        const uint32_t p0hi = __builtin_amdgcn_perm(v.w, v.z, 0x05040100u);
        const uint32_t p1lo = __builtin_amdgcn_perm(v.y, v.x, 0x07060302u);
        const uint32_t p1hi = __builtin_amdgcn_perm(v.w, v.z, 0x07060302u);
        o0[(size_t)h * 64] = make_uint2(p0lo, p0hi);
        o1[(size_t)h * 64] = make_uint2(p1lo, p1hi);
        uint32_t q0 = 0, q1 = 0;
        power_sums(p0lo, s1[0], s2[0], s3[0], q0);
        power_sums(p0hi, s1[0], s2[0], s3[0], q0);
        power_sums(p1lo, s1[1], s2[1], s3[1], q1);
        power_sums(p1hi, s1[1], s2[1], s3[1], q1);
        s4[0] += q0;
        s4[1] += q1;
    };
    if(a < nant_in) {  // wave-uniform
        int h = h0;
        for(; h + 4 <= h1; h += 4) {  // four loads in flight
            uint4 v[4];
#pragma unroll
            for(int k = 0; k < 4; ++k) v[k] = src[(size_t)(h + k) * heap_stride];
#pragma unroll
            for(int k = 0; k < 4; ++k) heap(h + k, v[k]);
        }
        for(; h < h1; ++h) heap(h, src[(size_t)h * heap_stride]);
    } else {  // padded antenna: zeros (and zero sums)
        for(int h = h0; h < h1; ++h) {
            o0[(size_t)h * 64] = make_uint2(0, 0);
            o1[(size_t)h * 64] = make_uint2(0, 0);
        }
    }
    const size_t ncols = (size_t)nchans * 2 * nant_out;
#pragma unroll
    for(int p = 0; p < 2; ++p) {
        const long long t1 = wave_sum(s1[p]), t2 = wave_sum(s2[p]), t3 = wave_sum(s3[p]);
        const long long t4 = wave_sum((long long)s4[p]);
        if(lane == 0) {
            long long* d = part + ((size_t)chunk * ncols + ((size_t)f * 2 + p) * nant_out + a) * 4;
            d[0] = t1;
            d[1] = t2;
            d[2] = t3;
            d[3] = t4;
        }
    }
}

__global__ void fpat_stats_finalize(const long long* __restrict__ part, float4* __restrict__ stats,
                                    int nchunks, int ncols, long long nvalues)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if(c >= ncols) return;
    long long S1 = 0, S2 = 0, S3 = 0, S4 = 0;
    for(int k = 0; k < nchunks; ++k) {
        const long long* p = part + ((size_t)k * ncols + c) * 4;
        S1 += p[0]; S2 += p[1]; S3 += p[2]; S4 += p[3];
    }
    const double n = (double)nvalues;
    const double d1 = (double)S1, d2 = (double)S2, d3 = (double)S3, d4 = (double)S4;
    const double mean = d1 / n;
    const double M2 = d2 - mean * d1;
    const double M3 = d3 - 3.0 * mean * d2 + 2.0 * mean * mean * d1;
    const double M4 = d4 - 4.0 * mean * d3 + 6.0 * mean * mean * d2 - 3.0 * mean * mean * mean * d1;
    float4 o;
    o.x = (float)mean;
    o.y = (float)sqrt(M2 / (n - 1.0));
    o.z = (float)(sqrt(n) * M3 / pow(M2, 1.5));
    o.w = (float)(n * M4 / (M2 * M2) - 3.0);
    stats[c] = o;
}

constexpr int kStatChunks = 8;

} // namespace

int corner_turn(const void* src, void* dst, int nchans, int R, int C, hipStream_t st, int src_ld,
                int dst_chan_cols)
{
    const int ld = src_ld > 0 ? src_ld : C;
    const int dcc = dst_chan_cols > 0 ? dst_chan_cols : C;
    const int tr = (R + 63) / 64, tc = (C + 63) / 64;
    const bool x4 = R % 4 == 0 && C % 4 == 0 && ld % 4 == 0 && ((uintptr_t)src & 7) == 0 &&
                    ((uintptr_t)dst & 7) == 0;
    hipLaunchKernelGGL(x4 ? corner_turn_x4_kernel : corner_turn_kernel<uint16_t>,
                       dim3((unsigned)((size_t)nchans * tr * tc)), dim3(256), 0, st,
                       static_cast<const uint16_t*>(src), static_cast<uint16_t*>(dst), R, C, ld, dcc, tr, tc);
    SW_LAUNCH_CHECK();
    return SW_OK;
}

int corner_turn_f64(const void* src, void* dst, int nchans, int R, int C, hipStream_t st, int src_ld,
                    int dst_chan_cols)
{
    const int ld = src_ld > 0 ? src_ld : C;
    const int dcc = dst_chan_cols > 0 ? dst_chan_cols : C;
    const int tr = (R + 63) / 64, tc = (C + 63) / 64;
    hipLaunchKernelGGL(corner_turn_kernel<float2>, dim3((unsigned)((size_t)nchans * tr * tc)), dim3(256), 0, st,
                       static_cast<const float2*>(src), static_cast<float2*>(dst), R, C, ld, dcc, tr, tc);
    SW_LAUNCH_CHECK();
    return SW_OK;
}

size_t fpat_stats_workspace(int nchans, int npol, int nant, int nsamples)
{
    const int fused_chunks = (nsamples / 256 + kFusedHeaps - 1) / kFusedHeaps;
    return (size_t)std::max(kStatChunks, fused_chunks) * nchans * npol * nant * 4 * sizeof(long long);
}

} // namespace sw

extern "C" int sw_transpose_fpat(const void* d_taftp, void* d_fpat, int nheap_groups, int nant_in,
                                 int nant_out, int nchans, void* stream)
{
    SW_CHECK_ARG(d_taftp && d_fpat, "null buffer");
    SW_CHECK_ARG(nheap_groups > 0 && nchans > 0, "empty input");
    SW_CHECK_ARG(nant_in > 0 && nant_in <= nant_out,
                 "Input number of antennas must be <= to the maximum nantennas");
    SW_CHECK_ARG((reinterpret_cast<uintptr_t>(d_taftp) & 15) == 0 &&
                     (reinterpret_cast<uintptr_t>(d_fpat) & 7) == 0,
                 "input must be 16-byte and output 8-byte aligned");
    const size_t total = (size_t)nheap_groups * nant_out * nchans * 64;
    hipLaunchKernelGGL(sw::taftp_to_fpat_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                       sw::as_stream(stream), static_cast<const uint32_t*>(d_taftp),
                       static_cast<uint8_t*>(d_fpat), nant_in, nant_out, nchans, nheap_groups);
    SW_LAUNCH_CHECK();
    return SW_OK;
}

namespace sw {
namespace {

int check_transpose_fpat_stats(const void* d_taftp, const void* d_fpat, const float* d_stats, int nheap_groups,
                               int nant_in, int nant_out, int nchans)
{
    SW_CHECK_ARG(d_taftp && d_fpat && d_stats, "null buffer");
    SW_CHECK_ARG(nheap_groups > 0 && nchans > 0, "empty input");
    SW_CHECK_ARG(nant_in > 0 && nant_in <= nant_out,
                 "Input number of antennas must be <= to the maximum nantennas");
    SW_CHECK_ARG((reinterpret_cast<uintptr_t>(d_taftp) & 15) == 0 &&
                     (reinterpret_cast<uintptr_t>(d_fpat) & 7) == 0,
                 "input must be 16-byte and output 8-byte aligned");
    SW_CHECK_ARG((size_t)nheap_groups * 256 * 2 < (1ull << 22), "too many samples per call for exact sums");
    return SW_OK;
}

// the fused pass with power-sum partials in ws, chunks of chunk_heaps heaps
int launch_transpose_fpat_stats(const void* d_taftp, void* d_fpat, float* d_stats, int nheap_groups, int nant_in,
                                int nant_out, int nchans, void* ws, int chunk_heaps, hipStream_t st)
{
    const int nchunks = (nheap_groups + chunk_heaps - 1) / chunk_heaps;
    const size_t nwaves = (size_t)nchunks * nant_out * nchans;
    hipLaunchKernelGGL(taftp_to_fpat_stats_kernel, dim3((unsigned)((nwaves + 3) / 4)), dim3(256), 0, st,
                       static_cast<const uint4*>(d_taftp), static_cast<uint2*>(d_fpat), static_cast<long long*>(ws),
                       nant_in, nant_out, nchans, nheap_groups, nchunks, chunk_heaps);
    SW_LAUNCH_CHECK();
    const int ncols = nchans * 2 * nant_out;
    hipLaunchKernelGGL(fpat_stats_finalize, dim3((ncols + 255) / 256), dim3(256), 0, st,
                       static_cast<const long long*>(ws), reinterpret_cast<float4*>(d_stats), nchunks, ncols,
                       (long long)nheap_groups * 256 * 2);
    SW_LAUNCH_CHECK();
    return SW_OK;
}

// Scratch of sw_transpose_fpat_stats_noalloc: the power-sum partials
// [chunk][col][4] int64.  One copy per GPU, held until the process ends;
// kernels get its address from hipGetSymbolAddress (noalloc_scratch_ptr),
// never from the host-side name.
constexpr size_t kNoallocScratchBytes = (size_t)16 << 20;
__device__ __align__(256) long long noalloc_scratch[kNoallocScratchBytes / sizeof(long long)];

constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
void* g_noalloc_ptr[kMaxDevices] = {};

// the scratch array's device address on the current device (looked up once per device)
int noalloc_scratch_ptr(void** out)
{
    int dev = 0;
    SW_HIP_TRY(hipGetDevice(&dev));
    SW_CHECK_ARG(dev >= 0 && dev < kMaxDevices, "device index out of range");
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if(!g_noalloc_ptr[dev]) {
        void* p = nullptr;
        SW_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(noalloc_scratch)));
        SW_CHECK_ARG(p != nullptr, "hipGetSymbolAddress returned a null address");
        g_noalloc_ptr[dev] = p;
    }
    *out = g_noalloc_ptr[dev];
    return SW_OK;
}

} // namespace
} // namespace sw

extern "C" int sw_transpose_fpat_stats(const void* d_taftp, void* d_fpat, float* d_stats, int nheap_groups,
                                      int nant_in, int nant_out, int nchans, void* d_workspace, void* stream)
{
    int rc = sw::check_transpose_fpat_stats(d_taftp, d_fpat, d_stats, nheap_groups, nant_in, nant_out, nchans);
    if(rc) return rc;
    const int nsamples = nheap_groups * 256;
    void* ws = d_workspace ? d_workspace : sw::scratch(sw::fpat_stats_workspace(nchans, 2, nant_out, nsamples), 2);
    if(!ws) return SW_ERR_ALLOC;
    return sw::launch_transpose_fpat_stats(d_taftp, d_fpat, d_stats, nheap_groups, nant_in, nant_out, nchans, ws,
                                           sw::kFusedHeaps, sw::as_stream(stream));
}

extern "C" int sw_transpose_fpat_stats_noalloc(const void* d_taftp, void* d_fpat, float* d_stats, int nheap_groups,
                                              int nant_in, int nant_out, int nchans, void* stream)
{
    int rc = sw::check_transpose_fpat_stats(d_taftp, d_fpat, d_stats, nheap_groups, nant_in, nant_out, nchans);
    if(rc) return rc;
    // the shortest chunks (most parallel waves) whose partials fit the scratch
    const size_t ncols = (size_t)nchans * 2 * nant_out;
    int chunk_heaps = 0;
    for(int h = sw::kFusedHeaps; h <= sw::kFusedHeapsMax; h *= 2) {
        const size_t nchunks = (size_t)(nheap_groups + h - 1) / h;
        if(nchunks * ncols * 4 * sizeof(long long) <= sw::kNoallocScratchBytes) {
            chunk_heaps = h;
            break;
        }
    }
    if(chunk_heaps == 0) {
        const size_t most = sw::kNoallocScratchBytes /
                            ((size_t)(nheap_groups + sw::kFusedHeapsMax - 1) / sw::kFusedHeapsMax * 4 * sizeof(long long));
        sw::set_error("sw_transpose_fpat_stats_noalloc: " + std::to_string(ncols) + " columns (nchans * 2 * nant_out) of " +
                      std::to_string(nheap_groups) + " heaps exceed its " +
                      std::to_string(sw::kNoallocScratchBytes >> 20) + " MiB scratch, which holds at most " +
                      std::to_string(most) + " columns at this gulp length");
        return SW_ERR_INVALID;
    }
    void* ws = nullptr;
    rc = sw::noalloc_scratch_ptr(&ws);
    if(rc) return rc;
    return sw::launch_transpose_fpat_stats(d_taftp, d_fpat, d_stats, nheap_groups, nant_in, nant_out, nchans, ws,
                                           chunk_heaps, sw::as_stream(stream));
}

extern "C" int sw_corner_turn(const void* d_src, void* d_dst, int nchans, int rows, int cols,
                              void* stream)
{
    SW_CHECK_ARG(d_src && d_dst && d_src != d_dst, "null or aliased buffer");
    SW_CHECK_ARG(nchans > 0 && rows > 0 && cols > 0, "empty input");
    return sw::corner_turn(d_src, d_dst, nchans, rows, cols, sw::as_stream(stream), 0, 0);
}

extern "C" int sw_statistics_fpat(const void* d_fpat, float* d_stats, int nchans, int nsamples,
                                  int npol, int nantennas, void* d_workspace, void* stream)
{
    SW_CHECK_ARG(d_fpat && d_stats, "null buffer");
    SW_CHECK_ARG(nchans > 0 && nsamples > 1 && nsamples % 2 == 0 && nantennas > 0, "bad dimensions");
    SW_CHECK_ARG((size_t)nsamples * 2 < (1ull << 22), "too many samples per call for exact sums");
    const int ncols = nchans * npol * nantennas;
    void* ws = d_workspace ? d_workspace : sw::scratch(sw::fpat_stats_workspace(nchans, npol, nantennas, nsamples), 2);
    if(!ws) return SW_ERR_ALLOC;
    const int chunk = (nsamples / sw::kStatChunks + 1) & ~1;
    hipLaunchKernelGGL(sw::fpat_stats_partial, dim3(ncols, sw::kStatChunks), dim3(256), 0,
                       sw::as_stream(stream), static_cast<const uint32_t*>(d_fpat),
                       static_cast<long long*>(ws), ncols, nsamples, chunk);
    SW_LAUNCH_CHECK();
    hipLaunchKernelGGL(sw::fpat_stats_finalize, dim3((ncols + 255) / 256), dim3(256), 0,
                       sw::as_stream(stream), static_cast<const long long*>(ws),
                       reinterpret_cast<float4*>(d_stats), sw::kStatChunks, ncols,
                       (long long)nsamples * 2);
    SW_LAUNCH_CHECK();
    return SW_OK;
}
