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

size_t fpat_stats_workspace(int nchans, int npol, int nant)
{
    return (size_t)kStatChunks * nchans * npol * nant * 4 * sizeof(long long);
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
    void* ws = d_workspace ? d_workspace : sw::scratch(sw::fpat_stats_workspace(nchans, npol, nantennas), 2);
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
