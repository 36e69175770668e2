// statistics.hip -- per-input (f, p, a) mean / std / skew / kurtosis.
//
// Reference: StatisticsCalculator::calculate_statistics and the
// calculate_statistics kernel, cpp/skyweaver/src/StatisticsCalculator.cu:20-130 (FP64
// Welford over the re and im parts of every sample), checked against the
// host RunningStats oracle (cpp/skyweaver/test/src/StatisticsCalculatorTester.cu:25-99)
// at 1e-5 relative.
//
// The inputs are int8, so the raw power sums S1..S4 of each column are
// computed exactly in integer arithmetic (one HBM read, bandwidth bound),
// and the central moments are formed once per column in FP64:
//   mean = S1/n, M2 = S2 - mean*S1, M3 = S3 - 3 mean S2 + 2 mean^2 S1,
//   M4 = S4 - 4 mean S3 + 6 mean^2 S2 - 3 mean^3 S1,
//   std = sqrt(M2/(n-1)), skew = sqrt(n) M3 / M2^1.5, kurt = n M4 / M2^2 - 3
// (n = 2T values per column).  The rounding of this FP64 evaluation is far
// below the float output precision for any int8 input.
#include "sw_common.hpp"

namespace sw {
namespace {

constexpr int kThreads = 256;

// Partial sums over the rows of one time chunk of one channel.
// part layout: [chunk][f][col][4] int64.
__global__ void __launch_bounds__(kThreads)
stats_partial(const uint4* __restrict__ ftpa, long long* __restrict__ part,
              int nchans, int nsamples, int pa, int rows_per_chunk)
{
    const int chunk = blockIdx.x;
    const int f     = blockIdx.y;
    const int G     = pa / 8;            // 16-byte column groups per row
    const int R     = kThreads / G;      // row lanes
    const int tid   = threadIdx.x;
    const int g     = tid % G;
    const int r     = tid / G;
    // int32 partials are flushed to int64 every 256 rows: |x+y| <= 2^8,
    // x^2+y^2 <= 2^15, |x^3+y^3| <= 2^22 per row.
    int s1[8], s2p[8], s3p[8];
    long long s2[8], s3[8], s4[8];
#pragma unroll
    for(int j = 0; j < 8; ++j) {
        s1[j] = 0; s2p[j] = 0; s3p[j] = 0; s2[j] = 0; s3[j] = 0; s4[j] = 0;
    }
    const int t0 = chunk * rows_per_chunk;
    const int t1 = min(nsamples, t0 + rows_per_chunk);
    if(r < R) {
        const uint4* base = ftpa + (size_t)f * nsamples * G + g;
        int iter = 0;
        for(int t = t0 + r; t < t1; t += R) {
            const uint4 v = base[(size_t)t * G];
            const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for(int q = 0; q < 4; ++q) {
#pragma unroll
                for(int h = 0; h < 2; ++h) {
                    const int j  = q * 2 + h;   // column within the group
                    const int re = (int)(int8_t)(w[q] >> (16 * h));
                    const int im = (int)(int8_t)(w[q] >> (16 * h + 8));
                    const int re2 = re * re, im2 = im * im;
                    s1[j] += re + im;
                    s2p[j] += re2 + im2;
                    s3p[j] += re * re2 + im * im2;
                    s4[j] += (long long)(re2 * re2 + im2 * im2);
                }
            }
            if(++iter == 256) {
#pragma unroll
                for(int j = 0; j < 8; ++j) {
                    s2[j] += s2p[j]; s2p[j] = 0;
                    s3[j] += s3p[j]; s3p[j] = 0;
                }
                iter = 0;
            }
        }
    }
#pragma unroll
    for(int j = 0; j < 8; ++j) { s2[j] += s2p[j]; s3[j] += s3p[j]; }

    // block reduction over the R row lanes, one moment at a time
    __shared__ long long red[kThreads * 8];
    long long* out = part + (((size_t)chunk * nchans + f) * pa + (size_t)g * 8) * 4;
#pragma unroll
    for(int m = 0; m < 4; ++m) {
#pragma unroll
        for(int j = 0; j < 8; ++j) {
            long long v = (m == 0) ? s1[j] : (m == 1) ? s2[j] : (m == 2) ? s3[j] : s4[j];
            red[tid * 8 + j] = v;
        }
        __syncthreads();
        if(r == 0) {
#pragma unroll
            for(int j = 0; j < 8; ++j) {
                long long acc = 0;
                for(int rr = 0; rr < R; ++rr) acc += red[(rr * G + g) * 8 + j];
                out[j * 4 + m] = acc;
            }
        }
        __syncthreads();
    }
}

__global__ void stats_finalize(const long long* __restrict__ part, float4* __restrict__ stats,
                               int nchunks, int ncols_total, long long nvalues)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;  // f*pa + col
    if(c >= ncols_total) return;
    long long S1 = 0, S2 = 0, S3 = 0, S4 = 0;
    for(int k = 0; k < nchunks; ++k) {
        const long long* p = part + ((size_t)k * ncols_total + c) * 4;
        S1 += p[0]; S2 += p[1]; S3 += p[2]; S4 += p[3];
    }
    const double n    = (double)nvalues;
    const double d1   = (double)S1, d2 = (double)S2, d3 = (double)S3, d4 = (double)S4;
    const double mean = d1 / n;
    const double M2   = d2 - mean * d1;
    const double M3   = d3 - 3.0 * mean * d2 + 2.0 * mean * mean * d1;
    const double M4   = d4 - 4.0 * mean * d3 + 6.0 * mean * mean * d2 - 3.0 * mean * mean * mean * d1;
    float4 o;
    o.x = (float)mean;
    o.y = (float)sqrt(M2 / (n - 1.0));
    o.z = (float)(sqrt(n) * M3 / pow(M2, 1.5));
    o.w = (float)(n * M4 / (M2 * M2) - 3.0);
    stats[c] = o;
}

int stats_chunks(int nchans, int nsamples)
{
    // aim for >= 2048 workgroups, chunks of >= 256 rows
    int chunks = (2048 + nchans - 1) / nchans;
    chunks = max(1, min(chunks, nsamples / 256));
    return chunks;
}

} // namespace
} // namespace sw

extern "C" size_t sw_statistics_workspace_size(int nchans, int npol, int nantennas)
{
    // upper bound over nsamples: chunks <= ceil(2048 / nchans)
    const size_t chunks = (2048 + (size_t)nchans - 1) / nchans;
    return chunks * (size_t)nchans * npol * nantennas * 4 * sizeof(long long);
}

extern "C" int sw_statistics(const void* d_ftpa, float* d_stats, int nchans, int nsamples,
                             int npol, int nantennas, void* d_workspace, void* stream)
{
    SW_CHECK_ARG(d_ftpa && d_stats, "null buffer");
    SW_CHECK_ARG(nchans > 0 && nsamples > 1 && nantennas > 0, "empty input");
    SW_CHECK_ARG(npol == 2, "only dual polarisation data is supported");
    SW_CHECK_ARG(nantennas % 4 == 0, "antenna count must be a multiple of 4");
    SW_CHECK_ARG((size_t)nsamples * 2 < (1ull << 22), "too many samples per call for exact sums");
    SW_CHECK_ARG((reinterpret_cast<uintptr_t>(d_ftpa) & 15) == 0, "input must be 16-byte aligned");
    const int pa = npol * nantennas;
    SW_CHECK_ARG(pa / 8 <= sw::kThreads, "too many antennas for one workgroup row");
    const int chunks = sw::stats_chunks(nchans, nsamples);
    const int rows   = (nsamples + chunks - 1) / chunks;
    void* ws = d_workspace ? d_workspace
                           : sw::scratch(sw_statistics_workspace_size(nchans, npol, nantennas), 0);
    if(!ws) return SW_ERR_ALLOC;
    hipLaunchKernelGGL(sw::stats_partial, dim3(chunks, nchans), dim3(sw::kThreads), 0,
                       sw::as_stream(stream), static_cast<const uint4*>(d_ftpa),
                       static_cast<long long*>(ws), nchans, nsamples, pa, rows);
    SW_LAUNCH_CHECK();
    const int ncols = nchans * pa;
    hipLaunchKernelGGL(sw::stats_finalize, dim3((ncols + 255) / 256), dim3(256), 0,
                       sw::as_stream(stream), static_cast<const long long*>(ws),
                       reinterpret_cast<float4*>(d_stats), chunks, ncols,
                       (long long)nsamples * 2);
    SW_LAUNCH_CHECK();
    return SW_OK;
}
