// codedisp.hip -- overlap-save coherent dedispersion: FFT, chirp multiply,
// inverse FFT in FP32 with FP64-generated chirps and twiddles.
//
// Reference: CoherentDedisperser (cpp/skyweaver/src/CoherentDedisperser.cu:19-353):
// int8 -> c64, FFT, x H[dm][f][k], IFFT, /N, round half-even, wrap to int8,
// keep [O/2, N-O/2).  The chirp is DMResponse (:313-335):
//   nu0 = f_low + c*bw_c - bw_c/2, nu = k*bw_c/N for all k in [0, N),
//   phase = nu^2 (-2 pi K DM) / ((nu0 + nu) nu0^2), H = exp(i phase).
// The FFT length N and the overlap O are per-DM parameters and the stream is
// processed as true overlap-save: the input of a channel is the previous
// gulp's last O samples followed by the current gulp; segment s covers input
// [s(N-O), s(N-O)+N) (zero past the end) and yields output samples
// [s(N-O), (s+1)(N-O)) from its positions [O/2, N-O/2).  Output sample j is
// the dedispersed input sample j - O/2.
//
// Data are time-major ([F][P*A][T] complex int8, one column per (pol,
// antenna)) so that every transform reads and writes contiguous runs:
//  * N <= 16384: one workgroup per (channel, segment, group of C columns)
//    keeps C x N points in LDS: load -> forward FFT -> per DM of the launch:
//    chirp, inverse FFT, quantise, store.  With two LDS buffers (C*N <= 8192)
//    the forward FFT is shared by every DM of the launch; with one (N = 16384)
//    it is recomputed per DM.
//  * N > 16384 = RP x L: a radix-RP decimation-in-frequency stage, the LDS
//    kernel on each of the RP sub-sequences (chirp permuted to match) and a
//    radix-RP decimation-in-time stage that quantises.
// Outputs are corner-turned to FTPA for the beamformers.
#include "fft.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <memory>
#include <vector>

namespace sw {
int corner_turn(const void* src, void* dst, int nchans, int R, int C, hipStream_t st, int src_ld = 0);
int corner_turn_f64(const void* src, void* dst, int nchans, int R, int C, hipStream_t st, int src_ld = 0);

namespace {

using namespace fft;

constexpr int kMaxL  = 16384;   // longest single-workgroup transform
constexpr int kMaxDm = 8;       // DMs per launch
constexpr double kDmConst = 4.148806423e15;   // Hz^2 pc^-1 cm^3 s

struct CdParams {
    const int8_t* cur;   // [F][PA][T] complex int8
    const int8_t* hist;  // [F][PA][Oh]
    float2* z;           // [F*nseg][RP][PA][L] (N > kMaxL)
    int T, O, Oh, hop, nseg, PA, F, RP, N;
    long in0;   // input index (relative to the gulp start) of segment 0's first sample
    long y0;    // output-buffer index of segment 0's first kept sample
    int ylen;   // output-buffer samples per column
    int ndm;
    const float2* chirp[kMaxDm];  // [F][RP][L] (RP == 1: natural order)
    int8_t* out[kMaxDm];          // [F][PA][ylen] complex int8
    float2* outf[kMaxDm];         // optional [F][PA][ylen] y/N
    float2* v[kMaxDm];            // [F*nseg][RP][PA][L] (N > kMaxL)
    const float2* tw;             // exp(-2 pi i m / L)
    const float2* twN;            // exp(-2 pi i m / N)
};

__device__ __forceinline__ float2 load_c8(const int8_t* p)
{
    const uint16_t v = *reinterpret_cast<const uint16_t*>(p);
    return make_float2((float)(int8_t)(v & 0xff), (float)(int8_t)(v >> 8));
}

__device__ __forceinline__ void store_c8(int8_t* p, float2 y)
{
    // __float2int_rn then static_cast<char>: round half even, wrap modulo 2^8
    const int re = (int)rintf(y.x), im = (int)rintf(y.y);
    *reinterpret_cast<uint16_t*>(p) = (uint16_t)((re & 0xff) | ((im & 0xff) << 8));
}

// input sample tau (relative to the gulp start) of column col of channel f
__device__ __forceinline__ float2 input_sample(const CdParams& p, int f, int col, int tau)
{
    if(tau < 0) return load_c8(p.hist + (((size_t)f * p.PA + col) * p.Oh + p.Oh + tau) * 2);
    if(tau < p.T) return load_c8(p.cur + (((size_t)f * p.PA + col) * p.T + tau) * 2);
    return make_float2(0.0f, 0.0f);
}

// raw complex int8 sample (re | im << 8), zero outside history + gulp
__device__ __forceinline__ uint32_t input_raw(const CdParams& p, int f, int col, int tau)
{
    if(tau < 0) return *reinterpret_cast<const uint16_t*>(p.hist + (((size_t)f * p.PA + col) * p.Oh + p.Oh + tau) * 2);
    if(tau < p.T) return *reinterpret_cast<const uint16_t*>(p.cur + (((size_t)f * p.PA + col) * p.T + tau) * 2);
    return 0;
}

__device__ __forceinline__ float2 unpack_c8(uint32_t v)
{
    return make_float2((float)(int8_t)(v & 0xff), (float)(int8_t)((v >> 8) & 0xff));
}

template <int L, int C, int NBUF, int NT, bool F32>
__global__ void __launch_bounds__(NT) codedisp_tm_kernel(CdParams p)
{
    __shared__ float2 lds[NBUF * C * padded(L) + TwiddleLds<L>::SIZE];
    float2* X = lds;
    float2* W = (NBUF == 2) ? lds + C * padded(L) : lds;
    const Twiddles tw = load_twiddles<L, NT>(lds + NBUF * C * padded(L), p.tw);
    const int G  = p.PA / C;
    const int bid = blockIdx.x;
    const int u = bid / G, g = bid % G;  // unit = (channel, segment, sub-sequence)
    const int fs = u / p.RP, r = u % p.RP;
    const int f = fs / p.nseg, s = fs % p.nseg;
    const int tid = threadIdx.x;
    const float inv_n = 1.0f / (float)p.N;  // exact: N is a power of two

    auto load = [&]() {
        for(int i = tid; i < C * L; i += NT) {
            const int c = i / L, n = i % L;
            const int col = g * C + c;
            float2 v;
            if constexpr(F32)
                v = p.z[(((size_t)fs * p.RP + r) * p.PA + col) * L + n];
            else
                v = input_sample(p, f, col, (int)(p.in0 + (long)s * p.hop + n));
            X[c * padded(L) + pad(n)] = v;
        }
        __syncthreads();
        fft_lds<false, L, C, NT>(X, tw);
    };

    load();
    for(int d = 0; d < p.ndm; ++d) {
        if(NBUF == 1 && d > 0) load();
        const float2* H = p.chirp[d] + ((size_t)f * p.RP + r) * L;
        for(int i = tid; i < C * L; i += NT) {
            const int c = i / L, k = i % L;
            W[c * padded(L) + pad(k)] = cmul(X[c * padded(L) + pad(k)], H[k]);
        }
        __syncthreads();
        fft_lds<true, L, C, NT>(W, tw);
        if constexpr(F32) {
            float2* V = p.v[d];
            for(int i = tid; i < C * L; i += NT) {
                const int c = i / L, n = i % L;
                const int col = g * C + c;
                const float2 w = cconj(p.twN[r * n]);  // exp(+2 pi i r n / N)
                V[(((size_t)fs * p.RP + r) * p.PA + col) * L + n] = cmul(W[c * padded(L) + pad(n)], w);
            }
        } else {
            int8_t* out  = p.out[d];
            float2* outf = p.outf[d];
            const int half = p.O / 2;
            for(int i = tid; i < C * p.hop; i += NT) {
                const int c = i / p.hop, m = i % p.hop;
                const long j = p.y0 + (long)s * p.hop + m;
                if(j < p.ylen) {
                    const int col = g * C + c;
                    const float2 y = cscale(W[c * padded(L) + pad(half + m)], inv_n);
                    const size_t o = ((size_t)f * p.PA + col) * p.ylen + j;
                    store_c8(out + o * 2, y);
                    if(outf) outf[o] = y;
                }
            }
        }
        __syncthreads();
    }
}

// N = 16384 held in registers: 512 threads x 32 points, one column per
// workgroup.  With n = 512 n1 + 16 n2 + n3 and k = k1 + 32 k2 + 1024 k3
// (n1, n2, k1, k2 < 32; n3, k3 < 16) the forward transform is
//   DFT32 over n1 (thread t = 16 n2 + n3), x W^(t k1)
//   -> exchange -> DFT32 over n2 (thread (k1, n3)), x W^(32 n3 k2)
//   -> exchange -> DFT16 over n3 (thread (k1, k2), two per thread),
// which leaves X[k] in registers in the order the chirp is read; the
// inverse runs the same steps backwards with conjugate twiddles and ends
// with each thread holding output samples 512 n1 + t.  Four LDS exchanges
// per DM instead of a read-write round trip per radix pass.
constexpr int k16 = 16384;
constexpr int kEx2Stride = 17;  // padded row of 16 points (conflict-free reads)
constexpr int k16Lds = 1024 * kEx2Stride;

// v[k] *= w1^k for k = 1..R-1 (R = 16 or 32), powers by the product tree of
// twiddle_powers, built 16 at a time to bound register use
template <int R>
__device__ __forceinline__ void apply_powers(float2* v, float2 w1)
{
    float2 w[16];
    twiddle_powers<16>(w1, w);
#pragma unroll
    for(int k = 1; k < 16; ++k) v[k] = cmul(v[k], w[k]);
    if constexpr(R == 32) {
        const float2 w16 = cmul(w[8], w[8]);
        v[16] = cmul(v[16], w16);
#pragma unroll
        for(int k = 1; k < 16; ++k) v[16 + k] = cmul(v[16 + k], cmul(w16, w[k]));
    }
}

// One DM per launch (d): forward FFT, chirp, inverse FFT of one column per
// unit u = (channel, segment, column).  Persistent: a workgroup (one per CU,
// the exchange buffer fills the LDS) walks units u = blockIdx.x + k gridDim.x
// and loads the next unit's 32 input samples per thread while transforming
// the current one.
__global__ void __launch_bounds__(512) codedisp_16k_kernel(CdParams p, int d, int nunits)
{
    __shared__ float2 lds[k16Lds + TwiddleLds<k16>::SIZE];
    const Twiddles tw = load_twiddles<k16, 512>(lds + k16Lds, p.tw);
    const int t = threadIdx.x;
    const float inv_n = 1.0f / (float)k16;
    uint32_t nxt[16];  // two raw samples per register
    auto fetch = [&](int u) {
        const int col = u % p.PA, fs = u / p.PA;
        const int f = fs / p.nseg, s = fs % p.nseg;
#pragma unroll
        for(int i = 0; i < 16; ++i)
            nxt[i] = input_raw(p, f, col, (int)(p.in0 + (long)s * p.hop + 512 * (2 * i) + t)) |
                     (input_raw(p, f, col, (int)(p.in0 + (long)s * p.hop + 512 * (2 * i + 1) + t)) << 16);
    };
    if(blockIdx.x < nunits) fetch(blockIdx.x);
    for(int u = blockIdx.x; u < nunits; u += gridDim.x) {
    const int col = u % p.PA;
    const int fs  = u / p.PA;
    const int f = fs / p.nseg, s = fs % p.nseg;
    float2 v[32];

    // ---- forward, step 1: thread t = (n2, n3) holds n1 = 0..31
#pragma unroll
    for(int i = 0; i < 16; ++i) {
        v[2 * i]     = unpack_c8(nxt[i]);
        v[2 * i + 1] = unpack_c8(nxt[i] >> 16);
    }
    if(u + (int)gridDim.x < nunits) fetch(u + gridDim.x);  // in flight during the transforms
    __syncthreads();  // twiddle tables; previous unit's readers of the exchange buffer
    dft32<false>(v);
    apply_powers<32>(v, tw(t));
#pragma unroll
    for(int k1 = 0; k1 < 32; ++k1) lds[k1 * 512 + t] = v[k1];
    __syncthreads();
    // ---- step 2: thread (k1, n3) holds n2 = 0..31
    {
        const int k1 = t >> 4, n3 = t & 15;
#pragma unroll
        for(int n2 = 0; n2 < 32; ++n2) v[n2] = lds[k1 * 512 + n2 * 16 + n3];
        __syncthreads();
        dft32<false>(v);
        apply_powers<32>(v, tw(32 * n3));
#pragma unroll
        for(int k2 = 0; k2 < 32; ++k2) lds[(k1 * 32 + k2) * kEx2Stride + n3] = v[k2];
    }
    __syncthreads();
    // ---- step 3: pairs q = t, t + 512 (q = 32 k2 + k1): DFT16 over n3 -> k3,
    // chirp at k = q + 1024 k3, then the inverse step 3': IDFT16 -> n3,
    // x W^(-32 k2 n3)
    const float2* H = p.chirp[d] + (size_t)f * k16;
#pragma unroll
    for(int h = 0; h < 2; ++h) {
        const int q = t + 512 * h;
        const int k1 = q & 31, k2 = q >> 5;
        float2* x = v + 16 * h;
#pragma unroll
        for(int n3 = 0; n3 < 16; ++n3) x[n3] = lds[(k1 * 32 + k2) * kEx2Stride + n3];
        dft16<false>(x);
#pragma unroll
        for(int k3 = 0; k3 < 16; ++k3) x[k3] = cmul(x[k3], H[q + 1024 * k3]);
        dft16<true>(x);
        apply_powers<16>(x, cconj(tw(32 * k2)));
    }
    __syncthreads();  // all step-3 reads done before the exchange buffer is rewritten
#pragma unroll
    for(int h = 0; h < 2; ++h) {
        const int q = t + 512 * h;
        const int k1 = q & 31, k2 = q >> 5;
#pragma unroll
        for(int n3 = 0; n3 < 16; ++n3) lds[(k1 * 32 + k2) * kEx2Stride + n3] = v[16 * h + n3];
    }
    __syncthreads();
    // ---- step 2': thread (k1, n3) holds k2 = 0..31; IDFT32 -> n2, x W^(-k1 (16 n2 + n3))
    {
        const int k1 = t >> 4, n3 = t & 15;
#pragma unroll
        for(int k2 = 0; k2 < 32; ++k2) v[k2] = lds[(k1 * 32 + k2) * kEx2Stride + n3];
        __syncthreads();
        dft32<true>(v);
        const float2 b = cconj(tw(k1 * n3));
        apply_powers<32>(v, cconj(tw(16 * k1)));
#pragma unroll
        for(int n2 = 0; n2 < 32; ++n2) v[n2] = cmul(v[n2], b);
#pragma unroll
        for(int n2 = 0; n2 < 32; ++n2) lds[k1 * 512 + n2 * 16 + n3] = v[n2];
    }
    __syncthreads();
    // ---- step 1': thread t = (n2, n3) holds k1 = 0..31; IDFT32 -> n1
#pragma unroll
    for(int k1 = 0; k1 < 32; ++k1) v[k1] = lds[k1 * 512 + t];
    dft32<true>(v);
    int8_t* out = p.out[d];
    float2* outf = p.outf[d];
    const int half = p.O / 2;
#pragma unroll
    for(int n1 = 0; n1 < 32; ++n1) {
        const int pos = 512 * n1 + t;
        const long j = p.y0 + (long)s * p.hop + pos - half;
        if(pos >= half && pos < k16 - half && j < p.ylen) {
            const float2 y = cscale(v[n1], inv_n);
            const size_t o = ((size_t)f * p.PA + col) * p.ylen + j;
            store_c8(out + o * 2, y);
            if(outf) outf[o] = y;
        }
    }
    }
}

// Decimation-in-frequency first stage for N = RP * L:
//   Z_r[n] = exp(-2 pi i n r / N) sum_q x[n + q L] exp(-2 pi i q r / RP)
template <int RP>
__global__ void __launch_bounds__(256) codedisp_dif_kernel(CdParams p, int L)
{
    const size_t idx   = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t total = (size_t)p.F * p.nseg * p.PA * L;
    if(idx >= total) return;
    const int n    = (int)(idx % L);
    const size_t t = idx / L;
    const int col  = (int)(t % p.PA);
    const int fs   = (int)(t / p.PA);
    const int f = fs / p.nseg, s = fs % p.nseg;
    float2 v[RP];
#pragma unroll
    for(int q = 0; q < RP; ++q) v[q] = input_sample(p, f, col, (int)(p.in0 + (long)s * p.hop + n + q * L));
    dft<RP, false>(v);
#pragma unroll
    for(int rr = 0; rr < RP; ++rr)
        p.z[(((size_t)fs * RP + rr) * p.PA + col) * L + n] = rr == 0 ? v[0] : cmul(v[rr], p.twN[n * rr]);
}

// Decimation-in-time last stage: y[n + q L] = sum_r V_r[n] exp(+2 pi i q r / RP),
// then /N, quantise and keep segment positions [O/2, N - O/2).
template <int RP>
__global__ void __launch_bounds__(256) codedisp_dit_kernel(CdParams p, int L, int d)
{
    const size_t idx   = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t total = (size_t)p.F * p.nseg * p.PA * L;
    if(idx >= total) return;
    const int n    = (int)(idx % L);
    const size_t t = idx / L;
    const int col  = (int)(t % p.PA);
    const int fs   = (int)(t / p.PA);
    const int f = fs / p.nseg, s = fs % p.nseg;
    float2 v[RP];
#pragma unroll
    for(int rr = 0; rr < RP; ++rr) v[rr] = p.v[d][(((size_t)fs * RP + rr) * p.PA + col) * L + n];
    dft<RP, true>(v);
    const float inv_n = 1.0f / (float)p.N;
    const int half = p.O / 2;
#pragma unroll
    for(int q = 0; q < RP; ++q) {
        const int pos = n + q * L;
        if(pos >= half && pos < p.N - half) {
            const long j = p.y0 + (long)s * p.hop + pos - half;
            if(j < p.ylen) {
                const float2 y = cscale(v[q], inv_n);
                const size_t o = ((size_t)f * p.PA + col) * p.ylen + j;
                store_c8(p.out[d] + o * 2, y);
                if(p.outf[d]) p.outf[d][o] = y;
            }
        }
    }
}

// exp(-2 pi i m / n) in FP64, rounded to FP32
__global__ void twiddle_kernel(float2* tw, int n)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if(m >= n) return;
    double s, c;
    sincos(-2.0 * 3.14159265358979323846 * (double)m / (double)n, &s, &c);
    tw[m] = make_float2((float)c, (float)s);
}

// DMResponse in FP64 (the only FP64 step), stored as [f][r][m] for k = RP m + r
__global__ void chirp_kernel(float2* out, int nchans, int N, int RP, double low_freq,
                             double chan_bw, double dm_prefactor)
{
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(idx >= (size_t)nchans * N) return;
    const int chan = (int)(idx / N), k = (int)(idx % N);
    const double fine_bw = chan_bw / N;
    const double nu_0 = low_freq + chan * chan_bw - 0.5 * chan_bw;
    const double nu   = k * fine_bw;
    const double phase_prefactor = nu * nu * dm_prefactor;
    const double phase = phase_prefactor / ((nu_0 + nu) * nu_0 * nu_0);
    double s, c;
    sincos(phase, &s, &c);
    const int L = N / RP, m = k / RP, r = k % RP;
    out[((size_t)chan * RP + r) * L + m] = make_float2((float)c, (float)s);
}

// new history = the last Oh samples of (old history ++ gulp), per column
__global__ void history_kernel(int8_t* dst, const int8_t* old, const int8_t* cur, int ncols, int T, int Oh)
{
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(idx >= (size_t)ncols * Oh) return;
    const int i = (int)(idx % Oh);
    const size_t col = idx / Oh;
    const long q = (long)i + T;  // position in old ++ gulp
    const uint16_t v = q < Oh ? *reinterpret_cast<const uint16_t*>(old + (col * Oh + q) * 2)
                              : *reinterpret_cast<const uint16_t*>(cur + (col * T + (q - Oh)) * 2);
    *reinterpret_cast<uint16_t*>(dst + idx * 2) = v;
}

// dst[col][i] = src[col][off + i] for i < n (row pitches sld, dld): carries
// outputs computed ahead of the emitted window to the front of the next
// gulp's buffer; when the regions overlap (latency > T) the copy goes
// through a scratch buffer in two steps
template <typename E>
__global__ void carry_kernel(E* dst, const E* src, int ncols, int dld, int sld, int off, int n)
{
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(idx >= (size_t)ncols * n) return;
    const int i = (int)(idx % n);
    const size_t col = idx / n;
    dst[col * dld + i] = src[col * sld + off + i];
}

template <typename E>
int carry(E* y, E* scratch, int ncols, int ylen, int T, int lat, hipStream_t st)
{
    const size_t n = (size_t)ncols * lat;
    const dim3 grid((unsigned)((n + 255) / 256));
    if(lat <= T) {
        hipLaunchKernelGGL(carry_kernel<E>, grid, dim3(256), 0, st, y, y, ncols, ylen, ylen, T, lat);
    } else {
        hipLaunchKernelGGL(carry_kernel<E>, grid, dim3(256), 0, st, scratch, y, ncols, lat, ylen, T, lat);
        hipLaunchKernelGGL(carry_kernel<E>, grid, dim3(256), 0, st, y, scratch, ncols, ylen, lat, 0, lat);
    }
    SW_LAUNCH_CHECK();
    return SW_OK;
}

// ------------------------------------------------------------------ host
struct LdsShape {
    int C, NBUF, NT;
};

LdsShape lds_shape(int L)
{
    if(L >= 16384) return {1, 1, 512};
    return {8192 / L, 2, 512};
}

// one persistent workgroup per CU of the current device
int persistent_blocks()
{
    int dev = 0, ncu = 0;
    if(hipGetDevice(&dev) != hipSuccess ||
       hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu <= 0)
        return 256;
    return ncu;
}

template <bool F32>
int launch_lds(int L, const CdParams& p, int blocks, hipStream_t st)
{
#define SW_CD(LL, CC, NB, NTT)                                                               \
    case LL:                                                                                 \
        hipLaunchKernelGGL((codedisp_tm_kernel<LL, CC, NB, NTT, F32>), dim3(blocks), dim3(NTT), \
                           0, st, p);                                                        \
        break;
    switch(L) {
        SW_CD(64, 128, 2, 512) SW_CD(128, 64, 2, 512) SW_CD(256, 32, 2, 512)
        SW_CD(512, 16, 2, 512) SW_CD(1024, 8, 2, 512) SW_CD(2048, 4, 2, 512)
        SW_CD(4096, 2, 2, 512) SW_CD(8192, 1, 2, 512)
    case 16384:
        if constexpr(F32)
            hipLaunchKernelGGL((codedisp_tm_kernel<16384, 1, 1, 512, true>), dim3(blocks), dim3(512), 0, st, p);
        else
            for(int d = 0; d < p.ndm; ++d)
                hipLaunchKernelGGL(codedisp_16k_kernel, dim3(std::min(blocks, persistent_blocks())), dim3(512), 0,
                                   st, p, d, blocks);
        break;
    default: set_error("codedisp: unsupported transform length"); return SW_ERR_UNSUPPORTED;
    }
#undef SW_CD
    SW_LAUNCH_CHECK();
    return SW_OK;
}

} // namespace

struct CodedispGroup {
    int N, O, L, RP, hop, nseg;  // nseg: most segments one gulp can need
    int lat = 0;                 // output latency (samples): 0 if hop divides T, else hop
    long long mdone = 0;         // segments of the stream transformed so far
    std::vector<int> dms;        // indices into the plan's DM list
};

} // namespace sw

struct sw_codedisp {
    int F, A, PA, T;
    int input_fpat;
    double low_freq, chan_bw, tsamp;
    std::vector<float> dms;
    std::vector<int> N, O;
    int Omax = 0;
    std::vector<sw::CodedispGroup> groups;
    std::vector<float2*> chirp;          // per DM
    std::map<int, float2*> twiddles;     // per length
    int Oh = 0;                          // history samples per column
    int8_t* hist = nullptr;              // [F][PA][Oh] complex int8
    int8_t* hist2 = nullptr;             // the other history buffer (ping-pong)
    int8_t* xtm  = nullptr;              // time-major copy of an FTPA input
    std::vector<int8_t*> ytm;            // per DM: [F][PA][T + lat] time-major outputs
    std::vector<float2*> ytf;            // per DM, on demand: float outputs, same layout
    std::vector<int> ylen;               // per DM: T + lat
    std::vector<int> lat;                // per DM output latency
    long long gulps = 0;                 // gulps completed (push_history calls)
    float2* zbuf = nullptr;              // DIF stage output (N > kMaxL)
    std::vector<float2*> vbuf;           // per DM slot (N > kMaxL)
    void* carry_tmp = nullptr;           // carries with latency > T
    ~sw_codedisp()
    {
        (void)hipFree(carry_tmp);
        for(auto* c : chirp) (void)hipFree(c);
        for(auto& kv : twiddles) (void)hipFree(kv.second);
        (void)hipFree(hist);
        (void)hipFree(hist2);
        (void)hipFree(xtm);
        for(auto* y : ytm) (void)hipFree(y);
        for(auto* y : ytf) (void)hipFree(y);
        (void)hipFree(zbuf);
        for(auto* v : vbuf) (void)hipFree(v);
    }
};

namespace sw {
namespace {

int even_ceil(double x)
{
    int v = (int)std::ceil(x - 1e-9);
    if(v < 0) v = 0;
    return v + (v & 1);
}

// Intra-channel dispersion smear at the bottom of the band, in samples:
// dm_delay(f_low, f_low + bw_c, DM) / tsamp (cpp/skyweaver/dedispersion_utils.cuh:20-23).
double smear_samples(double f_low, double chan_bw, double tsamp, double dm)
{
    const double f1 = f_low, f2 = f_low + chan_bw;
    return kDmConst * (1.0 / (f1 * f1) - 1.0 / (f2 * f2)) * dm / tsamp;
}

float2* twiddle_table(sw_codedisp* p, int n, hipStream_t st)
{
    auto it = p->twiddles.find(n);
    if(it != p->twiddles.end()) return it->second;
    float2* t = nullptr;
    if(hipMalloc(&t, sizeof(float2) * n) != hipSuccess) return nullptr;
    hipLaunchKernelGGL(twiddle_kernel, dim3((n + 255) / 256), dim3(256), 0, st, t, n);
    p->twiddles[n] = t;
    return t;
}

} // namespace
} // namespace sw

extern "C" int sw_codedisp_create(const sw_codedisp_config* cfg, void* stream, sw_codedisp** out)
{
    SW_CHECK_ARG(cfg && out, "null argument");
    *out = nullptr;
    SW_CHECK_ARG(cfg->nchans > 0 && cfg->nantennas > 0 && cfg->nsamples > 0, "empty dimensions");
    SW_CHECK_ARG(cfg->ndms > 0 && cfg->dms, "No coherent DMs specified");
    SW_CHECK_ARG(cfg->chan_bw > 0 && cfg->tsamp > 0 && cfg->low_freq > 0, "bad band description");
    hipStream_t st = sw::as_stream(stream);
    auto p = std::make_unique<sw_codedisp>();
    p->F = cfg->nchans;
    p->A = cfg->nantennas;
    p->PA = 2 * cfg->nantennas;
    p->T = cfg->nsamples;
    p->input_fpat = cfg->input_fpat;
    p->low_freq = cfg->low_freq;
    p->chan_bw = cfg->chan_bw;
    p->tsamp = cfg->tsamp;
    for(int d = 0; d < cfg->ndms; ++d) {
        const float dm = cfg->dms[d];
        SW_CHECK_ARG(dm >= 0.0f, "negative DM");
        int O = (cfg->overlap && cfg->overlap[d] >= 0)
                    ? cfg->overlap[d]
                    : sw::even_ceil(sw::smear_samples(cfg->low_freq, cfg->chan_bw, cfg->tsamp, dm));
        SW_CHECK_ARG(O % 2 == 0, "overlap must be even");
        int N = (cfg->fft_len && cfg->fft_len[d] > 0) ? cfg->fft_len[d] : 0;
        if(N == 0) {  // default: smallest power of two >= max(1024, 16 O)
            N = 1024;
            while(N < 16 * O) N *= 2;
        }
        SW_CHECK_ARG((N & (N - 1)) == 0 && N >= 64 && N <= 262144,
                     "FFT length must be a power of two in [64, 262144]");
        SW_CHECK_ARG(O < N, "overlap must be smaller than the FFT length");
        p->dms.push_back(dm);
        p->N.push_back(N);
        p->O.push_back(O);
        p->Omax = std::max(p->Omax, O);
    }
    // group DMs sharing (N, O): one forward transform serves the whole group
    std::map<std::pair<int, int>, int> gid;
    for(int d = 0; d < (int)p->dms.size(); ++d) {
        auto key = std::make_pair(p->N[d], p->O[d]);
        if(!gid.count(key)) {
            sw::CodedispGroup g;
            g.N = p->N[d];
            g.O = p->O[d];
            g.L = std::min(g.N, sw::kMaxL);
            g.RP = g.N / g.L;
            g.hop = g.N - g.O;
            // segments sit on one grid over the whole stream; when hop does
            // not divide the gulp, a gulp's last segment reaches into the next
            // gulp, so its outputs are emitted one hop later (latency = hop)
            g.lat = (p->T % g.hop == 0) ? 0 : g.hop;
            g.nseg = (p->T + g.hop - 1) / g.hop;
            SW_CHECK_ARG(p->PA % sw::lds_shape(g.L).C == 0,
                         "2*nantennas must be a multiple of the transform's column tile");
            gid[key] = (int)p->groups.size();
            p->groups.push_back(g);
        }
        p->groups[gid[key]].dms.push_back(d);
    }
    p->lat.assign(p->dms.size(), 0);
    p->ylen.assign(p->dms.size(), p->T);
    for(auto& g : p->groups) {
        // history: the input a gulp's first segment can reach back to
        p->Oh = std::max(p->Oh, g.lat > 0 ? g.N : g.O);
        for(int d : g.dms) {
            p->lat[d] = g.lat;
            p->ylen[d] = (p->T + g.lat + 3) / 4 * 4;  // 8-byte aligned rows for the corner turn
        }
    }
    size_t zmax = 0, slots = 1, vslots = 0;
    for(auto& g : p->groups) {
        if(!sw::twiddle_table(p.get(), g.L, st)) return SW_ERR_ALLOC;
        slots = std::max(slots, std::min<size_t>(sw::kMaxDm, g.dms.size()));
        if(g.RP > 1) {
            if(!sw::twiddle_table(p.get(), g.N, st)) return SW_ERR_ALLOC;
            zmax = std::max(zmax, (size_t)p->F * g.nseg * g.N * p->PA * sizeof(float2));
            vslots = std::max(vslots, std::min<size_t>(sw::kMaxDm, g.dms.size()));
        }
    }
    p->chirp.resize(p->dms.size(), nullptr);
    for(int d = 0; d < (int)p->dms.size(); ++d) {
        const int N = p->N[d], RP = N / std::min(N, sw::kMaxL);
        if(hipMalloc(&p->chirp[d], sizeof(float2) * (size_t)p->F * N) != hipSuccess) {
            sw::set_error("codedisp: chirp allocation failed");
            return SW_ERR_ALLOC;
        }
        // DMPrefactor: -1.0f * 2 pi * K * dm (cpp/skyweaver/dedispersion_utils.cuh:42-48)
        const double pref = -1.0f * (2 * 3.1415926535897932385) * sw::kDmConst * (double)p->dms[d];
        const size_t n = (size_t)p->F * N;
        hipLaunchKernelGGL(sw::chirp_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                           p->chirp[d], p->F, N, RP, p->low_freq, p->chan_bw, pref);
        SW_LAUNCH_CHECK();
    }
    const size_t vol = (size_t)p->F * p->PA * p->T * 2;
    if(p->Oh > 0) {
        const size_t hb = (size_t)p->F * p->PA * p->Oh * 2;
        SW_HIP_TRY(hipMalloc(&p->hist, hb));
        SW_HIP_TRY(hipMalloc(&p->hist2, hb));
        SW_HIP_TRY(hipMemsetAsync(p->hist, 0, hb, st));
    }
    if(!p->input_fpat) SW_HIP_TRY(hipMalloc(&p->xtm, vol));
    p->ytm.resize(p->dms.size(), nullptr);
    p->ytf.resize(p->dms.size(), nullptr);
    for(size_t d = 0; d < p->dms.size(); ++d) {
        const size_t yb = (size_t)p->F * p->PA * p->ylen[d] * 2;
        SW_HIP_TRY(hipMalloc(&p->ytm[d], yb));
        SW_HIP_TRY(hipMemsetAsync(p->ytm[d], 0, yb, st));  // outputs before the stream start are zero
    }
    (void)slots;
    int latmax_over_T = 0;
    for(int l: p->lat)
        if(l > p->T) latmax_over_T = std::max(latmax_over_T, l);
    if(latmax_over_T > 0)
        SW_HIP_TRY(hipMalloc(&p->carry_tmp, (size_t)p->F * p->PA * latmax_over_T * sizeof(float2)));
    if(zmax) {
        SW_HIP_TRY(hipMalloc(&p->zbuf, zmax));
        p->vbuf.resize(vslots, nullptr);
        for(auto& v : p->vbuf) SW_HIP_TRY(hipMalloc(&v, zmax));
    }
    SW_HIP_TRY(hipStreamSynchronize(st));
    *out = p.release();
    return SW_OK;
}

extern "C" int sw_codedisp_destroy(sw_codedisp* plan)
{
    if(plan) {
        (void)hipDeviceSynchronize();
        delete plan;
    }
    return SW_OK;
}

extern "C" int sw_codedisp_info(const sw_codedisp* plan, int dm_idx, int* fft_len, int* overlap,
                                int* latency)
{
    SW_CHECK_ARG(plan && dm_idx >= 0 && dm_idx < (int)plan->dms.size(), "bad DM index");
    if(fft_len) *fft_len = plan->N[dm_idx];
    if(overlap) *overlap = plan->O[dm_idx];
    if(latency) *latency = plan->lat[dm_idx];
    return SW_OK;
}

extern "C" int sw_codedisp_set_history(sw_codedisp* plan, const void* d_hist, void* stream)
{
    SW_CHECK_ARG(plan && d_hist, "null argument");
    if(plan->Oh == 0) return SW_OK;
    hipStream_t st = sw::as_stream(stream);
    if(plan->input_fpat) {
        SW_HIP_TRY(hipMemcpyAsync(plan->hist, d_hist, (size_t)plan->F * plan->PA * plan->Oh * 2,
                                  hipMemcpyDeviceToDevice, st));
        return SW_OK;
    }
    // FTPA [F][Oh][PA] -> [F][PA][Oh]
    return sw::corner_turn(d_hist, plan->hist, plan->F, plan->Oh, plan->PA, st);
}

static const int8_t* time_major_input(sw_codedisp* plan, const void* d_in, hipStream_t st, int& rc)
{
    rc = SW_OK;
    if(plan->input_fpat) return static_cast<const int8_t*>(d_in);
    rc = sw::corner_turn(d_in, plan->xtm, plan->F, plan->T, plan->PA, st);
    return plan->xtm;
}

extern "C" int sw_codedisp_push_history(sw_codedisp* plan, const void* d_in, void* stream)
{
    SW_CHECK_ARG(plan && d_in, "null argument");
    hipStream_t st = sw::as_stream(stream);
    const int ncols = plan->F * plan->PA;
    // outputs computed ahead of this gulp's window move to the front
    for(size_t d = 0; d < plan->dms.size(); ++d) {
        const int lat = plan->lat[d];
        if(lat == 0) continue;
        int rc = sw::carry(reinterpret_cast<uint16_t*>(plan->ytm[d]), static_cast<uint16_t*>(plan->carry_tmp),
                           ncols, plan->ylen[d], plan->T, lat, st);
        if(rc) return rc;
        if(plan->ytf[d]) {
            rc = sw::carry(plan->ytf[d], static_cast<float2*>(plan->carry_tmp), ncols, plan->ylen[d], plan->T, lat,
                           st);
            if(rc) return rc;
        }
    }
    for(auto& g : plan->groups) g.mdone = ((long long)(plan->gulps + 1) * plan->T) / g.hop;
    ++plan->gulps;
    if(plan->Oh == 0) return SW_OK;
    // an FTPA input was turned into xtm by the run of this gulp
    const int8_t* cur = plan->input_fpat ? static_cast<const int8_t*>(d_in) : plan->xtm;
    const size_t n = (size_t)ncols * plan->Oh;
    hipLaunchKernelGGL(sw::history_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                       plan->hist2, plan->hist, cur, ncols, plan->T, plan->Oh);
    SW_LAUNCH_CHECK();
    std::swap(plan->hist, plan->hist2);
    return SW_OK;
}

extern "C" int sw_codedisp_run(sw_codedisp* plan, const void* d_in, void* const* d_out,
                               void* const* d_out_f32, int dm_begin, int dm_count, int batched,
                               void* stream)
{
    SW_CHECK_ARG(plan && d_in && d_out, "null argument");
    const int nd = (int)plan->dms.size();
    SW_CHECK_ARG(dm_begin >= 0 && dm_count >= 0 && dm_begin + dm_count <= nd, "bad DM range");
    hipStream_t st = sw::as_stream(stream);
    int rc;
    const int8_t* cur = time_major_input(plan, d_in, st, rc);
    if(rc) return rc;
    for(auto& g : plan->groups) {
        std::vector<int> todo;
        for(int d : g.dms)
            if(d >= dm_begin && d < dm_begin + dm_count) todo.push_back(d);
        // segments of this gulp: all grid segments whose input ends within it
        const long long gT = plan->gulps * (long long)plan->T;
        const long long m0 = g.mdone;
        const long long m1 = (gT + plan->T) / g.hop;
        const int nseg = (int)(m1 - m0);
        if(nseg <= 0 && !todo.empty()) {
            // nothing new to transform (possible only when a gulp is shorter than hop)
        }
        const int per_launch = batched ? sw::kMaxDm : 1;
        for(size_t b0 = 0; b0 < todo.size(); b0 += per_launch) {
            const int nb = (int)std::min<size_t>(per_launch, todo.size() - b0);
            sw::CdParams p{};
            p.cur  = cur;
            p.hist = plan->hist;
            p.z    = plan->zbuf;
            p.T = plan->T;
            p.O = g.O;
            p.Oh = plan->Oh;
            p.hop = g.hop;
            p.nseg = std::max(nseg, 0);
            p.in0 = (long)(m0 * g.hop - g.O - gT);
            p.y0 = (long)(m0 * g.hop - gT + g.lat);
            p.ylen = plan->ylen[g.dms[0]];
            p.PA = plan->PA;
            p.F = plan->F;
            p.RP = g.RP;
            p.N = g.N;
            p.ndm = nb;
            p.tw  = plan->twiddles.at(g.L);
            p.twN = g.RP > 1 ? plan->twiddles.at(g.N) : nullptr;
            for(int k = 0; k < nb; ++k) {
                const int d = todo[b0 + k];
                SW_CHECK_ARG(d_out[d], "null output buffer");
                p.chirp[k] = plan->chirp[d];
                p.out[k]   = plan->ytm[d];
                if(d_out_f32 && d_out_f32[d] && !plan->ytf[d]) {
                    // float outputs (a diagnostic path): buffer allocated on first use
                    const size_t yb = (size_t)plan->F * plan->PA * plan->ylen[d] * sizeof(float2);
                    SW_HIP_TRY(hipMalloc(&plan->ytf[d], yb));
                    SW_HIP_TRY(hipMemsetAsync(plan->ytf[d], 0, yb, st));
                }
                p.outf[k]  = (d_out_f32 && d_out_f32[d]) ? plan->ytf[d] : nullptr;
                if(g.RP > 1) p.v[k] = plan->vbuf[k];
            }
            const int G = plan->PA / sw::lds_shape(g.L).C;
            if(p.nseg == 0) {
                // no segment completes in this gulp
            } else if(g.RP == 1) {
                rc = sw::launch_lds<false>(g.L, p, plan->F * p.nseg * G, st);
                if(rc) return rc;
            } else {
                const size_t items = (size_t)plan->F * p.nseg * plan->PA * g.L;
                const unsigned eb = (unsigned)((items + 255) / 256);
                switch(g.RP) {
                case 2: hipLaunchKernelGGL(sw::codedisp_dif_kernel<2>, dim3(eb), dim3(256), 0, st, p, g.L); break;
                case 4: hipLaunchKernelGGL(sw::codedisp_dif_kernel<4>, dim3(eb), dim3(256), 0, st, p, g.L); break;
                case 8: hipLaunchKernelGGL(sw::codedisp_dif_kernel<8>, dim3(eb), dim3(256), 0, st, p, g.L); break;
                case 16: hipLaunchKernelGGL(sw::codedisp_dif_kernel<16>, dim3(eb), dim3(256), 0, st, p, g.L); break;
                default: sw::set_error("codedisp: FFT length too large"); return SW_ERR_UNSUPPORTED;
                }
                SW_LAUNCH_CHECK();
                rc = sw::launch_lds<true>(g.L, p, plan->F * p.nseg * g.RP * G, st);
                if(rc) return rc;
                for(int k = 0; k < nb; ++k) {
                    switch(g.RP) {
                    case 2: hipLaunchKernelGGL(sw::codedisp_dit_kernel<2>, dim3(eb), dim3(256), 0, st, p, g.L, k); break;
                    case 4: hipLaunchKernelGGL(sw::codedisp_dit_kernel<4>, dim3(eb), dim3(256), 0, st, p, g.L, k); break;
                    case 8: hipLaunchKernelGGL(sw::codedisp_dit_kernel<8>, dim3(eb), dim3(256), 0, st, p, g.L, k); break;
                    default: hipLaunchKernelGGL(sw::codedisp_dit_kernel<16>, dim3(eb), dim3(256), 0, st, p, g.L, k); break;
                    }
                    SW_LAUNCH_CHECK();
                }
            }
            // emitted window [0, T) of the time-major result -> FTPA for the beamformers
            for(int k = 0; k < nb; ++k) {
                const int d = todo[b0 + k];
                rc = sw::corner_turn(plan->ytm[d], d_out[d], plan->F, plan->PA, plan->T, st, plan->ylen[d]);
                if(rc) return rc;
                if(p.outf[k]) {
                    rc = sw::corner_turn_f64(plan->ytf[d], d_out_f32[d], plan->F, plan->PA, plan->T, st,
                                             plan->ylen[d]);
                    if(rc) return rc;
                }
            }
        }
    }
    return SW_OK;
}
