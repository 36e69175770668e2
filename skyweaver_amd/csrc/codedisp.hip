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
// antenna)) so that every transform reads contiguous runs:
//  * N = 1024 ... 16384 (codedisp_reg_kernel<N/512>): a column is held in
//    the registers of 512 threads (32 points each) and transformed as
//    DFT-R1 x DFT-32 x DFT-16 with two LDS exchanges per direction; the
//    spectrum stays in registers in a permuted order that the chirp table
//    shares.  Persistent workgroups walk (channel, segment, column group).
//  * N = 32768 (codedisp_reg2_kernel): two 16384-point halves of a radix-2
//    decimation in frequency, fused in one persistent kernel.
//  * other lengths (shorter, or column counts the register kernels do not
//    divide into units): one workgroup per (channel, segment, group of C columns)
//    keeps C x N points in LDS (Stockham passes); with two LDS buffers the
//    forward FFT is shared by every DM of the launch.  N > 16384 there is a
//    radix-RP decimation-in-frequency stage, the LDS kernel on each of the
//    RP sub-sequences and a radix-RP decimation-in-time stage.
// Segments lie on one grid over the whole stream and outputs go to a
// per-DM time-major ring (stream sample x at ring index (x + latency) mod
// ylen), from which the beamformers read each gulp's window in place; a
// corner turn to FTPA is made only when the caller asks for an FTPA output.
#include "fft.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <vector>

namespace sw {
// [F][R][C] -> [F][C][R]; source rows src_ld >= C elements apart, channels
// R * src_ld apart; destination channels R * dst_chan_cols apart (>= C)
int corner_turn(const void* src, void* dst, int nchans, int R, int C, hipStream_t st, int src_ld = 0,
                int dst_chan_cols = 0);
int corner_turn_f64(const void* src, void* dst, int nchans, int R, int C, hipStream_t st, int src_ld = 0,
                    int dst_chan_cols = 0);

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
    long y0;    // ring index of segment 0's first kept sample (< ylen)
    int ylen;   // output ring samples per column
    const int8_t* prev;  // previous gulp ([F][PA][T]) holding the history, or null (hist used)
    int ndm;
    const float2* chirp[kMaxDm];  // [F][RP][L] (RP == 1: natural order)
    int8_t* out[kMaxDm];          // [F][PA][ylen] complex int8 output ring
    float2* outf[kMaxDm];         // optional [F][PA][ylen] y/N ring
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
    // __float2int_rn then static_cast<char>: round half even, wrap modulo 2^8.
    // Adding 1.5 * 2^23 rounds to an integer (the FP32 add rounds to nearest
    // even; 1.5 * 2^23 is even) and leaves it, two's complement, in the low
    // mantissa bits, exactly for |y| < 2^22.  A dedispersed sample is a
    // unit-gain filter of char2 input: |y| <= 128 * sqrt(2 N), < 2^22 for any
    // N below 2^29.
    const uint32_t re = __float_as_uint(y.x + 12582912.0f), im = __float_as_uint(y.y + 12582912.0f);
    *reinterpret_cast<uint16_t*>(p) = (uint16_t)((re & 0xffu) | ((im & 0xffu) << 8));
}

// input sample tau (relative to the gulp start) of column col of channel f
__device__ __forceinline__ float2 input_sample(const CdParams& p, int f, int col, int tau)
{
    if(tau < 0 && p.prev) return load_c8(p.prev + (((size_t)f * p.PA + col) * p.T + p.T + tau) * 2);
    if(tau < 0) return load_c8(p.hist + (((size_t)f * p.PA + col) * p.Oh + p.Oh + tau) * 2);
    if(tau < p.T) return load_c8(p.cur + (((size_t)f * p.PA + col) * p.T + tau) * 2);
    return make_float2(0.0f, 0.0f);
}

// raw complex int8 sample (re | im << 8), zero outside history + gulp
__device__ __forceinline__ uint32_t input_raw(const CdParams& p, int f, int col, int tau)
{
    if(tau < 0 && p.prev)
        return *reinterpret_cast<const uint16_t*>(p.prev + (((size_t)f * p.PA + col) * p.T + p.T + tau) * 2);
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
                long j = p.y0 + (long)s * p.hop + m;
                if(j >= p.ylen) j -= p.ylen;  // y0 < ylen and s*hop + m < ylen
                {
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

// N = 512 R1 (R1 = 2 ... 32: N = 1024 ... 16384) held in registers:
// 512 threads x 32 points per unit, a unit being C = 32 / R1 columns of one
// (channel, segment).  With n = 512 n1 + 16 n2 + n3 and
// k = k1 + R1 k2 + 32 R1 k3 (n1, k1 < R1; n2, k2 < 32; n3, k3 < 16) the
// forward transform is
//   DFT_R1 over n1 (thread t = 16 n2 + n3, each column), x W_N^(t k1)
//   -> exchange -> DFT32 over n2 (thread (c, k1, n3))
//   -> exchange -> x W_N^(R1 n3 k2), DFT16 over n3 (thread (c, k1, n3)
//      takes k2 = n3, n3 + 16),
// which leaves X[k] in registers; the chirp is stored permuted to the order
// the threads read it (index 32 R1 k3 + 32 k1 + k2).  The inverse runs the
// same steps backwards with conjugate twiddles and ends with each thread
// holding output samples 512 n1 + t of its columns.  Four LDS exchanges per
// transform pair instead of a read-write round trip per radix pass; the
// middle two (step 2 <-> step 3) stay inside the 16 lanes of a row, i.e.
// inside one wave.  Twiddles that multiply a DFT's inputs are folded into
// its first butterflies.
// Every exchange row j2 = (c, k1) has one pitch: the 512 points of
// exchanges 1/4 and the 32 rows of 17 of exchanges 2/3 both sit in
// [560 j2, 560 j2 + 560).  Row j2 is read in steps 2 and 2' and written in
// exchanges 2, 3 and 4 only by the 16 lanes (j2, n3) of one wave, so between
// the two column-wise exchanges (1 and 4) a wave touches only its own rows
// and needs no workgroup barrier.
constexpr int kEx1Row = 560;   // exchange 1/4 row (c, k1): 512 points + padding
constexpr int kEx2Row = 560;   // exchange 2/3 block (c, k1): 32 rows of 17
constexpr int kEx2Stride = 17; // padded row of 16 points
constexpr int kRegLds = 32 * kEx2Row;
// Lane groups of a wave that read or write one exchange row at a time start
// 32 banks apart (560 float2 are 32 dwords mod 64), and the rows of 17
// points put 32 consecutive k2 on distinct banks.

// An LDS exchange among the lanes of one wave: a wave's LDS operations are
// performed in program order, so a wave-scope fence (which also keeps the
// compiler from moving the accesses across it) is all the ordering needed.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The twiddles between steps 2 and 3 of an N = 512 R1 transform, W_N^(R1 n3
// k2) = W_512^(n3 k2) (n3 < 16, k2 < 32, tw[m] = W_N^m), as two tables in
// LDS: t1k[16 k2 + n3] for the inverse (a thread reads the row of its n3 at
// every k2) and t1k[512 + 32 n3 + k2] for the forward transform (a thread
// reads the row of its k2 at every n3).  Either way every read is the
// thread's base address plus an immediate offset, and the 16 addresses of a
// wave's lanes are 128 consecutive bytes.
template <int R1>
__device__ __forceinline__ void load_step2_tables(float2* t1k, const float2* __restrict__ tw, int t)
{
    t1k[t]       = tw[((t & 15) * (t >> 4)) * R1];  // R1 n3 k2 <= 465 R1 < N
    t1k[512 + t] = tw[((t >> 5) * (t & 31)) * R1];
}

// Chirp value k3 of the thread's column (k1, k2) of an N = 512 R1 chirp row
// (index 32 R1 k3 + 32 k1 + k2), read through a buffer resource over the
// row: one offset register per thread and a scalar offset per k3, so a
// run of these loads needs no address registers of its own.
template <int R1>
__device__ __forceinline__ float2 chirp_at(__amdgpu_buffer_rsrc_t rs, int voff, int k3)
{
    const auto v = __builtin_amdgcn_raw_buffer_load_b64(rs, voff, 256 * R1 * k3, 0);
    // (__builtin_bit_cast of an element of the returned vector reads element
    // 0 with this compiler; __uint_as_float takes the element's value)
    return make_float2(__uint_as_float(v[0]), __uint_as_float(v[1]));
}

// The step-1 twiddles W_N^(t k1), k1 = 1..R1-1, of thread t: copied from
// wk where a kernel holds them for its whole run (HELD), else built from
// W_N^t at the point of use (the same product tree, so the same values; the
// opaque copy keeps the compiler from hoisting the tree and holding it).
template <int R1, bool HELD>
__device__ __forceinline__ void step1_twiddles(const float2* wk, const Twiddles& tw, int t, float2 (&w)[R1])
{
    if constexpr(HELD) {
#pragma unroll
        for(int k = 1; k < R1; ++k) w[k] = wk[k];
    } else {
        float2 w1 = tw(t);
        asm volatile("" : "+v"(w1.x), "+v"(w1.y));
        twiddle_powers<R1>(w1, w);
    }
}

// Forward steps 1-3 of an N = 512 R1 transform of C = 32 / R1 columns: v
// holds the thread's inputs (value c R1 + n1 = sample 512 n1 + t of column
// c) and ends holding X[k] of the pairs q = t, t + 512 (16 values each,
// k3 = 0..15).  Step-1 twiddles as step1_twiddles<R1, HELD> gives them.
// The caller has synchronised since the exchange buffer was last read.
// With Hp given, the first half of the chirp values that reg_chirp_inverse
// multiplies by (k2 = n3, the chirp of the DM that follows) is loaded into
// hk0 during steps 2 and 3, so that the inverse does not wait for it.
template <int R1, bool HELD, bool PRE = false>
__device__ __forceinline__ void reg_forward(float2* v, float2* lds, const Twiddles& tw, const float2* t1k, int t,
                                            const float2* wk, const float2* __restrict__ Hp, float2 (&hk0)[16])
{
    constexpr int C = 32 / R1;
    const int n3 = t & 15, j2 = t >> 4;
#pragma unroll
    for(int c = 0; c < C; ++c) dft<R1, false>(v + c * R1);
    {
        float2 w[R1];
        step1_twiddles<R1, HELD>(wk, tw, t, w);
#pragma unroll
        for(int c = 0; c < C; ++c)
#pragma unroll
            for(int k = 1; k < R1; ++k) v[c * R1 + k] = cmul(v[c * R1 + k], w[k]);
    }
#pragma unroll
    for(int i = 0; i < 32; ++i) lds[i * kEx1Row + t] = v[i];  // row i = c R1 + k1
    __syncthreads();
    if constexpr(PRE) {
        const auto rs = buffer_rsrc(Hp, 512u * R1 * 8u);
        const int voff = 8 * (32 * (j2 % R1) + n3);
#pragma unroll
        for(int k3 = 0; k3 < 16; ++k3) hk0[k3] = chirp_at<R1>(rs, voff, k3);
    }
    // step 2: thread (j2 = c R1 + k1, n3) holds n2 = 0..31: DFT32 -> k2
#pragma unroll
    for(int n2 = 0; n2 < 32; ++n2) v[n2] = lds[j2 * kEx1Row + n2 * 16 + n3];
    wave_lds_sync();  // the wave's own rows are rewritten next
    dft32<false>(v);
#pragma unroll
    for(int k2 = 0; k2 < 32; ++k2) lds[j2 * kEx2Row + k2 * kEx2Stride + n3] = v[k2];
    wave_lds_sync();
    // step 3: thread (j2, n3) takes k2 = n3 and n3 + 16 of its own row j2,
    // which the 16 lanes of that row (in this wave) have just written:
    // x W_N^(R1 n3 k2) (the twiddle of value n3, folded into the first
    // butterflies), DFT16 over n3 -> k3
#pragma unroll
    for(int h = 0; h < 2; ++h) {
        const int k2 = n3 + 16 * h;
        float2* x = v + 16 * h;
        const float2* tk = t1k + 512 + k2;
        float2 w[16];
#pragma unroll
        for(int m = 0; m < 16; ++m) {
            x[m] = lds[j2 * kEx2Row + k2 * kEx2Stride + m];
            if(m > 0) w[m] = tk[32 * m];
        }
        dft16_mul<false, true>(x, w);
    }
}

// Y = X H with the chirp stored permuted (index 32 R1 k3 + 32 k1 + k2), then
// the inverse steps 3'-1': v ends holding y[512 n1 + t] (unnormalised) at
// c R1 + n1.  X (the spectrum reg_forward left) may be v itself, or be kept
// unchanged for further DMs.  The inverse's twiddle W_N^(-k1 (16 n2 + n3))
// between steps 2' and 1' is the conjugate of step-1 twiddle k1 of the
// thread t = 16 n2 + n3 that runs step 1'.
template <int R1, bool HELD, bool PRE = false>
__device__ __forceinline__ void reg_chirp_inverse(const float2* X, float2* v, float2* lds, const Twiddles& tw,
                                                  const float2* t1k, int t, const float2* __restrict__ H,
                                                  const float2* wk, const float2 (&hk0)[16])
{
    constexpr int C = 32 / R1;
    const int n3 = t & 15, j2 = t >> 4;
    const int k1_2 = j2 % R1;
    // step 3': IDFT16 of X H over k3 -> n3 (the chirp products folded into
    // the first butterflies)
    const auto rs = buffer_rsrc(H, 512u * R1 * 8u);
#pragma unroll
    for(int h = 0; h < 2; ++h) {
        const int k2 = n3 + 16 * h, k1 = k1_2;
        float2* x = v + 16 * h;
        float2 hk[16];
        const int voff = 8 * (32 * k1 + k2);
#pragma unroll
        for(int k3 = 0; k3 < 16; ++k3) {
            if constexpr(PRE)
                hk[k3] = h == 0 ? hk0[k3] : chirp_at<R1>(rs, voff, k3);
            else
                hk[k3] = H[32 * R1 * k3 + 32 * k1 + k2];
            x[k3]  = X[16 * h + k3];
        }
        dft16_mul<true>(x, hk);
    }
    // row j2 of the exchange buffer is this wave's alone between the two
    // workgroup barriers of a transform: written and read back by its own
    // 16 lanes
#pragma unroll
    for(int h = 0; h < 2; ++h) {
        const int k2 = n3 + 16 * h;
#pragma unroll
        for(int m = 0; m < 16; ++m) lds[j2 * kEx2Row + k2 * kEx2Stride + m] = v[16 * h + m];
    }
    wave_lds_sync();
    // step 2': thread (j2, n3) holds k2 = 0..31: x W_N^(-R1 n3 k2) (folded
    // into the first butterflies), IDFT32 -> n2
#pragma unroll
    for(int k2 = 0; k2 < 32; ++k2) v[k2] = lds[j2 * kEx2Row + k2 * kEx2Stride + n3];
    wave_lds_sync();  // the wave's own rows are rewritten next
    {
        const float2* tk = t1k + n3;
        float2 w[32];
#pragma unroll
        for(int k2 = 1; k2 < 32; ++k2) w[k2] = cconj(tk[16 * k2]);
        dft32_mul<true, true>(v, w);
    }
#pragma unroll
    for(int n2 = 0; n2 < 32; ++n2) lds[j2 * kEx1Row + n2 * 16 + n3] = v[n2];
    __syncthreads();
    // step 1': thread t holds k1 = 0..R1-1 of each column: x W_N^(-k1 t)
    // (folded into the first butterflies for R1 >= 16), IDFT_R1 -> n1
#pragma unroll
    for(int i = 0; i < 32; ++i) v[i] = lds[i * kEx1Row + t];
    float2 wc[R1];
    step1_twiddles<R1, HELD>(wk, tw, t, wc);
#pragma unroll
    for(int k = 1; k < R1; ++k) wc[k] = cconj(wc[k]);
#pragma unroll
    for(int c = 0; c < C; ++c) dft_mul<R1, true>(v + c * R1, wc);
}

// Raw complex int8 samples b + 512 i + t, i = 0..2 n - 1, of one column,
// two per register, where b is the first sample of a segment relative to
// the gulp start (negative: history).
template <int NP>
__device__ __forceinline__ void load_column(const CdParams& p, int f, int col, long b, int t, uint32_t* dst)
{
    constexpr int NS = 1024 * NP;  // samples spanned
    const size_t cc = (size_t)f * p.PA + col;
    const uint16_t* cur = reinterpret_cast<const uint16_t*>(p.cur) + cc * p.T;
    if(b >= 0 && b + NS <= p.T) {  // inside this gulp (uniform branch)
        const uint16_t* x = cur + b;
#pragma unroll
        for(int i = 0; i < NP; ++i)
            dst[i] = (uint32_t)x[1024 * i + t] | ((uint32_t)x[1024 * i + 512 + t] << 16);
    } else if(b < 0 && b + NS <= p.T) {  // history, then this gulp
        // lo[tau] for tau < 0: the previous gulp's tail or the history buffer
        const uint16_t* lo = p.prev ? reinterpret_cast<const uint16_t*>(p.prev) + cc * p.T + p.T
                                    : reinterpret_cast<const uint16_t*>(p.hist) + cc * p.Oh + p.Oh;
        const int b32 = (int)b;
#pragma unroll
        for(int i = 0; i < NP; ++i) {
            const int t0 = b32 + 1024 * i + t, t1 = t0 + 512;
            dst[i] = (uint32_t)(t0 < 0 ? lo : cur)[t0] | ((uint32_t)(t1 < 0 ? lo : cur)[t1] << 16);
        }
    } else {
#pragma unroll
        for(int i = 0; i < NP; ++i)
            dst[i] = input_raw(p, f, col, (int)(b + 1024 * i + t)) |
                     (input_raw(p, f, col, (int)(b + 1024 * i + 512 + t)) << 16);
    }
}

// Quantise y / N and write the kept positions [O/2, N - O/2) of a segment
// into the output ring (ring index j0 + position, wrapped once).
__device__ __forceinline__ void store_kept(const CdParams& p, int d, size_t colbase, int j0, int pos, float2 y)
{
    int jj = j0 + pos;
    if(jj >= p.ylen) jj -= p.ylen;
    store_c8(p.out[d] + (colbase + jj) * 2, y);
    if(p.outf[d]) p.outf[d][colbase + jj] = y;
}
__device__ __forceinline__ void store_output(const CdParams& p, int d, size_t colbase, int j0, int pos, int N,
                                             float2 y)
{
    const int half = p.O / 2;
    if(pos >= half && pos < N - half) store_kept(p, d, colbase, j0, pos, y);
}

// The stores of one unit of the register kernels: value c R1 + n1 of v is
// position 512 n1 + t of column col0 + c, and the kept positions
// [O/2, N - O/2) go to DM d's ring at index j0 + position, wrapped once.
// Ring offsets are carried in bytes and wrapped with one unsigned minimum
// (o - 2 ylen wraps around to a large value while o < 2 ylen); the column
// base is uniform, so a store needs no 64-bit address arithmetic.  When the
// overlap is at most 1024 only the first and last position rows (n1 = 0 and
// R1 - 1) hold dropped positions; the float ring, when asked for, is written
// in a separate pass so the int8 stores carry no test of it.
template <int R1>
__device__ __forceinline__ void store_unit(const CdParams& p, int d, size_t col0, int j0, int t, const float2* v)
{
    constexpr int N = 512 * R1, C = 32 / R1;
    int8_t* const o8 = p.out[d];
    float2* const of = p.outf[d];
    const int half = p.O / 2;
    const uint32_t ylen2 = 2u * (uint32_t)p.ylen;
    int tt = t;
    asm volatile("" : "+v"(tt));  // keeps the per-row offsets from being precomputed and held
    const uint32_t base = 2u * (uint32_t)(j0 + tt);  // byte offset of position tt (kept positions: >= 0)
    auto off = [&](int n1) {
        const uint32_t o = base + 1024u * n1;
        return min(o, o - ylen2);
    };
    auto put = [&](int8_t* ob, int n1, float2 y) {
        const uint32_t re = __float_as_uint(y.x + 12582912.0f), im = __float_as_uint(y.y + 12582912.0f);
        *reinterpret_cast<uint16_t*>(ob + off(n1)) = (uint16_t)((re & 0xffu) | ((im & 0xffu) << 8));
    };
    const bool lo_kept = tt >= half, hi_kept = tt < 512 - half;  // rows 0 and R1 - 1
    if(half <= 512) {
#pragma unroll
        for(int c = 0; c < C; ++c) {
            int8_t* const ob = o8 + (col0 + c) * (size_t)ylen2;
#pragma unroll
            for(int n1 = 1; n1 < R1 - 1; ++n1) put(ob, n1, v[c * R1 + n1]);
            if(lo_kept) put(ob, 0, v[c * R1]);
            if(hi_kept) put(ob, R1 - 1, v[c * R1 + R1 - 1]);
        }
    } else {
#pragma unroll
        for(int c = 0; c < C; ++c)
#pragma unroll
            for(int n1 = 0; n1 < R1; ++n1) {
                const int pos = 512 * n1 + tt;
                if(pos >= half && pos < N - half) store_kept(p, d, (col0 + c) * p.ylen, j0, pos, v[c * R1 + n1]);
            }
        return;  // store_kept wrote the float ring too
    }
    if(of) {
#pragma unroll
        for(int c = 0; c < C; ++c) {
            float2* const fb = of + (col0 + c) * (size_t)p.ylen;
#pragma unroll
            for(int n1 = 0; n1 < R1; ++n1) {
                const int pos = 512 * n1 + tt;
                if(pos >= half && pos < N - half) fb[off(n1) / 2] = v[c * R1 + n1];
            }
        }
    }
}

// Persistent: a workgroup (one per CU, the exchange buffer fills the LDS)
// walks units u = blockIdx.x + k gridDim.x, u = (channel, segment, column
// group), and loads the next unit's input while transforming the current
// one.  DM d of the launch's parameter block.
template <int R1>
__global__ void __launch_bounds__(512) codedisp_reg_kernel(CdParams p, int d, int nunits)
{
    constexpr int N = 512 * R1;
    constexpr int C = 32 / R1;
    __shared__ float2 lds[kRegLds + TwiddleLds<N>::SIZE + 1024];
    const Twiddles tw = load_twiddles<N, 512>(lds + kRegLds, p.tw);
    float2* t1k = lds + kRegLds + TwiddleLds<N>::SIZE;  // step 2/3 twiddle tables
    const int t = threadIdx.x;
    load_step2_tables<R1>(t1k, p.tw, t);
    const int ngrp = p.PA / C;
    uint32_t nxt[16];  // two raw samples per register: value i = c R1 + n1
    auto fetch = [&](int u) {
        const int cg = u % ngrp, fs = u / ngrp;
        const int f = fs / p.nseg, s = fs % p.nseg;
        const long b = p.in0 + (long)s * p.hop;
#pragma unroll
        for(int c = 0; c < C; ++c) load_column<R1 / 2>(p, f, cg * C + c, b, t, nxt + c * (R1 / 2));
    };
    if(blockIdx.x < nunits) fetch(blockIdx.x);
    __syncthreads();  // twiddle tables
    float2 wk[R1];    // step-1 twiddles W_N^(t k1), the same for every unit
    twiddle_powers<R1>(tw(t), wk);
    for(int u = blockIdx.x; u < nunits; u += gridDim.x) {
        const int cg = u % ngrp, fs = u / ngrp;
        const int f = fs / p.nseg, s = fs % p.nseg;
        float2 v[32];
#pragma unroll
        for(int i = 0; i < 16; ++i) {
            v[2 * i]     = unpack_c8(nxt[i]);
            v[2 * i + 1] = unpack_c8(nxt[i] >> 16);
        }
        if(u + (int)gridDim.x < nunits) fetch(u + gridDim.x);  // in flight during the transforms
        // no barrier: step 1 writes the column t that this thread alone read
        // in the previous unit's step 1'
        const float2* H = p.chirp[d] + (size_t)f * N;
        float2 hk0[16];
        reg_forward<R1, true, true>(v, lds, tw, t1k, t, wk, H, hk0);
        reg_chirp_inverse<R1, true, true>(v, v, lds, tw, t1k, t, H, wk, hk0);
        const int j0 = (int)(p.y0 + (long)s * p.hop) - p.O / 2;  // ring index of position 0
        store_unit<R1>(p, d, (size_t)f * p.PA + cg * C, j0, t, v);  // v already / N (chirp)
    }
}

// The DMs of one launch (p.ndm, sharing N and the overlap) from one forward
// transform: per unit the spectrum X stays in registers while each DM's
// chirp, inverse transform and stores run in turn, so a unit costs one
// forward and p.ndm inverse transforms instead of p.ndm of each.  The next
// unit's input is fetched during the last DM, once X is no longer needed.
template <int R1>
__global__ void __launch_bounds__(512) codedisp_regm_kernel(CdParams p, int nunits)
{
    constexpr int N = 512 * R1;
    constexpr int C = 32 / R1;
    __shared__ float2 lds[kRegLds + TwiddleLds<N>::SIZE + 1024];
    const Twiddles tw = load_twiddles<N, 512>(lds + kRegLds, p.tw);
    float2* t1k = lds + kRegLds + TwiddleLds<N>::SIZE;  // step 2/3 twiddle tables
    const int t = threadIdx.x;
    load_step2_tables<R1>(t1k, p.tw, t);
    const int ngrp = p.PA / C;
    uint32_t nxt[16];  // two raw samples per register: value i = c R1 + n1
    auto fetch = [&](int u) {
        const int cg = u % ngrp, fs = u / ngrp;
        const int f = fs / p.nseg, s = fs % p.nseg;
        const long b = p.in0 + (long)s * p.hop;
#pragma unroll
        for(int c = 0; c < C; ++c) load_column<R1 / 2>(p, f, cg * C + c, b, t, nxt + c * (R1 / 2));
    };
    if(blockIdx.x < nunits) fetch(blockIdx.x);
    __syncthreads();  // twiddle tables
    // step-1 twiddles W_N^(t k1), the same for every unit; from R1 = 16
    // they are rebuilt at each use (held beside X they would spill)
    constexpr bool HELD = R1 <= 8;
    float2 wk[R1];
    if constexpr(HELD) twiddle_powers<R1>(tw(t), wk);
    for(int u = blockIdx.x; u < nunits; u += gridDim.x) {
        const int cg = u % ngrp, fs = u / ngrp;
        const int f = fs / p.nseg, s = fs % p.nseg;
        float2 X[32];
#pragma unroll
        for(int i = 0; i < 16; ++i) {
            X[2 * i]     = unpack_c8(nxt[i]);
            X[2 * i + 1] = unpack_c8(nxt[i] >> 16);
        }
        // no barrier: step 1 writes the column t that this thread alone read
        // in the previous DM's step 1'
        float2 hk0[16];  // no chirp prefetch here: each DM has its own chirp
        reg_forward<R1, HELD>(X, lds, tw, t1k, t, wk, nullptr, hk0);
        const int j0 = (int)(p.y0 + (long)s * p.hop) - p.O / 2;  // ring index of position 0
        for(int d = 0; d < p.ndm; ++d) {
            if(d > 0) __syncthreads();  // the previous DM's step-1' readers of every row
            if(d == p.ndm - 1 && u + (int)gridDim.x < nunits) fetch(u + gridDim.x);
            float2 v[32];
            reg_chirp_inverse<R1, HELD>(X, v, lds, tw, t1k, t, p.chirp[d] + (size_t)f * N, wk, hk0);
            store_unit<R1>(p, d, (size_t)f * p.PA + cg * C, j0, t, v);  // v already / N (chirp)
        }
    }
}

// N = 32768 as two interleaved 16384-point halves (radix-2 decimation in
// frequency, RP = 2, L = N / 2), fused into one persistent kernel:
//   Z_r[n] = W_N^(n r) (x[n] + (-1)^r x[n + L]),  X[2m + r] = FFT_L(Z_r)[m]
//   V_r = IFFT_L(X[2m + r] H[2m + r])
//   y[n] = V_0[n] + W_N^(-n) V_1[n],  y[n + L] = V_0[n] - W_N^(-n) V_1[n]
// Each thread holds samples n = 512 n1 + t (and n + L) of one column.  V_0
// waits in a per-workgroup slice of `scratch` (each thread reads back only
// what it wrote) while V_1 is transformed; the chirp is stored [F][r][L] in
// the register order of the halves.  The raw samples are loaded for each
// half (the second load hits L2) rather than held across a transform, which
// would not fit the registers.
__global__ void __launch_bounds__(512) codedisp_reg2_kernel(CdParams p, int d, int nunits, float2* scratch)
{
    constexpr int L = 16384, N = 2 * L;
    __shared__ float2 lds[kRegLds + TwiddleLds<L>::SIZE + 1024];
    const Twiddles tw = load_twiddles<L, 512>(lds + kRegLds, p.tw);
    float2* t1k = lds + kRegLds + TwiddleLds<L>::SIZE;  // step 2/3 twiddle tables
    const int t = threadIdx.x;
    load_step2_tables<32>(t1k, p.tw, t);
    uint32_t raw[32];  // x[512 n1 + t] at 0..15, x[L + 512 n1 + t] at 16..31, two per register
    for(int u = blockIdx.x; u < nunits; u += gridDim.x) {
        const int col = u % p.PA, fs = u / p.PA;
        const int f = fs / p.nseg, s = fs % p.nseg;
        const long b = p.in0 + (long)s * p.hop;
        const float2* H = p.chirp[d] + (size_t)f * N;
        float2 v[32];
        // the two halves share one copy of the transform code (a loop that
        // is not unrolled): the kernel then fits the instruction cache
#pragma unroll 1
        for(int r = 0; r < 2; ++r) {
            // the thread's slots, formed afresh per half (an opaque offset
            // keeps the compiler from holding 32 precomputed addresses in
            // registers across the transforms)
            int off = t;
            asm volatile("" : "+v"(off));
            float2* mine = scratch + (size_t)blockIdx.x * L + off;  // [n1][512]
            const float2* wn = p.twN + off;                         // W_N^(512 n1 + t) at wn[512 n1]
            load_column<32>(p, f, col, b, off, raw);
            if(r == 0) {  // Z_0 = x[n] + x[n + L]
#pragma unroll
                for(int i = 0; i < 16; ++i) {
                    v[2 * i]     = cadd(unpack_c8(raw[i]), unpack_c8(raw[16 + i]));
                    v[2 * i + 1] = cadd(unpack_c8(raw[i] >> 16), unpack_c8(raw[16 + i] >> 16));
                }
            } else {  // Z_1 = W_N^n (x[n] - x[n + L])
#pragma unroll
                for(int i = 0; i < 16; ++i) {
                    v[2 * i]     = csub(unpack_c8(raw[i]), unpack_c8(raw[16 + i]));
                    v[2 * i + 1] = csub(unpack_c8(raw[i] >> 16), unpack_c8(raw[16 + i] >> 16));
                }
#pragma unroll
                for(int n1 = 0; n1 < 32; ++n1) v[n1] = cmul(v[n1], wn[512 * n1]);
            }
            __syncthreads();  // twiddle tables; previous readers of the exchange buffer
            // step-1 twiddles rebuilt at each use (held, they would not fit
            // the registers)
            float2 hk0[16];  // no chirp prefetch (registers)
            reg_forward<32, false>(v, lds, tw, t1k, t, nullptr, nullptr, hk0);
            reg_chirp_inverse<32, false>(v, v, lds, tw, t1k, t, H + r * L, nullptr, hk0);
            if(r == 0) {
#pragma unroll
                for(int n1 = 0; n1 < 32; ++n1) mine[512 * n1] = v[n1];
            }
        }
        // ---- y[n] = V_0 + W_N^(-n) V_1, y[n + L] = V_0 - W_N^(-n) V_1
        const int j0 = (int)(p.y0 + (long)s * p.hop) - p.O / 2;
        const size_t cb = ((size_t)f * p.PA + col) * p.ylen;
        int tt = t;  // opaque: addresses and position tests formed here, not held across units
        asm volatile("" : "+v"(tt));
        const float2* mine = scratch + (size_t)blockIdx.x * L + tt;
        const float2* wn = p.twN + tt;
        const int half = p.O / 2;
        if(half <= 512 && !p.outf[d]) {
            // as store_unit: byte offsets from the column's base wrapped by
            // one unsigned minimum; only rows 0 and 63 can hold dropped
            // positions
            int8_t* const ob = p.out[d] + cb * 2;
            const uint32_t ylen2 = 2u * (uint32_t)p.ylen;
            const uint32_t base = 2u * (uint32_t)(j0 + tt);
            auto put = [&](int row, float2 y) {
                const uint32_t o = base + 1024u * row;
                const uint32_t re = __float_as_uint(y.x + 12582912.0f), im = __float_as_uint(y.y + 12582912.0f);
                *reinterpret_cast<uint16_t*>(ob + min(o, o - ylen2)) = (uint16_t)((re & 0xffu) | ((im & 0xffu) << 8));
            };
            const bool lo_kept = tt >= half, hi_kept = tt < 512 - half;
#pragma unroll
            for(int n1 = 0; n1 < 32; ++n1) {
                const float2 a = mine[512 * n1];
                const float2 b = cmul(v[n1], cconj(wn[512 * n1]));  // V_0, V_1 already / N (chirp)
                if(n1 > 0 || lo_kept) put(n1, cadd(a, b));
                if(n1 < 31 || hi_kept) put(32 + n1, csub(a, b));
            }
        } else {
#pragma unroll
            for(int n1 = 0; n1 < 32; ++n1) {
                const float2 a = mine[512 * n1];
                const float2 b = cmul(v[n1], cconj(wn[512 * n1]));
                store_output(p, d, cb, j0, 512 * n1 + tt, N, cadd(a, b));
                store_output(p, d, cb, j0, L + 512 * n1 + tt, N, csub(a, b));
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
            long j = p.y0 + (long)s * p.hop + pos - half;
            if(j >= p.ylen) j -= p.ylen;
            {
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
// [RP][L = N / RP] with k = RP m + r at r L + m', where m' = m, or for the
// register kernels (R1 > 0) m = k1 + R1 k2 + 32 R1 k3 at m' = 32 R1 k3 + 32 k1 + k2
__global__ void chirp_kernel(float2* out, int nchans, int N, int RP, int R1, double low_freq,
                             double chan_bw, double dm_prefactor, int physical)
{
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(idx >= (size_t)nchans * N) return;
    const int chan = (int)(idx / N), k = (int)(idx % N);
    const double fine_bw = chan_bw / N;
    // reference: nu_0 half a channel below the channel's bottom edge and
    // nu = k dnu for every bin; physical: nu_0 the channel centre and the
    // upper half of the bins the negative baseband frequencies
    const double nu_0 = physical ? low_freq + chan * chan_bw + 0.5 * chan_bw : low_freq + chan * chan_bw - 0.5 * chan_bw;
    const double nu   = (physical && 2 * k >= N ? k - N : k) * fine_bw;
    const double phase_prefactor = nu * nu * dm_prefactor;
    const double phase = phase_prefactor / ((nu_0 + nu) * nu_0 * nu_0);
    double s, c;
    sincos(phase, &s, &c);
    const int L = N / RP, m = k / RP, r = k % RP;
    int mp = m;
    float scale = 1.0f;
    if(R1 > 0) {
        const int k1 = m % R1, k2 = (m / R1) % 32, k3 = m / (32 * R1);
        mp = 32 * R1 * k3 + 32 * k1 + k2;
        // the register kernels' chirp carries the 1/N of the inverse
        // transform: a power of two, so every value of the inverse is
        // scaled exactly and the result is bit-identical to dividing at the end
        scale = 1.0f / (float)N;
    }
    out[((size_t)chan * RP + r) * L + mp] = make_float2((float)c * scale, (float)s * scale);
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
        if constexpr(F32) {
            hipLaunchKernelGGL((codedisp_tm_kernel<16384, 1, 1, 512, true>), dim3(blocks), dim3(512), 0, st, p);
        } else {
            set_error("codedisp: N = 16384 runs in the register kernel");
            return SW_ERR_INVALID;
        }
        break;
    default: set_error("codedisp: unsupported transform length"); return SW_ERR_UNSUPPORTED;
    }
#undef SW_CD
    SW_LAUNCH_CHECK();
    return SW_OK;
}

// the register kernels take N = 1024 ... 16384 (R1 = N / 512) when the
// columns split into units of 32 / R1, and N = 32768 as two halves of 16384
// (RP = 2); 0: the LDS / global-memory kernels
int reg_r1(int N, int RP, int PA)
{
    if(RP == 2 && N == 32768) return 32;
    if(RP != 1 || N < 1024 || N > 16384 || (N & (N - 1)) != 0) return 0;
    const int R1 = N / 512;
    return PA % (32 / R1) == 0 ? R1 : 0;
}

// units = (channel, segment, group of 32 / R1 columns), or (channel,
// segment, column) for N = 32768.  N <= 16384: several DMs (p.ndm > 1) share
// one forward transform in codedisp_regm_kernel, one DM alone runs in
// codedisp_reg_kernel; N = 32768: one DM per launch.
int launch_reg(int R1, const CdParams& p, int nunits, float2* scratch, hipStream_t st)
{
    const dim3 grid((unsigned)std::min(nunits, persistent_blocks()));
    if(p.RP == 2) {
        for(int d = 0; d < p.ndm; ++d) {
            hipLaunchKernelGGL(codedisp_reg2_kernel, grid, dim3(512), 0, st, p, d, nunits, scratch);
            SW_LAUNCH_CHECK();
        }
        return SW_OK;
    }
    if(p.ndm > 1) {
#define SW_REGM(RR)                                                                             \
    case RR:                                                                                    \
        hipLaunchKernelGGL((codedisp_regm_kernel<RR>), grid, dim3(512), 0, st, p, nunits);     \
        break;
        switch(R1) {
            SW_REGM(2) SW_REGM(4) SW_REGM(8) SW_REGM(16) SW_REGM(32)
        default: set_error("codedisp: no register kernel for this length"); return SW_ERR_INVALID;
        }
#undef SW_REGM
        SW_LAUNCH_CHECK();
        return SW_OK;
    }
    for(int d = 0; d < p.ndm; ++d) {
#define SW_REG(RR)                                                                              \
    case RR:                                                                                    \
        hipLaunchKernelGGL((codedisp_reg_kernel<RR>), grid, dim3(512), 0, st, p, d, nunits);  \
        break;
        switch(R1) {
            SW_REG(2) SW_REG(4) SW_REG(8) SW_REG(16) SW_REG(32)
        default: set_error("codedisp: no register kernel for this length"); return SW_ERR_INVALID;
        }
#undef SW_REG
        SW_LAUNCH_CHECK();
    }
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
    std::vector<int8_t*> ytm;            // per DM: output ring [F][PA][ylen], time-major
    std::vector<float2*> ytf;            // per DM, on demand: float output ring, same layout
    std::vector<int> ylen;               // per DM ring length: T + lat rounded up to a multiple of T
    std::vector<int> wstart;             // per DM: ring index of the last run's emitted window
    bool hist_ref = false;               // history read in place from the previous gulp
    const int8_t* prev = nullptr;        // that gulp (null before the first)
    std::vector<int> lat;                // per DM output latency
    long long gulps = 0;                 // gulps completed (push_history calls)
    float2* zbuf = nullptr;              // DIF stage output (N > kMaxL)
    float2* reg2 = nullptr;              // N = 32768 kernel: one 16384-point half per workgroup
    std::vector<float2*> vbuf;           // per DM slot (N > kMaxL)
    ~sw_codedisp()
    {
        for(auto* c : chirp) (void)hipFree(c);
        for(auto& kv : twiddles) (void)hipFree(kv.second);
        (void)hipFree(hist);
        (void)hipFree(hist2);
        (void)hipFree(xtm);
        for(auto* y : ytm) (void)hipFree(y);
        for(auto* y : ytf) (void)hipFree(y);
        (void)hipFree(zbuf);
        (void)hipFree(reg2);
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

} // namespace

// The overlap a plan gives a DM when none is configured: the even ceiling
// of the intra-channel smear at the bottom of the band.
int codedisp_overlap(double low_freq, double chan_bw, double tsamp, float dm)
{
    return even_ceil(smear_samples(low_freq, chan_bw, tsamp, dm));
}

namespace {

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
    SW_CHECK_ARG(cfg->chirp == SW_CHIRP_REFERENCE || cfg->chirp == SW_CHIRP_PHYSICAL, "unknown chirp mapping");
    hipStream_t st = sw::as_stream(stream);
    auto p = std::make_unique<sw_codedisp>();
    p->F = cfg->nchans;
    p->A = cfg->nantennas;
    p->PA = 2 * cfg->nantennas;
    p->T = cfg->nsamples;
    p->input_fpat = cfg->input_fpat;
    p->hist_ref = cfg->history_in_place && cfg->input_fpat;
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
    }
    if(cfg->share_forward) {  // one overlap per FFT length: the largest
        std::map<int, int> omax;
        for(size_t d = 0; d < p->dms.size(); ++d) omax[p->N[d]] = std::max(omax[p->N[d]], p->O[d]);
        for(size_t d = 0; d < p->dms.size(); ++d) p->O[d] = omax[p->N[d]];
    }
    for(int O : p->O) p->Omax = std::max(p->Omax, O);
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
    p->wstart.assign(p->dms.size(), 0);
    for(auto& g : p->groups) {
        // history: the input a gulp's first segment can reach back to
        p->Oh = std::max(p->Oh, g.lat > 0 ? g.N : g.O);
        for(int d : g.dms) {
            p->lat[d] = g.lat;
            // a whole number of gulps (>= T + lat): every emitted window is one
            // contiguous, gulp-aligned slice [w0, w0 + T) of the ring
            p->ylen[d] = (p->T + g.lat + p->T - 1) / p->T * p->T;
        }
    }
    size_t zmax = 0, slots = 1, vslots = 0;
    bool reg2 = false;
    for(auto& g : p->groups) {
        if(!sw::twiddle_table(p.get(), g.L, st)) return SW_ERR_ALLOC;
        slots = std::max(slots, std::min<size_t>(sw::kMaxDm, g.dms.size()));
        if(g.RP > 1) {
            if(!sw::twiddle_table(p.get(), g.N, st)) return SW_ERR_ALLOC;
            if(sw::reg_r1(g.N, g.RP, p->PA) > 0) {
                reg2 = true;
            } else {
                zmax = std::max(zmax, (size_t)p->F * g.nseg * g.N * p->PA * sizeof(float2));
                vslots = std::max(vslots, std::min<size_t>(sw::kMaxDm, g.dms.size()));
            }
        }
    }
    if(reg2)
        SW_HIP_TRY(hipMalloc(&p->reg2, sizeof(float2) * 16384 * (size_t)sw::persistent_blocks()));
    p->chirp.resize(p->dms.size(), nullptr);
    for(int d = 0; d < (int)p->dms.size(); ++d) {
        const int N = p->N[d], RP = N / std::min(N, sw::kMaxL);
        const int R1 = sw::reg_r1(N, RP, p->PA);
        if(hipMalloc(&p->chirp[d], sizeof(float2) * (size_t)p->F * N) != hipSuccess) {
            sw::set_error("codedisp: chirp allocation failed");
            return SW_ERR_ALLOC;
        }
        // DMPrefactor: -1.0f * 2 pi * K * dm (cpp/skyweaver/dedispersion_utils.cuh:42-48)
        const double pref = -1.0f * (2 * 3.1415926535897932385) * sw::kDmConst * (double)p->dms[d];
        const size_t n = (size_t)p->F * N;
        hipLaunchKernelGGL(sw::chirp_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                           p->chirp[d], p->F, N, RP, R1, p->low_freq, p->chan_bw, pref,
                           cfg->chirp == SW_CHIRP_PHYSICAL ? 1 : 0);
        SW_LAUNCH_CHECK();
    }
    const size_t vol = (size_t)p->F * p->PA * p->T * 2;
    if(p->hist_ref && p->Oh > p->T) p->hist_ref = false;  // history reaches past the previous gulp
    if(p->Oh > 0) {
        // with in-place history only the zeros before the first gulp are needed
        const size_t hb = (size_t)p->F * p->PA * p->Oh * 2;
        SW_HIP_TRY(hipMalloc(&p->hist, hb));
        if(!p->hist_ref) SW_HIP_TRY(hipMalloc(&p->hist2, hb));
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
    if(zmax) {
        SW_HIP_TRY(hipMalloc(&p->zbuf, zmax));
        p->vbuf.resize(vslots, nullptr);
        for(auto& v : p->vbuf) SW_HIP_TRY(hipMalloc(&v, zmax));
    }
    SW_HIP_TRY(hipStreamSynchronize(st));
    *out = p.release();
    return SW_OK;
}

namespace sw {
namespace {
// deterministic pseudo-random complex int8 (a 32-bit mix of the index)
__global__ void tune_fill_kernel(uint32_t* x, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n) return;
    uint32_t h = (uint32_t)i * 0x9E3779B1u;
    h ^= h >> 15;
    h *= 0x85EBCA77u;
    h ^= h >> 13;
    x[i] = h & 0x3F3F3F3Fu;  // |value| < 64: no int8 saturation in the transforms
}
} // namespace
} // namespace sw

namespace sw {
namespace {
// Device time (ms) of one gulp of plan DMs [0, ndm) run together (batched) or
// one at a time: two warm-up gulps, three timed.
int time_gulps(sw_codedisp* plan, const void* d_in, int ndm, int dm_begin, int batched, hipStream_t st,
               hipEvent_t e0, hipEvent_t e1, float* ms_out)
{
    std::vector<void*> outs(plan->dms.size(), nullptr);  // outputs stay in the ring, as in the pipeline
    float acc = 0.0f;
    for(int it = 0; it < 5; ++it) {
        if(it >= 2) SW_HIP_TRY(hipEventRecord(e0, st));
        int rc = sw_codedisp_run(plan, d_in, outs.data(), nullptr, dm_begin, ndm, batched, st);
        if(rc) return rc;
        if(it >= 2) {
            SW_HIP_TRY(hipEventRecord(e1, st));
            SW_HIP_TRY(hipEventSynchronize(e1));
            float ms = 0.0f;
            SW_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
            acc += ms;
        }
        rc = sw_codedisp_push_history(plan, d_in, st);
        if(rc) return rc;
    }
    *ms_out = acc / 3.0f;
    return SW_OK;
}
} // namespace
} // namespace sw

extern "C" int sw_codedisp_autotune(const sw_codedisp_config* cfg, int tune_chans, int* fft_len_out,
                                    int* share_forward_out, float* ms_out, void* stream)
{
    SW_CHECK_ARG(cfg && fft_len_out && cfg->ndms > 0 && cfg->dms && tune_chans > 0, "bad argument");
    hipStream_t st = sw::as_stream(stream);
    const int nd = cfg->ndms;
    static const int kCand[] = {1024, 2048, 4096, 8192, 16384, 32768};
    constexpr int kNc = sizeof(kCand) / sizeof(kCand[0]);
    // the overlaps the plan would choose
    std::vector<int> O(nd);
    for(int d = 0; d < nd; ++d)
        O[d] = (cfg->overlap && cfg->overlap[d] >= 0)
                   ? cfg->overlap[d]
                   : sw::even_ceil(sw::smear_samples(cfg->low_freq, cfg->chan_bw, cfg->tsamp, cfg->dms[d]));
    sw_codedisp_config sub = *cfg;
    sub.nchans = std::min(tune_chans, cfg->nchans);
    sub.overlap = O.data();
    sub.history_in_place = 0;
    sub.share_forward = 0;
    const size_t words = (size_t)sub.nchans * 2 * cfg->nantennas * cfg->nsamples / 2;
    uint32_t* d_in = nullptr;
    SW_HIP_TRY(hipMalloc(&d_in, words * 4));
    hipLaunchKernelGGL(sw::tune_fill_kernel, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, st, d_in, words);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    const float scale = (float)cfg->nchans / (float)sub.nchans;
    std::vector<float> best(nd, 1e30f), times((size_t)(nd + 1) * kNc, -1.0f);
    std::vector<int> N(nd, 0);
    std::vector<std::vector<int>> fits(kNc);
    // a length fits DM d if it keeps the overlap waste below a quarter
    // (N >= 4 O) and is no longer than the gulp (a longer transform delays
    // the output by more than a gulp), or is the shortest length with
    // N >= 4 O when every such length is longer than the gulp
    auto fits_dm = [&](int c, int d) {
        if(kCand[c] < 4 * O[d]) return false;
        if(kCand[c] <= cfg->nsamples) return true;
        return c == 0 || kCand[c - 1] < 4 * O[d];
    };
    int rc = SW_OK;
    // one DM per launch, every length
    for(int c = 0; c < kNc && rc == SW_OK; ++c) {
        std::vector<int> fl(nd);
        for(int d = 0; d < nd; ++d) {
            const bool ok = fits_dm(c, d);
            fl[d] = ok ? kCand[c] : 0;
            if(ok) fits[c].push_back(d);
        }
        if(fits[c].empty()) continue;
        sub.fft_len = fl.data();
        sw_codedisp* plan = nullptr;
        rc = sw_codedisp_create(&sub, stream, &plan);
        if(rc) break;
        for(int d : fits[c]) {
            float ms = 0.0f;
            rc = sw::time_gulps(plan, d_in, 1, d, 0, st, e0, e1, &ms);
            if(rc) break;
            ms *= scale;
            times[(size_t)d * kNc + c] = ms;
            if(ms < best[d]) {
                best[d] = ms;
                N[d] = kCand[c];
            }
        }
        sw_codedisp_destroy(plan);
    }
    // every DM a length fits, batched at that length with one overlap
    float best_total = 0.0f;
    for(int d = 0; d < nd; ++d) best_total += best[d];
    int best_c = -1;
    for(int c = 0; c < kNc && rc == SW_OK; ++c) {
        if(fits[c].size() < 2) continue;
        std::vector<float> dm(fits[c].size());
        std::vector<int> fl(fits[c].size(), kCand[c]), ov(fits[c].size());
        for(size_t k = 0; k < fits[c].size(); ++k) {
            dm[k] = cfg->dms[fits[c][k]];
            ov[k] = O[fits[c][k]];
        }
        sw_codedisp_config bat = sub;
        bat.ndms = (int)dm.size();
        bat.dms = dm.data();
        bat.fft_len = fl.data();
        bat.overlap = ov.data();
        bat.share_forward = 1;
        sw_codedisp* plan = nullptr;
        rc = sw_codedisp_create(&bat, stream, &plan);
        if(rc) break;
        float ms = 0.0f;
        rc = sw::time_gulps(plan, d_in, bat.ndms, 0, 1, st, e0, e1, &ms);
        sw_codedisp_destroy(plan);
        if(rc) break;
        ms *= scale;
        times[(size_t)nd * kNc + c] = ms;
        float total = ms;
        for(int d = 0; d < nd; ++d)
            if(!fits_dm(c, d)) total += best[d];
        if(total < best_total) {
            best_total = total;
            best_c = c;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(d_in);
    if(rc) return rc;
    for(int d = 0; d < nd; ++d) {
        SW_CHECK_ARG(N[d] > 0, "autotune: no FFT length fits the overlap");
        fft_len_out[d] = (best_c >= 0 && fits_dm(best_c, d)) ? kCand[best_c] : N[d];
    }
    if(share_forward_out) *share_forward_out = best_c >= 0 ? 1 : 0;
    if(ms_out) std::memcpy(ms_out, times.data(), sizeof(float) * times.size());
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
    if(plan->hist_ref) {
        sw::set_error("sw_codedisp_set_history: not available with history_in_place");
        return SW_ERR_UNSUPPORTED;
    }
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

extern "C" int sw_codedisp_window(const sw_codedisp* plan, int dm, const void** d_ring, int* ld, int* offset)
{
    SW_CHECK_ARG(plan && dm >= 0 && dm < (int)plan->dms.size(), "bad plan or DM index");
    if(d_ring) *d_ring = plan->ytm[dm];
    if(ld) *ld = plan->ylen[dm];
    if(offset) *offset = plan->wstart[dm];
    return SW_OK;
}

extern "C" int sw_codedisp_push_history(sw_codedisp* plan, const void* d_in, void* stream)
{
    SW_CHECK_ARG(plan && d_in, "null argument");
    hipStream_t st = sw::as_stream(stream);
    const int ncols = plan->F * plan->PA;
    for(auto& g : plan->groups) g.mdone = ((long long)(plan->gulps + 1) * plan->T) / g.hop;
    ++plan->gulps;
    if(plan->Oh == 0) return SW_OK;
    if(plan->hist_ref) {  // the caller keeps this gulp intact until the next run
        plan->prev = static_cast<const int8_t*>(d_in);
        return SW_OK;
    }
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
            const long long R = plan->ylen[g.dms[0]];
            p.ylen = (int)R;
            // ring index of stream output x is (x + lat) mod R
            p.y0 = (long)((m0 * g.hop + g.lat) % R);
            p.prev = plan->hist_ref ? plan->prev : nullptr;
            p.PA = plan->PA;
            p.F = plan->F;
            p.RP = g.RP;
            p.N = g.N;
            p.ndm = nb;
            p.tw  = plan->twiddles.at(g.L);
            p.twN = g.RP > 1 ? plan->twiddles.at(g.N) : nullptr;
            const int R1 = sw::reg_r1(g.N, g.RP, plan->PA);
            for(int k = 0; k < nb; ++k) {
                const int d = todo[b0 + k];
                p.chirp[k] = plan->chirp[d];
                p.out[k]   = plan->ytm[d];
                if(d_out_f32 && d_out_f32[d] && !plan->ytf[d]) {
                    // float outputs (a diagnostic path): buffer allocated on first use
                    const size_t yb = (size_t)plan->F * plan->PA * plan->ylen[d] * sizeof(float2);
                    SW_HIP_TRY(hipMalloc(&plan->ytf[d], yb));
                    SW_HIP_TRY(hipMemsetAsync(plan->ytf[d], 0, yb, st));
                }
                p.outf[k]  = (d_out_f32 && d_out_f32[d]) ? plan->ytf[d] : nullptr;
                if(g.RP > 1 && R1 == 0) p.v[k] = plan->vbuf[k];  // global-memory DIF/DIT path
            }
            const int G = plan->PA / sw::lds_shape(g.L).C;
            if(p.nseg == 0) {
                // no segment completes in this gulp
            } else if(R1 > 0) {
                const int units = g.RP == 2 ? plan->F * p.nseg * plan->PA : plan->F * p.nseg * (plan->PA / (32 / R1));
                rc = sw::launch_reg(R1, p, units, plan->reg2, st);
                if(rc) return rc;
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
            // emitted window, ring [gT mod R, +T) (two pieces if it wraps, which
            // the gulp-multiple ring length rules out), -> FTPA for the
            // beamformers unless d_out[d] is null (left in the ring)
            for(int k = 0; k < nb; ++k) {
                const int d = todo[b0 + k];
                const long long R = plan->ylen[d];
                const int start = (int)(gT % R);
                plan->wstart[d] = start;
                if(!d_out[d]) continue;
                const int n1 = (int)std::min<long long>(plan->T, R - start), n2 = plan->T - n1;
                const size_t PA = plan->PA;
                rc = sw::corner_turn(plan->ytm[d] + (size_t)start * 2, d_out[d], plan->F, plan->PA, n1, st,
                                     (int)R, plan->T);
                if(rc) return rc;
                if(n2 > 0) {
                    rc = sw::corner_turn(plan->ytm[d], static_cast<int8_t*>(d_out[d]) + (size_t)n1 * PA * 2, plan->F,
                                         plan->PA, n2, st, (int)R, plan->T);
                    if(rc) return rc;
                }
                if(p.outf[k]) {
                    rc = sw::corner_turn_f64(plan->ytf[d] + start, d_out_f32[d], plan->F, plan->PA, n1, st, (int)R,
                                             plan->T);
                    if(rc) return rc;
                    if(n2 > 0) {
                        rc = sw::corner_turn_f64(plan->ytf[d], static_cast<float2*>(d_out_f32[d]) + (size_t)n1 * PA,
                                                 plan->F, plan->PA, n2, st, (int)R, plan->T);
                        if(rc) return rc;
                    }
                }
            }
        }
    }
    return SW_OK;
}
