// codedisp.hip -- overlap-save coherent dedispersion: FFT, chirp multiply,
// inverse FFT in FP32 with FP64-generated chirps and twiddles.
//
// Reference: CoherentDedisperser (sw/src/CoherentDedisperser.cu:19-353):
// int8 -> c64, FFT, x H[dm][f][k], IFFT, /N, round half-even, wrap to int8,
// keep [O/2, N-O/2).  The chirp is DMResponse (:313-335):
//   nu0 = f_low + c*bw_c - bw_c/2, nu = k*bw_c/N for all k in [0, N),
//   phase = nu^2 (-2 pi K DM) / ((nu0 + nu) nu0^2), H = exp(i phase).
// Unlike the reference pipeline (FFT length tied to the gulp, F4/F5 of the
// design notes), the FFT length N and the overlap O are per-DM parameters
// and the stream is processed as true overlap-save: the input of channel f
// is the previous gulp's last O samples followed by the current gulp;
// segment s covers input [s(N-O), s(N-O)+N) (zero past the end) and yields
// output samples [s(N-O), (s+1)(N-O)) from its positions [O/2, N-O/2).
// Output sample j is the dedispersed input sample j - O/2.
//
// Kernels (all columns = (pol, antenna) pairs of one channel):
//  * N <= 4096: one workgroup per (channel, segment, column group) keeps C
//    columns x N points in LDS: load int8 -> forward FFT -> for each DM of the
//    batch: multiply by the chirp, inverse FFT, quantise, store.  The forward
//    FFT is shared by every DM of a launch.
//  * N > 4096 = RP x L (L = 4096): a decimation-in-frequency radix-RP stage
//    (int8 in, FP32 out), the same LDS kernel on each of the RP sub-sequences
//    (chirp permuted to match), and a decimation-in-time radix-RP stage that
//    quantises.
#include "fft.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <memory>
#include <vector>

namespace sw {
namespace {

using namespace fft;

constexpr int kNT      = 256;
constexpr int kTile    = 4096;   // C * L complex values per LDS buffer
constexpr int kMaxL    = 4096;
constexpr int kMaxDm   = 8;      // DMs per launch
constexpr double kDmConst = 4.148806423e15;   // Hz^2 pc^-1 cm^3 s

struct CdParams {
    const int8_t* cur;   // [F][T][PA] complex int8
    const int8_t* hist;  // [F][Oh][PA]
    const float2* zin;   // [F*nseg][RP][L][PA] (float input stage)
    int T, O, Oh, hop, nseg, PA, F, RP, N;
    int ndm;
    const float2* chirp[kMaxDm];  // [F][RP][L] (RP == 1: natural order)
    int8_t* out[kMaxDm];          // [F][T][PA] complex int8
    float2* outf[kMaxDm];         // optional [F][T][PA] y/N
    float2* vout[kMaxDm];         // [F*nseg][RP][L][PA] (float output stage)
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

// Map a workgroup index to (unit, column group) so that all column groups of
// a unit (which read the same input rows) run on the same XCD.
__device__ __forceinline__ void decode_block(int units, int G, int& u, int& g)
{
    const int bid = blockIdx.x;
    if((units & 7) == 0) {
        const int xcd = bid & 7, k = bid >> 3;
        u = xcd + 8 * (k / G);
        g = k % G;
    } else {
        u = bid / G;
        g = bid % G;
    }
}

template <int L, int C, bool IN_F32, bool OUT_F32>
__global__ void __launch_bounds__(kNT) codedisp_lds_kernel(CdParams p)
{
    __shared__ float2 X[C * padded(L)];
    __shared__ float2 W[C * padded(L)];
    const int G     = p.PA / C;
    const int units = p.F * p.nseg * p.RP;
    int u, g;
    decode_block(units, G, u, g);
    const int fs = u / p.RP, r = u % p.RP;
    const int f = fs / p.nseg, s = fs % p.nseg;
    const int tid = threadIdx.x;

    // ---- load C columns x L points
    for(int i = tid; i < C * L; i += kNT) {
        const int n = i / C, c = i % C;
        const int col = g * C + c;
        float2 v;
        if constexpr(IN_F32) {
            v = p.zin[(((size_t)fs * p.RP + r) * L + n) * p.PA + col];
        } else {
            const int tau = s * p.hop - p.O + n;
            if(tau < 0)
                v = load_c8(p.hist + (((size_t)f * p.Oh + p.Oh + tau) * p.PA + col) * 2);
            else if(tau < p.T)
                v = load_c8(p.cur + (((size_t)f * p.T + tau) * p.PA + col) * 2);
            else
                v = make_float2(0.0f, 0.0f);
        }
        X[c * padded(L) + pad(n)] = v;
    }
    __syncthreads();
    fft_lds<false, L, C, kNT>(X, p.tw);

    const float inv_n = 1.0f / (float)p.N;   // exact: N is a power of two
    for(int d = 0; d < p.ndm; ++d) {
        const float2* H = p.chirp[d] + ((size_t)f * p.RP + r) * L;
        for(int i = tid; i < C * L; i += kNT) {
            const int c = i / L, k = i % L;
            W[c * padded(L) + pad(k)] = cmul(X[c * padded(L) + pad(k)], H[k]);
        }
        __syncthreads();
        fft_lds<true, L, C, kNT>(W, p.tw);
        if constexpr(OUT_F32) {
            float2* V = p.vout[d];
            for(int i = tid; i < C * L; i += kNT) {
                const int n = i / C, c = i % C;
                const int col = g * C + c;
                // exp(+2 pi i r n / N)
                const float2 w = cconj(p.twN[r * n]);
                V[(((size_t)fs * p.RP + r) * L + n) * p.PA + col] = cmul(W[c * padded(L) + pad(n)], w);
            }
        } else {
            int8_t* out = p.out[d];
            float2* outf = p.outf[d];
            const int half = p.O / 2;
            for(int i = tid; i < C * p.hop; i += kNT) {
                const int m = i / C, c = i % C;
                const int j = s * p.hop + m;
                if(j < p.T) {
                    const int col = g * C + c;
                    const float2 y = cscale(W[c * padded(L) + pad(half + m)], inv_n);
                    const size_t o = ((size_t)f * p.T + j) * p.PA + col;
                    store_c8(out + o * 2, y);
                    if(outf) outf[o] = y;
                }
            }
        }
        __syncthreads();
    }
}

// Decimation-in-frequency first stage for N = RP * L:
//   Z_r[n] = exp(-2 pi i n r / N) sum_q x[n + q L] exp(-2 pi i q r / RP)
template <int RP>
__global__ void __launch_bounds__(256) codedisp_dif_kernel(CdParams p, int L)
{
    const size_t idx   = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t total = (size_t)p.F * p.nseg * L * p.PA;
    if(idx >= total) return;
    const int col  = (int)(idx % p.PA);
    const size_t t = idx / p.PA;
    const int n    = (int)(t % L);
    const int fs   = (int)(t / L);
    const int f = fs / p.nseg, s = fs % p.nseg;
    float2 v[RP];
#pragma unroll
    for(int q = 0; q < RP; ++q) {
        const int tau = s * p.hop - p.O + n + q * L;
        if(tau < 0)
            v[q] = load_c8(p.hist + (((size_t)f * p.Oh + p.Oh + tau) * p.PA + col) * 2);
        else if(tau < p.T)
            v[q] = load_c8(p.cur + (((size_t)f * p.T + tau) * p.PA + col) * 2);
        else
            v[q] = make_float2(0.0f, 0.0f);
    }
    dft<RP, false>(v);
    float2* Z = const_cast<float2*>(p.zin);
#pragma unroll
    for(int rr = 0; rr < RP; ++rr) {
        const float2 w = (rr == 0) ? make_float2(1.0f, 0.0f) : p.twN[n * rr];
        Z[(((size_t)fs * RP + rr) * L + n) * p.PA + col] = rr == 0 ? v[rr] : cmul(v[rr], w);
    }
}

// Decimation-in-time last stage: y[n + q L] = sum_r V_r[n] exp(+2 pi i q r / RP),
// then /N, quantise and keep output positions [O/2, N - O/2).
template <int RP>
__global__ void __launch_bounds__(256) codedisp_dit_kernel(CdParams p, int L, int d)
{
    const size_t idx   = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t total = (size_t)p.F * p.nseg * L * p.PA;
    if(idx >= total) return;
    const int col  = (int)(idx % p.PA);
    const size_t t = idx / p.PA;
    const int n    = (int)(t % L);
    const int fs   = (int)(t / L);
    const int f = fs / p.nseg, s = fs % p.nseg;
    const float2* V = p.vout[d];
    float2 v[RP];
#pragma unroll
    for(int rr = 0; rr < RP; ++rr) v[rr] = V[(((size_t)fs * RP + rr) * L + n) * p.PA + col];
    dft<RP, true>(v);
    const float inv_n = 1.0f / (float)p.N;
    const int half = p.O / 2;
#pragma unroll
    for(int q = 0; q < RP; ++q) {
        const int pos = n + q * L;
        if(pos >= half && pos < p.N - half) {
            const int j = s * p.hop + pos - half;
            if(j < p.T) {
                const size_t o = ((size_t)f * p.T + j) * p.PA + col;
                const float2 y = cscale(v[q], inv_n);
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

__global__ void history_kernel(int8_t* hist, const int8_t* cur, int F, int T, int Oh, int PA)
{
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t total = (size_t)F * Oh * PA;
    if(idx >= total) return;
    const int col = (int)(idx % PA);
    const size_t t = idx / PA;
    const int i = (int)(t % Oh), f = (int)(t / Oh);
    const uint16_t v =
        *reinterpret_cast<const uint16_t*>(cur + (((size_t)f * T + T - Oh + i) * PA + col) * 2);
    *reinterpret_cast<uint16_t*>(hist + idx * 2) = v;
}

// ------------------------------------------------------------------ host
template <bool IN_F32, bool OUT_F32>
int launch_lds(int L, const CdParams& p, int blocks, hipStream_t st)
{
#define SW_CD(LL)                                                                           \
    case LL:                                                                                \
        hipLaunchKernelGGL((codedisp_lds_kernel<LL, kTile / LL, IN_F32, OUT_F32>), dim3(blocks), \
                           dim3(kNT), 0, st, p);                                            \
        break;
    switch(L) {
        SW_CD(16) SW_CD(32) SW_CD(64) SW_CD(128) SW_CD(256) SW_CD(512) SW_CD(1024) SW_CD(2048)
        SW_CD(4096)
    default: set_error("codedisp: unsupported sub-FFT length"); return SW_ERR_UNSUPPORTED;
    }
#undef SW_CD
    SW_LAUNCH_CHECK();
    return SW_OK;
}

} // namespace

struct CodedispGroup {
    int N, O, L, RP, hop, nseg;
    std::vector<int> dms;  // indices into the plan's DM list
};

} // namespace sw

struct sw_codedisp {
    int F, A, PA, T;
    double low_freq, chan_bw, tsamp;
    std::vector<float> dms;
    std::vector<int> N, O;
    int Omax = 0;
    std::vector<sw::CodedispGroup> groups;
    std::vector<float2*> chirp;          // per DM
    std::map<int, float2*> twiddles;     // per length
    int8_t* hist = nullptr;              // [F][Omax][PA] complex int8
    float2* zbuf = nullptr;              // DIF stage output (N > 4096)
    std::vector<float2*> vbuf;           // per DM slot of a launch (N > 4096)
    size_t zbytes = 0;
    ~sw_codedisp()
    {
        for(auto* c : chirp) (void)hipFree(c);
        for(auto& kv : twiddles) (void)hipFree(kv.second);
        (void)hipFree(hist);
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
// dm_delay(f_low, f_low + bw_c, DM) / tsamp (sw/dedispersion_utils.cuh:20-23).
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
        SW_CHECK_ARG((N & (N - 1)) == 0 && N >= 16 && N <= 65536, "FFT length must be a power of two in [16, 65536]");
        SW_CHECK_ARG(O < N, "overlap must be smaller than the FFT length");
        p->dms.push_back(dm);
        p->N.push_back(N);
        p->O.push_back(O);
        p->Omax = std::max(p->Omax, O);
    }
    // group DMs sharing (N, O): their forward transforms are shared
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
            g.nseg = (p->T + g.hop - 1) / g.hop;
            const int C = sw::kTile / g.L;
            SW_CHECK_ARG(p->PA % C == 0, "2*nantennas must be a multiple of the column tile");
            SW_CHECK_ARG(g.RP <= 16, "FFT length too large");
            gid[key] = (int)p->groups.size();
            p->groups.push_back(g);
        }
        p->groups[gid[key]].dms.push_back(d);
    }
    // device tables
    size_t zmax = 0;
    for(auto& g : p->groups) {
        if(!sw::twiddle_table(p.get(), g.L, st)) return SW_ERR_ALLOC;
        if(g.RP > 1) {
            if(!sw::twiddle_table(p.get(), g.N, st)) return SW_ERR_ALLOC;
            zmax = std::max(zmax, (size_t)p->F * g.nseg * g.N * p->PA * sizeof(float2));
        }
    }
    p->chirp.resize(p->dms.size(), nullptr);
    for(int d = 0; d < (int)p->dms.size(); ++d) {
        const int N = p->N[d], RP = N / std::min(N, sw::kMaxL);
        if(hipMalloc(&p->chirp[d], sizeof(float2) * (size_t)p->F * N) != hipSuccess) {
            sw::set_error("codedisp: chirp allocation failed");
            return SW_ERR_ALLOC;
        }
        // DMPrefactor: -1.0f * 2 pi * K * dm (sw/dedispersion_utils.cuh:42-48)
        const double pref = -1.0f * (2 * 3.1415926535897932385) * sw::kDmConst * (double)p->dms[d];
        const size_t n = (size_t)p->F * N;
        hipLaunchKernelGGL(sw::chirp_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                           p->chirp[d], p->F, N, RP, p->low_freq, p->chan_bw, pref);
        SW_LAUNCH_CHECK();
    }
    if(p->Omax > 0) {
        const size_t hb = (size_t)p->F * p->Omax * p->PA * 2;
        SW_HIP_TRY(hipMalloc(&p->hist, hb));
        SW_HIP_TRY(hipMemsetAsync(p->hist, 0, hb, st));
    }
    if(zmax) {
        // DIF-stage output plus one inverse-stage buffer per DM slot of a launch
        size_t slots = 0;
        for(auto& g : p->groups)
            if(g.RP > 1) slots = std::max(slots, std::min<size_t>(sw::kMaxDm, g.dms.size()));
        p->zbytes = zmax;
        SW_HIP_TRY(hipMalloc(&p->zbuf, zmax));
        p->vbuf.resize(sw::kMaxDm, nullptr);
        for(size_t k = 0; k < slots; ++k) SW_HIP_TRY(hipMalloc(&p->vbuf[k], zmax));
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

extern "C" int sw_codedisp_info(const sw_codedisp* plan, int dm_idx, int* fft_len, int* overlap)
{
    SW_CHECK_ARG(plan && dm_idx >= 0 && dm_idx < (int)plan->dms.size(), "bad DM index");
    if(fft_len) *fft_len = plan->N[dm_idx];
    if(overlap) *overlap = plan->O[dm_idx];
    return SW_OK;
}

extern "C" int sw_codedisp_set_history(sw_codedisp* plan, const void* d_hist, void* stream)
{
    SW_CHECK_ARG(plan && d_hist, "null argument");
    if(plan->Omax == 0) return SW_OK;
    SW_HIP_TRY(hipMemcpyAsync(plan->hist, d_hist, (size_t)plan->F * plan->Omax * plan->PA * 2,
                              hipMemcpyDeviceToDevice, sw::as_stream(stream)));
    return SW_OK;
}

extern "C" int sw_codedisp_push_history(sw_codedisp* plan, const void* d_ftpa, void* stream)
{
    SW_CHECK_ARG(plan && d_ftpa, "null argument");
    if(plan->Omax == 0) return SW_OK;
    SW_CHECK_ARG(plan->T >= plan->Omax, "gulp shorter than the overlap");
    const size_t n = (size_t)plan->F * plan->Omax * plan->PA;
    hipLaunchKernelGGL(sw::history_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       sw::as_stream(stream), plan->hist, static_cast<const int8_t*>(d_ftpa), plan->F,
                       plan->T, plan->Omax, plan->PA);
    SW_LAUNCH_CHECK();
    return SW_OK;
}

extern "C" int sw_codedisp_run(sw_codedisp* plan, const void* d_ftpa, void* const* d_out,
                               void* const* d_out_f32, int dm_begin, int dm_count, int batched,
                               void* stream)
{
    SW_CHECK_ARG(plan && d_ftpa && d_out, "null argument");
    const int nd = (int)plan->dms.size();
    SW_CHECK_ARG(dm_begin >= 0 && dm_count >= 0 && dm_begin + dm_count <= nd, "bad DM range");
    hipStream_t st = sw::as_stream(stream);
    for(auto& g : plan->groups) {
        // DMs of this group inside [dm_begin, dm_begin + dm_count)
        std::vector<int> todo;
        for(int d : g.dms)
            if(d >= dm_begin && d < dm_begin + dm_count) todo.push_back(d);
        const int per_launch = batched ? sw::kMaxDm : 1;
        for(size_t b0 = 0; b0 < todo.size(); b0 += per_launch) {
            const int nb = (int)std::min<size_t>(per_launch, todo.size() - b0);
            sw::CdParams p{};
            p.cur = static_cast<const int8_t*>(d_ftpa);
            p.hist = plan->hist;
            p.T = plan->T;
            p.O = g.O;
            p.Oh = plan->Omax;
            p.hop = g.hop;
            p.nseg = g.nseg;
            p.PA = plan->PA;
            p.F = plan->F;
            p.RP = g.RP;
            p.N = g.N;
            p.ndm = nb;
            p.tw = plan->twiddles.at(g.L);
            p.twN = g.RP > 1 ? plan->twiddles.at(g.N) : nullptr;
            for(int k = 0; k < nb; ++k) {
                const int d = todo[b0 + k];
                SW_CHECK_ARG(d_out[d], "null output buffer");
                p.chirp[k] = plan->chirp[d];
                p.out[k] = static_cast<int8_t*>(d_out[d]);
                p.outf[k] = d_out_f32 ? static_cast<float2*>(d_out_f32[d]) : nullptr;
            }
            const int C = sw::kTile / g.L;
            const int G = plan->PA / C;
            if(g.RP == 1) {
                const int blocks = plan->F * g.nseg * G;
                int rc = sw::launch_lds<false, false>(g.L, p, blocks, st);
                if(rc) return rc;
                continue;
            }
            // N > 4096: DIF stage, LDS stage per sub-sequence, DIT stage
            for(int k = 0; k < nb; ++k) p.vout[k] = plan->vbuf[k];
            p.zin = plan->zbuf;
            const size_t items = (size_t)plan->F * g.nseg * g.L * plan->PA;
            const unsigned eb = (unsigned)((items + 255) / 256);
            switch(g.RP) {
            case 2: hipLaunchKernelGGL(sw::codedisp_dif_kernel<2>, dim3(eb), dim3(256), 0, st, p, g.L); break;
            case 4: hipLaunchKernelGGL(sw::codedisp_dif_kernel<4>, dim3(eb), dim3(256), 0, st, p, g.L); break;
            case 8: hipLaunchKernelGGL(sw::codedisp_dif_kernel<8>, dim3(eb), dim3(256), 0, st, p, g.L); break;
            default: hipLaunchKernelGGL(sw::codedisp_dif_kernel<16>, dim3(eb), dim3(256), 0, st, p, g.L); break;
            }
            SW_LAUNCH_CHECK();
            const int blocks = plan->F * g.nseg * g.RP * G;
            int rc = sw::launch_lds<true, true>(g.L, p, blocks, st);
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
    }
    return SW_OK;
}
