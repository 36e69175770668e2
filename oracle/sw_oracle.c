/*
 * sw_oracle.c -- CPU oracle for the skyweaver BeamformerPipeline hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see sw_oracle.h): the checker for the HIP
 * library, never part of it.  Build with -ffp-contract=off.
 */
#include "sw_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#if defined(__FP_FAST_FMA) && !defined(SW_ORACLE_ALLOW_FMA_MACRO)
/* The macro only says the target has FMA; contraction itself is disabled by
 * -ffp-contract=off in the Makefile.  Nothing to do. */
#endif

/* ------------------------------------------------------------------------ */
/* Complex float helpers: cuCmulf / cuCaddf / cuConjf semantics (textbook),
 * cpp/skyweaver/beamformer_utils.cuh:103-106 for the squared magnitude. */
typedef struct { float x, y; } cf32;

static inline cf32 cf_mul(cf32 a, cf32 b)
{
    cf32 r;
    r.x = (a.x * b.x) - (a.y * b.y);
    r.y = (a.x * b.y) + (a.y * b.x);
    return r;
}
static inline cf32 cf_add(cf32 a, cf32 b)
{
    cf32 r;
    r.x = a.x + b.x;
    r.y = a.y + b.y;
    return r;
}
static inline cf32 cf_conj(cf32 a)
{
    cf32 r;
    r.x = a.x;
    r.y = -a.y;
    return r;
}
static inline float cf_mag2(cf32 a) { return a.x * a.x + a.y * a.y; }

/* calculate_stokes<S>, cpp/skyweaver/beamformer_utils.cuh:125-140 */
static inline float stokes_param(int s, cf32 p0, cf32 p1)
{
    switch(s) {
    case ORC_STOKES_I: return cf_mag2(p0) + cf_mag2(p1);
    case ORC_STOKES_Q: return cf_mag2(p0) - cf_mag2(p1);
    case ORC_STOKES_U: return 2 * cf_mul(p0, cf_conj(p1)).x;
    default: return 2 * cf_mul(p0, cf_conj(p1)).y;
    }
}

/* SingleStokesBeamformerTraits::clamp, cpp/skyweaver/beamformer_utils.cuh:186-195 */
static inline int8_t clamp_i8(float v)
{
    return (int8_t)fmaxf(-128.0f, fminf(127.0f, v));
}

/* rescale, cpp/skyweaver/beamformer_utils.cuh:174-184 */
static inline float rescale_param(int s, float power, float offset, float scale)
{
    if(s == ORC_STOKES_I)
        return rintf((power - offset) / scale);
    return rintf(power / scale);
}

/* ib_subtract, cpp/skyweaver/beamformer_utils.cuh:165-172 */
static inline float ib_subtract_param(float power, float ib_power, float mult,
                                      float scale)
{
    return rintf((power - ib_power * mult) / scale);
}

static inline cf32 load_c8(const int8_t* p)
{
    cf32 r;
    r.x = (float)p[0];
    r.y = (float)p[1];
    return r;
}

/* ------------------------------------------------------------------------ */
void orc_transpose(const int8_t* taftp, int8_t* ftpa, int nheaps, int nant_in,
                   int nant_out, int nchans, int nsamps_per_heap, int npol)
{
    /* cpp/skyweaver/test/src/TransposerTester.cu:37-71 */
    const size_t tp   = (size_t)nsamps_per_heap * npol;
    const size_t ftp  = (size_t)nchans * tp;
    const size_t aftp = (size_t)nant_in * ftp;
    const size_t pa   = (size_t)npol * nant_out;
    const size_t tpa  = (size_t)nsamps_per_heap * nheaps * pa;
    for(int h = 0; h < nheaps; ++h)
        for(int a = 0; a < nant_in; ++a)
            for(int c = 0; c < nchans; ++c)
                for(int s = 0; s < nsamps_per_heap; ++s)
                    for(int p = 0; p < npol; ++p) {
                        size_t in_idx = h * aftp + a * ftp + c * tp +
                                        (size_t)s * npol + p;
                        size_t t       = (size_t)h * nsamps_per_heap + s;
                        size_t out_idx = c * tpa + t * pa + (size_t)p * nant_out + a;
                        ftpa[2 * out_idx]     = taftp[2 * in_idx];
                        ftpa[2 * out_idx + 1] = taftp[2 * in_idx + 1];
                    }
}

/* ------------------------------------------------------------------------ */
void orc_ib_beamform(const int8_t* ftpa, float* raw, int8_t* out, int nchans,
                     int tscrunch, int nsamples, int nantennas,
                     const float* scale, const float* offset,
                     const float* beamset_weights, int nbeamsets, int stokes)
{
    /* cpp/skyweaver/test/src/IncoherentBeamformerTester.cu:55-102 (fscrunch = 1) */
    const int nvec         = (stokes == ORC_STOKES_IQUV) ? 4 : 1;
    const int nsamples_out = nsamples / tscrunch;
    const size_t a         = (size_t)nantennas;
    const size_t pa        = 2 * a;
    const size_t tpa       = (size_t)nsamples * pa;
    for(int bs = 0; bs < nbeamsets; ++bs)
        for(int f = 0; f < nchans; ++f)
            for(int T = 0; T < nsamples; T += tscrunch) {
                float power[4] = {0.0f, 0.0f, 0.0f, 0.0f};
                for(int t = T; t < T + tscrunch; ++t)
                    for(int ai = 0; ai < nantennas; ++ai) {
                        float w  = beamset_weights[bs * nantennas + ai];
                        cf32 p0  = load_c8(ftpa + 2 * (f * tpa + t * pa + ai));
                        cf32 p1  = load_c8(ftpa + 2 * (f * tpa + t * pa + a + ai));
                        if(nvec == 1) {
                            power[0] += stokes_param(stokes, p0, p1) * w;
                        } else {
                            for(int s = 0; s < 4; ++s)
                                power[s] += stokes_param(s, p0, p1) * w;
                        }
                    }
                size_t oidx = (size_t)bs * nsamples_out * nchans +
                              (size_t)(T / tscrunch) * nchans + f;
                int so = bs * nchans + f;
                for(int s = 0; s < nvec; ++s) {
                    int sp = (nvec == 1) ? stokes : s;
                    raw[oidx * nvec + s] = power[s];
                    out[oidx * nvec + s] = clamp_i8(
                        rescale_param(sp, power[s], offset[so], scale[so]));
                }
            }
}

/* ------------------------------------------------------------------------ */
void orc_cb_beamform(const int8_t* ftpa, const int8_t* weights, int8_t* out,
                     int nchans, int tscrunch, int nsamples, int nbeams,
                     int nantennas, const float* scales, const float* offsets,
                     const float* antenna_weights, const int* beamset_mapping,
                     int nbeamsets, int stokes, int ib_subtract)
{
    /* cpp/skyweaver/test/src/CoherentBeamformerTester.cu:59-151 (fscrunch = 1).
     * Per beam the float operations are performed in exactly the reference
     * oracle's order: complex sums over antennas in ascending order, Stokes
     * accumulated over the tscrunch window in ascending time.  The loops are
     * arranged so the beam index is innermost (independent lanes), and the
     * incoherent power, which the reference recomputes per beam with an
     * identical loop, is computed once per beamset. */
    const int nvec = (stokes == ORC_STOKES_IQUV) ? 4 : 1;
    const size_t A = (size_t)nantennas, B = (size_t)nbeams;
    float* wr = (float*)malloc(sizeof(float) * A * B);
    float* wi = (float*)malloc(sizeof(float) * A * B);
    float* acc0r = (float*)malloc(sizeof(float) * B);
    float* acc0i = (float*)malloc(sizeof(float) * B);
    float* acc1r = (float*)malloc(sizeof(float) * B);
    float* acc1i = (float*)malloc(sizeof(float) * B);
    float* power = (float*)malloc(sizeof(float) * B * 4);
    float* ibp   = (float*)malloc(sizeof(float) * (size_t)nbeamsets * 4);
    for(int f = 0; f < nchans; ++f) {
        for(size_t b = 0; b < B; ++b)
            for(size_t ai = 0; ai < A; ++ai) {
                const int8_t* w = weights + 2 * ((size_t)f * B * A + b * A + ai);
                wr[ai * B + b]  = (float)w[0];
                wi[ai * B + b]  = (float)w[1];
            }
        for(int s0 = 0; s0 < nsamples; s0 += tscrunch) {
            memset(power, 0, sizeof(float) * B * 4);
            memset(ibp, 0, sizeof(float) * (size_t)nbeamsets * 4);
            for(int so = 0; so < tscrunch; ++so) {
                const int8_t* v =
                    ftpa + 2 * (A * 2 * (size_t)nsamples * f + A * 2 * (size_t)(s0 + so));
                /* incoherent power of each beamset, a ascending */
                for(int bs = 0; bs < nbeamsets; ++bs)
                    for(size_t ai = 0; ai < A; ++ai) {
                        cf32 p0 = load_c8(v + 2 * ai);
                        cf32 p1 = load_c8(v + 2 * (A + ai));
                        float w = antenna_weights[bs * A + ai];
                        if(nvec == 1)
                            ibp[bs * 4] += stokes_param(stokes, p0, p1) * w;
                        else
                            for(int s = 0; s < 4; ++s)
                                ibp[bs * 4 + s] += stokes_param(s, p0, p1) * w;
                    }
                for(size_t b = 0; b < B; ++b) {
                    acc0r[b] = acc0i[b] = acc1r[b] = acc1i[b] = 0.0f;
                }
                for(size_t ai = 0; ai < A; ++ai) {
                    const float v0r = (float)v[2 * ai], v0i = (float)v[2 * ai + 1];
                    const float v1r = (float)v[2 * (A + ai)];
                    const float v1i = (float)v[2 * (A + ai) + 1];
                    const float* wra = wr + ai * B;
                    const float* wia = wi + ai * B;
                    for(size_t b = 0; b < B; ++b) {
                        /* cuCaddf(acc, cuCmulf(w, v)) */
                        float p0r = (wra[b] * v0r) - (wia[b] * v0i);
                        float p0i = (wra[b] * v0i) + (wia[b] * v0r);
                        float p1r = (wra[b] * v1r) - (wia[b] * v1i);
                        float p1i = (wra[b] * v1i) + (wia[b] * v1r);
                        acc0r[b] = acc0r[b] + p0r;
                        acc0i[b] = acc0i[b] + p0i;
                        acc1r[b] = acc1r[b] + p1r;
                        acc1i[b] = acc1i[b] + p1i;
                    }
                }
                for(size_t b = 0; b < B; ++b) {
                    cf32 p0 = {acc0r[b], acc0i[b]};
                    cf32 p1 = {acc1r[b], acc1i[b]};
                    if(nvec == 1)
                        power[b * 4] += stokes_param(stokes, p0, p1);
                    else
                        for(int s = 0; s < 4; ++s)
                            power[b * 4 + s] += stokes_param(s, p0, p1);
                }
            }
            const size_t out_base = (size_t)(s0 / tscrunch) * B * nchans + (size_t)f * B;
            for(size_t b = 0; b < B; ++b) {
                const int bs  = beamset_mapping[b];
                const int sco = bs * nchans + f;
                for(int s = 0; s < nvec; ++s) {
                    int sp = (nvec == 1) ? stokes : s;
                    float r;
                    if(ib_subtract)
                        r = ib_subtract_param(power[b * 4 + s], ibp[bs * 4 + s],
                                              16129.0f, scales[sco]);
                    else
                        r = rescale_param(sp, power[b * 4 + s], offsets[sco], scales[sco]);
                    out[(out_base + b) * nvec + s] = clamp_i8(r);
                }
            }
        }
    }
    free(wr); free(wi); free(acc0r); free(acc0i); free(acc1r); free(acc1i);
    free(power); free(ibp);
}

/* ------------------------------------------------------------------------ */
typedef struct {
    long long n;
    double M1, M2, M3, M4;
} running_stats;

/* RunningStats::push, cpp/skyweaver/test/src/StatisticsCalculatorTester.cu:54-69 */
static inline void rs_push(running_stats* s, double x)
{
    double delta, delta_n, delta_n2, term1;
    long long n1 = s->n;
    s->n++;
    long long n = s->n;
    delta    = x - s->M1;
    delta_n  = delta / n;
    delta_n2 = delta_n * delta_n;
    term1    = delta * delta_n * n1;
    s->M1 += delta_n;
    s->M4 += term1 * delta_n2 * (n * n - 3 * n + 3) + 6 * delta_n2 * s->M2 -
             4 * delta_n * s->M3;
    s->M3 += term1 * delta_n * (n - 2) - 3 * delta_n * s->M2;
    s->M2 += term1;
}

void orc_statistics(const int8_t* ftpa, double* out, int nchans, int nsamples,
                    int npol, int nantennas)
{
    /* cpp/skyweaver/test/src/StatisticsCalculatorTester.cu:120-153 */
    const size_t pa = (size_t)npol * nantennas;
    running_stats* st = (running_stats*)calloc(pa, sizeof(running_stats));
    for(int f = 0; f < nchans; ++f) {
        memset(st, 0, pa * sizeof(running_stats));
        const int8_t* base = ftpa + 2 * (size_t)f * nsamples * pa;
        for(int t = 0; t < nsamples; ++t)
            for(size_t i = 0; i < pa; ++i) {
                const int8_t* d = base + 2 * ((size_t)t * pa + i);
                rs_push(&st[i], (double)d[0]);
                rs_push(&st[i], (double)d[1]);
            }
        for(size_t i = 0; i < pa; ++i) {
            running_stats* s = &st[i];
            double* o = out + 4 * ((size_t)f * pa + i);
            /* RunningStats accessors, :71-99 */
            o[0] = s->M1;
            o[1] = sqrt(s->M2 / (s->n - 1.0));
            o[2] = sqrt((double)s->n) * s->M3 / pow(s->M2, 1.5);
            o[3] = (double)s->n * s->M4 / (s->M2 * s->M2) - 3.0;
        }
    }
    free(st);
}

/* ------------------------------------------------------------------------ */
void orc_update_scalings(const float* stds, const float* beamset_weights,
                         int nbeamsets, int nchans, int npol, int nantennas,
                         int cb_tscrunch, int ib_tscrunch, float output_level,
                         float* cb_offsets, float* cb_scaling,
                         float* ib_offsets, float* ib_scaling)
{
    /* cpp/skyweaver/src/StatisticsCalculator.cu:140-229.  std::pow(float, int) promotes
     * to double; std::sqrt(float) stays in float. */
    const float weights_amp = 127.0f;
    const int pa = npol * nantennas;
    for(int bs = 0; bs < nbeamsets; ++bs)
        for(int f = 0; f < nchans; ++f) {
            float sum = 0.0f, count = 0.0f;
            for(int p = 0; p < npol; ++p)
                for(int a = 0; a < nantennas; ++a) {
                    const float weight = beamset_weights[nantennas * bs + a];
                    sum += weight * stds[f * pa + p * nantennas + a];
                    count += weight;
                }
            const float avg_std = sum / count;
            const float eff_nant = count / (float)npol;
            const int idx = bs * nchans + f;
            {
                float x     = weights_amp * avg_std * sqrtf(eff_nant);
                float scale = (float)pow((double)x, 2);
                float dof   = (float)(2 * cb_tscrunch * npol);
                cb_offsets[idx] = scale * dof;
                cb_scaling[idx] = scale * sqrtf(2 * dof) / output_level;
            }
            {
                float scale = (float)pow((double)avg_std, 2);
                float dof   = (float)(2 * ib_tscrunch) * eff_nant * (float)npol;
                ib_offsets[idx] = scale * dof;
                ib_scaling[idx] = scale * sqrtf(2 * dof) / output_level;
            }
        }
}

/* ------------------------------------------------------------------------ */
#define ORC_DM_CONST_HZ_S 4.148806423e15 /* cpp/skyweaver/dedispersion_utils.cuh:14-16 */

int orc_id_delays(const float* dms, int ndms, const double* chan_freqs,
                  int nchans, double tsamp, int* delays)
{
    /* DMSampleDelay(dm, f_ref = lowest channel, tsamp)(f),
     * cpp/skyweaver/dedispersion_utils.cuh:19-40; dm_delay = K (1/f_ref^2 - 1/f^2) dm */
    int maxd = 0;
    const double f_ref = chan_freqs[0];
    for(int d = 0; d < ndms; ++d)
        for(int c = 0; c < nchans; ++c) {
            const double f  = chan_freqs[c];
            const double dd = (double)dms[d];
            double delay = ORC_DM_CONST_HZ_S * (1 / (f_ref * f_ref) - 1 / (f * f)) * dd;
            int v = (int)(delay / tsamp + 0.5);
            delays[d * nchans + c] = v;
            if(v > maxd)
                maxd = v;
        }
    return maxd;
}

int orc_incoherent_dedisperse(const int8_t* tfb, long nsamples, int nchans,
                              int nbeams, int nvec, const int* delays, int ndms,
                              int max_delay, int tscrunch, int8_t* out,
                              int nthreads)
{
    /* cpp/skyweaver/src/IncoherentDedisperser.cu:67-125.  Sums are exact integers in
     * float, so accumulation order is immaterial; the division by the float
     * sqrt(F*ts) and the clamp-then-truncate are reproduced exactly. */
    if(nsamples <= max_delay)
        return -1;
    if((nsamples - max_delay) % tscrunch != 0)
        return -2;
    const float scale_factor = (float)sqrt((double)((size_t)nchans * tscrunch));
    const size_t bf = (size_t)nbeams * nchans;
    const size_t bv = (size_t)nbeams * nvec;
#ifdef _OPENMP
    if(nthreads > 0)
        omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
    {
        float* powers = (float*)malloc(sizeof(float) * bv);
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for(long t = max_delay; t < nsamples; t += tscrunch) {
            size_t t_out = (size_t)(t - max_delay) / tscrunch * bv * ndms;
            for(int d = 0; d < ndms; ++d) {
                memset(powers, 0, sizeof(float) * bv);
                for(int ts = 0; ts < tscrunch; ++ts)
                    for(int f = 0; f < nchans; ++f) {
                        size_t idx = ((size_t)(t + ts - delays[d * nchans + f]) * bf +
                                      (size_t)f * nbeams) * nvec;
                        for(size_t b = 0; b < bv; ++b)
                            powers[b] += (float)tfb[idx + b];
                    }
                int8_t* o = out + t_out + (size_t)d * bv;
                for(size_t b = 0; b < bv; ++b)
                    o[b] = clamp_i8(powers[b] / scale_factor);
            }
        }
        free(powers);
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
void orc_fft(double* data, int n, int inverse)
{
    /* Iterative radix-2 decimation-in-time FFT: bit-reversal permutation,
     * then log2(n) butterfly passes with twiddles exp(-+2 pi i k / len). */
    for(int i = 1, j = 0; i < n; ++i) {
        int bit = n >> 1;
        for(; j & bit; bit >>= 1)
            j ^= bit;
        j ^= bit;
        if(i < j) {
            double tr = data[2 * i], ti = data[2 * i + 1];
            data[2 * i]     = data[2 * j];
            data[2 * i + 1] = data[2 * j + 1];
            data[2 * j]     = tr;
            data[2 * j + 1] = ti;
        }
    }
    const double sgn = inverse ? 1.0 : -1.0;
    for(int len = 2; len <= n; len <<= 1) {
        const int half = len >> 1;
        for(int k = 0; k < half; ++k) {
            /* twiddle from long double for a correctly rounded table */
            long double ang = (long double)sgn * 2.0L *
                              3.14159265358979323846264338327950288L *
                              (long double)k / (long double)len;
            const double wr = (double)cosl(ang), wi = (double)sinl(ang);
            for(int i = 0; i < n; i += len) {
                double* u = data + 2 * (i + k);
                double* v = data + 2 * (i + k + half);
                double xr = v[0] * wr - v[1] * wi;
                double xi = v[0] * wi + v[1] * wr;
                v[0] = u[0] - xr;
                v[1] = u[1] - xi;
                u[0] = u[0] + xr;
                u[1] = u[1] + xi;
            }
        }
    }
}

void orc_dm_chirp(float dm, double low_freq, double coarse_chan_bw, int nchans,
                  int fft_len, double* out)
{
    /* DMPrefactor: -1.0f * 2 pi * K * dm (double), dedispersion_utils.cuh:42-48;
     * DMResponse, cpp/skyweaver/src/CoherentDedisperser.cu:313-335. */
    const double two_pi = 2 * 3.1415926535897932385;
    const double pref   = -1.0f * two_pi * ORC_DM_CONST_HZ_S * (double)dm;
    const double fine_chan_bw = coarse_chan_bw / fft_len;
    for(int c = 0; c < nchans; ++c) {
        const double nu_0 = low_freq + c * coarse_chan_bw - 0.5f * coarse_chan_bw;
        for(int k = 0; k < fft_len; ++k) {
            const double nu = k * fine_chan_bw;
            const double phase_prefactor = nu * nu * pref;
            const double phase = phase_prefactor / ((nu_0 + nu) * nu_0 * nu_0);
            double* o = out + 2 * ((size_t)c * fft_len + k);
            o[0] = cos(phase);
            o[1] = sin(phase);
        }
    }
}

void orc_dm_chirp_physical(float dm, double low_freq, double coarse_chan_bw, int nchans,
                           int fft_len, double* out)
{
    /* same prefactor and phase expression as orc_dm_chirp; nu0 is the
     * channel centre and the upper half of the bins are negative offsets */
    const double two_pi = 2 * 3.1415926535897932385;
    const double pref   = -1.0f * two_pi * ORC_DM_CONST_HZ_S * (double)dm;
    const double fine_chan_bw = coarse_chan_bw / fft_len;
    for(int c = 0; c < nchans; ++c) {
        const double nu_0 = low_freq + c * coarse_chan_bw + 0.5 * coarse_chan_bw;
        for(int k = 0; k < fft_len; ++k) {
            const double nu = (2 * k >= fft_len ? k - fft_len : k) * fine_chan_bw;
            const double phase = nu * nu * pref / ((nu_0 + nu) * nu_0 * nu_0);
            double* o = out + 2 * ((size_t)c * fft_len + k);
            o[0] = cos(phase);
            o[1] = sin(phase);
        }
    }
}

static inline int8_t wrap8(long v)
{
    /* static_cast<char>(int): modulo 2^8, cpp/skyweaver/src/CoherentDedisperser.cu:217-221 */
    return (int8_t)(uint8_t)(v & 0xFF);
}

void orc_codedisp_channel(const int8_t* tpa, long nin, int pa,
                          const double* chirp, int N, int O, long nout,
                          double* out_f, int8_t* out_i8, int nthreads)
{
    const long hop  = N - O;
    const long nseg = (nout + hop - 1) / hop;
#ifdef _OPENMP
    if(nthreads > 0)
        omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
    {
        double* buf = (double*)malloc(sizeof(double) * 2 * N);
#ifdef _OPENMP
#pragma omp for schedule(dynamic)
#endif
        for(int col = 0; col < pa; ++col) {
            for(long s = 0; s < nseg; ++s) {
                const long start = s * hop;
                for(long k = 0; k < N; ++k) {
                    long t = start + k;
                    if(t < nin) {
                        buf[2 * k]     = (double)tpa[2 * (t * pa + col)];
                        buf[2 * k + 1] = (double)tpa[2 * (t * pa + col) + 1];
                    } else {
                        buf[2 * k] = buf[2 * k + 1] = 0.0;
                    }
                }
                orc_fft(buf, N, 0);
                for(long k = 0; k < N; ++k) {
                    double xr = buf[2 * k], xi = buf[2 * k + 1];
                    double hr = chirp[2 * k], hi = chirp[2 * k + 1];
                    buf[2 * k]     = hr * xr - hi * xi;
                    buf[2 * k + 1] = hr * xi + hi * xr;
                }
                orc_fft(buf, N, 1);
                for(long j = O / 2; j < N - O / 2; ++j) {
                    long o = start + j - O / 2;
                    if(o >= nout)
                        break;
                    double yr = buf[2 * j] / N, yi = buf[2 * j + 1] / N;
                    if(out_f) {
                        out_f[2 * (o * pa + col)]     = yr;
                        out_f[2 * (o * pa + col) + 1] = yi;
                    }
                    if(out_i8) {
                        out_i8[2 * (o * pa + col)]     = wrap8(lrint(yr));
                        out_i8[2 * (o * pa + col) + 1] = wrap8(lrint(yi));
                    }
                }
            }
        }
        free(buf);
    }
}

/* ------------------------------------------------------------------------ */
void orc_weights(const float* delays, int8_t* out, const double* chan_freqs,
                 int nantennas, int nbeams, int nchans, double current_epoch,
                 double delay_epoch)
{
    /* cpp/skyweaver/test/src/WeightsManagerTester.cu:45-70 loop; the product kernel's
     * rounding (__double2int_rn = half even) and clamp to [-128, 127]
     * (cpp/skyweaver/src/WeightsManager.cu:70-79). */
    const double TWOPI = 6.283185307179586;
    for(int a = 0; a < nantennas; ++a)
        for(int b = 0; b < nbeams; ++b) {
            const float* dm = delays + 3 * ((size_t)b * nantennas + a);
            for(int c = 0; c < nchans; ++c) {
                double frequency = chan_freqs[c];
                double t = (current_epoch - delay_epoch);
                double phase = (t * (double)dm[2] + (double)dm[1]) * frequency;
                double s = sin(TWOPI * phase), co = cos(TWOPI * phase);
                double re = co * 127.0 * (double)dm[0];
                double im = -1.0 * s * 127.0 * (double)dm[0];
                int ir = (int)nearbyint(re), ii = (int)nearbyint(im);
                int8_t* o = out + 2 * (((size_t)c * nbeams + b) * nantennas + a);
                o[0] = clamp_i8((float)ir);
                o[1] = clamp_i8((float)ii);
            }
        }
}

/* ------------------------------------------------------------------------ */
int orc_parse_beamsets(const float* model, int nbeams_model, int nant_model,
                       int nbeams_cfg, int nant_cfg, int* beamset_map,
                       float* beamset_weights, int max_beamsets)
{
    /* cpp/skyweaver/src/DelayManager.cu:162-232: beamsets are runs of contiguous beams
     * whose per-antenna scalar weights (x of the float3) are identical to the
     * previous beam's. */
    for(int b = 0; b < nbeams_cfg; ++b)
        beamset_map[b] = 0;
    for(int i = 0; i < max_beamsets * nant_cfg; ++i)
        beamset_weights[i] = 0.0f;
    int nsets = 1;
    for(int a = 0; a < nant_model; ++a)
        beamset_weights[a] = model[3 * a];
    for(int b = 1; b < nbeams_model; ++b) {
        int changed = 0;
        for(int a = 0; a < nant_model; ++a) {
            float w  = model[3 * ((size_t)b * nant_model + a)];
            float wp = model[3 * ((size_t)(b - 1) * nant_model + a)];
            if(w != wp)
                changed = 1;
        }
        if(changed) {
            if(nsets >= max_beamsets)
                return -1;
            for(int a = 0; a < nant_model; ++a)
                beamset_weights[nsets * nant_cfg + a] =
                    model[3 * ((size_t)b * nant_model + a)];
            beamset_map[b] = nsets;
            ++nsets;
        } else {
            beamset_map[b] = nsets - 1;
        }
    }
    return nsets;
}
