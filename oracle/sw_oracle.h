/*
 * sw_oracle.h -- CPU oracle for the skyweaver BeamformerPipeline hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the CPU
 * reference implementations that skyweaver's own unit tests use as their
 * oracles, plus an FP64 overlap-save coherent dedisperser (the reference
 * has no CPU version of that stage).  It is the checker, never the thing
 * measured or shipped: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.
 *
 * Every function cites the reference file:line it follows.  Paths are
 * relative to the reference repository root.
 *
 * Floating-point contract: compiled with -ffp-contract=off (no FMA
 * contraction) and IEEE semantics, like the host compilation of the
 * reference's test oracles.
 *
 * Pinning status of each function (what fixes it beyond this restatement):
 *   orc_transpose          known-answer i%113 / i%107 (TransposerTester.cu:95-122)
 *                          + brute-force index map in the tests
 *   orc_ib_beamform        closed forms (all-ones weights, single antenna,
 *                          zero input) + int64 brute force of the Stokes sums
 *   orc_cb_beamform        closed forms (single antenna w=127 => power =
 *                          127^2 |v|^2; zero input => 0) + int64 brute force
 *                          of the complex sums
 *   orc_statistics         exact integer power sums (int8 inputs) at 1e-5
 *   orc_update_scalings    closed form for unit std / uniform weights
 *   orc_incoherent_dedisperse  all-ones => trunc(sqrt(F*ts)); DM 0 => zero
 *                          delays; too-few-samples error
 *                          (IncoherentDedisperserTester.cu:39-125)
 *   orc_fft / orc_ifft     brute-force long-double DFT, Parseval, delta/tone
 *   orc_codedisp_channel   golden pair codedisp_{input,output}_DM1000_1chan.bin
 *                          at +-1 (CoherentDedisperserTester.cu:113-204);
 *                          DM 0 identity (CoherentDedisperserTester.cu:69-111)
 *   orc_weights            rint-half-even vs round-half-away agreement and the
 *                          closed form of zero delays (WeightsManagerTester.cu:99-145)
 *   orc_parse_beamsets     fixture test_delays.bin: 2 beamsets
 *                          (DelayManagerTester.cu:87-93)
 */
#ifndef SW_ORACLE_H
#define SW_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Stokes modes, cpp/skyweaver/beamformer_utils.cuh:108 plus Full (IQUV). */
enum { ORC_STOKES_I = 0, ORC_STOKES_Q = 1, ORC_STOKES_U = 2, ORC_STOKES_V = 3, ORC_STOKES_IQUV = 4 };

/* TAFTP -> FTPA transpose with zero padding of antennas >= nant_in.
 * cpp/skyweaver/test/src/TransposerTester.cu:30-72.  Arrays are interleaved complex
 * int8 (re, im).  The output must be pre-zeroed by the caller for the
 * padded antennas (reference: thrust resize value-initialises). */
void orc_transpose(const int8_t* taftp, int8_t* ftpa, int nheaps, int nant_in,
                   int nant_out, int nchans, int nsamps_per_heap, int npol);

/* Incoherent beamformer, cpp/skyweaver/test/src/IncoherentBeamformerTester.cu:41-103.
 * ftpa [F][T][2][A] complex int8; raw [S][T/ts][F] float (x4 for IQUV);
 * out [S][T/ts][F] int8 (x4 for IQUV); scale/offset [S*F];
 * beamset_weights [S*A]. */
void orc_ib_beamform(const int8_t* ftpa, float* raw, int8_t* out, int nchans,
                     int tscrunch, int nsamples, int nantennas,
                     const float* scale, const float* offset,
                     const float* beamset_weights, int nbeamsets, int stokes);

/* Coherent beamformer, cpp/skyweaver/test/src/CoherentBeamformerTester.cu:43-152.
 * weights [F][B][A] complex int8; out TFB [T/ts][F][B] int8 (x4 for IQUV);
 * scales/offsets [S*F]; antenna_weights [S*A]; beamset_mapping [B].
 * The incoherent power used for the IB subtraction is recomputed here in
 * the oracle's own loop order, as the reference oracle does. */
void orc_cb_beamform(const int8_t* ftpa, const int8_t* weights, int8_t* out,
                     int nchans, int tscrunch, int nsamples, int nbeams,
                     int nantennas, const float* scales, const float* offsets,
                     const float* antenna_weights, const int* beamset_mapping,
                     int nbeamsets, int stokes, int ib_subtract);

/* Per-(f,p,a) Welford statistics, cpp/skyweaver/test/src/StatisticsCalculatorTester.cu:25-99,
 * 120-171.  out [F][P][A][4] doubles = {mean, std, skew, kurtosis}. */
void orc_statistics(const int8_t* ftpa, double* out, int nchans, int nsamples,
                    int npol, int nantennas);

/* StatisticsCalculator::update_scalings, cpp/skyweaver/src/StatisticsCalculator.cu:132-239,
 * with the per-beamset index (beamset*F + f) the reference omits.
 * stds [F][P][A] float; beamset_weights [S][A]; outputs [S*F]. */
void orc_update_scalings(const float* stds, const float* beamset_weights,
                         int nbeamsets, int nchans, int npol, int nantennas,
                         int cb_tscrunch, int ib_tscrunch, float output_level,
                         float* cb_offsets, float* cb_scaling,
                         float* ib_offsets, float* ib_scaling);

/* IncoherentDedisperser::prepare, cpp/skyweaver/src/IncoherentDedisperser.cu:36-54 with
 * DMSampleDelay, cpp/skyweaver/dedispersion_utils.cuh:20-40.  delays [D][F]; returns
 * the maximum delay. */
int orc_id_delays(const float* dms, int ndms, const double* chan_freqs,
                  int nchans, double tsamp, int* delays);

/* IncoherentDedisperser::dedisperse, cpp/skyweaver/src/IncoherentDedisperser.cu:67-125.
 * tfb [T][F][B] int8 (x4 for IQUV, nvec=4); out [(T-maxd)/ts][D][B] (xnvec).
 * Returns 0, or -1 when T <= max_delay, -2 when (T-maxd) % ts != 0. */
int orc_incoherent_dedisperse(const int8_t* tfb, long nsamples, int nchans,
                              int nbeams, int nvec, const int* delays, int ndms,
                              int max_delay, int tscrunch, int8_t* out,
                              int nthreads);

/* Unnormalised in-place complex FFT in FP64 (forward: exp(-2 pi i jk/N)).
 * data [N][2]; N a power of two.  Iterative radix-2. */
void orc_fft(double* data, int n, int inverse);

/* Chirp H[c][k], cpp/skyweaver/src/CoherentDedisperser.cu:313-335 and
 * DMPrefactor cpp/skyweaver/dedispersion_utils.cuh:42-48.  out [nchans][N][2] FP64. */
void orc_dm_chirp(float dm, double low_freq, double coarse_chan_bw, int nchans,
                  int fft_len, double* out);
/* The physical mapping of the same chirp (no reference counterpart): channel
 * c's centre nu0 = low_freq + (c + 1/2) B, bin k at nu = k dnu for k < N/2 and
 * (k - N) dnu above.  Pinned by the cold-plasma delay law (its phase slope)
 * and by re-compressing a pulse dispersed in the time domain. */
void orc_dm_chirp_physical(float dm, double low_freq, double coarse_chan_bw, int nchans,
                           int fft_len, double* out);

/* FP64 overlap-save coherent dedispersion of one channel.
 * tpa [nin][pa] complex int8 (pa = P*A columns); chirp [N][2].
 * Output sample j corresponds to input sample j + O/2: segment s covers
 * input [s(N-O), s(N-O)+N) (zero beyond nin); its samples [O/2, N-O/2) are
 * kept (cpp/skyweaver/src/CoherentDedisperser.cu:196-221).  out_f [nout][pa][2]
 * doubles (y/N) may be NULL; out_i8 [nout][pa][2] = wrap8(rint(y/N)) may be
 * NULL. */
void orc_codedisp_channel(const int8_t* tpa, long nin, int pa,
                          const double* chirp, int N, int O, long nout,
                          double* out_f, int8_t* out_i8, int nthreads);

/* Weight generation, cpp/skyweaver/src/WeightsManager.cu:16-87 (the kernel semantics:
 * round-half-even and clamp to int8) with the host oracle's loop
 * (cpp/skyweaver/test/src/WeightsManagerTester.cu:33-71).
 * delays [B][A][3] floats {weight, offset, rate}; out [F][B][A] complex int8. */
void orc_weights(const float* delays, int8_t* out, const double* chan_freqs,
                 int nantennas, int nbeams, int nchans, double current_epoch,
                 double delay_epoch);

/* DelayManager::parse_beamsets, cpp/skyweaver/src/DelayManager.cu:162-232.
 * model [nbeams_model][nant_model][3]; beamset_map [nbeams_cfg];
 * beamset_weights [max_beamsets][nant_cfg].  Returns nbeamsets. */
int orc_parse_beamsets(const float* model, int nbeams_model, int nant_model,
                       int nbeams_cfg, int nant_cfg, int* beamset_map,
                       float* beamset_weights, int max_beamsets);

#ifdef __cplusplus
}
#endif
#endif /* SW_ORACLE_H */
