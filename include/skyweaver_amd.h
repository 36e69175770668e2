/*
 * skyweaver_amd.h -- C ABI of the MI355X (gfx950) implementation of the
 * skyweaver BeamformerPipeline hot path.
 *
 * Plain C: opaque handles, POD descriptors, raw device pointers and sizes.
 * Every entry point returns SW_OK (0) or a negative SW_ERR_* code; the
 * message of the last failure on the calling thread is available from
 * sw_last_error().  No exception crosses this boundary.
 *
 * Data layouts are the reference's (slowest -> fastest dimension), complex
 * int8 stored as interleaved (re, im) bytes:
 *   TAFTP  [T/256][A_in][F][256][P]     input voltages
 *   FTPA   [F][T][P][A]                 transposed / dedispersed voltages
 *   TFB    [T/ts][F][B]                 coherent beams (x4 bytes for IQUV)
 *   BTF    [S][T/ts][F]                 incoherent beam per beamset
 *   FPA    [F][P][A]                    statistics {mean, std, skew, kurt}
 *   TDB    [T'][D][B]                   incoherently dedispersed beams
 * Streams are hipStream_t passed as void* (NULL = default stream).
 *
 * Reference paths below are relative to the skyweaver repository root.
 */
#ifndef SKYWEAVER_AMD_H
#define SKYWEAVER_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SW_OK 0
#define SW_ERR_INVALID -1     /* bad argument / shape (std::runtime_error in the reference) */
#define SW_ERR_HIP -2         /* HIP runtime error (CUDA_ERROR_CHECK in the reference) */
#define SW_ERR_ALLOC -3       /* device allocation failed */
#define SW_ERR_IO -4          /* file open / read / write failure */
#define SW_ERR_DELAY_EPOCH -5 /* no valid delay model for the epoch (InvalidDelayEpoch) */
#define SW_ERR_EOF -6         /* end of delay-model file */
#define SW_ERR_UNSUPPORTED -7 /* valid request this build does not implement */

enum sw_stokes {
    SW_STOKES_I = 0,
    SW_STOKES_Q = 1,
    SW_STOKES_U = 2,
    SW_STOKES_V = 3,
    SW_STOKES_IQUV = 4
};

/* Coherent-beamformer kernel forms. */
enum sw_cb_form {
    SW_CB_PLAIN = 0, /* VALU v_dot4 int8 dot products (the reference's dp4a form) */
    SW_CB_MFMA  = 1  /* v_mfma_i32_32x32x32_i8 complex GEMM with fused epilogue */
};

/* Copies at most len-1 bytes of the last error message of this thread. */
int sw_last_error(char* buf, size_t len);
const char* sw_version(void);

/* ------------------------------------------------ device memory and streams
 * Thin wrappers over the HIP runtime this library is linked against, so a
 * caller (a test harness, a host language without HIP bindings) needs no
 * second GPU runtime in its process.  No reference counterpart: the
 * reference allocates through thrust containers (cpp/skyweaver/DescribedVector.hpp). */
int sw_device_count(int* count);
int sw_mem_info(size_t* free_bytes, size_t* total_bytes); /* current device */
int sw_malloc(void** d_ptr, size_t bytes);
int sw_free(void* d_ptr);
int sw_host_alloc(void** h_ptr, size_t bytes); /* pinned host memory */
int sw_host_free(void* h_ptr);
int sw_memcpy_htod(void* d_dst, const void* h_src, size_t bytes, void* stream);
int sw_memcpy_dtoh(void* h_dst, const void* d_src, size_t bytes, void* stream);
int sw_memcpy_dtod(void* d_dst, const void* d_src, size_t bytes, void* stream);
/* Synchronous device-to-host copy on a non-blocking stream of the calling
 * thread's own (created on the thread's first call, reused by later
 * threads once it ends); returns once the bytes are in h_dst.  It never
 * touches the legacy default stream, so a sink may call it while the
 * pipeline captures its next gulp into a hip graph.  No counterpart in the
 * reference (its writers copy with thrust on the default stream,
 * MultiFileWriter.cu:233-239). */
int sw_memcpy_dtoh_sync(void* h_dst, const void* d_src, size_t bytes);
int sw_memset(void* d_ptr, int value, size_t bytes, void* stream);
int sw_stream_create(void** stream); /* default priority */
int sw_stream_destroy(void* stream);
int sw_stream_synchronize(void* stream);
int sw_event_create(void** event);
int sw_event_destroy(void* event);
int sw_event_record(void* event, void* stream);
int sw_event_elapsed_ms(float* ms, void* start, void* stop);

/* ---------------------------------------------------------------- Transposer
 * Transposer::transpose(TAFTPVoltagesD const&, FTPAVoltagesD&, size_t
 * input_nantennas, cudaStream_t), cpp/skyweaver/Transposer.cuh:66-69,
 * cpp/skyweaver/src/Transposer.cu:117-184.  Antennas >= nant_in are written as zeros. */
int sw_transpose(const void* d_taftp, void* d_ftpa, int nheap_groups,
                 int nant_in, int nant_out, int nchans, void* stream);

/* ------------------------------------------------------ StatisticsCalculator
 * StatisticsCalculator::calculate_statistics(FTPAVoltagesD const&),
 * cpp/skyweaver/StatisticsCalculator.cuh:65, cpp/skyweaver/src/StatisticsCalculator.cu:54-130.
 * d_stats: FPA of struct {float mean, std, skew, kurtosis}.
 * d_workspace: sw_statistics_workspace_size() bytes of device scratch, or
 * NULL to use a library-owned buffer. */
size_t sw_statistics_workspace_size(int nchans, int npol, int nantennas);
int sw_statistics(const void* d_ftpa, float* d_stats, int nchans, int nsamples,
                  int npol, int nantennas, void* d_workspace, void* stream);

/* StatisticsCalculator::update_scalings(beamset_weights, nbeamsets),
 * cpp/skyweaver/src/StatisticsCalculator.cu:132-239 (host arithmetic).  All arrays are
 * host memory; outputs are [nbeamsets][nchans], indexed beamset*F + f. */
int sw_update_scalings(const float* h_stats, const float* h_beamset_weights,
                       int nbeamsets, int nchans, int npol, int nantennas,
                       int cb_tscrunch, int ib_tscrunch, float output_level,
                       float* h_cb_offsets, float* h_cb_scaling,
                       float* h_ib_offsets, float* h_ib_scaling);

/* ------------------------------------------------------ beamformer dimensions */
typedef struct sw_bf_dims {
    int nchans;      /* F */
    int nsamples;    /* T (input samples per channel) */
    int nantennas;   /* A, padded antenna count of the FTPA buffer, multiple of 4 */
    int nbeams;      /* B */
    int tscrunch;    /* CB and IB time scrunch (one value, as the reference requires) */
    int stokes;      /* enum sw_stokes */
    int nbeamsets;   /* S >= 1 */
    int ib_subtract; /* SKYWEAVER_IB_SUBTRACTION: 1 = subtract the IB from the CB */
} sw_bf_dims;

/* ------------------------------------------------------ IncoherentBeamformer
 * IncoherentBeamformer::beamform(input, output_raw, output, output_scale,
 * output_offset, antenna_weights, nbeamsets, stream),
 * cpp/skyweaver/IncoherentBeamformer.cuh:84-91, cpp/skyweaver/src/IncoherentBeamformer.cu:16-183.
 * Per-beamset antenna weights [S][A] are honoured for every beamset.
 * d_raw: BTF float (float4 for IQUV); d_out: BTF int8 (char4 for IQUV);
 * d_scale / d_offset: [S][F]; d_ant_weights: [S][A]. */
int sw_ib_beamform(const void* d_ftpa, float* d_raw, void* d_out,
                   const float* d_scale, const float* d_offset,
                   const float* d_ant_weights, const sw_bf_dims* dims,
                   void* stream);

/* -------------------------------------------------------- CoherentBeamformer
 * CoherentBeamformer::beamform(input, weights, output_scale, output_offset,
 * beamset_mapping, ib_powers, output, nbeamsets, stream),
 * cpp/skyweaver/CoherentBeamformer.cuh:84-92, cpp/skyweaver/src/CoherentBeamformer.cu:13-269.
 * d_weights: [F][B][A] complex int8; d_scale / d_offset: [S][F];
 * d_beamset_map: [B] int; d_ib_raw: BTF from sw_ib_beamform of the same
 * input; d_out: TFB int8 (char4 for IQUV).
 * form = enum sw_cb_form.  The MFMA form repacks the weights into
 * d_workspace (sw_cb_workspace_size() bytes; NULL = library-owned). */
size_t sw_cb_workspace_size(const sw_bf_dims* dims);
int sw_cb_beamform(const void* d_ftpa, const void* d_weights,
                   const float* d_scale, const float* d_offset,
                   const int* d_beamset_map, const float* d_ib_raw,
                   void* d_out, const sw_bf_dims* dims, int form,
                   void* d_workspace, void* stream);

/* The MFMA form in two steps: repack [F][B][A] weights once per weight
 * update (WeightsManager::weights, cpp/skyweaver/src/WeightsManager.cu:110-139 is where
 * the reference produces them), then beamform from the packed buffer. */
int sw_cb_pack_weights(const void* d_weights, void* d_packed, int nchans,
                       int nbeams, int nantennas, void* stream);
int sw_cb_beamform_packed(const void* d_ftpa, const void* d_packed,
                          const float* d_scale, const float* d_offset,
                          const int* d_beamset_map, const float* d_ib_raw,
                          void* d_out, const sw_bf_dims* dims, void* stream);
/* The same with the workgroup shape chosen: `waves` (2, 4 or 8) waves of 32
 * time rows per workgroup, each wave sweeping all beam tiles of its rows
 * (sw_cb_beamform_packed: 4).  The reference's launch shape is fixed
 * (cpp/skyweaver/src/CoherentBeamformer.cu:241-267); here it is a tuning
 * parameter (sw_pipeline_autotune).  Stokes I and IQUV take all three
 * shapes; Q, U and V only 4. */
int sw_cb_beamform_packed_tile(const void* d_ftpa, const void* d_packed,
                               const float* d_scale, const float* d_offset,
                               const int* d_beamset_map, const float* d_ib_raw,
                               void* d_out, const sw_bf_dims* dims, int waves, void* stream);

/* ----------------------------------------------------------- WeightsManager
 * WeightsManager::weights(delays, current_epoch, delay_epoch) and
 * generate_weights_k, cpp/skyweaver/src/WeightsManager.cu:16-87,109-139.
 * d_delays: [B][A] float3 {weight, offset (s), rate (s/s)};
 * d_chan_freqs: [F] Hz (FP64); d_weights out: [F][B][A] complex int8,
 * w = clamp(rint(127 wt (cos 2 pi phi, -sin 2 pi phi))),
 * phi = ((current - delay_epoch) rate + offset) nu_f. */
int sw_generate_weights(const float* d_delays, void* d_weights,
                        const double* d_chan_freqs, int nantennas, int nbeams,
                        int nchans, double current_epoch, double delay_epoch,
                        void* stream);

/* ------------------------------------------------------------ DelayManager
 * cpp/skyweaver/src/DelayManager.cu:16-250, file format
 * cpp/skyweaver/DelayManager.cuh:40-51: records of
 * {uint32 nbeams, uint32 nantennas, double start_epoch, double end_epoch}
 * followed by float3 [nbeams][nantennas].  open() reads the first model and
 * parses its beamsets (DelayManager::DelayManager / parse_beamsets);
 * get() = DelayManager::delays(epoch): advances to the model valid at epoch
 * (start <= epoch <= end) and writes it zero-padded to [nbeams][nantennas]
 * float3 on the host.  Errors: SW_ERR_DELAY_EPOCH (epoch before the current
 * model; InvalidDelayEpoch), SW_ERR_EOF (file exhausted), SW_ERR_INVALID
 * (too many beams/antennas, model dimensions changing), SW_ERR_IO. */
typedef struct sw_delay_model sw_delay_model;
int sw_delay_model_open(const char* path, int nbeams, int nantennas, sw_delay_model** out);
int sw_delay_model_close(sw_delay_model* model);
/* DelayManager::nbeamsets / beamset_mapping / beamset_weights: h_map [nbeams]
 * and h_weights [max_beamsets][nantennas] may be NULL. */
int sw_delay_model_beamsets(const sw_delay_model* model, int* nbeamsets, int* h_map,
                            float* h_weights, int max_beamsets);
int sw_delay_model_get(sw_delay_model* model, double epoch, float* h_delays,
                       double* model_epoch);

/* ---------------------------------------------------- IncoherentDedisperser
 * PipelineConfig::calculate_channel_frequencies,
 * cpp/skyweaver/src/PipelineConfig.cpp:188-203: centre - bw/2 + chbw/2 + c chbw. */
int sw_channel_frequencies(double centre_frequency, double bandwidth, int nchans,
                           double* h_freqs);
/* DMSampleDelay over a DM list, cpp/skyweaver/dedispersion_utils.cuh:19-40
 * (f_ref = chan_freqs[0]): h_delays [ndms][nchans] samples. */
int sw_id_delays(const float* dms, int ndms, const double* chan_freqs, int nchans,
                 double tsamp, int* h_delays, int* max_delay);
/* DedispersionPlan::add_block(std::string),
 * cpp/skyweaver/src/DedispersionPlan.cpp:17-78: "cdm", "cdm:tscrunch" or
 * "cdm:begin:end:step:tscrunch" -> coherent DM, tscrunch and the incoherent
 * DM list (h_dms may be NULL to query ndms). */
int sw_ddplan_parse_block(const char* spec, float* coherent_dm, int* tscrunch,
                          float* h_dms, int max_dms, int* ndms);

/* IncoherentDedisperser(config, dms, tscrunch) + dedisperse(tfb, tdb),
 * cpp/skyweaver/IncoherentDedisperser.cuh:11-39,
 * cpp/skyweaver/src/IncoherentDedisperser.cu:15-125.  Channel frequencies
 * are chan_freqs if given, else PipelineConfig's; the delay sampling time is
 * cb_tscrunch / (bandwidth / nchans). */
typedef struct {
    int nchans;
    double centre_frequency; /* Hz */
    double bandwidth;        /* Hz */
    const double* chan_freqs; /* optional [nchans] Hz */
    int cb_tscrunch;
    const float* dms;
    int ndms;
    int tscrunch;
    int kernel;   /* 0: matrix-core kernel where it applies (nchans % 4 == 0,
                     16 tscrunch <= 128), else the 16-bit-lane VALU kernel;
                     1: always the VALU kernel */
} sw_idisp_config;
typedef struct sw_idisp sw_idisp;
int sw_idisp_create(const sw_idisp_config* config, sw_idisp** out);
int sw_idisp_destroy(sw_idisp* plan);
/* the kernel the plan launches: "idisp_dma_kernel" (matrix cores, 32 output
 * samples per workgroup, tscrunch <= 2; rows whose width or strides are not
 * multiples of 16 bytes run idisp_mfma_kernel's 32-sample form instead),
 * "idisp_mfma_kernel" (matrix cores, 16 samples per workgroup) or
 * "idisp_kernel" (VALU) */
const char* sw_idisp_kernel(const sw_idisp* plan);
/* max_sample_delay(), delays() [ndms][nchans] and the scale sqrt(F ts);
 * any output may be NULL. */
int sw_idisp_info(const sw_idisp* plan, int* max_delay, int* h_delays, float* scale);
/* d_tfb: [nsamples][nchans][nbeams * nvec] int8 (nvec = 4: char4 IQUV);
 * d_tdb: [(nsamples - max_delay) / tscrunch][ndms][nbeams * nvec] int8.
 * nsamples must exceed max_delay and (nsamples - max_delay) be a multiple
 * of tscrunch (the reference's errors); nbeams * nvec % 4 == 0. */
int sw_idisp_dedisperse(sw_idisp* plan, const void* d_tfb, size_t nsamples, int nbeams,
                        int nvec, void* d_tdb, void* stream);
/* The same on channel-major input d_ftb: [nchans][nsamples][nbeams * nvec]
 * int8 (what the aggregation ring holds; a channel's rows are contiguous). */
int sw_idisp_dedisperse_ftb(sw_idisp* plan, const void* d_ftb, size_t nsamples, int nbeams,
                            int nvec, void* d_tdb, void* stream);

/* ------------------------------------------------------ CoherentDedisperser
 * create_coherent_dedisperser_config + CoherentDedisperser::dedisperse,
 * cpp/skyweaver/CoherentDedisperser.cuh:55-92, cpp/skyweaver/src/CoherentDedisperser.cu:19-353.
 * A plan holds, per coherent DM, the FFT length N, the overlap O (even),
 * the FP64-generated chirp tables and the overlap-save history (zeros
 * before the first gulp).  The overlap-save segments (hop N - O) lie on one
 * grid over the whole stream of gulps, so no transform is wasted on a
 * partial segment at a gulp's end.  Output sample j of the stream is the
 * dedispersed input sample j - O/2 (the reference's filter delay
 * tsamp*O/2); a gulp emits stream samples [gT - L, (g+1)T - L), where the
 * latency L = 0 when the hop divides the gulp length T and L = N - O
 * otherwise (sw_codedisp_info).  fft_len / overlap entries <= 0 (or NULL
 * arrays) select O = even-ceil of the intra-channel smear at the bottom of
 * the band and N = the smallest power of two >= max(1024, 16 O).
 * With share_forward, DMs of equal N get the largest O among them. */
/* Frequency of FFT bin k of channel c in the dedispersion chirp
 * H = exp(i pref nu^2 / ((nu0 + nu) nu0^2)), pref = -2 pi K DM, bandwidth B
 * per channel, low_freq the band's bottom edge, dnu = B / N:
 *  SW_CHIRP_REFERENCE: nu0 = low_freq + (c - 1/2) B, nu = k dnu for every k
 *    (DMResponse, cpp/skyweaver/src/CoherentDedisperser.cu:313-335; the
 *    default, and what the reference's golden output was made with);
 *  SW_CHIRP_PHYSICAL: nu0 = low_freq + (c + 1/2) B (the channel centre),
 *    nu = k dnu for k < N/2 and (k - N) dnu above (the negative baseband
 *    frequencies of an unshifted FFT), so every frequency of the channel is
 *    delayed to the centre's arrival time.  No counterpart in the reference. */
enum sw_chirp { SW_CHIRP_REFERENCE = 0, SW_CHIRP_PHYSICAL = 1 };
typedef struct sw_codedisp_config {
    int nchans;        /* F */
    int nantennas;     /* A (FTPA padded antenna count) */
    int nsamples;      /* T per gulp */
    double low_freq;   /* Hz, bottom edge of the band (centre - bw/2) */
    double chan_bw;    /* Hz, channel bandwidth */
    double tsamp;      /* s, channel sampling interval */
    int ndms;
    const float* dms;  /* [ndms] pc cm^-3 */
    const int* fft_len;/* [ndms] or NULL */
    const int* overlap;/* [ndms] or NULL */
    int input_fpat;    /* 0: inputs and history are FTPA; 1: time-major FPAT */
    int history_in_place; /* 1 (with input_fpat): the history is read from the
                             previous gulp's buffer, which the caller keeps
                             unchanged until the next sw_codedisp_run; no
                             history copy.  Ignored when the history spans
                             more than one gulp. */
    int share_forward;    /* 1: DMs of equal FFT length take one overlap, the
                             largest of their own, so that they lie on one
                             segment grid and a batched run computes each
                             segment's forward transform once for all of them */
    int chirp;            /* enum sw_chirp: which frequency FFT bin k stands for */
} sw_codedisp_config;
typedef struct sw_codedisp sw_codedisp;

int sw_codedisp_create(const sw_codedisp_config* cfg, void* stream, sw_codedisp** plan);
int sw_codedisp_destroy(sw_codedisp* plan);
/* FFT-length autotuner (no reference counterpart; the reference uses one
 * length for the whole gulp).  For every DM of cfg, with the overlap the
 * plan would choose, times one gulp of synthetic input through the
 * dedisperser on min(tune_chans, nchans) channels for each length in
 * {1024, ..., 32768} that fits it, one DM per launch; then, for each
 * length, all DMs it fits as one batched launch sharing the
 * forward transforms (share_forward), with the other DMs at their own best
 * length.  The faster plan is written out: fft_len_out[ndms] and
 * *share_forward_out (NULL allowed; 1 when the batched plan won).  ms_out
 * (NULL allowed) receives [ndms + 1][6] times per gulp scaled to nchans:
 * rows 0 .. ndms-1 the single-DM times, row ndms the batched time of each
 * length (-1: not tried).  A length fits a DM when N >= 4 O (at most a
 * quarter of each segment is overlap) and N <= nsamples (a longer transform
 * delays the output by more than a gulp); when every length with
 * N >= 4 O is longer than the gulp, the shortest of them fits. */
int sw_codedisp_autotune(const sw_codedisp_config* cfg, int tune_chans, int* fft_len_out,
                         int* share_forward_out, float* ms_out, void* stream);
int sw_codedisp_info(const sw_codedisp* plan, int dm_idx, int* fft_len, int* overlap,
                     int* latency);
/* Dedisperse the gulp d_ftpa (FTPA [F][T][2][A]) for the DMs
 * [dm_begin, dm_begin + dm_count): d_out[d] receives FTPA int8 for DM d
 * (NULL: the output stays in the plan's time-major ring, sw_codedisp_window);
 * d_out_f32 (NULL or per-DM NULL allowed) receives the unquantised y/N as
 * float (re, im) pairs.  batched = 1 shares each forward FFT between up to 8
 * DMs of equal (N, O) in one launch; batched = 0 launches one DM at a time.
 * Both give identical results.  Every DM of a gulp is run before
 * sw_codedisp_push_history, which ends the gulp. */
int sw_codedisp_run(sw_codedisp* plan, const void* d_ftpa, void* const* d_out,
                    void* const* d_out_f32, int dm_begin, int dm_count, int batched,
                    void* stream);
/* The last run's int8 output of DM dm in the plan's time-major ring:
 * sample t of column (f, p, a) at ring + ((f*2A + p*A + a) * ld + offset + t) * 2
 * (ld a multiple of the gulp length, offset a multiple of it).  Valid until
 * the next run. */
int sw_codedisp_window(const sw_codedisp* plan, int dm, const void** d_ring, int* ld, int* offset);
/* End the gulp d_ftpa: advance the overlap-save history and the segment
 * grid, and carry outputs computed ahead of this gulp's window. */
int sw_codedisp_push_history(sw_codedisp* plan, const void* d_ftpa, void* stream);
/* Replace the history ([F][H][2][A] complex int8; H = the largest overlap
 * when every DM's hop divides the gulp length, else the largest FFT length
 * among DMs with latency). */
int sw_codedisp_set_history(sw_codedisp* plan, const void* d_hist, void* stream);

/* ----------------------------------------------------- time-major layout
 * Inside the pipeline a gulp is held time-major, FPAT [F][P][A][T] complex
 * int8, so that the dedisperser's transforms read contiguous time runs.
 * No reference counterpart (the reference keeps FTPA and transposes per
 * channel inside cuFFT's strided plans, cpp/skyweaver/src/CoherentDedisperser.cu:106-131). */
int sw_transpose_fpat(const void* d_taftp, void* d_fpat, int nheap_groups, int nant_in,
                      int nant_out, int nchans, void* stream);
/* [F][rows][cols] -> [F][cols][rows] of 2-byte elements (FTPA <-> FPAT). */
int sw_corner_turn(const void* d_src, void* d_dst, int nchans, int rows, int cols, void* stream);
/* Statistics of an FPAT gulp; same results as sw_statistics on FTPA. */
int sw_statistics_fpat(const void* d_fpat, float* d_stats, int nchans, int nsamples, int npol,
                       int nantennas, void* d_workspace, void* stream);
/* sw_transpose_fpat and sw_statistics_fpat in one pass over the input: the
 * gulp is read once, transposed to FPAT and its exact power sums taken on
 * the way (results identical to the two calls).  Fuses
 * Transposer::transpose (cpp/skyweaver/Transposer.cuh:66-69) with
 * StatisticsCalculator::calculate_statistics
 * (cpp/skyweaver/StatisticsCalculator.cuh:65), which the reference runs as
 * two passes (cpp/skyweaver/detail/BeamformerPipeline.cu:157-167).
 * d_workspace: NULL (library scratch) or at least
 * max(8, ceil(nheap_groups/16)) * nchans * 2 * nant_out * 32 bytes. */
int sw_transpose_fpat_stats(const void* d_taftp, void* d_fpat, float* d_stats, int nheap_groups,
                            int nant_in, int nant_out, int nchans, void* d_workspace, void* stream);
/* sw_transpose_fpat_stats that allocates nothing: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The power-sum partials live in a fixed 16 MiB device array that each GPU
 * holds from library load until the process ends; results are identical to
 * sw_transpose_fpat_stats.  The partials take 32 bytes per column
 * (nchans * 2 * nant_out) and chunk of up to 64 heaps, i.e. at most
 * 524288 / ceil(nheap_groups / 64) columns (32768 columns at 1024 heaps);
 * larger shapes are rejected with SW_ERR_INVALID.  Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU (other streams,
 * other threads) race and must be serialised by the caller. */
int sw_transpose_fpat_stats_noalloc(const void* d_taftp, void* d_fpat, float* d_stats, int nheap_groups,
                                    int nant_in, int nant_out, int nchans, void* stream);

/* ---------------------------------------------------- BeamformerPipeline
 * BeamformerPipeline<CBH, IBH, StatsH, Traits>: ctor(PipelineConfig const&,
 * handlers), init(ObservationHeader const&, utc_offset),
 * operator()(TAFTPVoltagesH<char2> const&), cpp/skyweaver/BeamformerPipeline.cuh:48-67,
 * cpp/skyweaver/detail/BeamformerPipeline.cu:44-301.  Per gulp: transpose, statistics,
 * scalings (first gulp), coherent dedispersion of every DM, then per DM the
 * incoherent and coherent beams and the handler calls.  All buffers are
 * allocated at creation; processing runs on one stream owned by the
 * pipeline and returns after the gulp completes. */
enum sw_block_kind {
    SW_BLOCK_CB_TFB = 0,    /* coherent beams  [T/ts][F][B] int8 (x4 IQUV)   */
    SW_BLOCK_IB_BTF = 1,    /* incoherent beam [S][T/ts][F] int8 (x4 IQUV)   */
    SW_BLOCK_STATS_FPA = 2, /* statistics      [F][P][A] {mean,std,skew,kurt} */
    SW_BLOCK_TDB = 3        /* incoherently dedispersed [T'][D][B] int8      */
};

/* Metadata of a block handed to a sink (the DescribedVector fields the
 * reference handlers read, cpp/skyweaver/DescribedVector.hpp:97-573). */
typedef struct sw_block {
    int kind;              /* enum sw_block_kind */
    const void* data;      /* valid only during the sink call */
    int on_device;         /* 1: device pointer, 0: host pointer */
    size_t bytes;
    int ndims;
    size_t dims[4];        /* slowest -> fastest */
    char order[8];         /* "TFB", "BTF", "FPA", "TDB" */
    int dm_idx;            /* coherent DM index */
    float reference_dm;
    double tsamp;          /* s */
    double utc_offset;     /* s, relative to the observation's UTC_START */
    const float* dms;      /* incoherent DMs of a TDB block, else NULL */
    int ndms;
} sw_block;

typedef int (*sw_sink_fn)(void* user, const sw_block* block);

/* ------------------------------------------ IncoherentDedispersionPipeline
 * IncoherentDedispersionPipeline(config, handler) + AggregationBuffer,
 * cpp/skyweaver/detail/IncoherentDedispersionPipeline.cu:6-113,
 * cpp/skyweaver/detail/AggregationBuffer.cu:23-116.  One block per
 * dedispersion-plan entry (coherent DM, tscrunch, incoherent DMs).  Each
 * block aggregates pushed TFB samples ([T][F][B] int8, x4 for IQUV) in a
 * device ring of dispatch + max_delay samples, with dispatch =
 * max(10 max_delay, 8192) rounded down to a multiple of tscrunch (or
 * dispatch_size if > 0); a full ring is dedispersed on the GPU and handed to
 * the sink as a device TDB block, then its last max_delay samples move to
 * the front. */
typedef struct {
    int nchans;
    double centre_frequency;  /* Hz */
    double bandwidth;         /* Hz */
    const double* chan_freqs; /* optional [nchans] */
    int cb_tscrunch;
    int nbeams;
    int nvec;                 /* 1: int8, 4: char4 (IQUV) */
    int nblocks;
    const float* coherent_dms;   /* [nblocks] */
    const int* tscrunch;         /* [nblocks] */
    const float* const* dms;     /* [nblocks][ndms[i]] incoherent DMs */
    const int* ndms;             /* [nblocks] */
    int dispatch_size;           /* 0: the reference's rule */
} sw_idpipe_config;
typedef struct sw_idpipe sw_idpipe;
int sw_idpipe_create(const sw_idpipe_config* config, sw_idpipe** out);
int sw_idpipe_destroy(sw_idpipe* pipeline);
int sw_idpipe_set_sink(sw_idpipe* pipeline, sw_sink_fn sink, void* user);
int sw_idpipe_info(const sw_idpipe* pipeline, int block, int* max_delay, size_t* dispatch_size,
                   size_t* filled);
/* IncoherentDedispersionPipeline::operator()(TFB block, dm_idx): tsamp and
 * utc_offset describe the pushed TFB samples (the first one's time). */
int sw_idpipe_push(sw_idpipe* pipeline, int block, const void* tfb, size_t nsamples,
                   int on_device, double utc_offset, double tsamp, void* stream);

typedef struct sw_pipeline_config {
    int nantennas;          /* A (padded; multiple of 4) */
    int nchans;             /* F */
    int nbeams;             /* B */
    int nsamples;           /* gulp length T, multiple of 256 */
    int tscrunch;           /* coherent and incoherent beam tscrunch */
    int stokes;             /* enum sw_stokes */
    int ib_subtract;        /* 1: subtract the incoherent beam from the coherent beams */
    float output_level;     /* PipelineConfig::output_level (default 24) */
    double centre_frequency;/* Hz, centre of the subband */
    double bandwidth;       /* Hz, bandwidth of the subband */
    int total_nchans;       /* channels of the full band (timestamp step), e.g. 4096 */
    int ndms;
    const float* coherent_dms;
    const int* fft_len;     /* per DM or NULL (see sw_codedisp_config) */
    const int* overlap;     /* per DM or NULL */
    int cb_form;            /* enum sw_cb_form */
    int batched_dms;        /* 1: one launch per group of DMs with equal FFT length
                               and overlap, sharing each segment's forward FFT;
                               0: one DM per launch (identical results) */
    /* Incoherent dedispersion of the coherent beams (PipelineConfig::
     * enable_incoherent_dedispersion and the DedispersionPlan blocks): when
     * id_enable is set, every coherent DM's TFB output is also pushed into an
     * IncoherentDedispersionPipeline whose TDB blocks go to the TDB sink.
     * Per coherent DM: id_tscrunch[i] (NULL: 1) and the incoherent DM list
     * id_dms[i][id_ndms[i]] (NULL: the coherent DM alone). */
    int id_enable;
    const int* id_tscrunch;
    const float* const* id_dms;
    const int* id_ndms;
    int id_dispatch_size;   /* 0: max(10 max_delay, 8192) */
    /* 1: when fft_len is NULL, pick each coherent DM's FFT length (and
     * share_forward) with sw_codedisp_autotune (on 8 channels) at create time */
    int autotune_fft;
    /* sw_codedisp_config::share_forward for the pipeline's dedisperser */
    int share_forward;
    /* 1: the CB, IB and statistics sinks are called by a library thread, in
     * the same order as synchronous sinks, once the blocks are written, while
     * the pipeline goes on with the next DMs and gulps; each DM's outputs and
     * the statistics alternate between two buffers (a block stays valid until
     * its sink returns; the pipeline waits before reusing a buffer whose block
     * is not yet delivered).  A sink's failure is returned by the next
     * sw_pipeline_process / process_prefetched / synchronize call, which
     * also drops the blocks queued after it.  sw_pipeline_synchronize waits
     * until every queued block has been delivered.  0: the sinks are called
     * synchronously within sw_pipeline_process, as the reference's handlers
     * (cpp/skyweaver/detail/BeamformerPipeline.cu:240-257). */
    int async_sinks;
    /* 1 (needs async_sinks): after the first gulp, each gulp's device work
     * (transpose with the statistics' sums, dedispersion of every DM, every
     * DM's beams) is captured into a hip graph and launched as one; the
     * executable graph is updated in place from gulp to gulp.  The blocks are
     * released to the sinks, and the incoherent dedispersion is fed, after
     * the graph; results are identical to the stream launches.  Stage times
     * then report the whole graph as "other".  The sinks run on the
     * library's thread while the next gulp is captured, so they must not
     * use the legacy default (null) stream, which a capture forbids: copy
     * blocks out with sw_memcpy_dtoh_sync or on a non-blocking stream. */
    int graph;
    /* Coherent beamformer workgroup shape (sw_cb_beamform_packed_tile): 2, 4
     * or 8 waves of 32 time rows; 0: the tuning table's entry for this
     * workload, else 4.  When fft_len is NULL and autotune_fft is 0, a table
     * entry with the same DM list also gives the FFT lengths and
     * share_forward. */
    int cb_waves;
    const char* tuning_file; /* NULL: $SKYWEAVER_AMD_TUNING, else tuning/gfx950.json
                                in the library's package directory */
    int chirp;               /* enum sw_chirp of the coherent dedisperser */
} sw_pipeline_config;

typedef struct sw_obs_header {
    int nantennas;          /* antennas present in the TAFTP input (<= nantennas) */
    double utc_start;       /* unix seconds */
    double sample_clock;    /* Hz */
    double utc_offset;      /* s; BeamformerPipeline::init's utc_offset (file start offset) */
} sw_obs_header;

typedef struct sw_pipeline sw_pipeline;

int sw_pipeline_create(const sw_pipeline_config* cfg, sw_pipeline** pipeline);
int sw_pipeline_destroy(sw_pipeline* pipeline);
int sw_pipeline_set_sinks(sw_pipeline* pipeline, sw_sink_fn cb, void* cb_user, sw_sink_fn ib,
                          void* ib_user, sw_sink_fn stats, void* stats_user);
/* Static beamforming weights [F][B][A] complex int8 (host), used for every
 * gulp until replaced. */
int sw_pipeline_set_weights(sw_pipeline* pipeline, const void* h_weights);
/* Weights from a delay-model file instead (the reference's DelayManager +
 * WeightsManager path, cpp/skyweaver/detail/BeamformerPipeline.cu:133-147):
 * beamsets come from the first model; on every gulp the model valid at the
 * gulp's unix time utc_start + utc_offset + n 2 total_nchans T / sample_clock
 * is read and the weights are generated on the device for that time. */
int sw_pipeline_set_delay_file(sw_pipeline* pipeline, const char* path);
/* TDB blocks of the incoherent dedispersion (id_enable). */
int sw_pipeline_set_tdb_sink(sw_pipeline* pipeline, sw_sink_fn tdb, void* tdb_user);
/* Beamsets: nbeamsets, beam -> beamset map [B], antenna weights [S][A]. */
int sw_pipeline_set_beamsets(sw_pipeline* pipeline, int nbeamsets, const int* h_map,
                             const float* h_antenna_weights);
int sw_pipeline_init(sw_pipeline* pipeline, const sw_obs_header* header);
/* One gulp of TAFTP [T/256][A_in][F][256][2] complex int8; on_device = 1 if
 * taftp is a device pointer (no host transfer).  Returns 0 (the reference's
 * operator() always returns false). */
int sw_pipeline_process(sw_pipeline* pipeline, const void* taftp, size_t bytes, int on_device);
/* The same split in two so that the host-to-device copy of the next gulp
 * overlaps the processing of the current one (the reference copies inside
 * operator(), cpp/skyweaver/detail/BeamformerPipeline.cu:281-288, and
 * overlaps only the file read, cpp/skyweaver/src/skyweaver_cli.cu:196-220).
 * sw_pipeline_prefetch starts an asynchronous copy of a host gulp into one
 * of two device gulp buffers (allocated on the first call) on the
 * pipeline's copy stream and returns; at most two gulps may be pending.
 * The host buffer must stay unchanged until the sw_pipeline_process_prefetched
 * call that processes it returns (pinned memory, sw_host_alloc, makes the
 * copy asynchronous).  sw_pipeline_process_prefetched processes the oldest
 * pending gulp exactly as sw_pipeline_process does. */
int sw_pipeline_prefetch(sw_pipeline* pipeline, const void* h_taftp, size_t bytes);
int sw_pipeline_process_prefetched(sw_pipeline* pipeline);
/* Wait until every copy and kernel the pipeline has enqueued has finished
 * (for a caller about to free host buffers it prefetched from). */
int sw_pipeline_synchronize(sw_pipeline* pipeline);
/* Per-stage device times of the last gulp (ms), in the order transpose,
 * statistics, dedispersion, incoherent beam, coherent beam, incoherent
 * dedispersion (with its aggregation), weights (delay-model path), other;
 * n <= SW_NSTAGES entries are written.  Fused work is timed with the kernel
 * that does it: the statistics' power sums with the transpose and, for
 * Stokes I with 0/1 antenna weights, the incoherent beam with the coherent
 * beams (the statistics and incoherent-beam stages then read about 0). */
#define SW_NSTAGES 8
int sw_pipeline_stage_times(const sw_pipeline* pipeline, float* ms, int n);
/* The processing stream (hipStream_t) the pipeline enqueues work on. */
void* sw_pipeline_stream(const sw_pipeline* pipeline);
/* The tiles the pipeline runs with: the coherent beamformer's waves per
 * workgroup, each coherent DM's FFT length (fft_len [ndms], NULL allowed),
 * share_forward, and where they came from (*source: 0 defaults, 1 the
 * tuning table's beamformer entry, 2 its entry with FFT lengths too, 3 the
 * FFT autotuner at create). */
int sw_pipeline_tile(const sw_pipeline* pipeline, int* cb_waves, int* fft_len, int* share_forward, int* source);

/* ------------------------------------------------------------- tuning
 * Tile autotuner (no reference counterpart: the reference's launch shapes
 * are fixed, cpp/skyweaver/src/CoherentBeamformer.cu:241-267, and its FFT
 * length is the gulp).  For the workload of cfg on min(tune_chans, nchans)
 * channels of synthetic device input, times the coherent and incoherent
 * beams with 2, 4 and 8 waves per workgroup and, when cfg->fft_len is NULL,
 * runs sw_codedisp_autotune; writes the result as one JSON object (an
 * "entries" item of the tuning table, with the measured times) into json
 * (json_len bytes, NUL-terminated). */
int sw_pipeline_autotune(const sw_pipeline_config* cfg, int tune_chans, char* json, size_t json_len);
/* The tuning table's entry for a workload (path NULL: the default table):
 * *found = 0 none, 1 an entry for antennas, beams, nsamples, tscrunch and
 * stokes (cb_waves set), 2 one whose DM list equals dms (fft_len [ndms] and
 * share_forward set too).  Host only: needs no GPU. */
int sw_tuning_lookup(const char* path, int nantennas, int nbeams, int nsamples, int tscrunch, int stokes, int ndms,
                     const float* dms, int* found, int* cb_waves, int* fft_len, int* share_forward);

/* -------------------------------------------------------- microbenchmarks
 * Roofline peaks of the device, measured: which = 0 dense int8 MFMA (TOP/s),
 * 1 HBM copy of 1 GiB (GB/s, read + write), 2 FP32 FMA (TFLOP/s).  No
 * reference counterpart. */
int sw_microbench(int which, double* result, double* seconds, void* stream);

/* -------------------------------------------------- synthetic input (bench)
 * Seeded Gaussian complex int8: xoshiro256** seeded through splitmix64,
 * Box-Muller in FP64, value = clip(round_half_away(sigma*z), -127, 127).
 * Stream k of the generator starts at k xoshiro256** jumps (2^128 steps)
 * from the seeded state; chunk c of 2^20 complex values uses stream c. */
int sw_synth_gaussian_c8(int8_t* h_out, size_t ncomplex, uint64_t seed,
                         double sigma, int nthreads);
/* Random-phase weights (lround(127 cos th), lround(127 sin th)), th ~ U[0, 2pi). */
int sw_synth_phase_weights(int8_t* h_out, size_t ncomplex, uint64_t seed);

#ifdef __cplusplus
}
#endif
#endif /* SKYWEAVER_AMD_H */
