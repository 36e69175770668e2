// pipeline.cpp -- the per-gulp BeamformerPipeline behind sw_pipeline_*.
//
// Reference: BeamformerPipeline::process / operator(),
// cpp/skyweaver/detail/BeamformerPipeline.cu:127-301.  Stage order per gulp:
//   (H2D) -> transpose -> statistics -> scalings (first gulp only) ->
//   coherent dedispersion of all DMs -> per DM: incoherent beam, coherent
//   beam, CB handler, IB handler -> stats handler.
// With a delay-model file, the delays valid at the gulp's unix time are
// fetched and the weights regenerated on the device before the transpose
// (BeamformerPipeline.cu:133-147).  With incoherent dedispersion enabled,
// every coherent DM's TFB block is also pushed into the incoherent
// dedispersion pipeline (the reference installs it as the CB handler,
// cpp/skyweaver/detail/IncoherentDedispersionPipeline.cu:101-113).
// Differences by design: every buffer is allocated once at creation, the
// stages are enqueued back to back on one stream with no intermediate host
// synchronisation (only when a sink must read a block, and at the end of
// the gulp), and the scalings are indexed per beamset (beamset*F + f).
// Kernels are fused where the data allow: the statistics' power sums are
// taken by the transpose, and for Stokes I with 0/1 antenna weights the
// beamformers read the dedisperser's time-major ring directly, the
// incoherent beam being computed inside the coherent beamformer.  A host
// gulp can be copied on a second stream while the previous one is processed
// (sw_pipeline_prefetch / sw_pipeline_process_prefetched).
#include "sw_common.hpp"

#include <algorithm>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <utility>
#include <vector>

namespace sw {
std::string default_tuning_path();
int tuning_lookup(const std::string& path, int A, int B, int T, int ts, int stokes, int ndms, const float* dms,
                  int* found, int* cb_waves, int* fft_len, int* share_forward);
int cb_mfma_supported(const sw_bf_dims& d);
size_t fpat_stats_workspace(int nchans, int npol, int nant, int nsamples);
int ib_beamform_tm(const void* d_ring, long ld, long w0, float* d_raw, void* d_out, const float* d_scale,
                   const float* d_offset, const float* d_ant_weights, const sw_bf_dims& d, hipStream_t st);
int cb_beamform_packed_tm(const void* d_ring, long ld, long w0, const void* d_packed, const float* d_scale,
                          const float* d_offset, const int* d_beamset_map, const float* d_ib_raw, void* d_out,
                          const sw_bf_dims& d, hipStream_t st, float* d_ib_out_raw, void* d_ib_out,
                          const float* d_ib_scale, const float* d_ib_offset, const float* d_ant_weights, int waves,
                          bool fastdiv);
int cb_beamform_packed_waves(const void* d_ftpa, const void* d_packed, const float* d_scale, const float* d_offset,
                             const int* d_beamset_map, const float* d_ib_raw, void* d_out, const sw_bf_dims& d,
                             hipStream_t st, int waves);
int cb_fused_ib_supported(const sw_bf_dims& d);
}

namespace {

// Asynchronous sink delivery: blocks are handed to the sinks by one library
// thread, in the order the synchronous pipeline would call them, once the
// event recorded after the kernels that wrote them has completed.  A buffer
// counts its undelivered blocks; the pipeline waits for that count to reach
// zero before it writes the buffer again.
struct SinkWorker {
    struct Job {
        sw_sink_fn fn;
        void* user;
        sw_block blk;
        hipEvent_t ready;
        const void* buf;
    };
    std::thread th;
    std::mutex mu;
    std::condition_variable cv_job, cv_done;
    std::deque<Job> q;
    std::map<const void*, int> pending;  // buffer -> blocks not yet delivered
    bool stop = false;
    int rc = SW_OK;  // first failure (later blocks are dropped)
    std::string msg;

    void start() { th = std::thread([this] { run(); }); }
    void run()
    {
        std::unique_lock<std::mutex> lk(mu);
        for(;;) {
            cv_job.wait(lk, [this] { return stop || !q.empty(); });
            if(q.empty()) return;  // stop requested and drained
            Job j = q.front();
            q.pop_front();
            const bool failed = rc != SW_OK;
            lk.unlock();
            int r = SW_OK;
            std::string m;
            if(!failed) {
                const hipError_t e = hipEventSynchronize(j.ready);
                if(e != hipSuccess) {
                    r = SW_ERR_HIP;
                    m = std::string("sink worker: hipEventSynchronize failed: ") + hipGetErrorString(e);
                } else if(const int h = j.fn(j.user, &j.blk); h != 0) {
                    r = SW_ERR_IO;
                    m = "sw_pipeline_process: handler returned " + std::to_string(h);
                }
            }
            lk.lock();
            if(r != SW_OK && rc == SW_OK) {
                rc = r;
                msg = m;
            }
            if(--pending[j.buf] == 0) pending.erase(j.buf);
            cv_done.notify_all();
        }
    }
    void push(const Job& j)
    {
        std::lock_guard<std::mutex> lk(mu);
        q.push_back(j);
        ++pending[j.buf];
        cv_job.notify_one();
    }
    // wait until no block of `buf` is waiting for its sink
    void wait_free(const void* buf)
    {
        std::unique_lock<std::mutex> lk(mu);
        cv_done.wait(lk, [&] { return pending.find(buf) == pending.end(); });
    }
    void drain()
    {
        std::unique_lock<std::mutex> lk(mu);
        cv_done.wait(lk, [&] { return q.empty() && pending.empty(); });
    }
    // the first failure, reported once on the calling thread
    int take_error()
    {
        std::lock_guard<std::mutex> lk(mu);
        const int r = rc;
        if(r != SW_OK) {
            sw::set_error(msg);
            rc = SW_OK;
        }
        return r;
    }
    ~SinkWorker()
    {
        if(th.joinable()) {
            {
                std::lock_guard<std::mutex> lk(mu);
                stop = true;
            }
            cv_job.notify_all();
            th.join();
        }
    }
};

} // namespace

struct sw_pipeline {
    sw_pipeline_config cfg{};
    std::vector<float> dms;
    std::vector<int> fft_len, overlap;
    hipStream_t stream = nullptr;
    sw_codedisp* cd    = nullptr;
    // device buffers
    void* taftp_d      = nullptr;
    int8_t* ftpa       = nullptr;  // this gulp, FPAT
    int8_t* ftpa_prev  = nullptr;  // the previous gulp: the dedisperser reads its history there
    float* stats       = nullptr;
    void* stats_ws     = nullptr;
    std::vector<int8_t*> ftpa_dd;
    float* ib_raw      = nullptr;
    int8_t* ib_q       = nullptr;
    int8_t* cb_out     = nullptr;
    int8_t* weights    = nullptr;
    uint8_t* packed    = nullptr;
    float* scl         = nullptr;  // cb_offsets | cb_scaling | ib_offsets | ib_scaling, [4][S][F]
    int* bmap          = nullptr;
    float* antw        = nullptr;
    int nbeamsets      = 1;
    std::vector<int> h_bmap;
    std::vector<float> h_antw;
    std::vector<float> h_stats;
    bool weights_set   = false;
    bool use_mfma      = false;
    bool scalings_dirty = true;    // recompute the scalings from this gulp's statistics
    bool cb_fastdiv = false;       // every coherent-beam scale lies where the beamformer may divide with div_by
    int cb_waves = 4;              // MFMA beamformer workgroup shape (waves of 32 time rows)
    int tile_source = 0;           // sw_pipeline_tile's *source
    std::string tuning_file;
    // sinks
    sw_sink_fn cb_sink = nullptr, ib_sink = nullptr, st_sink = nullptr;
    void *cb_user = nullptr, *ib_user = nullptr, *st_user = nullptr;
    // delay-model path
    sw_delay_model* delays = nullptr;
    float* d_delays        = nullptr;  // [B][A] float3
    double* d_freqs        = nullptr;  // [F]
    std::vector<float> h_delays;
    double loaded_epoch    = NAN;
    // incoherent dedispersion of the coherent beams
    sw_idpipe* idp         = nullptr;
    std::vector<int> id_ts;
    std::vector<std::vector<float>> id_dms;
    // state
    sw_obs_header header{};
    bool inited        = false;
    size_t call_count  = 0;
    double tick_per_block = 0;
    hipEvent_t ev[6]{};
    hipEvent_t ev_w[2]{};            // weights generation
    std::vector<hipEvent_t> ev_dm;  // per DM: ib start, cb start, cb end, id end
    float stage_ms[SW_NSTAGES]{};
    // asynchronous sinks (cfg.async_sinks): the coherent and incoherent beam
    // outputs of each DM and the statistics have two buffers, used by
    // alternate gulps, and an event each recorded once they are written
    std::unique_ptr<SinkWorker> worker;
    std::vector<int8_t*> cb_bufs, ib_bufs;  // [dm][2]
    float* stats_bufs[2]{};
    std::vector<hipEvent_t> ev_ready;       // [dm][2], then the statistics' [2]
    // the ready events are recorded on this stream, never captured, after a
    // wait for the pipeline stream: the sink thread's hipEventSynchronize
    // then never names an event of a stream that is capturing a graph
    hipStream_t sig_stream = nullptr;
    hipEvent_t ev_link     = nullptr;
    // graph mode (cfg.graph): the executable graph of the last gulp
    hipGraphExec_t gexec = nullptr;
    int graph_instantiations = 0;
    // prefetched input (sw_pipeline_prefetch): two device gulp buffers filled
    // on a copy stream while the pipeline stream works on the previous gulp
    hipStream_t copy_stream = nullptr;
    void* pre_d[2]{};
    hipEvent_t pre_copied[2]{}, pre_consumed[2]{};
    bool pre_used[2]{};    // pre_consumed recorded at least once
    int pre_head = 0, pre_count = 0;

    size_t ftpa_bytes() const { return (size_t)cfg.nchans * cfg.nsamples * 2 * cfg.nantennas * 2; }
    int nv() const { return cfg.stokes == SW_STOKES_IQUV ? 4 : 1; }
    size_t tout() const { return (size_t)cfg.nsamples / cfg.tscrunch; }

    ~sw_pipeline()
    {
        if(stream) (void)hipStreamSynchronize(stream);
        if(gexec) (void)hipGraphExecDestroy(gexec);
        if(worker) {
            worker->drain();
            worker.reset();
        }
        if(!cb_bufs.empty()) {  // cb_out, ib_q and stats point into these
            for(auto* b : cb_bufs) (void)hipFree(b);
            for(auto* b : ib_bufs) (void)hipFree(b);
            for(auto* b : stats_bufs) (void)hipFree(b);
            cb_out = nullptr;
            ib_q = nullptr;
            stats = nullptr;
        }
        for(auto e : ev_ready) if(e) (void)hipEventDestroy(e);
        if(sig_stream) {
            (void)hipStreamSynchronize(sig_stream);
            (void)hipStreamDestroy(sig_stream);
        }
        if(ev_link) (void)hipEventDestroy(ev_link);
        if(cd) sw_codedisp_destroy(cd);
        if(idp) sw_idpipe_destroy(idp);
        if(delays) sw_delay_model_close(delays);
        void* bufs[] = {taftp_d, ftpa, ftpa_prev, stats, stats_ws, ib_raw, ib_q, cb_out, weights, packed, scl, bmap, antw,
                        d_delays, d_freqs};
        for(void* b : bufs) (void)hipFree(b);
        for(auto* b : ftpa_dd) (void)hipFree(b);
        for(auto e : ev) if(e) (void)hipEventDestroy(e);
        for(auto e : ev_w) if(e) (void)hipEventDestroy(e);
        for(auto e : ev_dm) if(e) (void)hipEventDestroy(e);
        if(copy_stream) (void)hipStreamSynchronize(copy_stream);
        for(int i = 0; i < 2; ++i) {
            if(pre_d[i]) (void)hipFree(pre_d[i]);
            if(pre_copied[i]) (void)hipEventDestroy(pre_copied[i]);
            if(pre_consumed[i]) (void)hipEventDestroy(pre_consumed[i]);
        }
        if(copy_stream) (void)hipStreamDestroy(copy_stream);
        if(stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

sw_bf_dims bf_dims(const sw_pipeline* p)
{
    sw_bf_dims d{};
    d.nchans      = p->cfg.nchans;
    d.nsamples    = p->cfg.nsamples;
    d.nantennas   = p->cfg.nantennas;
    d.nbeams      = p->cfg.nbeams;
    d.tscrunch    = p->cfg.tscrunch;
    d.stokes      = p->cfg.stokes;
    d.nbeamsets   = p->nbeamsets;
    d.ib_subtract = p->cfg.ib_subtract;
    return d;
}

int upload_beamsets(sw_pipeline* p)
{
    const int A = p->cfg.nantennas, B = p->cfg.nbeams;
    if(p->bmap) (void)hipFree(p->bmap);
    if(p->antw) (void)hipFree(p->antw);
    if(p->scl) (void)hipFree(p->scl);
    p->bmap = nullptr;
    p->antw = nullptr;
    p->scl  = nullptr;
    SW_HIP_TRY(hipMalloc(&p->bmap, sizeof(int) * B));
    SW_HIP_TRY(hipMalloc(&p->antw, sizeof(float) * p->nbeamsets * A));
    SW_HIP_TRY(hipMalloc(&p->scl, sizeof(float) * 4 * p->nbeamsets * p->cfg.nchans));
    SW_HIP_TRY(hipMemcpyAsync(p->bmap, p->h_bmap.data(), sizeof(int) * B, hipMemcpyHostToDevice, p->stream));
    SW_HIP_TRY(hipMemcpyAsync(p->antw, p->h_antw.data(), sizeof(float) * p->nbeamsets * A,
                              hipMemcpyHostToDevice, p->stream));
    SW_HIP_TRY(hipStreamSynchronize(p->stream));
    // the scalings depend on the beamset weights: recompute them from the
    // next gulp's statistics (the gulp count, the pipeline's time base, is
    // left alone)
    p->scalings_dirty = true;
    return SW_OK;
}

int call_sink(sw_pipeline* p, sw_sink_fn fn, void* user, sw_block& blk)
{
    if(!fn) return SW_OK;
    SW_HIP_TRY(hipStreamSynchronize(p->stream));
    const int rc = fn(user, &blk);
    if(rc != 0) {
        sw::set_error("sw_pipeline_process: handler returned " + std::to_string(rc));
        return SW_ERR_IO;
    }
    return SW_OK;
}

void fill_block(sw_block& b, int kind, const void* data, size_t bytes, const char* order,
                std::initializer_list<size_t> dims)
{
    std::memset(&b, 0, sizeof(b));
    b.kind      = kind;
    b.data      = data;
    b.on_device = 1;
    b.bytes     = bytes;
    b.ndims     = (int)dims.size();
    int i = 0;
    for(size_t v : dims) b.dims[i++] = v;
    std::strncpy(b.order, order, sizeof(b.order) - 1);
}

} // namespace

extern "C" int sw_pipeline_create(const sw_pipeline_config* cfg, sw_pipeline** out)
{
    SW_CHECK_ARG(cfg && out, "null argument");
    *out = nullptr;
    const sw_pipeline_config& c = *cfg;
    SW_CHECK_ARG(c.ndms > 0 && c.coherent_dms, "No DDPlan set for pipeline");
    SW_CHECK_ARG(c.nsamples > 0 && c.nsamples % 256 == 0,
                 "Gulp size is not a multiple of the number of samples per heap");
    SW_CHECK_ARG(c.nantennas > 0 && c.nantennas % 4 == 0, "nantennas must be a multiple of 4");
    SW_CHECK_ARG(c.nchans > 0 && c.nbeams > 0, "empty dimensions");
    SW_CHECK_ARG(c.tscrunch > 0 && c.nsamples % c.tscrunch == 0, "gulp must be a multiple of tscrunch");
    SW_CHECK_ARG(c.bandwidth > 0 && c.centre_frequency > 0, "bad band description");
    SW_CHECK_ARG(!c.graph || c.async_sinks, "graph mode needs asynchronous sinks (per-DM output buffers)");
    auto p = std::make_unique<sw_pipeline>();
    p->cfg = c;
    p->dms.assign(c.coherent_dms, c.coherent_dms + c.ndms);
    if(c.fft_len) p->fft_len.assign(c.fft_len, c.fft_len + c.ndms);
    if(c.overlap) p->overlap.assign(c.overlap, c.overlap + c.ndms);
    p->cfg.coherent_dms = p->dms.data();
    p->cfg.fft_len      = c.fft_len ? p->fft_len.data() : nullptr;
    p->cfg.overlap      = c.overlap ? p->overlap.data() : nullptr;
    if(p->cfg.output_level <= 0) p->cfg.output_level = 24.0f;
    if(p->cfg.total_nchans <= 0) p->cfg.total_nchans = 4096;
    // tiles: the configuration's, else the tuning table's, else defaults
    p->tuning_file = c.tuning_file ? c.tuning_file : sw::default_tuning_path();
    p->cfg.tuning_file = nullptr;
    {
        int found = 0, waves = 0, share = 0;
        std::vector<int> tl(c.ndms, 0);
        const bool want_fft = !c.fft_len && !c.autotune_fft;
        int rc = sw::tuning_lookup(p->tuning_file, c.nantennas, c.nbeams, c.nsamples, c.tscrunch, c.stokes, c.ndms,
                                   p->dms.data(), &found, &waves, tl.data(), &share);
        if(rc) return rc;
        p->cb_waves = c.cb_waves > 0 ? c.cb_waves : (found && waves > 0 ? waves : 4);
        if(found && c.cb_waves <= 0 && waves > 0) p->tile_source = 1;
        SW_CHECK_ARG(p->cb_waves == 2 || p->cb_waves == 4 || p->cb_waves == 8, "cb_waves must be 0, 2, 4 or 8");
        // the table's FFT lengths were timed for its own band: they are taken
        // only if each still leaves at most a quarter of a segment to the
        // overlap here and is no longer than the gulp (the autotuner's rule)
        bool fit = true;
        for(int d = 0; d < c.ndms && found == 2; ++d) {
            const int O = (c.overlap && c.overlap[d] >= 0)
                              ? c.overlap[d]
                              : sw::codedisp_overlap(c.centre_frequency - c.bandwidth / 2.0, c.bandwidth / c.nchans,
                                                     c.nchans / c.bandwidth, p->dms[d]);
            fit = fit && tl[d] >= 4 * O && (tl[d] <= c.nsamples || tl[d] < 8 * O || tl[d] == 1024);
        }
        if(found == 2 && want_fft && fit && std::all_of(tl.begin(), tl.end(), [](int n) { return n > 0; })) {
            p->fft_len = tl;
            p->cfg.fft_len = p->fft_len.data();
            p->cfg.share_forward = share;
            p->tile_source = 2;
        }
    }
    SW_HIP_TRY(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
    for(auto& e : p->ev) SW_HIP_TRY(hipEventCreate(&e));
    for(auto& e : p->ev_w) SW_HIP_TRY(hipEventCreate(&e));
    p->ev_dm.resize(4 * c.ndms);
    for(auto& e : p->ev_dm) SW_HIP_TRY(hipEventCreate(&e));

    // coherent dedispersion plan: tsamp = F / bw, band bottom = centre - bw/2
    // (create_coherent_dedisperser_config, cpp/skyweaver/src/CoherentDedisperser.cu:19-54)
    sw_codedisp_config cd{};
    cd.nchans    = c.nchans;
    cd.nantennas = c.nantennas;
    cd.nsamples  = c.nsamples;
    cd.low_freq  = c.centre_frequency - c.bandwidth / 2.0;
    cd.chan_bw   = c.bandwidth / c.nchans;
    cd.tsamp     = c.nchans / c.bandwidth;
    cd.ndms      = c.ndms;
    cd.dms       = p->dms.data();
    cd.fft_len   = p->cfg.fft_len;
    cd.overlap   = p->cfg.overlap;
    cd.share_forward = p->cfg.share_forward;
    cd.chirp     = c.chirp;
    std::vector<int> tuned(c.ndms, 0);
    if(c.autotune_fft && !c.fft_len) {
        cd.input_fpat = 1;
        int share = 0;
        int rc = sw_codedisp_autotune(&cd, 8, tuned.data(), &share, nullptr, p->stream);
        if(rc) return rc;
        cd.fft_len = tuned.data();
        cd.share_forward = share;
        p->cfg.share_forward = share;
        p->tile_source = 3;
    }
    cd.input_fpat = 1;
    cd.history_in_place = 1;
    int rc = sw_codedisp_create(&cd, p->stream, &p->cd);
    if(rc) return rc;

    // the two FPAT gulp buffers; the host-input staging buffer (taftp_d) and
    // the per-DM FTPA copies of the dedispersed voltages (ftpa_dd, read only
    // by the beamformers' FTPA forms) are allocated on first use
    const size_t fb = p->ftpa_bytes();
    SW_HIP_TRY(hipMalloc(&p->ftpa, fb));
    SW_HIP_TRY(hipMalloc(&p->ftpa_prev, fb));
    p->ftpa_dd.assign(c.ndms, nullptr);
    SW_HIP_TRY(hipMalloc(&p->stats, sizeof(float) * 4 * c.nchans * 2 * c.nantennas));
    SW_HIP_TRY(hipMalloc(&p->stats_ws, sw::fpat_stats_workspace(c.nchans, 2, c.nantennas, c.nsamples)));
    p->h_stats.resize((size_t)4 * c.nchans * 2 * c.nantennas);
    const int nv = p->nv();
    // incoherent beam outputs sized for the largest supported beamset count
    SW_HIP_TRY(hipMalloc(&p->ib_raw, sizeof(float) * 16 * p->tout() * c.nchans * nv));
    if(c.async_sinks) {
        p->cb_bufs.assign(2 * c.ndms, nullptr);
        p->ib_bufs.assign(2 * c.ndms, nullptr);
        for(auto& b : p->cb_bufs) SW_HIP_TRY(hipMalloc(&b, p->tout() * c.nchans * c.nbeams * nv));
        for(auto& b : p->ib_bufs) SW_HIP_TRY(hipMalloc(&b, (size_t)16 * p->tout() * c.nchans * nv));
        p->stats_bufs[0] = p->stats;
        SW_HIP_TRY(hipMalloc(&p->stats_bufs[1], sizeof(float) * 4 * c.nchans * 2 * c.nantennas));
        p->ev_ready.assign(2 * c.ndms + 2, nullptr);
        for(auto& e : p->ev_ready) SW_HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        SW_HIP_TRY(hipEventCreateWithFlags(&p->ev_link, hipEventDisableTiming));
        SW_HIP_TRY(hipStreamCreateWithFlags(&p->sig_stream, hipStreamNonBlocking));
        p->worker = std::make_unique<SinkWorker>();
        p->worker->start();
    } else {
        SW_HIP_TRY(hipMalloc(&p->cb_out, p->tout() * c.nchans * c.nbeams * nv));
        SW_HIP_TRY(hipMalloc(&p->ib_q, (size_t)16 * p->tout() * c.nchans * nv));
    }
    SW_HIP_TRY(hipMalloc(&p->weights, (size_t)c.nchans * c.nbeams * c.nantennas * 2));
    SW_HIP_TRY(hipMalloc(&p->packed, (size_t)c.nchans * c.nbeams * c.nantennas * 4));
    // one beamset of all antennas until told otherwise
    p->nbeamsets = 1;
    p->h_bmap.assign(c.nbeams, 0);
    p->h_antw.assign(c.nantennas, 1.0f);
    rc = upload_beamsets(p.get());
    if(rc) return rc;
    sw_bf_dims d = bf_dims(p.get());
    p->use_mfma = c.cb_form == SW_CB_MFMA && sw::cb_mfma_supported(d);

    if(c.id_enable) {
        // one incoherent dedispersion block per coherent DM
        p->id_ts.resize(c.ndms);
        p->id_dms.resize(c.ndms);
        std::vector<const float*> dm_ptrs(c.ndms);
        std::vector<int> nd(c.ndms);
        for(int i = 0; i < c.ndms; ++i) {
            p->id_ts[i] = c.id_tscrunch ? c.id_tscrunch[i] : 1;
            if(c.id_dms && c.id_ndms && c.id_dms[i] && c.id_ndms[i] > 0)
                p->id_dms[i].assign(c.id_dms[i], c.id_dms[i] + c.id_ndms[i]);
            else
                p->id_dms[i].assign(1, p->dms[i]);
            dm_ptrs[i] = p->id_dms[i].data();
            nd[i] = (int)p->id_dms[i].size();
        }
        sw_idpipe_config ic{};
        ic.nchans = c.nchans;
        ic.centre_frequency = c.centre_frequency;
        ic.bandwidth = c.bandwidth;
        ic.cb_tscrunch = c.tscrunch;
        ic.nbeams = c.nbeams;
        ic.nvec = nv;
        ic.nblocks = c.ndms;
        ic.coherent_dms = p->dms.data();
        ic.tscrunch = p->id_ts.data();
        ic.dms = dm_ptrs.data();
        ic.ndms = nd.data();
        ic.dispatch_size = c.id_dispatch_size;
        rc = sw_idpipe_create(&ic, &p->idp);
        if(rc) return rc;
    }
    SW_HIP_TRY(hipStreamSynchronize(p->stream));
    *out = p.release();
    return SW_OK;
}

extern "C" int sw_pipeline_destroy(sw_pipeline* p)
{
    delete p;
    return SW_OK;
}

extern "C" void* sw_pipeline_stream(const sw_pipeline* p) { return p ? p->stream : nullptr; }

extern "C" int sw_pipeline_tile(const sw_pipeline* p, int* cb_waves, int* fft_len, int* share_forward, int* source)
{
    SW_CHECK_ARG(p, "null pipeline");
    if(cb_waves) *cb_waves = p->cb_waves;
    if(fft_len)
        for(int d = 0; d < p->cfg.ndms; ++d) sw_codedisp_info(p->cd, d, &fft_len[d], nullptr, nullptr);
    if(share_forward) *share_forward = p->cfg.share_forward;
    if(source) *source = p->tile_source;
    return SW_OK;
}

extern "C" int sw_pipeline_set_sinks(sw_pipeline* p, sw_sink_fn cb, void* cb_user, sw_sink_fn ib,
                                     void* ib_user, sw_sink_fn stats, void* stats_user)
{
    SW_CHECK_ARG(p, "null pipeline");
    p->cb_sink = cb;
    p->cb_user = cb_user;
    p->ib_sink = ib;
    p->ib_user = ib_user;
    p->st_sink = stats;
    p->st_user = stats_user;
    return SW_OK;
}

extern "C" int sw_pipeline_set_tdb_sink(sw_pipeline* p, sw_sink_fn tdb, void* tdb_user)
{
    SW_CHECK_ARG(p, "null pipeline");
    SW_CHECK_ARG(p->idp, "incoherent dedispersion is not enabled in this pipeline");
    return sw_idpipe_set_sink(p->idp, tdb, tdb_user);
}

extern "C" int sw_pipeline_set_delay_file(sw_pipeline* p, const char* path)
{
    SW_CHECK_ARG(p && path, "null argument");
    const auto& c = p->cfg;
    sw_delay_model* m = nullptr;
    int rc = sw_delay_model_open(path, c.nbeams, c.nantennas, &m);
    if(rc) return rc;
    int ns = 0;
    sw_delay_model_beamsets(m, &ns, nullptr, nullptr, 0);
    if(ns > 16) {
        sw_delay_model_close(m);
        sw::set_error("delay model has more than 16 beamsets");
        return SW_ERR_UNSUPPORTED;
    }
    std::vector<int> map(c.nbeams);
    std::vector<float> bw((size_t)ns * c.nantennas);
    sw_delay_model_beamsets(m, &ns, map.data(), bw.data(), ns);
    if(p->delays) sw_delay_model_close(p->delays);
    p->delays = m;
    p->loaded_epoch = NAN;
    p->h_delays.assign((size_t)c.nbeams * c.nantennas * 3, 0.0f);
    if(!p->d_delays) SW_HIP_TRY(hipMalloc(&p->d_delays, sizeof(float) * p->h_delays.size()));
    if(!p->d_freqs) {
        std::vector<double> f(c.nchans);
        sw_channel_frequencies(c.centre_frequency, c.bandwidth, c.nchans, f.data());
        SW_HIP_TRY(hipMalloc(&p->d_freqs, sizeof(double) * c.nchans));
        SW_HIP_TRY(hipMemcpy(p->d_freqs, f.data(), sizeof(double) * c.nchans, hipMemcpyHostToDevice));
    }
    p->weights_set = true;  // generated per gulp
    return sw_pipeline_set_beamsets(p, ns, map.data(), bw.data());
}

extern "C" int sw_pipeline_set_weights(sw_pipeline* p, const void* h_weights)
{
    SW_CHECK_ARG(p && h_weights, "null argument");
    const auto& c = p->cfg;
    SW_HIP_TRY(hipMemcpyAsync(p->weights, h_weights, (size_t)c.nchans * c.nbeams * c.nantennas * 2,
                              hipMemcpyHostToDevice, p->stream));
    int rc = sw_cb_pack_weights(p->weights, p->packed, c.nchans, c.nbeams, c.nantennas, p->stream);
    if(rc) return rc;
    SW_HIP_TRY(hipStreamSynchronize(p->stream));
    p->weights_set = true;
    if(p->delays) {  // static weights replace the delay-model path
        sw_delay_model_close(p->delays);
        p->delays = nullptr;
    }
    return SW_OK;
}

extern "C" int sw_pipeline_set_beamsets(sw_pipeline* p, int nbeamsets, const int* h_map,
                                        const float* h_antenna_weights)
{
    SW_CHECK_ARG(p && h_map && h_antenna_weights, "null argument");
    SW_CHECK_ARG(nbeamsets > 0 && nbeamsets <= 16, "between 1 and 16 beamsets are supported");
    for(int b = 0; b < p->cfg.nbeams; ++b)
        SW_CHECK_ARG(h_map[b] >= 0 && h_map[b] < nbeamsets, "beamset map entry out of range");
    p->nbeamsets = nbeamsets;
    p->h_bmap.assign(h_map, h_map + p->cfg.nbeams);
    p->h_antw.assign(h_antenna_weights, h_antenna_weights + (size_t)nbeamsets * p->cfg.nantennas);
    return upload_beamsets(p);
}

extern "C" int sw_pipeline_init(sw_pipeline* p, const sw_obs_header* header)
{
    SW_CHECK_ARG(p && header, "null argument");
    SW_CHECK_ARG(header->nantennas > 0 && header->nantennas <= p->cfg.nantennas,
                 "Input data contains too many antennas for the configured nantennas");
    p->header = *header;
    p->inited = true;
    // timestamp step per block in sampler ticks (BeamformerPipeline.cu:67)
    p->tick_per_block = 2.0 * p->cfg.total_nchans * p->cfg.nsamples;
    return SW_OK;
}

extern "C" int sw_pipeline_stage_times(const sw_pipeline* p, float* ms, int n)
{
    SW_CHECK_ARG(p && ms && n >= 0, "null argument");
    for(int i = 0; i < n && i < SW_NSTAGES; ++i) ms[i] = p->stage_ms[i];
    return SW_OK;
}

namespace {

int check_gulp(const sw_pipeline* p, size_t bytes)
{
    SW_CHECK_ARG(p->inited, "pipeline not initialised (call sw_pipeline_init)");
    SW_CHECK_ARG(p->weights_set, "no beamforming weights set");
    const auto& c       = p->cfg;
    const size_t expect = (size_t)c.nsamples * p->header.nantennas * c.nchans * 2 * 2;
    if(bytes != expect) {
        sw::set_error("Unexpected buffer size, expected " + std::to_string(expect) + " but got " +
                      std::to_string(bytes));
        return SW_ERR_INVALID;
    }
    return SW_OK;
}

// One gulp from device memory `in` on the pipeline stream; `consumed`, if
// given, is recorded once the input has been read (after the transpose).
int run_gulp(sw_pipeline* p, const void* in, hipEvent_t consumed);

// Stokes I with 0/1 antenna weights (what delay models give): the MFMA
// beamformer (and the incoherent beam, inside it where supported, else its
// own kernel) reads the dedisperser's time-major output ring directly, with
// no corner turn to FTPA
bool beams_on_ring(const sw_pipeline* p)
{
    const auto& c = p->cfg;
    bool r = p->use_mfma && c.stokes == SW_STOKES_I && 2 * c.nantennas <= 256 && c.nsamples % 8 == 0;
    for(float w : p->h_antw) r = r && (w == 0.0f || w == 1.0f);
    return r;
}

// DM dm's incoherent and coherent beam kernels into p->ib_q / p->cb_out (the
// dedispersion of the gulp is already enqueued); `mid`, if given, is
// recorded between the two
int enqueue_beams(sw_pipeline* p, int dm, bool ring_direct, bool fuse_ib, hipStream_t st, hipEvent_t mid)
{
    const auto& c = p->cfg;
    const sw_bf_dims d = bf_dims(p);
    const int S = p->nbeamsets, F = c.nchans;
    float* cb_offsets = p->scl;
    float* cb_scaling = p->scl + S * F;
    float* ib_offsets = p->scl + 2 * S * F;
    float* ib_scaling = p->scl + 3 * S * F;
    int rc;
    const void* ring = nullptr;
    int ld = 0, w0 = 0;
    if(ring_direct) {
        rc = sw_codedisp_window(p->cd, dm, &ring, &ld, &w0);
        if(rc) return rc;
        // with the incoherent beam fused into the coherent beamformer (which
        // then reads the ring once) there is nothing to do here
        if(!fuse_ib) rc = sw::ib_beamform_tm(ring, ld, w0, p->ib_raw, p->ib_q, ib_scaling, ib_offsets, p->antw, d, st);
    } else {
        rc = sw_ib_beamform(p->ftpa_dd[dm], p->ib_raw, p->ib_q, ib_scaling, ib_offsets, p->antw, &d, st);
    }
    if(rc) return rc;
    if(mid) SW_HIP_TRY(hipEventRecord(mid, st));
    if(ring_direct)
        return sw::cb_beamform_packed_tm(ring, ld, w0, p->packed, cb_scaling, cb_offsets, p->bmap, p->ib_raw,
                                         p->cb_out, d, st, fuse_ib ? p->ib_raw : nullptr, p->ib_q, ib_scaling,
                                         ib_offsets, p->antw, p->cb_waves, p->cb_fastdiv);
    if(p->use_mfma)
        return sw::cb_beamform_packed_waves(p->ftpa_dd[dm], p->packed, cb_scaling, cb_offsets, p->bmap, p->ib_raw,
                                            p->cb_out, d, st, p->cb_waves);
    return sw_cb_beamform(p->ftpa_dd[dm], p->weights, cb_scaling, cb_offsets, p->bmap, p->ib_raw, p->cb_out, &d,
                          SW_CB_PLAIN, nullptr, st);
}

// The blocks of DM dm of this gulp (coherent beams, incoherent beam) with
// their metadata: the dedispersed blocks carry the filter delay tsamp O/2
// and, when the FFT hop does not divide the gulp, the dedisperser's latency
void dm_blocks(const sw_pipeline* p, int dm, double utc_offset_gulp, sw_block& cb, sw_block& ib, double& utc)
{
    const auto& c = p->cfg;
    int fl = 0, ov = 0, lat = 0;
    sw_codedisp_info(p->cd, dm, &fl, &ov, &lat);
    const double tsamp = c.nchans / c.bandwidth;
    utc = utc_offset_gulp + tsamp * (ov / 2.0 - lat);
    const size_t T2 = p->tout(), F = c.nchans;
    const int nv = p->nv();
    fill_block(cb, SW_BLOCK_CB_TFB, p->cb_out, T2 * F * c.nbeams * nv, "TFB", {T2, F, (size_t)c.nbeams});
    fill_block(ib, SW_BLOCK_IB_BTF, p->ib_q, (size_t)p->nbeamsets * T2 * F * nv, "BTF",
               {(size_t)p->nbeamsets, T2, F});
    for(sw_block* b : {&cb, &ib}) {
        b->dm_idx = dm;
        b->reference_dm = p->dms[dm];
        b->tsamp = tsamp * c.tscrunch;
        b->utc_offset = utc;
    }
}

void stats_block(const sw_pipeline* p, double utc_offset_gulp, sw_block& blk)
{
    const auto& c = p->cfg;
    fill_block(blk, SW_BLOCK_STATS_FPA, p->stats, sizeof(float) * 4 * c.nchans * 2 * c.nantennas, "FPA",
               {(size_t)c.nchans, 2, (size_t)c.nantennas});
    blk.tsamp = c.nchans / c.bandwidth * c.nsamples;
    blk.utc_offset = utc_offset_gulp;
}

// Record `ready` once the work enqueued so far on the pipeline stream is
// done, on the signal stream (the event the sink thread waits for).
int signal_ready(sw_pipeline* p, hipEvent_t ready)
{
    SW_HIP_TRY(hipEventRecord(p->ev_link, p->stream));
    SW_HIP_TRY(hipStreamWaitEvent(p->sig_stream, p->ev_link, 0));
    SW_HIP_TRY(hipEventRecord(ready, p->sig_stream));
    return SW_OK;
}

// The gulp's device work captured into a hip graph and launched: the
// transpose with the statistics' sums, the dedispersion of every DM and
// every DM's beams into this gulp's parity buffers.  The executable graph
// is updated in place from each gulp's capture (hipGraphExecUpdate: the
// kernels are the same, their arguments move with the segment grid and
// the buffer parity) and instantiated again only when the launches differ.
// The blocks are released to the sink thread, and the incoherent
// dedispersion fed, once the graph has run.
int run_gulp_graph(sw_pipeline* p, const void* in, hipEvent_t consumed, int par, double utc_offset_gulp);


} // namespace

extern "C" int sw_pipeline_process(sw_pipeline* p, const void* taftp, size_t bytes, int on_device)
{
    SW_CHECK_ARG(p && taftp, "null argument");
    const int rc = check_gulp(p, bytes);
    if(rc) return rc;
    const void* in = taftp;
    if(!on_device) {
        if(!p->taftp_d) SW_HIP_TRY(hipMalloc(&p->taftp_d, p->ftpa_bytes()));
        SW_HIP_TRY(hipMemcpyAsync(p->taftp_d, taftp, bytes, hipMemcpyHostToDevice, p->stream));
        in = p->taftp_d;
    }
    return run_gulp(p, in, nullptr);
}

extern "C" int sw_pipeline_prefetch(sw_pipeline* p, const void* h_taftp, size_t bytes)
{
    SW_CHECK_ARG(p && h_taftp, "null argument");
    const int rc = check_gulp(p, bytes);
    if(rc) return rc;
    SW_CHECK_ARG(p->pre_count < 2, "two gulps are already prefetched (process one first)");
    if(!p->copy_stream) {
        SW_HIP_TRY(hipStreamCreateWithFlags(&p->copy_stream, hipStreamNonBlocking));
        for(int i = 0; i < 2; ++i) {
            SW_HIP_TRY(hipEventCreateWithFlags(&p->pre_copied[i], hipEventDisableTiming));
            SW_HIP_TRY(hipEventCreateWithFlags(&p->pre_consumed[i], hipEventDisableTiming));
        }
    }
    const int slot = (p->pre_head + p->pre_count) % 2;
    if(!p->pre_d[slot]) {
        const auto& c = p->cfg;
        SW_HIP_TRY(hipMalloc(&p->pre_d[slot], (size_t)c.nsamples * c.nantennas * c.nchans * 2 * 2));
    }
    // the slot's previous gulp must have been read by its transpose
    if(p->pre_used[slot]) SW_HIP_TRY(hipStreamWaitEvent(p->copy_stream, p->pre_consumed[slot], 0));
    SW_HIP_TRY(hipMemcpyAsync(p->pre_d[slot], h_taftp, bytes, hipMemcpyHostToDevice, p->copy_stream));
    SW_HIP_TRY(hipEventRecord(p->pre_copied[slot], p->copy_stream));
    ++p->pre_count;
    return SW_OK;
}

extern "C" int sw_pipeline_process_prefetched(sw_pipeline* p)
{
    SW_CHECK_ARG(p, "null argument");
    SW_CHECK_ARG(p->pre_count > 0, "no prefetched gulp (call sw_pipeline_prefetch)");
    const int slot = p->pre_head;
    p->pre_head    = (p->pre_head + 1) % 2;
    --p->pre_count;
    SW_HIP_TRY(hipStreamWaitEvent(p->stream, p->pre_copied[slot], 0));
    p->pre_used[slot] = true;
    return run_gulp(p, p->pre_d[slot], p->pre_consumed[slot]);
}

extern "C" int sw_pipeline_synchronize(sw_pipeline* p)
{
    SW_CHECK_ARG(p, "null argument");
    if(p->copy_stream) SW_HIP_TRY(hipStreamSynchronize(p->copy_stream));
    SW_HIP_TRY(hipStreamSynchronize(p->stream));
    if(p->worker) {
        p->worker->drain();
        return p->worker->take_error();
    }
    return SW_OK;
}

namespace {

int run_gulp(sw_pipeline* p, const void* in, hipEvent_t consumed)
{
    const auto& c      = p->cfg;
    const int nant_in  = p->header.nantennas;
    hipStream_t st     = p->stream;
    const sw_bf_dims d = bf_dims(p);
    const int F = c.nchans, A = c.nantennas, S = p->nbeamsets;
    const double utc_offset_gulp =
        p->header.utc_offset + (double)p->call_count * p->tick_per_block / p->header.sample_clock;
    int rc;
    SinkWorker* wk = p->worker.get();
    const int par = (int)(p->call_count & 1);  // the buffer set of this gulp (asynchronous sinks)
    if(wk) {
        rc = wk->take_error();  // a sink of an earlier gulp failed
        if(rc) return rc;
        p->stats = p->stats_bufs[par];
        wk->wait_free(p->stats);
    }
    SW_HIP_TRY(hipEventRecord(p->ev_w[0], st));
    if(p->delays) {
        // DelayManager::delays + WeightsManager::weights at the gulp's unix time
        // (BeamformerPipeline.cu:292-295: utc_start + utc_offset + n ticks / clock)
        const long double unix_ts = (long double)p->header.utc_start + (long double)p->header.utc_offset +
                                    (long double)(p->call_count * (unsigned long long)p->tick_per_block) /
                                        (long double)p->header.sample_clock;
        double epoch = 0;
        rc = sw_delay_model_get(p->delays, (double)unix_ts, p->h_delays.data(), &epoch);
        if(rc) return rc;
        if(!(epoch == p->loaded_epoch)) {
            SW_HIP_TRY(hipMemcpyAsync(p->d_delays, p->h_delays.data(), sizeof(float) * p->h_delays.size(),
                                      hipMemcpyHostToDevice, st));
            SW_HIP_TRY(hipStreamSynchronize(st));  // the host copy is refilled on the next model
            p->loaded_epoch = epoch;
        }
        rc = sw_generate_weights(p->d_delays, p->weights, p->d_freqs, A, c.nbeams, F, (double)unix_ts, epoch, st);
        if(rc) return rc;
        if(p->use_mfma) {
            rc = sw_cb_pack_weights(p->weights, p->packed, F, c.nbeams, A, st);
            if(rc) return rc;
        }
    }
    SW_HIP_TRY(hipEventRecord(p->ev_w[1], st));
    if(c.graph && !p->scalings_dirty) return run_gulp_graph(p, in, consumed, par, utc_offset_gulp);
    SW_HIP_TRY(hipEventRecord(p->ev[0], st));
    // the gulp is held time-major (FPAT) up to the dedisperser; the two FPAT
    // buffers alternate so that the previous gulp stays readable as history
    std::swap(p->ftpa, p->ftpa_prev);
    // one pass over the gulp: transpose fused with the statistics' power
    // sums (the "transpose" stage time includes them; the "statistics"
    // stage is left with nothing to time)
    rc = sw_transpose_fpat_stats(in, p->ftpa, p->stats, c.nsamples / 256, nant_in, A, F, p->stats_ws, st);
    if(rc) return rc;
    if(consumed) SW_HIP_TRY(hipEventRecord(consumed, st));
    if(wk) {
        rc = signal_ready(p, p->ev_ready[2 * c.ndms + par]);
        if(rc) return rc;
    }
    SW_HIP_TRY(hipEventRecord(p->ev[1], st));
    SW_HIP_TRY(hipEventRecord(p->ev[2], st));
    if(p->scalings_dirty) {
        // update_scalings on the first gulp (BeamformerPipeline.cu:171-176)
        // and on the first gulp after the beamsets change
        SW_HIP_TRY(hipMemcpyAsync(p->h_stats.data(), p->stats, sizeof(float) * p->h_stats.size(),
                                  hipMemcpyDeviceToHost, st));
        SW_HIP_TRY(hipStreamSynchronize(st));
        std::vector<float> h(4 * (size_t)S * F);
        rc = sw_update_scalings(p->h_stats.data(), p->h_antw.data(), S, F, 2, A, c.tscrunch, c.tscrunch,
                                c.output_level, h.data(), h.data() + S * F, h.data() + 2 * S * F,
                                h.data() + 3 * S * F);
        if(rc) return rc;
        // the beamformer divides with a staged reciprocal when every coherent
        // scale lies in [2^-40, 2^60] (csrc/cb.hip, div_by)
        p->cb_fastdiv = true;
        for(size_t i = (size_t)S * F; i < 2 * (size_t)S * F; ++i)
            p->cb_fastdiv = p->cb_fastdiv && h[i] >= 0x1p-40f && h[i] <= 0x1p60f;
        SW_HIP_TRY(hipMemcpyAsync(p->scl, h.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice, st));
        SW_HIP_TRY(hipStreamSynchronize(st));  // h is freed at the end of this scope
        p->scalings_dirty = false;
    }
    const bool ring_direct = beams_on_ring(p);
    const bool fuse_ib = ring_direct && sw::cb_fused_ib_supported(d);
    if(!ring_direct)
        for(auto& b : p->ftpa_dd)
            if(!b) SW_HIP_TRY(hipMalloc(&b, p->ftpa_bytes()));
    std::vector<void*> outs(p->ftpa_dd.begin(), p->ftpa_dd.end());
    if(ring_direct) std::fill(outs.begin(), outs.end(), nullptr);
    rc = sw_codedisp_run(p->cd, p->ftpa, outs.data(), nullptr, 0, c.ndms, c.batched_dms, st);
    if(rc) return rc;
    rc = sw_codedisp_push_history(p->cd, p->ftpa, st);
    if(rc) return rc;
    SW_HIP_TRY(hipEventRecord(p->ev[3], st));

    const size_t T2 = p->tout();
    for(int dm = 0; dm < c.ndms; ++dm) {
        if(wk) {  // this DM's output buffers of this gulp parity, once their last blocks are delivered
            p->cb_out = p->cb_bufs[2 * dm + par];
            p->ib_q = p->ib_bufs[2 * dm + par];
            wk->wait_free(p->cb_out);
            wk->wait_free(p->ib_q);
        }
        SW_HIP_TRY(hipEventRecord(p->ev_dm[4 * dm], st));
        rc = enqueue_beams(p, dm, ring_direct, fuse_ib, st, p->ev_dm[4 * dm + 1]);
        if(rc) return rc;
        SW_HIP_TRY(hipEventRecord(p->ev_dm[4 * dm + 2], st));
        sw_block blk, iblk;
        double utc = 0;
        dm_blocks(p, dm, utc_offset_gulp, blk, iblk, utc);
        if(wk) {
            hipEvent_t ready = p->ev_ready[2 * dm + par];
            rc = signal_ready(p, ready);
            if(rc) return rc;
            if(p->cb_sink) wk->push({p->cb_sink, p->cb_user, blk, ready, p->cb_out});
            if(p->ib_sink) wk->push({p->ib_sink, p->ib_user, iblk, ready, p->ib_q});
        } else {
            rc = call_sink(p, p->cb_sink, p->cb_user, blk);
            if(rc) return rc;
            rc = call_sink(p, p->ib_sink, p->ib_user, iblk);
            if(rc) return rc;
        }
        if(p->idp) {
            rc = sw_idpipe_push(p->idp, dm, p->cb_out, T2, 1, utc, blk.tsamp, st);
            if(rc) return rc;
        }
        SW_HIP_TRY(hipEventRecord(p->ev_dm[4 * dm + 3], st));
    }
    SW_HIP_TRY(hipEventRecord(p->ev[4], st));
    {
        sw_block blk;
        stats_block(p, utc_offset_gulp, blk);
        if(wk) {
            if(p->st_sink) wk->push({p->st_sink, p->st_user, blk, p->ev_ready[2 * c.ndms + par], p->stats});
        } else {
            rc = call_sink(p, p->st_sink, p->st_user, blk);
            if(rc) return rc;
        }
    }
    SW_HIP_TRY(hipStreamSynchronize(st));
    // per-stage device times of this gulp
    float t = 0;
    for(int i = 0; i < SW_NSTAGES; ++i) p->stage_ms[i] = 0;
    SW_HIP_TRY(hipEventElapsedTime(&t, p->ev[0], p->ev[1]));
    p->stage_ms[0] = t;
    SW_HIP_TRY(hipEventElapsedTime(&t, p->ev[1], p->ev[2]));
    p->stage_ms[1] = t;
    SW_HIP_TRY(hipEventElapsedTime(&t, p->ev[2], p->ev[3]));
    p->stage_ms[2] = t;
    for(int dm = 0; dm < c.ndms; ++dm) {
        SW_HIP_TRY(hipEventElapsedTime(&t, p->ev_dm[4 * dm], p->ev_dm[4 * dm + 1]));
        p->stage_ms[3] += t;
        SW_HIP_TRY(hipEventElapsedTime(&t, p->ev_dm[4 * dm + 1], p->ev_dm[4 * dm + 2]));
        p->stage_ms[4] += t;
        if(p->idp) {
            SW_HIP_TRY(hipEventElapsedTime(&t, p->ev_dm[4 * dm + 2], p->ev_dm[4 * dm + 3]));
            p->stage_ms[5] += t;
        }
    }
    SW_HIP_TRY(hipEventElapsedTime(&t, p->ev_w[0], p->ev_w[1]));
    p->stage_ms[6] = t;
    SW_HIP_TRY(hipEventElapsedTime(&t, p->ev[0], p->ev[4]));
    p->stage_ms[7] = t - (p->stage_ms[0] + p->stage_ms[1] + p->stage_ms[2] + p->stage_ms[3] + p->stage_ms[4] +
                          p->stage_ms[5]);
    ++p->call_count;
    return SW_OK;
}

int run_gulp_graph(sw_pipeline* p, const void* in, hipEvent_t consumed, int par, double utc_offset_gulp)
{
    const auto& c      = p->cfg;
    hipStream_t st     = p->stream;
    SinkWorker* wk     = p->worker.get();
    const sw_bf_dims d = bf_dims(p);
    const bool ring_direct = beams_on_ring(p);
    const bool fuse_ib = ring_direct && sw::cb_fused_ib_supported(d);
    // allocations and buffer waits before the capture
    if(!ring_direct)
        for(auto& b : p->ftpa_dd)
            if(!b) SW_HIP_TRY(hipMalloc(&b, p->ftpa_bytes()));
    for(int dm = 0; dm < c.ndms; ++dm) {
        wk->wait_free(p->cb_bufs[2 * dm + par]);
        wk->wait_free(p->ib_bufs[2 * dm + par]);
    }
    std::vector<void*> outs(p->ftpa_dd.begin(), p->ftpa_dd.end());
    if(ring_direct) std::fill(outs.begin(), outs.end(), nullptr);
    SW_HIP_TRY(hipEventRecord(p->ev[0], st));
    SW_HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    int rc = SW_OK;
    {
        std::swap(p->ftpa, p->ftpa_prev);
        rc = sw_transpose_fpat_stats(in, p->ftpa, p->stats, c.nsamples / 256, p->header.nantennas, c.nantennas,
                                     c.nchans, p->stats_ws, st);
        if(!rc) rc = sw_codedisp_run(p->cd, p->ftpa, outs.data(), nullptr, 0, c.ndms, c.batched_dms, st);
        if(!rc) rc = sw_codedisp_push_history(p->cd, p->ftpa, st);
        for(int dm = 0; dm < c.ndms && !rc; ++dm) {
            p->cb_out = p->cb_bufs[2 * dm + par];
            p->ib_q = p->ib_bufs[2 * dm + par];
            rc = enqueue_beams(p, dm, ring_direct, fuse_ib, st, nullptr);
        }
    }
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(st, &g);
    if(rc) {
        if(g) (void)hipGraphDestroy(g);
        return rc;
    }
    if(ec != hipSuccess) {
        sw::set_error(std::string("sw_pipeline_process: hipStreamEndCapture failed: ") + hipGetErrorString(ec));
        return SW_ERR_HIP;
    }
    if(p->gexec) {
        hipGraphNode_t err_node = nullptr;
        hipGraphExecUpdateResult res{};
        if(hipGraphExecUpdate(p->gexec, g, &err_node, &res) != hipSuccess || res != hipGraphExecUpdateSuccess) {
            (void)hipGetLastError();
            (void)hipGraphExecDestroy(p->gexec);
            p->gexec = nullptr;
        }
    }
    if(!p->gexec) {
        const hipError_t e = hipGraphInstantiate(&p->gexec, g, nullptr, nullptr, 0);
        if(e != hipSuccess) {
            (void)hipGraphDestroy(g);
            p->gexec = nullptr;
            sw::set_error(std::string("sw_pipeline_process: hipGraphInstantiate failed: ") + hipGetErrorString(e));
            return SW_ERR_HIP;
        }
        ++p->graph_instantiations;
    }
    (void)hipGraphDestroy(g);
    SW_HIP_TRY(hipGraphLaunch(p->gexec, st));
    if(consumed) SW_HIP_TRY(hipEventRecord(consumed, st));
    hipEvent_t ready = p->ev_ready[2 * c.ndms + par];
    rc = signal_ready(p, ready);
    if(rc) return rc;
    for(int dm = 0; dm < c.ndms; ++dm) {
        p->cb_out = p->cb_bufs[2 * dm + par];
        p->ib_q = p->ib_bufs[2 * dm + par];
        sw_block blk, iblk;
        double utc = 0;
        dm_blocks(p, dm, utc_offset_gulp, blk, iblk, utc);
        if(p->cb_sink) wk->push({p->cb_sink, p->cb_user, blk, ready, p->cb_out});
        if(p->ib_sink) wk->push({p->ib_sink, p->ib_user, iblk, ready, p->ib_q});
        if(p->idp) {
            rc = sw_idpipe_push(p->idp, dm, p->cb_out, p->tout(), 1, utc, blk.tsamp, st);
            if(rc) return rc;
        }
    }
    {
        sw_block blk;
        stats_block(p, utc_offset_gulp, blk);
        if(p->st_sink) wk->push({p->st_sink, p->st_user, blk, ready, p->stats});
    }
    SW_HIP_TRY(hipEventRecord(p->ev[4], st));
    SW_HIP_TRY(hipStreamSynchronize(st));
    // the graph is timed as a whole: its time is reported as "other"
    float t = 0;
    for(int i = 0; i < SW_NSTAGES; ++i) p->stage_ms[i] = 0;
    SW_HIP_TRY(hipEventElapsedTime(&t, p->ev_w[0], p->ev_w[1]));
    p->stage_ms[6] = t;
    SW_HIP_TRY(hipEventElapsedTime(&t, p->ev[0], p->ev[4]));
    p->stage_ms[7] = t;
    ++p->call_count;
    return SW_OK;
}

} // namespace

namespace {
// deterministic pseudo-random complex int8 of moderate amplitude (|x| < 64)
__global__ void tune_voltages_kernel(uint32_t* x, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n) return;
    uint32_t h = (uint32_t)i * 0x9E3779B1u + 0x7F4A7C15u;
    h ^= h >> 15;
    h *= 0x85EBCA77u;
    h ^= h >> 13;
    x[i] = h & 0x3F3F3F3Fu;
}

std::string json_list(const std::vector<float>& v, int prec)
{
    std::string s = "[";
    char b[64];
    for(size_t i = 0; i < v.size(); ++i) {
        std::snprintf(b, sizeof(b), "%s%.*f", i ? ", " : "", prec, (double)v[i]);
        s += b;
    }
    return s + "]";
}
} // namespace

extern "C" int sw_pipeline_autotune(const sw_pipeline_config* cfg, int tune_chans, char* json, size_t json_len)
{
    SW_CHECK_ARG(cfg && json && json_len > 0 && tune_chans > 0, "bad argument");
    SW_CHECK_ARG(cfg->nchans > 0 && cfg->nsamples > 0 && cfg->nantennas > 0 && cfg->ndms > 0, "empty workload");
    const sw_pipeline_config& c = *cfg;
    const int Fs = std::min(tune_chans, c.nchans);
    sw_pipeline_config sub = c;
    sub.nchans = Fs;
    sub.bandwidth = c.bandwidth / c.nchans * Fs;  // same channel width
    sub.async_sinks = 0;
    sub.graph = 0;
    sub.id_enable = 0;
    sub.autotune_fft = 0;
    sub.tuning_file = "";  // no table: every tile as given here
    // the FFT plan: the configuration's, else the FFT autotuner's
    std::vector<int> fl(c.ndms, 0);
    int share = c.share_forward;
    if(c.fft_len) {
        fl.assign(c.fft_len, c.fft_len + c.ndms);
    } else {
        sw_codedisp_config cd{};
        cd.nchans = Fs;
        cd.nantennas = c.nantennas;
        cd.nsamples = c.nsamples;
        cd.low_freq = c.centre_frequency - c.bandwidth / 2.0;
        cd.chan_bw = c.bandwidth / c.nchans;
        cd.tsamp = c.nchans / c.bandwidth;
        cd.ndms = c.ndms;
        cd.dms = c.coherent_dms;
        cd.overlap = c.overlap;
        cd.input_fpat = 1;
        cd.chirp = c.chirp;
        int rc = sw_codedisp_autotune(&cd, Fs, fl.data(), &share, nullptr, nullptr);
        if(rc) return rc;
    }
    sub.fft_len = fl.data();
    sub.share_forward = share;
    // synthetic gulp and weights
    const size_t gulp = (size_t)c.nsamples * c.nantennas * Fs * 4;
    void* d_in = nullptr;
    SW_HIP_TRY(hipMalloc(&d_in, gulp));
    hipLaunchKernelGGL(tune_voltages_kernel, dim3((unsigned)((gulp / 4 + 255) / 256)), dim3(256), 0, nullptr,
                       static_cast<uint32_t*>(d_in), gulp / 4);
    SW_LAUNCH_CHECK();
    std::vector<int8_t> w((size_t)Fs * c.nbeams * c.nantennas * 2);
    sw_synth_phase_weights(w.data(), w.size() / 2, 0x5EED1000);
    static const int kWaves[] = {2, 4, 8};
    float ms[3] = {-1, -1, -1};
    int best = -1;
    int rc = SW_OK;
    for(int k = 0; k < 3 && rc == SW_OK; ++k) {
        sub.cb_waves = kWaves[k];
        sw_pipeline* p = nullptr;
        rc = sw_pipeline_create(&sub, &p);
        if(rc) break;
        sw_obs_header h{};
        h.nantennas = c.nantennas;
        h.utc_start = 1708082229.0;
        h.sample_clock = 1712e6;
        rc = sw_pipeline_set_weights(p, w.data());
        if(!rc) rc = sw_pipeline_init(p, &h);
        float acc = 0.0f;
        for(int it = 0; it < 5 && rc == SW_OK; ++it) {  // two warm-up gulps, three timed
            rc = sw_pipeline_process(p, d_in, gulp, 1);
            if(rc == SW_OK && it >= 2) acc += p->stage_ms[3] + p->stage_ms[4];
        }
        sw_pipeline_destroy(p);
        if(rc == SW_ERR_UNSUPPORTED) {  // this shape does not fit the workload (LDS): not a candidate
            rc = SW_OK;
            continue;
        }
        if(rc) break;
        ms[k] = acc / 3.0f * (float)c.nchans / (float)Fs;
        if(best < 0 || ms[k] < ms[best]) best = k;
    }
    (void)hipFree(d_in);
    if(rc) return rc;
    SW_CHECK_ARG(best >= 0, "autotune: no beamformer workgroup shape runs this workload");
    static const char* stokes[] = {"I", "Q", "U", "V", "IQUV"};
    std::vector<float> dms(c.coherent_dms, c.coherent_dms + c.ndms), fls(fl.begin(), fl.end());
    char buf[2048];
    const int n = std::snprintf(
        buf, sizeof(buf),
        "{\"antennas\": %d, \"beams\": %d, \"samples\": %d, \"tscrunch\": %d, \"stokes\": \"%s\", "
        "\"cb_waves\": %d, \"dms\": %s, \"fft_len\": %s, \"share_forward\": %d, \"channels\": %d, "
        "\"tuned_on_channels\": %d, \"beams_ms_per_gulp\": {\"2\": %.3f, \"4\": %.3f, \"8\": %.3f}}",
        c.nantennas, c.nbeams, c.nsamples, c.tscrunch, stokes[c.stokes < 0 || c.stokes > 4 ? 0 : c.stokes],
        kWaves[best], json_list(dms, 4).c_str(), json_list(fls, 0).c_str(), share, c.nchans, Fs, ms[0], ms[1],
        ms[2]);
    SW_CHECK_ARG(n > 0 && (size_t)n < json_len, "json buffer too small");
    std::memcpy(json, buf, (size_t)n + 1);
    return SW_OK;
}
