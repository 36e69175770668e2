// skyweaver_amd.hpp -- C++ facade over the C ABI with the reference's
// pipeline API: PipelineConfig, DedispersionPlan, ObservationHeader,
// BeamformerPipeline<CBHandler, IBHandler, StatsHandler, Traits>,
// IncoherentDedispersionPipeline<Handler>, MultiFileWriter and the
// InvalidDelayEpoch exception.
//
// Reference interfaces mirrored (paths relative to the skyweaver repository):
//   PipelineConfig            cpp/skyweaver/PipelineConfig.hpp:14-368
//   DedispersionPlan          cpp/skyweaver/DedispersionPlan.hpp:11-80
//   ObservationHeader         cpp/skyweaver/ObservationHeader.hpp
//   BeamformerPipeline        cpp/skyweaver/BeamformerPipeline.cuh:24-116
//   IncoherentDedispersionPipeline  cpp/skyweaver/IncoherentDedispersionPipeline.cuh
//   MultiFileWriter / FileStream    cpp/skyweaver/MultiFileWriter.cuh:20-86,
//                                   cpp/skyweaver/detail/MultiFileWriter.cu:34-248,
//                                   cpp/skyweaver/src/FileOutputStream.cpp:30-147
//   InvalidDelayEpoch         cpp/skyweaver/DelayManager.cuh:19-34
//   read_dada_header / validate_header / update_config
//                             cpp/skyweaver/src/ObservationHeader.cpp:6-69
//   MultiFileReader           cpp/skyweaver/MultiFileReader.cuh,
//                             cpp/skyweaver/src/MultiFileReader.cu:12-303
//   run_pipeline              cpp/skyweaver/src/skyweaver_cli.cu:109-240
//
// Differences that follow from the design: the reference's compile-time
// parameters (SKYWEAVER_NANTENNAS, _NCHANS, _NBEAMS, _CB_TSCRUNCH, ...) are
// runtime fields of PipelineConfig; handlers receive a BlockView (pointer,
// dimensions and the DescribedVector metadata) instead of thrust vectors;
// data handed to CB / IB / stats handlers lives in device memory and is only
// valid during the call (BlockView::copy_to_host copies it out).
#pragma once

#include "skyweaver_amd.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <ctime>
#include <filesystem>
#include <fstream>
#include <functional>
#include <iomanip>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace skyweaver_amd {

// ------------------------------------------------------------------ errors
class InvalidDelayEpoch : public std::exception {
  public:
    explicit InvalidDelayEpoch(double epoch_) : epoch(epoch_)
    {
        msg = "No valid delay solution for unix epoch " + std::to_string(epoch);
    }
    const char* what() const noexcept override { return msg.c_str(); }
    double epoch;

  private:
    std::string msg;
};

inline std::string last_error()
{
    char buf[1024];
    sw_last_error(buf, sizeof(buf));
    return buf;
}

// Message of an exception a handler threw inside a library callback (on the
// calling thread, or on the library's sink thread with asynchronous
// handlers); the callback returns an error code through the C ABI and
// check() rethrows it on the thread that receives the code.
namespace detail {
struct HandlerError {
    std::mutex mu;
    std::string msg;
};
inline HandlerError& handler_error()
{
    static HandlerError e;
    return e;
}
inline void set_handler_error(std::string const& m)
{
    auto& e = handler_error();
    std::lock_guard<std::mutex> lk(e.mu);
    if(e.msg.empty()) e.msg = m;
}
inline std::string take_handler_error()
{
    auto& e = handler_error();
    std::lock_guard<std::mutex> lk(e.mu);
    std::string m;
    m.swap(e.msg);
    return m;
}
} // namespace detail

// Convert a C ABI status into the reference's exception types.
inline void check(int rc, double epoch = 0.0)
{
    if(rc == SW_OK) return;
    const std::string h = detail::take_handler_error();
    if(!h.empty()) throw std::runtime_error("handler failed: " + h);
    if(rc == SW_ERR_DELAY_EPOCH) throw InvalidDelayEpoch(epoch);
    throw std::runtime_error(last_error());
}

// -------------------------------------------------------- DedispersionPlan
struct DedispersionPlanBlock {
    float coherent_dm = 0.0f;
    int tscrunch = 1;
    std::vector<float> incoherent_dms;
};

class DedispersionPlan {
  public:
    void add_block(float coherent_dm, float dm_begin, float dm_end, float dm_step, int tscrunch)
    {
        DedispersionPlanBlock b;
        b.coherent_dm = coherent_dm;
        b.tscrunch = tscrunch;
        for(float dm = dm_begin; dm < dm_end; dm += dm_step) b.incoherent_dms.push_back(dm);
        if(b.incoherent_dms.empty()) b.incoherent_dms.push_back(dm_begin);
        add_block(std::move(b));
    }
    void add_block(float coherent_dm, int tscrunch = 1)
    {
        DedispersionPlanBlock b;
        b.coherent_dm = coherent_dm;
        b.tscrunch = tscrunch;
        b.incoherent_dms = {coherent_dm};
        add_block(std::move(b));
    }
    void add_block(DedispersionPlanBlock&& b)
    {
        _coherent_dms.push_back(b.coherent_dm);
        _blocks.push_back(std::move(b));
    }
    // "cdm", "cdm:tscrunch" or "cdm:begin:end:step:tscrunch"
    void add_block(std::string const& spec)
    {
        float cdm = 0;
        int ts = 1, n = 0;
        check(sw_ddplan_parse_block(spec.c_str(), &cdm, &ts, nullptr, 0, &n));
        DedispersionPlanBlock b;
        b.coherent_dm = cdm;
        b.tscrunch = ts;
        b.incoherent_dms.resize(n);
        check(sw_ddplan_parse_block(spec.c_str(), &cdm, &ts, b.incoherent_dms.data(), n, &n));
        add_block(std::move(b));
    }
    std::vector<float> const& coherent_dms() const { return _coherent_dms; }
    std::vector<DedispersionPlanBlock> const& blocks() const { return _blocks; }
    DedispersionPlanBlock& operator[](std::size_t i) { return _blocks[i]; }
    DedispersionPlanBlock const& operator[](std::size_t i) const { return _blocks[i]; }

  private:
    std::vector<float> _coherent_dms;
    std::vector<DedispersionPlanBlock> _blocks;
};

// ---------------------------------------------------------- PipelineConfig
class PipelineConfig {
  public:
    // runtime forms of the reference's compile-time parameters
    std::size_t nantennas() const { return _nantennas; }
    void nantennas(std::size_t n) { _nantennas = n; }
    std::size_t nchans() const { return _nchans; }
    void nchans(std::size_t n) { _nchans = n; _freqs.clear(); }
    std::size_t nbeams() const { return _nbeams; }
    void nbeams(std::size_t n) { _nbeams = n; }
    std::size_t cb_tscrunch() const { return _tscrunch; }
    std::size_t ib_tscrunch() const { return _tscrunch; }  // one value for both (the CB indexes with it)
    void tscrunch(std::size_t t) { _tscrunch = t; }
    bool cb_ib_subtract() const { return _ib_subtract; }
    void cb_ib_subtract(bool v) { _ib_subtract = v; }
    std::size_t npol() const { return 2; }
    std::size_t nsamples_per_heap() const { return 256; }

    std::string const& delay_file() const { return _delay_file; }
    void delay_file(std::string const& f) { _delay_file = f; }
    std::string const& output_dir() const { return _output_dir; }
    void output_dir(std::string const& d) { _output_dir = d; }
    std::string const& output_file_prefix() const { return _prefix; }
    void output_file_prefix(std::string const& p) { _prefix = p; }
    std::size_t max_output_filesize() const { return _max_filesize; }
    void max_output_filesize(std::size_t s) { _max_filesize = s; }
    std::size_t dada_header_size() const { return _header_size; }
    void dada_header_size(std::size_t s) { _header_size = s; }

    double centre_frequency() const { return _cfreq; }
    void centre_frequency(double f) { _cfreq = f; _freqs.clear(); }
    double bandwidth() const { return _bw; }
    void bandwidth(double bw) { _bw = bw; _freqs.clear(); }
    std::vector<double> const& channel_frequencies() const
    {
        if(_freqs.size() != _nchans) {
            _freqs.resize(_nchans);
            check(sw_channel_frequencies(_cfreq, _bw, (int)_nchans, _freqs.data()));
        }
        return _freqs;
    }
    std::vector<float> const& coherent_dms() const { return _ddplan.coherent_dms(); }
    DedispersionPlan& ddplan() { return _ddplan; }
    DedispersionPlan const& ddplan() const { return _ddplan; }
    void enable_incoherent_dedispersion(bool e) { _enable_id = e; }
    bool enable_incoherent_dedispersion() const { return _enable_id; }
    std::size_t gulp_length_samps() const { return _gulp; }
    void gulp_length_samps(std::size_t n) { _gulp = n; }
    float output_level() const { return _level; }
    void output_level(float l) { _level = l; }
    std::size_t total_nchans() const { return _total_nchans; }
    void total_nchans(std::size_t n) { _total_nchans = n; }
    std::string stokes_mode() const { return _stokes; }
    void stokes_mode(std::string const& s) { _stokes = s; }

    // MI355X-specific knobs: coherent beamformer kernel form, FFT length and
    // overlap per coherent DM (0 = automatic)
    int cb_form() const { return _cb_form; }
    void cb_form(int f) { _cb_form = f; }
    std::vector<int> const& fft_lengths() const { return _fft_len; }
    void fft_lengths(std::vector<int> const& v) { _fft_len = v; }
    // coherent DMs of equal FFT length share one overlap and their forward
    // transforms (sw_codedisp_config::share_forward)
    bool share_forward() const { return _share_forward; }
    void share_forward(bool v) { _share_forward = v; }
    // FFT lengths (and share_forward) timed on the device at construction
    // when no fft_lengths are given
    bool autotune_fft() const { return _autotune_fft; }
    void autotune_fft(bool v) { _autotune_fft = v; }
    // handlers called by a library thread while the next DMs and gulps are
    // processed (sw_pipeline_config::async_sinks); synchronize() waits for
    // them and rethrows a handler's exception
    bool async_handlers() const { return _async; }
    void async_handlers(bool v) { _async = v; }
    // each gulp after the first captured into a hip graph and launched as one
    // (sw_pipeline_config::graph; needs async_handlers)
    bool graph_mode() const { return _graph; }
    void graph_mode(bool v) { _graph = v; }
    // the dedispersion chirp's bin-to-frequency mapping: the reference's
    // (false, the default) or the physical one, channel-centred with the
    // upper half of the bins at negative baseband frequencies (true)
    bool physical_chirp() const { return _physical_chirp; }
    void physical_chirp(bool v) { _physical_chirp = v; }

    // input (MultiFileReader, run_pipeline)
    std::vector<std::string> const& input_files() const { return _input_files; }
    void input_files(std::vector<std::string> const& f) { _input_files = f; }
    bool check_input_contiguity() const { return _check_contiguity; }
    void check_input_contiguity(bool c) { _check_contiguity = c; }
    float start_time() const { return _start_time; }
    void start_time(float t) { _start_time = t; }
    float duration() const { return _duration; }
    void duration(float d) { _duration = d; }

  private:
    std::size_t _nantennas = 64, _nchans = 64, _nbeams = 64, _tscrunch = 1;
    bool _ib_subtract = true;
    std::string _delay_file, _output_dir = ".", _prefix, _stokes = "I";
    std::size_t _max_filesize = 10000000000ull, _header_size = 4096;
    double _cfreq = 1284e6, _bw = 13375000.0;
    mutable std::vector<double> _freqs;
    DedispersionPlan _ddplan;
    bool _enable_id = false;
    std::size_t _gulp = 4096, _total_nchans = 4096;
    float _level = 24.0f;
    int _cb_form = SW_CB_MFMA;
    std::vector<int> _fft_len;
    bool _share_forward = false, _autotune_fft = false, _async = false, _graph = false, _physical_chirp = false;
    std::vector<std::string> _input_files;
    bool _check_contiguity = false;
    float _start_time = 0.0f, _duration = std::numeric_limits<float>::infinity();
};

// ------------------------------------------------------- ObservationHeader
struct ObservationHeader {
    std::size_t nchans = 0;
    std::size_t npol = 0;
    std::size_t nbits = 0;
    std::size_t nantennas = 0;
    std::size_t sample_clock_start = 0;
    std::size_t chan0_idx = 0;
    std::size_t obs_nchans = 0;
    long double bandwidth = 0;
    long double obs_bandwidth = 0;
    long double frequency = 0;
    long double obs_frequency = 0;
    long double tsamp = 0;
    long double sample_clock = 1712e6;
    long double sync_time = 0;
    long double utc_start = 0;
    long double mjd_start = 0;
    std::size_t obs_offset = 0;  // bytes of this file's data from the observation start
    std::string source_name, ra, dec, telescope, instrument;
};

namespace detail {
// psrdada's ascii_header_get(header, key, "%s", buf): the first line whose
// first word is `key`; the value is the next whitespace-delimited word.
inline std::string dada_header_value(char const* text, std::size_t n, char const* key)
{
    const std::string hdr(text, strnlen(text, n));
    const std::size_t klen = std::strlen(key);
    std::size_t pos = 0;
    while(pos < hdr.size()) {
        std::size_t end = hdr.find('\n', pos);
        if(end == std::string::npos) end = hdr.size();
        std::size_t b = pos;
        while(b < end && (hdr[b] == ' ' || hdr[b] == '\t')) ++b;
        if(end - b > klen && hdr.compare(b, klen, key) == 0 && std::isspace((unsigned char)hdr[b + klen])) {
            std::istringstream ss(hdr.substr(b + klen, end - b - klen));
            std::string v;
            if(ss >> v) {
                if(v == "unset") throw std::runtime_error(std::string("The header key ") + key + " was unset");
                return v;
            }
        }
        pos = end + 1;
    }
    throw std::runtime_error(std::string("Could not find ") + key + " key in DADA header.");
}
} // namespace detail

// read_dada_header (cpp/skyweaver/src/ObservationHeader.cpp:6-35): every key
// is required; integers parse as strtoul(.., 0), reals as strtold.
inline void read_dada_header(char const* text, std::size_t nbytes, ObservationHeader& h)
{
    auto u = [&](char const* k) { return (std::size_t)std::strtoul(detail::dada_header_value(text, nbytes, k).c_str(), nullptr, 0); };
    auto r = [&](char const* k) { return std::strtold(detail::dada_header_value(text, nbytes, k).c_str(), nullptr); };
    auto str = [&](char const* k) { return detail::dada_header_value(text, nbytes, k); };
    h.nchans = u("NCHAN");
    h.obs_nchans = u("OBS_NCHAN");
    h.npol = u("NPOL");
    h.nbits = u("NBIT");
    h.nantennas = u("NANT");
    h.sample_clock_start = u("SAMPLE_CLOCK_START");
    h.bandwidth = r("BW");
    h.obs_bandwidth = r("OBS_BW");
    h.frequency = r("FREQ");
    h.obs_frequency = r("OBS_FREQ");
    h.tsamp = r("TSAMP");
    h.sample_clock = r("SAMPLE_CLOCK");
    h.sync_time = r("SYNC_TIME");
    h.utc_start = r("UTC_START");
    h.mjd_start = r("MJD_START");
    h.source_name = str("SOURCE");
    h.ra = str("RA");
    h.dec = str("DEC");
    h.telescope = str("TELESCOPE");
    h.instrument = str("INSTRUMENT");
    h.chan0_idx = u("CHAN0_IDX");
    h.obs_offset = u("OBS_OFFSET");
}

// validate_header / update_config (cpp/skyweaver/src/ObservationHeader.cpp:37-69)
inline void validate_header(ObservationHeader const& header, PipelineConfig const& config)
{
    if(header.nantennas > config.nantennas())
        throw std::runtime_error("Input data contains too many antennas for current build, "
                                 "-DSKYWEAVER_NANTENNAS should be larger than the data nantennas");
    if(header.nchans != config.nchans())
        throw std::runtime_error("Input data contains incorrect number of channels for current build, "
                                 "-DSKYWEAVER_NCHANS should be equal to the data nchans");
    if(header.npol != config.npol())
        throw std::runtime_error("Input data contains incorrect number of polarisations for current build, "
                                 "-DSKYWEAVER_NPOL should be equal to the data npol");
    if(header.nbits != 8)
        throw std::runtime_error("Input data contains incorrect number of bits per sample, "
                                 "currently only 8-bit input supported");
}
inline void update_config(PipelineConfig& config, ObservationHeader const& header)
{
    config.bandwidth((double)header.bandwidth);
    config.centre_frequency((double)header.frequency);
}

// ------------------------------------------------------------- block views
// What a handler receives: the data pointer and the DescribedVector metadata
// the reference's handlers read (dims, tsamp, utc_offset, reference_dm, dms).
class BlockView {
  public:
    explicit BlockView(sw_block const& b) : _b(b) {}
    void const* data() const { return _b.data; }
    bool on_device() const { return _b.on_device != 0; }
    std::size_t size() const { return _b.bytes; }  // bytes (int8 / char4 elements)
    std::vector<std::size_t> dims() const { return {_b.dims, _b.dims + _b.ndims}; }
    std::string dims_as_string() const { return _b.order; }
    double tsamp() const { return _b.tsamp; }
    long double utc_offset() const { return _b.utc_offset; }
    float reference_dm() const { return _b.reference_dm; }
    std::vector<float> dms() const { return _b.dms ? std::vector<float>(_b.dms, _b.dms + _b.ndms) : std::vector<float>{}; }
    std::size_t ndms() const { return _b.dms ? (std::size_t)_b.ndms : 0; }
    int kind() const { return _b.kind; }
    std::size_t nsamples() const { return (_b.kind == SW_BLOCK_IB_BTF) ? _b.dims[1] : _b.dims[0]; }
    std::size_t nchannels() const
    {
        switch(_b.kind) {
        case SW_BLOCK_CB_TFB: return _b.dims[1];
        case SW_BLOCK_IB_BTF: return _b.dims[2];
        case SW_BLOCK_STATS_FPA: return _b.dims[0];
        default: return 1;
        }
    }
    std::size_t nbeams() const
    {
        switch(_b.kind) {
        case SW_BLOCK_CB_TFB: return _b.dims[2];
        case SW_BLOCK_IB_BTF: return _b.dims[0];
        case SW_BLOCK_TDB: return _b.dims[2];
        default: return 0;
        }
    }
    void copy_to_host(void* dst) const
    {
        if(on_device()) {
            check(sw_memcpy_dtoh_sync(dst, _b.data, _b.bytes));
        } else {
            std::memcpy(dst, _b.data, _b.bytes);
        }
    }
    std::vector<int8_t> to_host() const
    {
        std::vector<int8_t> v(_b.bytes);
        copy_to_host(v.data());
        return v;
    }
    sw_block const& raw() const { return _b; }

  private:
    sw_block const& _b;
};

// ------------------------------------------------------- beamformer traits
enum class StokesParameter { I = SW_STOKES_I, Q = SW_STOKES_Q, U = SW_STOKES_U, V = SW_STOKES_V };
template <StokesParameter S>
struct SingleStokesBeamformerTraits {
    static constexpr int stokes = static_cast<int>(S);
};
struct FullStokesBeamformerTraits {
    static constexpr int stokes = SW_STOKES_IQUV;
};

// --------------------------------------------------------- handler shims
namespace detail {
template <typename H>
int cb_trampoline(void* user, sw_block const* b)
{
    try {
        (*static_cast<H*>(user))(BlockView(*b), static_cast<std::size_t>(b->dm_idx));
        return 0;
    } catch(std::exception const& e) {
        set_handler_error(e.what());
        return 1;
    } catch(...) {
        set_handler_error("unknown exception");
        return 1;
    }
}
template <typename H>
int stats_trampoline(void* user, sw_block const* b)
{
    try {
        (*static_cast<H*>(user))(BlockView(*b));
        return 0;
    } catch(std::exception const& e) {
        set_handler_error(e.what());
        return 1;
    } catch(...) {
        set_handler_error("unknown exception");
        return 1;
    }
}
} // namespace detail

// ------------------------------------------ IncoherentDedispersionPipeline
// A CB handler that aggregates TFB blocks per coherent DM and hands TDB
// blocks (device memory) to Handler(BlockView, dm_idx).
template <typename Handler>
class IncoherentDedispersionPipeline {
  public:
    // dispatch_size 0: the reference's max(10 max_delay, 8192) samples
    IncoherentDedispersionPipeline(PipelineConfig const& config, Handler& handler, int dispatch_size = 0)
        : _config(config), _handler(handler)
    {
        auto const& blocks = config.ddplan().blocks();
        if(blocks.empty()) throw std::runtime_error("No DDPlan set for pipeline");
        std::vector<float> cdms;
        std::vector<int> ts, nd;
        std::vector<float const*> dms;
        for(auto const& b: blocks) {
            cdms.push_back(b.coherent_dm);
            ts.push_back(b.tscrunch);
            dms.push_back(b.incoherent_dms.data());
            nd.push_back((int)b.incoherent_dms.size());
        }
        sw_idpipe_config c{};
        c.nchans = (int)config.nchans();
        c.centre_frequency = config.centre_frequency();
        c.bandwidth = config.bandwidth();
        c.cb_tscrunch = (int)config.cb_tscrunch();
        c.nbeams = (int)config.nbeams();
        c.nvec = config.stokes_mode() == "IQUV" ? 4 : 1;
        c.nblocks = (int)blocks.size();
        c.coherent_dms = cdms.data();
        c.tscrunch = ts.data();
        c.dms = dms.data();
        c.ndms = nd.data();
        c.dispatch_size = dispatch_size;
        check(sw_idpipe_create(&c, &_p));
        check(sw_idpipe_set_sink(_p, &detail::cb_trampoline<Handler>, &_handler));
    }
    ~IncoherentDedispersionPipeline() { sw_idpipe_destroy(_p); }
    IncoherentDedispersionPipeline(IncoherentDedispersionPipeline const&) = delete;
    IncoherentDedispersionPipeline& operator=(IncoherentDedispersionPipeline const&) = delete;

    template <typename Header>
    void init(Header const& header) { _handler.init(header); }
    void operator()(BlockView const& tfb, std::size_t dm_idx)
    {
        auto const& b = tfb.raw();
        check(sw_idpipe_push(_p, (int)dm_idx, b.data, b.dims[0], b.on_device, b.utc_offset, b.tsamp, nullptr));
    }

  private:
    PipelineConfig const& _config;
    Handler& _handler;
    sw_idpipe* _p = nullptr;
};

// ------------------------------------------------------ BeamformerPipeline
template <typename CBHandler, typename IBHandler, typename StatsHandler,
          typename BeamformerTraits = SingleStokesBeamformerTraits<StokesParameter::I>>
class BeamformerPipeline {
  public:
    BeamformerPipeline(PipelineConfig const& config, CBHandler& cb_handler, IBHandler& ib_handler,
                       StatsHandler& stats_handler)
        : _config(config), _cb(cb_handler), _ib(ib_handler), _stats(stats_handler)
    {
        if(config.coherent_dms().empty()) throw std::runtime_error("No DDPlan set for pipeline");
        if(config.gulp_length_samps() % config.nsamples_per_heap() != 0)
            throw std::runtime_error("Gulp size is not a multiple of the number of samples per heap");
        _dms = config.coherent_dms();
        sw_pipeline_config c{};
        c.nantennas = (int)config.nantennas();
        c.nchans = (int)config.nchans();
        c.nbeams = (int)config.nbeams();
        c.nsamples = (int)config.gulp_length_samps();
        c.tscrunch = (int)config.cb_tscrunch();
        c.stokes = BeamformerTraits::stokes;
        c.ib_subtract = config.cb_ib_subtract() ? 1 : 0;
        c.output_level = config.output_level();
        c.centre_frequency = config.centre_frequency();
        c.bandwidth = config.bandwidth();
        c.total_nchans = (int)config.total_nchans();
        c.ndms = (int)_dms.size();
        c.coherent_dms = _dms.data();
        if(!config.fft_lengths().empty()) {
            if(config.fft_lengths().size() != _dms.size())
                throw std::runtime_error("one FFT length per coherent DM expected");
            c.fft_len = config.fft_lengths().data();
        }
        c.cb_form = config.cb_form();
        c.batched_dms = 1;
        c.share_forward = config.share_forward() ? 1 : 0;
        c.autotune_fft = config.autotune_fft() ? 1 : 0;
        c.async_sinks = config.async_handlers() ? 1 : 0;
        c.graph = config.graph_mode() ? 1 : 0;
        c.chirp = config.physical_chirp() ? SW_CHIRP_PHYSICAL : SW_CHIRP_REFERENCE;
        check(sw_pipeline_create(&c, &_p));
        try {
            check(sw_pipeline_set_sinks(_p, &detail::cb_trampoline<CBHandler>, &_cb,
                                        &detail::cb_trampoline<IBHandler>, &_ib,
                                        &detail::stats_trampoline<StatsHandler>, &_stats));
            if(!config.delay_file().empty()) check(sw_pipeline_set_delay_file(_p, config.delay_file().c_str()));
        } catch(...) {
            sw_pipeline_destroy(_p);
            throw;
        }
    }
    ~BeamformerPipeline() { sw_pipeline_destroy(_p); }
    BeamformerPipeline(BeamformerPipeline const&) = delete;
    BeamformerPipeline& operator=(BeamformerPipeline const&) = delete;

    void init(ObservationHeader const& header, long double utc_offset = 0.0)
    {
        _header = header;
        sw_obs_header h{};
        h.nantennas = (int)header.nantennas;
        h.utc_start = (double)header.utc_start;
        h.sample_clock = (double)header.sample_clock;
        h.utc_offset = (double)utc_offset;
        check(sw_pipeline_init(_p, &h));
        _cb.init(header);
        _ib.init(header);
        _stats.init(header);
        _utc_offset = utc_offset;
    }

    // Static weights [F][B][A] complex int8 (host) when no delay file is used.
    void set_weights(void const* h_weights) { check(sw_pipeline_set_weights(_p, h_weights)); }
    void set_beamsets(int nbeamsets, int const* h_map, float const* h_weights)
    {
        check(sw_pipeline_set_beamsets(_p, nbeamsets, h_map, h_weights));
    }

    // One gulp of TAFTP voltages from host memory (any contiguous container
    // of bytes / char2); always returns false like the reference.
    template <typename Container>
    bool operator()(Container const& taftp_on_host)
    {
        const std::size_t bytes = taftp_on_host.size() * sizeof(*taftp_on_host.data());
        const long double t = _header.utc_start + _utc_offset +
                              (long double)(_calls * 2ull * _config.total_nchans() * _config.gulp_length_samps()) /
                                  _header.sample_clock;
        check(sw_pipeline_process(_p, taftp_on_host.data(), bytes, 0), (double)t);
        ++_calls;
        return false;
    }
    // A gulp already in device memory (no host transfer).
    bool process_device(void const* d_taftp, std::size_t bytes)
    {
        check(sw_pipeline_process(_p, d_taftp, bytes, 1));
        ++_calls;
        return false;
    }
    // operator() in two halves: prefetch starts the host-to-device copy of a
    // gulp (at most two pending; the host buffer must stay unchanged until
    // the process_prefetched call that consumes it returns), and
    // process_prefetched processes the oldest pending gulp.
    template <typename Container>
    void prefetch(Container const& taftp_on_host)
    {
        check(sw_pipeline_prefetch(_p, taftp_on_host.data(), taftp_on_host.size() * sizeof(*taftp_on_host.data())));
    }
    bool process_prefetched()
    {
        check(sw_pipeline_process_prefetched(_p));
        ++_calls;
        return false;
    }
    // wait for every copy and kernel the pipeline has enqueued
    void synchronize() { check(sw_pipeline_synchronize(_p)); }
    std::vector<float> stage_times() const
    {
        std::vector<float> v(SW_NSTAGES);
        check(sw_pipeline_stage_times(_p, v.data(), SW_NSTAGES));
        return v;
    }

  private:
    PipelineConfig const& _config;
    CBHandler& _cb;
    IBHandler& _ib;
    StatsHandler& _stats;
    std::vector<float> _dms;
    sw_pipeline* _p = nullptr;
    ObservationHeader _header;
    long double _utc_offset = 0;
    std::size_t _calls = 0;
};

// ------------------------------------------------------------ file output
// FileStream: a sequence of files of at most bytes_per_file data bytes, each
// starting with a header produced by a callback; files are named
// <dir>/<base>_<16-digit byte offset><extension>.
class FileStream {
  public:
    using HeaderCallback = std::function<std::string(std::size_t bytes_written, std::size_t file_idx)>;
    FileStream(std::string dir, std::string base, std::string ext, std::size_t bytes_per_file, HeaderCallback cb)
        : _dir(std::move(dir)), _base(std::move(base)), _ext(std::move(ext)), _per_file(bytes_per_file),
          _cb(std::move(cb))
    {
        if(_per_file == 0) throw std::runtime_error("The number of bytes per file must be greater than zero");
        std::filesystem::create_directories(_dir);
    }
    void write(char const* ptr, std::size_t bytes)
    {
        while(bytes > 0) {
            if(!_file.is_open() || _in_file == _per_file) new_file();
            const std::size_t n = std::min(bytes, _per_file - _in_file);
            _file.write(ptr, (std::streamsize)n);
            if(!_file) throw std::runtime_error("Error while writing to " + _path);
            _in_file += n;
            _total += n;
            ptr += n;
            bytes -= n;
        }
    }
    std::vector<std::string> const& files() const { return _files; }

  private:
    void new_file()
    {
        if(_file.is_open()) _file.close();
        std::ostringstream p;
        p << _dir << "/" << _base << "_" << std::setfill('0') << std::setw(16) << _total << _ext;
        _path = p.str();
        _file.open(_path, std::ios::out | std::ios::binary);
        if(!_file.is_open()) throw std::runtime_error("Could not open file " + _path);
        const std::string h = _cb(_total, _files.size());
        _file.write(h.data(), (std::streamsize)h.size());
        _files.push_back(_path);
        _in_file = 0;
    }
    std::string _dir, _base, _ext, _path;
    std::size_t _per_file, _in_file = 0, _total = 0;
    HeaderCallback _cb;
    std::ofstream _file;
    std::vector<std::string> _files;
};

// DADA header text: "KEY value" lines in a zero-padded block of hdr_size
// bytes; set() replaces a key's line or appends one.
class DadaHeader {
  public:
    explicit DadaHeader(std::string text) : _text(std::move(text)) {}
    template <typename T>
    void set(std::string const& key, T const& value, int precision = -1)
    {
        std::ostringstream v;
        if(precision >= 0) v << std::fixed << std::setprecision(precision);
        v << value;
        set_str(key, v.str());
    }
    void set_list(std::string const& key, std::vector<float> const& values, int precision)
    {
        std::ostringstream v;
        v << std::fixed << std::setprecision(precision);
        for(std::size_t i = 0; i < values.size(); ++i) v << (i ? "," : "") << values[i];
        set_str(key, v.str());
    }
    std::string block(std::size_t hdr_size) const
    {
        if(_text.size() >= hdr_size) throw std::runtime_error("DADA header larger than HDR_SIZE");
        std::string b(hdr_size, '\0');
        std::memcpy(&b[0], _text.data(), _text.size());
        return b;
    }
    std::string const& text() const { return _text; }

  private:
    void set_str(std::string const& key, std::string const& value)
    {
        std::ostringstream line;
        line << std::left << std::setw(12) << key << " " << value;
        std::size_t pos = 0;
        while(pos < _text.size()) {
            std::size_t eol = _text.find('\n', pos);
            if(eol == std::string::npos) eol = _text.size();
            std::string l = _text.substr(pos, eol - pos);
            if(l.compare(0, key.size(), key) == 0 && (l.size() == key.size() || l[key.size()] == ' ' ||
                                                        l[key.size()] == '\t')) {
                _text.replace(pos, eol - pos, line.str());
                return;
            }
            pos = eol + 1;
        }
        if(!_text.empty() && _text.back() != '\n') _text += '\n';
        _text += line.str() + "\n";
    }
    std::string _text;
};

// MultiFileWriter: one DADA file stream per stream index (coherent DM),
// directory <output_dir>/<utc yyyy-mm-dd-hh:mm:ss>/<stream>/<freq:09d>,
// files <prefix>_<utc>_<stream>_<dm:09.3f>_<freq:09.0f>[_tag]_<offset:016d>.<order>,
// header keys as cpp/skyweaver/detail/MultiFileWriter.cu:34-175.
class MultiFileWriter {
  public:
    MultiFileWriter(PipelineConfig const& config, std::string tag = "") : _config(config), _tag(std::move(tag)) {}
    void init(ObservationHeader const& header) { _header = header; }
    MultiFileWriter(MultiFileWriter const&) = delete;
    MultiFileWriter& operator=(MultiFileWriter const&) = delete;
    ~MultiFileWriter()
    {
        if(_pinned) sw_host_free(_pinned);
    }
    bool operator()(BlockView const& block, std::size_t stream_idx)
    {
        auto it = _streams.find(stream_idx);
        if(it == _streams.end()) it = _streams.emplace(stream_idx, create_stream(block, stream_idx)).first;
        const std::size_t n = block.size();
        if(block.on_device()) {
            // device blocks come out through one reusable pinned buffer: the
            // copy runs at DMA rate instead of through a pageable staging copy
            if(n > _pinned_size) {
                if(_pinned) sw_host_free(_pinned);
                _pinned = nullptr;
                _pinned_size = 0;
                check(sw_host_alloc(&_pinned, n));
                _pinned_size = n;
            }
            block.copy_to_host(_pinned);
            it->second->write(static_cast<char const*>(_pinned), n);
        } else {
            _buf.resize(n);
            block.copy_to_host(_buf.data());
            it->second->write(reinterpret_cast<char const*>(_buf.data()), n);
        }
        return false;
    }
    FileStream const& stream(std::size_t idx) const { return *_streams.at(idx); }

    static std::string formatted_time(long double unix_timestamp)
    {
        char out[80];
        std::time_t t = (std::time_t)unix_timestamp;
        std::tm tmv{};
        gmtime_r(&t, &tmv);
        std::strftime(out, sizeof(out), "%Y-%m-%d-%H:%M:%S", &tmv);
        return out;
    }

  private:
    std::unique_ptr<FileStream> create_stream(BlockView const& b, std::size_t idx)
    {
        const std::size_t per_file = std::max<std::size_t>(1, _config.max_output_filesize() / b.size()) * b.size();
        std::ostringstream dir, base;
        dir << _config.output_dir() << "/" << formatted_time(_header.utc_start) << "/" << idx << "/" << std::fixed
            << std::setfill('0') << std::setw(9) << (int)_header.frequency;
        if(!_config.output_file_prefix().empty()) base << _config.output_file_prefix() << "_";
        base << formatted_time(_header.utc_start) << "_" << idx << "_" << std::fixed << std::setprecision(3)
             << std::setfill('0') << std::setw(9) << b.reference_dm() << "_" << std::setprecision(0)
             << std::setfill('0') << std::setw(9) << (double)_header.frequency;
        if(!_tag.empty()) base << "_" << _tag;
        std::string ext = b.dims_as_string();
        for(auto& c: ext) c = (char)std::tolower((unsigned char)c);
        // metadata of the first block of the stream
        const std::size_t nbeams = b.nbeams(), nchans = b.nchannels(), nsamp = b.nsamples();
        const std::vector<float> dms = b.dms();
        const float rdm = b.reference_dm();
        const double tsamp = b.tsamp();
        const long double utc0 = _header.utc_start + b.utc_offset();
        const std::string order = b.dims_as_string();
        auto cb = [=, this](std::size_t bytes_written, std::size_t file_idx) {
            DadaHeader h(default_header());
            h.set("NBEAM", nbeams);
            h.set("NCHAN", nchans);
            h.set("OBS_NCHAN", _header.obs_nchans);
            h.set("OBS_FREQUENCY", (double)_header.obs_frequency, 6);
            h.set("OBS_BW", (double)_header.obs_bandwidth, 6);
            h.set("NSAMP", nsamp);
            if(!dms.empty()) {
                h.set("NDMS", dms.size());
                h.set_list("DMS", dms, 7);
            }
            h.set("COHERENT_DM", (double)rdm, 6);
            h.set("FREQ", (double)_header.frequency, 6);
            h.set("BW", (double)_header.bandwidth, 6);
            h.set("TSAMP", tsamp * 1e6, 6);
            h.set("NPOL", _config.stokes_mode() == "IQUV" ? 4 : 1);
            h.set("STOKES_MODE", _config.stokes_mode());
            h.set("ORDER", order);
            h.set("CHAN0_IDX", _header.chan0_idx);
            h.set("FILE_SIZE", per_file);
            h.set("FILE_NUMBER", file_idx);
            h.set("OBS_OFFSET", bytes_written);
            h.set("OBS_OVERLAP", 0);
            h.set("UTC_START", (double)utc0, 6);
            return h.block(_config.dada_header_size());
        };
        return std::make_unique<FileStream>(dir.str(), base.str(), "." + ext, per_file, cb);
    }
    static std::string default_header()
    {
        return "HEADER       DADA\nHDR_VERSION  1.0\nHDR_SIZE     4096\nDADA_VERSION 1.0\n"
               "FILE_SIZE    100000000000\nFILE_NUMBER  0\nTELESCOPE    MeerKAT\nINSTRUMENT   CBF-Feng\n"
               "RECEIVER     L-band\nNBIT         8\nNDIM         1\nNPOL         1\n";
    }

    PipelineConfig const& _config;
    std::string _tag;
    ObservationHeader _header;
    std::map<std::size_t, std::unique_ptr<FileStream>> _streams;
    std::vector<int8_t> _buf;
    void* _pinned = nullptr;
    std::size_t _pinned_size = 0;
};

// A handler that accepts and discards blocks (for unused outputs).
struct NullHandler {
    template <typename H>
    void init(H const&) {}
    void operator()(BlockView const&, std::size_t = 0) {}
};

// ------------------------------------------------------------ file input
// MultiFileReader (cpp/skyweaver/src/MultiFileReader.cu:35-303): the data of
// a sequence of DADA files (each a header of dada_header_size bytes, then
// TAFTP voltages) as one byte stream.  Positions count data bytes from the
// start of the first file's data.  (The reference also resets its position
// to the header size whenever it opens a file; here tellg() is always the
// data offset, which the reference's own tests do not observe.)
class MultiFileReader {
  public:
    explicit MultiFileReader(PipelineConfig const& config)
        : _files(config.input_files()), _hsize(config.dada_header_size())
    {
        std::vector<char> hdr(_hsize);
        for(auto const& file : _files) {
            std::ifstream f(file, std::ios::in | std::ios::binary);
            if(!f.is_open()) throw std::runtime_error("Could not open file " + file);
            f.read(hdr.data(), (std::streamsize)_hsize);
            if((std::size_t)f.gcount() != _hsize) throw std::runtime_error("Reached end of file reading the header of " + file);
            ObservationHeader h;
            read_dada_header(hdr.data(), _hsize, h);
            _headers.push_back(h);
            f.seekg(0, std::ios::end);
            const std::size_t size = (std::size_t)f.tellg();
            _sizes.push_back(size - _hsize);
            _total += size - _hsize;
        }
        if(_files.empty()) throw std::runtime_error("No input files");
        if(config.check_input_contiguity()) check_contiguity();
        open(0);
    }
    MultiFileReader(MultiFileReader const&) = delete;
    MultiFileReader& operator=(MultiFileReader const&) = delete;

    ObservationHeader const& get_header() const { return _headers[0]; }
    ObservationHeader const& get_current_header() const { return _headers[_idx]; }
    std::size_t get_total_size() const { return _total; }
    bool can_read(std::size_t bytes) const { return _total - _pos >= bytes; }
    bool eof() const { return _eof; }
    bool good() const { return _in.good() && !_eof; }
    std::size_t tellg() const { return _pos; }
    bool is_open() const { return _in.is_open(); }
    void close()
    {
        if(_in.is_open()) _in.close();
        _eof = true;
    }

    // seekg(pos, beg | cur | end); a target past the data end throws, the
    // end itself sets eof
    void seekg(long pos, std::ios_base::seekdir dir = std::ios::beg)
    {
        long long target = pos;
        if(dir == std::ios::cur) target = (long long)_pos + pos;
        if(dir == std::ios::end) target = (long long)_total + pos;
        if(target < 0 || (std::size_t)target > _total)
            throw std::runtime_error("Invalid seek position: " + std::to_string(target));
        std::size_t cum = 0;
        for(std::size_t i = 0; i < _files.size(); ++i) {
            if((std::size_t)target < cum + _sizes[i]) {
                if(i != _idx || !_in.is_open()) open(i);
                _in.clear();
                _in.seekg((std::streamoff)(_hsize + (target - cum)), std::ios::beg);
                _pos = (std::size_t)target;
                _eof = false;
                return;
            }
            cum += _sizes[i];
        }
        _pos = _total;
        _eof = true;
    }

    // Reads up to bytes, continuing into the next files; returns the count
    std::streamsize read(char* dst, std::streamsize bytes)
    {
        std::streamsize total = 0;
        while(bytes > 0 && !_eof) {
            if(!_in.is_open()) throw std::runtime_error("File not open: " + _files[_idx]);
            _in.read(dst + total, bytes);
            const std::streamsize got = _in.gcount();
            total += got;
            bytes -= got;
            _pos += (std::size_t)got;
            if(bytes > 0) {  // this file is exhausted
                if(_idx + 1 < _files.size())
                    open(_idx + 1);
                else
                    _eof = true;
            }
        }
        if(_pos >= _total) _eof = true;
        return total;
    }

  private:
    void open(std::size_t i)
    {
        if(_in.is_open()) _in.close();
        _in.clear();
        _in.open(_files[i], std::ios::in | std::ios::binary);
        if(!_in.good()) throw std::runtime_error("Could not open file " + _files[i]);
        _in.seekg((std::streamoff)_hsize, std::ios::beg);
        _idx = i;
    }
    // each file's OBS_OFFSET must equal the data bytes of the files before it
    void check_contiguity() const
    {
        std::size_t so_far = _sizes[0];
        for(std::size_t i = 1; i < _sizes.size(); ++i) {
            if(_headers[i].obs_offset != so_far)
                throw std::runtime_error("These are not contiguous:" + _files[i - 1] + " and " + _files[i]);
            so_far += _sizes[i];
        }
        if(so_far != _total) throw std::runtime_error("Total size does not match the sum of individual sizes");
    }

    std::vector<std::string> _files;
    std::size_t _hsize;
    std::vector<ObservationHeader> _headers;
    std::vector<std::size_t> _sizes;
    std::size_t _total = 0, _pos = 0, _idx = 0;
    bool _eof = false;
    std::ifstream _in;
};

// run_pipeline (cpp/skyweaver/src/skyweaver_cli.cu:109-240): gulps of
// gulp_length_samps samples from the reader, starting start_time() seconds
// in (floored to whole heaps).  Three stages overlap: a thread reads gulp
// k + 2 from the files into pinned host memory while gulp k + 1 is copied to
// the device (prefetch) and gulp k is processed; three pinned buffers
// rotate.  As in the reference the gulp whose read reaches the end of the
// data is not processed, and processing stops once duration() of data has
// been done.  Returns the number of gulps processed.
template <typename Pipeline>
std::size_t run_pipeline(Pipeline& pipeline, PipelineConfig const& config, MultiFileReader& reader,
                         ObservationHeader const& header)
{
    const std::size_t bytes_per_sample = header.nantennas * config.nchans() * config.npol() * 2;
    const std::size_t input_bytes = bytes_per_sample * config.gulp_length_samps();
    const double tsamp = (double)(header.obs_nchans / header.obs_bandwidth);
    std::size_t offset = (std::size_t)(config.start_time() / tsamp);
    offset = offset / config.nsamples_per_heap() * config.nsamples_per_heap();
    pipeline.init(header, offset * tsamp);
    reader.seekg((long)(offset * bytes_per_sample), std::ios::beg);

    struct Pinned {
        void* p = nullptr;
        std::size_t n;
        explicit Pinned(std::size_t bytes) : n(bytes) { check(sw_host_alloc(&p, bytes)); }
        ~Pinned() { sw_host_free(p); }
        char* data() const { return static_cast<char*>(p); }
        std::size_t size() const { return n; }
    };
    Pinned b0(input_bytes), b1(input_bytes), b2(input_bytes);
    Pinned* buf[3] = {&b0, &b1, &b2};
    struct Drain {  // no copy from the buffers may be in flight when they are freed
        Pipeline& p;
        ~Drain()
        {
            try {
                p.synchronize();
            } catch(...) {
            }
        }
    } drain{pipeline};
    std::string read_error;
    auto read_into = [&](Pinned* b) {
        try {
            reader.read(b->data(), (std::streamsize)input_bytes);
        } catch(std::exception const& e) {
            read_error = e.what();
        }
    };
    auto check_read = [&] {
        if(!read_error.empty()) throw std::runtime_error("Error in input file read: " + read_error);
    };
    // gulp k is processed iff its read did not reach the end of the data
    read_into(buf[0]);
    check_read();
    if(reader.eof()) return 0;
    pipeline.prefetch(*buf[0]);
    read_into(buf[1]);
    check_read();
    bool next_ok = !reader.eof();
    std::size_t gulps = 0;
    double data_time = 0.0;
    for(std::size_t k = 0;; ++k) {
        if(next_ok) pipeline.prefetch(*buf[(k + 1) % 3]);
        std::thread next;
        // buffer (k + 2) % 3 held gulp k - 1, whose copy ended before it was processed
        if(next_ok) next = std::thread(read_into, buf[(k + 2) % 3]);
        try {
            pipeline.process_prefetched();
        } catch(...) {
            if(next.joinable()) next.join();
            throw;
        }
        ++gulps;
        const bool more = next_ok;
        if(next.joinable()) {
            next.join();
            check_read();
            next_ok = !reader.eof();
        }
        data_time += config.gulp_length_samps() * tsamp;
        if(data_time >= config.duration() || !more) break;
    }
    return gulps;
}

} // namespace skyweaver_amd
