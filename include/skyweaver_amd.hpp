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
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
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

// Message of an exception a handler threw inside a library callback; the
// callback returns an error code through the C ABI and check() rethrows it.
inline std::string& handler_error()
{
    static thread_local std::string msg;
    return msg;
}

// Convert a C ABI status into the reference's exception types.
inline void check(int rc, double epoch = 0.0)
{
    if(rc == SW_OK) return;
    if(!handler_error().empty()) {
        std::string m = "handler failed: " + handler_error();
        handler_error().clear();
        throw std::runtime_error(m);
    }
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
};

// ------------------------------------------------------- ObservationHeader
struct ObservationHeader {
    std::size_t nantennas = 0;
    std::size_t nchans = 0;
    std::size_t obs_nchans = 0;
    std::size_t chan0_idx = 0;
    long double sample_clock = 1712e6;
    long double utc_start = 0;
    long double frequency = 0;
    long double bandwidth = 0;
    long double obs_frequency = 0;
    long double obs_bandwidth = 0;
};

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
            check(sw_memcpy_dtoh(dst, _b.data, _b.bytes, nullptr));
            check(sw_stream_synchronize(nullptr));
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
        handler_error() = e.what();
        return 1;
    } catch(...) {
        handler_error() = "unknown exception";
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
        handler_error() = e.what();
        return 1;
    } catch(...) {
        handler_error() = "unknown exception";
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
    bool operator()(BlockView const& block, std::size_t stream_idx)
    {
        auto it = _streams.find(stream_idx);
        if(it == _streams.end()) it = _streams.emplace(stream_idx, create_stream(block, stream_idx)).first;
        _buf.resize(block.size());
        block.copy_to_host(_buf.data());
        it->second->write(reinterpret_cast<char const*>(_buf.data()), _buf.size());
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
};

// A handler that accepts and discards blocks (for unused outputs).
struct NullHandler {
    template <typename H>
    void init(H const&) {}
    void operator()(BlockView const&, std::size_t = 0) {}
};

} // namespace skyweaver_amd
