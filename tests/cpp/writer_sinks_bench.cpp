// writer_sinks_bench.cpp -- gulp time of the BeamformerPipeline with the
// reference's output path attached: MultiFileWriter handlers writing the
// coherent beams (TFB) and the incoherent beam (BTF) as DADA files
// (cpp/skyweaver/detail/MultiFileWriter.cu:103-249), called synchronously
// within each gulp as the reference does, and called by the library's sink
// thread while the next gulp is processed (async_handlers).  Also the gulp
// time with no handlers, for comparison.
//
// Workload: 64 antennas x 64 channels x 65536 samples (device-resident
// seeded Gaussian voltages), 800 beams, coherent DM 50 (FFT 16384),
// tscrunch 16, Stokes I: 210 MB of TFB and 26 KB of BTF per gulp.
//
// usage: writer_sinks_bench <scratch dir> [gulps]
// prints one JSON line with milliseconds per gulp (wall clock, every
// handler call included: the run ends with synchronize()).
#include "skyweaver_amd.hpp"

#include <chrono>
#include <cstdio>
#include <filesystem>
#include <string>
#include <vector>

using namespace skyweaver_amd;

namespace {

constexpr int A = 64, F = 64, T = 65536, B = 800, TS = 16;

PipelineConfig make_config(std::string const& outdir, bool async)
{
    PipelineConfig c;
    c.nantennas(A);
    c.nchans(F);
    c.nbeams(B);
    c.tscrunch(TS);
    c.gulp_length_samps(T);
    c.centre_frequency(1284e6);
    c.bandwidth(208984.375 * F);
    c.ddplan().add_block(50.0f);
    c.fft_lengths({16384});
    c.output_dir(outdir);
    c.async_handlers(async);
    return c;
}

ObservationHeader make_header()
{
    ObservationHeader h;
    h.nchans = F;
    h.npol = 2;
    h.nbits = 8;
    h.nantennas = A;
    h.obs_nchans = 4096;
    h.bandwidth = 208984.375 * F;
    h.obs_bandwidth = 856e6;
    h.frequency = 1284e6;
    h.obs_frequency = 1284e6;
    h.tsamp = 4096.0L / 856e6L;
    h.utc_start = 1708082229.0L;
    return h;
}

template <typename CBH, typename IBH>
double run(PipelineConfig const& c, CBH& cb, IBH& ib, void const* d_in, std::size_t bytes,
           std::vector<int8_t> const& w, int ngulps)
{
    NullHandler st;
    BeamformerPipeline<CBH, IBH, NullHandler> p(c, cb, ib, st);
    p.set_weights(w.data());
    p.init(make_header());
    p.process_device(d_in, bytes);  // warm-up (first-gulp scalings)
    p.synchronize();
    const auto t0 = std::chrono::steady_clock::now();
    for(int g = 0; g < ngulps; ++g) p.process_device(d_in, bytes);
    p.synchronize();
    const auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double, std::milli>(t1 - t0).count() / ngulps;
}

} // namespace

int main(int argc, char** argv)
{
    if(argc < 2) {
        std::fprintf(stderr, "usage: %s <scratch dir> [gulps]\n", argv[0]);
        return 2;
    }
    const std::string dir = argv[1];
    const int ngulps = argc > 2 ? std::atoi(argv[2]) : 6;
    try {
        const std::size_t bytes = (std::size_t)T * A * F * 2 * 2;
        std::vector<int8_t> h_in(bytes), w((std::size_t)F * B * A * 2);
        check(sw_synth_gaussian_c8(h_in.data(), bytes / 2, 0x5EED0002, 16.0, 16));
        check(sw_synth_phase_weights(w.data(), w.size() / 2, 0x5EED1002));
        void* d_in = nullptr;
        check(sw_malloc(&d_in, bytes));
        check(sw_memcpy_htod(d_in, h_in.data(), bytes, nullptr));
        check(sw_stream_synchronize(nullptr));

        double ms[3];
        {
            PipelineConfig c = make_config(dir, false);
            NullHandler cb, ib;
            ms[0] = run(c, cb, ib, d_in, bytes, w, ngulps);
        }
        const char* modes[2] = {"sync", "async"};
        for(int m = 0; m < 2; ++m) {
            const std::string out = dir + "/writer_sinks_" + modes[m];
            std::filesystem::create_directories(out);
            PipelineConfig c = make_config(out, m == 1);
            MultiFileWriter cb(c, "cb"), ib(c, "ib");
            ms[1 + m] = run(c, cb, ib, d_in, bytes, w, ngulps);
            std::filesystem::remove_all(out);
        }
        sw_free(d_in);
        std::printf("{\"workload\": \"64 antennas x 64 channels x 65536 samples, 800 beams, DM 50 (FFT 16384), "
                    "tscrunch 16, Stokes I; MultiFileWriter for the coherent and incoherent beams\", "
                    "\"gulps\": %d, \"tfb_bytes_per_gulp\": %zu, \"ms_per_gulp_no_handlers\": %.3f, "
                    "\"ms_per_gulp_writers_sync\": %.3f, \"ms_per_gulp_writers_async\": %.3f}\n",
                    ngulps, (std::size_t)(T / TS) * F * B, ms[0], ms[1], ms[2]);
    } catch(std::exception const& e) {
        std::fprintf(stderr, "writer_sinks_bench: %s\n", e.what());
        return 1;
    }
    return 0;
}
