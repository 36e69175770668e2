// facade_demo.cpp -- drives the C++ facade the way the reference CLI drives
// its pipeline (cpp/skyweaver/src/skyweaver_cli.cu:109-240): PipelineConfig +
// DedispersionPlan + delay-model file, a BeamformerPipeline whose coherent
// beams go both to a TFB MultiFileWriter and to an
// IncoherentDedispersionPipeline writing TDB files, an IB writer and a
// statistics writer.  Input gulps are seeded Gaussian voltages.
//
// usage: facade_demo <output_dir> <delay_file> [ngulps]
#include "skyweaver_amd.hpp"

#include <cstdio>
#include <vector>

using namespace skyweaver_amd;

// CB handler that writes the TFB stream and forwards it to the ID pipeline
template <typename A, typename B>
struct Tee {
    A& a;
    B& b;
    template <typename H>
    void init(H const& h)
    {
        a.init(h);
        b.init(h);
    }
    void operator()(BlockView const& v, std::size_t i)
    {
        a(v, i);
        b(v, i);
    }
};

int main(int argc, char** argv)
{
    if(argc < 3) {
        std::fprintf(stderr, "usage: %s <output_dir> <delay_file> [ngulps]\n", argv[0]);
        return 2;
    }
    const int ngulps = argc > 3 ? std::atoi(argv[3]) : 3;
    try {
        PipelineConfig config;
        config.nantennas(64);
        config.nchans(8);
        config.nbeams(72);
        config.tscrunch(16);
        config.gulp_length_samps(4096);
        config.centre_frequency(1284e6);
        config.bandwidth(13.375e6 * 8 / 64);
        config.total_nchans(4096);
        config.output_dir(argv[1]);
        config.output_file_prefix("demo");
        config.delay_file(argv[2]);
        config.ddplan().add_block("0:0:40:2.5:2");
        config.enable_incoherent_dedispersion(true);
        config.fft_lengths({1024});

        MultiFileWriter tfb_writer(config, "cb");
        MultiFileWriter tdb_writer(config, "tdb");
        MultiFileWriter ib_writer(config, "ib");
        MultiFileWriter stats_writer(config, "stats");
        IncoherentDedispersionPipeline<MultiFileWriter> id_pipeline(config, tdb_writer, 300);
        Tee<MultiFileWriter, IncoherentDedispersionPipeline<MultiFileWriter>> cb_handler{tfb_writer, id_pipeline};
        struct StatsHandler {
            MultiFileWriter& w;
            void init(ObservationHeader const& h) { w.init(h); }
            void operator()(BlockView const& v) { w(v, 0); }
        } stats_handler{stats_writer};

        BeamformerPipeline<decltype(cb_handler), MultiFileWriter, StatsHandler> pipeline(config, cb_handler,
                                                                                         ib_writer, stats_handler);
        ObservationHeader header;
        header.nantennas = 57;
        header.nchans = config.nchans();
        header.utc_start = 1708082300.0L;
        header.sample_clock = 1712e6L;
        header.frequency = config.centre_frequency();
        header.bandwidth = config.bandwidth();
        header.obs_nchans = 4096;
        header.obs_frequency = 1284e6L;
        header.obs_bandwidth = 856e6L;
        header.chan0_idx = 2688;
        pipeline.init(header);

        const std::size_t ncomplex = (std::size_t)config.gulp_length_samps() * header.nantennas * config.nchans() * 2;
        std::vector<int8_t> gulp(2 * ncomplex);
        for(int g = 0; g < ngulps; ++g) {
            check(sw_synth_gaussian_c8(gulp.data(), ncomplex, 0x5EED0200ull + g, 16.0, 4));
            pipeline(gulp);
        }
        std::printf("ok %d gulps\n", ngulps);
    } catch(InvalidDelayEpoch const& e) {
        std::fprintf(stderr, "InvalidDelayEpoch: %s\n", e.what());
        return 3;
    } catch(std::exception const& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
