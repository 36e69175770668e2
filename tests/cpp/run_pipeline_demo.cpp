// run_pipeline_demo.cpp -- the reference CLI's input path through the C++
// facade (cpp/skyweaver/src/skyweaver_cli.cu:109-240): DADA files made of the
// reference fixture's headers (4 antennas, 16 channels) and seeded Gaussian
// TAFTP voltages are read by MultiFileReader and fed to a BeamformerPipeline
// by the double-buffered run_pipeline.  The coherent beams must equal those
// of a second pipeline fed the same gulps directly, the gulp count must be
// the reference loop's (the gulp whose read reaches the end is dropped), and
// start_time must skip whole heaps.
//
// usage: run_pipeline_demo <dir with data_chunk_<i>.hdr> <scratch dir>
#include "skyweaver_amd.hpp"

#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <random>
#include <vector>

using namespace skyweaver_amd;

namespace {

struct Collect {
    std::vector<std::vector<int8_t>> blocks;
    void init(ObservationHeader const&) {}
    void operator()(BlockView const& v, std::size_t) { blocks.push_back(v.to_host()); }
    void operator()(BlockView const& v) { blocks.push_back(v.to_host()); }
};

std::vector<char> read_file(std::string const& path)
{
    std::ifstream f(path, std::ios::binary);
    return std::vector<char>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}

} // namespace

int main(int argc, char** argv)
{
    if(argc != 3) {
        std::fprintf(stderr, "usage: %s <header dir> <scratch dir>\n", argv[0]);
        return 2;
    }
    try {
        // ten files of 1000 samples (4 antennas x 16 channels x 2 pols x char2)
        const std::size_t nfiles = 10, per_file = 256000;
        std::mt19937 rng(0x5EED0011);
        std::normal_distribution<double> g(0.0, 16.0);
        std::vector<char> stream(nfiles * per_file);
        for(auto& c : stream) c = (char)std::max(-127.0, std::min(127.0, std::round(g(rng))));
        std::vector<std::string> files;
        for(std::size_t i = 0; i < nfiles; ++i) {
            const auto h = read_file(std::string(argv[1]) + "/data_chunk_" + std::to_string(i) + ".hdr");
            const std::string path = std::string(argv[2]) + "/run_" + std::to_string(i) + ".dada";
            std::ofstream f(path, std::ios::binary);
            f.write(h.data(), (std::streamsize)h.size());
            f.write(stream.data() + i * per_file, (std::streamsize)per_file);
            files.push_back(path);
        }

        PipelineConfig config;
        config.nantennas(16);
        config.nchans(16);
        config.nbeams(32);
        config.tscrunch(16);
        config.gulp_length_samps(2048);
        config.total_nchans(4096);
        config.ddplan().add_block("0");
        config.input_files(files);
        config.dada_header_size(4096);
        config.check_input_contiguity(true);

        MultiFileReader reader(config);
        ObservationHeader header = reader.get_header();
        validate_header(header, config);
        update_config(config, header);
        // the fixture's FREQ/BW are the full band's (856 MHz, "13.375"): use
        // the subband's frequencies for the beamformer
        config.centre_frequency(1284e6);
        config.bandwidth(13.375e6 * 16 / 64);

        std::vector<char> weights(16 * 32 * 16 * 2);
        for(auto& w : weights) w = (char)(rng() % 255 - 127);

        Collect cb1, ib1, st1, cb2, ib2, st2;
        BeamformerPipeline<Collect, Collect, Collect> p1(config, cb1, ib1, st1);
        p1.set_weights(weights.data());
        const std::size_t gulps = run_pipeline(p1, config, reader, header);

        // 10000 samples in gulps of 2048: reads 0..3 are whole, read 4 reaches
        // the end, so gulps 0..3 are processed
        bool ok = gulps == 4 && cb1.blocks.size() == 4;

        BeamformerPipeline<Collect, Collect, Collect> p2(config, cb2, ib2, st2);
        p2.set_weights(weights.data());
        p2.init(header, 0.0);
        const std::size_t gbytes = 2048 * 4 * 16 * 2 * 2;
        for(std::size_t k = 0; k < 4; ++k) {
            std::vector<char> gulp(stream.begin() + k * gbytes, stream.begin() + (k + 1) * gbytes);
            p2(gulp);
        }
        ok = ok && cb2.blocks.size() == 4;
        for(std::size_t k = 0; ok && k < 4; ++k) ok = cb1.blocks[k] == cb2.blocks[k] && ib1.blocks[k] == ib2.blocks[k];
        ok = ok && st1.blocks == st2.blocks;

        // start_time: 600 samples in, floored to 512 (two heaps)
        const double tsamp = (double)(header.obs_nchans / header.obs_bandwidth);
        config.start_time((float)(600.5 * tsamp));
        MultiFileReader reader3(config);
        Collect cb3, ib3, st3, cb4, ib4, st4;
        BeamformerPipeline<Collect, Collect, Collect> p3(config, cb3, ib3, st3);
        p3.set_weights(weights.data());
        const std::size_t gulps3 = run_pipeline(p3, config, reader3, header);
        BeamformerPipeline<Collect, Collect, Collect> p4(config, cb4, ib4, st4);
        p4.set_weights(weights.data());
        p4.init(header, 512 * tsamp);
        const std::size_t skip = 512 * 4 * 16 * 2 * 2;
        for(std::size_t k = 0; k < gulps3; ++k) {
            std::vector<char> gulp(stream.begin() + skip + k * gbytes, stream.begin() + skip + (k + 1) * gbytes);
            p4(gulp);
        }
        ok = ok && gulps3 == 4 && cb3.blocks.size() == 4 && cb3.blocks == cb4.blocks;

        if(!ok) {
            std::fprintf(stderr, "mismatch: gulps %zu, blocks %zu/%zu, offset run %zu\n", gulps, cb1.blocks.size(),
                         cb2.blocks.size(), gulps3);
            return 1;
        }
        std::printf("run_pipeline_demo passed (%zu gulps)\n", gulps);
        return 0;
    } catch(std::exception const& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
