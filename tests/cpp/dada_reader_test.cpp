// Host tests of the C++ facade's DADA input: read_dada_header on the
// reference's MultiFileReader fixture headers and MultiFileReader over ten
// files made of those headers and known payloads (the cases of
// cpp/skyweaver/test/src/MultiFileReaderTester.cu: multi-file read equals
// the concatenated stream, EOF at the end, a seek past it throws, zero-size
// read, missing file, contiguity).
//
// usage: dada_reader_test <dir with data_chunk_<i>.hdr> <scratch dir>
#include "skyweaver_amd.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

using namespace skyweaver_amd;

namespace {

int failures = 0;
#define EXPECT(cond)                                                                  \
    do {                                                                              \
        if(!(cond)) {                                                                 \
            std::cerr << __FILE__ << ":" << __LINE__ << ": expected " #cond "\n"; \
            ++failures;                                                               \
        }                                                                             \
    } while(0)

template <typename F>
bool throws(F f)
{
    try {
        f();
    } catch(std::runtime_error const&) {
        return true;
    }
    return false;
}

constexpr std::size_t kHeader = 4096, kData = 256000, kFiles = 10;

// payload byte at stream offset x
unsigned char payload(std::size_t x)
{
    return (unsigned char)(((x * 2654435761u) >> 13) & 0xff);
}

std::vector<char> read_file(std::string const& path)
{
    std::ifstream f(path, std::ios::binary);
    return std::vector<char>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}

} // namespace

int main(int argc, char** argv)
{
    if(argc != 3) {
        std::cerr << "usage: dada_reader_test <header dir> <scratch dir>\n";
        return 2;
    }
    const std::string hdir = argv[1], tmp = argv[2];

    // ---- header parsing (chunk 0 and 1 of the fixture)
    const auto h0 = read_file(hdir + "/data_chunk_0.hdr"), h1 = read_file(hdir + "/data_chunk_1.hdr");
    EXPECT(h0.size() == kHeader && h1.size() == kHeader);
    ObservationHeader a, b;
    read_dada_header(h0.data(), h0.size(), a);
    read_dada_header(h1.data(), h1.size(), b);
    EXPECT(a.nantennas == 4 && a.nchans == 16 && a.npol == 2 && a.nbits == 8 && a.obs_nchans == 4096);
    EXPECT(a.frequency == 856000000.0L && a.bandwidth == 13.375L && a.obs_bandwidth == 856000000.0L);
    EXPECT(a.obs_frequency == 1284000000.0L && a.sample_clock == 1712000000.0L && a.sync_time == 1708039531.0L);
    EXPECT(std::fabs((double)(a.utc_start - 1708082229.000020336L)) < 1e-6);
    EXPECT(std::fabs((double)(a.tsamp - 4.7850467290L)) < 1e-12);
    EXPECT(a.sample_clock_start == 73098976034816ull && a.chan0_idx == 2688);
    EXPECT(a.source_name == "J1644-4559" && a.telescope == "MeerKAT" && a.instrument == "CBF-Feng");
    EXPECT(a.ra == "16:44:49.27" && a.dec == "-45:59:09.7");
    EXPECT(a.obs_offset == 0 && b.obs_offset == kData);
    {
        std::string bad(h0.begin(), h0.end());
        const auto p = bad.find("\nNANT");
        bad.replace(p + 1, 4, "NXNT");  // key missing
        ObservationHeader c;
        EXPECT(throws([&] { read_dada_header(bad.data(), bad.size(), c); }));
    }

    // ---- validate_header / update_config
    PipelineConfig config;
    config.nchans(16);
    config.nantennas(16);
    EXPECT(!throws([&] { validate_header(a, config); }));
    config.nchans(64);
    EXPECT(throws([&] { validate_header(a, config); }));
    config.nchans(16);
    config.nantennas(2);
    EXPECT(throws([&] { validate_header(a, config); }));
    config.nantennas(16);
    update_config(config, a);
    EXPECT(config.centre_frequency() == 856000000.0 && config.bandwidth() == 13.375);

    // ---- ten files: fixture header + known payload
    std::vector<std::string> files;
    for(std::size_t i = 0; i < kFiles; ++i) {
        const auto h = read_file(hdir + "/data_chunk_" + std::to_string(i) + ".hdr");
        EXPECT(h.size() == kHeader);
        const std::string path = tmp + "/data_chunk_" + std::to_string(i) + ".dada";
        std::ofstream f(path, std::ios::binary);
        f.write(h.data(), (std::streamsize)h.size());
        std::vector<char> d(kData);
        for(std::size_t k = 0; k < kData; ++k) d[k] = (char)payload(i * kData + k);
        f.write(d.data(), (std::streamsize)d.size());
        files.push_back(path);
    }
    config.input_files(files);
    config.dada_header_size(kHeader);
    config.check_input_contiguity(true);
    {
        MultiFileReader r(config);
        EXPECT(r.get_total_size() == kFiles * kData);
        EXPECT(r.get_header().nantennas == 4 && r.get_header().obs_offset == 0);
        const std::size_t per_sample = 4 * 16 * 2 * 8 * 2 / 8;
        EXPECT(r.get_total_size() / per_sample == 10000);
        // the reference test: skip 4970 samples' worth of char2, read 4970 samples (crosses files)
        const std::size_t char2s = 4 * 16 * 2, gulp = 4970;
        r.seekg((long)(char2s * gulp), std::ios::beg);
        std::vector<char> buf(char2s * gulp * 2);
        EXPECT(r.read(buf.data(), (std::streamsize)buf.size()) == (std::streamsize)buf.size());
        bool same = true;
        for(std::size_t k = 0; k < buf.size(); ++k) same = same && (unsigned char)buf[k] == payload(char2s * gulp + k);
        EXPECT(same);
        EXPECT(r.tellg() == char2s * gulp + buf.size());
        // relative and end-relative seeks
        r.seekg(-1000, std::ios::cur);
        char c = 0;
        EXPECT(r.read(&c, 1) == 1 && (unsigned char)c == payload(char2s * gulp + buf.size() - 1000));
        r.seekg(-1, std::ios::end);
        EXPECT(r.read(&c, 1) == 1 && (unsigned char)c == payload(kFiles * kData - 1) && r.eof());
        // EOF at the end; past it throws
        r.seekg((long)r.get_total_size(), std::ios::beg);
        EXPECT(r.eof());
        EXPECT(throws([&] { r.seekg((long)r.get_total_size() + 1, std::ios::beg); }));
        // zero-size read
        r.seekg(0, std::ios::beg);
        EXPECT(r.read(buf.data(), 0) == 0 && !r.eof());
        // sequential gulps reproduce the stream and end at EOF
        std::size_t pos = 0;
        bool seq = true;
        std::vector<char> g(300000);
        while(!r.eof()) {
            const auto n = r.read(g.data(), (std::streamsize)g.size());
            for(std::streamsize k = 0; k < n; ++k) seq = seq && (unsigned char)g[k] == payload(pos + k);
            pos += (std::size_t)n;
        }
        EXPECT(seq && pos == kFiles * kData);
    }
    // missing file
    {
        PipelineConfig c2 = config;
        c2.input_files({tmp + "/non_existent_file.dada"});
        EXPECT(throws([&] { MultiFileReader r(c2); }));
    }
    // contiguity: the same files out of order fail, and pass unchecked
    {
        PipelineConfig c3 = config;
        auto f = files;
        std::swap(f[3], f[4]);
        c3.input_files(f);
        EXPECT(throws([&] { MultiFileReader r(c3); }));
        c3.check_input_contiguity(false);
        EXPECT(!throws([&] { MultiFileReader r(c3); }));
    }

    if(failures) {
        std::cerr << failures << " failure(s)\n";
        return 1;
    }
    std::cout << "dada_reader_test passed\n";
    return 0;
}
