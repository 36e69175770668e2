// facade_host_test.cpp -- host-only checks of the C++ facade (no GPU):
// DedispersionPlan parsing, channel frequencies, DADA header editing and the
// FileStream / MultiFileWriter file layout, the latter fed with host blocks.
#include "skyweaver_amd.hpp"

#include <cstdio>
#include <fstream>

using namespace skyweaver_amd;

#define EXPECT(c)                                                          \
    do {                                                                   \
        if(!(c)) {                                                         \
            std::fprintf(stderr, "%s:%d: EXPECT failed: %s\n", __FILE__, __LINE__, #c); \
            return 1;                                                      \
        }                                                                  \
    } while(0)

int main(int argc, char** argv)
{
    if(argc < 2) {
        std::fprintf(stderr, "usage: %s OUTPUT_DIRECTORY\n", argv[0]);
        return 2;
    }
    const std::string outdir = argv[1];
    // DedispersionPlan: the three block forms
    DedispersionPlan plan;
    plan.add_block("100:0:1000:0.9765625:1");
    plan.add_block("5:4");
    plan.add_block(7.0f, 0.0f, 1.0f, 0.25f, 2);
    EXPECT(plan.blocks().size() == 3);
    EXPECT(plan[0].incoherent_dms.size() == 1024 && plan[0].coherent_dm == 100.0f);
    EXPECT(plan[1].incoherent_dms.size() == 1 && plan[1].tscrunch == 4);
    EXPECT(plan[2].incoherent_dms.size() == 4 && plan[2].incoherent_dms[3] == 0.75f);
    EXPECT(plan.coherent_dms().size() == 3);
    bool threw = false;
    try {
        plan.add_block("1:2:3");
    } catch(std::runtime_error const&) {
        threw = true;
    }
    EXPECT(threw);

    // channel frequencies: centre - bw/2 + chbw/2 + i chbw
    PipelineConfig cfg;
    cfg.nchans(64);
    cfg.centre_frequency(1284e6);
    cfg.bandwidth(13.375e6);
    auto const& f = cfg.channel_frequencies();
    EXPECT(f.size() == 64);
    EXPECT(f[0] == 1284e6 - 13.375e6 / 2 + 13.375e6 / 64 / 2);

    // DADA header editing
    DadaHeader h("HEADER       DADA\nNBEAM        800\n");
    h.set("NBEAM", 72);
    h.set("TSAMP", 76.56, 6);
    EXPECT(h.text().find("NBEAM        72\n") != std::string::npos);
    EXPECT(h.text().find("NBEAM        800") == std::string::npos);
    EXPECT(h.text().find("TSAMP        76.560000\n") != std::string::npos);
    EXPECT(h.block(4096).size() == 4096);

    // MultiFileWriter over host blocks: rollover at max_output_filesize
    cfg.output_dir(outdir);
    cfg.max_output_filesize(3 * 1000);
    ObservationHeader oh;
    oh.utc_start = 1708082229.0L;
    oh.frequency = 1284e6L;
    std::vector<std::string> files;
    {
    MultiFileWriter w(cfg, "t");
    w.init(oh);
    std::vector<int8_t> data(1000);
    for(int i = 0; i < 1000; ++i) data[i] = (int8_t)(i % 127);
    float dms[2] = {1.5f, 2.5f};
    sw_block b{};
    b.kind = SW_BLOCK_TDB;
    b.data = data.data();
    b.on_device = 0;
    b.bytes = data.size();
    b.ndims = 3;
    b.dims[0] = 5;
    b.dims[1] = 2;
    b.dims[2] = 100;
    std::snprintf(b.order, sizeof(b.order), "TDB");
    b.reference_dm = 3.25f;
    b.tsamp = 7.6e-5;
    b.dms = dms;
    b.ndms = 2;
    for(int k = 0; k < 4; ++k) w(BlockView(b), 0);
    files = w.stream(0).files();
    }  // the writer closes its files
    EXPECT(files.size() == 2);
    std::ifstream in(files[0], std::ios::binary | std::ios::ate);
    EXPECT((std::size_t)in.tellg() == 4096 + 3000);
    std::ifstream in2(files[1], std::ios::binary | std::ios::ate);
    EXPECT((std::size_t)in2.tellg() == 4096 + 1000);
    EXPECT(files[0].find("2024-02-16-11:17:09/0/1284000000/2024-02-16-11:17:09_0_00003.250_1284000000_t_"
                         "0000000000000000.tdb") != std::string::npos);
    EXPECT(files[1].find("_0000000000003000.tdb") != std::string::npos);
    std::printf("facade host checks passed\n");
    return 0;
}
