// weights.hip -- beamforming weights from delay models, and the delay-model
// file reader.
//
// Reference: generate_weights_k and WeightsManager::weights,
// cpp/skyweaver/src/WeightsManager.cu:16-87,109-139: for channel frequency nu_f,
//   phase = ((t_now - t_model) * rate + offset) * nu_f            (FP64)
//   w = (clamp_i8(rint(cos(2 pi phase) * 127 * wt)),
//        clamp_i8(rint(-sin(2 pi phase) * 127 * wt)))
// and DelayManager (cpp/skyweaver/src/DelayManager.cu:16-250,
// cpp/skyweaver/DelayManager.cuh:40-51): a file of records
//   {uint32 nbeams, uint32 nantennas, double start_epoch, double end_epoch}
//   followed by float3 {weight, offset, rate}[nbeams][nantennas],
// a model being valid for start <= epoch <= end.  Beamsets are runs of
// contiguous beams whose per-antenna weights equal the previous beam's.
#include "sw_common.hpp"

#include <cerrno>
#include <cstring>
#include <fstream>
#include <memory>
#include <string>
#include <vector>

namespace sw {
namespace {

// one thread per (f, b, a)
__global__ void generate_weights_kernel(const float* __restrict__ delays, uint8_t* __restrict__ w,
                                        const double* __restrict__ freqs, int nant, int nbeams,
                                        int nchans, double dt)
{
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t total = (size_t)nchans * nbeams * nant;
    if(idx >= total) return;
    const int a = (int)(idx % nant);
    const size_t t = idx / nant;
    const int b = (int)(t % nbeams);
    const int f = (int)(t / nbeams);
    const float* dm = delays + 3 * ((size_t)b * nant + a);
    const double frequency = freqs[f];
    const double phase = (dt * (double)dm[2] + (double)dm[1]) * frequency;
    double s, c;
    sincos(6.283185307179586 * phase, &s, &c);
    const int re = __double2int_rn(c * 127.0 * (double)dm[0]);
    const int im = __double2int_rn(-1.0 * s * 127.0 * (double)dm[0]);
    w[2 * idx]     = (uint8_t)(int8_t)min(max(re, -128), 127);
    w[2 * idx + 1] = (uint8_t)(int8_t)min(max(im, -128), 127);
}

} // namespace
} // namespace sw

extern "C" int sw_generate_weights(const float* d_delays, void* d_weights, const double* d_chan_freqs,
                                   int nantennas, int nbeams, int nchans, double current_epoch,
                                   double delay_epoch, void* stream)
{
    SW_CHECK_ARG(d_delays && d_weights && d_chan_freqs, "null buffer");
    SW_CHECK_ARG(nantennas > 0 && nbeams > 0 && nchans > 0, "empty dimensions");
    const size_t n = (size_t)nchans * nbeams * nantennas;
    // t = (current_epoch - delay_epoch) + time_idx * tstep with one time step
    // (cpp/skyweaver/src/WeightsManager.cu:70,134-135)
    const double dt = current_epoch - delay_epoch;
    hipLaunchKernelGGL(sw::generate_weights_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       sw::as_stream(stream), d_delays, static_cast<uint8_t*>(d_weights), d_chan_freqs,
                       nantennas, nbeams, nchans, dt);
    SW_LAUNCH_CHECK();
    return SW_OK;
}

// ------------------------------------------------------------ delay models
struct sw_delay_model {
    std::ifstream in;
    int nbeams_cfg = 0, nant_cfg = 0;
    uint32_t nbeams = 0, nant = 0;  // of the current model
    uint32_t valid_nbeams = 0, valid_nant = 0;
    double start = 0, end = 0;
    std::vector<float> delays;      // [nbeams][nant][3] of the current model
    int nbeamsets = 0;
    std::vector<int> map;           // [nbeams_cfg]
    std::vector<float> bweights;    // [nbeamsets][nant_cfg]
};

namespace {

int read_next_model(sw_delay_model* m)
{
    struct {
        uint32_t nbeams, nantennas;
        double start_epoch, end_epoch;
    } hdr;
    m->in.read(reinterpret_cast<char*>(&hdr), sizeof(hdr));
    if(m->in.eof()) {
        sw::set_error("Reached end of delay model file");
        return SW_ERR_EOF;
    }
    if(!m->in) {
        sw::set_error(std::string("Error: Unable to read from delay file: ") + std::strerror(errno));
        return SW_ERR_IO;
    }
    if((int)hdr.nbeams > m->nbeams_cfg) {
        sw::set_error("Delay model contains too many beams for the configured beam count");
        return SW_ERR_INVALID;
    }
    m->nbeams = hdr.nbeams;
    m->nant   = hdr.nantennas;
    m->start  = hdr.start_epoch;
    m->end    = hdr.end_epoch;
    m->delays.resize((size_t)hdr.nbeams * hdr.nantennas * 3);
    m->in.read(reinterpret_cast<char*>(m->delays.data()), sizeof(float) * m->delays.size());
    if(m->in.eof()) {
        sw::set_error("Reached end of delay model file");
        return SW_ERR_EOF;
    }
    if(!m->in) {
        sw::set_error("Error: Unable to read from delay file");
        return SW_ERR_IO;
    }
    return SW_OK;
}

void parse_beamsets(sw_delay_model* m)
{
    // DelayManager::parse_beamsets, cpp/skyweaver/src/DelayManager.cu:162-232
    const int A = m->nant_cfg;
    m->map.assign(m->nbeams_cfg, 0);
    m->bweights.assign(A, 0.0f);
    for(uint32_t a = 0; a < m->nant; ++a) m->bweights[a] = m->delays[3 * a];
    int nsets = 1;
    for(uint32_t b = 1; b < m->nbeams; ++b) {
        bool changed = false;
        for(uint32_t a = 0; a < m->nant; ++a)
            if(m->delays[3 * ((size_t)b * m->nant + a)] != m->delays[3 * ((size_t)(b - 1) * m->nant + a)])
                changed = true;
        if(changed) {
            m->bweights.resize((size_t)(nsets + 1) * A, 0.0f);
            for(uint32_t a = 0; a < m->nant; ++a)
                m->bweights[(size_t)nsets * A + a] = m->delays[3 * ((size_t)b * m->nant + a)];
            m->map[b] = nsets;
            ++nsets;
        } else {
            m->map[b] = nsets - 1;
        }
    }
    m->nbeamsets = nsets;
}

} // namespace

extern "C" int sw_delay_model_open(const char* path, int nbeams, int nantennas, sw_delay_model** out)
{
    SW_CHECK_ARG(path && out && nbeams > 0 && nantennas > 0, "bad argument");
    *out = nullptr;
    auto m = std::make_unique<sw_delay_model>();
    m->nbeams_cfg = nbeams;
    m->nant_cfg = nantennas;
    m->in.open(path, std::ios::in | std::ios::binary);
    if(!m->in.is_open()) {
        sw::set_error(std::string("Error: Unable to open file delay file due to error: ") + std::strerror(errno));
        return SW_ERR_IO;
    }
    int rc = read_next_model(m.get());
    if(rc) return rc;
    if((int)m->nant > nantennas) {  // would overrun the beamset weight rows
        sw::set_error("Delay model contains too many antennas");
        return SW_ERR_INVALID;
    }
    parse_beamsets(m.get());
    m->valid_nbeams = m->nbeams;
    m->valid_nant = m->nant;
    *out = m.release();
    return SW_OK;
}

extern "C" int sw_delay_model_close(sw_delay_model* m)
{
    delete m;
    return SW_OK;
}

extern "C" int sw_delay_model_beamsets(const sw_delay_model* m, int* nbeamsets, int* h_map,
                                       float* h_weights, int max_beamsets)
{
    SW_CHECK_ARG(m && nbeamsets, "null argument");
    *nbeamsets = m->nbeamsets;
    if(h_map) std::memcpy(h_map, m->map.data(), sizeof(int) * m->map.size());
    if(h_weights) {
        SW_CHECK_ARG(max_beamsets >= m->nbeamsets, "beamset buffer too small");
        std::memcpy(h_weights, m->bweights.data(), sizeof(float) * m->bweights.size());
    }
    return SW_OK;
}

extern "C" int sw_delay_model_get(sw_delay_model* m, double epoch, float* h_delays, double* model_epoch)
{
    // DelayManager::delays, cpp/skyweaver/src/DelayManager.cu:56-99: the result is padded
    // to [nbeams_cfg][nantennas_cfg] with zeros.
    SW_CHECK_ARG(m && h_delays, "null argument");
    if(epoch < m->start) {
        sw::set_error("No valid delay solution for unix epoch " + std::to_string(epoch));
        return SW_ERR_DELAY_EPOCH;
    }
    while(!(epoch >= m->start && epoch <= m->end)) {
        int rc = read_next_model(m);
        if(rc) return rc;
        if(m->nbeams != m->valid_nbeams || m->nant != m->valid_nant) {
            sw::set_error("Variable delay model parameters are unsupported");
            return SW_ERR_INVALID;
        }
    }
    std::memset(h_delays, 0, sizeof(float) * 3 * (size_t)m->nbeams_cfg * m->nant_cfg);
    for(uint32_t b = 0; b < m->nbeams; ++b)
        std::memcpy(h_delays + 3 * ((size_t)b * m->nant_cfg), m->delays.data() + 3 * ((size_t)b * m->nant),
                    sizeof(float) * 3 * m->nant);
    if(model_epoch) *model_epoch = m->start;
    return SW_OK;
}
