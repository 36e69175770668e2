// runtime.cpp -- error state, library scratch buffers, host-side scalings and
// the synthetic input generator.
#include "sw_common.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

namespace sw {

namespace {
thread_local std::string g_last_error;

struct ScratchSlot {
    void* ptr   = nullptr;
    size_t size = 0;
};
std::mutex g_scratch_mu;
ScratchSlot g_scratch[8];

struct ScratchReleaser {
    ~ScratchReleaser()
    {
        // Device memory is reclaimed with the process; nothing to do here.
    }
} g_releaser;
} // namespace

void set_error(const std::string& msg) { g_last_error = msg; }
const std::string& get_error() { return g_last_error; }

void* scratch(size_t bytes, int slot)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if(slot < 0 || slot >= 8) return nullptr;
    ScratchSlot& s = g_scratch[slot];
    if(s.size >= bytes && s.ptr) return s.ptr;
    if(s.ptr) {
        (void)hipDeviceSynchronize();
        (void)hipFree(s.ptr);
        s.ptr  = nullptr;
        s.size = 0;
    }
    void* p = nullptr;
    if(hipMalloc(&p, std::max<size_t>(bytes, 256)) != hipSuccess) {
        set_error("scratch: hipMalloc of " + std::to_string(bytes) + " bytes failed");
        return nullptr;
    }
    s.ptr  = p;
    s.size = bytes;
    return p;
}

// ------------------------------------------------------ synthetic generator
namespace {
inline uint64_t splitmix64(uint64_t& x)
{
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

struct Xoshiro256ss {
    uint64_t s[4];
    explicit Xoshiro256ss(uint64_t seed)
    {
        uint64_t x = seed;
        for(auto& v : s) v = splitmix64(x);
    }
    uint64_t next()
    {
        const uint64_t result = rotl(s[1] * 5, 7) * 9;
        const uint64_t t      = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return result;
    }
    // advance by 2^128 steps (the generator's published jump polynomial)
    void jump()
    {
        static const uint64_t J[] = {0x180ec6d33cfd0abaull, 0xd5a61266f0c9392cull,
                                     0xa9582618e03fc9aaull, 0x39abdc4529b1661cull};
        uint64_t s0 = 0, s1 = 0, s2 = 0, s3 = 0;
        for(uint64_t j : J)
            for(int b = 0; b < 64; ++b) {
                if(j & (1ull << b)) {
                    s0 ^= s[0];
                    s1 ^= s[1];
                    s2 ^= s[2];
                    s3 ^= s[3];
                }
                next();
            }
        s[0] = s0;
        s[1] = s1;
        s[2] = s2;
        s[3] = s3;
    }
    // uniform double in [0, 1)
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
};

inline int8_t quantise(double v)
{
    double r = std::round(v);  // half away from zero (lround semantics)
    r        = std::min(127.0, std::max(-127.0, r));
    return (int8_t)r;
}

constexpr size_t kChunk = size_t(1) << 20;  // complex values per generator stream
} // namespace

} // namespace sw

extern "C" int sw_last_error(char* buf, size_t len)
{
    const std::string& e = sw::get_error();
    if(buf && len) {
        size_t n = std::min(len - 1, e.size());
        std::memcpy(buf, e.data(), n);
        buf[n] = 0;
    }
    return (int)e.size();
}

extern "C" const char* sw_version(void) { return "skyweaver_amd 0.1 (gfx950)"; }

extern "C" int sw_update_scalings(const float* h_stats, const float* h_beamset_weights,
                                  int nbeamsets, int nchans, int npol, int nantennas,
                                  int cb_tscrunch, int ib_tscrunch, float output_level,
                                  float* h_cb_offsets, float* h_cb_scaling,
                                  float* h_ib_offsets, float* h_ib_scaling)
{
    // StatisticsCalculator::update_scalings, cpp/skyweaver/src/StatisticsCalculator.cu:132-239,
    // with outputs indexed beamset*F + f for every beamset.
    SW_CHECK_ARG(h_stats && h_beamset_weights && h_cb_offsets && h_cb_scaling && h_ib_offsets &&
                     h_ib_scaling,
                 "null buffer");
    SW_CHECK_ARG(nbeamsets > 0 && nchans > 0 && npol > 0 && nantennas > 0, "bad dimensions");
    const float weights_amp = 127.0f;
    const int pa            = npol * nantennas;
    for(int bs = 0; bs < nbeamsets; ++bs)
        for(int f = 0; f < nchans; ++f) {
            float sum = 0.0f, count = 0.0f;
            for(int p = 0; p < npol; ++p)
                for(int a = 0; a < nantennas; ++a) {
                    const float weight = h_beamset_weights[nantennas * bs + a];
                    // Statistics struct {mean, std, skew, kurtosis}: std is field 1
                    sum += weight * h_stats[4 * (f * pa + p * nantennas + a) + 1];
                    count += weight;
                }
            const float avg_std  = sum / count;
            const float eff_nant = count / (float)npol;
            const int idx        = bs * nchans + f;
            {
                const float x     = weights_amp * avg_std * std::sqrt(eff_nant);
                const float scale = (float)std::pow((double)x, 2);
                const float dof   = (float)(2 * cb_tscrunch * npol);
                h_cb_offsets[idx] = scale * dof;
                h_cb_scaling[idx] = scale * std::sqrt(2 * dof) / output_level;
            }
            {
                const float scale = (float)std::pow((double)avg_std, 2);
                const float dof   = (float)(2 * ib_tscrunch) * eff_nant * (float)npol;
                h_ib_offsets[idx] = scale * dof;
                h_ib_scaling[idx] = scale * std::sqrt(2 * dof) / output_level;
            }
        }
    return SW_OK;
}

extern "C" int sw_synth_gaussian_c8(int8_t* h_out, size_t ncomplex, uint64_t seed,
                                    double sigma, int nthreads)
{
    SW_CHECK_ARG(h_out || ncomplex == 0, "null buffer");
    const size_t nchunks = (ncomplex + sw::kChunk - 1) / sw::kChunk;
    std::vector<sw::Xoshiro256ss> states;
    states.reserve(nchunks);
    sw::Xoshiro256ss g(seed);
    for(size_t c = 0; c < nchunks; ++c) {
        states.push_back(g);
        g.jump();
    }
    const double two_pi = 6.283185307179586476925286766559;
    auto work = [&](size_t c0, size_t c1) {
        for(size_t c = c0; c < c1; ++c) {
            sw::Xoshiro256ss st = states[c];
            const size_t beg = c * sw::kChunk, end = std::min(ncomplex, beg + sw::kChunk);
            for(size_t i = beg; i < end; ++i) {
                const double u1  = 1.0 - st.uniform();  // (0, 1]
                const double u2  = st.uniform();
                const double rad = std::sqrt(-2.0 * std::log(u1));
                h_out[2 * i]     = sw::quantise(sigma * rad * std::cos(two_pi * u2));
                h_out[2 * i + 1] = sw::quantise(sigma * rad * std::sin(two_pi * u2));
            }
        }
    };
    int nt = nthreads > 0 ? nthreads : (int)std::max(1u, std::thread::hardware_concurrency());
    nt     = (int)std::min<size_t>(nt, std::max<size_t>(nchunks, 1));
    std::vector<std::thread> th;
    for(int k = 0; k < nt; ++k) {
        const size_t c0 = nchunks * k / nt, c1 = nchunks * (k + 1) / nt;
        th.emplace_back(work, c0, c1);
    }
    for(auto& t : th) t.join();
    return SW_OK;
}

extern "C" int sw_synth_phase_weights(int8_t* h_out, size_t ncomplex, uint64_t seed)
{
    SW_CHECK_ARG(h_out || ncomplex == 0, "null buffer");
    sw::Xoshiro256ss g(seed);
    const double two_pi = 6.283185307179586476925286766559;
    for(size_t i = 0; i < ncomplex; ++i) {
        const double th = two_pi * g.uniform();
        h_out[2 * i]     = (int8_t)std::lround(127.0 * std::cos(th));
        h_out[2 * i + 1] = (int8_t)std::lround(127.0 * std::sin(th));
    }
    return SW_OK;
}

// ------------------------------------------------ device memory and streams
extern "C" int sw_device_count(int* count)
{
    SW_CHECK_ARG(count, "null pointer");
    int n = 0;
    if(hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
    return SW_OK;
}

extern "C" int sw_mem_info(size_t* free_bytes, size_t* total_bytes)
{
    SW_CHECK_ARG(free_bytes && total_bytes, "null pointer");
    SW_HIP_TRY(hipMemGetInfo(free_bytes, total_bytes));
    return SW_OK;
}

extern "C" int sw_malloc(void** d_ptr, size_t bytes)
{
    SW_CHECK_ARG(d_ptr, "null pointer");
    *d_ptr = nullptr;
    if(hipMalloc(d_ptr, bytes ? bytes : 1) != hipSuccess) {
        sw::set_error("sw_malloc: hipMalloc of " + std::to_string(bytes) + " bytes failed");
        return SW_ERR_ALLOC;
    }
    return SW_OK;
}

extern "C" int sw_free(void* d_ptr)
{
    if(d_ptr) SW_HIP_TRY(hipFree(d_ptr));
    return SW_OK;
}

extern "C" int sw_host_alloc(void** h_ptr, size_t bytes)
{
    SW_CHECK_ARG(h_ptr, "null pointer");
    *h_ptr = nullptr;
    if(hipHostMalloc(h_ptr, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) {
        sw::set_error("sw_host_alloc: hipHostMalloc of " + std::to_string(bytes) + " bytes failed");
        return SW_ERR_ALLOC;
    }
    return SW_OK;
}

extern "C" int sw_host_free(void* h_ptr)
{
    if(h_ptr) SW_HIP_TRY(hipHostFree(h_ptr));
    return SW_OK;
}

extern "C" int sw_memcpy_htod(void* d_dst, const void* h_src, size_t bytes, void* stream)
{
    SW_CHECK_ARG(d_dst && h_src, "null pointer");
    SW_HIP_TRY(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, sw::as_stream(stream)));
    return SW_OK;
}

extern "C" int sw_memcpy_dtoh(void* h_dst, const void* d_src, size_t bytes, void* stream)
{
    SW_CHECK_ARG(h_dst && d_src, "null pointer");
    SW_HIP_TRY(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, sw::as_stream(stream)));
    return SW_OK;
}

namespace sw {
namespace {
// Streams of sw_memcpy_dtoh_sync: one per live thread, returned to the pool
// when the thread ends (no HIP call at thread or process exit) and reused.
std::mutex g_copy_mu;
std::vector<hipStream_t> g_copy_free;
struct CopyStream {
    hipStream_t s = nullptr;
    ~CopyStream()
    {
        if(s) {
            std::lock_guard<std::mutex> lk(g_copy_mu);
            g_copy_free.push_back(s);
        }
    }
};
thread_local CopyStream g_copy_stream;

int thread_copy_stream(hipStream_t* out)
{
    if(!g_copy_stream.s) {
        {
            std::lock_guard<std::mutex> lk(g_copy_mu);
            if(!g_copy_free.empty()) {
                g_copy_stream.s = g_copy_free.back();
                g_copy_free.pop_back();
            }
        }
        if(!g_copy_stream.s) SW_HIP_TRY(hipStreamCreateWithFlags(&g_copy_stream.s, hipStreamNonBlocking));
    }
    *out = g_copy_stream.s;
    return SW_OK;
}
} // namespace
} // namespace sw

extern "C" int sw_memcpy_dtoh_sync(void* h_dst, const void* d_src, size_t bytes)
{
    SW_CHECK_ARG(h_dst && d_src, "null pointer");
    hipStream_t s = nullptr;
    const int rc = sw::thread_copy_stream(&s);
    if(rc) return rc;
    SW_HIP_TRY(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, s));
    SW_HIP_TRY(hipStreamSynchronize(s));
    return SW_OK;
}

extern "C" int sw_memcpy_dtod(void* d_dst, const void* d_src, size_t bytes, void* stream)
{
    SW_CHECK_ARG(d_dst && d_src, "null pointer");
    SW_HIP_TRY(hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, sw::as_stream(stream)));
    return SW_OK;
}

extern "C" int sw_memset(void* d_ptr, int value, size_t bytes, void* stream)
{
    SW_CHECK_ARG(d_ptr, "null pointer");
    SW_HIP_TRY(hipMemsetAsync(d_ptr, value, bytes, sw::as_stream(stream)));
    return SW_OK;
}

extern "C" int sw_stream_create(void** stream)
{
    SW_CHECK_ARG(stream, "null pointer");
    hipStream_t s = nullptr;
    SW_HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return SW_OK;
}

extern "C" int sw_stream_destroy(void* stream)
{
    if(stream) SW_HIP_TRY(hipStreamDestroy(sw::as_stream(stream)));
    return SW_OK;
}

extern "C" int sw_stream_synchronize(void* stream)
{
    SW_HIP_TRY(hipStreamSynchronize(sw::as_stream(stream)));
    return SW_OK;
}

extern "C" int sw_event_create(void** event)
{
    SW_CHECK_ARG(event, "null pointer");
    hipEvent_t e = nullptr;
    SW_HIP_TRY(hipEventCreate(&e));
    *event = e;
    return SW_OK;
}

extern "C" int sw_event_destroy(void* event)
{
    if(event) SW_HIP_TRY(hipEventDestroy(static_cast<hipEvent_t>(event)));
    return SW_OK;
}

extern "C" int sw_event_record(void* event, void* stream)
{
    SW_CHECK_ARG(event, "null event");
    SW_HIP_TRY(hipEventRecord(static_cast<hipEvent_t>(event), sw::as_stream(stream)));
    return SW_OK;
}

extern "C" int sw_event_elapsed_ms(float* ms, void* start, void* stop)
{
    SW_CHECK_ARG(ms && start && stop, "null pointer");
    SW_HIP_TRY(hipEventSynchronize(static_cast<hipEvent_t>(stop)));
    SW_HIP_TRY(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(stop)));
    return SW_OK;
}
