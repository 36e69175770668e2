// idpipe.cpp -- incoherent dedispersion pipeline: one aggregation ring and
// one incoherent dedisperser per dedispersion-plan block, fed with TFB powers
// and emitting TDB blocks to a sink.
//
// Reference: IncoherentDedispersionPipeline,
// cpp/skyweaver/detail/IncoherentDedispersionPipeline.cu:6-113, and
// AggregationBuffer, cpp/skyweaver/detail/AggregationBuffer.cu:23-116:
//   dispatch = max(10 * max_delay, 8192) rounded down to a multiple of tscrunch
//   ring     = (dispatch + max_delay) slots of nchans * nbeams values
//   push     = copy whole slots; when the ring is full, dedisperse it, hand the
//              TDB block to the handler and move the last max_delay slots to
//              the front.
// Unlike the reference (host vectors, OpenMP), the ring and the TDB block
// live in device memory and the dedispersion runs on the GPU.  The ring is
// held channel-major ([F][slots][B], "FTB"), so that the dedisperser reads
// each channel's rows from one contiguous region.  The TDB
// block's utc_offset is that of its first output sample (ring start plus
// max_delay samples); the reference stamps it from the most recently pushed
// block instead.
#include "sw_common.hpp"

#include <algorithm>
#include <cstdio>
#include <memory>
#include <vector>

struct sw_idpipe {
    struct Block {
        sw_idisp* plan = nullptr;
        float cdm = 0;
        int tscrunch = 1;
        std::vector<float> dms;
        int max_delay = 0;
        size_t dispatch = 0, slots = 0, filled = 0;
        uint8_t* ring = nullptr;
        uint8_t* tdb = nullptr;
        size_t tdb_bytes = 0;
        double ring_utc = 0;  // utc_offset of ring slot 0
        double tsamp = 0;
        size_t ndispatched = 0;
    };
    int nchans = 0, nbeams = 0, nvec = 1;
    size_t slot_bytes = 0;
    std::vector<Block> blocks;
    sw_sink_fn sink = nullptr;
    void* user = nullptr;
};

namespace {

// dst[f][d0 + i][w] = src[f][s0 + i][w] for f < F, i < k, w < nw (32-bit
// words), with independent channel and row strides
__global__ void __launch_bounds__(256)
rows_copy_kernel(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst, int F, long k, int nw,
                 long s_f, long s_t, long s0, long d_f, long d_t, long d0)
{
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t total = (size_t)F * k * nw;
    if(idx >= total) return;
    const int w = (int)(idx % nw);
    const size_t r = idx / nw;
    const long i = (long)(r % k);
    const int f = (int)(r / k);
    dst[f * d_f + (d0 + i) * d_t + w] = src[f * s_f + (s0 + i) * s_t + w];
}

int rows_copy(const void* src, void* dst, int F, long k, int nw, long s_f, long s_t, long s0, long d_f, long d_t,
              long d0, hipStream_t st)
{
    const size_t total = (size_t)F * k * nw;
    if(total == 0) return SW_OK;
    hipLaunchKernelGGL(rows_copy_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st,
                       static_cast<const uint32_t*>(src), static_cast<uint32_t*>(dst), F, k, nw, s_f, s_t, s0, d_f,
                       d_t, d0);
    SW_LAUNCH_CHECK();
    return SW_OK;
}

void release(sw_idpipe* p)
{
    for(auto& b: p->blocks) {
        sw_idisp_destroy(b.plan);
        (void)hipFree(b.ring);
        (void)hipFree(b.tdb);
    }
    p->blocks.clear();
}

int dispatch(sw_idpipe* p, int bi, hipStream_t st)
{
    auto& b = p->blocks[bi];
    int rc = sw_idisp_dedisperse_ftb(b.plan, b.ring, b.slots, p->nbeams, p->nvec, b.tdb, st);
    if(rc) return rc;
    SW_HIP_TRY(hipStreamSynchronize(st));
    const size_t nout = b.dispatch / b.tscrunch;
    if(p->sink) {
        sw_block blk{};
        blk.kind = SW_BLOCK_TDB;
        blk.data = b.tdb;
        blk.on_device = 1;
        blk.bytes = nout * b.dms.size() * p->nbeams * p->nvec;
        blk.ndims = 3;
        blk.dims[0] = nout;
        blk.dims[1] = b.dms.size();
        blk.dims[2] = p->nbeams;
        std::snprintf(blk.order, sizeof(blk.order), "TDB");
        blk.dm_idx = bi;
        blk.reference_dm = b.cdm;
        blk.tsamp = b.tsamp * b.tscrunch;
        blk.utc_offset = b.ring_utc + b.max_delay * b.tsamp;
        blk.dms = b.dms.data();
        blk.ndms = (int)b.dms.size();
        rc = p->sink(p->user, &blk);
        if(rc) {
            sw::set_error("incoherent dedispersion sink returned an error");
            return rc;
        }
    }
    ++b.ndispatched;
    // keep the last max_delay slots (AggregationBuffer::dispatch): slots
    // [dispatch, dispatch + max_delay) move to [0, max_delay), a left shift
    // by dispatch, done as ordered launches of at most dispatch rows each,
    // so that no launch reads a row another of its threads writes
    b.filled = 0;
    if(b.max_delay > 0) {
        const int nw = (int)(p->slot_bytes / p->nchans / 4);
        const long chan = (long)b.slots * nw;
        for(long r0 = 0; r0 < b.max_delay; r0 += (long)b.dispatch) {
            const long k = std::min<long>((long)b.dispatch, b.max_delay - r0);
            rc = rows_copy(b.ring, b.ring, p->nchans, k, nw, chan, nw, (long)b.dispatch + r0, chan, nw, r0, st);
            if(rc) return rc;
        }
        b.filled = b.max_delay;
    }
    b.ring_utc += b.dispatch * b.tsamp;
    return SW_OK;
}

} // namespace

extern "C" int sw_idpipe_create(const sw_idpipe_config* cfg, sw_idpipe** out)
{
    SW_CHECK_ARG(cfg && out, "null argument");
    *out = nullptr;
    SW_CHECK_ARG(cfg->nblocks > 0 && cfg->coherent_dms && cfg->tscrunch && cfg->dms && cfg->ndms,
                 "empty dedispersion plan");
    SW_CHECK_ARG(cfg->nbeams > 0 && (cfg->nvec == 1 || cfg->nvec == 4), "bad beam layout");
    SW_CHECK_ARG((cfg->nbeams * cfg->nvec) % 4 == 0, "nbeams * nvec must be a multiple of 4");
    auto p = std::make_unique<sw_idpipe>();
    p->nchans = cfg->nchans;
    p->nbeams = cfg->nbeams;
    p->nvec = cfg->nvec;
    p->slot_bytes = (size_t)cfg->nchans * cfg->nbeams * cfg->nvec;
    p->blocks.resize(cfg->nblocks);
    for(int i = 0; i < cfg->nblocks; ++i) {
        auto& b = p->blocks[i];
        b.cdm = cfg->coherent_dms[i];
        b.tscrunch = cfg->tscrunch[i];
        SW_CHECK_ARG(cfg->ndms[i] > 0 && cfg->dms[i], "block without DMs");
        b.dms.assign(cfg->dms[i], cfg->dms[i] + cfg->ndms[i]);
        sw_idisp_config ic{};
        ic.nchans = cfg->nchans;
        ic.centre_frequency = cfg->centre_frequency;
        ic.bandwidth = cfg->bandwidth;
        ic.chan_freqs = cfg->chan_freqs;
        ic.cb_tscrunch = cfg->cb_tscrunch;
        ic.dms = b.dms.data();
        ic.ndms = (int)b.dms.size();
        ic.tscrunch = b.tscrunch;
        int rc = sw_idisp_create(&ic, &b.plan);
        if(rc) {
            release(p.get());
            return rc;
        }
        sw_idisp_info(b.plan, &b.max_delay, nullptr, nullptr);
        size_t dispatch = cfg->dispatch_size > 0 ? (size_t)cfg->dispatch_size
                                                 : (size_t)std::max(b.max_delay * 10, 8192);
        dispatch -= dispatch % b.tscrunch;
        if(dispatch == 0) {
            release(p.get());
            sw::set_error("dispatch size must be at least tscrunch");
            return SW_ERR_INVALID;
        }
        b.dispatch = dispatch;
        b.slots = dispatch + b.max_delay;
        b.tdb_bytes = dispatch / b.tscrunch * b.dms.size() * p->slot_bytes / cfg->nchans;
        if(hipMalloc(&b.ring, b.slots * p->slot_bytes) != hipSuccess ||
           hipMalloc(&b.tdb, b.tdb_bytes) != hipSuccess) {
            release(p.get());
            sw::set_error("out of device memory for the incoherent dedispersion buffers");
            return SW_ERR_ALLOC;
        }
    }
    *out = p.release();
    return SW_OK;
}

extern "C" int sw_idpipe_destroy(sw_idpipe* p)
{
    if(p) {
        release(p);
        delete p;
    }
    return SW_OK;
}

extern "C" int sw_idpipe_set_sink(sw_idpipe* p, sw_sink_fn sink, void* user)
{
    SW_CHECK_ARG(p, "null pipeline");
    p->sink = sink;
    p->user = user;
    return SW_OK;
}

extern "C" int sw_idpipe_info(const sw_idpipe* p, int block, int* max_delay, size_t* dispatch_size,
                              size_t* filled)
{
    SW_CHECK_ARG(p && block >= 0 && block < (int)p->blocks.size(), "bad block index");
    const auto& b = p->blocks[block];
    if(max_delay) *max_delay = b.max_delay;
    if(dispatch_size) *dispatch_size = b.dispatch;
    if(filled) *filled = b.filled;
    return SW_OK;
}

extern "C" int sw_idpipe_push(sw_idpipe* p, int block, const void* tfb, size_t nsamples, int on_device,
                              double utc_offset, double tsamp, void* stream)
{
    // IncoherentDedispersionPipeline::operator() + AggregationBuffer::push_back
    SW_CHECK_ARG(p && block >= 0 && block < (int)p->blocks.size(), "bad block index");
    SW_CHECK_ARG(tfb || nsamples == 0, "null input");
    auto& b = p->blocks[block];
    hipStream_t st = sw::as_stream(stream);
    b.tsamp = tsamp;
    if(b.filled == 0 && b.ndispatched == 0) b.ring_utc = utc_offset;
    const auto* src = static_cast<const uint8_t*>(tfb);
    const int F = p->nchans;
    const size_t row = p->slot_bytes / F;  // bytes of one (t, f) row
    const int nw = (int)(row / 4);
    size_t done = 0;
    while(done < nsamples) {
        const size_t k = std::min(nsamples - done, b.slots - b.filled);
        // TFB [t][f][b] -> ring [f][slot][b]
        if(on_device) {
            int rc = rows_copy(src + done * p->slot_bytes, b.ring, F, (long)k, nw, nw, (long)F * nw, 0,
                               (long)b.slots * nw, nw, (long)b.filled, st);
            if(rc) return rc;
        } else {
            for(int f = 0; f < F; ++f)
                SW_HIP_TRY(hipMemcpy2DAsync(b.ring + ((size_t)f * b.slots + b.filled) * row, row,
                                            src + done * p->slot_bytes + f * row, p->slot_bytes, row, k,
                                            hipMemcpyHostToDevice, st));
        }
        b.filled += k;
        done += k;
        if(b.filled == b.slots) {
            int rc = dispatch(p, block, st);
            if(rc) return rc;
        }
    }
    if(!on_device) SW_HIP_TRY(hipStreamSynchronize(st));  // the host buffer may be reused on return
    return SW_OK;
}
