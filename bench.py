#!/usr/bin/env python3
"""Benchmark of the skyweaver BeamformerPipeline hot path on one MI355X.

Workload (one gulp = one step): 64 antennas x 64 channels (209 kHz around
1284 MHz) x 2 pols x 65536 samples of seeded Gaussian complex int8 (sigma
16, TAFTP order, resident in HBM), 800 random-phase int8 beams, one coherent
DM of 50 pc/cm^3 (FFT length 16384, overlap 10 = intra-channel smear),
tscrunch 16, Stokes I, with the incoherent beam and per-channel statistics.
Stages per gulp: transpose fused with the statistics' power sums, coherent
dedispersion, coherent beams with the incoherent beam computed in the same
kernel.

Two forms run on the same GPU and data:
  chosen: int8 MFMA coherent beamformer, DM-batched dedispersion (the value)
  plain:  VALU v_dot4 beamformer, one DM per dedispersion launch
and the CPU oracle (oracle/, -O3, OpenMP/threads) runs a bounded sample of
the same gulp on the host cores.

Prints one JSON line.  metric: beam-samples/s (beams x channels x input
samples per second, summed over coherent DMs); x-real-time = gulp duration
(65536 x 4096/856 MHz) / wall time per gulp.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# workload: 64 antennas / 64 channels / 65536 samples / 800 beams / DM 50
A, F, T, B, TS = 64, 64, 65536, 800, 16
DMS, FFT_LEN = [50.0], [16384]
CFREQ, BW = 1284e6, 13.375e6
TSAMP = 4096 / 856e6
SEED_V, SEED_W = 0x5EED0002, 0x5EED1002
INT8_PEAK_TOPS = 5000.0     # MI355X dense int8 MFMA
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3    # MI355X FP32 (packed); the unpacked VALU rate is half of it
VALU_ADD16_PEAK = 78.6e12   # 16-bit lane adds/s: 256 CUs x 64 lanes x 2 lanes per 32-bit add x 2.4 GHz
# HBM bytes per launch of the two large kernels at this workload, from
# rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE calibrated on a 1 GiB copy kernel
# (profiles/hbm_traffic_pipeline_64ant_64ch_65536samp_800beam_dm50.md)
TRAFFIC_BYTES = {"coherent_beam": 1.314e9, "dedispersion": 2.218e9}
INT8_LDS_BYTES_PER_CLK = 128   # LDS bytes per clock per CU
NUM_CU, PEAK_CLK_GHZ = 256, 2.4

# incoherent dedispersion workload: 256 channels (4 x 64 over 53.5 MHz at
# 1284 MHz), 512 beams, CB tscrunch 8, 1024 DM trials 0:1000:0.9765625, one
# dispatch of the aggregation buffer (max(10 max_delay, 8192) samples)
ID_F, ID_B, ID_CBTS, ID_BW = 256, 512, 8, 53.5e6
ID_DMS = np.arange(1024, dtype=np.float32) * np.float32(0.9765625)

# the 4x target workload: the same gulp with 8 coherent DMs of uneven FFT lengths
DMS8 = [0.0, 5.0, 20.0, 50.0, 100.0, 180.0, 260.0, 350.0]
FFT_LEN8 = [4096, 4096, 8192, 8192, 16384, 16384, 32768, 32768]


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def run_form(sw, d_taftp, weights, form, batched, steps, warmup, capture=False, dms=None, fft_len=None,
             share_forward=False, graph=False):
    """Times `steps` gulps after `warmup`; with capture, returns the TFB
    output of every coherent DM for one more gulp.  graph: each gulp's
    device work captured into a hip graph and launched as one (asynchronous
    sinks)."""
    dms = DMS if dms is None else dms
    fft_len = FFT_LEN if fft_len is None else fft_len
    p = sw.BeamformerPipeline(A, F, B, T, TS, dms, CFREQ, BW, stokes="I", fft_len=fft_len, form=form,
                              batched=batched, share_forward=share_forward, async_sinks=graph, graph=graph)
    p.set_weights(weights)
    p.init(A)
    captured = {}
    if capture:
        def cb_sink(blk):
            if captured.get("armed"):
                captured.setdefault("tfb", {})[blk.dm_idx] = sw.download_block(blk)
        p.set_sinks(cb=cb_sink)
    for _ in range(warmup):
        p(d_taftp)
    stage = {k: 0.0 for k in sw.STAGES}
    sw.synchronize()
    e0, e1 = sw.Event(), sw.Event()
    e0.record(p.stream)
    t0 = time.perf_counter()
    for _ in range(steps):
        p(d_taftp)
        for k, v in p.stage_times().items():
            stage[k] += v / steps
    e1.record(p.stream)
    wall = (time.perf_counter() - t0) / steps * 1e3
    ms = e0.elapsed_ms(e1) / steps
    if capture:
        captured["armed"] = True
        p(d_taftp)
    p.close()
    return ms, wall, stage, captured.get("tfb")


def bench_host_input(sw, taftp, weights, steps):
    """PCIe-inclusive rate: gulps from pinned host memory, (a) copied inside
    each process call (copy, then compute) and (b) prefetched, the next
    gulp's copy overlapping the current gulp's processing.  Wall-clock
    milliseconds per gulp."""
    from skyweaver_amd import HostBuffer
    hb = [HostBuffer(taftp.nbytes) for _ in range(2)]
    for h in hb:
        h.array[:] = taftp.reshape(-1).view(np.uint8)
    p = sw.BeamformerPipeline(A, F, B, T, TS, DMS, CFREQ, BW, stokes="I", fft_len=FFT_LEN)
    p.set_weights(weights)
    p.init(A)
    p(hb[0].array)
    p.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        p(hb[k % 2].array)
    serial = (time.perf_counter() - t0) / steps * 1e3
    p.prefetch(hb[0])
    p.process_prefetched()
    p.prefetch(hb[1])
    t0 = time.perf_counter()
    for k in range(steps):
        if k + 1 < steps:
            p.prefetch(hb[k % 2])
        p.process_prefetched()
    overlapped = (time.perf_counter() - t0) / steps * 1e3
    p.synchronize()
    p.close()
    for h in hb:
        h.close()
    return {"ms_per_step_copy_then_compute": round(serial, 3), "ms_per_step_prefetched": round(overlapped, 3),
            "xrealtime_prefetched": round(T * TSAMP / (overlapped * 1e-3), 2),
            "value_prefetched": B * F * T * len(DMS) / (overlapped * 1e-3),
            "gulp_bytes": int(taftp.nbytes), "steps": steps,
            "note": "wall clock per gulp with the input in pinned host memory; not the headline value"}


def _same_outputs(a, b):
    return bool(a is not None and b is not None and a.keys() == b.keys() and
                all(np.array_equal(a[k], b[k]) for k in a))


def bench_target_8dm(sw, d_taftp, weights, steps, warmup, fft_len=None, share_forward=True, label=""):
    """The same gulp with 8 coherent DMs: chosen form (batched dedispersion
    launches, MFMA beams) vs plain form (one DM per dedispersion launch,
    v_dot4 beams), with the same FFT lengths and overlaps for both (with
    share_forward, DMs of equal length share the larger overlap, so the
    chosen form's batched launches transform each segment forward once for
    all of them); outputs compared on the same gulp of the stream."""
    fft_len = FFT_LEN8 if fft_len is None else fft_len
    log(f"8 coherent DMs{label}: chosen form")
    ms_c, _, st_c, tfb_c = run_form(sw, d_taftp, weights, "mfma", True, steps, warmup, capture=True,
                                    dms=DMS8, fft_len=fft_len, share_forward=share_forward)
    log(f"8 coherent DMs{label}: plain form")
    ms_p, _, st_p, tfb_p = run_form(sw, d_taftp, weights, "plain", False, steps, warmup, capture=True,
                                    dms=DMS8, fft_len=fft_len, share_forward=share_forward)
    v_c = B * F * T * len(DMS8) / (ms_c * 1e-3)
    v_p = B * F * T * len(DMS8) / (ms_p * 1e-3)
    return {"workload": "64 antennas x 64 channels x 65536 samples, 800 beams, 8 coherent DMs "
                        f"{DMS8} (FFT {list(fft_len)}), tscrunch 16, Stokes I",
            "share_forward": bool(share_forward),
            "steps": steps, "warmup": warmup,
            "chosen": {"ms_per_step": round(ms_c, 4), "value": v_c, "xrealtime": round(T * TSAMP / (ms_c * 1e-3), 2),
                       "stage_ms": {k: round(v, 4) for k, v in st_c.items()}},
            "plain": {"ms_per_step": round(ms_p, 4), "value": v_p,
                      "stage_ms": {k: round(v, 4) for k, v in st_p.items()}},
            "speedup_vs_plain": round(v_c / v_p, 3), "target_speedup": 4.0,
            "outputs_bit_identical": _same_outputs(tfb_c, tfb_p)}


def host_cores():
    """Host CPU counts as seen by this process, and the threads the CPU
    baseline uses: the affinity set, capped at the CPU share a one-GPU job
    gets on the GPU box (OMP_NUM_THREADS, 16 there)."""
    import subprocess
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    try:
        nproc = int(subprocess.run(["nproc", "--all"], capture_output=True, text=True).stdout.strip())
    except (OSError, ValueError):
        nproc = os.cpu_count() or 1
    share = int(os.environ.get("OMP_NUM_THREADS", "16") or 16)
    return {"nproc_all": nproc, "os_cpu_count": os.cpu_count(), "affinity": aff, "share": share,
            "threads": max(1, min(aff, share))}


def cpu_baseline(sw, orc, taftp, weights, nchan_sample, nthreads):
    """The oracle's full per-gulp path on `nchan_sample` channels."""
    from concurrent.futures import ThreadPoolExecutor
    H = T // 256
    sub = np.ascontiguousarray(taftp[:, :, :nchan_sample])
    w = np.ascontiguousarray(weights[:nchan_sample])
    lowf = CFREQ - BW / 2
    chbw = BW / F
    O = 10
    chirp = orc.dm_chirp(DMS[0], lowf, chbw, F, FFT_LEN[0])
    ts = TS
    scale = np.full(nchan_sample, 1e8, np.float32)
    offset = np.full(nchan_sample, 1e9, np.float32)
    antw = np.ones((1, A), np.float32)
    bmap = np.zeros(B, np.int32)
    t0 = time.perf_counter()
    ftpa = orc.transpose(sub, A)                               # [c][T][2][A][2]

    def per_channel(c):
        x = ftpa[c:c + 1]
        orc.statistics(x)
        hist = np.zeros((O, 2, A, 2), np.int8)
        _, dd = orc.codedisp_channel(np.concatenate([hist, x[0]]), chirp[c], FFT_LEN[0], O, T,
                                     want_float=False, nthreads=1)
        dd = dd[None]
        raw, _ = orc.ib_beamform(dd, ts, scale[c:c + 1], offset[c:c + 1], antw, 1, "I")
        orc.cb_beamform(dd, w[c:c + 1], ts, scale[c:c + 1], offset[c:c + 1], antw, bmap, 1, "I", True)
        return c

    with ThreadPoolExecutor(max_workers=nthreads) as ex:
        list(ex.map(per_channel, range(nchan_sample)))
    return time.perf_counter() - t0


def bench_idisp(sw, steps):
    """One aggregation-buffer dispatch of the incoherent dedisperser, timed
    with HIP events on its stream; inputs resident in HBM."""
    from skyweaver_amd import DeviceBuffer
    idd = sw.IncoherentDedisperser(ID_F, CFREQ, ID_BW, ID_CBTS, ID_DMS, tscrunch=1)
    disp = max(10 * idd.max_delay, 8192)
    nsamp = disp + idd.max_delay
    nout = idd.nout(nsamp)
    # input: a 64 MiB seeded block tiled over the dispatch (sums are data-independent in cost)
    tile = sw.synth_voltages((1 << 25, 2), seed=0x5EED0005, sigma=20.0, nthreads=16).reshape(-1)
    d_in = DeviceBuffer(nsamp * ID_F * ID_B)
    d_tile = DeviceBuffer.from_host(tile)
    for off in range(0, d_in.nbytes, tile.nbytes):
        n = min(tile.nbytes, d_in.nbytes - off)
        sw.check(sw.lib().sw_memcpy_dtod(d_in.ptr + off, d_tile.ptr, n, None))
    d_out = DeviceBuffer(nout * len(ID_DMS) * ID_B)
    st = sw.Stream()
    # channel-major input, as held by the aggregation ring
    idd.dedisperse(d_in, nsamp, ID_B, 1, d_out, stream=st, layout="ftb")   # warm-up
    e0, e1 = sw.Event(), sw.Event()
    e0.record(st)
    for _ in range(steps):
        idd.dedisperse(d_in, nsamp, ID_B, 1, d_out, stream=st, layout="ftb")
    e1.record(st)
    st.synchronize()
    ms = e0.elapsed_ms(e1) / steps
    adds = float(nout) * len(ID_DMS) * ID_B * ID_F
    nbytes = d_in.nbytes + d_out.nbytes
    tsamp_cb = ID_CBTS * ID_F / ID_BW
    # the matrix-core form issues one 16x16x64 int8 product (16 useful adds
    # of 16384 MACs) per 16 beams x 16 samples x 4 channels, i.e. 64 int8
    # operations (MAC = 2 ops) per useful add
    mfma_ops = adds * 64 if idd.kernel in ("idisp_mfma_kernel", "idisp_dma_kernel") else 0.0
    lds_bound_ms = adds / (INT8_LDS_BYTES_PER_CLK * NUM_CU * PEAK_CLK_GHZ * 1e9) * 1e3
    return {"kernel": idd.kernel, "layout": "FTB (channel-major aggregation ring)", "ms_per_dispatch": round(ms, 3), "steps": steps,
            "dispatch_samples": disp, "max_delay": idd.max_delay, "ndms": len(ID_DMS), "nchans": ID_F,
            "nbeams": ID_B, "adds_per_dispatch": adds,
            "adds_per_s": adds / (ms * 1e-3), "valu_add16_peak": VALU_ADD16_PEAK,
            "frac_valu": round(adds / (ms * 1e-3) / VALU_ADD16_PEAK, 4),
            "int8_mfma_ops_per_dispatch": mfma_ops,
            "frac_int8_mfma_peak": round(mfma_ops / (ms * 1e-3) / (INT8_PEAK_TOPS * 1e12), 4),
            "lds_bound_ms": round(lds_bound_ms, 1), "frac_lds_bound": round(lds_bound_ms / ms, 4),
            "hbm_GB/s": round(nbytes / (ms * 1e-3) / 1e9, 1),
            "xrealtime": round(disp * tsamp_cb / (ms * 1e-3), 2)}


def bench_writer_sinks(ngulps=6):
    """Gulp time with the C++ facade's MultiFileWriter handlers attached,
    called synchronously (the reference's order, within each gulp) and by
    the library's sink thread (async handlers), beside no handlers
    (tests/cpp/writer_sinks_bench.cpp; files go to a scratch directory under
    the system temporary directory and are removed)."""
    import subprocess
    import tempfile
    exe = os.path.join(ROOT, "tests", "cpp", "_build", "writer_sinks_bench")
    if not os.path.exists(exe):
        return {"skipped": "tests/cpp/_build/writer_sinks_bench is not built"}
    with tempfile.TemporaryDirectory() as d:
        r = subprocess.run([exe, d, str(ngulps)], capture_output=True, text=True, timeout=300)
    if r.returncode != 0:
        return {"error": r.stderr.strip()[-400:]}
    return json.loads(r.stdout.strip().splitlines()[-1])


def bench_config5(sw, steps, warmup):
    """The full pipeline with incoherent dedispersion at its full size: 64
    antennas x 256 channels (53.5 MHz at 1284 MHz) x 262144-sample gulps,
    512 beams, tscrunch 8, coherent DMs {0, 100}, each feeding 1024
    incoherent DMs 0:1000:0.9765625 through the aggregation ring (dispatch
    max(10 max_delay, 8192) samples).  The 17 GB gulp is a 1 GiB seeded
    block tiled in device memory (the kernels' cost does not depend on the
    values).  Device time per gulp over `steps` gulps, incoherent
    dedispersion dispatches included as they fall."""
    from skyweaver_amd import DeviceBuffer
    A5, F5, T5, B5, TS5 = 64, 256, 262144, 512, 8
    BW5 = 208984.375 * F5
    cdms = [0.0, 100.0]
    nbytes = T5 * A5 * F5 * 4
    free0, total = sw.mem_info()
    tile = sw.synth_voltages((1 << 29, 2), seed=0x5EED0005, sigma=16.0, nthreads=16).reshape(-1)
    d_in = DeviceBuffer(nbytes)
    d_tile = DeviceBuffer.from_host(tile)
    for off in range(0, nbytes, tile.nbytes):
        sw.check(sw.lib().sw_memcpy_dtod(d_in.ptr + off, d_tile.ptr, min(tile.nbytes, nbytes - off), None))
    sw.synchronize()
    d_tile.close()
    del tile
    weights = sw.synth_weights((F5, B5, A5, 2), seed=0x5EED1005)
    free1, _ = sw.mem_info()
    p = sw.BeamformerPipeline(A5, F5, B5, T5, TS5, cdms, CFREQ, BW5, id_plan=[(1, ID_DMS), (1, ID_DMS)])
    p.set_weights(weights)
    p.init(A5)
    ndisp = [0]
    p.set_tdb_sink(lambda blk: ndisp.__setitem__(0, ndisp[0] + 1))
    for _ in range(warmup):
        p(d_in)
    free2, _ = sw.mem_info()
    ndisp[0] = 0
    stage = {k: 0.0 for k in sw.STAGES}
    e0, e1 = sw.Event(), sw.Event()
    e0.record(p.stream)
    for _ in range(steps):
        p(d_in)
        for k, v in p.stage_times().items():
            stage[k] += v / steps
    e1.record(p.stream)
    ms = e0.elapsed_ms(e1) / steps
    p.close()
    d_in.close()
    tsamp = F5 / BW5
    id_adds = (T5 / TS5) * len(ID_DMS) * B5 * F5 * len(cdms)
    return {"workload": "64 antennas x 256 channels (53.5 MHz) x 2 pols x 262144 samples, 512 beams, tscrunch 8, "
                        "coherent DMs [0, 100], each with 1024 incoherent DMs 0:1000:0.9765625 (TDB)",
            "steps": steps, "warmup": warmup, "ms_per_step": round(ms, 3),
            "value": B5 * F5 * T5 * len(cdms) / (ms * 1e-3), "unit": "beam-samples/s",
            "xrealtime": round(T5 * tsamp / (ms * 1e-3), 3),
            "stage_ms": {k: round(v, 3) for k, v in stage.items()},
            "tdb_dispatches_in_timed_gulps": ndisp[0],
            "incoherent_adds_per_gulp": id_adds,
            "hbm_bytes": {"input_gulp": nbytes, "pipeline": int(free1 - free2), "device_total": int(total)},
            "data": "1 GiB seeded Gaussian block (sigma 16, seed 0x5EED0005) tiled over the gulp"}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-channels", type=int, default=0, help="channels in the CPU sample (0 = auto)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-idisp", action="store_true", help="skip the incoherent dedispersion measurement")
    ap.add_argument("--no-8dm", action="store_true", help="skip the 8-DM chosen vs plain measurement")
    ap.add_argument("--no-config5", action="store_true", help="skip the full-size 256-channel pipeline with "
                    "incoherent dedispersion")
    ap.add_argument("--no-writers", action="store_true", help="skip the gulp time with DADA writers attached")
    ap.add_argument("--only-8dm", action="store_true",
                    help="time only the chosen form on the 8-DM workload (for kernel traces) and print its line")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (the path is built for one GPU); --gpus {args.gpus} is not "
              "supported", file=sys.stderr)
        sys.exit(2)

    import skyweaver_amd as sw
    from skyweaver_amd import DeviceBuffer
    if sw.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    H = T // 256
    log("generating synthetic input")
    taftp = sw.synth_voltages((H, A, F, 256, 2, 2), seed=SEED_V, sigma=16.0, nthreads=16)
    weights = sw.synth_weights((F, B, A, 2), seed=SEED_W)
    d_taftp = DeviceBuffer.from_host(taftp)
    if args.only_8dm:
        ms8, _, st8, _ = run_form(sw, d_taftp, weights, "mfma", True, args.steps, args.warmup, dms=DMS8,
                                  fft_len=FFT_LEN8, share_forward=True)
        print(json.dumps({"workload": f"8 coherent DMs {DMS8} (FFT {FFT_LEN8}, shared forward transforms), "
                                     "chosen form",
                          "ms_per_step": round(ms8, 4), "stage_ms": {k: round(v, 4) for k, v in st8.items()}}))
        return

    log("chosen form (MFMA beamformer, batched DMs)")
    ms, wall, stage, tfb_c = run_form(sw, d_taftp, weights, "mfma", True, args.steps, args.warmup,
                                      capture=not args.no_plain)
    plain = None
    if not args.no_plain:
        log("plain form (v_dot4 beamformer, one DM per launch)")
        psteps = max(2, args.steps // 4)
        # the dedisperser's segment grid moves with the gulp index, so both
        # forms are compared on the same gulp of the stream: the plain form
        # runs as many gulps before its capture as the chosen one
        pwarm = max(1, args.warmup + args.steps - psteps)
        pms, pwall, pstage, tfb_p = run_form(sw, d_taftp, weights, "plain", False, psteps, pwarm, capture=True)
        plain = {"ms_per_step": round(pms, 4), "value": B * F * T * len(DMS) / (pms * 1e-3),
                 "steps": psteps, "stage_ms": {k: round(v, 4) for k, v in pstage.items()},
                 "outputs_bit_identical_to_chosen": _same_outputs(tfb_c, tfb_p)}

    value = B * F * T * len(DMS) / (ms * 1e-3)
    xrt = T * TSAMP / (ms * 1e-3)
    cb_ms = stage["coherent_beam"]
    ops = 8.0 * F * B * T * 2 * A           # 4 int8 MACs per complex MAC, 2 ops per MAC
    achieved = ops / (cb_ms * 1e-3) / 1e12
    roofline_cb = {"kernel": "cb_mfma_kernel", "bound": "mfma", "achieved": round(achieved, 1),
                   "peak": INT8_PEAK_TOPS, "unit": "TFLOP/s", "frac": round(achieved / INT8_PEAK_TOPS, 4),
                   "traffic": TRAFFIC_BYTES["coherent_beam"], "ops_per_launch": ops, "avg_launch_ms": round(cb_ms, 4),
                   "note": "int8 operations (1 MAC = 2 ops); peak is the dense int8 MFMA rate; the launch also "
                           "computes the incoherent beam"}
    # the other large kernel, the coherent dedisperser (FP32 VALU): per column
    # and segment a forward and an inverse N-point transform and the chirp,
    # 10 N log2 N + 6 N flops; segments per gulp T / (N - O) on the stream's
    # segment grid.  The overlap O comes from a one-channel plan of the band.
    N = FFT_LEN[0]
    cd_plan = sw.CoherentDedisperser(1, 1, T, CFREQ - BW / 2, BW / F, TSAMP, DMS, fft_len=FFT_LEN)
    try:
        _, O = cd_plan.info(0)
    finally:
        cd_plan.close()
    cd_flops = F * 2 * A * (T / (N - O)) * (10 * N * np.log2(N) + 6 * N)
    cd_ms = stage["dedispersion"]
    cd_tf = cd_flops / (cd_ms * 1e-3) / 1e12
    roofline_dd = {"kernel": "codedisp_reg_kernel<32>", "bound": "valu", "achieved": round(cd_tf, 1),
                   "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(cd_tf / FP32_PEAK_TFLOPS, 4),
                   "traffic": TRAFFIC_BYTES["dedispersion"],
                   "flops_per_launch": cd_flops, "avg_launch_ms": round(cd_ms, 4), "fft_len": N, "overlap": O,
                   "note": "FP32 FFT flops (10 N log2 N + 6 N per column and segment); the peak counts packed "
                           "FP32, which this build does not use (unpacked rate 78.6 TFLOP/s)"}
    # the roofline line is for the kernel that takes the most time per gulp
    roofline = dict(roofline_dd if cd_ms >= cb_ms else roofline_cb)
    roofline["dominant_of"] = {"dedispersion_ms": round(cd_ms, 4), "coherent_beam_ms": round(cb_ms, 4)}
    hbm = {}
    # the transpose stage is taftp_to_fpat_stats_kernel: the statistics' power
    # sums are taken in the same pass (one read of the gulp, one write), and
    # only their finalisation is left in the statistics stage
    # the incoherent beam is computed inside cb_mfma_kernel from the rows it
    # stages (its own stage is left with nothing to time), so only a separate
    # incoherent-beam launch is reported here
    for k, nbytes in (("transpose", 2 * T * F * 2 * A * 2), ("incoherent_beam", T * F * 2 * A * 2)):
        if stage[k] > 0.05:
            gbs = nbytes / (stage[k] * 1e-3) / 1e9
            hbm[k] = {"GB/s": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 3)}
    hbm["transpose"]["kernel"] = "taftp_to_fpat_stats_kernel (transpose + statistics power sums)"
    if "incoherent_beam" not in hbm:
        hbm["incoherent_beam"] = {"kernel": "fused into cb_mfma_kernel (the coherent_beam stage)"}

    target = target_tuned = None
    if not args.no_8dm and not args.no_plain:
        target = bench_target_8dm(sw, d_taftp, weights, max(2, args.steps // 5), 2)
        # the same 8 DMs with FFT lengths picked by the autotuner (one set for both forms)
        log("autotuning the 8 DMs' FFT lengths")
        tuned, share, tms = sw.CoherentDedisperser.autotune(F, A, T, CFREQ - BW / 2, BW / F, TSAMP, DMS8,
                                                            tune_chans=8)
        target_tuned = bench_target_8dm(sw, d_taftp, weights, max(2, args.steps // 5), 2, fft_len=tuned,
                                        share_forward=share, label=" (autotuned FFT lengths)")
        target_tuned["autotune_ms_per_gulp"] = {
            "fft_len": [1024, 2048, 4096, 8192, 16384, 32768],
            "per_dm": {str(dm): [round(float(x), 3) for x in tms[d]] for d, dm in enumerate(DMS8)},
            "batched_all_that_fit": [round(float(x), 3) for x in tms[len(DMS8)]]}

    log("graph mode: each gulp's device work as one hip graph")
    gms, _, _, _ = run_form(sw, d_taftp, weights, "mfma", True, args.steps, args.warmup, graph=True)
    graph_mode = {"ms_per_step": round(gms, 4), "ms_per_step_stream_launches": round(ms, 4),
                  "note": "the headline gulp captured into a hip graph (updated in place each gulp) and launched as "
                          "one, against the same work launched kernel by kernel"}

    host_input = None
    if not args.no_plain:
        log("PCIe-inclusive: gulps from pinned host memory")
        host_input = bench_host_input(sw, taftp, weights, max(3, args.steps // 4))

    idisp = None
    if not args.no_idisp:
        log("incoherent dedispersion (1024 DMs x 256 channels x 512 beams)")
        idisp = bench_idisp(sw, max(1, args.steps // 5))

    config5 = None
    if not args.no_config5:
        log("full pipeline, 256 channels x 262144 samples, 2 coherent DMs x 1024 incoherent DMs")
        config5 = bench_config5(sw, max(3, args.steps // 5), 2)

    writers = None
    if not args.no_writers:
        log("DADA writers attached: synchronous vs asynchronous handlers")
        writers = bench_writer_sinks()

    cpu = None
    if not args.no_cpu:
        import oracle
        cores = host_cores()
        nthreads = cores["threads"]
        # three channels per thread: about 13 s of CPU work on 16 threads
        nchan = args.cpu_channels or max(1, min(F, 3 * nthreads))
        log(f"CPU oracle sample: {nchan} channels on {nthreads} threads")
        secs = cpu_baseline(sw, oracle, taftp, weights, nchan, nthreads)
        cpu_value = B * nchan * T * len(DMS) / secs
        log("CPU oracle sample: 2 channels on 1 thread")
        secs1 = cpu_baseline(sw, oracle, taftp, weights, 2, 1)
        cpu = {"value": cpu_value, "unit": "beam-samples/s", "cores": nthreads, "kind": "port",
               "sample": f"{nchan} of {F} channels of one gulp (transpose, statistics, FP64 dedispersion, "
                         f"incoherent and coherent beams), {secs:.1f} s on {nthreads} threads; work is linear in "
                         "channels",
               "seconds": round(secs, 2), "xrealtime": T * TSAMP / (secs * F / nchan),
               "host": cores,
               "one_core": {"value": B * 2 * T * len(DMS) / secs1, "seconds": round(secs1, 2),
                            "sample": f"2 of {F} channels on 1 thread"}}

    out = {
        "metric": "beam-samples/s",
        "value": value,
        "unit": "beam-samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8",
        "data": "synthetic",
        "config": {"workload": "64 antennas x 64 channels x 2 pols x 65536 samples, 800 beams, coherent DM 50 "
                               "(FFT 16384, overlap 10), tscrunch 16, Stokes I, incoherent beam + statistics",
                   "antennas": A, "channels": F, "samples_per_gulp": T, "beams": B, "tscrunch": TS,
                   "coherent_dms": DMS, "fft_len": FFT_LEN, "form": "int8 MFMA + batched DMs",
                   "parallelism": "single GPU"},
        "xrealtime": round(xrt, 2),
        "wall_ms_per_step": round(wall, 4),
        "stage_ms": {k: round(v, 4) for k, v in stage.items()},
        "roofline": roofline,
        "roofline_dedispersion": roofline_dd,
        "roofline_coherent_beam": roofline_cb,
        "hbm_kernels": hbm,
        "plain": plain,
        "speedup_vs_plain": round(value / plain["value"], 3) if plain else None,
        "target_8dm": target,
        "target_8dm_autotuned_fft": target_tuned,
        "incoherent_dedispersion": idisp,
        "config5_full_pipeline": config5,
        "writer_sinks": writers,
        "graph_mode": graph_mode,
        "host_input": host_input,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
