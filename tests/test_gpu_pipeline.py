"""End-to-end BeamformerPipeline on the GPU against the oracle chain.

Stage by stage: statistics at 1e-5; scalings from the GPU statistics through
the oracle's update_scalings; DM 0 (identity dedispersion) coherent and
incoherent beams bit-exact; DM > 0 beams within +-1 of the oracle fed with
FP64-dedispersed voltages (the int8 dedispersed voltages themselves agree
except for values within FP32 rounding of a half-integer)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

A, F, T, B, TS = 64, 8, 4096, 64, 16
CFREQ, BW = 1284e6, 13.375e6 * F / 64
TSAMP = F / BW


def _run_pipeline(sw, gulps, weights, dms, fft_len, form, batched, bmap=None, antw=None, stokes="I"):
    p = sw.BeamformerPipeline(A, F, B, T, TS, dms, CFREQ, BW, stokes=stokes, fft_len=fft_len, form=form,
                              batched=batched)
    p.set_weights(weights)
    if bmap is not None:
        p.set_beamsets(bmap, antw)
    p.init(A)
    got = {"cb": [], "ib": [], "stats": []}

    def grab(key):
        def f(blk):
            got[key].append((blk.dm_idx, sw.download_block(blk), blk.utc_offset, blk.tsamp))
        return f
    p.set_sinks(cb=grab("cb"), ib=grab("ib"), stats=grab("stats"))
    for g in gulps:
        p(g)
    p.close()
    return got


def test_pipeline_end_to_end(sw, orc, dev):
    from skyweaver_amd import DeviceBuffer as DB
    H = T // 256
    gulps = [sw.synth_voltages((H, A, F, 256, 2, 2), seed=0x5EED0001 + 16 * i, sigma=16.0) for i in range(2)]
    weights = sw.synth_weights((F, B, A, 2), seed=0x5EED1001)
    dms, fft_len = [0.0, 50.0], [1024, 2048]
    d_gulps = [DB.from_host(g) for g in gulps]
    got = _run_pipeline(sw, d_gulps, weights, dms, fft_len, "mfma", True)
    assert len(got["cb"]) == 4 and len(got["ib"]) == 4 and len(got["stats"]) == 2

    antw = np.ones((1, A), np.float32)
    bmap = np.zeros(B, np.int32)
    gpu_stats0 = got["stats"][0][1].view(np.float32).reshape(F, 2, A, 4)
    scal = orc.update_scalings(gpu_stats0[..., 1], antw, 1, TS, TS, 24.0)
    lowf = CFREQ - BW / 2
    ftpas = [orc.transpose(g, A) for g in gulps]
    # DM 50 (N 2048, O 10): hop 2038 does not divide T, so the dedisperser
    # runs on one segment grid over both gulps and emits each gulp's window
    # one hop late; the oracle dedisperses the whole stream
    O, lat = 10, fft_len[1] - 10
    chirp = orc.dm_chirp(dms[1], lowf, BW / F, F, fft_len[1])
    stream = np.concatenate(ftpas, 1)
    zero = np.zeros((O, 2, A, 2), np.int8)
    dstream = np.stack([orc.codedisp_channel(np.concatenate([zero, stream[f]]), chirp[f], fft_len[1], O,
                                             stream.shape[1], want_float=False, nthreads=16)[1] for f in range(F)])
    dstream = np.concatenate([np.zeros((F, lat, 2, A, 2), np.int8), dstream], 1)  # emitted = stream - lat
    exact = []
    for gi, g in enumerate(gulps):
        ftpa = ftpas[gi]
        st = got["stats"][gi][1].view(np.float32).reshape(F, 2, A, 4)
        np.testing.assert_allclose(st.astype(np.float64), orc.statistics(ftpa), rtol=1e-5, atol=1e-6)
        for d, dm in enumerate(dms):
            if dm == 0.0:
                dd = ftpa
            else:
                dd = np.ascontiguousarray(dstream[:, gi * T:(gi + 1) * T])
            _, ib_q = orc.ib_beamform(dd, TS, scal["ib_scaling"], scal["ib_offsets"], antw, 1, "I")
            cb = orc.cb_beamform(dd, weights, TS, scal["cb_scaling"], scal["cb_offsets"], antw, bmap, 1, "I", True)
            g_cb = got["cb"][2 * gi + d][1].view(np.int8).reshape(cb.shape)
            g_ib = got["ib"][2 * gi + d][1].view(np.int8).reshape(ib_q.shape)
            assert got["cb"][2 * gi + d][0] == d
            if dm == 0.0:
                np.testing.assert_array_equal(g_cb, cb)
                np.testing.assert_array_equal(g_ib, ib_q)
            else:
                dcb = g_cb.astype(int) - cb
                assert np.abs(dcb).max() <= 1 and np.mean(dcb == 0) > 0.995
                assert np.abs(g_ib.astype(int) - ib_q).max() <= 1
                exact.append(np.mean(dcb == 0))
    print(f"DM 50 coherent beams exact fraction: {exact}", flush=True)
    # metadata: the DM-50 blocks carry the filter delay O/2 minus the latency
    assert got["cb"][1][2] == pytest.approx(TSAMP * (5 - lat), rel=1e-9)
    assert got["cb"][1][3] == pytest.approx(TSAMP * TS, rel=1e-9)


@pytest.mark.parametrize("stokes,sigma", [("IQUV", 16.0), ("I", 16.0), ("I", 100.0)])
def test_pipeline_forms_identical_and_host_input(sw, dev, stokes, sigma):
    """plain (v_dot4, one DM per launch, FTPA beamformer input) == chosen
    (MFMA, batched; for Stokes I with 0/1 antenna weights the beamformers
    read the dedisperser's time-major ring) bit for bit, and a host-memory
    gulp gives the same result as a device-resident one.  sigma 100 takes the
    incoherent beam's sums past 2^24, where it follows the float chain."""
    from skyweaver_amd import DeviceBuffer as DB
    H = T // 256
    g = sw.synth_voltages((H, A, F, 256, 2, 2), seed=7, sigma=sigma)
    weights = sw.synth_weights((F, B, A, 2), seed=8)
    bmap = (np.arange(B) % 2).astype(np.int32)
    antw = np.ones((2, A), np.float32)
    antw[1, 40:] = 0
    dms, fft_len = [0.0, 20.0, 50.0], [1024, 1024, 1024]
    a = _run_pipeline(sw, [DB.from_host(g)], weights, dms, fft_len, "mfma", True, bmap, antw, stokes)
    b = _run_pipeline(sw, [g], weights, dms, fft_len, "plain", False, bmap, antw, stokes)
    for key in ("cb", "ib", "stats"):
        assert len(a[key]) == len(b[key])
        for x, y in zip(a[key], b[key]):
            np.testing.assert_array_equal(x[1], y[1])


def test_pipeline_prefetched_input(sw, dev):
    """Gulps copied ahead from pinned host memory (sw_pipeline_prefetch, up to
    two pending) and processed by sw_pipeline_process_prefetched give the same
    blocks as the same gulps processed from device memory."""
    from skyweaver_amd import DeviceBuffer as DB, HostBuffer
    H = T // 256
    gulps = [sw.synth_voltages((H, A, F, 256, 2, 2), seed=0x5EED0301 + i, sigma=16.0) for i in range(3)]
    weights = sw.synth_weights((F, B, A, 2), seed=0x5EED1301)
    dms, fft_len = [0.0, 50.0], [1024, 2048]
    a = _run_pipeline(sw, [DB.from_host(g) for g in gulps], weights, dms, fft_len, "mfma", True)

    hb = []
    for g in gulps:
        h = HostBuffer(g.nbytes)
        h.array[:] = g.reshape(-1).view(np.uint8)
        hb.append(h)
    p = sw.BeamformerPipeline(A, F, B, T, TS, dms, CFREQ, BW, fft_len=fft_len)
    p.set_weights(weights)
    p.init(A)
    got = {"cb": [], "ib": [], "stats": []}

    def grab(key):
        def f(blk):
            got[key].append((blk.dm_idx, sw.download_block(blk), blk.utc_offset, blk.tsamp))
        return f
    p.set_sinks(cb=grab("cb"), ib=grab("ib"), stats=grab("stats"))
    with pytest.raises(sw.SkyweaverError, match="no prefetched gulp"):
        p.process_prefetched()
    p.prefetch(hb[0])
    p.prefetch(hb[1])
    with pytest.raises(sw.SkyweaverError, match="already prefetched"):
        p.prefetch(hb[2])
    p.process_prefetched()
    p.prefetch(hb[2])
    p.process_prefetched()
    p.process_prefetched()
    p.synchronize()
    p.close()
    for key in ("cb", "ib", "stats"):
        assert len(got[key]) == len(a[key]) > 0
        for x, y in zip(a[key], got[key]):
            assert x[0] == y[0] and x[2] == y[2]
            np.testing.assert_array_equal(x[1], y[1])


def test_pipeline_rejects_wrong_gulp_size(sw, dev):
    p = sw.BeamformerPipeline(A, F, B, T, TS, [0.0], CFREQ, BW)
    p.set_weights(np.zeros((F, B, A, 2), np.int8))
    p.init(A)
    with pytest.raises(sw.SkyweaverError, match="Unexpected buffer size"):
        p(np.zeros(100, np.int8))


def _read_models(path):
    hdr = np.dtype([("nbeams", "<u4"), ("nantennas", "<u4"), ("start", "<f8"), ("end", "<f8")])
    raw, out, off = open(path, "rb").read(), [], 0
    while off < len(raw):
        h = np.frombuffer(raw, hdr, 1, off)[0]
        off += hdr.itemsize
        n = int(h["nbeams"]) * int(h["nantennas"]) * 3
        out.append((float(h["start"]), float(h["end"]),
                    np.frombuffer(raw, "<f4", n, off).reshape(h["nbeams"], h["nantennas"], 3)))
        off += 4 * n
    return out


def test_pipeline_delay_file(sw, orc, dev, golden_dir):
    """Weights and beamsets from the reference's delay-model fixture: per gulp
    the model valid at utc_start + n * 2 * total_nchans * T / sample_clock is
    turned into weights on the device; the beams equal the oracle chain fed
    with the oracle's weights for that time (DM 0: bit-exact)."""
    import os
    from skyweaver_amd import DeviceBuffer as DB
    Bd = 72                      # >= the 67 beams of the fixture
    H = T // 256
    utc_start, clock, total = 1708082300.0, 1712e6, 4096
    p = sw.BeamformerPipeline(A, F, Bd, T, TS, [0.0], CFREQ, BW, fft_len=[1024], total_nchans=total)
    p.set_delay_file(os.path.join(golden_dir, "test_delays.bin"))
    p.init(57, utc_start=utc_start, sample_clock=clock)
    got, stats = [], []
    p.set_sinks(cb=lambda blk: got.append(sw.download_block(blk)),
                stats=lambda blk: stats.append(sw.download_block(blk).view(np.float32).reshape(F, 2, A, 4)))
    gulps = [sw.synth_voltages((H, 57, F, 256, 2, 2), seed=0x5EED0009 + i, sigma=16.0) for i in range(2)]
    for g in gulps:
        p(DB.from_host(g))
    p.close()
    models = _read_models(os.path.join(golden_dir, "test_delays.bin"))
    freqs = sw.channel_frequencies(CFREQ, BW, F)
    nsets, bmap, bw = orc.parse_beamsets(models[0][2], Bd, A)
    assert nsets == 2
    scal = None
    for gi, g in enumerate(gulps):
        t_unix = utc_start + gi * 2.0 * total * T / clock
        k = next(i for i, m in enumerate(models) if m[0] <= t_unix <= m[1])
        delays = np.zeros((Bd, A, 3), np.float32)
        delays[:67, :57] = models[k][2]
        w = orc.weights(delays, freqs, t_unix, models[k][0])
        ftpa = orc.transpose(g, A)
        if scal is None:  # from the GPU statistics of the first gulp, as the pipeline does
            scal = orc.update_scalings(stats[0][..., 1], bw, nsets, TS, TS, 24.0)
        raw, _ = orc.ib_beamform(ftpa, TS, scal["ib_scaling"], scal["ib_offsets"], bw, nsets, "I")
        cb = orc.cb_beamform(ftpa, w, TS, scal["cb_scaling"], scal["cb_offsets"], bw, bmap, nsets, "I", True)
        g_cb = got[gi].view(np.int8).reshape(cb.shape)
        np.testing.assert_array_equal(g_cb, cb)


def test_pipeline_incoherent_dedispersion(sw, orc, dev):
    """The coherent beams of every gulp feed the incoherent dedisperser; each
    TDB block equals the oracle's dedispersion of the concatenated TFB stream."""
    from skyweaver_amd import DeviceBuffer as DB
    H = T // 256
    weights = sw.synth_weights((F, B, A, 2), seed=0x5EED100A)
    idms = np.arange(0, 40, 2.5, dtype=np.float32)
    p = sw.BeamformerPipeline(A, F, B, T, TS, [0.0], CFREQ, BW, fft_len=[1024], id_plan=[(2, idms)],
                              id_dispatch_size=300)
    p.set_weights(weights)
    p.init(A)
    tfb, tdb = [], []
    p.set_sinks(cb=lambda blk: tfb.append(sw.download_block(blk).view(np.int8).reshape(blk.dims[0], blk.dims[1],
                                                                                         blk.dims[2])))
    p.set_tdb_sink(lambda blk: tdb.append((sw.download_block(blk).view(np.int8).reshape(
        blk.dims[0], blk.dims[1], blk.dims[2]), blk.utc_offset, blk.tsamp)))
    for i in range(3):
        p(DB.from_host(sw.synth_voltages((H, A, F, 256, 2, 2), seed=0x5EED0100 + i, sigma=16.0)))
    times = p.stage_times()
    p.close()
    stream = np.concatenate(tfb)                       # 3 x 256 TFB samples
    tsamp_cb = TSAMP * TS
    d_o, mx = orc.id_delays(idms, sw.channel_frequencies(CFREQ, BW, F), tsamp_cb)
    assert mx > 0
    nblk = (len(stream) - mx) // 300
    assert len(tdb) == nblk >= 1
    for k, (a, utc, tsb) in enumerate(tdb):
        np.testing.assert_array_equal(a, orc.incoherent_dedisperse(stream[k * 300:k * 300 + 300 + mx], d_o, mx, 2))
        assert utc == pytest.approx((k * 300 + mx) * tsamp_cb, rel=1e-9)
        assert tsb == pytest.approx(tsamp_cb * 2, rel=1e-12)
    assert times["incoherent_dedispersion"] >= 0.0


def test_pipeline_full_size_beamsets_iquv(sw, orc, dev):
    """A whole subband with uneven beamsets: 80 antennas, 128 channels, 65536
    samples, 1024 beams in four beamsets over antennas 0-79, 0-63, 0-39 and
    0-19 (0/1 antenna weights; the complex weights of excluded antennas are
    zero, as delay models make them), full Stokes, tscrunch 16, DM 0.  The
    chosen form (MFMA) and the plain form (v_dot4) agree bit for bit on every
    coherent and incoherent beam and statistic; the first 256 output samples
    of channels 0 and 127 are checked against the oracle chain."""
    from skyweaver_amd import DeviceBuffer as DB
    A4, F4, T4, B4, TS4 = 80, 128, 65536, 1024, 16
    BW4 = 208984.375 * F4
    bounds = [(0, 256, 80), (256, 640, 64), (640, 896, 40), (896, 1024, 20)]
    S4 = len(bounds)
    print("full-size IQUV beamsets: synthesising 2.7 GB", flush=True)
    taftp = sw.synth_voltages((T4 // 256, A4, F4, 256, 2, 2), seed=0x5EED0004, sigma=16.0, nthreads=16)
    weights = sw.synth_weights((F4, B4, A4, 2), seed=0x5EED1004)
    bmap = np.zeros(B4, np.int32)
    antw = np.zeros((S4, A4), np.float32)
    for s, (b0, b1, na) in enumerate(bounds):
        bmap[b0:b1] = s
        antw[s, :na] = 1
        weights[:, b0:b1, na:] = 0
    d_in = DB.from_host(taftp)

    def run(form, batched):
        p = sw.BeamformerPipeline(A4, F4, B4, T4, TS4, [0.0], 1284e6, BW4, stokes="IQUV", form=form,
                                  batched=batched)
        p.set_weights(weights)
        p.set_beamsets(bmap, antw)
        p.init(A4)
        got = {}

        def grab(key):
            def f(blk):
                got[key] = sw.download_block(blk)
            return f
        p.set_sinks(cb=grab("cb"), ib=grab("ib"), stats=grab("stats"))
        p(d_in)
        p.close()
        return got

    print("chosen form", flush=True)
    a = run("mfma", True)
    print("plain form", flush=True)
    b = run("plain", False)
    for key in ("cb", "ib", "stats"):
        np.testing.assert_array_equal(a[key], b[key])
    cb = a["cb"].view(np.int8).reshape(T4 // TS4, F4, B4, 4)
    ib = a["ib"].view(np.int8).reshape(S4, T4 // TS4, F4, 4)
    assert np.unique(cb).size > 100

    stats = a["stats"].view(np.float32).reshape(F4, 2, A4, 4)
    scal = orc.update_scalings(stats[..., 1], antw, S4, TS4, TS4, 24.0)
    nt = 4096
    for f in (0, F4 - 1):
        ftpa = orc.transpose(np.ascontiguousarray(taftp[:, :, f:f + 1]), A4)[:, :nt]
        sel = np.arange(S4) * F4 + f
        _, ib_q = orc.ib_beamform(ftpa, TS4, scal["ib_scaling"][sel], scal["ib_offsets"][sel], antw, S4, "IQUV")
        cb_o = orc.cb_beamform(ftpa, weights[f:f + 1], TS4, scal["cb_scaling"][sel], scal["cb_offsets"][sel],
                               antw, bmap, S4, "IQUV", True)
        np.testing.assert_array_equal(cb[:nt // TS4, f:f + 1], cb_o)
        np.testing.assert_array_equal(ib[:, :nt // TS4, f:f + 1], ib_q)


def test_pipeline_full_size_8dm_forms_identical(sw, dev):
    """64 antennas x 64 channels x 131072 samples, 512 beams, 8 coherent DMs
    0-350 with their FFT lengths 4096-32768 (uneven), tscrunch 8, Stokes I,
    two gulps (the second reads its dedispersion history from the first):
    the chosen form (MFMA beams on the dedisperser's ring, batched DMs) and
    the plain form (v_dot4 beams on corner-turned FTPA, one DM per launch)
    give identical coherent beams, incoherent beams and statistics."""
    from skyweaver_amd import DeviceBuffer as DB
    A3, F3, T3, B3, TS3 = 64, 64, 131072, 512, 8
    dms = [0.0, 5.0, 20.0, 50.0, 100.0, 180.0, 260.0, 350.0]
    fft = [4096, 4096, 8192, 8192, 16384, 16384, 32768, 32768]
    BW3 = 208984.375 * F3
    print("full-size 8 DMs: synthesising 2 x 2.1 GB", flush=True)
    gulps = [DB.from_host(sw.synth_voltages((T3 // 256, A3, F3, 256, 2, 2), seed=0x5EED0003 + i, sigma=16.0,
                                            nthreads=16)) for i in range(2)]
    weights = sw.synth_weights((F3, B3, A3, 2), seed=0x5EED1003)

    def run(form, batched):
        p = sw.BeamformerPipeline(A3, F3, B3, T3, TS3, dms, 1284e6, BW3, fft_len=fft, form=form, batched=batched)
        p.set_weights(weights)
        p.init(A3)
        got = {"cb": [], "ib": [], "stats": []}

        def grab(key):
            def f(blk):
                got[key].append((blk.dm_idx, sw.download_block(blk)))
            return f
        p.set_sinks(cb=grab("cb"), ib=grab("ib"), stats=grab("stats"))
        for g in gulps:
            p(g)
        p.close()
        return got

    print("chosen form", flush=True)
    a = run("mfma", True)
    print("plain form", flush=True)
    b = run("plain", False)
    assert len(a["cb"]) == len(b["cb"]) == 16 and len(a["stats"]) == 2
    for key in ("cb", "ib", "stats"):
        for x, y in zip(a[key], b[key]):
            assert x[0] == y[0]
            np.testing.assert_array_equal(x[1], y[1])
    assert np.unique(a["cb"][-1][1]).size > 100


@pytest.mark.parametrize("nant,ts,sets", [(64, 4, 2), (32, 8, 2), (128, 32, 2), (16, 16, 2), (48, 16, 2), (64, 2, 2),
                                          (64, 4, 1), (128, 32, 1), (64, 16, -1)])
def test_pipeline_fused_incoherent_beam_shapes(sw, dev, nant, ts, sets):
    """The incoherent beam computed inside the coherent beamformer (antenna
    counts dividing 256, tscrunch * antennas >= 256) and, for the other
    shapes here (48 antennas; tscrunch 2), by its own kernel on the ring:
    both equal the plain form's incoherent beam on FTPA, for a DM with
    latency and sets = 2: two beamsets over uneven antenna subsets; 1: one
    beamset over every antenna (the fused form's single-dot-product path);
    -1: one beamset over a subset."""
    from skyweaver_amd import DeviceBuffer as DB
    Fq, Tq, Bq = 4, 4096, 96
    bw = 13.375e6 * Fq / 64
    H = Tq // 256
    g = [DB.from_host(sw.synth_voltages((H, nant, Fq, 256, 2, 2), seed=100 + nant + ts + i, sigma=24.0))
         for i in range(2)]
    weights = sw.synth_weights((Fq, Bq, nant, 2), seed=200 + nant)
    if sets == 2:
        bmap = (np.arange(Bq) % 2).astype(np.int32)
        antw = np.ones((2, nant), np.float32)
        antw[1, nant // 3:] = 0
        weights[:, 1::2, nant // 3:] = 0
    else:
        bmap = np.zeros(Bq, np.int32)
        antw = np.ones((1, nant), np.float32)
        if sets < 0:
            antw[0, ::5] = 0
            weights[:, :, ::5] = 0

    def run(form, batched):
        p = sw.BeamformerPipeline(nant, Fq, Bq, Tq, ts, [0.0, 30.0], 1284e6, bw, fft_len=[1024, 2048], form=form,
                                  batched=batched)
        p.set_weights(weights)
        p.set_beamsets(bmap, antw)
        p.init(nant)
        got = {"cb": [], "ib": []}

        def grab(key):
            def f(blk):
                got[key].append(sw.download_block(blk))
            return f
        p.set_sinks(cb=grab("cb"), ib=grab("ib"))
        for x in g:
            p(x)
        p.close()
        return got

    a, b = run("mfma", True), run("plain", False)
    for key in ("cb", "ib"):
        assert len(a[key]) == len(b[key]) == 4
        for x, y in zip(a[key], b[key]):
            np.testing.assert_array_equal(x, y)


def test_pipeline_set_beamsets_between_gulps_keeps_time_base(sw, orc, dev):
    """Changing the beamsets after a gulp recomputes the scalings from the
    next gulp's statistics and leaves the gulp count, the pipeline's time
    base, alone: the second gulp's blocks carry its own utc_offset and its
    beams equal the oracle's with scalings from its own statistics."""
    from skyweaver_amd import DeviceBuffer as DB
    H = T // 256
    gulps = [sw.synth_voltages((H, A, F, 256, 2, 2), seed=0x5EED0501 + i, sigma=16.0 + 8 * i) for i in range(2)]
    weights = sw.synth_weights((F, B, A, 2), seed=0x5EED1501)
    utc_start, clock, total = 1708082300.0, 1712e6, 4096
    p = sw.BeamformerPipeline(A, F, B, T, TS, [0.0], CFREQ, BW, fft_len=[1024], total_nchans=total)
    p.set_weights(weights)
    p.init(A, utc_start=utc_start, sample_clock=clock)
    got = {"cb": [], "ib": [], "stats": []}

    def grab(key):
        def f(blk):
            got[key].append((sw.download_block(blk), blk.utc_offset))
        return f
    p.set_sinks(cb=grab("cb"), ib=grab("ib"), stats=grab("stats"))
    p(DB.from_host(gulps[0]))
    bmap = (np.arange(B) % 2).astype(np.int32)
    antw = np.ones((2, A), np.float32)
    antw[1, 40:] = 0
    w2 = weights.copy()
    w2[:, 1::2, 40:] = 0
    p.set_beamsets(bmap, antw)
    p.set_weights(w2)
    p(DB.from_host(gulps[1]))
    p.close()
    step = 2.0 * total * T / clock
    for key in ("cb", "ib", "stats"):
        assert [u for _, u in got[key]] == pytest.approx([0.0, step], abs=1e-12)
    ftpa = orc.transpose(gulps[1], A)
    st = got["stats"][1][0].view(np.float32).reshape(F, 2, A, 4)
    scal = orc.update_scalings(st[..., 1], antw, 2, TS, TS, 24.0)
    _, ib_q = orc.ib_beamform(ftpa, TS, scal["ib_scaling"], scal["ib_offsets"], antw, 2, "I")
    cb = orc.cb_beamform(ftpa, w2, TS, scal["cb_scaling"], scal["cb_offsets"], antw, bmap, 2, "I", True)
    np.testing.assert_array_equal(got["cb"][1][0].view(np.int8).reshape(cb.shape), cb)
    np.testing.assert_array_equal(got["ib"][1][0].view(np.int8).reshape(ib_q.shape), ib_q)


def test_pipeline_sink_exception_propagates(sw, dev):
    """An exception raised by a Python sink stops the gulp (the C side sees a
    non-zero return) and is re-raised by the call that processed it."""
    from skyweaver_amd import DeviceBuffer as DB
    H = T // 256
    p = sw.BeamformerPipeline(A, F, B, T, TS, [0.0], CFREQ, BW, fft_len=[1024])
    p.set_weights(sw.synth_weights((F, B, A, 2), seed=3))
    p.init(A)

    class Boom(Exception):
        pass

    def bad(blk):
        raise Boom("sink failed")
    p.set_sinks(cb=bad)
    with pytest.raises(Boom, match="sink failed"):
        p(DB.from_host(sw.synth_voltages((H, A, F, 256, 2, 2), seed=4, sigma=16.0)))
    p.close()


def test_pipeline_async_sinks_match_synchronous(sw, dev):
    """Sinks called by the library's thread (async_sinks) receive the same
    blocks, in the same order and with the same metadata, as synchronous
    sinks, also when they are slow enough that the pipeline must wait for a
    block's delivery before reusing its buffer (four gulps, two buffers per
    DM); synchronize() returns once every block is delivered."""
    import time
    from skyweaver_amd import DeviceBuffer as DB
    H = T // 256
    gulps = [DB.from_host(sw.synth_voltages((H, A, F, 256, 2, 2), seed=0x5EED0601 + i, sigma=16.0)) for i in range(4)]
    weights = sw.synth_weights((F, B, A, 2), seed=0x5EED1601)
    dms, fft_len = [0.0, 20.0, 50.0], [1024, 2048, 1024]

    def run(async_sinks, delay):
        p = sw.BeamformerPipeline(A, F, B, T, TS, dms, CFREQ, BW, fft_len=fft_len, async_sinks=async_sinks)
        p.set_weights(weights)
        p.init(A)
        got = []

        def grab(key):
            def f(blk):
                if delay:
                    time.sleep(delay)
                got.append((key, blk.dm_idx, blk.utc_offset, blk.tsamp, sw.download_block(blk)))
            return f
        p.set_sinks(cb=grab("cb"), ib=grab("ib"), stats=grab("stats"))
        for g in gulps:
            p(g)
        p.synchronize()
        n = len(got)
        p.close()
        return got, n

    ref, _ = run(False, 0.0)
    for delay in (0.0, 0.02):
        got, n = run(True, delay)
        assert n == len(ref) == 4 * (2 * len(dms) + 1)
        for x, y in zip(ref, got):
            assert x[:4] == y[:4]
            np.testing.assert_array_equal(x[4], y[4])


def test_pipeline_async_sink_exception_propagates(sw, dev):
    """With asynchronous sinks a sink's exception is raised by a later call
    into the pipeline: the process call itself if the library's thread has
    already delivered the block by the time it returns, else synchronize."""
    from skyweaver_amd import DeviceBuffer as DB
    H = T // 256
    p = sw.BeamformerPipeline(A, F, B, T, TS, [0.0], CFREQ, BW, fft_len=[1024], async_sinks=True)
    p.set_weights(sw.synth_weights((F, B, A, 2), seed=5))
    p.init(A)

    class Boom(Exception):
        pass

    def bad(blk):
        raise Boom("late sink failed")
    p.set_sinks(ib=bad)
    g = DB.from_host(sw.synth_voltages((H, A, F, 256, 2, 2), seed=6, sigma=16.0))
    with pytest.raises(Boom, match="late sink failed"):
        p(g)
        p.synchronize()
    p.close()


@pytest.mark.parametrize("stokes,form", [("I", "mfma"), ("IQUV", "mfma"), ("I", "plain")])
def test_pipeline_graph_mode_matches_stream_launches(sw, dev, stokes, form):
    """Graph mode (each gulp after the first captured into a hip graph whose
    executable is updated in place and launched) against the same gulps
    launched on the stream: six gulps, three DMs whose segment grids move
    from gulp to gulp (the kernels' arguments change every gulp), the sinks
    downloading every block between the launches; every block of every
    replay is identical.  With Stokes I the coherent beams also feed the
    incoherent dedispersion, which runs after each graph."""
    from skyweaver_amd import DeviceBuffer as DB
    H = T // 256
    gulps = [DB.from_host(sw.synth_voltages((H, A, F, 256, 2, 2), seed=0x5EED0701 + i, sigma=16.0)) for i in range(6)]
    weights = sw.synth_weights((F, B, A, 2), seed=0x5EED1701)
    dms, fft_len = [0.0, 20.0, 50.0], [1024, 2048, 1024]
    idms = np.arange(0, 20, 2.5, dtype=np.float32)
    id_plan = [(2, idms)] * 3 if stokes == "I" else None

    def run(graph):
        p = sw.BeamformerPipeline(A, F, B, T, TS, dms, CFREQ, BW, stokes=stokes, fft_len=fft_len, form=form,
                                  batched=form == "mfma", async_sinks=True, graph=graph, id_plan=id_plan,
                                  id_dispatch_size=200 if id_plan else 0)
        p.set_weights(weights)
        p.init(A)
        got = []

        def grab(key):
            def f(blk):
                got.append((key, blk.dm_idx, blk.utc_offset, sw.download_block(blk)))
            return f
        p.set_sinks(cb=grab("cb"), ib=grab("ib"), stats=grab("stats"))
        if id_plan:
            p.set_tdb_sink(grab("tdb"))
        for g in gulps:
            p(g)
        p.synchronize()
        n = len(got)
        p.close()
        return got, n

    ref, nref = run(False)
    got, n = run(True)
    assert n == nref and nref >= 6 * (2 * len(dms) + 1)
    if id_plan:
        assert sum(1 for x in ref if x[0] == "tdb") >= 3
    key = lambda x: (x[0], x[1], x[2])
    # the TDB blocks come from the stream in both runs; the sink thread may
    # deliver them between other blocks, so compare per kind in order
    for kind in ("cb", "ib", "stats", "tdb"):
        a = [x for x in ref if x[0] == kind]
        b = [x for x in got if x[0] == kind]
        assert [key(x) for x in a] == [key(x) for x in b]
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x[3], y[3])


def test_pipeline_tiles_from_table_identical(sw, dev, tmp_path):
    """Workgroup shapes from a tuning table (2 and 8 waves, on the
    dedisperser's ring with the fused incoherent beam) and the table's FFT
    lengths: the pipeline reports where its tiles came from, and its blocks
    equal those of the default tiles."""
    import json
    from skyweaver_amd import DeviceBuffer as DB
    H = T // 256
    gulps = [DB.from_host(sw.synth_voltages((H, A, F, 256, 2, 2), seed=0x5EED0801 + i, sigma=16.0)) for i in range(2)]
    weights = sw.synth_weights((F, B, A, 2), seed=0x5EED1801)
    dms = [0.0, 50.0]

    def run(**kw):
        p = sw.BeamformerPipeline(A, F, B, T, TS, dms, CFREQ, BW, **kw)
        p.set_weights(weights)
        p.init(A)
        got = []
        p.set_sinks(cb=lambda blk: got.append(sw.download_block(blk)), ib=lambda blk: got.append(sw.download_block(blk)))
        for g in gulps:
            p(g)
        tile = p.tile()
        p.close()
        return got, tile

    ref, t0 = run(fft_len=[1024, 2048], tuning_file="")
    assert t0 == {"cb_waves": 4, "fft_len": [1024, 2048], "share_forward": False, "source": "default"}
    for waves in (2, 8):
        table = tmp_path / f"t{waves}.json"
        table.write_text(json.dumps({"entries": [
            {"antennas": A, "beams": B, "samples": T, "tscrunch": TS, "stokes": "I", "cb_waves": waves,
             "dms": dms, "fft_len": [1024, 2048], "share_forward": 0}]}))
        got, t = run(tuning_file=str(table))
        assert t == {"cb_waves": waves, "fft_len": [1024, 2048], "share_forward": False, "source": "table+fft"}
        assert len(got) == len(ref)
        for x, y in zip(ref, got):
            np.testing.assert_array_equal(x, y)


def test_pipeline_table_fft_lengths_must_fit_the_band(sw, dev, tmp_path):
    """A table's FFT lengths were timed for its own band: a pipeline takes
    them only while each leaves at most a quarter of a segment to its own
    overlap.  DM 350 with a 1024-point transform fits 1284 MHz (overlap 62);
    at 580 MHz the smear is about 650 samples, so the pipeline keeps the
    table's workgroup shape and falls back to its default length."""
    import json
    table = tmp_path / "t.json"
    table.write_text(json.dumps({"entries": [
        {"antennas": A, "beams": B, "samples": T, "tscrunch": TS, "stokes": "I", "cb_waves": 2,
         "dms": [350.0], "fft_len": [1024], "share_forward": 0}]}))
    p = sw.BeamformerPipeline(A, F, B, T, TS, [350.0], CFREQ, BW, tuning_file=str(table))
    assert p.tile() == {"cb_waves": 2, "fft_len": [1024], "share_forward": False, "source": "table+fft"}
    p.close()
    p = sw.BeamformerPipeline(A, F, B, T, TS, [350.0], 580e6, BW, tuning_file=str(table))
    t = p.tile()
    p.close()
    assert t["cb_waves"] == 2 and t["source"] == "table"
    assert t["fft_len"][0] > 1024
