"""The benchmark configurations through the BeamformerPipeline on the GPU,
checked against the oracle.

* 64 antennas x 64 channels x 8192 samples, 64 beams, tscrunch 4, DM 0,
  Stokes I: one whole gulp, coherent and incoherent beams bit-exact and the
  statistics within 1e-5 of the oracle (CoherentBeamformerTester.cu:43-152,
  IncoherentBeamformerTester.cu:41-103, StatisticsCalculatorTester.cu:25-99).
* 64 antennas x 64 channels x 65536 samples, 800 beams, DM 50 (FFT 16384,
  overlap 10), tscrunch 16: two whole gulps (the second reads its history
  from the first); the chosen form (MFMA beams on the dedisperser's ring)
  equals the plain form (v_dot4 beams on FTPA) bit for bit, and on two
  channels the dedispersed voltages are within 1e-5 relative L2 of the FP64
  overlap-save oracle and the beams equal the oracle's beams of those
  voltages bit for bit.
* 64 antennas x 256 channels, 512 beams, tscrunch 8, coherent DMs {0, 100},
  1024 incoherent DMs 0:1000:0.9765625 (IncoherentDedisperser.cu:67-125) on
  a 53.5 MHz band (256 channels of 208.984375 kHz at 1284 MHz): four
  16384-sample gulps and 512-sample incoherent-dedispersion dispatches
  (shorter than the 5462-sample maximum delay, so the aggregation ring's
  carried overlap is longer than a dispatch), five TDB blocks per coherent
  DM, each equal to the oracle's dedispersion of the GPU's TFB stream on a
  beam subset.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CHBW = 208984.375
CFREQ = 1284e6


def _collect(sw, p, keys=("cb", "ib", "stats")):
    got = {k: [] for k in keys}

    def grab(key):
        def f(blk):
            got[key].append((blk.dm_idx, sw.download_block(blk), blk.utc_offset, blk.tsamp,
                             tuple(blk.dims[:blk.ndims])))
        return f
    p.set_sinks(**{k: grab(k) for k in keys})
    return got


def _pool_map(fn, items, nthreads=16):
    with ThreadPoolExecutor(max_workers=nthreads) as ex:
        return list(ex.map(fn, items))


def test_config1_full_gulp_vs_oracle(sw, orc, dev):
    from skyweaver_amd import DeviceBuffer as DB
    A, F, T, B, TS = 64, 64, 8192, 64, 4
    BW = CHBW * F
    print("config 1: 64 ant x 64 ch x 8192 samples, 64 beams, ts 4, DM 0", flush=True)
    taftp = sw.synth_voltages((T // 256, A, F, 256, 2, 2), seed=0x5EED0001, sigma=16.0, nthreads=16)
    weights = sw.synth_weights((F, B, A, 2), seed=0x5EED1001)
    p = sw.BeamformerPipeline(A, F, B, T, TS, [0.0], CFREQ, BW, stokes="I")
    p.set_weights(weights)
    p.init(A)
    got = _collect(sw, p)
    p(DB.from_host(taftp))
    p.close()
    assert len(got["cb"]) == len(got["ib"]) == len(got["stats"]) == 1
    ftpa = orc.transpose(taftp, A)
    stats = got["stats"][0][1].view(np.float32).reshape(F, 2, A, 4)
    np.testing.assert_allclose(stats.astype(np.float64), orc.statistics(ftpa), rtol=1e-5, atol=1e-6)
    antw = np.ones((1, A), np.float32)
    bmap = np.zeros(B, np.int32)
    scal = orc.update_scalings(stats[..., 1], antw, 1, TS, TS, 24.0)
    _, ib = orc.ib_beamform(ftpa, TS, scal["ib_scaling"], scal["ib_offsets"], antw, 1, "I")

    def cb_chan(f):
        return orc.cb_beamform(ftpa[f:f + 1], weights[f:f + 1], TS, scal["cb_scaling"][f:f + 1],
                               scal["cb_offsets"][f:f + 1], antw, bmap, 1, "I", True)
    cb = np.concatenate(_pool_map(cb_chan, range(F)), axis=1)
    g_cb = got["cb"][0][1].view(np.int8).reshape(cb.shape)
    g_ib = got["ib"][0][1].view(np.int8).reshape(ib.shape)
    np.testing.assert_array_equal(g_ib, ib)
    np.testing.assert_array_equal(g_cb, cb)
    assert np.unique(g_cb).size > 50
    assert got["cb"][0][3] == pytest.approx(TS * F / BW, rel=1e-12)


def test_config2_two_gulps_forms_and_oracle(sw, orc, dev):
    from skyweaver_amd import DeviceBuffer as DB
    A, F, T, B, TS = 64, 64, 65536, 800, 16
    BW = CHBW * F
    TSAMP = F / BW
    N, O = 16384, 10
    dms = [50.0]
    print("config 2: 64 ant x 64 ch x 65536 samples, 800 beams, DM 50 (N 16384), 2 gulps", flush=True)
    gulps = [sw.synth_voltages((T // 256, A, F, 256, 2, 2), seed=0x5EED0002 + 0x100 * i, sigma=16.0, nthreads=16)
             for i in range(2)]
    weights = sw.synth_weights((F, B, A, 2), seed=0x5EED1002)
    d_gulps = [DB.from_host(g) for g in gulps]

    def run(form, batched):
        p = sw.BeamformerPipeline(A, F, B, T, TS, dms, CFREQ, BW, fft_len=[N], form=form, batched=batched)
        p.set_weights(weights)
        p.init(A)
        got = _collect(sw, p)
        for g in d_gulps:
            p(g)
        p.close()
        return got

    print("chosen form", flush=True)
    a = run("mfma", True)
    print("plain form", flush=True)
    b = run("plain", False)
    for key in ("cb", "ib", "stats"):
        assert len(a[key]) == len(b[key]) == 2
        for x, y in zip(a[key], b[key]):
            assert x[2] == y[2]
            np.testing.assert_array_equal(x[1], y[1])
    for d in d_gulps:
        d.close()

    # oracle chain on two channels (bottom and top of the band): the GPU
    # dedisperser run standalone on those channels (same kernel, same FFT
    # length and overlap, history from the first gulp)
    chans = [0, F - 1]
    lowf = CFREQ - BW / 2
    stats = a["stats"][0][1].view(np.float32).reshape(F, 2, A, 4)
    antw = np.ones((1, A), np.float32)
    bmap = np.zeros(B, np.int32)
    scal = orc.update_scalings(stats[..., 1], antw, 1, TS, TS, 24.0)
    for f in chans:
        print(f"oracle chain, channel {f}", flush=True)
        ftpas = [orc.transpose(np.ascontiguousarray(g[:, :, f:f + 1]), A) for g in gulps]   # [1][T][2][A][2]
        cd = sw.CoherentDedisperser(1, A, T, lowf + f * (BW / F), BW / F, TSAMP, dms, fft_len=[N], overlap=[O])
        lat = cd.latency(0)
        assert lat == N - O  # hop 16374 does not divide 65536
        dd_i8, dd_f = [], []
        for x in ftpas:
            d_in = DB.from_host(x)
            d_o = DB((T * 2 * A * 2))
            d_of = DB(T * 2 * A * 8)
            cd.run(d_in, [d_o], [d_of])
            cd.push_history(d_in)
            sw.synchronize()
            dd_i8.append(d_o.download().view(np.int8).reshape(1, T, 2, A, 2))
            dd_f.append(d_of.download().view(np.float32).reshape(T, 2, A, 2))
        cd.close()
        # FP64 overlap-save over the whole two-gulp stream; gulp g emits
        # stream samples [gT - lat, (g+1)T - lat)
        chirp = orc.dm_chirp(dms[0], lowf, BW / F, F, N)[f]
        stream = np.concatenate([x[0] for x in ftpas], 0)
        zero = np.zeros((O, 2, A, 2), np.int8)
        yf, yi = orc.codedisp_channel(np.concatenate([zero, stream]), chirp, N, O, 2 * T, want_float=True,
                                      nthreads=16)
        yf = np.concatenate([np.zeros((lat, 2, A), complex), yf])
        yi = np.concatenate([np.zeros((lat, 2, A, 2), np.int8), yi])
        for gi in range(2):
            ref = yf[gi * T:(gi + 1) * T]
            got_f = dd_f[gi][..., 0] + 1j * dd_f[gi][..., 1]
            rel = np.linalg.norm(got_f - ref) / np.linalg.norm(ref)
            print(f"channel {f} gulp {gi}: relative L2 vs FP64 {rel:.3g}", flush=True)
            assert rel <= 1e-5
            di = dd_i8[gi][0].astype(int) - yi[gi * T:(gi + 1) * T]
            assert np.abs(di).max() <= 1 and np.mean(di != 0) < 1e-4
            # the pipeline's beams of this channel are the oracle's beams of
            # the GPU's dedispersed voltages
            dd = dd_i8[gi]
            sel = slice(f, f + 1)
            _, ib_q = orc.ib_beamform(dd, TS, scal["ib_scaling"][sel], scal["ib_offsets"][sel], antw, 1, "I")
            cb = orc.cb_beamform(dd, weights[sel], TS, scal["cb_scaling"][sel], scal["cb_offsets"][sel], antw,
                                 bmap, 1, "I", True)
            g_cb = a["cb"][gi][1].view(np.int8).reshape(T // TS, F, B)[:, sel]
            g_ib = a["ib"][gi][1].view(np.int8).reshape(1, T // TS, F)[:, :, sel]
            np.testing.assert_array_equal(g_cb, cb)
            np.testing.assert_array_equal(g_ib, ib_q)
    # the blocks carry the filter delay O/2 and the latency
    assert a["cb"][1][2] == pytest.approx(2.0 * 4096 * T / 1712e6 + TSAMP * (O / 2 - (N - O)), rel=1e-9)


def test_config5_shape_incoherent_dedispersion(sw, orc, dev):
    from skyweaver_amd import DeviceBuffer as DB
    A, F, T, B, TS = 64, 256, 16384, 512, 8
    BW = CHBW * F
    cdms = [0.0, 100.0]
    idms = np.arange(1024, dtype=np.float32) * np.float32(0.9765625)
    disp = 512
    ngulps = 4
    print(f"config 5 shape: 64 ant x 256 ch, 512 beams, cDM {{0, 100}}, 1024 iDMs, {ngulps} x 16384 samples",
          flush=True)
    gulps = [sw.synth_voltages((T // 256, A, F, 256, 2, 2), seed=0x5EED0005 + 0x100 * i, sigma=16.0, nthreads=16)
             for i in range(ngulps)]
    weights = sw.synth_weights((F, B, A, 2), seed=0x5EED1005)
    p = sw.BeamformerPipeline(A, F, B, T, TS, cdms, CFREQ, BW, id_plan=[(1, idms), (1, idms)],
                              id_dispatch_size=disp)
    p.set_weights(weights)
    p.init(A)
    tfb = {0: [], 1: []}
    tfb_utc = {}
    tdb = []
    stats = []

    def cb_sink(blk):
        tfb_utc.setdefault(blk.dm_idx, blk.utc_offset)
        tfb[blk.dm_idx].append(sw.download_block(blk).view(np.int8).reshape(blk.dims[0], blk.dims[1], blk.dims[2]))

    def tdb_sink(blk):
        tdb.append((blk.dm_idx, sw.download_block(blk).view(np.int8).reshape(blk.dims[0], blk.dims[1], blk.dims[2]),
                    blk.utc_offset, blk.tsamp, blk.ndms))
    p.set_sinks(cb=cb_sink, stats=lambda blk: stats.append(sw.download_block(blk)))
    p.set_tdb_sink(tdb_sink)
    for g in gulps:
        p(DB.from_host(g))
    p.close()
    assert len(tfb[0]) == len(tfb[1]) == ngulps
    tsamp_cb = TS * F / BW
    d_o, mx = orc.id_delays(idms, sw.channel_frequencies(CFREQ, BW, F), tsamp_cb)
    print(f"max incoherent delay {mx} samples", flush=True)
    assert 5400 < mx < 5500
    nblk = (ngulps * T // TS - mx) // disp
    assert nblk == 5
    assert sorted(t[0] for t in tdb) == [0] * nblk + [1] * nblk
    beams = np.r_[0:8, 250:258, 504:512]
    streams = {k: np.concatenate(v) for k, v in tfb.items()}
    seen = {0: 0, 1: 0}
    for dm_idx, blk, utc, tsb, nd in tdb:
        k = seen[dm_idx]
        seen[dm_idx] += 1
        assert blk.shape == (disp, 1024, B) and nd == 1024
        sub = np.ascontiguousarray(streams[dm_idx][k * disp:k * disp + disp + mx][:, :, beams])
        want = orc.incoherent_dedisperse(sub, d_o, mx, 1, nthreads=16)
        np.testing.assert_array_equal(blk[:, :, beams], want, err_msg=f"cDM {dm_idx} block {k}")
        assert np.unique(blk[:, ::64, beams]).size > 10
        assert tsb == pytest.approx(tsamp_cb, rel=1e-12)
        assert utc == pytest.approx(tfb_utc[dm_idx] + (k * disp + mx) * tsamp_cb, rel=1e-9)
    # coherent DM 0 is the identity: its beams on two channels are the
    # oracle's beams of the transposed input
    st = stats[0].view(np.float32).reshape(F, 2, A, 4)
    antw = np.ones((1, A), np.float32)
    scal = orc.update_scalings(st[..., 1], antw, 1, TS, TS, 24.0)
    for f in (0, F - 1):
        ftpa = orc.transpose(np.ascontiguousarray(gulps[0][:, :, f:f + 1]), A)
        sel = slice(f, f + 1)
        cb = orc.cb_beamform(ftpa, weights[sel], TS, scal["cb_scaling"][sel], scal["cb_offsets"][sel], antw,
                             np.zeros(B, np.int32), 1, "I", True)
        np.testing.assert_array_equal(tfb[0][0][:, sel], cb)
