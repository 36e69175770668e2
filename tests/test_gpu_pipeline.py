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
    prev = None
    exact = []
    for gi, g in enumerate(gulps):
        ftpa = orc.transpose(g, A)
        st = got["stats"][gi][1].view(np.float32).reshape(F, 2, A, 4)
        np.testing.assert_allclose(st.astype(np.float64), orc.statistics(ftpa), rtol=1e-5, atol=1e-6)
        for d, dm in enumerate(dms):
            if dm == 0.0:
                dd = ftpa
            else:
                O = 10
                chirp = orc.dm_chirp(dm, lowf, BW / F, F, fft_len[d])
                hist = np.zeros((F, O, 2, A, 2), np.int8) if prev is None else prev[:, T - O:]
                dd = np.stack([orc.codedisp_channel(np.concatenate([hist[f], ftpa[f]]), chirp[f], fft_len[d], O, T,
                                                    want_float=False, nthreads=16)[1] for f in range(F)])
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
        prev = ftpa
    print(f"DM 50 coherent beams exact fraction: {exact}", flush=True)
    # metadata: the DM-50 blocks are delayed by the filter delay O/2 samples
    assert got["cb"][1][2] == pytest.approx(TSAMP * 5, rel=1e-9)
    assert got["cb"][1][3] == pytest.approx(TSAMP * TS, rel=1e-9)


def test_pipeline_forms_identical_and_host_input(sw, dev):
    """plain (v_dot4, one DM per launch) == chosen (MFMA, batched) bit for bit,
    and a host-memory gulp gives the same result as a device-resident one."""
    from skyweaver_amd import DeviceBuffer as DB
    H = T // 256
    g = sw.synth_voltages((H, A, F, 256, 2, 2), seed=7, sigma=16.0)
    weights = sw.synth_weights((F, B, A, 2), seed=8)
    bmap = (np.arange(B) % 2).astype(np.int32)
    antw = np.ones((2, A), np.float32)
    antw[1, 40:] = 0
    dms, fft_len = [0.0, 20.0, 50.0], [1024, 1024, 1024]
    a = _run_pipeline(sw, [DB.from_host(g)], weights, dms, fft_len, "mfma", True, bmap, antw, "IQUV")
    b = _run_pipeline(sw, [g], weights, dms, fft_len, "plain", False, bmap, antw, "IQUV")
    for key in ("cb", "ib", "stats"):
        assert len(a[key]) == len(b[key])
        for x, y in zip(a[key], b[key]):
            np.testing.assert_array_equal(x[1], y[1])


def test_pipeline_rejects_wrong_gulp_size(sw, dev):
    p = sw.BeamformerPipeline(A, F, B, T, TS, [0.0], CFREQ, BW)
    p.set_weights(np.zeros((F, B, A, 2), np.int8))
    p.init(A)
    with pytest.raises(sw.SkyweaverError, match="Unexpected buffer size"):
        p(np.zeros(100, np.int8))
