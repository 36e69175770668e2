"""GPU incoherent dedispersion against the oracle (bit-exact int8) and the
reference's closed forms (cpp/skyweaver/test/src/IncoherentDedisperserTester.cu:
all-ones input gives trunc(sqrt(F ts)); DM 0 has zero delays; too few samples
is an error)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(sw, idd, tfb, nvec, layout="tfb"):
    T, F = tfb.shape[:2]
    B = tfb.shape[2]
    nout = idd.nout(T)
    x = tfb if layout == "tfb" else tfb.swapaxes(0, 1)
    d_in = sw.DeviceBuffer.from_host(np.ascontiguousarray(x))
    shape = (nout, idd.ndms, B) + ((nvec,) if nvec > 1 else ())
    d_out = sw.DeviceBuffer.empty(shape, np.int8, fill=0x55)
    idd.dedisperse(d_in, T, B, nvec, d_out, layout=layout)
    sw.synchronize()
    return d_out.download()


def _rand(shape, seed, sigma=20.0):
    x = np.round(np.random.default_rng(seed).normal(0, sigma, shape))
    return np.clip(x, -128, 127).astype(np.int8)


@pytest.mark.parametrize("F,ts", [(64, 1), (64, 4), (37, 1)])
def test_all_ones(sw, dev, F, ts):
    idd = sw.IncoherentDedisperser(F, 1284e6, 13.375e6, 16, [0.0, 10.0, 100.0], tscrunch=ts)
    T = idd.max_delay + ts * 40
    out = _run(sw, idd, np.ones((T, F, 64), np.int8), 1)
    assert (out == int(np.sqrt(F * ts))).all()


def test_dm0_identity_sum(sw, orc, dev):
    F, B = 32, 48
    idd = sw.IncoherentDedisperser(F, 1284e6, 13.375e6, 16, [0.0])
    assert idd.max_delay == 0 and not idd.delays.any()
    tfb = _rand((100, F, B), 1)
    out = _run(sw, idd, tfb, 1)
    want = np.clip(tfb.astype(np.float32).sum(1) / np.float32(np.sqrt(F)), -128, 127).astype(np.int8)
    np.testing.assert_array_equal(out[:, 0], want)


CASES = [
    # F, B, nvec, ts, DMs, nout, band (centre, bw, cb_ts)
    (64, 64, 1, 1, np.arange(0, 200, 7.5, dtype=np.float32), 70, (1284e6, 13.375e6, 16)),
    (64, 80, 1, 2, np.float32([0, 5, 20, 50, 100, 180, 260, 350]), 45, (1284e6, 13.375e6, 16)),
    (64, 16, 4, 1, np.arange(0, 300, 3.0, dtype=np.float32), 33, (1284e6, 13.375e6, 16)),
    (128, 32, 1, 4, np.arange(0, 100, 2.5, dtype=np.float32), 20, (1284e6, 26.75e6, 16)),   # 512 adds/sum
    (300, 16, 1, 1, np.arange(0, 50, 1.0, dtype=np.float32), 19, (1284e6, 62.6953125e6, 8)),
    (256, 32, 1, 1, np.arange(1024, dtype=np.float32) * np.float32(0.9765625), 24, (1284e6, 53.5e6, 8)),
    (64, 72, 1, 2, np.arange(0, 40, 2.5, dtype=np.float32), 37, (1284e6, 13.375e6, 16)),      # 18-word rows
    (32, 300, 1, 1, np.arange(0, 80, 4.0, dtype=np.float32), 21, (1284e6, 6.6875e6, 16)),     # 75 words, 2 chunks
    (4, 48, 1, 1, np.arange(0, 400, 25.0, dtype=np.float32), 40, (1284e6, 13.375e6, 16)),    # one quad, part chunk
    (8, 144, 1, 2, np.arange(0, 300, 10.0, dtype=np.float32), 29, (1284e6, 26.75e6, 8)),     # two quads, 128 + 16 B
    (16, 64, 4, 1, np.float32([0, 1000]), 17, (1284e6, 53.5e6, 8)),                          # one DM per tile
]


@pytest.mark.parametrize("valu", [False, True], ids=["auto", "valu"])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_vs_oracle(sw, orc, dev, case, valu):
    F, B, nvec, ts, dms, nout, (fc, bw, cbts) = CASES[case]
    idd = sw.IncoherentDedisperser(F, fc, bw, cbts, dms, tscrunch=ts, valu_kernel=valu)
    if valu or F % 4 or 16 * ts > 128:
        want_kernel = "idisp_kernel"
    else:
        want_kernel = "idisp_dma_kernel" if 32 * ts <= 64 else "idisp_mfma_kernel"
    assert idd.kernel == want_kernel
    f = sw.channel_frequencies(fc, bw, F)
    d_o, mx_o = orc.id_delays(dms, f, cbts / (bw / F))
    np.testing.assert_array_equal(idd.delays, d_o)
    assert idd.max_delay == mx_o
    T = idd.max_delay + ts * nout
    shape = (T, F, B) + ((nvec,) if nvec > 1 else ())
    tfb = _rand(shape, 100 + case, sigma=40.0)
    tfb.reshape(-1)[::997] = -128  # extremes
    tfb.reshape(-1)[1::991] = 127
    want = orc.incoherent_dedisperse(tfb, d_o, mx_o, ts)
    for layout in ("tfb", "ftb"):
        out = _run(sw, idd, tfb, nvec, layout)
        assert out.shape == want.shape
        np.testing.assert_array_equal(out, want, err_msg=layout)


def test_saturating_sums(sw, orc, dev):
    # every input at the extremes: the quantised output clamps at -128 / 127
    F, B = 64, 32
    idd = sw.IncoherentDedisperser(F, 1284e6, 13.375e6, 16, [0.0, 30.0], tscrunch=2)
    T = idd.max_delay + 2 * 17
    tfb = np.where(np.random.default_rng(3).random((T, F, B)) < 0.5, -128, 127).astype(np.int8)
    out = _run(sw, idd, tfb, 1)
    np.testing.assert_array_equal(out, orc.incoherent_dedisperse(tfb, idd.delays, idd.max_delay, 2))


def test_errors(sw, dev):
    idd = sw.IncoherentDedisperser(64, 1284e6, 13.375e6, 16, [0.0, 100.0], tscrunch=4)
    assert idd.max_delay > 0
    buf = sw.DeviceBuffer.empty((idd.max_delay + 64, 64, 64), np.int8)
    out = sw.DeviceBuffer.empty((64, 2, 64), np.int8)
    with pytest.raises(sw.SkyweaverError, match="Fewer than max_sample_delay"):
        idd.dedisperse(buf, idd.max_delay, 64, 1, out)
    with pytest.raises(sw.SkyweaverError, match="multiple of tscrunch"):
        idd.dedisperse(buf, idd.max_delay + 6, 64, 1, out)
    with pytest.raises(sw.SkyweaverError):
        idd.dedisperse(buf, idd.max_delay + 8, 6, 1, out)  # 6 beams: rows not whole 32-bit words


def test_pipeline_aggregation(sw, orc, dev):
    """Pushed TFB blocks of uneven sizes are aggregated, dispatched every
    dispatch_size samples with max_delay samples of overlap, and each TDB
    block equals the oracle on the corresponding window of the stream."""
    F, B, fc, bw, cbts = 64, 32, 1284e6, 13.375e6, 16
    plan = [(0.0, 1, np.arange(0, 60, 6.0, dtype=np.float32)),
            (50.0, 2, np.float32([40.0, 50.0, 60.0]))]
    pipe = sw.IncoherentDedispersionPipeline(F, fc, bw, cbts, B, 1, plan, dispatch_size=300)
    got = {0: [], 1: []}

    def sink(blk):
        a = sw.download_block(blk).view(np.int8).reshape(blk.dims[0], blk.dims[1], blk.dims[2])
        got[blk.dm_idx].append((a, blk.utc_offset, blk.tsamp, [blk.dms[i] for i in range(blk.ndms)],
                                blk.reference_dm))

    pipe.set_sink(sink)
    tsamp = cbts / (bw / F)
    stream = _rand((1400, F, B), 77)
    sizes = [100, 257, 3, 400, 640]
    for blk_idx in (0, 1):
        pos = 0
        for n in sizes:
            chunk = stream[pos:pos + n]
            if blk_idx == 0:
                pipe.push(0, sw.DeviceBuffer.from_host(chunk), n, 10.0 + pos * tsamp, tsamp)
            else:
                pipe.push(1, chunk, n, 10.0 + pos * tsamp, tsamp)  # host input
            pos += n
    f = sw.channel_frequencies(fc, bw, F)
    for bi, (cdm, ts, dms) in enumerate(plan):
        info = pipe.info(bi)
        md, disp = info["max_delay"], info["dispatch_size"]
        assert disp == 300 - 300 % ts
        d_o, mx_o = orc.id_delays(dms, f, tsamp)
        assert md == mx_o
        nblocks = (len(stream) - md) // disp
        assert len(got[bi]) == nblocks and nblocks >= 3
        for k, (a, utc, tsb, bdms, rdm) in enumerate(got[bi]):
            window = stream[k * disp:k * disp + disp + md]
            np.testing.assert_array_equal(a, orc.incoherent_dedisperse(window, d_o, mx_o, ts))
            assert abs(utc - (10.0 + (k * disp + md) * tsamp)) < 1e-9
            assert tsb == tsamp * ts and rdm == cdm
            np.testing.assert_array_equal(np.float32(bdms), dms)
        assert info["filled"] == len(stream) - nblocks * disp
