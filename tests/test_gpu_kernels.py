"""GPU parity of the HIP kernels against the CPU oracle, through the C ABI.

int8 / index outputs must be bit-exact; float outputs of the beamformers
(raw incoherent powers) too, because the kernels keep the oracle's float
operation order; statistics use the reference's 1e-5 relative tolerance."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rng(seed):
    return np.random.default_rng(seed)


def _rand_c8(shape, seed, sigma=32.0):
    x = np.round(_rng(seed).normal(0, sigma, size=shape + (2,)))
    return np.clip(x, -127, 127).astype(np.int8)


def _phase_weights(shape, seed):
    th = _rng(seed).uniform(0, 2 * np.pi, shape)
    return np.stack([np.round(127 * np.cos(th)), np.round(127 * np.sin(th))], -1).astype(np.int8)


def _to_dev(a, dev):
    from skyweaver_amd import DeviceBuffer
    return DeviceBuffer.from_host(np.ascontiguousarray(a))


def _to_host(t):
    return t.download()


def _empty(shape, dtype, dev):
    # poison the output so unwritten elements are caught
    from skyweaver_amd import DeviceBuffer
    return DeviceBuffer.empty(shape, dtype, fill=0x55)


# ------------------------------------------------------------------ transpose
@pytest.mark.parametrize("H,A_in,A_out,F", [(12, 12, 64, 64), (16, 16, 64, 64), (3, 80, 80, 8),
                                            (2, 57, 64, 5), (2, 20, 20, 3), (1, 130, 132, 2)])
def test_transpose_parity(sw, orc, dev, H, A_in, A_out, F):
    print(f"transpose H={H} A_in={A_in} A_out={A_out} F={F}", flush=True)
    n = H * A_in * F * 256 * 2
    i = np.arange(n)
    taftp = np.stack([i % 113, i % 107], -1).astype(np.int8).reshape(H, A_in, F, 256, 2, 2)
    if H > 2:
        taftp = _rand_c8((H, A_in, F, 256, 2), 7)
    d_in = _to_dev(taftp, dev)
    d_out = _empty((F, H * 256, 2, A_out, 2), np.int8, dev)
    sw.transpose(d_in, d_out, H, A_in, A_out, F)
    np.testing.assert_array_equal(_to_host(d_out), orc.transpose(taftp, A_out))


def test_transpose_rejects_bad_shapes(sw, dev):
    x = _empty((4096,), np.int8, dev)
    with pytest.raises(sw.SkyweaverError):
        sw.transpose(x, x, 1, 8, 4, 1)  # nant_in > nant_out


# ----------------------------------------------------------------- statistics
@pytest.mark.parametrize("F,T,A", [(64, 1024, 64), (3, 300, 12), (2, 4096, 80)])
def test_statistics_parity(sw, orc, dev, F, T, A):
    print(f"statistics F={F} T={T} A={A}", flush=True)
    v = _rand_c8((F, T, 2, A), 1337, sigma=20.0)
    v[0, :, 1, 2] = np.clip(v[0, :, 1, 2].astype(int) + 40, -127, 127)
    d_v = _to_dev(v, dev)
    d_st = _empty((F, 2, A, 4), np.float32, dev)
    sw.statistics(d_v, d_st, F, T, A)
    got = _to_host(d_st).astype(np.float64)
    exp = orc.statistics(v)
    # StatisticsCalculatorTester.cu:157-169: relative 1e-5
    np.testing.assert_allclose(got, exp, rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("H,A_in,A_out,F", [(16, 16, 64, 64), (37, 57, 64, 5), (3, 80, 80, 8), (33, 4, 16, 16)])
def test_transpose_fpat_stats_fused(sw, orc, dev, H, A_in, A_out, F):
    """The fused TAFTP -> FPAT transpose + statistics equals the separate
    transpose and statistics (FPAT bytes and float statistics bit-identical)
    and the oracle; full-scale values (-128 everywhere in one column, +127
    in another) reach the kernel's 32-bit partial-sum bounds."""
    print(f"transpose_fpat_stats H={H} A_in={A_in} A_out={A_out} F={F}", flush=True)
    taftp = _rand_c8((H, A_in, F, 256, 2), 99 + H, sigma=30.0)
    taftp[:, 0, 0] = -128
    taftp[:, A_in - 1, F - 1, :, 1] = 127
    T = H * 256
    d_in = _to_dev(taftp, dev)
    d_fpat = _empty((F, 2, A_out, T, 2), np.int8, dev)
    d_st = _empty((F, 2, A_out, 4), np.float32, dev)
    sw.transpose_fpat_stats(d_in, d_fpat, d_st, H, A_in, A_out, F)
    d_fpat2 = _empty((F, 2, A_out, T, 2), np.int8, dev)
    d_st2 = _empty((F, 2, A_out, 4), np.float32, dev)
    sw.transpose_fpat(d_in, d_fpat2, H, A_in, A_out, F)
    sw.statistics_fpat(d_fpat2, d_st2, F, T, A_out)
    fpat = _to_host(d_fpat)
    ftpa = orc.transpose(taftp, A_out)  # [F][T][P][A][2]
    np.testing.assert_array_equal(fpat, ftpa.transpose(0, 2, 3, 1, 4))
    np.testing.assert_array_equal(fpat, _to_host(d_fpat2))
    st = _to_host(d_st)
    np.testing.assert_array_equal(st.view(np.uint32), _to_host(d_st2).view(np.uint32))
    live = np.zeros((F, 2, A_out), bool)
    live[:, :, :A_in] = True
    live[0, :, 0] = False  # constant column: std 0, skew/kurtosis undefined
    exp = orc.statistics(ftpa)
    np.testing.assert_allclose(st[live].astype(np.float64), exp[live], rtol=1e-5, atol=1e-6)
    np.testing.assert_array_equal(st[0, :, 0, :2], np.array([[-128.0, 0.0]] * 2, np.float32))


# --------------------------------------------------------- incoherent beamformer
def _ib_scales(F, S, A, ts, level=32.0):
    ib_scale = level ** 2
    dof = 2 * ts * A * 2
    return (np.full(S * F, ib_scale * np.sqrt(2 * dof) / level, np.float32),
            np.full(S * F, ib_scale * dof, np.float32))


def _run_ib(sw, dev, v, w, scale, offset, ts, stokes):
    F, T, _, A, _ = v.shape
    S = w.shape[0]
    nv = 4 if stokes == "IQUV" else 1
    shp = (S, T // ts, F) + ((4,) if nv == 4 else ())
    d_raw = _empty(shp, np.float32, dev)
    d_out = _empty(shp, np.int8, dev)
    d = sw.dims(F, T, A, 1, ts, stokes, S)
    sw.ib_beamform(_to_dev(v, dev), d_raw, d_out, _to_dev(scale, dev), _to_dev(offset, dev), _to_dev(w, dev), d)
    return _to_host(d_raw), _to_host(d_out)


@pytest.mark.parametrize("stokes", ["I", "Q", "U", "V", "IQUV"])
@pytest.mark.parametrize("S", [1, 2, 4])
def test_ib_parity(sw, orc, dev, stokes, S):
    print(f"ib stokes={stokes} S={S}", flush=True)
    F, T, A, ts = 8, 512, 64, 16
    v = _rand_c8((F, T, 2, A), 100 + S)
    w = np.ones((S, A), np.float32)
    if S > 1:
        w[1:] = (_rng(S).random((S - 1, A)) > 0.5)
    scale, offset = _ib_scales(F, S, A, ts)
    raw, out = _run_ib(sw, dev, v, w, scale, offset, ts, stokes)
    eraw, eout = orc.ib_beamform(v, ts, scale, offset, w, S, stokes)
    np.testing.assert_array_equal(raw, eraw)
    np.testing.assert_array_equal(out, eout)


@pytest.mark.parametrize("A,ts", [(80, 16), (12, 4), (200, 1)])
def test_ib_parity_shapes_and_float_weights(sw, orc, dev, A, ts):
    # non-integer weights make the float sum order-dependent: still bit-exact
    F, T, S = 3, 256, 2
    v = _rand_c8((F, T, 2, A), 5)
    w = (_rng(9).random((S, A)) * 1.7).astype(np.float32)
    scale, offset = _ib_scales(F, S, A, ts)
    raw, out = _run_ib(sw, dev, v, w, scale, offset, ts, "I")
    eraw, eout = orc.ib_beamform(v, ts, scale, offset, w, S, "I")
    np.testing.assert_array_equal(raw, eraw)
    np.testing.assert_array_equal(out, eout)


@pytest.mark.parametrize("A,ts,S", [(128, 32, 1), (64, 32, 3)])
def test_ib_parity_exact_path_overflow(sw, orc, dev, A, ts, S):
    # Stokes I with 0/1 weights takes the integer path; full-scale voltages
    # push the totals past 2^24 on some outputs, where the float chain rounds:
    # those waves must fall back to the oracle's float order
    print(f"ib overflow A={A} ts={ts} S={S}", flush=True)
    F, T = 2, 64 * 32 * 2
    v = _rand_c8((F, T, 2, A), 7, sigma=8.0)    # totals ~1e6: exact path
    v[0, : T // 2] = -128                       # channel 0, first half: totals ~1e8
    w = np.ones((S, A), np.float32)
    if S > 1:
        w[1:] = (_rng(11).random((S - 1, A)) > 0.5)
    scale, offset = _ib_scales(F, S, A, ts)
    raw, out = _run_ib(sw, dev, v, w, scale, offset, ts, "I")
    eraw, eout = orc.ib_beamform(v, ts, scale, offset, w, S, "I")
    assert (eraw >= 2 ** 24).any() and (eraw < 2 ** 24).any()
    np.testing.assert_array_equal(raw, eraw)
    np.testing.assert_array_equal(out, eout)


# --------------------------------------------------------- coherent beamformer
def _cb_scales(F, S, A, ts, level=32.0):
    # CoherentBeamformerTester.cu:224-242
    scale = (127.0 * level * np.sqrt(A)) ** 2
    dof = 2 * ts * 1 * 2
    return (np.full(S * F, scale * np.sqrt(2 * dof) / level, np.float32),
            np.full(S * F, scale * dof, np.float32))


def _run_cb(sw, dev, v, w, bmap, antw, ts, stokes, ib_sub, form, ib_raw=None):
    F, T, _, A, _ = v.shape
    B = w.shape[1]
    S = antw.shape[0]
    cbs, cbo = _cb_scales(F, S, A, ts)
    ibs, ibo = _ib_scales(F, S, A, ts)
    nv = 4 if stokes == "IQUV" else 1
    d_v = _to_dev(v, dev)
    d_antw = _to_dev(antw, dev)
    if ib_raw is None:
        shp = (S, T // ts, F) + ((4,) if nv == 4 else ())
        ib_raw = _empty(shp, np.float32, dev)
        ib_q = _empty(shp, np.int8, dev)
        sw.ib_beamform(d_v, ib_raw, ib_q, _to_dev(ibs, dev), _to_dev(ibo, dev), d_antw, sw.dims(F, T, A, B, ts, stokes, S))
    d_out = _empty((T // ts, F, B) + ((4,) if nv == 4 else ()), np.int8, dev)
    d = sw.dims(F, T, A, B, ts, stokes, S, ib_sub)
    sw.cb_beamform(d_v, _to_dev(w, dev), _to_dev(cbs, dev), _to_dev(cbo, dev), _to_dev(bmap, dev), ib_raw, d_out, d,
                   form=form)
    return _to_host(d_out), (cbs, cbo)


@pytest.mark.parametrize("form", ["plain", "mfma"])
@pytest.mark.parametrize("stokes", ["I", "Q", "U", "V", "IQUV"])
def test_cb_parity_stokes(sw, orc, dev, form, stokes):
    # CoherentBeamformerTester.cu:211-361: 2 beamsets, beam 0 -> beamset 1
    print(f"cb form={form} stokes={stokes}", flush=True)
    F, T, A, B, ts, S = 4, 256, 64, 64, 16, 2
    v = _rand_c8((F, T, 2, A), 200)
    w = _phase_weights((F, B, A), 201)
    bmap = np.zeros(B, np.int32)
    bmap[0] = 1
    antw = np.ones((S, A), np.float32)
    out, (cbs, cbo) = _run_cb(sw, dev, v, w, bmap, antw, ts, stokes, True, form)
    exp = orc.cb_beamform(v, w, ts, cbs, cbo, antw, bmap, S, stokes, True)
    np.testing.assert_array_equal(out, exp)


@pytest.mark.parametrize("form", ["plain", "mfma"])
@pytest.mark.parametrize("A,B,ts,ib_sub", [(64, 96, 4, True), (80, 37, 8, True), (32, 33, 32, False),
                                            (64, 64, 1, True), (16, 40, 2, False)])
def test_cb_parity_shapes(sw, orc, dev, form, A, B, ts, ib_sub):
    print(f"cb form={form} A={A} B={B} ts={ts} ib_sub={ib_sub}", flush=True)
    F, T, S = 3, 256, 3
    v = _rand_c8((F, T, 2, A), 300 + A)
    w = _phase_weights((F, B, A), 301 + B)
    bmap = (np.arange(B) * 7 % S).astype(np.int32)
    antw = (_rng(3).random((S, A)) > 0.3).astype(np.float32)
    antw[0] = 1
    out, (cbs, cbo) = _run_cb(sw, dev, v, w, bmap, antw, ts, "I", ib_sub, form)
    exp = orc.cb_beamform(v, w, ts, cbs, cbo, antw, bmap, S, "I", ib_sub)
    np.testing.assert_array_equal(out, exp)


@pytest.mark.parametrize("form", ["plain", "mfma"])
@pytest.mark.parametrize("stokes,ts,ib_sub", [("I", 1, True), ("IQUV", 1, True), ("IQUV", 2, False),
                                               ("I", 32, True), ("IQUV", 32, True)])
def test_cb_parity_partial_chunk(sw, orc, dev, form, stokes, ts, ib_sub):
    # T = 160 is not a multiple of the MFMA form's 128-sample workgroup chunk:
    # the last workgroup has idle waves that still take part in weight staging
    print(f"cb partial chunk form={form} stokes={stokes} ts={ts} ib_sub={ib_sub}", flush=True)
    F, T, A, B, S = 2, 160, 48, 50, 2
    v = _rand_c8((F, T, 2, A), 500 + ts)
    w = _phase_weights((F, B, A), 501 + ts)
    bmap = (np.arange(B) % S).astype(np.int32)
    antw = np.ones((S, A), np.float32)
    antw[1, ::3] = 0
    out, (cbs, cbo) = _run_cb(sw, dev, v, w, bmap, antw, ts, stokes, ib_sub, form)
    exp = orc.cb_beamform(v, w, ts, cbs, cbo, antw, bmap, S, stokes, ib_sub)
    np.testing.assert_array_equal(out, exp)


@pytest.mark.parametrize("form", ["plain", "mfma"])
def test_cb_parity_extreme_int8(sw, orc, dev, form):
    # -128 in weights and voltages (the MFMA form's ~x + 1 negation path) and
    # full-scale values whose powers round in float
    F, T, A, B, ts, S = 2, 128, 64, 64, 8, 1
    v = _rng(401).integers(-128, 128, (F, T, 2, A, 2)).astype(np.int8)
    v[0, :5] = -128
    w = _rng(402).integers(-128, 128, (F, B, A, 2)).astype(np.int8)
    w[:, :3, :, 1] = -128
    bmap = np.zeros(B, np.int32)
    antw = np.ones((S, A), np.float32)
    out, (cbs, cbo) = _run_cb(sw, dev, v, w, bmap, antw, ts, "IQUV", True, form)
    exp = orc.cb_beamform(v, w, ts, cbs, cbo, antw, bmap, S, "IQUV", True)
    np.testing.assert_array_equal(out, exp)


def test_cb_plain_equals_mfma_at_scale(sw, dev):
    # 800 beams (25 tiles of 32), 4096 samples: the two forms agree bit for bit
    print("cb plain vs mfma at 800 beams", flush=True)
    F, T, A, B, ts, S = 4, 4096, 64, 800, 16, 1
    v = sw.synth_voltages((F, T, 2, A, 2), seed=0x5EED0002, sigma=16.0)
    w = sw.synth_weights((F, B, A, 2), seed=0x5EED1002)
    bmap = np.zeros(B, np.int32)
    antw = np.ones((S, A), np.float32)
    o1, _ = _run_cb(sw, dev, v, w, bmap, antw, ts, "I", True, "plain")
    o2, _ = _run_cb(sw, dev, v, w, bmap, antw, ts, "I", True, "mfma")
    np.testing.assert_array_equal(o1, o2)
    assert np.unique(o1).size > 50  # non-degenerate output


@pytest.mark.parametrize("F,R,C", [(2, 128, 4096), (3, 4096, 128), (2, 130, 66), (1, 7, 9), (2, 64, 68)])
def test_corner_turn(sw, dev, F, R, C):
    # [F][R][C] -> [F][C][R] of 16-bit (complex int8) elements, both kernel paths
    x = _rng(F * R + C).integers(-128, 128, (F, R, C, 2)).astype(np.int8)
    d_x = _to_dev(x, dev)
    d_y = _empty((F, C, R, 2), np.int8, dev)
    sw.corner_turn(d_x, d_y, F, R, C)
    np.testing.assert_array_equal(_to_host(d_y), x.transpose(0, 2, 1, 3))


@pytest.mark.parametrize("stokes,A,ts", [("I", 64, 16), ("I", 128, 8), ("IQUV", 80, 16), ("I", 16, 1),
                                         ("IQUV", 32, 32)])
def test_cb_workgroup_shapes_identical(sw, orc, dev, stokes, A, ts):
    """The MFMA beamformer with 2, 4 and 8 waves (of 32 time rows) per
    workgroup (the tuning table's beamformer tile) gives the same beams,
    equal to the oracle's; T = 800 leaves every shape a partial last
    workgroup."""
    print(f"cb workgroup shapes stokes={stokes} A={A} ts={ts}", flush=True)
    F, T, B, S = 2, 800, 72, 2
    v = _rand_c8((F, T, 2, A), 600 + A + ts)
    w = _phase_weights((F, B, A), 601 + A)
    bmap = (np.arange(B) % S).astype(np.int32)
    antw = np.ones((S, A), np.float32)
    antw[1, ::4] = 0
    cbs, cbo = _cb_scales(F, S, A, ts)
    ibs, ibo = _ib_scales(F, S, A, ts)
    nv = 4 if stokes == "IQUV" else 1
    d_v, d_antw = _to_dev(v, dev), _to_dev(antw, dev)
    shp = (S, T // ts, F) + ((4,) if nv == 4 else ())
    ib_raw, ib_q = _empty(shp, np.float32, dev), _empty(shp, np.int8, dev)
    sw.ib_beamform(d_v, ib_raw, ib_q, _to_dev(ibs, dev), _to_dev(ibo, dev), d_antw, sw.dims(F, T, A, B, ts, stokes, S))
    d = sw.dims(F, T, A, B, ts, stokes, S, True)
    packed = _empty((F * B * A * 4,), np.int8, dev)
    sw.cb_pack_weights(_to_dev(w, dev), packed, F, B, A)
    outs = []
    for waves in (2, 4, 8):
        d_out = _empty((T // ts, F, B) + ((4,) if nv == 4 else ()), np.int8, dev)
        sw.cb_beamform_packed_tile(d_v, packed, _to_dev(cbs, dev), _to_dev(cbo, dev), _to_dev(bmap, dev), ib_raw,
                                   d_out, d, waves)
        outs.append(_to_host(d_out))
    exp = orc.cb_beamform(v, w, ts, cbs, cbo, antw, bmap, S, stokes, True)
    for o in outs:
        np.testing.assert_array_equal(o, exp)
