"""GPU coherent dedispersion against the FP64 overlap-save oracle and the
reference's golden vectors.

The dedisperser's segments lie on one grid over the stream of gulps; a gulp
g emits stream samples [gT - L, (g+1)T - L) with the plan's latency L (0
when the hop N - O divides T).  The oracle transforms the whole stream with
the same grid, and each gulp is compared with its window of the oracle's
output (zeros before the stream start).

Tolerances: dedispersed voltages (y/N before quantisation) within 1e-5
relative (L2 norm over each DM's output) of the FP64 oracle; int8 outputs
within +-1 (modulo the int8 wrap) of the oracle, with mismatches confined to
the rare values that lie within FP32 rounding of a half-integer.  The golden
pair is checked at +-1, the reference test's own tolerance
(CoherentDedisperserTester.cu:195-203)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TSAMP = 4096 / 856e6
BAND_BW = 13.375e6
CHAN_BW = BAND_BW / 64
F_LOW = 1284e6 - BAND_BW / 2


def _wrapdiff(a, b):
    return ((a.astype(np.int64) - b.astype(np.int64) + 128) % 256) - 128


def _check(got_i8, got_f, exp_i8, exp_f, label):
    gf = got_f[..., 0].astype(np.float64) + 1j * got_f[..., 1]
    rel = np.linalg.norm(gf - exp_f) / np.linalg.norm(exp_f)
    mx = np.abs(gf - exp_f).max()
    d = _wrapdiff(got_i8, exp_i8)
    frac = np.mean(d != 0)
    print(f"{label}: rel L2 {rel:.3e} max abs {mx:.3e} int8 mismatch {frac:.2e}", flush=True)
    assert rel <= 1e-5, label
    assert np.abs(d).max() <= 1, label
    assert frac <= 1e-3, label
    return rel


def _window(ei, ef, start, T):
    """Samples [start, start + T) of stream outputs [F][Ts][...] (zeros before 0)."""
    F = ei.shape[0]
    pad = max(0, -start)
    zi = np.zeros((F, pad) + ei.shape[2:], ei.dtype)
    zf = np.zeros((F, pad) + ef.shape[2:], ef.dtype)
    a = max(0, start)
    return (np.concatenate([zi, ei[:, a:start + T]], 1), np.concatenate([zf, ef[:, a:start + T]], 1))


def _oracle_stream(orc, hist, cur, chirps, N, O, nthreads=16):
    """hist [F][O][P][A][2], cur [F][T][P][A][2] -> (int8, complex) [F][T][P][A]."""
    F, T = cur.shape[:2]
    outs_i, outs_f = [], []
    for f in range(F):
        x = np.concatenate([hist[f, hist.shape[1] - O:], cur[f]], 0)
        yf, yi = orc.codedisp_channel(x, chirps[f], N, O, T, nthreads=nthreads)
        outs_i.append(yi)
        outs_f.append(yf)
    return np.stack(outs_i), np.stack(outs_f)


def test_codedisp_golden_pair(sw, orc, dev, golden_dir):
    from skyweaver_amd import DeviceBuffer as DB
    inp = np.fromfile(os.path.join(golden_dir, "codedisp_input_DM1000_1chan.bin"), np.int8).reshape(1, 1024, 2, 64, 2)
    gold = np.fromfile(os.path.join(golden_dir, "codedisp_output_DM1000_1chan.bin"), np.int8).reshape(1, 446, 2, 64, 2)
    O, T = 578, 446
    plan = sw.CoherentDedisperser(1, 64, T, 856e6, BAND_BW / 64, TSAMP, [1000.0], fft_len=[1024], overlap=[O])
    hist = DB.from_host(np.ascontiguousarray(inp[:, :O]))
    cur = DB.from_host(np.ascontiguousarray(inp[:, O:]))
    out = DB.empty((1, T, 2, 64, 2), np.int8, fill=0x55)
    outf = DB.empty((1, T, 2, 64, 2), np.float32)
    plan.set_history(hist)
    plan.run(cur, [out], [outf])
    got = out.download()
    d = _wrapdiff(got, gold)
    print(f"golden: exact {np.mean(d == 0):.5f} max |d| {np.abs(d).max()}", flush=True)
    assert np.abs(d).max() <= 1
    chirp = orc.dm_chirp(1000.0, 856e6, BAND_BW / 64, 64, 1024)[:1]
    ei, ef = _oracle_stream(orc, inp[:, :O], inp[:, O:], chirp, 1024, O)
    _check(got, outf.download(), ei, ef, "golden DM1000 N1024")


@pytest.mark.parametrize("dm,N", [(50.0, 16384), (50.0, 1024), (350.0, 32768), (100.0, 4096)])
def test_codedisp_vs_fp64_oracle(sw, orc, dev, dm, N):
    from skyweaver_amd import DeviceBuffer as DB
    F, A, T = 2, 64, 40000
    cur = sw.synth_voltages((F, T, 2, A, 2), seed=0x5EED0002, sigma=16.0)
    plan = sw.CoherentDedisperser(F, A, T, F_LOW, CHAN_BW, TSAMP, [dm], fft_len=[N])
    _, O = plan.info(0)
    lat = plan.latency(0)
    assert lat == (0 if T % (N - O) == 0 else N - O)
    print(f"codedisp dm={dm} N={N} O={O} latency={lat}", flush=True)
    d_cur = DB.from_host(cur)
    out = DB.empty(cur.shape, np.int8, fill=0x55)
    outf = DB.empty(cur.shape, np.float32)
    plan.run(d_cur, [out], [outf])
    chirp = orc.dm_chirp(dm, F_LOW, CHAN_BW, F, N)
    ei, ef = _oracle_stream(orc, np.zeros((F, O, 2, A, 2), np.int8), cur, chirp, N, O)
    ei, ef = _window(ei, ef, -lat, T)
    _check(out.download(), outf.download(), ei, ef, f"DM {dm} N {N}")


@pytest.mark.parametrize("N,O", [(4096, 1024), (8192, 1200), (32768, 1100)])
def test_codedisp_large_overlap(sw, orc, dev, N, O):
    """Overlap 1024 (the first and last position rows of a register unit
    dropped whole) and overlaps above it (more rows dropped: the kernels'
    general store path): against the oracle, and int8 outputs identical
    with and without the float ring."""
    from skyweaver_amd import DeviceBuffer as DB
    F, A, T, dm = 2, 64, 40000, 50.0
    cur = sw.synth_voltages((F, T, 2, A, 2), seed=0x5EED0012, sigma=16.0)
    plan = sw.CoherentDedisperser(F, A, T, F_LOW, CHAN_BW, TSAMP, [dm], fft_len=[N], overlap=[O])
    plan_i = sw.CoherentDedisperser(F, A, T, F_LOW, CHAN_BW, TSAMP, [dm], fft_len=[N], overlap=[O])
    assert plan.info(0)[1] == O
    lat = plan.latency(0)
    print(f"codedisp large overlap N={N} O={O} latency={lat}", flush=True)
    d_cur = DB.from_host(cur)
    out = DB.empty(cur.shape, np.int8, fill=0x55)
    outf = DB.empty(cur.shape, np.float32)
    out_i = DB.empty(cur.shape, np.int8, fill=0x33)
    plan.run(d_cur, [out], [outf])
    plan_i.run(d_cur, [out_i])
    got = out.download()
    np.testing.assert_array_equal(out_i.download(), got)
    chirp = orc.dm_chirp(dm, F_LOW, CHAN_BW, F, N)
    ei, ef = _oracle_stream(orc, np.zeros((F, O, 2, A, 2), np.int8), cur, chirp, N, O)
    ei, ef = _window(ei, ef, -lat, T)
    _check(got, outf.download(), ei, ef, f"DM {dm} N {N} O {O}")


@pytest.mark.parametrize("N", [512, 4096, 32768])
def test_codedisp_physical_chirp_vs_fp64_oracle(sw, orc, dev, N):
    """The physical bin-to-frequency mapping (channel-centred, upper bins
    negative) through the LDS, register and two-half register kernels, against
    the FP64 oracle with the same mapping."""
    from skyweaver_amd import DeviceBuffer as DB
    F, A, T, dm = 2, 64, 40000, 100.0
    cur = sw.synth_voltages((F, T, 2, A, 2), seed=0x5EED0042, sigma=16.0)
    plan = sw.CoherentDedisperser(F, A, T, F_LOW, CHAN_BW, TSAMP, [dm], fft_len=[N], chirp="physical")
    _, O = plan.info(0)
    lat = plan.latency(0)
    print(f"codedisp physical chirp dm={dm} N={N} O={O} latency={lat}", flush=True)
    out = DB.empty(cur.shape, np.int8, fill=0x55)
    outf = DB.empty(cur.shape, np.float32)
    plan.run(DB.from_host(cur), [out], [outf])
    chirp = orc.dm_chirp(dm, F_LOW, CHAN_BW, F, N, physical=True)
    ei, ef = _oracle_stream(orc, np.zeros((F, O, 2, A, 2), np.int8), cur, chirp, N, O)
    ei, ef = _window(ei, ef, -lat, T)
    _check(out.download(), outf.download(), ei, ef, f"physical chirp DM {dm} N {N}")


def test_codedisp_physical_chirp_recompresses_a_dispersed_pulse(sw, dev):
    """A pulse dispersed in the time domain by the cold-plasma law over ~580
    samples of one channel (DM 1000 at 856 MHz) comes out of the GPU
    dedisperser with the physical chirp gathered into a few samples at its
    centre-frequency arrival time."""
    from skyweaver_amd import DeviceBuffer as DB
    from test_oracle import _dispersed_pulse
    dm, low, bwc, N, O, A = 1000.0, 856e6, 856e6 / 4096, 4096, 580, 4
    T, n_c = 8192, 3000
    x, width = _dispersed_pulse(n_c, T, dm, low + 0.5 * bwc, bwc, 40.0)
    cur = np.zeros((1, T, 2, A, 2), np.int8)
    cur[0, :, 1, 2] = x[:, 0, 0]                      # pol 1, antenna 2; the other columns stay zero
    plan = sw.CoherentDedisperser(1, A, T, low, bwc, 1 / bwc, [dm], fft_len=[N], overlap=[O], chirp="physical")
    lat = plan.latency(0)
    outf = DB.empty(cur.shape, np.float32)
    plan.run(DB.from_host(cur), [DB.empty(cur.shape, np.int8)], [outf])
    y = outf.download()
    p = y[0, :, 1, 2, 0].astype(np.float64) ** 2 + y[0, :, 1, 2, 1].astype(np.float64) ** 2
    j = n_c + O // 2 + lat                            # output i is input i - lat - O/2 (the DM 0 identity's shift)
    assert j + 5 <= T
    print(f"pulse: width {width} latency {lat} peak {int(np.argmax(p))} expected {j} "
          f"fraction {p[j - 4:j + 5].sum() / p.sum():.4f}", flush=True)
    assert abs(int(np.argmax(p)) - j) <= 1
    assert p[j - 4:j + 5].sum() / p.sum() > 0.9
    assert np.abs(y[0, :, 0]).max() == 0 and np.abs(y[0, :, 1, :2]).max() == 0


def test_codedisp_dm0_is_identity(sw, dev):
    from skyweaver_amd import DeviceBuffer as DB
    F, A, T = 3, 64, 9000
    cur = sw.synth_voltages((F, T, 2, A, 2), seed=11, sigma=30.0)
    plan = sw.CoherentDedisperser(F, A, T, F_LOW, CHAN_BW, TSAMP, [0.0, 0.0], fft_len=[4096, 16384],
                                  overlap=[0, 20])
    outs = [DB.empty(cur.shape, np.int8, fill=0x55) for _ in range(2)]
    plan.run(DB.from_host(cur), outs)
    # stream sample j is input j - O/2, emitted with the plan's latency
    for k, O in enumerate((0, 20)):
        lat = plan.latency(k)
        sh = lat + O // 2
        exp = np.concatenate([np.zeros((F, min(sh, T), 2, A, 2), np.int8), cur[:, :max(0, T - sh)]], 1)
        np.testing.assert_array_equal(outs[k].download(), exp)


@pytest.mark.parametrize("mode", ["own_overlap", "paired_overlap", "share_forward"])
def test_codedisp_uneven_batch_and_history(sw, orc, dev, mode):
    """Eight DMs with uneven FFT lengths over two gulps: batched == one DM per
    launch bit for bit, and both match the oracle on the continuous stream.
    paired_overlap: the DMs of each FFT length are given one overlap (the
    larger); share_forward: the plan chooses that overlap itself.  Then the
    batched launch of each pair of length 4096 ... 16384 transforms every
    segment forward once for both DMs (codedisp_regm_kernel); the 32768-point
    pair runs one DM per launch.  own_overlap: every DM has its own overlap,
    hence its own segment grid, and runs alone."""
    from skyweaver_amd import DeviceBuffer as DB
    F, A, T = 2, 64, 24576
    dms = [0.0, 5.0, 20.0, 50.0, 100.0, 180.0, 260.0, 350.0]
    Ns = [4096, 4096, 8192, 8192, 16384, 16384, 32768, 32768]
    paired = [2, 2, 10, 10, 32, 32, 62, 62]
    ov = paired if mode == "paired_overlap" else None
    share = mode == "share_forward"
    plan_b = sw.CoherentDedisperser(F, A, T, F_LOW, CHAN_BW, TSAMP, dms, fft_len=Ns, overlap=ov, share_forward=share)
    plan_u = sw.CoherentDedisperser(F, A, T, F_LOW, CHAN_BW, TSAMP, dms, fft_len=Ns, overlap=ov, share_forward=share)
    Os = [plan_b.info(d)[1] for d in range(8)]
    assert Os == (paired if mode != "own_overlap" else [0, 2, 4, 10, 18, 32, 46, 62])
    lats = [plan_b.latency(d) for d in range(8)]
    assert lats == [0 if T % (Ns[d] - Os[d]) == 0 else Ns[d] - Os[d] for d in range(8)]
    g1 = sw.synth_voltages((F, T, 2, A, 2), seed=0x5EED0003, sigma=16.0)
    g2 = sw.synth_voltages((F, T, 2, A, 2), seed=0x5EED0013, sigma=16.0)
    for gi, g in enumerate((g1, g2)):
        d_g = DB.from_host(g)
        ob = [DB.empty(g.shape, np.int8, fill=0x55) for _ in dms]
        obf = [DB.empty(g.shape, np.float32) for _ in dms]
        ou = [DB.empty(g.shape, np.int8, fill=0x33) for _ in dms]
        plan_b.run(d_g, ob, obf, batched=True)
        for d in range(8):
            plan_u.run(d_g, ou, dm_begin=d, dm_count=1, batched=False)
        plan_b.push_history(d_g)
        plan_u.push_history(d_g)
        for d in range(8):
            got = ob[d].download()
            np.testing.assert_array_equal(got, ou[d].download())
            if gi == 1:
                chirp = orc.dm_chirp(dms[d], F_LOW, CHAN_BW, F, Ns[d])
                ei, ef = _oracle_stream(orc, np.zeros((F, Os[d], 2, A, 2), np.int8), np.concatenate([g1, g2], 1),
                                        chirp, Ns[d], Os[d])
                ei, ef = _window(ei, ef, T - lats[d], T)
                _check(got, obf[d].download(), ei, ef, f"gulp 2 DM {dms[d]} N {Ns[d]}")


@pytest.mark.parametrize("T,dms,Ns", [
    (40960, [0.0, 20.0, 100.0, 180.0], [4096, 8192, 16384, 16384]),  # history read in place
    (8192, [50.0, 180.0], [8192, 16384]),  # history longer than a gulp: copied
])
def test_codedisp_history_in_place(sw, orc, dev, T, dms, Ns):
    """Time-major (FPAT) input with the history read from the previous gulp's
    buffer: bit for bit the plan that copies its history, over three gulps
    whose output windows wrap the output rings, and the oracle on the
    continuous stream."""
    from skyweaver_amd import DeviceBuffer as DB
    F, A = 2, 64
    kw = dict(fft_len=Ns, input_fpat=True)
    plan_c = sw.CoherentDedisperser(F, A, T, F_LOW, CHAN_BW, TSAMP, dms, **kw)
    plan_r = sw.CoherentDedisperser(F, A, T, F_LOW, CHAN_BW, TSAMP, dms, history_in_place=True, **kw)
    gulps = [sw.synth_voltages((F, T, 2, A, 2), seed=0x5EED0040 + g, sigma=16.0) for g in range(3)]
    kept = []  # every gulp's buffer stays alive (and unchanged) until the next run
    for gi, g in enumerate(gulps):
        d_g = DB.from_host(np.ascontiguousarray(g.transpose(0, 2, 3, 1, 4)))  # FPAT
        kept.append(d_g)
        oc = [DB.empty(g.shape, np.int8, fill=0x55) for _ in dms]
        orr = [DB.empty(g.shape, np.int8, fill=0x33) for _ in dms]
        orf = [DB.empty(g.shape, np.float32) for _ in dms]
        plan_c.run(d_g, oc)
        plan_r.run(d_g, orr, orf)
        plan_c.push_history(d_g)
        plan_r.push_history(d_g)
        for d in range(len(dms)):
            np.testing.assert_array_equal(orr[d].download(), oc[d].download(), err_msg=f"gulp {gi} DM {dms[d]}")
        if gi == 2:
            d = len(dms) - 1
            N, O = plan_r.info(d)
            chirp = orc.dm_chirp(dms[d], F_LOW, CHAN_BW, F, N)
            ei, ef = _oracle_stream(orc, np.zeros((F, O, 2, A, 2), np.int8), np.concatenate(gulps, 1), chirp, N, O)
            ei, ef = _window(ei, ef, 2 * T - plan_r.latency(d), T)
            _check(orr[d].download(), orf[d].download(), ei, ef, f"in place gulp 3 DM {dms[d]} N {N}")


def test_codedisp_autotune(sw, dev):
    """The FFT-length autotuner returns, per DM, a timed power-of-two length
    of at least 4 O (the fastest of those it timed), and a pipeline built
    with autotune_fft runs."""
    F, A, T = 4, 16, 32768
    dms = [0.0, 50.0, 350.0]
    nd = len(dms)
    lens, share, ms = sw.CoherentDedisperser.autotune(F, A, T, F_LOW, CHAN_BW, TSAMP, dms, tune_chans=2)
    cands = [1024, 2048, 4096, 8192, 16384, 32768]
    plan = sw.CoherentDedisperser(F, A, T, F_LOW, CHAN_BW, TSAMP, dms)
    Os = [plan.info(d)[1] for d in range(nd)]
    single = []
    for d in range(nd):
        assert lens[d] in cands and lens[d] >= 4 * Os[d]
        tried = [c for c in range(6) if ms[d][c] >= 0]
        assert all(cands[c] >= 4 * Os[d] for c in tried)
        single.append(min(tried, key=lambda c: ms[d][c]))
    best_single = sum(ms[d][single[d]] for d in range(nd))
    # the batched row: every length that fits at least two DMs was timed
    for c in range(6):
        fit = [d for d in range(nd) if cands[c] >= 4 * Os[d]]
        assert (ms[nd][c] >= 0) == (len(fit) >= 2)
    if not share:
        assert [cands.index(x) for x in lens] == single
    else:
        # one length c for every DM it fits, the others at their own best,
        # and that plan is faster than every DM at its own best
        def plan_of(c):
            fit = [d for d in range(nd) if cands[c] >= 4 * Os[d]]
            return [cands[c] if d in fit else cands[single[d]] for d in range(nd)], fit
        won = [c for c in range(6) if ms[nd][c] >= 0 and plan_of(c)[0] == lens]
        assert won
        fit = plan_of(won[0])[1]
        total = ms[nd][won[0]] + sum(ms[d][single[d]] for d in range(nd) if d not in fit)
        assert total < best_single
    # lengths longer than the gulp (which would delay the output by more
    # than a gulp) are not candidates while a shorter one fits the overlap
    short, _, ms_short = sw.CoherentDedisperser.autotune(F, A, 4096, F_LOW, CHAN_BW, TSAMP, dms, tune_chans=2)
    assert all(n <= 4096 for n in short), short
    assert all(x < 0 for x in ms_short[:, 3:].ravel())
    from skyweaver_amd import DeviceBuffer as DB
    p = sw.BeamformerPipeline(A, F, 32, 4096, 16, dms, 1284e6, CHAN_BW * F, autotune_fft=True)
    p.set_weights(sw.synth_weights((F, 32, A, 2), seed=1))
    p.init(A)
    got = []
    p.set_sinks(cb=lambda blk: got.append(blk.dm_idx))
    p(DB.from_host(sw.synth_voltages((16, A, F, 256, 2, 2), seed=2, sigma=16.0)))
    p.close()
    assert got == [0, 1, 2]
