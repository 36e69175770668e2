"""Pins the CPU oracle (oracle/) against the reference's golden vectors,
known-answer tests and closed forms, and against independent numpy
restatements (exact integer arithmetic, numpy's pocketfft) chosen so that a
dropped term, a wrong sign, index or operand order fails.  CPU only."""
import os

import numpy as np
import pytest

F32 = np.float32


def _rng(seed):
    return np.random.default_rng(seed)


def _rand_c8(shape, seed, sigma=20.0):
    x = np.round(_rng(seed).normal(0, sigma, size=shape + (2,)))
    return np.clip(x, -127, 127).astype(np.int8)


# ------------------------------------------------------------------ transpose
@pytest.mark.parametrize("H,A_in,A_out,F", [(12, 12, 12, 4), (16, 16, 16, 3), (2, 10, 16, 5)])
def test_transpose_known_answer(orc, H, A_in, A_out, F):
    # TransposerTester.cu:95-122: x = i % 113, y = i % 107
    n = H * A_in * F * 256 * 2
    i = np.arange(n)
    taftp = np.stack([i % 113, i % 107], -1).astype(np.int8).reshape(H, A_in, F, 256, 2, 2)
    out = orc.transpose(taftp, A_out)
    exp = np.zeros((F, H, 256, 2, A_out, 2), np.int8)
    exp[:, :, :, :, :A_in, :] = taftp.transpose(2, 0, 3, 4, 1, 5)
    np.testing.assert_array_equal(out, exp.reshape(F, H * 256, 2, A_out, 2))


# ------------------------------------------------------------------ helpers
def _stokes_int(v0r, v0i, v1r, v1i):
    """Exact integer Stokes I, Q, U, V (numpy int64)."""
    m0 = v0r * v0r + v0i * v0i
    m1 = v1r * v1r + v1i * v1i
    return [m0 + m1, m0 - m1, 2 * (v0r * v1r + v0i * v1i), 2 * (v0i * v1r - v0r * v1i)]


def _stokes_f32(r0, i0, r1, i1):
    """Float32 Stokes with one rounding per operation, in the reference's order."""
    m0 = r0 * r0 + i0 * i0
    m1 = r1 * r1 + i1 * i1
    return [m0 + m1, m0 - m1, F32(2) * (r0 * r1 + i0 * i1), F32(2) * (i0 * r1 - r0 * i1)]


def _clamp_i8(x):
    return np.maximum(F32(-128), np.minimum(F32(127), x)).astype(np.int8)


# ------------------------------------------------------------- incoherent beam
@pytest.mark.parametrize("stokes", ["I", "Q", "U", "V", "IQUV"])
@pytest.mark.parametrize("S", [1, 3])
def test_ib_exact_integer_sums(orc, stokes, S):
    F, T, A, ts = 3, 64, 12, 8
    v = _rand_c8((F, T, 2, A), 11)
    w = (_rng(5).random((S, A)) > 0.3).astype(np.float32)
    w[0] = 1.0
    scale = (_rng(6).random(S * F) * 50 + 10).astype(np.float32)
    offset = (_rng(7).random(S * F) * 1000).astype(np.float32)
    raw, out = orc.ib_beamform(v, ts, scale, offset, w, S, stokes)
    vi = v.astype(np.int64)
    st = _stokes_int(vi[:, :, 0, :, 0], vi[:, :, 0, :, 1], vi[:, :, 1, :, 0], vi[:, :, 1, :, 1])  # [F][T][A]
    comps = [0, 1, 2, 3] if stokes == "IQUV" else [orc.STOKES[stokes]]
    for k, c in enumerate(comps):
        P = np.einsum("fta,sa->sft", st[c].astype(np.float64), w.astype(np.float64))
        P = P.reshape(S, F, T // ts, ts).sum(-1).transpose(0, 2, 1)  # [S][T'][F], exact ints
        r = raw[..., k] if stokes == "IQUV" else raw
        o = out[..., k] if stokes == "IQUV" else out
        np.testing.assert_array_equal(r, P.astype(np.float32))
        sc = scale.reshape(S, 1, F)
        of = offset.reshape(S, 1, F)
        q = np.rint((r - of) / sc) if c == 0 else np.rint(r / sc)
        np.testing.assert_array_equal(o, _clamp_i8(q.astype(np.float32)))


def test_ib_single_antenna_closed_form(orc):
    F, T, A, ts = 2, 32, 8, 4
    v = _rand_c8((F, T, 2, A), 3)
    w = np.zeros((1, A), np.float32)
    w[0, 5] = 1.0
    raw, _ = orc.ib_beamform(v, ts, np.ones(F, np.float32), np.zeros(F, np.float32), w, 1, "I")
    a = v[:, :, :, 5, :].astype(np.int64)
    p = (a ** 2).sum(axis=(2, 3)).reshape(F, T // ts, ts).sum(-1).T
    np.testing.assert_array_equal(raw[0], p.astype(np.float32))


# --------------------------------------------------------------- coherent beam
def _cb_numpy(v, w, ts, scales, offsets, antw, bmap, stokes, ib_sub):
    """Independent restatement: exact int64 complex sums, float32 epilogue."""
    vi = v.astype(np.int64)
    wi = w.astype(np.int64)
    # [F][T][P][A] x [F][B][A]
    re = np.einsum("fba,ftpa->fbtp", wi[..., 0], vi[..., 0]) - np.einsum("fba,ftpa->fbtp", wi[..., 1], vi[..., 1])
    im = np.einsum("fba,ftpa->fbtp", wi[..., 0], vi[..., 1]) + np.einsum("fba,ftpa->fbtp", wi[..., 1], vi[..., 0])
    re = re.astype(np.float32)
    im = im.astype(np.float32)
    st = _stokes_f32(re[..., 0], im[..., 0], re[..., 1], im[..., 1])  # [F][B][T]
    Fn, B, T = re.shape[:3]
    ist = _stokes_int(vi[:, :, 0, :, 0], vi[:, :, 0, :, 1], vi[:, :, 1, :, 0], vi[:, :, 1, :, 1])  # [F][T][A]
    comps = [0, 1, 2, 3] if stokes == "IQUV" else [{"I": 0, "Q": 1, "U": 2, "V": 3}[stokes]]
    outs = []
    for c in comps:
        power = np.zeros((Fn, B, T // ts), np.float32)
        for q in range(ts):  # sequential float accumulation over the window
            power = power + st[c][:, :, q::ts]
        ibp = np.einsum("fta,sa->sft", ist[c].astype(np.float64), antw.astype(np.float64))
        ibp = ibp.reshape(antw.shape[0], Fn, T // ts, ts).sum(-1).astype(np.float32)  # [S][F][T']
        s = bmap  # [B]
        sc = scales.reshape(-1, Fn)[s].T[:, :, None]  # [F][B][1]
        of = offsets.reshape(-1, Fn)[s].T[:, :, None]
        ib = ibp[s].transpose(1, 0, 2)  # [F][B][T']
        if ib_sub:
            r = np.rint((power - ib * F32(16129.0)) / sc)
        elif c == 0:
            r = np.rint((power - of) / sc)
        else:
            r = np.rint(power / sc)
        outs.append(_clamp_i8(r.astype(np.float32)).transpose(2, 0, 1))  # TFB
    return np.stack(outs, -1) if stokes == "IQUV" else outs[0]


@pytest.mark.parametrize("stokes", ["I", "Q", "U", "V", "IQUV"])
@pytest.mark.parametrize("ib_sub", [True, False])
def test_cb_matches_exact_integer_restatement(orc, stokes, ib_sub):
    F, T, A, B, ts, S = 2, 32, 8, 5, 4, 2
    v = _rand_c8((F, T, 2, A), 21, sigma=32)
    th = _rng(22).uniform(0, 2 * np.pi, (F, B, A))
    w = np.stack([np.round(127 * np.cos(th)), np.round(127 * np.sin(th))], -1).astype(np.int8)
    bmap = np.array([1, 0, 0, 1, 0], np.int32)
    antw = np.ones((S, A), np.float32)
    scales = (_rng(23).random(S * F) * 4e6 + 1e6).astype(np.float32)
    offsets = (_rng(24).random(S * F) * 1e7).astype(np.float32)
    out = orc.cb_beamform(v, w, ts, scales, offsets, antw, bmap, S, stokes, ib_sub)
    exp = _cb_numpy(v, w, ts, scales, offsets, antw, bmap, stokes, ib_sub)
    np.testing.assert_array_equal(out, exp)


def test_cb_single_antenna_closed_form(orc):
    # one antenna with w = 127: power = 127^2 |v|^2 summed over the window
    F, T, A, B, ts = 1, 16, 4, 1, 4
    v = _rand_c8((F, T, 2, A), 31)
    w = np.zeros((F, B, A, 2), np.int8)
    w[0, 0, 2, 0] = 127
    scale = np.ones(1, np.float32)
    out = orc.cb_beamform(v, w, ts, scale, np.zeros(1, np.float32), np.zeros((1, A), np.float32),
                          np.zeros(B, np.int32), 1, "I", True)
    p = 16129 * (v[0, :, :, 2, :].astype(np.int64) ** 2).sum(axis=(1, 2)).reshape(T // ts, ts).sum(-1)
    np.testing.assert_array_equal(out[:, 0, 0], np.clip(p, -128, 127).astype(np.int8))
    # zero voltages: zero power, zero output
    out0 = orc.cb_beamform(np.zeros_like(v), w, ts, scale, np.zeros(1, np.float32), np.ones((1, A), np.float32),
                           np.zeros(B, np.int32), 1, "I", True)
    assert not out0.any()


# ----------------------------------------------------------------- statistics
def test_statistics_exact_moments(orc):
    F, T, A = 2, 300, 4
    v = _rand_c8((F, T, 2, A), 41, sigma=20)
    v[1, :, 1, 3, :] = np.clip(v[1, :, 1, 3, :] + 50, -127, 127)  # non-zero mean column
    out = orc.statistics(v)
    x = v.astype(np.int64).transpose(0, 2, 3, 1, 4).reshape(F, 2, A, 2 * T)  # both re and im samples
    n = 2 * T
    S1 = x.sum(-1)
    for f in range(F):
        for p in range(2):
            for a in range(A):
                xs = [int(q) for q in x[f, p, a]]
                from fractions import Fraction as Fr
                mean = Fr(sum(xs), n)
                M2 = sum((Fr(q) - mean) ** 2 for q in xs)
                M3 = sum((Fr(q) - mean) ** 3 for q in xs)
                M4 = sum((Fr(q) - mean) ** 4 for q in xs)
                exp = [float(mean), float(M2 / (n - 1)) ** 0.5, np.sqrt(n) * float(M3) / float(M2) ** 1.5,
                       n * float(M4) / float(M2) ** 2 - 3.0]
                np.testing.assert_allclose(out[f, p, a], exp, rtol=1e-10, atol=1e-12)
    assert S1.shape == (F, 2, A)


def test_update_scalings_closed_form(orc):
    F, P, A, ts = 3, 2, 64, 16
    stds = np.full((F, P, A), 16.0, np.float32)
    bw = np.ones((2, A), np.float32)
    bw[1, 32:] = 0.0  # second beamset: 32 antennas
    out = orc.update_scalings(stds, bw, 2, ts, ts, 24.0)
    for bs, nant in ((0, 64), (1, 32)):
        sl = slice(bs * F, (bs + 1) * F)
        scale_cb = (127 * 16.0 * np.sqrt(nant)) ** 2
        np.testing.assert_allclose(out["cb_offsets"][sl], scale_cb * 2 * ts * P, rtol=1e-6)
        np.testing.assert_allclose(out["cb_scaling"][sl], scale_cb * np.sqrt(4 * ts * P) / 24.0, rtol=1e-6)
        dof = 2 * ts * nant * P
        np.testing.assert_allclose(out["ib_offsets"][sl], 256.0 * dof, rtol=1e-6)
        np.testing.assert_allclose(out["ib_scaling"][sl], 256.0 * np.sqrt(2 * dof) / 24.0, rtol=1e-6)


# ---------------------------------------------------- incoherent dedispersion
def _ref_config_freqs(bw=32e6, cfreq=580e6, nchans=64):
    chbw = bw / nchans
    fbottom = cfreq - bw / 2
    return fbottom + chbw / 2 + chbw * np.arange(nchans), chbw


def test_id_zero_dm_delays(orc):
    freqs, chbw = _ref_config_freqs(13.375e6, 1284e6)
    d, maxd = orc.id_delays([0.0], freqs, 16 / chbw)
    assert maxd == 0 and not d.any()


@pytest.mark.parametrize("ts", [1, 2])
def test_id_ones(orc, ts):
    # IncoherentDedisperserTester.cu:52-79 and :96-126 (64 channels, 32 beams,
    # cb_tscrunch 16, 32 MHz at 580 MHz, DMs 0..40)
    freqs, chbw = _ref_config_freqs()
    delays, maxd = orc.id_delays([0, 10, 20, 30, 40], freqs, 16 / chbw)
    assert maxd > 0
    nsamples = maxd * 2 if ts == 1 else maxd * 3
    data = np.ones((nsamples, 64, 32), np.int8)
    out = orc.incoherent_dedisperse(data, delays, maxd, ts)
    assert out.shape == ((nsamples - maxd) // ts, 5, 32)
    assert (out == int(np.sqrt(64 * ts))).all()


def test_id_too_few_samples(orc):
    freqs, chbw = _ref_config_freqs()
    delays, maxd = orc.id_delays([0, 10, 20, 30, 40], freqs, 16 / chbw)
    with pytest.raises(RuntimeError):
        orc.incoherent_dedisperse(np.zeros((maxd, 64, 32), np.int8), delays, maxd, 1)


@pytest.mark.parametrize("nvec", [1, 4])
def test_id_random_brute_force(orc, nvec):
    F, B, ts = 16, 8, 2
    freqs, chbw = _ref_config_freqs(13.375e6, 1284e6, F)
    dms = np.array([0, 100, 500, 1000], np.float32)
    delays, maxd = orc.id_delays(dms, freqs, 4 / chbw)
    # DMSampleDelay: int(K (1/f0^2 - 1/f^2) DM / tsamp + 0.5)
    K = 4.148806423e15
    exp_d = (K * (1 / freqs[0] ** 2 - 1 / freqs[None, :] ** 2) * dms[:, None].astype(np.float64) / (4 / chbw)
             + 0.5).astype(np.int64)
    np.testing.assert_array_equal(delays, exp_d)
    T = maxd + 10 * ts
    shape = (T, F, B) + ((nvec,) if nvec > 1 else ())
    data = _rng(51).integers(-128, 128, shape).astype(np.int8)
    out = orc.incoherent_dedisperse(data, delays, maxd, ts)
    scale = np.float32(np.sqrt(F * ts))
    for to in range(out.shape[0]):
        t = maxd + to * ts
        for d in range(len(dms)):
            acc = np.zeros((B,) + shape[3:], np.int64)
            for q in range(ts):
                for f in range(F):
                    acc += data[t + q - delays[d, f], f].astype(np.int64)
            v = acc.astype(np.float32) / scale
            np.testing.assert_array_equal(out[to, d], np.maximum(-128, np.minimum(127, v)).astype(np.int8))


# ------------------------------------------------------------------------ FFT
@pytest.mark.parametrize("n", [8, 64, 256])
def test_fft_brute_force_dft(orc, n):
    x = _rng(n).normal(size=n) + 1j * _rng(n + 1).normal(size=n)
    k = np.arange(n)
    for inv, sgn in ((False, -1), (True, 1)):
        ang = sgn * 2 * np.pi * np.outer(k, k).astype(np.longdouble) / n
        dft = (np.cos(ang) + 1j * np.sin(ang)) @ x
        np.testing.assert_allclose(orc.fft(x, inv), dft, rtol=0, atol=1e-12 * np.sqrt(n) * np.abs(x).max())


def test_fft_numpy_and_parseval(orc):
    n = 4096
    x = _rng(1).normal(size=n) + 1j * _rng(2).normal(size=n)
    X = orc.fft(x)
    np.testing.assert_allclose(X, np.fft.fft(x), rtol=0, atol=1e-10)
    np.testing.assert_allclose(np.sum(np.abs(X) ** 2) / n, np.sum(np.abs(x) ** 2), rtol=1e-12)
    d = np.zeros(n, complex)
    d[3] = 1
    np.testing.assert_allclose(orc.fft(d), np.exp(-2j * np.pi * 3 * np.arange(n) / n), atol=1e-13)


# --------------------------------------------------------- coherent dedispersion
GOLD_F_LOW = 856e6
GOLD_BW = 13.375e6
GOLD_TSAMP = 4096 / 856e6


def _gold_overlap():
    # CoherentDedisperserTester.cu:35-48
    K = 4.148806423e15
    d = K * (1 / GOLD_F_LOW ** 2 - 1 / (GOLD_F_LOW + 856e6 / 4096) ** 2) * 1000
    return int(np.ceil(d / GOLD_TSAMP))


def test_codedisp_golden(orc, golden_dir):
    # CoherentDedisperserTester.cu:113-204: DM 1000, N 1024, channel 0, +-1
    inp = np.fromfile(os.path.join(golden_dir, "codedisp_input_DM1000_1chan.bin"), np.int8).reshape(1024, 2, 64, 2)
    gold = np.fromfile(os.path.join(golden_dir, "codedisp_output_DM1000_1chan.bin"), np.int8).reshape(446, 2, 64, 2)
    O = _gold_overlap()
    assert O == 578
    chirp = orc.dm_chirp(1000.0, GOLD_F_LOW, GOLD_BW / 64, 64, 1024)
    _, out = orc.codedisp_channel(inp, chirp[0], 1024, O, 1024 - O)
    diff = out.astype(int) - gold.astype(int)
    assert np.abs(diff).max() <= 1
    assert np.mean(diff == 0) > 0.999


def test_chirp_closed_forms(orc):
    H = orc.dm_chirp(50.0, 1284e6 - 13.375e6 / 2, 13.375e6 / 64, 4, 256)
    np.testing.assert_allclose(np.abs(H), 1, atol=1e-15)
    np.testing.assert_array_equal(H[:, 0], 1)
    assert np.all(orc.dm_chirp(0.0, 856e6, 2e5, 2, 64) == 1)
    # phase(k) = -2 pi K DM nu^2 / ((nu0 + nu) nu0^2), nu = k bw_c / N, nu0 = f_low + (c - 1/2) bw_c
    K, dm, f_low, bwc, N = 4.148806423e15, 50.0, 1284e6 - 13.375e6 / 2, 13.375e6 / 64, 256
    nu = np.arange(N) * bwc / N
    nu0 = f_low + 2 * bwc - 0.5 * bwc
    ph = -2 * np.pi * K * dm * nu ** 2 / ((nu0 + nu) * nu0 ** 2)
    np.testing.assert_allclose(np.angle(H[2] * np.exp(-1j * ph)), 0, atol=1e-9)


def test_physical_chirp_follows_the_delay_law(orc):
    """The physical chirp's group delay -(1/2 pi) dphi/dnu is the cold-plasma
    law K DM (1/nu0^2 - 1/(nu0 + nu)^2) at every baseband frequency nu of the
    channel (negative in the upper half of the bins), from adjacent-bin phase
    differences (midpoint rule); the reference mapping's upper half follows
    it at the wrong frequencies."""
    K, dm, N = 4.148806423e15, 1000.0, 8192
    low, bwc = 856e6, 856e6 / 4096
    Hp = orc.dm_chirp(dm, low, bwc, 3, N, physical=True)
    Hr = orc.dm_chirp(dm, low, bwc, 3, N)
    dnu = bwc / N
    k = np.arange(N)
    for c in range(3):
        nu0 = low + (c + 0.5) * bwc
        nu = np.where(2 * k >= N, k - N, k) * dnu
        order = np.argsort(nu)                        # natural frequency order
        h, f = Hp[c][order], nu[order]
        dphi = np.angle(h[1:] * np.conj(h[:-1]))      # small steps: no wrap
        assert np.abs(dphi).max() < 1.0
        tau = -dphi / (2 * np.pi * dnu)
        mid = f[:-1] + dnu / 2
        law = K * dm * (1 / nu0 ** 2 - 1 / (nu0 + mid) ** 2)
        np.testing.assert_allclose(tau, law, rtol=2e-6, atol=1e-9)
        assert abs(Hp[c][0] - 1) < 1e-15               # zero delay and phase at the centre
    # the reference chirp's bins N/2..N-1 carry the delays of frequencies
    # above the channel: its slope there disagrees with the law at nu < 0
    dr = -np.angle(Hr[0][N // 2 + 1:] * np.conj(Hr[0][N // 2:-1])) / (2 * np.pi * dnu)
    nu0 = low + 0.5 * bwc
    law_neg = K * dm * (1 / nu0 ** 2 - 1 / (nu0 + (np.arange(N // 2, N - 1) - N + 0.5) * dnu) ** 2)
    assert np.max(np.abs(dr - law_neg)) > 1e-4


def _dispersed_pulse(n_c, T, dm, f_c, bwc, amp):
    """A pulse arriving at sample n_c at the channel centre f_c, dispersed by
    the cold-plasma law: its instantaneous baseband frequency at time t is
    the nu whose arrival time t_c + K DM (1/(f_c + nu)^2 - 1/f_c^2) is t
    (nu within the channel), phase integrated sample by sample; complex int8."""
    K = 4.148806423e15
    tsamp = 1 / bwc
    t = (np.arange(T) - n_c) * tsamp
    inv2 = 1 / f_c ** 2 + t / (K * dm)
    nu = np.full(T, np.nan)
    ok = inv2 > 0
    nu[ok] = 1 / np.sqrt(inv2[ok]) - f_c
    on = ok & (np.abs(nu) < bwc / 2)
    phase = 2 * np.pi * np.cumsum(np.where(on, nu, 0.0)) * tsamp
    x = np.where(on, amp * np.exp(1j * phase), 0)
    out = np.zeros((T, 1, 1, 2), np.int8)
    out[:, 0, 0, 0] = np.clip(np.rint(x.real), -127, 127)
    out[:, 0, 0, 1] = np.clip(np.rint(x.imag), -127, 127)
    return out, int(on.sum())


def test_physical_chirp_recompresses_a_dispersed_pulse(orc):
    """Overlap-save dedispersion with the physical chirp gathers a pulse
    dispersed over ~580 samples of one channel (DM 1000 at 856 MHz) back
    into a few samples at its centre-frequency arrival time; with the
    reference chirp half the band stays spread."""
    dm, low, bwc, N, O = 1000.0, 856e6, 856e6 / 4096, 4096, 580
    f_c = low + 0.5 * bwc
    T, n_c = 6000, 3000
    x, width = _dispersed_pulse(n_c, T, dm, f_c, bwc, 40.0)
    assert 500 < width < 700
    fracs = {}
    for physical in (True, False):
        chirp = orc.dm_chirp(dm, low, bwc, 1, N, physical=physical)[0]
        yf, _ = orc.codedisp_channel(x, chirp, N, O, T - O)
        p = np.abs(yf[:, 0, 0]) ** 2                  # output j is input sample j + O/2
        j = n_c - O // 2
        fracs[physical] = p[j - 4:j + 5].sum() / p.sum()
        if physical:
            assert abs(int(np.argmax(p)) - j) <= 1
    assert fracs[True] > 0.9, fracs
    assert fracs[False] < 0.6, fracs


@pytest.mark.parametrize("N,O", [(256, 0), (256, 20), (128, 2)])
def test_codedisp_dm0_identity(orc, N, O):
    T = 1000
    x = _rand_c8((T + O, 2, 3), 61)
    chirp = np.ones(N, complex)
    yf, yi = orc.codedisp_channel(x, chirp, N, O, T)
    np.testing.assert_array_equal(yi, x[O // 2:O // 2 + T])
    np.testing.assert_allclose(yf, x[O // 2:O // 2 + T, ..., 0] + 1j * x[O // 2:O // 2 + T, ..., 1], atol=1e-10)


def test_codedisp_framing_vs_numpy(orc):
    N, O, T = 128, 14, 700
    x = _rand_c8((T + O, 2, 2), 71)
    chirp = orc.dm_chirp(200.0, 1000e6, 1e6, 1, N)[0]
    yf, yi = orc.codedisp_channel(x, chirp, N, O, T)
    xc = x[..., 0].astype(float) + 1j * x[..., 1]
    hop = N - O
    exp = np.zeros((T, 2, 2), complex)
    for j in range(T):
        s, k = divmod(j, hop)
        seg = np.zeros((N, 2, 2), complex)
        chunk = xc[s * hop:s * hop + N]
        seg[:len(chunk)] = chunk
        exp[j] = np.fft.ifft(np.fft.fft(seg, axis=0) * chirp[:, None, None], axis=0)[O // 2 + k]
    np.testing.assert_allclose(yf, exp, rtol=0, atol=1e-9)
    ref_i8 = np.stack([np.rint(exp.real), np.rint(exp.imag)], -1).astype(np.int64)
    np.testing.assert_array_equal(yi, ((ref_i8 + 128) % 256 - 128).astype(np.int8))


def test_codedisp_linearity(orc):
    N, O, T = 256, 30, 600
    a = _rand_c8((T + O, 2, 2), 81, sigma=10)
    b = _rand_c8((T + O, 2, 2), 82, sigma=10)
    chirp = orc.dm_chirp(500.0, 900e6, 2e5, 1, N)[0]
    ya, _ = orc.codedisp_channel(a, chirp, N, O, T)
    yb, _ = orc.codedisp_channel(b, chirp, N, O, T)
    yab, _ = orc.codedisp_channel((a.astype(np.int16) + b).astype(np.int8), chirp, N, O, T)
    np.testing.assert_allclose(yab, ya + yb, atol=1e-9)


# -------------------------------------------------------------------- weights
def test_weights_zero_delays(orc):
    # WeightsManagerTester.cu:99-112: {1, 0, 0} delays -> (127, 0)
    d = np.zeros((4, 8, 3), np.float32)
    d[..., 0] = 1.0
    w = orc.weights(d, np.linspace(1.2e9, 1.3e9, 3), 10.0, 9.0)
    assert (w[..., 0] == 127).all() and (w[..., 1] == 0).all()


# phase phi = ((t - t0) rate + offset) nu in turns; w = (rint(127 cos 2 pi phi),
# rint(-127 sin 2 pi phi)) (WeightsManager.cu:70-79).  Offsets and rates are
# chosen to put phi on eighths of a turn, where the weight is known without
# computing a cosine: the cases separate the sign of the imaginary part, the
# factor 2 pi, and the sign and scaling of the rate term.
PHASE_CASES = [
    # (offset turns, rate turns over dt, expected (re, im))
    (0.0, 0.0, (127, 0)),
    (0.25, 0.0, (0, -127)),
    (0.5, 0.0, (-127, 0)),
    (0.75, 0.0, (0, 127)),
    (0.125, 0.0, (90, -90)),      # 127 / sqrt(2) = 89.80
    (0.375, 0.0, (-90, -90)),
    (0.0, 0.25, (0, -127)),       # rate only
    (0.0, -0.25, (0, 127)),
    (0.125, 0.125, (0, -127)),    # offset and rate add
    (0.375, -0.125, (0, -127)),
    (1.25, 0.0, (0, -127)),       # whole turns drop out
]


@pytest.mark.parametrize("off_turns,rate_turns,expect", PHASE_CASES)
def test_weights_phase_closed_forms(orc, off_turns, rate_turns, expect):
    nu, t0, dt = 1.2841e9, 1708082168.0, 16.0
    d = np.zeros((2, 3, 3), np.float32)
    d[..., 0] = 1.0
    d[..., 1] = np.float32(off_turns / nu)
    d[..., 2] = np.float32(rate_turns / (nu * dt))
    w = orc.weights(d, np.array([nu]), t0 + dt, t0)
    assert (w[..., 0] == expect[0]).all() and (w[..., 1] == expect[1]).all(), w[0, 0]


def test_weights_vs_numpy(orc):
    B, A = 6, 5
    d = np.zeros((B, A, 3), np.float32)
    idx = np.arange(B * A).reshape(B, A)
    d[..., 0] = (idx % 3 != 0).astype(np.float32)
    d[..., 1] = idx * 1e-11
    d[..., 2] = idx * 1e-15
    freqs = 1284e6 - 13.375e6 / 2 + 13.375e6 / 64 * (np.arange(64) + 0.5)
    w = orc.weights(d, freqs, 1708082200.0, 1708082168.957)
    t = 1708082200.0 - 1708082168.957
    ph = (t * d[..., 2].astype(np.float64) + d[..., 1].astype(np.float64))[None] * freqs[:, None, None]
    re = np.cos(6.283185307179586 * ph) * 127.0 * d[..., 0]
    im = -1.0 * np.sin(6.283185307179586 * ph) * 127.0 * d[..., 0]
    np.testing.assert_array_equal(w[..., 0], np.clip(np.rint(re), -128, 127).astype(np.int8))
    np.testing.assert_array_equal(w[..., 1], np.clip(np.rint(im), -128, 127).astype(np.int8))
    # the host oracle's round-half-away agrees (no exact .5 values)
    np.testing.assert_array_equal(w[..., 0], np.trunc(re + np.sign(re) * 0.5).astype(np.int8))


# ---------------------------------------------------------------- delay model
def _read_delay_models(path):
    raw = np.fromfile(path, np.uint8)
    off, models = 0, []
    while off < len(raw):
        nb, na = np.frombuffer(raw[off:off + 8].tobytes(), np.uint32)
        st, en = np.frombuffer(raw[off + 8:off + 24].tobytes(), np.float64)
        off += 24
        d = np.frombuffer(raw[off:off + nb * na * 12].tobytes(), np.float32).reshape(nb, na, 3)
        off += nb * na * 12
        models.append((st, en, d))
    return models


def test_parse_beamsets_fixture(orc, golden_dir):
    # DelayManagerTester.cu:87-93: test_delays.bin holds 2 beamsets
    models = _read_delay_models(os.path.join(golden_dir, "test_delays.bin"))
    assert len(models) == 5 and models[0][2].shape == (67, 57, 3)
    n, bmap, bw = orc.parse_beamsets(models[0][2], 96, 64)
    assert n == 2
    assert (bmap[:58] == 0).all() and (bmap[58:67] == 1).all() and (bmap[67:] == 0).all()
    assert bw[0].sum() == 57 and bw[1].sum() == 37 and not bw[:, 57:].any()
