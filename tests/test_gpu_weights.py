"""GPU weights generation against the oracle, bit-exact int8.

The cases mirror cpp/skyweaver/test/src/WeightsManagerTester.cu:99-145
(zero delays, one real delay everywhere, per-element values) plus the
reference's delay-model fixture with realistic L-band channel frequencies and
random delays whose phases span many turns.  Channel frequencies follow
PipelineConfig::calculate_channel_frequencies
(cpp/skyweaver/src/PipelineConfig.cpp:188-203)."""
import pathlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def chan_freqs(centre, bw, nchans):
    chbw = bw / nchans
    fbottom = centre - bw / 2.0
    return np.array([fbottom + chbw / 2.0 + chbw * i for i in range(nchans)], np.float64)


def run(sw, delays, freqs, cur, ep):
    B, A, _ = delays.shape
    F = len(freqs)
    d_del = sw.DeviceBuffer.from_host(np.ascontiguousarray(delays, np.float32))
    d_f = sw.DeviceBuffer.from_host(freqs)
    d_w = sw.DeviceBuffer.empty((F, B, A, 2), np.int8, fill=0x55)
    sw.generate_weights(d_del, d_w, d_f, A, B, F, cur, ep)
    sw.synchronize()
    return d_w.download()


def check(sw, orc, delays, freqs, cur=10.0, ep=9.0):
    got = run(sw, delays, freqs, cur, ep)
    want = orc.weights(delays, freqs, cur, ep)
    np.testing.assert_array_equal(got, want)
    return got


FREQS = chan_freqs(1284e6, 13.375e6, 64)


def test_zero_value(sw, orc, dev):
    d = np.zeros((32, 64, 3), np.float32)
    d[..., 0] = 1.0
    w = check(sw, orc, d, FREQS)
    assert (w[..., 0] == 127).all() and (w[..., 1] == 0).all()


def test_real_value(sw, orc, dev):
    d = np.broadcast_to(np.array([1.0, 1e-11, 1e-10], np.float32), (32, 64, 3)).copy()
    check(sw, orc, d, FREQS)


def test_real_values(sw, orc, dev):
    n = 32 * 64
    d = np.zeros((n, 3), np.float32)
    d[:, 0] = np.arange(n, dtype=np.float32) * np.float32(1e-11)
    d[:, 1] = np.arange(n, dtype=np.float32) * np.float32(1e-15)
    check(sw, orc, d.reshape(32, 64, 3), FREQS)


def test_delay_fixture(sw, orc, dev, golden_dir):
    dm = sw.DelayModelFile(pathlib.Path(golden_dir) / "test_delays.bin", 72, 64)
    delays, ep = dm.delays(1708082300.0)
    w = check(sw, orc, delays, chan_freqs(1284e6, 856e6, 64), cur=1708082300.0, ep=ep)
    # padded beams and antennas carry weight 0
    assert not w[:, 67:].any() and not w[:, :, 57:].any()


def test_phase_closed_forms(sw, orc, dev):
    """Offsets and rates that put the phase on eighths of a turn give the
    weights the closed forms fix (test_oracle.PHASE_CASES): the GPU and the
    oracle agree with them."""
    from test_oracle import PHASE_CASES
    nu, t0, dt = 1.2841e9, 1708082168.0, 16.0
    d = np.zeros((len(PHASE_CASES), 4, 3), np.float32)
    d[..., 0] = 1.0
    for b, (off, rate, _) in enumerate(PHASE_CASES):
        d[b, :, 1] = np.float32(off / nu)
        d[b, :, 2] = np.float32(rate / (nu * dt))
    w = check(sw, orc, d, np.array([nu]), cur=t0 + dt, ep=t0)
    for b, (_, _, expect) in enumerate(PHASE_CASES):
        assert (w[0, b, :, 0] == expect[0]).all() and (w[0, b, :, 1] == expect[1]).all(), (b, w[0, b, 0])


@pytest.mark.parametrize("scale", [1.0, 0.5, -1.0])
def test_random_delays(sw, orc, dev, scale):
    rng = np.random.default_rng(17)
    d = np.empty((96, 80, 3), np.float32)
    d[..., 0] = scale * rng.integers(0, 2, (96, 80))
    d[..., 1] = rng.uniform(-1e-6, 1e-6, (96, 80))   # up to ~1300 turns at 1.3 GHz
    d[..., 2] = rng.uniform(-1e-11, 1e-11, (96, 80))
    w = check(sw, orc, d, chan_freqs(1284e6, 53.5e6, 256), cur=1708082300.5, ep=1708082168.957)
    # unit-magnitude weights: |w|^2 close to (127 wt)^2
    mag = np.hypot(w[..., 0].astype(float), w[..., 1].astype(float))
    on = np.broadcast_to(d[..., 0] != 0, mag.shape)
    assert np.abs(mag[on] - 127 * abs(scale)).max() < 1.0
