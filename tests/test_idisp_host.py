"""Host-side parts of incoherent dedispersion (no GPU): channel frequencies,
DM sample delays and DedispersionPlan block parsing, against the oracle and
closed forms (cpp/skyweaver/src/PipelineConfig.cpp:188-203,
cpp/skyweaver/dedispersion_utils.cuh:19-40,
cpp/skyweaver/src/DedispersionPlan.cpp:17-78)."""
import numpy as np
import pytest


def test_channel_frequencies(sw):
    f = sw.channel_frequencies(1284e6, 13.375e6, 64)
    chbw = 13.375e6 / 64
    assert f[0] == 1284e6 - 13.375e6 / 2 + chbw / 2
    np.testing.assert_allclose(np.diff(f), chbw, rtol=1e-12)
    assert abs(f.mean() - 1284e6) < 1e-3


@pytest.mark.parametrize("nchans,bw,cbts", [(256, 53.5e6, 8), (64, 13.375e6, 16), (64, 856e6, 1)])
def test_id_delays_match_oracle(sw, orc, nchans, bw, cbts):
    f = sw.channel_frequencies(1284e6, bw, nchans)
    dms = (np.arange(1024, dtype=np.float32) * np.float32(0.9765625))
    tsamp = cbts / (bw / nchans)
    d, mx = sw.id_delays(dms, f, tsamp)
    d_o, mx_o = orc.id_delays(dms, f, tsamp)
    np.testing.assert_array_equal(d, d_o)
    assert mx == mx_o == d.max()
    assert (d[0] == 0).all() and (d[:, 0] == 0).all()
    # monotone in DM and in channel (delays are relative to the lowest channel)
    assert (np.diff(d, axis=0) >= 0).all() and (np.diff(d, axis=1) >= 0).all()


def test_config5_max_delay(sw):
    # 256 channels centred on 1284 MHz over 53.5 MHz, CB tscrunch 8, DM < 1000
    f = sw.channel_frequencies(1284e6, 53.5e6, 256)
    dms = np.arange(1024, dtype=np.float32) * np.float32(0.9765625)
    d, mx = sw.id_delays(dms, f, 8 / (53.5e6 / 256))
    k = 4.148806423e15 * (1 / f[0] ** 2 - 1 / f[-1] ** 2) * float(dms[-1]) / (8 / (53.5e6 / 256))
    assert mx == int(k + 0.5)


def test_ddplan_block_forms(sw):
    assert sw.ddplan_parse_block("5")[:2] == (5.0, 1)
    cdm, ts, dms = sw.ddplan_parse_block("5:4")
    assert (cdm, ts) == (5.0, 4) and list(dms) == [5.0]
    cdm, ts, dms = sw.ddplan_parse_block("100:0:1000:0.9765625:1")
    assert cdm == 100.0 and ts == 1 and len(dms) == 1024
    np.testing.assert_array_equal(dms, np.arange(1024, dtype=np.float32) * np.float32(0.9765625))
    # float accumulation, as the reference's for-loop
    cdm, ts, dms = sw.ddplan_parse_block("0:0:1:0.1:2")
    acc, want = np.float32(0), []
    while acc < np.float32(1):
        want.append(acc)
        acc = np.float32(acc + np.float32(0.1))
    np.testing.assert_array_equal(dms, np.array(want, np.float32))
    # an empty range keeps the start DM
    assert list(sw.ddplan_parse_block("1:2:2:1:1")[2]) == [2.0]


@pytest.mark.parametrize("bad", ["1:2:3", "a", "1:2:3:4:5:6", ""])
def test_ddplan_block_errors(sw, bad):
    with pytest.raises(sw.SkyweaverError):
        sw.ddplan_parse_block(bad)
