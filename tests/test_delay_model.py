"""DelayManager behaviour of the library's delay-model reader (host code, no GPU).

Mirrors cpp/skyweaver/test/src/DelayManagerTester.cu:36-93 on the reference's
own fixture test_delays.bin (5 models, 67 beams x 57 antennas, 2 beamsets,
first valid epoch 1708082168.957, last 1708082468.957).  Expected models are
read straight from the file with numpy; beamsets are checked against the
oracle's parse_beamsets and against the fixture's documented count.
"""
import pathlib

import numpy as np
import pytest

HDR = np.dtype([("nbeams", "<u4"), ("nantennas", "<u4"), ("start", "<f8"), ("end", "<f8")])


def read_models(path):
    raw = open(path, "rb").read()
    models, off = [], 0
    while off < len(raw):
        h = np.frombuffer(raw, HDR, 1, off)[0]
        off += HDR.itemsize
        n = int(h["nbeams"]) * int(h["nantennas"]) * 3
        d = np.frombuffer(raw, "<f4", n, off).reshape(h["nbeams"], h["nantennas"], 3)
        off += 4 * n
        models.append((float(h["start"]), float(h["end"]), d))
    return models


@pytest.fixture(scope="module")
def delay_file(golden_dir):
    return pathlib.Path(golden_dir) / "test_delays.bin"


def test_fixture_layout(delay_file):
    models = read_models(delay_file)
    assert len(models) == 5
    assert all(m[2].shape == (67, 57, 3) for m in models)
    assert delay_file.stat().st_size == 5 * (24 + 67 * 57 * 12)


def test_valid_read_first_block(sw, delay_file):
    models = read_models(delay_file)
    dm = sw.DelayModelFile(delay_file, nbeams=72, nantennas=64)
    d, ep = dm.delays(1708082169.0)
    assert d.shape == (72, 64, 3)
    assert ep == models[0][0]
    np.testing.assert_array_equal(d[:67, :57], models[0][2])
    assert not d[67:].any() and not d[:, 57:].any()  # zero padding


def test_valid_read_nth_block(sw, delay_file):
    models = read_models(delay_file)
    epoch = 1708082409.957
    k = next(i for i, m in enumerate(models) if m[0] <= epoch <= m[1])
    assert k > 0
    dm = sw.DelayModelFile(delay_file, nbeams=67, nantennas=57)
    d, ep = dm.delays(epoch)
    assert ep == models[k][0]
    np.testing.assert_array_equal(d, models[k][2])
    # the reader moves forward only: an earlier epoch is now invalid
    with pytest.raises(sw.SkyweaverError) as e:
        dm.delays(1708082169.0)
    assert e.value.code == -5


def test_too_early_epoch(sw, delay_file):
    dm = sw.DelayModelFile(delay_file, nbeams=67, nantennas=57)
    with pytest.raises(sw.SkyweaverError) as e:
        dm.delays(1708082165.0)
    assert e.value.code == -5  # InvalidDelayEpoch


def test_too_late_epoch(sw, delay_file):
    dm = sw.DelayModelFile(delay_file, nbeams=67, nantennas=57)
    with pytest.raises(sw.SkyweaverError) as e:
        dm.delays(1708082470.957)
    assert e.value.code == -6  # end of file
    assert "end of delay model file" in str(e.value)


def test_beamset_count(sw, orc, delay_file):
    models = read_models(delay_file)
    dm = sw.DelayModelFile(delay_file, nbeams=72, nantennas=64)
    n, bmap, bw = dm.beamsets()
    assert n == 2
    n_o, bmap_o, bw_o = orc.parse_beamsets(models[0][2], 72, 64)
    assert n_o == n
    np.testing.assert_array_equal(bmap, bmap_o)
    np.testing.assert_array_equal(bw, bw_o)
    # by construction: beams 0..57 use all 57 antennas, 58..66 use 37
    assert (bmap[:58] == 0).all() and (bmap[58:67] == 1).all()
    assert bw[0, :57].sum() == 57 and bw[1].sum() == 37


def test_too_many_beams_or_antennas(sw, delay_file):
    with pytest.raises(sw.SkyweaverError):
        sw.DelayModelFile(delay_file, nbeams=64, nantennas=64)
    with pytest.raises(sw.SkyweaverError):
        sw.DelayModelFile(delay_file, nbeams=67, nantennas=32)


def test_missing_file(sw, tmp_path):
    with pytest.raises(sw.SkyweaverError) as e:
        sw.DelayModelFile(tmp_path / "nope.bin", 8, 8)
    assert e.value.code == -4


def test_variable_dimensions_rejected(sw, tmp_path):
    p = tmp_path / "var.bin"
    with open(p, "wb") as f:
        for k, (nb, na) in enumerate([(4, 8), (5, 8)]):
            np.array([(nb, na, 10.0 * k, 10.0 * k + 9.0)], HDR).tofile(f)
            np.ones((nb, na, 3), "<f4").tofile(f)
    dm = sw.DelayModelFile(p, 8, 8)
    dm.delays(1.0)
    with pytest.raises(sw.SkyweaverError) as e:
        dm.delays(15.0)
    assert "Variable delay model" in str(e.value)
