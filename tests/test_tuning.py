"""The tuning table reader (sw_tuning_lookup, host only) and the committed
table for this device."""
import json
import os

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(os.path.dirname(HERE), "skyweaver_amd", "tuning", "gfx950.json")


def _table(tmp_path, entries, name="t.json"):
    p = tmp_path / name
    p.write_text(json.dumps({"device": "test", "entries": entries}))
    return str(p)


def test_lookup_matches_workload_and_dms(sw, tmp_path):
    path = _table(tmp_path, [
        {"antennas": 64, "beams": 800, "samples": 65536, "tscrunch": 16, "stokes": "I", "cb_waves": 8,
         "dms": [50.0], "fft_len": [16384], "share_forward": 0, "beams_ms_per_gulp": {"2": 1.0}},
        {"antennas": 64, "beams": 800, "samples": 65536, "tscrunch": 16, "stokes": "IQUV", "cb_waves": 2},
    ])
    assert sw.tuning_lookup(64, 800, 65536, 16, "I", [50.0], path=path) == \
        {"cb_waves": 8, "fft_len": [16384], "share_forward": False}
    # the beamformer entry applies to other DM lists, the FFT lengths do not
    assert sw.tuning_lookup(64, 800, 65536, 16, "I", [50.0, 100.0], path=path) == {"cb_waves": 8}
    # an entry whose DM list matches is preferred to the workload's first entry
    path2 = _table(tmp_path, [
        {"antennas": 64, "beams": 800, "samples": 65536, "tscrunch": 16, "stokes": "I", "cb_waves": 8,
         "dms": [50.0], "fft_len": [16384], "share_forward": 0},
        {"antennas": 64, "beams": 800, "samples": 65536, "tscrunch": 16, "stokes": "I", "cb_waves": 4,
         "dms": [0.0, 5.0], "fft_len": [4096, 4096], "share_forward": 1}], name="t2.json")
    assert sw.tuning_lookup(64, 800, 65536, 16, "I", [0.0, 5.0], path=path2) == \
        {"cb_waves": 4, "fft_len": [4096, 4096], "share_forward": True}
    assert sw.tuning_lookup(64, 800, 65536, 16, "I", [0.0], path=path2) == {"cb_waves": 8}
    assert sw.tuning_lookup(64, 800, 65536, 16, "IQUV", [0.0], path=path) == {"cb_waves": 2}
    for args in [(64, 800, 65536, 8, "I"), (80, 800, 65536, 16, "I"), (64, 512, 65536, 16, "I"),
                 (64, 800, 32768, 16, "I"), (64, 800, 65536, 16, "Q")]:
        assert sw.tuning_lookup(*args, [50.0], path=path) is None
    assert sw.tuning_lookup(64, 800, 65536, 16, "I", [50.0], path=str(tmp_path / "absent.json")) is None


@pytest.mark.parametrize("text", ['{"entries": [', '{"entries": 3}', '[1, 2', '{"entries": [{"antennas": }]}'])
def test_lookup_rejects_malformed_table(sw, tmp_path, text):
    p = tmp_path / "bad.json"
    p.write_text(text)
    with pytest.raises(sw.SkyweaverError):
        sw.tuning_lookup(64, 800, 65536, 16, "I", [50.0], path=str(p))


def test_committed_table_entries(sw):
    """The committed table parses, names a valid workgroup shape for every
    workload it holds, and those entries are what the lookup returns."""
    if not os.path.exists(TABLE):
        pytest.skip("no tuning table committed for this device")
    with open(TABLE) as f:
        tab = json.load(f)
    assert tab["entries"]
    for e in tab["entries"]:
        assert e["cb_waves"] in (2, 4, 8)
        times = e["beams_ms_per_gulp"]
        valid = {int(k): v for k, v in times.items() if v >= 0}
        assert valid[e["cb_waves"]] == min(valid.values())
        got = sw.tuning_lookup(e["antennas"], e["beams"], e["samples"], e["tscrunch"], e["stokes"], e["dms"])
        assert got == {"cb_waves": e["cb_waves"], "fft_len": e["fft_len"], "share_forward": bool(e["share_forward"])}
        assert all(n in (1024, 2048, 4096, 8192, 16384, 32768) for n in e["fft_len"])
        # no transform longer than the gulp: it would delay the output by more than a gulp
        assert max(e["fft_len"]) <= e["samples"], e["workload"]
