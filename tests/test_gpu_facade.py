"""The C++ facade end to end on the GPU: tests/cpp/facade_demo.cpp drives
BeamformerPipeline with a delay-model file, a TFB MultiFileWriter, an
IncoherentDedispersionPipeline writing TDB files, and IB / statistics
writers -- the reference CLI's arrangement.  The written DADA files are read
back and checked against the oracle: coherent beams bit-exact (DM 0), TDB
bit-exact against the oracle's dedispersion of the written TFB stream."""
import glob
import os
import re
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
A, A_IN, F, B, T, TS = 64, 57, 8, 72, 4096, 16
CFREQ, BW = 1284e6, 13.375e6 * 8 / 64
UTC0, CLOCK, TOTAL = 1708082300.0, 1712e6, 4096
HDR = np.dtype([("nbeams", "<u4"), ("nantennas", "<u4"), ("start", "<f8"), ("end", "<f8")])


def _models(path):
    raw, out, off = open(path, "rb").read(), [], 0
    while off < len(raw):
        h = np.frombuffer(raw, HDR, 1, off)[0]
        off += HDR.itemsize
        n = int(h["nbeams"]) * int(h["nantennas"]) * 3
        out.append((float(h["start"]), float(h["end"]),
                    np.frombuffer(raw, "<f4", n, off).reshape(h["nbeams"], h["nantennas"], 3)))
        off += 4 * n
    return out


def _stream(root, tag, ext):
    files = sorted(glob.glob(os.path.join(root, "**", f"*_{tag}_*.{ext}"), recursive=True),
                   key=lambda p: int(re.search(r"_(\d{16})\." + ext + "$", p).group(1)))
    assert files, (tag, ext)
    hdr = open(files[0], "rb").read(4096).split(b"\0")[0].decode()
    keys = dict(line.split(None, 1) for line in hdr.strip().splitlines() if " " in line)
    data = b"".join(open(p, "rb").read()[4096:] for p in files)
    return keys, np.frombuffer(data, np.int8)


def test_facade_demo(sw, orc, dev, golden_dir, tmp_path):
    exe = os.path.join(HERE, "cpp", "_build", "facade_demo")
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is not built: run __graft_entry__.build()")
    delay_file = os.path.join(golden_dir, "test_delays.bin")
    ngulps = 3
    r = subprocess.run([exe, str(tmp_path), delay_file, str(ngulps)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = str(tmp_path)

    kcb, cb = _stream(out, "cb", "tfb")
    _, stats = _stream(out, "stats", "fpa")
    ktdb, tdb = _stream(out, "tdb", "tdb")
    assert kcb["ORDER"].strip() == "TFB" and int(kcb["NBEAM"]) == B and int(kcb["NCHAN"]) == F
    assert ktdb["ORDER"].strip() == "TDB" and int(ktdb["NDMS"]) == 16
    cb = cb.reshape(ngulps, T // TS, F, B)
    stats = stats.view(np.float32).reshape(ngulps, F, 2, A, 4)

    models = _models(delay_file)
    freqs = sw.channel_frequencies(CFREQ, BW, F)
    nsets, bmap, bw = orc.parse_beamsets(models[0][2], B, A)
    scal = orc.update_scalings(stats[0][..., 1], bw, nsets, TS, TS, 24.0)
    for g in range(ngulps):
        volt = sw.synth_voltages((T // 256, A_IN, F, 256, 2, 2), seed=0x5EED0200 + g, sigma=16.0)
        t_unix = UTC0 + g * 2.0 * TOTAL * T / CLOCK
        k = next(i for i, m in enumerate(models) if m[0] <= t_unix <= m[1])
        delays = np.zeros((B, A, 3), np.float32)
        delays[:67, :57] = models[k][2]
        w = orc.weights(delays, freqs, t_unix, models[k][0])
        ftpa = orc.transpose(volt, A)
        want = orc.cb_beamform(ftpa, w, TS, scal["cb_scaling"], scal["cb_offsets"], bw, bmap, nsets, "I", True)
        np.testing.assert_array_equal(cb[g], want)

    idms = np.arange(0, 40, 2.5, dtype=np.float32)
    d_o, mx = orc.id_delays(idms, freqs, TS * F / BW)
    stream = cb.reshape(-1, F, B)
    nblk = (len(stream) - mx) // 300
    tdb = tdb.reshape(-1, len(idms), B)
    assert len(tdb) == nblk * 150
    for k in range(nblk):
        want = orc.incoherent_dedisperse(stream[k * 300:k * 300 + 300 + mx], d_o, mx, 2)
        np.testing.assert_array_equal(tdb[k * 150:(k + 1) * 150], want)


def test_run_pipeline_from_dada_files(tmp_path):
    """DADA files (the reference fixture's headers, 4 antennas x 16 channels,
    seeded voltages) read by MultiFileReader and fed by the double-buffered
    run_pipeline give the same coherent beams, incoherent beam and statistics
    as the same gulps fed directly; the gulp reaching the end of the data is
    dropped and start_time skips whole heaps, as in
    cpp/skyweaver/src/skyweaver_cli.cu:109-240."""
    exe = os.path.join(HERE, "cpp", "_build", "run_pipeline_demo")
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is not built: run __graft_entry__.build()")
    r = subprocess.run([exe, os.path.join(HERE, "golden", "dada"), str(tmp_path)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "run_pipeline_demo passed (4 gulps)" in r.stdout
