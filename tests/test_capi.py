"""CPU-side checks of the C-ABI library: it loads, exports every declared
entry point, and its host-only functions (scalings, synthetic generator)
behave.  No GPU calls."""
import ctypes

import numpy as np
import pytest


def test_library_exports_every_header_symbol(sw):
    names = sw.exported_symbols()
    for core in ("sw_transpose", "sw_statistics", "sw_ib_beamform", "sw_cb_beamform", "sw_last_error"):
        assert core in names
    lib = sw.lib()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert b"gfx950" in lib.sw_version()


def test_argument_errors_do_not_touch_the_gpu(sw):
    lib = sw.lib()
    assert lib.sw_transpose(None, None, 1, 4, 4, 1, None) == -1
    assert "null buffer" in sw.last_error()
    d = sw.dims(4, 64, 8, 32, 16, "I", 0)
    assert lib.sw_ib_beamform(1, 1, 1, 1, 1, 1, ctypes.byref(d), None) == -1
    assert "beamsets" in sw.last_error()
    d = sw.dims(4, 64, 6, 32, 16)
    assert lib.sw_cb_beamform(16, 16, 16, 16, 16, 16, 16, ctypes.byref(d), 0, None, None) == -1
    assert "multiple of 4" in sw.last_error()
    assert lib.sw_transpose_fpat_stats(16, 16, 16, 1, 8, 4, 1, None, None) == -1
    assert "nantennas" in sw.last_error()
    assert lib.sw_transpose_fpat_stats(8, 16, 16, 1, 4, 4, 1, None, None) == -1
    assert "aligned" in sw.last_error()
    assert lib.sw_transpose_fpat_stats(16, 16, 16, 1 << 14, 4, 4, 1, None, None) == -1
    assert "exact sums" in sw.last_error()
    assert lib.sw_pipeline_prefetch(None, None, 0) == -1
    assert "null" in sw.last_error()
    assert lib.sw_pipeline_process_prefetched(None) == -1
    assert lib.sw_pipeline_synchronize(None) == -1


def test_update_scalings_matches_oracle_bitwise(sw, orc):
    rng = np.random.default_rng(3)
    F, P, A, S = 8, 2, 64, 3
    stats = rng.random((F, P, A, 4)).astype(np.float32) * 30
    bw = (rng.random((S, A)) > 0.4).astype(np.float32)
    got = sw.update_scalings(stats, bw, S, 16, 16, 24.0)
    exp = orc.update_scalings(stats[..., 1], bw, S, 16, 16, 24.0)
    for k in exp:
        np.testing.assert_array_equal(got[k], exp[k], err_msg=k)


def test_synth_deterministic_and_gaussian(sw):
    a = sw.synth_voltages((3, 1 << 19, 2), seed=0x5EED0001, sigma=16.0, nthreads=4)
    b = sw.synth_voltages((3, 1 << 19, 2), seed=0x5EED0001, sigma=16.0, nthreads=1)
    np.testing.assert_array_equal(a, b)
    c = sw.synth_voltages((1 << 20, 2), seed=0x5EED0001, sigma=16.0)
    np.testing.assert_array_equal(c, a.reshape(-1, 2)[: 1 << 20])  # streams are chunk-aligned
    assert not np.array_equal(a, sw.synth_voltages((3, 1 << 19, 2), seed=0x5EED0002))
    x = a.astype(np.float64)
    assert abs(x.mean()) < 0.05 and abs(x.std() - 16.0) < 0.05
    assert x.min() >= -127 and x.max() <= 127


def test_synth_weights_amplitude(sw):
    w = sw.synth_weights((64, 64, 2), seed=0x5EED1001).astype(np.float64)
    amp = np.hypot(w[..., 0], w[..., 1])
    assert np.all(np.abs(amp - 127) <= 1.0)
    assert np.abs(w).max() <= 127
