"""sw_transpose_fpat_stats_noalloc (the entry point that allocates nothing)
against sw_transpose_fpat_stats, through the C ABI: the transposed gulp and
the statistics are identical on two calls in a row, the second with a
smaller input, and a shape whose partial sums exceed the fixed scratch array
is rejected before any launch."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _tiled_input(sw, DB, nbytes, seed):
    """A device gulp of nbytes filled by tiling a 64 MiB seeded block."""
    tile = sw.synth_voltages((1 << 25, 2), seed=seed, sigma=16.0, nthreads=16).reshape(-1)
    d = DB(nbytes)
    d_tile = DB.from_host(tile)
    for off in range(0, nbytes, tile.nbytes):
        sw.check(sw.lib().sw_memcpy_dtod(d.ptr + off, d_tile.ptr, min(tile.nbytes, nbytes - off), None))
    sw.synchronize()
    d_tile.close()
    return d


def _compare(sw, DB, d_in, H, nant_in, nant_out, F):
    T = H * 256
    fb = F * 2 * nant_out * T * 2
    out = [DB(fb) for _ in range(2)]
    st = [DB(F * 2 * nant_out * 16) for _ in range(2)]
    for o in out:
        o.fill_bytes(0x5A)
    sw.transpose_fpat_stats(d_in, out[0], st[0], H, nant_in, nant_out, F)
    sw.transpose_fpat_stats_noalloc(d_in, out[1], st[1], H, nant_in, nant_out, F)
    sw.synchronize()
    s0, s1 = st[0].download().view(np.float32), st[1].download().view(np.float32)
    np.testing.assert_array_equal(s1, s0)
    assert np.isfinite(s0.reshape(F, 2, nant_out, 4)[:, :, :nant_in]).all()
    # the transposed gulps, compared on the device in 1 GiB pieces
    step = 1 << 30
    for off in range(0, fb, step):
        n = min(step, fb - off)
        a, b = np.empty(n, np.uint8), np.empty(n, np.uint8)
        sw.check(sw.lib().sw_memcpy_dtoh(a.ctypes.data, out[0].ptr + off, n, None))
        sw.check(sw.lib().sw_memcpy_dtoh(b.ctypes.data, out[1].ptr + off, n, None))
        sw.synchronize()
        assert np.array_equal(a, b), f"FPAT differs in bytes [{off}, {off + n})"
    for b in out + st:
        b.close()


def test_noalloc_matches_allocating_entry_point(sw, orc, dev):
    from skyweaver_amd import DeviceBuffer as DB
    # call 1: 264 heaps x 128 channels x 120 of 128 antennas: 32768 columns
    # take more than the 16 MiB scratch in 16-heap chunks (17 x 32768 x 32 B),
    # so the partials use 32-heap chunks
    H, F, A_in, A = 264, 128, 120, 128
    print(f"noalloc call 1: {H} heaps, {F} channels, {A_in}/{A} antennas", flush=True)
    d_in = _tiled_input(sw, DB, H * A_in * F * 256 * 4, seed=0x5EED0901)
    _compare(sw, DB, d_in, H, A_in, A, F)
    d_in.close()
    # call 2, smaller: the scratch holds the first call's partials, which
    # must not leak into this one
    H, F, A_in, A = 32, 8, 64, 64
    print(f"noalloc call 2: {H} heaps, {F} channels, {A_in}/{A} antennas", flush=True)
    g = sw.synth_voltages((H, A_in, F, 256, 2, 2), seed=0x5EED0902, sigma=16.0)
    d_in = DB.from_host(g)
    _compare(sw, DB, d_in, H, A_in, A, F)
    # and against the oracle's statistics of the same gulp
    st = DB(F * 2 * A * 16)
    fp = DB(F * 2 * A * H * 256 * 2)
    sw.transpose_fpat_stats_noalloc(d_in, fp, st, H, A_in, A, F)
    sw.synchronize()
    got = st.download().view(np.float32).reshape(F, 2, A, 4)
    want = orc.statistics(orc.transpose(g, A))
    np.testing.assert_allclose(got[:, :, :A_in].astype(np.float64), want[:, :, :A_in], rtol=1e-5, atol=1e-6)


def test_noalloc_rejects_oversize(sw, dev):
    # 1024 heaps hold at most 32768 columns in the scratch: 40960 are refused
    # by the host-side check, before any launch (null-free dummy pointers)
    with pytest.raises(sw.SkyweaverError, match="scratch"):
        sw.check(sw.lib().sw_transpose_fpat_stats_noalloc(ctypes.c_void_p(256), ctypes.c_void_p(256),
                                                          ctypes.c_void_p(256), 1024, 128, 128, 160, None))
