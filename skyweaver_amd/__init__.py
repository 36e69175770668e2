"""skyweaver_amd -- MI355X (gfx950) implementation of the skyweaver
BeamformerPipeline hot path.

The product is the C-ABI library ``_build/libskyweaver_amd.so`` (hand-written
HIP kernels, declared in ``include/skyweaver_amd.h``).  This module is a thin
ctypes binding used by the tests and the benchmark.  Device memory, streams
and events also come from the C ABI (one HIP runtime per process), every
compute call goes through it, and a missing library is an error -- there is
no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libskyweaver_amd.so")
HEADER = os.path.join(os.path.dirname(_HERE), "include", "skyweaver_amd.h")

SW_OK = 0
STOKES = {"I": 0, "Q": 1, "U": 2, "V": 3, "IQUV": 4}
FORMS = {"plain": 0, "mfma": 1}

_lib = None


class SkyweaverError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"skyweaver_amd error {code}: {msg}")
        self.code = code


def build(jobs: int = 8, force: bool = True) -> str:
    """Compile the HIP library for gfx950 (hipcc, in-tree); by default every
    source is compiled again, also where a built library exists."""
    cmd = ["make", "-s", "-C", _HERE, f"-j{jobs}"] + (["-B"] if force else [])
    subprocess.run(cmd, check=True)
    return LIB_PATH


def lib():
    """Load the C-ABI library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


class BfDims(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in (
        "nchans", "nsamples", "nantennas", "nbeams", "tscrunch", "stokes", "nbeamsets", "ib_subtract")]


class IdispConfig(ctypes.Structure):
    _fields_ = [("nchans", ctypes.c_int), ("centre_frequency", ctypes.c_double), ("bandwidth", ctypes.c_double),
                ("chan_freqs", ctypes.c_void_p), ("cb_tscrunch", ctypes.c_int), ("dms", ctypes.c_void_p),
                ("ndms", ctypes.c_int), ("tscrunch", ctypes.c_int), ("kernel", ctypes.c_int)]


class CodedispConfig(ctypes.Structure):
    _fields_ = [("nchans", ctypes.c_int), ("nantennas", ctypes.c_int), ("nsamples", ctypes.c_int),
                ("low_freq", ctypes.c_double), ("chan_bw", ctypes.c_double), ("tsamp", ctypes.c_double),
                ("ndms", ctypes.c_int), ("dms", ctypes.POINTER(ctypes.c_float)),
                ("fft_len", ctypes.POINTER(ctypes.c_int)), ("overlap", ctypes.POINTER(ctypes.c_int)),
                ("input_fpat", ctypes.c_int), ("history_in_place", ctypes.c_int), ("share_forward", ctypes.c_int),
                ("chirp", ctypes.c_int)]

# enum sw_chirp
CHIRP = {"reference": 0, "physical": 1}


_vp = ctypes.c_void_p
_c = ctypes.c_int
_sz = ctypes.c_size_t
_fp = ctypes.POINTER(ctypes.c_float)


def _declare(L):
    L.sw_last_error.argtypes = [ctypes.c_char_p, _sz]
    L.sw_last_error.restype = _c
    L.sw_version.restype = ctypes.c_char_p
    L.sw_transpose.argtypes = [_vp, _vp, _c, _c, _c, _c, _vp]
    L.sw_statistics_workspace_size.argtypes = [_c, _c, _c]
    L.sw_statistics_workspace_size.restype = _sz
    L.sw_statistics.argtypes = [_vp, _vp, _c, _c, _c, _c, _vp, _vp]
    L.sw_update_scalings.argtypes = [_vp, _vp, _c, _c, _c, _c, _c, _c, ctypes.c_float, _vp, _vp, _vp, _vp]
    L.sw_ib_beamform.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, ctypes.POINTER(BfDims), _vp]
    L.sw_cb_workspace_size.argtypes = [ctypes.POINTER(BfDims)]
    L.sw_cb_workspace_size.restype = _sz
    L.sw_cb_beamform.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.POINTER(BfDims), _c, _vp, _vp]
    L.sw_cb_pack_weights.argtypes = [_vp, _vp, _c, _c, _c, _vp]
    L.sw_cb_beamform_packed.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.POINTER(BfDims), _vp]
    L.sw_cb_beamform_packed_tile.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.POINTER(BfDims), _c, _vp]
    L.sw_tuning_lookup.argtypes = [ctypes.c_char_p, _c, _c, _c, _c, _c, _c, _vp, ctypes.POINTER(_c),
                                   ctypes.POINTER(_c), _vp, ctypes.POINTER(_c)]
    L.sw_synth_gaussian_c8.argtypes = [_vp, _sz, ctypes.c_uint64, ctypes.c_double, _c]
    L.sw_synth_phase_weights.argtypes = [_vp, _sz, ctypes.c_uint64]
    L.sw_microbench.argtypes = [_c, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double), _vp]
    L.sw_transpose_fpat.argtypes = [_vp, _vp, _c, _c, _c, _c, _vp]
    L.sw_corner_turn.argtypes = [_vp, _vp, _c, _c, _c, _vp]
    L.sw_statistics_fpat.argtypes = [_vp, _vp, _c, _c, _c, _c, _vp, _vp]
    L.sw_transpose_fpat_stats.argtypes = [_vp, _vp, _vp, _c, _c, _c, _c, _vp, _vp]
    L.sw_transpose_fpat_stats_noalloc.argtypes = [_vp, _vp, _vp, _c, _c, _c, _c, _vp]
    _pvp = ctypes.POINTER(ctypes.c_void_p)
    L.sw_codedisp_create.argtypes = [ctypes.POINTER(CodedispConfig), _vp, _pvp]
    L.sw_codedisp_destroy.argtypes = [_vp]
    L.sw_codedisp_info.argtypes = [_vp, _c, ctypes.POINTER(_c), ctypes.POINTER(_c), ctypes.POINTER(_c)]
    L.sw_codedisp_autotune.argtypes = [ctypes.POINTER(CodedispConfig), _c, _vp, _vp, _vp, _vp]
    L.sw_codedisp_run.argtypes = [_vp, _vp, _pvp, _pvp, _c, _c, _c, _vp]
    L.sw_codedisp_push_history.argtypes = [_vp, _vp, _vp]
    L.sw_codedisp_set_history.argtypes = [_vp, _vp, _vp]
    L.sw_device_count.argtypes = [ctypes.POINTER(ctypes.c_int)]
    L.sw_mem_info.argtypes = [ctypes.POINTER(_sz), ctypes.POINTER(_sz)]
    L.sw_malloc.argtypes = [_pvp, _sz]
    L.sw_free.argtypes = [_vp]
    L.sw_host_alloc.argtypes = [_pvp, _sz]
    L.sw_host_free.argtypes = [_vp]
    L.sw_memcpy_htod.argtypes = [_vp, _vp, _sz, _vp]
    L.sw_memcpy_dtoh.argtypes = [_vp, _vp, _sz, _vp]
    L.sw_memcpy_dtoh_sync.argtypes = [_vp, _vp, _sz]
    L.sw_memcpy_dtod.argtypes = [_vp, _vp, _sz, _vp]
    L.sw_memset.argtypes = [_vp, _c, _sz, _vp]
    L.sw_stream_create.argtypes = [_pvp]
    L.sw_stream_destroy.argtypes = [_vp]
    L.sw_stream_synchronize.argtypes = [_vp]
    L.sw_event_create.argtypes = [_pvp]
    L.sw_event_destroy.argtypes = [_vp]
    L.sw_event_record.argtypes = [_vp, _vp]
    L.sw_event_elapsed_ms.argtypes = [ctypes.POINTER(ctypes.c_float), _vp, _vp]
    _d = ctypes.c_double
    L.sw_generate_weights.argtypes = [_vp, _vp, _vp, _c, _c, _c, _d, _d, _vp]
    L.sw_delay_model_open.argtypes = [ctypes.c_char_p, _c, _c, _pvp]
    L.sw_delay_model_close.argtypes = [_vp]
    L.sw_delay_model_beamsets.argtypes = [_vp, ctypes.POINTER(_c), _vp, _vp, _c]
    L.sw_delay_model_get.argtypes = [_vp, _d, _vp, ctypes.POINTER(_d)]
    L.sw_channel_frequencies.argtypes = [_d, _d, _c, _vp]
    L.sw_id_delays.argtypes = [_vp, _c, _vp, _c, _d, _vp, ctypes.POINTER(_c)]
    L.sw_ddplan_parse_block.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(_c), _vp, _c,
                                        ctypes.POINTER(_c)]
    L.sw_idisp_create.argtypes = [ctypes.POINTER(IdispConfig), _pvp]
    L.sw_idisp_destroy.argtypes = [_vp]
    L.sw_idisp_info.argtypes = [_vp, ctypes.POINTER(_c), _vp, ctypes.POINTER(ctypes.c_float)]
    L.sw_idisp_kernel.argtypes = [_vp]
    L.sw_idisp_kernel.restype = ctypes.c_char_p
    L.sw_idisp_dedisperse.argtypes = [_vp, _vp, _sz, _c, _c, _vp, _vp]
    L.sw_idisp_dedisperse_ftb.argtypes = [_vp, _vp, _sz, _c, _c, _vp, _vp]


def last_error() -> str:
    buf = ctypes.create_string_buffer(1024)
    lib().sw_last_error(buf, 1024)
    return buf.value.decode()


def check(rc: int) -> None:
    if rc != SW_OK:
        raise SkyweaverError(rc, last_error())


def exported_symbols() -> list:
    """Names of every function declared in include/skyweaver_amd.h."""
    import re
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|size_t|const char\*)\s+(sw_\w+)\s*\(", text, re.M)))


# ------------------------------------------------------ memory and streams
def device_count() -> int:
    n = ctypes.c_int(0)
    check(lib().sw_device_count(ctypes.byref(n)))
    return n.value


def mem_info():
    """(free, total) bytes of device memory on the current device."""
    f, t = ctypes.c_size_t(), ctypes.c_size_t()
    check(lib().sw_mem_info(ctypes.byref(f), ctypes.byref(t)))
    return f.value, t.value


class DeviceBuffer:
    """Device allocation owned by this object (freed on close / GC)."""

    def __init__(self, nbytes: int, shape=None, dtype=None):
        import numpy as np
        p = ctypes.c_void_p()
        check(lib().sw_malloc(ctypes.byref(p), int(nbytes)))
        self.ptr = p.value
        self.nbytes = int(nbytes)
        self.shape = tuple(shape) if shape is not None else (int(nbytes),)
        self.dtype = np.dtype(dtype) if dtype is not None else np.dtype(np.uint8)

    @classmethod
    def empty(cls, shape, dtype, fill=None):
        import numpy as np
        n = int(np.prod(shape)) * np.dtype(dtype).itemsize
        b = cls(n, shape, dtype)
        if fill is not None:
            b.fill_bytes(fill)
        return b

    @classmethod
    def from_host(cls, arr, stream=None):
        import numpy as np
        arr = np.ascontiguousarray(arr)
        b = cls(arr.nbytes, arr.shape, arr.dtype)
        b.upload(arr, stream)
        return b

    def upload(self, arr, stream=None):
        import numpy as np
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        check(lib().sw_memcpy_htod(self.ptr, arr.ctypes.data, arr.nbytes, _stream(stream)))
        check(lib().sw_stream_synchronize(_stream(stream)))

    def download(self, stream=None):
        import numpy as np
        out = np.empty(self.shape, self.dtype)
        check(lib().sw_memcpy_dtoh(out.ctypes.data, self.ptr, out.nbytes, _stream(stream)))
        check(lib().sw_stream_synchronize(_stream(stream)))
        return out

    def fill_bytes(self, value: int, stream=None):
        check(lib().sw_memset(self.ptr, int(value), self.nbytes, _stream(stream)))
        check(lib().sw_stream_synchronize(_stream(stream)))

    def close(self):
        if self.ptr:
            lib().sw_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HostBuffer:
    """Pinned host allocation (sw_host_alloc) with a numpy view."""

    def __init__(self, nbytes: int):
        import numpy as np
        p = ctypes.c_void_p()
        check(lib().sw_host_alloc(ctypes.byref(p), int(nbytes)))
        self.ptr = p.value
        self.nbytes = int(nbytes)
        self.array = np.ctypeslib.as_array((ctypes.c_uint8 * self.nbytes).from_address(self.ptr))

    def close(self):
        if self.ptr:
            self.array = None
            lib().sw_host_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Stream:
    def __init__(self):
        p = ctypes.c_void_p()
        check(lib().sw_stream_create(ctypes.byref(p)))
        self.handle = p.value

    def synchronize(self):
        check(lib().sw_stream_synchronize(self.handle))

    def __del__(self):
        try:
            if self.handle:
                lib().sw_stream_destroy(self.handle)
        except Exception:
            pass


class Event:
    def __init__(self):
        p = ctypes.c_void_p()
        check(lib().sw_event_create(ctypes.byref(p)))
        self.handle = p.value

    def record(self, stream=None):
        check(lib().sw_event_record(self.handle, _stream(stream)))

    def elapsed_ms(self, end: "Event") -> float:
        ms = ctypes.c_float()
        check(lib().sw_event_elapsed_ms(ctypes.byref(ms), self.handle, end.handle))
        return ms.value

    def __del__(self):
        try:
            if self.handle:
                lib().sw_event_destroy(self.handle)
        except Exception:
            pass


def synchronize(stream=None):
    check(lib().sw_stream_synchronize(_stream(stream)))


def _ptr(t) -> int:
    if t is None:
        return 0
    return int(t.ptr) if isinstance(t, DeviceBuffer) else int(t)


def _stream(stream) -> int:
    if stream is None:
        return 0
    return int(stream.handle) if isinstance(stream, Stream) else int(stream)


def dims(nchans, nsamples, nantennas, nbeams, tscrunch, stokes="I", nbeamsets=1, ib_subtract=True) -> BfDims:
    return BfDims(nchans, nsamples, nantennas, nbeams, tscrunch, STOKES[stokes] if isinstance(stokes, str) else stokes,
                  nbeamsets, 1 if ib_subtract else 0)


def transpose(taftp, ftpa, nheaps, nant_in, nant_out, nchans, stream=None):
    check(lib().sw_transpose(_ptr(taftp), _ptr(ftpa), nheaps, nant_in, nant_out, nchans, _stream(stream)))


def microbench(which: str) -> float:
    """Measured device peak: 'mfma_i8' (TOP/s), 'hbm_copy' (GB/s), 'fma_f32' (TFLOP/s)."""
    code = {"mfma_i8": 0, "hbm_copy": 1, "fma_f32": 2}[which]
    r, s = ctypes.c_double(), ctypes.c_double()
    check(lib().sw_microbench(code, ctypes.byref(r), ctypes.byref(s), None))
    return r.value


def transpose_fpat(taftp, fpat, nheaps, nant_in, nant_out, nchans, stream=None):
    check(lib().sw_transpose_fpat(_ptr(taftp), _ptr(fpat), nheaps, nant_in, nant_out, nchans, _stream(stream)))


def transpose_fpat_stats(taftp, fpat, stats, nheaps, nant_in, nant_out, nchans, workspace=None, stream=None):
    """TAFTP -> FPAT fused with the statistics (one pass over the gulp)."""
    check(lib().sw_transpose_fpat_stats(_ptr(taftp), _ptr(fpat), _ptr(stats), nheaps, nant_in, nant_out, nchans,
                                        _ptr(workspace), _stream(stream)))


def transpose_fpat_stats_noalloc(taftp, fpat, stats, nheaps, nant_in, nant_out, nchans, stream=None):
    """sw_transpose_fpat_stats with its scratch in the library's fixed device
    array: no device allocation."""
    check(lib().sw_transpose_fpat_stats_noalloc(_ptr(taftp), _ptr(fpat), _ptr(stats), nheaps, nant_in, nant_out,
                                                nchans, _stream(stream)))


def corner_turn(src, dst, nchans, rows, cols, stream=None):
    check(lib().sw_corner_turn(_ptr(src), _ptr(dst), nchans, rows, cols, _stream(stream)))


def statistics_fpat(fpat, stats, nchans, nsamples, nantennas, workspace=None, stream=None):
    check(lib().sw_statistics_fpat(_ptr(fpat), _ptr(stats), nchans, nsamples, 2, nantennas, _ptr(workspace),
                                   _stream(stream)))


def statistics(ftpa, stats, nchans, nsamples, nantennas, workspace=None, stream=None):
    check(lib().sw_statistics(_ptr(ftpa), _ptr(stats), nchans, nsamples, 2, nantennas, _ptr(workspace),
                              _stream(stream)))


def ib_beamform(ftpa, raw, out, scale, offset, ant_weights, d: BfDims, stream=None):
    check(lib().sw_ib_beamform(_ptr(ftpa), _ptr(raw), _ptr(out), _ptr(scale), _ptr(offset), _ptr(ant_weights),
                               ctypes.byref(d), _stream(stream)))


def cb_beamform(ftpa, weights, scale, offset, bmap, ib_raw, out, d: BfDims, form="mfma", workspace=None,
                stream=None):
    check(lib().sw_cb_beamform(_ptr(ftpa), _ptr(weights), _ptr(scale), _ptr(offset), _ptr(bmap), _ptr(ib_raw),
                               _ptr(out), ctypes.byref(d), FORMS[form], _ptr(workspace), _stream(stream)))


def cb_workspace_size(d: BfDims) -> int:
    return int(lib().sw_cb_workspace_size(ctypes.byref(d)))


def cb_pack_weights(weights, packed, nchans, nbeams, nantennas, stream=None):
    check(lib().sw_cb_pack_weights(_ptr(weights), _ptr(packed), nchans, nbeams, nantennas, _stream(stream)))


def cb_beamform_packed(ftpa, packed, scale, offset, bmap, ib_raw, out, d: BfDims, stream=None):
    check(lib().sw_cb_beamform_packed(_ptr(ftpa), _ptr(packed), _ptr(scale), _ptr(offset), _ptr(bmap),
                                      _ptr(ib_raw), _ptr(out), ctypes.byref(d), _stream(stream)))


def cb_beamform_packed_tile(ftpa, packed, scale, offset, bmap, ib_raw, out, d: BfDims, waves, stream=None):
    """sw_cb_beamform_packed with `waves` (2, 4, 8) waves of 32 time rows per workgroup."""
    check(lib().sw_cb_beamform_packed_tile(_ptr(ftpa), _ptr(packed), _ptr(scale), _ptr(offset), _ptr(bmap),
                                           _ptr(ib_raw), _ptr(out), ctypes.byref(d), int(waves), _stream(stream)))


def tuning_lookup(nantennas, nbeams, nsamples, tscrunch, stokes="I", dms=(), path=None):
    """The tuning table's entry for a workload: None, or {"cb_waves"} and, when its DM
    list equals dms, {"fft_len", "share_forward"} too.  Host only."""
    import numpy as np
    d = np.ascontiguousarray(dms, np.float32)
    found, waves, share = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    fl = np.zeros(max(1, len(d)), np.int32)
    check(lib().sw_tuning_lookup(None if path is None else str(path).encode(), nantennas, nbeams, nsamples, tscrunch,
                                 STOKES[stokes] if isinstance(stokes, str) else stokes, len(d),
                                 d.ctypes.data if len(d) else None, ctypes.byref(found), ctypes.byref(waves),
                                 fl.ctypes.data, ctypes.byref(share)))
    if found.value == 0:
        return None
    out = {"cb_waves": waves.value}
    if found.value == 2:
        out.update(fft_len=[int(x) for x in fl[:len(d)]], share_forward=bool(share.value))
    return out


def pipeline_autotune(nantennas, nchans, nbeams, nsamples, tscrunch, dms, centre_frequency, bandwidth,
                      stokes="I", fft_len=None, tune_chans=8):
    """sw_pipeline_autotune: the tuning-table entry (a dict) measured on the device for
    this workload: beamformer waves per workgroup and, when fft_len is None, the FFT plan."""
    import json
    import numpy as np
    L = lib()
    if not hasattr(L, "_pipeline_declared"):
        _declare_pipeline(L)
        L._pipeline_declared = True
    d = np.ascontiguousarray(dms, np.float32)
    f = np.ascontiguousarray(fft_len, np.int32) if fft_len is not None else None
    ip = ctypes.POINTER(ctypes.c_int)
    cfg = PipelineConfigC(nantennas, nchans, nbeams, nsamples, tscrunch, STOKES[stokes], 1, 24.0, centre_frequency,
                          bandwidth, 4096, len(d), d.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                          f.ctypes.data_as(ip) if f is not None else None, None, FORMS["mfma"], 1)
    buf = ctypes.create_string_buffer(4096)
    check(L.sw_pipeline_autotune(ctypes.byref(cfg), int(tune_chans), buf, 4096))
    return json.loads(buf.value.decode())


def update_scalings(stats_np, beamset_weights_np, nbeamsets, cb_tscrunch, ib_tscrunch, output_level):
    """Host scalings from an FPA statistics array [F][P][A][4] float32."""
    import numpy as np
    st = np.ascontiguousarray(stats_np, np.float32)
    F, P, A = st.shape[:3]
    bw = np.ascontiguousarray(beamset_weights_np, np.float32)
    outs = [np.zeros(nbeamsets * F, np.float32) for _ in range(4)]
    check(lib().sw_update_scalings(st.ctypes.data, bw.ctypes.data, nbeamsets, F, P, A, cb_tscrunch, ib_tscrunch,
                                   float(output_level), *[o.ctypes.data for o in outs]))
    return dict(zip(("cb_offsets", "cb_scaling", "ib_offsets", "ib_scaling"), outs))


def synth_voltages(shape, seed, sigma=16.0, nthreads=0):
    """Seeded Gaussian complex int8 host array of `shape` (last axis = re/im)."""
    import numpy as np
    out = np.empty(shape, np.int8)
    assert out.shape[-1] == 2
    check(lib().sw_synth_gaussian_c8(out.ctypes.data, out.size // 2, seed, float(sigma), nthreads))
    return out


def synth_weights(shape, seed):
    """Random-phase complex int8 weights of amplitude 127 (last axis = re/im)."""
    import numpy as np
    out = np.empty(shape, np.int8)
    assert out.shape[-1] == 2
    check(lib().sw_synth_phase_weights(out.ctypes.data, out.size // 2, seed))
    return out


def generate_weights(delays, weights, chan_freqs, nantennas, nbeams, nchans, current_epoch, delay_epoch,
                     stream=None):
    """WeightsManager::weights: device [B][A] float3 delays -> [F][B][A] complex int8."""
    check(lib().sw_generate_weights(_ptr(delays), _ptr(weights), _ptr(chan_freqs), nantennas, nbeams, nchans,
                                    float(current_epoch), float(delay_epoch), _stream(stream)))


def channel_frequencies(centre_frequency, bandwidth, nchans):
    """PipelineConfig::channel_frequencies (Hz, FP64)."""
    import numpy as np
    out = np.zeros(nchans, np.float64)
    check(lib().sw_channel_frequencies(float(centre_frequency), float(bandwidth), nchans, out.ctypes.data))
    return out


def id_delays(dms, chan_freqs, tsamp):
    """DMSampleDelay table [ndms][nchans] and its maximum."""
    import numpy as np
    dms = np.ascontiguousarray(dms, np.float32)
    cf = np.ascontiguousarray(chan_freqs, np.float64)
    out = np.zeros((len(dms), len(cf)), np.int32)
    mx = ctypes.c_int()
    check(lib().sw_id_delays(dms.ctypes.data, len(dms), cf.ctypes.data, len(cf), float(tsamp), out.ctypes.data,
                             ctypes.byref(mx)))
    return out, mx.value


def ddplan_parse_block(spec):
    """DedispersionPlan::add_block(std::string) -> (coherent_dm, tscrunch, incoherent DMs float32)."""
    import numpy as np
    cdm, ts, n = ctypes.c_float(), ctypes.c_int(), ctypes.c_int()
    check(lib().sw_ddplan_parse_block(spec.encode(), ctypes.byref(cdm), ctypes.byref(ts), None, 0, ctypes.byref(n)))
    dms = np.zeros(n.value, np.float32)
    check(lib().sw_ddplan_parse_block(spec.encode(), ctypes.byref(cdm), ctypes.byref(ts), dms.ctypes.data, n.value,
                                      ctypes.byref(n)))
    return cdm.value, ts.value, dms


class IncoherentDedisperser:
    """IncoherentDedisperser(config, dms, tscrunch) over TFB device buffers (sw_idisp_* in the C ABI)."""

    def __init__(self, nchans, centre_frequency, bandwidth, cb_tscrunch, dms, tscrunch=1, chan_freqs=None,
                 valu_kernel=False):
        """valu_kernel: always use the 16-bit-lane VALU kernel instead of the matrix-core one."""
        import numpy as np
        self._dms = np.ascontiguousarray(dms, np.float32)
        self._cf = None if chan_freqs is None else np.ascontiguousarray(chan_freqs, np.float64)
        cfg = IdispConfig(nchans, float(centre_frequency), float(bandwidth),
                          None if self._cf is None else self._cf.ctypes.data, cb_tscrunch, self._dms.ctypes.data,
                          len(self._dms), tscrunch, 1 if valu_kernel else 0)
        h = ctypes.c_void_p()
        check(lib().sw_idisp_create(ctypes.byref(cfg), ctypes.byref(h)))
        self.kernel = lib().sw_idisp_kernel(h).decode()
        self._h = h
        self.nchans, self.ndms, self.tscrunch = nchans, len(self._dms), tscrunch
        md, sc = ctypes.c_int(), ctypes.c_float()
        self.delays = np.zeros((self.ndms, nchans), np.int32)
        check(lib().sw_idisp_info(h, ctypes.byref(md), self.delays.ctypes.data, ctypes.byref(sc)))
        self.max_delay, self.scale = md.value, sc.value

    def nout(self, nsamples):
        return (nsamples - self.max_delay) // self.tscrunch

    def dedisperse(self, tfb, nsamples, nbeams, nvec, tdb, stream=None, layout="tfb"):
        """layout "tfb": [T][F][B]; "ftb": channel-major [F][T][B]."""
        fn = lib().sw_idisp_dedisperse if layout == "tfb" else lib().sw_idisp_dedisperse_ftb
        check(fn(self._h, _ptr(tfb), nsamples, nbeams, nvec, _ptr(tdb), _stream(stream)))

    def close(self):
        if getattr(self, "_h", None):
            lib().sw_idisp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DelayModelFile:
    """DelayManager over a delay-model file (sw_delay_model_* in the C ABI)."""

    def __init__(self, path, nbeams, nantennas):
        self.nbeams, self.nantennas = nbeams, nantennas
        h = ctypes.c_void_p()
        check(lib().sw_delay_model_open(str(path).encode(), nbeams, nantennas, ctypes.byref(h)))
        self._h = h

    def beamsets(self):
        """(nbeamsets, beamset map [B] int32, beamset weights [S][A] float32)."""
        import numpy as np
        n = ctypes.c_int()
        check(lib().sw_delay_model_beamsets(self._h, ctypes.byref(n), None, None, 0))
        bmap = np.zeros(self.nbeams, np.int32)
        bw = np.zeros((n.value, self.nantennas), np.float32)
        check(lib().sw_delay_model_beamsets(self._h, ctypes.byref(n), bmap.ctypes.data, bw.ctypes.data, n.value))
        return n.value, bmap, bw

    def delays(self, epoch):
        """DelayManager::delays(epoch) -> ([B][A][3] float32 on the host, model start epoch)."""
        import numpy as np
        out = np.zeros((self.nbeams, self.nantennas, 3), np.float32)
        ep = ctypes.c_double()
        check(lib().sw_delay_model_get(self._h, float(epoch), out.ctypes.data, ctypes.byref(ep)))
        return out, ep.value

    def close(self):
        if getattr(self, "_h", None):
            lib().sw_delay_model_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class CoherentDedisperser:
    """Overlap-save coherent dedispersion plan (sw_codedisp_* in the C ABI)."""

    def __init__(self, nchans, nantennas, nsamples, low_freq, chan_bw, tsamp, dms, fft_len=None,
                 overlap=None, stream=None, input_fpat=False, history_in_place=False, share_forward=False,
                 chirp="reference"):
        """history_in_place (with input_fpat): the history is read from the
        gulp last given to push_history, which must stay unchanged until the
        next run.  share_forward: DMs of equal FFT length take the largest of
        their overlaps, so a batched run transforms each segment once for
        all of them.  chirp: "reference" (the reference's bin-to-frequency
        mapping) or "physical" (channel-centred, negative upper bins)."""
        import numpy as np
        self._dms = np.ascontiguousarray(dms, np.float32)
        n = len(self._dms)
        self._fft = np.ascontiguousarray(fft_len if fft_len is not None else [0] * n, np.int32)
        self._ov = np.ascontiguousarray(overlap if overlap is not None else [-1] * n, np.int32)
        cfg = CodedispConfig(nchans, nantennas, nsamples, low_freq, chan_bw, tsamp, n,
                             self._dms.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                             self._fft.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                             self._ov.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), 1 if input_fpat else 0,
                             1 if history_in_place else 0, 1 if share_forward else 0, CHIRP[chirp])
        h = ctypes.c_void_p()
        check(lib().sw_codedisp_create(ctypes.byref(cfg), _stream(stream), ctypes.byref(h)))
        self.handle = h.value
        self.ndms = n

    def info(self, d):
        """(FFT length, overlap) of DM d."""
        N, O, L = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(lib().sw_codedisp_info(self.handle, d, ctypes.byref(N), ctypes.byref(O), ctypes.byref(L)))
        return N.value, O.value

    @staticmethod
    def autotune(nchans, nantennas, nsamples, low_freq, chan_bw, tsamp, dms, overlap=None, tune_chans=8,
                 input_fpat=True, stream=None):
        """sw_codedisp_autotune: (FFT length per DM, share_forward, times [ndms + 1][6] in ms
        per gulp for lengths 1024..32768: one row per DM run alone, then the batched run of
        every DM a length fits; -1 where not tried)."""
        import numpy as np
        dm = np.ascontiguousarray(dms, np.float32)
        n = len(dm)
        ov = np.ascontiguousarray(overlap if overlap is not None else [-1] * n, np.int32)
        cfg = CodedispConfig(nchans, nantennas, nsamples, low_freq, chan_bw, tsamp, n,
                             dm.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), None,
                             ov.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), 1 if input_fpat else 0, 0)
        out = np.zeros(n, np.int32)
        share = ctypes.c_int(0)
        ms = np.zeros((n + 1, 6), np.float32)
        check(lib().sw_codedisp_autotune(ctypes.byref(cfg), tune_chans, out.ctypes.data, ctypes.byref(share),
                                         ms.ctypes.data, _stream(stream)))
        return [int(x) for x in out], bool(share.value), ms

    def latency(self, d):
        """Output latency of DM d in samples: a gulp g emits stream samples [gT - L, (g+1)T - L)."""
        L = ctypes.c_int()
        check(lib().sw_codedisp_info(self.handle, d, None, None, ctypes.byref(L)))
        return L.value

    def run(self, ftpa, outs, outs_f32=None, dm_begin=0, dm_count=None, batched=True, stream=None):
        n = self.ndms
        dm_count = n - dm_begin if dm_count is None else dm_count
        arr = (ctypes.c_void_p * n)(*[_ptr(o) if o is not None else None for o in outs])
        farr = None
        if outs_f32 is not None:
            farr = (ctypes.c_void_p * n)(*[_ptr(o) if o is not None else None for o in outs_f32])
        check(lib().sw_codedisp_run(self.handle, _ptr(ftpa), arr, farr, dm_begin, dm_count, 1 if batched else 0,
                                    _stream(stream)))

    def push_history(self, ftpa, stream=None):
        check(lib().sw_codedisp_push_history(self.handle, _ptr(ftpa), _stream(stream)))

    def set_history(self, hist, stream=None):
        check(lib().sw_codedisp_set_history(self.handle, _ptr(hist), _stream(stream)))

    def close(self):
        if getattr(self, "handle", None):
            lib().sw_codedisp_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PipelineConfigC(ctypes.Structure):
    _fields_ = [("nantennas", ctypes.c_int), ("nchans", ctypes.c_int), ("nbeams", ctypes.c_int),
                ("nsamples", ctypes.c_int), ("tscrunch", ctypes.c_int), ("stokes", ctypes.c_int),
                ("ib_subtract", ctypes.c_int), ("output_level", ctypes.c_float),
                ("centre_frequency", ctypes.c_double), ("bandwidth", ctypes.c_double),
                ("total_nchans", ctypes.c_int), ("ndms", ctypes.c_int),
                ("coherent_dms", ctypes.POINTER(ctypes.c_float)), ("fft_len", ctypes.POINTER(ctypes.c_int)),
                ("overlap", ctypes.POINTER(ctypes.c_int)), ("cb_form", ctypes.c_int), ("batched_dms", ctypes.c_int),
                ("id_enable", ctypes.c_int), ("id_tscrunch", ctypes.c_void_p), ("id_dms", ctypes.c_void_p),
                ("id_ndms", ctypes.c_void_p), ("id_dispatch_size", ctypes.c_int), ("autotune_fft", ctypes.c_int),
                ("share_forward", ctypes.c_int), ("async_sinks", ctypes.c_int), ("graph", ctypes.c_int),
                ("cb_waves", ctypes.c_int), ("tuning_file", ctypes.c_char_p), ("chirp", ctypes.c_int)]


class ObsHeaderC(ctypes.Structure):
    _fields_ = [("nantennas", ctypes.c_int), ("utc_start", ctypes.c_double), ("sample_clock", ctypes.c_double),
                ("utc_offset", ctypes.c_double)]


class BlockC(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("data", ctypes.c_void_p), ("on_device", ctypes.c_int),
                ("bytes", ctypes.c_size_t), ("ndims", ctypes.c_int), ("dims", ctypes.c_size_t * 4),
                ("order", ctypes.c_char * 8), ("dm_idx", ctypes.c_int), ("reference_dm", ctypes.c_float),
                ("tsamp", ctypes.c_double), ("utc_offset", ctypes.c_double),
                ("dms", ctypes.POINTER(ctypes.c_float)), ("ndms", ctypes.c_int)]


SINK_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(BlockC))


class _SinkErrors:
    """Exceptions raised by Python sinks inside C callbacks.  A callback that
    raises returns 1, so the library stops the gulp with SW_ERR_IO; the
    recorded exception is then re-raised by the call that ran the gulp."""

    def __init__(self):
        self.exc = None

    def wrap(self, fn):
        if fn is None:
            return SINK_FN()

        def _f(user, blk):
            try:
                fn(blk.contents)
                return 0
            except BaseException as e:  # noqa: BLE001 -- re-raised by check()
                if self.exc is None:
                    self.exc = e
                return 1
        return SINK_FN(_f)

    def check(self, rc):
        """check(rc), but a sink's own exception takes precedence."""
        exc, self.exc = self.exc, None
        if exc is not None:
            raise exc
        check(rc)


def _declare_pipeline(L):
    _pvp = ctypes.POINTER(ctypes.c_void_p)
    L.sw_pipeline_create.argtypes = [ctypes.POINTER(PipelineConfigC), _pvp]
    L.sw_pipeline_destroy.argtypes = [_vp]
    L.sw_pipeline_set_sinks.argtypes = [_vp, SINK_FN, _vp, SINK_FN, _vp, SINK_FN, _vp]
    L.sw_pipeline_set_weights.argtypes = [_vp, _vp]
    L.sw_pipeline_set_beamsets.argtypes = [_vp, _c, _vp, _vp]
    L.sw_pipeline_init.argtypes = [_vp, ctypes.POINTER(ObsHeaderC)]
    L.sw_pipeline_process.argtypes = [_vp, _vp, _sz, _c]
    L.sw_pipeline_prefetch.argtypes = [_vp, _vp, _sz]
    L.sw_pipeline_process_prefetched.argtypes = [_vp]
    L.sw_pipeline_synchronize.argtypes = [_vp]
    L.sw_pipeline_stage_times.argtypes = [_vp, ctypes.POINTER(ctypes.c_float), _c]
    L.sw_pipeline_set_delay_file.argtypes = [_vp, ctypes.c_char_p]
    L.sw_pipeline_set_tdb_sink.argtypes = [_vp, SINK_FN, _vp]
    L.sw_pipeline_stream.argtypes = [_vp]
    L.sw_pipeline_stream.restype = _vp
    L.sw_pipeline_tile.argtypes = [_vp, ctypes.POINTER(_c), _vp, ctypes.POINTER(_c), ctypes.POINTER(_c)]
    L.sw_pipeline_autotune.argtypes = [ctypes.POINTER(PipelineConfigC), _c, ctypes.c_char_p, _sz]


STAGES = ("transpose", "statistics", "dedispersion", "incoherent_beam", "coherent_beam", "incoherent_dedispersion",
          "weights", "other")


class BeamformerPipeline:
    """Native per-gulp pipeline (sw_pipeline_* in the C ABI)."""

    def __init__(self, nantennas, nchans, nbeams, nsamples, tscrunch, dms, centre_frequency, bandwidth,
                 stokes="I", ib_subtract=True, output_level=24.0, fft_len=None, overlap=None, form="mfma",
                 batched=True, total_nchans=4096, id_plan=None, id_dispatch_size=0, autotune_fft=False,
                 share_forward=False, async_sinks=False, graph=False, cb_waves=0, tuning_file=None,
                 chirp="reference"):
        """id_plan: None (no incoherent dedispersion) or, per coherent DM, (tscrunch, incoherent DMs).
        autotune_fft (with fft_len None): FFT lengths and share_forward chosen by sw_codedisp_autotune.
        share_forward: DMs of equal FFT length share one overlap (and, batched, their forward FFTs).
        async_sinks: sinks are called by a library thread while the pipeline goes on (blocks of each
        DM alternate between two buffers); a sink's exception is raised by the next call.
        graph (with async_sinks): each gulp after the first is captured into a hip graph and launched.
        cb_waves: MFMA beamformer waves per workgroup (2, 4, 8; 0: the tuning table's, else 4);
        tuning_file: the tuning table (None: the default one, "" none).
        chirp: the dedisperser's bin-to-frequency mapping, "reference" or "physical"."""
        import numpy as np
        L = lib()
        if not hasattr(L, "_pipeline_declared"):
            _declare_pipeline(L)
            L._pipeline_declared = True
        n = len(dms)
        self._dms = np.ascontiguousarray(dms, np.float32)
        self._fft = np.ascontiguousarray(fft_len, np.int32) if fft_len is not None else None
        self._ov = np.ascontiguousarray(overlap, np.int32) if overlap is not None else None
        ip = ctypes.POINTER(ctypes.c_int)
        cfg = PipelineConfigC(nantennas, nchans, nbeams, nsamples, tscrunch, STOKES[stokes], 1 if ib_subtract else 0,
                              output_level, centre_frequency, bandwidth, total_nchans, n,
                              self._dms.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                              self._fft.ctypes.data_as(ip) if self._fft is not None else None,
                              self._ov.ctypes.data_as(ip) if self._ov is not None else None,
                              FORMS[form], 1 if batched else 0)
        cfg.autotune_fft = 1 if autotune_fft else 0
        cfg.share_forward = 1 if share_forward else 0
        cfg.async_sinks = 1 if async_sinks else 0
        cfg.graph = 1 if graph else 0
        cfg.cb_waves = int(cb_waves)
        cfg.chirp = CHIRP[chirp]
        self._tuning_file = None if tuning_file is None else str(tuning_file).encode()
        cfg.tuning_file = self._tuning_file
        if id_plan is not None:
            assert len(id_plan) == n
            self._id_ts = np.array([b[0] for b in id_plan], np.int32)
            self._id_dms = [np.ascontiguousarray(b[1], np.float32) for b in id_plan]
            self._id_ptrs = (ctypes.c_void_p * n)(*[d.ctypes.data for d in self._id_dms])
            self._id_nd = np.array([len(d) for d in self._id_dms], np.int32)
            cfg.id_enable = 1
            cfg.id_tscrunch = self._id_ts.ctypes.data
            cfg.id_dms = ctypes.cast(self._id_ptrs, ctypes.c_void_p)
            cfg.id_ndms = self._id_nd.ctypes.data
            cfg.id_dispatch_size = id_dispatch_size
        h = ctypes.c_void_p()
        check(L.sw_pipeline_create(ctypes.byref(cfg), ctypes.byref(h)))
        self.handle = h.value
        self._errors = _SinkErrors()
        self._sinks = []
        self._tdb_sink = SINK_FN()

    def set_weights(self, weights_np):
        import numpy as np
        w = np.ascontiguousarray(weights_np, np.int8)
        check(lib().sw_pipeline_set_weights(self.handle, w.ctypes.data))

    def set_beamsets(self, bmap, antenna_weights):
        import numpy as np
        m = np.ascontiguousarray(bmap, np.int32)
        w = np.ascontiguousarray(antenna_weights, np.float32)
        check(lib().sw_pipeline_set_beamsets(self.handle, w.shape[0], m.ctypes.data, w.ctypes.data))

    def set_sinks(self, cb=None, ib=None, stats=None):
        """Python callables f(block: BlockC) -> None, invoked synchronously.
        An exception raised by a sink ends the gulp and is re-raised by the
        call that processed it."""
        wrap = self._errors.wrap
        self._sinks = [wrap(cb), wrap(ib), wrap(stats)]
        check(lib().sw_pipeline_set_sinks(self.handle, self._sinks[0], None, self._sinks[1], None,
                                          self._sinks[2], None))

    def set_delay_file(self, path):
        """Weights from a delay-model file, regenerated on the device every gulp."""
        check(lib().sw_pipeline_set_delay_file(self.handle, str(path).encode()))

    def set_tdb_sink(self, fn):
        self._tdb_sink = self._errors.wrap(fn)
        check(lib().sw_pipeline_set_tdb_sink(self.handle, self._tdb_sink, None))

    def init(self, nantennas, utc_start=1708082229.0, sample_clock=1712e6, utc_offset=0.0):
        hdr = ObsHeaderC(nantennas, utc_start, sample_clock, utc_offset)
        check(lib().sw_pipeline_init(self.handle, ctypes.byref(hdr)))

    def __call__(self, taftp, nbytes=None):
        """Process one gulp: a DeviceBuffer (device-resident) or a numpy array (host)."""
        if isinstance(taftp, DeviceBuffer):
            self._errors.check(lib().sw_pipeline_process(self.handle, taftp.ptr, nbytes or taftp.nbytes, 1))
        else:
            import numpy as np
            a = np.ascontiguousarray(taftp)
            self._errors.check(lib().sw_pipeline_process(self.handle, a.ctypes.data, a.nbytes, 0))
        return False

    def prefetch(self, taftp):
        """Start the host-to-device copy of a host gulp (a HostBuffer, or a numpy
        array kept alive and unchanged by the caller until it is processed)."""
        if isinstance(taftp, HostBuffer):
            check(lib().sw_pipeline_prefetch(self.handle, taftp.ptr, taftp.nbytes))
        else:
            check(lib().sw_pipeline_prefetch(self.handle, taftp.ctypes.data, taftp.nbytes))

    def process_prefetched(self):
        """Process the oldest prefetched gulp."""
        self._errors.check(lib().sw_pipeline_process_prefetched(self.handle))
        return False

    def synchronize(self):
        """Wait for every copy and kernel enqueued and, with asynchronous sinks,
        every block's delivery."""
        self._errors.check(lib().sw_pipeline_synchronize(self.handle))

    def stage_times(self):
        buf = (ctypes.c_float * len(STAGES))()
        check(lib().sw_pipeline_stage_times(self.handle, buf, len(STAGES)))
        return dict(zip(STAGES, list(buf)))

    @property
    def stream(self):
        return lib().sw_pipeline_stream(self.handle)

    def tile(self):
        """{"cb_waves", "fft_len", "share_forward", "source"} the pipeline runs with
        (source: "default", "table", "table+fft", "autotune")."""
        w, sh, src = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        fl = (ctypes.c_int * len(self._dms))()
        check(lib().sw_pipeline_tile(self.handle, ctypes.byref(w), fl, ctypes.byref(sh), ctypes.byref(src)))
        return {"cb_waves": w.value, "fft_len": list(fl), "share_forward": bool(sh.value),
                "source": ("default", "table", "table+fft", "autotune")[src.value]}

    def close(self):
        if getattr(self, "handle", None):
            lib().sw_pipeline_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def download_block(blk):
    """Copy a sink block (device memory) to a numpy array of its shape."""
    import numpy as np
    out = np.empty(blk.bytes, np.uint8)
    check(lib().sw_memcpy_dtoh_sync(out.ctypes.data, blk.data, blk.bytes))
    return out


class IdPipelineConfigC(ctypes.Structure):
    _fields_ = [("nchans", ctypes.c_int), ("centre_frequency", ctypes.c_double), ("bandwidth", ctypes.c_double),
                ("chan_freqs", ctypes.c_void_p), ("cb_tscrunch", ctypes.c_int), ("nbeams", ctypes.c_int),
                ("nvec", ctypes.c_int), ("nblocks", ctypes.c_int), ("coherent_dms", ctypes.c_void_p),
                ("tscrunch", ctypes.c_void_p), ("dms", ctypes.c_void_p), ("ndms", ctypes.c_void_p),
                ("dispatch_size", ctypes.c_int)]


class IncoherentDedispersionPipeline:
    """IncoherentDedispersionPipeline over device TFB blocks (sw_idpipe_* in the C ABI).

    blocks: list of (coherent_dm, tscrunch, incoherent_dms) -- one per plan block."""

    def __init__(self, nchans, centre_frequency, bandwidth, cb_tscrunch, nbeams, nvec, blocks, dispatch_size=0,
                 chan_freqs=None):
        import numpy as np
        L = lib()
        _c, _vp, _sz, _d = ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_double
        L.sw_idpipe_create.argtypes = [ctypes.POINTER(IdPipelineConfigC), ctypes.POINTER(ctypes.c_void_p)]
        L.sw_idpipe_destroy.argtypes = [_vp]
        L.sw_idpipe_set_sink.argtypes = [_vp, SINK_FN, _vp]
        L.sw_idpipe_info.argtypes = [_vp, _c, ctypes.POINTER(_c), ctypes.POINTER(_sz), ctypes.POINTER(_sz)]
        L.sw_idpipe_push.argtypes = [_vp, _c, _vp, _sz, _c, _d, _d, _vp]
        self._cdms = np.array([b[0] for b in blocks], np.float32)
        self._ts = np.array([b[1] for b in blocks], np.int32)
        self._dms = [np.ascontiguousarray(b[2], np.float32) for b in blocks]
        self._ptrs = (ctypes.c_void_p * len(blocks))(*[d.ctypes.data for d in self._dms])
        self._nd = np.array([len(d) for d in self._dms], np.int32)
        self._cf = None if chan_freqs is None else np.ascontiguousarray(chan_freqs, np.float64)
        cfg = IdPipelineConfigC(nchans, float(centre_frequency), float(bandwidth),
                                None if self._cf is None else self._cf.ctypes.data, cb_tscrunch, nbeams, nvec,
                                len(blocks), self._cdms.ctypes.data, self._ts.ctypes.data,
                                ctypes.cast(self._ptrs, ctypes.c_void_p), self._nd.ctypes.data, dispatch_size)
        h = ctypes.c_void_p()
        check(L.sw_idpipe_create(ctypes.byref(cfg), ctypes.byref(h)))
        self.handle = h
        self._errors = _SinkErrors()
        self._sink = SINK_FN()

    def info(self, block):
        md, ds, fl = ctypes.c_int(), ctypes.c_size_t(), ctypes.c_size_t()
        check(lib().sw_idpipe_info(self.handle, block, ctypes.byref(md), ctypes.byref(ds), ctypes.byref(fl)))
        return {"max_delay": md.value, "dispatch_size": ds.value, "filled": fl.value}

    def set_sink(self, fn):
        """fn(block: BlockC) -> None, invoked synchronously for every TDB block;
        an exception it raises is re-raised by the push that produced the block."""
        self._sink = self._errors.wrap(fn)
        check(lib().sw_idpipe_set_sink(self.handle, self._sink, None))

    def push(self, block, tfb, nsamples, utc_offset, tsamp, stream=None):
        import numpy as np
        if isinstance(tfb, np.ndarray):
            arr = np.ascontiguousarray(tfb)
            self._errors.check(lib().sw_idpipe_push(self.handle, block, arr.ctypes.data, nsamples, 0,
                                                    float(utc_offset), float(tsamp), _stream(stream)))
        else:
            self._errors.check(lib().sw_idpipe_push(self.handle, block, _ptr(tfb), nsamples, 1, float(utc_offset),
                                                    float(tsamp), _stream(stream)))

    def close(self):
        if getattr(self, "handle", None):
            lib().sw_idpipe_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
