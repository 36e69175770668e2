"""CPU oracle for the skyweaver BeamformerPipeline hot path (test infrastructure).

This package is the *checker*: a plain-C restatement of the reference's CPU
test oracles (see ``sw_oracle.h`` for the file:line each function follows and
what pins it).  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it; the product library never
does.

Arrays follow the reference layouts; complex int8 data is held as ``int8``
numpy arrays whose last axis is (re, im).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libsworacle.so")
_lib = None

STOKES = {"I": 0, "Q": 1, "U": 2, "V": 3, "IQUV": 4}


def build() -> str:
    """Compile the oracle shared library (gcc, -ffp-contract=off), always from
    the sources."""
    subprocess.run(["make", "-B", "-s", "-C", _HERE, "CC=gcc"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_i8p = ctypes.POINTER(ctypes.c_int8)
_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int)
_c = ctypes.c_int
_l = ctypes.c_long
_d = ctypes.c_double
_f = ctypes.c_float


def _declare(L):
    L.orc_transpose.argtypes = [_i8p, _i8p, _c, _c, _c, _c, _c, _c]
    L.orc_ib_beamform.argtypes = [_i8p, _f32p, _i8p, _c, _c, _c, _c, _f32p, _f32p, _f32p, _c, _c]
    L.orc_cb_beamform.argtypes = [_i8p, _i8p, _i8p, _c, _c, _c, _c, _c, _f32p, _f32p, _f32p, _i32p, _c, _c, _c]
    L.orc_statistics.argtypes = [_i8p, _f64p, _c, _c, _c, _c]
    L.orc_update_scalings.argtypes = [_f32p, _f32p, _c, _c, _c, _c, _c, _c, _f, _f32p, _f32p, _f32p, _f32p]
    L.orc_id_delays.argtypes = [_f32p, _c, _f64p, _c, _d, _i32p]
    L.orc_id_delays.restype = _c
    L.orc_incoherent_dedisperse.argtypes = [_i8p, _l, _c, _c, _c, _i32p, _c, _c, _c, _i8p, _c]
    L.orc_incoherent_dedisperse.restype = _c
    L.orc_fft.argtypes = [_f64p, _c, _c]
    L.orc_dm_chirp.argtypes = [_f, _d, _d, _c, _c, _f64p]
    L.orc_dm_chirp_physical.argtypes = [_f, _d, _d, _c, _c, _f64p]
    L.orc_codedisp_channel.argtypes = [_i8p, _l, _c, _f64p, _c, _c, _l, _f64p, _i8p, _c]
    L.orc_weights.argtypes = [_f32p, _i8p, _f64p, _c, _c, _c, _d, _d]
    L.orc_parse_beamsets.argtypes = [_f32p, _c, _c, _c, _c, _i32p, _f32p, _c]
    L.orc_parse_beamsets.restype = _c


def _p(a: np.ndarray, t):
    assert a.flags["C_CONTIGUOUS"], "oracle arrays must be C-contiguous"
    return a.ctypes.data_as(t)


def _as(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def transpose(taftp: np.ndarray, nant_out: int) -> np.ndarray:
    """taftp [H][A_in][F][256][P][2] int8 -> ftpa [F][H*256][P][A_out][2]."""
    taftp = _as(taftp, np.int8)
    H, A, F, S, P, _ = taftp.shape
    out = np.zeros((F, H * S, P, nant_out, 2), np.int8)
    lib().orc_transpose(_p(taftp, _i8p), _p(out, _i8p), H, A, nant_out, F, S, P)
    return out


def ib_beamform(ftpa, tscrunch, scale, offset, beamset_weights, nbeamsets, stokes="I"):
    """ftpa [F][T][2][A][2] -> (raw [S][T/ts][F](x4), out int8 same shape)."""
    ftpa = _as(ftpa, np.int8)
    F, T, P, A, _ = ftpa.shape
    nvec = 4 if stokes == "IQUV" else 1
    shape = (nbeamsets, T // tscrunch, F) + ((4,) if nvec == 4 else ())
    raw = np.zeros(shape, np.float32)
    out = np.zeros(shape, np.int8)
    scale, offset, bw = _as(scale, np.float32), _as(offset, np.float32), _as(beamset_weights, np.float32)
    lib().orc_ib_beamform(_p(ftpa, _i8p), _p(raw, _f32p), _p(out, _i8p), F, tscrunch, T, A,
                          _p(scale, _f32p), _p(offset, _f32p), _p(bw, _f32p), nbeamsets, STOKES[stokes])
    return raw, out


def cb_beamform(ftpa, weights, tscrunch, scales, offsets, antenna_weights, beamset_mapping,
                nbeamsets, stokes="I", ib_subtract=True):
    """ftpa [F][T][2][A][2], weights [F][B][A][2] -> TFB int8 [T/ts][F][B](x4)."""
    ftpa = _as(ftpa, np.int8)
    weights = _as(weights, np.int8)
    F, T, P, A, _ = ftpa.shape
    B = weights.shape[1]
    nvec = 4 if stokes == "IQUV" else 1
    out = np.zeros((T // tscrunch, F, B) + ((4,) if nvec == 4 else ()), np.int8)
    sc, of = _as(scales, np.float32), _as(offsets, np.float32)
    aw, bm = _as(antenna_weights, np.float32), _as(beamset_mapping, np.int32)
    lib().orc_cb_beamform(_p(ftpa, _i8p), _p(weights, _i8p), _p(out, _i8p), F, tscrunch, T, B, A,
                          _p(sc, _f32p), _p(of, _f32p), _p(aw, _f32p), _p(bm, _i32p), nbeamsets,
                          STOKES[stokes], 1 if ib_subtract else 0)
    return out


def statistics(ftpa) -> np.ndarray:
    """ftpa [F][T][P][A][2] -> [F][P][A][4] float64 {mean, std, skew, kurtosis}."""
    ftpa = _as(ftpa, np.int8)
    F, T, P, A, _ = ftpa.shape
    out = np.zeros((F, P, A, 4), np.float64)
    lib().orc_statistics(_p(ftpa, _i8p), _p(out, _f64p), F, T, P, A)
    return out


def update_scalings(stds, beamset_weights, nbeamsets, cb_tscrunch, ib_tscrunch, output_level):
    """stds [F][P][A] float32, beamset_weights [S][A] -> dict of [S*F] float32."""
    stds = _as(stds, np.float32)
    F, P, A = stds.shape
    bw = _as(beamset_weights, np.float32)
    outs = {k: np.zeros(nbeamsets * F, np.float32) for k in ("cb_offsets", "cb_scaling", "ib_offsets", "ib_scaling")}
    lib().orc_update_scalings(_p(stds, _f32p), _p(bw, _f32p), nbeamsets, F, P, A, cb_tscrunch, ib_tscrunch,
                              float(output_level), _p(outs["cb_offsets"], _f32p), _p(outs["cb_scaling"], _f32p),
                              _p(outs["ib_offsets"], _f32p), _p(outs["ib_scaling"], _f32p))
    return outs


def id_delays(dms, chan_freqs, tsamp):
    dms = _as(dms, np.float32)
    cf = _as(chan_freqs, np.float64)
    delays = np.zeros((len(dms), len(cf)), np.int32)
    maxd = lib().orc_id_delays(_p(dms, _f32p), len(dms), _p(cf, _f64p), len(cf), float(tsamp), _p(delays, _i32p))
    return delays, int(maxd)


def incoherent_dedisperse(tfb, delays, max_delay, tscrunch, nthreads=0):
    """tfb [T][F][B](x4) int8 -> tdb [(T-maxd)/ts][D][B](x4) int8."""
    tfb = _as(tfb, np.int8)
    T, F, B = tfb.shape[:3]
    nvec = tfb.shape[3] if tfb.ndim == 4 else 1
    delays = _as(delays, np.int32)
    D = delays.shape[0]
    nout = max((T - max_delay) // tscrunch, 0)
    out = np.zeros((nout, D, B) + ((nvec,) if nvec > 1 else ()), np.int8)
    rc = lib().orc_incoherent_dedisperse(_p(tfb, _i8p), T, F, B, nvec, _p(delays, _i32p), D, max_delay,
                                         tscrunch, _p(out, _i8p), nthreads)
    if rc == -1:
        raise RuntimeError("Fewer than max_sample_delay samples passed to dedisperse method")
    if rc == -2:
        raise RuntimeError("(nsamples - max_sample_delay) must be a multiple of tscrunch")
    return out


def fft(x: np.ndarray, inverse=False) -> np.ndarray:
    """Unnormalised FP64 FFT of a complex vector (power-of-two length)."""
    buf = np.ascontiguousarray(np.stack([np.real(x), np.imag(x)], -1), np.float64)
    lib().orc_fft(_p(buf, _f64p), len(x), 1 if inverse else 0)
    return buf[:, 0] + 1j * buf[:, 1]


def dm_chirp(dm, low_freq, coarse_chan_bw, nchans, fft_len, physical=False) -> np.ndarray:
    """FP64 chirp [nchans][N] complex128 (physical: orc_dm_chirp_physical's mapping)."""
    out = np.zeros((nchans, fft_len, 2), np.float64)
    fn = lib().orc_dm_chirp_physical if physical else lib().orc_dm_chirp
    fn(float(dm), float(low_freq), float(coarse_chan_bw), nchans, fft_len, _p(out, _f64p))
    return out[..., 0] + 1j * out[..., 1]


def codedisp_channel(tpa, chirp, N, O, nout, want_float=True, nthreads=0):
    """tpa [T][P][A][2] int8 -> (float [nout][P][A] complex128 or None, int8 [nout][P][A][2])."""
    tpa = _as(tpa, np.int8)
    nin = tpa.shape[0]
    pa = int(np.prod(tpa.shape[1:-1]))
    ch = np.ascontiguousarray(np.stack([np.real(chirp), np.imag(chirp)], -1), np.float64)
    out_i8 = np.zeros((nout,) + tpa.shape[1:], np.int8)
    out_f = np.zeros((nout,) + tpa.shape[1:], np.float64) if want_float else None
    lib().orc_codedisp_channel(_p(tpa, _i8p), nin, pa, _p(ch, _f64p), N, O, nout,
                               _p(out_f, _f64p) if want_float else None, _p(out_i8, _i8p), nthreads)
    if want_float:
        out_f = out_f[..., 0] + 1j * out_f[..., 1]
    return out_f, out_i8


def weights(delays, chan_freqs, current_epoch, delay_epoch):
    """delays [B][A][3] float32 -> [F][B][A][2] int8."""
    delays = _as(delays, np.float32)
    B, A, _ = delays.shape
    cf = _as(chan_freqs, np.float64)
    out = np.zeros((len(cf), B, A, 2), np.int8)
    lib().orc_weights(_p(delays, _f32p), _p(out, _i8p), _p(cf, _f64p), A, B, len(cf),
                      float(current_epoch), float(delay_epoch))
    return out


def parse_beamsets(model, nbeams_cfg, nant_cfg, max_beamsets=64):
    """model [B_m][A_m][3] float32 -> (nbeamsets, beamset_map [B_cfg], weights [S][A_cfg])."""
    model = _as(model, np.float32)
    Bm, Am, _ = model.shape
    bmap = np.zeros(nbeams_cfg, np.int32)
    bw = np.zeros((max_beamsets, nant_cfg), np.float32)
    n = lib().orc_parse_beamsets(_p(model, _f32p), Bm, Am, nbeams_cfg, nant_cfg, _p(bmap, _i32p),
                                 _p(bw, _f32p), max_beamsets)
    if n < 0:
        raise RuntimeError("too many beamsets")
    return n, bmap, bw[:n].copy()
