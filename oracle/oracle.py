"""ctypes binding of the CPU oracle (oracle/liboracle.so) and, where it has
been built, of the reference CPU core (oracle/_ref/libbfref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.  The product package
never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libbfref.so")

_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_i8p = np.ctypeslib.ndpointer(np.int8, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_sz = C.c_size_t


def build(reference: str | None = None) -> None:
    """Compile liboracle.so (and _ref/libbfref.so when the reference exists)."""
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    ref = reference or "/root/reference"
    if os.path.isdir(os.path.join(ref, "src")):
        subprocess.run(["make", "-s", "-C", HERE, "ref", f"REFERENCE={ref}"], check=True)


def _load():
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    sig = {
        "or_decode_nibble": (C.c_int, [C.c_uint]),
        "or_unpack_bytes": (None, [_u8p, _sz, _i8p, _i8p]),
        "or_validate_dimensions": (C.c_int, [_sz, _sz, _sz, _sz]),
        "or_default_positions": (C.c_int, [_sz, C.c_float, _f32p]),
        "or_regular_array": (C.c_int, [_sz, _sz, C.c_float, _f32p]),
        "or_channelized_frequencies": (C.c_int, [_sz, C.c_float, C.c_float, _f32p]),
        "or_direction_from_lm": (C.c_int, [C.c_float, C.c_float, _f32p]),
        "or_default_beam_grid": (C.c_int, [_sz, C.c_float, C.c_float, _f32p]),
        "or_rectangular_beam_grid": (C.c_int, [_sz, C.c_float, C.c_float, _f32p]),
        "or_generate_weights": (C.c_int, [_sz, _sz, _sz, _f32p, _f32p, _f32p, _f32p]),
        "or_generate_weights_f64": (C.c_int, [_sz, _sz, _sz, _f32p, _f32p, _f32p, _f64p]),
        "or_generate_tiled_weights": (C.c_int, [_sz, _sz, _sz, _f32p, _f32p, _f32p, _f32p]),
        "or_beamform_packed": (None, [_u8p, _f32p, _sz, _sz, _sz, _sz, _f32p]),
        "or_beamform_packed_omp": (None, [_u8p, _f32p, _sz, _sz, _sz, _sz, _f32p]),
        "or_beamform_packed_f64": (None, [_u8p, _f32p, _sz, _sz, _sz, _sz, _f64p]),
        "or_beamform_packed_f64w": (None, [_u8p, _f64p, _sz, _sz, _sz, _sz, _f64p]),
        "or_integrated_time_count": (_sz, [_sz, _sz]),
        "or_integrate": (None, [_f32p, _sz, _sz, _sz, _sz, _f32p]),
        "or_integrate_f64": (None, [_f64p, _sz, _sz, _sz, _sz, _f64p]),
        "or_quant_param_count": (_sz, [_sz, _sz, _sz]),
        "or_quantize": (None, [_f32p, _sz, _sz, _sz, _i8p, _f32p]),
        "or_dequantize": (None, [_i8p, _f32p, _sz, _sz, _sz, _f32p]),
        "or_make_noise": (None, [_sz, C.c_uint32, _u8p]),
        "or_make_point_source": (C.c_int, [_sz, _sz, _sz, _f32p, _f32p, _f32p, C.c_float, _u8p]),
        "or_hash64": (C.c_uint64, [C.c_uint64, C.c_uint64]),
        "or_make_uniform_int4": (None, [_sz, C.c_uint64, _u8p]),
        "or_tracker_window_count": (_sz, [_sz, _sz]),
        "or_tracker_direction": (C.c_int, [C.c_float, C.c_float, C.c_float, C.c_float, _sz, _f32p]),
        "or_beam_tracker": (C.c_int, [_u8p, _sz, _sz, _sz, C.c_float, C.c_float, C.c_float,
                                      C.c_float, _sz, _f32p]),
        "or_beam_tracker_omp": (C.c_int, [_u8p, _sz, _sz, _sz, C.c_float, C.c_float, C.c_float,
                                          C.c_float, _sz, _f32p]),
        "or_make_moving_point_source": (C.c_int, [_sz, _sz, _sz, _f32p, _f32p, C.c_float,
                                                  C.c_float, C.c_float, C.c_float, C.c_float,
                                                  _u8p]),
        "or_multibeam_tracker": (C.c_int, [_u8p, _sz, _sz, _sz, _sz, _sz, _f32p, _f32p, _f32p,
                                           _f32p, C.c_void_p]),
        "or_multibeam_tracker_spectra": (C.c_int, [_u8p, _sz, _sz, _sz, _sz, _sz, _f32p, _f32p,
                                                   _f32p, _f32p, C.c_void_p]),
        "or_covariance_i64": (None, [_u8p, _sz, _sz, _sz, _sz, _sz, _sz, _i64p]),
        "or_smooth_forward_i64": (None, [_i64p, _sz, _sz, _i64p]),
        "or_diagonal_load_f64": (None, [_f64p, _sz, C.c_double]),
        "or_cholesky_f64": (C.c_int, [_f64p, _sz]),
        "or_capon_from_factor_f64": (C.c_double, [_f64p, _f64p, _sz]),
        "or_bartlett_f64": (C.c_double, [_f64p, _f64p, _sz]),
        "or_quadratic_peak_offset": (C.c_double, [C.c_double, C.c_double, C.c_double]),
        "or_doa_update": (C.c_int, [_u8p, _sz, _sz, _sz, _sz, _sz, _sz, C.c_double, C.c_double,
                                    C.c_int, _f32p, _f32p, _sz, _f64p, _f64p, _f64p, _f64p,
                                    _i32p]),
        "or_max_rel_err_floored": (C.c_double, [_f32p, _f64p, _sz, C.c_double]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


LIB = _load()


def _check(rc, what):
    if rc != 0:
        raise ValueError(f"{what}: oracle rejected the input (code {rc})")


# ----------------------------------------------------------------- helpers
def default_positions(n_ant: int, spacing: float = 0.6) -> np.ndarray:
    out = np.zeros((n_ant, 3), np.float32)
    _check(LIB.or_default_positions(n_ant, spacing, out), "default_positions")
    return out


def regular_array(rows: int, cols: int, spacing: float) -> np.ndarray:
    out = np.zeros((rows * cols, 3), np.float32)
    _check(LIB.or_regular_array(rows, cols, spacing, out), "regular_array")
    return out


def channelized_frequencies(n: int = 336, start: float = 300e6, width: float = 300e3):
    out = np.zeros(n, np.float32)
    _check(LIB.or_channelized_frequencies(n, start, width, out), "channelized_frequencies")
    return out


def direction_from_lm(l: float, m: float) -> np.ndarray:
    out = np.zeros(3, np.float32)
    _check(LIB.or_direction_from_lm(l, m, out), "direction_from_lm")
    return out


def default_beam_grid(n: int, step: float = 0.02, m: float = 0.0) -> np.ndarray:
    out = np.zeros((n, 3), np.float32)
    _check(LIB.or_default_beam_grid(n, step, m, out), "default_beam_grid")
    return out


def rectangular_beam_grid(n_ant: int, spacing: float = 0.6, design: float = 400e6):
    out = np.zeros((n_ant, 3), np.float32)
    _check(LIB.or_rectangular_beam_grid(n_ant, spacing, design, out), "rectangular_beam_grid")
    return out


def generate_weights(pos, freqs, dirs) -> np.ndarray:
    pos = np.ascontiguousarray(pos, np.float32)
    freqs = np.ascontiguousarray(freqs, np.float32)
    dirs = np.ascontiguousarray(dirs, np.float32)
    B, F, A = len(dirs), len(freqs), len(pos)
    out = np.zeros((B, F, A, 2), np.float32)
    _check(LIB.or_generate_weights(B, F, A, pos, freqs, dirs, out), "generate_weights")
    return out


def generate_weights_f64(pos, freqs, dirs) -> np.ndarray:
    pos = np.ascontiguousarray(pos, np.float32)
    freqs = np.ascontiguousarray(freqs, np.float32)
    dirs = np.ascontiguousarray(dirs, np.float32)
    B, F, A = len(dirs), len(freqs), len(pos)
    out = np.zeros((B, F, A, 2), np.float64)
    LIB.or_generate_weights_f64(B, F, A, pos, freqs, dirs, out)
    return out


def generate_tiled_weights(pos, freqs, dirs) -> np.ndarray:
    pos = np.ascontiguousarray(pos, np.float32)
    freqs = np.ascontiguousarray(freqs, np.float32)
    dirs = np.ascontiguousarray(dirs, np.float32)
    B, F, A = len(dirs), len(freqs), len(pos)
    tiles = (B + 31) // 32
    out = np.zeros((F, tiles, A, 32, 2), np.float32)
    _check(LIB.or_generate_tiled_weights(B, F, A, pos, freqs, dirs, out), "tiled")
    return out


def beamform(x: np.ndarray, w: np.ndarray, T, F, A, B, threads: bool = False):
    x = np.ascontiguousarray(x, np.uint8).reshape(-1)
    w = np.ascontiguousarray(w, np.float32).reshape(-1)
    out = np.zeros((T, F, B), np.float32)
    (LIB.or_beamform_packed_omp if threads else LIB.or_beamform_packed)(x, w, T, F, A, B, out)
    return out


def beamform_f64(x, w, T, F, A, B):
    x = np.ascontiguousarray(x, np.uint8).reshape(-1)
    out = np.zeros((T, F, B), np.float64)
    if w.dtype == np.float64:
        LIB.or_beamform_packed_f64w(x, np.ascontiguousarray(w).reshape(-1), T, F, A, B, out)
    else:
        LIB.or_beamform_packed_f64(x, np.ascontiguousarray(w, np.float32).reshape(-1),
                                   T, F, A, B, out)
    return out


def integrated_time_count(T: int, S: int) -> int:
    return int(LIB.or_integrated_time_count(T, S))


def integrate(i: np.ndarray, T, F, B, S):
    W = integrated_time_count(T, S)
    out = np.zeros((W, F, B), np.float32)
    LIB.or_integrate(np.ascontiguousarray(i, np.float32).reshape(-1), T, F, B, S, out)
    return out


def integrate_f64(i: np.ndarray, T, F, B, S):
    W = integrated_time_count(T, S)
    out = np.zeros((W, F, B), np.float64)
    LIB.or_integrate_f64(np.ascontiguousarray(i, np.float64).reshape(-1), T, F, B, S, out)
    return out


def quantize(i: np.ndarray, Tw, F, B):
    codes = np.zeros((Tw, F, B), np.int8)
    params = np.zeros((int(LIB.or_quant_param_count(Tw, F, B)), 2), np.float32)
    LIB.or_quantize(np.ascontiguousarray(i, np.float32).reshape(-1), Tw, F, B, codes, params)
    return codes, params


def dequantize(codes, params, Tw, F, B):
    out = np.zeros((Tw, F, B), np.float32)
    LIB.or_dequantize(np.ascontiguousarray(codes, np.int8).reshape(-1),
                      np.ascontiguousarray(params, np.float32).reshape(-1), Tw, F, B, out)
    return out


def make_noise(T, F, A, seed: int) -> np.ndarray:
    out = np.zeros(T * F * A, np.uint8)
    LIB.or_make_noise(T * F * A, seed, out)
    return out.reshape(T, F, A)


def make_point_source(T, F, A, pos, freqs, direction, amplitude=4.0):
    out = np.zeros(T * F * A, np.uint8)
    _check(LIB.or_make_point_source(T, F, A, np.ascontiguousarray(pos, np.float32),
                                    np.ascontiguousarray(freqs, np.float32),
                                    np.ascontiguousarray(direction, np.float32),
                                    amplitude, out), "make_point_source")
    return out.reshape(T, F, A)


def make_uniform_int4(count: int, seed: int) -> np.ndarray:
    out = np.zeros(count, np.uint8)
    LIB.or_make_uniform_int4(count, seed, out)
    return out


def tracker_direction(l0, m0, rl, rm, t):
    out = np.zeros(3, np.float32)
    _check(LIB.or_tracker_direction(l0, m0, rl, rm, t, out), "tracker_direction")
    return out


def beam_tracker(x, T, F, A, l0, m0, rl, rm, S, threads=False):
    out = np.zeros((T, F, 1), np.float32)
    fn = LIB.or_beam_tracker_omp if threads else LIB.or_beam_tracker
    _check(fn(np.ascontiguousarray(x, np.uint8).reshape(-1), T, F, A, l0, m0, rl, rm, S, out),
           "beam_tracker")
    return out


def make_moving_point_source(T, F, A, pos, freqs, l0, m0, rl, rm, amplitude=4.0):
    out = np.zeros(T * F * A, np.uint8)
    _check(LIB.or_make_moving_point_source(T, F, A, np.ascontiguousarray(pos, np.float32),
                                           np.ascontiguousarray(freqs, np.float32),
                                           l0, m0, rl, rm, amplitude, out), "moving source")
    return out.reshape(T, F, A)


def multibeam_tracker(x, T, F, A, B, S, pos, freqs, dirs, want_f64=False, spectra=False):
    """dirs: [W][B][3] float32.  Returns float32 (and float64 oracle)."""
    W = (T + S - 1) // S
    shape = (T, F, B) if spectra else (W, F, B)
    out = np.zeros(shape, np.float32)
    out64 = np.zeros(shape, np.float64) if want_f64 else None
    fn = LIB.or_multibeam_tracker_spectra if spectra else LIB.or_multibeam_tracker
    _check(fn(np.ascontiguousarray(x, np.uint8).reshape(-1), T, F, A, B, S,
              np.ascontiguousarray(pos, np.float32), np.ascontiguousarray(freqs, np.float32),
              np.ascontiguousarray(dirs, np.float32).reshape(-1), out,
              out64.ctypes.data if want_f64 else None), "multibeam_tracker")
    return (out, out64) if want_f64 else out


def covariance_i64(x, T, F, A, t0, K, f):
    out = np.zeros((A, A, 2), np.int64)
    LIB.or_covariance_i64(np.ascontiguousarray(x, np.uint8).reshape(-1), T, F, A, t0, K, f, out)
    return out


def smooth_forward_i64(R, A, M):
    out = np.zeros((M, M, 2), np.int64)
    LIB.or_smooth_forward_i64(np.ascontiguousarray(R, np.int64), A, M, out)
    return out


def cholesky_f64(R):
    L = np.ascontiguousarray(R, np.float64).copy()
    ok = LIB.or_cholesky_f64(L, L.shape[0])
    return L, bool(ok)


def doa_update(x, T, F, A, t0, K, M, alpha, load, first, freqs, xpos, theta, R_state):
    nth = len(theta)
    capon = np.zeros((F, nth), np.float64)
    bart = np.zeros((F, nth), np.float64)
    ok = np.zeros(F, np.int32)
    _check(LIB.or_doa_update(np.ascontiguousarray(x, np.uint8).reshape(-1), T, F, A, t0, K, M,
                             alpha, load, int(first), np.ascontiguousarray(freqs, np.float32),
                             np.ascontiguousarray(xpos, np.float32), nth,
                             np.ascontiguousarray(theta, np.float64), R_state, capon, bart, ok),
           "doa_update")
    return capon, bart, ok


# ------------------------------------------------------- reference (_ref)
def ref_lib():
    """The reference CPU core compiled from its own sources, or None."""
    if not os.path.exists(REF_PATH):
        return None
    lib = C.CDLL(REF_PATH)
    sigs = {
        "ref_default_positions": [_sz, C.c_float, _f32p],
        "ref_channelized_frequencies": [_sz, C.c_float, C.c_float, _f32p],
        "ref_direction_from_lm": [C.c_float, C.c_float, _f32p],
        "ref_default_beam_grid": [_sz, C.c_float, C.c_float, _f32p],
        "ref_rectangular_beam_grid": [_sz, C.c_float, C.c_float, _f32p],
        "ref_fft_beam_grid": [_sz, _sz, _f32p],
        "ref_generate_weights": [_sz, _sz, _sz, _f32p, _f32p, _f32p, _f32p],
        "ref_generate_tiled_weights": [_sz, _sz, _sz, _f32p, _f32p, _f32p, _f32p],
        "ref_beamform_packed": [_u8p, _f32p, _sz, _sz, _sz, _sz, _f32p],
        "ref_integrate": [_f32p, _sz, _sz, _sz, _sz, _sz, _f32p],
        "ref_quantize": [_f32p, _sz, _sz, _sz, _sz, _i8p, _f32p],
        "ref_make_noise": [_sz, _sz, _sz, C.c_uint32, _u8p],
        "ref_make_point_source": [_sz, _sz, _sz, _f32p, _f32p, _f32p, C.c_float, _u8p],
        "ref_make_moving_point_source": [_sz, _sz, _sz, _f32p, _f32p, C.c_float, C.c_float,
                                         C.c_float, C.c_float, C.c_float, _u8p],
        "ref_beam_tracker": [_u8p, _sz, _sz, _sz, C.c_float, C.c_float, C.c_float, C.c_float,
                             _sz, _f32p],
        "ref_beam_tracker_opt_v2": [_u8p, _sz, _sz, _sz, C.c_float, C.c_float, C.c_float,
                                    C.c_float, _sz, _f32p],
        "ref_tracker_direction": [C.c_float, C.c_float, C.c_float, C.c_float, _sz, _f32p],
        "ref_write_quantized_metadata": [C.c_char_p, _sz, _sz, _sz, _sz, _sz, _sz, _sz,
                                         C.c_char_p],
        "ref_read_quantized_metadata": [C.c_char_p, C.POINTER(C.c_uint64)],
        "ref_write_packed_shard": [C.c_char_p, C.c_char_p, C.c_char_p, _u8p, _u8p, _sz, _sz,
                                   _sz, _sz, _sz, C.c_uint64, C.c_uint64],
        "ref_read_packed_shard": [C.c_char_p, C.c_char_p, C.c_char_p, _sz, _sz, _sz, _sz,
                                  _u8p, _u8p, C.POINTER(C.c_uint64)],
        "ref_write_tiled_weights_file": [C.c_char_p, _f32p, _sz, _sz, _sz, _sz],
    }
    for name, args in sigs.items():
        fn = getattr(lib, name)
        fn.restype = C.c_int
        fn.argtypes = args
    return lib
