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
        "or_doa_peaks": (_sz, [_f64p, _f64p, _sz, _sz, _f64p]),
        "or_doa_update": (C.c_int, [_u8p, _sz, _sz, _sz, _sz, _sz, _sz, C.c_double, C.c_double,
                                    C.c_int, _f32p, _f32p, _sz, _f64p, _f64p, _f64p, _f64p,
                                    _f64p, _i32p]),
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


def doa_update(x, T, F, A, t0, K, M, alpha, load, first, freqs, xpos, theta, R_state,
               load_state):
    """One C5 update.  R_state [F][M][M][2] (the unloaded recursion) and
    load_state [F] (the folded diagonal load) are float64 arrays updated in
    place; Capon uses R + load_state * I."""
    nth = len(theta)
    capon = np.zeros((F, nth), np.float64)
    bart = np.zeros((F, nth), np.float64)
    ok = np.zeros(F, np.int32)
    assert R_state.dtype == np.float64 and R_state.flags.c_contiguous and R_state.size == F * M * M * 2
    assert load_state.dtype == np.float64 and load_state.flags.c_contiguous and load_state.size == F
    _check(LIB.or_doa_update(np.ascontiguousarray(x, np.uint8).reshape(-1), T, F, A, t0, K, M,
                             alpha, load, int(first), np.ascontiguousarray(freqs, np.float32),
                             np.ascontiguousarray(xpos, np.float32), nth,
                             np.ascontiguousarray(theta, np.float64), R_state, load_state,
                             capon, bart, ok),
           "doa_update")
    return capon, bart, ok


def doa_peaks(P, theta, max_peaks):
    """[count][2] (angle, power) peak list of one broadband spectrum."""
    P = np.ascontiguousarray(P, np.float64)
    out = np.zeros((max(max_peaks, 1), 2), np.float64)
    n = LIB.or_doa_peaks(P, np.ascontiguousarray(theta, np.float64), len(P), max_peaks, out)
    return out[:n]


# ------------------------------------------------------- reference (_ref)
def host_has_avx512() -> bool:
    """True where /proc/cpuinfo lists the AVX-512 subsets x86-64-v4 needs."""
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("flags"):
                    have = set(ln.split(":", 1)[1].split())
                    return {"avx512f", "avx512bw", "avx512cd", "avx512dq",
                            "avx512vl"} <= have
    except OSError:
        pass
    return False


def host_cpu() -> dict:
    """CPU model, logical CPUs usable by this process, and sockets."""
    model, sockets = "unknown", set()
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name") and model == "unknown":
                    model = ln.split(":", 1)[1].strip()
                elif ln.startswith("physical id"):
                    sockets.add(ln.split(":", 1)[1].strip())
    except OSError:
        pass
    return {"model": model, "nproc": len(os.sched_getaffinity(0)),
            "sockets": max(1, len(sockets))}


def _ref_path(name: str, native: bool) -> str:
    """oracle/_ref/<name>, or its x86-64-v4 build when `native` and the host
    has AVX-512 (the nearest to the reference perf core's -march=native)."""
    if native and host_has_avx512():
        p = os.path.join(HERE, "_ref", "v4", name)
        if os.path.exists(p):
            return p
    return os.path.join(HERE, "_ref", name)


def ref_doa_lib(native: bool = True):
    """The reference's FP32 DOA primitives (ref_doa_shim.cpp), or None."""
    path = _ref_path("libbfref_doa.so", native)
    if not os.path.exists(path):
        return None
    lib = C.CDLL(path)
    fn = lib.ref_doa_update_f32
    fn.restype = C.c_int
    fn.argtypes = [_u8p, _sz, _sz, _sz, _sz, _sz, _sz, C.c_float, C.c_float, C.c_int, _f32p,
                   _f32p, _sz, _f64p, _f32p, _f32p, _f32p, _f32p]
    return lib


def ref_lib(native: bool = False):
    """The reference CPU core compiled from its own sources, or None.  The
    tests load the portable x86-64-v3 build; `native` (the CPU baselines)
    takes the AVX-512 build where the host supports it."""
    path = _ref_path("libbfref.so", native)
    if not os.path.exists(path):
        return None
    lib = C.CDLL(path)
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
        "ref_make_loss_mask": [_sz, _sz, _sz, C.c_uint32, C.c_float, _u8p],
        "ref_hex_targets": [_sz, C.c_float, C.c_float, _f32p, C.POINTER(_sz)],
        "ref_estimate_hex": [C.c_float, _sz, _sz, C.c_float, C.c_float, C.c_float,
                             C.POINTER(_sz)],
        "ref_select_fft": [C.c_float, C.c_float, _sz, _sz, _sz, _f32p, C.POINTER(_sz)],
        "ref_fft_beam_grid_at": [_sz, _sz, C.c_float, C.c_float, _f32p],
        "ref_cpu_opt_tracker": [_u8p, _sz, _sz, _sz, _sz, _f32p, _f32p, _f32p, _f32p, _sz,
                                C.c_int, _sz, _sz, C.c_float, C.c_float, _sz, C.c_float,
                                C.c_float, C.c_int, _f32p, _f32p],
    }
    for name, args in sigs.items():
        fn = getattr(lib, name)
        fn.restype = C.c_int
        fn.argtypes = args
    return lib


def ref_opt_tracker(lib, x, n_calls, T, F, A, pos, freqs, start, rate, S, estimator=0, G=1,
                    levels=0, fov_l=0.2, fov_m=0.2, p_sub=0, lam=1.0, load=1e-3, interp=True):
    """The reference's CpuOptBeamTracker (src/cpu_opt_beam_tracker.cpp:1619-2457)
    over n_calls consecutive chunks of T spectra: (intensity [n][T][F],
    directions [n][W][3])."""
    W = (T + S - 1) // S
    out = np.zeros((n_calls, T, F), np.float32)
    dirs = np.zeros((n_calls, W, 3), np.float32)
    _check(lib.ref_cpu_opt_tracker(np.ascontiguousarray(x, np.uint8).reshape(-1), n_calls, T, F,
                                   A, np.ascontiguousarray(pos, np.float32),
                                   np.ascontiguousarray(freqs, np.float32),
                                   np.ascontiguousarray(start, np.float32),
                                   np.ascontiguousarray(rate, np.float32), S, int(estimator), G,
                                   levels, fov_l, fov_m, p_sub, lam, load, int(interp), out,
                                   dirs), "ref_cpu_opt_tracker")
    return out, dirs


# ------------------------------------------ adaptive planar DOA search
# Restatement of CpuOptBeamTracker::run_into's searching path
# (src/cpu_opt_beam_tracker.cpp:1619-2457) in FP64 numpy, with the
# reference's float32 grid and decision arithmetic.  fb_in_place=True
# reproduces the reference's forward-backward step, which reads entries it
# has already averaged (:647-656); False is the textbook average
# (R + J conj(R) J) / 2 of :486-498.
def _f32(v):
    return np.float32(v)


def _clamp_dir(l, m):
    """Horizon clamp and direction_from_lm (:2350-2357, src/geometry.cpp:97-106)."""
    l, m = _f32(l), _f32(m)
    r2 = l * l + m * m
    if r2 > _f32(0.999):
        s = np.sqrt(_f32(0.999) / r2)
        l, m = l * s, m * s
    return np.array([l, m, np.sqrt(_f32(1.0) - (l * l + m * m))], np.float32)


def _coarse_cells(c, G, fov_l, fov_m, coplanar):
    """build_centred_grid (:798-828) and the steered directions (:1971-1993)."""
    fl, fm = _f32(fov_l), _f32(fov_m)
    step_l = _f32(0.0) if G == 1 else (_f32(2.0) * fl) / _f32(G - 1)
    step_m = _f32(0.0) if G == 1 else (_f32(2.0) * fm) / _f32(G - 1)
    sl = 0.0 if G == 1 else (2.0 * float(fl)) / (G - 1)
    sm = 0.0 if G == 1 else (2.0 * float(fm)) / (G - 1)
    mid = (G - 1) * 0.5
    dirs, steer = [], []
    for v in range(G):
        for u in range(G):
            d = _clamp_dir((c[0] - fl) + _f32(u) * step_l, (c[1] - fm) + _f32(v) * step_m)
            st = d.astype(np.float64)
            if coplanar:
                nl = float(c[0]) + (u - mid) * sl
                nm = float(c[1]) + (v - mid) * sm
                if not (nl * nl + nm * nm > 0.999):
                    st = np.array([nl, nm, 0.0])
            dirs.append(d)
            steer.append(st)
    return np.array(dirs), np.array(steer)


def _refine_cells(c, G, fov_l, fov_m, level, coplanar):
    """3 x 3 candidates of refinement `level` (:2179-2233)."""
    hl, hm = _f32(fov_l) / _f32(G), _f32(fov_m) / _f32(G)
    for _ in range(level):
        hl, hm = hl * _f32(0.5), hm * _f32(0.5)
    dirs, steer = [], []
    for dv in (-1, 0, 1):
        for du in (-1, 0, 1):
            l, m = c[0] + _f32(du) * hl, c[1] + _f32(dv) * hm
            clamped = l * l + m * m > _f32(0.999)
            d = _clamp_dir(l, m)
            if coplanar and not clamped:
                st = np.array([float(c[0]) + du * float(hl), float(c[1]) + dv * float(hm), 0.0])
            else:
                st = d.astype(np.float64)
            dirs.append(d)
            steer.append(st)
    return np.array(dirs), np.array(steer), hl, hm


def _quadratic_interp(pc, p_up, p_down, p_vp, p_vm):
    """quadratic_interp (:830-853), float32."""
    pc, p_up, p_down, p_vp, p_vm = map(_f32, (pc, p_up, p_down, p_vp, p_vm))
    du = dv = _f32(0.0)
    den = p_down - _f32(2.0) * pc + p_up
    if den > 0:
        du = np.clip((_f32(0.5) * (p_up - p_down)) / den, _f32(-0.5), _f32(0.5))
    den = p_vm - _f32(2.0) * pc + p_vp
    if den > 0:
        dv = np.clip((_f32(0.5) * (p_vp - p_vm)) / den, _f32(-0.5), _f32(0.5))
    return _f32(du), _f32(dv)


def _decode(x):
    """int4 pairs -> complex128 (low nibble real, high nibble imaginary)."""
    x = np.asarray(x, np.uint8).astype(np.int16)
    return (((x & 15) ^ 8) - 8) + 1j * ((((x >> 4) & 15) ^ 8) - 8)


def opt_tracker_search(x, S, pos, freqs, start, rate, estimator=0, G=16, levels=0,
                       fov_l=0.2, fov_m=0.2, p_sub=0, lam=1.0, load=1e-3, interp=True,
                       state=None, fb_in_place=False):
    """Window directions [W][3] of one run_into call over x [T][F][A], and
    the covariance state {"R": [F][M][M] complex} carried to the next call
    (pass it back as `state`)."""
    x = np.asarray(x)
    T, F, A = x.shape
    M = p_sub if (p_sub != 0 and p_sub < A) else A
    L = A - M + 1
    smoothing = M < A
    capon = estimator == 1
    pos = np.asarray(pos, np.float64)
    coplanar = bool(np.all(np.abs(np.asarray(pos, np.float32)[:M, 2]) <= np.float32(1e-6)))
    k = 2.0 * np.pi * np.asarray(freqs, np.float32).astype(np.float64) / 299792458.0
    state = {} if state is None else state
    W = (T + S - 1) // S
    prior = _clamp_dir(start[0], start[1])  # window 0 of the trajectory
    prev = prior
    dirs = np.zeros((W, 3), np.float32)
    lam32 = _f32(lam)
    for w in range(W):
        xs = _decode(x[w * S:min(T, w * S + S)])  # [K][F][A]
        K = xs.shape[0]
        Rf = np.einsum("kfa,kfb->fab", xs, np.conj(xs))  # [F][A][A]
        Rs = sum(Rf[:, l:l + M, l:l + M] for l in range(L)) / (K * L)
        if smoothing:
            if fb_in_place:
                for f in range(F):
                    R = Rs[f]
                    for r in range(M):
                        for c in range(M):
                            R[r, c] = 0.5 * (R[r, c] + np.conj(R[M - 1 - r, M - 1 - c]))
            else:
                Rs = 0.5 * (Rs + np.conj(Rs[:, ::-1, ::-1]))
        if capon and load > 0:
            tr = np.trace(Rs, axis1=1, axis2=2).real
            Rs = Rs + (float(load) * tr / M)[:, None, None] * np.eye(M)
        if "R" not in state or lam32 >= 1:
            state["R"] = Rs
        else:
            state["R"] = float(lam32) * state["R"] + float(_f32(1.0) - lam32) * Rs
        R = state["R"]
        chol = []
        if capon:
            for f in range(F):
                try:
                    Lf = np.linalg.cholesky(R[f])
                    ok = bool(np.all(np.diag(Lf).real ** 2 > 1e-12))
                except np.linalg.LinAlgError:
                    ok = False
                chol.append(np.linalg.inv(Lf) if ok else None)

        def powers(steer):
            ph = k[:, None, None] * (steer[None, :, :] @ pos[:M].T)  # [F][cells][M]
            a = np.exp(-1j * ph)
            if not capon:
                p = np.einsum("fca,fab,fcb->fc", np.conj(a), R, a, optimize=True).real
            else:
                p = np.zeros(a.shape[:2])
                for f in range(F):
                    if chol[f] is None:
                        p[f] = np.einsum("ca,ab,cb->c", np.conj(a[f]), R[f], a[f]).real
                    else:
                        v = a[f] @ chol[f].T
                        n2 = np.sum(np.abs(v) ** 2, axis=1)
                        p[f] = np.where(n2 > 1e-12, 1.0 / n2, 0.0)
            p[~np.isfinite(p)] = 0.0
            return p.sum(axis=0).astype(np.float32)

        cd, cs = _coarse_cells(prev, G, fov_l, fov_m, coplanar)
        P = powers(cs)
        best = int(np.argmax(P))
        centre = cd[best]
        if interp and levels == 0:
            bu, bv = best % G, best // G

            def at(u, v):
                uu, vv = bu + u, bv + v
                return P[vv * G + uu] if 0 <= uu < G and 0 <= vv < G else 0.0
            du, dv = _quadratic_interp(P[best], at(1, 0), at(-1, 0), at(0, 1), at(0, -1))
            step_l = _f32(0.0) if G == 1 else (_f32(2.0) * _f32(fov_l)) / _f32(G - 1)
            step_m = _f32(0.0) if G == 1 else (_f32(2.0) * _f32(fov_m)) / _f32(G - 1)
            centre = _clamp_dir(centre[0] + du * step_l, centre[1] + dv * step_m)
        for level in range(1, levels + 1):
            rd, rs, hl, hm = _refine_cells(centre, G, fov_l, fov_m, level, coplanar)
            P = powers(rs)
            centre = rd[int(np.argmax(P))]
            if interp:
                du, dv = _quadratic_interp(P[4], P[5], P[3], P[7], P[1])
                centre = _clamp_dir(centre[0] + du * hl, centre[1] + dv * hm)
        prev = centre
        dirs[w] = centre
    return dirs, state
