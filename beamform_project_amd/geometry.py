"""Python mirror of the reference's geometry and weight producers
(include/beamformer/geometry.hpp, weights.hpp): array layouts, frequency
plans, beam grids (line, FFT-bin, rectangular, hexagonal), their text
loaders and the steering weights, through the C-ABI (bf_* geometry entry
points).  Directions and positions are float32 [n][3], (u, v) centres
[n][2]; weights are float32 [..][2] (interleaved complex).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from . import _capi
from ._capi import call
from .api import Dimensions, _ptr, _sh

_vp, _sz, _f = C.c_void_p, C.c_size_t, C.c_float
_pd = C.POINTER(_capi.Dims)
for _name, _args in {
    "bf_regular_array": [_sz, _sz, _f, _vp],
    "bf_default_positions": [_sz, _f, _vp],
    "bf_constant_frequencies": [_sz, _f, _vp],
    "bf_channelized_frequencies": [_sz, _f, _f, _vp],
    "bf_direction_from_lm": [_f, _f, _vp],
    "bf_default_beam_grid": [_sz, _f, _f, _vp],
    "bf_centered_integer_range": [_sz, _vp],
    "bf_select_fft_beam_centers_rectangular": [_f, _f, _sz, _sz, _sz, _vp, _vp],
    "bf_fft_beam_grid": [_sz, _sz, _f, _f, _vp],
    "bf_estimate_nbeams_hex_formula": [_f, _sz, _sz, _f, _f, _f, C.POINTER(_sz)],
    "bf_generate_hex_targets_cropped_fov": [_sz, _f, _f, _vp, C.POINTER(_sz)],
    "bf_rectangular_beam_grid": [_sz, _f, _f, _vp],
    "bf_load_positions": [C.c_char_p, _sz, _vp],
    "bf_load_frequencies": [C.c_char_p, _sz, _vp],
    "bf_load_directions": [C.c_char_p, _sz, _vp],
    "bf_generate_weights": [_pd, _vp, _vp, _vp, _vp, _sz],
    "bf_generate_tiled_weights": [_pd, _vp, _vp, _vp, _vp, _sz],
    "bf_generate_weights_device": [_vp, _pd, _vp, _vp, _vp, _vp, C.c_int],
}.items():
    _capi.register(_name, _args)

DEFAULT_SPACING_M = 0.6              # config.hpp:27
BEAM_GRID_DESIGN_FREQUENCY_HZ = 400e6  # config.hpp:26
TILED_WEIGHT_BEAM_TILE = 32          # indexing.hpp:52


def _f32(*shape):
    return np.zeros(shape, np.float32)


def regular_array(rows: int, columns: int, spacing_m: float = DEFAULT_SPACING_M) -> np.ndarray:
    out = _f32(rows * columns, 3)
    call("bf_regular_array", rows, columns, spacing_m, out.ctypes.data)
    return out


def default_positions(n_ant: int, spacing_m: float = DEFAULT_SPACING_M) -> np.ndarray:
    out = _f32(max(n_ant, 1), 3)
    call("bf_default_positions", n_ant, spacing_m, out.ctypes.data)
    return out


def constant_frequencies(n_freq: int = 336,
                         frequency_hz: float = BEAM_GRID_DESIGN_FREQUENCY_HZ) -> np.ndarray:
    out = _f32(max(n_freq, 1))
    call("bf_constant_frequencies", n_freq, frequency_hz, out.ctypes.data)
    return out


def channelized_frequencies(n_freq: int = 336, start_hz: float = 300e6,
                            channel_width_hz: float = 300e3) -> np.ndarray:
    out = _f32(max(n_freq, 1))
    call("bf_channelized_frequencies", n_freq, start_hz, channel_width_hz, out.ctypes.data)
    return out


def direction_from_lm(l: float, m: float) -> np.ndarray:
    out = _f32(3)
    call("bf_direction_from_lm", l, m, out.ctypes.data)
    return out


def default_beam_grid(n_beams: int, l_step: float = 0.02, m: float = 0.0) -> np.ndarray:
    out = _f32(max(n_beams, 1), 3)
    call("bf_default_beam_grid", n_beams, l_step, m, out.ctypes.data)
    return out


def centered_integer_range(count: int) -> list:
    out = np.zeros(max(count, 1), np.int64)
    call("bf_centered_integer_range", count, out.ctypes.data)
    return [int(v) for v in out]


@dataclass
class FftBeamSelection:
    """Reference geometry.hpp:15-21."""
    centers: np.ndarray
    n_u: int
    n_v: int
    n_bank_u: int
    n_bank_v: int


def select_fft_beam_centers_rectangular(du_fft_u: float, dv_fft_v: float, n_beams: int,
                                        m_side: int, n_side: int) -> FftBeamSelection:
    cen = _f32(max(n_beams, 1), 2)
    shape = np.zeros(4, np.uint64)
    call("bf_select_fft_beam_centers_rectangular", du_fft_u, dv_fft_v, n_beams, m_side, n_side,
         cen.ctypes.data, shape.ctypes.data)
    return FftBeamSelection(cen, *[int(v) for v in shape])


def fft_beam_grid(n_ant: int, n_beams: int, spacing_m: float = DEFAULT_SPACING_M,
                  design_frequency_hz: float = BEAM_GRID_DESIGN_FREQUENCY_HZ) -> np.ndarray:
    out = _f32(max(n_beams, 1), 3)
    call("bf_fft_beam_grid", n_ant, n_beams, spacing_m, design_frequency_hz, out.ctypes.data)
    return out


def estimate_nbeams_hex_formula(spacing_m: float, m_side: int, n_side: int, bw_e_deg: float,
                                bw_h_deg: float, wavelength_m: float) -> int:
    out = _sz()
    call("bf_estimate_nbeams_hex_formula", spacing_m, m_side, n_side, bw_e_deg, bw_h_deg,
         wavelength_m, C.byref(out))
    return out.value


def generate_hex_targets_cropped_fov(n_beams: int, u_max: float, v_max: float) -> np.ndarray:
    out = _f32(max(n_beams, 1), 2)
    n = _sz()
    call("bf_generate_hex_targets_cropped_fov", n_beams, u_max, v_max, out.ctypes.data,
         C.byref(n))
    return out[:n.value]


def rectangular_beam_grid(n_ant: int, spacing_m: float = DEFAULT_SPACING_M,
                          design_frequency_hz: float = BEAM_GRID_DESIGN_FREQUENCY_HZ):
    out = _f32(max(n_ant, 1), 3)
    call("bf_rectangular_beam_grid", n_ant, spacing_m, design_frequency_hz, out.ctypes.data)
    return out


def load_positions(path, expected_count: int) -> np.ndarray:
    out = _f32(max(expected_count, 1), 3)
    call("bf_load_positions", os.fsencode(path), expected_count, out.ctypes.data)
    return out[:expected_count]


def load_frequencies(path, expected_count: int) -> np.ndarray:
    out = _f32(max(expected_count, 1))
    call("bf_load_frequencies", os.fsencode(path), expected_count, out.ctypes.data)
    return out[:expected_count]


def load_directions(path, expected_count: int) -> np.ndarray:
    out = _f32(max(expected_count, 1), 3)
    call("bf_load_directions", os.fsencode(path), expected_count, out.ctypes.data)
    return out[:expected_count]


def _weight_args(positions, frequencies, directions):
    p = np.ascontiguousarray(positions, np.float32)
    f = np.ascontiguousarray(frequencies, np.float32)
    d = np.ascontiguousarray(directions, np.float32)
    return p, f, d


def generate_weights(dims: Dimensions, positions, frequencies, directions) -> np.ndarray:
    """[B][F][A][2] float32 (src/weights.cpp:11-60)."""
    p, f, d = _weight_args(positions, frequencies, directions)
    out = _f32(dims.n_beams, dims.n_freq, dims.n_ant, 2)
    c = dims.c()
    call("bf_generate_weights", C.byref(c), p.ctypes.data, f.ctypes.data, d.ctypes.data,
         out.ctypes.data, out.size // 2)
    return out


def tiled_weight_count(dims: Dimensions) -> int:
    tiles = -(-dims.n_beams // TILED_WEIGHT_BEAM_TILE)
    return dims.n_freq * tiles * dims.n_ant * TILED_WEIGHT_BEAM_TILE


def generate_tiled_weights(dims: Dimensions, positions, frequencies, directions) -> np.ndarray:
    """[F][ceil(B/32)][A][32][2] float32 (src/weights.cpp:63-122)."""
    p, f, d = _weight_args(positions, frequencies, directions)
    tiles = -(-dims.n_beams // TILED_WEIGHT_BEAM_TILE)
    out = _f32(dims.n_freq, tiles, dims.n_ant, TILED_WEIGHT_BEAM_TILE, 2)
    c = dims.c()
    call("bf_generate_tiled_weights", C.byref(c), p.ctypes.data, f.ctypes.data, d.ctypes.data,
         out.ctypes.data, out.size // 2)
    return out


def generate_weights_device(dims: Dimensions, d_positions, d_frequencies, d_directions,
                            d_weights, tiled: bool = False, stream=None) -> None:
    """The same weights written by a device kernel into d_weights."""
    c = dims.c()
    call("bf_generate_weights_device", _sh(stream), C.byref(c), _ptr(d_positions),
         _ptr(d_frequencies), _ptr(d_directions), _ptr(d_weights), int(tiled))
