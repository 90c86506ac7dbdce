"""Python mirror of the reference's adaptive DOA tracker
(include/beamformer/cpu_opt_beam_tracker.hpp: TrackerEstimator,
CpuOptTrackerConfig, CpuOptBeamTracker, cpu_opt_beam_tracker_packed_intensity)
on the GPU, through the C-ABI bf_opt_tracker_*.

Names: OptTrackerConfig (CpuOptTrackerConfig), OptBeamTracker
(CpuOptBeamTracker), opt_beam_tracker_packed_intensity
(cpu_opt_beam_tracker_packed_intensity).  The search runs on the device; the
Python layer marshals arguments only.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from enum import IntEnum
from typing import Optional

import numpy as np

from . import _capi
from ._capi import call
from .api import DeviceArray, Dimensions, Kernel, _ptr, _sh
from .tracker import TrackerConfig, TrackerConfigC

_vp, _sz = C.c_void_p, C.c_size_t


class TrackerEstimator(IntEnum):
    """Reference cpu_opt_beam_tracker.hpp:39-43."""
    BARTLETT = 0
    CAPON = 1


class OptTrackerConfigC(C.Structure):
    _fields_ = [("estimator", C.c_int), ("coarse_grid_resolution", _sz),
                ("refinement_levels", _sz), ("search_fov_l", C.c_float),
                ("search_fov_m", C.c_float), ("spatial_smoothing_subarray_size", _sz),
                ("forgetting_factor", C.c_float), ("diagonal_load", C.c_float),
                ("enable_quadratic_peak_interp", C.c_int), ("integration_spectra", _sz),
                ("kernel", C.c_int)]


_pd = C.POINTER(_capi.Dims)
_po = C.POINTER(OptTrackerConfigC)
for _name, _args in {
    "bf_opt_tracker_default_config": [_po],
    "bf_opt_tracker_create": [_vp, _vp, _pd, _po, C.POINTER(_vp)],
    "bf_opt_tracker_destroy": [_vp],
    "bf_opt_tracker_seed_trajectory": [_vp, _vp, _vp],
    "bf_opt_tracker_reset": [_vp],
    "bf_opt_tracker_run_device": [_vp, _vp, _vp, _vp],
    "bf_opt_tracker_run": [_vp, _vp, _sz, _vp, _sz],
    "bf_opt_tracker_window_directions": [_vp, _vp, _sz, C.POINTER(_sz)],
    "bf_opt_tracker_packed_intensity": [_vp, _sz, _pd, C.POINTER(TrackerConfigC), _po, _vp,
                                        _sz],
}.items():
    _capi.register(_name, _args)


@dataclass
class OptTrackerConfig:
    """Reference CpuOptTrackerConfig (cpu_opt_beam_tracker.hpp:50-88); the
    defaults reproduce the open-loop tracker.  `kernel` computes the emitted
    intensities (AUTO = SPLIT_F16)."""
    estimator: TrackerEstimator = TrackerEstimator.BARTLETT
    coarse_grid_resolution: int = 1
    refinement_levels: int = 0
    search_fov_l: float = 0.2
    search_fov_m: float = 0.2
    spatial_smoothing_subarray_size: int = 0
    forgetting_factor: float = 1.0
    diagonal_load: float = 1.0e-3
    enable_quadratic_peak_interp: bool = True
    integration_spectra: int = 320
    kernel: Kernel = Kernel.AUTO

    def c(self) -> OptTrackerConfigC:
        return OptTrackerConfigC(int(self.estimator), self.coarse_grid_resolution,
                                 self.refinement_levels, self.search_fov_l, self.search_fov_m,
                                 self.spatial_smoothing_subarray_size, self.forgetting_factor,
                                 self.diagonal_load, int(self.enable_quadratic_peak_interp),
                                 self.integration_spectra, int(self.kernel))


class OptBeamTracker:
    """Reference CpuOptBeamTracker (cpu_opt_beam_tracker.hpp:94-196)."""

    def __init__(self, positions: Optional[np.ndarray], frequencies: Optional[np.ndarray],
                 dims: Dimensions, config: OptTrackerConfig):
        self._h = None
        self.dims = dims
        self.config = config
        pos = None if positions is None else np.ascontiguousarray(positions, np.float32)
        fr = None if frequencies is None else np.ascontiguousarray(frequencies, np.float32)
        d, c = dims.c(), config.c()
        h = C.c_void_p()
        call("bf_opt_tracker_create", None if pos is None else pos.ctypes.data,
             None if fr is None else fr.ctypes.data, C.byref(d), C.byref(c), C.byref(h))
        self._h = h.value

    def close(self):
        if self._h is not None:
            call("bf_opt_tracker_destroy", self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def seed_trajectory(self, direction_start, direction_rate_per_sample=(0.0, 0.0)) -> None:
        s = (C.c_float * 3)(*direction_start)
        r = (C.c_float * 2)(*direction_rate_per_sample)
        call("bf_opt_tracker_seed_trajectory", self._h, s, r)

    def reset(self) -> None:
        call("bf_opt_tracker_reset", self._h)

    def run_into(self, packed) -> np.ndarray:
        """run_into on host data: packed [T][F][A] uint8 -> [T][F][1] f32."""
        x = np.ascontiguousarray(packed, np.uint8)
        out = np.zeros((self.dims.n_time, self.dims.n_freq, 1), np.float32)
        call("bf_opt_tracker_run", self._h, x.ctypes.data, x.size, out.ctypes.data, out.size)
        return out

    def run_device(self, d_packed: DeviceArray, d_intensity: DeviceArray, stream=None) -> None:
        call("bf_opt_tracker_run_device", self._h, _sh(stream), _ptr(d_packed),
             _ptr(d_intensity))

    def window_directions(self) -> np.ndarray:
        n = _sz()
        call("bf_opt_tracker_window_directions", self._h, None, 0, C.byref(n))
        out = np.zeros((n.value, 3), np.float32)
        call("bf_opt_tracker_window_directions", self._h, out.ctypes.data, n.value, C.byref(n))
        return out

    def window_direction(self, window: int) -> tuple:
        d = self.window_directions()
        return tuple(float(v) for v in d[window]) if window < len(d) else (0.0, 0.0, 1.0)


def default_config() -> OptTrackerConfig:
    c = OptTrackerConfigC()
    call("bf_opt_tracker_default_config", C.byref(c))
    return OptTrackerConfig(TrackerEstimator(c.estimator), c.coarse_grid_resolution,
                            c.refinement_levels, c.search_fov_l, c.search_fov_m,
                            c.spatial_smoothing_subarray_size, c.forgetting_factor,
                            c.diagonal_load, bool(c.enable_quadratic_peak_interp),
                            c.integration_spectra, Kernel(c.kernel))


def opt_beam_tracker_packed_intensity(packed, dims: Dimensions, trajectory: TrackerConfig,
                                      opt: OptTrackerConfig | None = None) -> np.ndarray:
    """Reference cpu_opt_beam_tracker_packed_intensity (src/cpu_opt_beam_tracker.cpp:
    2459-2486): default geometry, a fresh covariance recursion."""
    x = np.ascontiguousarray(packed, np.uint8)
    out = np.zeros((dims.n_time, dims.n_freq, 1), np.float32)
    d, t, c = dims.c(), trajectory.c(), (opt or OptTrackerConfig()).c()
    call("bf_opt_tracker_packed_intensity", x.ctypes.data, x.size, C.byref(d), C.byref(t),
         C.byref(c), out.ctypes.data, out.size)
    return out
