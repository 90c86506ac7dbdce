"""Python mirror of the reference's tracker API
(include/beamformer/beam_tracker.hpp, cuda_beam_tracker_v5.hpp) and of the
build's multi-beam tracker, driving libbeamform_hip.so through its C-ABI.

Names: beam_tracker_into (cuda_beam_tracker_v5_into),
beam_tracker_device_resident (cuda_beam_tracker_v5_device_resident),
BatchedTrackerStream (BatchedTrackerStreamV5).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _capi
from ._capi import call
from .api import DeviceArray, Dimensions, Kernel, _ptr, _sh


class TrackerConfigC(C.Structure):
    _fields_ = [("direction_start", C.c_float * 3),
                ("direction_rate_per_sample", C.c_float * 2),
                ("integration_spectra", C.c_size_t)]


class TrackerExecC(C.Structure):
    _fields_ = [("kernel", C.c_int), ("enable_graph", C.c_int)]


class MultibeamArgsC(C.Structure):
    _fields_ = [("n_time", C.c_size_t), ("n_freq", C.c_size_t), ("n_ant", C.c_size_t),
                ("n_beams", C.c_size_t), ("integration_spectra", C.c_size_t),
                ("positions", C.c_void_p), ("frequencies", C.c_void_p),
                ("directions", C.c_void_p), ("trajectories", C.c_void_p),
                ("window_offset", C.c_size_t), ("integrate", C.c_int), ("kernel", C.c_int)]


_vp, _sz = C.c_void_p, C.c_size_t
_pd = C.POINTER(_capi.Dims)
_pc = C.POINTER(TrackerConfigC)
_pe = C.POINTER(TrackerExecC)
for _name, _args in {
    "bf_beam_tracker_device_resident": [_vp, _vp, _pd, _pc, _pe, _vp],
    "bf_beam_tracker_into": [_vp, _sz, _pd, _pc, _vp, _sz, _pe],
    "bf_tracker_stream_create": [_pd, _pc, _sz, _pe, C.POINTER(_vp)],
    "bf_tracker_stream_destroy": [_vp],
    "bf_tracker_stream_process_batch": [_vp, _sz, _vp, _vp],
    "bf_tracker_stream_process_batch_kernel_only": [_vp, _sz],
    "bf_tracker_stream_last_kernel_time_ms": [_vp, C.POINTER(C.c_float)],
    "bf_tracker_stream_device_packed_buffer": [_vp, C.POINTER(_vp)],
    "bf_tracker_stream_device_intensity_buffer": [_vp, C.POINTER(_vp)],
    "bf_tracker_stream_device_stream": [_vp, C.POINTER(_vp)],
    "bf_tracker_stream_sizes": [_vp, C.POINTER(_sz), C.POINTER(_sz), C.POINTER(_sz),
                                C.POINTER(_sz)],
    "bf_multibeam_tracker_launch": [_vp, _vp, _vp, C.POINTER(MultibeamArgsC)],
}.items():
    _capi.register(_name, _args)


def direction_from_lm(l: float, m: float):
    l32, m32 = np.float32(l), np.float32(m)
    return (float(l32), float(m32), float(np.sqrt(np.float32(1.0) - (l32 * l32 + m32 * m32))))


@dataclass
class TrackerConfig:
    """Reference include/beamformer/beam_tracker.hpp:42-57."""
    direction_start: tuple = (0.0, 0.0, 1.0)
    direction_rate_per_sample: tuple = (0.0, 0.0)
    integration_spectra: int = 320

    def c(self) -> TrackerConfigC:
        return TrackerConfigC((C.c_float * 3)(*self.direction_start),
                              (C.c_float * 2)(*self.direction_rate_per_sample),
                              self.integration_spectra)


@dataclass
class TrackerExec:
    kernel: Kernel = Kernel.AUTO
    enable_graph: bool = False

    def c(self) -> TrackerExecC:
        return TrackerExecC(int(self.kernel), int(self.enable_graph))


def beam_tracker_into(packed, dims: Dimensions, tracker: TrackerConfig,
                      exec_cfg: TrackerExec | None = None) -> np.ndarray:
    """Reference src/cuda_beam_tracker_v5.cu:422-473."""
    packed = np.ascontiguousarray(packed, np.uint8)
    out = np.zeros((dims.n_time, dims.n_freq, 1), np.float32)
    d, t, e = dims.c(), tracker.c(), (exec_cfg or TrackerExec()).c()
    call("bf_beam_tracker_into", packed.ctypes.data, packed.size, C.byref(d), C.byref(t),
         out.ctypes.data, out.size, C.byref(e))
    return out


def beam_tracker_device_resident(d_packed, d_intensity, dims: Dimensions,
                                 tracker: TrackerConfig, exec_cfg: TrackerExec | None = None,
                                 stream=None) -> None:
    """Reference src/cuda_beam_tracker_v5.cu:383-420."""
    d, t, e = dims.c(), tracker.c(), (exec_cfg or TrackerExec()).c()
    call("bf_beam_tracker_device_resident", _ptr(d_packed), _ptr(d_intensity), C.byref(d),
         C.byref(t), C.byref(e), _sh(stream))


class BatchedTrackerStream:
    """Reference include/beamformer/cuda_beam_tracker_v5.hpp:90-125."""

    def __init__(self, single_window_dims: Dimensions, tracker: TrackerConfig,
                 batch_size: int, exec_cfg: TrackerExec | None = None):
        self._h = None
        d, t, e = single_window_dims.c(), tracker.c(), (exec_cfg or TrackerExec()).c()
        h = C.c_void_p()
        call("bf_tracker_stream_create", C.byref(d), C.byref(t), batch_size, C.byref(e),
             C.byref(h))
        self._h = h.value
        self.single_window_dims = single_window_dims

    def close(self):
        if self._h:
            _capi.LIB.bf_tracker_stream_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def process_batch(self, first_window_index: int, host_packed: np.ndarray,
                      host_intensity: np.ndarray) -> None:
        host_packed = np.ascontiguousarray(host_packed, np.uint8)
        assert host_packed.size == self.batch_voltage_bytes()
        assert host_intensity.dtype == np.float32 and host_intensity.size == \
            self.batch_output_floats()
        call("bf_tracker_stream_process_batch", self._h, first_window_index,
             host_packed.ctypes.data, host_intensity.ctypes.data)

    def process_batch_kernel_only(self, first_window_index: int) -> None:
        call("bf_tracker_stream_process_batch_kernel_only", self._h, first_window_index)

    def last_kernel_time_ms(self) -> float:
        v = C.c_float()
        call("bf_tracker_stream_last_kernel_time_ms", self._h, C.byref(v))
        return v.value

    def _sizes(self):
        a, b, c_, d = (C.c_size_t() for _ in range(4))
        call("bf_tracker_stream_sizes", self._h, C.byref(a), C.byref(b), C.byref(c_),
             C.byref(d))
        return a.value, b.value, c_.value, d.value

    def batch_size(self):
        return self._sizes()[0]

    def window_bytes(self):
        return self._sizes()[1]

    def batch_voltage_bytes(self):
        return self._sizes()[2]

    def batch_output_floats(self):
        return self._sizes()[3]

    def device_packed_buffer(self) -> int:
        p = C.c_void_p()
        call("bf_tracker_stream_device_packed_buffer", self._h, C.byref(p))
        return p.value

    def device_intensity_buffer(self) -> int:
        p = C.c_void_p()
        call("bf_tracker_stream_device_intensity_buffer", self._h, C.byref(p))
        return p.value

    def device_stream(self) -> int:
        p = C.c_void_p()
        call("bf_tracker_stream_device_stream", self._h, C.byref(p))
        return p.value


def multibeam_tracker_launch(stream, d_packed, d_out, n_time, n_freq, n_ant, n_beams,
                             integration_spectra, positions=None, frequencies=None,
                             directions=None, trajectories=None, window_offset=0,
                             integrate=True, kernel=Kernel.SPLIT_F16) -> None:
    """The build's multi-beam tracker (C-ABI bf_multibeam_tracker_launch)."""
    a = MultibeamArgsC(n_time, n_freq, n_ant, n_beams, integration_spectra, _ptr(positions),
                       _ptr(frequencies), _ptr(directions), _ptr(trajectories), window_offset,
                       int(bool(integrate)), int(kernel))
    call("bf_multibeam_tracker_launch", _sh(stream), _ptr(d_packed), _ptr(d_out), C.byref(a))
