"""Python mirror of the reference's C++ host API for the fixed-grid path
(include/beamformer/cuda_stage_api.hpp, cuda_beamformer.hpp, cuda_frame.hpp),
driving libbeamform_hip.so through its C-ABI.

Names follow the reference with the vendor prefix dropped or replaced:
launch_packed_beamformer, launch_packed_integrated_beamformer,
launch_quantize_integrated_intensity, BeamformerWorkspace
(CudaBeamformerWorkspace), beamform_packed_intensity
(cuda_beamform_packed_intensity) and so on.  Errors raise the exception
types of _capi, which mirror the reference's std exception types.
"""
from __future__ import annotations

import ctypes as C
import enum
from dataclasses import dataclass

import numpy as np

from . import _capi
from ._capi import Dims, FrameView, LIB, Shard, Timings, call

LOCAL_FREQUENCY_CHANNELS = 336
MAXIMUM_BEAMS = 128


class Kernel(enum.IntEnum):
    """bf_kernel: DIRECT/TILED are the reference's selectors."""
    DIRECT = 0
    TILED = 1
    FP32 = 2
    SPLIT_F16 = 3
    F16 = 4
    AUTO = 5


class Output(enum.IntEnum):
    FLOAT32 = 0
    INT8 = 1


class DataType(enum.IntEnum):
    PACKED_INT4X2 = 0
    COMPLEX_FLOAT32 = 1
    FLOAT32 = 2
    SIGNED_INT8 = 3


@dataclass
class Dimensions:
    """Reference include/beamformer/config.hpp:74-79."""
    n_time: int = 15360
    n_freq: int = LOCAL_FREQUENCY_CHANNELS
    n_ant: int = 64
    n_beams: int = 64

    def c(self) -> Dims:
        return Dims(self.n_time, self.n_freq, self.n_ant, self.n_beams)


def validate_dimensions(dims: Dimensions) -> None:
    d = dims.c()
    call("bf_validate_dimensions", C.byref(d))


def integrated_time_count(n_time: int, integration_spectra: int) -> int:
    out = C.c_size_t()
    call("bf_integrated_time_count", n_time, integration_spectra, C.byref(out))
    return out.value


def quantization_parameter_count(integrated_dims: Dimensions) -> int:
    out = C.c_size_t()
    d = integrated_dims.c()
    call("bf_quantization_parameter_count", C.byref(d), C.byref(out))
    return out.value


def resolve_kernel(kernel: Kernel, dims: Dimensions, integration_spectra: int = 0) -> Kernel:
    out = C.c_int()
    d = dims.c()
    call("bf_resolve_kernel", int(kernel), C.byref(d), integration_spectra, C.byref(out))
    return Kernel(out.value)


# -------------------------------------------------------------- device memory
class DeviceArray:
    """A HIP device allocation owned from Python (freed on close/GC)."""

    def __init__(self, nbytes: int, dtype=np.uint8, shape=None):
        self.nbytes = int(nbytes)
        self.dtype = np.dtype(dtype)
        self.shape = tuple(shape) if shape is not None else (self.nbytes // self.dtype.itemsize,)
        p = C.c_void_p()
        call("bf_device_malloc", self.nbytes, C.byref(p))
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, arr: np.ndarray) -> "DeviceArray":
        arr = np.ascontiguousarray(arr)
        d = cls(arr.nbytes, arr.dtype, arr.shape)
        if arr.nbytes:
            call("bf_memcpy_htod", d.ptr, arr.ctypes.data, arr.nbytes)
        return d

    @classmethod
    def empty(cls, shape, dtype) -> "DeviceArray":
        n = int(np.prod(shape)) * np.dtype(dtype).itemsize
        return cls(n, dtype, shape)

    def to_numpy(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            call("bf_memcpy_dtoh", out.ctypes.data, self.ptr, self.nbytes)
        return out

    def fill(self, byte: int) -> None:
        call("bf_memset", self.ptr, byte, self.nbytes)

    def close(self) -> None:
        if getattr(self, "ptr", None):
            LIB.bf_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Stream:
    def __init__(self):
        p = C.c_void_p()
        call("bf_stream_create", C.byref(p))
        self.handle = p.value

    def synchronize(self) -> None:
        call("bf_stream_synchronize", self.handle)

    def close(self) -> None:
        if self.handle:
            LIB.bf_stream_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Event:
    def __init__(self):
        p = C.c_void_p()
        call("bf_event_create", C.byref(p))
        self.handle = p.value

    def record(self, stream: Stream | None = None) -> None:
        call("bf_event_record", self.handle, stream.handle if stream else None)

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float()
        call("bf_event_elapsed_ms", self.handle, end.handle, C.byref(ms))
        return ms.value

    def __del__(self):
        try:
            if self.handle:
                LIB.bf_event_destroy(self.handle)
        except Exception:
            pass


def device_synchronize() -> None:
    call("bf_device_synchronize")


def device_count() -> int:
    n = C.c_int()
    call("bf_device_count", C.byref(n))
    return n.value


def device_info() -> dict:
    name = C.create_string_buffer(128)
    mem = C.c_size_t()
    cus = C.c_int()
    clk = C.c_int()
    call("bf_device_info", name, 128, C.byref(mem), C.byref(cus), C.byref(clk))
    return {"arch": name.value.decode(), "global_memory_bytes": mem.value,
            "compute_units": cus.value, "clock_khz": clk.value}


def _ptr(x):
    if x is None:
        return None
    if isinstance(x, DeviceArray):
        return x.ptr
    return int(x)


def _sh(stream):
    return stream.handle if isinstance(stream, Stream) else stream


# ------------------------------------------------------------- stage launchers
def launch_packed_beamformer(kernel: Kernel, stream, packed_voltage, weights, intensity,
                             dims: Dimensions) -> None:
    """Reference include/beamformer/cuda_stage_api.hpp:19-24."""
    d = dims.c()
    call("bf_launch_packed_beamformer", int(kernel), _sh(stream), _ptr(packed_voltage),
         _ptr(weights), _ptr(intensity), C.byref(d))


def launch_packed_integrated_beamformer(kernel: Kernel, stream, packed_voltage, weights,
                                        integrated_intensity, dims: Dimensions,
                                        integration_spectra: int) -> None:
    """Reference include/beamformer/cuda_stage_api.hpp:26-34."""
    d = dims.c()
    call("bf_launch_packed_integrated_beamformer", int(kernel), _sh(stream),
         _ptr(packed_voltage), _ptr(weights), _ptr(integrated_intensity), C.byref(d),
         integration_spectra)


def launch_packed_beamformer_masked(kernel: Kernel, stream, packed_voltage, loss_mask, weights,
                                    intensity, dims: Dimensions) -> None:
    """launch_packed_beamformer with a [T][F] device loss mask (0 = lost
    frame, contributes nothing); MFMA kernels only (bf_launch_packed_beamformer_masked)."""
    d = dims.c()
    call("bf_launch_packed_beamformer_masked", int(kernel), _sh(stream), _ptr(packed_voltage),
         _ptr(loss_mask), _ptr(weights), _ptr(intensity), C.byref(d))


def launch_packed_integrated_beamformer_masked(kernel: Kernel, stream, packed_voltage,
                                               loss_mask, weights, integrated_intensity,
                                               dims: Dimensions,
                                               integration_spectra: int) -> None:
    """launch_packed_integrated_beamformer with a [T][F] device loss mask."""
    d = dims.c()
    call("bf_launch_packed_integrated_beamformer_masked", int(kernel), _sh(stream),
         _ptr(packed_voltage), _ptr(loss_mask), _ptr(weights), _ptr(integrated_intensity),
         C.byref(d), integration_spectra)


def launch_quantize_integrated_intensity(stream, integrated_intensity, quantized_intensity,
                                         parameters, integrated_dims: Dimensions) -> None:
    """Reference include/beamformer/cuda_stage_api.hpp:36-42."""
    d = integrated_dims.c()
    call("bf_launch_quantize_integrated_intensity", _sh(stream), _ptr(integrated_intensity),
         _ptr(quantized_intensity), _ptr(parameters), C.byref(d))


NOALLOC_SCRATCH_BYTES = 64 << 20


def launch_packed_quantized_integrated_beamformer_noalloc(
        kernel: Kernel, stream, packed_voltage, weights, quantized_intensity, parameters,
        dims: Dimensions, integration_spectra: int) -> None:
    """Beamform + integrate + int8-quantise with no device allocation: the
    f32 intermediate lives in the library's per-GPU scratch of
    NOALLOC_SCRATCH_BYTES (calls overlapping on one GPU race)."""
    d = dims.c()
    call("bf_launch_packed_quantized_integrated_beamformer_noalloc", int(kernel), _sh(stream),
         _ptr(packed_voltage), _ptr(weights), _ptr(quantized_intensity), _ptr(parameters),
         C.byref(d), integration_spectra)


def autotune_packed_beamformer(kernel: Kernel, stream, packed_voltage, weights, output,
                               dims: Dimensions, integration_spectra: int = 0,
                               repeats: int = 3) -> list:
    """Times every MFMA tile shape on these device buffers and keeps the
    fastest for later launches of the same problem (bf_autotune_packed_beamformer).
    Returns one dict per candidate."""
    n = C.c_size_t()
    entries = (_capi.TuneEntry * 64)()
    call("bf_autotune_packed_beamformer", int(kernel), _sh(stream), _ptr(packed_voltage),
         _ptr(weights), _ptr(output), C.byref(dims.c()), integration_spectra, repeats, entries,
         64, C.byref(n))
    return [{f: getattr(e, f) for f, _ in e._fields_} for e in entries[:min(n.value, 64)]]


def autotune_clear() -> None:
    call("bf_autotune_clear")


def weight_storage_count(dims: Dimensions, kernel: Kernel) -> int:
    if kernel == Kernel.TILED:
        return dims.n_freq * ((dims.n_beams + 31) // 32) * dims.n_ant * 32
    return dims.n_beams * dims.n_freq * dims.n_ant


# ------------------------------------------------------------------ workspace
class BeamformerWorkspace:
    """Reference include/beamformer/cuda_beamformer.hpp:44-92
    (CudaBeamformerWorkspace).  Synchronous methods returning milliseconds."""

    def __init__(self, capacity: Dimensions, kernel: Kernel = Kernel.DIRECT,
                 temporal_integration: int | None = None, output: Output = Output.FLOAT32):
        self._h = None
        d = capacity.c()
        h = C.c_void_p()
        call("bf_workspace_create", C.byref(d), int(kernel), int(temporal_integration or 0),
             int(output), C.byref(h))
        self._h = h.value
        self.capacity = capacity
        self._kernel = Kernel(kernel)
        self._spectra = int(temporal_integration or 0)
        self._output = Output(output)

    def close(self):
        if self._h:
            LIB.bf_workspace_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _int(self, name):
        v = C.c_int()
        call(name, self._h, C.byref(v))
        return v.value

    def setup_ms(self) -> float:
        v = C.c_double()
        call("bf_workspace_setup_ms", self._h, C.byref(v))
        return v.value

    def kernel(self) -> Kernel:
        return Kernel(self._int("bf_workspace_kernel"))

    def has_temporal_integration(self) -> bool:
        return bool(self._int("bf_workspace_has_temporal_integration"))

    def has_int8_quantization(self) -> bool:
        return bool(self._int("bf_workspace_has_int8_quantization"))

    def packed_voltage_capacity_bytes(self) -> int:
        v = C.c_size_t()
        call("bf_workspace_packed_voltage_capacity_bytes", self._h, C.byref(v))
        return v.value

    def upload_packed_voltage(self, packed: np.ndarray, dims: Dimensions) -> float:
        packed = np.ascontiguousarray(packed, np.uint8)
        d, ms = dims.c(), C.c_double()
        call("bf_workspace_upload_packed_voltage", self._h, packed.ctypes.data, packed.size,
             C.byref(d), C.byref(ms))
        return ms.value

    def upload_weights(self, weights: np.ndarray, dims: Dimensions) -> float:
        weights = np.ascontiguousarray(weights, np.float32)
        d, ms = dims.c(), C.c_double()
        call("bf_workspace_upload_weights", self._h, weights.ctypes.data, weights.size // 2,
             C.byref(d), C.byref(ms))
        return ms.value

    def run_kernel(self, dims: Dimensions) -> float:
        d, ms = dims.c(), C.c_double()
        call("bf_workspace_run_kernel", self._h, C.byref(d), C.byref(ms))
        return ms.value

    def download_intensity(self, intensity: np.ndarray, dims: Dimensions) -> float:
        assert intensity.dtype == np.float32 and intensity.flags.c_contiguous
        d, ms = dims.c(), C.c_double()
        call("bf_workspace_download_intensity", self._h, intensity.ctypes.data, intensity.size,
             C.byref(d), C.byref(ms))
        return ms.value

    def run_pipeline(self, packed: np.ndarray, intensity: np.ndarray, dims: Dimensions):
        packed = np.ascontiguousarray(packed, np.uint8)
        d, t = dims.c(), Timings()
        call("bf_workspace_run_pipeline", self._h, packed.ctypes.data, packed.size,
             intensity.ctypes.data, intensity.size, C.byref(d), C.byref(t))
        return t.as_dict()

    def run_integrated_pipeline(self, packed: np.ndarray, integrated: np.ndarray,
                                dims: Dimensions):
        packed = np.ascontiguousarray(packed, np.uint8)
        d, t = dims.c(), Timings()
        call("bf_workspace_run_integrated_pipeline", self._h, packed.ctypes.data, packed.size,
             integrated.ctypes.data, integrated.size, C.byref(d), C.byref(t))
        return t.as_dict()

    def run_quantized_integrated_pipeline(self, packed: np.ndarray, codes: np.ndarray,
                                          parameters: np.ndarray, dims: Dimensions):
        packed = np.ascontiguousarray(packed, np.uint8)
        d, t = dims.c(), Timings()
        call("bf_workspace_run_quantized_integrated_pipeline", self._h, packed.ctypes.data,
             packed.size, codes.ctypes.data, codes.size, parameters.ctypes.data,
             parameters.size // 2, C.byref(d), C.byref(t))
        return t.as_dict()

    def run_device_frame(self, voltage: FrameView, weights: FrameView, output: FrameView,
                         quantization_parameters: FrameView | None = None):
        t = Timings()
        call("bf_workspace_run_device_frame", self._h, C.byref(voltage), C.byref(weights),
             C.byref(output),
             C.byref(quantization_parameters) if quantization_parameters is not None else None,
             C.byref(t))
        return t.as_dict()


class OfflineFrameRunner:
    """Reference include/beamformer/cuda_offline_runner.hpp:17-52
    (CudaOfflineFrameRunner): one offline frame from host arrays through
    temporary device copies and the workspace's device-frame path."""

    def __init__(self, capacity: Dimensions, kernel: Kernel = Kernel.DIRECT,
                 temporal_integration: int | None = None, output: Output = Output.FLOAT32):
        self._h = None
        h = C.c_void_p()
        call("bf_offline_frame_runner_create", C.byref(capacity.c()), int(kernel),
             int(temporal_integration or 0), int(output), C.byref(h))
        self._h = h.value
        self._kernel, self._output = Kernel(kernel), Output(output)
        self._spectra = int(temporal_integration or 0)

    def kernel(self) -> Kernel:
        return self._kernel

    def output(self) -> Output:
        return self._output

    def has_temporal_integration(self) -> bool:
        return self._spectra != 0

    def run(self, packed, weights, dims: Dimensions, frame_id: int, shard: Shard) -> dict:
        """Returns {"output_dims", "timings", "float32_output"} or, for int8
        runners, {"output_dims", "timings", "codes", "parameters"}."""
        packed = np.ascontiguousarray(packed, np.uint8)
        weights = np.ascontiguousarray(weights, np.float32)
        rows = integrated_time_count(dims.n_time, self._spectra) if self._spectra else dims.n_time
        shape = (rows, dims.n_freq, dims.n_beams)
        od, t = Dims(), Timings()
        res = {}
        if self._output == Output.INT8:
            codes = np.zeros(shape, np.int8)
            npar = quantization_parameter_count(Dimensions(rows, dims.n_freq, dims.n_ant,
                                                           dims.n_beams))
            par = np.zeros((npar, 2), np.float32)
            call("bf_offline_frame_runner_run", self._h, packed.ctypes.data, packed.size,
                 weights.ctypes.data, weights.size // 2, C.byref(dims.c()), frame_id,
                 C.byref(shard), None, 0, codes.ctypes.data, codes.size, par.ctypes.data,
                 npar, C.byref(od), C.byref(t))
            res.update(codes=codes, parameters=par)
        else:
            out = np.zeros(shape, np.float32)
            call("bf_offline_frame_runner_run", self._h, packed.ctypes.data, packed.size,
                 weights.ctypes.data, weights.size // 2, C.byref(dims.c()), frame_id,
                 C.byref(shard), out.ctypes.data, out.size, None, 0, None, 0, C.byref(od),
                 C.byref(t))
            res.update(float32_output=out)
        res["output_dims"] = Dimensions(od.n_time, od.n_freq, od.n_ant, od.n_beams)
        res["timings"] = t.as_dict()
        return res

    def close(self) -> None:
        if self._h:
            LIB.bf_offline_frame_runner_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# --------------------------------------------------------------- frame views
def default_shard_descriptors():
    """Reference include/beamformer/config.hpp:65-72."""
    return [Shard(0, 2, 336, 0, 0, 1, 0, 1), Shard(1, 2, 336, 336, 0, 1, 1, 1)]


_DTYPE_SIZE = {DataType.PACKED_INT4X2: 1, DataType.SIGNED_INT8: 1,
               DataType.COMPLEX_FLOAT32: 8, DataType.FLOAT32: 4}


def make_contiguous_frame_view(device_data, nbytes: int, dtype: DataType, shape, frame_id: int,
                               shard: Shard) -> FrameView:
    """Reference include/beamformer/cuda_frame.hpp:151-167."""
    v = FrameView()
    v.device_data = _ptr(device_data)
    v.bytes = nbytes
    v.type = int(dtype)
    v.rank = len(shape)
    stride = _DTYPE_SIZE[DataType(dtype)]
    strides = [0] * len(shape)
    for ax in range(len(shape) - 1, -1, -1):
        strides[ax] = stride
        stride *= shape[ax]
    for i in range(4):
        v.shape[i] = shape[i] if i < len(shape) else 0
        v.strides_bytes[i] = strides[i] if i < len(shape) else 0
    v.frame_id = frame_id
    v.shard = shard
    return v


# ------------------------------------------------------------ one-shot API
def beamform_packed_intensity(packed, weights, dims: Dimensions, kernel=Kernel.DIRECT,
                              with_timings=False):
    """Reference src/cuda_beamformer.cu:1081-1105."""
    packed = np.ascontiguousarray(packed, np.uint8)
    weights = np.ascontiguousarray(weights, np.float32)
    out = np.zeros((dims.n_time, dims.n_freq, dims.n_beams), np.float32)
    d, t = dims.c(), Timings()
    call("bf_beamform_packed_intensity", packed.ctypes.data, packed.size, weights.ctypes.data,
         weights.size // 2, C.byref(d), int(kernel), out.ctypes.data, out.size, C.byref(t))
    return (out, t.as_dict()) if with_timings else out


def beamform_packed_integrated_intensity(packed, weights, dims: Dimensions,
                                         integration_spectra: int, kernel=Kernel.TILED,
                                         with_timings=False):
    """Reference src/cuda_beamformer.cu:1107-1136."""
    packed = np.ascontiguousarray(packed, np.uint8)
    weights = np.ascontiguousarray(weights, np.float32)
    W = integrated_time_count(dims.n_time, integration_spectra)
    out = np.zeros((W, dims.n_freq, dims.n_beams), np.float32)
    d, t = dims.c(), Timings()
    call("bf_beamform_packed_integrated_intensity", packed.ctypes.data, packed.size,
         weights.ctypes.data, weights.size // 2, C.byref(d), integration_spectra, int(kernel),
         out.ctypes.data, out.size, C.byref(t))
    return (out, t.as_dict()) if with_timings else out


def beamform_packed_quantized_integrated_intensity(packed, weights, dims: Dimensions,
                                                   integration_spectra: int,
                                                   kernel=Kernel.TILED, with_timings=False):
    """Reference src/cuda_beamformer.cu:1138-1173.  Returns (codes, params)."""
    packed = np.ascontiguousarray(packed, np.uint8)
    weights = np.ascontiguousarray(weights, np.float32)
    W = integrated_time_count(dims.n_time, integration_spectra)
    codes = np.zeros((W, dims.n_freq, dims.n_beams), np.int8)
    npar = W * ((dims.n_freq + 15) // 16) * ((dims.n_beams + 15) // 16)
    params = np.zeros((npar, 2), np.float32)
    d, t = dims.c(), Timings()
    call("bf_beamform_packed_quantized_integrated_intensity", packed.ctypes.data, packed.size,
         weights.ctypes.data, weights.size // 2, C.byref(d), integration_spectra, int(kernel),
         codes.ctypes.data, codes.size, params.ctypes.data, npar, C.byref(t))
    return ((codes, params), t.as_dict()) if with_timings else (codes, params)


# --------------------------------------------------------------- generators
def generate_uniform_int4(stream, out: DeviceArray, count: int, seed: int) -> None:
    call("bf_generate_uniform_int4", _sh(stream), _ptr(out), count, seed)


def generate_point_source_noise(stream, out: DeviceArray, n_time, n_freq, n_ant,
                                positions: DeviceArray, frequencies: DeviceArray, direction,
                                amplitude: float, sigma: float, seed: int) -> None:
    d3 = (C.c_float * 3)(*[float(v) for v in direction])
    call("bf_generate_point_source_noise", _sh(stream), _ptr(out), n_time, n_freq, n_ant,
         _ptr(positions), _ptr(frequencies), d3, amplitude, sigma, seed)
