"""On-disk formats of the reference, read and written by libbeamform_hip.so
(include/beamformer/io.hpp, src/io.cpp; C-ABI bf_read_* / bf_write_*).

Every file is byte-compatible with the reference's: raw little-endian
arrays whose size must match the dimensions exactly, and key=value metadata
text (beamformer-quantized-int8-v1, beamformer-packed-shard-v1).  IO
failures raise HipRuntimeError (std::runtime_error in the reference), size
and contract violations InvalidArgument (std::invalid_argument).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from . import _capi
from ._capi import Shard, call
from .api import Dimensions, integrated_time_count, quantization_parameter_count

_vp, _sz, _cp = C.c_void_p, C.c_size_t, C.c_char_p
_pd = C.POINTER(_capi.Dims)
_psh = C.POINTER(Shard)


class QuantizedOutputMetadataC(C.Structure):
    _fields_ = [("output_dims", _capi.Dims), ("input_n_time", C.c_size_t),
                ("integration_spectra", C.c_size_t), ("shard", Shard),
                ("parameters_file", C.c_char * 1024)]


_pm = C.POINTER(QuantizedOutputMetadataC)
for _name, _args in {
    "bf_default_shard_descriptor": [_sz, _psh],
    "bf_validate_shard_descriptor": [_psh],
    "bf_read_packed_voltage": [_cp, _pd, _vp, _sz],
    "bf_write_packed_voltage": [_cp, _vp, _sz, _pd],
    "bf_read_weights": [_cp, _pd, _vp, _sz],
    "bf_write_weights": [_cp, _vp, _sz, _pd],
    "bf_read_tiled_weights": [_cp, _pd, _vp, _sz],
    "bf_write_tiled_weights": [_cp, _vp, _sz, _pd],
    "bf_write_intensities": [_cp, _vp, _sz, _pd],
    "bf_read_quantized_intensities": [_cp, _pd, _vp, _sz],
    "bf_write_quantized_intensities": [_cp, _vp, _sz, _pd],
    "bf_read_quantization_parameters": [_cp, _pd, _vp, _sz],
    "bf_write_quantization_parameters": [_cp, _vp, _sz, _pd],
    "bf_write_quantized_output_metadata": [_cp, _pm],
    "bf_read_quantized_output_metadata": [_cp, _pm],
    "bf_write_packed_shard": [_cp, _cp, _cp, _psh, _vp, _sz, _vp, _sz, _pd],
    "bf_read_packed_shard": [_cp, _cp, _cp, _pd, _psh, _vp, _sz, _vp, _sz],
    "bf_validate_packed_shard": [_psh, _vp, _sz, _vp, _sz, _pd],
    "bf_validate_packed_shards": [_psh, C.POINTER(_vp), _sz, C.POINTER(_vp), _sz, _pd],
    "bf_make_loss_mask": [_pd, C.c_uint32, C.c_float, _vp, _sz],
    "bf_shard_frequencies": [_psh, C.c_float, C.c_float, _vp],
}.items():
    _capi.register(_name, _args)


def _path(p) -> bytes:
    return os.fsencode(os.fspath(p))


@dataclass
class ShardDescriptor:
    """Reference include/beamformer/config.hpp:29-41."""
    shard_id: int = 0
    shard_count: int = 2
    local_frequency_count: int = 336
    absolute_frequency_start: int = 0
    timestamp_start: int = 0
    timestamp_step: int = 1
    loss_mask_id: int = 0
    loss_mask_independent: bool = True

    def c(self) -> Shard:
        return Shard(self.shard_id, self.shard_count, self.local_frequency_count,
                     self.absolute_frequency_start, self.timestamp_start,
                     self.timestamp_step, self.loss_mask_id,
                     1 if self.loss_mask_independent else 0)

    @staticmethod
    def from_c(s: Shard) -> "ShardDescriptor":
        return ShardDescriptor(s.shard_id, s.shard_count, s.local_frequency_count,
                               s.absolute_frequency_start, s.timestamp_start,
                               s.timestamp_step, s.loss_mask_id,
                               bool(s.loss_mask_independent))


def default_shard_descriptor(shard_id: int) -> ShardDescriptor:
    """default_shard_descriptors()[shard_id] (config.hpp:65-72)."""
    out = Shard()
    call("bf_default_shard_descriptor", shard_id, C.byref(out))
    return ShardDescriptor.from_c(out)


def validate_shard_descriptor(shard: ShardDescriptor) -> None:
    s = shard.c()
    call("bf_validate_shard_descriptor", C.byref(s))


@dataclass
class QuantizedOutputMetadata:
    """Reference include/beamformer/quantization.hpp:75-81."""
    output_dims: Dimensions
    input_n_time: int
    integration_spectra: int
    shard: ShardDescriptor
    parameters_file: str

    def c(self) -> QuantizedOutputMetadataC:
        m = QuantizedOutputMetadataC()
        m.output_dims = self.output_dims.c()
        m.input_n_time = self.input_n_time
        m.integration_spectra = self.integration_spectra
        m.shard = self.shard.c()
        name = self.parameters_file.encode()
        if len(name) >= 1024:
            raise _capi.InvalidArgument("parameters_file is longer than 1023 bytes")
        m.parameters_file = name
        return m


def _array(a, dtype, count: int, what: str) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=dtype)
    if a.size != count:
        raise _capi.InvalidArgument(f"{what} count does not match dimensions")
    return a


def read_packed_voltage(path, dims: Dimensions) -> np.ndarray:
    """[T][F][A] packed int4 bytes (io.cpp:60-64)."""
    out = np.empty((dims.n_time, dims.n_freq, dims.n_ant), np.uint8)
    d = dims.c()
    call("bf_read_packed_voltage", _path(path), C.byref(d), out.ctypes.data, out.size)
    return out


def write_packed_voltage(path, voltage, dims: Dimensions) -> None:
    v = np.ascontiguousarray(voltage, dtype=np.uint8)
    d = dims.c()
    call("bf_write_packed_voltage", _path(path), v.ctypes.data, v.size, C.byref(d))


def read_weights(path, dims: Dimensions) -> np.ndarray:
    """[B][F][A] complex64 (io.cpp:74-77)."""
    out = np.empty((dims.n_beams, dims.n_freq, dims.n_ant), np.complex64)
    d = dims.c()
    call("bf_read_weights", _path(path), C.byref(d), out.ctypes.data, out.size)
    return out


def write_weights(path, weights, dims: Dimensions) -> None:
    w = np.ascontiguousarray(weights, dtype=np.complex64)
    d = dims.c()
    call("bf_write_weights", _path(path), w.ctypes.data, w.size, C.byref(d))


def tiled_weight_count(dims: Dimensions) -> int:
    """formats.hpp:58-61: F x ceil(B/32) x A x 32."""
    return dims.n_freq * ((dims.n_beams + 31) // 32) * dims.n_ant * 32


def read_tiled_weights(path, dims: Dimensions) -> np.ndarray:
    """[F][ceil(B/32)][A][32] complex64 (io.cpp:88-92)."""
    out = np.empty((dims.n_freq, (dims.n_beams + 31) // 32, dims.n_ant, 32), np.complex64)
    d = dims.c()
    call("bf_read_tiled_weights", _path(path), C.byref(d), out.ctypes.data, out.size)
    return out


def write_tiled_weights(path, weights, dims: Dimensions) -> None:
    w = np.ascontiguousarray(weights, dtype=np.complex64)
    d = dims.c()
    call("bf_write_tiled_weights", _path(path), w.ctypes.data, w.size, C.byref(d))


def write_intensities(path, intensity, dims: Dimensions) -> None:
    """[T][F][B] float32 (io.cpp:103-111)."""
    a = np.ascontiguousarray(intensity, dtype=np.float32)
    d = dims.c()
    call("bf_write_intensities", _path(path), a.ctypes.data, a.size, C.byref(d))


def read_intensities(path, dims: Dimensions) -> np.ndarray:
    """[T][F][B] float32; the reference writes these files only, and the
    size check matches its readers'."""
    nbytes = dims.n_time * dims.n_freq * dims.n_beams * 4
    if not os.path.exists(path):
        raise _capi.HipRuntimeError(f"cannot inspect input file: {os.fspath(path)}")
    have = os.path.getsize(path)
    if have != nbytes:
        raise _capi.HipRuntimeError(f"{os.fspath(path)} has {have} bytes; expected {nbytes}")
    return np.fromfile(path, np.float32).reshape(dims.n_time, dims.n_freq, dims.n_beams)


def read_quantized_intensities(path, dims: Dimensions) -> np.ndarray:
    """[W][F][B] int8 codes (io.cpp:113-117)."""
    out = np.empty((dims.n_time, dims.n_freq, dims.n_beams), np.int8)
    d = dims.c()
    call("bf_read_quantized_intensities", _path(path), C.byref(d), out.ctypes.data, out.size)
    return out


def write_quantized_intensities(path, codes, dims: Dimensions) -> None:
    a = np.ascontiguousarray(codes, dtype=np.int8)
    d = dims.c()
    call("bf_write_quantized_intensities", _path(path), a.ctypes.data, a.size, C.byref(d))


def read_quantization_parameters(path, integrated_dims: Dimensions) -> np.ndarray:
    """[n][2] float32 {offset, scale} (io.cpp:131-135)."""
    n = quantization_parameter_count(integrated_dims)
    out = np.empty((n, 2), np.float32)
    d = integrated_dims.c()
    call("bf_read_quantization_parameters", _path(path), C.byref(d), out.ctypes.data, n)
    return out


def write_quantization_parameters(path, params, integrated_dims: Dimensions) -> None:
    a = np.ascontiguousarray(params, dtype=np.float32).reshape(-1)
    if a.size % 2:
        raise _capi.InvalidArgument("parameters are {offset, scale} pairs")
    d = integrated_dims.c()
    call("bf_write_quantization_parameters", _path(path), a.ctypes.data, a.size // 2,
         C.byref(d))


def write_quantized_output_metadata(path, metadata: QuantizedOutputMetadata) -> None:
    """beamformer-quantized-int8-v1 (io.cpp:149-183)."""
    m = metadata.c()
    call("bf_write_quantized_output_metadata", _path(path), C.byref(m))


def read_quantized_output_metadata(path) -> QuantizedOutputMetadata:
    """io.cpp:185-242."""
    m = QuantizedOutputMetadataC()
    call("bf_read_quantized_output_metadata", _path(path), C.byref(m))
    d = m.output_dims
    return QuantizedOutputMetadata(Dimensions(d.n_time, d.n_freq, d.n_ant, d.n_beams),
                                   m.input_n_time, m.integration_spectra,
                                   ShardDescriptor.from_c(m.shard),
                                   m.parameters_file.decode())


def quantized_output_metadata(input_dims: Dimensions, integration_spectra: int,
                              shard_id: int, parameters_file: str) -> QuantizedOutputMetadata:
    """The metadata beamformer_cuda writes for an int8 run (tools/beamformer_cuda.cpp:235-240)."""
    w = integrated_time_count(input_dims.n_time, integration_spectra)
    return QuantizedOutputMetadata(
        Dimensions(w, input_dims.n_freq, input_dims.n_ant, input_dims.n_beams),
        input_dims.n_time, integration_spectra, default_shard_descriptor(shard_id),
        parameters_file)


def write_packed_shard(payload_path, metadata_path, loss_mask_path, shard: ShardDescriptor,
                       payload, loss_mask, dims: Dimensions) -> None:
    """beamformer-packed-shard-v1 (io.cpp:244-278)."""
    p = np.ascontiguousarray(payload, dtype=np.uint8)
    m = np.ascontiguousarray(loss_mask, dtype=np.uint8)
    s = shard.c()
    d = dims.c()
    call("bf_write_packed_shard", _path(payload_path), _path(metadata_path),
         _path(loss_mask_path), C.byref(s), p.ctypes.data, p.size, m.ctypes.data, m.size,
         C.byref(d))


def read_packed_shard(payload_path, metadata_path, loss_mask_path, dims: Dimensions):
    """io.cpp:280-342.  Returns (descriptor, payload [T][F][A], loss_mask [T][F])."""
    payload = np.empty((dims.n_time, dims.n_freq, dims.n_ant), np.uint8)
    mask = np.empty((dims.n_time, dims.n_freq), np.uint8)
    s = Shard()
    d = dims.c()
    call("bf_read_packed_shard", _path(payload_path), _path(metadata_path),
         _path(loss_mask_path), C.byref(d), C.byref(s), payload.ctypes.data, payload.size,
         mask.ctypes.data, mask.size)
    return ShardDescriptor.from_c(s), payload, mask


def validate_packed_shard(shard: ShardDescriptor, payload, loss_mask, dims: Dimensions) -> None:
    """Reference src/synthetic_data.cpp:139-153."""
    p = np.ascontiguousarray(payload, np.uint8).reshape(-1)
    m = np.ascontiguousarray(loss_mask, np.uint8).reshape(-1)
    s, d = shard.c(), dims.c()
    call("bf_validate_packed_shard", C.byref(s), p.ctypes.data, p.size, m.ctypes.data, m.size,
         C.byref(d))


def validate_packed_shards(shards, payloads, loss_masks, dims: Dimensions) -> None:
    """Reference src/synthetic_data.cpp:155-171: a two-shard set."""
    ps = [np.ascontiguousarray(p, np.uint8).reshape(-1) for p in payloads]
    ms = [np.ascontiguousarray(m, np.uint8).reshape(-1) for m in loss_masks]
    if len(shards) != 2 or len(ps) != 2 or len(ms) != 2 or ps[0].size != ps[1].size \
            or ms[0].size != ms[1].size:
        raise ValueError("a shard set holds two shards of equal sizes")
    sh = (Shard * 2)(shards[0].c(), shards[1].c())
    pp = (C.c_void_p * 2)(ps[0].ctypes.data, ps[1].ctypes.data)
    mm = (C.c_void_p * 2)(ms[0].ctypes.data, ms[1].ctypes.data)
    d = dims.c()
    call("bf_validate_packed_shards", sh, pp, ps[0].size, mm, ms[0].size, C.byref(d))


def make_loss_mask(dims: Dimensions, seed: int, loss_probability: float) -> np.ndarray:
    """Reference src/synthetic_data.cpp:121-137: [T][F] validity bytes."""
    out = np.zeros((dims.n_time, dims.n_freq), np.uint8)
    d = dims.c()
    call("bf_make_loss_mask", C.byref(d), seed, loss_probability, out.ctypes.data, out.size)
    return out


def shard_frequencies(shard: ShardDescriptor, start_hz: float = 300e6,
                      channel_width_hz: float = 300e3) -> np.ndarray:
    """The shard's 336 channel frequencies of the full-band plan."""
    out = np.zeros(shard.local_frequency_count, np.float32)
    s = shard.c()
    call("bf_shard_frequencies", C.byref(s), start_hz, channel_width_hz, out.ctypes.data)
    return out
