"""Python binding of the streaming tracker server (C-ABI bf_server_*).

The server owns a device voltage ring, a request table and an HBM arena for
request outputs; `step()` batches every window that completed since the
previous step, across all admitted requests, into one launch.  The Python
layer only marshals arguments: scheduling, admission and timing are native.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import _capi
from ._capi import call
from .api import Kernel

_vp, _sz = C.c_void_p, C.c_size_t


class ServerConfigC(C.Structure):
    _fields_ = [("n_freq", _sz), ("n_ant", _sz), ("ring_spectra", _sz), ("max_window", _sz),
                ("max_step_spectra", _sz), ("hbm_budget_bytes", _sz), ("max_requests", _sz),
                ("positions", _vp), ("frequencies", _vp), ("kernel", C.c_int),
                ("enable_graph", C.c_int), ("loss_masks", C.c_int), ("shard", _vp),
                ("max_active_beams", _sz)]


class RingSpanC(C.Structure):
    _fields_ = [("first_spectrum", C.c_longlong), ("n_spectra", _sz), ("voltages", _vp),
                ("valid", _vp)]


class RequestC(C.Structure):
    _fields_ = [("n_beams", _sz), ("window", _sz), ("n_spectra", _sz), ("start", C.c_longlong),
                ("output_windows", _sz), ("trajectories", _vp)]


class RequestStatusC(C.Structure):
    _fields_ = [("state", C.c_int), ("start", C.c_longlong), ("windows_total", _sz),
                ("windows_issued", _sz), ("windows_done", _sz), ("output_windows", _sz),
                ("output_bytes", _sz)]


class StatsC(C.Structure):
    _fields_ = [("windows_timed", _sz), ("latency_p50_ms", C.c_double),
                ("latency_p90_ms", C.c_double), ("latency_p99_ms", C.c_double),
                ("latency_max_ms", C.c_double), ("latency_mean_ms", C.c_double),
                ("steps", _sz), ("windows_processed", _sz), ("beam_samples", C.c_double),
                ("gpu_busy_ms", C.c_double), ("requests_submitted", _sz),
                ("requests_admitted", _sz), ("requests_completed", _sz),
                ("requests_rejected", _sz), ("requests_pending", _sz),
                ("requests_in_slots", _sz), ("arena_bytes_in_use", _sz),
                ("arena_peak_bytes", _sz), ("head", C.c_longlong),
                ("processed_head", C.c_longlong), ("mean_admission_wait_spectra", C.c_double)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class SimConfigC(C.Structure):
    _fields_ = [("server", ServerConfigC), ("seed", C.c_uint64), ("cadence_s", C.c_double),
                ("run_seconds", C.c_double), ("chunk_spectra", _sz),
                ("arrival_rate_hz", C.c_double), ("min_beams", _sz), ("max_beams", _sz),
                ("min_duration_s", C.c_double), ("max_duration_s", C.c_double),
                ("window_a", _sz), ("window_b", _sz), ("output_windows_cap", _sz),
                ("paced", C.c_int), ("pool_spectra", _sz), ("source_amplitude", C.c_float),
                ("noise_sigma", C.c_float)]


class SimReportC(C.Structure):
    _fields_ = [("stats", StatsC), ("wall_seconds", C.c_double),
                ("stream_seconds", C.c_double), ("beam_samples_per_s", C.c_double),
                ("mean_requests_in_slots", C.c_double), ("requests_generated", _sz),
                ("max_pacing_lag_ms", C.c_double), ("graph_used", C.c_int)]

    def as_dict(self):
        d = {n: getattr(self, n) for n, _ in self._fields_ if n != "stats"}
        d["stats"] = self.stats.as_dict()
        return d


for _name, _args in {
    "bf_server_create": [C.POINTER(ServerConfigC), C.POINTER(_vp)],
    "bf_server_destroy": [_vp],
    "bf_server_submit": [_vp, C.POINTER(RequestC), C.POINTER(C.c_uint64)],
    "bf_server_ingest": [_vp, _vp, _sz, C.c_int],
    "bf_server_ingest_masked": [_vp, _vp, _vp, _sz, C.c_int],
    "bf_server_acquire": [_vp, _sz, C.POINTER(RingSpanC), C.POINTER(_sz)],
    "bf_server_commit": [_vp, _sz, _vp],
    "bf_server_step": [_vp, C.POINTER(_sz)],
    "bf_server_request_status": [_vp, C.c_uint64, C.POINTER(RequestStatusC)],
    "bf_server_read_output": [_vp, C.c_uint64, _sz, _sz, _vp],
    "bf_server_release": [_vp, C.c_uint64],
    "bf_server_synchronize": [_vp],
    "bf_server_get_stats": [_vp, C.POINTER(StatsC)],
    "bf_server_reset_stats": [_vp],
    "bf_server_simulate": [C.POINTER(SimConfigC), C.POINTER(SimReportC)],
}.items():
    _capi.register(_name, _args)

PENDING, ACTIVE, COMPLETE, REJECTED = 0, 1, 2, 3


@dataclass
class ServerConfig:
    n_freq: int = 336
    n_ant: int = 64
    ring_spectra: int = 1 << 20
    max_window: int = 320
    max_step_spectra: int = 8192
    hbm_budget_bytes: int = 4 << 30
    max_requests: int = 256
    positions: Optional[np.ndarray] = None     # [A][3] metres
    frequencies: Optional[np.ndarray] = None   # [F] Hz
    kernel: Kernel = Kernel.SPLIT_F16
    enable_graph: bool = True
    loss_masks: bool = False                   # validity ring [ring][F]
    shard: Optional[object] = None             # formats.ShardDescriptor of the stream
    max_active_beams: int = 0                  # compute budget for admission; 0: none

    def to_c(self, keep: list) -> ServerConfigC:
        pos = fr = None
        if self.positions is not None:
            pos = np.ascontiguousarray(self.positions, np.float32)
            keep.append(pos)
        if self.frequencies is not None:
            fr = np.ascontiguousarray(self.frequencies, np.float32)
            keep.append(fr)
        sh = None
        if self.shard is not None:
            sh = self.shard.c()
            keep.append(sh)
        return ServerConfigC(self.n_freq, self.n_ant, self.ring_spectra, self.max_window,
                             self.max_step_spectra, self.hbm_budget_bytes, self.max_requests,
                             None if pos is None else pos.ctypes.data,
                             None if fr is None else fr.ctypes.data, int(self.kernel),
                             int(self.enable_graph), int(self.loss_masks),
                             None if sh is None else C.addressof(sh),
                             int(self.max_active_beams))


@dataclass
class Request:
    trajectories: np.ndarray          # [B][4] l0, m0, rate_l, rate_m (per spectrum)
    window: int
    n_spectra: int
    start: int = -1                   # -1: the head at admission
    output_windows: int = 0           # 0: keep every window


class TrackerServer:
    """Continuous-batching tracker server (bf_server)."""

    def __init__(self, cfg: ServerConfig):
        self._h = None
        self.cfg = cfg
        keep: list = []
        c = cfg.to_c(keep)
        h = _vp()
        call("bf_server_create", C.byref(c), C.byref(h))
        self._h = h

    def close(self):
        if self._h is not None:
            call("bf_server_destroy", self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, req: Request) -> int:
        traj = np.ascontiguousarray(req.trajectories, np.float32).reshape(-1, 4)
        q = RequestC(traj.shape[0], req.window, req.n_spectra, req.start, req.output_windows,
                     traj.ctypes.data)
        rid = C.c_uint64()
        call("bf_server_submit", self._h, C.byref(q), C.byref(rid))
        return rid.value

    def ingest(self, voltages, valid=None) -> None:
        """Append spectra [n][F][A]: a host uint8 array or an api.DeviceArray,
        with optional validity bytes [n][F] of the same kind (0 = lost frame;
        needs a server with loss_masks)."""
        row = self.cfg.n_freq * self.cfg.n_ant
        if isinstance(voltages, np.ndarray):
            v = np.ascontiguousarray(voltages, np.uint8)
            n, on_dev, src = v.size // row, 0, v.ctypes.data
        else:
            n, on_dev, src = voltages.nbytes // row, 1, voltages.ptr
        if valid is None:
            call("bf_server_ingest", self._h, src, n, on_dev)
            return
        if on_dev:
            if valid.nbytes != n * self.cfg.n_freq:
                raise ValueError("valid must hold [n_spectra][n_freq] bytes")
            call("bf_server_ingest_masked", self._h, src, valid.ptr, n, 1)
        else:
            m = np.ascontiguousarray(valid, np.uint8)
            if m.size != n * self.cfg.n_freq:
                raise ValueError("valid must hold [n_spectra][n_freq] bytes")
            call("bf_server_ingest_masked", self._h, src, m.ctypes.data, n, 0)

    def acquire(self, n_spectra: int) -> list:
        """Ring rows for the next n_spectra spectra, written in place by the
        caller: a list of one or two dicts (first_spectrum, n_spectra,
        voltages, valid) with device addresses."""
        spans = (RingSpanC * 2)()
        n = _sz()
        call("bf_server_acquire", self._h, n_spectra, spans, C.byref(n))
        return [{"first_spectrum": spans[i].first_spectrum, "n_spectra": spans[i].n_spectra,
                 "voltages": spans[i].voltages or 0, "valid": spans[i].valid or 0}
                for i in range(n.value)]

    def commit(self, n_spectra: int, stream: int = 0) -> None:
        """Publishes acquired spectra once the work queued on `stream`
        (a hipStream_t address, 0 = null stream) has finished."""
        call("bf_server_commit", self._h, n_spectra, stream or None)

    def ingest_device(self, ptr: int, n_spectra: int) -> None:
        call("bf_server_ingest", self._h, ptr, n_spectra, 1)

    def step(self) -> int:
        n = _sz()
        call("bf_server_step", self._h, C.byref(n))
        return n.value

    def drain(self) -> int:
        """Steps until every ingested spectrum has been released to a step."""
        total = 0
        while True:
            st = self.stats()
            if st["processed_head"] >= st["head"] and st["requests_pending"] == 0:
                break
            before = st["processed_head"]
            total += self.step()
            if self.stats()["processed_head"] == before:
                break
        return total

    def status(self, rid: int) -> dict:
        s = RequestStatusC()
        call("bf_server_request_status", self._h, rid, C.byref(s))
        return {n: getattr(s, n) for n, _ in s._fields_}

    def read(self, rid: int, first: int, n: int, n_beams: int) -> np.ndarray:
        out = np.empty((n, self.cfg.n_freq, n_beams), np.float32)
        call("bf_server_read_output", self._h, rid, first, n, out.ctypes.data)
        return out

    def release(self, rid: int) -> None:
        call("bf_server_release", self._h, rid)

    def synchronize(self) -> None:
        call("bf_server_synchronize", self._h)

    def stats(self) -> dict:
        s = StatsC()
        call("bf_server_get_stats", self._h, C.byref(s))
        return s.as_dict()

    def reset_stats(self) -> None:
        call("bf_server_reset_stats", self._h)


@dataclass
class SimConfig:
    server: ServerConfig = field(default_factory=ServerConfig)
    seed: int = 2024
    cadence_s: float = 2.56e-6
    run_seconds: float = 4.0
    chunk_spectra: int = 128
    arrival_rate_hz: float = 6.0
    min_beams: int = 1
    max_beams: int = 64
    min_duration_s: float = 0.5
    max_duration_s: float = 5.0
    window_a: int = 10
    window_b: int = 320
    output_windows_cap: int = 1024
    paced: bool = True
    pool_spectra: int = 16384
    source_amplitude: float = 4.33
    noise_sigma: float = 2.5


def simulate(cfg: SimConfig) -> dict:
    """Runs the native load generator (bf_server_simulate) and returns its report."""
    keep: list = []
    c = SimConfigC(cfg.server.to_c(keep), cfg.seed, cfg.cadence_s, cfg.run_seconds,
                   cfg.chunk_spectra, cfg.arrival_rate_hz, cfg.min_beams, cfg.max_beams,
                   cfg.min_duration_s, cfg.max_duration_s, cfg.window_a, cfg.window_b,
                   cfg.output_windows_cap, int(cfg.paced), cfg.pool_spectra,
                   cfg.source_amplitude, cfg.noise_sigma)
    r = SimReportC()
    call("bf_server_simulate", C.byref(c), C.byref(r))
    return r.as_dict()
