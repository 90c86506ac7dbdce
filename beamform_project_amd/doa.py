"""Python binding of the adaptive DOA path (C-ABI bf_doa_*)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _capi
from ._capi import call
from .api import _ptr, _sh


class DoaConfigC(C.Structure):
    _fields_ = [("n_freq", C.c_size_t), ("n_ant", C.c_size_t), ("snapshots", C.c_size_t),
                ("subarray", C.c_size_t), ("n_theta", C.c_size_t), ("max_peaks", C.c_size_t),
                ("forgetting", C.c_double), ("diagonal_load", C.c_double),
                ("theta_min_rad", C.c_double), ("theta_max_rad", C.c_double),
                ("element_spacing_m", C.c_double), ("frequencies", C.c_void_p)]


class DoaTimingsC(C.Structure):
    _fields_ = [("covariance_ms", C.c_double), ("fold_scan_ms", C.c_double),
                ("estimate_ms", C.c_double)]


_vp, _sz = C.c_void_p, C.c_size_t
for _name, _args in {
    "bf_doa_create": [C.POINTER(DoaConfigC), C.POINTER(_vp)],
    "bf_doa_destroy": [_vp],
    "bf_doa_reset": [_vp],
    "bf_doa_update": [_vp, _vp, _vp, _sz, _sz],
    "bf_doa_read_results": [_vp, _vp, _vp, _vp],
    "bf_doa_read_channel_spectra": [_vp, _vp, _vp, _vp],
    "bf_doa_read_covariance": [_vp, _vp],
    "bf_doa_get_state": [_vp, _vp, C.POINTER(C.c_int)],
    "bf_doa_set_state": [_vp, _vp, C.c_int],
    "bf_doa_last_timings": [_vp, C.POINTER(DoaTimingsC)],
}.items():
    _capi.register(_name, _args)


@dataclass
class DoaConfig:
    frequencies: np.ndarray
    n_ant: int = 64
    snapshots: int = 1024
    subarray: int = 48
    n_theta: int = 721
    max_peaks: int = 4
    forgetting: float = 0.98
    diagonal_load: float = 1e-3
    theta_min_rad: float = -np.pi / 2
    theta_max_rad: float = np.pi / 2
    element_spacing_m: float = 0.3


class DoaEstimator:
    """Stateful adaptive DOA estimator on the GPU (bf_doa)."""

    def __init__(self, cfg: DoaConfig):
        self._h = None
        self.cfg = cfg
        self._fr = np.ascontiguousarray(cfg.frequencies, np.float32)
        c = DoaConfigC(len(self._fr), cfg.n_ant, cfg.snapshots, cfg.subarray, cfg.n_theta,
                       cfg.max_peaks, cfg.forgetting, cfg.diagonal_load, cfg.theta_min_rad,
                       cfg.theta_max_rad, cfg.element_spacing_m, self._fr.ctypes.data)
        h = C.c_void_p()
        call("bf_doa_create", C.byref(c), C.byref(h))
        self._h = h.value
        self.F = len(self._fr)
        self.M = cfg.subarray
        self.theta = np.linspace(cfg.theta_min_rad, cfg.theta_max_rad, cfg.n_theta)

    def close(self):
        if self._h:
            _capi.LIB.bf_doa_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reset(self):
        call("bf_doa_reset", self._h)

    def update(self, d_packed, n_time: int, t0: int, stream=None):
        call("bf_doa_update", self._h, _sh(stream), _ptr(d_packed), n_time, t0)

    def results(self):
        n, P = self.cfg.n_theta, self.cfg.max_peaks
        bb = np.zeros((2, n), np.float64)
        pk = np.zeros((2, P, 2), np.float64)
        npk = np.zeros(2, np.int32)
        call("bf_doa_read_results", self._h, bb.ctypes.data, pk.ctypes.data, npk.ctypes.data)
        return {"capon": bb[0], "bartlett": bb[1],
                "capon_peaks": pk[0, :npk[0]], "bartlett_peaks": pk[1, :npk[1]]}

    def last_timings(self) -> dict:
        t = DoaTimingsC()
        call("bf_doa_last_timings", self._h, C.byref(t))
        return {n: getattr(t, n) for n, _ in t._fields_}

    def channel_spectra(self):
        n = self.cfg.n_theta
        cp = np.zeros((self.F, n), np.float64)
        bt = np.zeros((self.F, n), np.float64)
        ok = np.zeros(self.F, np.int32)
        call("bf_doa_read_channel_spectra", self._h, cp.ctypes.data, bt.ctypes.data,
             ok.ctypes.data)
        return cp, bt, ok

    def covariance(self):
        rs = np.zeros((self.F, self.M, self.M, 2), np.int32)
        call("bf_doa_read_covariance", self._h, rs.ctypes.data)
        return rs

    def get_state(self):
        st = np.zeros((self.F, self.M, self.M, 2), np.float64)
        ini = C.c_int()
        call("bf_doa_get_state", self._h, st.ctypes.data, C.byref(ini))
        return st, bool(ini.value)

    def set_state(self, state: np.ndarray, initialised: bool):
        state = np.ascontiguousarray(state, np.float64)
        call("bf_doa_set_state", self._h, state.ctypes.data, int(initialised))
