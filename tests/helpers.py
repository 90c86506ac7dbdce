"""Shared helpers for the parity tests (tolerances written out here)."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
F = 336

# Reference CPU-vs-GPU tolerance: |d| <= 1e-3 + 1e-5 |cpu|
# (tests/cuda/test_cuda_packed.cu:24-25, 113-120).
REF_ABS = 1.0e-3
REF_REL = 1.0e-5


def golden(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)


def assert_ref_close(got, ref, abs_tol=REF_ABS, rel_tol=REF_REL, what=""):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    assert got.shape == ref.shape, (got.shape, ref.shape)
    d = np.abs(got - ref)
    bad = d > abs_tol + rel_tol * np.abs(ref)
    if bad.any():
        i = np.unravel_index(np.argmax(d - (abs_tol + rel_tol * np.abs(ref))), d.shape)
        raise AssertionError(f"{what}: {bad.sum()} mismatches; worst at {i}: got {got[i]} "
                             f"ref {ref[i]}")


def tracker_close(got, ref, rel_tol, abs_tol, what=""):
    """Tracker tolerance of tests/cuda/test_cuda_tracker_v2.cpp:23-36:
    a cell fails when |d| > abs_tol and |d| > rel_tol * ref."""
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    d = np.abs(got - ref)
    bad = (d > abs_tol) & (d > rel_tol * ref)
    assert not bad.any(), f"{what}: {bad.sum()} cells outside tolerance, max diff {d.max()}"


def rel_err_floored(got, ref64, floor):
    """Max relative error vs an FP64 oracle with the denominator floored at
    `floor` (the incoherent per-beam power), as defined for this build."""
    got = np.asarray(got, np.float64)
    den = np.maximum(np.abs(ref64), floor)
    return float(np.max(np.abs(got - ref64) / den))


def incoherent_floor(x, n_ant, spectra=1):
    """tau = A * E|x|^2 (times S for integrated outputs)."""
    b = np.asarray(x, np.uint8).reshape(-1)
    re = ((b & 0xF).astype(np.int16) ^ 8) - 8
    im = ((b >> 4).astype(np.int16) ^ 8) - 8
    ex2 = float(np.mean(re.astype(np.float64) ** 2 + im.astype(np.float64) ** 2))
    return n_ant * ex2 * spectra
