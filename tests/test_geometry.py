"""Geometry, beam grids and weight producers (host functions of the C-ABI;
no GPU).  The cases of the reference's tests/cpu/test_geometry.cpp, and
bit-for-bit equality with the reference's own outputs: the geometry_grids
and weights_a32_b5 fixtures written by its src/geometry.cpp and
src/weights.cpp (tests/golden/make_golden.py), and the live reference
library where it is built."""
import os

import numpy as np
import pytest

import oracle as O
from beamform_project_amd import _capi, api, geometry as G
from helpers import F, golden

LAMBDA_400 = np.float32(299792458.0 / 400e6)


def test_reference_geometry_cases(tmp_path):
    # tests/cpu/test_geometry.cpp, case by case
    p32 = G.default_positions(32)
    assert p32.shape == (32, 3) and np.all(p32[0] == 0)
    assert abs(p32[7, 0] - 4.2) < 1e-6 and abs(p32[8, 1] - 0.6) < 1e-6
    assert abs(p32[31, 0] - 4.2) < 1e-6 and abs(p32[31, 1] - 1.8) < 1e-6
    assert np.array_equal(G.default_positions(64, 0.5)[63], [3.5, 3.5, 0.0])
    assert G.default_positions(128, 0.5).shape == (128, 3)
    assert G.default_positions(256, 0.5).shape == (256, 3)
    fr = G.channelized_frequencies()
    assert fr.shape == (336,) and fr[0] == 300e6 and fr[1] == 300.3e6 and fr[-1] == 400.5e6
    b1 = G.default_beam_grid(1)
    assert np.array_equal(b1[0], [0.0, 0.0, 1.0])
    b5 = G.default_beam_grid(5)
    assert b5[0, 0] == np.float32(-0.04) and b5[2, 0] == 0 and b5[4, 0] == np.float32(0.04)
    b10 = G.default_beam_grid(10)
    assert b10[7, 0] == np.float32(0.04)
    assert np.all(np.abs(np.linalg.norm(b10, axis=1) - 1) < 1e-6)
    assert G.default_beam_grid(128, 0.005).shape == (128, 3)
    r32 = G.rectangular_beam_grid(32)
    dl = LAMBDA_400 / (np.float32(8) * np.float32(0.6))
    dm = LAMBDA_400 / (np.float32(4) * np.float32(0.6))
    assert abs(r32[1, 0] - r32[0, 0] - dl) < 1e-6 and abs(r32[8, 1] - r32[0, 1] - dm) < 1e-6
    assert abs(r32[12, 0] - 0.5 * dl) < 1e-6 and abs(r32[12, 1] + 0.5 * dm) < 1e-6
    assert r32[0, 0] < 0 and r32[0, 1] < 0 and r32[-1, 0] > 0 and r32[-1, 1] > 0
    assert G.rectangular_beam_grid(64).shape == (64, 3)
    assert G.centered_integer_range(4) == [-2, -1, 0, 1]
    du = LAMBDA_400 / (np.float32(16) * np.float32(0.6))
    dv32 = LAMBDA_400 / (np.float32(8) * np.float32(0.6))
    s32 = G.select_fft_beam_centers_rectangular(du, dv32, 128, 8, 4)
    assert (s32.n_u, s32.n_v, s32.n_bank_u, s32.n_bank_v) == (16, 8, 16, 8)
    assert np.array_equal(s32.centers[0], [0, 0])
    assert np.all(np.abs(np.linalg.norm(G.fft_beam_grid(32, 128), axis=1) - 1) < 1e-6)
    s64 = G.select_fft_beam_centers_rectangular(du, du, 128, 8, 8)
    assert (s64.n_u, s64.n_v, s64.n_bank_u, s64.n_bank_v) == (12, 11, 16, 16)
    assert G.estimate_nbeams_hex_formula(0.6, 8, 8, 108.0, 74.0, LAMBDA_400) > 0
    hx = G.generate_hex_targets_cropped_fov(128, 0.8, 0.6)
    assert hx.shape == (128, 2)
    assert np.all(hx[:, 0] ** 2 / 0.64 + hx[:, 1] ** 2 / 0.36 <= 1 + 1e-6)
    pp, fp = tmp_path / "p.txt", tmp_path / "f.txt"
    pp.write_text("# x, y, z in metres\n1.0, 2.0, 3.0\n4.0 5.0 6.0\n")
    fp.write_text("# Hz\n400000000\n401000000\n")
    assert np.array_equal(G.load_positions(pp, 2), [[1, 2, 3], [4, 5, 6]])
    assert np.array_equal(G.load_frequencies(fp, 2), [400e6, 401e6])
    with pytest.raises(_capi.InvalidArgument, match="32, 64, 128, or 256"):
        G.default_positions(48)


def test_grids_equal_reference_outputs():
    g = golden("geometry_grids")
    for A in (32, 64):
        assert np.array_equal(G.rectangular_beam_grid(A), g[f"rect_a{A}"])
        assert np.array_equal(G.rectangular_beam_grid(A, 0.7, 450e6), g[f"rect_a{A}_s07_f450"])
        for B in (1, 7, 128):
            assert np.array_equal(G.fft_beam_grid(A, B), g[f"fft_a{A}_b{B}"]), (A, B)
        assert np.array_equal(G.fft_beam_grid(A, 50, 0.5, 450e6), g[f"fft_a{A}_b50_s05_f450"])
    for n, um, vm in ((128, 0.8, 0.6), (37, 0.3, 0.5), (1, 0.2, 0.2)):
        assert np.array_equal(G.generate_hex_targets_cropped_fov(n, um, vm),
                              g[f"hex_{n}_{um}_{vm}"]), n
    est = [G.estimate_nbeams_hex_formula(*a) for a in
           ((0.6, 8, 8, 108.0, 74.0, 0.75), (0.6, 8, 4, 60.0, 40.0, 0.9993),
            (0.3, 16, 16, 20.0, 30.0, 1.5))]
    assert est == g["hex_estimates"].tolist()
    du = np.float32(0.0390303)
    for B, m_side, n_side in ((128, 8, 4), (128, 8, 8), (13, 8, 8)):
        s = G.select_fft_beam_centers_rectangular(
            du, du * np.float32(2.0 if n_side == 4 else 1.0), B, m_side, n_side)
        assert np.array_equal(s.centers, g[f"select_{B}_{m_side}x{n_side}"])
        assert [s.n_u, s.n_v, s.n_bank_u, s.n_bank_v] == \
            g[f"select_{B}_{m_side}x{n_side}_shape"].tolist()
    assert np.array_equal(G.default_beam_grid(7, 0.015, 0.05), g["line_7_0.015_0.05"])


def test_weights_equal_reference_bits():
    g = golden("weights_a32_b5")
    dims = api.Dimensions(1, F, 32, 5)
    w = G.generate_weights(dims, g["positions"], g["frequencies"], g["directions"])
    assert np.array_equal(w, g["weights"])
    t = G.generate_tiled_weights(dims, g["positions"], g["frequencies"], g["directions"])
    assert np.array_equal(t, g["tiled"])


def test_grids_equal_live_reference_where_built():
    R = O.ref_lib()
    if R is None:
        pytest.skip("oracle/_ref not built")
    fr, pos = G.channelized_frequencies(), G.default_positions(64)
    dirs = G.fft_beam_grid(64, 128)
    want = np.zeros((128, F, 64, 2), np.float32)
    assert R.ref_generate_weights(128, F, 64, pos, fr, dirs, want) == 0
    assert np.array_equal(G.generate_weights(api.Dimensions(1, F, 64, 128), pos, fr, dirs), want)


def test_validation_messages(tmp_path):
    with pytest.raises(_capi.InvalidArgument, match="between 1 and 128"):
        G.default_beam_grid(129)
    with pytest.raises(_capi.InvalidArgument, match="only for 32 or 64 antennas"):
        G.fft_beam_grid(128, 8)
    with pytest.raises(_capi.InvalidArgument, match="zero-padded"):
        G.fft_beam_grid(32, 129)
    with pytest.raises(_capi.InvalidArgument, match="l\\*l \\+ m\\*m <= 1"):
        G.direction_from_lm(0.8, 0.8)
    with pytest.raises(_capi.InvalidArgument, match="unit vectors"):
        G.generate_weights(api.Dimensions(1, F, 32, 1), G.default_positions(32),
                           G.channelized_frequencies(), np.array([[0.5, 0.5, 0.5]], np.float32))
    bad = tmp_path / "bad.txt"
    bad.write_text("1.0 2.0\n")
    with pytest.raises(_capi.HipRuntimeError, match="expected three coordinates"):
        G.load_positions(bad, 1)
    with pytest.raises(_capi.HipRuntimeError, match="contains 1 rows; expected 2"):
        G.load_frequencies(_one_freq(tmp_path), 2)
    nd = tmp_path / "d.txt"
    nd.write_text("0.5 0.5 0.5\n")
    with pytest.raises(_capi.HipRuntimeError, match="finite unit vectors"):
        G.load_directions(nd, 1)


def _one_freq(tmp_path):
    p = tmp_path / "one.txt"
    p.write_text("400000000\n")
    return p
