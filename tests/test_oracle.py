"""Pins the CPU oracle before it is trusted as the checker.

Each oracle function is checked against (a) the golden vectors written by
the reference CPU core (tests/golden/make_golden.py), (b) the live reference
library where oracle/_ref was built in this container, and (c) closed forms
and invariants chosen so that a plausible mistake (dropped term, wrong sign,
transposed operand) fails at least one of them.
"""
import os

import numpy as np
import pytest

import oracle as O
from helpers import F, GOLDEN, golden

A32 = 32


def decode(b):
    b = np.asarray(b, np.int16)
    re = (b & 0xF) - 16 * ((b & 0xF) >= 8)
    im = ((b >> 4) & 0xF) - 16 * (((b >> 4) & 0xF) >= 8)
    return re, im


# ------------------------------------------------------------------ int4
def test_int4_exhaustive_closed_form():
    # tests/cpu/test_contract.cpp:29-42: every byte decodes to two's-complement nibbles.
    raw = np.arange(256, dtype=np.uint8)
    re = np.zeros(256, np.int8)
    im = np.zeros(256, np.int8)
    O.LIB.or_unpack_bytes(raw, 256, re, im)
    ere, eim = decode(raw)
    assert np.array_equal(re, ere) and np.array_equal(im, eim)
    assert O.LIB.or_decode_nibble(0x7) == 7 and O.LIB.or_decode_nibble(0x8) == -8
    # pack/unpack round trip over the whole [-8, 7]^2 square
    out = np.zeros(1, np.uint8).ctypes
    import ctypes
    b = ctypes.c_uint8()
    for r in range(-8, 8):
        for i in range(-8, 8):
            assert O.LIB.or_pack(r, i, ctypes.byref(b)) == 0
            assert decode(np.array([b.value]))[0][0] == r
            assert decode(np.array([b.value]))[1][0] == i
    assert O.LIB.or_pack(8, 0, ctypes.byref(b)) != 0
    del out


# -------------------------------------------------------------- geometry
def test_geometry_closed_forms():
    p = O.default_positions(64, 0.6)
    assert p.shape == (64, 3)
    assert np.allclose(p[9], [0.6, 0.6, 0.0])  # row 1, column 1 of the 8x8 grid
    assert np.allclose(O.default_positions(32)[8], [0.0, 0.6, 0.0])  # 4 rows x 8 columns
    assert np.allclose(O.default_positions(128)[16], [0.0, 0.6, 0.0])  # 8 x 16
    fr = O.channelized_frequencies()
    assert fr[0] == np.float32(300e6) and fr[1] == np.float32(300.3e6)
    d = O.direction_from_lm(0.6, 0.0)
    assert np.isclose(d[2], 0.8)
    with pytest.raises(ValueError):
        O.direction_from_lm(0.9, 0.9)


# --------------------------------------------------------------- weights
def test_weights_match_reference_fixture():
    g = golden("weights_a32_b5")
    w = O.generate_weights(g["positions"], g["frequencies"], g["directions"])
    assert np.array_equal(w, g["weights"])  # bit-exact float32
    t = O.generate_tiled_weights(g["positions"], g["frequencies"], g["directions"])
    assert np.array_equal(t, g["tiled"])
    # tiled padding is zero and the tiled/canonical views agree
    assert np.all(t[:, 0, :, 5:, :] == 0)
    assert np.array_equal(t[:, 0, :, :5, :].transpose(2, 0, 1, 3), w)


def test_weights_unit_modulus_and_phase_closed_form():
    pos = O.default_positions(32)
    fr = O.channelized_frequencies()
    dirs = O.default_beam_grid(5)
    w = O.generate_weights(pos, fr, dirs).astype(np.float64)
    mag = w[..., 0] ** 2 + w[..., 1] ** 2
    assert np.abs(mag - 1).max() < 1e-6
    k = 2 * np.pi * fr.astype(np.float64) / 299792458.0
    ph = k[None, :, None] * (pos.astype(np.float64) @ dirs.astype(np.float64).T).T[:, None, :]
    assert np.abs(w[..., 0] - np.cos(ph)).max() < 1e-7
    assert np.abs(w[..., 1] - np.sin(ph)).max() < 1e-7  # exp(+j phi), not exp(-j phi)


# ----------------------------------------------------------- beamformer
def test_beamformer_matches_reference_fixtures():
    g = golden("weights_a32_b5")
    for name, A in (("beamform_noise_a32_b5", 32), ("point_source_a32", 32)):
        c = golden(name)
        w = O.generate_weights(O.default_positions(A), O.channelized_frequencies(),
                               c["directions"] if "directions" in c else g["directions"])
        T = c["x"].shape[0]
        B = w.shape[0]
        out = O.beamform(c["x"], w, T, F, A, B)
        assert np.array_equal(out, c["intensity"]), name
    c = golden("beamform_noise_a64_b7")
    w = O.generate_weights(O.default_positions(64), O.channelized_frequencies(), c["directions"])
    out = O.beamform(c["x"], w, 10, F, 64, 7)
    assert np.array_equal(out, c["intensity"])
    out_mt = O.beamform(c["x"], w, 10, F, 64, 7, threads=True)
    assert np.array_equal(out_mt, c["intensity"])


def test_noise_bytes_match_reference():
    c = golden("noise_bytes")
    x = O.make_noise(*[int(v) for v in c["shape"]], int(c["seed"]))
    assert np.array_equal(x, c["bytes"])
    # make_noise never emits -8 (src/synthetic_data.cpp:113-115)
    re, im = decode(x)
    assert re.min() == -7 and im.min() == -7 and re.max() == 7


def test_mt19937_standard_value():
    # C++ [rand.predef]: the 10000th output of a default-seeded mt19937 is 4123659995.
    import ctypes
    g = O.or_mt = None
    st = (ctypes.c_uint32 * 625)()
    O.LIB.or_mt_seed.argtypes = [ctypes.c_void_p, ctypes.c_uint32]
    O.LIB.or_mt_next.argtypes = [ctypes.c_void_p]
    O.LIB.or_mt_next.restype = ctypes.c_uint32
    O.LIB.or_mt_seed(st, 5489)
    for _ in range(9999):
        O.LIB.or_mt_next(st)
    assert O.LIB.or_mt_next(st) == 4123659995
    del g


def test_beamformer_closed_forms():
    pos = O.default_positions(32)
    fr = O.channelized_frequencies()
    dirs = O.default_beam_grid(5)
    w = O.generate_weights(pos, fr, dirs)
    # one-hot {3,-2} at (t=0, f=17, a=7) -> |w|^2 |x|^2 = 13 at f=17 only
    x = np.zeros((1, F, 32), np.uint8)
    x[0, 17, 7] = (3 & 0xF) | ((-2 & 0xF) << 4)
    out = O.beamform(x, w, 1, F, 32, 5)
    assert np.allclose(out[0, 17], 13.0, atol=1e-3)
    assert np.all(out[0, np.arange(F) != 17] == 0)
    # broadside constant {1,0} -> A^2;  {7,-8} -> A^2 * 113 exactly
    bw = O.generate_weights(pos, fr, O.direction_from_lm(0.0, 0.0)[None])
    x1 = np.full((2, F, 32), 1, np.uint8)
    assert np.allclose(O.beamform(x1, bw, 2, F, 32, 1), 32 * 32, atol=1e-3)
    x78 = np.full((2, F, 32), (7 & 0xF) | ((-8 & 0xF) << 4), np.uint8)
    assert np.all(O.beamform(x78, bw, 2, F, 32, 1) == 32 * 32 * 113)
    # point source at l = 0.04 peaks at beam 4 of the 5-beam grid
    ps = O.make_point_source(1, F, 32, pos, fr, O.direction_from_lm(0.04, 0.0), 4.0)
    assert np.argmax(O.beamform(ps, w, 1, F, 32, 5)[0].sum(axis=0)) == 4
    # rectangular grid: the injected beam 12 wins
    rg = O.rectangular_beam_grid(32)
    wr = O.generate_weights(pos, fr, rg)
    ps12 = O.make_point_source(1, F, 32, pos, fr, rg[12], 4.0)
    assert np.argmax(O.beamform(ps12, wr, 1, F, 32, 32)[0].sum(axis=0)) == 12


def test_fp64_oracle_brute_force_and_invariants():
    rng = np.random.default_rng(1)
    T, A, B = 2, 32, 3
    x = rng.integers(0, 256, size=(T, F, A), dtype=np.uint8)
    w = rng.standard_normal((B, F, A, 2)).astype(np.float32)
    out = O.beamform_f64(x, w, T, F, A, B)
    re, im = decode(x)
    xc = re + 1j * im
    wc = w[..., 0].astype(np.float64) + 1j * w[..., 1]
    y = np.einsum("tfa,bfa->tfb", xc, wc)  # sum_a W X, no conjugate
    assert np.allclose(out, np.abs(y) ** 2, rtol=1e-12, atol=1e-9)
    # a conjugated contraction differs: the oracle is not conjugation-blind
    yc = np.einsum("tfa,bfa->tfb", xc, np.conj(wc))
    assert not np.allclose(out, np.abs(yc) ** 2, rtol=1e-3)
    # linearity in |W|: I(cW) = |c|^2 I(W)
    out2 = O.beamform_f64(x, (2 * w).astype(np.float32), T, F, A, B)
    assert np.allclose(out2, 4 * out, rtol=1e-12)
    # float32 oracle agrees with FP64 to float rounding
    out32 = O.beamform(x, w, T, F, A, B)
    assert np.allclose(out32, out, rtol=1e-4, atol=1e-3)


# ---------------------------------------------------------- integration
def test_integration_fixture_and_counts():
    c = golden("integrate_ramp")
    T, _, B = c["input"].shape
    out = O.integrate(c["input"], T, F, B, int(c["spectra"]))
    assert np.array_equal(out, c["output"])
    # tests/cpu/test_temporal_integration.cpp:48-58
    assert O.integrated_time_count(480, 10) == 48
    assert O.integrated_time_count(15360, 320) == 48
    assert O.integrated_time_count(1, 320) == 1
    assert O.integrated_time_count(301, 320) == 1
    assert O.integrated_time_count(321, 320) == 2
    # partial final window: S = 10 over 8 spectra of ones -> 8
    ones = np.ones((18, F, 1), np.float32)
    o = O.integrate(ones, 18, F, 1, 10)
    assert np.all(o[0] == 10.0) and np.all(o[1] == 8.0)


# ------------------------------------------------------------ quantiser
def test_quantizer_fixture_and_properties():
    c = golden("quantize_ramp")
    T, _, B = c["input"].shape
    codes, params = O.quantize(c["input"], T, F, B)
    assert np.array_equal(codes, c["codes"])
    assert np.array_equal(params, c["params"])
    assert params.shape[0] == 84  # tests/cpu/test_quantization.cpp:50-53
    assert codes[0, 0, 0] == -127 and codes[0, 15, 15] == 127
    assert codes[0, 0, 16] == -127 and codes[0, 15, 16] == 127
    rec = O.dequantize(codes, params, T, F, B)
    scale = params.reshape(T, 21, 2, 2)[..., 1]
    sc = np.repeat(np.repeat(scale, 16, axis=1)[:, :F], 16, axis=2)[:, :, :B]
    assert np.all(np.abs(rec - c["input"]) <= sc * 0.5 + 5e-4)
    # constant chunk -> scale 0, offset = value, code 0; NaN -> -128 and zero params
    const = np.full((1, F, 1), 42.0, np.float32)
    cc, cp = O.quantize(const, 1, F, 1)
    assert np.all(cp[:, 1] == 0) and np.all(cp[:, 0] == 42) and np.all(cc == 0)
    const[0, 0, 0] = np.nan
    nc, npar = O.quantize(const, 1, F, 1)
    assert nc[0, 0, 0] == -128 and npar[0, 0] == 0 and npar[0, 1] == 0


# --------------------------------------------------------------- tracker
def test_tracker_fixture_and_properties():
    c = golden("tracker_moving_a32")
    l0, m0, rl, rm = [float(v) for v in c["trajectory"]]
    x = c["x"]
    T = x.shape[0]
    S = int(c["spectra"])
    out = O.beam_tracker(x, T, F, 32, l0, m0, rl, rm, S)
    assert np.array_equal(out, c["intensity"])
    out_mt = O.beam_tracker(x, T, F, 32, l0, m0, rl, rm, S, threads=True)
    assert np.array_equal(out_mt, c["intensity"])  # naive == opt_v2, bit-exact
    # tests/cpu/test_beam_tracker.cpp:79-114
    assert O.LIB.or_tracker_window_count(0, 320) == 0
    assert O.LIB.or_tracker_window_count(321, 320) == 2
    assert O.LIB.or_tracker_window_count(641, 320) == 3
    d10 = O.tracker_direction(0.02, -0.03, 1e-4, -2e-4, 10)
    assert np.array_equal(d10, O.direction_from_lm(np.float32(0.02) + np.float32(10) * np.float32(1e-4),
                                                   np.float32(-0.03) + np.float32(10) * np.float32(-2e-4)))
    with pytest.raises(ValueError):
        O.tracker_direction(0.9, 0.0, 0.01, 0.0, 20)


def test_stationary_tracker_equals_fixed_grid_bit_exact():
    # tests/cpu/test_beam_tracker.cpp:120-139
    A, T = 32, 4
    pos = O.default_positions(A)
    fr = O.channelized_frequencies()
    d = O.direction_from_lm(0.04, 0.0)
    x = O.make_point_source(T, F, A, pos, fr, d, 4.0)
    grid = O.beamform(x, O.generate_weights(pos, fr, d[None]), T, F, A, 1)
    trk = O.beam_tracker(x, T, F, A, 0.04, 0.0, 0.0, 0.0, T)
    assert np.array_equal(grid, trk)


def test_multibeam_tracker_reduces_to_fixed_grid():
    A, T, S, B = 32, 20, 8, 3
    pos = O.default_positions(A)
    fr = O.channelized_frequencies()
    dirs = O.default_beam_grid(B)
    W = (T + S - 1) // S
    x = O.make_noise(T, F, A, 7)
    table = np.repeat(dirs[None], W, axis=0)
    out, out64 = O.multibeam_tracker(x, T, F, A, B, S, pos, fr, table, want_f64=True)
    ref = O.integrate(O.beamform(x, O.generate_weights(pos, fr, dirs), T, F, A, B), T, F, B, S)
    assert np.array_equal(out, ref)
    assert np.allclose(out64, ref, rtol=1e-5)
    spec = O.multibeam_tracker(x, T, F, A, B, S, pos, fr, table, spectra=True)
    assert np.array_equal(spec, O.beamform(x, O.generate_weights(pos, fr, dirs), T, F, A, B))


# ------------------------------------------------------------------- DOA
def test_covariance_exact_and_hermitian():
    A, T = 16 * 4, 50
    x = O.make_uniform_int4(T * F * A, 3).reshape(T, F, A)
    R = O.covariance_i64(x, T, F, A, 5, 40, 11)
    re, im = decode(x[5:45, 11])
    xc = re + 1j * im
    ref = xc.T @ np.conj(xc)  # sum_k x_r conj(x_c)
    assert np.array_equal(R[..., 0], ref.real.astype(np.int64))
    assert np.array_equal(R[..., 1], ref.imag.astype(np.int64))
    assert np.array_equal(R[..., 0], R[..., 0].T) and np.array_equal(R[..., 1], -R[..., 1].T)
    assert np.trace(R[..., 0]) == int(np.sum(re.astype(np.int64) ** 2 + im.astype(np.int64) ** 2))
    Rs = O.smooth_forward_i64(R, A, 48)
    L = A - 48 + 1
    exp = sum(ref[l:l + 48, l:l + 48] for l in range(L))
    assert np.array_equal(Rs[..., 0], exp.real.astype(np.int64))


def test_cholesky_capon_bartlett_against_linear_algebra():
    rng = np.random.default_rng(5)
    M = 12
    Z = rng.standard_normal((M, 40)) + 1j * rng.standard_normal((M, 40))
    Rc = Z @ Z.conj().T / 40
    R = np.stack([Rc.real, Rc.imag], -1).copy()
    L, ok = O.cholesky_f64(R)
    assert ok
    Lc = np.tril(L[..., 0] + 1j * L[..., 1])
    assert np.abs(Lc @ Lc.conj().T - Rc).max() < 1e-12
    a = np.exp(-1j * rng.uniform(0, 2 * np.pi, M))
    av = np.stack([a.real, a.imag], -1).copy()
    capon = O.LIB.or_capon_from_factor_f64(L, av, M)
    assert np.isclose(capon, 1.0 / np.real(a.conj() @ np.linalg.solve(Rc, a)), rtol=1e-10)
    bart = O.LIB.or_bartlett_f64(R, av, M)
    assert np.isclose(bart, np.real(a.conj() @ Rc @ a), rtol=1e-12)
    # a non-positive-definite matrix is rejected
    bad = R.copy()
    bad[0, 0, 0] = -1.0
    assert not O.cholesky_f64(bad)[1]


def test_quadratic_peak_offset_closed_form():
    # samples of p(u) = 5 - (u - 0.3)^2 at u = -1, 0, 1 -> vertex at +0.3
    p = lambda u: 5 - (u - 0.3) ** 2
    assert np.isclose(O.LIB.or_quadratic_peak_offset(p(-1), p(0), p(1)), 0.3)
    # a minimum (convex) gives no offset; clamping at +-0.5
    assert O.LIB.or_quadratic_peak_offset(1.0, 0.0, 1.0) == 0.0
    q = lambda u: -(u - 0.9) ** 2
    assert O.LIB.or_quadratic_peak_offset(q(-1), q(0), q(1)) == 0.5


def test_doa_peaks_against_find_peaks_and_closed_forms():
    from scipy.signal import find_peaks
    # the local maxima of tie-free data are scipy's peaks; the list holds
    # the strongest, in descending power
    rng = np.random.default_rng(5)
    theta = np.linspace(-1.5, 1.5, 721)
    for max_peaks in (1, 4, 1000):
        P = rng.random(721)
        idx, _ = find_peaks(P)
        want = idx[np.argsort(-P[idx], kind="stable")][:max_peaks]
        got = O.doa_peaks(P, theta, max_peaks)
        assert len(got) == len(want)
        assert np.array_equal(got[:, 1], P[want])
        assert np.all(np.abs(got[:, 0] - theta[want]) <= 0.5 * (theta[1] - theta[0]) + 1e-15)
    # exact parabolas: the vertex comes back, whatever the grid offset
    th = np.linspace(-1.0, 1.0, 201)
    vert = [(-0.4031, 3.0), (0.1207, 5.0), (0.6655, 4.0)]
    P = np.max([h - 50.0 * (th - c) ** 2 for c, h in vert], axis=0)
    got = O.doa_peaks(P, th, 3)
    assert np.allclose(got[:, 0], [0.1207, 0.6655, -0.4031], atol=1e-12)
    # equal powers: the earlier angle first; a plateau counts at its left edge;
    # no peak at either end; fewer peaks than slots
    P = np.array([9.0, 0.0, 2.0, 0.0, 3.0, 3.0, 0.0, 2.0, 0.0, 9.0])
    got = O.doa_peaks(P, np.arange(10.0), 8)
    assert np.array_equal(got[:, 1], [3.0, 2.0, 2.0])
    assert got[1, 0] < got[2, 0] and got[0, 0] <= 4.5
    assert len(O.doa_peaks(np.zeros(5), np.arange(5.0), 3)) == 0


def test_doa_update_recovers_single_source():
    A, M, K = 64, 48, 256
    fr = np.linspace(400e6, 800e6, F).astype(np.float32)
    xpos = (0.3 * np.arange(A)).astype(np.float32)
    theta0 = np.deg2rad(17.3)
    k = 2 * np.pi * fr.astype(np.float64) / 299792458.0
    rng = np.random.default_rng(2)
    sig = np.exp(1j * rng.uniform(0, 2 * np.pi, K))
    ph = k[None, :, None] * xpos[None, None, :] * np.sin(theta0)
    v = 4 * sig[:, None, None] * np.exp(-1j * ph) + rng.normal(0, 1.0, (K, F, A)) * (1 + 1j)
    re = np.clip(np.rint(v.real), -8, 7).astype(np.int16)
    im = np.clip(np.rint(v.imag), -8, 7).astype(np.int16)
    x = ((re & 0xF) | ((im & 0xF) << 4)).astype(np.uint8)
    theta = np.deg2rad(np.linspace(-90, 90, 721))
    R_state = np.zeros((F, M, M, 2), np.float64)
    capon, bart, ok = O.doa_update(x, K, F, A, 0, K, M, 0.98, 1e-3, True, fr, xpos, theta,
                                   R_state, np.zeros(F, np.float64))
    assert ok.all()
    for P in (capon.sum(0), bart.sum(0)):
        i = int(np.argmax(P))
        d = O.LIB.or_quadratic_peak_offset(P[i - 1], P[i], P[i + 1])
        est = np.rad2deg(theta[i]) + d * 0.25
        assert abs(est - 17.3) < 0.1


def _doa_recursions(x, Fs, A, M, K, n_up, alpha, load, variant="reference"):
    """Capon / Bartlett operands of the reference's per-window recursion,
    built here from decoded voltages with numpy (src/cpu_opt_beam_tracker.cpp
    :1864-1892): R_block = (1/(K L)) sum_l sum_k x_l x_l^H over the forward
    sub-arrays; Capon folds R_block + load tr(R_block)/M I, Bartlett folds
    R_block.  `variant` plants a plausible mistake instead."""
    L = A - M + 1
    re, im = decode(x)
    xc = (re + 1j * im).astype(np.complex128)
    RL = [None] * Fs
    RU = [None] * Fs
    out = []
    for u in range(n_up):
        blk = xc[u * K:(u + 1) * K]
        for f in range(Fs):
            X = blk[:, f, :]
            Rb = sum(X[:, l:l + M].T @ np.conj(X[:, l:l + M]) for l in range(L))
            Rb = Rb / (K * (L if variant != "scale_k_only" else 1))
            eps = load * np.trace(Rb).real / M
            a_old, a_new = (1 - alpha, alpha) if variant == "alpha_swap" else (alpha, 1 - alpha)
            if variant == "load_compounds":
                # the loaded state folded, then loaded again by its own trace
                RL[f] = Rb.copy() if u == 0 else a_old * RL[f] + a_new * Rb
                RL[f] = RL[f] + load * np.trace(RL[f]).real / M * np.eye(M)
            else:
                Rl = Rb + eps * np.eye(M)
                RL[f] = Rl if u == 0 else a_old * RL[f] + a_new * Rl
            RU[f] = Rb if u == 0 else a_old * RU[f] + a_new * Rb
        out.append(([r.copy() for r in RL],
                    [(RL[f] if variant == "bartlett_loaded" else RU[f]).copy()
                     for f in range(Fs)]))
    return out


def test_doa_update_recursion_matches_reference_definition():
    """or_doa_update over four updates on distinct blocks against the
    reference's recursion: each block is loaded (Capon only) and then folded
    with alpha, Bartlett scans the unloaded fold.  The blocks' powers differ
    by up to 16x, so an alpha <-> 1-alpha swap, a load compounded into the
    state, a Bartlett taken from the loaded R and a 1/K scale without the
    sub-array count each move some output far past the tolerance (asserted
    below on the same data)."""
    A, M, K, Fs, n_up = 16, 12, 48, 3, 4
    alpha, load = 0.98, 0.05
    rng = np.random.default_rng(21)
    amps = [3.0, 0.75, 2.0, 1.2]
    v = np.concatenate([a * (rng.standard_normal((K, Fs, A)) + 1j * rng.standard_normal((K, Fs, A)))
                        for a in amps])
    re = np.clip(np.rint(v.real), -8, 7).astype(np.int16)
    im = np.clip(np.rint(v.imag), -8, 7).astype(np.int16)
    x = ((re & 0xF) | ((im & 0xF) << 4)).astype(np.uint8)
    T = n_up * K
    fr = np.array([400e6, 555.5e6, 800e6], np.float32)
    xpos = (0.25 * np.arange(A)).astype(np.float32)
    theta = np.linspace(-np.pi / 2, np.pi / 2, 37)
    k = 2 * np.pi * fr.astype(np.float64) / 299792458.0

    def spectra(RL, RU):
        cp = np.zeros((Fs, len(theta)))
        bt = np.zeros((Fs, len(theta)))
        for f in range(Fs):
            for q, th in enumerate(theta):
                a = np.exp(-1j * k[f] * xpos[:M].astype(np.float64) * np.sin(th))
                cp[f, q] = 1.0 / np.real(a.conj() @ np.linalg.solve(RL[f], a))
                bt[f, q] = np.real(a.conj() @ RU[f] @ a)
        return cp, bt

    ref = _doa_recursions(x, Fs, A, M, K, n_up, alpha, load)
    R_state = np.zeros((Fs, M, M, 2), np.float64)
    load_state = np.zeros(Fs, np.float64)
    got = []
    for u in range(n_up):
        cp, bt, ok = O.doa_update(x, T, Fs, A, u * K, K, M, alpha, load, u == 0, fr, xpos,
                                  theta, R_state, load_state)
        assert ok.all()
        RL, RU = ref[u]
        for f in range(Fs):
            Rc = R_state[f, ..., 0] + 1j * R_state[f, ..., 1]
            assert np.abs(Rc - RU[f]).max() <= 1e-13 * np.abs(RU[f]).max(), (u, f)
            assert np.abs(Rc + load_state[f] * np.eye(M) - RL[f]).max() <= \
                1e-13 * np.abs(RL[f]).max(), (u, f)
        rc, rb = spectra(RL, RU)
        assert np.max(np.abs(cp - rc) / np.abs(rc)) < 1e-10, u
        assert np.max(np.abs(bt - rb) / np.abs(rb)) < 1e-10, u
        got.append((cp, bt))
    # each planted mistake is visible on this data at 1e4 x the tolerance
    for variant in ("alpha_swap", "load_compounds", "bartlett_loaded", "scale_k_only"):
        bad = _doa_recursions(x, Fs, A, M, K, n_up, alpha, load, variant)
        worst = 0.0
        for u in range(n_up):
            bc, bb = spectra(*bad[u])
            cp, bt = got[u]
            worst = max(worst, np.max(np.abs(cp - bc) / np.abs(bc)),
                        np.max(np.abs(bt - bb) / np.abs(bb)))
        assert worst > 1e-6, variant


def test_reference_doa_primitives_agree_with_the_fp64_update():
    """The C5 CPU baseline (the reference's FP32 DOA primitives from
    src/cpu_opt_beam_tracker.cpp, composed by oracle/ref_doa_shim.cpp)
    computes the same broadband spectra as the FP64 restatement over three
    updates with forgetting and a diagonal load, to FP32 accuracy.  Without
    smoothing (M = A) the reference's covariance has no forward-backward
    step, so both sides define the same quantity."""
    lib = O.ref_doa_lib(native=False)
    if lib is None:
        pytest.skip("oracle/_ref was not built (no reference tree)")
    A = M = 32
    K, Fs, n_up = 256, 4, 3
    alpha, load = 0.98, 1e-3
    x = O.make_uniform_int4(n_up * K * Fs * A, 17)
    T = n_up * K
    fr = np.linspace(400e6, 800e6, Fs).astype(np.float32)
    xpos = (0.3 * np.arange(A)).astype(np.float32)
    theta = np.linspace(-np.pi / 2, np.pi / 2, 181)
    R64 = np.zeros((Fs, M, M, 2), np.float64)
    L64 = np.zeros(Fs, np.float64)
    Rb = np.zeros(Fs * M * M * 2, np.float32)
    Rc = np.zeros(Fs * M * M * 2, np.float32)
    for u in range(n_up):
        cp, bt, ok = O.doa_update(x, T, Fs, A, u * K, K, M, alpha, load, u == 0, fr, xpos,
                                  theta, R64, L64)
        pb = np.zeros(len(theta), np.float32)
        pc = np.zeros(len(theta), np.float32)
        assert lib.ref_doa_update_f32(x, T, Fs, A, u * K, K, M, alpha, load, int(u == 0), fr,
                                      xpos, len(theta), theta, Rb, Rc, pb, pc) == 0
        assert ok.all()
        eb = np.max(np.abs(pb - bt.sum(0)) / bt.sum(0).max())
        ec = np.max(np.abs(pc - cp.sum(0)) / np.abs(cp.sum(0)))
        assert eb < 1e-5 and ec < 1e-3, (u, eb, ec)


def test_opt_tracker_restatement_matches_reference_fixtures():
    """The numpy restatement of the adaptive search (oracle.opt_tracker_search)
    against the reference's CpuOptBeamTracker directions, with the
    reference's in-place forward-backward step for the smoothing case."""
    import glob
    import zlib
    for path in sorted(glob.glob(os.path.join(GOLDEN, "opt_tracker_*.npz"))):
        c = np.load(path, allow_pickle=False)
        if int(c["G"]) <= 1:
            continue
        A, T, n = int(c["A"]), int(c["T"]), int(c["calls"])
        pos, fr = O.default_positions(A), O.channelized_frequencies(F)
        x = O.make_moving_point_source(n * T, F, A, pos, fr, *[float(v) for v in c["source"]])
        assert zlib.crc32(x.tobytes()) == int(c["x_crc"])
        x = x.reshape(n, T, F, A)
        state = None
        for k in range(n):
            d, state = O.opt_tracker_search(
                x[k], int(c["S"]), pos, fr, c["start"], c["rate"], int(c["estimator"]),
                int(c["G"]), int(c["levels"]), float(c["fov"]), float(c["fov"]),
                int(c["p_sub"]), float(c["lam"]), 1e-3, True, state, fb_in_place=True)
            assert np.array_equal(d, c["directions"][k]), (path, k)
