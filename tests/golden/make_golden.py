"""Writes the golden fixtures under tests/golden/ from the reference CPU core.

The reference's C++ CPU sources are compiled in place by oracle/Makefile
(`make -C oracle ref`) into oracle/_ref/libbfref.so; this script calls that
library through oracle/ref_shim.cpp and stores small input/output vectors.
Each case mirrors a case of the reference's own tests:

  weights_a32_b5        tests/cpu/test_cpu_beamformer.cpp:76-110
  noise_bytes           src/synthetic_data.cpp:107-119 (make_noise, seed 4242)
  beamform_noise_a32_b5 tests/cpu/test_cpu_beamformer.cpp:207-237 (seeded noise)
  beamform_noise_a64_b7 tests/cuda/test_cuda_packed.cu:198-202 (fft grid, seed 9876)
  point_source_a32      tests/cpu/test_cpu_beamformer.cpp:162-181
  integrate_ramp        tests/cpu/test_temporal_integration.cpp:59-90
  quantize_ramp         tests/cpu/test_quantization.cpp:38-62
  tracker_moving_a32    tests/cuda/test_cuda_tracker_v2.cpp:45-58 (shortened)
  io_quantized_meta.txt src/io.cpp:149-183 (write_quantized_output_metadata)
  io_shard_*.bin/.txt   src/io.cpp:244-278 (write_packed_shard, T=2 A=32)
  loss_mask_seed7       src/synthetic_data.cpp:121-137 (make_loss_mask, T=64)
  geometry_grids        tests/cpu/test_geometry.cpp (rectangular, FFT-bin, hex and
                        line grids, FFT bin selection, hex beam estimates)
  opt_tracker_*         tests/cpu/test_cpu_opt_beam_tracker.cpp:290-645 (the
                        adaptive tracker: Bartlett, Capon, coarse-only with
                        interpolation, A=64, forgetting on a moving source,
                        state carried across calls, scan disabled, smoothing);
                        inputs are regenerated from the stored source and
                        checked against the stored CRC-32

Run:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys
import zlib

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as O  # noqa: E402

F = 336


# start: the prior's direction_start; rate its direction_rate_per_sample;
# source: the (moving) point source l0, m0, rate_l, rate_m of the input.
_Z = (0.0, 0.0, 1.0)
OPT_CASES = {
    "bartlett_static": dict(A=32, T=4, S=4, source=(0.05, -0.03, 0.0, 0.0), start=_Z,
                            rate=(0.0, 0.0), estimator=0, G=16, levels=2, fov=0.2, p_sub=0,
                            lam=1.0),
    "capon_static": dict(A=32, T=4, S=4, source=(0.05, -0.03, 0.0, 0.0), start=_Z,
                         rate=(0.0, 0.0), estimator=1, G=16, levels=2, fov=0.2, p_sub=0,
                         lam=1.0),
    "coarse_interp": dict(A=32, T=4, S=4, source=(0.031, 0.017, 0.0, 0.0), start=_Z,
                          rate=(0.0, 0.0), estimator=0, G=12, levels=0, fov=0.2, p_sub=0,
                          lam=1.0),
    "a64_bartlett": dict(A=64, T=4, S=4, source=(-0.04, 0.06, 0.0, 0.0), start=_Z,
                         rate=(0.0, 0.0), estimator=0, G=16, levels=2, fov=0.2, p_sub=0,
                         lam=1.0),
    "moving_forgetting": dict(A=32, T=8, S=2, source=(0.02, -0.01, 5e-3, 2e-3), start=_Z,
                              rate=(0.0, 0.0), estimator=0, G=16, levels=1, fov=0.3, p_sub=0,
                              lam=0.9),
    "capon_stateful": dict(A=32, T=4, S=2, calls=2, source=(0.01, 0.02, 4e-3, -3e-3),
                           start=(0.01, 0.0, 0.99995), rate=(0.0, 0.0), estimator=1, G=12,
                           levels=1, fov=0.2, p_sub=0, lam=0.95),
    "scan_off": dict(A=32, T=8, S=2, source=(0.02, -0.01, 1e-3, 5e-4),
                     start=(0.02, -0.01, 0.99975), rate=(1e-3, 5e-4), estimator=0, G=1,
                     levels=0, fov=0.2, p_sub=0, lam=1.0),
    "partial_odd_capon": dict(A=64, T=7, S=3, source=(-0.02, 0.03, 2e-3, -1e-3),
                              start=(0.0, 0.0, 1.0), rate=(0.0, 0.0), estimator=1, G=5,
                              levels=3, fov=0.15, p_sub=0, lam=0.8),
    "edge_peak": dict(A=32, T=4, S=4, source=(0.17, -0.12, 0.0, 0.0), start=_Z,
                      rate=(0.0, 0.0), estimator=0, G=8, levels=0, fov=0.1, p_sub=0, lam=1.0),
    "smoothing_bartlett": dict(A=32, T=4, S=4, source=(0.05, -0.03, 0.0, 0.0), start=_Z,
                               rate=(0.0, 0.0), estimator=0, G=16, levels=1, fov=0.2, p_sub=24,
                               lam=1.0),
}


def main() -> None:
    R = O.ref_lib()
    if R is None:
        raise SystemExit("oracle/_ref/libbfref.so missing: run `make -C oracle ref`")

    def pos(A, spacing=0.6):
        p = np.zeros((A, 3), np.float32)
        assert R.ref_default_positions(A, spacing, p) == 0
        return p

    freqs = np.zeros(F, np.float32)
    assert R.ref_channelized_frequencies(F, 300e6, 300e3, freqs) == 0

    def beam_grid(n):
        d = np.zeros((n, 3), np.float32)
        assert R.ref_default_beam_grid(n, 0.02, 0.0, d) == 0
        return d

    def weights(A, dirs):
        B = len(dirs)
        w = np.zeros((B, F, A, 2), np.float32)
        assert R.ref_generate_weights(B, F, A, pos(A), freqs, dirs, w) == 0
        return w

    def noise(T, A, seed):
        x = np.zeros(T * F * A, np.uint8)
        assert R.ref_make_noise(T, F, A, seed, x) == 0
        return x.reshape(T, F, A)

    def beamform(x, w):
        T, _, A = x.shape
        B = w.shape[0]
        out = np.zeros((T, F, B), np.float32)
        assert R.ref_beamform_packed(x.reshape(-1), w.reshape(-1), T, F, A, B, out) == 0
        return out

    save = lambda name, **kw: np.savez_compressed(os.path.join(HERE, name + ".npz"), **kw)

    # weights + tiled weights on the default 5-beam grid (A = 32)
    dirs5 = beam_grid(5)
    w5 = weights(32, dirs5)
    tiled = np.zeros((F, 1, 32, 32, 2), np.float32)
    assert R.ref_generate_tiled_weights(5, F, 32, pos(32), freqs, dirs5, tiled) == 0
    save("weights_a32_b5", positions=pos(32), frequencies=freqs, directions=dirs5,
         weights=w5, tiled=tiled)

    x = noise(2, 32, 4242)
    save("noise_bytes", seed=np.uint32(4242), shape=np.array(x.shape), bytes=x)
    save("beamform_noise_a32_b5", x=x, directions=dirs5, intensity=beamform(x, w5))

    dirs7 = np.zeros((7, 3), np.float32)
    assert R.ref_fft_beam_grid(64, 7, dirs7) == 0
    x64 = noise(10, 64, 9876)
    save("beamform_noise_a64_b7", x=x64, directions=dirs7,
         intensity=beamform(x64, weights(64, dirs7)))

    d004 = np.zeros(3, np.float32)
    assert R.ref_direction_from_lm(0.04, 0.0, d004) == 0
    ps = np.zeros(F * 32, np.uint8)
    assert R.ref_make_point_source(1, F, 32, pos(32), freqs, d004, 4.0, ps) == 0
    ps = ps.reshape(1, F, 32)
    save("point_source_a32", direction=d004, x=ps, intensity=beamform(ps, w5))

    T, B, S = 7, 3, 3
    t, f, b = np.meshgrid(np.arange(T), np.arange(F), np.arange(B), indexing="ij")
    ramp = (100 * t + 10 * f + b).astype(np.float32)
    integ = np.zeros((3, F, B), np.float32)
    assert R.ref_integrate(ramp.reshape(-1), T, F, 32, B, S, integ) == 0
    save("integrate_ramp", input=ramp, spectra=np.int64(S), output=integ)

    T, B = 2, 17
    t, f, b = np.meshgrid(np.arange(T), np.arange(F), np.arange(B), indexing="ij")
    qin = (1000 * t + 10 * f + b).astype(np.float32)
    codes = np.zeros((T, F, B), np.int8)
    params = np.zeros((T * 21 * 2, 2), np.float32)
    assert R.ref_quantize(qin.reshape(-1), T, F, 32, B, codes, params) == 0
    save("quantize_ramp", input=qin, codes=codes, params=params)

    T, A, S = 40, 32, 16
    l0, m0, rl, rm = 0.02, -0.01, 1.0e-4, 5.0e-5
    xm = np.zeros(T * F * A, np.uint8)
    assert R.ref_make_moving_point_source(T, F, A, pos(A), freqs, l0, m0, rl, rm, 4.0, xm) == 0
    out = np.zeros((T, F, 1), np.float32)
    assert R.ref_beam_tracker(xm, T, F, A, l0, m0, rl, rm, S, out) == 0
    out2 = np.zeros_like(out)
    assert R.ref_beam_tracker_opt_v2(xm, T, F, A, l0, m0, rl, rm, S, out2) == 0
    assert np.array_equal(out, out2)
    save("tracker_moving_a32", x=xm.reshape(T, F, A),
         trajectory=np.array([l0, m0, rl, rm], np.float32), spectra=np.int64(S),
         intensity=out)
    # On-disk formats written by the reference's own io.cpp.
    meta = os.path.join(HERE, "io_quantized_meta.txt")
    # C2-like output: T=32768 -> 3277 windows of 10, A=64, B=128, shard 1
    assert R.ref_write_quantized_metadata(meta.encode(), 3277, F, 64, 128, 32768, 10, 1,
                                          b"c2_params.bin") == 0
    T, A = 2, 32
    payload = np.arange(T * F * A, dtype=np.uint32).astype(np.uint8)
    mask = (np.arange(T * F) % 7 != 3).astype(np.uint8)
    assert R.ref_write_packed_shard(os.path.join(HERE, "io_shard_payload.bin").encode(),
                                    os.path.join(HERE, "io_shard_meta.txt").encode(),
                                    os.path.join(HERE, "io_shard_mask.bin").encode(),
                                    payload, mask, T, F, A, 5, 0, 1234567890123, 2560) == 0
    # Beam-grid generators (src/geometry.cpp:108-366), cases of
    # tests/cpu/test_geometry.cpp plus off-default spacings.
    import ctypes as C
    grids = {}
    for A in (32, 64):
        r = np.zeros((A, 3), np.float32)
        assert R.ref_rectangular_beam_grid(A, 0.6, 400e6, r) == 0
        grids[f"rect_a{A}"] = r
        r2 = np.zeros((A, 3), np.float32)
        assert R.ref_rectangular_beam_grid(A, 0.7, 450e6, r2) == 0
        grids[f"rect_a{A}_s07_f450"] = r2
        for B in (1, 7, 128):
            g = np.zeros((B, 3), np.float32)
            assert R.ref_fft_beam_grid(A, B, g) == 0
            grids[f"fft_a{A}_b{B}"] = g
        g = np.zeros((50, 3), np.float32)
        assert R.ref_fft_beam_grid_at(A, 50, 0.5, 450e6, g) == 0
        grids[f"fft_a{A}_b50_s05_f450"] = g
    for n, um, vm in ((128, 0.8, 0.6), (37, 0.3, 0.5), (1, 0.2, 0.2)):
        h = np.zeros((n, 2), np.float32)
        cnt = C.c_size_t()
        assert R.ref_hex_targets(n, um, vm, h, C.byref(cnt)) == 0
        grids[f"hex_{n}_{um}_{vm}"] = h[:cnt.value]
    est = []
    for args in ((0.6, 8, 8, 108.0, 74.0, 0.75), (0.6, 8, 4, 60.0, 40.0, 0.9993),
                 (0.3, 16, 16, 20.0, 30.0, 1.5)):
        e = C.c_size_t()
        assert R.ref_estimate_hex(*args, C.byref(e)) == 0
        est.append(e.value)
    grids["hex_estimates"] = np.array(est, np.int64)
    for B, m_side, n_side in ((128, 8, 4), (128, 8, 8), (13, 8, 8)):
        cen = np.zeros((B, 2), np.float32)
        shape = (C.c_size_t * 4)()
        du = np.float32(0.0390303)
        assert R.ref_select_fft(du, du * np.float32(2.0 if n_side == 4 else 1.0), B, m_side,
                                n_side, cen, shape) == 0
        grids[f"select_{B}_{m_side}x{n_side}"] = cen
        grids[f"select_{B}_{m_side}x{n_side}_shape"] = np.array(list(shape), np.int64)
    d7 = np.zeros((7, 3), np.float32)
    assert R.ref_default_beam_grid(7, 0.015, 0.05, d7) == 0
    grids["line_7_0.015_0.05"] = d7
    save("geometry_grids", **grids)
    lm = np.zeros((64, F), np.uint8)
    assert R.ref_make_loss_mask(64, F, 32, 7, 0.125, lm) == 0
    save("loss_mask_seed7", seed=np.int64(7), probability=np.float32(0.125), mask=lm)
    # Adaptive DOA tracker (CpuOptBeamTracker::run_into) cases.
    for name, case in OPT_CASES.items():
        A, T, n = case["A"], case["T"], case.get("calls", 1)
        src = case["source"]  # l0, m0, rate_l, rate_m
        x = np.zeros(n * T * F * A, np.uint8)
        assert R.ref_make_moving_point_source(n * T, F, A, pos(A), freqs, *src, 4.0, x) == 0
        out, dirs = O.ref_opt_tracker(R, x, n, T, F, A, pos(A), freqs, case["start"],
                                      case["rate"], case["S"], case["estimator"], case["G"],
                                      case["levels"], case["fov"], case["fov"], case["p_sub"],
                                      case["lam"], 1e-3, True)
        save("opt_tracker_" + name, A=np.int64(A), T=np.int64(T), calls=np.int64(n),
             S=np.int64(case["S"]), source=np.array(src, np.float32),
             start=np.array(case["start"], np.float32), rate=np.array(case["rate"], np.float32),
             estimator=np.int64(case["estimator"]), G=np.int64(case["G"]),
             levels=np.int64(case["levels"]), fov=np.float32(case["fov"]),
             p_sub=np.int64(case["p_sub"]), lam=np.float32(case["lam"]),
             x_crc=np.int64(zlib.crc32(x.tobytes())), intensity=out, directions=dirs)
    for name in sorted(os.listdir(HERE)):
        if name.endswith(".npz") or name.startswith("io_"):
            print(f"{name:32s} {os.path.getsize(os.path.join(HERE, name)):>9d} bytes")


if __name__ == "__main__":
    main()
