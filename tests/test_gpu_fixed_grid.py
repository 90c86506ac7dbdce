"""GPU parity of the fixed-grid path (stage launchers, workspace, device
frames, one-shot API) against the CPU oracle and the reference's golden
vectors.

Tolerances:
  * reference FP32 kernels (DIRECT/TILED/FP32) vs the CPU FP32 path:
    |d| <= 1e-3 + 1e-5 |cpu|  (tests/cuda/test_cuda_packed.cu:24-25)
  * SPLIT_F16 vs FP64 oracle: max floored relative error <= 1e-5
  * F16 (single) vs FP64 oracle: max floored relative error <= 2e-3
  * int8 codes: exactly equal to the CPU quantiser applied to the GPU floats
    (tests/cuda/test_cuda_quantization.cpp:29)
"""
import numpy as np
import pytest

import oracle as O
from beamform_project_amd import _capi, api
from helpers import F, assert_ref_close, golden, incoherent_floor, rel_err_floored

pytestmark = pytest.mark.gpu

K = api.Kernel
SPLIT_TOL = 1.0e-5
F16_TOL = 2.0e-3


def run_gpu(kernel, x, w, dims, S=0):
    """Runs one stage launch on device copies; returns the host result."""
    dx = api.DeviceArray.from_numpy(x)
    dw = api.DeviceArray.from_numpy(w)
    if S:
        W = api.integrated_time_count(dims.n_time, S)
        out = api.DeviceArray.empty((W, dims.n_freq, dims.n_beams), np.float32)
        out.fill(0xFF)
        api.launch_packed_integrated_beamformer(kernel, None, dx, dw, out, dims, S)
    else:
        out = api.DeviceArray.empty((dims.n_time, dims.n_freq, dims.n_beams), np.float32)
        out.fill(0xFF)
        api.launch_packed_beamformer(kernel, None, dx, dw, out, dims)
    api.device_synchronize()
    return out.to_numpy()


def random_case(T, A, B, seed):
    rng = np.random.default_rng(seed)
    x = O.make_uniform_int4(T * F * A, seed).reshape(T, F, A)
    pos = O.regular_array(A // 8, 8, 0.3)
    fr = np.linspace(400e6, 800e6, F).astype(np.float32)
    l = rng.uniform(-0.5, 0.5, B)
    m = rng.uniform(-0.5, 0.5, B)
    dirs = np.stack([O.direction_from_lm(a, b) for a, b in zip(l, m)])
    return x, O.generate_weights(pos, fr, dirs)


def cpu_reference(x, w, T, A, B, S):
    i = O.beamform(x, w, T, F, A, B, threads=True)
    return O.integrate(i, T, F, B, S) if S else i


def fp64_reference(x, w, T, A, B, S):
    i = O.beamform_f64(x, w, T, F, A, B)
    return O.integrate_f64(i, T, F, B, S) if S else i


# ------------------------------------------------------------- fixtures
@pytest.mark.parametrize("kernel", [K.DIRECT, K.FP32, K.SPLIT_F16, K.F16])
@pytest.mark.parametrize("case", ["beamform_noise_a32_b5", "beamform_noise_a64_b7",
                                  "point_source_a32"])
def test_golden_fixtures(kernel, case):
    c = golden(case)
    x = c["x"]
    T, _, A = x.shape
    dirs = c["directions"] if "directions" in c else golden("weights_a32_b5")["directions"]
    w = O.generate_weights(O.default_positions(A), O.channelized_frequencies(), dirs)
    B = w.shape[0]
    got = run_gpu(kernel, x, w, api.Dimensions(T, F, A, B))
    if kernel == K.F16:
        ref64 = fp64_reference(x, w, T, A, B, 0)
        assert rel_err_floored(got, ref64, incoherent_floor(x, A)) <= F16_TOL
    else:
        assert_ref_close(got, c["intensity"], what=f"{case}/{kernel.name}")


def test_tiled_layout_fixture():
    g = golden("weights_a32_b5")
    c = golden("beamform_noise_a32_b5")
    got = run_gpu(K.TILED, c["x"], g["tiled"], api.Dimensions(2, F, 32, 5))
    assert_ref_close(got, c["intensity"], what="tiled")


# ---------------------------------------------------------- known answers
@pytest.mark.parametrize("kernel", [K.DIRECT, K.FP32, K.SPLIT_F16])
def test_known_answers(kernel):
    pos = O.default_positions(32)
    fr = O.channelized_frequencies()
    dims = api.Dimensions(1, F, 32, 5)
    w = O.generate_weights(pos, fr, O.default_beam_grid(5))
    x = np.zeros((1, F, 32), np.uint8)
    x[0, 17, 7] = (3 & 0xF) | ((-2 & 0xF) << 4)
    out = run_gpu(kernel, x, w, dims)
    assert np.allclose(out[0, 17], 13.0, atol=1e-3)
    assert np.all(np.abs(out[0, np.arange(F) != 17]) <= 1e-6)
    # constant {7,-8} on broadside weights -> A^2 * 113 exactly; boundary nibbles
    bw = O.generate_weights(pos, fr, O.direction_from_lm(0.0, 0.0)[None])
    for re, im in ((7, -8), (-8, -8), (-8, 7), (7, 7)):
        xc = np.full((2, F, 32), (re & 0xF) | ((im & 0xF) << 4), np.uint8)
        out = run_gpu(kernel, xc, bw, api.Dimensions(2, F, 32, 1))
        assert np.all(out == 32 * 32 * (re * re + im * im)), (re, im)


def test_exhaustive_byte_decode_through_the_mfma_path():
    # Every byte value at every antenna slot: with a one-hot beam weight per
    # antenna the output of beam b is |x_b|^2 = re^2 + im^2 exactly.
    A, B = 32, 32
    T = 8
    bytes_ = np.arange(256, dtype=np.uint8)
    x = bytes_[np.arange(T * F * A) % 256].reshape(T, F, A)
    w = np.zeros((B, F, A, 2), np.float32)
    for b in range(B):
        w[b, :, b, 0] = 1.0
    re = ((x.astype(np.int16) & 0xF) ^ 8) - 8
    im = (((x.astype(np.int16) >> 4) & 0xF) ^ 8) - 8
    expect = (re * re + im * im).astype(np.float32)  # [T][F][A] with B == A
    for kernel in (K.SPLIT_F16, K.F16, K.FP32):
        out = run_gpu(kernel, x, w, api.Dimensions(T, F, A, B))
        assert np.array_equal(out, expect), kernel.name


# ------------------------------------------------------------ random parity
CASES = [
    # (T, A, B, S)
    (37, 32, 16, 0),
    (50, 64, 5, 0),
    (96, 64, 128, 10),
    (333, 64, 17, 10),
    (640, 64, 40, 320),
    (700, 32, 3, 320),
    (64, 128, 33, 8),
    (40, 256, 9, 0),
    (100, 256, 128, 20),
    (21, 128, 1, 10),
]


@pytest.mark.parametrize("T,A,B,S", CASES)
def test_fp32_kernels_match_cpu(T, A, B, S):
    x, w = random_case(T, A, B, seed=T * 7 + A + B)
    dims = api.Dimensions(T, F, A, B)
    ref = cpu_reference(x, w, T, A, B, S)
    got = run_gpu(K.FP32, x, w, dims, S)
    assert_ref_close(got, ref, what=f"FP32 T={T} A={A} B={B} S={S}")


@pytest.mark.parametrize("T,A,B,S", CASES)
def test_mfma_kernels_match_fp64(T, A, B, S):
    x, w = random_case(T, A, B, seed=T * 3 + A + B)
    dims = api.Dimensions(T, F, A, B)
    ref64 = fp64_reference(x, w, T, A, B, S)
    floor = incoherent_floor(x, A, S if S else 1)
    got = run_gpu(K.SPLIT_F16, x, w, dims, S)
    err = rel_err_floored(got, ref64, floor)
    assert err <= SPLIT_TOL, f"split-f16 error {err:.3e}"
    # and against the reference FP32 semantics at the reference tolerance
    assert_ref_close(got, cpu_reference(x, w, T, A, B, S), what="split vs cpu")
    got16 = run_gpu(K.F16, x, w, dims, S)
    err16 = rel_err_floored(got16, ref64, floor)
    assert err16 <= F16_TOL, f"f16 error {err16:.3e}"


def test_reference_tiled_integration_paths():
    T, A, B = 650, 64, 33
    x, w = random_case(T, A, B, seed=11)
    pos = O.regular_array(8, 8, 0.3)
    ref10 = cpu_reference(x, w, T, A, B, 10)
    dims = api.Dimensions(T, F, A, B)
    # Tiled layout with S = 10 and 320, Direct with 320.
    tiles = (B + 31) // 32
    tw = np.zeros((F, tiles, A, 32, 2), np.float32)
    for b in range(B):
        tw[:, b // 32, :, b % 32, :] = w[b]
    assert_ref_close(run_gpu(K.TILED, x, tw, dims, 10), ref10, what="tiled S=10")
    ref320 = cpu_reference(x, w, T, A, B, 320)
    assert_ref_close(run_gpu(K.TILED, x, tw, dims, 320), ref320, what="tiled S=320")
    assert_ref_close(run_gpu(K.DIRECT, x, w, dims, 320), ref320, what="direct S=320")
    del pos


# ------------------------------------------------------------- quantiser
@pytest.mark.parametrize("kernel", [K.TILED, K.SPLIT_F16])
def test_int8_codes_exact(kernel):
    T, A, B = 1010, 64, 37
    x, w = random_case(T, A, B, seed=23)
    if kernel == K.TILED:
        tiles = (B + 31) // 32
        wk = np.zeros((F, tiles, A, 32, 2), np.float32)
        for b in range(B):
            wk[:, b // 32, :, b % 32, :] = w[b]
    else:
        wk = w
    dims = api.Dimensions(T, F, A, B)
    (codes, params), t = api.beamform_packed_quantized_integrated_intensity(
        x, wk, dims, 10, kernel, with_timings=True)
    floats = api.beamform_packed_integrated_intensity(x, wk, dims, 10, kernel)
    W = floats.shape[0]
    rc, rp = O.quantize(floats, W, F, B)
    assert np.array_equal(codes, rc)
    assert np.array_equal(params, rp)
    assert t["kernel_ms"] > 0 and t["quantization_ms"] > 0


def test_quantizer_special_values():
    Tw, B = 3, 20
    rng = np.random.default_rng(4)
    v = rng.uniform(0, 1000, (Tw, F, B)).astype(np.float32)
    v[0, :16, :16] = 42.0          # constant chunk
    v[1, 3, 5] = np.nan            # NaN chunk
    v[2, 20, 17] = np.inf          # inf chunk
    d_in = api.DeviceArray.from_numpy(v)
    codes = api.DeviceArray.empty((Tw, F, B), np.int8)
    npar = api.quantization_parameter_count(api.Dimensions(Tw, F, 32, B))
    params = api.DeviceArray.empty((npar, 2), np.float32)
    api.launch_quantize_integrated_intensity(None, d_in, codes, params,
                                             api.Dimensions(Tw, F, 32, B))
    api.device_synchronize()
    rc, rp = O.quantize(v, Tw, F, B)
    assert np.array_equal(codes.to_numpy(), rc)
    assert np.array_equal(params.to_numpy(), rp)
    assert codes.to_numpy()[1, 3, 5] == -128


@pytest.mark.parametrize("B", [16, 48, 128])
def test_quantizer_16_beam_rows(B):
    """B a multiple of 16 takes the 16-byte-row quantiser, whose quotients
    are formed by a reciprocal and re-divided near rounding boundaries: the
    codes equal the CPU quantiser's on random chunks, on chunks whose
    quotients are exact half-integers (ties to even), on a chunk with a
    subnormal scale, and on constant, NaN and inf chunks."""
    Tw = 7
    rng = np.random.default_rng(B)
    v = rng.uniform(0, 1000, (Tw, F, B)).astype(np.float32)
    v[1] = rng.exponential(50.0, (F, B)).astype(np.float32)
    # min 0, max 255 -> scale 1, offset 127.5: every integer value is a tie
    v[2, :16, :16] = rng.integers(0, 256, (16, 16)).astype(np.float32)
    v[2, 0, 0], v[2, 1, 1] = 0.0, 255.0
    # span 510 -> scale 2: odd integers give ties, even ones integers
    v[2, 16:32, :16] = rng.integers(0, 511, (16, 16)).astype(np.float32)
    v[2, 16, 0], v[2, 17, 1] = 0.0, 510.0
    v[3, :16, :16] = rng.uniform(0, 1e-38, (16, 16)).astype(np.float32)  # subnormal scale
    v[4, :16, :16] = 42.0          # constant chunk
    v[5, 3, 5] = np.nan            # NaN chunk
    v[6, 20, B - 1] = np.inf       # inf chunk
    v[6, 40, 2] = -np.inf
    d_in = api.DeviceArray.from_numpy(v)
    codes = api.DeviceArray.empty((Tw, F, B), np.int8)
    npar = api.quantization_parameter_count(api.Dimensions(Tw, F, 32, B))
    params = api.DeviceArray.empty((npar, 2), np.float32)
    api.launch_quantize_integrated_intensity(None, d_in, codes, params,
                                             api.Dimensions(Tw, F, 32, B))
    api.device_synchronize()
    rc, rp = O.quantize(v, Tw, F, B)
    assert np.array_equal(params.to_numpy(), rp)
    assert np.array_equal(codes.to_numpy(), rc)
    got = codes.to_numpy()
    assert got[5, 3, 5] == -128 and got[6, 20, B - 1] == -128 and got[6, 40, 2] == -128
    assert np.all(got[4, :16, :16] == 0)


# ------------------------------------------------------------- workspace
def test_workspace_contract():
    cap = api.Dimensions(128, F, 64, 16)
    ws = api.BeamformerWorkspace(cap, K.DIRECT)
    assert not ws.has_temporal_integration() and not ws.has_int8_quantization()
    assert ws.packed_voltage_capacity_bytes() == 128 * F * 64
    with pytest.raises(_capi.LogicError):
        ws.run_kernel(cap)  # src/cuda_beamformer.cu:581-589
    x, w = random_case(100, 64, 12, seed=5)
    d = api.Dimensions(100, F, 64, 12)  # shrink T and B
    with pytest.raises(_capi.InvalidArgument):
        ws.upload_packed_voltage(x, api.Dimensions(100, F, 32, 12))
    with pytest.raises(_capi.InvalidArgument):
        ws.upload_packed_voltage(x, api.Dimensions(200, F, 64, 12))
    assert ws.upload_packed_voltage(x, d) >= 0
    assert ws.upload_weights(w, d) >= 0
    assert ws.run_kernel(d) > 0
    out = np.full(100 * F * 12 + 3, -1.0, np.float32)
    ws.download_intensity(out, d)
    assert np.all(out[-3:] == -1.0)  # the tail is left untouched
    assert_ref_close(out[:-3].reshape(100, F, 12), cpu_reference(x, w, 100, 64, 12, 0))
    out2 = np.zeros((100, F, 12), np.float32)
    t = ws.run_pipeline(x, out2, d)
    assert t["kernel_ms"] > 0 and t["host_to_device_ms"] > 0
    assert np.array_equal(out2, out[:-3].reshape(100, F, 12))
    with pytest.raises(_capi.LogicError):
        ws.run_integrated_pipeline(x, out2, d)
    ws.close()


def test_workspace_integrated_and_quantized():
    cap = api.Dimensions(700, F, 64, 20)
    x, w = random_case(700, 64, 20, seed=9)
    ws = api.BeamformerWorkspace(cap, K.SPLIT_F16, 10, api.Output.INT8)
    assert ws.has_temporal_integration() and ws.has_int8_quantization()
    ws.upload_weights(w, cap)
    W = 70
    integ = np.zeros((W, F, 20), np.float32)
    ws.run_integrated_pipeline(x, integ, cap)
    ref64 = fp64_reference(x, w, 700, 64, 20, 10)
    assert rel_err_floored(integ, ref64, incoherent_floor(x, 64, 10)) <= SPLIT_TOL
    codes = np.zeros((W, F, 20), np.int8)
    params = np.zeros((W * 21 * 2, 2), np.float32)
    ws.run_quantized_integrated_pipeline(x, codes, params, cap)
    rc, rp = O.quantize(integ, W, F, 20)
    assert np.array_equal(codes, rc) and np.array_equal(params, rp)


def test_device_frame_path():
    T, A, B = 320, 64, 8
    x, w = random_case(T, A, B, seed=17)
    shard = api.default_shard_descriptors()[1]
    dx = api.DeviceArray.from_numpy(x)
    dw = api.DeviceArray.from_numpy(w)
    dims = api.Dimensions(T, F, A, B)
    ws = api.BeamformerWorkspace(dims, K.DIRECT, 320, api.Output.INT8)
    vin = api.make_contiguous_frame_view(dx, x.nbytes, api.DataType.PACKED_INT4X2, (T, F, A),
                                         7, shard)
    vw = api.make_contiguous_frame_view(dw, w.nbytes, api.DataType.COMPLEX_FLOAT32,
                                        (B, F, A), 7, shard)
    out = api.DeviceArray.empty((1, F, B), np.int8)
    vout = api.make_contiguous_frame_view(out, out.nbytes, api.DataType.SIGNED_INT8,
                                          (1, F, B), 7, shard)
    par = api.DeviceArray.empty((1, 21, 1, 2), np.float32)
    vpar = api.make_contiguous_frame_view(par, par.nbytes, api.DataType.FLOAT32,
                                          (1, 21, 1, 2), 7, shard)
    t = ws.run_device_frame(vin, vw, vout, vpar)
    assert t["kernel_ms"] > 0 and t["host_to_device_ms"] == 0
    integ = api.beamform_packed_integrated_intensity(x, w, dims, 320, K.DIRECT)
    rc, rp = O.quantize(integ, 1, F, B)
    assert np.array_equal(out.to_numpy(), rc)
    assert np.array_equal(par.to_numpy().reshape(-1, 2), rp)
    # mismatched frame id / shard are rejected before launch
    vout.frame_id = 8
    with pytest.raises(_capi.InvalidArgument, match="frame_id"):
        ws.run_device_frame(vin, vw, vout, vpar)
    vout.frame_id = 7
    vw.shard = api.default_shard_descriptors()[0]
    with pytest.raises(_capi.InvalidArgument, match="shard"):
        ws.run_device_frame(vin, vw, vout, vpar)


# ------------------------------------------------- full-size properties
def test_full_shard_sampled_frequencies():
    """C2 shape (A=64, T=32768, B=128, S=10): the full-size launch checked on
    sampled channels against the FP64 oracle, plus window-sum consistency."""
    T, A, B, S = 32768, 64, 128, 10
    pos = O.regular_array(8, 8, 0.3)
    fr = np.linspace(400e6, 800e6, F).astype(np.float32)
    ls = np.linspace(-0.35, 0.35, 8)
    ms = np.linspace(-0.35, 0.35, 16)
    dirs = np.stack([O.direction_from_lm(l, m) for l in ls for m in ms])
    w = O.generate_weights(pos, fr, dirs)
    x = O.make_uniform_int4(T * F * A, 99).reshape(T, F, A)
    dims = api.Dimensions(T, F, A, B)
    got = run_gpu(K.SPLIT_F16, x, w, dims, S)
    assert got.shape == (3277, F, B)
    sample = [0, 101, 222, 335]
    xs = np.ascontiguousarray(x[:, sample, :])
    ws_ = np.ascontiguousarray(w[:, sample])
    i64 = np.zeros((T, len(sample), B), np.float64)
    O.LIB.or_beamform_packed_f64(xs.reshape(-1), ws_.reshape(-1), T, len(sample), A, B, i64)
    ref = O.integrate_f64(i64, T, len(sample), B, S)
    err = rel_err_floored(got[:, sample], ref, incoherent_floor(x[:, sample], A, S))
    assert err <= SPLIT_TOL, err


@pytest.mark.parametrize("T,A,B,S,kernel", [
    (1600, 64, 128, 10, K.SPLIT_F16),
    (640, 64, 40, 0, K.F16),
    (960, 128, 33, 320, K.SPLIT_F16),
    (256, 256, 9, 16, K.SPLIT_F16),
    (800, 32, 128, 10, K.F16),
])
def test_autotuned_shapes_compute_identical_sums(T, A, B, S, kernel):
    """Every tile shape the autotuner tries computes the same sums in the
    same order, so the tuned launch is bit-identical to the default one."""
    x, w = random_case(T, A, B, seed=T + A + B)
    dims = api.Dimensions(T, F, A, B)
    api.autotune_clear()
    default = run_gpu(kernel, x, w, dims, S)
    dx, dw = api.DeviceArray.from_numpy(x), api.DeviceArray.from_numpy(w)
    rows = T if S == 0 else api.integrated_time_count(T, S)
    out = api.DeviceArray.empty((rows, F, B), np.float32)
    table = api.autotune_packed_beamformer(kernel, None, dx, dw, out, dims, S, repeats=2)
    n_shapes = {32: 5, 64: 5, 128: 3, 256: 1}[A]
    assert len(table) == n_shapes * 2 * 3
    assert all(r["median_ms"] > 0 for r in table)
    assert np.array_equal(out.to_numpy(), default)  # the winner's output
    tuned = run_gpu(kernel, x, w, dims, S)             # a later launch takes the winner
    assert np.array_equal(tuned, default)
    api.autotune_clear()
    with pytest.raises(_capi.InvalidArgument, match="autotuner"):
        api.autotune_packed_beamformer(K.FP32, None, dx, dw, out, dims, S)
