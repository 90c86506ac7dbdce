"""GPU parity of the dynamic tracker.

Tolerances (the reference's own tracker tests,
tests/cuda/test_cuda_tracker_v2.cpp:23-36 and
tests/cuda/test_cuda_beam_tracker_v5.cpp:34-45): a cell fails when
|d| > abs_tol and |d| > rel_tol * ref, with (rel, abs) = (1e-4, 1e-5) for
the split-FP16 and FP32 paths against the CPU tracker; the multi-beam paths
are also held to the FP64 oracle: floored relative error <= 1e-5 (split),
<= 2e-3 (single FP16).
"""
import numpy as np
import pytest

import oracle as O
from beamform_project_amd import _capi, api, tracker as TR
from helpers import F, golden, incoherent_floor, rel_err_floored, tracker_close

pytestmark = pytest.mark.gpu
K = api.Kernel


def cfg(l0, m0, rl, rm, S):
    return TR.TrackerConfig(TR.direction_from_lm(l0, m0), (rl, rm), S)


@pytest.mark.parametrize("kernel", [K.SPLIT_F16, K.FP32, K.AUTO])
def test_golden_moving_source(kernel):
    c = golden("tracker_moving_a32")
    l0, m0, rl, rm = [float(v) for v in c["trajectory"]]
    x = c["x"]
    T = x.shape[0]
    out = TR.beam_tracker_into(x, api.Dimensions(T, F, 32, 1),
                               cfg(l0, m0, rl, rm, int(c["spectra"])), TR.TrackerExec(kernel))
    tracker_close(out, c["intensity"], 1e-4, 1e-5, what=kernel.name)


def test_reference_tracker_fixture_a64():
    # tests/cuda/test_cuda_tracker_v2.cpp:45-58: A=64, T=3200, source drifting
    # 1e-5 per sample in l from broadside, 320-spectrum windows.
    T, A = 3200, 64
    pos = O.default_positions(A)
    fr = O.channelized_frequencies()
    x = O.make_moving_point_source(T, F, A, pos, fr, 0.0, 0.0, 1e-5, 0.0)
    ref = O.beam_tracker(x, T, F, A, 0.0, 0.0, 1e-5, 0.0, 320, threads=True)
    for kernel in (K.SPLIT_F16, K.FP32):
        dx = api.DeviceArray.from_numpy(x)
        dout = api.DeviceArray.empty((T, F, 1), np.float32)
        TR.beam_tracker_device_resident(dx, dout, api.Dimensions(T, F, A, 1),
                                        cfg(0.0, 0.0, 1e-5, 0.0, 320), TR.TrackerExec(kernel))
        tracker_close(dout.to_numpy(), ref, 1e-4, 1e-5, what=kernel.name)
    # the F16 single path is held to the reference V4 half tolerance (5e-3, 1e-3)
    out16 = TR.beam_tracker_into(x, api.Dimensions(T, F, A, 1), cfg(0.0, 0.0, 1e-5, 0.0, 320),
                                 TR.TrackerExec(K.F16))
    tracker_close(out16, ref, 5e-3, 1e-3, what="F16")


@pytest.mark.parametrize("A", [32, 128, 256])
def test_tracker_antenna_counts_and_partial_window(A):
    """Against the CPU tracker at the reference's tolerances: V2's
    (rel 1e-4, abs 1e-5; tests/cuda/test_cuda_tracker_v2.cpp:23-36) for
    32 and 64 antennas, the sizes V2 covers, and V5's (rel 1e-3, abs 1e-4;
    tests/cuda/test_cuda_beam_tracker_v5.cpp:34-35) for 128 and 256, where
    the FP32 sum over more antennas in another order differs more; and
    against FP64 sums of the same weights at the split-FP16 bar."""
    T, S = 333, 100  # last window holds 33 spectra
    x = O.make_noise(T, F, A, 31 + A)
    ref = O.beam_tracker(x, T, F, A, 0.05, -0.02, 2e-5, 1e-5, S, threads=True)
    rel, ab = (1e-4, 1e-5) if A <= 64 else (1e-3, 1e-4)
    pos, fr = O.default_positions(A), O.channelized_frequencies()
    W = (T + S - 1) // S
    dirs = np.stack([O.tracker_direction(0.05, -0.02, 2e-5, 1e-5, w * S)
                     for w in range(W)])[:, None, :].astype(np.float32)
    _, ref64 = O.multibeam_tracker(x, T, F, A, 1, S, pos, fr, dirs, want_f64=True, spectra=True)
    for kernel in (K.AUTO, K.SPLIT_F16, K.FP32):
        out = TR.beam_tracker_into(x, api.Dimensions(T, F, A, 1),
                                   cfg(0.05, -0.02, 2e-5, 1e-5, S), TR.TrackerExec(kernel))
        tracker_close(out, ref, rel, ab, what=f"A={A} {kernel.name}")
        assert rel_err_floored(out, ref64, incoherent_floor(x, A)) <= 1e-5, (A, kernel.name)


def test_stationary_tracker_equals_fixed_grid():
    # tests/cpu/test_beam_tracker.cpp:120-139, on the GPU: one window covering
    # the whole run is the one-beam fixed grid.  The two are different FP32
    # kernels (the tracker's weights are reduced-phase sincos within about
    # one float ulp, its antenna sums run in another order), so they agree
    # to FP32 accuracy, not bit for bit as the reference's one CPU loop does.
    T, A = 48, 64
    pos, fr = O.default_positions(A), O.channelized_frequencies()
    d = O.direction_from_lm(0.04, 0.0)
    x = O.make_point_source(T, F, A, pos, fr, d, 4.0)
    grid = api.beamform_packed_intensity(x, O.generate_weights(pos, fr, d[None]),
                                         api.Dimensions(T, F, A, 1), K.FP32)
    trk = TR.beam_tracker_into(x, api.Dimensions(T, F, A, 1), cfg(0.04, 0.0, 0.0, 0.0, T),
                               TR.TrackerExec(K.FP32))
    np.testing.assert_allclose(trk, grid, rtol=1e-5, atol=1e-4)


def test_tracker_validation():
    with pytest.raises(_capi.InvalidArgument, match="n_beams"):
        TR.beam_tracker_into(np.zeros(10 * F * 64, np.uint8), api.Dimensions(10, F, 64, 2),
                             cfg(0, 0, 0, 0, 320))
    with pytest.raises(_capi.InvalidArgument, match="l\\*l"):
        TR.beam_tracker_into(np.zeros(700 * F * 64, np.uint8), api.Dimensions(700, F, 64, 1),
                             cfg(0.9, 0.0, 1e-3, 0.0, 320))
    with pytest.raises(_capi.InvalidArgument, match="unit vector"):
        TR.beam_tracker_into(np.zeros(10 * F * 64, np.uint8), api.Dimensions(10, F, 64, 1),
                             TR.TrackerConfig((0.5, 0.5, 0.5), (0, 0), 320))


@pytest.mark.parametrize("graph", [False, True])
def test_batched_stream_honours_first_window_index(graph):
    A, S, nb = 64, 320, 3
    T = S * 5
    x = O.make_noise(T, F, A, 77)
    full = O.beam_tracker(x, T, F, A, 0.01, 0.02, 3e-5, -2e-5, S, threads=True)
    st = TR.BatchedTrackerStream(api.Dimensions(S, F, A, 1), cfg(0.01, 0.02, 3e-5, -2e-5, S),
                                 nb, TR.TrackerExec(K.SPLIT_F16, graph))
    assert st.batch_size() == nb and st.window_bytes() == S * F * A
    for first in (0, 2, 2, 0):  # replays, and a re-capture when the position changes
        out = np.zeros(st.batch_output_floats(), np.float32)
        st.process_batch(first, x[first * S:(first + nb) * S], out)
        tracker_close(out.reshape(nb * S, F, 1), full[first * S:(first + nb) * S], 1e-4, 1e-5,
                      what=f"first={first}")
        assert st.last_kernel_time_ms() > 0
    st.process_batch_kernel_only(1)
    st.close()


def multibeam_case(T, A, B, S, seed):
    rng = np.random.default_rng(seed)
    x = O.make_uniform_int4(T * F * A, seed).reshape(T, F, A)
    pos = O.regular_array(A // 8, 8, 0.3)
    fr = np.linspace(400e6, 800e6, F).astype(np.float32)
    W = (T + S - 1) // S
    l0 = rng.uniform(-0.4, 0.4, B)
    m0 = rng.uniform(-0.4, 0.4, B)
    dl = rng.uniform(-0.01, 0.01, B)
    dirs = np.zeros((W, B, 3), np.float32)
    for w in range(W):
        for b in range(B):
            dirs[w, b] = O.direction_from_lm(l0[b] + w * dl[b], m0[b] - w * dl[b])
    return x, pos, fr, dirs


@pytest.mark.parametrize("T,A,B,S,integrate", [
    (1000, 64, 32, 320, True),
    (333, 64, 5, 10, True),
    (96, 32, 40, 16, False),
    (250, 128, 9, 100, True),
])
def test_multibeam_tracker_vs_fp64(T, A, B, S, integrate):
    x, pos, fr, dirs = multibeam_case(T, A, B, S, seed=T + B)
    ref, ref64 = O.multibeam_tracker(x, T, F, A, B, S, pos, fr, dirs, want_f64=True,
                                     spectra=not integrate)
    floor = incoherent_floor(x, A, S if integrate else 1)
    dx = api.DeviceArray.from_numpy(x)
    dpos, dfr, ddir = (api.DeviceArray.from_numpy(a) for a in (pos, fr, dirs))
    for kernel, tol in ((K.SPLIT_F16, 1e-5), (K.F16, 2e-3), (K.FP32, 1e-5)):
        dout = api.DeviceArray.empty(ref.shape, np.float32)
        TR.multibeam_tracker_launch(None, dx, dout, T, F, A, B, S, dpos, dfr, directions=ddir,
                                    integrate=integrate, kernel=kernel)
        api.device_synchronize()
        got = dout.to_numpy()
        err = rel_err_floored(got, ref64, floor)
        assert err <= tol, f"{kernel.name}: {err:.3e}"
        if kernel != K.F16:
            tracker_close(got, ref, 1e-4, 1e-5, what=kernel.name)


def test_multibeam_trajectories_match_table_and_window_offset():
    T, A, B, S = 640, 64, 3, 64
    x = O.make_uniform_int4(T * F * A, 5).reshape(T, F, A)
    traj = np.array([[0.1, 0.0, 1e-4, 0.0], [-0.2, 0.1, 0.0, -1e-4], [0.0, 0.3, 5e-5, 5e-5]],
                    np.float32)
    off = 3
    W = T // S
    dirs = np.zeros((W, B, 3), np.float32)
    for w in range(W):
        for b in range(B):
            dirs[w, b] = O.tracker_direction(*traj[b], (off + w) * S)
    dx = api.DeviceArray.from_numpy(x)
    dtraj, ddir = api.DeviceArray.from_numpy(traj), api.DeviceArray.from_numpy(dirs)
    a = api.DeviceArray.empty((W, F, B), np.float32)
    b = api.DeviceArray.empty((W, F, B), np.float32)
    TR.multibeam_tracker_launch(None, dx, a, T, F, A, B, S, trajectories=dtraj,
                                window_offset=off, kernel=K.SPLIT_F16)
    TR.multibeam_tracker_launch(None, dx, b, T, F, A, B, S, directions=ddir, kernel=K.SPLIT_F16)
    api.device_synchronize()
    assert np.array_equal(a.to_numpy(), b.to_numpy())
    ref = O.multibeam_tracker(x, T, F, A, B, S, O.default_positions(A),
                              O.channelized_frequencies(), dirs)
    tracker_close(a.to_numpy(), ref, 1e-4, 1e-5, what="trajectory")
