"""GPU tests of the reference-named host API: the offline frame runner
(CudaOfflineFrameRunner, include/beamformer/cuda_offline_runner.hpp:17-52)
through the C-ABI against the CPU oracle, and the header-only C++ layer
include/beamformer/reference_api.hpp (workspace, one-shot API, offline
runner, the V2/FWS/V3/V4/V5 tracker entry points and the batched V5 stream)
built and run as a C++ program."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle as O
from beamform_project_amd import _capi, api
from helpers import F, assert_ref_close, incoherent_floor, rel_err_floored

pytestmark = pytest.mark.gpu
K = api.Kernel
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def case(T, A, B, seed):
    rng = np.random.default_rng(seed)
    x = O.make_uniform_int4(T * F * A, seed).reshape(T, F, A)
    pos = O.regular_array(A // 8, 8, 0.3)
    fr = np.linspace(400e6, 800e6, F).astype(np.float32)
    dirs = np.stack([O.direction_from_lm(l, m)
                     for l, m in zip(rng.uniform(-0.4, 0.4, B), rng.uniform(-0.4, 0.4, B))])
    return x, O.generate_weights(pos, fr, dirs)


def test_offline_frame_runner_float_and_int8():
    cap = api.Dimensions(700, F, 64, 24)
    shard = api.default_shard_descriptors()[1]
    x, w = case(650, 64, 17, 5)
    d = api.Dimensions(650, F, 64, 17)
    r = api.OfflineFrameRunner(cap, K.SPLIT_F16, 10)
    res = r.run(x, w, d, 9, shard)
    assert (res["output_dims"].n_time, res["output_dims"].n_beams) == (65, 17)
    ref64 = O.integrate_f64(O.beamform_f64(x, w, 650, F, 64, 17), 650, F, 17, 10)
    assert rel_err_floored(res["float32_output"], ref64, incoherent_floor(x, 64, 10)) <= 1e-5
    assert res["timings"]["kernel_ms"] > 0 and res["timings"]["host_to_device_ms"] > 0
    rq = api.OfflineFrameRunner(cap, K.SPLIT_F16, 10, api.Output.INT8)
    q = rq.run(x, w, d, 9, shard)
    rc, rp = O.quantize(res["float32_output"], 65, F, 17)
    assert np.array_equal(q["codes"], rc) and np.array_equal(q["parameters"], rp)
    # the reference's Direct kernel without integration: per-spectrum FP32
    rd = api.OfflineFrameRunner(cap, K.DIRECT)
    res_d = rd.run(x[:100], w, api.Dimensions(100, F, 64, 17), 1, shard)
    assert_ref_close(res_d["float32_output"], O.beamform(x[:100], w, 100, F, 64, 17))
    with pytest.raises(_capi.InvalidArgument, match="capacity"):
        r.run(x, w, api.Dimensions(650, F, 64, 30), 1, shard)
    with pytest.raises(_capi.InvalidArgument):
        api.OfflineFrameRunner(cap, K.DIRECT, None, api.Output.INT8)
    for h in (r, rq, rd):
        h.close()


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_reference_named_cpp_api(tmp_path):
    exe = tmp_path / "test_reference_api"
    libdir = os.path.dirname(_capi.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror",
                    "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "test_reference_api.cpp"),
                    "-L", libdir, "-lbeamform_hip", f"-Wl,-rpath,{libdir}", "-o", str(exe)],
                   check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=110)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "OK" in r.stdout
