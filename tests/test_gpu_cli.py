"""The command-line tools on the GPU: run_tracker_stream (the reference's
tools/run_tracker_stream.cpp, as tools/astronomical_validation/runner.py
calls it) and beamformer_amd (tools/beamformer_cuda.cpp).  Their files must
equal the library API's results bit for bit, and the tracker's the CPU
oracle within the reference tracker tolerance (1e-4 rel, 1e-5 abs).
"""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from beamform_project_amd import api, formats as FM, tracker as TR
from helpers import F, golden, tracker_close

pytestmark = pytest.mark.gpu
BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   "beamform_project_amd", "bin")


def run(tool, *args):
    exe = os.path.join(BIN, tool)
    assert os.path.exists(exe), f"{exe} missing: build with python -m beamform_project_amd.build"
    res = subprocess.run([exe, *map(str, args)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    return res.stdout


@pytest.mark.parametrize("engine", ["amd_split", "cuda_v5", "amd_fp32"])
def test_run_tracker_stream_as_the_validation_runner_calls_it(tmp_path, engine):
    c = golden("tracker_moving_a32")
    x = c["x"]
    T, A = x.shape[0], x.shape[2]
    l0, m0, rl, rm = [float(v) for v in c["trajectory"]]
    S = int(c["spectra"])
    inp, out = tmp_path / "packed_input.bin", tmp_path / "intensity_output.bin"
    inp.write_bytes(x.tobytes())
    # the argument list of runner.run_beam_tracker (runner.py:66-79)
    msg = run("run_tracker_stream", "--engine", engine, "--n-time", T, "--n-freq", F,
              "--n-ant", A, "--spectra", S, "--source-l0", l0, "--source-m0", m0,
              "--source-dl", rl, "--source-dm", rm, "--input", inp, "--output", out)
    assert f"({T * F * 4} bytes)" in msg
    got = np.fromfile(out, np.float32).reshape(T, F, 1)
    kernel = api.Kernel.FP32 if engine == "amd_fp32" else api.Kernel.SPLIT_F16
    lib = TR.beam_tracker_into(x, api.Dimensions(T, F, A, 1),
                               TR.TrackerConfig(TR.direction_from_lm(l0, m0), (rl, rm), S),
                               TR.TrackerExec(kernel))
    assert np.array_equal(got, lib)
    tracker_close(got, c["intensity"], 1e-4, 1e-5, what=engine)


def test_beamformer_cli_float32_and_int8(tmp_path):
    T, A, B, S = 40, 64, 7, 10
    pos = O.default_positions(A)
    fr = O.channelized_frequencies()
    w = O.generate_weights(pos, fr, O.default_beam_grid(B))  # [B][F][A][2] f32
    x = O.make_noise(T, F, A, 77)
    dims = api.Dimensions(T, F, A, B)
    FM.write_packed_voltage(tmp_path / "x.bin", x, dims)
    FM.write_weights(tmp_path / "w.bin", w.view(np.complex64).reshape(B, F, A), dims)
    wflat = np.ascontiguousarray(w, np.float32).reshape(-1)
    # float32 per spectrum, split-FP16 MFMA kernel
    run("beamformer_amd", "--input", tmp_path / "x.bin", "--weights", tmp_path / "w.bin",
        "--output", tmp_path / "i.bin", "--n-time", T, "--n-ant", A, "--n-beams", B,
        "--kernel", "split-f16", "--metrics", tmp_path / "m.csv")
    got = FM.read_intensities(tmp_path / "i.bin", dims)
    lib = api.beamform_packed_intensity(x, wflat, dims, api.Kernel.SPLIT_F16)
    assert np.array_equal(got, lib.reshape(got.shape))
    rows = (tmp_path / "m.csv").read_text().splitlines()
    assert rows[0].startswith("backend,n_time,n_freq,n_ant,n_beams") and rows[1].startswith("hip,40,")
    # int8 with 10-spectrum integration (the MFMA kernels; the reference's
    # Direct kernel integrates 320 only): codes, parameters and metadata files
    run("beamformer_amd", "--input", tmp_path / "x.bin", "--weights", tmp_path / "w.bin",
        "--output", tmp_path / "c.bin", "--n-time", T, "--n-ant", A, "--n-beams", B,
        "--kernel", "split-f16", "--integration-spectra", S, "--output-format", "int8",
        "--quantization-params", tmp_path / "q.bin", "--quantization-metadata",
        tmp_path / "q.txt", "--shard-id", 1)
    meta = FM.read_quantized_output_metadata(tmp_path / "q.txt")
    assert meta == FM.quantized_output_metadata(dims, S, 1, "q.bin")
    idims = meta.output_dims
    codes = FM.read_quantized_intensities(tmp_path / "c.bin", idims)
    params = FM.read_quantization_parameters(tmp_path / "q.bin", idims)
    lc, lp = api.beamform_packed_quantized_integrated_intensity(x, wflat, dims, S,
                                                                api.Kernel.SPLIT_F16)
    assert np.array_equal(codes.reshape(-1), np.asarray(lc).reshape(-1))
    assert np.array_equal(params.reshape(-1), np.asarray(lp).reshape(-1))
