"""Argument handling of the command-line tools, which finishes before any
GPU call: usage text, required options, rejected values, and the tracker
bridge's leniency towards flags it does not know (the reference's
tools/run_tracker_stream.cpp:70-87 passes over them)."""
import os
import subprocess

import pytest

BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   "beamform_project_amd", "bin")


def tool(name, *args):
    exe = os.path.join(BIN, name)
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built")
    return subprocess.run([exe, *map(str, args)], capture_output=True, text=True, timeout=60)


def test_beamformer_usage_lists_every_flag():
    r = tool("beamformer_amd", "--help")
    assert r.returncode == 0
    for flag in ("--input", "--weights", "--output", "--n-time", "--n-ant", "--n-beams",
                 "--kernel", "--integration-spectra", "--output-format",
                 "--quantization-params", "--quantization-metadata", "--shard-id", "--metrics"):
        assert flag in r.stdout, flag


@pytest.mark.parametrize("args,message", [
    (["--input", "x"], "required: --weights, --output"),
    (["--input", "x", "--weights", "w", "--output", "o", "--bogus", "1"], "unknown option: --bogus"),
    (["--input", "x", "--weights", "w", "--output"], "missing value after --output"),
    (["--input", "x", "--weights", "w", "--output", "o", "--n-time", "12x"],
     "invalid integer for --n-time"),
    (["--input", "x", "--weights", "w", "--output", "o", "--kernel", "wmma"],
     "--kernel must be one of direct, tiled, split-f16, f16, auto"),
    (["--input", "x", "--weights", "w", "--output", "o", "--output-format", "int8"],
     "int8 output needs --integration-spectra"),
    (["--input", "x", "--weights", "w", "--output", "o", "--output-format", "int8",
      "--integration-spectra", "10"], "needs --quantization-params and --quantization-metadata"),
    (["--input", "x", "--weights", "w", "--output", "o", "--shard-id", "2"],
     "--shard-id must be 0 or 1"),
    (["--input", "x", "--weights", "w", "--output", "o", "--n-ant", "48"], "n_ant"),
    (["--input", "/nonexistent/x.bin", "--weights", "w", "--output", "o"], "x.bin"),
])
def test_beamformer_rejects_bad_arguments(args, message):
    r = tool("beamformer_amd", *args)
    assert r.returncode == 1
    assert r.stderr.startswith("beamformer_amd: ") and message in r.stderr, r.stderr


def test_tracker_bridge_requires_files_and_skips_unknown_flags():
    r = tool("run_tracker_stream", "--n-time", "10")
    assert r.returncode == 1 and "--input and --output are required" in r.stderr
    assert "Usage:" in r.stderr
    r = tool("run_tracker_stream", "-h")
    assert r.returncode == 0 and "--source-dm" in r.stdout
    # an unknown flag is passed over; the engine is checked before any file
    r = tool("run_tracker_stream", "--verbose", "--engine", "cpu_v2", "--input", "a",
             "--output", "b")
    assert r.returncode == 1 and "reference's CPU tracker" in r.stderr
    r = tool("run_tracker_stream", "--engine", "nonsense", "--input", "a", "--output", "b")
    assert r.returncode == 1 and "Unknown engine: nonsense" in r.stderr
    # the last flag without its value is passed over too
    r = tool("run_tracker_stream", "--input", "a", "--output", "b", "--engine", "amd_fp32",
             "--n-ant")
    assert r.returncode == 1 and "a" in r.stderr  # fails reading the input file
