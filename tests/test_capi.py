"""CPU-side checks of the C-ABI boundary: the library loads, exports every
symbol include/*.h declares, and rejects contract violations with the
reference's error classes before touching the GPU."""
import ctypes as C
import os
import re

import pytest

from beamform_project_amd import (_capi, api, doa, formats, geometry,  # noqa: F401
                                  opt_tracker, serve, tracker)  # (register the ABI)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(inc, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(bf_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_every_declared_symbol_is_exported():
    names = declared_symbols()
    assert len(names) > 40
    lib = C.CDLL(_capi.LIB_PATH)
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    assert declared_symbols() <= set(_capi.SIGNATURES)


def test_dimension_contract():
    # include/beamformer/config.hpp:81-94
    api.validate_dimensions(api.Dimensions(1, 336, 32, 1))
    for bad in (api.Dimensions(0, 336, 64, 1), api.Dimensions(1, 672, 64, 1),
                api.Dimensions(1, 336, 48, 1), api.Dimensions(1, 336, 64, 0),
                api.Dimensions(1, 336, 64, 129)):
        with pytest.raises(_capi.InvalidArgument):
            api.validate_dimensions(bad)


def test_integration_rules_of_reference_selectors():
    d = api.Dimensions(640, 336, 64, 4)
    # src/cuda_beamformer.cu:418-433: Tiled 10/320, Direct 320 only.
    with pytest.raises(_capi.InvalidArgument, match="Direct"):
        api.launch_packed_integrated_beamformer(api.Kernel.DIRECT, None, 256, 256, 256, d, 10)
    with pytest.raises(_capi.InvalidArgument, match="10 or 320"):
        api.launch_packed_integrated_beamformer(api.Kernel.TILED, None, 256, 256, 256, d, 20)
    with pytest.raises(_capi.InvalidArgument):
        api.launch_packed_integrated_beamformer(api.Kernel.SPLIT_F16, None, 256, 256, 256, d, 7)
    with pytest.raises(_capi.InvalidArgument):
        api.launch_packed_integrated_beamformer(api.Kernel.FP32, None, 256, 256, 256, d, 0)
    with pytest.raises(_capi.InvalidArgument, match="null"):
        api.launch_packed_beamformer(api.Kernel.FP32, None, None, 256, 256, d)
    with pytest.raises(_capi.InvalidArgument, match="aligned"):
        api.launch_packed_beamformer(api.Kernel.FP32, None, 257, 256, 256, d)
    with pytest.raises(_capi.InvalidArgument, match="selector"):
        api.launch_packed_beamformer(99, None, 256, 256, 256, d)


def test_counts_and_kernel_resolution():
    assert api.integrated_time_count(32768, 10) == 3277
    assert api.integrated_time_count(65536, 320) == 205
    with pytest.raises(_capi.InvalidArgument):
        api.integrated_time_count(10, 0)
    assert api.quantization_parameter_count(api.Dimensions(2, 336, 32, 17)) == 84
    d = api.Dimensions(32768, 336, 64, 128)
    assert api.resolve_kernel(api.Kernel.AUTO, d, 10) == api.Kernel.SPLIT_F16
    assert api.resolve_kernel(api.Kernel.AUTO, d, 7) == api.Kernel.FP32
    assert api.resolve_kernel(api.Kernel.TILED, d, 10) == api.Kernel.TILED


def test_workspace_rejects_int8_without_integration():
    with pytest.raises(_capi.InvalidArgument, match="int8"):
        api.BeamformerWorkspace(api.Dimensions(10, 336, 64, 4), api.Kernel.TILED, None,
                                api.Output.INT8)


def test_error_message_is_thread_local_and_cleared():
    with pytest.raises(_capi.InvalidArgument):
        api.validate_dimensions(api.Dimensions(0, 336, 64, 1))
    assert b"n_time" in _capi.LIB.bf_last_error_message()
    api.validate_dimensions(api.Dimensions(1, 336, 64, 1))
    assert _capi.LIB.bf_last_error_message() == b""


def test_server_config_is_validated_before_any_device_call():
    base = dict(n_freq=336, n_ant=64, ring_spectra=1 << 12, max_window=320,
                max_step_spectra=4096, hbm_budget_bytes=1 << 20, max_requests=8)
    for bad, msg in ((dict(n_ant=128), "antennas"), (dict(ring_spectra=1000), "power of two"),
                     (dict(max_window=4096), "max_window"), (dict(max_requests=300), "max_requests"),
                     (dict(kernel=api.Kernel.FP32), "MFMA")):
        cfg = serve.ServerConfig(**{**base, **bad})
        with pytest.raises(_capi.InvalidArgument, match=msg):
            serve.TrackerServer(cfg)


def test_default_mfma_shapes_are_instantiated():
    # Every (beams per wave, waves) the dispatcher may pick by default must
    # exist for its antenna count: checked by the tuner's candidate lists in
    # k_mfma.hip against the instantiation macros in mfma_fused.hpp.
    src = open(os.path.join(ROOT, "beamform_project_amd", "csrc", "mfma_fused.hpp")).read()
    tuner = open(os.path.join(ROOT, "beamform_project_amd", "csrc", "k_mfma.hip")).read()
    inst = {}
    for name, body in re.findall(r"#define (BF_MFMA_SHAPES_\w+)(?:\(A_\))?\s*((?:.*\\\n)*.*)",
                                 src):
        inst[name] = set(re.findall(r"BF_SHAPE\(\w+, (\d), (\d)\)", body))
    small = inst["BF_MFMA_SHAPES_32_64"]
    assert {("1", "4"), ("2", "4"), ("2", "8"), ("4", "4")} <= small
    assert ("2", "4") in inst["BF_MFMA_SHAPES_128"] and ("1", "4") in inst["BF_MFMA_SHAPES_128"]
    assert ("1", "4") in inst["BF_MFMA_SHAPES_256"]
    lists = re.findall(r"return \{(\{\d, \d\}(?:, \{\d, \d\})*)\};", tuner)
    tuned = [set(re.findall(r"\{(\d), (\d)\}", l)) for l in lists]
    assert tuned[0] == small and tuned[1] == inst["BF_MFMA_SHAPES_128"]
    assert tuned[2] == inst["BF_MFMA_SHAPES_256"]


def test_opt_tracker_rejects_bad_configs():
    """CpuOptBeamTracker constructor checks (src/cpu_opt_beam_tracker.cpp:1084-1113,
    tests/cpu/test_cpu_opt_beam_tracker.cpp:163-215), raised before any GPU work."""
    from beamform_project_amd.opt_tracker import OptBeamTracker, OptTrackerConfig, default_config
    d = api.Dimensions(4, 336, 32, 1)
    cases = [
        (dict(forgetting_factor=0.0), "forgetting_factor"),
        (dict(forgetting_factor=1.5), "forgetting_factor"),
        (dict(spatial_smoothing_subarray_size=100), "spatial_smoothing_subarray_size"),
        (dict(integration_spectra=0), "integration_spectra"),
        (dict(coarse_grid_resolution=65), "coarse_grid_resolution"),
    ]
    for kw, msg in cases:
        with pytest.raises(_capi.InvalidArgument, match=msg):
            OptBeamTracker(None, None, d, OptTrackerConfig(**kw))
    with pytest.raises(_capi.InvalidArgument, match="n_beams == 1"):
        OptBeamTracker(None, None, api.Dimensions(4, 336, 32, 2), OptTrackerConfig())
    with pytest.raises(_capi.InvalidArgument, match="32 or 64"):
        OptBeamTracker(None, None, api.Dimensions(4, 336, 128, 1), OptTrackerConfig())
    c = default_config()
    assert (c.coarse_grid_resolution, c.refinement_levels, c.forgetting_factor) == (1, 0, 1.0)
    assert abs(c.diagonal_load - 1e-3) < 1e-9 and c.enable_quadratic_peak_interp
