"""MI355X (gfx950) build of the CHARTS voltage beamformer / dynamic tracker
hot path.  The product is libbeamform_hip.so (C-ABI in include/bf_capi.h);
this package holds its HIP sources (csrc/), the build script and a ctypes
mirror of the reference host API used by tests and the benchmark."""

__all__ = ["api", "build"]
