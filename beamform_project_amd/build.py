"""Builds libbeamform_hip.so in-tree with hipcc for gfx950.

    python -m beamform_project_amd.build

Objects are compiled in parallel into beamform_project_amd/_build/ and
linked into beamform_project_amd/libbeamform_hip.so; the command-line tools
in cli/ are then built into beamform_project_amd/bin/.  Every call compiles
(the driver's build check relies on that); object files are reused only
when their source and the shared headers are older than the object.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libbeamform_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    # Packed FP32 VALU forms are kept out of the generated code.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-I", os.path.join(ROOT, "include"),
]


# Translation units whose float arithmetic must stay separately rounded
# (bit-exact parity with the reference CPU code): no a*b+c contraction.
PER_FILE_FLAGS = {
    "k_quant.hip": ["-ffp-contract=off"],
    "k_tracker.hip": ["-ffp-contract=off"],
    "k_opt_tracker.hip": ["-ffp-contract=off"],
}


def _sources():
    return sorted(
        f for f in os.listdir(CSRC) if f.endswith(".hip") or f.endswith(".cpp")
    )


def _headers():
    hs = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hpp")]
    hs.append(os.path.join(ROOT, "include", "bf_capi.h"))
    return hs


def _compile(src: str, force: bool) -> str:
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, src + ".o")
    newest = max([os.path.getmtime(path), os.path.getmtime(os.path.abspath(__file__))]
                 + [os.path.getmtime(h) for h in _headers()])
    if not force and os.path.exists(obj) and os.path.getmtime(obj) >= newest:
        return obj
    lang = ["-x", "hip"]
    cmd = [HIPCC, *lang, *COMMON_FLAGS, *PER_FILE_FLAGS.get(src, []), "-c", path, "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed on {src}")
    return obj


def build(force: bool = False, jobs: int | None = None) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = _sources()
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, force), srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("link of libbeamform_hip.so failed")
    shutil.move(tmp, LIB)
    build_cli()
    return LIB


CLI = os.path.join(PKG, "cli")
BIN = os.path.join(PKG, "bin")
CXX = os.environ.get("CXX", "g++")


def build_cli() -> list:
    """Compiles the command-line tools in cli/ against the library (host
    compiler; they use the C++ layer in include/beamformer/ only)."""
    os.makedirs(BIN, exist_ok=True)
    out = []
    for src in sorted(f for f in os.listdir(CLI) if f.endswith(".cpp")):
        exe = os.path.join(BIN, src[:-4])
        cmd = [CXX, "-O2", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wextra",
               "-I", os.path.join(ROOT, "include"), os.path.join(CLI, src), "-o", exe + ".tmp",
               "-L", PKG, "-lbeamform_hip", "-Wl,-rpath,$ORIGIN/.."]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError(f"build of {src} failed")
        shutil.move(exe + ".tmp", exe)
        out.append(exe)
    return out


if __name__ == "__main__":
    print(build(force="--force" in sys.argv))
