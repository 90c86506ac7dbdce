"""Benchmark of the fixed-grid hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c2|c1]

Workload (default c2, the full-shard fixed grid): A=64 antennas on an 8x8
grid at 0.3 m, F=336 channels over 400-800 MHz, T=32768 spectra, B=128 beams
on an 8x16 (l, m) grid spanning +-0.35, one point source at a seeded
direction with per-element SNR 3 (amplitude^2 / (2 sigma^2)) plus complex
Gaussian noise sigma=2.5, rounded and clipped to int4.  One step = the fused
split-FP16 MFMA beamform + |.|^2 + 10-spectrum integration (f32 output)
followed by the int8 quantiser (codes + parameters), inputs resident in HBM.

Prints one JSON line (see the keys below).  The plain FP32 fused kernel is
timed beside it for the speedup figure.  `cpu_baseline` times the reference
CPU path (oracle/_ref, compiled from the reference's sources; the restated
oracle when _ref is absent) on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

F = 336


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--workload", default="c2", choices=["c2", "c1"])
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-seconds", type=float, default=10.0)
    return p.parse_args()


def workload(name):
    if name == "c1":
        return dict(name="C1 plain fixed grid: A=32 ULA 0.3 m, F=336, T=1024, B=16 over "
                         "+-60 deg, uniform int4, f32 [T][F][B] output",
                    A=32, T=1024, B=16, S=0, kernel="FP32", ula=True)
    return dict(name="C2 full-shard fixed grid: A=64 8x8 0.3 m, F=336, T=32768, B=128 "
                     "8x16 grid, point source SNR 3 + noise sigma 2.5, S=10, f32 + int8",
                A=64, T=32768, B=128, S=10, kernel="SPLIT_F16", ula=False)


def host_geometry(wl):
    """Positions, frequencies and beam directions (host float32)."""
    A, B = wl["A"], wl["B"]
    fr = np.linspace(400e6, 800e6, F).astype(np.float32)
    if wl["ula"]:
        pos = np.zeros((A, 3), np.float32)
        pos[:, 0] = 0.3 * np.arange(A)
        th = np.deg2rad(np.linspace(-60, 60, B))
        dirs = np.stack([np.sin(th), np.zeros(B), np.cos(th)], 1).astype(np.float32)
    else:
        c, r = np.meshgrid(np.arange(8), np.arange(8))
        pos = np.stack([0.3 * c.ravel(), 0.3 * r.ravel(), np.zeros(64)], 1).astype(np.float32)
        ls = np.linspace(-0.35, 0.35, 8)
        ms = np.linspace(-0.35, 0.35, 16)
        dirs = np.array([[l, m, np.sqrt(1 - l * l - m * m)] for l in ls for m in ms],
                        np.float32)
    return pos, fr, dirs


def weights_from(pos, fr, dirs):
    """w[b][f][a] = exp(+j 2 pi f/c p.d): phase in float64, narrowed once
    (the weight formula of the reference's generate_weights)."""
    k = 2 * np.pi * fr.astype(np.float64) / 299792458.0
    delay = dirs.astype(np.float64) @ pos.astype(np.float64).T  # [B][A]
    ph = k[None, :, None] * delay[:, None, :]
    w = np.empty(ph.shape + (2,), np.float32)
    w[..., 0] = np.cos(ph)
    w[..., 1] = np.sin(ph)
    return w


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the path "
              "shards by frequency with no exchange, see DESIGN.md", file=sys.stderr)
        sys.exit(2)
    from beamform_project_amd import api

    wl = workload(args.workload)
    A, T, B, S = wl["A"], wl["T"], wl["B"], wl["S"]
    dims = api.Dimensions(T, F, A, B)
    pos, fr, dirs = host_geometry(wl)
    w = weights_from(pos, fr, dirs)
    stream = api.Stream()
    x = api.DeviceArray.empty((T, F, A), np.uint8)
    if wl["ula"]:
        api.generate_uniform_int4(stream, x, x.nbytes, 1234)
    else:
        rng = np.random.default_rng(2024)
        l, m = rng.uniform(-0.3, 0.3, 2)
        src = np.array([l, m, np.sqrt(1 - l * l - m * m)], np.float32)
        sigma = 2.5
        amp = float(np.sqrt(3.0 * 2 * sigma * sigma))
        dpos = api.DeviceArray.from_numpy(pos)
        dfr = api.DeviceArray.from_numpy(fr)
        api.generate_point_source_noise(stream, x, T, F, A, dpos, dfr, src, amp, sigma, 77)
    dw = api.DeviceArray.from_numpy(w)
    kernel = api.Kernel[wl["kernel"]]
    if S:
        W = api.integrated_time_count(T, S)
        out = api.DeviceArray.empty((W, F, B), np.float32)
        npar = api.quantization_parameter_count(api.Dimensions(W, F, A, B))
        codes = api.DeviceArray.empty((W, F, B), np.int8)
        params = api.DeviceArray.empty((npar, 2), np.float32)
        idims = api.Dimensions(W, F, A, B)
    else:
        out = api.DeviceArray.empty((T, F, B), np.float32)
    stream.synchronize()

    ev_k0, ev_k1, ev_end = api.Event(), api.Event(), api.Event()

    def step(k, quantize=True):
        ev_k0.record(stream)
        if S:
            api.launch_packed_integrated_beamformer(k, stream, x, dw, out, dims, S)
        else:
            api.launch_packed_beamformer(k, stream, x, dw, out, dims)
        ev_k1.record(stream)
        if S and quantize:
            api.launch_quantize_integrated_intensity(stream, out, codes, params, idims)
        ev_end.record(stream)

    def timed(k, steps, warmup, quantize=True):
        for _ in range(warmup):
            step(k, quantize)
        stream.synchronize()
        api.device_synchronize()
        kern_ms = []
        t0 = time.perf_counter()
        for _ in range(steps):
            step(k, quantize)
            kern_ms.append(ev_k0.elapsed_ms(ev_k1))  # waits for this step's kernel
        stream.synchronize()
        api.device_synchronize()
        wall = time.perf_counter() - t0
        return wall * 1e3 / steps, float(np.mean(kern_ms))

    ms_step, kern_ms = timed(kernel, args.steps, args.warmup)
    beam_samples = T * F * B
    value = beam_samples / (ms_step * 1e-3)

    # Plain FP32 fused kernel on the same inputs (same outputs, no int8 leg).
    fp32_steps = max(3, args.steps // 4)
    fp32_ms, fp32_kern = timed(api.Kernel.FP32, fp32_steps, 1, quantize=False)
    split_ms_only, split_kern = timed(kernel, fp32_steps, 1, quantize=False)

    # Algorithmic counts: 8 flops per complex MAC + 3 per
    # output power + 1 per integrated sample.
    cmac = T * F * B * A
    flops = 8 * cmac + 3 * T * F * B + (T * F * B if S else 0)
    out_bytes = (W * F * B * 4) if S else T * F * B * 4
    bytes_alg = T * F * A + B * F * A * 8 + out_bytes
    is_mfma = kernel in (api.Kernel.SPLIT_F16, api.Kernel.F16)
    if is_mfma:
        achieved = flops / (kern_ms * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": round(achieved, 2), "peak": 2500.0,
                "unit": "TFLOP/s", "frac": round(achieved / 2500.0, 4)}
    else:
        achieved = flops / (kern_ms * 1e-3) / 1e12
        roof = {"bound": "valu-fp32", "achieved": round(achieved, 2), "peak": 157.3,
                "unit": "TFLOP/s", "frac": round(achieved / 157.3, 4)}
    roof["traffic"] = read_traffic(wl)
    roof["kernel_avg_ms"] = round(kern_ms, 4)
    roof["algorithmic_bytes"] = bytes_alg

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(wl, pos, fr, dirs, args.cpu_seconds)

    line = {
        "metric": "beam-samples/s (T·F·B)",
        "value": round(value, 1),
        "unit": "beam-samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f16 split hi+lo, f32 accumulate" if is_mfma else "f32",
        "data": "synthetic (device-generated, seeded)",
        "config": {"workload": wl["name"], "n_ant": A, "n_freq": F, "n_time": T,
                   "n_beams": B, "integration_spectra": S, "kernel": wl["kernel"],
                   "outputs": "f32 integrated + int8 codes/params" if S else "f32"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "fp32_fused_kernel": {"ms_per_step": round(fp32_ms, 4),
                              "kernel_avg_ms": round(fp32_kern, 4),
                              "beam_samples_per_s": round(beam_samples / (fp32_ms * 1e-3), 1),
                              "split_f16_kernel_avg_ms": round(split_kern, 4),
                              "split_over_fp32_kernel": round(fp32_kern / split_kern, 2)},
    }
    print(json.dumps(line), flush=True)


def read_traffic(wl):
    """HBM bytes per launch from the committed rocprofv3 PMC summary, or None."""
    name = "traffic_" + ("c2_split_f16" if not wl["ula"] else "c1_fp32") + ".txt"
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    vals = {}
    for ln in open(path):
        if ":" in ln and not ln.startswith("#"):
            k, v = ln.split(":", 1)
            try:
                vals[k.strip()] = float(v.split()[0])
            except (ValueError, IndexError):
                pass
    return vals.get("hbm_bytes_per_launch")


def cpu_baseline(wl, pos, fr, dirs, seconds):
    """The reference's single-threaded CPU fixed-grid path on a bounded
    prefix of the workload (T_sample spectra, all channels and beams),
    scaled to beam-samples/s."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # the checker/baseline library, never the product path

    A, B, S = wl["A"], wl["B"], wl["S"]
    w = weights_from(pos, fr, dirs)
    R = O.ref_lib()
    kind = "reference" if R is not None else "port"

    def run(T):
        x = O.make_uniform_int4(T * F * A, 5).reshape(T, F, A)
        out = np.zeros((T, F, B), np.float32)
        t0 = time.perf_counter()
        if R is not None:
            R.ref_beamform_packed(x.reshape(-1), w.reshape(-1), T, F, A, B, out)
            if S:
                Wn = (T + S - 1) // S
                integ = np.zeros((Wn, F, B), np.float32)
                R.ref_integrate(out.reshape(-1), T, F, A, B, S, integ)
                codes = np.zeros((Wn, F, B), np.int8)
                par = np.zeros((Wn * 21 * ((B + 15) // 16), 2), np.float32)
                R.ref_quantize(integ.reshape(-1), Wn, F, A, B, codes, par)
        else:
            o = O.beamform(x, w, T, F, A, B)
            if S:
                Wn = (T + S - 1) // S
                O.quantize(O.integrate(o, T, F, B, S), Wn, F, B)
        return time.perf_counter() - t0

    t_cal = 10 if S else 8
    dt = run(t_cal)
    T = max(t_cal, int(t_cal * seconds / max(dt, 1e-6)))
    if S:
        T = max(S, (T // S) * S)
    times = [run(T) for _ in range(3)]
    med = float(np.median(times))
    return {"value": round(T * F * B / med, 1), "unit": "beam-samples/s", "cores": 1,
            "kind": kind,
            "sample": f"T={T} spectra x F={F} x B={B} (A={A}) through "
                      f"cpu_beamform_packed_intensity_into"
                      + (" + cpu_integrate_intensity + cpu_quantize_integrated_intensity"
                         if S else "")
                      + f"; median of 3 runs, {med:.2f} s each; single thread as in the "
                        "reference"}


if __name__ == "__main__":
    main()
