"""Benchmark of the hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c2|c1|c3|c4|c5|b1|d2]

Default (c2, the full-shard fixed grid, BASELINE's headline): A=64 antennas on
an 8x8 grid at 0.3 m, F=336 channels over 400-800 MHz, T=32768 spectra,
B=128 beams on an 8x16 (l, m) grid spanning +-0.35, one point source at a
seeded direction with per-element SNR 3 (amplitude^2 / (2 sigma^2)) plus
complex Gaussian noise sigma=2.5, rounded and clipped to int4.  One step =
the fused split-FP16 MFMA beamform + |.|^2 + 10-spectrum integration (f32
output) followed by the int8 quantiser (codes + parameters), inputs
resident in HBM.

The default line also carries, under `legs`, the other throughput
workloads, each with its own value, roofline and CPU baseline:
  c1  plain fixed grid (A=32 ULA, T=1024, B=16, f32 per-spectrum output),
  c3  dynamic tracker over the full band (F=672, T=65536, S=320, B=32),
  c5  adaptive DOA update (K=1024 snapshots x 336 channels, M=48, 721 angles),
  b1  the reference's published single-beam tracker case (A=64, T=15360,
      S=320, per-spectrum output), the one figure the reference publishes,
and under `serving` the streaming server under Poisson request arrivals at
the stream's real-time rate (C4).  `--workload X` prints one of them alone.

CPU baselines time the reference's own CPU code (oracle/_ref, compiled from
the reference's sources; the AVX-512 build where the host has it) on a
bounded sample of the same workload: one warm-up run, then the median of
five, with the host CPU model and core count in the record.
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
    p.add_argument("--workload", default="c2",
                   choices=["c2", "c1", "c3", "c4", "c5", "b1", "d2"])
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-seconds", type=float, default=10.0)
    p.add_argument("--no-serving", action="store_true")
    p.add_argument("--no-legs", action="store_true",
                   help="default line without the c1/c3/c5 legs")
    p.add_argument("--serve-seconds", type=float, default=6.0)
    p.add_argument("--calibrate-copy", action="store_true",
                   help="only run three 1 GiB device-to-device copies (known bytes for "
                        "scaling rocprofv3 FETCH_SIZE/WRITE_SIZE counters) and exit")
    return p.parse_args()


def workload(name):
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

    if args.calibrate_copy:
        n = 1 << 30
        a, b = api.DeviceArray(n), api.DeviceArray(n)
        a.fill(1)
        api.device_synchronize()
        for _ in range(3):
            api._capi.call("bf_memcpy_dtod", b.ptr, a.ptr, n)
        api.device_synchronize()
        print(json.dumps({"calibration": "3 x 1 GiB device-to-device copies"}), flush=True)
        return
    if args.workload in ("c1", "c3", "b1", "c4", "c5", "d2"):
        line = {"c1": bench_c1, "c3": bench_c3, "b1": bench_b1, "c4": bench_c4,
                "c5": bench_c5, "d2": bench_d2}[args.workload](args)
        line["host"] = host_info()
        print(json.dumps(line), flush=True)
        return
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

    # Tile shape for this exact problem from the library's autotuner
    # (untimed; it leaves the winner cached for the launches below).
    tile_shape = None
    if kernel in (api.Kernel.SPLIT_F16, api.Kernel.F16):
        tab = api.autotune_packed_beamformer(kernel, stream, x, dw, out, dims, S, repeats=5)
        best = min(tab, key=lambda r: r["median_ms"])
        tile_shape = {k: best[k] for k in ("beams_per_wave", "waves_per_workgroup",
                                           "double_buffered", "resident_workgroups")}

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
        kern_ms, quant_ms = [], []
        t0 = time.perf_counter()
        for _ in range(steps):
            step(k, quantize)
            kern_ms.append(ev_k0.elapsed_ms(ev_k1))  # waits for this step's kernel
            if S and quantize:
                quant_ms.append(ev_k1.elapsed_ms(ev_end))
        stream.synchronize()
        api.device_synchronize()
        wall = time.perf_counter() - t0
        timed.quant_ms = float(np.mean(quant_ms)) if quant_ms else None
        return wall * 1e3 / steps, float(np.mean(kern_ms))

    ms_step, kern_ms = timed(kernel, args.steps, args.warmup)
    quant_ms = timed.quant_ms
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

    err = None
    if S:
        err = fp64_error(x, out, w, A, B, S)

    quant = None
    if S and quant_ms:
        qbytes = W * F * B * 4 + W * F * B + npar * 8
        qa = qbytes / (quant_ms * 1e-3) / 1e9
        quant = {"kernel_avg_ms": round(quant_ms, 4), "bytes_per_launch": qbytes,
                 "achieved_GBps": round(qa, 1), "frac_of_hbm": round(qa / 8000.0, 4)}

    serving = sweep = None
    if not args.no_serving:
        serving = serving_run(args.serve_seconds)
        sweep = serving_sweep(3.0) + serving_sweep(3.0, (12.0, 24.0), max_active_beams=320)

    cpu = cpu_all = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(wl, pos, fr, dirs, args.cpu_seconds)
        cpu_all = cpu_baseline_all_cores(wl, pos, fr, dirs, args.cpu_seconds)

    legs = {}
    if not args.no_legs:
        for name, fn in (("c1", bench_c1), ("c3", bench_c3), ("c5", bench_c5), ("b1", bench_b1)):
            legs[name] = fn(args)

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
                   "outputs": "f32 integrated + int8 codes/params" if S else "f32",
                   "tile_shape": tile_shape},
        "max_rel_err_vs_fp64": err,
        "p99_window_latency_ms": serving["latency_p99_ms"] if serving else None,
        "roofline": roof,
        "cpu_baseline": cpu,
        "cpu_baseline_all_cores": cpu_all,
        "quantizer": quant,
        "serving": serving,
        "serving_sweep": sweep,
        "fp32_fused_kernel": {"ms_per_step": round(fp32_ms, 4),
                              "kernel_avg_ms": round(fp32_kern, 4),
                              "beam_samples_per_s": round(beam_samples / (fp32_ms * 1e-3), 1),
                              "split_f16_kernel_avg_ms": round(split_kern, 4),
                              "split_over_fp32_kernel": round(fp32_kern / split_kern, 2)},
        "legs": legs,
        "host": host_info(),
    }
    print(json.dumps(line), flush=True)


def common_line(args, metric, value, unit, ms, dtype, config, **extra):
    line = {"metric": metric, "value": round(value, 1), "unit": unit, "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
            "data": "synthetic (device-generated, seeded)", "config": config}
    line.update(extra)
    return line


def host_info():
    """The host the CPU baselines ran on."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # the checker/baseline library, never the product path
    h = O.host_cpu()
    h["omp_num_threads"] = int(os.environ.get("OMP_NUM_THREADS", h["nproc"]))
    h["avx512"] = O.host_has_avx512()
    return h


def median_of_five(run, size0, step, seconds, hi=None):
    """CPU timing protocol of the reference's benchmarks
    (benchmarks/benchmark_beam_tracker_opt_v2.cpp:16-21,59-60): the sample
    size is calibrated so one run takes about seconds/6, then one warm-up
    run and the median of five timed runs.  Returns (size, median_s)."""
    dt = run(size0)
    size = max(size0, int(size0 * (seconds / 6.0) / max(dt, 1e-6)) // step * step)
    if hi is not None:
        size = min(size, hi)
    run(size)
    return size, float(np.median([run(size) for _ in range(5)]))


def timed_launches(stream, launch, steps, warmup):
    """Average wall ms per launch, and the average HIP-event ms per launch:
    the events bracket `steps` back-to-back launches on `stream`, so a short
    kernel's figure is its own duration, not the host's launch gap."""
    from beamform_project_amd import api
    e0, e1 = api.Event(), api.Event()
    for _ in range(warmup):
        launch()
    stream.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(steps):
        launch()
    e1.record(stream)
    stream.synchronize()
    wall = time.perf_counter() - t0
    return wall * 1e3 / steps, e0.elapsed_ms(e1) / steps


def tracker_fp64_error(x_dev, out, pos, fr, dirs, T0, F_, A, B, S):
    """Floored max relative error of the first T0/S integrated windows of
    the tracker against complex128 sums with weights (float)cos/sin of the
    FP64 phase of each window's direction."""
    from beamform_project_amd import _capi
    W0 = T0 // S
    xb = np.empty((T0, F_, A), np.uint8)
    _capi.call("bf_memcpy_dtoh", xb.ctypes.data, x_dev.ptr, xb.nbytes)
    got = out.to_numpy()[:W0]
    re = ((xb & 0xF).astype(np.int16) ^ 8) - 8
    im = ((xb >> 4).astype(np.int16) ^ 8) - 8
    xc = (re.astype(np.float64) + 1j * im.astype(np.float64)).transpose(1, 0, 2)  # [F][T][A]
    k = 2 * np.pi * fr.astype(np.float64) / 299792458.0
    ref = np.zeros((W0, F_, B))
    for w in range(W0):
        delay = dirs[w].astype(np.float64) @ pos.astype(np.float64).T  # [B][A]
        ph = k[:, None, None] * delay[None]  # [F][B][A]
        wc = np.cos(ph).astype(np.float32).astype(np.float64) + \
            1j * np.sin(ph).astype(np.float32).astype(np.float64)
        v = np.matmul(xc[:, w * S:(w + 1) * S], wc.transpose(0, 2, 1))  # [F][S][B]
        ref[w] = (v.real ** 2 + v.imag ** 2).sum(1)
    floor = A * float(np.mean(re.astype(np.float64) ** 2 + im.astype(np.float64) ** 2)) * S
    return float(np.max(np.abs(got - ref) / np.maximum(ref, floor)))


def bench_c3(args):
    """Dynamic multi-source tracker over the full band on one GPU."""
    from beamform_project_amd import api, tracker as TR
    A, F_, T, S, B = 64, 672, 65536, 320, 32
    W = (T + S - 1) // S
    c, r = np.meshgrid(np.arange(8), np.arange(8))
    pos = np.stack([0.3 * c.ravel(), 0.3 * r.ravel(), np.zeros(64)], 1).astype(np.float32)
    fr = np.linspace(400e6, 800e6, F_).astype(np.float32)
    rng = np.random.default_rng(33)
    az0 = rng.uniform(0, 2 * np.pi, B)
    el0 = np.deg2rad(rng.uniform(30, 90, B))
    rate = np.deg2rad(rng.uniform(0, 15, (2, B))) / 3600.0  # rad/s
    tw = np.arange(W) * S * 2.56e-6
    az = az0[None] + rate[0][None] * tw[:, None]
    el = np.minimum(el0[None] + rate[1][None] * tw[:, None], np.pi / 2)
    dirs = np.stack([np.cos(el) * np.sin(az), np.cos(el) * np.cos(az), np.sin(el)],
                    -1).astype(np.float32)  # [W][B][3]
    stream = api.Stream()
    x = api.DeviceArray.empty((T, F_, A), np.uint8)
    dpos, dfr, ddir = (api.DeviceArray.from_numpy(a) for a in (pos, fr, dirs))
    sigma = 2.5
    api.generate_point_source_noise(stream, x, T, F_, A, dpos, dfr, dirs[0, 0],
                                    float(np.sqrt(3.0 * 2 * sigma * sigma)), sigma, 303)
    out = api.DeviceArray.empty((W, F_, B), np.float32)
    stream.synchronize()
    res = {}
    for name in ("SPLIT_F16", "F16", "FP32"):
        kern = api.Kernel[name]

        def launch():
            TR.multibeam_tracker_launch(stream, x, out, T, F_, A, B, S, dpos, dfr,
                                        directions=ddir, kernel=kern)
        steps = args.steps if name != "FP32" else max(3, args.steps // 4)
        ms, kms = timed_launches(stream, launch, steps, args.warmup)
        err = tracker_fp64_error(x, out, pos, fr, dirs, 2 * S, F_, A, B, S)
        res[name] = {"ms_per_step": round(ms, 4), "kernel_avg_ms": round(kms, 4),
                     "beam_samples_per_s": round(T * F_ * B / (ms * 1e-3), 1),
                     "max_rel_err_vs_fp64": err}
    split = res["SPLIT_F16"]
    cmac = T * F_ * B * A
    flops = 8 * cmac + 4 * T * F_ * B
    ach = flops / (split["kernel_avg_ms"] * 1e-3) / 1e12
    cpu = None if args.no_cpu_baseline else tracker_cpu_baseline(args.cpu_seconds, B, F_)
    return common_line(
        args, "beam-samples/s (T·F·B)", T * F_ * B / (split["ms_per_step"] * 1e-3),
        "beam-samples/s", split["ms_per_step"], "f16 split hi+lo, f32 accumulate",
        {"workload": "C3 dynamic tracker: A=64 8x8 0.3 m, F=672 (400-800 MHz), T=65536, "
                     "S=320 (W=205), B=32 sources drifting 0-15 deg/h in az/el, weights "
                     "regenerated per window on the device, point source + noise sigma 2.5",
         "n_ant": A, "n_freq": F_, "n_time": T, "n_beams": B, "integration_spectra": S,
         "kernel": "SPLIT_F16"},
        max_rel_err_vs_fp64=split["max_rel_err_vs_fp64"],
        roofline={"bound": "mfma", "achieved": round(ach, 2), "peak": 2500.0,
                  "unit": "TFLOP/s", "frac": round(ach / 2500.0, 4),
                  "traffic": read_profile_value("traffic_c3_tracker_split_f16.txt",
                                                "hbm_bytes_per_launch"),
                  "kernel_avg_ms": split["kernel_avg_ms"],
                  "algorithmic_bytes": T * F_ * A + W * F_ * B * 4},
        cpu_baseline=cpu, paths=res)


def tracker_cpu_baseline(seconds, B, F_):
    """The reference's OpenMP CPU tracker (beam_tracker_opt_v2, single beam,
    F=336 as it requires) on all host cores over a bounded T sample, scaled
    to beam-samples/s of B beams over F_ channels (the tracker runs once per
    source and per 336-channel shard)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # the checker/baseline library, never the product path
    R = O.ref_lib(native=True)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    A, F0, S = 64, 336, 320
    x_all = O.make_uniform_int4(S * 4 * F0 * A, 9).reshape(S * 4, F0, A)

    def run(T):
        x = np.ascontiguousarray(np.resize(x_all, (T, F0, A)))
        out = np.zeros((T, F0, 1), np.float32)
        t0 = time.perf_counter()
        if R is not None:
            R.ref_beam_tracker_opt_v2(x, T, F0, A, 0.0, 0.0, 1e-5, 0.0, S, out)
        else:
            O.beam_tracker(x, T, F0, A, 0.0, 0.0, 1e-5, 0.0, S, threads=True)
        return time.perf_counter() - t0

    T, med = median_of_five(run, S, S, seconds)
    return {"value": round(T * F0 / med, 1), "unit": "beam-samples/s", "cores": threads,
            "kind": "reference" if R is not None else "port",
            "sample": f"beam_tracker_opt_v2 (OpenMP), one beam, T={T} x F={F0}, A={A}, "
                      f"S={S}; 1 warm-up, median of 5 runs of {med:.3f} s; the C3 job is "
                      f"B={B} such beams over F={F_} channels, so its CPU time is "
                      f"{B * F_ // F0} times this per beam-sample"}


def bench_b1(args):
    """The reference's published tracker case: one beam, A=64, T=15360,
    F=336, S=320, device-resident (weights regenerated on the device)."""
    from beamform_project_amd import api, tracker as TR
    A, F_, T, S = 64, 336, 15360, 320
    x = api.DeviceArray.empty((T, F_, A), np.uint8)
    stream = api.Stream()
    api.generate_uniform_int4(stream, x, x.nbytes, 77)
    out = api.DeviceArray.empty((T, F_, 1), np.float32)
    stream.synchronize()
    dims = api.Dimensions(T, F_, A, 1)
    cfg = TR.TrackerConfig(TR.direction_from_lm(0.0, 0.0), (1e-5, 0.0), S)
    res = {}
    for name in ("AUTO", "SPLIT_F16", "FP32"):
        ex = TR.TrackerExec(api.Kernel[name])

        def launch():
            TR.beam_tracker_device_resident(x, out, dims, cfg, ex, stream)
        ms, kms = timed_launches(stream, launch, args.steps, args.warmup)
        res[name] = {"ms_per_step": round(ms, 4), "kernel_avg_ms": round(kms, 4),
                     "beam_samples_per_s": round(T * F_ / (ms * 1e-3), 1)}
    best = min(res, key=lambda k: res[k]["ms_per_step"])
    ms = res[best]["ms_per_step"]
    value = T * F_ / (ms * 1e-3)
    return common_line(
        args, "beam-samples/s (T·F·B)", value, "beam-samples/s", ms,
        "f16 split hi+lo, f32 accumulate" if best == "SPLIT_F16" else "f32",
        {"workload": "B=1 dynamic tracker, device-resident: A=64 (reference grid 0.6 m), "
                     "F=336, T=15360, S=320, per-spectrum output, uniform int4",
         "n_ant": A, "n_freq": F_, "n_time": T, "n_beams": 1, "integration_spectra": S,
         "kernel": best},
        roofline={"bound": "hbm",
                  "achieved": round((T * F_ * A + T * F_ * 4) / (res[best]["kernel_avg_ms"] * 1e-3) / 1e9, 1),
                  "peak": 8000.0, "unit": "GB/s",
                  "frac": round((T * F_ * A + T * F_ * 4) / (res[best]["kernel_avg_ms"] * 1e-3) / 8e12, 4),
                  "kernel_avg_ms": res[best]["kernel_avg_ms"],
                  "algorithmic_bytes": T * F_ * A + T * F_ * 4},
        vs_baseline=round(value / 3.59e8, 2),
        baseline_note="reference CUDA V5 device-resident A=64: 14.38 ms = 3.59e8 "
                      "beam-samples/s on a Quadro P1000",
        paths=res)


def bench_c4(args):
    """The streaming server alone (serving leg of the default line)."""
    sv = serving_run(args.serve_seconds)
    return common_line(
        args, "beam-samples/s (T·F·B); p99 per-window latency", sv["beam_samples_per_s"],
        "beam-samples/s", sv["wall_seconds"] * 1e3, "f16 split hi+lo, f32 accumulate",
        {"workload": sv["workload"]}, p99_window_latency_ms=sv["latency_p99_ms"], serving=sv)


def bench_c5(args):
    """Adaptive DOA: one update = K=1024 snapshots x 336 channels through
    the int8-MFMA covariance, FP64 fold/Cholesky/scans and the estimates."""
    from beamform_project_amd import api, doa as D
    A, F_, K, M, NTH = 64, 336, 1024, 48, 721
    n_upd = 8
    T = K * n_upd
    pos = np.zeros((A, 3), np.float32)
    pos[:, 0] = 0.3 * np.arange(A)
    fr = np.linspace(400e6, 800e6, F_).astype(np.float32)
    th0 = np.deg2rad(17.3)
    stream = api.Stream()
    x = api.DeviceArray.empty((T, F_, A), np.uint8)
    dpos, dfr = api.DeviceArray.from_numpy(pos), api.DeviceArray.from_numpy(fr)
    sigma = 2.5
    api.generate_point_source_noise(stream, x, T, F_, A, dpos, dfr,
                                    np.array([np.sin(th0), 0.0, np.cos(th0)], np.float32),
                                    float(np.sqrt(3.0 * 2 * sigma * sigma)), sigma, 505)
    est = D.DoaEstimator(D.DoaConfig(fr, n_ant=A, snapshots=K, subarray=M, n_theta=NTH,
                                     forgetting=0.98, diagonal_load=1e-3,
                                     element_spacing_m=0.3))
    for i in range(max(args.warmup, 1)):
        est.update(x, T, (i % n_upd) * K, stream)
    stream.synchronize()
    e0, e1 = api.Event(), api.Event()
    wall, kern = [], {"covariance_ms": [], "fold_scan_ms": [], "estimate_ms": []}
    for i in range(args.steps):
        e0.record(stream)
        est.update(x, T, (i % n_upd) * K, stream)
        e1.record(stream)
        wall.append(e0.elapsed_ms(e1))
        for k_, v in est.last_timings().items():
            kern[k_].append(v)
    ms = float(np.mean(wall))
    kt = {k_: round(float(np.mean(v)), 4) for k_, v in kern.items()}
    peaks = np.rad2deg(est.results()["capon_peaks"][:, 0]).tolist()
    cov_bytes = K * F_ * A + F_ * M * M * 2 * 4
    ach = cov_bytes / (kt["covariance_ms"] * 1e-3) / 1e9
    cpu = None if args.no_cpu_baseline else doa_cpu_baseline(args.cpu_seconds, A, K, M, NTH)
    return common_line(
        args, "DOA snapshot-channels/s (K·F per update)", K * F_ / (ms * 1e-3),
        "snapshot-channels/s", ms, "int8 MFMA covariance (int32 exact), f64 fold/Cholesky/scan",
        {"workload": "C5 adaptive DOA: A=64 ULA 0.3 m, F=336 (400-800 MHz), K=1024 snapshots "
                     "per update, forgetting 0.98, forward smoothing M=48 (L=17), 721 angles "
                     "over +-90 deg, diagonal load 1e-3 tr(R)/M, one source at 17.3 deg + noise",
         "n_ant": A, "n_freq": F_, "snapshots": K, "subarray": M, "n_theta": NTH},
        updates_per_s=round(1e3 / ms, 1), kernel_ms=kt, capon_peaks_deg=peaks,
        roofline={"bound": "hbm", "achieved": round(ach, 1), "peak": 8000.0, "unit": "GB/s",
                  "frac": round(ach / 8000.0, 4),
                  "traffic": read_profile_value("traffic_c5_covariance.txt",
                                                "hbm_bytes_per_launch"),
                  "kernel": "covariance", "kernel_avg_ms": kt["covariance_ms"],
                  "algorithmic_bytes": cov_bytes},
        cpu_baseline=cpu)


def bench_d2(args):
    """Adaptive planar DOA tracker (CpuOptBeamTracker's searching path) on
    the reference benchmark's configuration
    (benchmarks/benchmark_cpu_opt_beam_tracker.cpp:56-84): A=64, F=336,
    T=15360, S=320 (48 windows), 12 x 12 coarse grid over +-0.2, 2
    refinement levels, prior at zenith, source near l=0.03.  One step is one
    run over the shard: the 48 window searches back to back, then the
    emission; per-frame time = step time / windows."""
    from beamform_project_amd import api, opt_tracker as OT
    A, F_, T, S, G, LV = 64, 336, 15360, 320, 12, 2
    W = T // S
    stream = api.Stream()
    x = api.DeviceArray.empty((T, F_, A), np.uint8)
    pos = np.zeros((A, 3), np.float32)
    pos[:, 0] = (np.arange(A) % 8) * np.float32(0.6)
    pos[:, 1] = (np.arange(A) // 8) * np.float32(0.6)
    fr = (np.float32(300e6) + np.arange(F_, dtype=np.float32) * np.float32(300e3))
    dpos, dfr = api.DeviceArray.from_numpy(pos), api.DeviceArray.from_numpy(fr)
    src = np.array([0.03, 0.0, np.sqrt(1 - 0.03 ** 2)], np.float32)
    api.generate_point_source_noise(stream, x, T, F_, A, dpos, dfr, src, 4.0, 1.0, 2024)
    out = api.DeviceArray.empty((T, F_, 1), np.float32)
    stream.synchronize()
    dims = api.Dimensions(T, F_, A, 1)
    res, dirs = {}, {}
    for name, est in (("bartlett", OT.TrackerEstimator.BARTLETT),
                      ("capon", OT.TrackerEstimator.CAPON)):
        cfg = OT.OptTrackerConfig(est, G, LV, 0.2, 0.2, 0, 1.0, 1e-3, True, S)
        t = OT.OptBeamTracker(pos, fr, dims, cfg)
        t.seed_trajectory((0.0, 0.0, 1.0))

        def launch():
            t.run_device(x, out, stream)
        ms, kms = timed_launches(stream, launch, args.steps, args.warmup)
        d = t.window_directions()
        dirs[name] = [round(float(d[-1, 0]), 5), round(float(d[-1, 1]), 5)]
        t.close()
        res[name] = {"ms_per_run": round(ms, 4), "gpu_ms_per_run": round(kms, 4),
                     "ms_per_frame": round(kms / W, 4), "last_window_lm": dirs[name]}
    ms = res["bartlett"]["gpu_ms_per_run"]
    cpu = None if args.no_cpu_baseline else opt_tracker_cpu_baseline(args.cpu_seconds)
    return common_line(
        args, "windows/s (adaptive DOA tracker frames)", W / (ms * 1e-3), "windows/s", ms,
        "int8 MFMA covariance (int32 exact), f64 fold/Cholesky, f32 scan, split-f16 emission",
        {"workload": "adaptive planar DOA tracker: A=64 (reference grid 0.6 m), F=336, "
                     "T=15360, S=320 (48 frames), 12x12 coarse grid +-0.2, 2 refinement "
                     "levels, Bartlett; point source at l=0.03 in noise",
         "n_ant": A, "n_freq": F_, "n_time": T, "integration_spectra": S,
         "coarse_grid_resolution": G, "refinement_levels": LV},
        ms_per_frame=res["bartlett"]["ms_per_frame"], target_ms_per_frame=0.5, paths=res,
        cpu_baseline=cpu)


def opt_tracker_cpu_baseline(seconds):
    """The reference's CpuOptBeamTracker (oracle/_ref, OpenMP over the host's
    cores) on a bounded sample of the same configuration: whole frames of
    S=320 spectra, scaled to windows/s."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # the checker/baseline library, never the product path
    lib = O.ref_lib()
    if lib is None:
        return None
    A, F_, S = 64, 336, 320
    pos, fr = O.default_positions(A), O.channelized_frequencies(F_)

    def run(n_frames):
        T = n_frames * S
        x = O.make_moving_point_source(T, F_, A, pos, fr, 0.03, 0.0, 1e-5, 0.0)
        t0 = time.perf_counter()
        O.ref_opt_tracker(lib, x, 1, T, F_, A, pos, fr, (0.0, 0.0, 1.0), (0.0, 0.0), S, 0, 12,
                          2, 0.2, 0.2, 0, 1.0, 1e-3, True)
        return time.perf_counter() - t0

    dt = run(1)
    n = int(max(1, min(48, seconds / 2 / max(dt, 1e-6))))
    med = float(np.median([run(n) for _ in range(2)]))
    return {"value": round(n / med, 2), "unit": "windows/s",
            "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)),
            "kind": "reference",
            "sample": f"{n} frames of S=320 (A=64, F=336, G=12, 2 levels, Bartlett) through "
                      f"the reference's CpuOptBeamTracker; median of 2 runs of {med:.2f} s, "
                      f"{1e3 * med / n:.1f} ms per frame"}


def doa_cpu_baseline(seconds, A, K, M, NTH):
    """The reference's FP32 DOA primitives (src/cpu_opt_beam_tracker.cpp,
    composed per update by oracle/ref_doa_shim.cpp: covariance with
    sub-array smoothing, diagonal load, fold, Cholesky, Bartlett and Capon
    over the angle grid, OpenMP over channels as in run_into) on all host
    cores, over a bounded subset of the channels, scaled to
    snapshot-channels/s.  The FP64 C restatement stands in where
    oracle/_ref was not built."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # the checker/baseline library, never the product path
    lib = O.ref_doa_lib(native=True)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    theta = np.linspace(-np.pi / 2, np.pi / 2, NTH)
    xpos = (0.3 * np.arange(A)).astype(np.float32)

    def run(Fs):
        x = O.make_uniform_int4(K * Fs * A, 3)
        fr = np.linspace(400e6, 800e6, Fs).astype(np.float32)
        t0 = time.perf_counter()
        if lib is not None:
            Rb = np.zeros(Fs * M * M * 2, np.float32)
            Rc = np.zeros(Fs * M * M * 2, np.float32)
            pb = np.zeros(NTH, np.float32)
            pc = np.zeros(NTH, np.float32)
            assert lib.ref_doa_update_f32(x, K, Fs, A, 0, K, M, 0.98, 1e-3, 1, fr, xpos, NTH,
                                          theta, Rb, Rc, pb, pc) == 0
        else:
            state = np.zeros((Fs, M, M, 2), np.float64)
            load = np.zeros(Fs, np.float64)
            O.doa_update(x, K, Fs, A, 0, K, M, 0.98, 1e-3, True, fr, xpos, theta, state, load)
        return time.perf_counter() - t0

    step = max(1, threads)
    Fs, med = median_of_five(run, step, step, seconds, hi=336)
    return {"value": round(K * Fs / med, 1), "unit": "snapshot-channels/s",
            "cores": threads if lib is not None else 1,
            "kind": "reference" if lib is not None else "port",
            "sample": f"one update (K={K}, A={A}, M={M} with the reference's smoothing "
                      f"and forward-backward step, {NTH} angles, Bartlett + Capon) over "
                      f"{Fs} channels through the reference's FP32 DOA primitives; 1 "
                      f"warm-up, median of 5 runs of {med:.3f} s"}


def fp64_error(x, out, w, A, B, S, n_windows=32):
    """Floored max relative error of the first n_windows integrated windows
    against the same sums in FP64 (f32 weights, complex128 arithmetic);
    floor = A * E|x|^2 * S, the incoherent per-beam power."""
    from beamform_project_amd import _capi
    T = n_windows * S
    xb = np.empty((T, F, A), np.uint8)
    _capi.call("bf_memcpy_dtoh", xb.ctypes.data, x.ptr, xb.nbytes)
    got = np.empty((n_windows, F, B), np.float32)
    _capi.call("bf_memcpy_dtoh", got.ctypes.data, out.ptr, got.nbytes)
    re = ((xb & 0xF).astype(np.int16) ^ 8) - 8
    im = ((xb >> 4).astype(np.int16) ^ 8) - 8
    xc = re.astype(np.float64) + 1j * im.astype(np.float64)  # [T][F][A]
    wc = w[..., 0].astype(np.float64) + 1j * w[..., 1].astype(np.float64)  # [B][F][A]
    v = np.matmul(xc.transpose(1, 0, 2), wc.transpose(1, 2, 0))  # [F][T][B]
    p = (v.real ** 2 + v.imag ** 2).reshape(F, n_windows, S, B).sum(2).transpose(1, 0, 2)
    floor = A * float(np.mean(np.abs(xc) ** 2)) * S
    return float(np.max(np.abs(got - p) / np.maximum(p, floor)))


def serving_sweep(seconds, rates=(6.0, 12.0, 24.0, 48.0), max_active_beams=0):
    """The serving loop at rising offered loads (requests/s), `seconds` of
    stream each, until p99 passes the 0.5 ms budget (that load is the last
    one reported); with max_active_beams, admission also caps the beams in
    flight (requests then wait, reported as the mean admission wait)."""
    rows = []
    for r in rates:
        sv = serving_run(seconds, r, max_active_beams)
        rows.append({"arrival_rate_hz": r, "max_active_beams": max_active_beams,
                     "mean_admission_wait_ms": sv["mean_admission_wait_ms"],
                     "latency_p50_ms": sv["latency_p50_ms"],
                     "latency_p99_ms": sv["latency_p99_ms"],
                     "latency_max_ms": sv["latency_max_ms"],
                     "gpu_busy_frac": sv["gpu_busy_frac"],
                     "beam_samples_per_s": sv["beam_samples_per_s"],
                     "mean_requests_in_slots": sv["mean_requests_in_slots"],
                     "windows_timed": sv["windows_timed"]})
        if sv["latency_p99_ms"] > 0.5:
            break
    return rows


def serving_run(seconds, rate_hz=6.0, max_active_beams=0):
    """The streaming server under Poisson arrivals, paced at 2.56 us per
    spectrum (the native load generator bf_server_simulate)."""
    from beamform_project_amd import serve as SV

    c, r = np.meshgrid(np.arange(8), np.arange(8))
    pos = np.stack([0.3 * c.ravel(), 0.3 * r.ravel(), np.zeros(64)], 1).astype(np.float32)
    fr = np.linspace(400e6, 800e6, F).astype(np.float32)
    cfg = SV.SimConfig(
        server=SV.ServerConfig(n_freq=F, n_ant=64, ring_spectra=1 << 20, max_window=320,
                               max_step_spectra=8192, hbm_budget_bytes=2 << 30,
                               max_requests=256, positions=pos, frequencies=fr,
                               enable_graph=True, max_active_beams=max_active_beams),
        seed=4242, cadence_s=2.56e-6, run_seconds=seconds, chunk_spectra=128,
        arrival_rate_hz=rate_hz, min_beams=1, max_beams=64, min_duration_s=0.5,
        max_duration_s=5.0, window_a=10, window_b=320, output_windows_cap=1024, paced=True,
        pool_spectra=65536, source_amplitude=float(np.sqrt(3.0 * 2 * 2.5 * 2.5)),
        noise_sigma=2.5)
    rep = SV.simulate(cfg)
    st = rep["stats"]
    return {
        "workload": f"C4 streaming: Poisson arrivals {rate_hz:g}/s, sources U{{1..64}}, windows 10|320, "
                    "durations U[0.5,5] s, A=64, F=336, ring 2^20 spectra (22.5 GB), "
                    "HBM budget 2 GiB for outputs (1024-window output rings), "
                    f"{seconds:g} s of stream paced at 2.56 us/spectrum, 128-spectrum ingest",
        "latency_p50_ms": round(st["latency_p50_ms"], 4),
        "latency_p90_ms": round(st["latency_p90_ms"], 4),
        "latency_p99_ms": round(st["latency_p99_ms"], 4),
        "latency_max_ms": round(st["latency_max_ms"], 4),
        "latency_budget_ms": 0.5,
        "windows_timed": st["windows_timed"],
        "beam_samples_per_s": round(rep["beam_samples_per_s"], 1),
        "gpu_busy_frac": round(st["gpu_busy_ms"] / 1e3 / rep["wall_seconds"], 4),
        "steps": st["steps"],
        "requests_generated": rep["requests_generated"],
        "requests_admitted": st["requests_admitted"],
        "requests_completed": st["requests_completed"],
        "mean_requests_in_slots": round(rep["mean_requests_in_slots"], 2),
        "arena_peak_bytes": st["arena_peak_bytes"],
        "mean_admission_wait_ms": round(st["mean_admission_wait_spectra"] * 2.56e-3, 3),
        "max_pacing_lag_ms": round(rep["max_pacing_lag_ms"], 3),
        "wall_seconds": round(rep["wall_seconds"], 3),
    }


def read_traffic(wl):
    """HBM bytes per launch from the committed rocprofv3 PMC summary, or None."""
    name = "traffic_" + ("c2_split_f16" if not wl["ula"] else "c1_fp32") + ".txt"
    return read_profile_value(name, "hbm_bytes_per_launch")


def read_profile_value(name, key):
    """A `key: value` figure from a committed profile summary, or None."""
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
    return vals.get(key)


def cpu_baseline(wl, pos, fr, dirs, seconds, per_spectrum=False):
    """The reference's single-threaded CPU fixed-grid path
    (cpu_beamform_packed_intensity_into, then cpu_integrate_intensity and
    cpu_quantize_integrated_intensity when the workload integrates) on a
    bounded prefix of the workload (T_sample spectra, all channels and
    beams), scaled to beam-samples/s."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # the checker/baseline library, never the product path

    A, B, S = wl["A"], wl["B"], (0 if per_spectrum else wl["S"])
    w = weights_from(pos, fr, dirs)
    R = O.ref_lib(native=True)
    kind = "reference" if R is not None else "port"
    x_all = O.make_uniform_int4(1024 * F * A, 5).reshape(1024, F, A)

    def run(T):
        x = np.ascontiguousarray(np.resize(x_all, (T, F, A)))
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
                O.quantize(O.integrate(o, T, F, B, S), (T + S - 1) // S, F, B)
        return time.perf_counter() - t0

    step = S if S else 1
    T, med = median_of_five(run, max(step, 10), step, seconds, hi=wl["T"])
    stages = ("cpu_beamform_packed_intensity_into"
              + (" + cpu_integrate_intensity + cpu_quantize_integrated_intensity" if S else ""))
    return {"value": round(T * F * B / med, 1), "unit": "beam-samples/s", "cores": 1,
            "kind": kind,
            "sample": f"T={T} spectra x F={F} x B={B} (A={A}) through {stages}; 1 warm-up, "
                      f"median of 5 runs of {med:.3f} s; single thread as in the reference"}


def cpu_baseline_all_cores(wl, pos, fr, dirs, seconds):
    """Not the reference: its fixed-grid loop parallelised naively over
    (t, f) on all host cores (the oracle's OpenMP restatement, same
    arithmetic), then integration and quantisation, as an upper bound for a
    CPU implementation of the same loop."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # the checker/baseline library, never the product path

    A, B, S = wl["A"], wl["B"], wl["S"]
    w = weights_from(pos, fr, dirs)
    x_all = O.make_uniform_int4(1024 * F * A, 5).reshape(1024, F, A)

    def run(T):
        x = np.ascontiguousarray(np.resize(x_all, (T, F, A)))
        t0 = time.perf_counter()
        o = O.beamform(x, w, T, F, A, B, threads=True)
        if S:
            O.quantize(O.integrate(o, T, F, B, S), (T + S - 1) // S, F, B)
        return time.perf_counter() - t0

    step = S if S else 1
    T, med = median_of_five(run, max(step, 10), step, seconds)
    return {"value": round(T * F * B / med, 1), "unit": "beam-samples/s",
            "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)),
            "kind": "port",
            "sample": f"NOT the reference (which is single-threaded): its loop over (t, f) "
                      f"in OpenMP, T={T} x F={F} x B={B} (A={A}), then integration and "
                      f"quantisation; 1 warm-up, median of 5 runs of {med:.3f} s"}


def bench_c1(args):
    """Plain fixed grid: A=32 ULA at 0.3 m, F=336 over 400-800 MHz,
    T=1024, B=16 beams over +-60 deg, uniform int4, f32 [T][F][B] output."""
    from beamform_project_amd import api
    A, T, B = 32, 1024, 16
    wl = dict(A=A, T=T, B=B, S=0, ula=True)
    pos, fr, dirs = host_geometry(wl)
    w = weights_from(pos, fr, dirs)
    stream = api.Stream()
    x = api.DeviceArray.empty((T, F, A), np.uint8)
    api.generate_uniform_int4(stream, x, x.nbytes, 1234)
    dw = api.DeviceArray.from_numpy(w)
    out = api.DeviceArray.empty((T, F, B), np.float32)
    stream.synchronize()
    dims = api.Dimensions(T, F, A, B)
    steps = max(args.steps, 50)
    res = {}
    tiles = {}
    for name in ("SPLIT_F16", "F16"):
        # the library's autotuner picks this problem's tile shape (untimed)
        tab = api.autotune_packed_beamformer(api.Kernel[name], stream, x, dw, out, dims, 0,
                                             repeats=5)
        best = min(tab, key=lambda r: r["median_ms"])
        tiles[name] = {k: best[k] for k in ("beams_per_wave", "waves_per_workgroup",
                                            "double_buffered", "resident_workgroups")}
    for name in ("SPLIT_F16", "F16", "FP32"):
        kern = api.Kernel[name]

        def launch():
            api.launch_packed_beamformer(kern, stream, x, dw, out, dims)
        ms, kms = timed_launches(stream, launch, steps, args.warmup)
        err = fp64_error_spectra(x, out, w, T, A, B)
        res[name] = {"ms_per_step": round(ms, 4), "kernel_avg_ms": round(kms, 4),
                     "beam_samples_per_s": round(T * F * B / (ms * 1e-3), 1),
                     "max_rel_err_vs_fp64": err}
    split = res["SPLIT_F16"]
    bytes_alg = T * F * A + B * F * A * 8 + T * F * B * 4
    ach = bytes_alg / (split["kernel_avg_ms"] * 1e-3) / 1e9
    flops = 8 * T * F * B * A + 3 * T * F * B
    fp32_ach = flops / (res["FP32"]["kernel_avg_ms"] * 1e-3) / 1e12
    cpu = None if args.no_cpu_baseline else cpu_baseline(wl, pos, fr, dirs, args.cpu_seconds,
                                                         per_spectrum=True)
    return common_line(
        args, "beam-samples/s (T·F·B)", T * F * B / (split["ms_per_step"] * 1e-3),
        "beam-samples/s", split["ms_per_step"], "f16 split hi+lo, f32 accumulate",
        {"workload": "C1 plain fixed grid: A=32 ULA 0.3 m, F=336 (400-800 MHz), T=1024, "
                     "B=16 over +-60 deg, uniform int4, f32 [T][F][B] output",
         "n_ant": A, "n_freq": F, "n_time": T, "n_beams": B, "integration_spectra": 0,
         "kernel": "SPLIT_F16", "tile_shape": tiles["SPLIT_F16"]},
        max_rel_err_vs_fp64=split["max_rel_err_vs_fp64"],
        roofline={"bound": "hbm", "achieved": round(ach, 1), "peak": 8000.0, "unit": "GB/s",
                  "frac": round(ach / 8000.0, 4),
                  "traffic": read_profile_value("traffic_c1_split_f16.txt",
                                                "hbm_bytes_per_launch"),
                  "kernel_avg_ms": split["kernel_avg_ms"], "algorithmic_bytes": bytes_alg},
        fp32_roofline={"bound": "valu-fp32", "achieved": round(fp32_ach, 2), "peak": 157.3,
                       "unit": "TFLOP/s", "frac": round(fp32_ach / 157.3, 4)},
        cpu_baseline=cpu, paths=res)


def fp64_error_spectra(x, out, w, T, A, B):
    """Floored max relative error of a per-spectrum output [T][F][B] against
    complex128 sums with the same f32 weights; floor = A * E|x|^2."""
    from beamform_project_amd import _capi
    xb = np.empty((T, F, A), np.uint8)
    _capi.call("bf_memcpy_dtoh", xb.ctypes.data, x.ptr, xb.nbytes)
    got = out.to_numpy()
    re = ((xb & 0xF).astype(np.int16) ^ 8) - 8
    im = ((xb >> 4).astype(np.int16) ^ 8) - 8
    xc = re.astype(np.float64) + 1j * im.astype(np.float64)
    wc = w[..., 0].astype(np.float64) + 1j * w[..., 1].astype(np.float64)  # [B][F][A]
    v = np.matmul(xc.transpose(1, 0, 2), wc.transpose(1, 2, 0))  # [F][T][B]
    p = (v.real ** 2 + v.imag ** 2).transpose(1, 0, 2)
    floor = A * float(np.mean(np.abs(xc) ** 2))
    return float(np.max(np.abs(got - p) / np.maximum(p, floor)))


if __name__ == "__main__":
    main()
