"""Benchmark: SAR backprojection + forward projection throughput on one MI355X.

A step is one backprojection (A^H) and one forward projection (A) of the
large single-scene workload (1024 x 1024 pixels at 0.10 m, 2048 pulses x 4096
samples, complex64 phase history, range upsampling 8, linear interpolation,
monostatic circular path R = 10 km, h = 5 km, 4 degree aperture), with
device-resident inputs.  The reported value is (2 * M * P) pixel-pulse
updates per step / step time, in Gpixel-pulse/s.

Also reported:
  roofline      the dominant kernel (the longer of the forward scatter and the
                backprojection gather) against the MI355X FP32 vector peak:
                algorithmic flop per pixel-pulse (40 scatter / 39 gather,
                linear interpolation, 8 less on monostatic pulses) x M x P per
                launch / mean launch time, measured with HIP events around that
                launch on the library's stream; the other kernel as
                roofline_other_kernel;
  spectral_stage  each operator's spectral stage against HBM: 8 P (N + N_w)
                bytes per application / its mean device time;
  cpu_baseline  the FP64 CPU oracle (a restatement of the reference algorithm:
                per-pulse scatter/gather + full N_w FFT per scene block, threads
                over scene blocks as the reference's joblib pool) on bounded
                samples (pulse subsets of this workload on all cores and on one
                core, and the whole 128^2 workload), medians of 5 runs;
  pcg           time to relative residual 1e-3 for (A^H A + 1e-3 I) x = A^H y
                (A normalised by its 2-norm) on held-out 20 dB scenes of the
                512^2 workload (1024 pulses x 2048 samples, 3 deg): plain CGLS,
                PCG with the analytic demodulated circulant base, PCG with the
                learned spectral filter alone, and PCG with the learned
                preconditioner (learned filter + bf16 CNN, committed weights);
                medians over the scenes and their time ratios to CGLS.
Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c2_large]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

# Algorithmic flop per pixel-pulse of the two main kernels, as the hot-path
# work model counts them: two distances 18 (3 sub + 3 mul + 2 add + 1 sqrt
# each), minus the SRP range 2, bin scale and fraction 2, complex lerp 6,
# phase and sin/cos 3, complex multiply 6, accumulate 2 = 39 for the linear
# backprojection; 32 without interpolation (nearest rules); 40 / 32 for the
# forward.  A monostatic pulse evaluates one distance, counted twice above:
# 8 less (the plan reports `monostatic`).
FLOP_PER_PX_PULSE = {"linear": 39.0, "nearest": 32.0, "nearest_ref": 32.0}
FWD_FLOP_PER_PX_PULSE = {"linear": 40.0, "nearest": 32.0, "nearest_ref": 32.0}
MONOSTATIC_SAVING = 8.0
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c2_large")
    ap.add_argument("--interp", default="linear")
    ap.add_argument("--precision", default="series")
    ap.add_argument("--cpu-sample-pulses", type=int, default=128)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="copy a 512 MiB buffer once before timing (HBM counter scaling)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-scene-forward", action="store_true",
                    help="skip the informational forward of the sparse scene (profiling runs: "
                         "every scatter launch then has the dense input)")
    ap.add_argument("--pcg-scenes", type=int, default=64,
                    help="held-out scenes of the learned-PCG vs CGLS comparison (0: skip)")
    ap.add_argument("--resample-pulses", type=int, default=16384,
                    help="input pulses of the dataset-resampling measurement (0: skip)")
    return ap.parse_args()


def _oracle_pass(O, ap, pos, x, y, U, interp, pulses, threads):
    """One forward + one adjoint of the FP64 oracle on the first `pulses` pulses
    (all pixels); returns wall seconds."""
    sl = slice(0, pulses)
    tx = ap.transmitter_positions[sl]
    srp = ap.stab_ref_posns[sl]
    W = np.fft.ifftshift(ap.waveform_fft)
    wts = ap.slow_time_weighting[sl]
    args = (ap.centre_frequency, ap.sample_bandwidth)
    t0 = time.perf_counter()
    O.forward(tx, tx, srp, pos, x, W, wts, *args, ap.num_freqs, c_eff=ap.c_eff,
              upsample_ratio=U, interp=interp, num_threads=threads)
    O.adjoint(tx, tx, srp, pos, np.fft.ifftshift(y[sl], axes=1), W, wts, *args, c_eff=ap.c_eff,
              upsample_ratio=U, interp=interp, num_threads=threads)
    return time.perf_counter() - t0


def cpu_baseline(w, ap, pos, x, y, interp, pulses, repeats=5):
    """The FP64 CPU oracle (a restatement of the reference algorithm: per-pulse
    scatter / gather + a full N_w FFT per pulse in every scene block, T = cores
    - 1 scene-block threads as the reference's joblib pool) on bounded samples,
    each the median of `repeats` runs of one forward + one adjoint:
      value        the benchmarked workload, first `pulses` pulses, all pixels, T threads;
      single_thread  the same workload, first pulses // 8 pulses, 1 thread (per-core rate);
      c1_small     the whole 128^2 / 256 x 512 workload, T threads.
    Rates in Gpixel-pulse/s (2 M P' / median wall)."""
    import ctypes.util
    import oracle.oracle as O
    from sarsolver_amd import synthetic
    O.build()
    ncores = len(os.sched_getaffinity(0))
    threads = max(1, ncores - 1)
    fft = ("FFTW present on the host but not used: the oracle's own radix-2 / Bluestein FFT"
           if ctypes.util.find_library("fftw3") else
           "the oracle's own radix-2 / Bluestein FFT (no FFTW on the host)")

    def med(args, n):
        ts = [_oracle_pass(O, *args) for _ in range(n)]
        return float(np.median(ts)), ts

    t_main, ts_main = med((ap, pos, x, y, w.upsample, interp, pulses, threads), repeats)
    p1 = max(1, pulses // 8)
    t_one, _ = med((ap, pos, x, y, w.upsample, interp, p1, 1), repeats)
    w1 = synthetic.WORKLOADS["c1_small"]
    ap1, sc1 = synthetic.build(w1)
    g = np.random.default_rng(w1.seed)
    x1 = synthetic.scene_amplitudes(w1, g).reshape(-1)
    y1 = synthetic.complex_normal(g, (w1.pulses, w1.samples))
    t_c1, _ = med((ap1, sc1.positions, x1, y1, w1.upsample, interp, w1.pulses, threads), repeats)
    m = pos.shape[0]
    rate = lambda mm, pp, t: 2.0 * mm * pp / t / 1e9
    return {"value": rate(m, pulses, t_main), "unit": "Gpixel-pulse/s", "cores": threads,
            "kind": "port",
            "sample": f"{w.name} geometry, first {pulses} of {w.pulses} pulses, all {m} pixels, "
                      f"one forward + one adjoint ({interp}, FP64), median of {repeats} runs "
                      f"({', '.join(f'{t:.2f}' for t in ts_main)} s) on {threads} threads "
                      f"(host has {ncores}); FFT: {fft}",
            "single_thread": {"value": rate(m, p1, t_one), "cores": 1,
                              "sample": f"{w.name}, first {p1} pulses, median of {repeats}"},
            "c1_small": {"value": rate(sc1.num_scatterers, w1.pulses, t_c1), "cores": threads,
                         "sample": f"whole c1_small workload ({w1.n}^2 px, {w1.pulses} x "
                                   f"{w1.samples}), median of {repeats}, "
                                   f"{t_c1 * 1e3:.1f} ms per forward + adjoint"}}


def traffic_from_profile(kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    FETCH_SIZE / WRITE_SIZE runs of this benchmark (profiles/traffic.json):
    static, measured on the build the file names, not in this run (null if
    absent)."""
    f = ROOT / "profiles" / "traffic.json"
    if not f.exists():
        return None
    try:
        return json.loads(f.read_text()).get(kernel, {}).get("bytes_per_launch")
    except Exception:
        return None


def resample_measurement(dev, P: int):
    """Dataset resampling (`sarhip_resample`) of a synthetic P x 4096 phase
    history on the circular 4 deg aperture: slow factor 4 (20-tap kernel), 4096
    -> 2048 samples, new SRP 5 m off.  The whole call is timed after one
    warm-up call (which builds the cached work buffers and FFT plans): the
    per-call uploads of the small host arrays and the closing stream
    synchronisation are inside; bytes = input read once + output written once.
    cpu_baseline: the FP64 oracle's window loop (the reference's order) on 64
    windows, scaled to all windows."""
    import torch
    from oracle import resample as rs
    from sarsolver_amd import _native, synthetic
    from sarsolver_amd.utils import get_downsample_kernel, srp_range_change
    from sarsolver_amd.utils.signal import downsample_window
    N, N2, f = 4096, 2048, 4
    ap = synthetic.circular_aperture(P, N, 4.0)
    g = np.random.default_rng(17)
    ph = (g.standard_normal((P, N)) + 1j * g.standard_normal((P, N))).astype(np.complex64)
    new_srp = np.array([3.0, 4.0, 0.0])
    tx, srps = ap.transmitter_positions, ap.stab_ref_posns
    dr = srp_range_change(srps, new_srp, tx, tx)
    kern = get_downsample_kernel(f)
    K = kern.shape[0]
    starts = np.arange(((P - K) // f) - 1, dtype=np.int32) * f
    win = downsample_window(N2, 1.2) / N
    phd = torch.from_numpy(ph).to(dev)
    run = lambda: _native.resample(phd, dr, ap.centre_frequency, ap.sample_bandwidth, 1.0,
                                   ap.c_eff, N2, win, kern, starts)
    run()
    torch.cuda.synchronize()
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        run()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / reps * 1e3
    nbytes = 8.0 * (P * N + starts.size * N2)
    sub = 64
    c0 = time.perf_counter()
    rs.resample(ph[: sub * f + K].astype(np.complex128), tx[: sub * f + K], tx[: sub * f + K],
                srps[: sub * f + K], np.ones(sub * f + K, complex), np.ones(N, complex),
                ap.centre_frequency, ap.sample_bandwidth, 1.0, ap.c_eff, [], new_srp, N2, f)
    cpu_s = (time.perf_counter() - c0) * starts.size / max(((sub * f + K - K) // f) - 1, 1)
    return {"metric": "dataset resampling (SRP move + fast/slow downsample)",
            "workload": f"{P} pulses x {N} samples -> {starts.size} x {N2}, slow factor {f}",
            "ms": round(ms, 3), "algorithmic_gb_s": round(nbytes / (ms * 1e-3) / 1e9, 1),
            "hbm_peak_gb_s": HBM_PEAK_GBS,
            "cpu_baseline_ms": round(cpu_s * 1e3, 1), "cpu_kind": "port (FP64 oracle window loop, 1 thread, 64 windows scaled)"}


def pcg_comparison(dev, scenes: int):
    """Learned-PCG vs CGLS on the 512^2 workload with the committed weights."""
    from sarsolver_amd import learned
    from sarsolver_amd.train_preconditioner import default_weights, workload_operator
    path = default_weights("c5_scene")
    if not path.exists():
        return {"skipped": f"no trained weights at {path.relative_to(ROOT)}"}
    w, op = workload_operator("c5_scene", dev)
    n = w.n
    pre = learned.load_preconditioner(path, op, n)
    analytic = learned.CirculantPreconditioner(op, n, pre.base.tau, pad=pre.base.pad)
    res = learned.evaluate(op, {"circulant": analytic, "learned_filter": pre.base,
                                "learned": lambda r: pre(r, (n, n))}, n, scenes=scenes)
    s = res["solvers"]
    return {
        "metric": "time to 1e-3 residual, learned-PCG vs CGLS",
        "value": round(s["learned"]["time_ratio_vs_cgls"], 4), "unit": "ratio (lower is better)",
        "workload": f"{w.name}: {n}x{n} px, {w.pulses} pulses x {w.samples} samples, "
                    f"{w.aperture_deg} deg, lam 1e-3, 20 dB, held-out seed 12345",
        "scenes": scenes,
        "median_ms": {k: round(v["median_ms"], 3) for k, v in s.items()},
        "median_iters": {k: v["median_iters"] for k, v in s.items()},
        "time_ratio_vs_cgls": {k: round(v["time_ratio_vs_cgls"], 4)
                               for k, v in s.items() if k != "cgls"},
        "max_true_residual": {k: float(f"{v['max_true_residual']:.3e}") for k, v in s.items()},
        "fallbacks": {k: v["fallbacks"] for k, v in s.items()},
        "tau": pre.base.tau, "weights": str(path.relative_to(ROOT)),
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU (asked for {args.gpus}): sharding the "
              "pulses or pixels over several GPUs is designed but not built or run.",
              file=sys.stderr)
        sys.exit(2)
    import torch
    from sarsolver_amd import synthetic

    w = synthetic.WORKLOADS[args.workload]
    rng = np.random.default_rng(w.seed)
    ap, sc = synthetic.build(w)
    pos = sc.positions
    x = synthetic.scene_amplitudes(w, rng).reshape(-1)
    dev = torch.device("cuda", 0)
    pl = synthetic.workload_plan(w, ap, pos, dev, interp=args.interp, precision=args.precision)
    xd = torch.from_numpy(x.astype(np.complex64)).to(dev)
    # The timed forward runs on a dense image (complex Gaussian, every pixel
    # non-zero), as the iterates of a solver are: the forward skips regions of
    # the image that are exactly zero, so the sparse scene would credit it
    # with pixel-pulses it never processed.  The scene's own rate is reported
    # separately.
    xdense = synthetic.dense_image(w, pos.shape[0])
    xdd = torch.from_numpy(xdense.astype(np.complex64)).to(dev)
    if args.calibrate_copy:
        src = torch.ones(512 * 2 ** 20 // 4, dtype=torch.float32, device=dev)
        dst = src.clone()
        torch.cuda.synchronize()
        del src, dst
    # the adjoint's input: the phase history of the dense image (its cost does
    # not depend on the content; every forward launch then has the same input)
    yd = pl.forward(xdd)
    y = yd.cpu().numpy()
    img = torch.empty_like(xd)
    ph = torch.empty_like(yd)
    for _ in range(args.warmup):
        pl.adjoint(yd, out=img)
        pl.forward(xdd, out=ph)
    torch.cuda.synchronize()
    pl.set_timing(True)
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    start.record()
    for _ in range(args.steps):
        pl.adjoint(yd, out=img)
        pl.forward(xdd, out=ph)
    stop.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = start.elapsed_time(stop) / args.steps
    t_adj = pl.timing("adjoint")
    t_fwd = pl.timing("forward")
    # forward of the sparse scene itself (informational, not in `value`)
    t_fwd_scene = None
    if not args.no_scene_forward:
        pl.set_timing(True)
        for _ in range(max(1, args.steps // 2)):
            pl.forward(xd, out=ph)
        t_fwd_scene = pl.timing("forward")
    pl.set_timing(False)
    M, P = pos.shape[0], w.pulses
    pp = float(M) * P
    value = 2.0 * pp / (ms_step * 1e-3) / 1e9
    # roofline of the dominant kernel (longer mean launch, HIP events around
    # the launch on the library stream); the other main kernel alongside
    mono_save = MONOSTATIC_SAVING if pl.layout["monostatic"] else 0.0
    kern = {
        "bp_gather_kernel": (FLOP_PER_PX_PULSE[args.interp] - mono_save, t_adj["main_kernel"]),
        "fp_scatter_kernel": (FWD_FLOP_PER_PX_PULSE[args.interp] - mono_save,
                              t_fwd["main_kernel"]),
    }
    def roof(name):
        fpp, ms = kern[name]
        achieved = fpp * pp / (ms * 1e-3) / 1e12
        return {"bound": "valu", "kernel": name, "achieved": round(achieved, 3),
                "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
                "traffic": traffic_from_profile(name),
                "traffic_source": "static: profiles/traffic.json (rocprofv3 --pmc FETCH_SIZE / "
                                  "WRITE_SIZE of this benchmark, scaled by a copy kernel)",
                "launch_ms": round(ms, 4),
                "launch_ms_source": "mean of HIP events around each launch on the library "
                                    "stream over the timed steps",
                "algorithmic_flop_per_px_pulse": fpp}
    # spectral stage of each operator against HBM: 8 P (N + N_w) bytes per
    # application (read the band, write the upsampled rows, or the reverse)
    nw = int(w.upsample * w.samples)
    spec_bytes = 8.0 * P * (w.samples + nw)
    def spectral(t):
        gbs = spec_bytes / (t["fft"] * 1e-3) / 1e9
        return {"bound": "hbm", "algorithmic_bytes": spec_bytes, "ms": round(t["fft"], 4),
                "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(gbs / HBM_PEAK_GBS, 4)}
    kernel = max(kern, key=lambda n: kern[n][1])
    other = [n for n in kern if n != kernel][0]
    line = {
        "metric": "Gpixel-pulse/s backproj+forward",
        "value": round(value, 3),
        "unit": "Gpixel-pulse/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 (f64 range)",
        "data": "synthetic",
        "config": {"workload": w.name, "pixels": f"{w.n}x{w.n}", "spacing_m": w.spacing,
                   "pulses": w.pulses, "samples": w.samples, "upsample": w.upsample,
                   "interp": args.interp, "precision": args.precision,
                   "aperture_deg": w.aperture_deg, "global_batch": 1, "parallelism": "none"},
        "backprojection_gpixel_pulse_s": round(pp / (t_adj["total"] * 1e-3) / 1e9, 3),
        "forward_gpixel_pulse_s": round(pp / (t_fwd["total"] * 1e-3) / 1e9, 3),
        "forward_input": "dense complex Gaussian image (seeded); every pixel processed",
        "forward_scene_gpixel_pulse_s": (None if t_fwd_scene is None else
                                         round(pp / (t_fwd_scene["total"] * 1e-3) / 1e9, 3)),
        "stage_ms": {"adjoint": {k: round(v, 4) for k, v in t_adj.items()},
                     "forward": {k: round(v, 4) for k, v in t_fwd.items()}},
        "roofline": roof(kernel),
        "roofline_other_kernel": roof(other),
        "spectral_stage": {"adjoint": spectral(t_adj), "forward": spectral(t_fwd)},
        "wall_s_timed": round(wall, 3),
        "plan_layout": pl.layout,
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(w, ap, pos, x, y, args.interp, args.cpu_sample_pulses)
    if args.pcg_scenes > 0:
        line["pcg"] = pcg_comparison(dev, args.pcg_scenes)
    if args.resample_pulses > 0:
        line["resample"] = resample_measurement(dev, args.resample_pulses)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
