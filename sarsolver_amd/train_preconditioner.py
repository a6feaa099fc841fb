"""Train and evaluate the learned CG preconditioner for one synthetic workload.

    python -m sarsolver_amd.train_preconditioner [--workload c5_scene]
        [--out sarsolver_amd/weights/precond_c5_scene.npz] [--eval-scenes 64]

Steps (all on the GPU, through the project's own SAR kernels):
1. plan for the workload (512^2 scene, 1024 pulses x 2048 samples, 3 deg,
   U = 8, linear interpolation, series ranges) and Â = A / ||A||_2;
2. tau of the demodulated circulant base chosen on 8 training problems;
3. the base's spectral log-gain trained through 12 unrolled PCG iterations
   (--filter-steps problems of --batch scenes);
4. the six-layer bf16 CNN trained on top of it through 12 unrolled PCG
   iterations (--steps problems of --batch scenes);
5. weights written to --out (.npz: FP32 network, float16 log-gain, tau, pad
   and training settings);
6. CGLS, PCG with the analytic circulant base, PCG with the learned filter
   alone and PCG with the learned filter + CNN on --eval-scenes held-out 20 dB
   scenes, to relative residual 1e-3; the summary (medians, time ratios,
   worst true residuals) printed as JSON.
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np
import torch

from . import learned, synthetic
from ._native import SarPlan

PKG = Path(__file__).resolve().parent
TAUS = (0.03, 0.05, 0.08, 0.12, 0.2)
PAD_FACTOR = 2


def workload_operator(name: str, device):
    w = synthetic.WORKLOADS[name]
    ap, sc = synthetic.build(w)
    plan = SarPlan(ap.transmitter_positions, ap.receiver_positions, ap.stab_ref_posns,
                   sc.positions, np.fft.ifftshift(ap.waveform_fft), ap.slow_time_weighting,
                   ap.centre_frequency, ap.sample_bandwidth, ap.c_eff, w.upsample, 1.0,
                   interp="linear", precision="series", device=device)
    return w, learned.NormalisedOperator(plan)


def default_weights(name: str) -> Path:
    return PKG / "weights" / f"precond_{name}.npz"


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--workload", default="c5_scene")
    ap.add_argument("--out", default=None)
    ap.add_argument("--steps", type=int, default=2500, help="CNN training steps")
    ap.add_argument("--filter-steps", type=int, default=2500, help="spectral-gain training steps")
    ap.add_argument("--filter-lr", type=float, default=0.02)
    ap.add_argument("--iters", type=int, default=12)
    ap.add_argument("--batch", type=int, default=2)
    ap.add_argument("--lr", type=float, default=3e-4)
    ap.add_argument("--eval-scenes", type=int, default=64)
    ap.add_argument("--summary", default=None, help="also write the summary JSON here")
    ap.add_argument("--init", default=None,
                    help="continue training from these weights (the circulant base keeps their tau)")
    ap.add_argument("--seed", type=int, default=5, help="seed of the training scenes")
    ap.add_argument("--no-eval", action="store_true", help="skip the held-out evaluation")
    a = ap.parse_args(argv)
    if not torch.cuda.is_available():
        print("a GPU is required", file=sys.stderr)
        return 2
    dev = torch.device("cuda", 0)
    w, op = workload_operator(a.workload, dev)
    n, lam = w.n, 1e-3
    if a.init:
        pre = learned.load_preconditioner(a.init, op, n)
        base, tau = pre.base, pre.base.tau
        print(f"continuing from {a.init} (tau {tau})", flush=True)
    else:
        gens = learned.make_gens(777, dev)
        with torch.no_grad():
            probs = [learned.noisy_problem(op, x, 20.0, gens.dev_gen)
                     for x in learned.random_scenes(n, 8, gens, dev)]
        tau, scores = learned.choose_tau(op, n, TAUS, probs, lam=lam, pad=PAD_FACTOR * n)
        print(f"tau scores (median PCG iterations): {scores} -> tau {tau}", flush=True)
        base = learned.CirculantPreconditioner(op, n, tau, pad=PAD_FACTOR * n)
        pre = learned.CnnPreconditioner(dev, base=base)
    analytic = learned.CirculantPreconditioner(op, n, tau, pad=PAD_FACTOR * n)
    cval = learned.problem_batch(op, n, 4, learned.make_gens(4242, dev))
    prog = lambda st, l, el: print(f"  step {st} loss {l:.4f} ({el:.1f} s)", flush=True)
    with torch.no_grad():
        la = float(learned.unrolled_pcg_loss(op, analytic, cval, lam, a.iters))
    finfo = {"seconds": 0.0}
    if a.filter_steps > 0:
        print(f"spectral gain: {a.filter_steps} steps", flush=True)
        finfo = learned.train_filter(op, base, n, lam=lam, iters=a.iters, steps=a.filter_steps,
                                     batch=a.batch, lr=a.filter_lr, seed=a.seed + 1000,
                                     log_every=max(1, a.filter_steps // 10), progress=prog)
    with torch.no_grad():
        lb = float(learned.unrolled_pcg_loss(op, base, cval, lam, a.iters))
        l0 = float(learned.unrolled_pcg_loss(op, lambda r: pre(r, (n, n)), cval, lam, a.iters))
    print(f"CNN: {a.steps} steps", flush=True)
    pre, info = learned.train_unrolled(op, pre, n, lam=lam, iters=a.iters, steps=a.steps,
                                       batch=a.batch, lr=a.lr, seed=a.seed,
                                       log_every=max(1, a.steps // 20), progress=prog)
    with torch.no_grad():
        l1 = float(learned.unrolled_pcg_loss(op, lambda r: pre(r, (n, n)), cval, lam, a.iters))
    print(f"held-out unrolled loss (mean log10 residual over {a.iters} its): analytic "
          f"circulant {la:.4f}; learned filter {lb:.4f}; network before {l0:.4f} -> after "
          f"{l1:.4f}", flush=True)
    out = Path(a.out) if a.out else default_weights(a.workload)
    out.parent.mkdir(parents=True, exist_ok=True)
    learned.save_preconditioner(out, pre, {
        "tau": tau, "pad": PAD_FACTOR * n, "steps": a.steps, "iters": a.iters, "lr": a.lr,
        "batch": a.batch, "seed": a.seed, "lam": lam, "train_seconds": info["seconds"],
        "filter_steps": a.filter_steps, "filter_lr": a.filter_lr,
        "filter_seconds": finfo["seconds"], "loss_analytic": la,
        "loss_base": lb, "loss_start": l0, "loss_learned": l1})
    print(f"weights -> {out}", flush=True)
    if a.no_eval:
        return 0
    t = time.perf_counter()
    res = learned.evaluate(op, {"circulant": analytic, "learned_filter": base,
                                "learned": lambda r: pre(r, (n, n))}, n,
                           scenes=a.eval_scenes, lam=lam)
    summary = {k: v for k, v in res.items() if k != "rows"}
    summary.update({"workload": a.workload, "tau": tau, "train_seconds": info["seconds"],
                    "filter_seconds": finfo["seconds"], "loss_analytic": la,
                    "loss_base": lb, "loss_start": l0, "loss_learned": l1,
                    "eval_seconds": time.perf_counter() - t})
    print(json.dumps(summary), flush=True)
    if a.summary:
        Path(a.summary).write_text(json.dumps(summary, indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main())
