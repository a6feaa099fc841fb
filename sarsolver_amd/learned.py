"""Learned preconditioner for CG on the SAR normal equations.

System: (Â^H Â + lam I) x = Â^H y with Â = A / ||A||_2 (the operator normalised
by its power-iteration norm, so that lam = 1e-3 is a relative shift).

Network: six 3x3 convolutions, 32 channels, 2 (real, imag) channels in and
out, ReLU between layers; bfloat16 weights and activations, FP32 accumulation,
on the hand-written gfx950 kernels of csrc/conv_kernels.hip (matrix-core
forward and data gradient, vector-ALU weight gradient).  FP32 master weights
are kept for the optimiser.  The preconditioner is

    M(r) = r + s * net(r / s),      s = rms(r),

so it is scale-equivariant (CG residuals shrink by orders of magnitude) and is
the identity (plain CG) when the last layer is zero.

Training pairs come from the operator itself: random scenes x (point
scatterers plus Gaussian-smoothed clutter), r = K x with K = Â^H Â + lam I, and
the loss is the K-energy norm of the error of M as an approximate inverse,
    L = e^H K e / x^H K x,   e = M(r) - x,
computed as ||Â e||^2 + lam ||e||^2, so its gradient flows through the
project's own forward and backprojection kernels (torch_ops.sar_forward).
"""
from __future__ import annotations

import ctypes
import json
import math
import time
from dataclasses import dataclass
from typing import List, Optional

import numpy as np
import torch

from ._native import SarPlan, check, load_library
from .solvers import cgls, fpcg, true_residual
from .torch_ops import sar_adjoint, sar_forward

C = 32  # activation channels


def _lib():
    L = load_library()
    if not hasattr(L, "_conv_bound"):
        vp = ctypes.c_void_p
        L.sarhip_conv3x3_bf16.argtypes = [vp, vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, vp]
        L.sarhip_conv3x3_bf16.restype = ctypes.c_int
        L.sarhip_conv3x3_wgrad.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_int, vp, vp, ctypes.c_size_t, vp]
        L.sarhip_conv3x3_wgrad.restype = ctypes.c_int
        L.sarhip_conv3x3_wgrad_workspace.argtypes = [ctypes.c_int] * 4
        L.sarhip_conv3x3_wgrad_workspace.restype = ctypes.c_size_t
        L.sarhip_conv_last_error.restype = ctypes.c_char_p
        L._conv_bound = True
    return L


def _conv_check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed ({rc}): {_lib().sarhip_conv_last_error().decode()}")


def _stream(dev):
    return ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


def _ptr(t: torch.Tensor):
    return ctypes.c_void_p(t.data_ptr())


_frag_index_cache = {}


def pack_weights(w: torch.Tensor) -> torch.Tensor:
    """[9, Cin, Cout] float -> bf16 B-fragments [9][NB][64 lanes][8]: lane l, element i
    holds W[tap][ci = 8 (l // 16) + i][co = 16 nb + l % 16] (zero beyond Cin / Cout)."""
    taps, cin, cout = w.shape
    nb = 2 if cout > 16 else 1
    key = (cin, cout, w.device)
    if key not in _frag_index_cache:
        lane = torch.arange(64, device=w.device)
        i = torch.arange(8, device=w.device)
        ci = (8 * (lane[:, None] // 16) + i[None, :])                      # [64, 8]
        co = torch.stack([16 * b + (lane % 16) for b in range(nb)])       # [nb, 64]
        _frag_index_cache[key] = (ci, co)
    ci, co = _frag_index_cache[key]
    full = torch.zeros(9, C, 32, dtype=torch.float32, device=w.device)
    full[:, :cin, :cout] = w
    frag = full[:, ci[None, :, :].expand(nb, 64, 8), co[:, :, None].expand(nb, 64, 8)]
    return frag.to(torch.bfloat16).contiguous()                           # [9, nb, 64, 8]


def flip_transpose(w: torch.Tensor) -> torch.Tensor:
    """Weights of the data gradient: W'[tap][co][ci] = W[8 - tap][ci][co]."""
    return torch.flip(w, dims=[0]).transpose(1, 2).contiguous()


def conv3x3(x: torch.Tensor, wfrag: torch.Tensor, bias: torch.Tensor, out_c: int, relu: bool,
            out_f32: bool) -> torch.Tensor:
    """y = act(conv3x3(x) + bias): x [B,H,W,32] bf16 -> y [B,H,W,out_c] (bf16 or f32)."""
    B, H, W, c = x.shape
    assert c == C and x.dtype == torch.bfloat16 and x.is_contiguous()
    y = torch.empty((B, H, W, out_c), dtype=torch.float32 if out_f32 else torch.bfloat16,
                    device=x.device)
    rc = _lib().sarhip_conv3x3_bf16(_ptr(x), _ptr(wfrag), _ptr(bias), B, H, W, out_c,
                                    int(relu), int(out_f32), _ptr(y), _stream(x.device))
    _conv_check(rc, "sarhip_conv3x3_bf16")
    return y


def conv3x3_wgrad(x: torch.Tensor, dy: torch.Tensor, out_c: int):
    """(dW [9, 32, co], db [co]) with co = 16 or 32 the padded output width."""
    B, H, W, _ = x.shape
    co = 32 if out_c > 16 else 16
    L = _lib()
    ws = torch.empty(int(L.sarhip_conv3x3_wgrad_workspace(B, H, W, out_c)) // 4 + 1,
                     dtype=torch.float32, device=x.device)
    out = torch.empty(9 * C * co + co, dtype=torch.float32, device=x.device)
    rc = L.sarhip_conv3x3_wgrad(_ptr(x), _ptr(dy), B, H, W, out_c, _ptr(out), _ptr(ws),
                                ws.numel() * 4, _stream(x.device))
    _conv_check(rc, "sarhip_conv3x3_wgrad")
    return out[: 9 * C * co].view(9, C, co), out[9 * C * co:]


def _pad_channels(t: torch.Tensor, c: int) -> torch.Tensor:
    if t.shape[-1] == c:
        return t.contiguous()
    out = torch.zeros(t.shape[:-1] + (c,), dtype=t.dtype, device=t.device)
    out[..., : t.shape[-1]] = t
    return out


class _ConvLayer(torch.autograd.Function):
    """One 3x3 layer on the gfx950 kernels; x is bf16 [B,H,W,32]."""

    @staticmethod
    def forward(ctx, x, w, b, relu: bool, last: bool, need_dx: bool):
        cout = w.shape[2]
        y = conv3x3(x, pack_weights(w), b, cout, relu, out_f32=last)
        ctx.save_for_backward(x, w, y)
        ctx.relu, ctx.last, ctx.need_dx = relu, last, need_dx
        return y

    @staticmethod
    def backward(ctx, gy):
        x, w, y = ctx.saved_tensors
        cin, cout = w.shape[1], w.shape[2]
        g = gy.float()
        if ctx.relu:
            g = g * (y > 0).float()
        co = 32 if cout > 16 else 16
        g_pad = _pad_channels(g, co)
        dw, db = conv3x3_wgrad(x, g_pad, cout)
        dw = dw[:, :cin, :cout].contiguous()
        db = db[:cout].contiguous()
        dx = None
        if ctx.need_dx:
            gin = _pad_channels(g.to(torch.bfloat16), C)
            zero_b = torch.zeros(32, dtype=torch.float32, device=x.device)
            dx = conv3x3(gin, pack_weights(flip_transpose(w)), zero_b, cin, False, False)
            dx = _pad_channels(dx, C)
        return dx, dw, db, None, None, None


class CnnPreconditioner:
    """M(r) = r + s * net(r / s) on complex images [H, W] or [B, H, W]."""

    def __init__(self, device, layers: int = 6, width: int = C, seed: int = 0):
        g = torch.Generator(device="cpu").manual_seed(seed)
        chans = [2] + [width] * (layers - 1) + [2]
        self.weights: List[torch.Tensor] = []
        self.biases: List[torch.Tensor] = []
        for i in range(layers):
            cin, cout = chans[i], chans[i + 1]
            std = math.sqrt(2.0 / (9 * cin))
            wi = torch.randn(9, cin, cout, generator=g) * std
            if i == layers - 1:
                wi.zero_()
            self.weights.append(wi.to(device).requires_grad_(True))
            self.biases.append(torch.zeros(cout, device=device).requires_grad_(True))
        self.device = device

    def parameters(self):
        return self.weights + self.biases

    def _packed(self):
        """B-fragments of every layer, repacked only when a weight has changed."""
        key = tuple(w._version for w in self.weights)
        if getattr(self, "_pack_key", None) != key:
            with torch.no_grad():
                self._pack = [pack_weights(w) for w in self.weights]
                self._bias = [b.detach().contiguous() for b in self.biases]
            self._pack_key = key
        return self._pack, self._bias

    def net(self, z: torch.Tensor) -> torch.Tensor:
        """z complex [B,H,W] -> complex [B,H,W]."""
        x = torch.zeros(z.shape + (C,), dtype=torch.bfloat16, device=z.device)
        x[..., 0] = z.real.to(torch.bfloat16)
        x[..., 1] = z.imag.to(torch.bfloat16)
        n = len(self.weights)
        if torch.is_grad_enabled():
            for i in range(n):
                last = i == n - 1
                x = _ConvLayer.apply(x, self.weights[i], self.biases[i], not last, last, i > 0)
        else:
            packs, biases = self._packed()
            for i in range(n):
                last = i == n - 1
                x = conv3x3(x, packs[i], biases[i], self.weights[i].shape[2], not last, last)
        return torch.complex(x[..., 0], x[..., 1])

    def __call__(self, r: torch.Tensor, shape) -> torch.Tensor:
        """Apply to a flat residual [M] (or [B, M]) of images of `shape` (H, W)."""
        flat = r.dim() == 1
        z = r.reshape((-1,) + tuple(shape))
        s = torch.sqrt(torch.mean(torch.abs(z) ** 2, dim=(1, 2), keepdim=True)).clamp_min(1e-30)
        out = z + s * self.net(z / s)
        return out.reshape(-1) if flat else out.reshape(r.shape)

    def state_dict(self):
        return {f"w{i}": w.detach().cpu() for i, w in enumerate(self.weights)} | \
               {f"b{i}": b.detach().cpu() for i, b in enumerate(self.biases)}

    def load_state_dict(self, sd):
        with torch.no_grad():
            for i, w in enumerate(self.weights):
                w.copy_(sd[f"w{i}"].to(w.device))
            for i, b in enumerate(self.biases):
                b.copy_(sd[f"b{i}"].to(b.device))


class NormalisedOperator:
    """Â = A / ||A||_2 on device vectors (||A||_2 by power iteration on A^H A)."""

    def __init__(self, plan: SarPlan, iters: int = 30, seed: int = 0):
        self.plan = plan
        g = torch.Generator(device=plan.device).manual_seed(seed)
        v = torch.randn(plan.num_scatterers, dtype=torch.complex64, device=plan.device,
                        generator=g)
        lam = 1.0
        with torch.no_grad():
            for _ in range(iters):
                v = plan.adjoint(plan.forward(v))
                lam = float(torch.linalg.vector_norm(v))
                v = v / lam
        self.norm = math.sqrt(lam)
        self.scale = 1.0 / self.norm

    def forward(self, x):
        return sar_forward(x, self.plan) * self.scale

    def adjoint(self, y):
        return sar_adjoint(y, self.plan) * self.scale


def random_scenes(side: int, batch: int, gen: "Gens", device) -> torch.Tensor:
    """[batch, side*side] complex64: point scatterers + Gaussian-smoothed complex clutter
    with random weights (generated on the device)."""
    out = []
    k = torch.fft.fftfreq(side, device=device)
    for _ in range(batch):
        sigma = float(torch.empty(1).uniform_(1.0, 4.0, generator=gen.cpu_gen))
        f = torch.complex(torch.randn(side, side, device=device, generator=gen.dev_gen),
                          torch.randn(side, side, device=device, generator=gen.dev_gen))
        g = torch.exp(-2.0 * (math.pi * sigma) ** 2 * (k[:, None] ** 2 + k[None, :] ** 2))
        clutter = torch.fft.ifft2(torch.fft.fft2(f) * g)
        clutter = clutter / torch.sqrt(torch.mean(torch.abs(clutter) ** 2))
        pts = torch.zeros(side * side, dtype=torch.complex64, device=device)
        npts = int(torch.randint(10, 200, (1,), generator=gen.cpu_gen))
        idx = torch.randint(0, side * side, (npts,), device=device, generator=gen.dev_gen)
        amp = torch.complex(torch.randn(npts, device=device, generator=gen.dev_gen),
                            torch.randn(npts, device=device, generator=gen.dev_gen)) * 3.0
        pts[idx] = amp
        wc = float(torch.empty(1).uniform_(0.0, 1.0, generator=gen.cpu_gen))
        out.append((wc * clutter.reshape(-1) + pts).to(torch.complex64))
    return torch.stack(out)


@dataclass
class Gens:
    cpu_gen: torch.Generator
    dev_gen: torch.Generator


def make_gens(seed: int, device) -> Gens:
    return Gens(torch.Generator(device="cpu").manual_seed(seed),
                torch.Generator(device=device).manual_seed(seed))


def train(op: NormalisedOperator, side: int, lam: float = 1e-3, steps: int = 2000,
          batch: int = 4, lr: float = 1e-3, seed: int = 1, log_every: int = 100,
          pre: Optional[CnnPreconditioner] = None, progress=None):
    """Train M on `steps * batch` random scenes (each used once)."""
    dev = op.plan.device
    pre = pre or CnnPreconditioner(dev, seed=seed)
    opt = torch.optim.Adam(pre.parameters(), lr=lr)
    sched = torch.optim.lr_scheduler.CosineAnnealingLR(opt, steps)
    gens = make_gens(seed, dev)
    history = []
    t0 = time.perf_counter()
    for step in range(steps):
        x = random_scenes(side, batch, gens, dev)
        with torch.no_grad():
            r = torch.stack([op.adjoint(op.forward(x[b])) + lam * x[b] for b in range(batch)])
            den = torch.stack([torch.vdot(x[b], r[b]).real for b in range(batch)]).clamp_min(1e-30)
        e = pre(r, (side, side)) - x
        loss = 0.0
        for b in range(batch):
            Ae = op.forward(e[b])
            num = torch.sum(torch.abs(Ae) ** 2) + lam * torch.sum(torch.abs(e[b]) ** 2)
            loss = loss + num / den[b]
        loss = loss / batch
        opt.zero_grad()
        loss.backward()
        opt.step()
        sched.step()
        if step % log_every == 0 or step == steps - 1:
            lv = float(loss.detach())
            history.append((step, lv))
            if progress is not None:
                progress(step, lv, time.perf_counter() - t0)
    torch.cuda.synchronize(dev)
    return pre, {"history": history, "seconds": time.perf_counter() - t0,
                 "scenes": steps * batch}


def _kapply(op, v, lam):
    return op.adjoint(op.forward(v)) + lam * v


def pcg_iterates(op, c: torch.Tensor, lam: float, precond=None, iters: int = 30,
                 tol: float = 0.0):
    """Iterates x_0 = 0, x_1, ... of (flexible, Polak-Ribiere) PCG on K x = c with
    K = Â^H Â + lam I; plain CG when precond is None.  Stops after `iters`
    steps or at relative residual <= tol.  Returns (iterates, final x)."""
    x = torch.zeros_like(c)
    r = c.clone()
    z = r if precond is None else precond(r)
    p = z.clone()
    rz = torch.vdot(r, z)
    cn = float(torch.linalg.vector_norm(c))
    xs = [x]
    for _ in range(iters):
        Kp = _kapply(op, p, lam)
        alpha = rz / torch.vdot(p, Kp)
        x = x + alpha * p
        r_new = r - alpha * Kp
        xs.append(x)
        if tol and float(torch.linalg.vector_norm(r_new)) <= tol * cn:
            break
        z_new = r_new if precond is None else precond(r_new)
        beta = torch.vdot(r_new - r, z_new) / rz
        rz = torch.vdot(r_new, z_new)
        p = z_new + beta * p
        r = r_new
    return xs, x


def noisy_problem(op, x_true: torch.Tensor, snr_db: float, gen: torch.Generator):
    """c = Â^H y with y = Â x_true + complex Gaussian noise at `snr_db`."""
    y = op.forward(x_true)
    noise = torch.complex(torch.randn(y.shape, device=y.device, generator=gen),
                          torch.randn(y.shape, device=y.device, generator=gen))
    noise = noise * torch.sqrt(torch.mean(torch.abs(y) ** 2) / 2 / 10 ** (snr_db / 10))
    return (y + noise.to(torch.complex64)).contiguous()


def trajectory_pairs(op, side: int, scenes: int, lam: float = 1e-3, snr_db: float = 20.0,
                     iters: int = 30, seed: int = 7, precond=None):
    """Training pairs from solver trajectories: for each random scene, the exact
    solution x* of K x = Â^H y (CG to 1e-5), then for every iterate x_k of a
    solve (plain CG, or PCG with `precond`) the error e_k = x* - x_k and its
    residual r_k = K e_k.  Returns (R, E) stacked [n, side*side] complex64."""
    dev = op.plan.device
    gens = make_gens(seed, dev)
    R, E = [], []
    with torch.no_grad():
        for _ in range(scenes):
            x_true = random_scenes(side, 1, gens, dev)[0]
            c = op.adjoint(noisy_problem(op, x_true, snr_db, gens.dev_gen))
            _, x_star = pcg_iterates(op, c, lam, None, iters=400, tol=1e-5)
            xs, _ = pcg_iterates(op, c, lam, precond, iters=iters, tol=1e-4)
            for xk in xs:
                e = x_star - xk
                R.append(_kapply(op, e, lam))
                E.append(e)
    return torch.stack(R), torch.stack(E)


def train_pairs(op: NormalisedOperator, pre: CnnPreconditioner, R: torch.Tensor,
                E: torch.Tensor, side: int, lam: float = 1e-3, steps: int = 1000,
                batch: int = 8, lr: float = 1e-3, seed: int = 3, log_every: int = 100,
                progress=None):
    """Minimise the K-energy error of M as an approximate inverse on (r, e) pairs:
    mean over the batch of (||Â d||^2 + lam ||d||^2) / (e^H r), d = M(r) - e.
    Adam on FP32 master weights, cosine learning-rate decay."""
    dev = op.plan.device
    opt = torch.optim.Adam(pre.parameters(), lr=lr)
    sched = torch.optim.lr_scheduler.CosineAnnealingLR(opt, steps)
    g = torch.Generator(device="cpu").manual_seed(seed)
    den_all = torch.stack([torch.vdot(E[i], R[i]).real for i in range(E.shape[0])])
    history = []
    t0 = time.perf_counter()
    for step in range(steps):
        idx = torch.randint(0, R.shape[0], (batch,), generator=g).to(dev)
        r, e, den = R[idx], E[idx], den_all[idx].clamp_min(1e-30)
        d = pre(r, (side, side)) - e
        loss = 0.0
        for b in range(batch):
            Ad = op.forward(d[b])
            loss = loss + (torch.sum(torch.abs(Ad) ** 2)
                           + lam * torch.sum(torch.abs(d[b]) ** 2)) / den[b]
        loss = loss / batch
        opt.zero_grad()
        loss.backward()
        opt.step()
        sched.step()
        if step % log_every == 0 or step == steps - 1:
            lv = float(loss.detach())
            history.append((step, lv))
            if progress is not None:
                progress(step, lv, time.perf_counter() - t0)
    torch.cuda.synchronize(dev)
    return pre, {"history": history, "seconds": time.perf_counter() - t0}


def evaluate(op: NormalisedOperator, pre: CnnPreconditioner, side: int, scenes: int = 64,
             lam: float = 1e-3, tol: float = 1e-3, snr_db: float = 20.0, seed: int = 12345,
             max_iter: int = 300):
    """Solve (Â^H Â + lam I) x = Â^H y on held-out scenes with plain CGLS and with
    the learned flexible PCG; every result is judged by the true residual."""
    dev = op.plan.device
    gens = make_gens(seed, dev)
    rows = []
    shape = (side, side)
    with torch.no_grad():
        for i in range(scenes):
            x = random_scenes(side, 1, gens, dev)[0]
            y = op.forward(x)
            noise = torch.complex(torch.randn(y.shape, device=dev, generator=gens.dev_gen),
                                  torch.randn(y.shape, device=dev, generator=gens.dev_gen))
            noise = noise * torch.sqrt(torch.mean(torch.abs(y) ** 2) / 2 / 10 ** (snr_db / 10))
            y = (y + noise.to(torch.complex64)).contiguous()
            torch.cuda.synchronize(dev)
            rc = cgls(op, y, lam=lam, tol=tol, max_iter=max_iter)
            rp = fpcg(op, y, lambda r: pre(r, shape), lam=lam, tol=tol, max_iter=max_iter)
            rows.append({"cgls_iters": rc.iterations, "cgls_s": rc.seconds,
                         "cgls_true": true_residual(op, rc.x, y, lam),
                         "pcg_iters": rp.iterations, "pcg_s": rp.seconds,
                         "pcg_true": true_residual(op, rp.x, y, lam),
                         "pcg_fallback": rp.fallback})
    med = lambda k: float(np.median([r[k] for r in rows]))
    return {"scenes": scenes, "median_cgls_iters": med("cgls_iters"),
            "median_pcg_iters": med("pcg_iters"), "median_cgls_ms": med("cgls_s") * 1e3,
            "median_pcg_ms": med("pcg_s") * 1e3,
            "time_ratio_pcg_over_cgls": med("pcg_s") / med("cgls_s"),
            "max_true_residual_cgls": max(r["cgls_true"] for r in rows),
            "max_true_residual_pcg": max(r["pcg_true"] for r in rows),
            "fallbacks": sum(r["pcg_fallback"] is not None for r in rows), "rows": rows}
