"""Learned preconditioner for CG on the SAR normal equations.

System: (Â^H Â + lam I) x = Â^H y with Â = A / ||A||_2 (the operator normalised
by its power-iteration norm, so that lam = 1e-3 is a relative shift), solved
to relative residual 1e-3.

Preconditioner: M(r) = D [ z + s net(z / s) ],  z = P D^H r,  s = rms(z).

* D = diag(exp(i phi)) removes the carrier phase phi of the middle pulse.
  In these demodulated coordinates the normal operator is close to
  shift-invariant (in pixel coordinates it is not: the near-field look
  direction moves the k-space band across the scene).
* P is the circulant inverse filter 1 / (sigma(k) + tau) of the point-spread
  function measured through the operator itself, on a zero-padded grid
  (`CirculantPreconditioner`), with tau chosen on training problems.
* net: six 3x3 convolutions, 32 channels, 2 (real, imag) channels in and
  out, ReLU between layers; bfloat16 weights and activations, FP32
  accumulation, on the hand-written gfx950 kernels of csrc/conv_kernels.hip
  (matrix-core forward and data gradient, vector-ALU weight gradient); FP32
  master weights for Adam.  Its last layer starts at zero, so the untrained
  M is exactly D P D^H; s makes M scale-equivariant.

Training (`train_unrolled`): backpropagation through `iters` unrolled
flexible-PCG iterations on fresh random problems, minimising the mean log10
residual; the operator's gradient flows through the project's own forward
and backprojection kernels (torch_ops).  Losses that score M one step at a
time (approximate-inverse error, K-angle of the search direction) improved
that score while making PCG converge more slowly, because a nonlinear M
breaks conjugacy between steps; the unrolled loss judges M by the
convergence it produces.
"""
from __future__ import annotations

import ctypes
import math
import time
from dataclasses import dataclass
from pathlib import Path
from typing import Dict, List, Optional

import numpy as np
import torch

from ._native import SarPlan, load_library
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
        L.sarhip_precond_net_workspace.argtypes = [ctypes.c_int] * 3
        L.sarhip_precond_net_workspace.restype = ctypes.c_size_t
        L.sarhip_precond_net.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int, vp, vp, vp, vp, ctypes.c_size_t, vp, vp]
        L.sarhip_precond_net.restype = ctypes.c_int
        L.sarhip_precond_net_noalloc.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                 ctypes.c_int, vp, vp, vp, vp, vp]
        L.sarhip_precond_net_noalloc.restype = ctypes.c_int
        L.sarhip_spectral_filter_workspace.argtypes = [ctypes.c_int] * 3
        L.sarhip_spectral_filter_workspace.restype = ctypes.c_size_t
        L.sarhip_spectral_filter.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp,
                                             vp, ctypes.c_size_t, vp, vp, vp]
        L.sarhip_spectral_filter.restype = ctypes.c_int
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

    def __init__(self, device, layers: int = 6, width: int = C, seed: int = 0,
                 base: Optional["CirculantPreconditioner"] = None):
        self.base = base
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

    def _fused(self) -> bool:
        widths = [w.shape[2] for w in self.weights]
        return all(c == C for c in widths[:-1]) and widths[-1] == 2

    def _infer(self, z: torch.Tensor, demod: Optional[torch.Tensor]) -> torch.Tensor:
        """Inference in one library call (sarhip_precond_net): z [B, H, W]
        complex64 -> (z + s net(z / s)) * demod."""
        B, H, W = z.shape
        L = _lib()
        packs, biases = self._packed()
        key = (B, H, W, z.device)
        if getattr(self, "_work_key", None) != key:
            nbytes = int(L.sarhip_precond_net_workspace(B, H, W))
            self._work = torch.empty(nbytes, dtype=torch.uint8, device=z.device)
            self._work_key = key
        n = len(packs)
        vp = ctypes.c_void_p
        wf = (vp * n)(*[p.data_ptr() for p in packs])
        bs = (vp * n)(*[b.data_ptr() for b in biases])
        oc = (ctypes.c_int * n)(*[w.shape[2] for w in self.weights])
        out = torch.empty_like(z)
        rc = L.sarhip_precond_net(_ptr(z), None if demod is None else _ptr(demod), B, H, W, n,
                                  ctypes.cast(wf, vp), ctypes.cast(bs, vp), ctypes.cast(oc, vp),
                                  _ptr(self._work), self._work.numel(), _ptr(out),
                                  _stream(z.device))
        _conv_check(rc, "sarhip_precond_net")
        return out

    def __call__(self, r: torch.Tensor, shape) -> torch.Tensor:
        """Apply to a flat residual [M] (or [B, M]) of images of `shape` (H, W).
        With a base preconditioner, the network refines its output in the
        base's demodulated coordinates.  Without gradients the network runs as
        one library call (scale, bf16 packing, the layers, output and phasor)."""
        flat = r.dim() == 1
        b = self.base
        zb = r if b is None else b.base_coords(r)
        z = zb.reshape((-1,) + tuple(shape))
        if not torch.is_grad_enabled() and self._fused():
            demod = None if b is None else b.demod
            out = self._infer(z.to(torch.complex64).contiguous(), demod)
            return out.reshape(-1) if flat else out.reshape(r.shape)
        s = torch.sqrt(torch.mean(torch.abs(z) ** 2, dim=(1, 2), keepdim=True)).clamp_min(1e-30)
        out = (z + s * self.net(z / s)).reshape(r.shape)
        if b is not None:
            out = b.from_base(out)
        return out.reshape(-1) if flat else out

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
        self.device = plan.device
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


def carrier_phase(plan: SarPlan, pulse: Optional[int] = None) -> np.ndarray:
    """k_c dR_c(x) for every pixel: the phase of the centre-frequency echo of
    one pulse (default: the middle one), dR_c = |Tx - x| + |Rx - x| minus the
    same for the stabilisation reference point (FP64, reduced mod 2 pi)."""
    g = plan.geometry
    p = plan.num_slow_times // 2 if pulse is None else pulse
    tx, rx, srp, pos = g["tx"][p], g["rx"][p], g["srp"][p], g["pos"]
    dr = (np.linalg.norm(pos - tx, axis=1) + np.linalg.norm(pos - rx, axis=1)
          - np.linalg.norm(srp - tx) - np.linalg.norm(srp - rx))
    return np.mod(plan.centre_wavenumber * dr, 2.0 * np.pi)


class CirculantPreconditioner:
    """P r = D F^H [ F(D^H r) / (sigma(k) + tau) ] (D^H r zero-padded to
    `pad` x `pad`): the inverse filter of the operator's point-spread function
    in demodulated coordinates.

    D = diag(exp(i s phi)) removes the carrier phase phi of the middle pulse
    (`carrier_phase`); the sign s is the one that moves the PSF's band to
    zero frequency.  In those coordinates the normal operator D^H K D (same
    spectrum as K = Â^H Â) is close to shift-invariant, whereas K itself is not
    (the near-field look direction moves the band across the scene).
    sigma(k) is the spectrum of D^H K D delta for a delta at the image centre,
    re-centred at the origin.  P is Hermitian positive definite; tau >= lam
    regularises the band edges and out-of-band frequencies.

    Learned spectral gain (optional): the filter becomes
    exp(g(k)) / (sigma(k) + tau) with g a real log-gain held on a (pad/2)^2
    grid and interpolated bilinearly to the padded grid; exp keeps P
    Hermitian positive definite.  g is trained by backpropagating through
    unrolled PCG (`train_filter`); it corrects what a single shift-invariant
    filter gets wrong -- the band edges move across the scene (the look
    direction changes) and the image borders truncate the point-spread
    function -- as well as one filter can, at no cost per application."""

    def __init__(self, op, side: int, tau: float, pad: Optional[int] = None,
                 demodulate: bool = True, phase: Optional[np.ndarray] = None):
        dev = op.device
        self.side, self.pad, self.tau = side, pad or side, tau
        n, m = side, self.pad
        with torch.no_grad():
            if demodulate:
                if phase is None:
                    phase = carrier_phase(op.plan)
                phi = torch.from_numpy(np.asarray(phase, dtype=np.float64)).to(dev)
                best = None
                for sgn in (1.0, -1.0):
                    d = torch.polar(torch.ones_like(phi), sgn * phi).to(torch.complex64)
                    h = self._psf(op, d, n)
                    spec = torch.abs(torch.fft.fft2(h)) ** 2
                    f = torch.fft.fftfreq(n, device=dev)
                    # mean squared distance of the band from zero frequency
                    spread = float((spec * (f[:, None] ** 2 + f[None, :] ** 2)).sum() / spec.sum())
                    if best is None or spread < best[0]:
                        best = (spread, d, h)
                self.demod = best[1]
                h = best[2]
            else:
                self.demod = None
                h = self._psf(op, None, n)
            hp = torch.zeros(m, m, dtype=torch.complex64, device=dev)
            hp[:n, :n] = h
            hp = torch.roll(hp, shifts=(-(n // 2), -(n // 2)), dims=(0, 1))
            sig = torch.fft.fft2(hp).real.clamp_min(0.0)
            self.base_filter = (1.0 / (sig + tau)).to(torch.float32)
            self.sigma = sig
        self.log_gain: Optional[torch.Tensor] = None  # [pad/2, pad/2] or None
        self._filter_cache = None

    @staticmethod
    def _psf(op, d, n):
        delta = torch.zeros(n * n, dtype=torch.complex64, device=op.device)
        delta[(n // 2) * n + n // 2] = 1.0
        if d is not None:
            delta = delta * d
        h = op.adjoint(op.forward(delta))
        if d is not None:
            h = h * d.conj()
        return h.reshape(n, n)

    # -- learned spectral gain ------------------------------------------------
    def learn_gain(self) -> torch.Tensor:
        """Make the log-gain a trainable tensor (zeros: P unchanged) and return it."""
        if self.log_gain is None:
            h = (self.pad + 1) // 2
            self.log_gain = torch.zeros(h, h, dtype=torch.float32, device=self.base_filter.device)
        self.log_gain.requires_grad_(True)
        self._filter_cache = None
        return self.log_gain

    def set_gain(self, log_gain) -> None:
        """Install a (pad/2)^2 log-gain (numpy or tensor), frozen."""
        t = torch.as_tensor(np.asarray(log_gain, dtype=np.float32)).to(self.base_filter.device)
        h = (self.pad + 1) // 2
        if tuple(t.shape) != (h, h):
            raise ValueError(f"log-gain shape {tuple(t.shape)}, expected {(h, h)}")
        self.log_gain = t
        self._filter_cache = None

    def _gain(self) -> torch.Tensor:
        g = self.log_gain[None, None]
        m = self.pad
        return torch.nn.functional.interpolate(g, size=(m, m), mode="bilinear",
                                               align_corners=False)[0, 0]

    @property
    def filter(self) -> torch.Tensor:
        """The spectral filter on the padded grid, [pad, pad] float32."""
        if self.log_gain is None:
            return self.base_filter
        if self.log_gain.requires_grad and torch.is_grad_enabled():
            return self.base_filter * torch.exp(self._gain())
        key = self.log_gain._version
        if self._filter_cache is None or self._filter_cache[0] != key:
            with torch.no_grad():
                self._filter_cache = (key, (self.base_filter * torch.exp(self._gain())).contiguous())
        return self._filter_cache[1]

    # -- application ----------------------------------------------------------
    def _fused(self, r: torch.Tensor, phasor_out: bool) -> Optional[torch.Tensor]:
        """The whole filtering in one library call (sarhip_spectral_filter:
        demodulation, zero padding, 2D FFT, filter, inverse FFT, crop and the
        optional output phasor) when no gradient is needed; None otherwise."""
        if torch.is_grad_enabled() or not r.is_cuda or r.dtype != torch.complex64:
            return None
        n, m = self.side, self.pad
        if not (m >= 64 and m <= 1024 and (m & (m - 1)) == 0 and n <= m):
            return None
        z = r.reshape(-1, n, n).contiguous()
        B = z.shape[0]
        L = _lib()
        key = (B, n, m, z.device)
        if getattr(self, "_sf_key", None) != key:
            self._sf_work = torch.empty(int(L.sarhip_spectral_filter_workspace(B, n, m)),
                                        dtype=torch.uint8, device=z.device)
            self._sf_key = key
        f = self.filter.contiguous()
        out = torch.empty_like(z)
        d = None if self.demod is None else _ptr(self.demod)
        rc = L.sarhip_spectral_filter(_ptr(z), d, B, n, m, _ptr(f), _ptr(self._sf_work),
                                      self._sf_work.numel(), _ptr(out), d if phasor_out else None,
                                      _stream(z.device))
        _conv_check(rc, "sarhip_spectral_filter")
        return out.reshape(r.shape)

    def base_coords(self, r: torch.Tensor) -> torch.Tensor:
        """P's output before the final phasor: apply_base(to_base(r))."""
        out = self._fused(r, phasor_out=False)
        return out if out is not None else self.apply_base(self.to_base(r))

    def to_base(self, r: torch.Tensor) -> torch.Tensor:
        return r if self.demod is None else r * self.demod.conj()

    def from_base(self, z: torch.Tensor) -> torch.Tensor:
        return z if self.demod is None else z * self.demod

    def apply_base(self, z: torch.Tensor) -> torch.Tensor:
        """Circulant inverse in demodulated coordinates: [..., M] -> [..., M]."""
        n, m = self.side, self.pad
        zz = z.reshape((-1, n, n))
        if m != n:
            zp = torch.zeros((zz.shape[0], m, m), dtype=zz.dtype, device=zz.device)
            zp[:, :n, :n] = zz
        else:
            zp = zz
        out = torch.fft.ifft2(torch.fft.fft2(zp) * self.filter)[:, :n, :n]
        return out.reshape(z.shape).to(z.dtype)

    def __call__(self, r: torch.Tensor, shape=None) -> torch.Tensor:
        out = self._fused(r, phasor_out=True)
        return out if out is not None else self.from_base(self.apply_base(self.to_base(r)))


@dataclass
class Gens:
    cpu_gen: torch.Generator
    dev_gen: torch.Generator


def make_gens(seed: int, device) -> Gens:
    return Gens(torch.Generator(device="cpu").manual_seed(seed),
                torch.Generator(device=device).manual_seed(seed))


def random_scenes(side: int, batch: int, gen: Gens, device) -> torch.Tensor:
    """[batch, side*side] complex64: 10-200 point scatterers (complex Gaussian
    amplitudes, std 3) plus complex Gaussian clutter smoothed by a Gaussian of
    1-4 px and weighted by U(0, 1), generated on the device."""
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


def noisy_problem(op, x_true: torch.Tensor, snr_db: float, gen: torch.Generator):
    """y = Â x_true + complex Gaussian noise at `snr_db` (per-sample power ratio)."""
    y = op.forward(x_true)
    noise = torch.complex(torch.randn(y.shape, device=y.device, generator=gen),
                          torch.randn(y.shape, device=y.device, generator=gen))
    noise = noise * torch.sqrt(torch.mean(torch.abs(y) ** 2) / 2 / 10 ** (snr_db / 10))
    return (y + noise.to(torch.complex64)).contiguous()


def problem_batch(op, side: int, count: int, gens: Gens, snr_db: float = 20.0) -> torch.Tensor:
    """[count, M] right-hand sides c = Â^H y of random noisy scenes."""
    with torch.no_grad():
        xs = random_scenes(side, count, gens, op.device)
        return torch.stack([op.adjoint(noisy_problem(op, xs[b], snr_db, gens.dev_gen))
                            for b in range(count)])


def unrolled_pcg_loss(op, precond, c: torch.Tensor, lam: float, iters: int) -> torch.Tensor:
    """Mean over t = 1..iters of log10(||r_t|| / ||c||) for the flexible PCG
    recursion (Polak-Ribiere, as solvers.fpcg) on K x = c, K = Â^H Â + lam I,
    differentiable end to end.  c: [B, M]; returns the batch mean."""
    cn = torch.linalg.vector_norm(c, dim=1)
    r = c
    z = precond(r)
    p = z
    rz = torch.sum(r.conj() * z, dim=1)
    total = 0.0
    for _ in range(iters):
        Kp = torch.stack([op.adjoint(op.forward(p[b])) for b in range(p.shape[0])]) + lam * p
        alpha = rz / torch.sum(p.conj() * Kp, dim=1)
        r_new = r - alpha[:, None] * Kp
        total = total + torch.log10(torch.linalg.vector_norm(r_new, dim=1) / cn)
        z_new = precond(r_new)
        beta = torch.sum((r_new - r).conj() * z_new, dim=1) / rz
        rz = torch.sum(r_new.conj() * z_new, dim=1)
        p = z_new + beta[:, None] * p
        r = r_new
    return torch.mean(total) / iters


def train_unrolled(op, pre: CnnPreconditioner, side: int, lam: float = 1e-3, iters: int = 12,
                   steps: int = 200, batch: int = 2, lr: float = 1e-4, seed: int = 5,
                   clip: float = 1.0, log_every: int = 25, progress=None):
    """Train M by backpropagating through `iters` unrolled PCG iterations on
    fresh random problems (loss `unrolled_pcg_loss`), Adam with cosine decay
    and gradient-norm clipping."""
    opt = torch.optim.Adam(pre.parameters(), lr=lr)
    sched = torch.optim.lr_scheduler.CosineAnnealingLR(opt, steps)
    gens = make_gens(seed, op.device)
    shape = (side, side)
    history = []
    t0 = time.perf_counter()
    for step in range(steps):
        c = problem_batch(op, side, batch, gens)
        loss = unrolled_pcg_loss(op, lambda r: pre(r, shape), c, lam, iters)
        opt.zero_grad()
        loss.backward()
        torch.nn.utils.clip_grad_norm_(pre.parameters(), clip)
        opt.step()
        sched.step()
        if step % log_every == 0 or step == steps - 1:
            lv = float(loss.detach())
            history.append((step, lv))
            if progress is not None:
                progress(step, lv, time.perf_counter() - t0)
    if op.device.type == "cuda":
        torch.cuda.synchronize(op.device)
    return pre, {"history": history, "seconds": time.perf_counter() - t0}


def train_filter(op, base: CirculantPreconditioner, side: int, lam: float = 1e-3,
                 iters: int = 12, steps: int = 1000, batch: int = 2, lr: float = 0.02,
                 seed: int = 5, log_every: int = 100, progress=None):
    """Train the base's spectral log-gain through `iters` unrolled PCG
    iterations on fresh random problems (loss `unrolled_pcg_loss` with the
    base alone as M), Adam with cosine decay."""
    g = base.learn_gain()
    opt = torch.optim.Adam([g], lr=lr)
    sched = torch.optim.lr_scheduler.CosineAnnealingLR(opt, steps)
    gens = make_gens(seed, op.device)
    history = []
    t0 = time.perf_counter()
    for step in range(steps):
        c = problem_batch(op, side, batch, gens)
        loss = unrolled_pcg_loss(op, base, c, lam, iters)
        opt.zero_grad()
        loss.backward()
        opt.step()
        sched.step()
        if step % log_every == 0 or step == steps - 1:
            lv = float(loss.detach())
            history.append((step, lv))
            if progress is not None:
                progress(step, lv, time.perf_counter() - t0)
    g.requires_grad_(False)
    base._filter_cache = None
    if op.device.type == "cuda":
        torch.cuda.synchronize(op.device)
    return {"history": history, "seconds": time.perf_counter() - t0}


def choose_tau(op, side: int, taus, problems, lam: float = 1e-3, tol: float = 1e-3,
               pad: Optional[int] = None, max_iter: int = 300):
    """The tau of `CirculantPreconditioner` with the fewest median PCG
    iterations on `problems` (a list of y); returns (tau, {tau: median})."""
    scores = {}
    with torch.no_grad():
        for tau in taus:
            P = CirculantPreconditioner(op, side, tau, pad=pad)
            its = [fpcg(op, y, P, lam=lam, tol=tol, max_iter=max_iter).iterations
                   for y in problems]
            scores[tau] = float(np.median(its))
    best = min(scores, key=lambda t: (scores[t], -t))
    return best, scores


def evaluate(op, preconds: Dict[str, object], side: int, scenes: int = 64,
             lam: float = 1e-3, tol: float = 1e-3, snr_db: float = 20.0, seed: int = 12345,
             max_iter: int = 300):
    """Solve (Â^H Â + lam I) x = Â^H y on `scenes` held-out noisy scenes with
    plain CGLS and with PCG under each named preconditioner (callables on
    flat residuals).  Every solve is scored by its recomputed true residual.
    Returns medians of iterations and wall time (ms) per solver, the ratio of
    each PCG's median time to CGLS's, and the per-scene rows."""
    dev = op.device
    gens = make_gens(seed, dev)
    rows = []
    with torch.no_grad():
        for _ in range(scenes):
            x = random_scenes(side, 1, gens, dev)[0]
            y = noisy_problem(op, x, snr_db, gens.dev_gen)
            if dev.type == "cuda":
                torch.cuda.synchronize(dev)
            rc = cgls(op, y, lam=lam, tol=tol, max_iter=max_iter)
            row = {"cgls": {"iters": rc.iterations, "ms": rc.seconds * 1e3,
                            "true_residual": true_residual(op, rc.x, y, lam),
                            "fallback": None}}
            for name, M in preconds.items():
                rp = fpcg(op, y, M, lam=lam, tol=tol, max_iter=max_iter)
                row[name] = {"iters": rp.iterations, "ms": rp.seconds * 1e3,
                             "true_residual": true_residual(op, rp.x, y, lam),
                             "fallback": rp.fallback}
            rows.append(row)
    out = {"scenes": scenes, "tol": tol, "lam": lam, "snr_db": snr_db, "solvers": {}}
    for name in ["cgls", *preconds]:
        its = [r[name]["iters"] for r in rows]
        ms = [r[name]["ms"] for r in rows]
        out["solvers"][name] = {
            "median_iters": float(np.median(its)), "median_ms": float(np.median(ms)),
            "max_true_residual": max(r[name]["true_residual"] for r in rows),
            "fallbacks": sum(r[name]["fallback"] is not None for r in rows)}
    base_ms = out["solvers"]["cgls"]["median_ms"]
    for name in preconds:
        out["solvers"][name]["time_ratio_vs_cgls"] = out["solvers"][name]["median_ms"] / base_ms
    out["rows"] = rows
    return out


def save_preconditioner(path, pre: CnnPreconditioner, meta: dict) -> None:
    """Network weights (FP32 master copies), the base's learned spectral
    log-gain (float16, when trained) and the base's settings as .npz."""
    arrays = {k: v.numpy() for k, v in pre.state_dict().items()}
    if pre.base is not None and pre.base.log_gain is not None:
        arrays["log_gain"] = pre.base.log_gain.detach().cpu().numpy().astype(np.float16)
    arrays["meta_keys"] = np.array(list(meta.keys()))
    arrays["meta_vals"] = np.array([float(v) for v in meta.values()])
    np.savez(path, **arrays)


def load_preconditioner(path, op, side: int, phase: Optional[np.ndarray] = None) -> CnnPreconditioner:
    """Rebuild M = D [z + s net(z / s)], z = P D^H r, for `op` from
    `save_preconditioner` output (P with its learned gain when stored; the
    carrier phase of D as `CirculantPreconditioner` takes it)."""
    with np.load(Path(path), allow_pickle=False) as f:
        meta = dict(zip([str(k) for k in f["meta_keys"]], f["meta_vals"].tolist()))
        sd = {k: torch.from_numpy(f[k]) for k in f.files if k[0] in "wb" and k[1:].isdigit()}
        gain = f["log_gain"].astype(np.float32) if "log_gain" in f.files else None
    pad = int(meta.get("pad", side))
    base = CirculantPreconditioner(op, side, meta["tau"], pad=pad, phase=phase)
    if gain is not None:
        base.set_gain(gain)
    pre = CnnPreconditioner(op.device, base=base)
    pre.load_state_dict(sd)
    return pre
