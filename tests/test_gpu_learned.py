"""The bf16 3x3 convolution kernels of the learned preconditioner against a
plain PyTorch FP32 reference, on ragged image sizes (not multiples of the
16 x 16 tile): forward with bias / ReLU / bf16 or FP32 output, the weight
gradient, and the layer backward (data gradient through the flipped,
channel-transposed forward kernel); the fused spectral filter of the
circulant base against torch's FFTs; the preconditioners at work in PCG."""
import ctypes

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
F = torch.nn.functional


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda", 0)


def _rand_input(dev, B, H, W, cin, seed):
    g = torch.Generator(device="cpu").manual_seed(seed)
    x = torch.zeros(B, H, W, 32)
    x[..., :cin] = torch.randn(B, H, W, cin, generator=g)
    return x.to(torch.bfloat16).to(dev)


def _torch_conv(x_bf16, w, b):
    """FP32 reference on the same bf16-rounded operands. w [9, cin, cout]."""
    cin, cout = w.shape[1], w.shape[2]
    wr = w.to(torch.bfloat16).float().reshape(3, 3, cin, cout).permute(3, 2, 0, 1)
    xin = x_bf16.float()[..., :cin].permute(0, 3, 1, 2)
    return F.conv2d(xin, wr, b, padding=1).permute(0, 2, 3, 1)


@pytest.mark.parametrize("cin,cout,relu,out_f32", [
    (32, 32, True, False), (2, 32, True, False), (32, 2, False, True), (32, 17, True, True),
    (32, 16, False, False)])
def test_conv_forward_matches_torch(dev, cin, cout, relu, out_f32):
    from sarsolver_amd.learned import conv3x3, pack_weights
    print(f"conv fwd cin={cin} cout={cout} relu={relu} f32={out_f32}", flush=True)
    g = torch.Generator(device="cpu").manual_seed(cin + cout)
    w = (torch.randn(9, cin, cout, generator=g) * 0.1).to(dev)
    b = torch.randn(cout, generator=g).to(dev)
    x = _rand_input(dev, 2, 37, 50, cin, 3)
    y = conv3x3(x, pack_weights(w), b, cout, relu, out_f32)
    ref = _torch_conv(x, w, b)
    if relu:
        ref = torch.relu(ref)
    assert y.shape == (2, 37, 50, cout)
    err = rel_l2(y.float().cpu().numpy(), ref.cpu().numpy())
    print(f"  rel L2 {err:.2e}")
    # FP32 accumulation over 288 products (order differs from torch): ~1e-7;
    # bf16 output rounding: 2^-9 relative per element at most.
    assert err <= (1e-5 if out_f32 else 4e-3)


@pytest.mark.parametrize("cout", [32, 2])
def test_conv_weight_gradient_matches_torch(dev, cout):
    from sarsolver_amd.learned import conv3x3_wgrad
    print(f"conv wgrad cout={cout}", flush=True)
    x = _rand_input(dev, 2, 37, 50, 32, 5)
    co = 32 if cout > 16 else 16
    g = torch.Generator(device="cpu").manual_seed(11)
    dy = torch.zeros(2, 37, 50, co)
    dy[..., :cout] = torch.randn(2, 37, 50, cout, generator=g)
    dy = dy.to(dev)
    dw, db = conv3x3_wgrad(x, dy, cout)
    # reference: autograd of the FP64 conv in the same layout
    w = torch.zeros(9, 32, cout, dtype=torch.float64, device=dev, requires_grad=True)
    b = torch.zeros(cout, dtype=torch.float64, device=dev, requires_grad=True)
    wr = w.reshape(3, 3, 32, cout).permute(3, 2, 0, 1)
    y = F.conv2d(x.double().permute(0, 3, 1, 2), wr, b, padding=1).permute(0, 2, 3, 1)
    (y * dy[..., :cout].double()).sum().backward()
    e1 = rel_l2(dw[:, :, :cout].cpu().numpy(), w.grad.cpu().numpy())
    e2 = rel_l2(db[:cout].cpu().numpy(), b.grad.cpu().numpy())
    print(f"  dW rel L2 {e1:.2e}, db rel L2 {e2:.2e}")
    assert e1 <= 1e-5 and e2 <= 1e-5
    if cout < co:
        assert float(dw[:, :, cout:].abs().max()) == 0.0


def test_layer_backward_matches_torch(dev):
    from sarsolver_amd.learned import _ConvLayer
    print("conv layer backward", flush=True)
    g = torch.Generator(device="cpu").manual_seed(21)
    w0 = (torch.randn(9, 32, 32, generator=g) * 0.1).to(dev)
    b0 = (torch.randn(32, generator=g) * 0.1).to(dev)
    x = _rand_input(dev, 1, 45, 29, 32, 8)
    xr = x.clone().requires_grad_(True)
    w = w0.clone().requires_grad_(True)
    b = b0.clone().requires_grad_(True)
    y = _ConvLayer.apply(xr, w, b, True, False, True)
    gy = torch.randn(y.shape, generator=g).to(dev)
    (y.float() * gy).sum().backward()
    # reference: FP32 conv on the bf16-rounded weights the kernel uses
    wq = w0.to(torch.bfloat16).float().requires_grad_(True)
    bq = b0.clone().requires_grad_(True)
    xq = x.float().requires_grad_(True)
    wr = wq.reshape(3, 3, 32, 32).permute(3, 2, 0, 1)
    yr = torch.relu(F.conv2d(xq.permute(0, 3, 1, 2), wr, bq, padding=1)).permute(0, 2, 3, 1)
    (yr * gy).sum().backward()
    ex = rel_l2(xr.grad.float().cpu().numpy(), xq.grad.cpu().numpy())
    ew = rel_l2(w.grad.cpu().numpy(), wq.grad.cpu().numpy())
    eb = rel_l2(b.grad.cpu().numpy(), bq.grad.cpu().numpy())
    print(f"  dx {ex:.2e}  dW {ew:.2e}  db {eb:.2e}")
    # the ReLU mask comes from the bf16 output and dx is rounded to bf16
    assert ex <= 1e-2 and ew <= 1e-2 and eb <= 1e-2


def test_learned_preconditioner_gain(dev):
    """What the network gains, on the 128^2 workload (256 pulses x 512 samples,
    2 deg, U = 8): CGLS, PCG with the demodulated circulant base alone, and PCG
    with the base plus the CNN after a short unrolled training, on 8 held-out
    20 dB scenes to relative residual 1e-3.  Every solve must reach the
    tolerance (true residual, recomputed); the trained network must not lose
    to the base on the held-out unrolled loss by more than training noise and
    the learned PCG must not need more iterations than CGLS.  The gains are
    printed."""
    from sarsolver_amd import learned
    from sarsolver_amd.train_preconditioner import workload_operator
    print("learned preconditioner gain on c1_small", flush=True)
    w, op = workload_operator("c1_small", dev)
    n, lam = w.n, 1e-3
    gens = learned.make_gens(11, dev)
    with torch.no_grad():
        probs = [learned.noisy_problem(op, x, 20.0, gens.dev_gen)
                 for x in learned.random_scenes(n, 4, gens, dev)]
    tau, scores = learned.choose_tau(op, n, (0.03, 0.08, 0.2), probs, lam=lam, pad=2 * n)
    base = learned.CirculantPreconditioner(op, n, tau, pad=2 * n)
    pre = learned.CnnPreconditioner(dev, base=base)
    cval = learned.problem_batch(op, n, 4, learned.make_gens(12, dev))
    M = lambda r: pre(r, (n, n))
    with torch.no_grad():
        l0 = float(learned.unrolled_pcg_loss(op, M, cval, lam, 8))
    learned.train_unrolled(op, pre, n, lam=lam, iters=8, steps=40, batch=2, lr=1e-4)
    with torch.no_grad():
        l1 = float(learned.unrolled_pcg_loss(op, M, cval, lam, 8))
    res = learned.evaluate(op, {"circulant": base, "learned": M}, n, scenes=8, lam=lam)
    s = res["solvers"]
    print(f"  tau {tau} (scores {scores}); unrolled loss base {l0:.4f} -> learned {l1:.4f}")
    for name, v in s.items():
        print(f"  {name:9s} median iters {v['median_iters']:5.1f}  median ms {v['median_ms']:7.2f}"
              f"  worst true residual {v['max_true_residual']:.2e}  fallbacks {v['fallbacks']}")
    print(f"  network gain over base: {s['circulant']['median_iters'] - s['learned']['median_iters']:+.1f}"
          f" iterations; learned/CGLS time {s['learned']['time_ratio_vs_cgls']:.2f}")
    for v in s.values():
        assert v["max_true_residual"] <= 1.05e-3
    assert l1 <= l0 + 0.05
    assert s["learned"]["median_iters"] <= s["cgls"]["median_iters"]


@pytest.mark.parametrize("with_base,batch", [(False, 1), (True, 1), (True, 2)])
def test_fused_inference_matches_layer_path(dev, with_base, batch):
    """sarhip_precond_net (scale, bf16 packing, six layers, output and phasor
    in one call) against the layer-by-layer path the training uses, with
    random weights in every layer (the last one included), on 128^2 images,
    flat and batched, with and without the demodulated circulant base."""
    from sarsolver_amd import learned
    from sarsolver_amd.train_preconditioner import workload_operator
    w, op = workload_operator("c1_small", dev)
    n = w.n
    base = learned.CirculantPreconditioner(op, n, 0.08, pad=2 * n) if with_base else None
    pre = learned.CnnPreconditioner(dev, base=base, seed=3)
    g = torch.Generator(device="cpu").manual_seed(5)
    with torch.no_grad():
        pre.weights[-1].copy_(0.05 * torch.randn(pre.weights[-1].shape, generator=g))
        for b in pre.biases:
            b.copy_(0.02 * torch.randn(b.shape, generator=g))
    shape = (n * n,) if batch == 1 else (batch, n * n)
    r = torch.complex(torch.randn(shape, generator=g), torch.randn(shape, generator=g)).to(dev)
    print(f"case: fused preconditioner inference base={with_base} batch={batch}", flush=True)
    assert pre._fused()
    with torch.no_grad():
        fused = pre(r, (n, n))
    with torch.enable_grad():
        ref = pre(r, (n, n)).detach()
    assert fused.shape == ref.shape and fused.dtype == ref.dtype
    z = r if base is None else base.apply_base(base.to_base(r))
    e = rel_l2(fused.cpu().numpy(), ref.cpu().numpy())
    # the network's share of the output, to show the comparison is not of z alone
    share = rel_l2((ref if base is None else base.to_base(ref)).cpu().numpy(), z.cpu().numpy())
    print(f"  fused vs layer path {e:.2e} (network share of the output {share:.2e})")
    assert share > 1e-2
    assert e <= 2e-3 * max(share, 1.0)


def test_c5_learned_pcg_on_held_out_scenes(dev):
    """The 512^2 workload of the learned preconditioner (1024 pulses x 2048
    samples, 3 deg, U = 8) with the committed weights: on 8 held-out 20 dB
    scenes (seed 2024, not the training seeds) every solver -- CGLS, PCG with
    the analytic demodulated circulant, PCG with the learned spectral filter
    alone, PCG with the learned filter + CNN -- reaches relative residual 1e-3
    by its recomputed true residual; the linear preconditioners never fall
    back to CGLS (the nonlinear one may, and is then finished by CGLS from its
    best iterate); the learned preconditioner takes fewer median iterations
    than the analytic circulant and than CGLS, and what the CNN gains over
    the learned filter alone is printed."""
    from sarsolver_amd import learned
    from sarsolver_amd.train_preconditioner import default_weights, workload_operator
    path = default_weights("c5_scene")
    print(f"case: c5 learned PCG with {path.name}", flush=True)
    w, op = workload_operator("c5_scene", dev)
    n = w.n
    pre = learned.load_preconditioner(path, op, n)
    assert pre.base.log_gain is not None
    analytic = learned.CirculantPreconditioner(op, n, pre.base.tau, pad=pre.base.pad)
    res = learned.evaluate(op, {"circulant": analytic, "learned_filter": pre.base,
                                "learned": lambda r: pre(r, (n, n))}, n, scenes=8, seed=2024)
    s = res["solvers"]
    for name, v in s.items():
        print(f"  {name:14s} median iters {v['median_iters']:5.1f}  median ms {v['median_ms']:7.2f}"
              f"  worst true residual {v['max_true_residual']:.3e}  fallbacks {v['fallbacks']}"
              + (f"  time/CGLS {v['time_ratio_vs_cgls']:.3f}" if name != "cgls" else ""))
    print(f"  CNN gain over the learned filter alone: "
          f"{s['learned_filter']['median_iters'] - s['learned']['median_iters']:+.1f} iterations")
    for name, v in s.items():
        assert v["max_true_residual"] <= 1e-3
        if name != "learned":
            assert v["fallbacks"] == 0
    assert s["learned"]["median_iters"] < s["circulant"]["median_iters"]
    assert s["learned"]["median_iters"] < s["cgls"]["median_iters"]


@pytest.mark.parametrize("n,m,B", [(128, 256, 1), (512, 1024, 1), (100, 128, 2), (64, 64, 1)])
def test_spectral_filter_matches_fft_reference(dev, n, m, B):
    """sarhip_spectral_filter (rows / columns / rows passes, one wave per
    transform) against the same filtering written with torch's FFTs in FP32:
    demodulate, zero-pad to m x m, FFT2, real filter / m^2, inverse FFT2, crop,
    phasor -- with and without the phasors, batched."""
    from sarsolver_amd import learned
    L = learned._lib()
    g = torch.Generator(device="cpu").manual_seed(n + m)
    z = torch.complex(torch.randn(B, n, n, generator=g), torch.randn(B, n, n, generator=g))
    z = z.to(torch.complex64).to(dev)
    f = (torch.rand(m, m, generator=g) * 3.0 + 0.1).to(dev)
    ph = torch.rand(n * n, generator=g, dtype=torch.float64) * 2 * np.pi
    d = torch.polar(torch.ones_like(ph), ph).to(torch.complex64).to(dev)
    work = torch.empty(int(L.sarhip_spectral_filter_workspace(B, n, m)), dtype=torch.uint8,
                       device=dev)
    vp = ctypes.c_void_p
    stream = vp(torch.cuda.current_stream(dev).cuda_stream)
    print(f"case: spectral filter n={n} m={m} B={B}", flush=True)
    for phasors in (False, True):
        out = torch.empty_like(z)
        dp = vp(d.data_ptr()) if phasors else None
        rc = L.sarhip_spectral_filter(vp(z.data_ptr()), dp, B, n, m, vp(f.data_ptr()),
                                      vp(work.data_ptr()), work.numel(), vp(out.data_ptr()), dp,
                                      stream)
        assert rc == 0, L.sarhip_conv_last_error()
        zz = z * d.reshape(n, n).conj() if phasors else z
        zp = torch.zeros(B, m, m, dtype=torch.complex64, device=dev)
        zp[:, :n, :n] = zz
        ref = torch.fft.ifft2(torch.fft.fft2(zp) * f)[:, :n, :n]
        if phasors:
            ref = ref * d.reshape(n, n)
        e = rel_l2(out.cpu().numpy(), ref.cpu().numpy())
        print(f"  phasors {phasors}: rel L2 {e:.2e}")
        assert e <= 2e-6
