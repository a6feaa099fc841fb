"""sarhip_precond_net_noalloc (the preconditioner network's entry point that
allocates nothing: its workspace is a static device array of the library)
against sarhip_precond_net with a caller workspace, through the C ABI.

Two calls in a row: one 512 x 512 image (the largest the static array takes)
with a demodulation phasor, then two 128 x 128 images without one.  Both entry
points run the same kernels on the same inputs, so the outputs must agree to
FP32 rounding (they are expected to be bit-identical).  A size beyond the
static array is rejected with an error and no GPU work.
"""
import ctypes

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda", 0)


def _net(dev):
    from sarsolver_amd import learned
    pre = learned.CnnPreconditioner(dev, seed=7)
    g = torch.Generator(device="cpu").manual_seed(9)
    with torch.no_grad():
        pre.weights[-1].copy_(0.05 * torch.randn(pre.weights[-1].shape, generator=g))
        for b in pre.biases:
            b.copy_(0.02 * torch.randn(b.shape, generator=g))
    return pre


def _args(pre):
    packs, biases = pre._packed()
    n = len(packs)
    vp = ctypes.c_void_p
    wf = (vp * n)(*[p.data_ptr() for p in packs])
    bs = (vp * n)(*[b.data_ptr() for b in biases])
    oc = (ctypes.c_int * n)(*[w.shape[2] for w in pre.weights])
    return n, wf, bs, oc


def _call_both(pre, z, demod):
    from sarsolver_amd import learned
    L = learned._lib()
    B, H, W = z.shape
    n, wf, bs, oc = _args(pre)
    vp = ctypes.c_void_p
    stream = vp(torch.cuda.current_stream(z.device).cuda_stream)
    dm = None if demod is None else vp(demod.data_ptr())
    work = torch.empty(int(L.sarhip_precond_net_workspace(B, H, W)), dtype=torch.uint8,
                       device=z.device)
    out_ws = torch.empty_like(z)
    rc = L.sarhip_precond_net(vp(z.data_ptr()), dm, B, H, W, n, ctypes.cast(wf, vp),
                              ctypes.cast(bs, vp), ctypes.cast(oc, vp), vp(work.data_ptr()),
                              work.numel(), vp(out_ws.data_ptr()), stream)
    assert rc == 0, L.sarhip_conv_last_error()
    out_na = torch.full_like(z, float("nan"))
    rc = L.sarhip_precond_net_noalloc(vp(z.data_ptr()), dm, B, H, W, n, ctypes.cast(wf, vp),
                                      ctypes.cast(bs, vp), ctypes.cast(oc, vp),
                                      vp(out_na.data_ptr()), stream)
    assert rc == 0, L.sarhip_conv_last_error()
    torch.cuda.synchronize(z.device)
    return out_ws.cpu().numpy(), out_na.cpu().numpy()


def test_noalloc_matches_workspace_entry_point(dev):
    pre = _net(dev)
    g = torch.Generator(device="cpu").manual_seed(3)
    cases = [((1, 512, 512), True), ((2, 128, 128), False)]
    for shape, with_demod in cases:
        print(f"case: precond_net_noalloc {shape} demod={with_demod}", flush=True)
        z = torch.complex(torch.randn(shape, generator=g), torch.randn(shape, generator=g))
        z = z.to(torch.complex64).to(dev).contiguous()
        demod = None
        if with_demod:
            ph = torch.rand(shape[1] * shape[2], generator=g, dtype=torch.float64) * 2 * np.pi
            demod = torch.polar(torch.ones_like(ph), ph).to(torch.complex64).to(dev)
        a, b = _call_both(pre, z, demod)
        e = rel_l2(b, a)
        print(f"  rel L2 {e:.2e}, bit-identical {np.array_equal(a, b)}")
        assert np.all(np.isfinite(b))
        assert e <= 1e-6


def test_noalloc_rejects_sizes_beyond_its_scratch(dev):
    from sarsolver_amd import learned
    L = learned._lib()
    pre = _net(dev)
    n, wf, bs, oc = _args(pre)
    vp = ctypes.c_void_p
    z = torch.zeros((1, 8, 8), dtype=torch.complex64, device=dev)
    # 1 x 1024 x 1024 needs four times the static array; the check precedes any GPU work
    rc = L.sarhip_precond_net_noalloc(vp(z.data_ptr()), None, 1, 1024, 1024, n,
                                      ctypes.cast(wf, vp), ctypes.cast(bs, vp),
                                      ctypes.cast(oc, vp), vp(z.data_ptr()), None)
    msg = L.sarhip_conv_last_error().decode()
    print(f"  oversize: rc {rc}: {msg}")
    assert rc == -1 and "static scratch" in msg
