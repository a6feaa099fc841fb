"""Host-side pieces of the learned preconditioner (CPU only): the weight
fragment packing against an emulation of the matrix-core conv kernel, the
data-gradient identity (flipped, channel-transposed weights), and argument
checks of the convolution entry points."""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")
F = torch.nn.functional


def _emulate_conv_kernel(x, wfrag, bias, out_c, relu):
    """What conv3x3_fwd_kernel computes, written from its operand layouts:
    for each tap, D += A B with A[m][k] = input pixel m (shifted by the tap),
    channel k, and B[k][n] read from the fragment as lane l, element i ->
    B[8 (l // 16) + i][l % 16] of column block nb."""
    Bn, H, W, Cc = x.shape
    taps, NB = wfrag.shape[0], wfrag.shape[1]
    Bmat = torch.zeros(taps, 32, 16 * NB)
    for tap in range(taps):
        for nb in range(NB):
            for lane in range(64):
                for i in range(8):
                    Bmat[tap, 8 * (lane // 16) + i, 16 * nb + lane % 16] = wfrag[tap, nb, lane, i]
    xp = F.pad(x.permute(0, 3, 1, 2), (1, 1, 1, 1)).permute(0, 2, 3, 1)
    y = torch.zeros(Bn, H, W, 16 * NB)
    for tap in range(9):
        kh, kw = divmod(tap, 3)
        y += xp[:, kh:kh + H, kw:kw + W, :] @ Bmat[tap]
    y = y[..., :out_c] + bias[:out_c]
    return torch.relu(y) if relu else y


@pytest.mark.parametrize("cin,cout", [(32, 32), (2, 32), (32, 2), (32, 17)])
def test_fragment_packing_matches_torch_conv(cin, cout):
    from sarsolver_amd.learned import pack_weights
    g = torch.Generator().manual_seed(cin * 100 + cout)
    w = torch.randn(9, cin, cout, generator=g)
    b = torch.randn(cout, generator=g)
    x = torch.zeros(2, 7, 9, 32)
    x[..., :cin] = torch.randn(2, 7, 9, cin, generator=g)
    frag = pack_weights(w).float()
    assert frag.shape == (9, 2 if cout > 16 else 1, 64, 8)
    got = _emulate_conv_kernel(x, frag, b, cout, relu=False)
    # torch reference on the same bf16-rounded weights: conv2d weight [co, ci, kh, kw]
    wr = w.to(torch.bfloat16).float().reshape(3, 3, cin, cout).permute(3, 2, 0, 1)
    ref = F.conv2d(x[..., :cin].permute(0, 3, 1, 2), wr, b, padding=1).permute(0, 2, 3, 1)
    assert torch.allclose(got, ref, rtol=1e-5, atol=1e-4)


def test_data_gradient_is_flipped_transposed_conv():
    from sarsolver_amd.learned import flip_transpose
    g = torch.Generator().manual_seed(7)
    w = torch.randn(9, 5, 6, generator=g, dtype=torch.float64)
    x = torch.randn(1, 5, 8, 11, generator=g, dtype=torch.float64, requires_grad=True)
    wt = w.reshape(3, 3, 5, 6).permute(3, 2, 0, 1)
    y = F.conv2d(x, wt, padding=1)
    gy = torch.randn(y.shape, generator=g, dtype=torch.float64)
    (y * gy).sum().backward()
    w2 = flip_transpose(w).reshape(3, 3, 6, 5).permute(3, 2, 0, 1)
    dx = F.conv2d(gy, w2, padding=1)
    assert torch.allclose(dx, x.grad, rtol=1e-12, atol=1e-12)


def test_conv_entry_points_reject_bad_arguments_without_a_gpu():
    from sarsolver_amd import _native
    L = _native.load_library()
    L.sarhip_conv3x3_bf16.restype = ctypes.c_int
    L.sarhip_conv_last_error.restype = ctypes.c_char_p
    vp = ctypes.c_void_p
    rc = L.sarhip_conv3x3_bf16(vp(0), vp(0), vp(0), 1, 16, 16, 32, 1, 0, vp(0), vp(0))
    assert rc == -1 and b"bad argument" in L.sarhip_conv_last_error()
    L.sarhip_conv3x3_wgrad_workspace.restype = ctypes.c_size_t
    assert L.sarhip_conv3x3_wgrad_workspace(2, 33, 17, 32) == 2 * 3 * 2 * (9 * 32 * 32 + 32) * 4
    assert L.sarhip_conv3x3_wgrad_workspace(1, 16, 16, 2) == (9 * 32 * 16 + 16) * 4
