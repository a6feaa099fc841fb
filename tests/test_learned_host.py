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


class _CirculantOp:
    """A = S F D^H on n x n images (F the unitary 2-D DFT, S = diag(sqrt(sigma)),
    D = diag(exp(i phi))): A^H A = D F^H diag(sigma) F D^H, a demodulated
    circulant.  CPU, complex128."""

    def __init__(self, n, sigma, phi):
        self.n, self.device = n, torch.device("cpu")
        self.s = torch.sqrt(torch.as_tensor(sigma, dtype=torch.float64))
        self.d = torch.polar(torch.ones(n * n, dtype=torch.float64),
                             torch.as_tensor(phi, dtype=torch.float64))

    def forward(self, x):
        z = (x.to(torch.complex128) * self.d.conj()).reshape(self.n, self.n)
        return (self.s * torch.fft.fft2(z, norm="ortho")).reshape(-1)

    def adjoint(self, y):
        z = torch.fft.ifft2(self.s * y.reshape(self.n, self.n), norm="ortho").reshape(-1)
        return z * self.d


def _band_sigma(n):
    f = torch.fft.fftfreq(n, dtype=torch.float64)
    return 0.05 + torch.exp(-((f[:, None] / 0.15) ** 2 + (f[None, :] / 0.1) ** 2))


def test_circulant_preconditioner_inverts_a_demodulated_circulant():
    """For an operator whose normal matrix is exactly D C D^H, the base with
    tau = lam and no padding is (A^H A + lam I)^-1: the PSF measured through the
    operator, the re-centring roll, the demodulation sign and the filter all
    have to be right for K P r = r."""
    from sarsolver_amd.learned import CirculantPreconditioner
    n, lam = 32, 1e-3
    yy, xx = np.mgrid[0:n, 0:n]
    phi = (0.9 * yy + 0.03 * xx ** 2 + 0.4 * xx).reshape(-1)   # carrier with curvature
    op = _CirculantOp(n, _band_sigma(n), phi)
    P = CirculantPreconditioner(op, n, lam, pad=n, phase=phi)
    assert torch.allclose(P.demod.to(torch.complex128), op.d, atol=1e-6)
    g = torch.Generator().manual_seed(3)
    r = torch.complex(torch.randn(n * n, generator=g, dtype=torch.float64),
                      torch.randn(n * n, generator=g, dtype=torch.float64))
    z = P(r.to(torch.complex64)).to(torch.complex128)
    Kz = op.adjoint(op.forward(z)) + lam * z
    assert float(torch.linalg.vector_norm(Kz - r) / torch.linalg.vector_norm(r)) < 1e-5


def test_circulant_preconditioner_picks_the_demodulation_sign():
    """With the conjugate carrier the base must select the opposite sign."""
    from sarsolver_amd.learned import CirculantPreconditioner
    n = 32
    yy, xx = np.mgrid[0:n, 0:n]
    phi = (1.3 * yy + 0.02 * yy ** 2 + 0.7 * xx).reshape(-1)
    op = _CirculantOp(n, _band_sigma(n), -phi)
    P = CirculantPreconditioner(op, n, 1e-3, pad=n, phase=phi)
    assert torch.allclose(P.demod.to(torch.complex128), op.d, atol=1e-6)


def test_unrolled_loss_with_identity_is_plain_cg():
    """With M = I the unrolled recursion is CG on the normal equations, whose
    residuals CGLS reports; the loss is the mean log10 of the first `iters`."""
    from sarsolver_amd.learned import unrolled_pcg_loss
    from sarsolver_amd.solvers import DenseOperator, cgls
    g = torch.Generator().manual_seed(5)
    A = torch.complex(torch.randn(60, 40, generator=g, dtype=torch.float64),
                      torch.randn(60, 40, generator=g, dtype=torch.float64))
    b = torch.complex(torch.randn(60, generator=g, dtype=torch.float64),
                      torch.randn(60, generator=g, dtype=torch.float64))
    op = DenseOperator(A)
    lam, T = 0.01, 6
    ref = cgls(op, b, lam=lam, tol=0.0, max_iter=T)
    c = op.adjoint(b)[None, :]
    loss = float(unrolled_pcg_loss(op, lambda r: r, c, lam, T))
    assert abs(loss - float(np.mean(np.log10(ref.residuals[1:T + 1])))) < 1e-8


def test_weights_round_trip(tmp_path):
    from sarsolver_amd.learned import CnnPreconditioner, save_preconditioner
    pre = CnnPreconditioner(torch.device("cpu"), seed=4)
    with torch.no_grad():
        pre.weights[-1].normal_()
    save_preconditioner(tmp_path / "m.npz", pre, {"tau": 0.05, "pad": 64})
    with np.load(tmp_path / "m.npz", allow_pickle=False) as f:
        assert dict(zip(f["meta_keys"].tolist(), f["meta_vals"].tolist())) == {"tau": 0.05, "pad": 64.0}
        for i, w in enumerate(pre.weights):
            assert np.array_equal(f[f"w{i}"], w.detach().numpy())


def _circ_phase(n=32):
    yy, xx = np.mgrid[0:n, 0:n]
    return (0.9 * yy + 0.03 * xx ** 2 + 0.4 * xx).reshape(-1)


def _circ_setup(n=32, pad=64, tau=0.3):
    from sarsolver_amd.learned import CirculantPreconditioner
    phi = _circ_phase(n)
    op = _CirculantOp(n, _band_sigma(n), phi)
    return op, CirculantPreconditioner(op, n, tau, pad=pad, phase=phi)


def test_spectral_gain_scales_the_filter():
    """The learned log-gain multiplies the analytic filter by exp(g): g = 0
    leaves P unchanged, a constant g = c scales P r by e^c, and the gain lives
    on a (pad/2)^2 grid interpolated to the pad x pad filter grid."""
    op, P = _circ_setup()
    g = torch.Generator().manual_seed(7)
    r = torch.complex(torch.randn(32 * 32, generator=g), torch.randn(32 * 32, generator=g))
    with torch.no_grad():
        r0 = P(r)
        P.set_gain(np.zeros((32, 32)))
        assert torch.equal(P.filter, P.base_filter * 1.0)
        assert torch.allclose(P(r), r0)
        P.set_gain(np.full((32, 32), 0.5))
        assert torch.allclose(P.filter, P.base_filter * np.exp(0.5), rtol=1e-6)
        assert torch.allclose(P(r), r0 * np.exp(0.5), rtol=1e-5, atol=1e-6)
    with pytest.raises(ValueError):
        P.set_gain(np.zeros((64, 64)))


def test_spectral_gain_training_lowers_the_unrolled_loss():
    """train_filter on an operator whose normal matrix is a demodulated
    circulant: with a badly regularised base (tau = 0.3 against lam = 1e-3)
    the learned gain lowers the held-out unrolled PCG loss."""
    from sarsolver_amd import learned
    op, P = _circ_setup(tau=0.3)
    cval = learned.problem_batch(op, 32, 2, learned.make_gens(11, torch.device("cpu")))
    with torch.no_grad():
        l0 = float(learned.unrolled_pcg_loss(op, P, cval, 1e-3, 6))
    learned.train_filter(op, P, 32, lam=1e-3, iters=6, steps=30, batch=2, lr=0.05, seed=3)
    assert not P.log_gain.requires_grad
    with torch.no_grad():
        l1 = float(learned.unrolled_pcg_loss(op, P, cval, 1e-3, 6))
    assert l1 < l0 - 0.05, (l0, l1)


def test_spectral_gain_round_trip(tmp_path):
    from sarsolver_amd import learned
    op, P = _circ_setup()
    gain = np.random.default_rng(2).standard_normal((32, 32)) * 0.1
    P.set_gain(gain)
    pre = learned.CnnPreconditioner(torch.device("cpu"), base=P)
    learned.save_preconditioner(tmp_path / "g.npz", pre, {"tau": 0.3, "pad": 64})
    back = learned.load_preconditioner(tmp_path / "g.npz", op, 32, phase=_circ_phase())
    # stored as float16
    assert np.array_equal(back.base.log_gain.numpy(), gain.astype(np.float16).astype(np.float32))
    assert back.base.pad == 64 and back.base.tau == 0.3
