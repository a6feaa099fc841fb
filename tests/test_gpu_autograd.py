"""Differentiation through the gfx950 operators (workload c4: a batch of 16 scenes,
512 x 512 complex64 each, own circular trajectory of 1024 pulses x 2048 samples,
3 degree aperture with jittered start angle, U = 8, linear interpolation).

Loss L = 0.5 ||A x - y||^2 with y from a perturbed scene.  The autograd backward
(the hand-written adjoint kernel) is checked (1) by central finite differences
on 100 random coordinates -- exact for a quadratic loss up to rounding, and
(2) against the FP64 oracle's adjoint on one scene.
"""
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


@pytest.fixture(scope="module")
def c4(dev):
    from sarsolver_amd import SarPlan, synthetic
    w = synthetic.WORKLOADS["c4_scene"]
    rng = np.random.default_rng(w.seed)
    plans, xs, ys, aps = [], [], [], []
    sc = synthetic.grid_scene(w.n, w.spacing)
    pos = sc.positions
    for b in range(16):
        start = float(rng.uniform(-180.0, 180.0))
        ap = synthetic.circular_aperture(w.pulses, w.samples, w.aperture_deg, start_deg=start)
        pl = SarPlan(ap.transmitter_positions, ap.receiver_positions, ap.stab_ref_posns, pos,
                     np.fft.ifftshift(ap.waveform_fft), ap.slow_time_weighting,
                     ap.centre_frequency, ap.sample_bandwidth, ap.c_eff, w.upsample, 1.0,
                     interp="linear", precision="series", device=dev)
        x = synthetic.scene_amplitudes(w, rng).reshape(-1)
        delta = 0.1 * synthetic.smooth_field(rng, w.n, 3.0).reshape(-1)
        xt = torch.from_numpy(x.astype(np.complex64)).to(dev)
        y = pl.forward(torch.from_numpy((x + delta).astype(np.complex64)).to(dev))
        plans.append(pl)
        xs.append(xt)
        ys.append(y)
        aps.append(ap)
    return w, pos, plans, torch.stack(xs), torch.stack(ys), aps


def test_batched_gradient_is_adjoint_of_residual(c4):
    from sarsolver_amd.torch_ops import BatchedSarOperator, data_fidelity
    w, pos, plans, X, Y, _ = c4
    A = BatchedSarOperator(plans)
    x = X.clone().requires_grad_(True)
    loss = data_fidelity(A, x, Y)
    loss.backward()
    with torch.no_grad():
        for b in (0, 7, 15):
            r = plans[b].forward(X[b]) - Y[b]
            g = plans[b].adjoint(r)
            assert rel_l2(x.grad[b].cpu().numpy(), g.cpu().numpy()) <= 1e-6


def test_finite_differences_100_coordinates(c4):
    from sarsolver_amd.torch_ops import BatchedSarOperator, data_fidelity, sar_forward
    w, pos, plans, X, Y, _ = c4
    A = BatchedSarOperator(plans)
    x = X.clone().requires_grad_(True)
    data_fidelity(A, x, Y).backward()
    grad = x.grad.detach()
    rng = np.random.default_rng(99)
    fd, an = [], []
    h = float(torch.sqrt(torch.mean(torch.abs(X) ** 2)))
    with torch.no_grad():
        for _ in range(100):
            b = int(rng.integers(0, 16))
            m = int(rng.integers(0, X.shape[1]))
            part = int(rng.integers(0, 2))  # 0: real direction, 1: imaginary direction
            step = h if part == 0 else 1j * h

            def loss_at(s):
                xb = X[b].clone()
                xb[m] += s
                r = sar_forward(xb, plans[b]) - Y[b]
                return 0.5 * float(torch.sum(r.real.double() ** 2 + r.imag.double() ** 2))
            fd.append((loss_at(step) - loss_at(-step)) / (2 * h))
            an.append(float(grad[b, m].real if part == 0 else grad[b, m].imag))
    fd, an = np.array(fd), np.array(an)
    err = np.linalg.norm(fd - an) / np.linalg.norm(an)
    print(f"finite differences vs autograd on 100 coordinates: rel error {err:.2e}")
    assert err <= 1e-3


def test_gradient_matches_oracle_adjoint(c4, oracle_lib):
    """One scene: A^H (A x - y) from the GPU against the FP64 oracle applied to the
    same residual, and the residual itself against the oracle forward."""
    w, pos, plans, X, Y, aps = c4
    b = 3
    ap = aps[b]
    with torch.no_grad():
        r = (plans[b].forward(X[b]) - Y[b])
        g = plans[b].adjoint(r).cpu().numpy()
        Ax = plans[b].forward(X[b]).cpu().numpy()
    tx = ap.transmitter_positions
    Wf = np.fft.ifftshift(ap.waveform_fft)
    go = oracle_lib.adjoint(tx, tx, ap.stab_ref_posns, pos, r.cpu().numpy(), Wf,
                            ap.slow_time_weighting, ap.centre_frequency, ap.sample_bandwidth,
                            upsample_ratio=w.upsample, interp="linear", num_threads=15)
    Axo = oracle_lib.forward(tx, tx, ap.stab_ref_posns, pos, X[b].cpu().numpy(), Wf,
                             ap.slow_time_weighting, ap.centre_frequency, ap.sample_bandwidth,
                             w.samples, upsample_ratio=w.upsample, interp="linear",
                             num_threads=15)
    eg, ef = rel_l2(g, go), rel_l2(Ax, Axo)
    # For this scene's look direction the smoothed field is nearly invisible to
    # the sensor: its forward is a sum that cancels heavily (the band holds
    # about 1/60 of the profile's norm), which amplifies per-term errors about
    # 60-fold.  The plain relative error is held to 1e-5 all the same, and the
    # normwise error ||dY|| / (||A|| ||x||), ||A||^2 = P N for unit weights,
    # to 1e-6.
    normwise = np.linalg.norm(Ax - Axo) / (np.sqrt(w.pulses * w.samples)
                                           * np.linalg.norm(X[b].cpu().numpy()))
    print(f"c4 scene {b}: gradient vs oracle {eg:.2e}; forward vs oracle {ef:.2e} "
          f"(normwise {normwise:.2e})")
    assert eg <= 1e-5
    assert ef <= 1e-5
    assert normwise <= 1e-6
