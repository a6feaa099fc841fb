"""The fused polyphase spectral kernels (fft_kernels.hip) against the rocFFT
stages they replace and against the FP64 oracle: every supported band length,
integer upsampling ratios 1-16, FFT and centred order, 1/2/4 phases per
workgroup."""
import os

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


def _plan(N, U, P, fused, interp="linear"):
    from sarsolver_amd import synthetic
    from sarsolver_amd._native import SarPlan
    from sarsolver_amd.numpy import GridSarScene
    ap = synthetic.circular_aperture(P, N, 2.0, fs=60e6 * N / 256)
    sc = GridSarScene((1, 24, 20), (0.0, 11.5, 9.5), grid_centre=np.array([0.5, -0.3, 0.0]))
    g = np.random.default_rng(N + U)
    W = g.standard_normal(N) + 1j * g.standard_normal(N)
    w = g.standard_normal(P) + 1j * g.standard_normal(P)
    os.environ["SARHIP_FUSED_FFT"] = "1" if fused else "0"
    try:
        pl = SarPlan(ap.transmitter_positions, ap.receiver_positions, ap.stab_ref_posns,
                     sc.positions, W, w, ap.centre_frequency, ap.sample_bandwidth, ap.c_eff,
                     float(U), 1.0, interp=interp, precision="series")
    finally:
        os.environ.pop("SARHIP_FUSED_FFT", None)
    return pl, ap, sc, W, w


@pytest.mark.parametrize("N,U", [(64, 4), (128, 8), (256, 1), (256, 3), (512, 2), (1024, 16),
                                 (2048, 8), (4096, 8), (8192, 2)])
def test_fused_matches_rocfft(dev, N, U):
    P = 48
    print(f"case: fused spectral kernels N={N} U={U}", flush=True)
    fu, ap, sc, W, w = _plan(N, U, P, True)
    rf, *_ = _plan(N, U, P, False)
    assert fu.layout["fused_fft_ratio"] == U and rf.layout["fused_fft_ratio"] == 0
    g = torch.Generator().manual_seed(N)
    x = torch.complex(torch.randn(sc.num_scatterers, generator=g),
                      torch.randn(sc.num_scatterers, generator=g)).to(dev)
    y = torch.complex(torch.randn(P, N, generator=g), torch.randn(P, N, generator=g)).to(dev)
    for order in ("fft", "centred"):
        ef = rel_l2(fu.forward(x, order=order).cpu().numpy(), rf.forward(x, order=order).cpu().numpy())
        ea = rel_l2(fu.adjoint(y, order=order).cpu().numpy(), rf.adjoint(y, order=order).cpu().numpy())
        print(f"  {order}: forward {ef:.2e} adjoint {ea:.2e}")
        assert ef <= 2e-6 and ea <= 2e-6


@pytest.mark.parametrize("phases", ["1", "2", "4"])
def test_fused_phase_batches_against_oracle(dev, oracle_lib, phases):
    """N = 1024, U = 8 with 1, 2 and 4 phases per workgroup, against the FP64
    oracle (linear interpolation, both directions)."""
    N, U, P = 1024, 8, 40
    os.environ["SARHIP_FFT_PHASES"] = phases
    try:
        print(f"case: fused spectral kernels, {phases} phases per batch", flush=True)
        pl, ap, sc, W, w = _plan(N, U, P, True)
        g = np.random.default_rng(3)
        x = g.standard_normal(sc.num_scatterers) + 1j * g.standard_normal(sc.num_scatterers)
        y = g.standard_normal((P, N)) + 1j * g.standard_normal((P, N))
        Yd = pl.forward(torch.from_numpy(x.astype(np.complex64)).to(dev)).cpu().numpy()
        Xd = pl.adjoint(torch.from_numpy(y.astype(np.complex64)).to(dev)).cpu().numpy()
        torch.cuda.synchronize()
    finally:
        os.environ.pop("SARHIP_FFT_PHASES", None)
    tx, rx, srp = ap.transmitter_positions, ap.receiver_positions, ap.stab_ref_posns
    Yo = oracle_lib.forward(tx, rx, srp, sc.positions, x, W, w, ap.centre_frequency,
                            ap.sample_bandwidth, N, upsample_ratio=float(U), interp="linear")
    Xo = oracle_lib.adjoint(tx, rx, srp, sc.positions, y, W, w, ap.centre_frequency,
                            ap.sample_bandwidth, upsample_ratio=float(U), interp="linear")
    ef, ea = rel_l2(Yd, Yo), rel_l2(Xd, Xo)
    print(f"  forward vs oracle {ef:.2e}, adjoint vs oracle {ea:.2e}")
    assert ef <= 1e-5 and ea <= 1e-5


@pytest.mark.parametrize("extent,interp,precision,bistatic", [
    (6.0, "linear", "series", False), (40.0, "linear", "series", False),
    (6.0, "nearest_ref", "fp64", False), (6.0, "linear", "series", True),
    (40.0, "nearest", "series", True)])
def test_row_spans(dev, oracle_lib, extent, interp, precision, bistatic):
    """The fused stages touch only part of each profile row: the forward merge
    stores the bins its windows cover and the forward spectral stage reads only
    those; the adjoint spectral stage stores only the bins the backprojection
    can read.  A scene whose range extent (~16 bins per metre, N_w = 256)
    covers part of the row and one that covers it more than twice (the row
    wraps onto itself), monostatic and bistatic, against the oracle and
    against the rocFFT stages (which write and read whole rows); each
    direction is called again after calls with other data have left their
    values in the rows, and must repeat its first result exactly."""
    from sarsolver_amd import synthetic
    from sarsolver_amd._native import SarPlan
    from sarsolver_amd.numpy import GridSarScene
    N, U, P, fs = 64, 4, 40, 1.2e9
    ap = synthetic.circular_aperture(P, N, 3.0, fs=fs)
    tx, srp = ap.transmitter_positions, ap.stab_ref_posns
    rx = tx + np.array([300.0, -800.0, 100.0]) if bistatic else ap.receiver_positions
    n = 128  # 16 regions of 32 x 32 pixels, so that each region's window fits the LDS
    sc = GridSarScene((1, n, n), (0.0, extent, extent))
    # range extent of the scene in profile bins at the first pulse
    dr = np.linalg.norm(sc.positions - tx[0], axis=1) + np.linalg.norm(sc.positions - rx[0], axis=1)
    bins = np.ptp(dr) * fs * U / ap.c_eff
    assert (bins > 2 * N * U) if extent > 20 else (bins < N * U)
    g = np.random.default_rng(int(extent))
    x = g.standard_normal(n * n) + 1j * g.standard_normal(n * n)
    y = g.standard_normal((P, N)) + 1j * g.standard_normal((P, N))
    W = np.fft.ifftshift(ap.waveform_fft)
    args = (W, ap.slow_time_weighting, ap.centre_frequency, fs)
    want_y = oracle_lib.forward(tx, rx, srp, sc.positions, x, *args, N, upsample_ratio=float(U),
                                interp=interp)
    want_x = oracle_lib.adjoint(tx, rx, srp, sc.positions, y, *args, upsample_ratio=float(U),
                                interp=interp)
    xd = torch.from_numpy(x.astype(np.complex64)).to(dev)
    yd = torch.from_numpy(y.astype(np.complex64)).to(dev)
    other_x = 1e3 * torch.randn(n * n, dtype=torch.complex64, device=dev)
    other_y = 1e3 * torch.randn(P, N, dtype=torch.complex64, device=dev)
    out = {}
    for fused in (True, False):
        os.environ["SARHIP_FUSED_FFT"] = "1" if fused else "0"
        try:
            pl = SarPlan(tx, rx, srp, sc.positions, W, ap.slow_time_weighting, ap.centre_frequency,
                         fs, ap.c_eff, float(U), 1.0, interp=interp, precision=precision,
                         device=dev)
        finally:
            os.environ.pop("SARHIP_FUSED_FFT", None)
        assert pl.layout["fused_fft_ratio"] == (U if fused else 0)
        assert pl.layout["fwd_window_bins"] > 0
        print(f"case: row spans extent {extent} {interp} {precision} bistatic={bistatic} "
              f"fused={fused}", flush=True)
        f1 = pl.forward(xd).cpu().numpy()
        a1 = pl.adjoint(yd).cpu().numpy()
        pl.adjoint(other_y)
        f2 = pl.forward(xd).cpu().numpy()
        pl.forward(other_x)
        a2 = pl.adjoint(yd).cpu().numpy()
        assert np.array_equal(f1, f2) and np.array_equal(a1, a2)
        out[fused] = (f2, a2)
        ef, ea = rel_l2(f2, want_y), rel_l2(a2, want_x)
        print(f"  vs oracle: forward {ef:.2e}, adjoint {ea:.2e}")
        assert ef <= 1e-5 and ea <= 1e-5
    assert rel_l2(out[True][0], out[False][0]) <= 2e-6
    assert rel_l2(out[True][1], out[False][1]) <= 2e-6
