"""Ragged and tiny sizes against the FP64 oracle: one pixel, a pixel count
that leaves a partial tile and a partial forward region, one pulse, pulse
counts that leave a partial group of 64, and both spectral paths (N = 64:
fused polyphase kernels; N = 33: rocFFT)."""
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


@pytest.mark.parametrize("M,P,N", [(1, 16, 64), (1500, 65, 64), (3333, 1, 64), (2100, 130, 33),
                                   (700, 7, 128)])
@pytest.mark.parametrize("interp", ["linear", "nearest_ref"])
def test_edge_sizes_match_oracle(dev, oracle_lib, M, P, N, interp):
    from sarsolver_amd import SarPlan, synthetic
    ap = synthetic.circular_aperture(P, N, 2.5)
    g = np.random.default_rng(M + P + N)
    pos = np.zeros((M, 3))
    pos[:, :2] = g.uniform(-6.0, 6.0, (M, 2))  # scattered points (no grid)
    x = g.standard_normal(M) + 1j * g.standard_normal(M)
    y = g.standard_normal((P, N)) + 1j * g.standard_normal((P, N))
    tx, rx, srp = ap.transmitter_positions, ap.receiver_positions, ap.stab_ref_posns
    W = np.fft.ifftshift(ap.waveform_fft)
    print(f"case: M={M} P={P} N={N} {interp}", flush=True)
    pl = SarPlan(tx, rx, srp, pos, W, ap.slow_time_weighting, ap.centre_frequency,
                 ap.sample_bandwidth, ap.c_eff, 8.0, 1.0, interp=interp, precision="series",
                 device=dev)
    Yd = pl.forward(torch.from_numpy(x.astype(np.complex64)).to(dev)).cpu().numpy()
    Xd = pl.adjoint(torch.from_numpy(y.astype(np.complex64)).to(dev)).cpu().numpy()
    Yo = oracle_lib.forward(tx, rx, srp, pos, x, W, ap.slow_time_weighting, ap.centre_frequency,
                            ap.sample_bandwidth, N, upsample_ratio=8.0, interp=interp)
    Xo = oracle_lib.adjoint(tx, rx, srp, pos, y, W, ap.slow_time_weighting, ap.centre_frequency,
                            ap.sample_bandwidth, upsample_ratio=8.0, interp=interp)
    ef, ea = rel_l2(Yd, Yo), rel_l2(Xd, Xo)
    print(f"  layout {pl.layout['fwd_window_bins']} bins, {pl.layout['fwd_region_tiles']} tiles, "
          f"fused {pl.layout['fused_fft_ratio']}; forward {ef:.2e} adjoint {ea:.2e}")
    assert pl.layout["fused_fft_ratio"] == (8 if N in (64, 128) else 0)
    assert ef <= 1e-5 and ea <= 1e-5
