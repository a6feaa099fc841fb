"""The backprojection's pair path: demodulated pair rows (h[n], h[n+1] - h[n])
e^{-2 pi i rho (n + 1/2)} formed after the spectral stage, the per-pixel phase
rho f + phi' in FP32 (rho = f_c / (f_s U), revolutions per bin).  It is taken
for monostatic planar series plans with fused spectral stages when rho is an
FP32 number of magnitude at most 2.5.  Each case is held to the FP64 oracle at
1e-5, and the pair path to the plain gather (SARHIP_BP_PAIRS=0) at 1e-6:
rho = 2 (the X-band workloads: 9.6 GHz / (600 MHz x 8)), 1.5 and 0.25 take the
pair path; rho = 25/12 (not an FP32 number) and 4 (above the bound) do not."""
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


@pytest.mark.parametrize("fc,fs,U,pairs", [
    (9.6e9, 600e6, 8.0, True), (7.2e9, 600e6, 8.0, True), (1.2e9, 600e6, 8.0, True),
    (10e9, 300e6, 16.0, False), (9.6e9, 600e6, 4.0, False)])
def test_pair_path_matches_oracle_and_plain_gather(dev, oracle_lib, fc, fs, U, pairs):
    from sarsolver_amd import SarPlan, synthetic
    P, N = 128, 512
    ap = synthetic.circular_aperture(P, N, 2.0, fc=fc, fs=fs)
    sc = synthetic.grid_scene(128, 0.1)
    tx, rx, srp = ap.transmitter_positions, ap.receiver_positions, ap.stab_ref_posns
    g = np.random.default_rng(int(fc / 1e8) + int(U))
    y = g.standard_normal((P, N)) + 1j * g.standard_normal((P, N))
    W = np.fft.ifftshift(ap.waveform_fft)
    out = {}
    for flag in ("1", "0"):
        print(f"case: gather pairs fc {fc} fs {fs} U {U} SARHIP_BP_PAIRS={flag}", flush=True)
        os.environ["SARHIP_BP_PAIRS"] = flag
        try:
            pl = SarPlan(tx, rx, srp, sc.positions, W, ap.slow_time_weighting, fc, fs, ap.c_eff,
                         U, 1.0, interp="linear", precision="series", device=dev)
        finally:
            os.environ.pop("SARHIP_BP_PAIRS", None)
        print(f"  layout {pl.layout}")
        assert pl.layout["bp_series_order"] == 3
        assert pl.layout["bp_pairs"] == (pairs and flag == "1")
        X = pl.adjoint(torch.from_numpy(y.astype(np.complex64)).to(dev)).cpu().numpy()
        out[flag] = X
        del pl
    Xo = oracle_lib.adjoint(tx, rx, srp, sc.positions, y, W, ap.slow_time_weighting, fc, fs,
                            upsample_ratio=U, interp="linear")
    e1, e0, d = rel_l2(out["1"], Xo), rel_l2(out["0"], Xo), rel_l2(out["1"], out["0"])
    print(f"  rho {fc / (fs * U):.4f}: pair path {e1:.2e}, plain {e0:.2e}, between {d:.2e}")
    assert e1 <= 1e-5 and e0 <= 1e-5 and d <= 1e-6


@pytest.mark.parametrize("fc", [9.6e9, 4.8e9])
@pytest.mark.parametrize("interp", ["linear", "nearest_ref"])
def test_pairs_formed_by_the_spectral_kernel(dev, oracle_lib, fc, interp):
    """For an integer rho (2 for the X-band workloads, 1 here with f_c = 4.8 GHz)
    the demodulating phasor is (-1)^rho, and the adjoint spectral kernel forms
    the demodulated rows itself as it stores the profile row; with
    SARHIP_PAIRS_FUSED=0 pair_rows_kernel forms them after it.  The two give
    the same rows, so the same image bit for bit, and both match the oracle."""
    from sarsolver_amd import SarPlan, synthetic
    P, N, U, fs = 128, 512, 8.0, 600e6
    ap = synthetic.circular_aperture(P, N, 2.0, fc=fc, fs=fs)
    sc = synthetic.grid_scene(128, 0.1)
    tx, rx, srp = ap.transmitter_positions, ap.receiver_positions, ap.stab_ref_posns
    g = np.random.default_rng(int(fc / 1e8) + 3)
    y = g.standard_normal((P, N)) + 1j * g.standard_normal((P, N))
    W = np.fft.ifftshift(ap.waveform_fft)
    out = {}
    for flag in ("1", "0"):
        print(f"case: pairs from the spectral kernel fc {fc} {interp} SARHIP_PAIRS_FUSED={flag}",
              flush=True)
        os.environ["SARHIP_PAIRS_FUSED"] = flag
        try:
            pl = SarPlan(tx, rx, srp, sc.positions, W, ap.slow_time_weighting, fc, fs, ap.c_eff,
                         U, 1.0, interp=interp, precision="series", device=dev)
        finally:
            os.environ.pop("SARHIP_PAIRS_FUSED", None)
        assert pl.layout["bp_pairs"], pl.layout
        out[flag] = pl.adjoint(torch.from_numpy(y.astype(np.complex64)).to(dev)).cpu().numpy()
        del pl
    Xo = oracle_lib.adjoint(tx, rx, srp, sc.positions, y, W, ap.slow_time_weighting, fc, fs,
                            upsample_ratio=U, interp=interp)
    e = rel_l2(out["1"], Xo)
    print(f"  rho {fc / (fs * U):.1f}: {e:.2e} against the oracle; identical "
          f"{np.array_equal(out['1'], out['0'])}")
    assert np.array_equal(out["1"], out["0"]) and e <= 1e-5
