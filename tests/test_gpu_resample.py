"""Dataset resampling on the GPU (`NumpySimpleSarDataset.from_resampling` ->
`sarhip_resample`) against the FP64 oracle restatement of the reference's
window loop (oracle/resample.py)."""
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


def _dataset(oracle_lib, P, N, bad=()):
    """Circular 2 deg aperture with P pulses x N samples; phase history = the
    oracle forward of five scatterers near (4, -3, 0) m plus complex noise
    (centred order); pulses in `bad` get a non-finite transmitter."""
    from sarsolver_amd import synthetic
    from sarsolver_amd.numpy import NumpySimpleSarDataset
    ap = synthetic.circular_aperture(P, N, 2.0)
    tx = ap.transmitter_positions.copy()
    for b in bad:
        tx[b, 0] = np.nan
    g = np.random.default_rng(P + N)
    pos = np.array([4.0, -3.0, 0.0]) + g.uniform(-6, 6, size=(5, 3)) * [1, 1, 0]
    amps = g.standard_normal(5) + 1j * g.standard_normal(5)
    W = np.ones(N, complex)
    Y = oracle_lib.forward(tx, tx, ap.stab_ref_posns, pos, amps, W, np.ones(P, complex),
                           ap.centre_frequency, ap.sample_bandwidth, N, upsample_ratio=8.0,
                           interp="linear")
    ph = np.fft.fftshift(Y, axes=1)
    ph = ph + 0.05 * np.abs(ph).max() * (g.standard_normal(ph.shape) + 1j * g.standard_normal(ph.shape))
    ds = NumpySimpleSarDataset.from_scratch(tx, tx.copy(), ap.stab_ref_posns, ap.centre_frequency,
                                            ap.sample_bandwidth, phase_history=ph,
                                            waveform_fft=W, slow_time_weighting=np.ones(P, complex),
                                            c_eff=ap.c_eff)
    return ds


def test_resampling_matches_oracle(dev, oracle_lib):
    """slow factor 3 (K = 15 taps), fast factor 2 (512 -> 257 samples, odd),
    new SRP off the old one, a bad pulse whose windows must be dropped."""
    from oracle import resample as rs
    from sarsolver_amd.numpy import NumpySimpleSarDataset
    print("case: resampling 512 x 512, slow 3, fast 2, bad pulse 200", flush=True)
    ds = _dataset(oracle_lib, 512, 512, bad=(200,))
    ap = ds.aperture
    new_srp = np.array([4.0, -3.0, 0.0])
    out = NumpySimpleSarDataset.from_resampling(ds, fast_downsample_factor=2.0,
                                                slow_downsample_factor=3, new_srp=new_srp,
                                                device=dev)
    T, R, S, Wt, Ph, wave = rs.resample(ds.phase_history(), ap.transmitter_positions,
                                        ap.receiver_positions, ap.stab_ref_posns,
                                        ap.slow_time_weighting, ap.waveform_fft,
                                        ap.centre_frequency, ap.sample_bandwidth,
                                        ap.sign_multiplier, ap.c_eff, ap.bad_slow_times, new_srp,
                                        257, 3, safety_factor=1.2)
    na = out.aperture
    assert na.num_slow_times == T.shape[0] and na.num_freqs == 257
    assert T.shape[0] < ((512 - 15) // 3) - 1  # the windows around pulse 200 are gone
    assert np.array_equal(na.transmitter_positions, T) and np.array_equal(na.stab_ref_posns, S)
    assert np.allclose(na.slow_time_weighting, Wt, atol=1e-12)
    assert np.allclose(na.waveform_fft, wave, atol=1e-12)
    e = rel_l2(out.phase_history(), Ph)
    print(f"  {T.shape[0]} output pulses, phase history rel L2 {e:.2e}")
    assert e <= 1e-5


def test_resampling_more_pulses_than_one_grid_column(dev, oracle_lib):
    """70000 input pulses (row launches are chunked at 65535), slow factor 1
    (K = 5), 64 -> 32 samples; against the oracle's per-pulse stages applied
    to all pulses and the kernel summed over each window."""
    from oracle import resample as rs
    from sarsolver_amd import _native, synthetic
    from sarsolver_amd.utils.signal import downsample_window
    print("case: resampling 70000 x 64", flush=True)
    P, N, N2 = 70000, 64, 32
    ap = synthetic.circular_aperture(P, N, 2.0)
    g = np.random.default_rng(11)
    ph = g.standard_normal((P, N)) + 1j * g.standard_normal((P, N))
    new_srp = np.array([2.0, 1.0, 0.0])
    tx, srps = ap.transmitter_positions, ap.stab_ref_posns
    dr = rs.srp_range_change(srps, new_srp, tx, tx)
    kern = rs.downsample_kernel(1)
    starts = np.arange(0, P - 5, dtype=np.int32)[: ((P - 5) // 1) - 1]
    out = _native.resample(torch.from_numpy(ph.astype(np.complex64)).to(dev), dr,
                           ap.centre_frequency, ap.sample_bandwidth, 1.0, ap.c_eff, N2,
                           downsample_window(N2, 1.2) / N, kern, starts).cpu().numpy()
    ds = rs.fast_freq_downsample(rs.move_srps(ph, srps, new_srp, tx, tx, ap.centre_frequency,
                                              ap.sample_bandwidth, 1.0, ap.c_eff), N2, 1.2)
    ref = sum(kern[k] * ds[starts + k] for k in range(5))
    e = rel_l2(out, ref)
    print(f"  {starts.size} output pulses, rel L2 {e:.2e}")
    assert out.shape == (starts.size, N2) and e <= 1e-5
