"""Dataset resampling: the FP64 oracle restatement (oracle/resample.py) pinned
by closed forms and by the forward model, and the package's host utilities
against it (CPU only)."""
import numpy as np
import pytest

from conftest import rel_l2


@pytest.fixture(scope="module")
def rs():
    from oracle import resample
    return resample


def test_downsample_kernel_closed_form(rs):
    """round(5 r) = 5 taps for r = 1: fftshift(ifft([1, 1, 0, 0, 1])) is
    (1 + 2 cos(2 pi k / 5)) / 5 re-centred, times hann(5) = [0, .5, 1, .5, 0]."""
    c1 = (1 + 2 * np.cos(2 * np.pi / 5)) / 5
    want = np.array([0.0, 0.5 * c1, 0.6, 0.5 * c1, 0.0])
    got = rs.downsample_kernel(1)
    assert np.allclose(got.real, want, atol=1e-7) and np.allclose(got.imag, 0, atol=1e-7)
    assert rs.downsample_kernel(4).shape == (20,)


@pytest.mark.parametrize("n", [64, 65])
def test_fast_freq_downsample_identity(rs, n):
    """Keeping every sample with a flat taper (safety 1: Tukey alpha 0) returns
    the input: the shifts, the crop indices and the FFT pair are consistent."""
    g = np.random.default_rng(n)
    ph = g.standard_normal((3, n)) + 1j * g.standard_normal((3, n))
    assert np.allclose(rs.fast_freq_downsample(ph, n, safety_factor=1.0), ph, atol=1e-12)


def test_fast_freq_downsample_keeps_a_short_range_profile(rs):
    """A range profile confined to the kept samples survives the crop: the
    cropped spectrum is the N2-point DFT of those samples (taper at 1 there)."""
    n, n2 = 64, 32
    prof = np.zeros(n, complex)
    prof[[0, 1, 2, 3, 60, 61, 62, 63]] = np.arange(1, 9) * (1 + 0.5j)
    ph = np.fft.fftshift(np.fft.fft(prof))[None, :]
    got = rs.fast_freq_downsample(ph, n2, safety_factor=1.0)
    short = np.zeros(n2, complex)
    short[[0, 1, 2, 3]] = prof[[0, 1, 2, 3]]
    short[[28, 29, 30, 31]] = prof[[60, 61, 62, 63]]
    assert np.allclose(got[0], np.fft.fftshift(np.fft.fft(short)), atol=1e-12)


def test_move_srps_round_trip(rs):
    g = np.random.default_rng(3)
    P, n = 20, 32
    tx = g.standard_normal((P, 3)) * 100 + [8000, 0, 4000]
    rx = tx + g.standard_normal((P, 3))
    srps = np.zeros((P, 3))
    ph = g.standard_normal((P, n)) + 1j * g.standard_normal((P, n))
    new = np.array([3.0, -2.0, 0.5])
    there = rs.move_srps(ph, srps, new, tx, rx, 9.6e9, 6e8)
    back = rs.move_srps(there, np.repeat(new[None], P, 0), np.zeros(3), tx, rx, 9.6e9, 6e8)
    assert np.allclose(back, ph, atol=1e-9)
    assert not np.allclose(there, ph)


def test_resampled_point_target_matches_the_new_aperture(rs, oracle_lib):
    """Physics: the resampled data of a scatterer near the new SRP is what the
    oracle forward measures with the resampled aperture (its positions, SRP,
    kernel-summed weights and cropped waveform): equal for a scatterer at the
    SRP, correlation > 0.995 two metres away (the slow-time kernel is a
    low-pass in Doppler).  The same check with the modulation sign flipped in
    the SRP move must fail, so the test can tell the conventions apart."""
    from sarsolver_amd import synthetic
    ap = synthetic.circular_aperture(256, 256, 2.0)
    tx, rx, srp = ap.transmitter_positions, ap.receiver_positions, ap.stab_ref_posns
    W = np.ones(256, complex)
    w = np.ones(256, complex)
    new_srp = np.array([4.0, -3.0, 0.0])
    def corr(offset, sign):
        x0 = (new_srp + offset)[None, :]
        Y = oracle_lib.forward(tx, rx, srp, x0, np.ones(1, complex), W, w, ap.centre_frequency,
                               ap.sample_bandwidth, 256, upsample_ratio=128.0, interp="linear")
        T, R, S, Wt, Ph, wave = rs.resample(np.fft.fftshift(Y, axes=1), tx, rx, srp, w, W,
                                            ap.centre_frequency, ap.sample_bandwidth, sign,
                                            ap.c_eff, [], new_srp, 128, 2, safety_factor=1.2)
        Yn = oracle_lib.forward(T, R, S, x0, np.ones(1, complex), np.fft.ifftshift(wave), Wt,
                                ap.centre_frequency, ap.sample_bandwidth, 128,
                                upsample_ratio=128.0, interp="linear")
        a, b = Ph.reshape(-1), np.fft.fftshift(Yn, axes=1).reshape(-1)
        return abs(np.vdot(a, b)) / (np.linalg.norm(a) * np.linalg.norm(b))

    c_at, c_off = corr(np.zeros(3), 1.0), corr(np.array([1.5, 1.2, 0.0]), 1.0)
    c_flip = corr(np.array([1.5, 1.2, 0.0]), -1.0)
    print(f"correlation at the SRP {c_at:.9f}, 1.9 m away {c_off:.6f}, sign flipped {c_flip:.3f}")
    assert c_at >= 1 - 1e-9 and c_off >= 0.995 and c_flip < 0.5


def test_package_utilities_match_the_oracle(rs):
    from sarsolver_amd.utils import fast_freq_downsample, get_downsample_kernel, move_srps
    g = np.random.default_rng(9)
    ph = g.standard_normal((5, 96)) + 1j * g.standard_normal((5, 96))
    for r in (1, 2, 3):
        assert np.allclose(get_downsample_kernel(r), rs.downsample_kernel(r), atol=1e-15)
    for n2, sf in ((48, 1.2), (31, 1.5), (96, 1.0)):
        assert rel_l2(fast_freq_downsample(ph, n2, sf), rs.fast_freq_downsample(ph, n2, sf)) < 1e-14
    tx = g.standard_normal((5, 3)) * 50 + [9000, 100, 5000]
    srps = g.standard_normal((5, 3))
    a = move_srps(ph, srps, np.array([1.0, 2.0, 0.0]), tx, tx, 9.6e9, 6e8, -1.0, 2.9e8)
    b = rs.move_srps(ph, srps, np.array([1.0, 2.0, 0.0]), tx, tx, 9.6e9, 6e8, -1.0, 2.9e8)
    assert rel_l2(a, b) < 1e-12
