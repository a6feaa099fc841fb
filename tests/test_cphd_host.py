"""CPHD readers (sarsolver_amd.io; reference io/cphd101.py, io/cphd03.py) on
files written by tests/cphd_files.py (CPU only).

sarpy, which the reference reads CPHD with, is absent and the reference holds
no CPHD file (its test downloads one), so the readers' parity with sarpy is
unpinned; these tests pin what does not depend on it: the layout round trip
in every sample format, the reference's derived quantities (centre frequency
f0 + N df / 2, bandwidth N df, sign), and that a point target measured by the
FP64 oracle, stored in a file and read back, backprojects to its position."""
import numpy as np
import pytest

from cphd_files import write_cphd03, write_cphd101


def _geometry(P, seed=0):
    g = np.random.default_rng(seed)
    th = np.deg2rad(np.linspace(-1.5, 1.5, P))
    tx = np.stack([10e3 * np.cos(th), 10e3 * np.sin(th), np.full(P, 5e3)], 1)
    rx = tx + g.normal(0, 0.5, (P, 3))
    srp = g.normal(0, 0.1, (P, 3))
    return tx, rx, srp


@pytest.mark.parametrize("fmt", ["CF8", "CI4", "CI2"])
@pytest.mark.parametrize("with_sf", [False, True])
def test_cphd101_round_trip(tmp_path, fmt, with_sf):
    from sarsolver_amd.io import Cphd101Dataset
    P, N = 37, 24
    g = np.random.default_rng(1)
    if fmt == "CF8":
        sig = (g.standard_normal((P, N)) + 1j * g.standard_normal((P, N))).astype(np.complex64)
    else:
        hi = 30000 if fmt == "CI4" else 120
        sig = g.integers(-hi, hi, (P, N)) + 1j * g.integers(-hi, hi, (P, N))
    sf = g.uniform(0.5, 2.0, P) if with_sf else None
    tx, rx, srp = _geometry(P)
    path = write_cphd101(tmp_path / "a.cphd", [dict(signal=sig, tx=tx, rx=rx, srp=srp, sc0=9.3e9,
                                                     scss=1.25e6, amp_sf=sf)], sgn=-1, fmt=fmt)
    ds = Cphd101Dataset(path)
    assert ds.num_slow_times == P and ds.num_freqs == N
    assert np.array_equal(ds.transmitter_positions, tx)
    assert np.array_equal(ds.receiver_positions, rx) and np.array_equal(ds.stab_ref_posns, srp)
    assert ds.sample_bandwidth == N * 1.25e6
    assert ds.centre_frequency == 9.3e9 + 0.5 * N * 1.25e6
    assert ds.sign_multiplier == -1.0
    want = sig.astype(np.complex64)
    if sf is not None:
        want = (want * sf[:, None].astype(np.float32)).astype(np.complex64)
    ph = ds.phase_history()
    assert ph.dtype == np.complex64 and np.array_equal(ph, want)
    assert np.array_equal(ds.phase_history(slice(5, 20, 3)), want[5:20:3])
    ap = ds.aperture
    assert ap.num_slow_times == P and np.allclose(ap.waveform_fft, 1.0)
    sub = ds.get_subset(slice(10, 20))
    assert sub.aperture.num_slow_times == 10 and np.array_equal(sub.phase_history(), want[10:20])
    blocks = ds.to_blocks(3)
    assert [b.aperture.num_slow_times for b in blocks] == [12, 12, 13]


def test_cphd101_channels_and_sign(tmp_path):
    from sarsolver_amd.io import Cphd101Dataset, open_cphd
    chans = []
    for i, (P, N) in enumerate(((10, 8), (13, 16))):
        tx, rx, srp = _geometry(P, seed=i)
        sig = np.full((P, N), i + 1j * (i + 1), dtype=np.complex64)
        chans.append(dict(signal=sig, tx=tx, rx=rx, srp=srp, sc0=1e9 * (i + 1), scss=1e6))
    path = write_cphd101(tmp_path / "b.cphd", chans, sgn=+1)
    f = open_cphd(path)
    assert f.version == "1.0.1" and len(f.channels) == 2
    ds = Cphd101Dataset(path, channel_index=1)
    assert ds.num_slow_times == 13 and ds.num_freqs == 16 and ds.sign_multiplier == 1.0
    assert np.array_equal(ds.transmitter_positions, chans[1]["tx"])
    assert np.all(ds.phase_history() == np.complex64(1 + 2j))
    with pytest.raises(IndexError):
        Cphd101Dataset(path, channel_index=2)


@pytest.mark.parametrize("fmt", ["RE32F_IM32F", "RE16I_IM16I", "RE08I_IM08I"])
def test_cphd03_round_trip(tmp_path, fmt):
    from sarsolver_amd.io import Cphd03Dataset, Cphd101Dataset
    P, N = 21, 12
    g = np.random.default_rng(2)
    sig = g.integers(-100, 100, (P, N)) + 1j * g.integers(-100, 100, (P, N))
    sf = g.uniform(0.5, 2.0, P)
    tx, rx, srp = _geometry(P, seed=3)
    path = write_cphd03(tmp_path / "c.cphd", [dict(signal=sig, tx=tx, rx=rx, srp=srp, fx0=9.0e9,
                                                    fx_ss=2e6, amp_sf=sf)], sgn=+1, fmt=fmt)
    ds = Cphd03Dataset(path)
    assert ds.num_slow_times == P and ds.num_freqs == N and ds.sign_multiplier == 1.0
    assert np.array_equal(ds.stab_ref_posns, srp) and np.array_equal(ds.receiver_positions, rx)
    assert ds.centre_frequency == 9.0e9 + 0.5 * N * 2e6
    want = (sig.astype(np.complex64) * sf[:, None].astype(np.float32)).astype(np.complex64)
    assert np.array_equal(ds.phase_history(), want)
    with pytest.raises(ValueError):
        Cphd101Dataset(path)


def test_reader_rejects_bad_files(tmp_path):
    from sarsolver_amd.io import open_cphd
    p = tmp_path / "x.cphd"
    p.write_bytes(b"NITF02.10\n")
    with pytest.raises(ValueError):
        open_cphd(p)
    p.write_bytes(b"CPHD/2.5.0\nXML_BLOCK_SIZE := 1\n\f\n")
    with pytest.raises(ValueError):
        open_cphd(p)
    P, N = 3, 4
    tx, rx, srp = _geometry(P)
    good = write_cphd101(tmp_path / "g.cphd", [dict(signal=np.zeros((P, N), np.complex64), tx=tx,
                                                     rx=rx, srp=srp, sc0=1e9, scss=1e6)])
    with pytest.raises(ValueError):  # XML cut short by a wrong block size
        p.write_bytes(good.read_bytes().replace(b"XML_BLOCK_SIZE := ", b"XML_BLOCK_SIZE := 1"))
        open_cphd(p)
    comp = write_cphd101(tmp_path / "z.cphd", [dict(signal=np.zeros((P, N), np.complex64), tx=tx,
                                                     rx=rx, srp=srp, sc0=1e9, scss=1e6)],
                         compressed=True)
    with pytest.raises(ValueError):
        open_cphd(comp)


def test_point_target_through_a_file(tmp_path, oracle_lib):
    """The oracle's phase history of one scatterer, written as CPHD (SC0 chosen
    so that the reference's centre-frequency rule returns fc) and read back,
    backprojects (FP64 oracle, linear) to a peak at the scatterer's pixel."""
    from sarsolver_amd import synthetic
    from sarsolver_amd.io import Cphd101Dataset
    from sarsolver_amd.numpy import GridSarScene
    ap = synthetic.circular_aperture(128, 128, 3.0)
    fc, fs, N = ap.centre_frequency, ap.sample_bandwidth, 128
    sc = GridSarScene((1, 41, 41), (0.0, 8.0, 8.0))
    x0 = sc.positions[16 * 41 + 26][None]  # the pixel at (1.2, -0.8)
    tx, srp = ap.transmitter_positions, ap.stab_ref_posns
    ones = np.ones(N, complex)
    Y = oracle_lib.forward(tx, tx, srp, x0, np.ones(1, complex), ones, np.ones(128, complex), fc, fs,
                           N, upsample_ratio=8.0, interp="linear")
    ph = np.fft.fftshift(Y, axes=1).astype(np.complex64)
    path = write_cphd101(tmp_path / "pt.cphd", [dict(signal=ph, tx=tx, rx=tx, srp=srp,
                                                      sc0=fc - 0.5 * fs, scss=fs / N)], sgn=+1)
    ds = Cphd101Dataset(path)
    rap = ds.aperture
    assert rap.centre_frequency == fc and rap.sample_bandwidth == fs
    img = oracle_lib.adjoint(rap.transmitter_positions, rap.receiver_positions, rap.stab_ref_posns,
                             sc.positions, np.fft.ifftshift(ds.phase_history(), axes=1), ones,
                             np.ones(128, complex), rap.centre_frequency, rap.sample_bandwidth,
                             upsample_ratio=8.0, interp="linear")
    k = int(np.argmax(np.abs(img)))
    assert k == 16 * 41 + 26
    assert np.abs(img[k]) > 0.95 * 128 * N
