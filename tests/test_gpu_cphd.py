"""CPHD signal arrays decoded on the GPU (`sarhip_decode_signal`) against the
host decode, and dataset resampling fed from a CPHD file."""
import numpy as np
import pytest

from cphd_files import write_cphd03, write_cphd101
from conftest import rel_l2

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda", 0)


def _geometry(P):
    th = np.deg2rad(np.linspace(-1.0, 1.0, P))
    tx = np.stack([10e3 * np.cos(th), 10e3 * np.sin(th), np.full(P, 5e3)], 1)
    return tx, tx.copy(), np.zeros((P, 3))


@pytest.mark.parametrize("fmt", ["CF8", "CI4", "CI2"])
@pytest.mark.parametrize("with_sf", [False, True])
def test_device_decode_equals_host(tmp_path, dev, fmt, with_sf):
    """Bit-identical complex64 for every sample format, with and without
    per-vector amplitude scales, on a slice and on the whole channel."""
    from sarsolver_amd.io import Cphd101Dataset
    P, N = 300, 100
    g = np.random.default_rng(4)
    if fmt == "CF8":
        sig = (g.standard_normal((P, N)) + 1j * g.standard_normal((P, N))).astype(np.complex64)
        sig[3, 7] = np.nan
    else:
        hi = 32767 if fmt == "CI4" else 127
        sig = g.integers(-hi - 1, hi + 1, (P, N)) + 1j * g.integers(-hi - 1, hi + 1, (P, N))
    sf = g.uniform(0.1, 3.0, P) if with_sf else None
    tx, rx, srp = _geometry(P)
    path = write_cphd101(tmp_path / "d.cphd", [dict(signal=sig, tx=tx, rx=rx, srp=srp, sc0=9e9,
                                                     scss=1e6, amp_sf=sf)], fmt=fmt)
    ds = Cphd101Dataset(path)
    print(f"case: decode {fmt} amp_sf={with_sf}", flush=True)
    for item in (slice(None), slice(17, 250)):
        d = ds.phase_history_device(item, dev).cpu().numpy()
        h = ds.phase_history(item)
        assert d.dtype == np.complex64 and d.shape == h.shape
        assert np.array_equal(d.view(np.uint32), h.view(np.uint32))


def test_device_decode_many_vectors(tmp_path, dev):
    """70 000 vectors (row launches chunked at 65 535), CPHD 0.3 int16 samples."""
    from sarsolver_amd.io import Cphd03Dataset
    P, N = 70000, 6
    g = np.random.default_rng(5)
    sig = g.integers(-3000, 3000, (P, N)) + 1j * g.integers(-3000, 3000, (P, N))
    tx, rx, srp = _geometry(P)
    path = write_cphd03(tmp_path / "m.cphd", [dict(signal=sig, tx=tx, rx=rx, srp=srp, fx0=9e9,
                                                    fx_ss=1e6, amp_sf=g.uniform(0.5, 2, P))],
                        fmt="RE16I_IM16I")
    ds = Cphd03Dataset(path)
    print("case: decode 70000 x 6 RE16I_IM16I", flush=True)
    d = ds.phase_history_device(slice(None), dev).cpu().numpy()
    assert np.array_equal(d, ds.phase_history())


def test_resampling_from_a_cphd_file(tmp_path, dev):
    """from_resampling on a Cphd101Dataset (device decode of int8 samples)
    equals from_resampling on the same data held in memory."""
    from sarsolver_amd.io import Cphd101Dataset
    from sarsolver_amd.numpy import NumpySimpleSarDataset
    P, N = 400, 256
    g = np.random.default_rng(6)
    sig = g.integers(-100, 100, (P, N)) + 1j * g.integers(-100, 100, (P, N))
    tx, rx, srp = _geometry(P)
    path = write_cphd101(tmp_path / "r.cphd", [dict(signal=sig, tx=tx, rx=rx, srp=srp,
                                                     sc0=9.6e9 - 3e8, scss=6e8 / N)], sgn=+1, fmt="CI2")
    ds = Cphd101Dataset(path)
    mem = NumpySimpleSarDataset(ds.aperture, ds.phase_history())
    print("case: resampling from CPHD 400 x 256", flush=True)
    kw = dict(fast_downsample_factor=2.0, slow_downsample_factor=2, new_srp=np.array([1.0, 2.0, 0.0]),
              device=dev)
    a = NumpySimpleSarDataset.from_resampling(ds, **kw)
    b = NumpySimpleSarDataset.from_resampling(mem, **kw)
    assert a.aperture.num_slow_times == b.aperture.num_slow_times
    assert rel_l2(a.phase_history(), b.phase_history()) == 0.0
