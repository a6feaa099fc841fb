import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (MI355X)")


def rel_l2(a, b) -> float:
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle.oracle as O
    O.build()
    return O


@pytest.fixture(scope="session")
def little():
    """The reference's 'little' problem (tests/conftest.py:46-72 of the reference):
    P=249 linear monostatic track at x=2000 m, N=251, fc=10 GHz, fs=300 MHz,
    scene from GridSarScene.from_aperture(..., [0, 40, 40], safety_factor=3)."""
    from sarsolver_amd.numpy import GridSarScene, NumpySimpleSarAperture
    P, N = 249, 251
    fc, fs, L, R = 10.0e9, 300.0e6, 60.0, 2000.0
    tx = np.linspace(np.array([R, -0.5 * L, 0.0]), np.array([R, 0.5 * L, 0.0]), P)
    ap = NumpySimpleSarAperture(trans_posns=tx, rec_posns=tx.copy(), srps=np.zeros([P, 3]),
                                centre_freq=fc, sample_freq=fs, num_freqs=N)
    sc = GridSarScene.from_aperture(ap, np.array([0.0, 0.0, 0.0]), np.array([0.0, 40.0, 40.0]),
                                    safety_factor=3.0)
    return ap, sc


@pytest.fixture(scope="session")
def little_arrays():
    return {k: np.load(GOLDEN / f"little_{k}.npy", allow_pickle=False)
            for k in ("backproj", "noisy_meas", "gt", "soln")}


LITTLE_NORM = 756.6231381576064  # reference tests/conftest.py:95
