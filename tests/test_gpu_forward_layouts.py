"""Every forward-scatter layout against the FP64 oracle.

The plan picks one layout per geometry: 64-pulse workgroups of 512 threads
with wave-uniform pixel streams when the LDS windows fit in 64 KiB, else 32
or 16 pulses per 256-thread workgroup; regions of 1..16 tiles; global float
atomics when no window fits.  The workload-driven tests see only the layout
their geometry selects, so each one is forced here (through the plan's
environment overrides, read at plan creation) on a 128^2 grid at 0.1 m under
the 256-pulse x 512-sample, 2 deg aperture (U = 8; windows of ~100 bins, so
that every layout fits) with a dense complex Gaussian image and with a sparse
one (so that all-zero regions are skipped), in the linear and
reference-nearest modes."""
import numpy as np
import pytest

from conftest import rel_l2
from gpu_helpers import oracle_forward, plan

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

LAYOUTS = {
    "default": {},
    "g64": {"SARHIP_FP_PULSES": "64"},
    "g32": {"SARHIP_FP_PULSES": "32"},
    "g16": {"SARHIP_FP_PULSES": "16"},
    "g32_r2": {"SARHIP_FP_PULSES": "32", "SARHIP_FP_REGION": "2", "SARHIP_FP_LDS_KB": "64"},
    "g16_r4": {"SARHIP_FP_PULSES": "16", "SARHIP_FP_REGION": "4", "SARHIP_FP_LDS_KB": "64"},
    "atomics": {"SARHIP_FP_NO_WINDOWS": "1"},
}
ENV = ("SARHIP_FP_PULSES", "SARHIP_FP_REGION", "SARHIP_FP_LDS_KB", "SARHIP_FP_NO_WINDOWS")


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda", 0)


@pytest.fixture(scope="module")
def case(oracle_lib):
    from sarsolver_amd import synthetic
    w = synthetic.WORKLOADS["c1_small"]
    ap = synthetic.circular_aperture(w.pulses, w.samples, w.aperture_deg)
    sc = synthetic.grid_scene(w.n, 0.1)
    rng = np.random.default_rng(w.seed + 50)
    dense = synthetic.complex_normal(rng, sc.positions.shape[0])
    sparse = np.zeros_like(dense)
    idx = rng.choice(dense.size, size=40, replace=False)
    sparse[idx] = dense[idx]
    ref = {}
    for interp in ("linear", "nearest_ref"):
        for name, x in (("dense", dense), ("sparse", sparse)):
            ref[(interp, name)] = oracle_forward(oracle_lib, ap, sc.positions, x, w.upsample, interp)
    return w, ap, sc, {"dense": dense, "sparse": sparse}, ref


@pytest.mark.parametrize("layout", list(LAYOUTS))
@pytest.mark.parametrize("interp", ["linear", "nearest_ref"])
def test_forward_layout_matches_oracle(dev, case, monkeypatch, layout, interp):
    w, ap, sc, xs, ref = case
    for k in ENV:
        monkeypatch.delenv(k, raising=False)
    for k, v in LAYOUTS[layout].items():
        monkeypatch.setenv(k, v)
    print(f"case: forward layout {layout} {interp}", flush=True)
    pl = plan(ap, sc.positions, w.upsample, interp, "series", dev=dev)
    lay = pl.layout
    want = LAYOUTS[layout]
    if "SARHIP_FP_NO_WINDOWS" in want:
        assert lay["fwd_window_bins"] == 0
    else:
        assert lay["fwd_window_bins"] > 0
        if "SARHIP_FP_PULSES" in want:
            assert lay["fwd_pulses_per_group"] == int(want["SARHIP_FP_PULSES"])
        if "SARHIP_FP_REGION" in want:
            assert lay["fwd_region_tiles"] == int(want["SARHIP_FP_REGION"])
    for name, x in xs.items():
        Y = pl.forward(torch.from_numpy(x.astype(np.complex64)).to(dev)).cpu().numpy()
        e = rel_l2(Y, ref[(interp, name)])
        print(f"  {layout} {interp} {name}: rel L2 {e:.2e} (layout {lay})")
        assert e <= 1e-5


@pytest.mark.parametrize("splits", ["1", "3", "7", "default"])
def test_backprojection_pulse_splits_match_oracle(dev, case, oracle_lib, monkeypatch, splits):
    """The backprojection splits each tile's pulses over several workgroups
    (FP64 partial images summed in split order); every split count, including
    ones that do not divide the pulse count, gives the oracle's image."""
    from gpu_helpers import oracle_adjoint
    w, ap, sc, xs, ref = case
    monkeypatch.delenv("SARHIP_BP_SPLITS", raising=False)
    if splits != "default":
        monkeypatch.setenv("SARHIP_BP_SPLITS", splits)
    print(f"case: backprojection splits {splits}", flush=True)
    rng = np.random.default_rng(w.seed + 51)
    y = rng.standard_normal((w.pulses, w.samples)) + 1j * rng.standard_normal((w.pulses, w.samples))
    for interp in ("linear", "nearest_ref"):
        pl = plan(ap, sc.positions, w.upsample, interp, "series", dev=dev)
        X = pl.adjoint(torch.from_numpy(y.astype(np.complex64)).to(dev)).cpu().numpy()
        Xo = oracle_adjoint(oracle_lib, ap, sc.positions, y, w.upsample, interp)
        e = rel_l2(X, Xo)
        print(f"  splits {splits} {interp}: rel L2 {e:.2e}")
        assert e <= 1e-5
