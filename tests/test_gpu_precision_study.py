"""Precision study on the 2048^2 workload (0.05 m pixels, 4096 pulses x 8192
samples, 8 degree aperture at 10 km standoff and 5 km height, U = 8, linear
interpolation, 2000 point scatterers with CN(0,1) amplitudes, seed 1003).

Each GPU precision mode is run on the whole problem and compared with the
FP64 oracle on exact subsets the oracle finishes in seconds: every output
pixel of A^H is independent of the others, so the oracle evaluates 4048 of
them (the 2000 scatterer pixels, where the image peaks, and 2048 random
ones); every output pulse of A is independent of the others, so the oracle
evaluates 32 pulses spread over the aperture for the full image.  Reported
per mode: relative L2 error and peak phase error (radians; scatterer pixels
for A^H, samples above 10% of the peak magnitude for A).

fp64   : FP64 ranges as the reference evaluates them, FP32 products and sums
         (64-pulse FP32 blocks folded into FP64)           -> rel L2 <= 1e-5
series : FP64 series about 2^-10 m-snapped tile centres    -> rel L2 <= 1e-5
hilo   : FP64 tile-centre range + FP32 differential range  -> reported (<= 1e-3)
fp32   : FP32 absolute coordinates and ranges              -> reported only
"""
import numpy as np
import pytest

from conftest import rel_l2
from gpu_helpers import geometry, oracle_adjoint, oracle_forward, plan

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

MODES = ("fp64", "series", "hilo", "fp32")
LIMITS = {"fp64": 1e-5, "series": 1e-5, "hilo": 1e-3, "fp32": None}


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda", 0)


@pytest.fixture(scope="module")
def study(dev, oracle_lib):
    from sarsolver_amd import synthetic
    w = synthetic.WORKLOADS["c3_precision"]
    print(f"precision study: {w}", flush=True)
    rng = np.random.default_rng(w.seed)
    ap, sc = synthetic.build(w)
    pos = sc.positions
    x = synthetic.scene_amplitudes(w, rng).reshape(-1)
    scat = np.flatnonzero(x)
    others = rng.choice(np.setdiff1d(np.arange(x.size), scat), size=2048, replace=False)
    pix = np.concatenate([scat, others])
    rows = np.linspace(0, w.pulses - 1, 32).round().astype(int)
    # common adjoint input: the FP64-mode forward of the scene
    p64 = plan(ap, pos, w.upsample, "linear", "fp64", dev=dev)
    xd = torch.from_numpy(x.astype(np.complex64)).to(dev)
    y = p64.forward(xd).cpu().numpy()
    del p64
    torch.cuda.empty_cache()
    print("oracle subsets ...", flush=True)
    tx, rx, srp = geometry(ap)
    X_or = oracle_lib.adjoint(tx, rx, srp, pos[pix], y.astype(np.complex128),
                              np.fft.ifftshift(ap.waveform_fft), ap.slow_time_weighting,
                              ap.centre_frequency, ap.sample_bandwidth, c_eff=ap.c_eff,
                              upsample_ratio=w.upsample, interp="linear", num_threads=15)
    Y_or = oracle_lib.forward(tx[rows], rx[rows], srp[rows], pos, x,
                              np.fft.ifftshift(ap.waveform_fft),
                              np.asarray(ap.slow_time_weighting)[rows], ap.centre_frequency,
                              ap.sample_bandwidth, w.samples, c_eff=ap.c_eff,
                              upsample_ratio=w.upsample, interp="linear", num_threads=15)
    print("oracle done", flush=True)
    return dict(w=w, ap=ap, pos=pos, x=x, y=y, pix=pix, nscat=scat.size, rows=rows,
                X_or=X_or, Y_or=Y_or)


def _peak_phase(a, b):
    return float(np.max(np.abs(np.angle(a * np.conj(b))))) if a.size else 0.0


@pytest.mark.parametrize("mode", MODES)
def test_precision_mode_against_fp64_oracle(dev, study, mode):
    s = study
    w, ap = s["w"], s["ap"]
    print(f"case: precision study {mode}", flush=True)
    pl = plan(ap, s["pos"], w.upsample, "linear", mode, dev=dev)
    assert pl.effective_precision == mode
    xd = torch.from_numpy(s["x"].astype(np.complex64)).to(dev)
    yd = torch.from_numpy(s["y"].astype(np.complex64)).to(dev)
    X = pl.adjoint(yd).cpu().numpy()[s["pix"]]
    Y = pl.forward(xd).cpu().numpy()[s["rows"]]
    del pl
    torch.cuda.empty_cache()
    ea, ef = rel_l2(X, s["X_or"]), rel_l2(Y, s["Y_or"])
    ns = s["nscat"]
    pa = _peak_phase(X[:ns], s["X_or"][:ns])
    strong = np.abs(s["Y_or"]) >= 0.1 * np.abs(s["Y_or"]).max()
    pf = _peak_phase(Y[strong], s["Y_or"][strong])
    print(f"| {mode:6s} | A^H rel L2 {ea:.2e} | A^H peak phase {pa:.2e} rad "
          f"| A rel L2 {ef:.2e} | A peak phase {pf:.2e} rad |")
    lim = LIMITS[mode]
    assert np.isfinite(ea) and np.isfinite(ef)
    if lim is not None:
        assert ea <= lim and ef <= lim
