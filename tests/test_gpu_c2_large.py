"""Parity on the benchmarked workload (1024^2 px at 0.1 m, 2048 pulses x 4096
samples, 4 degree circular monostatic aperture, U = 8, linear interpolation,
series ranges, scene of 500 point scatterers plus a 64 x 64 clutter patch,
seed 1002), with the plan the benchmark builds.

The plan's layout is the one the benchmark times (63-bin forward windows of
2-tile regions, 64-pulse groups, third-order gather records read with
demodulated pair rows, fused spectral stages at U = 8); the test asserts it,
so a layout change shows up here
before it reaches the benchmark unchecked.

The FP64 oracle evaluates exact subsets of the outputs: every output pixel of
A^H depends on no other output pixel (the 500 scatterer pixels, the four
corners and centre of the clutter patch, and 2048 random pixels), and every
output pulse of A depends on no other pulse (32 pulses spread over the
aperture, for the whole image).  Bar: relative L2 <= 1e-5 for both operators,
on the scene's phase history (A^H), on the dense complex Gaussian image the
benchmark's forward is timed on (A), and on the scene itself (A); and the
dot-product test of the whole operator pair (linear interpolation is an exact
adjoint pair) at 1e-6.
"""
import numpy as np
import pytest

from conftest import rel_l2
from gpu_helpers import geometry

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

BENCH_LAYOUT = {"precision": "series", "monostatic": True, "valid_pulses": 2048,
                "fwd_window_bins": 63, "fwd_pulses_per_group": 64, "fwd_region_tiles": 2,
                "bp_series_order": 3, "fused_fft_ratio": 8, "bp_pairs": True}


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda", 0)


@pytest.fixture(scope="module")
def c2(dev, oracle_lib):
    from sarsolver_amd import synthetic
    w = synthetic.WORKLOADS["c2_large"]
    print(f"case: c2_large parity {w}", flush=True)
    rng = np.random.default_rng(w.seed)
    ap, sc = synthetic.build(w)
    pos = sc.positions
    x = synthetic.scene_amplitudes(w, rng).reshape(-1)
    xdense = synthetic.dense_image(w, pos.shape[0])
    pl = synthetic.workload_plan(w, ap, pos, dev)
    lay = pl.layout
    xs = torch.from_numpy(x.astype(np.complex64)).to(dev)
    y = pl.forward(xs)                              # the scene's phase history
    X = pl.adjoint(y).cpu().numpy()
    Yd = pl.forward(torch.from_numpy(xdense.astype(np.complex64)).to(dev)).cpu().numpy()
    # the dense image does not cancel (kappa ~0.8): no row needs the FP64 pass
    dense_fp64_rows = pl.forward_fp64_rows()
    Ys = y.cpu().numpy()
    # dot-product test at the full size (linear interpolation is an exact
    # adjoint pair): <A x, y> against <x, A^H y> in FP64 for the dense image
    # and a seeded complex Gaussian phase history
    yr = synthetic.complex_normal(np.random.default_rng(w.seed + 7), (w.pulses, w.samples))
    Ayr = pl.adjoint(torch.from_numpy(yr.astype(np.complex64)).to(dev)).cpu().numpy()
    xd64 = xdense.astype(np.complex64).astype(np.complex128)
    lhs = np.vdot(yr.astype(np.complex64).astype(np.complex128), Yd.astype(np.complex128))
    rhs = np.vdot(Ayr.astype(np.complex128), xd64)
    dot = abs(lhs - rhs) / (np.linalg.norm(Yd) * np.linalg.norm(yr.astype(np.complex64)))
    del pl
    torch.cuda.empty_cache()
    # pixel subset: the 500 scatterers (the workload's first draw), the clutter
    # patch's corners and centre (its rows are the only ones with >= 64
    # non-zero pixels), and 2048 random pixels
    n = w.n
    point_pix = np.random.default_rng(w.seed).choice(n * n, size=500, replace=False)
    nzr = np.count_nonzero(x.reshape(n, n), axis=1)
    prow = np.flatnonzero(nzr >= 64)
    r0, r1 = prow.min(), prow.max()
    pcol = np.flatnonzero(np.count_nonzero(x.reshape(n, n)[r0:r1 + 1], axis=0) >= 32)
    c0, c1 = pcol.min(), pcol.max()
    assert r1 - r0 == 63 and c1 - c0 == 63
    corners = [r0 * n + c0, r0 * n + c1, r1 * n + c0, r1 * n + c1,
               ((r0 + r1) // 2) * n + (c0 + c1) // 2]
    others = rng.choice(np.setdiff1d(np.arange(x.size), point_pix), size=2048, replace=False)
    pix = np.unique(np.concatenate([point_pix, corners, others]))
    rows = np.linspace(0, w.pulses - 1, 32).round().astype(int)
    tx, rx, srp = geometry(ap)
    W = np.fft.ifftshift(ap.waveform_fft)
    wts = np.asarray(ap.slow_time_weighting)
    args = (ap.centre_frequency, ap.sample_bandwidth)
    print("oracle subsets ...", flush=True)
    y_fft = Ys.astype(np.complex128)  # the plan's default order is FFT order
    X_or = oracle_lib.adjoint(tx, rx, srp, pos[pix], y_fft, W, wts, *args, c_eff=ap.c_eff,
                              upsample_ratio=w.upsample, interp="linear", num_threads=4)
    Yd_or = oracle_lib.forward(tx[rows], rx[rows], srp[rows], pos,
                               xdense.astype(np.complex64).astype(np.complex128), W, wts[rows],
                               *args, w.samples, c_eff=ap.c_eff, upsample_ratio=w.upsample,
                               interp="linear", num_threads=8)
    Ys_or = oracle_lib.forward(tx[rows], rx[rows], srp[rows], pos,
                               x.astype(np.complex64).astype(np.complex128), W, wts[rows],
                               *args, w.samples, c_eff=ap.c_eff, upsample_ratio=w.upsample,
                               interp="linear", num_threads=8)
    print("oracle done", flush=True)
    return dict(w=w, lay=lay, X=X[pix], X_or=X_or, Yd=Yd[rows], Yd_or=Yd_or,
                Ys=Ys[rows], Ys_or=Ys_or, dense_fp64_rows=dense_fp64_rows, dot=dot)


def test_c2_plan_layout_is_the_benchmarked_one(c2):
    lay = c2["lay"]
    got = {k: lay[k] for k in BENCH_LAYOUT}
    print(f"c2 layout: {lay}")
    assert got == BENCH_LAYOUT
    assert lay["series_err_bins"] <= 1e-9 and lay["bp_series_err_bins"] <= 1e-9
    # the adaptive spectral stage is on, and the benchmarked dense forward
    # runs every row in FP32
    assert lay["fwd_fft_kappa"] == 2.0
    print(f"c2 dense forward: {c2['dense_fp64_rows']} rows redone in FP64")
    assert c2["dense_fp64_rows"] == 0


def test_c2_adjoint_pixel_subset_against_oracle(c2):
    e = rel_l2(c2["X"], c2["X_or"])
    print(f"c2 A^H on {c2['X'].size} pixels: rel L2 {e:.2e}")
    assert e <= 1e-5


def test_c2_forward_dense_image_pulse_rows_against_oracle(c2):
    e = rel_l2(c2["Yd"], c2["Yd_or"])
    print(f"c2 A (dense image) on 32 pulses: rel L2 {e:.2e}")
    assert e <= 1e-5


def test_c2_forward_scene_pulse_rows_against_oracle(c2):
    e = rel_l2(c2["Ys"], c2["Ys_or"])
    print(f"c2 A (scene) on 32 pulses: rel L2 {e:.2e}")
    assert e <= 1e-5


def test_c2_dot_product_at_full_size(c2):
    """|<A x, y> - <x, A^H y>| / (||A x|| ||y||) <= 1e-6 on the benchmarked plan."""
    print(f"c2 dot-product test: {c2['dot']:.2e}")
    assert c2["dot"] <= 1e-6


def test_c2_reference_rules_against_oracle(dev, oracle_lib):
    """The reference's own bin rules (`nearest_ref`: forward round(u), adjoint
    round(1/2 + u)) on the same workload, which run on the nearest-rule paths
    of the demodulated rows and the window-relative scatter: A^H of a seeded
    complex Gaussian phase history on 2048 random pixels and A of the dense
    image on 32 pulse rows against the FP64 oracle at 1e-5."""
    from sarsolver_amd import synthetic
    w = synthetic.WORKLOADS["c2_large"]
    print("case: c2_large nearest_ref parity", flush=True)
    ap, sc = synthetic.build(w)
    pos = sc.positions
    xdense = synthetic.dense_image(w, pos.shape[0])
    yr = synthetic.complex_normal(np.random.default_rng(w.seed + 11), (w.pulses, w.samples))
    pl = synthetic.workload_plan(w, ap, pos, dev, interp="nearest_ref")
    lay = pl.layout
    print(f"c2 nearest_ref layout: {lay}")
    assert lay["bp_pairs"] and lay["bp_series_order"] == 3 and lay["fwd_window_bins"] == 63
    X = pl.adjoint(torch.from_numpy(yr.astype(np.complex64)).to(dev)).cpu().numpy()
    Yd = pl.forward(torch.from_numpy(xdense.astype(np.complex64)).to(dev)).cpu().numpy()
    del pl
    torch.cuda.empty_cache()
    pix = np.sort(np.random.default_rng(w.seed + 12).choice(pos.shape[0], size=2048, replace=False))
    rows = np.linspace(0, w.pulses - 1, 32).round().astype(int)
    tx, rx, srp = geometry(ap)
    W = np.fft.ifftshift(ap.waveform_fft)
    wts = np.asarray(ap.slow_time_weighting)
    args = (ap.centre_frequency, ap.sample_bandwidth)
    print("oracle subsets ...", flush=True)
    X_or = oracle_lib.adjoint(tx, rx, srp, pos[pix], yr.astype(np.complex64).astype(np.complex128),
                              W, wts, *args, c_eff=ap.c_eff, upsample_ratio=w.upsample,
                              interp="nearest_ref", num_threads=4)
    Yd_or = oracle_lib.forward(tx[rows], rx[rows], srp[rows], pos,
                               xdense.astype(np.complex64).astype(np.complex128), W, wts[rows],
                               *args, w.samples, c_eff=ap.c_eff, upsample_ratio=w.upsample,
                               interp="nearest_ref", num_threads=8)
    ea, ef = rel_l2(X[pix], X_or), rel_l2(Yd[rows], Yd_or)
    print(f"c2 nearest_ref: A^H on {pix.size} pixels {ea:.2e}, A (dense) on 32 pulses {ef:.2e}")
    assert ea <= 1e-5 and ef <= 1e-5
