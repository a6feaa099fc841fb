"""GPU precision modes against the FP64 oracle.

fp64   : ranges as the reference evaluates them -> rel L2 <= 1e-5, bins identical.
series : FP64 series about tile centres, exact re-evaluation at bin ties ->
         rel L2 <= 1e-5 in every interpolation mode, including tie-heavy inputs.
hilo   : FP64 tile-centre range + FP32 differential range (precision study).
fp32   : all-FP32 absolute ranges (precision study; fails at 10 km standoff).
"""
import numpy as np
import pytest

from conftest import rel_l2
from gpu_helpers import c1_inputs, oracle_adjoint, oracle_forward, plan

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda", 0)


def to_dev(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.complex64)).to(dev)


@pytest.fixture(scope="module")
def c1_oracle(oracle_lib):
    w, ap, sc, x, y = c1_inputs()
    pos = sc.positions
    out = {}
    for interp in ("nearest_ref", "nearest", "linear"):
        out[interp] = (oracle_forward(oracle_lib, ap, pos, x, w.upsample, interp),
                       oracle_adjoint(oracle_lib, ap, pos, y, w.upsample, interp))
    return w, ap, sc, x, y, out


@pytest.mark.parametrize("interp", ["nearest_ref", "nearest", "linear"])
@pytest.mark.parametrize("precision", ["fp64", "series"])
def test_exact_modes_match_oracle(dev, c1_oracle, interp, precision):
    w, ap, sc, x, y, ref = c1_oracle
    print(f"case: c1 {interp} {precision}", flush=True)
    pl = plan(ap, sc.positions, w.upsample, interp, precision, dev=dev)
    assert pl.effective_precision == precision
    Y = pl.forward(to_dev(x, dev)).cpu().numpy()
    X = pl.adjoint(to_dev(y, dev)).cpu().numpy()
    ef, ea = rel_l2(Y, ref[interp][0]), rel_l2(X, ref[interp][1])
    print(f"c1 {interp} {precision}: forward {ef:.2e} adjoint {ea:.2e}")
    assert ef <= 1e-5 and ea <= 1e-5


@pytest.mark.parametrize("precision", ["hilo", "fp32"])
def test_study_modes_report_error(dev, c1_oracle, precision):
    """The reduced-precision study modes run in both directions; their error against
    the FP64 oracle is reported (hilo stays usable, all-FP32 does not)."""
    w, ap, sc, x, y, ref = c1_oracle
    pl = plan(ap, sc.positions, w.upsample, "linear", precision, dev=dev)
    Y = pl.forward(to_dev(x, dev)).cpu().numpy()
    X = pl.adjoint(to_dev(y, dev)).cpu().numpy()
    ef, ea = rel_l2(Y, ref["linear"][0]), rel_l2(X, ref["linear"][1])
    print(f"c1 linear {precision}: forward {ef:.2e} adjoint {ea:.2e}")
    assert np.all(np.isfinite(Y)) and np.all(np.isfinite(X))
    if precision == "hilo":
        assert ef <= 1e-2 and ea <= 1e-2
    else:
        assert ef > 1e-2 and ea > 1e-2


def _tie_problem():
    """Monostatic pulses at (1000, 0, 0), SRP at the origin, c = fs = 1, U = 4, so
    by_dx = 4 and pixels on the x-axis at multiples of 1/16 m have u = -8x: every
    other pixel sits exactly on a rounding tie of the forward or adjoint rule."""
    P, N, U = 4, 16, 4.0
    tx = np.tile(np.array([[1000.0, 0.0, 0.0]]), (P, 1))
    xs = np.arange(-64, 64) / 16.0
    pos = np.stack([xs, np.zeros_like(xs), np.zeros_like(xs)], 1)
    rng = np.random.default_rng(17)
    a = rng.standard_normal(len(xs)) + 1j * rng.standard_normal(len(xs))
    y = rng.standard_normal((P, N)) + 1j * rng.standard_normal((P, N))
    W = np.ones(N, complex)
    w = np.exp(1j * np.arange(P))
    return tx, pos, a, y, W, w, N, U


@pytest.mark.parametrize("precision", ["fp64", "series"])
@pytest.mark.parametrize("interp", ["nearest_ref", "nearest"])
def test_exact_ties_follow_the_reference_rule(dev, oracle_lib, precision, interp):
    from sarsolver_amd import SarPlan
    tx, pos, a, y, W, w, N, U = _tie_problem()
    fc, fs, c = 0.37, 1.0, 1.0
    pl = SarPlan(tx, tx, np.zeros_like(tx), pos, W, w, fc, fs, c, U, 1.0, interp=interp,
                 precision=precision, device=dev)
    Y = pl.forward(to_dev(a, dev)).cpu().numpy()
    X = pl.adjoint(to_dev(y, dev)).cpu().numpy()
    Yo = oracle_lib.forward(tx, tx, np.zeros_like(tx), pos, a, W, w, fc, fs, N, c_eff=c,
                            upsample_ratio=U, interp=interp)
    Xo = oracle_lib.adjoint(tx, tx, np.zeros_like(tx), pos, y, W, w, fc, fs, c_eff=c,
                            upsample_ratio=U, interp=interp)
    print(f"ties {interp} {precision} ({pl.effective_precision}): "
          f"forward {rel_l2(Y, Yo):.2e} adjoint {rel_l2(X, Xo):.2e}")
    assert rel_l2(Y, Yo) <= 1e-5 and rel_l2(X, Xo) <= 1e-5


@pytest.mark.parametrize("interp", ["nearest_ref", "nearest"])
def test_exact_ties_on_the_pair_path(dev, oracle_lib, interp):
    """The tie problem moved to 100 km (third-order series bound ~1e-12 bin) with
    N = 64 samples (fused spectral stages, N_w = 256) and f_c = 1/2, so that
    rho = f_c / (f_s U) = 1/8 is an FP32 number: the series plan runs the
    nearest rules on the gather's pair rows and the scatter's window-relative
    path.  Pixels at multiples of 1/16 m have u = -8x exactly, so every pixel
    sits on a tie of the forward rule round(u) or of the adjoint rule
    (round(1/2 + u) or round(u)), and each tie must be decided as the
    reference decides it."""
    from sarsolver_amd import SarPlan
    P, N, U = 4, 64, 4.0
    tx = np.tile(np.array([[1.0e5, 0.0, 0.0]]), (P, 1))
    xs = np.arange(-64, 64) / 16.0
    pos = np.stack([xs, np.zeros_like(xs), np.zeros_like(xs)], 1)
    rng = np.random.default_rng(29)
    a = rng.standard_normal(len(xs)) + 1j * rng.standard_normal(len(xs))
    y = rng.standard_normal((P, N)) + 1j * rng.standard_normal((P, N))
    W = np.ones(N, complex)
    w = np.exp(1j * np.arange(P))
    fc, fs, c = 0.5, 1.0, 1.0
    pl = SarPlan(tx, tx, np.zeros_like(tx), pos, W, w, fc, fs, c, U, 1.0, interp=interp,
                 precision="series", device=dev)
    lay = pl.layout
    assert lay["precision"] == "series" and lay["bp_series_order"] == 3 and lay["bp_pairs"], lay
    Y = pl.forward(to_dev(a, dev)).cpu().numpy()
    X = pl.adjoint(to_dev(y, dev)).cpu().numpy()
    Yo = oracle_lib.forward(tx, tx, np.zeros_like(tx), pos, a, W, w, fc, fs, N, c_eff=c,
                            upsample_ratio=U, interp=interp)
    Xo = oracle_lib.adjoint(tx, tx, np.zeros_like(tx), pos, y, W, w, fc, fs, c_eff=c,
                            upsample_ratio=U, interp=interp)
    # a tie decided the other way takes a neighbouring, independent sample: an
    # error of the order of the pixel's value
    ex = np.abs(X - Xo) / np.abs(Xo).max()
    print(f"pair-path ties {interp}: forward {rel_l2(Y, Yo):.2e} adjoint {rel_l2(X, Xo):.2e} "
          f"(worst pixel {ex.max():.2e})")
    assert rel_l2(Y, Yo) <= 1e-5 and rel_l2(X, Xo) <= 1e-5 and ex.max() <= 1e-5


def test_series_falls_back_when_unbounded(dev, little):
    """At U = 128 on the little problem the series bound still holds (5th order) or
    the plan reports fp64; either way the reported mode is one of the exact ones."""
    ap, sc = little
    pl = plan(ap, sc.positions, 128.0, "nearest_ref", "series", dev=dev)
    assert pl.effective_precision in ("series", "fp64")
    print("little series layout:", pl.layout)


@pytest.mark.parametrize("N", [32, 64])
def test_nonfinite_pixels_and_pulses(dev, oracle_lib, N):
    """NaN pixel positions contribute nothing and backproject to 0; pulses with NaN
    geometry are skipped (their phase-history rows are 0), as the oracle does.
    N = 32 runs the rocFFT stages, N = 64 the fused ones (whose row spans leave
    a skipped pulse's row untouched); the forward is repeated after the adjoint
    has written the rows."""
    from sarsolver_amd import SarPlan, synthetic
    ap = synthetic.circular_aperture(16, N, 2.0)
    sc = synthetic.grid_scene(16, 0.25)
    pos = sc.positions.copy()
    pos[5] = np.nan
    tx = ap.transmitter_positions.copy()
    tx[3, 1] = np.nan
    rng = np.random.default_rng(3)
    a = rng.standard_normal(len(pos)) + 1j * rng.standard_normal(len(pos))
    y = rng.standard_normal((16, N)) + 1j * rng.standard_normal((16, N))
    W, w = np.ones(N, complex), np.ones(16, complex)
    for precision in ("fp64", "series"):
        pl = SarPlan(tx, tx, ap.stab_ref_posns, pos, W, w, ap.centre_frequency,
                     ap.sample_bandwidth, ap.c_eff, 8.0, 1.0, interp="linear",
                     precision=precision, device=dev)
        assert pl.layout["fused_fft_ratio"] == (8 if N == 64 else 0)
        Y = pl.forward(to_dev(a, dev)).cpu().numpy()
        X = pl.adjoint(to_dev(y, dev)).cpu().numpy()
        Y2 = pl.forward(to_dev(a, dev)).cpu().numpy()
        Yo = oracle_lib.forward(tx, tx, ap.stab_ref_posns, pos, a, W, w, ap.centre_frequency,
                                ap.sample_bandwidth, N, upsample_ratio=8.0, interp="linear")
        Xo = oracle_lib.adjoint(tx, tx, ap.stab_ref_posns, pos, y, W, w, ap.centre_frequency,
                                ap.sample_bandwidth, upsample_ratio=8.0, interp="linear")
        assert np.all(Y[3] == 0) and X[5] == 0 and np.array_equal(Y, Y2)
        assert rel_l2(Y, Yo) <= 1e-5 and rel_l2(X, Xo) <= 1e-5


@pytest.mark.parametrize("interp", ["nearest_ref", "linear"])
def test_dense_scene_forward_fixed_point(dev, oracle_lib, interp):
    """The forward scatter accumulates LDS windows in 32-bit fixed point; a dense
    random-phase scene (smoothed complex Gaussian field with point scatterers, the
    hardest case for the fixed-point bounds) must still match the oracle."""
    from sarsolver_amd import synthetic
    w = synthetic.WORKLOADS["c1_small"]
    ap, sc = synthetic.build(w)
    rng = np.random.default_rng(44)
    x = (synthetic.smooth_field(rng, w.n, 3.0) + 30 * synthetic.point_scatterers(rng, w.n, 5))
    x = x.reshape(-1)
    pl = plan(ap, sc.positions, w.upsample, interp, "series", dev=dev)
    Y = pl.forward(to_dev(x, dev)).cpu().numpy()
    Yo = oracle_forward(oracle_lib, ap, sc.positions, x, w.upsample, interp)
    print(f"dense {interp}: forward {rel_l2(Y, Yo):.2e}")
    assert rel_l2(Y, Yo) <= 1e-5


@pytest.mark.parametrize("fused", [True, False])
def test_nonfinite_amplitudes_poison_the_forward(dev, oracle_lib, fused):
    """A non-finite amplitude makes every row of the reference's forward
    non-finite (its FFT spreads the value over the row).  The scatter cannot
    window such a region: with the fused spectral stage (row spans, rows not
    cleared) it poisons the row and the output row is NaN; with rocFFT it adds
    global atomics into the cleared rows.  Either way every valid row comes
    out non-finite, and the next call with finite amplitudes is exact again
    (the poison flags are cleared)."""
    import os
    from sarsolver_amd import SarPlan, synthetic
    ap = synthetic.circular_aperture(16, 64, 2.0)
    sc = synthetic.grid_scene(64, 0.25)
    rng = np.random.default_rng(8)
    a = rng.standard_normal(sc.num_scatterers) + 1j * rng.standard_normal(sc.num_scatterers)
    W, w = np.ones(64, complex), np.ones(16, complex)
    os.environ["SARHIP_FUSED_FFT"] = "1" if fused else "0"
    try:
        pl = SarPlan(ap.transmitter_positions, ap.receiver_positions, ap.stab_ref_posns,
                     sc.positions, W, w, ap.centre_frequency, ap.sample_bandwidth, ap.c_eff, 8.0,
                     1.0, interp="linear", precision="series", device=dev)
    finally:
        os.environ.pop("SARHIP_FUSED_FFT", None)
    assert pl.layout["fused_fft_ratio"] == (8 if fused else 0) and pl.layout["fwd_window_bins"] > 0
    print(f"case: non-finite amplitudes fused={fused}", flush=True)
    bad = a.copy()
    bad[100] = np.nan
    Yb = pl.forward(to_dev(bad, dev)).cpu().numpy()
    assert not np.any(np.isfinite(Yb).all(axis=1))
    Y = pl.forward(to_dev(a, dev)).cpu().numpy()
    Yo = oracle_lib.forward(ap.transmitter_positions, ap.receiver_positions, ap.stab_ref_posns,
                            sc.positions, a, W, w, ap.centre_frequency, ap.sample_bandwidth, 64,
                            upsample_ratio=8.0, interp="linear")
    assert rel_l2(Y, Yo) <= 1e-5


@pytest.mark.parametrize("interp", ["nearest_ref", "linear"])
@pytest.mark.parametrize("exp2", [-110, -60, 60, 100])
def test_forward_is_linear_at_extreme_scales(dev, interp, exp2):
    """A(c x) = c A(x) for c = 2^k far from 1.  The fixed-point unit of the
    scatter's windows and of the merge follows the input's magnitude (scale
    m 2^e with m in (1, 2]), so a power-of-two c only moves e: the integer sums
    are identical and the outputs differ by c exactly.  (A scale formed as
    2^31 / bound overflows FP32 for bounds below ~6e-30, which is where
    c = 2^-110 puts the regions of this image.)"""
    from sarsolver_amd import synthetic
    w = synthetic.WORKLOADS["c1_small"]
    ap, sc = synthetic.build(w)
    x = synthetic.dense_image(w, sc.num_scatterers).astype(np.complex64)
    pl = plan(ap, sc.positions, w.upsample, interp, "series", dev=dev)
    c = np.float32(2.0 ** exp2)
    print(f"case: linearity at 2^{exp2} {interp}", flush=True)
    Y = pl.forward(to_dev(x, dev)).cpu().numpy()
    Yc = pl.forward(to_dev(x * c, dev)).cpu().numpy()
    assert np.all(np.isfinite(Yc))
    e = rel_l2(Yc.astype(np.complex128) / float(c), Y)
    print(f"2^{exp2} {interp}: rel L2 of A(cx)/c - A(x) = {e:.2e}")
    assert e <= 1e-6


@pytest.mark.parametrize("interp", ["nearest_ref", "linear"])
def test_forward_after_adjoint_odd_profile_length(dev, oracle_lib, interp):
    """P N_w odd (P = 7, N = 33, U = 3: 693 profile bins, rocFFT stages with
    rows cleared before the scatter): the forward that follows an adjoint on
    the same plan must not see the adjoint's values in the last bin."""
    from sarsolver_amd import synthetic
    P, N, U = 7, 33, 3.0
    ap = synthetic.circular_aperture(P, N, 2.0)
    sc = synthetic.grid_scene(16, 0.25)
    rng = np.random.default_rng(21)
    a = rng.standard_normal(sc.num_scatterers) + 1j * rng.standard_normal(sc.num_scatterers)
    y = rng.standard_normal((P, N)) + 1j * rng.standard_normal((P, N))
    pl = plan(ap, sc.positions, U, interp, "fp64", dev=dev)
    assert pl.layout["fused_fft_ratio"] == 0 and (P * pl.working_num_fast_times) % 2 == 1
    print(f"case: adjoint then forward, odd P N_w, {interp}", flush=True)
    X = pl.adjoint(to_dev(y * 1e3, dev)).cpu().numpy()
    Y = pl.forward(to_dev(a, dev)).cpu().numpy()
    Yo = oracle_forward(oracle_lib, ap, sc.positions, a, U, interp)
    Xo = oracle_adjoint(oracle_lib, ap, sc.positions, y * 1e3, U, interp)
    print(f"odd P N_w {interp}: forward {rel_l2(Y, Yo):.2e} adjoint {rel_l2(X, Xo):.2e}")
    assert rel_l2(Y, Yo) <= 1e-5 and rel_l2(X, Xo) <= 1e-5
