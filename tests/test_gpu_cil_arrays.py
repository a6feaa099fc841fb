"""The CIL operator's array logic on the GPU (CIL itself is not installed, so
the container classes are not exercised).  The reference's CIL adjoint test
(tests/test_cil.py:12-15: adjoint of little_noisy_meas equals
little_backproj, elementwise rel 1e-3) is run through `SarArrayOperator`,
which is what CpuSarOperator.adjoint calls: the [P, N] input and the grid-
shaped (1, 244, 241) output.  The plan-bound variant (GpuSarOperator's) must
equal it, and direct must give the [P, N] phase history of
multi_forward_evaluate for an image of grid shape."""
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


@pytest.mark.parametrize("bind", [False, True])
def test_cil_array_operator_little_problem(dev, little, little_arrays, bind):
    from sarsolver_amd.cil import SarArrayOperator
    from sarsolver_amd.cxx_binding import multi_forward_evaluate
    ap, sc = little
    print(f"case: CIL array operator, plans bound {bind}", flush=True)
    op = SarArrayOperator(sc, ap, num_calculation_blocks=2 if bind else 1, bind_plans=bind)
    assert op.domain_shape == (1, 244, 241) and op.range_shape == (249, 251)
    h = op.adjoint_array(little_arrays["noisy_meas"])
    ref = little_arrays["backproj"]
    assert h.shape == ref.shape and h.dtype == np.complex128
    assert rel_l2(h, ref) <= 1e-5
    assert h.ravel().tolist() == pytest.approx(ref.ravel().tolist(), rel=1e-3)
    gt = little_arrays["gt"].reshape(op.domain_shape)
    m = op.direct_array(gt)
    assert m.shape == (249, 251) and m.dtype == np.complex128
    m0 = multi_forward_evaluate(sc, ap, gt.reshape(-1))
    assert rel_l2(m, m0) <= 1e-6
