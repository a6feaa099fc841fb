"""Plans are independent of each other: two plans driven from two host
threads, each on its own HIP stream, give the same results as the same calls
made one after the other.  (One plan holds mutable scratch -- profile rows,
window partials, flags -- so calls on one plan must be ordered on one stream;
that is the documented rule and is not exercised here.)"""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda", 0)


def _problem(dev, k):
    from sarsolver_amd import synthetic
    w = synthetic.WORKLOADS["c1_small"]
    ap = synthetic.circular_aperture(w.pulses, w.samples, w.aperture_deg, start_deg=37.0 * k)
    sc = synthetic.grid_scene(w.n, w.spacing)
    pl = synthetic.workload_plan(w, ap, sc.positions, dev, interp="linear", precision="series")
    g = np.random.default_rng(50 + k)
    x = torch.from_numpy(synthetic.complex_normal(g, sc.num_scatterers).astype(np.complex64))
    y = torch.from_numpy(synthetic.complex_normal(g, (w.pulses, w.samples)).astype(np.complex64))
    return pl, x.to(dev), y.to(dev)


def _run(pl, x, y, reps):
    outs = []
    for _ in range(reps):
        outs.append((pl.forward(x), pl.adjoint(y)))
    return outs


def test_two_plans_two_threads_two_streams(dev):
    print("case: two plans on two threads and streams", flush=True)
    probs = [_problem(dev, k) for k in range(2)]
    reps = 20
    seq = [[(a.cpu().numpy(), b.cpu().numpy()) for a, b in _run(*p, reps)] for p in probs]
    torch.cuda.synchronize(dev)
    streams = [torch.cuda.Stream(device=dev) for _ in probs]
    results = [None, None]
    errors = []

    def worker(i):
        try:
            with torch.cuda.stream(streams[i]):
                outs = _run(*probs[i], reps)
                streams[i].synchronize()
                results[i] = [(a.cpu().numpy(), b.cpu().numpy()) for a, b in outs]
        except Exception as e:  # reported below
            errors.append(repr(e))

    th = [threading.Thread(target=worker, args=(i,)) for i in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=120)
    assert not errors, errors
    assert all(not t.is_alive() for t in th)
    for i in range(2):
        for (a, b), (sa, sb) in zip(results[i], seq[i]):
            assert np.array_equal(a, sa) and np.array_equal(b, sb)
