"""The serving loop (taxila-lbm_amd/serve.py) on the GPU.

Each request's result (steps taken and the convergence quantity it stopped
on: permeability, invading mass or steady-state norm) must be the same when
served in the mix, with continuous batching and HBM-budget admission, as
when served alone; the tolerance is exact equality, because a request's
steps are the single-domain steps in both and its checks fall on the same
step counts.
"""
import pytest

import serve

pytestmark = pytest.mark.gpu


def small_stream(n, seed):
    return serve.make_stream(n, rate=500.0, seed=seed,
                             edges={"perm": (16, 28), "drain": (20, 28), "bubble": (48, 80)},
                             budgets={"perm": (100, 300), "drain": (50, 150), "bubble": (50, 150)})


@pytest.mark.parametrize("budget", [64 << 30, 3 << 20])
def test_served_results_equal_alone(budget):
    """A large budget admits everything at once; a 3 MiB budget forces the
    queue to wait for memory (requests are admitted as others finish)."""
    stream = small_stream(10, seed=5)
    kinds = {r.kind for r in stream}
    assert {"perm", "drain"} <= kinds
    inputs = serve.prepare(stream)
    served = serve.Server(hbm_budget=budget, quantum=50).run(serve.clone(stream), inputs)
    plain = serve.run_plain(serve.clone(stream), inputs, quantum=50)
    assert len(served) == len(plain) == len(stream)
    a = {r.rid: r for r in served}
    for r in plain:
        s = a[r.rid]
        assert s.steps == r.steps and s.steps <= r.budget + 49
        assert repr(s.result) == repr(r.result), (r.kind, s.result, r.result)  # NaN-aware
        assert all(v == v for v in s.result.values() if isinstance(v, float)), (r.kind, s.result)
        assert s.done >= s.start >= s.arrival - 1e-9
