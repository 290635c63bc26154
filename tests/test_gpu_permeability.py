"""Batched permeability solves (taxila-lbm_amd/permeability.py) on the GPU.

Samples advanced together in a batch, dropping out as they converge, must
reach the same permeability after the same number of steps as the same
samples solved one at a time (the batched step is the single-domain step and
the FP64 flow sums are deterministic).  Physics check: in the bulk of a
periodic pack the Darcy velocity is positive along the force.
"""
import numpy as np
import pytest

import permeability as pm
import txg

pytestmark = pytest.mark.gpu


def small_specs(n):
    out = []
    for i in range(n):
        rng = np.random.default_rng([7, i])
        edge = int(rng.integers(20, 33))
        walls = pm.workloads.sphere_pack((edge,) * 3, float(rng.uniform(0.3, 0.5)), 4.0, 10.0, seed=100 + i,
                                         minerals=(1.0,))
        out.append((edge, None, walls))
    return out


@pytest.mark.parametrize("prec", [txg.FP64, txg.FP32])
def test_batched_equals_plain(prec):
    specs = small_specs(5)
    kb, _, _ = pm.solve_batched(specs, check=50, precision=prec, tol=1e-5, max_steps=2000)
    kp, _, _ = pm.solve_plain(specs, check=50, precision=prec, tol=1e-5, max_steps=2000)
    assert [s for _, s in kb] == [s for _, s in kp]
    assert [k for k, _ in kb] == [k for k, _ in kp]
    # samples converge at different steps (continuous batching exercised)
    assert len(set(s for _, s in kb)) > 1 or max(s for _, s in kb) == 2000
    assert all(k >= 0 for k, _ in kb)
