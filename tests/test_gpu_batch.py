"""Batched steps of many uneven domains (txg_batch_*), on the GPU.

A member of a batch takes exactly the step it takes alone, so a batched run
must equal the members' individual runs bit for bit in every precision, and
in FP64 the oracle as well; members may join (fresh from txg_init) and leave
between batch steps.
"""
import numpy as np
import pytest

import cases
import oracle
import txg

pytestmark = pytest.mark.gpu

SHAPES = [(20, 16, 18), (37, 23, 19), (16, 16, 40), (9, 30, 11), (64, 8, 6)]


def member(shape, prec, seed, walls_frac=0.12):
    d = txg.make_desc(ndims=3, ncomponents=2, shape=shape, periodic=(1, 1, 1), isotropy_order=4,
                      g=[[0, 0.1], [0.1, 0]], tau=[1.0, 0.8], gw=[[0.05, -0.05]], gvt=[1e-5, 0, 0],
                      precision=prec, rho0=[0.5, 0.5])
    walls = cases.random_walls(shape, walls_frac, seed=seed, minerals=(1.0,))
    rho = cases.ball_ic(shape, min(shape) // 4, (0.03, 0.97), (0.97, 0.03))
    u = np.zeros(d.nnodes * 2 * 3)
    return d, walls, rho, u


def start(d, walls, rho, u):
    dom = txg.Domain(d)
    dom.set_walls(walls)
    dom.set_state(rho, u)
    dom.init()
    return dom


@pytest.mark.parametrize("prec", [txg.FP64, txg.FP32, txg.FP32_FP16])
def test_batch_equals_individual_runs(prec):
    steps = 9
    specs = [member(s, prec, seed=i + 1) for i, s in enumerate(SHAPES)]
    b = txg.Batch()
    doms = [start(*sp) for sp in specs]
    for dom in doms:
        b.add(dom)
    assert len(b) == len(SHAPES)
    b.step(4)
    b.step(steps - 4)
    tx, ty, zs, ms = b.tune_info()
    assert (tx, ty) in ((32, 8), (64, 4)) and zs > 0 and ms > 0
    for sp, dom in zip(specs, doms):
        alone = start(*sp)
        alone.tune("onepass", (32, 8), 5)
        alone.step(steps)
        assert np.array_equal(dom.get_fi(), alone.get_fi()), "shape %s" % (sp[0].nx,)
    if prec == txg.FP64:
        d, walls, rho, u = specs[1]
        o = cases.run(oracle.Oracle, d, walls, rho, u, steps).get_fi()
        m = np.repeat(walls == 0, d.Q * d.ncomponents)
        assert np.array_equal(doms[1].get_fi()[m], o[m])


def test_batch_members_join_and_leave():
    """Continuous batching: a member joins after 3 batch steps (its first
    step is its own), another leaves after 5 and is stepped alone; every
    member ends equal to an individual run of the same number of steps."""
    prec = txg.FP32
    specs = [member(s, prec, seed=10 + i) for i, s in enumerate(SHAPES[:4])]
    doms = [start(*sp) for sp in specs]
    b = txg.Batch()
    b.tune((64, 4), 8)
    for dom in doms[:3]:
        b.add(dom)
    b.step(3)
    b.add(doms[3])            # joins fresh from txg_init
    b.step(2)
    b.remove(doms[0])         # leaves after 5 steps
    doms[0].step(2)           # continues alone
    b.step(4)
    counts = [7, 9, 9, 6]
    for sp, dom, n in zip(specs, doms, counts):
        alone = start(*sp)
        alone.step(n)
        assert np.array_equal(dom.get_fi(), alone.get_fi())


def test_batch_rejects_mixed_configurations():
    a = start(*member((16, 16, 16), txg.FP32, 1))
    c = start(*member((16, 16, 16), txg.FP64, 2))
    b = txg.Batch()
    b.add(a)
    with pytest.raises(txg.TxgError, match="share"):
        b.add(c)
    b2 = txg.Batch()
    with pytest.raises(txg.TxgError, match="already"):
        b2.add(a)
