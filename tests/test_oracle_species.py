"""The oracle's passive-scalar species distributions (CPU).

No reference output covers the transport, so the restatement of
lbm_transport.F90 in oracle/taxila_oracle.c (collide_species, init_species,
step_species) is pinned by what the reference's routines fix:
  - the equilibrium (DiscretizationEquilf with RelaxationCreate's d_k = 0,
    lbm_relaxation.F90:81) has moments rho, rho j and rho (I/2 + j j) where
    j is the species' own momentum sum g c (TransportUpdateMoments) and not
    a velocity: a tau = 1 collision leaves exactly these moments, and the
    rest population -rho j.j/2;
  - on a uniform periodic state one SRT step turns j into
    j (1 - (1 - rho)/tau), which an equilibrium built on the velocity j/rho
    (the flow's) would not give;
  - TransportCollisionD2 collides only nodes with walls > 0 (as the
    reference's test reads), so on D2Q9 fluid populations stream untouched:
    after k steps on a periodic box g is the initial g shifted by k c;
  - collisions and bounce-back conserve each species' mass in a closed,
    porous box, for SRT and MRT, on D3Q19 and D2Q9;
  - MRT with every rate equal to 1/tau equals SRT.
"""
import numpy as np
import pytest

import cases
import oracle
import txg

from cases import C2, C3


def species_desc(**kw):
    kw.setdefault("isotropy_order", 0)
    return txg.make_desc(transport=1, **kw)


def fi_of(o, d):
    return o.get_fi().reshape(d.nnodes, d.Q, d.ncomponents)


def grid(fi, d):
    """(node, Q, C) -> (Q, C, z, y, x) for rolls along the lattice axes."""
    shape = (d.nz, d.ny, d.nx) if d.ndims == 3 else (d.ny, d.nx)
    return np.moveaxis(fi, 0, -1).reshape(d.Q, d.ncomponents, *shape)


def run(d, walls, rho, u, steps, fi=None):
    o = oracle.Oracle(d)
    o.set_walls(walls)
    o.set_state(rho, u)
    if fi is not None:
        o.set_fi(fi)
    o.init()
    o.step(steps)
    return o


def test_equilibrium_moments_and_rest_population():
    """tau = 1: g after the collision is feq(psi, j); un-stream one step and
    check its moments node by node."""
    d = species_desc(ndims=3, ncomponents=2, shape=(8, 7, 6), tau=[1.0, 1.0])
    rng = np.random.default_rng(1)
    rho = 0.2 + rng.random(d.nnodes * 2)
    u = 0.03 * rng.standard_normal(d.nnodes * 3 * 2)
    o = run(d, np.zeros(d.nnodes), rho, u, 1)
    # the state after init is feq(rho, u); its sums are the moments the
    # collision sees
    o0 = run(d, np.zeros(d.nnodes), rho, u, 0)
    r0, j0, _ = o0.get_moments()
    r0 = r0.reshape(d.nnodes, 2)
    j0 = j0.reshape(d.nnodes, 3, 2)
    g = grid(fi_of(o, d), d)
    back = np.empty_like(g)
    for n, c in enumerate(C3):
        back[n] = np.roll(g[n], shift=(-c[2], -c[1], -c[0]), axis=(-3, -2, -1))
    post = np.moveaxis(back.reshape(19, 2, -1), -1, 0)  # (node, Q, C)
    for m in range(2):
        rm, jm = r0[:, m], j0[:, :, m]
        assert np.allclose(post[:, :, m].sum(1), rm, rtol=0, atol=1e-14)
        assert np.allclose(post[:, :, m] @ C3, rm[:, None] * jm, rtol=0, atol=1e-14)
        second = np.einsum("nq,qa,qb->nab", post[:, :, m], C3, C3)
        want = rm[:, None, None] * (0.5 * np.eye(3)[None] + jm[:, :, None] * jm[:, None, :])
        assert np.allclose(second, want, rtol=0, atol=1e-14)
        assert np.allclose(post[:, 0, m], -0.5 * rm * (jm * jm).sum(1), rtol=0, atol=1e-15)


@pytest.mark.parametrize("rho0,tau", [(0.5, 1.0), (0.25, 0.8), (1.0, 1.3)])
def test_uniform_state_momentum_relaxes_to_rho_j(rho0, tau):
    d = species_desc(ndims=3, ncomponents=1, shape=(5, 5, 5), tau=[tau])
    j = np.array([0.01, -0.004, 0.002])
    o = run(d, np.zeros(d.nnodes), np.full(d.nnodes, rho0), np.tile(j, d.nnodes), 0)
    r_a, j_a, _ = o.get_moments()
    o.step(1)
    r_b, j_b, _ = o.get_moments()
    # init sums feq(rho0, j): its density is rho0 and its momentum rho0 j
    assert np.allclose(r_a, rho0, rtol=0, atol=1e-15)
    ja = rho0 * j
    assert np.allclose(j_a.reshape(-1, 3), ja, rtol=0, atol=1e-15)
    want = ja - (ja - rho0 * ja) / tau
    assert np.allclose(r_b, rho0, rtol=0, atol=1e-15)
    assert np.allclose(j_b.reshape(-1, 3), want, rtol=0, atol=1e-15)


def test_d2_fluid_populations_only_stream():
    d = species_desc(ndims=2, ncomponents=2, shape=(9, 7), tau=[0.7, 1.2])
    rng = np.random.default_rng(4)
    fi = 0.1 + rng.random(d.nnodes * 9 * 2)
    o = run(d, np.zeros(d.nnodes), np.ones(d.nnodes * 2), np.zeros(d.nnodes * 4), 0, fi=fi)
    g0 = grid(fi.reshape(d.nnodes, 9, 2), d)
    k = 5
    o.step(k)
    g = grid(fi_of(o, d), d)
    for n, c in enumerate(C2):
        assert np.array_equal(g[n], np.roll(g0[n], shift=(k * c[1], k * c[0]), axis=(-2, -1)))


@pytest.mark.parametrize("ndims,mode", [(3, txg.SRT), (3, txg.MRT), (2, txg.SRT), (2, txg.MRT)])
def test_mass_conserved_in_closed_porous_box(ndims, mode):
    shape = (14, 12, 10) if ndims == 3 else (30, 24)
    kw = dict(ndims=ndims, ncomponents=2, shape=shape, periodic=(0, 0, 0), relaxation_mode=mode,
              tau=[0.8, 1.4], s_c=[1.0, 1.0], s_e=[1.2, 1.1], s_e2=[1.3, 0.9], s_q=[1.1, 1.5],
              s_nu=[1 / 0.8, 1 / 1.4], s_pi=[1.2, 1.2], s_m=[1.1, 1.3])
    d = species_desc(**kw)
    walls = cases.random_walls(shape, 0.15, seed=2)
    rng = np.random.default_rng(6)
    rho = (0.05 + rng.random(d.nnodes * 2)) * np.repeat(walls == 0, 2)
    u = 0.01 * rng.standard_normal(d.nnodes * ndims * 2)
    o = run(d, walls, rho, u, 0)
    m0 = o.get_moments()[0].reshape(-1, 2).sum(0)
    o.step(40)
    r, j, _ = o.get_moments()
    assert np.isfinite(r).all()
    m1 = r.reshape(-1, 2).sum(0)
    assert np.allclose(m1, m0, rtol=1e-13, atol=0)


def test_mrt_with_equal_rates_is_srt():
    shape = (10, 9, 8)
    walls = cases.random_walls(shape, 0.1, seed=8)
    rng = np.random.default_rng(9)
    rho = 0.3 + rng.random(np.prod(shape))
    u = 0.02 * rng.standard_normal(np.prod(shape) * 3)
    tau = 0.9
    a = run(species_desc(ndims=3, shape=shape, tau=[tau]), walls, rho, u, 12)
    r = [1 / tau]
    b = run(species_desc(ndims=3, shape=shape, relaxation_mode=txg.MRT, s_c=r, s_e=r, s_e2=r, s_q=r, s_nu=r,
                         s_pi=r, s_m=r), walls, rho, u, 12)
    fl = np.repeat(walls == 0, 19)
    assert np.allclose(a.get_fi()[fl], b.get_fi()[fl], rtol=0, atol=1e-13)


def test_species_refuse_forces_bcs_and_flow_calls():
    with pytest.raises(ValueError, match="forces"):
        oracle.Oracle(species_desc(ndims=3, ncomponents=2, shape=(6, 6, 6), g=[[0, 0.1], [0.1, 0]]))
    with pytest.raises(ValueError, match="boundary"):
        oracle.Oracle(species_desc(ndims=3, shape=(6, 6, 6), periodic=(0, 1, 1), bc={"xm": [1.0]}))
    d = species_desc(ndims=3, shape=(6, 6, 6))
    o = run(d, np.zeros(d.nnodes), np.ones(d.nnodes), np.zeros(d.nnodes * 3), 1)
    with pytest.raises(RuntimeError):
        o.diagnostics()
