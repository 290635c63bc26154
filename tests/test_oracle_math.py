"""The oracle against what the mathematics fixes (CPU).

These pin the parts no reference output covers (D3Q19, MRT, the isotropy-8
and -10 stencils, walls, wettability, body force, Dirichlet faces), each
chosen so that a plausible mistake (a dropped term, a wrong sign, index or
weight, a transposed matrix) fails:
  - equilibrium moments: sum feq = rho, sum feq c = rho u, and for mm = 1
    sum feq c c = rho/3 I + rho u u (brute force over random u);
  - MRT collision equals the textbook moment-space relaxation
    f - M^-1 S (M f - M feq) computed with numpy, and equals SRT when every
    rate is 1/tau;
  - per-component mass conservation with walls (bounce-back) and total
    momentum conservation in a periodic two-component run;
  - the psi-gradient of a linear density field is its exact slope, so the
    fluid-fluid force is -c0 rho_m sum_n g_mn slope_n;
  - the fluid-solid force of one mineral voxel next to one fluid node equals
    the hand-written k_n rho gw c_n (with the default-real weights);
  - a 90-degree rotation of the domain rotates the solution;
  - Dirichlet faces hold each component's density after the node correction.
"""
import numpy as np
import pytest

import cases
import oracle
import txg

from cases import C2, C3, W2, W3  # noqa: E402


def m_d3q19():
    c = C3
    M = np.zeros((19, 19))
    e2 = (c ** 2).sum(1)
    M[0] = 1
    M[1] = 19 * e2 - 30
    M[2] = (21 * e2 ** 2 - 53 * e2 + 24) / 2
    M[3], M[5], M[7] = c[:, 0], c[:, 1], c[:, 2]
    M[4] = (5 * e2 - 9) * c[:, 0]
    M[6] = (5 * e2 - 9) * c[:, 1]
    M[8] = (5 * e2 - 9) * c[:, 2]
    M[9] = 3 * c[:, 0] ** 2 - e2
    M[10] = (3 * e2 - 5) * (3 * c[:, 0] ** 2 - e2)
    M[11] = c[:, 1] ** 2 - c[:, 2] ** 2
    M[12] = (3 * e2 - 5) * (c[:, 1] ** 2 - c[:, 2] ** 2)
    M[13] = c[:, 0] * c[:, 1]
    M[14] = c[:, 1] * c[:, 2]
    M[15] = c[:, 0] * c[:, 2]
    M[16] = (c[:, 1] ** 2 - c[:, 2] ** 2) * c[:, 0]
    M[17] = (c[:, 2] ** 2 - c[:, 0] ** 2) * c[:, 1]
    M[18] = (c[:, 0] ** 2 - c[:, 1] ** 2) * c[:, 2]
    return M


def m_d2q9():
    c = C2
    e2 = (c ** 2).sum(1)
    M = np.zeros((9, 9))
    M[0] = 1
    M[1] = 3 * e2 - 4
    M[2] = (9 * e2 ** 2 - 21 * e2 + 8) / 2
    M[3], M[5] = c[:, 0], c[:, 1]
    M[4] = (3 * e2 - 5) * c[:, 0]
    M[6] = (3 * e2 - 5) * c[:, 1]
    M[7] = c[:, 0] ** 2 - c[:, 1] ** 2
    M[8] = c[:, 0] * c[:, 1]
    return M


def test_mrt_matrices_are_the_textbook_ones():
    """The oracle's M (lbm_discretization_*: mt_mrt) is d'Humieres' D3Q19 /
    Lallemand-Luo D2Q9 moment basis, and M M^T = diag(mmt)."""
    import ctypes
    for dims, M, mmt in ((3, m_d3q19(), [19, 2394, 252, 10, 40, 10, 40, 10, 40, 36, 72, 12, 24, 4, 4, 4, 8, 8, 8]),
                         (2, m_d2q9(), [9, 36, 36, 6, 12, 6, 12, 4, 4])):
        assert np.array_equal(M @ M.T, np.diag(mmt))


def single_node_desc(dims, **kw):
    shape = (3, 3, 3) if dims == 3 else (3, 3)
    return txg.make_desc(ndims=dims, ncomponents=kw.pop("ncomponents", 1), shape=shape, isotropy_order=0, **kw)


@pytest.mark.parametrize("dims", [2, 3])
@pytest.mark.parametrize("mm", [1.0, 2.5])
def test_equilibrium_moments(dims, mm):
    rng = np.random.default_rng(3)
    d = single_node_desc(dims, mm=[mm])
    n = d.nnodes
    c, w = (C3, W3) if dims == 3 else (C2, W2)
    for _ in range(4):
        rho = rng.uniform(0.2, 2.0, n)
        u = rng.uniform(-0.1, 0.1, (n, dims))
        o = oracle.Oracle(d)
        o.set_walls(np.zeros(n))
        o.set_state(rho, u.reshape(-1))
        o.phase(12)  # fi = feq(rho, u)
        f = o.get_fi().reshape(n, d.Q)
        assert np.allclose(f.sum(1), rho, rtol=1e-14, atol=0)
        assert np.allclose(f @ c, rho[:, None] * u, rtol=1e-13, atol=1e-16)
        if mm == 1.0:
            pi = np.einsum("kn,na,nb->kab", f, c, c)
            exp = rho[:, None, None] * (np.eye(dims) / 3 + u[:, :, None] * u[:, None, :])
            assert np.allclose(pi, exp, rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("dims", [2, 3])
def test_mrt_collision_in_moment_space(dims):
    """No forces: f' = f - M^-1 S (M f - M feq(rho, j/rho))."""
    rng = np.random.default_rng(11)
    rates = dict(s_c=[1.0], s_e=[1.19], s_e2=[1.4], s_q=[1.2], s_nu=[1.6], s_pi=[0.9], s_m=[1.1])
    d = single_node_desc(dims, relaxation_mode=txg.MRT, **rates)
    n, Q = d.nnodes, d.Q
    c, w = (C3, W3) if dims == 3 else (C2, W2)
    M = m_d3q19() if dims == 3 else m_d2q9()
    rmap = [0, 1, 2, 0, 3, 0, 3, 0, 3, 4, 5, 4, 5, 4, 4, 4, 6, 6, 6] if dims == 3 else [0, 1, 2, 0, 3, 0, 3, 4, 4]
    s = np.array([[1.0, 1.19, 1.4, 1.2, 1.6, 0.9, 1.1][k] for k in rmap])
    f0 = (w[None, :] * rng.uniform(0.8, 1.2, n)[:, None]) * (1 + 0.05 * rng.standard_normal((n, Q)))
    o = oracle.Oracle(d)
    o.set_walls(np.zeros(n))
    o.set_fi(f0.reshape(-1))
    o.init()  # rho, j, u from f0 (no forces: g = 0, no walls)
    o.phase(0)  # feq + collide
    got = o.get_fi().reshape(n, Q)
    rho = f0.sum(1)
    u = (f0 @ c) / rho[:, None]
    cu = u @ c.T
    feq = w[None, :] * rho[:, None] * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * (u ** 2).sum(1)[:, None])
    mom = (f0 - feq) @ M.T
    exp = f0 - (mom * s[None, :]) @ np.linalg.inv(M).T
    assert np.allclose(got, exp, rtol=1e-13, atol=1e-16)


@pytest.mark.parametrize("dims", [2, 3])
def test_mrt_equal_rates_is_srt(dims):
    tau = 0.8
    shape = (14, 12, 10) if dims == 3 else (20, 18)
    common = dict(ndims=dims, ncomponents=2, shape=shape, isotropy_order=4, g=[[0, 0.1], [0.1, 0]])
    srt = txg.make_desc(tau=[tau, tau], **common)
    mrt = txg.make_desc(relaxation_mode=txg.MRT, **{k: [1 / tau] * 2 for k in
                                                      ("s_c", "s_e", "s_e2", "s_q", "s_nu", "s_pi", "s_m")},
                        **common)
    rho = cases.ball_ic(shape, 4, (0.03, 0.97), (0.97, 0.03))
    walls = np.zeros(srt.nnodes)
    a = cases.run(oracle.Oracle, srt, walls, rho, None, 20).get_fi()
    b = cases.run(oracle.Oracle, mrt, walls, rho, None, 20).get_fi()
    assert np.allclose(a, b, rtol=1e-11, atol=1e-14)
    assert not np.array_equal(a, b) or dims == 2  # different arithmetic, same physics


def porous_case(dims, periodic, bc=None, gw=None, gvt=None, iso=8, mode=txg.MRT):
    shape = (18, 14, 12) if dims == 3 else (30, 24)
    kw = dict(ndims=dims, ncomponents=2, shape=shape, periodic=periodic, isotropy_order=iso,
              relaxation_mode=mode, g=[[0, 0.1], [0.1, 0]], tau=[1.0, 0.7])
    if mode == txg.MRT:
        kw.update(s_c=[1.0, 1 / 0.7], s_e=[1.19] * 2, s_e2=[1.4] * 2, s_q=[1.2] * 2, s_nu=[1.0, 1 / 0.7])
    if gw is not None:
        kw["gw"] = gw
    if gvt is not None:
        kw["gvt"] = gvt
    if bc is not None:
        kw["bc"] = bc
    d = txg.make_desc(**kw)
    walls = cases.random_walls(shape, 0.2, seed=5, minerals=(1.0, 2.0) if gw and len(gw) > 1 else (1.0,))
    rho = cases.ball_ic(shape, min(shape) // 4, (0.03, 0.97), (0.97, 0.03))
    return d, walls, rho


@pytest.mark.parametrize("dims", [2, 3])
def test_mass_conserved_with_walls(dims):
    d, walls, rho = porous_case(dims, periodic=(0, 0, 0), gw=[[0.05, -0.05]], gvt=[1e-5, 0, 0][:dims])
    o = cases.run(oracle.Oracle, d, walls, rho, None, 0)
    m0 = o.get_moments()[0].reshape(-1, 2).sum(0)
    o.step(60)
    m1 = o.get_moments()[0].reshape(-1, 2).sum(0)
    assert np.allclose(m0, m1, rtol=1e-12, atol=0)


@pytest.mark.parametrize("dims", [2, 3])
def test_momentum_conserved_periodic(dims):
    """Symmetric g, no walls, periodic: sum over nodes and components of
    (j + F/2) stays zero (the interaction forces cancel pairwise)."""
    shape = (16, 14, 12) if dims == 3 else (32, 28)
    d = txg.make_desc(ndims=dims, ncomponents=2, shape=shape, isotropy_order=8, g=[[0, 0.1], [0.1, 0]],
                      tau=[1.0, 0.8])
    rng = np.random.default_rng(2)
    rho = np.clip(cases.ball_ic(shape, 4, (0.03, 0.97), (0.97, 0.03)) + 0.01 * rng.standard_normal(
        d.nnodes * 2), 0.01, None)
    o = cases.run(oracle.Oracle, d, np.zeros(d.nnodes), rho, None, 40)
    f = o.get_fi().reshape(d.nnodes, d.Q, 2)
    _, _, F = o.get_moments()
    c = C3 if dims == 3 else C2
    p = np.einsum("nqm,qd->d", f, c) + 0.5 * F.reshape(d.nnodes, dims, 2).sum((0, 2))
    assert np.abs(p).max() < 1e-12


@pytest.mark.parametrize("dims,iso", [(3, 4), (3, 8), (2, 4), (2, 8), (2, 10)])
def test_gradient_of_linear_field_is_exact(dims, iso):
    shape = (14, 14, 14) if dims == 3 else (20, 20)
    g = [[0.0, 0.1], [0.07, 0.0]]
    d = txg.make_desc(ndims=dims, ncomponents=2, shape=shape, periodic=(0, 0, 0), isotropy_order=iso, g=g)
    slope = np.array([[0.003, -0.002, 0.001][:dims], [-0.001, 0.0025, 0.0015][:dims]])
    grids = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    pos = np.stack([np.transpose(gr, tuple(range(dims))[::-1]).reshape(-1) for gr in grids], 1)
    rho = np.stack([0.5 + pos @ slope[0], 0.4 + pos @ slope[1]], 1)
    o = oracle.Oracle(d)
    o.set_walls(np.zeros(d.nnodes))
    o.set_state(rho.reshape(-1), None)
    o.init()
    _, _, F = o.get_moments()
    F = F.reshape(d.nnodes, dims, 2)
    r = {4: 1, 8: 2, 10: 3}[iso] + 1
    interior = np.all((pos >= r) & (pos < np.array(shape) - r), axis=1)
    for m in range(2):
        exp = -6.0 * rho[:, m:m + 1] * (np.array(g)[m] @ slope)[None, :]
        assert np.allclose(F[interior, :, m], exp[interior], rtol=1e-10, atol=1e-17)


@pytest.mark.parametrize("dims", [2, 3])
def test_fluid_solid_force_by_hand(dims):
    """One mineral voxel at offset c_n of a fluid node: F_m = -k_n rho_m gw_m c_n
    with k = 1./6. (1./3. in 2-D) on axes and 1./12. on diagonals, as
    default-real (single precision) literals (lbm_forcing.F90:1355-1414)."""
    shape = (7, 7, 7) if dims == 3 else (7, 7)
    gw = [[0.2, -0.3]]
    d = txg.make_desc(ndims=dims, ncomponents=2, shape=shape, isotropy_order=0, gw=gw)
    c = C3 if dims == 3 else C2
    kax = float(np.float32(1.0) / np.float32(6.0 if dims == 3 else 3.0))
    kdg = float(np.float32(1.0) / np.float32(12.0))
    for n in (1, 2, 3 if dims == 2 else 5, 5 if dims == 2 else 7, 8 if dims == 2 else 16):
        walls = np.zeros(d.nnodes)
        centre = np.array([3] * dims)
        wpos = centre + c[n]
        stride = [1, shape[0], shape[0] * shape[1]][:dims]
        walls[int(wpos @ np.array(stride))] = 1.0
        rho = np.tile([0.6, 0.4], d.nnodes)
        o = oracle.Oracle(d)
        o.set_walls(walls)
        o.set_state(rho, None)
        o.init()
        _, _, F = o.get_moments()
        F = F.reshape(d.nnodes, dims, 2)[int(centre @ np.array(stride))]
        k = kax if n <= 2 * dims else kdg
        r_after = o.get_moments()[0].reshape(-1, 2)[int(centre @ np.array(stride))]
        for m in range(2):
            exp = -(k * r_after[m] * gw[0][m]) * c[n]
            assert np.allclose(F[:, m], exp, rtol=1e-15, atol=1e-18), (n, F[:, m], exp)
        # the 1./6. literal matters: a double 1/6 would differ in the 8th digit
        assert k != (1 / 6 if dims == 3 else 1 / 3) or n > 2 * dims


def test_rotation_equivariance_d3():
    """Rotating the domain (x,y,z) -> (y,z,x) rotates the fields."""
    d, walls, rho = porous_case(3, periodic=(1, 1, 1), gw=[[0.04, -0.04]], gvt=None, iso=8)
    shape = (d.nx, d.ny, d.nz)
    o = cases.run(oracle.Oracle, d, walls, rho, None, 15)
    r1 = o.get_moments()[0].reshape(shape[2], shape[1], shape[0], 2)
    # rotated problem: new axes (X,Y,Z) = old (y,z,x)
    d2 = txg.make_desc(ndims=3, ncomponents=2, shape=(shape[1], shape[2], shape[0]), isotropy_order=8,
                       relaxation_mode=txg.MRT, g=[[0, 0.1], [0.1, 0]], tau=[1.0, 0.7], s_c=[1.0, 1 / 0.7],
                       s_e=[1.19] * 2, s_e2=[1.4] * 2, s_q=[1.2] * 2, s_nu=[1.0, 1 / 0.7], gw=[[0.04, -0.04]])
    w3 = walls.reshape(shape[2], shape[1], shape[0])       # [z,y,x]
    rr = rho.reshape(shape[2], shape[1], shape[0], 2)
    # new array index [Z,Y,X] = [x, z, y]
    w_rot = np.transpose(w3, (2, 0, 1)).copy()
    r_rot = np.transpose(rr, (2, 0, 1, 3)).copy()
    o2 = cases.run(oracle.Oracle, d2, w_rot.reshape(-1), r_rot.reshape(-1), None, 15)
    r2 = o2.get_moments()[0].reshape(shape[0], shape[2], shape[1], 2)
    assert np.allclose(np.transpose(r1, (2, 0, 1, 3)), r2, rtol=1e-11, atol=1e-14)


@pytest.mark.parametrize("dims", [2, 3])
def test_dirichlet_faces_hold_density(dims):
    bcv = {"xm": [0.985, 0.02], "xp": [0.975, 0.02]}
    d, walls, rho = porous_case(dims, periodic=(0, 1, 1), bc=bcv, iso=4, mode=txg.SRT)
    o = cases.run(oracle.Oracle, d, walls, rho, None, 30)
    r = o.get_moments()[0].reshape(-1, 2)
    nx = d.nx
    x = np.arange(d.nnodes) % nx
    fl = walls == 0
    for face, xv in (("xm", 0), ("xp", nx - 1)):
        sel = fl & (x == xv)
        assert np.allclose(r[sel], np.array(bcv[face])[None, :], rtol=1e-13, atol=0)


def test_permeability_samples_are_seeded():
    """The porous samples of the batched permeability workload are a pure
    function of their index: edge in 32..96, porosity drawn in [0.3, 0.5]
    and reached to within one sphere's volume."""
    import permeability as pm
    for i in (0, 5):
        e1, p1, w1 = pm.sample(i)
        e2, p2, w2 = pm.sample(i)
        assert e1 == e2 and p1 == p2 and np.array_equal(w1, w2)
        assert 32 <= e1 <= 96 and 0.3 <= p1 <= 0.5
        por = float((w1 == 0).mean())
        assert p1 - 4.0 / 3.0 * np.pi * 10.0 ** 3 / e1 ** 3 <= por <= p1 + 1e-12


def test_request_stream_is_seeded_and_mixed():
    """The serving workload's request stream: seeded, Poisson arrivals
    (mean gap 1/rate), the 70/20/10 kind mix and the per-kind edge ranges."""
    import serve
    a = serve.make_stream(2000, rate=50.0, seed=3)
    b = serve.make_stream(2000, rate=50.0, seed=3)
    assert [(r.kind, r.seed, r.edge, r.budget, r.arrival) for r in a] == \
           [(r.kind, r.seed, r.edge, r.budget, r.arrival) for r in b]
    gaps = np.diff([0.0] + [r.arrival for r in a])
    assert abs(gaps.mean() - 1 / 50.0) < 0.1 / 50.0
    frac = {k: sum(r.kind == k for r in a) / len(a) for k in serve.KINDS}
    assert abs(frac["perm"] - 0.7) < 0.04 and abs(frac["drain"] - 0.2) < 0.03 and abs(frac["bubble"] - 0.1) < 0.03
    for r in a:
        lo, hi = {"perm": (32, 128), "drain": (64, 160), "bubble": (256, 1024)}[r.kind]
        assert lo <= r.edge <= hi
