"""Synthetic inputs of the benchmark configurations.

Initial conditions and wall fields are generated on the host with numpy's
seeded PCG64 generator, so every consumer (the HIP path, the oracle, the
batch and serving loops) sees bit-identical inputs for a given seed.

Arrays use the C ABI layouts of include/taxila_lbm.h (natural order, x
fastest, DOF innermost).
"""
from __future__ import annotations

import numpy as np

import txg

# G = 3 c_0 g = 18 g in the reference's convention, so G = 1.8 is g = 0.1
# (tests/bubble_2D/input_data: -g_12 0.1 -g_21 0.1)
G_BUBBLE = 0.1
RHO_MAJOR, RHO_MINOR = 0.97, 0.03


def ball_ic(shape, radius, inner, outer, center=None):
    """Disc / sphere of `inner` densities in `outer` densities, centred at
    (N+1)/2 in the reference's 1-based coordinates (integer division, the
    initialize_state_bubble.F90:187-190 convention)."""
    dims = len(shape)
    c = center or tuple((n + 1) // 2 for n in shape)
    grids = np.meshgrid(*[np.arange(1, n + 1) for n in shape], indexing="ij")
    r2 = sum((g - cc) ** 2 for g, cc in zip(grids, c))
    ins = r2 <= radius * radius
    ins = np.transpose(ins, tuple(range(dims))[::-1])  # (x,y,z) -> natural (z,y,x)
    rho = np.where(ins[..., None], np.array(inner, dtype=np.float64), np.array(outer, dtype=np.float64))
    return rho.reshape(-1)


def sphere_pack(shape, porosity, rmin, rmax, seed, minerals=(1.0, 2.0), periodic=True):
    """Overlapping-sphere pack.  Spheres with radii U[rmin, rmax] at uniform
    random centres are added until the solid fraction reaches 1 - porosity;
    each sphere's mineral tag is drawn uniformly from `minerals` and the first
    sphere covering a voxel keeps it.  With periodic=True spheres wrap."""
    rng = np.random.default_rng(seed)
    if len(shape) == 2:
        nx, ny, nz = shape[0], shape[1], 1
    else:
        nx, ny, nz = shape
    tags = np.zeros((nz, ny, nx))
    target = 1.0 - porosity
    solid = 0
    total = nx * ny * nz
    while solid < target * total:
        r = rng.uniform(rmin, rmax)
        c = (rng.uniform(0, nx), rng.uniform(0, ny), rng.uniform(0, nz) if nz > 1 else 0.0)
        mineral = float(rng.choice(np.asarray(minerals, dtype=np.float64)))
        ri = int(np.ceil(r))
        xs = np.arange(int(np.floor(c[0])) - ri, int(np.floor(c[0])) + ri + 2)
        ys = np.arange(int(np.floor(c[1])) - ri, int(np.floor(c[1])) + ri + 2)
        zs = np.arange(int(np.floor(c[2])) - ri, int(np.floor(c[2])) + ri + 2) if nz > 1 else np.array([0])
        Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
        inside = (X - c[0]) ** 2 + (Y - c[1]) ** 2 + ((Z - c[2]) ** 2 if nz > 1 else 0) <= r * r
        if periodic:
            X, Y, Z = X % nx, Y % ny, Z % nz
        else:
            keep = (X >= 0) & (X < nx) & (Y >= 0) & (Y < ny) & (Z >= 0) & (Z < nz)
            inside &= keep
            X, Y, Z = np.clip(X, 0, nx - 1), np.clip(Y, 0, ny - 1), np.clip(Z, 0, nz - 1)
        zz, yy, xx = Z[inside], Y[inside], X[inside]
        free = tags[zz, yy, xx] == 0
        tags[zz[free], yy[free], xx[free]] = mineral
        solid += int(free.sum())
    return tags.reshape(-1)


# --------------------------------------------------------------- configs
def bubble_desc(ndims, n, precision=txg.FP32, relaxation_mode=txg.SRT, tau=1.0, rho0=RHO_MINOR):
    """Two-component Shan-Chen bubble: G = 1.8, tau_A = tau_B, periodic.
    MRT uses the standard D3Q19 rates s_e = 1.19, s_eps = 1.4, s_q = 1.2,
    s_nu = s_c = 1/tau, s_pi = s_m = 1 (reference defaults)."""
    shape = (n, n, n) if ndims == 3 else (n, n)
    kw = dict(ndims=ndims, ncomponents=2, shape=shape, periodic=(1, 1, 1), isotropy_order=4,
              relaxation_mode=relaxation_mode, precision=precision, tau=[tau, tau], mm=[1.0, 1.0],
              g=[[0.0, G_BUBBLE], [G_BUBBLE, 0.0]], rho0=[rho0, rho0])
    if relaxation_mode == txg.MRT:
        kw.update(s_c=[1 / tau] * 2, s_e=[1.19] * 2, s_e2=[1.4] * 2, s_q=[1.2] * 2, s_nu=[1 / tau] * 2)
    return txg.make_desc(**kw)


def bubble_ic(ndims, n, radius):
    """Component A (rho = 0.97, 0.03 of B dissolved) in a disc/sphere of the
    given radius centred in the domain; B fills the rest."""
    shape = (n, n, n) if ndims == 3 else (n, n)
    rho = ball_ic(shape, radius, (RHO_MAJOR, RHO_MINOR), (RHO_MINOR, RHO_MAJOR))
    nn = int(np.prod(shape))
    return np.zeros(nn), rho, np.zeros(nn * 2 * ndims)


# ---------------------------------------------------------------- drainage
def wall_adhesion(theta_deg, g=G_BUBBLE, rho_major=RHO_MAJOR, rho_minor=RHO_MINOR):
    """(gw_A, gw_B) giving contact angle theta (measured through component
    A) by the relation of Huang, Thorne, Schaap & Sukop, Phys. Rev. E 76,
    066701 (2007): cos(theta) = (G_ads,B - G_ads,A) / (G_c (rho_1 - rho_2)/2).
    The reference's fluid-fluid force is -c_0 g rho_m sum ffw rho' e with
    c_0 = 6 (lbm_forcing.F90:51-958) and its fluid-solid force
    -gw rho_m sum k_n e with k_n = 1/6, 1/12 (lbm_forcing.F90:1326-1373), the
    same weights as the isotropy-4 stencil, so G_c = 6 g and G_ads = gw.
    With the isotropy-8 fluid-fluid stencil the relation is approximate."""
    d = np.cos(np.radians(theta_deg)) * 6.0 * g * (rho_major - rho_minor) / 2.0
    return -d / 2.0, d / 2.0


DRAIN_TAU = (1.0, 0.6)
DRAIN_DRHO = 0.01
DRAIN_ANGLES = (45.0, 120.0)


def drainage_desc(n, precision=txg.FP32, isotropy_order=8, shape=None):
    """Two-phase drainage through an n^3 sphere pack: MRT (s_e = 1.19,
    s_eps = 1.4, s_q = 1.2, s_nu = s_c = 1/tau), tau_A = 1.0, tau_B = 0.6,
    g = 0.1, isotropy 8; component A (invading) enters at the x- face with
    rho_A = 0.97 + 0.01 over the outlet (Dirichlet densities, BCApplyDirichlet,
    lbm_bc.F90:1273-1333); the other faces are walls (ghost 999 bounce-back).
    Mineral 1 wets A at 45 degrees, mineral 2 at 120 degrees."""
    shape = shape or (n, n, n)
    ta, tb = DRAIN_TAU
    gw1 = wall_adhesion(DRAIN_ANGLES[0])
    gw2 = wall_adhesion(DRAIN_ANGLES[1])
    return txg.make_desc(ndims=3, ncomponents=2, shape=shape, periodic=(0, 0, 0), isotropy_order=isotropy_order,
                         relaxation_mode=txg.MRT, precision=precision, tau=[ta, tb], mm=[1.0, 1.0],
                         g=[[0.0, G_BUBBLE], [G_BUBBLE, 0.0]], s_c=[1 / ta, 1 / tb], s_nu=[1 / ta, 1 / tb],
                         s_e=[1.19] * 2, s_e2=[1.4] * 2, s_q=[1.2] * 2, gw=[list(gw1), list(gw2)],
                         bc={"xm": [RHO_MAJOR + DRAIN_DRHO, RHO_MINOR], "xp": [RHO_MINOR, RHO_MAJOR]},
                         rho0=[RHO_MINOR, RHO_MINOR])


def connected_to_inlet(walls, shape):
    """Pore nodes 6-connected to a pore node of the x = 0 plane."""
    from scipy import ndimage
    nx, ny, nz = shape
    pore = (walls.reshape(nz, ny, nx) == 0)
    lab, _ = ndimage.label(pore)
    keep = np.unique(lab[:, :, 0])
    keep = keep[keep > 0]
    return np.isin(lab, keep).reshape(-1)


def drainage_ic(n, seed, shape=None, inlet_layer=3, reservoir=3):
    """(walls, rho, u): porosity-0.40 pack of radii U[6, 12] (mineral per
    sphere with probability 1/2, non-periodic), filled with B except an
    inlet layer of A.  The `reservoir` planes next to the inlet and outlet
    faces are kept open (a density boundary node inside a pore throat, with
    bounce-back links on most sides, drives an ever faster jet through it),
    and pores not connected to the inlet are filled with mineral 1 (they
    take no part in the drainage)."""
    shape = shape or (n, n, n)
    walls = sphere_pack(shape, 0.40, 6.0, 12.0, seed, minerals=(1.0, 2.0), periodic=False)
    nx = shape[0]
    xs = np.tile(np.arange(nx), int(np.prod(shape)) // nx)
    walls = np.where((xs < reservoir) | (xs >= nx - reservoir), 0.0, walls)
    walls = np.where((walls == 0) & ~connected_to_inlet(walls, shape), 1.0, walls)
    nx = shape[0]
    x = np.tile(np.arange(nx), int(np.prod(shape)) // nx)
    a = np.where(x < inlet_layer, RHO_MAJOR, RHO_MINOR)
    b = np.where(x < inlet_layer, RHO_MINOR, RHO_MAJOR)
    rho = np.stack([a, b], axis=1).reshape(-1)
    return walls, rho, np.zeros(int(np.prod(shape)) * 2 * 3)


WORKLOADS = {
    # name: (description, desc builder, ic builder, nodes)
    "c1_bubble128_fp32_bgk": "D3Q19 two-component bubble 128^3, sphere r=24, periodic, G=1.8, "
                             "tau=1 BGK, FP32 compute and storage",
    "c1_bubble128_fp32_mrt": "D3Q19 two-component bubble 128^3, MRT s_e=1.19 s_eps=1.4 s_q=1.2 "
                             "s_nu=1/tau, FP32",
    "c1_bubble128_fp64_bgk": "D3Q19 two-component bubble 128^3, BGK, FP64",
    "c1_bubble128_fp16_bgk": "D3Q19 two-component bubble 128^3, BGK, FP32 compute, shifted FP16 storage",
    "c0_bubble200_fp64": "D2Q9 two-component bubble 200^2, disc r=40, FP64 BGK",
    "c2_drainage256_fp32": "D3Q19 two-phase drainage 256^3, sphere pack porosity 0.40 radii 6-12, two minerals "
                           "(45/120 degrees), Dirichlet drho=0.01 on x, walls elsewhere, MRT isotropy 8, "
                           "tau 1.0/0.6, FP32",
    "c2_drainage256_fp16": "c2_drainage256 with FP32 compute and shifted-FP16 storage",
    "c2_drainage256_fp64": "c2_drainage256 in FP64",
    "c2_drainage64_fp32": "c2 recipe at 64^3, FP32",
    "t1_species128_fp32": "D3Q19 two passive-scalar species 128^3 through the c2 sphere pack (porosity 0.40), "
                          "tau 0.8 / 1.2, a sphere of species A in species B, FP32",
    "t1_species128_fp64": "t1_species128 in FP64",
}


def workload(name):
    """(txg.Desc, walls, rho, u, description) of a named workload."""
    if name.startswith("c1_bubble128"):
        prec = {"fp32": txg.FP32, "fp64": txg.FP64, "fp16": txg.FP32_FP16}[name.split("_")[2]]
        mode = txg.MRT if name.endswith("mrt") else txg.SRT
        d = bubble_desc(3, 128, precision=prec, relaxation_mode=mode)
        walls, rho, u = bubble_ic(3, 128, 24)
        return d, walls, rho, u, WORKLOADS[name]
    if name.startswith("c2_drainage"):
        n = int(name.split("_")[1][len("drainage"):])
        prec = {"fp32": txg.FP32, "fp64": txg.FP64, "fp16": txg.FP32_FP16}[name.split("_")[2]]
        d = drainage_desc(n, precision=prec)
        walls, rho, u = drainage_ic(n, seed=2024)
        return d, walls, rho, u, WORKLOADS[name]
    if name.startswith("t1_species"):
        prec = {"fp32": txg.FP32, "fp64": txg.FP64}[name.split("_")[2]]
        d = txg.make_desc(ndims=3, ncomponents=2, shape=(128, 128, 128), periodic=(1, 1, 1), isotropy_order=0,
                          tau=[0.8, 1.2], precision=prec, transport=1)
        walls = sphere_pack((128, 128, 128), 0.40, 6, 12, seed=7)
        psi = ball_ic((128, 128, 128), 24, (1.0, 0.0), (0.0, 1.0)) * np.repeat(walls == 0, 2)
        return d, walls, psi, np.zeros(d.nnodes * 6), WORKLOADS[name]
    if name == "c0_bubble200_fp64":
        d = bubble_desc(2, 200, precision=txg.FP64)
        walls, rho, u = bubble_ic(2, 200, 40)
        return d, walls, rho, u, WORKLOADS[name]
    raise KeyError(name)
