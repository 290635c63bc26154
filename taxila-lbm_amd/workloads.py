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
def bubble_desc(ndims, n, precision=txg.FP32, relaxation_mode=txg.SRT, tau=1.0, rho0=0.5):
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


WORKLOADS = {
    # name: (description, desc builder, ic builder, nodes)
    "c1_bubble128_fp32_bgk": "D3Q19 two-component bubble 128^3, sphere r=24, periodic, G=1.8, "
                             "tau=1 BGK, FP32 compute and storage",
    "c1_bubble128_fp32_mrt": "D3Q19 two-component bubble 128^3, MRT s_e=1.19 s_eps=1.4 s_q=1.2 "
                             "s_nu=1/tau, FP32",
    "c1_bubble128_fp64_bgk": "D3Q19 two-component bubble 128^3, BGK, FP64",
    "c1_bubble128_fp16_bgk": "D3Q19 two-component bubble 128^3, BGK, FP32 compute, shifted FP16 storage",
    "c0_bubble200_fp64": "D2Q9 two-component bubble 200^2, disc r=40, FP64 BGK",
}


def workload(name):
    """(txg.Desc, walls, rho, u, description) of a named workload."""
    if name.startswith("c1_bubble128"):
        prec = {"fp32": txg.FP32, "fp64": txg.FP64, "fp16": txg.FP32_FP16}[name.split("_")[2]]
        mode = txg.MRT if name.endswith("mrt") else txg.SRT
        d = bubble_desc(3, 128, precision=prec, relaxation_mode=mode)
        walls, rho, u = bubble_ic(3, 128, 24)
        return d, walls, rho, u, WORKLOADS[name]
    if name == "c0_bubble200_fp64":
        d = bubble_desc(2, 200, precision=txg.FP64)
        walls, rho, u = bubble_ic(2, 200, 40)
        return d, walls, rho, u, WORKLOADS[name]
    raise KeyError(name)
