"""Synthetic inputs shared by the tests: initial conditions and wall fields.

All random fields come from numpy's seeded PCG64 so that the oracle and the
HIP path see bit-identical inputs.
"""
import lzma
import os

import numpy as np

import txg

# lattice velocities and weights in the reference's direction order
# (lbm_discretization_d3q19.F90:20-38, lbm_discretization_d2q9.F90:20-29)
C3 = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [-1, 0, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1], [1, 1, 0],
               [-1, 1, 0], [-1, -1, 0], [1, -1, 0], [1, 0, 1], [-1, 0, 1], [-1, 0, -1], [1, 0, -1], [0, 1, 1],
               [0, -1, 1], [0, -1, -1], [0, 1, -1]])
W3 = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
C2 = np.array([[0, 0], [1, 0], [0, 1], [-1, 0], [0, -1], [1, 1], [-1, 1], [-1, -1], [1, -1]])
W2 = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)


def lattice_c(ndims):
    return C3 if ndims == 3 else C2

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden_vec(name):
    with open(os.path.join(GOLDEN, name), "rb") as f:
        return txg.read_petsc_vec(lzma.decompress(f.read()))


def bubble_2d_fixture():
    """tests/bubble_2D of the reference: D2Q9, SRT tau=1, g_12=g_21=0.1,
    128x128 periodic, a 53x53 square [38..90]^2 (tests/bubble_2D/initialize_state.F90:187-203)
    with rho_inner=(0.03,0.97) inside and rho_outer=(0.97,0.03) outside
    (tests/bubble_2D/input_data)."""
    d = txg.make_desc(ndims=2, ncomponents=2, shape=(128, 128), periodic=(1, 1), isotropy_order=4,
                      relaxation_mode=txg.SRT, tau=[1.0, 1.0], mm=[1.0, 1.0], g=[[0.0, 0.1], [0.1, 0.0]])
    rho = square_ic((128, 128), 26, (0.03, 0.97), (0.97, 0.03))
    return d, np.zeros(128 * 128), rho, np.zeros(128 * 128 * 4)


def square_ic(shape, half, inner, outer):
    nx, ny = shape
    lx, rx = (nx + 1) // 2 - half, (nx + 1) // 2 + half
    ly, ry = (ny + 1) // 2 - half, (ny + 1) // 2 + half
    i = np.arange(1, nx + 1)[None, :]
    j = np.arange(1, ny + 1)[:, None]
    ins = (i >= lx) & (i <= rx) & (j >= ly) & (j <= ry)
    rho = np.where(ins[..., None], np.array(inner), np.array(outer))
    return rho.reshape(-1)


def ball_ic(shape, radius, inner, outer, center=None):
    """Disc / sphere of `inner` densities in `outer`, centre (N+1)/2 in 1-based
    coordinates (initialize_state_bubble.F90:187-190 convention)."""
    dims = len(shape)
    c = center or tuple((n + 1) // 2 for n in shape)
    grids = np.meshgrid(*[np.arange(1, n + 1) for n in shape], indexing="ij")
    r2 = sum((g - cc) ** 2 for g, cc in zip(grids, c))
    ins = r2 <= radius * radius
    # natural order: x fastest -> transpose (x,y,z) index arrays to (z,y,x)
    ins = np.transpose(ins, tuple(range(dims))[::-1])
    rho = np.where(ins[..., None], np.array(inner), np.array(outer))
    return rho.reshape(-1)


def random_walls(shape, fraction, seed, minerals=(1,), ghost_tags=False):
    """Random solid voxels (tag = a mineral id) at the given volume fraction."""
    rng = np.random.default_rng(seed)
    n = int(np.prod(shape))
    solid = rng.random(n) < fraction
    tags = rng.choice(np.array(minerals, dtype=np.float64), size=n)
    return np.where(solid, tags, 0.0)


def sphere_pack(shape, porosity, rmin, rmax, seed, minerals=(1.0, 2.0)):
    """Overlapping-sphere pack: spheres with radii U[rmin,rmax] at random
    centres (periodic images) are added until the pore fraction drops to
    `porosity`; each sphere's mineral is chosen uniformly from `minerals`, the
    first sphere covering a voxel wins.  Natural order, x fastest."""
    rng = np.random.default_rng(seed)
    nz, ny, nx = shape[::-1] if len(shape) == 3 else (1,) + tuple(shape[::-1])
    tags = np.zeros((nz, ny, nx))
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    target = 1.0 - porosity
    while (tags != 0).mean() < target:
        r = rng.uniform(rmin, rmax)
        c = [rng.uniform(0, n) for n in (nx, ny, nz)]
        mineral = rng.choice(np.array(minerals))
        dx = np.abs(x - c[0]); dx = np.minimum(dx, nx - dx)
        dy = np.abs(y - c[1]); dy = np.minimum(dy, ny - dy)
        if len(shape) == 3:
            dz = np.abs(z - c[2]); dz = np.minimum(dz, nz - dz)
        else:
            dz = 0
        ins = (dx * dx + dy * dy + dz * dz <= r * r) & (tags == 0)
        tags[ins] = mineral
    return tags.reshape(-1)


def run(cls, desc, walls, rho, u, steps, fi=None):
    """Drive a Domain (HIP) or an Oracle through set-up, init and steps."""
    dom = cls(desc)
    dom.set_walls(walls)
    if fi is not None:
        dom.set_fi(fi)
    else:
        dom.set_state(rho, u)
    dom.init()
    if steps:
        dom.step(steps)
    return dom
