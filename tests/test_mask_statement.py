"""The numpy statement of the link masks (used by the GPU mask test) equals
a direct per-node loop over the reference's wall semantics on small grids
(CPU only): ghost 999 beyond non-periodic faces, periodic images, pull
sources that are walls, mineral neighbours, a wall-free stencil reach and
Dirichlet faces."""
import numpy as np
import pytest

import cases
import txg
from test_gpu_masks import FACE, FFCLEAN, FSNEAR, SOLID, numpy_mask


def loop_mask(d, walls):
    D = d.ndims
    nx, ny, nz = d.nx, d.ny, d.nz if D == 3 else 1
    reach = {0: 1, 4: 1, 8: 2, 10: 3}[d.isotropy_order]

    def cell(x, y, z):
        c = [x, y, z]
        n3 = [nx, ny, nz]
        for k in range(D):
            if c[k] < 0 or c[k] >= n3[k]:
                if not d.periodic[k]:
                    return 999.0
                c[k] %= n3[k]
        return walls[c[0] + nx * (c[1] + ny * c[2])]

    C = cases.lattice_c(D)
    out = np.zeros(nx * ny * nz, dtype=np.uint32)
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                i = x + nx * (y + ny * z)
                if walls[i] != 0:
                    out[i] = SOLID
                    continue
                mk = 0
                for n in range(1, len(C)):
                    c = list(C[n]) + [0] * (3 - D)
                    if cell(x - c[0], y - c[1], z - c[2]) > 0:
                        mk |= 1 << n
                    t = cell(x + c[0], y + c[1], z + c[2])
                    if 0 < t < 998 and 1 <= int(t) <= d.nminerals:
                        mk |= FSNEAR
                zr = range(-reach, reach + 1) if D == 3 else [0]
                r = range(-reach, reach + 1)
                if all(cell(x + dx, y + dy, z + dz) == 0 for dz in zr for dy in r for dx in r):
                    mk |= FFCLEAN
                pos, hi = [x, y, z], [nx - 1, ny - 1, nz - 1]
                for f in range(2 * D):
                    if d.bc_flags[f] == txg.BC_DENSITY and pos[f // 2] == (hi[f // 2] if f % 2 else 0):
                        mk |= 1 << (FACE + f)
                out[i] = mk
    return out


@pytest.mark.parametrize("kw,shape", [
    (dict(ndims=3, ncomponents=2, shape=(7, 6, 5), isotropy_order=4, gw=[[0.1, -0.1], [0, 0]],
          g=[[0, .1], [.1, 0]]), (7, 6, 5)),
    (dict(ndims=3, ncomponents=2, shape=(8, 7, 6), periodic=(0, 1, 0), isotropy_order=8,
          gw=[[0.1, -0.1], [-0.1, 0.1]], g=[[0, .1], [.1, 0]], bc={"xm": [0.98, 0.03], "xp": [0.97, 0.03]}),
     (8, 7, 6)),
    (dict(ndims=2, ncomponents=2, shape=(12, 9), periodic=(1, 0), isotropy_order=10,
          gw=[[0.1, -0.1], [-0.1, 0.1]], g=[[0, .1], [.1, 0]]), (12, 9)),
])
def test_numpy_mask_statement_matches_loop(kw, shape):
    d = txg.make_desc(**kw)
    w = cases.random_walls(shape, 0.25, seed=3, minerals=(1.0, 2.0))
    assert np.array_equal(numpy_mask(d, w), loop_mask(d, w))
