"""Wall/fluid mask and boundary links built on the GPU (txg_set_walls),
bit for bit against an independent numpy statement of the reference's wall
semantics: ghost layers of non-periodic faces are 999 walls
(WallsSetGhostNodesD3, lbm_walls.F90:208-231), periodic faces wrap
(WallsCommunicate, :233-244), a pull from a wall is a bounce-back link
(lbm_distribution_function.F90:669-734), a mineral neighbour is a
wettability link (lbm_forcing.F90:1326-1373)."""
import numpy as np
import pytest

import cases
import txg

pytestmark = pytest.mark.gpu

SOLID, FSNEAR, FFCLEAN, FACE = 1 << 31, 1 << 19, 1 << 20, 24


def numpy_mask(d, walls):
    D = d.ndims
    shape = (d.nz, d.ny, d.nx) if D == 3 else (d.ny, d.nx)
    w = walls.reshape(shape)
    reach = {0: 1, 4: 1, 8: 2, 10: 3}[d.isotropy_order]
    pad = reach + 1
    # ghosted field: periodic images or 999 (axis order z, y, x)
    per = [d.periodic[k] for k in range(D)][::-1]
    g = w
    for ax in range(D):
        width = [(pad, pad) if a == ax else (0, 0) for a in range(D)]
        g = np.pad(g, width, mode="wrap") if per[ax] else np.pad(g, width, mode="constant", constant_values=999.0)

    def at(off):  # value at x + off (off in x, y, z order)
        sl = []
        for ax in range(D):
            o = off[D - 1 - ax]
            sl.append(slice(pad + o, pad + o + shape[ax]))
        return g[tuple(sl)]

    C = cases.lattice_c(D)
    mask = np.zeros(shape, dtype=np.uint64)
    for n in range(1, len(C)):
        mask |= np.where(at(-C[n]) > 0, 1 << n, 0).astype(np.uint64)
        t = at(C[n])
        mineral = (t > 0) & (t < 998) & (np.floor(t) >= 1) & (np.floor(t) <= d.nminerals)
        mask |= np.where(mineral, FSNEAR, 0).astype(np.uint64)
    clean = np.ones(shape, dtype=bool)
    rng = range(-reach, reach + 1)
    for dz in (rng if D == 3 else [0]):
        for dy in rng:
            for dx in rng:
                clean &= at((dx, dy, dz)[:D] if D == 3 else (dx, dy)) == 0
    mask |= np.where(clean, FFCLEAN, 0).astype(np.uint64)
    idx = np.indices(shape)
    for f in range(2 * D):
        if d.bc_flags[f] == txg.BC_DENSITY:
            ax = D - 1 - f // 2
            edge = 0 if f % 2 == 0 else shape[ax] - 1
            mask |= np.where(idx[ax] == edge, 1 << (FACE + f), 0).astype(np.uint64)
    mask = np.where(w != 0, SOLID, mask)
    return mask.reshape(-1).astype(np.uint32)


@pytest.mark.parametrize("case", ["periodic_iso4", "walls_iso8_dirichlet", "d2_mixed_iso10"])
def test_link_masks_bit_exact(case):
    if case == "periodic_iso4":
        d = txg.make_desc(ndims=3, ncomponents=2, shape=(21, 17, 13), isotropy_order=4, gw=[[0.1, -0.1], [0, 0]],
                          g=[[0, 0.1], [0.1, 0]])
    elif case == "walls_iso8_dirichlet":
        d = txg.make_desc(ndims=3, ncomponents=2, shape=(24, 19, 16), periodic=(0, 1, 0), isotropy_order=8,
                          g=[[0, 0.1], [0.1, 0]], gw=[[0.1, -0.1], [-0.1, 0.1]],
                          bc={"xm": [0.98, 0.03], "xp": [0.97, 0.03]})
    else:
        d = txg.make_desc(ndims=2, ncomponents=2, shape=(40, 33), periodic=(1, 0), isotropy_order=10,
                          g=[[0, 0.1], [0.1, 0]], gw=[[0.1, -0.1], [-0.1, 0.1]])
    shape = (d.nx, d.ny, d.nz) if d.ndims == 3 else (d.nx, d.ny)
    walls = cases.random_walls(shape, 0.2, seed=11, minerals=(1.0, 2.0) if d.nminerals >= 2 else (1.0,))
    dom = txg.Domain(d)
    dom.set_walls(walls)
    got = dom.get_mask()
    want = numpy_mask(d, walls)
    diff = np.nonzero(got != want)[0]
    assert diff.size == 0, (diff[:5], got[diff[:5]], want[diff[:5]])
