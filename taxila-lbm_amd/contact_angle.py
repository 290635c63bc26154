"""Contact angles of the drainage recipe's wettability, measured on sessile
droplets, and the calibration of its two minerals' -gw values.

    python taxila-lbm_amd/contact_angle.py [--g 0.15] [--steps N] [--out FILE]

A droplet of component A (rho_A = 0.97) sits on a flat mineral wall in B,
with the drainage recipe's fluid physics (D3Q19, two components, MRT with
s_e = 1.19, s_eps = 1.4, s_q = 1.2, s_nu = s_c = 1/tau, tau_A = 1.0,
tau_B = 0.6, isotropy-8 psi-gradient, FP32) and a coupling -g_12 = -g_21 = g.

The drainage recipe's own g = 0.1 (G = 18 g = 1.8) at total density 1 lies
below the critical point of the two-component Shan-Chen mixture: its
pressure p = rho/3 + (G/3) rho_A rho_B makes the mixed state stable unless
G^2 rho_A rho_B > 1, i.e. G rho > 2 at equal parts, so a droplet of A
dissolves into B (its peak density falls from 0.97 to 0.6 within 600 steps
at radius 14) and has no contact angle.  The default g = 0.15 (G = 2.7)
separates the phases; the measured relation between -gw and the angle is
reported for it.  The wall is the
plane z = 0 (mineral 1, fluid-solid force LBMAddFluidSolidForcesD3,
lbm_forcing.F90:1326-1373), x and y are periodic, z+ is a ghost-999 wall.
After the run, the A/B interface (rho_A = rho_B, linear interpolation along
the lattice lines) is fitted by a circle in the two axial slices through the
droplet's centre (points more than 2.5 nodes above the wall plane, the
halfway bounce-back plane z = 0.5); the contact angle through A is
cos(theta) = (z_wall - z_centre) / R.

The Huang et al. relation (Phys. Rev. E 76, 066701 (2007)) used by
workloads.wall_adhesion assumes the isotropy-4 stencil and bulk densities;
the calibration scales its -gw pair by bisection until the measured angle is
within 1 degree of each target (a film counts as 0 degrees, a lifted-off
droplet as 180).
"""
from __future__ import annotations

import argparse
import json
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

import txg  # noqa: E402
import workloads  # noqa: E402

SHAPE = (96, 96, 56)
RADIUS = 20.0


def droplet_desc(gw, precision=txg.FP32, shape=SHAPE, g=0.15):
    ta, tb = workloads.DRAIN_TAU
    return txg.make_desc(ndims=3, ncomponents=2, shape=shape, periodic=(1, 1, 0), isotropy_order=8,
                         relaxation_mode=txg.MRT, precision=precision, tau=[ta, tb], mm=[1.0, 1.0],
                         g=[[0.0, g], [g, 0.0]], s_c=[1 / ta, 1 / tb], s_nu=[1 / ta, 1 / tb],
                         s_e=[1.19] * 2, s_e2=[1.4] * 2, s_q=[1.2] * 2, gw=[list(gw)],
                         rho0=[workloads.RHO_MINOR] * 2)


def coexistence(g, steps=20000):
    """Bulk densities (major, minor) of the two coexisting phases at this g:
    two flat slabs, A and B, on a periodic 4 x 4 x 64 lattice relaxed to
    equilibrium with the droplet's fluid physics; each phase read at its
    slab's centre."""
    shape = (4, 4, 64)
    d = droplet_desc((0.0, 0.0), txg.FP64, shape, g)
    d.periodic[2] = 1
    d.nminerals = 0
    n = int(np.prod(shape))
    z = np.repeat(np.arange(64), 16)
    inA = (z >= 16) & (z < 48)
    hi, lo = workloads.RHO_MAJOR, workloads.RHO_MINOR
    rho = np.stack([np.where(inA, hi, lo), np.where(inA, lo, hi)], axis=1).reshape(-1)
    dom = txg.Domain(d)
    dom.set_walls(np.zeros(n))
    dom.set_state(rho, np.zeros(n * 6))
    dom.init()
    dom.step(steps)
    r = dom.get_moments()[0].reshape(64, 16, 2)
    dom.close()
    return float(r[32, 0, 0]), float(r[0, 0, 0])


def droplet_ic(shape=SHAPE, radius=RADIUS, major=workloads.RHO_MAJOR, minor=workloads.RHO_MINOR):
    """Mineral-1 plane at z = 0; a hemisphere of A (densities major/minor of
    A/B inside, minor/major outside) centred on the wall plane."""
    nx, ny, nz = shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    walls = np.where(z == 0, 1.0, 0.0).reshape(-1)
    r2 = (x - nx // 2) ** 2 + (y - ny // 2) ** 2 + (z - 0.5) ** 2
    ins = (r2 <= radius * radius) & (z > 0)
    hi, lo = major, minor
    a = np.where(ins, hi, lo).reshape(-1)
    b = np.where(ins, lo, hi).reshape(-1)
    fl = walls == 0
    rho = (np.stack([a, b], axis=1) * fl[:, None]).reshape(-1)
    return walls, rho, np.zeros(nx * ny * nz * 6)


def interface_points(phi):
    """Zero crossings of phi (2-D array [z, s]) along rows and columns."""
    pts = []
    nz, ns = phi.shape
    for axis in (0, 1):
        a = phi if axis == 0 else phi.T
        p, q = a[:-1], a[1:]
        i, j = np.nonzero((p > 0) != (q > 0))
        t = p[i, j] / (p[i, j] - q[i, j])
        c = i + t
        if axis == 0:
            pts.append(np.stack([j.astype(float), c], axis=1))   # (s, z)
        else:
            pts.append(np.stack([c, j.astype(float)], axis=1))
    return np.concatenate(pts)


def fit_circle(pts):
    """Algebraic least-squares circle (Kasa): returns (s_c, z_c, R)."""
    s, z = pts[:, 0], pts[:, 1]
    A = np.stack([s, z, np.ones_like(s)], axis=1)
    b = s * s + z * z
    (cs, cz, c0), *_ = np.linalg.lstsq(A, b, rcond=None)
    sc, zc = cs / 2, cz / 2
    return sc, zc, float(np.sqrt(c0 + sc * sc + zc * zc))


def measure(rho, shape=SHAPE, zwall=0.5):
    """Contact angle (degrees, through A) from the two axial slices."""
    nx, ny, nz = shape
    r = rho.reshape(nz, ny, nx, 2)
    phi = r[..., 0] - r[..., 1]
    phi[0] = -1.0  # the wall plane counts as B for the crossings
    cy, cx = ny // 2, nx // 2
    angles = []
    for sl in (phi[:, cy, :], phi[:, :, cx]):
        pts = interface_points(sl)
        pts = pts[pts[:, 1] > zwall + 2.5]
        if len(pts) < 8:
            raise ValueError("droplet too flat to fit: %d interface points above the wall" % len(pts))
        sc, zc, R = fit_circle(pts)
        angles.append(float(np.degrees(np.arccos(np.clip((zwall - zc) / R, -1, 1)))))
    return float(np.mean(angles)), angles


def angle_or_limit(rho, shape=SHAPE):
    """The measured angle, or 0 / 180 degrees when A has spread into a film
    on the wall / lifted off it (no cap left to fit)."""
    try:
        return measure(rho, shape)
    except ValueError:
        nx, ny, nz = shape
        r = rho.reshape(nz, ny, nx, 2)
        film = bool((r[1, :, :, 0] > r[1, :, :, 1]).any())
        return (0.0 if film else 180.0), []


COEX = {}


def run(scale, target, steps, g, precision=txg.FP32):
    """A droplet with -gw = scale x the Huang relation's values for target,
    started at the coexisting densities of g."""
    if g not in COEX:
        COEX[g] = coexistence(g)
    major, minor = COEX[g]
    gw = tuple(scale * v for v in workloads.wall_adhesion(target, g=g, rho_major=major, rho_minor=minor))
    return droplet_angle(gw, steps, g, precision)


def droplet_angle(gw, steps, g, precision=txg.FP32):
    """(gw, angle, per-slice angles) of a droplet on a wall with these -gw."""
    if g not in COEX:
        COEX[g] = coexistence(g)
    major, minor = COEX[g]
    d = droplet_desc(gw, precision, g=g)
    walls, rho, u = droplet_ic(major=major, minor=minor)
    dom = txg.Domain(d)
    dom.set_walls(walls)
    dom.set_state(rho, u)
    dom.init()
    dom.step(steps)
    r, _, _ = dom.get_moments()
    dom.close()
    theta, parts = angle_or_limit(r)
    return gw, theta, parts


def calibrate(target, steps, g, log, iters=8):
    """Bisection on the scale of the relation's -gw values (the angle moves
    monotonically away from 90 degrees as the scale grows) until the
    measured angle is within 1 degree of the target."""
    lo, hi = 0.0, 1.0
    best = None
    for _ in range(iters):
        sc = 0.5 * (lo + hi)
        gw, th, parts = run(sc, target, steps, g)
        log.append({"target": target, "scale": sc, "gw": list(gw), "measured": th, "slices": parts})
        print(json.dumps(log[-1]), flush=True)
        if best is None or abs(th - target) < abs(best[2] - target):
            best = (sc, gw, th)
        if abs(th - target) <= 1.0:
            break
        # too close to 90 degrees: more adhesion contrast; past the target: less
        if abs(th - 90.0) < abs(target - 90.0):
            lo = sc
        else:
            hi = sc
    return best


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=8000)
    ap.add_argument("--g", type=float, default=0.15)
    ap.add_argument("--out", default="")
    args = ap.parse_args()
    log = []
    result = {"shape": SHAPE, "radius0": RADIUS, "steps": args.steps, "g": args.g, "targets": {}}
    for target in workloads.DRAIN_ANGLES:
        _, th_rel, _ = run(1.0, target, args.steps, args.g)  # the relation as it stands
        sc, gw, th = calibrate(target, args.steps, args.g, log)
        result["targets"][str(target)] = {"relation_measured": th_rel, "scale": sc, "gw": list(gw), "measured": th}
    # convergence: the calibrated values measured again after twice the steps
    for target in workloads.DRAIN_ANGLES:
        sc = result["targets"][str(target)]["scale"]
        _, th2, _ = run(sc, target, 2 * args.steps, args.g)
        result["targets"][str(target)]["measured_at_2x_steps"] = th2
    result["coexistence"] = COEX.get(args.g)
    result["log"] = log
    print(json.dumps(result))
    if args.out:
        with open(args.out, "w") as f:
            json.dump(result, f, indent=1)


if __name__ == "__main__":
    main()
