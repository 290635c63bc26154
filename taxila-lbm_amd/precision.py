"""Error of the FP32-compute modes against the FP64 mode, per workload.

    python taxila-lbm_amd/precision.py [workload] [steps]

runs a workload in FP64 (which reproduces the reference arithmetic bit for
bit), FP32 storage and shifted-FP16 storage on the GPU from the same initial
state and prints, for each reduced-precision mode, the max and RMS error of
the component densities, of the velocity u = sum_m mm_m (j_m + F_m/2) /
rho_t (FlowUpdateDiagnosticsD3, lbm_flow.F90:654-708), of the interface
position (bubble radius, or saturation and front position of a drainage)
and the relative mass change of each component over the run.

The interface position of a bubble is its equivalent radius: with
phi = clip((rho_A - rho_out) / (rho_in - rho_out), 0, 1) the bubble volume
is sum(phi) and R = (3 V / 4 pi)^(1/3) in 3-D (sqrt(A / pi) in 2-D).  For a
drainage run it is the invading-phase volume (saturation) and the mean
x of the invading front cells.
"""
from __future__ import annotations

import json
import math
import sys

import numpy as np

import txg
import workloads


def run(desc, walls, rho, u, steps):
    """(rho[C, N], u[D, N], mass0[C], mass1[C]) after `steps` steps."""
    dom = txg.Domain(desc)
    dom.set_walls(walls)
    dom.set_state(rho, u)
    dom.init()
    m0 = dom.reduce()[: desc.ncomponents]
    dom.step(steps)
    r, _, _ = dom.get_moments()
    _, _, v = dom.diagnostics()
    m1 = dom.reduce()[: desc.ncomponents]
    dom.close()
    C, D = desc.ncomponents, desc.ndims
    return r.reshape(-1, C).T.copy(), v.reshape(-1, D).T.copy(), m0, m1


def bubble_radius(rho_a, fluid, ndims, rho_in=workloads.RHO_MAJOR, rho_out=workloads.RHO_MINOR):
    phi = np.clip((rho_a - rho_out) / (rho_in - rho_out), 0.0, 1.0)[fluid]
    v = float(phi.sum())
    return (3.0 * v / (4.0 * math.pi)) ** (1.0 / 3.0) if ndims == 3 else math.sqrt(v / math.pi)


def invading(rho_a, fluid, shape, rho_in=workloads.RHO_MAJOR, rho_out=workloads.RHO_MINOR):
    """(saturation, mean x) of the invading component A: phi as for the
    bubble radius, saturation = mean phi over pore nodes, mean x = the
    phi-weighted x of the pore nodes (the front position of a drainage)."""
    phi = np.clip((rho_a - rho_out) / (rho_in - rho_out), 0.0, 1.0)
    x = np.tile(np.arange(shape[0]), int(np.prod(shape)) // shape[0])
    w = phi[fluid]
    return float(w.mean()), float((w * x[fluid]).sum() / max(w.sum(), 1e-300))


def errors(ref, out, fluid, ndims, shape=None, drainage=False):
    """Error figures of `out` against `ref` (both from run())."""
    r0, u0, _, _ = ref
    r1, u1, m0, m1 = out
    e = {}
    for m in range(r0.shape[0]):
        d = np.abs(r1[m] - r0[m])[fluid]
        e["rho%d_max" % m] = float(d.max())
        e["rho%d_rms" % m] = float(np.sqrt(np.mean(d * d)))
        e["rho%d_max_rel" % m] = float((d / np.abs(r0[m][fluid])).max())
        e["mass%d_rel_change" % m] = float((m1[m] - m0[m]) / m0[m])
    du = np.sqrt(((u1 - u0) ** 2).sum(axis=0))[fluid]
    e["u_max"] = float(du.max())
    e["u_rms"] = float(np.sqrt(np.mean(du * du)))
    e["u_ref_max"] = float(np.sqrt((u0 ** 2).sum(axis=0))[fluid].max())
    if drainage:
        (sa, xa), (sb, xb) = invading(r0[0], fluid, shape), invading(r1[0], fluid, shape)
        e["saturation_ref"], e["saturation_err"] = sa, sb - sa
        e["front_x_ref"], e["front_x_err"] = xa, xb - xa
    else:
        ra, rb = bubble_radius(r0[0], fluid, ndims), bubble_radius(r1[0], fluid, ndims)
        e["radius_ref"] = ra
        e["radius_err"] = rb - ra
    return e


def compare(name: str, steps: int, fp16_shifts=None):
    """{mode name: error dict} of the reduced-precision modes of a workload.
    fp16_shifts: storage shifts rho0 to try with FP16 storage (default: the
    workload's own)."""
    d, walls, rho, u, _ = workloads.workload(name)
    fluid = walls == 0
    C = d.ncomponents
    shape = (d.nx, d.ny, d.nz) if d.ndims == 3 else (d.nx, d.ny)
    kw = dict(shape=shape, drainage=name.startswith("c2_drainage"))
    own = [d.rho0[m] for m in range(C)]
    d.precision = txg.FP64
    ref = run(d, walls, rho, u, steps)
    out = {"workload": name, "steps": steps,
           "fp64_mass_rel_change": [float((ref[3][m] - ref[2][m]) / ref[2][m]) for m in range(C)]}
    d.precision = txg.FP32
    out["fp32"] = errors(ref, run(d, walls, rho, u, steps), fluid, d.ndims, **kw)
    d.precision = txg.FP32_FP16
    for s in (fp16_shifts or [None]):
        for m in range(C):
            d.rho0[m] = own[m] if s is None else s
        key = "fp16" if s is None else "fp16_shift%g" % s
        out[key] = errors(ref, run(d, walls, rho, u, steps), fluid, d.ndims, **kw)
        out[key]["shift"] = [d.rho0[m] for m in range(C)]
    return out


if __name__ == "__main__":
    name = sys.argv[1] if len(sys.argv) > 1 else "c1_bubble128_fp32_bgk"
    steps = int(sys.argv[2]) if len(sys.argv) > 2 else 1000
    shifts = [float(x) for x in sys.argv[3].split(",")] if len(sys.argv) > 3 else None
    print(json.dumps(compare(name, steps, shifts), indent=1))
