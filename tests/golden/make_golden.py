"""Writes tests/golden/oracle_small.npz: outputs of the CPU oracle on small
seeded inputs (test_constants.jl geometry).  The reference ships no data
files for this path, so these fixtures freeze the oracle once its pins in
tests/test_oracle.py pass; tests compare against them to catch any later
change of the oracle.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (os.path.join(ROOT, "erebus.jl_amd"), os.path.join(ROOT, "oracle"),
          os.path.dirname(HERE)):
    if p not in sys.path:
        sys.path.insert(0, p)

import cases  # noqa: E402
from ereb.constants import test_constants  # noqa: E402
from oracle import Oracle  # noqa: E402

OUT = os.path.join(HERE, "oracle_small.npz")


def compute() -> dict:
    g, mat = test_constants()
    orc = Oracle(g, mat)
    out = {}
    mk = cases.ref_markers(g, mat, seed=1234, extra=500)
    hrs, hrf = np.array([2e-6, 1e-6, 0.0]), np.array([4e-7, 0.0, 0.0])
    for mode in (1, 9):
        props, sums = orc.marker_to_nodes(mk, hrs, hrf, mode)
        for k, v in props.items():
            out[f"m{mode}_{k}"] = v
        out[f"m{mode}_sums"] = np.stack([s.ravel(order="F") for s in sums[8:]])
        out[f"m{mode}_sums_basic"] = np.stack([s.ravel(order="F") for s in sums[:8]])
    rng = np.random.default_rng(99)
    S = (g.Ny1, g.Nx1)
    vx, vy, vxf, vyf = (cases.rand_grid(rng, S, -1, 1) for _ in range(4))
    wyx = cases.rand_grid(rng, (g.Ny, g.Nx), -1e-3, 1e-3)
    tk2 = cases.rand_grid(rng, S, 263, 310)
    r = orc.move_markers_rk4(mk["xm"], mk["ym"], mk["tm"], mk["tkm"], mk["phim"], mk["sxym"],
                             mk["sxxm"], vx, vy, vxf, vyf, wyx, tk2, 0.9 * g.dx, 1)
    for k, v in zip(["xm", "ym", "tkm", "sxym", "sxxm"], r):
        out[f"rk4_{k}"] = v
    DSXX = cases.rand_grid(rng, S, -1e5, 1e5)
    DSXY = cases.rand_grid(rng, (g.Ny, g.Nx), -1e5, 1e5)
    out["stress_sxx"], out["stress_sxy"] = orc.update_marker_stress(
        mk["xm"], mk["ym"], mk["sxxm"], mk["sxym"], DSXX, DSXY)
    out["temp_t2"] = orc.update_marker_temperature(mk["xm"], mk["ym"], mk["tkm"], DSXX / 1e5,
                                                   tk2, 2)
    cp, rv, nz, R = orc.assemble_hydromechanical(cases.stokes_inputs(g), 1e11)
    out.update(sd_colptr=cp, sd_rowval=rv, sd_nzval=nz, sd_R=R)
    cp, rv, nz, R = orc.assemble_thermal(cases.thermal_inputs(g), 3e10)
    out.update(t_colptr=cp, t_rowval=rv, t_nzval=nz, t_R=R)
    perm, key, start = orc.fine_sort(mk["xm"], mk["ym"])
    out.update(sort_perm=perm, sort_key=key, sort_start=start)
    return out


if __name__ == "__main__":
    np.savez_compressed(OUT, **compute())
    print(OUT, os.path.getsize(OUT), "bytes")
