"""Seeded inputs shared by the CPU and GPU tests.

They mirror the shapes of the reference's own tests (test/runtests.jl):
the test_constants.jl geometry (15x15 nodes, 4x4 markers per cell, 3,136
markers) plus 10,000 extra markers placed uniformly in [-dx, xsize+dx] x
[-dy, ysize+dy] (T:1597-1600, T:1806-1812) so that clamped cells and
extrapolated weights are exercised.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse

from ereb.constants import Geometry, Materials


def lattice(g: Geometry):
    Nxm, Nym = g.Nxm, g.Nym
    jm, im = np.meshgrid(np.arange(Nxm), np.arange(Nym), indexing="ij")
    return g.dxm / 2 + jm.ravel() * g.dxm, g.dym / 2 + im.ravel() * g.dym


def ref_markers(g: Geometry, mat: Materials, seed: int = 42, extra: int = 10_000,
                jitter: bool = True) -> dict:
    """Random marker state on the lattice plus `extra` markers anywhere in
    the extended box; every property random (positive where the reference
    divides by it)."""
    rng = np.random.default_rng(seed)
    xl, yl = lattice(g)
    if jitter:
        xl = xl + (rng.random(xl.size) - 0.5) * g.dxm
        yl = yl + (rng.random(yl.size) - 0.5) * g.dym
    xe = -g.dx + rng.random(extra) * (g.xsize + 2 * g.dx)
    ye = -g.dy + rng.random(extra) * (g.ysize + 2 * g.dy)
    xm = np.concatenate([xl, xe])
    ym = np.concatenate([yl, ye])
    M = xm.size
    tm = rng.integers(1, 4, M).astype(np.int64)
    d = {
        "xm": xm, "ym": ym, "tm": tm,
        "tkm": rng.choice(np.arange(150.0, 480.0, 0.1), M),  # spans all mode-1 branches
        "phim": mat.phimin + rng.random(M) * (0.5 - mat.phimin),
        "sxxm": rng.random(M) * 1e6 - 5e5, "sxym": rng.random(M) * 1e6 - 5e5,
        "XWsolidm0": rng.random(M),
        "etavpm": 1e16 + rng.random(M) * 1e19,
        "inv_gggtotalm": 1e-10 * (1 + rng.random(M)),
        "fricttotalm": rng.random(M),
        "cohestotalm": 1e8 * rng.random(M),
        "tenstotalm": 6e7 * rng.random(M),
        "rhofluidcur": 900 + 200 * rng.random(M),
        "alphasolidcur": 3e-5 * rng.random(M),
        "alphafluidcur": 5e-5 * rng.random(M),
        "rhototalm": 1 + 3000 * rng.random(M),
        "rhocptotalm": 1e6 + 3e6 * rng.random(M),
        "etatotalm": 1e16 * (1 + rng.random(M)),
        "hrtotalm": 1e-6 * rng.random(M),
        "ktotalm": 1 + 3000 * rng.random(M),
        "tkm_rhocptotalm": np.zeros(M),
        "etafluidcur_inv_kphim": np.zeros(M),
    }
    return {k: np.ascontiguousarray(v) for k, v in d.items()}


def rand_grid(rng, shape, lo=0.0, hi=1.0):
    return np.asfortranarray(lo + (hi - lo) * rng.random(shape))


def stokes_inputs(g: Geometry, seed: int = 7, air_fraction: float = 0.3) -> list:
    """The 20 grids of assemble_hydromechanical_lse! scaled as T:4097-4116.
    A block of constant "air" values makes some stencil coefficients exactly
    zero, which exercises ExtendableSparse's zero skip."""
    rng = np.random.default_rng(seed)
    B, S = (g.Ny, g.Nx), (g.Ny1, g.Nx1)
    ETA = rand_grid(rng, B) * 1e16
    ETAP = rand_grid(rng, S) * 1e16
    GGG = rand_grid(rng, B) * 1e10
    GGGP = rand_grid(rng, S) * 1e10
    SXY0 = rand_grid(rng, B) * 1e4
    SXX0 = rand_grid(rng, S) * 1e4
    RHOX = rand_grid(rng, S) * 1e4
    RHOY = rand_grid(rng, S) * 1e4
    RHOFX = rand_grid(rng, S) * 1e4
    RHOFY = rand_grid(rng, S) * 1e4
    RX = rand_grid(rng, S) * 1e39
    RY = rand_grid(rng, S) * 1e39
    ETAPHI = rand_grid(rng, S) * 1e15
    BETAPHI = rand_grid(rng, S) * 1e-10
    PHI = rand_grid(rng, S) * 0.9
    gx = rand_grid(rng, S)
    gy = rand_grid(rng, S)
    pr0 = rand_grid(rng, S) * 1e5
    pf0 = rand_grid(rng, S) * 1e5
    DMP = rand_grid(rng, S) * 1e5
    # constant region: equal viscosities and flat density -> zero cross terms
    na = max(2, int(g.Nx * air_fraction))
    for a, v in ((ETA, 1e16), (GGG, 1e10), (SXY0, 0.0)):
        a[:, :na] = v
    for a, v in ((ETAP, 1e16), (GGGP, 1e10), (RHOX, 1.0), (RHOY, 1.0), (SXX0, 0.0)):
        a[:, :na] = v
    RX[:3, :3] = 0.0  # zero Darcy coefficients must vanish from the pattern
    return [ETA, ETAP, GGG, GGGP, SXY0, SXX0, RHOX, RHOY, RHOFX, RHOFY, RX, RY, ETAPHI,
            BETAPHI, PHI, gx, gy, pr0, pf0, DMP]


def thermal_inputs(g: Geometry, seed: int = 9) -> list:
    rng = np.random.default_rng(seed)
    S = (g.Ny1, g.Nx1)
    tk1 = rand_grid(rng, S, 150, 400)
    RHOCP = rand_grid(rng, S, 1e6, 4e6)
    KX = rand_grid(rng, S, 1, 3000)
    KY = rand_grid(rng, S, 1, 3000)
    KX[:2, :] = 0.0  # zero conductivities drop their off-diagonals
    HR = rand_grid(rng, S) * 1e-6
    HA = rand_grid(rng, S) * 1e-7
    HS = rand_grid(rng, S) * 1e-7
    DHP = rand_grid(rng, S) * 1e-7
    return [tk1, RHOCP, KX, KY, HR, HA, HS, DHP]


# --- error bound for reordered sums -----------------------------------------

def _cells(x, y, x0, y0, dx, dy, jmin, jmax, imin, imax):
    tx = (x - x0) * (1.0 / dx)
    ty = (y - y0) * (1.0 / dy)
    j = np.clip(np.trunc(np.nan_to_num(tx, nan=-1e18)).astype(np.int64) + 1, jmin, jmax)
    i = np.clip(np.trunc(np.nan_to_num(ty, nan=-1e18)).astype(np.int64) + 1, imin, imax)
    return i, j


def abs_sums(g: Geometry, mk: dict, props: dict) -> list:
    """Per node sum of |property * weight| for the 30 SUM grids (the scale
    against which a reordered floating-point sum is compared)."""
    a = g.axes()
    Nx, Ny = g.Nx, g.Ny
    x, y = mk["xm"], mk["ym"]
    spec = {
        "basic": (a["x"], a["y"], (1, Nx - 1, 1, Ny - 1), (Ny, Nx)),
        "vx": (a["xvx"], a["yvx"], (1, Nx - 1, 1, Ny), (Ny + 1, Nx + 1)),
        "vy": (a["xvy"], a["yvy"], (1, Nx, 1, Ny - 1), (Ny + 1, Nx + 1)),
        "p": (a["xp"], a["yp"], (1, Nx, 1, Ny), (Ny + 1, Nx + 1)),
    }
    fields = {
        "basic": ["etatotalm", "etavpm", "inv_gggtotalm", "sxym", "cohestotalm",
                  "tenstotalm", "fricttotalm", None],
        "vx": ["rhototalm", "rhofluidcur", "ktotalm", "phim", "etafluidcur_inv_kphim", None],
        "vy": ["rhototalm", "rhofluidcur", "ktotalm", "phim", "etafluidcur_inv_kphim", None],
        "p": ["rhototalm", "rhocptotalm", "alphasolidcur", "alphafluidcur", "hrtotalm",
              "inv_gggtotalm", "sxxm", "tkm_rhocptotalm", "phim", None],
    }
    allv = dict(mk)
    allv.update(props)
    out = []
    M = x.size
    for gname in ("basic", "vx", "vy", "p"):
        ax, ay, lim, shp = spec[gname]
        i, j = _cells(x, y, ax[0], ay[0], g.dx, g.dy, *lim)
        A = np.abs((x - ax[j - 1]) / g.dx)
        B = np.abs((y - ay[i - 1]) / g.dy)
        # |weights| as a sparse (node x marker) matrix, 4 entries per marker
        # (column-major node index), applied to all |property| columns at once
        w = np.stack([np.abs(1 - B) * np.abs(1 - A), B * np.abs(1 - A), np.abs(1 - B) * A,
                      B * A], axis=1)
        node = np.stack([(j - 1 + dj) * shp[0] + (i - 1 + di)
                         for di, dj in ((0, 0), (1, 0), (0, 1), (1, 1))], axis=1)
        W = scipy.sparse.csc_matrix((w.ravel(), node.ravel(), np.arange(0, 4 * M + 1, 4)),
                                    shape=(shp[0] * shp[1], M))
        P = np.stack([np.ones_like(x) if f is None else np.abs(allv[f])
                      for f in fields[gname]], axis=1)
        S = W @ P
        out.extend(S[:, k].reshape(shp, order="F") for k in range(P.shape[1]))
    return out
