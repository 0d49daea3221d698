"""Seeded synthetic inputs for the marker-in-cell transfer workloads.

Marker layout follows define_markers! (src/Erebus.jl:593-668): markers sit
at sub-cell centres ``dxm/2 + (jm-1)*dxm`` with a uniform jitter of
``(U - 0.5)*dxm`` and are numbered ``m = (jm-1)*Nym + im`` (E:617).  The
random stream is numpy's PCG64 seeded with the workload seed (the
reference's MersenneTwister stream cannot be reproduced without Julia).

Workloads (140 km box, planetesimal centred in sticky air):
  cpu41   41x41 basic nodes, 4x4 markers per cell, R = 50 km (25,600 markers)
  fp64    281x281, 6x6 per cell, R = 50 km (2,822,400 markers)
  adj     201x201, 5x5 per cell, R = 50 km (1,000,000 markers)
  large40 / large65  561x561, 8x8 per cell, R = 40 / 65 km (20,070,400)

Planet markers: porosity U[0.05, 0.3], T U[150, 300] K, wet molar fraction
0 (solid density MD/VD = 3300 via E:721), fluid density 1000.  Air markers:
T = 273 K and porosity phimin = 1e-4 as define_markers! sets for air
(E:647); a porosity of exactly 0 would make etafluidcur_inv_kphim infinite
(E:1274).  Crust is r > 0.96 R (the 48/50 ratio of prod_constants.jl).

The velocity field is the stream function psi = A sin(pi x/L) sin(pi y/L)
sampled at basic nodes, vx = dpsi/dy at Vx nodes and vy = -dpsi/dx at Vy
nodes (discretely divergence free), with A set for a Courant number of 0.5
at dt = 1e11 s.
"""
from __future__ import annotations

import numpy as np

from .constants import DT_LONGEST, Geometry, Materials

WORKLOADS = {
    "cpu41": dict(N=41, mc=4, R=50_000.0),
    "fp64": dict(N=281, mc=6, R=50_000.0),
    "mixed": dict(N=281, mc=6, R=50_000.0),
    "adj": dict(N=201, mc=5, R=50_000.0),
    "adj41": dict(N=41, mc=5, R=50_000.0),
    "large40": dict(N=561, mc=8, R=40_000.0),
    "large65": dict(N=561, mc=8, R=65_000.0),
}

MARKER_FIELDS = [
    "xm", "ym", "tm", "tkm", "phim", "sxxm", "sxym", "XWsolidm0", "etavpm",
    "inv_gggtotalm", "fricttotalm", "cohestotalm", "tenstotalm", "rhofluidcur",
    "alphasolidcur", "alphafluidcur", "rhototalm", "rhocptotalm", "etatotalm",
    "hrtotalm", "ktotalm", "tkm_rhocptotalm", "etafluidcur_inv_kphim",
]


def workload_geometry(name: str, box: float = 140_000.0,
                      mode: int = 1) -> tuple[Geometry, Materials]:
    w = WORKLOADS[name]
    g = Geometry(Nx=w["N"], Ny=w["N"], xsize=box, ysize=box, Nxmc=w["mc"],
                 Nymc=w["mc"], rplanet=w["R"], rcrust=0.96 * w["R"])
    return g, Materials(marker_property_mode=mode)


def make_markers(g: Geometry, mat: Materials, seed: int = 0,
                 stress: bool = False) -> dict:
    """Marker SoA of a planetesimal workload (see module docstring)."""
    rng = np.random.default_rng(seed)
    Nxm, Nym = g.Nxm, g.Nym
    M = Nxm * Nym
    jm, im = np.meshgrid(np.arange(Nxm), np.arange(Nym), indexing="ij")
    jm = jm.ravel()
    im = im.ravel()  # m = jm*Nym + im  (0-based form of E:617)
    xm = g.dxm / 2 + jm * g.dxm + (rng.random(M) - 0.5) * g.dxm
    ym = g.dym / 2 + im * g.dym + (rng.random(M) - 0.5) * g.dym
    r = np.sqrt((xm - g.xcenter) ** 2 + (ym - g.ycenter) ** 2)
    planet = r < g.rplanet
    tm = np.where(planet, np.where(r > g.rcrust, 2, 1), 3).astype(np.int64)
    t = tm - 1
    tab = lambda name: np.asarray(getattr(mat, name), dtype=np.float64)[t]
    phim = np.where(planet, 0.05 + 0.25 * rng.random(M), mat.phimin)
    tkm = np.where(planet, 150.0 + 150.0 * rng.random(M), 273.0)
    d = {
        "xm": xm, "ym": ym, "tm": tm, "tkm": tkm, "phim": phim,
        "sxxm": np.zeros(M), "sxym": np.zeros(M),
        "XWsolidm0": np.zeros(M),
        "etavpm": tab("etasolidm"),
        "inv_gggtotalm": 1.0 / tab("gggsolidm"),
        "fricttotalm": tab("frictsolidm"),
        "cohestotalm": tab("cohessolidm"),
        "tenstotalm": tab("tenssolidm"),
        "rhofluidcur": tab("rhofluidm"),
        "alphasolidcur": tab("alphasolidm"),
        "alphafluidcur": tab("alphafluidm"),
        # static air properties (E:651-655); rocks are recomputed each step
        "rhototalm": np.where(planet, 0.0, tab("rhosolidm")),
        "rhocptotalm": np.where(planet, 0.0, tab("rhocpsolidm")),
        "etatotalm": np.where(planet, 0.0, tab("etasolidm")),
        "hrtotalm": np.where(planet, 0.0, tab("start_hrsolidm")),
        "ktotalm": np.where(planet, 0.0, tab("ksolidm")),
        "tkm_rhocptotalm": np.zeros(M),
        "etafluidcur_inv_kphim": np.zeros(M),
    }
    if stress:
        d["sxxm"] = 1.0e6 * (rng.random(M) - 0.5)
        d["sxym"] = 1.0e6 * (rng.random(M) - 0.5)
    for k, v in d.items():
        d[k] = np.ascontiguousarray(v)
    return d


def _F(shape):
    return np.zeros(shape, dtype=np.float64, order="F")


def stream_velocity(g: Geometry, courant: float = 0.5,
                    dt: float = DT_LONGEST) -> tuple[np.ndarray, np.ndarray]:
    """vx (Vx nodes) and vy (Vy nodes) from psi = A sin(pi x/L) sin(pi y/L).
    Ghost rows/columns copy their neighbour (free slip)."""
    a = g.axes()
    X, Y = np.meshgrid(a["x"], a["y"])           # basic nodes (Ny, Nx)
    psi = np.sin(np.pi * X / g.xsize) * np.sin(np.pi * Y / g.ysize)
    vx = _F(g.stag_shape)
    vy = _F(g.stag_shape)
    # Vx node (i, j) (1-based) lies between basic (i-1, j) and (i, j)
    vx[1:g.Ny, 0:g.Nx] = (psi[1:, :] - psi[:-1, :]) / g.dy
    vx[0, :] = vx[1, :]
    vx[g.Ny, :] = vx[g.Ny - 1, :]
    # Vy node (i, j) lies between basic (i, j-1) and (i, j)
    vy[0:g.Ny, 1:g.Nx] = -(psi[:, 1:] - psi[:, :-1]) / g.dx
    vy[:, 0] = vy[:, 1]
    vy[:, g.Nx] = vy[:, g.Nx - 1]
    vmax = max(np.abs(vx).max(), np.abs(vy).max())
    A = courant * min(g.dx, g.dy) / (vmax * dt)
    return np.asfortranarray(vx * A), np.asfortranarray(vy * A)


def rotation_rate(g: Geometry, vx: np.ndarray, vy: np.ndarray) -> np.ndarray:
    """compute_rotation_rate! (E:4946-4956) at basic nodes."""
    Nx, Ny = g.Nx, g.Ny
    w = 0.5 * ((vy[:Ny, 1:Nx + 1] - vy[:Ny, :Nx]) * (1.0 / g.dx)
               - (vx[1:Ny + 1, :Nx] - vx[:Ny, :Nx]) * (1.0 / g.dy))
    return np.asfortranarray(w)


def bulk_viscosity(ETA: np.ndarray, PHI: np.ndarray, etaphikoef: float = 1.0,
                   ETAP: np.ndarray | None = None):
    """recompute_bulk_viscosity! (E:3634-3647): ETAP at interior P nodes
    from the four surrounding basic-node viscosities, ETAPHI = ETAP/PHI."""
    if ETAP is None:
        ETAP = np.full(PHI.shape, 1.0e16, order="F")
    ETAP = np.array(ETAP, order="F", copy=True)
    ETAP[1:-1, 1:-1] = 4.0 / (1.0 / ETA[:-1, :-1] + 1.0 / ETA[1:, :-1]
                              + 1.0 / ETA[:-1, 1:] + 1.0 / ETA[1:, 1:])
    ETAPHI = etaphikoef * ETAP * (1.0 / PHI)
    return ETAP, np.asfortranarray(ETAPHI)


def disk_gravity(g: Geometry, rho: float = 3300.0,
                 G: float = 6.672e-11) -> tuple[np.ndarray, np.ndarray]:
    """Gravity of a uniform infinite cylinder of radius rplanet, at Vx (gx)
    and Vy (gy) nodes; stands in for the host Poisson solve (E:2715)."""
    a = g.axes()

    def field(xs, ys):
        X, Y = np.meshgrid(xs, ys)
        dxc, dyc = X - g.xcenter, Y - g.ycenter
        r = np.hypot(dxc, dyc)
        R = g.rplanet
        mag = np.where(r < R, 2 * np.pi * G * rho * r,
                       2 * np.pi * G * rho * R * R / np.maximum(r, 1e-30))
        with np.errstate(invalid="ignore", divide="ignore"):
            ux = np.where(r > 0, dxc / r, 0.0)
            uy = np.where(r > 0, dyc / r, 0.0)
        return -mag * ux, -mag * uy

    gx, _ = field(a["xvx"], a["yvx"])
    _, gy = field(a["xvy"], a["yvy"])
    return np.asfortranarray(gx), np.asfortranarray(gy)


def smooth_field(g: Geometry, shape, amp: float, kx: int, ky: int,
                 phase: float = 0.0) -> np.ndarray:
    ny, nx = shape
    X, Y = np.meshgrid(np.arange(nx) / max(nx - 1, 1), np.arange(ny) / max(ny - 1, 1))
    return np.asfortranarray(amp * np.sin(np.pi * kx * X + phase) * np.cos(np.pi * ky * Y))
