"""GPU: geometries whose axis nodes are not all first node + k*spacing
rounded once (a box of 10 km / 3 per cell).  The kernels then read the node
coordinates from the axis arrays (the workloads' exact axes let them compute
them instead); the bars are those of the small parity tests: gathers
bit-exact, RK4 within 1e-14 of |x| + dt*max|v|, scatter sums within 1e-12 of
the per-node sum of |terms|.
"""
from __future__ import annotations

from fractions import Fraction

import numpy as np
import pytest

import cases
import ereb
from ereb import api
from ereb.constants import Geometry, Materials
from oracle import Oracle

pytestmark = pytest.mark.gpu

START_TIME = 2.25e6 * 365.25 * 24 * 3600  # C:339


def _geometry():
    g = Geometry(Nx=31, Ny=25, xsize=1.0e5, ysize=8.0e4, Nxmc=4, Nymc=4,
                 rplanet=3.0e4, rcrust=2.88e4)
    return g, Materials(marker_property_mode=1)


def _fma_exact(a, h):
    """every node equals fl(a[0] + k*h) rounded once (what the kernels compute)"""
    return all(float(Fraction(k) * Fraction(h) + Fraction(float(a[0]))) == a[k]
               for k in range(len(a)))


def test_geometry_has_inexact_axes():
    g, _ = _geometry()
    ax = g.axes()
    assert not all(_fma_exact(ax[n], h) for n, h in
                   (("x", g.dx), ("y", g.dy), ("xp", g.dx), ("yp", g.dy)))


def test_gathers_bit_exact_with_loaded_axes():
    g, mat = _geometry()
    mk = cases.ref_markers(g, mat, seed=31)
    rng = np.random.default_rng(32)
    S, B = (g.Ny1, g.Nx1), (g.Ny, g.Nx)
    DT = cases.rand_grid(rng, S, -5, 5)
    tk2 = cases.rand_grid(rng, S, 263, 310)
    DSXX = cases.rand_grid(rng, S, -1e5, 1e5)
    DSXY = cases.rand_grid(rng, B, -1e5, 1e5)
    orc = Oracle(g, mat)
    ref_t = orc.update_marker_temperature(mk["xm"], mk["ym"], mk["tkm"], DT, tk2, 2)
    ref_s = orc.update_marker_stress(mk["xm"], mk["ym"], mk["sxxm"], mk["sxym"], DSXX, DSXY)
    tkm, sxx, sxy = mk["tkm"].copy(), mk["sxxm"].copy(), mk["sxym"].copy()
    with ereb.Context(g, mat, ereb.FP64) as ctx:
        api.update_marker_temperature_(ctx, mk["xm"], mk["ym"], tkm, DT, tk2, 2, tkm.size)
        api.update_marker_stress_(ctx, mk["xm"], mk["ym"], sxx, sxy, DSXX, DSXY, sxx.size)
    np.testing.assert_array_equal(tkm, ref_t)
    np.testing.assert_array_equal(sxx, ref_s[0])
    np.testing.assert_array_equal(sxy, ref_s[1])


def test_rk4_and_scatter_with_loaded_axes():
    g, mat = _geometry()
    mk = cases.ref_markers(g, mat, seed=33)
    rng = np.random.default_rng(34)
    S = (g.Ny1, g.Nx1)
    vx, vy, vxf, vyf = (cases.rand_grid(rng, S, -1, 1) for _ in range(4))
    wyx = cases.rand_grid(rng, (g.Ny, g.Nx), -1e-12, 1e-12)
    tk2 = cases.rand_grid(rng, S, 263, 310)
    dt = 0.9 * g.dx
    orc = Oracle(g, mat)
    ref = orc.move_markers_rk4(mk["xm"], mk["ym"], mk["tm"], mk["tkm"], mk["phim"],
                               mk["sxym"], mk["sxxm"], vx, vy, vxf, vyf, wyx, tk2, dt, 1)
    xm, ym, tkm = mk["xm"].copy(), mk["ym"].copy(), mk["tkm"].copy()
    sxy, sxx = mk["sxym"].copy(), mk["sxxm"].copy()
    hrs, hrf = mat.radioactive_heating(START_TIME)
    props_o, sums_o = orc.marker_to_nodes(mk, hrs, hrf, 1)
    with ereb.Context(g, mat, ereb.FP64) as ctx:
        ctx.set_markers(mk)
        ctx.step_scatter(hrs, hrf, 1)
        sums = [ctx.get_grid(n) for n in ereb.SUM_FIELDS]
    with ereb.Context(g, mat, ereb.FP64) as ctx:
        api.move_markers_rk4_(ctx, xm, ym, mk["tm"], tkm, mk["phim"], sxy, sxx, vx, vy, vxf,
                              vyf, wyx, tk2, xm.size, dt, 1)
    scale = dt * max(np.abs(v).max() for v in (vx, vy, vxf, vyf))
    for name, a, b in (("xm", xm, ref[0]), ("ym", ym, ref[1])):
        assert np.all(np.abs(a - b) <= 1e-14 * (np.abs(b) + scale)), name
    np.testing.assert_allclose(tkm, ref[2], rtol=1e-13, atol=0)
    absb = cases.abs_sums(g, mk, props_o)
    for n, a, b, c in zip(ereb.SUM_FIELDS, sums, sums_o, absb):
        assert np.all(np.abs(a - b) <= 1e-12 * c + 1e-300), n
