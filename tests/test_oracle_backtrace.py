"""Pins of the oracle's backtrace_pressures_rk4! (E:5284-5496).

  (a) a port of the reference test's restatement T:6112-6394 (from
      madcph.m: MATLAB operation orders, `dtm/2*v`, `1/6*(...)`, weights
      (1-dxmj/dx)(1-dymi/dy)), in the reference test's regime (dtm = 0.9 s,
      |v| <= 1e-9 m/s: sub-nanometre paths) and with paths of a third of a
      cell (dt = 3e11 s), at rtol 1e-12 (the reference test: 1e-9);
  (b) zero velocity: every pressure is returned unchanged, bit for bit;
  (c) uniform velocity (a, b) and affine pressures: the second differences
      vanish, every stage moves by exactly (a, b) and the bilinear weights
      reproduce the affine field, so an interior node takes
      p(xA - dt a, yA - dt b) (rtol 1e-12), boundary nodes keep p.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from ereb.constants import test_constants as tconsts
from oracle import Oracle


@pytest.fixture(scope="module")
def tc():
    g, mat = tconsts()
    return g, mat, Oracle(g, mat)


def _trace_ver(g, vx, vy, xA, yA, dtm):
    """T:6142-6249: the MATLAB-form RK4 back-trace of one node"""
    a = g.axes()
    Nx, Ny, dx, dy = g.Nx, g.Ny, g.dx, g.dy
    xcur, ycur = xA, yA
    vxm, vym = [0.0] * 4, [0.0] * 4
    for rk in range(4):
        j = int(math.trunc((xcur - a["xvx"][0]) / dx)) + 1
        i = int(math.trunc((ycur - a["yvx"][0]) / dy)) + 1
        j = min(max(j, 1), Nx - 1)
        i = min(max(i, 1), Ny)
        dxmj = xcur - a["xvx"][j - 1]
        dymi = ycur - a["yvx"][i - 1]
        V = lambda ii, jj: vx[ii - 1, jj - 1]  # noqa: E731
        v13 = V(i, j) * (1 - dxmj / dx) + V(i, j + 1) * dxmj / dx
        v24 = V(i + 1, j) * (1 - dxmj / dx) + V(i + 1, j + 1) * dxmj / dx
        if dxmj / dx >= 0.5:
            if j < Nx - 1:
                v13 = v13 + 1 / 2 * ((dxmj / dx - 0.5) ** 2) * (V(i, j) - 2 * V(i, j + 1) + V(i, j + 2))
                v24 = v24 + 1 / 2 * ((dxmj / dx - 0.5) ** 2) * (V(i + 1, j) - 2 * V(i + 1, j + 1) + V(i + 1, j + 2))
        elif j > 1:
            v13 = v13 + 1 / 2 * ((dxmj / dx - 0.5) ** 2) * (V(i, j - 1) - 2 * V(i, j) + V(i, j + 1))
            v24 = v24 + 1 / 2 * ((dxmj / dx - 0.5) ** 2) * (V(i + 1, j - 1) - 2 * V(i + 1, j) + V(i + 1, j + 1))
        vxm[rk] = (1 - dymi / dy) * v13 + (dymi / dy) * v24
        j = int(math.trunc((xcur - a["xvy"][0]) / dx)) + 1
        i = int(math.trunc((ycur - a["yvy"][0]) / dy)) + 1
        j = min(max(j, 1), Nx)
        i = min(max(i, 1), Ny - 1)
        dxmj = xcur - a["xvy"][j - 1]
        dymi = ycur - a["yvy"][i - 1]
        W = lambda ii, jj: vy[ii - 1, jj - 1]  # noqa: E731
        v12 = W(i, j) * (1 - dymi / dy) + W(i + 1, j) * dymi / dy
        v34 = W(i, j + 1) * (1 - dymi / dy) + W(i + 1, j + 1) * dymi / dy
        if dymi / dy >= 0.5:
            if i < Ny - 1:
                v12 = v12 + 1 / 2 * ((dymi / dy - 0.5) ** 2) * (W(i, j) - 2 * W(i + 1, j) + W(i + 2, j))
                v34 = v34 + 1 / 2 * ((dymi / dy - 0.5) ** 2) * (W(i, j + 1) - 2 * W(i + 1, j + 1) + W(i + 2, j + 1))
        elif i > 1:
            v12 = v12 + 1 / 2 * ((dymi / dy - 0.5) ** 2) * (W(i - 1, j) - 2 * W(i, j) + W(i + 1, j))
            v34 = v34 + 1 / 2 * ((dymi / dy - 0.5) ** 2) * (W(i - 1, j + 1) - 2 * W(i, j + 1) + W(i + 1, j + 1))
        vym[rk] = (1 - dxmj / dx) * v12 + (dxmj / dx) * v34
        if rk in (0, 1):
            xcur = xA - dtm / 2 * vxm[rk]
            ycur = yA - dtm / 2 * vym[rk]
        elif rk == 2:
            xcur = xA - dtm * vxm[rk]
            ycur = yA - dtm * vym[rk]
    vxeff = 1 / 6 * (vxm[0] + 2 * vxm[1] + 2 * vxm[2] + vxm[3])
    vyeff = 1 / 6 * (vym[0] + 2 * vym[1] + 2 * vym[2] + vym[3])
    return xA - dtm * vxeff, yA - dtm * vyeff


def _interp_ver(g, p, x, y):
    """T:6251-6281: P-grid cell, MATLAB weights, dot of the four nodes"""
    a = g.axes()
    j = min(max(int(math.trunc((x - a["xp"][0]) / g.dx)) + 1, 1), g.Nx)
    i = min(max(int(math.trunc((y - a["yp"][0]) / g.dy)) + 1, 1), g.Ny)
    dxmj = x - a["xp"][j - 1]
    dymi = y - a["yp"][i - 1]
    w = ((1 - dxmj / g.dx) * (1 - dymi / g.dy), (1 - dxmj / g.dx) * (dymi / g.dy),
         (dxmj / g.dx) * (1 - dymi / g.dy), (dxmj / g.dx) * (dymi / g.dy))
    return (p[i - 1, j - 1] * w[0] + p[i, j - 1] * w[1] + p[i - 1, j] * w[2] + p[i, j] * w[3])


def _fields(g, seed):
    rng = np.random.default_rng(seed)
    S = (g.Ny1, g.Nx1)
    v = [np.asfortranarray(rng.random(S) * 2e-9 - 1e-9) for _ in range(4)]   # T:6113-6116
    p = [np.asfortranarray(rng.random(S) * 1e6) for _ in range(3)]           # T:6117-6122
    return v, p


@pytest.mark.parametrize("dtm", [0.9, 3e11])
def test_backtrace_restatement(tc, dtm):
    g, _, orc = tc
    (vx, vy, vxf, vyf), (pr, pf, ps) = _fields(g, 17)
    pr0, ps0, pf0 = orc.backtrace_pressures(pr, ps, pf, vx, vy, vxf, vyf, dtm)
    a = g.axes()
    moved = 0.0
    for jj in range(1, g.Nx1 + 1):
        for ii in range(1, g.Ny1 + 1):
            if 2 <= ii <= g.Ny and 2 <= jj <= g.Nx:
                xA, yA = a["xp"][jj - 1], a["yp"][ii - 1]
                x, y = _trace_ver(g, vx, vy, xA, yA, dtm)
                moved = max(moved, abs(x - xA), abs(y - yA))
                er = (_interp_ver(g, pr, x, y), _interp_ver(g, ps, x, y))
                x, y = _trace_ver(g, vxf, vyf, xA, yA, dtm)
                ef = _interp_ver(g, pf, x, y)
                for got, want in ((pr0, er[0]), (ps0, er[1]), (pf0, ef)):
                    assert got[ii - 1, jj - 1] == pytest.approx(want, rel=1e-12)
            else:
                assert pr0[ii - 1, jj - 1] == pr[ii - 1, jj - 1]
                assert ps0[ii - 1, jj - 1] == ps[ii - 1, jj - 1]
                assert pf0[ii - 1, jj - 1] == pf[ii - 1, jj - 1]
    if dtm > 1:
        assert moved > 0.1 * g.dx


def test_backtrace_zero_velocity(tc):
    g, _, orc = tc
    _, (pr, pf, ps) = _fields(g, 18)
    z = np.zeros((g.Ny1, g.Nx1), order="F")
    pr0, ps0, pf0 = orc.backtrace_pressures(pr, ps, pf, z, z, z, z, 3e11)
    assert np.array_equal(pr0, pr) and np.array_equal(ps0, ps) and np.array_equal(pf0, pf)


def test_backtrace_uniform_velocity_affine_pressure(tc):
    g, _, orc = tc
    S = (g.Ny1, g.Nx1)
    a = g.axes()
    X, Y = np.meshgrid(a["xp"], a["yp"])
    aff = lambda x, y, c: c[0] + c[1] * x + c[2] * y  # noqa: E731
    cs = [(2e5, 3.0, -7.0), (1e6, -0.5, 2.5), (-3e5, 11.0, 1.0)]
    pr, ps, pf = (np.asfortranarray(aff(X, Y, c)) for c in cs)
    va, vb, wa, wb = 1.3e-9, -0.7e-9, -2.1e-9, 0.4e-9
    dt = 3e11
    vx, vy = np.full(S, va, order="F"), np.full(S, vb, order="F")
    vxf, vyf = np.full(S, wa, order="F"), np.full(S, wb, order="F")
    pr0, ps0, pf0 = orc.backtrace_pressures(pr, ps, pf, vx, vy, vxf, vyf, dt)
    inner = np.zeros(S, bool)
    inner[1:g.Ny, 1:g.Nx] = True
    for got, p, c, (u, v) in ((pr0, pr, cs[0], (va, vb)), (ps0, ps, cs[1], (va, vb)),
                              (pf0, pf, cs[2], (wa, wb))):
        want = aff(X - dt * u, Y - dt * v, c)
        scale = np.abs(p).max()
        assert np.all(np.abs(got - want)[inner] <= 1e-12 * scale)
        assert np.array_equal(got[~inner], p[~inner])
