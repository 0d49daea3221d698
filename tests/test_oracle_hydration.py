"""Pins of the oracle's viscosity, compaction and (de)hydration transfers
(update_marker_viscosity! E:774, update_marker_porosity! E:4822,
update_p_nodes_melt_composition! E:1929, perform_thermochemical_reaction!
E:2558 and its helpers E:2382-2524).

Each is checked against a port of the reference test's own restatement
(T: = test/runtests.jl, restated there from Gerya's HTM-planetary.m and
HTM-hydration.m with MATLAB operation orders: (1-a)(1-b) weights, /dx,
8.314 as a literal), at the reference test's tolerance or tighter, and
against closed forms (partition of unity, constant fields, conservation of
the reacting-marker weight).  The reference's random streams
(MersenneTwister(42)) cannot be reproduced here; the inputs follow the same
distributions from seeded numpy generators.

  oracle function                          pins
  orc_update_marker_viscosity              T:841-909 restatement (rtol 1e-12; T: 1e-9),
                                           etatotalm exact, all-yielding / none-yielding
  orc_update_marker_porosity               T:5602-5653 restatement (rtol 1e-12; T: 1e-9)
                                           with exp() saturated (as T) and unsaturated
  orc_update_p_nodes_melt_composition      T:3356-3421 restatement, constant XW, sum WT = M
  orc_dtreaction                           T:3481-3498 (all four modes, mode 3 in the
                                           test's own form), unknown mode
  orc_gibbs_free_energy / relative_enthalpy
  / reaction_constant                      T:1062-1125
  orc_perform_thermochemical_reaction      T:3512-3731 restatement, both timestep/titer
                                           branches, weight conservation; see below

In the reference test's regime (rate mode 1, dt = 0.9 Δtreaction << the
rate-mode-1 reaction time) the reaction barely proceeds, RV = 1 - O(1e-11)
and ΔM = (1 - RV)/dt keeps only ~5 significant digits: one ulp of KWD (the
restatement divides by 8.314 then by T, the function by RG*T) moves ΔM by
~1e-5 relative.  The reference compares DMP and DHP normwise (Julia's ≈ on
arrays) at 1e-6; so does this test, with per-element 1e-9 on the
well-conditioned outputs (XWsolidm, phinewm, pfm0, WTPSUM).  A second case
(rate mode 9, dt = 1.5 Δtreaction: full equilibrium, ΔG = 0) is well
conditioned and is compared per element at 1e-9.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from ereb.constants import Hydration, test_constants as tconsts
from oracle import Oracle

DT_LONGEST = 1.0e11  # C:319


@pytest.fixture(scope="module")
def tc():
    g, mat = tconsts()
    return g, mat, Oracle(g, mat)


def _markers(g, rng, n):
    """xm, ym uniform in [-dx, x[end]+dx] as T:843-844"""
    xm = -g.dx + rng.random(n) * (g.xsize + 2 * g.dx)
    ym = -g.dy + rng.random(n) * (g.ysize + 2 * g.dy)
    return xm, ym


def _cell(g, x, y, grid):
    """T:861-884 / T:3388-3401: MATLAB-form cell, distances and weights"""
    a = g.axes()
    if grid == "basic":
        ax, ay, jmax, imax = a["x"], a["y"], g.Nx - 1, g.Ny - 1
    else:
        ax, ay, jmax, imax = a["xp"], a["yp"], g.Nx, g.Ny
    j = int(math.trunc((x - ax[0]) / g.dx)) + 1
    i = int(math.trunc((y - ay[0]) / g.dy)) + 1
    j = min(max(j, 1), jmax)
    i = min(max(i, 1), imax)
    dxmj = x - ax[j - 1]
    dymi = y - ay[i - 1]
    w = ((1 - dxmj / g.dx) * (1 - dymi / g.dy), (1 - dxmj / g.dx) * (dymi / g.dy),
         (dxmj / g.dx) * (1 - dymi / g.dy), (dxmj / g.dx) * (dymi / g.dy))
    return i - 1, j - 1, w  # 0-based upper-left node


# ------------------------------------------------------------------ viscosity

def _viscosity_ver(g, mat, xm, ym, tm, tkm, ETA, YNY):
    """T:859-907"""
    n = xm.size
    out = np.zeros(n)
    for m in range(n):
        i, j, w = _cell(g, xm[m], ym[m], "basic")
        t = tm[m] - 1
        if tm[m] < 3:
            es = mat.etasolidmm[t] if tkm[m] > mat.tmsolidphase else mat.etasolidm[t]
            eta = es
        else:
            eta = mat.etasolidm[t]
        if YNY[i, j] or YNY[i + 1, j] or YNY[i, j + 1] or YNY[i + 1, j + 1]:
            v = 1 / (YNY[i, j] / ETA[i, j] * w[0] + YNY[i + 1, j] / ETA[i + 1, j] * w[1] +
                     YNY[i, j + 1] / ETA[i, j + 1] * w[2] +
                     YNY[i + 1, j + 1] / ETA[i + 1, j + 1] * w[3])
            out[m] = eta if v >= eta else v
        else:
            out[m] = eta
    return out


@pytest.mark.parametrize("yny", ["random", "all", "none"])
def test_update_marker_viscosity(tc, yny):
    g, mat, orc = tc
    rng = np.random.default_rng({"random": 1, "all": 2, "none": 3}[yny])
    n = 4000
    xm, ym = _markers(g, rng, n)
    tm = rng.integers(1, 4, n).astype(np.int64)
    tkm = mat.tmfluidphase - 100 + rng.random(n) * (mat.tmsolidphase - mat.tmfluidphase + 200)
    etatotalm = np.array([mat.etasolidm[t - 1] for t in tm])
    # ETA scaled to viscosities so that both branches of the cap occur
    ETA = np.asfortranarray(10.0 ** (14 + 7 * rng.random((g.Ny, g.Nx))))
    YNY = {"random": rng.random((g.Ny, g.Nx)) < 0.5,
           "all": np.ones((g.Ny, g.Nx), bool), "none": np.zeros((g.Ny, g.Nx), bool)}[yny]
    YNY = np.asfortranarray(YNY)
    YIE = np.asfortranarray(YNY / ETA)
    et, ev = orc.update_marker_viscosity(xm, ym, tm, tkm, etatotalm, YNY, YIE)
    ver = _viscosity_ver(g, mat, xm, ym, tm, tkm, ETA, YNY)
    np.testing.assert_allclose(ev, ver, rtol=1e-12)
    for m in range(n):  # etatotal_rocks for rocks (E:786), unchanged air
        if tm[m] < 3:
            t = tm[m] - 1
            es = mat.etasolidmm[t] if tkm[m] > mat.tmsolidphase else mat.etasolidm[t]
            ef = mat.etafluidmm[t] if tkm[m] > mat.tmfluidphase else mat.etafluidm[t]
            assert et[m] == max(mat.etamin, es, ef)
        else:
            assert et[m] == etatotalm[m]
    if yny == "none":
        assert np.array_equal(ev, et)
    else:
        assert np.any(ev < et) and np.any(ev == et)


# ------------------------------------------------------------------- porosity

@pytest.mark.parametrize("scale", ["reference", "unsaturated"])
def test_update_marker_porosity(tc, scale):
    g, mat, orc = tc
    rng = np.random.default_rng(11)
    n = 4000
    xm, ym = _markers(g, rng, n)
    tm = rng.integers(1, 4, n).astype(np.int64)
    # the reference test's APHI = rand saturates exp(aphim*dt) at dt_longest;
    # a second case keeps aphim*dt in [-3, 3]
    APHI = rng.random((g.Ny1, g.Nx1))
    if scale == "unsaturated":
        APHI = (APHI - 0.5) * 6.0 / DT_LONGEST
    APHI = np.asfortranarray(APHI)
    phim = rng.random(n)
    H = Hydration()
    out = orc.update_marker_porosity(xm, ym, tm, phim, APHI, DT_LONGEST)
    ver = phim.copy()
    for m in range(n):  # T:5616-5649
        if tm[m] < 3:
            i, j, w = _cell(g, xm[m], ym[m], "p")
            aphim = (APHI[i, j] * w[0] + APHI[i + 1, j] * w[1] + APHI[i, j + 1] * w[2] +
                     APHI[i + 1, j + 1] * w[3])
            with np.errstate(over="ignore"):
                v = ver[m] / ((1 - ver[m]) * np.exp(aphim * DT_LONGEST) + ver[m])
            ver[m] = H.phimin if v < H.phimin else (H.phimax if v > H.phimax else v)
    np.testing.assert_allclose(out, ver, rtol=1e-12)
    assert np.array_equal(out[tm == 3], phim[tm == 3])
    if scale == "unsaturated":
        inner = (out > H.phimin) & (out < H.phimax) & (tm < 3)
        assert inner.sum() > n // 4


# ---------------------------------------------------------- melt composition

def test_melt_composition(tc):
    g, mat, orc = tc
    rng = np.random.default_rng(5)
    n = 6000
    a = g.axes()
    xm = -a["xp"][0] + rng.random(n) * (a["xp"][-1] + g.dx + a["xp"][0])  # T:3368
    ym = -a["yp"][0] + rng.random(n) * (a["yp"][-1] + g.dy + a["yp"][0])
    XW = rng.random(n)
    XWS, XWSSUM, WTPSUM = orc.melt_composition(xm, ym, XW)
    sv = np.zeros((g.Ny1, g.Nx1))
    wv = np.zeros((g.Ny1, g.Nx1))
    for m in range(n):  # T:3373-3408
        i, j, w = _cell(g, xm[m], ym[m], "p")
        for (di, dj), wk in zip(((0, 0), (1, 0), (0, 1), (1, 1)), w):
            sv[i + di, j + dj] += XW[m] * wk
            wv[i + di, j + dj] += wk
    ver = np.where(wv > 0, sv / np.where(wv > 0, wv, 1), 0.0)
    np.testing.assert_allclose(XWS, ver, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(WTPSUM, wv, rtol=1e-12, atol=1e-13)
    # partition of unity of the weights
    assert abs(WTPSUM.sum() - n) <= 1e-10 * n
    # a constant composition is reproduced wherever a weight is positive
    XWS_c, _, W_c = orc.melt_composition(xm, ym, np.full(n, 0.375))
    pos = W_c > 1e-3
    np.testing.assert_allclose(XWS_c[pos], 0.375, rtol=1e-13)
    assert np.all(XWS_c[W_c <= 0] == 0)


# ------------------------------------------------- reaction helper functions

def test_dtreaction(tc):
    g, mat, orc = tc
    H = Hydration()
    lcr = math.log(0.01)
    phi = 0.3177
    for T in [170.0, 263.0, 270.0, 273.0, 276.0, 283.0, 370.0, 470.0]:  # T:3482
        assert orc.dtreaction(T, phi, 1) == pytest.approx(
            -lcr / (H.A_I * phi) * math.exp(H.b_I * (T - H.c_I) ** 2), rel=1e-12)
        assert orc.dtreaction(T, phi, 2) == pytest.approx(
            -lcr / (H.Sxo_B * phi) * 2.0 ** (-(T - H.To_B) / H.Tscl_B), rel=1e-12)
        assert orc.dtreaction(T, phi, 3) == pytest.approx(
            -lcr / (H.Sxo_T * phi * math.exp(H.Ea_T / H.RG * (1 / H.To_T - 1 / T))), rel=1e-12)
        assert orc.dtreaction(T, phi, 9) == H.dtreaction
    with pytest.raises(ValueError):
        orc.dtreaction(300.0, phi, 4)


def test_gibbs_enthalpy_reaction_constant(tc):
    g, mat, orc = tc
    H = Hydration()
    hc = H.to_c(mat)
    rng = np.random.default_rng(9)
    for _ in range(50):  # T:1062-1125
        phi, XW0 = rng.random(2)
        tknm, pfnm, XD0 = rng.random(3)
        dtr = orc.dtreaction(tknm, phi, H.rate_mode)
        for dt in (0.5 * dtr, 1.5 * dtr):
            g0 = (hc.dHWD - tknm * hc.dSWD + hc.dVWD * pfnm +
                  8.314 * tknm * math.log(XD0 / XW0))
            ver = g0 * (1 - dt / dtr) if dt < dtr else 0.0
            assert orc.gibbs_free_energy(tknm, pfnm, XD0, XW0, dt, dtr) == pytest.approx(
                ver, rel=1e-12, abs=1e-300)
        Xs0 = rng.random()
        assert orc.relative_enthalpy(Xs0, XW0) == pytest.approx(
            -Xs0 * XW0 * hc.dHWD / (mat.MD + mat.MH2O), rel=1e-14)
        dG = rng.random()
        assert orc.reaction_constant(tknm, pfnm, dG) == pytest.approx(
            math.exp(-(hc.dHWD - tknm * hc.dSWD + hc.dVWD * pfnm - dG) / 8.314 / tknm),
            rel=1e-12)
    # derived constants (C:250-258)
    assert hc.VH2Of == mat.MH2O / mat.rhoH2Of and hc.VH2Ofi == mat.MH2O / mat.rhoH2Ofi
    assert hc.dVWD == mat.VDs + hc.VH2Of - mat.VWs
    assert hc.phimax == 1.0 - hc.phimin


# ------------------------------------------------- thermochemical reaction

def _reaction_inputs(g, rng, n):
    xm, ym = _markers(g, rng, n)
    return {
        "tm": rng.integers(1, 4, n).astype(np.int64), "xm": xm, "ym": ym,
        "XWsolidm0": rng.random(n), "XWsolidm": rng.random(n),
        "phim": rng.random(n), "phinewm": rng.random(n),
        "pfm0": 2700 + 600 * rng.random(n),
    }


def _reaction_ver(g, mat, mk, pf, tk2, dt, timestep, titer, mode=1):
    """T:3566-3720 (HTM-hydration.m form)"""
    H = Hydration()
    hc = H.to_c(mat)
    mk = {k: np.array(v, copy=True) for k, v in mk.items()}
    S = (g.Ny1, g.Nx1)
    DMS, DHS, WTS = np.zeros(S), np.zeros(S), np.zeros(S)
    VDs, VWs, MD, MH2O = mat.VDs, mat.VWs, mat.MD, mat.MH2O
    for m in range(mk["xm"].size):
        if not mk["tm"][m] < 3:
            continue
        i, j, w = _cell(g, mk["xm"][m], mk["ym"][m], "p")
        pfnm = pf[i, j] * w[0] + pf[i + 1, j] * w[1] + pf[i, j + 1] * w[2] + pf[i + 1, j + 1] * w[3]
        tknm = tk2[i, j] * w[0] + tk2[i + 1, j] * w[1] + tk2[i, j + 1] * w[2] + tk2[i + 1, j + 1] * w[3]
        pfnm = max(pfnm, 0)
        if titer > 2:
            pfnm = pfnm * (1 - hc.pfcoeff) + mk["pfm0"][m] * hc.pfcoeff
        mk["pfm0"][m] = pfnm
        VH2O = hc.VH2Of if tknm > mat.tmfluidphase else hc.VH2Ofi
        XW0, ph = mk["XWsolidm0"][m], mk["phim"][m]
        XD0 = 1 - XW0
        Xf0 = ph * (XW0 * VWs + XD0 * VDs) / ((1 - ph) * VH2O + ph * (XW0 * VWs + XD0 * VDs))
        Xs0 = 1 - Xf0
        XH2Ot = (XW0 * Xs0 + Xf0) / (1 + XW0 * Xs0)
        XDt = 1 - XH2Ot
        rhos0 = (MD + MH2O * XW0) / (VWs * XW0 + VDs * XD0)
        rhof = mat.rhoH2Of if tknm > mat.tmfluidphase else mat.rhoH2Ofi
        Ht0 = -Xs0 * XW0 * hc.dHWD / (MD + MH2O)
        dG0 = hc.dHWD - tknm * hc.dSWD + hc.dVWD * pfnm + 8.314 * tknm * math.log(XD0 / XW0)
        dtr = (-hc.log_completion_rate / (H.A_I * ph) * math.exp(H.b_I * (tknm - H.c_I) ** 2)
               if mode == 1 else H.dtreaction)
        dG = dG0 * (1 - dt / dtr) if dt < dtr else 0
        KWD = math.exp(-(hc.dHWD - tknm * hc.dSWD + hc.dVWD * pfnm - dG) / 8.314 / tknm)
        XW1 = 1 / (KWD + 1)
        XD1 = 1 - XW1
        Xs1 = XDt / (1 - XDt * XW1)
        Xf1 = 1 - Xs1
        if 0 < Xf1 < 1:
            phi1 = Xf1 * VH2O / (Xf1 * VH2O + Xs1 * (XW1 * VWs + XD1 * VDs))
            rhos1 = (MD + MH2O * XW1) / (VWs * XW1 + VDs * XD1)
            Ht1 = -Xs1 * XW1 * hc.dHWD / (MD + MH2O)
            RV = (rhos1 * (1 - phi1) + rhof * phi1) / (rhos0 * (1 - ph) + rhof * ph)
            Gm = (rhos0 * RV * (1 - ph) - rhos1 * (1 - phi1)) / dt
            DMm = (1 - RV) / dt
            DHm = Gm * (Ht1 - Ht0)
            mk["XWsolidm"][m] = XW1
            mk["phinewm"][m] = phi1
            if timestep == 1:
                mk["XWsolidm0"][m] = XW1
                mk["phim"][m] = phi1
            for (di, dj), wk in zip(((0, 0), (1, 0), (0, 1), (1, 1)), w):
                DMS[i + di, j + dj] += DMm * wk
                DHS[i + di, j + dj] += DHm * wk
                WTS[i + di, j + dj] += wk
    pos = WTS > 0
    DMP = np.where(pos, DMS / np.where(pos, WTS, 1), 0.0)
    DHP = np.where(pos, DHS / np.where(pos, WTS, 1), 0.0)
    return mk, DMP, DHP, DMS, DHS, WTS


def _normwise(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(a), np.linalg.norm(b)))


@pytest.mark.parametrize("timestep,titer,mode", [(1, 3, 1), (2, 1, 1), (1, 3, 9), (2, 4, 9)])
def test_thermochemical_reaction(tc, timestep, titer, mode):
    g, mat, orc = tc
    rng = np.random.default_rng(13 + timestep + mode)
    n = 4000
    mk = _reaction_inputs(g, rng, n)
    pf = np.asfortranarray(2700 + 600 * rng.random((g.Ny1, g.Nx1)))   # T:3519
    tk2 = np.asfortranarray(263 + 47 * rng.random((g.Ny1, g.Nx1)))    # T:3520
    H = Hydration(rate_mode=mode)
    dt = (0.9 if mode == 1 else 1.5) * H.dtreaction                    # T:3531
    io, DMP, DHP, DMS, DHS, WTS = orc.thermochemical_reaction(mk, pf, tk2, dt, timestep,
                                                              titer, hyd=H)
    vmk, vDMP, vDHP, vDMS, vDHS, vWTS = _reaction_ver(g, mat, mk, pf, tk2, dt, timestep,
                                                      titer, mode)
    for name in io:
        np.testing.assert_allclose(io[name], vmk[name], rtol=1e-9, err_msg=name)
    np.testing.assert_allclose(WTS, vWTS, rtol=1e-12, atol=1e-12)
    for a, b, name in ((DMP, vDMP, "DMP"), (DHP, vDHP, "DHP"), (DMS, vDMS, "DMPSUM"),
                       (DHS, vDHS, "DHPSUM")):
        if mode == 1:   # ill-conditioned: normwise as T:3721-3722
            assert _normwise(a, b) <= 1e-6, name
        else:
            np.testing.assert_allclose(a, b, rtol=1e-9, atol=1e-9 * np.abs(b).max(),
                                       err_msg=name)
    # the weight sums to the number of reacting rock markers
    reacting = np.flatnonzero(io["XWsolidm"] != mk["XWsolidm"])
    assert reacting.size > n // 10
    assert abs(WTS.sum() - reacting.size) <= 1e-10 * n
    # air untouched, pfm0 written for every rock marker
    air = mk["tm"] == 3
    for name in io:
        if name != "pfm0":
            assert np.array_equal(io[name][air], mk[name][air]), name
    assert np.all(io["pfm0"][~air] != mk["pfm0"][~air])


def test_thermochemical_unknown_mode(tc):
    g, mat, orc = tc
    rng = np.random.default_rng(3)
    mk = _reaction_inputs(g, rng, 50)
    pf = np.asfortranarray(np.full((g.Ny1, g.Nx1), 3000.0))
    tk2 = np.asfortranarray(np.full((g.Ny1, g.Nx1), 280.0))
    with pytest.raises(ValueError):
        orc.thermochemical_reaction(mk, pf, tk2, 1e9, 1, 1, hyd=Hydration(rate_mode=7))
