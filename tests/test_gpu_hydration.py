"""GPU parity of the viscosity, compaction and (de)hydration transfers
(update_marker_viscosity! E:774, update_marker_porosity! E:4822,
update_p_nodes_melt_composition! E:1929, perform_thermochemical_reaction!
E:2558) against the CPU oracle, through the C ABI.

Bars: marker outputs of the gathers bit-for-bit where no libm call is
involved (viscosity) and within 1e-12 relative otherwise (exp in the
porosity update); P-node sums (summed per fine cell and tile instead of in
marker order) within 1e-12 of the sum of |terms| per node; normalised node
values within 1e-12 relative.  The thermochemical ΔM and ΔH of the
reference test's regime (rate mode 1, dt = 0.9 Δtreaction, T = 263..310 K)
are ill-conditioned (1 - RV ~ 1e-11, see tests/test_oracle_hydration.py): there the
bar is the reference test's own, normwise 1e-6 (T:3721-3722); a
well-conditioned regime (rate mode 9 past the reaction time) is held to
1e-12 of the sum of |terms|.  The mixed form stores etatotalm, etavpm and
XWsolidm0 in FP32: 1e-6 relative.
"""
from __future__ import annotations

import numpy as np
import pytest

import cases
import ereb
from ereb import api
from ereb.constants import Hydration, test_constants as tconsts
from oracle import Oracle

pytestmark = pytest.mark.gpu

DT_LONGEST = 1.0e11


def _exact(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    same = (a == b) | (np.isnan(a) & np.isnan(b))
    assert same.all(), f"{what}: {(~same).sum()} of {a.size} differ"


def _markers(g, mat, seed):
    mk = cases.ref_markers(g, mat, seed=seed)
    rng = np.random.default_rng(seed + 100)
    mk["XWsolidm"] = rng.random(mk["xm"].size)
    mk["phinewm"] = rng.random(mk["xm"].size)
    mk["pfm0"] = 2700 + 600 * rng.random(mk["xm"].size)
    mk["phim"] = rng.random(mk["xm"].size)
    return mk


@pytest.mark.parametrize("precision", [ereb.FP64, ereb.MIXED])
def test_marker_viscosity(precision):
    g, mat = tconsts()
    mk = _markers(g, mat, 21)
    rng = np.random.default_rng(4)
    n = mk["xm"].size
    tkm = mat.tmfluidphase - 100 + rng.random(n) * (mat.tmsolidphase - mat.tmfluidphase + 200)
    eta0 = np.array([mat.etasolidm[t - 1] for t in mk["tm"]])
    ETA = 10.0 ** (14 + 7 * rng.random((g.Ny, g.Nx)))
    YNY = np.asfortranarray(rng.random((g.Ny, g.Nx)) < 0.5)
    YIE = np.asfortranarray(YNY / ETA)
    et_o, ev_o = Oracle(g, mat).update_marker_viscosity(mk["xm"], mk["ym"], mk["tm"], tkm,
                                                         eta0, YNY, YIE)
    et, ev = eta0.copy(), np.zeros(n)
    with ereb.Context(g, mat, precision) as ctx:
        api.update_marker_viscosity_(ctx, mk["xm"], mk["ym"], mk["tm"], tkm, et, ev,
                                     YNY, YIE, n)
    if precision == ereb.FP64:
        _exact(et, et_o, "etatotalm")
        _exact(ev, ev_o, "etavpm")
    else:
        np.testing.assert_allclose(et, et_o, rtol=1e-6)
        np.testing.assert_allclose(ev, ev_o, rtol=1e-6)


@pytest.mark.parametrize("scale", ["reference", "unsaturated"])
def test_marker_porosity(scale):
    g, mat = tconsts()
    mk = _markers(g, mat, 22)
    rng = np.random.default_rng(5)
    APHI = rng.random((g.Ny1, g.Nx1))
    if scale == "unsaturated":
        APHI = (APHI - 0.5) * 6.0 / DT_LONGEST
    APHI = np.asfortranarray(APHI)
    ph_o = Oracle(g, mat).update_marker_porosity(mk["xm"], mk["ym"], mk["tm"], mk["phim"],
                                                 APHI, DT_LONGEST)
    ph = mk["phim"].copy()
    with ereb.Context(g, mat, ereb.FP64) as ctx:
        api.update_marker_porosity_(ctx, mk["xm"], mk["ym"], mk["tm"], ph, APHI, DT_LONGEST,
                                    ph.size)
    np.testing.assert_allclose(ph, ph_o, rtol=1e-12)
    _exact(ph[mk["tm"] == 3], mk["phim"][mk["tm"] == 3], "air porosity")


@pytest.mark.parametrize("precision", [ereb.FP64, ereb.MIXED])
def test_melt_composition(precision):
    g, mat = tconsts()
    mk = _markers(g, mat, 23)
    orc = Oracle(g, mat)
    XWS_o, XS_o, WT_o = orc.melt_composition(mk["xm"], mk["ym"], mk["XWsolidm0"])
    S = (g.Ny1, g.Nx1)
    XWS, XS, WT = (np.zeros(S, order="F") for _ in range(3))
    with ereb.Context(g, mat, precision) as ctx:
        api.update_p_nodes_melt_composition_(ctx, mk["xm"], mk["ym"], mk["XWsolidm0"], XWS,
                                             XS, WT, mk["xm"].size)
    tol = 1e-12 if precision == ereb.FP64 else 1e-6
    # positive terms: the sum of |terms| is the sum itself
    assert np.all(np.abs(WT - WT_o) <= 1e-12 * WT_o + 1e-300)
    assert np.all(np.abs(XS - XS_o) <= tol * XS_o + 1e-300)
    np.testing.assert_allclose(XWS, XWS_o, rtol=2 * tol, atol=0)
    assert abs(WT.sum() - mk["xm"].size) <= 1e-10 * mk["xm"].size


def test_melt_composition_follows_moved_markers():
    """The P-grid scatters reuse the fine-key sort; after an RK4 move (and
    after new positions are set) it must be redone, not reused."""
    g, mat = tconsts()
    mk = _markers(g, mat, 24)
    rng = np.random.default_rng(6)
    # |v| dt up to one cell
    vx = np.asfortranarray(2e-8 * (rng.random((g.Ny1, g.Nx1)) - 0.5))
    vy = np.asfortranarray(2e-8 * (rng.random((g.Ny1, g.Nx1)) - 0.5))
    orc = Oracle(g, mat)
    with ereb.Context(g, mat, ereb.FP64) as ctx:
        ctx.set_markers(mk)
        for n, a in (("vx", vx), ("vy", vy), ("vxf", vx), ("vyf", vy),
                     ("wyx", np.zeros((g.Ny, g.Nx))), ("tk2", np.zeros((g.Ny1, g.Nx1)))):
            ctx.set_grid(n, a)
        ctx.step_melt_composition()                     # sorts
        ctx.step_advect_rk4(DT_LONGEST, 9)              # moves
        ctx.step_melt_composition()                     # must re-sort
        xm, ym = ctx.get_marker("xm"), ctx.get_marker("ym")
        moved = np.abs(xm - mk["xm"]).max()
        XWS = ctx.get_grid("XWS")
        XWS_o = orc.melt_composition(xm, ym, mk["XWsolidm0"])[0]
        np.testing.assert_allclose(XWS, XWS_o, rtol=2e-12)
        # new positions set from the host
        x2 = mk["xm"][::-1].copy()
        ctx.set_marker("xm", x2)
        ctx.step_melt_composition()
        np.testing.assert_allclose(ctx.get_grid("XWS"),
                                   orc.melt_composition(x2, ym, mk["XWsolidm0"])[0], rtol=2e-12)
    assert moved > 0.1 * g.dx


@pytest.mark.parametrize("regime", ["reference", "equilibrium"])
@pytest.mark.parametrize("timestep,titer", [(1, 3), (2, 1)])
def test_thermochemical_reaction(regime, timestep, titer):
    g, mat = tconsts()
    mk = _markers(g, mat, 25 + timestep)
    rng = np.random.default_rng(7 + timestep)
    pf = np.asfortranarray(2700 + 600 * rng.random((g.Ny1, g.Nx1)))
    tk2 = np.asfortranarray(263 + 47 * rng.random((g.Ny1, g.Nx1)))
    H = Hydration(rate_mode=1 if regime == "reference" else 9)
    dt = (0.9 if regime == "reference" else 1.5) * H.dtreaction
    orc = Oracle(g, mat)
    io_o, DMP_o, DHP_o, DMS_o, DHS_o, WTS_o = orc.thermochemical_reaction(
        mk, pf, tk2, dt, timestep, titer, hyd=H)
    names = ["XWsolidm0", "XWsolidm", "phim", "phinewm", "pfm0"]
    io = {k: mk[k].copy() for k in names}
    S = (g.Ny1, g.Nx1)
    DMP, DHP, DMS, DHS, WTS = (np.zeros(S, order="F") for _ in range(5))
    with ereb.Context(g, mat, ereb.FP64) as ctx:
        api.perform_thermochemical_reaction_(ctx, DMP, DHP, DMS, DHS, WTS, pf, tk2, mk["tm"],
                                             mk["xm"], mk["ym"], io["XWsolidm0"],
                                             io["XWsolidm"], io["phim"], io["phinewm"],
                                             io["pfm0"], mk["xm"].size, dt, timestep, titer,
                                             hyd=H)
    _exact(io["pfm0"], io_o["pfm0"], "pfm0")
    for k in names:
        np.testing.assert_allclose(io[k], io_o[k], rtol=1e-12, err_msg=k)
    assert np.all(np.abs(WTS - WTS_o) <= 1e-12 * WTS_o + 1e-300)
    if regime == "reference":
        for a, b, k in ((DMP, DMP_o, "DMP"), (DHP, DHP_o, "DHP"), (DMS, DMS_o, "DMPSUM"),
                        (DHS, DHS_o, "DHPSUM")):
            err = np.linalg.norm(a - b) / np.linalg.norm(b)
            assert err <= 1e-6, (k, err)
    else:
        # |ΔM| <= max|ΔM| per reacting marker, so WTPSUM * max|ΔM| bounds the
        # sum of |terms| at each node
        for a, b, avg, k in ((DMS, DMS_o, DMP_o, "DMPSUM"), (DHS, DHS_o, DHP_o, "DHPSUM")):
            bound = 1e-12 * WTS_o * np.abs(avg).max() + 1e-300
            assert np.all(np.abs(a - b) <= bound), k
        pos = WTS_o > 1e-3
        for a, b, k in ((DMP, DMP_o, "DMP"), (DHP, DHP_o, "DHP")):
            np.testing.assert_allclose(a[pos], b[pos], rtol=1e-9,
                                       atol=1e-12 * np.abs(b).max(), err_msg=k)
            assert np.all(a[WTS_o == 0] == 0)


def test_thermochemical_unknown_mode_and_resident_form():
    g, mat = tconsts()
    mk = _markers(g, mat, 31)
    rng = np.random.default_rng(8)
    pf = 2700 + 600 * rng.random((g.Ny1, g.Nx1))
    tk2 = 263 + 47 * rng.random((g.Ny1, g.Nx1))
    H = Hydration(rate_mode=9)
    dt = 1.5 * H.dtreaction
    orc = Oracle(g, mat)
    io_o, DMP_o, DHP_o, _, _, WTS_o = orc.thermochemical_reaction(mk, pf, tk2, dt, 1, 1, hyd=H)
    with ereb.Context(g, mat, ereb.FP64) as ctx:
        ctx.set_markers(mk)
        ctx.set_grid("pf", pf)
        ctx.set_grid("tk2", tk2)
        before = ctx.get_marker("pfm0")
        with pytest.raises(ereb.EREBModeError):
            ctx.step_thermochemical_reaction(dt, 1, 1, hyd=Hydration(rate_mode=4))
        _exact(ctx.get_marker("pfm0"), before, "pfm0 untouched by a rejected call")
        ctx.step_thermochemical_reaction(dt, 1, 1, hyd=H)
        for k in ("XWsolidm0", "XWsolidm", "phim", "phinewm", "pfm0"):
            np.testing.assert_allclose(ctx.get_marker(k), io_o[k], rtol=1e-12, err_msg=k)
        np.testing.assert_allclose(ctx.get_grid("WTPSUM"), WTS_o, rtol=1e-12, atol=1e-14)
        pos = WTS_o > 1e-3
        np.testing.assert_allclose(ctx.get_grid("DMP")[pos], DMP_o[pos], rtol=1e-9,
                                   atol=1e-12 * np.abs(DMP_o).max())


def test_hydration_rows_at_workload_size():
    """Size-independent checks at the 281x281-node, 2.82M-marker workload:
    the melt-composition weights sum to M, a constant composition is
    reproduced at every node with weight, and the porosity update leaves
    air markers untouched and rocks within [phimin, phimax]."""
    from ereb import synth
    g, mat = synth.workload_geometry("fp64", mode=1)
    mk = synth.make_markers(g, mat, seed=0)
    M = mk["xm"].size
    H = Hydration()
    with ereb.Context(g, mat, ereb.FP64) as ctx:
        ctx.set_markers(mk)
        ctx.set_marker("XWsolidm0", np.full(M, 0.4375))
        ctx.step_melt_composition()
        WT = ctx.get_grid("WTPSUM")
        XWS = ctx.get_grid("XWS")
        assert abs(WT.sum() - M) <= 1e-9 * M
        np.testing.assert_allclose(XWS[WT > 1e-3], 0.4375, rtol=1e-13)
        rng = np.random.default_rng(9)
        ctx.set_grid("APHI", (rng.random((g.Ny1, g.Nx1)) - 0.5) * 6.0 / DT_LONGEST)
        ctx.step_marker_porosity(DT_LONGEST, hyd=H)
        ph = ctx.get_marker("phim")
    air = mk["tm"] == 3
    _exact(ph[air], mk["phim"][air], "air porosity")
    assert np.all((ph[~air] >= H.phimin) & (ph[~air] <= H.phimax))
