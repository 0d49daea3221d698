"""GPU parity: every step of the HIP library against the CPU oracle on the
same seeded inputs, through the C ABI.

Bars: cell indices, sort order and sparsity pattern bit-exact; Float64
values bit-exact where the operation sequence is the reference's
(properties, node normalisation, gathers, matrix coefficients); reordered
sums (the scatter) within 1e-12 of the sum of |terms| per node; RK4, which
multiplies by 1/dx and fuses products, within 1e-14 of |x| and 1e-13 of T;
sin/cos-dependent stresses within 1e-12 relative.
"""
from __future__ import annotations

import numpy as np
import pytest

import cases
import ereb
from ereb import api
from ereb.constants import test_constants as tconsts
from oracle import Oracle

pytestmark = pytest.mark.gpu

START_TIME = 2.25e6 * 365.25 * 24 * 3600  # C:339


def _ctx(g, mat):
    return ereb.Context(g, mat, ereb.FP64)


def _exact(a, b, what):
    a = np.asarray(a)
    b = np.asarray(b)
    same = (a == b) | (np.isnan(a) & np.isnan(b))
    if not same.all():
        k = np.flatnonzero(~same.ravel())[:5]
        raise AssertionError(f"{what}: {(~same).sum()} of {a.size} differ, e.g. "
                             f"{a.ravel()[k]} vs {b.ravel()[k]}")


def _sum_close(gpu, orc, absb, what, rtol=1e-12):
    err = np.abs(gpu - orc)
    bad = err > rtol * absb + 1e-300
    if bad.any():
        k = np.flatnonzero(bad.ravel())[:5]
        raise AssertionError(f"{what}: {bad.sum()} nodes exceed {rtol} of sum|terms|, "
                             f"err {err.ravel()[k]} bound {absb.ravel()[k]}")


@pytest.mark.parametrize("mode", [9, 1])
def test_scatter_dropin_vs_oracle(mode):
    g, mat = tconsts()
    mk = cases.ref_markers(g, mat, seed=42 + mode)
    hrs, hrf = mat.radioactive_heating(START_TIME)
    hrf = np.array([1e-7, 0.0, 0.0])  # exercise the fluid heat production term
    props_o, sums_o = Oracle(g, mat).marker_to_nodes(mk, hrs, hrf, mode)
    M = mk["xm"].size
    out = {k: mk[k].copy() for k in ["rhototalm", "rhocptotalm", "etatotalm", "hrtotalm",
                                     "ktotalm", "tkm_rhocptotalm", "etafluidcur_inv_kphim"]}
    sums = [np.zeros(s.shape, order="F") for s in sums_o]
    with _ctx(g, mat) as ctx:
        api.marker_to_nodes_(ctx, mk["xm"], mk["ym"], mk["tm"], mk["tkm"], mk["phim"],
                             mk["XWsolidm0"], mk["etavpm"], mk["inv_gggtotalm"],
                             mk["fricttotalm"], mk["cohestotalm"], mk["tenstotalm"],
                             mk["rhofluidcur"], mk["alphasolidcur"], mk["alphafluidcur"],
                             mk["sxxm"], mk["sxym"], out["rhototalm"], out["rhocptotalm"],
                             out["etatotalm"], out["hrtotalm"], out["ktotalm"],
                             out["tkm_rhocptotalm"], out["etafluidcur_inv_kphim"], hrs, hrf,
                             mode, sums, M)
        perm, key, nt = ctx.sort_order()
    for k in out:
        _exact(out[k], props_o[k], k)
    absb = cases.abs_sums(g, mk, props_o)
    for s, (a, b, c) in enumerate(zip(sums, sums_o, absb)):
        _sum_close(a, b, c, ereb.SUM_FIELDS[s])
    operm, okey, _ = Oracle(g, mat).fine_sort(mk["xm"], mk["ym"])
    _exact(key, okey, "fine keys")
    _exact(perm, operm, "sort order")


def test_scatter_material_vectors_by_type_and_fallback():
    """The seven material vectors that the reference derives from the marker
    type (E:659-665: 1/G, cohesion, tension, friction, fluid density and the
    two expansivities) are read through tm while every marker holds its
    type's values (checked on the device after any of them changes); one
    marker that does not makes the scatter read the vectors again.  Each
    state against the oracle on its own data."""
    g, mat = ereb.synth.workload_geometry("cpu41", mode=1)
    mk = ereb.synth.make_markers(g, mat, seed=1, stress=True)
    hrs, hrf = mat.radioactive_heating(START_TIME)
    orc = Oracle(g, mat)
    coh = mk["cohestotalm"].copy()
    coh[123] += 7.5e7
    rho = mk["rhofluidcur"].copy()
    rho[-5] *= 1.5
    states = [("typed", {}), ("one marker differs", {"cohestotalm": coh}),
              ("another differs", {"cohestotalm": mk["cohestotalm"], "rhofluidcur": rho}),
              ("typed again", {"rhofluidcur": mk["rhofluidcur"]})]
    data = dict(mk)
    with _ctx(g, mat) as ctx:
        ctx.set_markers(mk)
        for name, change in states:
            for k, v in change.items():
                ctx.set_marker(k, v)
                data[k] = v
            ctx.step_scatter(hrs, hrf, 1)
            sums = [ctx.get_grid(n) for n in ereb.SUM_FIELDS]
            props_o, sums_o = orc.marker_to_nodes(data, hrs, hrf, 1)
            absb = cases.abs_sums(g, data, props_o)
            for n, a, b, c in zip(ereb.SUM_FIELDS, sums, sums_o, absb):
                _sum_close(a, b, c, f"{name}: {n}")


def test_scatter_unknown_mode_raises():
    g, mat = tconsts()
    mk = cases.ref_markers(g, mat, extra=10)
    with _ctx(g, mat) as ctx:
        ctx.set_markers(mk)
        with pytest.raises(ereb.EREBModeError, match="unknown mode 3"):
            ctx.step_scatter(np.zeros(3), np.zeros(3), 3)


def test_scatter_deterministic_and_conservative_planetesimal():
    g, mat = ereb.synth.workload_geometry("cpu41", mode=1)
    mk = ereb.synth.make_markers(g, mat, seed=0, stress=True)
    hrs, hrf = mat.radioactive_heating(START_TIME)
    props_o, sums_o = Oracle(g, mat).marker_to_nodes(mk, hrs, hrf, 1)
    with _ctx(g, mat) as ctx:
        ctx.set_markers(mk)
        ctx.step_scatter(hrs, hrf, 1)
        s1 = [ctx.get_grid(n) for n in ereb.SUM_FIELDS]
        ctx.step_scatter(hrs, hrf, 1)
        s2 = [ctx.get_grid(n) for n in ereb.SUM_FIELDS]
    for n, a, b in zip(ereb.SUM_FIELDS, s1, s2):
        _exact(a, b, f"run-to-run {n}")
    absb = cases.abs_sums(g, mk, props_o)
    for n, a, b, c in zip(ereb.SUM_FIELDS, s1, sums_o, absb):
        _sum_close(a, b, c, n)
    M = mk["xm"].size
    for wt in ("WTSUM", "WTXSUM", "WTYSUM", "WTPSUM"):
        tot = s1[ereb.SUM_FIELDS.index(wt)].sum()
        assert abs(tot - M) <= 1e-9 * M, (wt, tot, M)


def test_node_properties_bit_exact():
    g, mat = tconsts()
    rng = np.random.default_rng(3)
    B, S = (g.Ny, g.Nx), (g.Ny1, g.Nx1)
    orc = Oracle(g, mat)
    sums_b = [cases.rand_grid(rng, B) for _ in range(8)]
    sums_b[7][rng.random(B) < 0.3] = 0.0  # zero weights keep old values
    outs_b = [cases.rand_grid(rng, B) for _ in range(7)]
    yny = (rng.random(B) < 0.2).astype(np.uint8, order="F")
    ob, oy = orc.basic_node_properties(sums_b, outs_b, yny)
    gb = [np.array(a, order="F", copy=True) for a in outs_b]
    gy = np.array(yny, order="F", copy=True)
    sums_v = [cases.rand_grid(rng, S) for _ in range(6)]
    sums_v[5][rng.random(S) < 0.3] = 0.0
    outs_v = [cases.rand_grid(rng, S) for _ in range(5)]
    ov = orc.v_node_properties(sums_v, outs_v)
    sums_p = [cases.rand_grid(rng, S) for _ in range(10)]
    sums_p[9][rng.random(S) < 0.3] = 0.0
    outs_p = [cases.rand_grid(rng, S) for _ in range(10)]
    op = orc.p_node_properties(sums_p, outs_p)
    with _ctx(g, mat) as ctx:
        api.compute_basic_node_properties_(ctx, *sums_b, *gb, gy)
        gvx = [np.array(a, order="F", copy=True) for a in outs_v]
        api.compute_vx_node_properties_(ctx, *sums_v, *gvx)
        gvy = [np.array(a, order="F", copy=True) for a in outs_v]
        api.compute_vy_node_properties_(ctx, *sums_v, *gvy)
        gp = [np.array(a, order="F", copy=True) for a in outs_p]
        api.compute_p_node_properties_(ctx, *sums_p, *gp)
        t = cases.rand_grid(rng, S)
        tt = np.array(t, order="F", copy=True)
        api.apply_insulating_boundary_conditions_(ctx, tt)
    for k, (a, b) in enumerate(zip(gb, ob)):
        _exact(a, b, f"basic out {k}")
    _exact(gy, oy, "YNY")
    for k, (a, b, c) in enumerate(zip(gvx, gvy, ov)):
        _exact(a, c, f"vx out {k}")
        _exact(b, c, f"vy out {k}")
    for k, (a, b) in enumerate(zip(gp, op)):
        _exact(a, b, f"p out {k}")
    _exact(tt, orc.insulating_bc(t), "insulating BC")


P_SUMS = ["RHOSUM", "RHOCPSUM", "ALPHASUM", "ALPHAFSUM", "HRSUM", "GGGPSUM", "SXXSUM", "TKSUM",
          "PHISUM", "WTPSUM"]
P_OUTS = ["RHO", "RHOCP", "ALPHA", "ALPHAF", "HR", "GGGP", "SXX0", "tk1", "PHI", "BETAPHI"]


def test_step_node_pass_applies_insulating_bc():
    """The resident node pass (ereb_step_node_properties) applies the
    insulating conditions to tk1 inside its P-grid pass: the P outputs equal
    compute_p_node_properties! followed by
    apply_insulating_boundary_conditions! on tk1 (E:2432-2449, oracle) bit
    for bit, with zero weights on part of the inner ring, whose nodes keep
    their previous tk1, which the boundary then copies (corners included)."""
    g, mat = tconsts()
    rng = np.random.default_rng(21)
    S = (g.Ny1, g.Nx1)
    sums_p = [cases.rand_grid(rng, S) for _ in range(10)]
    w = sums_p[9]
    w[rng.random(S) < 0.3] = 0.0
    w[1, 1::3] = 0.0
    w[-2, 2::4] = 0.0
    w[1::5, 1] = 0.0
    w[2::2, -2] = 0.0
    w[1, 1] = w[-2, -2] = 0.0
    outs_p = [cases.rand_grid(rng, S) for _ in range(10)]
    orc = Oracle(g, mat)
    op = orc.p_node_properties(sums_p, outs_p)
    op[7] = orc.insulating_bc(op[7])
    with _ctx(g, mat) as ctx:
        for n, a in zip(P_SUMS + P_OUTS, sums_p + outs_p):
            ctx.set_grid(n, a)
        ctx.step_node_properties()
        got = [ctx.get_grid(n) for n in P_OUTS]
    for n, a, b in zip(P_OUTS, got, op):
        _exact(a, b, n)


@pytest.mark.parametrize("timestep", [1, 2])
def test_gather_temperature_bit_exact(timestep):
    g, mat = tconsts()
    mk = cases.ref_markers(g, mat, seed=5)
    rng = np.random.default_rng(11)
    DT = cases.rand_grid(rng, (g.Ny1, g.Nx1), -5, 5)
    tk2 = cases.rand_grid(rng, (g.Ny1, g.Nx1), 263, 310)
    ref = Oracle(g, mat).update_marker_temperature(mk["xm"], mk["ym"], mk["tkm"], DT, tk2,
                                                   timestep)
    tkm = mk["tkm"].copy()
    with _ctx(g, mat) as ctx:
        api.update_marker_temperature_(ctx, mk["xm"], mk["ym"], tkm, DT, tk2, timestep,
                                       tkm.size)
    _exact(tkm, ref, "tkm")


def test_gather_stress_bit_exact():
    g, mat = tconsts()
    mk = cases.ref_markers(g, mat, seed=6)
    rng = np.random.default_rng(12)
    DSXX = cases.rand_grid(rng, (g.Ny1, g.Nx1), -1e5, 1e5)
    DSXY = cases.rand_grid(rng, (g.Ny, g.Nx), -1e5, 1e5)
    rx, ry = Oracle(g, mat).update_marker_stress(mk["xm"], mk["ym"], mk["sxxm"], mk["sxym"],
                                                 DSXX, DSXY)
    sx, sy = mk["sxxm"].copy(), mk["sxym"].copy()
    with _ctx(g, mat) as ctx:
        api.update_marker_stress_(ctx, mk["xm"], mk["ym"], sx, sy, DSXX, DSXY, sx.size)
    _exact(sx, rx, "sxxm")
    _exact(sy, ry, "sxym")


@pytest.mark.parametrize("mode", [9, 1])
def test_rk4_vs_oracle(mode):
    g, mat = tconsts()
    mk = cases.ref_markers(g, mat, seed=21)
    rng = np.random.default_rng(22)
    S = (g.Ny1, g.Nx1)
    vx, vy, vxf, vyf = (cases.rand_grid(rng, S, -1, 1) for _ in range(4))
    wyx = cases.rand_grid(rng, (g.Ny, g.Nx), -1, 1)
    tk2 = cases.rand_grid(rng, S, 263, 310)
    dt = 0.9 * g.dx  # Courant ~0.9 for |v| <= 1 (T:5760 uses dtm = 0.9)
    ref = Oracle(g, mat).move_markers_rk4(mk["xm"], mk["ym"], mk["tm"], mk["tkm"], mk["phim"],
                                          mk["sxym"], mk["sxxm"], vx, vy, vxf, vyf, wyx, tk2,
                                          dt, mode)
    xm, ym, tkm = mk["xm"].copy(), mk["ym"].copy(), mk["tkm"].copy()
    sxy, sxx = mk["sxym"].copy(), mk["sxxm"].copy()
    with _ctx(g, mat) as ctx:
        api.move_markers_rk4_(ctx, xm, ym, mk["tm"], tkm, mk["phim"], sxy, sxx, vx, vy, vxf,
                              vyf, wyx, tk2, xm.size, dt, mode)
    # the kernel multiplies by 1/dx, 1/dy and fuses a*b+c where the reference
    # divides and rounds separately: positions within 1e-14 of |x| + dt*max|v|
    # (the RK4 displacement scale), temperatures within 1e-13 relative
    scale = dt * max(np.abs(v).max() for v in (vx, vy, vxf, vyf))
    for name, a, b in (("xm", xm, ref[0]), ("ym", ym, ref[1])):
        err = np.abs(a - b)
        assert np.all(err <= 1e-14 * (np.abs(b) + scale)), (name, err.max())
    np.testing.assert_allclose(tkm, ref[2], rtol=1e-13, atol=0, err_msg="tkm")
    scale = np.abs(mk["sxxm"]) + np.abs(mk["sxym"])
    assert np.all(np.abs(sxy - ref[3]) <= 1e-12 * scale), "sxym"
    assert np.all(np.abs(sxx - ref[4]) <= 1e-12 * scale), "sxxm"


def test_assemble_stokes_darcy_bit_exact():
    g, mat = tconsts()
    grids = cases.stokes_inputs(g)
    dt = 1e11
    cp, rv, nz, R = Oracle(g, mat).assemble_hydromechanical(grids, dt)
    Rg = np.zeros(R.size)
    with _ctx(g, mat) as ctx:
        L = api.assemble_hydromechanical_lse_(ctx, *grids, dt, Rg)
    _exact(L.colptr, cp, "colptr")
    _exact(L.rowval, rv, "rowval")
    _exact(L.nzval, nz, "nzval")
    _exact(Rg, R, "R")
    assert L.nnz < 40 * g.Nx1 * g.Ny1


def test_assemble_thermal_bit_exact():
    g, mat = tconsts()
    grids = cases.thermal_inputs(g)
    dt = 3e10
    cp, rv, nz, R = Oracle(g, mat).assemble_thermal(grids, dt)
    Rg = np.zeros(R.size)
    with _ctx(g, mat) as ctx:
        L = api.assemble_thermal_lse_(ctx, *grids, Rg, dt)
    _exact(L.colptr, cp, "colptr")
    _exact(L.rowval, rv, "rowval")
    _exact(L.nzval, nz, "nzval")
    _exact(Rg, R, "RT")
