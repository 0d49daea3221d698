"""Mixed-precision form (EREB_MIXED): FP32 marker properties, weights,
interpolation and per-fine-cell sums with compensated FP32 per-node
accumulation; Float64 marker state, node normalisation and matrix
coefficients.

Bar: every nodal field (the 30 SUM grids and the 27 normalised node fields)
and every marker property within 1e-6 of the Float64 result, normwise
(max|a - ref| / max|ref| per field); errors are reported per field by
ereb.precision.  Checked against the CPU oracle for one pass and against the
Float64 GPU path (itself within 1e-12 of the oracle) over 10 full steps.
"""
from __future__ import annotations

import numpy as np
import pytest

import cases
import ereb
from ereb import api, precision, synth
from ereb.constants import test_constants as tconsts
from oracle import Oracle

pytestmark = pytest.mark.gpu

TOL = 1e-6
START_TIME = 2.25e6 * 365.25 * 24 * 3600  # C:339


def _check(errs, names, tol=TOL):
    bad = {n: errs[n] for n in names if errs[n][0] > tol}
    assert not bad, f"fields above {tol} normwise: {bad}"


def test_mixed_properties_at_temperature_breakpoints():
    """Regression: the FP32 property laws must branch on the Float64
    temperature.  Rounded to float first, a marker within half a float
    spacing (1.5e-5 K at 273 K) of a breakpoint of compute_rhocpfluidm
    (E:2272-2293: 268, 273, 278, 410 K), compute_kfluidm (273, 410 K) or the
    fluid/solid phase switches (273, 1416 K) took the other branch: rhocp
    off by up to 40 %.  Every marker property per marker within 1e-6."""
    g, mat = synth.workload_geometry("cpu41", mode=1)
    mk = synth.make_markers(g, mat, seed=4)
    rng = np.random.default_rng(8)
    rock = np.flatnonzero(mk["tm"] < 3)
    br = np.array([268.0, 273.0, 278.0, 410.0, 1416.0])
    off = np.array([-1e-5, -3e-6, -1e-9, 0.0, 1e-9, 3e-6, 1e-5])
    T = mk["tkm"].copy()
    T[rock] = (br[rng.integers(0, br.size, rock.size)] +
               off[rng.integers(0, off.size, rock.size)])
    mk["tkm"] = T
    hrs, hrf = mat.radioactive_heating(START_TIME)
    props_o, _ = Oracle(g, mat).marker_to_nodes(mk, hrs, hrf, 1)
    with ereb.Context(g, mat, ereb.MIXED) as ctx:
        ctx.set_markers(mk)
        ctx.step_scatter(hrs, hrf, 1)
        for n in precision.MARKER_PROPS:
            got, ref = ctx.get_marker(n)[rock], props_o[n][rock]
            # k_fluid is negative far above 410 K, so ktotal is NaN there in
            # both (Julia's sqrt would throw): NaN must match NaN
            both = np.isnan(got) & np.isnan(ref)
            err = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300)
            assert np.all(both | (err <= TOL)), (n, np.nanmax(err))


@pytest.mark.parametrize("mode", [1, 9])
def test_mixed_scatter_and_nodes_vs_oracle(mode):
    g, mat = synth.workload_geometry("cpu41", mode=mode)
    mk = synth.make_markers(g, mat, seed=3, stress=True)
    hrs, hrf = mat.radioactive_heating(START_TIME)
    orc = Oracle(g, mat)
    props_o, sums_o = orc.marker_to_nodes(mk, hrs, hrf, mode)
    B, S = g.basic_shape, g.stag_shape
    z = lambda shp, n: [np.zeros(shp, order="F") for _ in range(n)]
    ob, _ = orc.basic_node_properties(sums_o[:8], z(B, 7), np.zeros(B, np.uint8, order="F"))
    ovx = orc.v_node_properties(sums_o[8:14], z(S, 5))
    ovy = orc.v_node_properties(sums_o[14:20], z(S, 5))
    op = orc.p_node_properties(sums_o[20:30], z(S, 10))
    op[7] = orc.insulating_bc(op[7])
    with ereb.Context(g, mat, ereb.MIXED) as ctx:
        ctx.set_markers(mk)
        ctx.step_scatter(hrs, hrf, mode)
        ctx.step_node_properties()
        errs = {}
        for n, ref in zip(ereb.SUM_FIELDS, sums_o):
            errs[n] = precision.rel_errors(ctx.get_grid(n), ref)
        refs = dict(zip(precision.NODE_FIELDS, list(ob) + list(ovx) + list(ovy) + list(op)))
        for n, ref in refs.items():
            errs[n] = precision.rel_errors(ctx.get_grid(n), ref)
        for n in precision.MARKER_PROPS:
            errs[n] = precision.rel_errors(ctx.get_marker(n), props_o[n])
    _check(errs, list(errs))
    # the weights sum to the marker count on every grid (conservation)
    for n in ("WTSUM", "WTXSUM", "WTYSUM", "WTPSUM"):
        assert errs[n][0] < 1e-7, (n, errs[n])


def test_mixed_gathers_vs_oracle():
    g, mat = tconsts()
    mk = cases.ref_markers(g, mat, seed=8)
    rng = np.random.default_rng(9)
    DT = cases.rand_grid(rng, (g.Ny1, g.Nx1), -5, 5)
    tk2 = cases.rand_grid(rng, (g.Ny1, g.Nx1), 263, 310)
    DSXX = cases.rand_grid(rng, (g.Ny1, g.Nx1), -1e5, 1e5)
    DSXY = cases.rand_grid(rng, (g.Ny, g.Nx), -1e5, 1e5)
    orc = Oracle(g, mat)
    rT = orc.update_marker_temperature(mk["xm"], mk["ym"], mk["tkm"], DT, tk2, 2)
    rx, ry = orc.update_marker_stress(mk["xm"], mk["ym"], mk["sxxm"], mk["sxym"], DSXX, DSXY)
    tkm, sx, sy = mk["tkm"].copy(), mk["sxxm"].copy(), mk["sxym"].copy()
    with ereb.Context(g, mat, ereb.MIXED) as ctx:
        api.update_marker_temperature_(ctx, mk["xm"], mk["ym"], tkm, DT, tk2, 2, tkm.size)
        api.update_marker_stress_(ctx, mk["xm"], mk["ym"], sx, sy, DSXX, DSXY, sx.size)
    # the temperature gather is Float64 in both forms (the property laws
    # switch branch at fixed temperatures); the stress gather interpolates
    # the increment in FP32 and adds it to the Float64 state
    assert np.array_equal(tkm, rT)
    assert np.max(np.abs(sx - rx)) <= TOL * np.abs(DSXX).max()
    assert np.max(np.abs(sy - ry)) <= TOL * np.abs(DSXY).max()


def _run(g, mat, mk, grids, prec, steps):
    hrs, hrf = mat.radioactive_heating(START_TIME)
    mode = mat.marker_property_mode
    ctx = ereb.Context(g, mat, prec)
    ctx.set_markers(mk)
    for n, a in grids.items():
        ctx.set_grid(n, a)
    for s in range(steps):
        ctx.step_scatter(hrs, hrf, mode)
        ctx.step_node_properties()
        ctx.step_assemble_stokes_darcy(ereb.DT_LONGEST)
        ctx.step_assemble_thermal(ereb.DT_LONGEST)
        ctx.step_gather_stress()
        ctx.step_gather_temperature(s + 2)
        ctx.step_advect_rk4(ereb.DT_LONGEST, mode)
    ctx.step_scatter(hrs, hrf, mode)
    ctx.step_node_properties()
    ctx.sync()
    return ctx


def test_mixed_ten_steps_vs_fp64():
    g, mat = synth.workload_geometry("cpu41", mode=1)
    mk = synth.make_markers(g, mat, seed=0, stress=True)
    vx, vy = synth.stream_velocity(g)
    S = g.stag_shape
    grids = {"vx": vx, "vy": vy, "vxf": vx, "vyf": vy, "wyx": synth.rotation_rate(g, vx, vy),
             "DT": synth.smooth_field(g, S, 0.05, 2, 1),
             "tk2": np.full(S, 280.0, order="F") + synth.smooth_field(g, S, 5.0, 1, 1),
             "DSXX": synth.smooth_field(g, S, 1e3, 3, 2),
             "DSXY": synth.smooth_field(g, g.basic_shape, 1e3, 2, 3),
             "ETAP": np.full(S, 1e21, order="F"), "ETAPHI": np.full(S, 1e22, order="F"),
             "pr0": np.full(S, mat.psurface, order="F"),
             "pf0": np.full(S, mat.psurface, order="F")}
    grids["gx"], grids["gy"] = synth.disk_gravity(g)
    c64 = _run(g, mat, mk, grids, ereb.FP64, 10)
    cmx = _run(g, mat, mk, grids, ereb.MIXED, 10)
    try:
        errs = precision.field_errors(cmx, c64)
    finally:
        cmx.close()
        c64.close()
    _check(errs, ereb.SUM_FIELDS + precision.NODE_FIELDS + precision.MARKER_PROPS)
    for n in ("stokes_darcy.nzval", "thermal.nzval", "stokes_darcy.rhs", "thermal.rhs"):
        assert n in errs and errs[n][0] <= TOL, (n, errs.get(n))


def test_precision_switch_converts_properties():
    g, mat = tconsts()
    mk = cases.ref_markers(g, mat, seed=2, extra=100)
    with ereb.Context(g, mat, ereb.FP64) as ctx:
        ctx.set_markers(mk)
        ctx.set_precision(ereb.MIXED)
        for n in ("etavpm", "rhototalm", "xm", "tkm"):
            a = ctx.get_marker(n)
            if n in ("xm", "tkm"):  # state stays Float64
                assert np.array_equal(a, mk[n]), n
            else:
                assert np.array_equal(a, mk[n].astype(np.float32).astype(np.float64)), n
        ctx.set_precision(ereb.FP64)
        a = ctx.get_marker("etavpm")
        assert np.array_equal(a, mk["etavpm"].astype(np.float32).astype(np.float64))
