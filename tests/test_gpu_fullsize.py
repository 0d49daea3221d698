"""GPU parity at the benchmark's sizes: the 281x281-node workload with
2,822,400 markers and the 561x561-node one with 20,070,400 markers
(planetesimal of R = 40 and 65 km), all with jittered lattice markers in the
reference's order.  The C oracle finishes these in seconds, so the bars are
those of the small parity tests: fine keys, sort order, marker properties,
sparsity patterns and matrix values bit-exact; scatter sums within 1e-12 of
the per-node sum of |terms| (which the oracle also computes, in C); RK4
positions within 1e-14 of |x| + dt*max|v|, temperatures within 1e-13
relative, stresses within 1e-12 of |sxx| + |sxy|.
Size-independent checks ride along: the weights of every grid sum to the
marker count and a second scatter reproduces the first bit for bit.
"""
from __future__ import annotations

import numpy as np
import pytest

import cases
import ereb
from ereb import api, synth
from oracle import Oracle

pytestmark = pytest.mark.gpu

START_TIME = 2.25e6 * 365.25 * 24 * 3600  # C:339


def _exact(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    same = (a == b) | (np.isnan(a) & np.isnan(b))
    if not same.all():
        k = np.flatnonzero(~same.ravel())[:5]
        raise AssertionError(f"{what}: {(~same).sum()} of {a.size} differ, e.g. "
                             f"{a.ravel()[k]} vs {b.ravel()[k]}")


@pytest.mark.parametrize("name", ["fp64", "large40", "large65"])
def test_sort_properties_and_scatter_at_workload_size(name):
    g, mat = synth.workload_geometry(name, mode=1)
    mk = synth.make_markers(g, mat, seed=0, stress=True)
    M = mk["xm"].size
    print(f"[{name}] {M:,} markers, {g.Nx}x{g.Ny} basic nodes", flush=True)
    hrs, hrf = mat.radioactive_heating(START_TIME)
    orc = Oracle(g, mat)
    props_o, sums_o = orc.marker_to_nodes(mk, hrs, hrf, 1)
    perm_o, key_o, _ = orc.fine_sort(mk["xm"], mk["ym"])
    with ereb.Context(g, mat, ereb.FP64) as ctx:
        ctx.set_markers(mk)
        ctx.step_scatter(hrs, hrf, 1)
        perm, key, _ = ctx.sort_order()
        s1 = [ctx.get_grid(n) for n in ereb.SUM_FIELDS]
        props = {n: ctx.get_marker(n) for n in props_o}
        ctx.step_scatter(hrs, hrf, 1)
        s2 = [ctx.get_grid(n) for n in ereb.SUM_FIELDS]
    _exact(key, key_o, "fine key")
    _exact(perm, perm_o, "sort order")
    for n in props_o:
        _exact(props[n], props_o[n], n)
    for n, a, b in zip(ereb.SUM_FIELDS, s1, s2):
        _exact(a, b, f"run-to-run {n}")
    _, absb = orc.marker_to_nodes(mk, hrs, hrf, 1, absval=True)
    for n, a, b, c in zip(ereb.SUM_FIELDS, s1, sums_o, absb):
        err = np.abs(a - b)
        bad = err > 1e-12 * c + 1e-300
        assert not bad.any(), (n, int(bad.sum()), float(err.max()))
    for wt in ("WTSUM", "WTXSUM", "WTYSUM", "WTPSUM"):
        tot = s1[ereb.SUM_FIELDS.index(wt)].sum()
        assert abs(tot - M) <= 1e-9 * M, (wt, tot, M)


def test_rk4_at_workload_size():
    g, mat = synth.workload_geometry("fp64", mode=1)
    mk = synth.make_markers(g, mat, seed=0, stress=True)
    rng = np.random.default_rng(3)
    S = (g.Ny1, g.Nx1)
    vx, vy = synth.stream_velocity(g)
    # fluid velocities differ from the solid ones, so the E:5173 read of vx
    # inside the fluid loop is exercised
    vxf = vx * (1.0 + 0.2 * (rng.random(S) - 0.5))
    vyf = vy * (1.0 + 0.2 * (rng.random(S) - 0.5))
    wyx = synth.rotation_rate(g, vx, vy)
    tk2 = 150.0 + 250.0 * rng.random(S)
    dt = ereb.DT_LONGEST
    ref = Oracle(g, mat).move_markers_rk4(mk["xm"], mk["ym"], mk["tm"], mk["tkm"], mk["phim"],
                                          mk["sxym"], mk["sxxm"], vx, vy, vxf, vyf, wyx, tk2,
                                          dt, 1)
    xm, ym, tkm = mk["xm"].copy(), mk["ym"].copy(), mk["tkm"].copy()
    sxy, sxx = mk["sxym"].copy(), mk["sxxm"].copy()
    with ereb.Context(g, mat, ereb.FP64) as ctx:
        api.move_markers_rk4_(ctx, xm, ym, mk["tm"], tkm, mk["phim"], sxy, sxx, vx, vy, vxf,
                              vyf, wyx, tk2, xm.size, dt, 1)
    scale = dt * max(np.abs(v).max() for v in (vx, vy, vxf, vyf))
    for name, a, b in (("xm", xm, ref[0]), ("ym", ym, ref[1])):
        err = np.abs(a - b)
        assert np.all(err <= 1e-14 * (np.abs(b) + scale)), (name, err.max())
    np.testing.assert_allclose(tkm, ref[2], rtol=1e-13, atol=0, err_msg="tkm")
    sc = np.abs(mk["sxxm"]) + np.abs(mk["sxym"])
    assert np.all(np.abs(sxy - ref[3]) <= 1e-12 * sc), "sxym"
    assert np.all(np.abs(sxx - ref[4]) <= 1e-12 * sc), "sxxm"


@pytest.mark.parametrize("name", ["fp64", "large40"])
def test_assembly_at_workload_size(name):
    g, mat = synth.workload_geometry(name, mode=1)
    orc = Oracle(g, mat)
    sd = cases.stokes_inputs(g)
    th = cases.thermal_inputs(g)
    cp, rv, nz, R = orc.assemble_hydromechanical(sd, 1e11)
    cpt, rvt, nzt, RT = orc.assemble_thermal(th, 3e10)
    Rg, RTg = np.zeros(R.size), np.zeros(RT.size)
    with ereb.Context(g, mat, ereb.FP64) as ctx:
        L = api.assemble_hydromechanical_lse_(ctx, *sd, 1e11, Rg)
        LT = api.assemble_thermal_lse_(ctx, *th, RTg, 3e10)
    print(f"[{name}] Stokes-Darcy n={R.size:,} nnz={nz.size:,}", flush=True)
    for what, a, b in (("colptr", L.colptr, cp), ("rowval", L.rowval, rv),
                       ("nzval", L.nzval, nz), ("R", Rg, R), ("thermal colptr", LT.colptr, cpt),
                       ("thermal rowval", LT.rowval, rvt), ("thermal nzval", LT.nzval, nzt),
                       ("RT", RTg, RT)):
        _exact(a, b, what)


def test_mixed_form_at_workload_size():
    """The mixed form against the Float64 form over 5 transfer steps of the
    2.8M-marker workload: every SUM and node field within 1e-6 normwise
    (max|a - ref| / max|ref|), the bar of the mixed form."""
    from ereb import precision
    g, mat = synth.workload_geometry("fp64", mode=1)
    mk = synth.make_markers(g, mat, seed=0, stress=True)
    hrs, hrf = mat.radioactive_heating(START_TIME)
    vx, vy = synth.stream_velocity(g)
    rng = np.random.default_rng(4)
    S = (g.Ny1, g.Nx1)
    grids = {"vx": vx, "vy": vy, "vxf": vx, "vyf": vy, "wyx": synth.rotation_rate(g, vx, vy),
             "ETAP": np.full(S, 1e18), "ETAPHI": np.full(S, 1e19), "gx": np.full(S, -0.1),
             "gy": np.full(S, -0.2), "DT": 0.01 * rng.random(S), "tk2": 250 + 50 * rng.random(S),
             "pr0": np.full(S, 1e5), "pf0": np.full(S, 9e4), "DSXX": 1e3 * rng.random(S),
             "DSXY": 1e3 * rng.random(g.basic_shape)}
    ctxs = []
    try:
        for prec in (ereb.FP64, ereb.MIXED):
            c = ereb.Context(g, mat, prec)
            ctxs.append(c)
            c.set_markers(mk)
            for n, a in grids.items():
                c.set_grid(n, a)
            for ts in range(1, 6):
                c.step_transfer(hrs, hrf, 1, ereb.DT_LONGEST, ts)
        errs = precision.field_errors(ctxs[1], ctxs[0])
        nodal = ereb.SUM_FIELDS + precision.NODE_FIELDS
        bad = {k: errs[k][0] for k in nodal if errs[k][0] > 1e-6}
        assert not bad, bad
    finally:
        for c in ctxs:
            c.close()
