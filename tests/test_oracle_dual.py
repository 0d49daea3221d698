"""Pins of the dual-number oracle (oracle/ereb_dual.cpp), the exact
derivatives the GPU pullbacks and adjoint are checked against.

  - with S = double the templates reproduce the oracle: move_markers_rk4!
    and the marker properties bit for bit, the marker contributions dotted
    with random SUM cotangents summing to the cotangents dotted with the
    oracle's SUM grids (to rounding);
  - their derivatives agree with central finite differences of the oracle's
    own functions (scatter, both gathers, RK4 in every marker field and
    grid direction, the multi-step loss), which catches a wrong sign, index
    or dropped term in the templates' tangent propagation;
  - closed forms: the derivative of the temperature gather of an affine DT
    field along x is its slope; the RK4 step in a uniform velocity field
    moves every marker rigidly (d x_new / d x = 1, d x_new / d y = 0); the
    weight sum of the scatter does not depend on the position.
"""
from __future__ import annotations

import numpy as np

import cases
import ereb
from ereb import synth
from ereb.constants import test_constants as tconsts
from oracle import DualOracle, Oracle

START_TIME = 2.25e6 * 365.25 * 24 * 3600
DT = ereb.DT_LONGEST


def _setup(seed=3, extra=300):
    g, mat = tconsts()
    mk = cases.ref_markers(g, mat, seed=seed, extra=extra)
    hrs, hrf = mat.radioactive_heating(7.1e13)
    return g, mat, mk, hrs, hrf


def _rk4_grids(g, rng):
    S, B = g.stag_shape, g.basic_shape
    return {"vx": cases.rand_grid(rng, S, -1e-9, 1e-9), "vy": cases.rand_grid(rng, S, -1e-9, 1e-9),
            "vxf": cases.rand_grid(rng, S, -1e-9, 1e-9), "vyf": cases.rand_grid(rng, S, -1e-9, 1e-9),
            "wyx": cases.rand_grid(rng, B, -1e-12, 1e-12), "tk2": cases.rand_grid(rng, S, 200, 300)}


def _fd(f, x0, h):
    return (f(x0 + h) - f(x0 - h)) / (2 * h)


def test_rk4_template_equals_oracle_bit_for_bit():
    g, mat, mk, _, _ = _setup()
    rng = np.random.default_rng(1)
    gr = _rk4_grids(g, rng)
    D, O = DualOracle(g, mat), Oracle(g, mat)
    for mode in (1, 9):
        a = D.rk4_double(mk, gr, 3e11, mode)
        b = O.move_markers_rk4(mk["xm"], mk["ym"], mk["tm"], mk["tkm"], mk["phim"], mk["sxym"],
                               mk["sxxm"], gr["vx"], gr["vy"], gr["vxf"], gr["vyf"], gr["wyx"],
                               gr["tk2"], 3e11, mode)
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


def test_scatter_template_equals_oracle():
    g, mat, mk, hrs, hrf = _setup()
    rng = np.random.default_rng(2)
    bars = [cases.rand_grid(rng, g.basic_shape if k < 8 else g.stag_shape, -1, 1)
            for k in range(30)]
    D, O = DualOracle(g, mat), Oracle(g, mat)
    for mode in (1, 9):
        dots, props = D.marker_to_nodes_dot(mk, hrs, hrf, mode, bars)
        props_o, sums = O.marker_to_nodes(mk, hrs, hrf, mode)
        names = ["rhototalm", "rhocptotalm", "etatotalm", "hrtotalm", "ktotalm",
                 "tkm_rhocptotalm", "etafluidcur_inv_kphim"]
        for k, n in enumerate(names):
            assert np.array_equal(props[k], props_o[n]), n
        want = sum(float(np.sum(b * s)) for b, s in zip(bars, sums))
        scale = sum(float(np.sum(np.abs(b * s))) for b, s in zip(bars, sums))
        assert abs(dots.sum() - want) <= 1e-12 * scale


def test_scatter_derivatives_against_oracle_differences():
    g, mat, mk, hrs, hrf = _setup(extra=50)
    rng = np.random.default_rng(4)
    bars = [cases.rand_grid(rng, g.basic_shape if k < 8 else g.stag_shape, -1, 1)
            for k in range(30)]
    pbars = [rng.random(mk["xm"].size) for _ in range(7)]
    D, O = DualOracle(g, mat), Oracle(g, mat)
    names = ["rhototalm", "rhocptotalm", "etatotalm", "hrtotalm", "ktotalm", "tkm_rhocptotalm",
             "etafluidcur_inv_kphim"]
    # markers away from the property kinks and cell edges
    pick = [m for m in rng.choice(mk["xm"].size, 200, replace=False)
            if min(abs(mk["tkm"][m] - b) for b in (263, 268, 273, 278, 410)) > 0.5][:25]

    def out(field, m, v):
        sub = {k: a[[m]].copy() for k, a in mk.items()}
        sub[field][0] = v
        props, sums = O.marker_to_nodes(sub, hrs, hrf, 1)
        return sums, [props[n][0] for n in names]

    def fd_obj(field, m, h):
        # the difference of the outputs first, then weighted: the outputs that
        # do not depend on the field cancel exactly instead of adding rounding
        # noise of their (large) values
        (sa, pa), (sb, pb_) = out(field, m, mk[field][m] + h), out(field, m, mk[field][m] - h)
        s = sum(float(np.sum(b * (x - y))) for b, x, y in zip(bars, sa, sb))
        s += sum(float(pb[m] * (x - y)) for pb, x, y in zip(pbars, pa, pb_))
        return s / (2 * h)

    for field in ("xm", "ym", "tkm", "phim", "XWsolidm0", "sxxm", "etavpm", "rhofluidcur"):
        ad = D.marker_to_nodes_vjp(mk, hrs, hrf, 1, bars, pbars, field)
        for m in pick:
            x0 = mk[field][m]
            h = 1e-6 * max(abs(x0), 1.0 if field in ("xm", "ym") else 1e-3)
            fd = fd_obj(field, m, h)
            assert abs(fd - ad[m]) <= 1e-6 * max(abs(ad[m]), abs(fd)) + 1e-9, (field, m, fd, ad[m])


def test_gather_derivatives_against_oracle_differences():
    g, mat, mk, _, _ = _setup(extra=100)
    rng = np.random.default_rng(5)
    S, B = g.stag_shape, g.basic_shape
    DTg, tk2 = cases.rand_grid(rng, S, -1, 1), cases.rand_grid(rng, S, 200, 300)
    DSXX, DSXY = cases.rand_grid(rng, S, -1e3, 1e3), cases.rand_grid(rng, B, -1e3, 1e3)
    M = mk["xm"].size
    tb, sb, yb = rng.random(M) - 0.5, rng.random(M) - 0.5, rng.random(M) - 0.5
    D, O = DualOracle(g, mat), Oracle(g, mat)
    pick = rng.choice(M, 30, replace=False)
    for ts in (1, 2):
        for which, key in (("xm", "xm"), ("ym", "ym"), ("tkm", "tkm")):
            ad = D.gather_T_vjp(mk["xm"], mk["ym"], mk["tkm"], DTg, tk2, ts, tb, which)
            for m in pick:
                def f(v):
                    p = {k: mk[k][[m]].copy() for k in ("xm", "ym", "tkm")}
                    p[key][0] = v
                    return tb[m] * O.update_marker_temperature(p["xm"], p["ym"], p["tkm"], DTg,
                                                               tk2, ts)[0]
                # the interpolation is linear in x, y within a cell: steps of
                # 0.5 m carry no truncation error, only the rounding of |f|
                h = 0.5 if key != "tkm" else 1e-3
                fd = _fd(f, mk[key][m], h)
                tol = 1e-8 * abs(ad[m]) + 4e-16 * abs(f(mk[key][m])) / h
                assert abs(fd - ad[m]) <= tol, (ts, which, m, fd, ad[m])
    for which in ("xm", "ym", "sxxm", "sxym"):
        ad = D.gather_S_vjp(mk["xm"], mk["ym"], mk["sxxm"], mk["sxym"], DSXX, DSXY, sb, yb, which)
        for m in pick:
            def f(v):
                p = {k: mk[k][[m]].copy() for k in ("xm", "ym", "sxxm", "sxym")}
                p[which][0] = v
                a, b = O.update_marker_stress(p["xm"], p["ym"], p["sxxm"], p["sxym"], DSXX, DSXY)
                return sb[m] * a[0] + yb[m] * b[0]
            h = 0.5 if which in ("xm", "ym") else 1.0
            fd = _fd(f, mk[which][m], h)
            tol = 1e-8 * abs(ad[m]) + 4e-16 * (abs(sb[m] * mk["sxxm"][m]) +
                                               abs(yb[m] * mk["sxym"][m]) + 1e3) / h
            assert abs(fd - ad[m]) <= tol, (which, m, fd, ad[m])
    # grid directions: the gathers are linear in their grids, so the derivative
    # along a direction is the gather of the direction itself
    gd = cases.rand_grid(rng, S, -1, 1)
    ad = D.gather_T_vjp(mk["xm"], mk["ym"], mk["tkm"], DTg, tk2, 2, tb, "grid", gd)
    want = tb * O.update_marker_temperature(mk["xm"], mk["ym"], np.zeros(M), gd, tk2, 2)
    np.testing.assert_allclose(ad, want, rtol=1e-13, atol=1e-13)


def test_gather_closed_form_affine_field():
    """d/dx of tkm + <DT> for DT = a + b x + c y is b, at interior markers."""
    g, mat, mk, _, _ = _setup(extra=0)
    xp, yp = g.axes()["xp"], g.axes()["yp"]
    X, Y = np.meshgrid(xp, yp)
    DTg = np.asfortranarray(3.0 + 2e-5 * X - 7e-6 * Y)
    inner = (mk["xm"] > xp[1]) & (mk["xm"] < xp[-2]) & (mk["ym"] > yp[1]) & (mk["ym"] < yp[-2])
    D = DualOracle(g, mat)
    one = np.ones(mk["xm"].size)
    dx = D.gather_T_vjp(mk["xm"], mk["ym"], mk["tkm"], DTg, DTg, 2, one, "xm")
    dy = D.gather_T_vjp(mk["xm"], mk["ym"], mk["tkm"], DTg, DTg, 2, one, "ym")
    np.testing.assert_allclose(dx[inner], 2e-5, rtol=1e-9)
    np.testing.assert_allclose(dy[inner], -7e-6, rtol=1e-9)


def test_rk4_derivatives_against_oracle_differences():
    g, mat, mk, _, _ = _setup(extra=60)
    rng = np.random.default_rng(6)
    gr = _rk4_grids(g, rng)
    M = mk["xm"].size
    bars = {k: rng.random(M) - 0.5 for k in ("xm", "ym", "tkm", "sxym", "sxxm")}
    dt = 3e11
    D, O = DualOracle(g, mat), Oracle(g, mat)
    pick = [m for m in rng.choice(M, 60, replace=False)
            if min(abs(mk["tkm"][m] - b) for b in (263, 268, 273, 278, 410)) > 0.5][:20]

    outs = ("xm", "ym", "tkm", "sxym", "sxxm")

    def run(sel, field, v, grids=gr):
        p = {k: mk[k][sel].copy() for k in ("xm", "ym", "tm", "tkm", "phim", "sxym", "sxxm")}
        if field is not None:
            p[field][0] = v
        return O.move_markers_rk4(p["xm"], p["ym"], p["tm"], p["tkm"], p["phim"], p["sxym"],
                                  p["sxxm"], grids["vx"], grids["vy"], grids["vxf"], grids["vyf"],
                                  grids["wyx"], grids["tk2"], dt, 1)

    def fd_of(sel, a, b, h):
        # per-output differences first (outputs that do not move cancel
        # exactly), then weighted by the output cotangents
        return sum(float(np.sum(bars[k][sel] * (x - y)))
                   for k, x, y in zip(outs, a, b)) / (2 * h)

    steps = {"xm": 1e-2, "ym": 1e-2, "tkm": 1e-4, "phim": 1e-7, "sxym": 1.0, "sxxm": 1.0}
    for field, h in steps.items():
        ad = D.rk4_vjp(mk, gr, dt, 1, bars, field)
        for m in pick:
            fd = fd_of([m], run([m], field, mk[field][m] + h), run([m], field, mk[field][m] - h), h)
            assert abs(fd - ad[m]) <= 1e-6 * (abs(ad[m]) + 1e-6), (field, m, fd, ad[m])
    # grid directions, summed over the markers
    allm = np.arange(M)
    for name in ("vx", "vy", "vxf", "vyf", "wyx", "tk2"):
        gd = cases.rand_grid(rng, gr[name].shape, -1, 1) * np.abs(gr[name]).max()
        ad = D.rk4_vjp(mk, gr, dt, 1, bars, name, gd).sum()
        e = 1e-5
        ga, gb = dict(gr), dict(gr)
        ga[name] = gr[name] + e * gd
        gb[name] = gr[name] - e * gd
        fd = fd_of(allm, run(allm, None, 0.0, ga), run(allm, None, 0.0, gb), e)
        assert abs(fd - ad) <= 1e-6 * abs(ad) + 1e-9, (name, fd, ad)


def test_rk4_uniform_velocity_moves_rigidly():
    g, mat, mk, _, _ = _setup(extra=0)
    S, B = g.stag_shape, g.basic_shape
    gr = {"vx": np.full(S, 2e-10, order="F"), "vy": np.full(S, -1e-10, order="F"),
          "vxf": np.full(S, 2e-10, order="F"), "vyf": np.full(S, -1e-10, order="F"),
          "wyx": np.zeros(B, order="F"), "tk2": np.full(S, 250.0, order="F")}
    D = DualOracle(g, mat)
    one = {"xm": np.ones(mk["xm"].size)}
    np.testing.assert_allclose(D.rk4_vjp(mk, gr, 1e11, 1, one, "xm"), 1.0, rtol=1e-15)
    np.testing.assert_allclose(D.rk4_vjp(mk, gr, 1e11, 1, one, "ym"), 0.0, atol=1e-15)


def test_loss_gradient_against_oracle_differences():
    """The multi-step loss map of the dual oracle equals the oracle's
    composed map and its exact gradient entries agree with central
    differences of it."""
    g, mat = synth.workload_geometry("cpu41", mode=1)
    mk = synth.make_markers(g, mat, seed=11)
    hrs, hrf = mat.radioactive_heating(START_TIME)
    vx, vy = synth.stream_velocity(g)
    O, D = Oracle(g, mat), DualOracle(g, mat)
    nsteps = 2
    S = g.stag_shape
    z = [np.zeros(S, order="F") for _ in range(10)]
    wyx = np.zeros(g.basic_shape, order="F")

    def loss(m0):
        mk2 = {k: np.array(v, copy=True) for k, v in m0.items()}

        def tk1_of():
            props, sums = O.marker_to_nodes(mk2, hrs, hrf, 1)
            mk2.update(props)
            op = O.p_node_properties(sums[20:30], z)
            return O.insulating_bc(op[7])

        for k in range(nsteps):
            tk2 = tk1_of()
            if k == 0:
                mk2["tkm"] = O.update_marker_temperature(mk2["xm"], mk2["ym"], mk2["tkm"],
                                                         np.zeros(S, order="F"), tk2, 1)
            r = O.move_markers_rk4(mk2["xm"], mk2["ym"], mk2["tm"], mk2["tkm"], mk2["phim"],
                                   mk2["sxym"], mk2["sxxm"], vx, vy, vx, vy, wyx, tk2, DT, 1)
            mk2["xm"], mk2["ym"], mk2["tkm"], mk2["sxym"], mk2["sxxm"] = r
        t = tk1_of()
        return float(np.sum(t * t))

    rng = np.random.default_rng(3)
    ok = np.flatnonzero((mk["tm"] < 3) & np.all(np.abs(
        mk["tkm"][:, None] - np.array([263.0, 268.0, 273.0, 278.0, 410.0])[None, :]) > 0.5,
        axis=1))
    idx = rng.choice(ok, 4, replace=False)
    seeds = [(int(m), 0) for m in idx] + [(int(m), 1) for m in idx]
    L, grad = D.tk1_loss_grad(mk, vx, vy, nsteps, DT, 1, hrs, hrf, seeds)
    assert abs(L - loss(mk)) <= 1e-12 * L
    for (m, kind), gv in zip(seeds, grad):
        f = "tkm" if kind == 0 else "phim"
        h = 1e-5 * abs(mk[f][m])

        def lk(v):
            p = dict(mk)
            p[f] = mk[f].copy()
            p[f][m] = v
            return loss(p)
        fd = _fd(lk, mk[f][m], h)
        assert abs(fd - gv) <= 1e-5 * abs(gv), (m, kind, fd, gv)
