"""GPU pullbacks (the reverse-mode kernels behind the rrules) against the
dual-number oracle (oracle/ereb_dual.cpp), entry by entry, through the C ABI:

  - marker_to_nodes pullback: every marker-field cotangent, from random SUM
    and property-output cotangents, modes 1 and 9, markers inside and
    outside the box (extrapolated weights);
  - update_marker_temperature! (timestep 1 and later) and
    update_marker_stress! pullbacks: marker and position cotangents per
    marker, grid cotangents per node (the dual oracle seeded on single
    nodes) and along random directions;
  - move_markers_rk4! pullback: the six marker-input cotangents per marker
    with vxf != vx (the E:5173 centre quirk) and wyx != 0, the six grid
    cotangents per node and along random directions;
  - the grid cotangents are summed in fixed point: two calls give the same
    bits.
Bars: per entry within 1e-10 of the dual value plus 1e-12 of the field's
largest entry (the kernels evaluate the derivative formulas in their own
operation order)."""
from __future__ import annotations

import numpy as np
import pytest

import cases
import ereb
from ereb import api
from ereb.constants import test_constants as tconsts
from oracle import DualOracle

pytestmark = pytest.mark.gpu

DT = 3e11


def _close(gpu, ref, what, rtol=1e-10, atol_rel=1e-12):
    gpu, ref = np.asarray(gpu), np.asarray(ref)
    scale = max(np.abs(ref).max(), 1e-300)
    err = np.abs(gpu - ref) - (rtol * np.abs(ref) + atol_rel * scale)
    assert np.all(err <= 0), (what, float(np.max(np.abs(gpu - ref))), scale)


@pytest.fixture(scope="module")
def setup():
    g, mat = tconsts()
    mk = cases.ref_markers(g, mat, seed=21, extra=400)
    hrs, hrf = mat.radioactive_heating(7.1e13)
    return g, mat, mk, hrs, hrf


@pytest.mark.parametrize("mode", [1, 9])
def test_marker_to_nodes_pullback_vs_dual(setup, mode):
    g, mat, mk, hrs, hrf = setup
    rng = np.random.default_rng(31 + mode)
    M = mk["xm"].size
    sum_bar = [cases.rand_grid(rng, g.basic_shape if k < 8 else g.stag_shape, -1, 1)
               for k in range(30)]
    sum_bar[13] = None  # a missing cotangent counts as zero
    prop_bar = [rng.random(M) - 0.5 for _ in range(7)]
    D = DualOracle(g, mat)
    with ereb.Context(g, mat, ereb.FP64) as ctx:
        out = api.marker_to_nodes_pullback(ctx, mk, hrs, hrf, mode, sum_bar, prop_bar)
    for f, v in out.items():
        _close(v, D.marker_to_nodes_vjp(mk, hrs, hrf, mode, sum_bar, prop_bar, f), f)


def _node_seeds(shape, rng, n=24):
    picks = rng.choice(shape[0] * shape[1], n, replace=False)
    for k in picks:
        e = np.zeros(shape[0] * shape[1])
        e[k] = 1.0
        yield k, e.reshape(shape, order="F")


def test_gather_pullbacks_vs_dual(setup):
    g, mat, mk, _, _ = setup
    rng = np.random.default_rng(41)
    S, B = g.stag_shape, g.basic_shape
    M = mk["xm"].size
    DTg, tk2 = cases.rand_grid(rng, S, -1, 1), cases.rand_grid(rng, S, 200, 300)
    DSXX, DSXY = cases.rand_grid(rng, S, -1e3, 1e3), cases.rand_grid(rng, B, -1e3, 1e3)
    tb, sxb, syb = rng.random(M) - 0.5, rng.random(M) - 0.5, rng.random(M) - 0.5
    D = DualOracle(g, mat)
    with ereb.Context(g, mat, ereb.FP64) as ctx:
        for ts in (1, 2):
            tin, dtb, tkb, xb, yb = api.update_marker_temperature_pullback(
                ctx, mk["xm"], mk["ym"], DTg, tk2, ts, M, tb)
            grid_b = tkb if ts == 1 else dtb
            assert not np.any(dtb if ts == 1 else tkb)
            args = (mk["xm"], mk["ym"], mk["tkm"], DTg, tk2, ts, tb)
            _close(xb, D.gather_T_vjp(*args, "xm"), ("T xm", ts))
            _close(yb, D.gather_T_vjp(*args, "ym"), ("T ym", ts))
            _close(tin, D.gather_T_vjp(*args, "tkm"), ("T tkm", ts))
            for k, e in _node_seeds(S, rng):
                ref = D.gather_T_vjp(*args, "grid", e).sum()
                got = grid_b.ravel(order="F")[k]
                assert abs(got - ref) <= 1e-10 * abs(ref) + 1e-12 * np.abs(grid_b).max(), (ts, k)
            gd = cases.rand_grid(rng, S, -1, 1)
            ref = D.gather_T_vjp(*args, "grid", gd).sum()
            _close([np.sum(grid_b * gd)], [ref], ("T grid direction", ts))
        bxx, bxy, xb, yb = api.update_marker_stress_pullback(ctx, mk["xm"], mk["ym"], DSXX, DSXY,
                                                             M, sxb, syb)
    args = (mk["xm"], mk["ym"], mk["sxxm"], mk["sxym"], DSXX, DSXY, sxb, syb)
    _close(xb, D.gather_S_vjp(*args, "xm"), "S xm")
    _close(yb, D.gather_S_vjp(*args, "ym"), "S ym")
    for name, gb, shape in (("DSXX", bxx, S), ("DSXY", bxy, B)):
        gd = cases.rand_grid(rng, shape, -1, 1)
        _close([np.sum(gb * gd)], [D.gather_S_vjp(*args, name, gd).sum()], name)
        for k, e in _node_seeds(shape, rng, 12):
            ref = D.gather_S_vjp(*args, name, e).sum()
            assert abs(gb.ravel(order="F")[k] - ref) <= 1e-10 * abs(ref) + 1e-12 * np.abs(gb).max()


def _rk4_case(g, rng):
    S, B = g.stag_shape, g.basic_shape
    vx, vy = cases.rand_grid(rng, S, -1e-9, 1e-9), cases.rand_grid(rng, S, -1e-9, 1e-9)
    return {"vx": vx, "vy": vy,
            "vxf": vx + cases.rand_grid(rng, S, -3e-10, 3e-10),
            "vyf": vy + cases.rand_grid(rng, S, -3e-10, 3e-10),
            "wyx": cases.rand_grid(rng, B, -1e-12, 1e-12), "tk2": cases.rand_grid(rng, S, 200, 300)}


def _rk4_pullback(ctx, mk, gr, bars, mode):
    return api.move_markers_rk4_pullback(ctx, mk["xm"], mk["ym"], mk["tm"], mk["tkm"], mk["phim"],
                                         mk["sxym"], mk["sxxm"], gr["vx"], gr["vy"], gr["vxf"],
                                         gr["vyf"], gr["wyx"], gr["tk2"], mk["xm"].size, DT, mode,
                                         bars)


@pytest.mark.parametrize("mode", [1, 9])
def test_rk4_pullback_vs_dual(setup, mode):
    g, mat, mk, _, _ = setup
    rng = np.random.default_rng(51 + mode)
    M = mk["xm"].size
    gr = _rk4_case(g, rng)
    bars = {k: rng.random(M) - 0.5 for k in ("xm", "ym", "tkm", "sxym", "sxxm")}
    bars["xm"] *= 1e-3
    bars["ym"] *= 1e-3
    D = DualOracle(g, mat)
    with ereb.Context(g, mat, ereb.FP64) as ctx:
        mo, go = _rk4_pullback(ctx, mk, gr, bars, mode)
        mo2, go2 = _rk4_pullback(ctx, mk, gr, bars, mode)
    for f in api.RK4_IN:
        _close(mo[f], D.rk4_vjp(mk, gr, DT, mode, bars, f), ("rk4", f))
        assert np.array_equal(mo[f], mo2[f]), f
    for n in api.RK4_GRIDS:
        assert np.array_equal(go[n], go2[n]), ("fixed-point grid sums differ between calls", n)
        gd = cases.rand_grid(rng, go[n].shape, -1, 1)
        ref = D.rk4_vjp(mk, gr, DT, mode, bars, n, gd).sum()
        scale = np.abs(go[n]).sum() + 1e-300
        assert abs(np.sum(go[n] * gd) - ref) <= 1e-10 * scale, (n, np.sum(go[n] * gd), ref)
        for k, e in _node_seeds(go[n].shape, rng, 10):
            ref = D.rk4_vjp(mk, gr, DT, mode, bars, n, e).sum()
            got = go[n].ravel(order="F")[k]
            assert abs(got - ref) <= 1e-10 * abs(ref) + 1e-12 * np.abs(go[n]).max(), (n, k, got,
                                                                                       ref)
