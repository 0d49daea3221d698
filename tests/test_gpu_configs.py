"""The benchmark's larger configurations at their own workloads.

  - 20,070,400 markers on 561x561 basic nodes (8x8 per cell), planetesimal
    radius 40 and 65 km: five transfer steps of the mixed-precision form
    against the Float64 form, every SUM and node field within 1e-6 normwise
    (max|mixed - fp64| / max|fp64|); one checkpointed adjoint step with its
    gradient checked by fourth-order central differences of the GPU forward
    map on four planet markers per input, each within 1e-6 of the difference
    quotient;
  - 1,000,000 markers on 201x201 basic nodes (5x5 per cell), 20 steps with
    checkpoints every 5: the same differences on 64 planet markers per input,
    each within 1e-6 relative of its own difference quotient (a per-marker
    bar).  The markers are picked with their temperature after the first
    step's gather at least 2 K away from the kinks of the property laws, so
    that a difference step does not cross a branch of the piecewise map.
"""
from __future__ import annotations

import numpy as np
import pytest

import ereb
from ereb import precision, synth

pytestmark = pytest.mark.gpu

START_TIME = 2.25e6 * 365.25 * 24 * 3600  # C:339
DT = ereb.DT_LONGEST
BREAKS = np.array([263.0, 268.0, 273.0, 278.0, 410.0])


def _dL(a, b):
    return float(np.sum((a - b) * (a + b)))


def _setup_flow(ctx, g, mk):
    vx, vy = synth.stream_velocity(g)
    ctx.set_markers(mk)
    for n, a in (("vx", vx), ("vy", vy), ("vxf", vx), ("vyf", vy),
                 ("wyx", np.zeros(g.basic_shape, order="F"))):
        ctx.set_grid(n, a)


def _pick(ctx, g, mk, hrs, hrf, n, rng, margin):
    """Planet markers whose temperature after the first step's gather
    (tk2 := tk1, as in the differentiated map) and whose initial temperature
    both lie `margin` K or more from every kink."""
    ctx.step_scatter(hrs, hrf, 1)
    ctx.step_node_properties()
    ctx.set_grid("tk2", ctx.get_grid("tk1"))
    ctx.step_gather_temperature(1)
    T1 = ctx.get_marker("tkm")
    ctx.set_markers(mk)
    far = lambda T: np.all(np.abs(T[:, None] - BREAKS[None, :]) >= margin, axis=1)  # noqa: E731
    ok = np.flatnonzero((mk["tm"] < 3) & far(T1) & far(mk["tkm"]))
    return rng.choice(ok, n, replace=False)


# Fourth-order central differences,
#   (8 (L(x+h) - L(x-h)) - (L(x+2h) - L(x-2h))) / 12h,
# truncation error O(h^4), with relative steps h.  The quotient's rounding
# noise is ~1.5 u / h, u the relative rounding error of a loss difference:
# one marker moves tk1 through node sums of ~100 markers, so at a relative
# step of 1e-4 (second order) the change of tk1 is ~1e-8 K, ~1e5 ulps of tk1,
# and single markers measured 1e-7 to 1.4e-6 of noise, depending on the
# summation order of the node sums.  Steps of 1e-5 (tkm: 2h = 6e-3 K, far
# inside the 2 K margin from the kinks) and 1e-3 (phim, which enters
# smoothly) keep the noise near 1e-7 and the truncation below 1e-12.
FD_STEP = {"tkm": 1e-5, "phim": 1e-3}


def _fd_check(ctx, mk, idx, nsteps, every, hrs, hrf, gT, gP, bar):
    for name, grad in (("tkm", gT), ("phim", gP)):
        for m in idx:
            h = FD_STEP[name] * abs(mk[name][m])
            tk = {}
            for k in (1, -1, 2, -2):
                a = mk[name].copy()
                a[m] += k * h
                ctx.set_marker(name, a)
                tk[k] = ctx.adjoint_tk1_loss(nsteps, every, DT, 1, hrs, hrf, grad=False,
                                             want_tk1=True)[3]
            ctx.set_marker(name, mk[name])
            fd = (8.0 * _dL(tk[1], tk[-1]) - _dL(tk[2], tk[-2])) / (12.0 * h)
            assert abs(fd - grad[m]) <= bar * abs(fd), (name, int(m), fd, grad[m])


@pytest.mark.parametrize("name", ["large40", "large65"])
def test_large_mixed_form_and_adjoint_step(name):
    g, mat = synth.workload_geometry(name, mode=1)
    mk = synth.make_markers(g, mat, seed=0, stress=True)
    M = mk["xm"].size
    print(f"[{name}] {M:,} markers, {g.Nx}x{g.Ny} basic nodes", flush=True)
    hrs, hrf = mat.radioactive_heating(START_TIME)
    vx, vy = synth.stream_velocity(g)
    rng = np.random.default_rng(4)
    S = g.stag_shape
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
                c.step_transfer(hrs, hrf, 1, DT, ts)
        errs = precision.field_errors(ctxs[1], ctxs[0])
        nodal = ereb.SUM_FIELDS + precision.NODE_FIELDS
        bad = {k: errs[k][0] for k in nodal if errs[k][0] > 1e-6}
        assert not bad, bad
        print(f"[{name}] mixed form: worst nodal field "
              f"{precision.worst(errs, set(nodal))}", flush=True)
        ctxs[1].close()
        ctxs.pop()
        # one checkpointed adjoint step on the Float64 context
        ctx = ctxs[0]
        _setup_flow(ctx, g, mk)
        idx = _pick(ctx, g, mk, hrs, hrf, 4, np.random.default_rng(5), 2.0)
        L, gT, gP, _ = ctx.adjoint_tk1_loss(1, 1, DT, 1, hrs, hrf)
        assert np.all(np.isfinite(gT)) and np.all(np.isfinite(gP))
        _fd_check(ctx, mk, idx, 1, 1, hrs, hrf, gT, gP, 1e-6)
    finally:
        for c in ctxs:
            c.close()


def test_adjoint_config_workload():
    """1M markers, 201x201 nodes, 20 steps, checkpoints every 5: per-marker
    central differences on 64 markers for each input."""
    g, mat = synth.workload_geometry("adj", mode=1)
    mk = synth.make_markers(g, mat, seed=0)
    hrs, hrf = mat.radioactive_heating(START_TIME)
    with ereb.Context(g, mat, ereb.FP64) as ctx:
        _setup_flow(ctx, g, mk)
        idx = _pick(ctx, g, mk, hrs, hrf, 64, np.random.default_rng(12345), 2.0)
        L, gT, gP, _ = ctx.adjoint_tk1_loss(20, 5, DT, 1, hrs, hrf)
        L2, gT2, gP2, _ = ctx.adjoint_tk1_loss(20, 5, DT, 1, hrs, hrf)
        assert L == L2 and np.array_equal(gT, gT2) and np.array_equal(gP, gP2)
        _fd_check(ctx, mk, idx, 20, 5, hrs, hrf, gT, gP, 1e-6)
