"""Adjoint kernels: gradient of L = sum(tk1^2) after n steps of the transfer
map (scatter, node properties, tk2 := tk1, temperature gather, RK4) with
respect to the initial marker temperature and porosity.

Pins:
  - the exact gradient of the dual-number oracle (oracle/ereb_dual.cpp, a
    forward-mode restatement of the oracle's map, in place of ForwardDiff of
    the reference) at 41x41 nodes, 40,000 markers, 5 steps: every checked
    entry within 1e-10 of it;
  - central finite differences of the GPU forward map (h = 1e-6 relative),
    agreement within 1e-6 relative to the largest gradient entry checked;
  - central finite differences of the CPU oracle's forward map (the
    restatement of the reference functions, composed the same way), the
    same bar, so the adjoint is checked against the reference's arithmetic
    and not only against the GPU forward;
  - the tk2 cotangents are summed in fixed point: the gradient is the same
    bits on every run and for every checkpoint spacing.
The loss difference of a finite-difference pair is summed node by node as
sum((a - b) * (a + b)) so that it keeps its relative accuracy.
"""
from __future__ import annotations

import numpy as np
import pytest

import ereb
from ereb import synth
from oracle import DualOracle, Oracle

pytestmark = pytest.mark.gpu

START_TIME = 2.25e6 * 365.25 * 24 * 3600  # C:339
DT = ereb.DT_LONGEST
BREAKS = (263.0, 268.0, 273.0, 278.0, 410.0)  # kinks and steps of the T-dependent properties


def _setup(ctx, g, mk):
    vx, vy = synth.stream_velocity(g)
    ctx.set_markers(mk)
    for n, a in (("vx", vx), ("vy", vy), ("vxf", vx), ("vyf", vy),
                 ("wyx", np.zeros(g.basic_shape, order="F"))):
        ctx.set_grid(n, a)
    return vx, vy


def _pick(mk, n, rng):
    ok = np.flatnonzero((mk["tm"] < 3) & np.all(
        np.abs(mk["tkm"][:, None] - np.array(BREAKS)[None, :]) > 0.5, axis=1))
    return rng.choice(ok, n, replace=False)


def _dL(a, b):
    return float(np.sum((a - b) * (a + b)))


@pytest.fixture(scope="module")
def adj41():
    g, mat = synth.workload_geometry("adj41", mode=1)
    mk = synth.make_markers(g, mat, seed=7)
    hrs, hrf = mat.radioactive_heating(START_TIME)
    return g, mat, mk, hrs, hrf


def test_adjoint_vs_gpu_finite_differences(adj41):
    g, mat, mk, hrs, hrf = adj41
    nsteps, every, mode = 4, 2, 1
    rng = np.random.default_rng(1)
    idx = _pick(mk, 8, rng)
    with ereb.Context(g, mat, ereb.FP64) as ctx:
        _setup(ctx, g, mk)
        L, gT, gP, _ = ctx.adjoint_tk1_loss(nsteps, every, DT, mode, hrs, hrf)
        assert np.all(np.isfinite(gT)) and np.all(np.isfinite(gP))
        for name, grad in (("tkm", gT), ("phim", gP)):
            fd = []
            for m in idx:
                h = 1e-6 * abs(mk[name][m])
                tk = []
                for sgn in (1.0, -1.0):
                    a = mk[name].copy()
                    a[m] += sgn * h
                    ctx.set_marker(name, a)
                    tk.append(ctx.adjoint_tk1_loss(nsteps, every, DT, mode, hrs, hrf,
                                                   grad=False, want_tk1=True)[3])
                fd.append(_dL(tk[0], tk[1]) / (2 * h))
            ctx.set_marker(name, mk[name])
            fd = np.array(fd)
            scale = np.abs(grad[idx]).max()
            err = np.abs(fd - grad[idx]).max() / scale
            assert err <= 1e-6, (name, err, fd, grad[idx])


def test_adjoint_checkpoint_spacing_and_state(adj41):
    g, mat, mk, hrs, hrf = adj41
    with ereb.Context(g, mat, ereb.FP64) as ctx:
        _setup(ctx, g, mk)
        L1, gT1, gP1, _ = ctx.adjoint_tk1_loss(5, 1, DT, 1, hrs, hrf)
        L5, gT5, gP5, _ = ctx.adjoint_tk1_loss(5, 5, DT, 1, hrs, hrf)
        L2, gT2, gP2, _ = ctx.adjoint_tk1_loss(5, 2, DT, 1, hrs, hrf)
        # the markers are left as they were
        for n in ("xm", "ym", "tkm", "sxxm", "sxym"):
            assert np.array_equal(ctx.get_marker(n), mk[n]), n
    assert L1 == L5 == L2
    for a in (gT5, gT2):
        assert np.array_equal(a, gT1)
    for a in (gP5, gP2):
        assert np.array_equal(a, gP1)


def test_adjoint_grid_budget_paths_agree(adj41):
    """While every step's grid fields fit the budget, a recomputed step runs
    only its sort, temperature gather and RK4; with a budget of 0 it runs the
    scatter and the node pass again.  Same loss, tk1 and gradient bits, in
    both property modes; the markers are left as they were."""
    g, mat, mk, hrs, hrf = adj41
    with ereb.Context(g, mat, ereb.FP64) as ctx:
        _setup(ctx, g, mk)
        for mode in (1, 9):
            ctx.adjoint_set_grid_budget(4 << 30)
            La, gTa, gPa, tka = ctx.adjoint_tk1_loss(6, 2, DT, mode, hrs, hrf, want_tk1=True)
            ctx.adjoint_set_grid_budget(0)
            Lb, gTb, gPb, tkb = ctx.adjoint_tk1_loss(6, 2, DT, mode, hrs, hrf, want_tk1=True)
            assert La == Lb, mode
            assert np.array_equal(tka, tkb), mode
            assert np.array_equal(gTa, gTb) and np.array_equal(gPa, gPb), mode
            assert np.all(np.isfinite(gTa)) and np.abs(gTa).max() > 0
        for n in ("xm", "ym", "tkm", "sxxm", "sxym"):
            assert np.array_equal(ctx.get_marker(n), mk[n]), n


def test_adjoint_vs_dual_exact_gradient(adj41):
    """Per entry against the exact (dual-number) gradient of the oracle's map,
    and bit-identical on a second run."""
    g, mat, mk, hrs, hrf = adj41
    nsteps, every, mode = 5, 2, 1
    rng = np.random.default_rng(9)
    idx = _pick(mk, 48, rng)
    with ereb.Context(g, mat, ereb.FP64) as ctx:
        vx, vy = _setup(ctx, g, mk)
        L, gT, gP, _ = ctx.adjoint_tk1_loss(nsteps, every, DT, mode, hrs, hrf)
        L2, gT2, gP2, _ = ctx.adjoint_tk1_loss(nsteps, every, DT, mode, hrs, hrf)
    assert L == L2 and np.array_equal(gT, gT2) and np.array_equal(gP, gP2)
    seeds = [(int(m), 0) for m in idx] + [(int(m), 1) for m in idx]
    Ld, gd = DualOracle(g, mat).tk1_loss_grad(mk, vx, vy, nsteps, DT, mode, hrs, hrf, seeds)
    assert abs(L - Ld) <= 1e-12 * Ld
    n = len(idx)
    for name, got, ref in (("tkm", gT[idx], gd[:n]), ("phim", gP[idx], gd[n:])):
        err = np.abs(got - ref) / (np.abs(ref) + 1e-6 * np.abs(ref).max())
        assert err.max() <= 1e-10, (name, float(err.max()))


def _oracle_tk1(orc, g, mk, vx, vy, nsteps, mode, hrs, hrf):
    """The differentiated map composed from the oracle's restatements."""
    mk = {k: np.array(v, copy=True) for k, v in mk.items()}
    S = g.stag_shape
    z = [np.zeros(S, order="F") for _ in range(10)]
    wyx = np.zeros(g.basic_shape, order="F")
    DT0 = np.zeros(S, order="F")

    def tk1_of(mk):
        props, sums = orc.marker_to_nodes(mk, hrs, hrf, mode)
        mk.update(props)
        op = orc.p_node_properties(sums[20:30], z)
        return orc.insulating_bc(op[7])

    for k in range(nsteps):
        tk2 = tk1_of(mk)
        if k == 0:
            mk["tkm"] = orc.update_marker_temperature(mk["xm"], mk["ym"], mk["tkm"], DT0, tk2, 1)
        r = orc.move_markers_rk4(mk["xm"], mk["ym"], mk["tm"], mk["tkm"], mk["phim"],
                                 mk["sxym"], mk["sxxm"], vx, vy, vx, vy, wyx, tk2, DT, mode)
        mk["xm"], mk["ym"], mk["tkm"], mk["sxym"], mk["sxxm"] = r
    return tk1_of(mk)


def test_adjoint_vs_oracle_finite_differences(adj41):
    g, mat, mk, hrs, hrf = adj41
    nsteps, mode = 2, 1
    orc = Oracle(g, mat)
    rng = np.random.default_rng(2)
    idx = _pick(mk, 5, rng)
    with ereb.Context(g, mat, ereb.FP64) as ctx:
        vx, vy = _setup(ctx, g, mk)
        L, gT, gP, tk1 = ctx.adjoint_tk1_loss(nsteps, 1, DT, mode, hrs, hrf, want_tk1=True)
    # the GPU forward map is the oracle's (to rounding)
    tk1_o = _oracle_tk1(orc, g, mk, vx, vy, nsteps, mode, hrs, hrf)
    np.testing.assert_allclose(tk1, tk1_o, rtol=1e-12)
    for name, grad in (("tkm", gT), ("phim", gP)):
        fd = []
        for m in idx:
            h = 1e-6 * abs(mk[name][m])
            tk = []
            for sgn in (1.0, -1.0):
                p = dict(mk)
                p[name] = mk[name].copy()
                p[name][m] += sgn * h
                tk.append(_oracle_tk1(orc, g, p, vx, vy, nsteps, mode, hrs, hrf))
            fd.append(_dL(tk[0], tk[1]) / (2 * h))
        fd = np.array(fd)
        err = np.abs(fd - grad[idx]).max() / np.abs(grad[idx]).max()
        assert err <= 1e-6, (name, err, fd, grad[idx])
