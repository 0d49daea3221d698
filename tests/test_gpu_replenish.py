"""GPU parity of replenish_markers! (E:5586-5698) against the CPU oracle
through the C ABI: the appended markers, their order, every copied marker
vector, the new count and the lattice arrays mdis and mnum, bit for bit,
in the reference test's mode (randomized = false); the jittered mode keeps
the same picks and places each new marker within half a lattice cell of its
anchor, reproducibly for a given seed."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import ereb
from ereb import api, synth
from ereb.constants import test_constants as tconsts
from ereb.lib import MARKER_FIELDS
from ereb.precision import MARKER_PROPS
from oracle import Oracle
from test_oracle_replenish import quarter_markers

pytestmark = pytest.mark.gpu


def _exact(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, (what, a.shape, b.shape)
    d = a != b
    assert not d.any(), f"{what}: {d.sum()} of {a.size} differ"


# marker vectors stored as float in the mixed form
F32_FIELDS = ["XWsolidm0"] + MARKER_FIELDS[MARKER_FIELDS.index("etavpm"):
                                           MARKER_FIELDS.index("etafluidcur_inv_kphim") + 1]


def _args(mk):
    return [mk[f] for f in api.REPLENISH_FIELDS]


@pytest.mark.parametrize("precision", [ereb.FP64, ereb.MIXED])
def test_replenish_dropin_vs_oracle(precision):
    g, mat = tconsts()
    mk = quarter_markers(g, 31)
    want, n, mdis_o, mnum_o = Oracle(g, mat).replenish_markers(mk)
    mdis = np.zeros((g.Nym, g.Nxm), order="F")
    mnum = np.zeros((g.Nym, g.Nxm), dtype=np.int64, order="F")
    with ereb.Context(g, mat, precision) as ctx:
        out, nn = api.replenish_markers_(ctx, *_args(mk), mdis, mnum)
    assert nn == n
    for name, v in zip(api.REPLENISH_FIELDS, out):
        w = want[name]
        if precision == ereb.MIXED and name in F32_FIELDS:  # stored as float
            w = w.astype(np.float32).astype(np.float64)
        _exact(v, w, name)
    _exact(mnum, mnum_o, "mnum")
    _exact(mdis, mdis_o, "mdis")


def test_replenish_randomized_and_seeded():
    g, mat = tconsts()
    mk = quarter_markers(g, 32)
    want, n, _, mnum_o = Oracle(g, mat).replenish_markers(mk)
    xxm, yym = g.marker_axes()
    neg = np.argwhere(mnum_o.T < 0)
    runs = []
    with ereb.Context(g, mat, ereb.FP64) as ctx:
        for seed in (7, 7, 8):
            ctx.set_markers(mk)
            nn, _, mnum = ctx.step_replenish_markers(randomized=True, seed=seed)
            assert nn == n
            _exact(mnum, mnum_o, "mnum")
            runs.append({f: ctx.get_marker(f) for f in MARKER_FIELDS})
    n0 = mk["xm"].size
    for f in MARKER_FIELDS:
        if f not in ("xm", "ym"):
            _exact(runs[0][f], want[f], f)
    dx = runs[0]["xm"][n0:] - xxm[neg[:, 0]]
    dy = runs[0]["ym"][n0:] - yym[neg[:, 1]]
    assert np.all(np.abs(dx) <= 0.5 * g.dxm) and np.all(np.abs(dy) <= 0.5 * g.dym)
    assert np.std(dx) > 0.2 * g.dxm and np.std(dy) > 0.2 * g.dym
    _exact(runs[0]["xm"], runs[1]["xm"], "same seed")
    assert not np.array_equal(runs[0]["xm"][n0:], runs[2]["xm"][n0:])


def test_replenish_resident_workload_hole():
    """2.82M markers of the 281x281-node workload with a 30x20-cell hole:
    the resident arrays grow and the result equals the oracle's."""
    g, mat = synth.workload_geometry("fp64", mode=1)
    mk = synth.make_markers(g, mat, seed=0)
    rng = np.random.default_rng(1)
    for f in ("XWsolidm", "phinewm", "pfm0"):
        mk[f] = rng.random(mk["xm"].size)
    hole = ((mk["xm"] > 40e3) & (mk["xm"] < 55e3) & (mk["ym"] > 60e3) & (mk["ym"] < 70e3))
    mk = {f: np.ascontiguousarray(v[~hole]) for f, v in mk.items()}
    want, n, mdis_o, mnum_o = Oracle(g, mat).replenish_markers(mk)
    assert n > mk["xm"].size
    with ereb.Context(g, mat, ereb.FP64) as ctx:
        ctx.set_markers(mk)
        nn, mdis, mnum = ctx.step_replenish_markers()
        assert nn == n == ctx.marknum
        for f in MARKER_FIELDS:
            _exact(ctx.get_marker(f), want[f], f)
    _exact(mnum, mnum_o, "mnum")
    _exact(mdis, mdis_o, "mdis")


NEW_MAX = 4096  # most empty anchors that the listed append handles


@pytest.mark.parametrize("spare,hole_km,move", [(0, (15, 10), 0.6), (10_000, (15, 10), 0.6),
                                                (10_000, (1.5, 1.0), 0.1)])
def test_replenish_after_advection_in_sort_order(spare, hole_km, move):
    """The timestep's order: a scatter sorts the markers, the advection moves
    them (by up to `move` cells), then the replenishment visits them in the
    now stale sort order.  The result equals the oracle's on the moved
    markers, bit for bit.  A 15 x 10 km hole and moves of 0.6 cells leave
    tens of thousands of empty anchors, more than the listed append takes
    (the flag scan over all anchors runs); a 1.5 x 1 km hole and moves of
    0.1 cells with spare capacity for 10,000 markers is the case of every
    later timestep: about a thousand empty anchors, appended from the
    compacted list ahead of the host's read of the count."""
    g, mat = synth.workload_geometry("fp64", mode=1)
    mk = synth.make_markers(g, mat, seed=2)
    rng = np.random.default_rng(3)
    for f in ("XWsolidm", "phinewm", "pfm0"):
        mk[f] = rng.random(mk["xm"].size)
    hw, hh = hole_km[0] * 1e3, hole_km[1] * 1e3
    hole = ((mk["xm"] > 40e3) & (mk["xm"] < 40e3 + hw) & (mk["ym"] > 60e3)
            & (mk["ym"] < 60e3 + hh))
    mk = {f: np.ascontiguousarray(v[~hole]) for f, v in mk.items()}
    M = mk["xm"].size
    moved = dict(mk)
    moved["xm"] = mk["xm"] + move * g.dx * (2 * rng.random(M) - 1)
    moved["ym"] = mk["ym"] + move * g.dy * (2 * rng.random(M) - 1)
    hrs, hrf = mat.radioactive_heating(2.25e6 * 365.25 * 24 * 3600)
    with ereb.Context(g, mat, ereb.FP64) as ctx:
        if spare:
            ctx.check(ctx.lib.ereb_ctx_reserve(ctx.h, M + spare))
        ctx.set_markers(mk)
        ctx.step_scatter(hrs, hrf, 1)
        # the scatter's key pass rewrote the marker properties
        for f in MARKER_PROPS:
            moved[f] = ctx.get_marker(f)
        want, n, mdis_o, mnum_o = Oracle(g, mat).replenish_markers(moved)
        assert n > M
        listed = spare >= NEW_MAX and n - M <= NEW_MAX
        assert listed == (hole_km[0] < 10), "case does not reach the intended append path"
        ctx.set_marker("xm", moved["xm"])
        ctx.set_marker("ym", moved["ym"])
        nn, mdis, mnum = ctx.step_replenish_markers()
        assert nn == n == ctx.marknum
        for f in MARKER_FIELDS:
            _exact(ctx.get_marker(f), want[f], f)
    _exact(mnum, mnum_o, "mnum")
    _exact(mdis, mdis_o, "mdis")


def test_replenish_dropin_capacity_too_small():
    g, mat = tconsts()
    mk = quarter_markers(g, 33)
    want, n, _, _ = Oracle(g, mat).replenish_markers(mk)
    n0 = mk["xm"].size
    bufs = [np.array(mk[f], dtype=np.int64 if f == "tm" else np.float64)
            for f in api.REPLENISH_FIELDS]
    before = [b.copy() for b in bufs]
    out = ctypes.c_int64()
    mg = g.marker_grid_c()
    with ereb.Context(g, mat, ereb.FP64) as ctx:
        rc = ctx.lib.ereb_replenish_markers(ctx.h, *[b.ctypes.data for b in bufs], None, None,
                                            n0, n0, ctypes.addressof(mg), 0, 0,
                                            ctypes.addressof(out))
        assert rc == ereb.lib.ERR_ARG and out.value == n
        for b, c in zip(bufs, before):
            _exact(b, c, "host arrays untouched")
        assert ctx.marknum == n
        _exact(ctx.get_marker("xm"), want["xm"], "resident result")
