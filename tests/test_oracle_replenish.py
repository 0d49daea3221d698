"""Pins of the oracle's replenish_markers! (E:5586-5698, randomized =
false, the reference test's mode).

  (a) a port of the reference test's restatement T:6396-6598 (from
      HTM-planetary.m, distances as (dx^2+dy^2)^0.5, its own window loop
      with bounds tests inside), on its inputs' distribution: start_marknum
      markers confined to the upper-left quarter of the box, so three
      quarters of the lattice is empty; every marker vector, the count and
      mnum compared exactly (the reference test compares ten of the vectors
      with ==), mdis to an ulp (the restatement's (.)^0.5 is pow, the
      reference's distance() is sqrt);
  (b) invariants: a full lattice (one marker per anchor) adds nothing; new
      markers sit exactly on their anchors; an anchor left empty has no
      first-pass marker within its window.
"""
from __future__ import annotations

import numpy as np
import pytest

from ereb.constants import test_constants as tconsts
from ereb.lib import MARKER_FIELDS
from oracle import Oracle


@pytest.fixture(scope="module")
def tc():
    g, mat = tconsts()
    return g, mat, Oracle(g, mat)


def quarter_markers(g, seed, n=None):
    """T:6397-6424: xm in -dx:0.1:x[end]/2, every property rand"""
    rng = np.random.default_rng(seed)
    n = g.start_marknum if n is None else n
    a = g.axes()
    mk = {f: rng.random(n) for f in MARKER_FIELDS}
    mk["xm"] = -g.dx + np.round(rng.random(n) * (a["x"][-1] / 2 + g.dx) * 10) / 10
    mk["ym"] = -g.dy + np.round(rng.random(n) * (a["y"][-1] / 2 + g.dy) * 10) / 10
    mk["tm"] = rng.integers(1, 4, n).astype(np.int64)
    return mk


def replenish_ver(g, mk):
    """T:6478-6584"""
    xxm, yym = g.marker_axes()
    Nxm, Nym, dxm, dym = g.Nxm, g.Nym, g.dxm, g.dym
    xm, ym = list(mk["xm"]), list(mk["ym"])
    out = {f: list(mk[f]) for f in MARKER_FIELDS}
    mdis = np.full((Nym, Nxm), 1e30)
    mnum = np.zeros((Nym, Nxm), dtype=np.int64)
    n0 = len(xm)
    for m in range(1, n0 + 1):
        j = int(np.trunc((xm[m - 1] - xxm[0]) / dxm)) + 1
        i = int(np.trunc((ym[m - 1] - yym[0]) / dym)) + 1
        j = 1 if j < 1 else (Nxm - 1 if j > Nxm - 1 else j)
        i = 1 if i < 1 else (Nym - 1 if i > Nym - 1 else i)
        for ii, jj in ((i, j), (i + 1, j), (i, j + 1), (i + 1, j + 1)):
            d = ((xm[m - 1] - xxm[jj - 1]) ** 2 + (ym[m - 1] - yym[ii - 1]) ** 2) ** 0.5
            if d < mdis[ii - 1, jj - 1]:
                mdis[ii - 1, jj - 1] = d
                mnum[ii - 1, jj - 1] = m
    dii, djj = 5 * g.Nxmc, 5 * g.Nymc
    first = mnum.copy()
    for j in range(1, Nxm + 1):
        for i in range(1, Nym + 1):
            if mnum[i - 1, j - 1] != 0:
                continue
            # the window in the reference's scan order (jj outer, ii inner)
            j0, j1 = max(j - djj, 1), min(j + djj, Nxm)
            i0, i1 = max(i - dii, 1), min(i + dii, Nym)
            win = first[i0 - 1:i1, j0 - 1:j1].ravel(order="F")
            cand = win[win > 0]
            if cand.size:
                d = ((np.array(xm)[cand - 1] - xxm[j - 1]) ** 2 +
                     (np.array(ym)[cand - 1] - yym[i - 1]) ** 2) ** 0.5
                k = int(np.argmin(d))          # first minimum = the strict-< loop's pick
                if d[k] < mdis[i - 1, j - 1]:
                    mdis[i - 1, j - 1] = d[k]
                    mnum[i - 1, j - 1] = -cand[k]
            if mnum[i - 1, j - 1] < 0:
                m = -mnum[i - 1, j - 1]
                for f in MARKER_FIELDS:
                    out[f].append(out[f][m - 1])
                out["xm"][-1] = xxm[j - 1]
                out["ym"][-1] = yym[i - 1]
    return {f: np.array(v) for f, v in out.items()}, len(out["xm"]), mdis, mnum


def test_replenish_restatement(tc):
    g, _, orc = tc
    mk = quarter_markers(g, 31)
    got, n, mdis, mnum = orc.replenish_markers(mk)
    want, nv, mdis_v, mnum_v = replenish_ver(g, mk)
    assert n == nv > mk["xm"].size
    for f in MARKER_FIELDS:
        assert np.array_equal(got[f], want[f]), f
    assert np.array_equal(mnum, mnum_v)
    # the restatement's (.)^0.5 is pow, the reference's distance() sqrt
    # (E:1098): an ulp apart now and then
    np.testing.assert_allclose(mdis, mdis_v, rtol=4e-16, atol=0)
    # new markers sit on their anchors; anchors left empty saw no marker
    xxm, yym = g.marker_axes()
    neg = np.argwhere(mnum.T < 0)                 # (j, i) in column-major order
    assert np.array_equal(got["xm"][mk["xm"].size:], xxm[neg[:, 0]])
    assert np.array_equal(got["ym"][mk["xm"].size:], yym[neg[:, 1]])
    empty = np.argwhere(mnum == 0)
    first = np.where(mnum > 0, mnum, 0)
    for i, j in empty[:200]:
        w = first[max(i - 5 * g.Nymc, 0):i + 5 * g.Nymc + 1, max(j - 5 * g.Nxmc, 0):j + 5 * g.Nxmc + 1]
        assert not (w > 0).any()


def test_replenish_full_lattice_adds_nothing(tc):
    g, _, orc = tc
    xxm, yym = g.marker_axes()
    X, Y = np.meshgrid(xxm, yym, indexing="ij")
    rng = np.random.default_rng(3)
    n = X.size
    mk = {f: rng.random(n) for f in MARKER_FIELDS}
    mk["xm"] = X.ravel() + (rng.random(n) - 0.5) * g.dxm * 0.9
    mk["ym"] = Y.ravel() + (rng.random(n) - 0.5) * g.dym * 0.9
    mk["tm"] = rng.integers(1, 4, n).astype(np.int64)
    got, nn, mdis, mnum = orc.replenish_markers(mk)
    assert nn == n and (mnum > 0).all() and (mdis < np.hypot(g.dxm, g.dym)).all()
