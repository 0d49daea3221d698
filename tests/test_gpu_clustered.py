"""GPU parity on crowded, empty and degenerate marker sets, through the C ABI.

Markers pile up where the flow converges; the code paths that only crowded
cells reach are held to the bars of the uniform parity tests:
  - a fine cell of 20,000 markers: the sort orders it by a bitonic sort in
    global memory (more than 16,384 markers), the scatter walks its column in
    chunks of 256 staged markers, the P-grid scatter reads the tile's sort
    order from memory (more than 4,096 markers in the tile);
  - a fine cell of 3,000 markers (bitonic sort in LDS), one of 30 (ordered
    by one thread), 400 markers spread over one basic column (a column split
    into chunks across many fine cells);
  - markers exactly on basic and P node lines, on the box edges and outside
    the box (the clamps of the cell index, C:66-101).
Bars: fine keys, sort order and marker properties bit-exact; Float64 node
sums within 1e-12 of the per-node sum of |terms|; mixed-form node sums within
1e-6 of the field's largest sum of |terms| (max|error| <= 1e-6 max|terms|
sums, per field): the normwise bar of the mixed form, max|error| / max|sum|,
wherever the terms share a sign.  The random-sign stresses of 20,000 markers
in one cell cancel, which the plain normwise bar does not allow for
(SXXSUM's error is 1.8e-6 of its largest sum); a per-node bar does not fit
the mixed form either, whose float stencil fractions carry an absolute
error of ~6e-8, large against a node's own sum where it gets only small
weights.  Melt composition and the thermochemical sums as in the uniform
tests.  No
marker and a single marker run every step without error and give the
oracle's sums.
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

START_TIME = 2.25e6 * 365.25 * 24 * 3600  # C:339
PROPS = ["rhototalm", "rhocptotalm", "etatotalm", "hrtotalm", "ktotalm", "tkm_rhocptotalm",
         "etafluidcur_inv_kphim"]


def _exact(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    same = (a == b) | (np.isnan(a) & np.isnan(b))
    if not same.all():
        k = np.flatnonzero(~same.ravel())[:5]
        raise AssertionError(f"{what}: {(~same).sum()} of {a.size} differ, e.g. "
                             f"{a.ravel()[k]} vs {b.ravel()[k]}")


def _box(rng, n, x0, y0, h):
    """n markers uniform in the square of half-width h around (x0, y0)."""
    return x0 + h * (2 * rng.random(n) - 1), y0 + h * (2 * rng.random(n) - 1)


def clustered_markers(g, mat, seed=5):
    """Lattice markers, 2,000 anywhere in the extended box, and the crowded
    and degenerate sets of the module docstring (1 km cells, 14 x 14 cells,
    tiles of 6 x 6 cells)."""
    rng = np.random.default_rng(seed)
    d = g.dx
    parts = [
        _box(rng, 20_000, 2 * d + d / 4, 3 * d + d / 4, 0.1 * d),   # cell (4, 3), lower quadrant
        _box(rng, 3_000, 9 * d + 3 * d / 4, 10 * d + 3 * d / 4, 0.1 * d),
        _box(rng, 30, 12 * d + d / 4, 1 * d + 3 * d / 4, 0.1 * d),
        (7 * d + 0.98 * d * rng.random(400) + 0.01 * d, g.ysize * rng.random(400)),
    ]
    k = rng.integers(0, g.Nx, 150).astype(float)
    parts.append((k * d, g.ysize * rng.random(150)))             # on basic node columns
    parts.append((g.xsize * rng.random(150), k * d))             # on basic node rows
    parts.append(((k - 0.5) * d, (rng.integers(0, g.Ny, 150) + 0.5) * d))  # on P nodes
    parts.append((np.array([0.0, g.xsize, 0.0, g.xsize, -d, g.xsize + d]),
                  np.array([0.0, 0.0, g.ysize, g.ysize, 0.5 * d, 0.5 * d])))
    xc = np.concatenate([p[0] for p in parts])
    yc = np.concatenate([p[1] for p in parts])
    mk = cases.ref_markers(g, mat, seed=seed, extra=2_000 + xc.size)
    mk["xm"][-xc.size:] = xc
    mk["ym"][-yc.size:] = yc
    M = mk["xm"].size
    mk["XWsolidm"] = rng.random(M)
    mk["phinewm"] = rng.random(M)
    mk["pfm0"] = 2700 + 600 * rng.random(M)
    return mk


@pytest.fixture(scope="module")
def crowded():
    g, mat = tconsts()
    mk = clustered_markers(g, mat)
    # the crowded paths are reached: one fine key holds more than 16,384
    # markers, one between 49 and 16,384, one between 17 and 48
    _, key, _ = Oracle(g, mat).fine_sort(mk["xm"], mk["ym"])
    counts = np.bincount(np.asarray(key, dtype=np.int64))
    assert counts.max() > 16_384
    assert ((counts > 48) & (counts <= 16_384)).any()
    assert ((counts > 16) & (counts <= 48)).any()
    return g, mat, mk


def _scatter(ctx, mk, hrs, hrf, mode, shapes):
    out = {k: mk[k].copy() for k in PROPS}
    sums = [np.zeros(s, order="F") for s in shapes]
    api.marker_to_nodes_(ctx, mk["xm"], mk["ym"], mk["tm"], mk["tkm"], mk["phim"],
                         mk["XWsolidm0"], mk["etavpm"], mk["inv_gggtotalm"], mk["fricttotalm"],
                         mk["cohestotalm"], mk["tenstotalm"], mk["rhofluidcur"],
                         mk["alphasolidcur"], mk["alphafluidcur"], mk["sxxm"], mk["sxym"],
                         out["rhototalm"], out["rhocptotalm"], out["etatotalm"],
                         out["hrtotalm"], out["ktotalm"], out["tkm_rhocptotalm"],
                         out["etafluidcur_inv_kphim"], hrs, hrf, mode, sums, mk["xm"].size)
    return out, sums


@pytest.mark.parametrize("mode", [1, 9])
def test_crowded_sort_and_scatter(crowded, mode):
    g, mat, mk = crowded
    hrs, hrf = mat.radioactive_heating(START_TIME)
    orc = Oracle(g, mat)
    props_o, sums_o = orc.marker_to_nodes(mk, hrs, hrf, mode)
    perm_o, key_o, _ = orc.fine_sort(mk["xm"], mk["ym"])
    with ereb.Context(g, mat, ereb.FP64) as ctx:
        shapes = [s.shape for s in sums_o]
        out, sums = _scatter(ctx, mk, hrs, hrf, mode, shapes)
        perm, key, _ = ctx.sort_order()
        _, again = _scatter(ctx, mk, hrs, hrf, mode, shapes)
    _exact(key, key_o, "fine keys")
    _exact(perm, perm_o, "sort order")
    for k in PROPS:
        _exact(out[k], props_o[k], k)
    absb = cases.abs_sums(g, mk, props_o)
    for n, a, b, c, a2 in zip(ereb.SUM_FIELDS, sums, sums_o, absb, again):
        _exact(a, a2, f"run-to-run {n}")
        err = np.abs(a - b)
        bad = err > 1e-12 * c + 1e-300
        assert not bad.any(), (n, int(bad.sum()), float(err.max()))
    M = mk["xm"].size
    for wt in ("WTSUM", "WTXSUM", "WTYSUM", "WTPSUM"):
        assert abs(sums[ereb.SUM_FIELDS.index(wt)].sum() - M) <= 1e-9 * M, wt


def test_crowded_scatter_mixed_form(crowded):
    g, mat, mk = crowded
    hrs, hrf = mat.radioactive_heating(START_TIME)
    props_o, sums_o = Oracle(g, mat).marker_to_nodes(mk, hrs, hrf, 1)
    with ereb.Context(g, mat, ereb.MIXED) as ctx:
        ctx.set_markers(mk)
        ctx.step_scatter(hrs, hrf, 1)
        got = [ctx.get_grid(n) for n in ereb.SUM_FIELDS]
    absb = cases.abs_sums(g, mk, props_o)
    for n, a, b, c in zip(ereb.SUM_FIELDS, got, sums_o, absb):
        err = np.abs(a - b).max() / max(c.max(), 1e-300)
        assert err <= 1e-6, (n, err)


def test_crowded_melt_composition_and_reaction(crowded):
    g, mat, mk = crowded
    orc = Oracle(g, mat)
    M = mk["xm"].size
    S = (g.Ny1, g.Nx1)
    XWS_o, XS_o, WT_o = orc.melt_composition(mk["xm"], mk["ym"], mk["XWsolidm0"])
    XWS, XS, WT = (np.zeros(S, order="F") for _ in range(3))
    rng = np.random.default_rng(11)
    pf = np.asfortranarray(2700 + 600 * rng.random(S))
    tk2 = np.asfortranarray(263 + 47 * rng.random(S))
    H = Hydration(rate_mode=9)  # past the reaction time: well conditioned
    dt = 1.5 * H.dtreaction
    io_o, DMP_o, DHP_o, DMS_o, DHS_o, WTS_o = orc.thermochemical_reaction(mk, pf, tk2, dt, 2, 3,
                                                                          hyd=H)
    names = ["XWsolidm0", "XWsolidm", "phim", "phinewm", "pfm0"]
    io = {k: mk[k].copy() for k in names}
    DMP, DHP, DMS, DHS, WTS = (np.zeros(S, order="F") for _ in range(5))
    with ereb.Context(g, mat, ereb.FP64) as ctx:
        api.update_p_nodes_melt_composition_(ctx, mk["xm"], mk["ym"], mk["XWsolidm0"], XWS,
                                             XS, WT, M)
        api.perform_thermochemical_reaction_(ctx, DMP, DHP, DMS, DHS, WTS, pf, tk2, mk["tm"],
                                             mk["xm"], mk["ym"], io["XWsolidm0"],
                                             io["XWsolidm"], io["phim"], io["phinewm"],
                                             io["pfm0"], M, dt, 2, 3, hyd=H)
    # positive terms: the sum of |terms| is the sum itself
    assert np.all(np.abs(WT - WT_o) <= 1e-12 * WT_o + 1e-300)
    assert np.all(np.abs(XS - XS_o) <= 1e-12 * XS_o + 1e-300)
    np.testing.assert_allclose(XWS, XWS_o, rtol=2e-12, atol=0)
    _exact(io["pfm0"], io_o["pfm0"], "pfm0")
    for k in names:
        np.testing.assert_allclose(io[k], io_o[k], rtol=1e-12, err_msg=k)
    assert np.all(np.abs(WTS - WTS_o) <= 1e-12 * WTS_o + 1e-300)
    # the bound of the uniform test: WTPSUM times the largest node average
    for a, b, avg, k in ((DMS, DMS_o, DMP_o, "DMPSUM"), (DHS, DHS_o, DHP_o, "DHPSUM")):
        bound = 1e-12 * WTS_o * np.abs(avg).max() + 1e-300
        assert np.all(np.abs(a - b) <= bound), k


@pytest.mark.parametrize("M", [0, 1])
def test_no_marker_and_one_marker(M):
    g, mat = tconsts()
    mk = cases.ref_markers(g, mat, seed=8, extra=0, jitter=True)
    pick = np.arange(M) * 977 % mk["xm"].size
    mk = {k: np.ascontiguousarray(v[pick]) for k, v in mk.items()}
    hrs, hrf = mat.radioactive_heating(START_TIME)
    props_o, sums_o = Oracle(g, mat).marker_to_nodes(mk, hrs, hrf, 1)
    S = (g.Ny1, g.Nx1)
    vx, vy = ereb.synth.stream_velocity(g)
    wyx = ereb.synth.rotation_rate(g, vx, vy)
    tk2 = np.asfortranarray(np.full(S, 300.0))
    with ereb.Context(g, mat, ereb.FP64) as ctx:
        out, sums = _scatter(ctx, mk, hrs, hrf, 1, [s.shape for s in sums_o])
        ctx.step_node_properties()
        xm, ym, tkm = mk["xm"].copy(), mk["ym"].copy(), mk["tkm"].copy()
        sxy, sxx = mk["sxym"].copy(), mk["sxxm"].copy()
        api.move_markers_rk4_(ctx, xm, ym, mk["tm"], tkm, mk["phim"], sxy, sxx, vx, vy, vx, vy,
                              wyx, tk2, M, ereb.DT_LONGEST, 1)
        XWS, XS, WT = (np.zeros(S, order="F") for _ in range(3))
        api.update_p_nodes_melt_composition_(ctx, mk["xm"], mk["ym"], mk["XWsolidm0"], XWS,
                                             XS, WT, M)
    for k in PROPS:
        _exact(out[k], props_o[k], k)
    for n, a, b in zip(ereb.SUM_FIELDS, sums, sums_o):
        np.testing.assert_allclose(a, b, rtol=1e-12, atol=0, err_msg=n)
    assert abs(WT.sum() - M) <= 1e-12
    if M == 0:
        assert all(not a.any() for a in sums)
    else:
        # one marker moves by dt times a velocity of at most max|v|
        vmax = max(np.abs(vx).max(), np.abs(vy).max())
        assert np.hypot(xm - mk["xm"], ym - mk["ym"]).max() <= 2 * ereb.DT_LONGEST * vmax
