"""The device side of save_state (E:5718-5957): ereb_save_state_fields fills
every array of the record that the context holds in one call, bit for bit
equal to the per-field reads (ereb_grid_get, ereb_yny_get,
ereb_markers_get), on a state that a transfer step, the hydration rows and
random grids have filled (so no two fields hold the same values), in both
precision forms (the mixed form's FP32 property vectors and int32 tm are
widened on the device), for all fields at once, a subset, and fewer markers
than the resident count."""
from __future__ import annotations

import numpy as np
import pytest

import cases
import ereb
from ereb import lib as L
from ereb.constants import test_constants as tconsts

pytestmark = pytest.mark.gpu


def _state(ctx, g, mat, mk, seed):
    rng = np.random.default_rng(seed)
    ctx.set_markers(mk)
    for name in L.GRID_FIELDS:
        ctx.set_grid(name, cases.rand_grid(rng, ctx.grid_shape(name), -1, 1))
    ctx.set_yny(rng.random((g.Ny, g.Nx)) < 0.5)
    hrs, hrf = mat.radioactive_heating(7.1e13)
    ctx.step_scatter(hrs, hrf, 1)
    ctx.step_node_properties()
    ctx.step_marker_viscosity()


def _expected(ctx, table, n):
    exp = {}
    for key, kind, field in table:
        if kind == L.SAVE_GRID:
            exp[key] = ctx.get_grid(L.GRID_FIELDS[field])
        elif kind == L.SAVE_YNY:
            exp[key] = ctx.get_yny()
        else:
            exp[key] = ctx.get_marker(L.MARKER_FIELDS[field], n)
    return exp


@pytest.mark.parametrize("precision", [ereb.FP64, ereb.MIXED])
def test_save_state_fields_equal_per_field_reads(precision):
    g, mat = tconsts()
    mk = cases.ref_markers(g, mat, seed=5)
    mk["XWsolidm"] = np.full(mk["xm"].size, 0.3)
    mk["phinewm"] = np.full(mk["xm"].size, 0.1)
    mk["pfm0"] = np.full(mk["xm"].size, 1e5)
    table = L.save_fields()
    with ereb.Context(g, mat, precision) as ctx:
        _state(ctx, g, mat, mk, 17)
        M = ctx.marknum
        for n in (M, M // 3):
            got = ctx.save_state_fields(marknum=n)
            exp = _expected(ctx, table, n)
            assert list(got) == [k for k, _, _ in table]
            for key in exp:
                assert got[key].dtype == exp[key].dtype and got[key].shape == exp[key].shape, key
                assert np.array_equal(got[key], exp[key], equal_nan=True), key
        sub = ["tk1", "YNY", "tm", "xm", "etatotalm", "XWS"]
        got = ctx.save_state_fields(sub)
        assert sorted(got) == sorted(sub)
        exp = _expected(ctx, [t for t in table if t[0] in sub], M)
        for key in sub:
            assert np.array_equal(got[key], exp[key]), key
        with pytest.raises(KeyError):
            ctx.save_state_fields(["EXY"])  # produced by the host solves
