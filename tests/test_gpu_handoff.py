"""GPU: the solver hand-off of the assembled systems (E:6719-6749, E:6925).

The reference hands L to LinearSolve's UMFPACK cache at every plastic
iteration; UMFPACK keeps its symbolic analysis while the pattern is
unchanged.  The library reports whether the pattern of the last assembly
differs from the one last fetched, and fetches values alone.  Bars: the
change flag equals the oracle's pattern comparison of the same two
assemblies; values-only and page-locked fetches are bit-identical to the
full fetch.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import cases
import ereb
from ereb import api
from ereb.constants import test_constants as tconsts
from oracle import Oracle

pytestmark = pytest.mark.gpu


def _same_pattern(a, b):
    return np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def _assemble(ctx, which, grids, R, dt):
    """assemble_thermal_lse! / assemble_hydromechanical_lse! through the C
    ABI, leaving the matrix on the device (no fetch)."""
    keep = []
    ptrs = [api._in(a, grid=True, keep=keep) for a in grids]
    nnz = ctypes.c_int64()
    if which == ereb.SYS_THERMAL:
        rc = ctx.lib.ereb_assemble_thermal_lse(ctx.h, *ptrs, api._io(R), float(dt),
                                               ctypes.addressof(nnz))
    else:
        rc = ctx.lib.ereb_assemble_hydromechanical_lse(ctx.h, *ptrs, float(dt), api._io(R),
                                                       ctypes.addressof(nnz))
    ctx.check(rc, "assemble")
    return nnz.value


def test_thermal_pattern_change_follows_oracle():
    g, mat = tconsts()
    orc = Oracle(g, mat)
    dt = 3e10
    grids = cases.thermal_inputs(g)
    ref0 = orc.assemble_thermal(grids, dt)
    # a zero conductivity at one interior x face drops two off-diagonals
    grids2 = [a.copy() for a in grids]
    grids2[2][g.Ny1 // 2, g.Nx1 // 2] = 0.0
    ref2 = orc.assemble_thermal(grids2, dt)
    assert not _same_pattern(ref0, ref2)
    # only the values change: same pattern
    grids3 = [a.copy() for a in grids]
    grids3[1] *= 1.5
    ref3 = orc.assemble_thermal(grids3, dt)
    assert _same_pattern(ref0, ref3)
    RT = np.zeros(g.Ny1 * g.Nx1)
    with ereb.Context(g, mat, ereb.FP64) as ctx:
        _assemble(ctx, ereb.SYS_THERMAL, grids, RT, dt)
        assert ctx.csc_pattern_changed(ereb.SYS_THERMAL)  # nothing fetched yet
        full = ctx.csc(ereb.SYS_THERMAL)
        for k in range(3):
            np.testing.assert_array_equal(full[k], ref0[k])
        assert not ctx.csc_pattern_changed(ereb.SYS_THERMAL)
        for inputs, ref in ((grids3, ref3), (grids2, ref2), (grids2, ref2), (grids, ref0)):
            prev = full
            assert _assemble(ctx, ereb.SYS_THERMAL, inputs, RT, dt) == ref[2].size
            changed = ctx.csc_pattern_changed(ereb.SYS_THERMAL)
            assert changed == (not _same_pattern(prev, ref)), "pattern flag vs oracle"
            vals = ctx.csc(ereb.SYS_THERMAL, pattern=False)
            assert vals[0] is None and vals[1] is None
            np.testing.assert_array_equal(vals[2], ref[2])
            np.testing.assert_array_equal(vals[3], ref[3])
            # a values-only fetch leaves the fetched pattern as it was
            assert ctx.csc_pattern_changed(ereb.SYS_THERMAL) == changed
            full = ctx.csc(ereb.SYS_THERMAL)
            for k in range(3):
                np.testing.assert_array_equal(full[k], ref[k])
            assert not ctx.csc_pattern_changed(ereb.SYS_THERMAL)


def test_stokes_darcy_values_only_and_pinned_fetch():
    g, mat = tconsts()
    grids = cases.stokes_inputs(g)
    dt = 1e11
    cp, rv, nz, R = Oracle(g, mat).assemble_hydromechanical(grids, dt)
    Rg = np.zeros(R.size)
    with ereb.Context(g, mat, ereb.FP64) as ctx:
        api.assemble_hydromechanical_lse_(ctx, *grids, dt, Rg)  # fetches the pattern
        assert not ctx.csc_pattern_changed(ereb.SYS_STOKES_DARCY)
        # assembling the same inputs again keeps the pattern
        _assemble(ctx, ereb.SYS_STOKES_DARCY, grids, Rg, dt)
        assert not ctx.csc_pattern_changed(ereb.SYS_STOKES_DARCY)
        n, nnz = ctx.csc_size(ereb.SYS_STOKES_DARCY)
        assert (n, nnz) == (R.size, nz.size)
        pinned = (ctx.host_array(n + 1, np.int64), ctx.host_array(nnz, np.int64),
                  ctx.host_array(nnz, np.float64), ctx.host_array(n, np.float64))
        got = ctx.csc(ereb.SYS_STOKES_DARCY, out=pinned)
        for a, b, what in zip(got, (cp, rv, nz, R), ("colptr", "rowval", "nzval", "R")):
            np.testing.assert_array_equal(a, b, err_msg=what)
        pinned[2][:] = 0.0
        got = ctx.csc(ereb.SYS_STOKES_DARCY, pattern=False, out=(None, None) + pinned[2:])
        np.testing.assert_array_equal(got[2], nz)
        np.testing.assert_array_equal(got[3], R)
        with pytest.raises(ValueError):
            ctx.csc(ereb.SYS_STOKES_DARCY, out=(None, None, pinned[3], pinned[3]))
        # new Darcy coefficients (diagonal entries only): the pattern is
        # unchanged and the assembly refills nzval alone (the values-only
        # path); then a pattern change takes the full path again
        g2 = [a.copy() for a in grids]
        g2[10] = g2[10] * 1.75  # RX
        ref2 = Oracle(g, mat).assemble_hydromechanical(g2, dt)
        assert _same_pattern(ref2[:2], (cp, rv))
        for inputs, ref in ((g2, ref2), (grids, (cp, rv, nz, R)), (g2, ref2)):
            _assemble(ctx, ereb.SYS_STOKES_DARCY, inputs, Rg, dt)
            got = ctx.csc(ereb.SYS_STOKES_DARCY)
            for a, b, what in zip(got, ref, ("colptr", "rowval", "nzval", "R")):
                np.testing.assert_array_equal(a, b, err_msg=what)


def test_pattern_query_before_assembly_is_an_error():
    g, mat = tconsts()
    with ereb.Context(g, mat, ereb.FP64) as ctx:
        with pytest.raises(ereb.EREBError):
            ctx.csc_pattern_changed(ereb.SYS_THERMAL)
        assert ctx.host_array(0, np.float64).size == 0
