"""Drop-in mirror of the reference's transfer-path functions.

Each function keeps the argument list of the Erebus.jl function it replaces
(src/Erebus.jl), with a leading ``ctx`` (the GPU context) and a trailing
underscore in place of Julia's ``!``.  Arrays are numpy arrays in Julia's
layout: marker vectors 1-D, grids Fortran-ordered ``(rows, cols)`` so that
``A[i-1, j-1]`` is Julia's ``A[i, j]``.  Outputs are written in place into
the arrays passed, as the ``!`` functions do.  Each call is one C-ABI call
into libereb_hip.so; all arithmetic runs on the GPU.
"""
from __future__ import annotations

import ctypes

import numpy as np

from .lib import OK, SYS_STOKES_DARCY, SYS_THERMAL, Context

__all__ = [
    "marker_to_nodes_", "compute_basic_node_properties_",
    "compute_vx_node_properties_", "compute_vy_node_properties_",
    "compute_p_node_properties_", "apply_insulating_boundary_conditions_",
    "update_marker_temperature_", "update_marker_stress_", "move_markers_rk4_",
    "assemble_hydromechanical_lse_", "assemble_thermal_lse_", "CSC",
    "update_marker_viscosity_", "update_marker_porosity_",
    "update_p_nodes_melt_composition_", "perform_thermochemical_reaction_",
    "backtrace_pressures_rk4_", "replenish_markers_",
]


def _io(a, dtype=np.float64, grid=False):
    """In/out array: must already have the dtype and layout (no copies)."""
    if not isinstance(a, np.ndarray) or a.dtype != dtype:
        raise TypeError(f"expected a numpy {np.dtype(dtype).name} array")
    ok = a.flags.f_contiguous if grid else a.flags.c_contiguous
    if not ok:
        raise TypeError("array must be contiguous in Julia (column-major) layout")
    return a.ctypes.data


def _in(a, dtype=np.float64, grid=False, keep=None):
    a = np.asarray(a, dtype=dtype)
    a = np.asfortranarray(a) if grid else np.ascontiguousarray(a)
    if keep is not None:
        keep.append(a)
    return a.ctypes.data


class CSC:
    """SparseMatrixCSC{Float64,Int64}: 1-based colptr/rowval as Julia."""

    def __init__(self, n, colptr, rowval, nzval):
        self.m = self.n = n
        self.colptr, self.rowval, self.nzval = colptr, rowval, nzval

    @property
    def nnz(self):
        return int(self.nzval.size)

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csc_matrix((self.nzval, self.rowval - 1, self.colptr - 1),
                             shape=(self.n, self.n))


def marker_to_nodes_(ctx: Context, xm, ym, tm, tkm, phim, XWsolidm0, etavpm,
                     inv_gggtotalm, fricttotalm, cohestotalm, tenstotalm,
                     rhofluidcur, alphasolidcur, alphafluidcur, sxxm, sxym,
                     rhototalm, rhocptotalm, etatotalm, hrtotalm, ktotalm,
                     tkm_rhocptotalm, etafluidcur_inv_kphim, hrsolidm, hrfluidm,
                     mode, sums, marknum):
    """The marker loop E:6419-6517 (compute_marker_properties! E:700 and
    marker_to_{basic,vx,vy,p}_nodes! E:1605-1863 for m = 1:marknum) after
    reset_interpolated_properties! (E:968).  ``sums`` is the list of the 30
    SUM grids in reset order; they and the 7 property vectors are written."""
    keep = []
    ins = [_in(a, np.int64 if k == 2 else np.float64, keep=keep) for k, a in enumerate(
        [xm, ym, tm, tkm, phim, XWsolidm0, etavpm, inv_gggtotalm, fricttotalm,
         cohestotalm, tenstotalm, rhofluidcur, alphasolidcur, alphafluidcur, sxxm, sxym])]
    outs = [_io(a) for a in [rhototalm, rhocptotalm, etatotalm, hrtotalm, ktotalm,
                             tkm_rhocptotalm, etafluidcur_inv_kphim]]
    if len(sums) != 30:
        raise ValueError("sums must hold the 30 SUM grids")
    sp = (ctypes.c_void_p * 30)(*[_io(s, grid=True) for s in sums])
    hs = _in(hrsolidm, keep=keep)
    hf = _in(hrfluidm, keep=keep)
    ctx.check(ctx.lib.ereb_marker_to_nodes(ctx.h, *ins, *outs, hs, hf, int(mode),
                                           ctypes.addressof(sp), int(marknum)),
              "marker_to_nodes!")


def compute_basic_node_properties_(ctx: Context, ETA0SUM, ETASUM, GGGSUM, SXYSUM,
                                   COHSUM, TENSUM, FRISUM, WTSUM, ETA0, ETA, GGG,
                                   SXY0, COH, TEN, FRI, YNY):
    """compute_basic_node_properties! (E:1970).  YNY is a uint8 (Bool) grid."""
    keep = []
    ins = [_in(a, grid=True, keep=keep) for a in
           [ETA0SUM, ETASUM, GGGSUM, SXYSUM, COHSUM, TENSUM, FRISUM, WTSUM]]
    outs = [_io(a, grid=True) for a in [ETA0, ETA, GGG, SXY0, COH, TEN, FRI]]
    ctx.check(ctx.lib.ereb_compute_basic_node_properties(
        ctx.h, *ins, *outs, _io(YNY, np.uint8, grid=True)), "compute_basic_node_properties!")


def _v_node(fn, ctx, sums, outs, what):
    keep = []
    ins = [_in(a, grid=True, keep=keep) for a in sums]
    o = [_io(a, grid=True) for a in outs]
    ctx.check(fn(ctx.h, *ins, *o), what)


def compute_vx_node_properties_(ctx: Context, RHOXSUM, RHOFXSUM, KXSUM, PHIXSUM, RXSUM,
                                WTXSUM, RHOX, RHOFX, KX, PHIX, RX):
    """compute_vx_node_properties! (E:2033)"""
    _v_node(ctx.lib.ereb_compute_vx_node_properties, ctx,
            [RHOXSUM, RHOFXSUM, KXSUM, PHIXSUM, RXSUM, WTXSUM], [RHOX, RHOFX, KX, PHIX, RX],
            "compute_vx_node_properties!")


def compute_vy_node_properties_(ctx: Context, RHOYSUM, RHOFYSUM, KYSUM, PHIYSUM, RYSUM,
                                WTYSUM, RHOY, RHOFY, KY, PHIY, RY):
    """compute_vy_node_properties! (E:2086)"""
    _v_node(ctx.lib.ereb_compute_vy_node_properties, ctx,
            [RHOYSUM, RHOFYSUM, KYSUM, PHIYSUM, RYSUM, WTYSUM], [RHOY, RHOFY, KY, PHIY, RY],
            "compute_vy_node_properties!")


def compute_p_node_properties_(ctx: Context, RHOSUM, RHOCPSUM, ALPHASUM, ALPHAFSUM, HRSUM,
                               GGGPSUM, SXXSUM, TKSUM, PHISUM, WTPSUM, RHO, RHOCP, ALPHA,
                               ALPHAF, HR, GGGP, SXX0, tk1, PHI, BETAPHI):
    """compute_p_node_properties! (E:2148)"""
    _v_node(ctx.lib.ereb_compute_p_node_properties, ctx,
            [RHOSUM, RHOCPSUM, ALPHASUM, ALPHAFSUM, HRSUM, GGGPSUM, SXXSUM, TKSUM, PHISUM,
             WTPSUM],
            [RHO, RHOCP, ALPHA, ALPHAF, HR, GGGP, SXX0, tk1, PHI, BETAPHI],
            "compute_p_node_properties!")


def apply_insulating_boundary_conditions_(ctx: Context, t):
    """apply_insulating_boundary_conditions! (E:2432)"""
    Nyy, Nxx = t.shape
    ctx.check(ctx.lib.ereb_apply_insulating_boundary_conditions(
        ctx.h, _io(t, grid=True), Nyy, Nxx), "apply_insulating_boundary_conditions!")


def update_marker_temperature_(ctx: Context, xm, ym, tkm, DT, tk2, timestep, marknum):
    """update_marker_temperature! (E:4782)"""
    keep = []
    ctx.check(ctx.lib.ereb_update_marker_temperature(
        ctx.h, _in(xm, keep=keep), _in(ym, keep=keep), _io(tkm),
        _in(DT, grid=True, keep=keep), _in(tk2, grid=True, keep=keep),
        int(timestep), int(marknum)), "update_marker_temperature!")


def update_marker_stress_(ctx: Context, xm, ym, sxxm, sxym, DSXX, DSXY, marknum):
    """update_marker_stress! (E:4354)"""
    keep = []
    ctx.check(ctx.lib.ereb_update_marker_stress(
        ctx.h, _in(xm, keep=keep), _in(ym, keep=keep), _io(sxxm), _io(sxym),
        _in(DSXX, grid=True, keep=keep), _in(DSXY, grid=True, keep=keep), int(marknum)),
        "update_marker_stress!")


def move_markers_rk4_(ctx: Context, xm, ym, tm, tkm, phim, sxym, sxxm, vx, vy, vxf, vyf,
                      wyx, tk2, marknum, dt, mode):
    """move_markers_rk4! (E:4990), argument order as the reference"""
    keep = []
    ctx.check(ctx.lib.ereb_move_markers_rk4(
        ctx.h, _io(xm), _io(ym), _in(tm, np.int64, keep=keep), _io(tkm),
        _in(phim, keep=keep), _io(sxym), _io(sxxm),
        *[_in(a, grid=True, keep=keep) for a in (vx, vy, vxf, vyf, wyx, tk2)],
        int(marknum), float(dt), int(mode)), "move_markers_rk4!")


def assemble_hydromechanical_lse_(ctx: Context, ETA, ETAP, GGG, GGGP, SXY0, SXX0, RHOX,
                                  RHOY, RHOFX, RHOFY, RX, RY, ETAPHI, BETAPHI, PHI, gx, gy,
                                  pr0, pf0, DMP, dt, R) -> CSC:
    """assemble_hydromechanical_lse! (E:3089): fills R, returns L (CSC)."""
    keep = []
    grids = [_in(a, grid=True, keep=keep) for a in
             [ETA, ETAP, GGG, GGGP, SXY0, SXX0, RHOX, RHOY, RHOFX, RHOFY, RX, RY, ETAPHI,
              BETAPHI, PHI, gx, gy, pr0, pf0, DMP]]
    nnz = ctypes.c_int64()
    ctx.check(ctx.lib.ereb_assemble_hydromechanical_lse(
        ctx.h, *grids, float(dt), _io(R), ctypes.addressof(nnz)),
        "assemble_hydromechanical_lse!")
    cp, rv, nz, _ = ctx.csc(SYS_STOKES_DARCY)
    return CSC(R.size, cp, rv, nz)


def assemble_thermal_lse_(ctx: Context, tk1, RHOCP, KX, KY, HR, HA, HS, DHP, RT, dt) -> CSC:
    """assemble_thermal_lse! (E:4526): fills RT, returns LT (CSC)."""
    keep = []
    grids = [_in(a, grid=True, keep=keep) for a in [tk1, RHOCP, KX, KY, HR, HA, HS, DHP]]
    nnz = ctypes.c_int64()
    ctx.check(ctx.lib.ereb_assemble_thermal_lse(
        ctx.h, *grids, _io(RT), float(dt), ctypes.addressof(nnz)), "assemble_thermal_lse!")
    cp, rv, nz, _ = ctx.csc(SYS_THERMAL)
    return CSC(RT.size, cp, rv, nz)


def update_marker_viscosity_(ctx: Context, xm, ym, tm, tkm, etatotalm, etavpm, YNY,
                             YNY_inv_ETA, marknum):
    """update_marker_viscosity! (E:774) for m = 1..marknum (the loop
    E:6950-6953); YNY is the Bool (Ny, Nx) yielding mask."""
    keep = []
    ctx.check(ctx.lib.ereb_update_marker_viscosity(
        ctx.h, _in(xm, keep=keep), _in(ym, keep=keep), _in(tm, np.int64, keep=keep),
        _in(tkm, keep=keep), _io(etatotalm), _io(etavpm),
        _in(YNY, np.uint8, grid=True, keep=keep), _in(YNY_inv_ETA, grid=True, keep=keep),
        int(marknum)), "update_marker_viscosity!")


def update_marker_porosity_(ctx: Context, xm, ym, tm, phim, APHI, dt, marknum, hyd=None):
    """update_marker_porosity! (E:4822); hyd: the constants of C:227-279
    (ereb.Hydration, defaults as src/constants.jl)."""
    keep = []
    ctx.check(ctx.lib.ereb_update_marker_porosity(
        ctx.h, ctx._hyd(hyd), _in(xm, keep=keep), _in(ym, keep=keep),
        _in(tm, np.int64, keep=keep), _io(phim), _in(APHI, grid=True, keep=keep), float(dt),
        int(marknum)), "update_marker_porosity!")


def update_p_nodes_melt_composition_(ctx: Context, xm, ym, XWsolidm0, XWS, XWSSUM, WTPSUM,
                                     marknum):
    """update_p_nodes_melt_composition! (E:1929)"""
    keep = []
    ctx.check(ctx.lib.ereb_update_p_nodes_melt_composition(
        ctx.h, _in(xm, keep=keep), _in(ym, keep=keep), _in(XWsolidm0, keep=keep),
        _io(XWS, grid=True), _io(XWSSUM, grid=True), _io(WTPSUM, grid=True), int(marknum)),
        "update_p_nodes_melt_composition!")


def perform_thermochemical_reaction_(ctx: Context, DMP, DHP, DMPSUM, DHPSUM, WTPSUM, pf, tk2,
                                     tm, xm, ym, XWsm0, XWsm, phim, phinewm, pfm0, marknum,
                                     dt, timestep, titer, hyd=None):
    """perform_thermochemical_reaction! (E:2558), without its @info lines"""
    keep = []
    ctx.check(ctx.lib.ereb_perform_thermochemical_reaction(
        ctx.h, ctx._hyd(hyd), _io(DMP, grid=True), _io(DHP, grid=True),
        _io(DMPSUM, grid=True), _io(DHPSUM, grid=True), _io(WTPSUM, grid=True),
        _in(pf, grid=True, keep=keep), _in(tk2, grid=True, keep=keep),
        _in(tm, np.int64, keep=keep), _in(xm, keep=keep), _in(ym, keep=keep), _io(XWsm0),
        _io(XWsm), _io(phim), _io(phinewm), _io(pfm0), int(marknum), float(dt),
        int(timestep), int(titer)), "perform_thermochemical_reaction!")


def backtrace_pressures_rk4_(ctx: Context, pr, pr0, ps, ps0, pf, pf0, vx, vy, vxf, vyf, dt):
    """backtrace_pressures_rk4! (E:5284): fills pr0, ps0, pf0"""
    keep = []
    ctx.check(ctx.lib.ereb_backtrace_pressures_rk4(
        ctx.h, _in(pr, grid=True, keep=keep), _io(pr0, grid=True),
        _in(ps, grid=True, keep=keep), _io(ps0, grid=True), _in(pf, grid=True, keep=keep),
        _io(pf0, grid=True), *[_in(a, grid=True, keep=keep) for a in (vx, vy, vxf, vyf)],
        float(dt)), "backtrace_pressures_rk4!")


# argument order of replenish_markers! (E:5586-5614)
REPLENISH_FIELDS = [
    "xm", "ym", "tm", "tkm", "phim", "sxxm", "sxym", "etavpm", "phinewm", "pfm0", "XWsolidm",
    "XWsolidm0", "rhototalm", "rhocptotalm", "etatotalm", "hrtotalm", "ktotalm",
    "inv_gggtotalm", "fricttotalm", "cohestotalm", "tenstotalm", "rhofluidcur",
    "alphasolidcur", "alphafluidcur", "tkm_rhocptotalm", "etafluidcur_inv_kphim",
]


def replenish_markers_(ctx: Context, xm, ym, tm, tkm, phim, sxxm, sxym, etavpm, phinewm, pfm0,
                       XWsolidm, XWsolidm0, rhototalm, rhocptotalm, etatotalm, hrtotalm,
                       ktotalm, inv_gggtotalm, fricttotalm, cohestotalm, tenstotalm,
                       rhofluidcur, alphasolidcur, alphafluidcur, tkm_rhocptotalm,
                       etafluidcur_inv_kphim, mdis, mnum, randomized=False, seed=0):
    """replenish_markers! (E:5586).  Julia grows the vectors with push!;
    here each argument is a numpy vector and the function returns the list
    of the 26 grown vectors (in argument order) and the new marker count,
    filling mdis and mnum ((Nym, Nxm), Fortran order) in place."""
    vecs = [xm, ym, tm, tkm, phim, sxxm, sxym, etavpm, phinewm, pfm0, XWsolidm, XWsolidm0,
            rhototalm, rhocptotalm, etatotalm, hrtotalm, ktotalm, inv_gggtotalm, fricttotalm,
            cohestotalm, tenstotalm, rhofluidcur, alphasolidcur, alphafluidcur,
            tkm_rhocptotalm, etafluidcur_inv_kphim]
    n0 = len(xm)
    g = ctx.geom
    cap = n0 + g.Nxm * g.Nym  # at most one new marker per anchor
    buf = []
    for name, v in zip(REPLENISH_FIELDS, vecs):
        dt = np.int64 if name == "tm" else np.float64
        b = np.zeros(cap, dtype=dt)
        b[:n0] = v
        buf.append(b)
    mg = g.marker_grid_c()
    n = ctypes.c_int64()
    ctx.check(ctx.lib.ereb_replenish_markers(
        ctx.h, *[b.ctypes.data for b in buf], _io(mdis, grid=True),
        _io(mnum, np.int64, grid=True), n0, cap, ctypes.addressof(mg), int(bool(randomized)),
        int(seed), ctypes.addressof(n)), "replenish_markers!")
    return [b[:n.value].copy() for b in buf], n.value


# ---- pullbacks (the reverse-mode kernels behind ChainRulesCore rrules) -------

def _opt_in(a, n, keep, grid=False):
    return None if a is None else _in(a, grid=grid, keep=keep)


def _opt_out(want, shape, grid=False):
    if not want:
        return None
    return np.zeros(shape, order="F") if grid else np.zeros(shape)


def marker_to_nodes_pullback(ctx: Context, markers: dict, hrsolidm, hrfluidm, mode,
                             sum_bar, prop_bar=None, want=None):
    """Pullback of marker_to_nodes_ (compute_marker_properties! E:700 +
    marker_to_*_nodes! E:1605-1863): sum_bar is a list of 30 SUM-grid
    cotangents (None = 0), prop_bar a list of 7 property-output cotangents or
    None.  Returns {field: cotangent} for the marker fields in `want`
    (default: every differentiable input)."""
    from .lib import MARKER_FIELDS
    keep = []
    M = markers["xm"].size
    mptr = (ctypes.c_void_p * len(MARKER_FIELDS))()
    for k, f in enumerate(MARKER_FIELDS):
        if f in markers:
            mptr[k] = _in(markers[f], np.int64 if f == "tm" else np.float64, keep=keep)
    sb = (ctypes.c_void_p * 30)(*[None if a is None else _in(a, grid=True, keep=keep)
                                  for a in sum_bar])
    pb = None if prop_bar is None else (ctypes.c_void_p * 7)(
        *[None if a is None else _in(a, keep=keep) for a in prop_bar])
    want = want or ["xm", "ym", "tkm", "phim", "XWsolidm0", "sxxm", "sxym", "etavpm",
                    "inv_gggtotalm", "fricttotalm", "cohestotalm", "tenstotalm", "rhofluidcur",
                    "alphasolidcur", "alphafluidcur", "rhototalm", "rhocptotalm", "etatotalm",
                    "hrtotalm", "ktotalm"]
    out = {f: np.zeros(M) for f in want}
    optr = (ctypes.c_void_p * len(MARKER_FIELDS))(
        *[out[f].ctypes.data if f in out else None for f in MARKER_FIELDS])
    hs = np.ascontiguousarray(hrsolidm, dtype=np.float64)
    hf = np.ascontiguousarray(hrfluidm, dtype=np.float64)
    ctx.check(ctx.lib.ereb_marker_to_nodes_pullback(ctx.h, mptr, M, hs.ctypes.data,
                                                    hf.ctypes.data, int(mode), sb, pb, optr),
              "marker_to_nodes pullback")
    return out


def update_marker_temperature_pullback(ctx: Context, xm, ym, DT, tk2, timestep, marknum,
                                       tkm_bar):
    """Pullback of update_marker_temperature! (E:4782): returns (tkm_bar_in,
    DT_bar, tk2_bar, xm_bar, ym_bar)."""
    keep = []
    M = int(marknum)
    g = ctx.grid_shape("tk2")
    outs = [np.zeros(M), np.zeros(g, order="F"), np.zeros(g, order="F"), np.zeros(M),
            np.zeros(M)]
    ctx.check(ctx.lib.ereb_update_marker_temperature_pullback(
        ctx.h, _in(xm, keep=keep), _in(ym, keep=keep), _opt_in(DT, M, keep, grid=True),
        _opt_in(tk2, M, keep, grid=True), int(timestep), M, _in(tkm_bar, keep=keep),
        *[o.ctypes.data for o in outs]), "update_marker_temperature pullback")
    return tuple(outs)


def update_marker_stress_pullback(ctx: Context, xm, ym, DSXX, DSXY, marknum, sxxm_bar,
                                  sxym_bar):
    """Pullback of update_marker_stress! (E:4354): returns (DSXX_bar,
    DSXY_bar, xm_bar, ym_bar); the stress cotangents pass through."""
    keep = []
    M = int(marknum)
    outs = [np.zeros(ctx.grid_shape("DSXX"), order="F"),
            np.zeros(ctx.grid_shape("DSXY"), order="F"), np.zeros(M), np.zeros(M)]
    ctx.check(ctx.lib.ereb_update_marker_stress_pullback(
        ctx.h, _in(xm, keep=keep), _in(ym, keep=keep), _in(DSXX, grid=True, keep=keep),
        _in(DSXY, grid=True, keep=keep), M, _in(sxxm_bar, keep=keep), _in(sxym_bar, keep=keep),
        *[o.ctypes.data for o in outs]), "update_marker_stress pullback")
    return tuple(outs)


RK4_IN = ["xm", "ym", "tkm", "phim", "sxym", "sxxm"]
RK4_GRIDS = ["vx", "vy", "vxf", "vyf", "wyx", "tk2"]


def move_markers_rk4_pullback(ctx: Context, xm, ym, tm, tkm, phim, sxym, sxxm, vx, vy, vxf,
                              vyf, wyx, tk2, marknum, dt, mode, bars: dict, grids=True):
    """Pullback of move_markers_rk4! (E:4990): bars holds the cotangents of
    the outputs (xm, ym, tkm, sxym, sxxm; missing = 0).  Returns
    ({input marker field: cotangent}, {grid: cotangent} or {})."""
    keep = []
    M = int(marknum)
    b = [_opt_in(bars.get(f), M, keep) for f in ("xm", "ym", "tkm", "sxym", "sxxm")]
    mo = {f: np.zeros(M) for f in RK4_IN}
    go = {}
    if grids:
        for n in RK4_GRIDS:
            go[n] = np.zeros(ctx.grid_shape(n), order="F")
    ctx.check(ctx.lib.ereb_move_markers_rk4_pullback(
        ctx.h, _in(xm, keep=keep), _in(ym, keep=keep), _in(tm, np.int64, keep=keep),
        _in(tkm, keep=keep), _in(phim, keep=keep), _in(sxym, keep=keep), _in(sxxm, keep=keep),
        *[_in(a, grid=True, keep=keep) for a in (vx, vy, vxf, vyf, wyx, tk2)], M, float(dt),
        int(mode), *b, *[mo[f].ctypes.data for f in RK4_IN],
        *[go[n].ctypes.data if n in go else None for n in RK4_GRIDS]), "move_markers_rk4 pullback")
    return mo, go
