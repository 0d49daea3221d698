"""ctypes loader of the CPU oracle (oracle/ereb_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by the CPU-baseline leg of bench.py, always as the checker or the CPU
baseline, never as the product path.  Every wrapper takes numpy arrays in
Julia layout (Fortran-ordered grids) and follows the reference function
named in its docstring (E: = src/Erebus.jl).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# EREB_ORACLE_BUILD selects another build of the same sources (the tests'
# sanitizer run sets it to "_build_san")
_BUILD = os.path.join(HERE, os.environ.get("EREB_ORACLE_BUILD", "_build"))
LIB = os.path.join(_BUILD, "libereb_oracle.so")
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "erebus.jl_amd"))

from ereb.constants import CGeometry, CMaterials, Geometry, Hydration, Materials  # noqa: E402
from ereb.lib import MARKER_FIELDS  # noqa: E402

_P = ctypes.c_void_p
_D = ctypes.c_double
_I64 = ctypes.c_int64
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        sig = {
            "orc_total": (_D, [_D, _D, _D]),
            "orc_ktotal": (_D, [_D, _D, _D]),
            "orc_kphi": (_D, [_D, _D, _D]),
            "orc_etaf_inv_kphi": (_D, [_D, _D, _D, _D]),
            "orc_compute_rhocpfluidm": (_D, [_P, _D, _I64, _P]),
            "orc_compute_ksolidm": (_D, [_P, _D, _I64, _P]),
            "orc_compute_kfluidm": (_D, [_P, _D, _I64, _P]),
            "orc_etatotal_rocks": (_D, [_P, _D, _I64]),
            "orc_q_radiogenic": (_D, [_D, _D, _D, _D, _D]),
            "orc_fix": (None, [_D, _D, _P, _P, _D, _D, _I64, _I64, _I64, _I64, _P, _P]),
            "orc_fix_distances": (None, [_D, _D, _P, _P, _D, _D, _I64, _I64, _I64, _I64,
                                         _P, _P, _P, _P]),
            "orc_fix_weights": (None, [_D, _D, _P, _P, _D, _D, _I64, _I64, _I64, _I64,
                                       _P, _P, _P]),
            "orc_compute_marker_properties": (ctypes.c_int, [_P, _I64] + [_P] * 13 + [_I64]),
            "orc_marker_to_nodes": (ctypes.c_int, [_P, _P] + [_P] * 23 + [_P, _P, _I64, _P,
                                                                         _I64]),
            "orc_marker_to_nodes_abs": (ctypes.c_int, [_P, _P] + [_P] * 23 + [_P, _P, _I64, _P,
                                                                             _I64]),
            "orc_compute_basic_node_properties": (None, [_P] * 17),
            "orc_compute_v_node_properties": (None, [_P] * 12),
            "orc_compute_p_node_properties": (None, [_P] * 21),
            "orc_apply_insulating_bc": (None, [_P, _I64, _I64]),
            "orc_update_marker_temperature": (None, [_P] * 6 + [_I64, _I64]),
            "orc_update_marker_stress": (None, [_P] * 7 + [_I64]),
            "orc_move_markers_rk4": (ctypes.c_int, [_P, _P] + [_P] * 13 + [_I64, _D, _I64]),
            "orc_assemble_hydromechanical": (_I64, [_P, _P] + [_P] * 20 + [_D, _P, _P, _P, _P,
                                                                           _I64]),
            "orc_assemble_thermal": (_I64, [_P] + [_P] * 8 + [_P, _D, _P, _P, _P, _I64]),
            "orc_fine_keys": (None, [_P, _I64, _P, _P, _I64, _P]),
            "orc_key_sort": (None, [_P, _I64, _I64, _P, _P]),
            "orc_compute_rotation_rate": (None, [_P, _P, _P, _P]),
            "orc_update_marker_viscosity": (None, [_P, _P] + [_P] * 8 + [_I64]),
            "orc_update_marker_porosity": (None, [_P, _P] + [_P] * 5 + [_D, _I64]),
            "orc_update_p_nodes_melt_composition": (None, [_P] + [_P] * 6 + [_I64]),
            "orc_dtreaction": (_D, [_P, _D, _D, _I64, _P]),
            "orc_gibbs_free_energy": (_D, [_P, _D, _D, _D, _D, _D, _D]),
            "orc_reaction_constant": (_D, [_P, _D, _D, _D]),
            "orc_relative_enthalpy": (_D, [_P, _P, _D, _D]),
            "orc_perform_thermochemical_reaction": (ctypes.c_int, [_P, _P, _P] + [_P] * 15 +
                                                    [_I64, _D, _I64, _I64]),
            "orc_backtrace_pressures_rk4": (None, [_P] + [_P] * 10 + [_D]),
            "orc_replenish_markers": (_I64, [_P, _P, _I64, _I64, _P, _P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def _c(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


def _F(a):
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


class Oracle:
    """Oracle bound to one geometry and material set."""

    def __init__(self, geom: Geometry, mat: Materials):
        self.L = lib()
        self.geom = geom
        self.mat = mat
        self.cg = geom.to_c()
        self.cm = mat.to_c(geom.xsize, geom.ysize)
        self.pg = ctypes.addressof(self.cg)
        self.pm = ctypes.addressof(self.cm)

    # ---- scalar helpers -----------------------------------------------------
    def fix(self, x, y, grid):
        ax, ay, lim = self._grid(grid)
        i, j = ctypes.c_int64(), ctypes.c_int64()
        self.L.orc_fix(x, y, _p(ax), _p(ay), self.geom.dx, self.geom.dy, *lim,
                       ctypes.addressof(i), ctypes.addressof(j))
        return i.value, j.value

    def fix_distances(self, x, y, grid):
        ax, ay, lim = self._grid(grid)
        i, j, a, b = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_double(), ctypes.c_double()
        self.L.orc_fix_distances(x, y, _p(ax), _p(ay), self.geom.dx, self.geom.dy, *lim,
                                 ctypes.addressof(i), ctypes.addressof(j),
                                 ctypes.addressof(a), ctypes.addressof(b))
        return i.value, j.value, a.value, b.value

    def fix_weights(self, x, y, grid):
        ax, ay, lim = self._grid(grid)
        i, j = ctypes.c_int64(), ctypes.c_int64()
        w = np.zeros(4)
        self.L.orc_fix_weights(x, y, _p(ax), _p(ay), self.geom.dx, self.geom.dy, *lim,
                               ctypes.addressof(i), ctypes.addressof(j), _p(w))
        return i.value, j.value, w

    def _grid(self, grid):
        a = self.geom.axes()
        Nx, Ny = self.geom.Nx, self.geom.Ny
        return {
            "basic": (a["x"], a["y"], (1, Nx - 1, 1, Ny - 1)),
            "vx": (a["xvx"], a["yvx"], (1, Nx - 1, 1, Ny)),
            "vy": (a["xvy"], a["yvy"], (1, Nx, 1, Ny - 1)),
            "p": (a["xp"], a["yp"], (1, Nx, 1, Ny)),
            "p_stress": (a["xp"], a["yp"], (2, Nx - 1, 2, Ny - 1)),
        }[grid]

    def rhocpfluidm(self, T, mode):
        return self.L.orc_compute_rhocpfluidm(self.pm, T, mode, None)

    def ksolidm(self, T, mode):
        return self.L.orc_compute_ksolidm(self.pm, T, mode, None)

    def kfluidm(self, T, mode):
        return self.L.orc_compute_kfluidm(self.pm, T, mode, None)

    # ---- step 1 + 2 ---------------------------------------------------------
    def marker_to_nodes(self, mk: dict, hrs, hrf, mode, absval: bool = False):
        """Loop E:6419-6517; returns (props dict, list of 30 SUM grids).  With
        absval, the grids hold the per-node sums of |value * weight| instead
        (the scale of a sum taken in another order)."""
        g = self.geom
        M = mk["xm"].size
        ins = [_c(mk[k], np.int64 if k == "tm" else np.float64) for k in
               ["xm", "ym", "tm", "tkm", "phim", "XWsolidm0", "etavpm", "inv_gggtotalm",
                "fricttotalm", "cohestotalm", "tenstotalm", "rhofluidcur", "alphasolidcur",
                "alphafluidcur", "sxxm", "sxym"]]
        names = ["rhototalm", "rhocptotalm", "etatotalm", "hrtotalm", "ktotalm",
                 "tkm_rhocptotalm", "etafluidcur_inv_kphim"]
        outs = [np.array(mk.get(n, np.zeros(M)), dtype=np.float64, copy=True) for n in names]
        sums = [np.zeros((g.Ny, g.Nx), order="F") for _ in range(8)] + \
               [np.zeros((g.Ny1, g.Nx1), order="F") for _ in range(22)]
        sp = (ctypes.c_void_p * 30)(*[s.ctypes.data for s in sums])
        hs, hf = _c(hrs), _c(hrf)
        f = self.L.orc_marker_to_nodes_abs if absval else self.L.orc_marker_to_nodes
        rc = f(self.pg, self.pm, *[_p(a) for a in ins], *[_p(a) for a in outs], _p(hs), _p(hf),
               int(mode), ctypes.addressof(sp), M)
        if rc != 0:
            raise ValueError(f"unknown mode {mode}")
        return dict(zip(names, outs)), sums

    # ---- node properties ------------------------------------------------------
    def basic_node_properties(self, sums8, outs7, YNY):
        s = [_F(a) for a in sums8]
        o = [np.array(a, dtype=np.float64, order="F", copy=True) for a in outs7]
        y = np.array(YNY, dtype=np.uint8, order="F", copy=True)
        self.L.orc_compute_basic_node_properties(self.pg, *[_p(a) for a in s],
                                                 *[_p(a) for a in o], _p(y))
        return o, y

    def v_node_properties(self, sums6, outs5):
        s = [_F(a) for a in sums6]
        o = [np.array(a, dtype=np.float64, order="F", copy=True) for a in outs5]
        self.L.orc_compute_v_node_properties(self.pg, *[_p(a) for a in s], *[_p(a) for a in o])
        return o

    def p_node_properties(self, sums10, outs10):
        s = [_F(a) for a in sums10]
        o = [np.array(a, dtype=np.float64, order="F", copy=True) for a in outs10]
        self.L.orc_compute_p_node_properties(self.pg, *[_p(a) for a in s], *[_p(a) for a in o])
        return o

    def insulating_bc(self, t):
        t = np.array(t, dtype=np.float64, order="F", copy=True)
        self.L.orc_apply_insulating_bc(_p(t), t.shape[0], t.shape[1])
        return t

    # ---- step 3 / 4 -------------------------------------------------------------
    def update_marker_temperature(self, xm, ym, tkm, DT, tk2, timestep):
        t = np.array(tkm, dtype=np.float64, copy=True)
        xm, ym, DT, tk2 = _c(xm), _c(ym), _F(DT), _F(tk2)
        self.L.orc_update_marker_temperature(self.pg, _p(xm), _p(ym), _p(t), _p(DT), _p(tk2),
                                             int(timestep), t.size)
        return t

    def update_marker_stress(self, xm, ym, sxxm, sxym, DSXX, DSXY):
        sx = np.array(sxxm, dtype=np.float64, copy=True)
        sy = np.array(sxym, dtype=np.float64, copy=True)
        xm, ym, DSXX, DSXY = _c(xm), _c(ym), _F(DSXX), _F(DSXY)
        self.L.orc_update_marker_stress(self.pg, _p(xm), _p(ym), _p(sx), _p(sy), _p(DSXX),
                                        _p(DSXY), sx.size)
        return sx, sy

    def move_markers_rk4(self, xm, ym, tm, tkm, phim, sxym, sxxm, vx, vy, vxf, vyf, wyx, tk2,
                         dt, mode):
        xm = np.array(xm, dtype=np.float64, copy=True)
        ym = np.array(ym, dtype=np.float64, copy=True)
        tkm = np.array(tkm, dtype=np.float64, copy=True)
        sxym = np.array(sxym, dtype=np.float64, copy=True)
        sxxm = np.array(sxxm, dtype=np.float64, copy=True)
        tm = _c(tm, np.int64)
        phim = _c(phim)
        grids = [_F(a) for a in (vx, vy, vxf, vyf, wyx, tk2)]
        rc = self.L.orc_move_markers_rk4(self.pg, self.pm, _p(xm), _p(ym), _p(tm), _p(tkm),
                                         _p(phim), _p(sxym), _p(sxxm),
                                         *[_p(a) for a in grids], xm.size, float(dt), int(mode))
        if rc != 0:
            raise ValueError(f"unknown mode {mode}")
        return xm, ym, tkm, sxym, sxxm

    # ---- step 5 -----------------------------------------------------------------
    def assemble_hydromechanical(self, grids20, dt):
        g = self.geom
        n = 6 * g.Nx1 * g.Ny1
        cap = 40 * g.Nx1 * g.Ny1
        gs = [_F(a) for a in grids20]
        R = np.zeros(n)
        cp = np.zeros(n + 1, dtype=np.int64)
        rv = np.zeros(cap, dtype=np.int64)
        nz = np.zeros(cap)
        nnz = self.L.orc_assemble_hydromechanical(self.pg, self.pm, *[_p(a) for a in gs],
                                                  float(dt), _p(R), _p(cp), _p(rv), _p(nz), cap)
        assert nnz >= 0
        return cp, rv[:nnz].copy(), nz[:nnz].copy(), R

    def assemble_thermal(self, grids8, dt):
        g = self.geom
        n = g.Nx1 * g.Ny1
        cap = 5 * n
        gs = [_F(a) for a in grids8]
        R = np.zeros(n)
        cp = np.zeros(n + 1, dtype=np.int64)
        rv = np.zeros(cap, dtype=np.int64)
        nz = np.zeros(cap)
        nnz = self.L.orc_assemble_thermal(self.pg, *[_p(a) for a in gs], _p(R), float(dt),
                                          _p(cp), _p(rv), _p(nz), cap)
        assert nnz >= 0
        return cp, rv[:nnz].copy(), nz[:nnz].copy(), R

    # ---- sort specification -------------------------------------------------------
    def fine_sort(self, xm, ym, tile=6):
        """The library's sort specification: stable counting sort by fine key
        (see orc_fine_keys); returns perm, key, start."""
        g = self.geom
        xm, ym = _c(xm), _c(ym)
        key = np.zeros(xm.size, dtype=np.int32)
        self.L.orc_fine_keys(self.pg, tile, _p(xm), _p(ym), xm.size, _p(key))
        nty = (g.Ny - 1 + tile - 1) // tile
        ntx = (g.Nx - 1 + tile - 1) // tile
        nkeys = ntx * nty * (2 * tile) ** 2
        perm = np.zeros(xm.size, dtype=np.int32)
        start = np.zeros(nkeys + 1, dtype=np.int64)
        self.L.orc_key_sort(_p(key), xm.size, nkeys, _p(perm), _p(start))
        return perm, key, start

    # ---- viscosity, compaction and (de)hydration -----------------------------------
    def _h(self, hyd):
        self._hc = (Hydration() if hyd is None else hyd).to_c(self.mat)
        return ctypes.addressof(self._hc)

    def update_marker_viscosity(self, xm, ym, tm, tkm, etatotalm, YNY, YNY_inv_ETA):
        """update_marker_viscosity! (E:774) for every marker: (etatotalm, etavpm)"""
        et = np.array(etatotalm, dtype=np.float64, copy=True)
        ev = np.zeros(et.size)
        xm, ym, tkm = _c(xm), _c(ym), _c(tkm)
        tm = _c(tm, np.int64)
        Y = np.asfortranarray(np.asarray(YNY, dtype=np.uint8))
        YI = _F(YNY_inv_ETA)
        self.L.orc_update_marker_viscosity(self.pg, self.pm, _p(xm), _p(ym), _p(tm), _p(tkm),
                                           _p(et), _p(ev), _p(Y), _p(YI), et.size)
        return et, ev

    def update_marker_porosity(self, xm, ym, tm, phim, APHI, dt, hyd=None):
        """update_marker_porosity! (E:4822): new phim"""
        ph = np.array(phim, dtype=np.float64, copy=True)
        xm, ym, A = _c(xm), _c(ym), _F(APHI)
        tm = _c(tm, np.int64)
        self.L.orc_update_marker_porosity(self.pg, self._h(hyd), _p(xm), _p(ym), _p(tm),
                                          _p(ph), _p(A), float(dt), ph.size)
        return ph

    def melt_composition(self, xm, ym, XWsolidm0):
        """update_p_nodes_melt_composition! (E:1929): (XWS, XWSSUM, WTPSUM)"""
        g = self.geom
        xm, ym, xw = _c(xm), _c(ym), _c(XWsolidm0)
        out = [np.zeros((g.Ny1, g.Nx1), order="F") for _ in range(3)]
        self.L.orc_update_p_nodes_melt_composition(self.pg, _p(xm), _p(ym), _p(xw),
                                                   *[_p(a) for a in out], xm.size)
        return tuple(out)

    def dtreaction(self, T, phi, mode, hyd=None):
        err = ctypes.c_int(0)
        v = self.L.orc_dtreaction(self._h(hyd), T, phi, int(mode), ctypes.addressof(err))
        if err.value:
            raise ValueError(f"unknown mode {mode}")
        return v

    def gibbs_free_energy(self, T, pf, XDs, XWs, dt, dtr, hyd=None):
        return self.L.orc_gibbs_free_energy(self._h(hyd), T, pf, XDs, XWs, dt, dtr)

    def reaction_constant(self, T, pf, dG, hyd=None):
        return self.L.orc_reaction_constant(self._h(hyd), T, pf, dG)

    def relative_enthalpy(self, Xs, XWs, hyd=None):
        return self.L.orc_relative_enthalpy(self.pm, self._h(hyd), Xs, XWs)

    def thermochemical_reaction(self, mk: dict, pf, tk2, dt, timestep, titer, hyd=None):
        """perform_thermochemical_reaction! (E:2558).  mk holds tm, xm, ym,
        XWsolidm0, XWsolidm, phim, phinewm, pfm0; returns (dict of the five
        updated marker vectors, DMP, DHP, DMPSUM, DHPSUM, WTPSUM)."""
        g = self.geom
        names = ["XWsolidm0", "XWsolidm", "phim", "phinewm", "pfm0"]
        io = {n: np.array(mk[n], dtype=np.float64, copy=True) for n in names}
        xm, ym = _c(mk["xm"]), _c(mk["ym"])
        tm = _c(mk["tm"], np.int64)
        grids = [np.zeros((g.Ny1, g.Nx1), order="F") for _ in range(5)]
        pf, tk2 = _F(pf), _F(tk2)
        rc = self.L.orc_perform_thermochemical_reaction(
            self.pg, self.pm, self._h(hyd), *[_p(a) for a in grids], _p(pf), _p(tk2), _p(tm),
            _p(xm), _p(ym), *[_p(io[n]) for n in names], xm.size, float(dt), int(timestep),
            int(titer))
        if rc != 0:
            raise ValueError("unknown mode")
        return (io, *grids)

    def backtrace_pressures(self, pr, ps, pf, vx, vy, vxf, vyf, dt):
        """backtrace_pressures_rk4! (E:5284): (pr0, ps0, pf0)"""
        g = self.geom
        ins = [_F(a) for a in (pr, ps, pf, vx, vy, vxf, vyf)]
        out = [np.zeros((g.Ny1, g.Nx1), order="F") for _ in range(3)]
        self.L.orc_backtrace_pressures_rk4(self.pg, _p(ins[0]), _p(out[0]), _p(ins[1]),
                                           _p(out[1]), _p(ins[2]), _p(out[2]),
                                           *[_p(a) for a in ins[3:]], float(dt))
        return tuple(out)

    def replenish_markers(self, mk: dict):
        """replenish_markers! (E:5586) with randomized = false: (dict of the
        grown marker vectors, new marknum, mdis, mnum)"""
        g = self.geom
        n0 = mk["xm"].size
        cap = n0 + g.Nxm * g.Nym
        buf = {}
        for f in MARKER_FIELDS:
            dt = np.int64 if f == "tm" else np.float64
            b = np.zeros(cap, dtype=dt)
            b[:n0] = mk[f]
            buf[f] = b
        ptrs = (ctypes.c_void_p * len(MARKER_FIELDS))(*[buf[f].ctypes.data for f in MARKER_FIELDS])
        mg = g.marker_grid_c()
        mdis = np.zeros((g.Nym, g.Nxm), order="F")
        mnum = np.zeros((g.Nym, g.Nxm), dtype=np.int64, order="F")
        n = self.L.orc_replenish_markers(ctypes.addressof(mg), ctypes.addressof(ptrs), n0, cap,
                                         _p(mdis), _p(mnum))
        assert n >= 0
        return {f: buf[f][:n].copy() for f in MARKER_FIELDS}, n, mdis, mnum

    def rotation_rate(self, vx, vy):
        g = self.geom
        w = np.zeros((g.Ny, g.Nx), order="F")
        vx, vy = _F(vx), _F(vy)
        self.L.orc_compute_rotation_rate(self.pg, _p(vx), _p(vy), _p(w))
        return w


BENCH_LIB = os.path.join(_BUILD, "libereb_oracle_bench.so")
# marker vectors of orc_bench_transfer, in its order (ereb_cpu_step.c)
BENCH_MARKERS = ["xm", "ym", "tkm", "phim", "sxxm", "sxym", "XWsolidm0", "etavpm",
                 "inv_gggtotalm", "fricttotalm", "cohestotalm", "tenstotalm", "rhofluidcur",
                 "alphasolidcur", "alphafluidcur", "rhototalm", "rhocptotalm", "etatotalm",
                 "hrtotalm", "ktotalm", "tkm_rhocptotalm", "etafluidcur_inv_kphim"]
BENCH_GRIDS = ["vx", "vy", "wyx", "tk2", "DT", "DSXX", "DSXY", "ETAP", "ETAPHI", "gx", "gy"]


def bench_transfer(geom: Geometry, mat: Materials, mk: dict, grids: dict, hrs, hrf, mode: int,
                   dt: float, max_steps: int, budget_s: float):
    """CPU baseline: whole transfer timesteps of the C restatement at -O3,
    with OpenMP in the loops that the reference threads, timed inside C
    (ereb_cpu_step.c).  `mk` is updated in place.  Returns (steps done,
    seconds, threads)."""
    if not os.path.exists(BENCH_LIB):
        build()
    L = ctypes.CDLL(BENCH_LIB)
    L.orc_bench_transfer.restype = _I64
    L.orc_bench_transfer.argtypes = [_P, _P, _P, _P, _I64, _P, _P, _P, _I64, _D, _I64, _D, _P]
    L.orc_bench_threads.restype = ctypes.c_int
    cg = geom.to_c()
    cm = mat.to_c(geom.xsize, geom.ysize)
    M = mk["xm"].size
    for f in BENCH_MARKERS:
        assert mk[f].dtype == np.float64 and mk[f].flags.c_contiguous and mk[f].size == M, f
    tm = _c(mk["tm"], np.int64)
    mptrs = (ctypes.c_void_p * len(BENCH_MARKERS))(*[mk[f].ctypes.data for f in BENCH_MARKERS])
    gs = [_F(grids[n]) for n in BENCH_GRIDS] + [np.zeros(geom.stag_shape, order="F")]
    gptrs = (ctypes.c_void_p * len(gs))(*[a.ctypes.data for a in gs])
    hs, hf = _c(hrs), _c(hrf)
    sec = ctypes.c_double()
    done = L.orc_bench_transfer(ctypes.addressof(cg), ctypes.addressof(cm), mptrs, _p(tm), M,
                                gptrs, _p(hs), _p(hf), int(mode), float(dt), int(max_steps),
                                float(budget_s), ctypes.byref(sec))
    if done < 0:
        raise RuntimeError(f"orc_bench_transfer failed with status {-done}")
    return int(done), sec.value, int(L.orc_bench_threads())


DUAL_LIB = os.path.join(_BUILD, "libereb_oracle_dual.so")
_dual = None

# marker seeds of the dual functions
RK4_SEEDS = {"xm": 0, "ym": 1, "tkm": 2, "phim": 3, "sxym": 4, "sxxm": 5,
             "vx": 6, "vy": 7, "vxf": 8, "vyf": 9, "wyx": 10, "tk2": 11}
GATHER_S_SEEDS = {"xm": 0, "ym": 1, "sxxm": 2, "sxym": 3, "DSXX": 4, "DSXY": 5}
GATHER_T_SEEDS = {"xm": 0, "ym": 1, "tkm": 2, "grid": 3}


def dual_lib():
    global _dual
    if _dual is None:
        if not os.path.exists(DUAL_LIB):
            build()
        L = ctypes.CDLL(DUAL_LIB)
        ci = ctypes.c_int
        sig = {
            "dual_marker_to_nodes_vjp": (ci, [_P, _P, _P, _I64, _P, _P, _I64, _P, _P, ci, _P]),
            "dual_marker_to_nodes_dot": (ci, [_P, _P, _P, _I64, _P, _P, _I64, _P, _P, _P]),
            "dual_gather_T_vjp": (ci, [_P, _P, _P, _P, _P, _P, _I64, _I64, _P, ci, _P, _P]),
            "dual_gather_S_vjp": (ci, [_P] * 7 + [_I64, _P, _P, ci, _P, _P]),
            "dual_rk4_vjp": (ci, [_P, _P] + [_P] * 13 + [_I64, _D, _I64] + [_P] * 5 +
                             [ci, _P, _P]),
            "dual_rk4_double": (ci, [_P, _P] + [_P] * 13 + [_I64, _D, _I64]),
            "dual_tk1_loss_grad": (ci, [_P, _P, _P, _I64, _P, _P, _I64, _D, _I64, _P, _P,
                                        _I64, _P, _P, _P, _P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _dual = L
    return _dual


class DualOracle:
    """Exact (forward-mode) derivatives of the oracle's transfer functions
    (oracle/ereb_dual.cpp), bound to one geometry and material set."""

    def __init__(self, geom: Geometry, mat: Materials):
        self.L = dual_lib()
        self.geom = geom
        self.cg = geom.to_c()
        self.cm = mat.to_c(geom.xsize, geom.ysize)
        self.pg = ctypes.addressof(self.cg)
        self.pm = ctypes.addressof(self.cm)

    def _markers(self, mk):
        keep = {}
        ptrs = (ctypes.c_void_p * len(MARKER_FIELDS))()
        for k, f in enumerate(MARKER_FIELDS):
            if f in mk:
                a = _c(mk[f], np.int64 if f == "tm" else np.float64)
                keep[f] = a
                ptrs[k] = a.ctypes.data
        return ptrs, keep

    @staticmethod
    def _grid_ptrs(arrs, n):
        keep = [None if a is None else _F(a) for a in arrs]
        ptrs = (ctypes.c_void_p * n)(*[None if a is None else a.ctypes.data for a in keep])
        return ptrs, keep

    def marker_to_nodes_vjp(self, mk, hrs, hrf, mode, sum_bar, prop_bar, field):
        """d/d(field of marker m) of sum_bar . (marker m's SUM contributions)
        + prop_bar . (marker m's properties), for every marker."""
        ptrs, keep = self._markers(mk)
        sb, ks = self._grid_ptrs(sum_bar, 30)
        pb_keep = None if prop_bar is None else [None if a is None else _c(a) for a in prop_bar]
        pb = None if prop_bar is None else (ctypes.c_void_p * 7)(
            *[None if a is None else a.ctypes.data for a in pb_keep])
        M = mk["xm"].size
        out = np.empty(M)
        hs, hf = _c(hrs), _c(hrf)
        rc = self.L.dual_marker_to_nodes_vjp(self.pg, self.pm, ptrs, M, _p(hs), _p(hf),
                                             int(mode), sb, pb, MARKER_FIELDS.index(field),
                                             _p(out))
        assert rc == 0, rc
        return out

    def marker_to_nodes_dot(self, mk, hrs, hrf, mode, sum_bar):
        """(per-marker sum_bar . contributions, the 7 property vectors) with
        the template at S = double."""
        ptrs, keep = self._markers(mk)
        sb, ks = self._grid_ptrs(sum_bar, 30)
        M = mk["xm"].size
        out = np.empty(M)
        props = np.empty((7, M))
        hs, hf = _c(hrs), _c(hrf)
        rc = self.L.dual_marker_to_nodes_dot(self.pg, self.pm, ptrs, M, _p(hs), _p(hf), int(mode),
                                             sb, _p(out), _p(props))
        assert rc == 0, rc
        return out, props

    def gather_T_vjp(self, xm, ym, tkm, DT, tk2, timestep, tkm_bar, which, gdot=None):
        M = xm.size
        out = np.empty(M)
        a = [_c(xm), _c(ym), _c(tkm), _F(DT), _F(tk2), _c(tkm_bar)]
        gd = None if gdot is None else _F(gdot)
        rc = self.L.dual_gather_T_vjp(self.pg, *[_p(v) for v in a[:5]], int(timestep), M,
                                      _p(a[5]), GATHER_T_SEEDS[which], _p(gd), _p(out))
        assert rc == 0
        return out

    def gather_S_vjp(self, xm, ym, sxxm, sxym, DSXX, DSXY, sxx_bar, sxy_bar, which, gdot=None):
        M = xm.size
        out = np.empty(M)
        a = [_c(xm), _c(ym), _c(sxxm), _c(sxym), _F(DSXX), _F(DSXY)]
        b = [_c(sxx_bar), _c(sxy_bar)]
        gd = None if gdot is None else _F(gdot)
        rc = self.L.dual_gather_S_vjp(self.pg, *[_p(v) for v in a], M, _p(b[0]), _p(b[1]),
                                      GATHER_S_SEEDS[which], _p(gd), _p(out))
        assert rc == 0
        return out

    def rk4_vjp(self, mk, grids, dt, mode, bars, which, gdot=None):
        """bars: dict of output cotangents (xm ym tkm sxym sxxm; missing = 0);
        grids: vx vy vxf vyf wyx tk2."""
        M = mk["xm"].size
        m = [_c(mk[f]) for f in ("xm", "ym")] + [_c(mk["tm"], np.int64)] + \
            [_c(mk[f]) for f in ("tkm", "phim", "sxym", "sxxm")]
        gr = [_F(grids[n]) for n in ("vx", "vy", "vxf", "vyf", "wyx", "tk2")]
        bb = [None if bars.get(f) is None else _c(bars[f])
              for f in ("xm", "ym", "tkm", "sxym", "sxxm")]
        gd = None if gdot is None else _F(gdot)
        out = np.empty(M)
        rc = self.L.dual_rk4_vjp(self.pg, self.pm, *[_p(v) for v in m], *[_p(v) for v in gr], M,
                                 float(dt), int(mode), *[_p(v) for v in bb], RK4_SEEDS[which],
                                 _p(gd), _p(out))
        assert rc == 0
        return out

    def rk4_double(self, mk, grids, dt, mode):
        m = [_c(mk[f]).copy() for f in ("xm", "ym")] + [_c(mk["tm"], np.int64)] + \
            [_c(mk[f]).copy() for f in ("tkm", "phim", "sxym", "sxxm")]
        gr = [_F(grids[n]) for n in ("vx", "vy", "vxf", "vyf", "wyx", "tk2")]
        rc = self.L.dual_rk4_double(self.pg, self.pm, *[_p(v) for v in m], *[_p(v) for v in gr],
                                    m[0].size, float(dt), int(mode))
        assert rc == 0
        return m[0], m[1], m[3], m[5], m[6]

    def tk1_loss_grad(self, mk, vx, vy, nsteps, dt, mode, hrs, hrf, seeds):
        """Exact dL/d(tkm or phim) of the listed (marker, kind) seeds, kind 0
        tkm, 1 phim, for L = sum(tk1^2) after nsteps steps of the map of
        ereb_adjoint_tk1_loss; returns (L, grad)."""
        ptrs, keep = self._markers(mk)
        sm = np.ascontiguousarray([s[0] for s in seeds], dtype=np.int64)
        sk = np.ascontiguousarray([s[1] for s in seeds], dtype=np.int32)
        grad = np.empty(len(seeds))
        L = ctypes.c_double()
        a, b = _F(vx), _F(vy)
        hs, hf = _c(hrs), _c(hrf)
        rc = self.L.dual_tk1_loss_grad(self.pg, self.pm, ptrs, mk["xm"].size, _p(a), _p(b),
                                       int(nsteps), float(dt), int(mode), _p(hs), _p(hf),
                                       len(seeds), _p(sm), _p(sk), _p(grad), ctypes.byref(L))
        assert rc == 0
        return L.value, grad
