"""ctypes binding of libereb_hip.so (the C ABI in include/ereb.h).

The library is built in-tree by erebus.jl_amd/build.py.  Loading fails
loudly when it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .constants import CGeometry, CMaterials, Geometry, Hydration, Materials

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.normpath(os.path.join(HERE, "..", "lib", "libereb_hip.so"))

OK, ERR_ARG, ERR_MODE, ERR_HIP, ERR_DOMAIN, ERR_NOMEM, ERR_STATE = range(7)
FP64, MIXED = 0, 1
SYS_STOKES_DARCY, SYS_THERMAL = 0, 1

MARKER_FIELDS = [
    "xm", "ym", "tm", "tkm", "phim", "sxxm", "sxym", "XWsolidm0", "etavpm",
    "inv_gggtotalm", "fricttotalm", "cohestotalm", "tenstotalm", "rhofluidcur",
    "alphasolidcur", "alphafluidcur", "rhototalm", "rhocptotalm", "etatotalm",
    "hrtotalm", "ktotalm", "tkm_rhocptotalm", "etafluidcur_inv_kphim",
    # hydration state (E:2558-2694)
    "XWsolidm", "phinewm", "pfm0",
]
MARKER_ID = {n: k for k, n in enumerate(MARKER_FIELDS)}

GRID_FIELDS = [
    # basic (Ny, Nx)
    "ETA0SUM", "ETASUM", "GGGSUM", "SXYSUM", "COHSUM", "TENSUM", "FRISUM", "WTSUM",
    "ETA0", "ETA", "GGG", "SXY0", "COH", "TEN", "FRI", "wyx", "DSXY",
    # Vx (Ny1, Nx1)
    "RHOXSUM", "RHOFXSUM", "KXSUM", "PHIXSUM", "RXSUM", "WTXSUM",
    "RHOX", "RHOFX", "KX", "PHIX", "RX", "vx", "vxf", "gx",
    # Vy
    "RHOYSUM", "RHOFYSUM", "KYSUM", "PHIYSUM", "RYSUM", "WTYSUM",
    "RHOY", "RHOFY", "KY", "PHIY", "RY", "vy", "vyf", "gy",
    # P
    "RHOSUM", "RHOCPSUM", "ALPHASUM", "ALPHAFSUM", "HRSUM", "GGGPSUM", "SXXSUM",
    "TKSUM", "PHISUM", "WTPSUM",
    "RHO", "RHOCP", "ALPHA", "ALPHAF", "HR", "GGGP", "SXX0", "tk1", "PHI", "BETAPHI",
    "ETAP", "ETAPHI", "pr0", "pf0", "DMP", "DT", "tk2", "DSXX", "HA", "HS", "DHP",
    # basic (Ny, Nx)
    "YNY_inv_ETA",
    # P
    "APHI", "XWS", "XWSSUM", "pf", "DMPSUM", "DHPSUM",
    "pr", "ps", "ps0",
]
GRID_ID = {n: k for k, n in enumerate(GRID_FIELDS)}
BASIC_GRIDS = set(GRID_FIELDS[:GRID_FIELDS.index("DSXY") + 1]) | {"YNY_inv_ETA"}

# SUM grids in the order of reset_interpolated_properties! (E:968-999)
SUM_FIELDS = [
    "ETA0SUM", "ETASUM", "GGGSUM", "SXYSUM", "COHSUM", "TENSUM", "FRISUM", "WTSUM",
    "RHOXSUM", "RHOFXSUM", "KXSUM", "PHIXSUM", "RXSUM", "WTXSUM",
    "RHOYSUM", "RHOFYSUM", "KYSUM", "PHIYSUM", "RYSUM", "WTYSUM",
    "RHOSUM", "RHOCPSUM", "ALPHASUM", "ALPHAFSUM", "HRSUM", "GGGPSUM", "SXXSUM",
    "TKSUM", "PHISUM", "WTPSUM",
]

KERNEL_CLASSES = ["sort", "scatter", "combine", "nodes", "gather_T", "gather_S",
                  "rk4", "asm_sd", "asm_t", "viscosity", "porosity", "melt", "reaction",
                  "backtrace", "replenish"]

_D = ctypes.c_double
_I64 = ctypes.c_int64
_P = ctypes.c_void_p
_DP = ctypes.POINTER(ctypes.c_double)
_I = ctypes.c_int

# exported C entry points: name -> argtypes (restype int unless noted)
PROTOTYPES = {
    "ereb_version": [],
    "ereb_host_selfcheck": [_P, _P],
    "ereb_last_error": [_P],
    "ereb_ctx_create": [ctypes.POINTER(CGeometry), ctypes.POINTER(CMaterials), _I, _I,
                        ctypes.POINTER(_P)],
    "ereb_ctx_destroy": [_P],
    "ereb_ctx_set_precision": [_P, _I],
    "ereb_ctx_reserve": [_P, _I64],
    "ereb_sync": [_P],
    "ereb_marknum": [_P],
    "ereb_markers_set": [_P, _I, _P, _I64],
    "ereb_markers_get": [_P, _I, _P, _I64],
    "ereb_grid_numel": [_P, _I],
    "ereb_save_field": [_I, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(_I),
                        ctypes.POINTER(_I)],
    "ereb_save_state_fields": [_P, _P, _I64],
    "ereb_grid_set": [_P, _I, _P],
    "ereb_grid_get": [_P, _I, _P],
    "ereb_yny_set": [_P, _P],
    "ereb_yny_get": [_P, _P],
    "ereb_step_scatter": [_P, _P, _P, _I64],
    "ereb_step_node_properties": [_P],
    "ereb_step_gather_temperature": [_P, _I64],
    "ereb_step_gather_stress": [_P],
    "ereb_step_advect_rk4": [_P, _D, _I64],
    "ereb_step_assemble_stokes_darcy": [_P, _D, _P],
    "ereb_step_assemble_thermal": [_P, _D, _P],
    "ereb_csc_get": [_P, _I, _P, _P, _P, _P],
    "ereb_csc_size": [_P, _I, _P, _P],
    "ereb_csc_pattern_changed": [_P, _I, _P],
    "ereb_host_alloc": [_P, _I64, _P],
    "ereb_host_free": [_P, _P],
    "ereb_sort_get": [_P, _P, _P, _P],
    "ereb_marker_to_nodes": [_P] + [_P] * 23 + [_P, _P, _I64, _P, _I64],
    "ereb_compute_basic_node_properties": [_P] + [_P] * 16,
    "ereb_compute_vx_node_properties": [_P] + [_P] * 11,
    "ereb_compute_vy_node_properties": [_P] + [_P] * 11,
    "ereb_compute_p_node_properties": [_P] + [_P] * 20,
    "ereb_apply_insulating_boundary_conditions": [_P, _P, _I64, _I64],
    "ereb_update_marker_temperature": [_P, _P, _P, _P, _P, _P, _I64, _I64],
    "ereb_update_marker_stress": [_P, _P, _P, _P, _P, _P, _P, _I64],
    "ereb_move_markers_rk4": [_P] + [_P] * 13 + [_I64, _D, _I64],
    "ereb_assemble_hydromechanical_lse": [_P] + [_P] * 20 + [_D, _P, _P],
    "ereb_assemble_thermal_lse": [_P] + [_P] * 8 + [_P, _D, _P],
    "ereb_assemble_thermal_lse_noalloc": [_P] + [_P] * 8 + [_P, _D, _P, _P, _P, _I64, _P],
    "ereb_adjoint_tk1_loss": [_P, _I64, _I64, _D, _I64, _P, _P, _P, _P, _P, _P],
    "ereb_adjoint_set_grid_budget": [_P, _I64],
    "ereb_step_transfer": [_P, _P, _P, _I64, _D, _I64],
    "ereb_ctx_set_graphs": [_P, _I],
    "ereb_step_graph_stats": [_P, _P, _P],
    "ereb_marker_to_nodes_pullback": [_P, _P, _I64, _P, _P, _I64, _P, _P, _P],
    "ereb_update_marker_temperature_pullback": [_P, _P, _P, _P, _P, _I64, _I64] + [_P] * 6,
    "ereb_update_marker_stress_pullback": [_P, _P, _P, _P, _P, _I64] + [_P] * 6,
    "ereb_move_markers_rk4_pullback": [_P] + [_P] * 13 + [_I64, _D, _I64] + [_P] * 17,
    "ereb_step_marker_viscosity": [_P],
    "ereb_step_marker_porosity": [_P, _P, _D],
    "ereb_step_melt_composition": [_P],
    "ereb_step_thermochemical_reaction": [_P, _P, _D, _I64, _I64],
    "ereb_update_marker_viscosity": [_P] + [_P] * 8 + [_I64],
    "ereb_update_marker_porosity": [_P, _P] + [_P] * 5 + [_D, _I64],
    "ereb_update_p_nodes_melt_composition": [_P] + [_P] * 6 + [_I64],
    "ereb_perform_thermochemical_reaction": [_P, _P] + [_P] * 15 + [_I64, _D, _I64, _I64],
    "ereb_step_backtrace_pressures": [_P, _D],
    "ereb_backtrace_pressures_rk4": [_P] + [_P] * 10 + [_D],
    "ereb_step_replenish_markers": [_P, _P, _I, ctypes.c_uint64, _P, _P, _P],
    "ereb_replenish_markers": [_P] + [_P] * 28 + [_I64, _I64, _P, _I, ctypes.c_uint64, _P],
    "ereb_profile_enable": [_P, _I],
    "ereb_profile_get": [_P, _I, _P, _P],
    "ereb_profile_reset": [_P],
}
RESTYPES = {"ereb_last_error": ctypes.c_char_p, "ereb_marknum": _I64,
            "ereb_grid_numel": _I64}

_lib = None


class EREBError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[ereb {code}] {msg}")
        self.code = code


class EREBModeError(EREBError, ValueError):
    """Unknown property mode; the reference throws "unknown mode" (E:2289)."""


def load(path: str | None = None) -> ctypes.CDLL:
    """Load libereb_hip.so; raises when the HIP library is not built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            f"libereb_hip.so not found at {p}: build it with "
            "`python erebus.jl_amd/build.py` (there is no CPU fallback)")
    lib = ctypes.CDLL(p)
    for name, args in PROTOTYPES.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = RESTYPES.get(name, _I)
    if path is None:
        _lib = lib
    return lib


SAVE_GRID, SAVE_YNY, SAVE_MARKER = 0, 1, 2


def save_fields(lib: ctypes.CDLL | None = None):
    """The library's save_state field set: [(key, kind, field enum)] in the
    order of the record's keyword list (E:5864-5957)."""
    lib = lib or load()
    name, kind, field = ctypes.c_char_p(), _I(), _I()
    out, k = [], 0
    while lib.ereb_save_field(k, ctypes.byref(name), ctypes.byref(kind),
                              ctypes.byref(field)) == OK:
        out.append((name.value.decode(), kind.value, field.value))
        k += 1
    return out


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data


def _f64(a, n=None) -> np.ndarray:
    a = np.asarray(a, dtype=np.float64)
    a = np.require(a, requirements=["F_CONTIGUOUS", "ALIGNED"]) if a.ndim > 1 else \
        np.ascontiguousarray(a)
    if n is not None and a.size != n:
        raise ValueError(f"expected {n} elements, got {a.size}")
    return a


class Context:
    """Owns one ereb_ctx: geometry, materials and the HBM-resident state."""

    def __init__(self, geom: Geometry, mat: Materials, precision: int = FP64,
                 device: int = 0):
        self.lib = load()
        self._pinned = []
        self.geom = geom
        self.mat = mat
        self._cg = geom.to_c()
        self._cm = mat.to_c(geom.xsize, geom.ysize)
        h = _P()
        rc = self.lib.ereb_ctx_create(ctypes.byref(self._cg), ctypes.byref(self._cm),
                                      precision, device, ctypes.byref(h))
        self.h = h.value
        if rc != OK:
            msg = self.lib.ereb_last_error(self.h).decode() if self.h else "invalid geometry"
            if self.h:
                self.lib.ereb_ctx_destroy(self.h)
            self.h = None
            raise EREBError(rc, f"ereb_ctx_create: {msg}")

    def close(self):
        if self.h:
            for p in self._pinned:
                self.lib.ereb_host_free(self.h, p)
            self._pinned = []
            self.lib.ereb_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- status -----------------------------------------------------------
    def check(self, rc: int, what: str = ""):
        if rc == OK:
            return
        msg = self.lib.ereb_last_error(self.h).decode()
        cls = EREBModeError if rc == ERR_MODE else EREBError
        raise cls(rc, f"{what}: {msg}" if what else msg)

    def sync(self):
        self.check(self.lib.ereb_sync(self.h), "sync")

    def set_precision(self, p: int):
        self.check(self.lib.ereb_ctx_set_precision(self.h, p))

    @property
    def marknum(self) -> int:
        return int(self.lib.ereb_marknum(self.h))

    # ---- markers ----------------------------------------------------------
    def set_marker(self, name: str, a: np.ndarray):
        fid = MARKER_ID[name]
        if name == "tm":
            a = np.ascontiguousarray(a, dtype=np.int64)
        else:
            a = np.ascontiguousarray(a, dtype=np.float64)
        self.check(self.lib.ereb_markers_set(self.h, fid, _ptr(a), a.size), f"set {name}")

    def get_marker(self, name: str, n: int | None = None) -> np.ndarray:
        n = self.marknum if n is None else n
        out = np.empty(n, dtype=np.int64 if name == "tm" else np.float64)
        self.check(self.lib.ereb_markers_get(self.h, MARKER_ID[name], _ptr(out), n),
                   f"get {name}")
        return out

    def set_markers(self, d: dict):
        for k, v in d.items():
            if k in MARKER_ID:
                self.set_marker(k, v)

    # ---- grids ------------------------------------------------------------
    def grid_shape(self, name: str):
        g = self.geom
        return (g.Ny, g.Nx) if name in BASIC_GRIDS else (g.Ny1, g.Nx1)

    def set_grid(self, name: str, a: np.ndarray):
        shp = self.grid_shape(name)
        a = np.asfortranarray(np.asarray(a, dtype=np.float64))
        if a.shape != shp:
            raise ValueError(f"{name}: shape {a.shape} != {shp}")
        self.check(self.lib.ereb_grid_set(self.h, GRID_ID[name], _ptr(a)), f"set {name}")

    def get_grid(self, name: str) -> np.ndarray:
        out = np.empty(self.grid_shape(name), dtype=np.float64, order="F")
        self.check(self.lib.ereb_grid_get(self.h, GRID_ID[name], _ptr(out)), f"get {name}")
        return out

    def set_yny(self, a: np.ndarray):
        a = np.asfortranarray(np.asarray(a, dtype=np.uint8))
        self.check(self.lib.ereb_yny_set(self.h, _ptr(a)))

    def get_yny(self) -> np.ndarray:
        out = np.empty((self.geom.Ny, self.geom.Nx), dtype=np.uint8, order="F")
        self.check(self.lib.ereb_yny_get(self.h, _ptr(out)))
        return out

    def save_state_fields(self, names=None, marknum: int | None = None) -> dict:
        """The device-resident arrays of save_state's record (E:5718-5957) in
        one library call: {key: array} for `names` (default: all 76), grids
        and YNY shaped as the reference's matrices, marker vectors of
        `marknum` entries (default: the resident count)."""
        n = self.marknum if marknum is None else int(marknum)
        table = save_fields(self.lib)
        want = [k for k, _, _ in table] if names is None else list(names)
        out, ptrs = {}, (ctypes.c_void_p * len(table))()
        for i, (key, kind, field) in enumerate(table):
            if key not in want:
                continue
            if kind == SAVE_GRID:
                a = np.empty(self.grid_shape(GRID_FIELDS[field]), dtype=np.float64, order="F")
            elif kind == SAVE_YNY:
                a = np.empty((self.geom.Ny, self.geom.Nx), dtype=np.uint8, order="F")
            else:
                a = np.empty(n, dtype=np.int64 if field == MARKER_ID["tm"] else np.float64)
            out[key] = a
            ptrs[i] = a.ctypes.data
        missing = set(want) - set(out)
        if missing:
            raise KeyError(f"not in the save_state field set: {sorted(missing)}")
        self.check(self.lib.ereb_save_state_fields(self.h, ptrs, n), "save_state_fields")
        return out

    # ---- resident steps ---------------------------------------------------
    def step_scatter(self, hrsolidm, hrfluidm, mode: int):
        hs = np.ascontiguousarray(hrsolidm, dtype=np.float64)
        hf = np.ascontiguousarray(hrfluidm, dtype=np.float64)
        self.check(self.lib.ereb_step_scatter(self.h, _ptr(hs), _ptr(hf), int(mode)),
                   "step_scatter")

    def step_transfer(self, hrsolidm, hrfluidm, mode: int, dt: float, timestep: int):
        """The whole resident transfer step (on the context's two streams; graph
        replay when set_graphs(True))."""
        hs = np.ascontiguousarray(hrsolidm, dtype=np.float64)
        hf = np.ascontiguousarray(hrfluidm, dtype=np.float64)
        self.check(self.lib.ereb_step_transfer(self.h, _ptr(hs), _ptr(hf), int(mode), float(dt),
                                               int(timestep)), "step_transfer")

    def set_graphs(self, on: bool):
        self.check(self.lib.ereb_ctx_set_graphs(self.h, int(bool(on))), "set_graphs")

    def graph_stats(self):
        """(captures, replays) of the transfer-step graph so far."""
        cap, rep = ctypes.c_int64(), ctypes.c_int64()
        self.check(self.lib.ereb_step_graph_stats(self.h, ctypes.byref(cap), ctypes.byref(rep)),
                   "graph_stats")
        return cap.value, rep.value

    def step_node_properties(self):
        self.check(self.lib.ereb_step_node_properties(self.h), "step_node_properties")

    def step_gather_temperature(self, timestep: int):
        self.check(self.lib.ereb_step_gather_temperature(self.h, int(timestep)))

    def step_gather_stress(self):
        self.check(self.lib.ereb_step_gather_stress(self.h))

    def step_advect_rk4(self, dt: float, mode: int):
        self.check(self.lib.ereb_step_advect_rk4(self.h, float(dt), int(mode)), "rk4")

    def step_assemble_stokes_darcy(self, dt: float, want_nnz: bool = False):
        n = ctypes.c_int64(-1)
        self.check(self.lib.ereb_step_assemble_stokes_darcy(
            self.h, float(dt), ctypes.addressof(n) if want_nnz else None))
        return n.value if want_nnz else None

    def step_assemble_thermal(self, dt: float, want_nnz: bool = False):
        n = ctypes.c_int64(-1)
        self.check(self.lib.ereb_step_assemble_thermal(
            self.h, float(dt), ctypes.addressof(n) if want_nnz else None))
        return n.value if want_nnz else None

    # ---- viscosity, compaction and (de)hydration ---------------------------
    def _hyd(self, hyd):
        hyd = Hydration() if hyd is None else hyd
        c = hyd.to_c(self.mat)
        self._hyd_keep = c
        return ctypes.addressof(c)

    def step_marker_viscosity(self):
        self.check(self.lib.ereb_step_marker_viscosity(self.h), "step_marker_viscosity")

    def step_marker_porosity(self, dt: float, hyd: Hydration | None = None):
        self.check(self.lib.ereb_step_marker_porosity(self.h, self._hyd(hyd), float(dt)),
                   "step_marker_porosity")

    def step_melt_composition(self):
        self.check(self.lib.ereb_step_melt_composition(self.h), "step_melt_composition")

    def step_thermochemical_reaction(self, dt: float, timestep: int, titer: int,
                                     hyd: Hydration | None = None):
        self.check(self.lib.ereb_step_thermochemical_reaction(
            self.h, self._hyd(hyd), float(dt), int(timestep), int(titer)),
            "step_thermochemical_reaction")

    def step_backtrace_pressures(self, dt: float):
        self.check(self.lib.ereb_step_backtrace_pressures(self.h, float(dt)),
                   "step_backtrace_pressures")

    def step_replenish_markers(self, randomized: bool = False, seed: int = 0,
                               maps: bool = True):
        """replenish_markers! on the resident markers: (new marknum, mdis,
        mnum), mdis and mnum (Nym, Nxm) Fortran-ordered as Julia's; with
        maps=False they stay on the device (None is returned for them)."""
        g = self.geom
        mg = g.marker_grid_c()
        mdis = np.empty((g.Nym, g.Nxm), order="F") if maps else None
        mnum = np.empty((g.Nym, g.Nxm), dtype=np.int64, order="F") if maps else None
        n = ctypes.c_int64()
        self.check(self.lib.ereb_step_replenish_markers(
            self.h, ctypes.addressof(mg), int(bool(randomized)), int(seed), _ptr(mdis),
            _ptr(mnum), ctypes.addressof(n)), "step_replenish_markers")
        return n.value, mdis, mnum

    def csc_size(self, which: int):
        """(n, nnz) of the last assembled system."""
        n = ctypes.c_int64()
        nnz = ctypes.c_int64()
        self.check(self.lib.ereb_csc_size(self.h, which, ctypes.addressof(n),
                                          ctypes.addressof(nnz)), "csc_size")
        return n.value, nnz.value

    def csc(self, which: int, pattern: bool = True, out=None):
        """(colptr, rowval, nzval, rhs) of the last assembly, 1-based Int64
        indices as in SparseMatrixCSC{Float64,Int64}.  pattern=False fetches
        only nzval and rhs (colptr, rowval returned as None).  out: arrays
        (colptr, rowval, nzval, rhs) to fill instead of new ones, e.g. from
        host_array; entries may be None to skip that part."""
        n, nnz = self.csc_size(which)
        if out is None:
            out = (np.empty(n + 1, dtype=np.int64) if pattern else None,
                   np.empty(nnz, dtype=np.int64) if pattern else None,
                   np.empty(nnz, dtype=np.float64), np.empty(n, dtype=np.float64))
        want = [(np.int64, n + 1), (np.int64, nnz), (np.float64, nnz), (np.float64, n)]
        for a, (dt, size) in zip(out, want):
            if a is not None and (a.dtype != dt or a.size < size or not a.flags.c_contiguous):
                raise ValueError(f"csc: output array needs {size} contiguous {np.dtype(dt)}")
        self.check(self.lib.ereb_csc_get(self.h, which, *[_ptr(a) for a in out]), "csc_get")
        return tuple(None if a is None else a[:size] for a, (_, size) in zip(out, want))

    def csc_pattern_changed(self, which: int) -> bool:
        """True when the last assembled pattern differs from the one last
        fetched with colptr/rowval (or none was fetched)."""
        ch = ctypes.c_int()
        self.check(self.lib.ereb_csc_pattern_changed(self.h, which, ctypes.addressof(ch)),
                   "csc_pattern_changed")
        return bool(ch.value)

    def host_array(self, n: int, dtype) -> np.ndarray:
        """Page-locked host array of n elements (ereb_host_alloc); valid until
        the context closes, which frees it."""
        dt = np.dtype(dtype)
        nbytes = int(n) * dt.itemsize
        p = _P()
        self.check(self.lib.ereb_host_alloc(self.h, nbytes, ctypes.byref(p)), "host_alloc")
        if not p.value:
            return np.empty(0, dtype=dt)
        self._pinned.append(p.value)
        return np.frombuffer((ctypes.c_char * nbytes).from_address(p.value), dtype=dt,
                             count=int(n))

    def sort_order(self):
        n = self.marknum
        perm = np.empty(n, dtype=np.int32)
        key = np.empty(n, dtype=np.int32)
        nt = ctypes.c_int64()
        self.check(self.lib.ereb_sort_get(self.h, _ptr(perm), _ptr(key),
                                          ctypes.addressof(nt)), "sort_get")
        return perm, key, nt.value

    # ---- adjoint ----------------------------------------------------------
    def adjoint_tk1_loss(self, nsteps: int, checkpoint_every: int, dt: float, mode: int,
                         hrsolidm, hrfluidm, grad: bool = True, want_tk1: bool = False,
                         out=None):
        """L = sum(tk1^2) after nsteps steps of the transfer map and, with
        grad, dL/dtkm and dL/dphim of the initial markers (adjoint kernels,
        checkpoints every checkpoint_every steps).  Returns (L, gT, gphi,
        tk1) with None for what was not asked; the markers are unchanged.
        out: optional (gT, gphi) arrays of marknum float64 to write the
        gradient into (page-locked ones from host_array take the device copy
        without staging)."""
        hs = np.ascontiguousarray(hrsolidm, dtype=np.float64)
        hf = np.ascontiguousarray(hrfluidm, dtype=np.float64)
        n = self.marknum
        L = ctypes.c_double()
        if grad and out is not None:
            gT, gP = out
            for a in (gT, gP):
                if a.dtype != np.float64 or a.shape != (n,) or not a.flags.c_contiguous:
                    raise ValueError("adjoint out arrays must be contiguous float64 of marknum")
        else:
            gT = np.empty(n) if grad else None
            gP = np.empty(n) if grad else None
        tk1 = np.empty(self.grid_shape("tk1"), order="F") if want_tk1 else None
        self.check(self.lib.ereb_adjoint_tk1_loss(
            self.h, int(nsteps), int(checkpoint_every), float(dt), int(mode), _ptr(hs),
            _ptr(hf), ctypes.addressof(L), _ptr(tk1), _ptr(gT), _ptr(gP)), "adjoint")
        return L.value, gT, gP, tk1

    def adjoint_set_grid_budget(self, nbytes: int):
        """Most bytes of per-step grid fields the adjoint keeps (0: recompute
        every step in full)."""
        self.check(self.lib.ereb_adjoint_set_grid_budget(self.h, int(nbytes)), "grid budget")

    # ---- profiling --------------------------------------------------------
    def profile(self, on: bool = True):
        self.check(self.lib.ereb_profile_enable(self.h, 1 if on else 0))

    def profile_reset(self):
        self.check(self.lib.ereb_profile_reset(self.h))

    def profile_get(self, kclass: str):
        ms = ctypes.c_double()
        cnt = ctypes.c_int64()
        self.check(self.lib.ereb_profile_get(self.h, KERNEL_CLASSES.index(kclass),
                                             ctypes.addressof(ms), ctypes.addressof(cnt)))
        return ms.value, cnt.value
