"""Geometry and material constants of Erebus.jl as runtime objects.

The reference bakes these in as compile-time ``const`` values
(src/constants.jl, src/test_constants.jl, src/prod_constants.jl); here they
are plain data handed to the C ABI (``ereb_geometry``, ``ereb_materials``
in include/ereb.h).  Line citations are into src/constants.jl (C:) unless
stated otherwise.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field

import numpy as np

_DP = ctypes.POINTER(ctypes.c_double)


class CGeometry(ctypes.Structure):
    """Mirror of ``ereb_geometry`` (include/ereb.h)."""
    _fields_ = [
        ("Nx", ctypes.c_int64), ("Ny", ctypes.c_int64),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double),
        ("x", _DP), ("y", _DP), ("xvx", _DP), ("yvx", _DP),
        ("xvy", _DP), ("yvy", _DP), ("xp", _DP), ("yp", _DP),
    ]


_MAT_TABLES = [
    "rhosolidm", "rhofluidm", "etasolidm", "etasolidmm", "etafluidm",
    "etafluidmm", "rhocpsolidm", "rhocpfluidm", "alphasolidm", "alphafluidm",
    "ksolidm", "kfluidm", "gggsolidm", "frictsolidm", "cohessolidm",
    "tenssolidm", "kphim0", "tkm0",
]
_MAT_SCALARS = [
    "phim0", "tmsolidphase", "tmfluidphase", "Lf", "MH2O", "MD", "rhoH2Of",
    "rhoH2Ofi", "VDs", "VWs", "etamin", "Kcont", "psurface",
    "bcleft", "bcright", "bctop", "bcbottom",
    "bcfleft", "bcfright", "bcftop", "bcfbottom",
    "vxleft", "vxright", "vytop", "vybottom",
]


class CMaterials(ctypes.Structure):
    """Mirror of ``ereb_materials`` (include/ereb.h)."""
    _fields_ = ([(n, ctypes.c_double * 3) for n in _MAT_TABLES]
                + [(n, ctypes.c_double) for n in _MAT_SCALARS]
                + [("brk4", ctypes.c_double * 4), ("crk4", ctypes.c_double * 3)])


@dataclass
class Materials:
    """Material tables (C:146-186) and scalars used on the transfer path."""
    rhosolidm: tuple = (3300.0, 3300.0, 1.0)          # C:146
    rhofluidm: tuple = (1000.0, 1000.0, 1.0)          # C:148
    etasolidm: tuple = (1.0e19, 1.0e19, 1.0e16)       # C:150
    etasolidmm: tuple = (1.0e19, 1.0e19, 1.0e16)      # C:152
    etafluidm: tuple = (1.0e12, 1.0e12, 1.0e-3)       # C:154
    etafluidmm: tuple = (1.0e-3, 1.0e-3, 1.0e-3)      # C:156
    rhocpsolidm: tuple = (3.3e6, 3.3e6, 3.0e6)        # C:158
    rhocpfluidm: tuple = (1.0e6, 1.0e6, 3.0e6)        # C:160
    alphasolidm: tuple = (3.0e-5, 3.0e-5, 0.0)        # C:162
    alphafluidm: tuple = (5.0e-5, 5.0e-5, 0.0)        # C:164
    ksolidm: tuple = (3.0, 3.0, 3000.0)               # C:166
    kfluidm: tuple = (50.0, 50.0, 3000.0)             # C:168
    start_hrsolidm: tuple = (0.0, 0.0, 0.0)           # C:170
    start_hrfluidm: tuple = (0.0, 0.0, 0.0)           # C:172
    gggsolidm: tuple = (1.0e10, 1.0e10, 1.0e10)       # C:174
    frictsolidm: tuple = (0.6, 0.6, 0.0)              # C:176
    cohessolidm: tuple = (1.0e8, 1.0e8, 1.0e8)        # C:178
    tenssolidm: tuple = (6.0e7, 6.0e7, 6.0e7)         # C:180
    kphim0: tuple = (1.0e-13, 1.0e-13, 1.0e-17)       # C:182
    tkm0: tuple = (170.0, 170.0, 170.0)               # C:184
    XWsolidm_init: tuple = (0.5, 0.5, float("nan"))   # C:186
    phim0: float = 0.2                                # C:225
    phimin: float = 1.0e-4                            # C:227
    tmsolidphase: float = 1416.0                      # C:218
    tmfluidphase: float = 273.0                       # C:220
    Lf: float = 333.55e3                              # C:222
    MH2O: float = 0.018                               # C:234
    MD: float = 0.120                                 # C:236
    rhoDs: float = 3300.0                             # C:238
    rhoWs: float = 2600.0                             # C:240
    rhoH2Of: float = 1000.0                           # C:242
    rhoH2Ofi: float = 917.0                           # C:244
    etamin: float = 1.0e12                            # C:343
    Kcont: float = 1.0e20                             # C:143
    psurface: float = 1.0e3                           # C:12
    bc: tuple = (-1.0, -1.0, -1.0, -1.0)              # C:283-289 l r t b
    bcf: tuple = (-1.0, -1.0, -1.0, -1.0)             # C:292-298
    strainrate: float = 0.0                           # C:301
    # radiogenic heating, C:1-5 and C:194-215
    hr_al: bool = True
    hr_fe: bool = False
    t_half_al: float = 717_000 * 31_540_000
    ratio_al: float = 5.0e-5
    E_al: float = 5.0470e-13
    f_al: float = 1.9e23
    t_half_fe: float = 2_620_000 * 31_540_000
    ratio_fe: float = 1.0e-6
    E_fe: float = 4.34e-13
    f_fe: float = 1.957e24
    marker_property_mode: int = 9                     # C:188 (prod: 1)

    @property
    def VDs(self) -> float:        # C:246
        return self.MD / self.rhoDs

    @property
    def VWs(self) -> float:        # C:248
        return (self.MD + self.MH2O) / self.rhoWs

    @property
    def brk4(self) -> tuple:       # C:312, Float64 of 1//6, 2//6 = 1//3
        return (1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0)

    @property
    def crk4(self) -> tuple:       # C:314
        return (0.5, 0.5, 1.0)

    def boundary_velocities(self, xsize: float, ysize: float) -> tuple:
        """vxleft, vxright, vytop, vybottom (C:303-309)."""
        s = self.strainrate
        return (s * xsize / 2, -s * xsize / 2, -s * ysize / 2, s * ysize / 2)

    def to_c(self, xsize: float, ysize: float) -> CMaterials:
        c = CMaterials()
        for n in _MAT_TABLES:
            setattr(c, n, (ctypes.c_double * 3)(*getattr(self, n)))
        for n in ["phim0", "tmsolidphase", "tmfluidphase", "Lf", "MH2O", "MD",
                  "rhoH2Of", "rhoH2Ofi", "etamin", "Kcont", "psurface"]:
            setattr(c, n, float(getattr(self, n)))
        c.VDs = self.VDs
        c.VWs = self.VWs
        c.bcleft, c.bcright, c.bctop, c.bcbottom = self.bc
        c.bcfleft, c.bcfright, c.bcftop, c.bcfbottom = self.bcf
        c.vxleft, c.vxright, c.vytop, c.vybottom = self.boundary_velocities(xsize, ysize)
        c.brk4 = (ctypes.c_double * 4)(*self.brk4)
        c.crk4 = (ctypes.c_double * 3)(*self.crk4)
        return c

    def radioactive_heating(self, timesum: float):
        """calculate_radioactive_heating (E:1336-1357) with Q_radiogenic
        (E:1294-1296): per-type solid and fluid heat production [W/m^3]."""
        def q(f, ratio, E, tau, t):
            return f * ratio * E * math.exp(-t * (1.0 / tau)) * (1.0 / tau)
        tau_al = self.t_half_al / math.log(2)
        tau_fe = self.t_half_fe / math.log(2)
        if self.hr_al:
            qa = q(self.f_al, self.ratio_al, self.E_al, tau_al, timesum)
            hrs = (qa * self.rhosolidm[0], qa * self.rhosolidm[1], 0.0)
        else:
            hrs = (0.0, 0.0, 0.0)
        if self.hr_fe:
            qf = q(self.f_fe, self.ratio_fe, self.E_fe, tau_fe, timesum)
            hrf = (qf * self.rhofluidm[0], 0.0, 0.0)
        else:
            hrf = (0.0, 0.0, 0.0)
        return np.array(hrs, dtype=np.float64), np.array(hrf, dtype=np.float64)


class CMarkerGrid(ctypes.Structure):
    """Mirror of ``ereb_marker_grid`` (include/ereb.h)."""
    _fields_ = [("Nxm", ctypes.c_int64), ("Nym", ctypes.c_int64),
                ("dxm", ctypes.c_double), ("dym", ctypes.c_double),
                ("xxm", _DP), ("yym", _DP), ("dii", ctypes.c_int64), ("djj", ctypes.c_int64),
                ("mdis_init", ctypes.c_double)]


class CHydration(ctypes.Structure):
    """Mirror of ``ereb_hydration`` (include/ereb.h)."""
    _fields_ = [(n, ctypes.c_double) for n in [
        "phimin", "phimax", "RG", "VH2Of", "VH2Ofi", "dHWD", "dSWD", "dVWD", "pfcoeff",
        "dtreaction", "log_completion_rate", "A_I", "b_I", "c_I", "Sxo_B", "Tscl_B", "To_B",
        "Sxo_T", "To_T", "Ea_T"]] + [("rate_mode", ctypes.c_int64)]


@dataclass
class Hydration:
    """Compaction and silicate (de)hydration constants (C:227-279)."""
    phimin: float = 1.0e-4                            # C:227
    RG: float = 8.314                                 # C:232
    dHWD: float = 40000.0                             # C:254
    dSWD: float = 60.0                                # C:256
    pfcoeff: float = 0.5                              # C:260
    dtreaction: float = 1.0e10                        # C:266
    completion_rate: float = 0.01                     # C:268 (log taken)
    A_I: float = 1.0e-11                              # C:273
    b_I: float = 2.5e-4
    c_I: float = 543.0
    Sxo_B: float = 2.0e-11                            # C:276
    Tscl_B: float = 10.0
    To_B: float = 293.0
    Sxo_T: float = 2.0e-11                            # C:279
    To_T: float = 293.0
    Ea_T: float = 63.8e3
    rate_mode: int = 1                                # C:271

    @property
    def phimax(self) -> float:     # C:229
        return 1.0 - self.phimin

    def to_c(self, mat: "Materials") -> CHydration:
        """Derived constants from the material set: VH2O (C:250-252), dVWD
        = VDs + VH2Of - VWs (C:258)."""
        c = CHydration()
        for n in ["phimin", "RG", "dHWD", "dSWD", "pfcoeff", "dtreaction", "A_I", "b_I",
                  "c_I", "Sxo_B", "Tscl_B", "To_B", "Sxo_T", "To_T", "Ea_T"]:
            setattr(c, n, float(getattr(self, n)))
        c.phimax = self.phimax
        c.VH2Of = mat.MH2O / mat.rhoH2Of
        c.VH2Ofi = mat.MH2O / mat.rhoH2Ofi
        c.dVWD = mat.VDs + c.VH2Of - mat.VWs
        c.log_completion_rate = math.log(self.completion_rate)
        c.rate_mode = int(self.rate_mode)
        return c


@dataclass
class Geometry:
    """Staggered-grid geometry (C:14-65) and marker lattice (C:102-124)."""
    Nx: int
    Ny: int
    xsize: float
    ysize: float
    Nxmc: int = 4
    Nymc: int = 4
    rplanet: float = 5_000.0
    rcrust: float = 4_800.0
    _axes: dict = field(default_factory=dict, repr=False)

    @property
    def Nx1(self) -> int:
        return self.Nx + 1

    @property
    def Ny1(self) -> int:
        return self.Ny + 1

    @property
    def dx(self) -> float:          # C:35
        return self.xsize / (self.Nx - 1)

    @property
    def dy(self) -> float:          # C:37
        return self.ysize / (self.Ny - 1)

    @property
    def Nxm(self) -> int:           # C:108
        return (self.Nx - 1) * self.Nxmc

    @property
    def Nym(self) -> int:           # C:110
        return (self.Ny - 1) * self.Nymc

    @property
    def dxm(self) -> float:         # C:112
        return self.xsize / self.Nxm

    @property
    def dym(self) -> float:         # C:114
        return self.ysize / self.Nym

    @property
    def start_marknum(self) -> int:  # C:124
        return self.Nxm * self.Nym

    @property
    def xcenter(self) -> float:     # C:21
        return self.xsize / 2

    @property
    def ycenter(self) -> float:     # C:23
        return self.ysize / 2

    def axes(self) -> dict:
        """Node axes of C:40-64 as Float64 arrays.  Every element is
        x0 + k*step; for the supported geometries the steps are whole or
        half metres, so the values are exact and equal Julia's range
        elements."""
        if "x" not in self._axes:
            dx, dy = self.dx, self.dy
            k = lambda n: np.arange(n, dtype=np.float64)
            node_axes = {
                "x": 0.0 + k(self.Nx) * dx,                 # C:40
                "y": 0.0 + k(self.Ny) * dy,                 # C:43
                "xvx": 0.0 + k(self.Nx1) * dx,              # C:47
                "yvx": -dy / 2 + k(self.Ny1) * dy,          # C:50
                "xvy": -dx / 2 + k(self.Nx1) * dx,          # C:54
                "yvy": 0.0 + k(self.Ny1) * dy,              # C:57
                "xp": -dx / 2 + k(self.Nx1) * dx,           # C:61
                "yp": -dy / 2 + k(self.Ny1) * dy,           # C:64
            }
            for a in node_axes.values():
                a.setflags(write=False)
            self._axes.update(node_axes)
        return {n: self._axes[n] for n in ("x", "y", "xvx", "yvx", "xvy", "yvy", "xp", "yp")}

    def marker_axes(self) -> tuple:
        """Marker launch lattice xxm = dxm/2:dxm:xsize-dxm/2 and yym (C:116,
        C:119): Julia lifts these float ranges to the rationals
        (k - 1/2) * xsize / Nxm, so each element is that rational rounded
        once."""
        if "xxm" not in self._axes:  # exact rationals are slow: computed once
            from fractions import Fraction
            fx, fy = Fraction(self.xsize) / self.Nxm, Fraction(self.ysize) / self.Nym
            xxm = np.array([float((k + Fraction(1, 2)) * fx) for k in range(self.Nxm)])
            yym = np.array([float((k + Fraction(1, 2)) * fy) for k in range(self.Nym)])
            for a in (xxm, yym):
                a.setflags(write=False)
            self._axes["xxm"], self._axes["yym"] = xxm, yym
        return self._axes["xxm"], self._axes["yym"]

    def marker_grid_c(self) -> CMarkerGrid:
        """ereb_marker_grid of this geometry: dii = 5*Nymc, djj = 5*Nxmc
        (E:5638-5639), mdis_init = 1e30 (C:122)."""
        xxm, yym = self.marker_axes()
        c = CMarkerGrid()
        c.Nxm, c.Nym, c.dxm, c.dym = self.Nxm, self.Nym, self.dxm, self.dym
        c.xxm, c.yym = xxm.ctypes.data_as(_DP), yym.ctypes.data_as(_DP)
        c.dii, c.djj, c.mdis_init = 5 * self.Nymc, 5 * self.Nxmc, 1.0e30
        c._keep = (xxm, yym)
        return c

    def to_c(self) -> CGeometry:
        a = self.axes()
        g = CGeometry()
        g.Nx, g.Ny, g.dx, g.dy = self.Nx, self.Ny, self.dx, self.dy
        for n in ("x", "y", "xvx", "yvx", "xvy", "yvy", "xp", "yp"):
            setattr(g, n, a[n].ctypes.data_as(_DP))
        g._keep = a  # keep the arrays alive with the struct
        return g

    # grid shapes, column-major as Julia
    @property
    def basic_shape(self):
        return (self.Ny, self.Nx)

    @property
    def stag_shape(self):
        return (self.Ny1, self.Nx1)


def test_constants() -> tuple[Geometry, Materials]:
    """src/test_constants.jl: 15x15 grid, 14 km box, 4x4 markers, mode 9."""
    return (Geometry(Nx=15, Ny=15, xsize=14_000.0, ysize=14_000.0,
                     Nxmc=4, Nymc=4, rplanet=5_000.0, rcrust=4_800.0),
            Materials(marker_property_mode=9))


def prod_constants() -> tuple[Geometry, Materials]:
    """src/prod_constants.jl: 141x141 grid, 140 km box, mode 1."""
    return (Geometry(Nx=141, Ny=141, xsize=140_000.0, ysize=140_000.0,
                     Nxmc=4, Nymc=4, rplanet=50_000.0, rcrust=48_000.0),
            Materials(marker_property_mode=1))


DT_LONGEST = 1.0e11  # C:319
