"""The C ABI: libereb_hip.so loads without a GPU and exports every function
that include/ereb.h declares; the ctypes prototypes cover the same set.
No compute call is made here."""
from __future__ import annotations

import ctypes
import os
import re

import pytest

import ereb
from ereb import lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ereb.h")


def declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ereb_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = declared()
    assert "ereb_ctx_create" in names and "ereb_move_markers_rk4" in names
    assert len(names) >= 35


def test_library_exports_every_declared_symbol():
    lib = ereb.load()
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_prototypes_match_header():
    assert set(L.PROTOTYPES) == set(declared())


def test_enums_match_header():
    src = open(HEADER).read()
    m = re.search(r"enum ereb_marker_field \{(.*?)\};", src, re.S).group(1)
    names = re.findall(r"EREB_M_([A-Z0-9_]+)", m)
    assert [n.lower() for n in names[:-1]] == [f.lower() for f in L.MARKER_FIELDS]
    g = re.search(r"enum ereb_grid_field \{(.*?)\};", src, re.S).group(1)
    gn = re.findall(r"EREB_G_([A-Z0-9_]+)", g)
    assert [n.lower() for n in gn[:-1]] == [f.lower() for f in L.GRID_FIELDS]
    k = re.search(r"enum ereb_kernel_class \{(.*?)\};", src, re.S).group(1)
    assert len(re.findall(r"EREB_K_([A-Z0-9_]+)", k)) - 1 == len(L.KERNEL_CLASSES)


def test_version_and_null_handling():
    lib = ereb.load()
    assert lib.ereb_version() == 1
    assert lib.ereb_last_error(None) == b"null context"
    assert lib.ereb_ctx_destroy(None) == 0
    out = ctypes.c_void_p()
    # invalid geometry is rejected before any device call
    g, mat = ereb.test_constants()
    cg = g.to_c()
    cg.Nx = 2
    cm = mat.to_c(g.xsize, g.ysize)
    assert lib.ereb_ctx_create(ctypes.byref(cg), ctypes.byref(cm), 0, 0,
                               ctypes.byref(out)) == L.ERR_ARG
    assert out.value is None


def test_library_path_is_in_tree():
    assert os.path.commonpath([ereb.LIB_PATH, ROOT]) == ROOT
    assert os.path.exists(ereb.LIB_PATH)


# The arrays among save_state's jldsave keywords (E:5864-5957), in order.
SAVE_STATE_ARRAYS = """
ETA ETA0 GGG EXY SXY SXY0 wyx COH TEN FRI YNY RHOX RHOFX KX PHIX vx vxf RX qxD gx
RHOY RHOFY KY PHIY vy vyf RY qyD gy RHO RHOCP ALPHA ALPHAF HR HA HS ETAP GGGP EXX
SXX SXX0 tk1 tk2 vxp vyp vxpf vypf pr pf ps pr0 pf0 ps0 ETAPHI BETAPHI PHI APHI FI
ETA5 ETA00 YNY5 YNY00 YNY_inv_ETA DSXY EII SII DSXX DMP DHP XWS XWsolidm0 xm ym tm
tkm sxxm sxym etavpm phim rhototalm rhocptotalm etatotalm hrtotalm ktotalm
tkm_rhocptotalm etafluidcur_inv_kphim inv_gggtotalm fricttotalm cohestotalm
tenstotalm rhofluidcur alphasolidcur alphafluidcur""".split()
# produced by the host solves and the plastic iterations, not held on the device
NOT_RESIDENT = set("EXY SXY qxD qyD EXX SXX vxp vyp vxpf vypf FI ETA5 ETA00 YNY5 YNY00 "
                   "EII SII".split())


def test_save_state_field_set():
    """The library's save_state field set is the record's arrays minus the 17
    the device does not hold, in the record's order, each key naming the
    resident field of the same name."""
    table = L.save_fields()
    assert len(table) == 76
    assert L.load().ereb_save_field(-1, None, None, None) == L.ERR_ARG
    assert [k for k, _, _ in table] == [a for a in SAVE_STATE_ARRAYS if a not in NOT_RESIDENT]
    for key, kind, field in table:
        if kind == L.SAVE_GRID:
            assert L.GRID_FIELDS[field] == key
        elif kind == L.SAVE_MARKER:
            assert L.MARKER_FIELDS[field] == key
        else:
            assert (kind, key) == (L.SAVE_YNY, "YNY")
    assert L.load().ereb_save_field(76, None, None, None) == L.ERR_ARG
