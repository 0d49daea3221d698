"""Host-side checks of libereb_hip_san.so (the library's host code built with
ASan + UBSan, no device code), run in a child process that preloads the
sanitizer runtime (tests/test_sanitize.py).  Any sanitizer report aborts the
process with a non-zero status; a failed expectation raises.  Needs no GPU:

  - ereb_host_selfcheck: the stencil slot discovery of both systems, which
    runs the assembly's equation code over a small grid on the host;
  - the save_state field table;
  - every exported entry point called with a null context (and null or zero
    arguments): EREB_ERR_ARG, -1 for the counts, "null context" for the
    error text, EREB_OK for ereb_ctx_destroy (as free(NULL));
  - ereb_ctx_create's validation: null arguments, an unknown precision form,
    too few nodes, staggered axes that do not coincide, grids too large;
    then a valid geometry, which without a device returns the context with
    a HIP error message (with one, a working context) and is destroyed.
"""
from __future__ import annotations

import ctypes
import sys

from ereb import lib as L
from ereb.constants import test_constants

# entry points whose first argument is not a context
NOT_CTX = {"ereb_version", "ereb_host_selfcheck", "ereb_save_field", "ereb_ctx_create"}


def _zero(t):
    if t in (ctypes.c_double,):
        return 0.0
    if t in (ctypes.c_int, ctypes.c_int64, ctypes.c_uint64):
        return 0
    return None


def main(path: str) -> None:
    lib = L.load(path)
    sd, th = ctypes.c_int64(), ctypes.c_int64()
    assert lib.ereb_host_selfcheck(ctypes.byref(sd), ctypes.byref(th)) == L.OK
    assert (sd.value, th.value) == (44, 5), (sd.value, th.value)
    table = L.save_fields(lib)
    assert len(table) == 76 and table[0][0] == "ETA" and table[-1][0] == "alphafluidcur"

    for name, args in L.PROTOTYPES.items():
        if name in NOT_CTX:
            continue
        rc = getattr(lib, name)(*[_zero(t) for t in args])
        if name == "ereb_last_error":
            assert rc == b"null context", rc
        elif name == "ereb_ctx_destroy":
            assert rc == L.OK, rc  # like free(NULL)
        elif name in ("ereb_marknum", "ereb_grid_numel"):
            assert rc == -1, (name, rc)
        else:
            assert rc == L.ERR_ARG, (name, rc)

    g, mat = test_constants()
    cg, cm = g.to_c(), mat.to_c(g.xsize, g.ysize)
    h = ctypes.c_void_p()
    create = lib.ereb_ctx_create
    assert create(ctypes.byref(cg), ctypes.byref(cm), 0, 0, None) == L.ERR_ARG
    assert create(None, ctypes.byref(cm), 0, 0, ctypes.byref(h)) == L.ERR_ARG and not h.value
    assert create(ctypes.byref(cg), ctypes.byref(cm), 7, 0, ctypes.byref(h)) == L.ERR_ARG
    for field, value in (("Nx", 2), ("Ny", 1), ("Nx", 2_000_000), ("dx", 0.0)):
        bad = g.to_c()
        setattr(bad, field, value)
        assert create(ctypes.byref(bad), ctypes.byref(cm), 0, 0, ctypes.byref(h)) == L.ERR_ARG
        assert not h.value
    a = g.axes()
    xvx = a["xvx"].copy()
    xvx[3] += 1.0
    bad = g.to_c()
    bad.xvx = xvx.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    assert create(ctypes.byref(bad), ctypes.byref(cm), 0, 0, ctypes.byref(h)) == L.ERR_ARG
    rc = create(ctypes.byref(cg), ctypes.byref(cm), 0, 0, ctypes.byref(h))
    if rc != L.OK:
        assert rc == L.ERR_HIP and h.value, rc
        msg = lib.ereb_last_error(h.value)
        assert msg, msg
    assert lib.ereb_ctx_destroy(h.value) == L.OK
    print("host checks passed", flush=True)


if __name__ == "__main__":
    main(sys.argv[1])
