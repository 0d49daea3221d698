"""ereb_assemble_thermal_lse_noalloc: assemble_thermal_lse! (E:4526-4611)
into caller-owned device buffers with no device allocation, its scratch in
a static device array.  Through the C ABI its RT and 1-based Int64 CSC equal
those of ereb_assemble_thermal_lse bit for bit, on two calls in a row on one
device, the second on a smaller grid (so the scratch holds the first call's
larger contents), then once more on the first grid with other values."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import ereb
from ereb import api, synth
from ereb.constants import test_constants as tconsts

pytestmark = pytest.mark.gpu

_H2D, _D2H = 1, 2


class _Hip:
    """The few HIP runtime calls a test needs to hand device buffers to the
    library (the runtime the library links)."""

    def __init__(self):
        self.L = ctypes.CDLL("libamdhip64.so.7")
        self.bufs = []

    def to_dev(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), ctypes.c_size_t(max(a.nbytes, 8))) == 0
        self.bufs.append(p)
        if a.nbytes:
            assert self.L.hipMemcpy(p, ctypes.c_void_p(a.ctypes.data),
                                    ctypes.c_size_t(a.nbytes), _H2D) == 0
        return p

    def empty(self, n, dtype):
        return self.to_dev(np.full(n, -7, dtype=dtype))

    def get(self, p, n, dtype):
        out = np.empty(n, dtype=dtype)
        assert self.L.hipDeviceSynchronize() == 0
        if out.nbytes:
            assert self.L.hipMemcpy(ctypes.c_void_p(out.ctypes.data), p,
                                    ctypes.c_size_t(out.nbytes), _D2H) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def _grids(g, seed):
    rng = np.random.default_rng(seed)
    S = g.stag_shape
    f = lambda lo, hi: np.asfortranarray(lo + (hi - lo) * rng.random(S))  # noqa: E731
    tk1, rc = f(150, 300), f(1e6, 4e6)
    KX, KY = f(0.5, 5), f(0.5, 5)
    # zero conductivities leave entries out: the pattern depends on the values
    KX[rng.random(S) < 0.05] = 0.0
    KY[rng.random(S) < 0.05] = 0.0
    return [tk1, rc, KX, KY, f(0, 1e-6), f(0, 1e-7), f(0, 1e-7), f(-1e-8, 1e-8)]


def _check(ctx, g, grids, dt, hip):
    n = g.Ny1 * g.Nx1
    RT = np.zeros(n)
    L = api.assemble_thermal_lse_(ctx, *grids, RT, dt)
    d_in = [hip.to_dev(a.ravel(order="F")) for a in grids]
    cap = 5 * n
    d_RT = hip.empty(n, np.float64)
    d_cp = hip.empty(n + 1, np.int64)
    d_rv = hip.empty(cap, np.int64)
    d_nz = hip.empty(cap, np.float64)
    nnz = ctypes.c_int64(-1)
    rc = ctx.lib.ereb_assemble_thermal_lse_noalloc(ctx.h, *d_in, d_RT, ctypes.c_double(dt),
                                                   d_cp, d_rv, d_nz, ctypes.c_int64(cap),
                                                   ctypes.byref(nnz))
    ctx.check(rc, "assemble_thermal_lse_noalloc")
    k = nnz.value
    assert k == L.nzval.size, (k, L.nzval.size)
    np.testing.assert_array_equal(hip.get(d_RT, n, np.float64), RT)
    np.testing.assert_array_equal(hip.get(d_cp, n + 1, np.int64), L.colptr)
    np.testing.assert_array_equal(hip.get(d_rv, k, np.int64), L.rowval)
    np.testing.assert_array_equal(hip.get(d_nz, k, np.float64), L.nzval)
    return k


def test_noalloc_thermal_equals_allocating_entry_point():
    hip = _Hip()
    big = synth.workload_geometry("adj41", mode=1)
    small = tconsts()
    try:
        sizes = []
        for (g, mat), seed, dt in ((big, 1, 1e11), (small, 2, 3e10), (big, 3, 7e10)):
            with ereb.Context(g, mat, ereb.FP64) as ctx:
                sizes.append((g.Ny1 * g.Nx1, _check(ctx, g, _grids(g, seed), dt, hip)))
            hip.free()
        assert sizes[1][0] < sizes[0][0]  # the second call is the smaller grid
    finally:
        hip.free()


def test_noalloc_rejects_grids_beyond_its_scratch():
    """A grid of more P cells than the static scratch holds is refused before
    any launch."""
    g, mat = synth.workload_geometry("large40", mode=1)  # 562 x 562 fits
    assert g.Ny1 * g.Nx1 <= 1 << 19
    from ereb.constants import Geometry
    gg = Geometry(Nx=800, Ny=800, xsize=140e3, ysize=140e3, Nxmc=1, Nymc=1,
                  rplanet=50e3, rcrust=48e3)
    with ereb.Context(gg, mat, ereb.FP64) as ctx:
        lib = ctx.lib
        one = ctypes.c_void_p(8)  # never dereferenced: the size check comes first
        rc = lib.ereb_assemble_thermal_lse_noalloc(ctx.h, *([one] * 9), ctypes.c_double(1.0),
                                                   one, one, one, ctypes.c_int64(1 << 30), None)
        assert rc == ereb.lib.ERR_ARG
