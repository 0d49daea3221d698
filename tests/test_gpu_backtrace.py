"""GPU parity of backtrace_pressures_rk4! (E:5284-5496) against the CPU
oracle through the C ABI: bit for bit (the kernel keeps the reference's
operation sequence: divisions, no fused products), at the test geometry in
the reference test's regime and with paths of a third of a cell, and at
the 282x282-node workload grid."""
from __future__ import annotations

import numpy as np
import pytest

import ereb
from ereb import api, synth
from ereb.constants import test_constants as tconsts
from oracle import Oracle

pytestmark = pytest.mark.gpu


def _fields(g, seed, amp=1e-9):
    rng = np.random.default_rng(seed)
    S = (g.Ny1, g.Nx1)
    v = [np.asfortranarray((rng.random(S) * 2 - 1) * amp) for _ in range(4)]
    p = [np.asfortranarray(rng.random(S) * 1e6) for _ in range(3)]
    return v, p


def _exact(a, b, what):
    d = a != b
    assert not d.any(), f"{what}: {d.sum()} of {a.size} nodes differ"


@pytest.mark.parametrize("dt", [0.9, 3e11])
def test_backtrace_dropin_bit_exact(dt):
    g, mat = tconsts()
    (vx, vy, vxf, vyf), (pr, ps, pf) = _fields(g, 41)
    want = Oracle(g, mat).backtrace_pressures(pr, ps, pf, vx, vy, vxf, vyf, dt)
    S = (g.Ny1, g.Nx1)
    got = [np.full(S, np.nan, order="F") for _ in range(3)]
    with ereb.Context(g, mat, ereb.FP64) as ctx:
        api.backtrace_pressures_rk4_(ctx, pr, got[0], ps, got[1], pf, got[2], vx, vy, vxf,
                                     vyf, dt)
    for a, b, k in zip(got, want, ("pr0", "ps0", "pf0")):
        _exact(a, b, k)


def test_backtrace_resident_at_workload_grid():
    g, mat = synth.workload_geometry("fp64", mode=1)
    # velocities of up to 0.5 cell per dt, as the workload's stream function
    (vx, vy, vxf, vyf), (pr, ps, pf) = _fields(g, 42, amp=0.5 * g.dx / ereb.DT_LONGEST)
    want = Oracle(g, mat).backtrace_pressures(pr, ps, pf, vx, vy, vxf, vyf, ereb.DT_LONGEST)
    with ereb.Context(g, mat, ereb.FP64) as ctx:
        for n, a in (("pr", pr), ("ps", ps), ("pf", pf), ("vx", vx), ("vy", vy),
                     ("vxf", vxf), ("vyf", vyf)):
            ctx.set_grid(n, a)
        ctx.step_backtrace_pressures(ereb.DT_LONGEST)
        got = [ctx.get_grid(n) for n in ("pr0", "ps0", "pf0")]
    for a, b, k in zip(got, want, ("pr0", "ps0", "pf0")):
        _exact(a, b, k)
