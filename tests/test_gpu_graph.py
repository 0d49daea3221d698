"""The graph-replayed transfer step (ereb_step_transfer) equals the same
launches run eagerly, bit for bit, step after step: marker state and
properties, SUM and node grids, both assembled systems.  Between replays the
host does its own work (reads, a melt-composition scatter that relies on the
sort bookkeeping of the replayed step).  dt and the heating rates change at
every step, as in the reference loop (dt control E:6619, E:6790, E:6933;
calculate_radioactive_heating(timesum) E:6413), and the graph is captured
once: the kernels read them from a device block written before each replay.
A mode change forces a new capture."""
from __future__ import annotations

import numpy as np
import pytest

import ereb
from ereb import synth

pytestmark = pytest.mark.gpu

START_TIME = 2.25e6 * 365.25 * 24 * 3600


def _ctx(g, mat, mk, graphs, precision=ereb.FP64):
    ctx = ereb.Context(g, mat, precision)
    ctx.set_graphs(graphs)
    ctx.set_markers(mk)
    vx, vy = synth.stream_velocity(g)
    S = g.stag_shape
    rng = np.random.default_rng(3)
    grids = {"vx": vx, "vy": vy, "vxf": vx, "vyf": vy,
             "wyx": synth.rotation_rate(g, vx, vy), "ETAP": np.full(S, 1e18),
             "ETAPHI": np.full(S, 1e19), "gx": np.full(S, -1.0), "gy": np.full(S, -2.0),
             "DT": 0.01 * rng.random(S), "tk2": 250 + 50 * rng.random(S),
             "pr0": np.full(S, 1e5), "pf0": np.full(S, 9e4), "DSXX": 1e3 * rng.random(S),
             "DSXY": 1e3 * rng.random(g.basic_shape)}
    for n, a in grids.items():
        ctx.set_grid(n, a)
    return ctx


def _scalars(mat, ts):
    """dt and heating rates of timestep ts: dt shrinks and grows as the
    reference's step control would make it, timesum advances by dt."""
    dt = ereb.DT_LONGEST * (1.0 - 0.37 * ((ts * 5) % 7) / 7.0)
    timesum = START_TIME + ts * ereb.DT_LONGEST
    hrs, hrf = mat.radioactive_heating(timesum)
    return dt, hrs, hrf


def _state(ctx):
    out = {n: ctx.get_marker(n) for n in ereb.MARKER_FIELDS}
    for n in ereb.SUM_FIELDS + ["ETA", "RHO", "tk1", "RX", "KY", "XWS"]:
        out[n] = ctx.get_grid(n)
    for which in (0, 1):
        for k, a in enumerate(ctx.csc(which)):
            out[f"csc{which}.{k}"] = a
    return out


@pytest.mark.parametrize("precision", [ereb.FP64, ereb.MIXED])
def test_graph_replay_equals_eager(precision):
    """Also: the transfer step's fused gathers + RK4 kernel equals the
    separate gather and RK4 kernels (the third context), in both forms."""
    g, mat = synth.workload_geometry("adj41", mode=1)
    mk = synth.make_markers(g, mat, seed=5)
    A = _ctx(g, mat, mk, graphs=False, precision=precision)
    B = _ctx(g, mat, mk, graphs=True, precision=precision)
    try:
        for ts in range(1, 9):
            dt, hrs, hrf = _scalars(mat, ts)
            mode = 9 if ts == 7 else 1  # a mode change captures again
            for c in (A, B):
                c.step_transfer(hrs, hrf, mode, dt, ts)
            if ts == 3:  # reuses or redoes the sort: must match
                for c in (A, B):
                    c.step_melt_composition()
            sa, sb = _state(A), _state(B)
            for k in sa:
                assert np.array_equal(sa[k], sb[k], equal_nan=True), (ts, k)
        # step 1 eager; captures at steps 2, 7 (mode 9) and 8 (mode 1 again)
        assert B.graph_stats() == (3, 7), B.graph_stats()
        # the replayed steps also equal the separate resident calls
        C = _ctx(g, mat, mk, graphs=False, precision=precision)
        for ts in range(1, 9):
            dt, hrs, hrf = _scalars(mat, ts)
            mode = 9 if ts == 7 else 1
            C.step_scatter(hrs, hrf, mode)
            C.step_node_properties()
            C.step_assemble_stokes_darcy(dt)
            C.step_assemble_thermal(dt)
            C.step_gather_stress()
            C.step_gather_temperature(ts)
            C.step_advect_rk4(dt, mode)
            if ts == 3:
                C.step_melt_composition()
        sc = _state(C)
        for k in sc:
            assert np.array_equal(sc[k], sb[k], equal_nan=True), k
        C.close()
    finally:
        A.close()
        B.close()


def test_graph_replay_across_replenishment():
    """A replenishment that appends markers (and grows the resident arrays)
    between replays: the next step captures again and still equals the
    eager launches bit for bit."""
    g, mat = synth.workload_geometry("adj41", mode=1)
    mk = synth.make_markers(g, mat, seed=6)
    # an empty patch for the replenishment to fill
    hole = (np.abs(mk["xm"] - 0.3 * g.xsize) < 0.05 * g.xsize) & \
           (np.abs(mk["ym"] - 0.6 * g.ysize) < 0.05 * g.ysize)
    mk = {k: v[~hole] for k, v in mk.items()}
    hrs, hrf = mat.radioactive_heating(START_TIME)
    dt = ereb.DT_LONGEST
    A = _ctx(g, mat, mk, graphs=False)
    B = _ctx(g, mat, mk, graphs=True)
    try:
        M0 = A.marknum
        for ts in range(1, 7):
            for c in (A, B):
                c.step_transfer(hrs, hrf, 1, dt, ts)
            if ts == 3:
                na, mda, mna = A.step_replenish_markers()
                nb, mdb, mnb = B.step_replenish_markers()
                assert na == nb and na > M0, (na, nb, M0)
                np.testing.assert_array_equal(mna, mnb)
                np.testing.assert_array_equal(mda, mdb)
            sa, sb = _state(A), _state(B)
            for k in sa:
                assert np.array_equal(sa[k], sb[k], equal_nan=True), (ts, k)
    finally:
        A.close()
        B.close()


@pytest.mark.parametrize("M", [0, 1])
def test_step_with_no_marker_and_one_marker(M):
    """The two-stream step and its graph replays with 0 and 1 markers (the
    marker branch then launches nothing, or one workgroup's worth): equal to
    the separate resident calls bit for bit, step after step."""
    g, mat = synth.workload_geometry("adj41", mode=1)
    mk = synth.make_markers(g, mat, seed=9)
    pick = np.arange(M) * 977 % mk["xm"].size
    mk = {k: np.ascontiguousarray(v[pick]) for k, v in mk.items()}
    A = _ctx(g, mat, mk, graphs=False)
    B = _ctx(g, mat, mk, graphs=True)
    C = _ctx(g, mat, mk, graphs=False)
    try:
        for ts in range(1, 5):
            dt, hrs, hrf = _scalars(mat, ts)
            for c in (A, B):
                c.step_transfer(hrs, hrf, 1, dt, ts)
            C.step_scatter(hrs, hrf, 1)
            C.step_node_properties()
            C.step_assemble_stokes_darcy(dt)
            C.step_assemble_thermal(dt)
            C.step_gather_stress()
            C.step_gather_temperature(ts)
            C.step_advect_rk4(dt, 1)
            sa, sb, sc = _state(A), _state(B), _state(C)
            for k in sa:
                assert np.array_equal(sa[k], sb[k], equal_nan=True), (ts, k)
                assert np.array_equal(sa[k], sc[k], equal_nan=True), (ts, k)
        assert A.marknum == M and B.marknum == M
    finally:
        A.close()
        B.close()
        C.close()


@pytest.mark.parametrize("precision", [ereb.FP64, ereb.MIXED])
def test_two_stream_step_at_workload_size(precision):
    """At 2,822,400 markers, where the marker branch and the node pass and
    assemblies really run side by side: the eager two-stream steps and the
    graph replays equal the separate resident calls bit for bit after each
    of four steps with varying dt (marker state, SUM and node grids, both
    assembled systems), in both forms."""
    g, mat = synth.workload_geometry("fp64", mode=1)
    mk = synth.make_markers(g, mat, seed=0, stress=True)
    A = _ctx(g, mat, mk, graphs=False, precision=precision)
    B = _ctx(g, mat, mk, graphs=True, precision=precision)
    C = _ctx(g, mat, mk, graphs=False, precision=precision)
    try:
        for ts in range(1, 5):
            dt, hrs, hrf = _scalars(mat, ts)
            for c in (A, B):
                c.step_transfer(hrs, hrf, 1, dt, ts)
            C.step_scatter(hrs, hrf, 1)
            C.step_node_properties()
            C.step_assemble_stokes_darcy(dt)
            C.step_assemble_thermal(dt)
            C.step_gather_stress()
            C.step_gather_temperature(ts)
            C.step_advect_rk4(dt, 1)
            for n in ("xm", "ym", "tkm", "sxxm", "sxym"):
                a, b, c_ = A.get_marker(n), B.get_marker(n), C.get_marker(n)
                assert np.array_equal(a, c_) and np.array_equal(b, c_), (ts, n)
        sa, sb, sc = _state(A), _state(B), _state(C)
        for k in sc:
            assert np.array_equal(sa[k], sc[k], equal_nan=True), k
            assert np.array_equal(sb[k], sc[k], equal_nan=True), k
        assert B.graph_stats()[1] == 3  # steps 2-4 replayed
    finally:
        A.close()
        B.close()
        C.close()
