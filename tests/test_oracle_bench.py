"""The benchmark's CPU baseline (oracle/ereb_cpu_step.c, the oracle built at
-O3 with OpenMP in the two gathers, where the reference has @threads) runs
the same arithmetic as the oracle's functions called one by one: two
timesteps of the transfer path on the 41x41-node workload give the same
marker state bit for bit, whatever the thread count."""
from __future__ import annotations

import numpy as np

import ereb
import oracle as O
from ereb import synth

START_TIME = 2.25e6 * 365.25 * 24 * 3600


def _inputs(g, mat, orc, mk, hrs, hrf):
    S, B = g.stag_shape, g.basic_shape
    props, sums = orc.marker_to_nodes(mk, hrs, hrf, 1)
    ob, _ = orc.basic_node_properties(sums[:8], [np.zeros(B, order="F")] * 7,
                                      np.zeros(B, np.uint8, order="F"))
    op = orc.p_node_properties(sums[20:30], [np.zeros(S, order="F")] * 10)
    vx, vy = synth.stream_velocity(g)
    grids = {"vx": vx, "vy": vy, "wyx": synth.rotation_rate(g, vx, vy)}
    grids["ETAP"], grids["ETAPHI"] = synth.bulk_viscosity(ob[1], np.maximum(op[8], 1e-4))
    grids["gx"], grids["gy"] = synth.disk_gravity(g)
    grids["DT"] = synth.smooth_field(g, S, 0.05, 2, 1)
    grids["tk2"] = np.asfortranarray(op[7] + grids["DT"])
    grids["DSXX"] = synth.smooth_field(g, S, 1e3, 3, 2)
    grids["DSXY"] = synth.smooth_field(g, B, 1e3, 2, 3)
    return props, grids


def test_cpu_baseline_harness_equals_oracle_steps():
    g, mat = synth.workload_geometry("adj41", mode=1)
    mk = synth.make_markers(g, mat, seed=3)
    hrs, hrf = mat.radioactive_heating(START_TIME)
    orc = O.Oracle(g, mat)
    props, grids = _inputs(g, mat, orc, mk, hrs, hrf)
    mk.update(props)
    for f in O.BENCH_MARKERS:
        mk[f] = np.ascontiguousarray(mk[f], dtype=np.float64)
    dt = ereb.DT_LONGEST
    a = {k: v.copy() for k, v in mk.items()}
    done, sec, threads = O.bench_transfer(g, mat, a, grids, hrs, hrf, 1, dt, 2, 1e9)
    assert done == 2 and sec > 0 and threads >= 1
    # the same two steps through the oracle's functions
    b = {k: v.copy() for k, v in mk.items()}
    for _ in range(2):
        p, _ = orc.marker_to_nodes(b, hrs, hrf, 1)
        b.update(p)
        b["sxxm"], b["sxym"] = orc.update_marker_stress(b["xm"], b["ym"], b["sxxm"], b["sxym"],
                                                        grids["DSXX"], grids["DSXY"])
        b["tkm"] = orc.update_marker_temperature(b["xm"], b["ym"], b["tkm"], grids["DT"],
                                                 grids["tk2"], 2)
        r = orc.move_markers_rk4(b["xm"], b["ym"], b["tm"], b["tkm"], b["phim"], b["sxym"],
                                 b["sxxm"], grids["vx"], grids["vy"], grids["vx"], grids["vy"],
                                 grids["wyx"], grids["tk2"], dt, 1)
        b["xm"], b["ym"], b["tkm"], b["sxym"], b["sxxm"] = r
    for f in O.BENCH_MARKERS:
        assert np.array_equal(a[f], b[f]), f
