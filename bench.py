"""Benchmark: marker-steps/s of the marker-in-cell transfer pipeline on one
MI355X.

Workload (default "fp64"): 281x281 basic nodes, 140 km box, 6x6 jittered
markers per cell = 2,822,400 markers, planetesimal R = 50 km in sticky air,
stream-function velocity at Courant 0.5, dt near 1e11 s, property mode 1.
A step is one timestep of the transfer path in the reference loop order
(src/Erebus.jl:6376-7053) with every input resident in HBM:
  tile sort with the marker properties in its key pass, scatter (E:6419-6517)
  node normalisation + insulating BC (E:6522-6604)
  Stokes-Darcy assembly to CSC (E:6680) and thermal assembly (E:6922)
  stress gather (E:6980), temperature gather (E:7004), RK4 advection (E:7036)
dt and the heating rates change at every step as in the reference loop (dt
control E:6619, E:6790, E:6933; calculate_radioactive_heating(timesum)
E:6413).  The step's launches go to two streams: the gathers + RK4 kernel
runs beside the node pass and the assemblies once the scatter is done.
The host LU solves are outside the step (they stay on the host in the
reference's design); the CSC matrices stay in HBM.  The grids that the
solves and grid-only operations would produce (DT, DSXX, DSXY, tk2,
pressures, gravity, ETAP/ETAPHI) are fixed synthetic fields
(erebus.jl_amd/ereb/synth.py).

Prints one JSON line.  `value` is the Float64 form.  `roofline` is for the
kernel class with the longest device time per step (HIP events on the
library's stream), with the algorithmic bytes of the reference functions it
implements; `rooflines` gives every class, `step_roofline` the whole step's
algorithmic bytes against HBM peak.  `mixed` runs the same steps in the
mixed-precision form on identical inputs and reports its throughput and the
per-field error against the Float64 state after the same steps.  `adjoint`
times the gradient of sum(tk1^2) after 20 steps (1,000,000 markers, 201x201
nodes, checkpoints every 5 steps) against the forward map and checks it
against central finite differences on 64 markers.  `next_rows` gives the
mean device time and algorithmic bandwidth of the transfers outside the
step.  `large` runs the 20,070,400-marker workloads (R = 40 and 65 km) in
both forms and times one checkpointed adjoint step there.  `host_lu` times the host factorisation and solve of the assembled
systems (scipy's SuperLU; the reference uses UMFPACK).  `cpu_baseline` times
the C restatement of the reference CPU path (the oracle, -O3, OpenMP where
the reference threads) on a bounded sample of the Float64 workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "erebus.jl_amd"))

import numpy as np  # noqa: E402

import ereb  # noqa: E402
from ereb import precision, synth  # noqa: E402

METRIC = "marker-steps/s, 2.8M markers: FP64 vs mixed-precision; adjoint/forward ratio"
# how ereb_step_transfer issues a step (the library's default)
STEP_LAUNCHES = ("eager on two streams: gathers + RK4 beside the node pass and assemblies "
                 "once the scatter is done (graph replay off)")
HBM_PEAK_GBS = 8000.0
START_TIME = 2.25e6 * 365.25 * 24 * 3600  # C:339

# Algorithmic bytes per kernel class: every Julia-visible input of the
# reference functions a class implements read once and every output written
# once (Float64 8 B, mixed-form properties 4 B, tm 4 B, YNY 1 B), per marker
# and per node of the Ny1 x Nx1 grid, plus 8 B per stored CSC value.
#   sort    compute_marker_properties! (E:700): r tm tkm phim X_W rhocp, w 7
#   scatter marker_to_*_nodes! (E:1605-1863): r xm ym + 18 staged values,
#           w 30 SUM grids
#   combine the SUM grids' tile overlap (no reference counterpart: 0; in the
#           transfer step part of the nodes class, k_combine_nodes)
#   nodes   compute_*_node_properties! + insulating tk1 (E:1970-2189, E:2432):
#           r 30 SUM, w 27 grids + YNY
#   rk4     update_marker_stress! (E:4354: r xm ym sxx sxy, w sxx sxy; r DSXX
#           DSXY) + update_marker_temperature! (E:4782: r xm ym tkm, w tkm;
#           r DT) + move_markers_rk4! (E:4990: r 7, w 5; r vx vy vxf vyf wyx
#           tk2): 32 + 48 + 92 B per marker, 8 + 16 + 48 B per node
#   asm_sd  assemble_hydromechanical_lse! (E:3089): r 20 grids, w R (6 per
#           cell) and nzval
#   asm_t   assemble_thermal_lse! (E:4526): r 8 grids, w RT and nzval
ALG_BYTES = {
    "fp64": {"sort": (92, 0), "scatter": (160, 240), "combine": (0, 0), "nodes": (0, 457),
             "rk4": (172, 72), "asm_sd": (0, 208), "asm_t": (0, 72)},
    # mixed form: the 16 material and property vectors 4 B, the marker state
    # (xm ym tkm phim sxxm sxym) and every grid 8 B
    #   sort    r tm 4 + tkm phim 16 + X_W rhocp 8, w 7 x 4
    #   scatter r xm ym 16 + phim sxx sxy 24 + 15 properties x 4
    "mixed": {"sort": (56, 0), "scatter": (100, 240), "combine": (0, 0), "nodes": (0, 457),
              "rk4": (172, 72), "asm_sd": (0, 208), "asm_t": (0, 72)},
}
STEP_CLASSES = ["sort", "scatter", "combine", "nodes", "rk4", "asm_sd", "asm_t"]
# kernel-name prefix of each class's main kernel, per form, in the rocprofv3
# traffic summary (profiles/hbm_traffic_<workload>.json, made by
# erebus.jl_amd/ereb/rocprof_summary.py from FETCH_SIZE / WRITE_SIZE runs)
TRAFFIC_KERNEL = {
    "fp64": {"scatter": "void ereb::k_scatter<double>", "rk4": "void ereb::k_rk4<1, true>",
             "sort": "void ereb::k_fine_count<ereb::PropArgs<double>",
             "nodes": "ereb::k_combine_nodes", "asm_sd": "void ereb::k_asm_rows<0>",
             "asm_t": "void ereb::k_asm_rows<1>"},
    "mixed": {"scatter": "void ereb::k_scatter<float>", "rk4": "void ereb::k_rk4<2, true>",
              "sort": "void ereb::k_fine_count<ereb::PropArgs<float>",
              "nodes": "ereb::k_combine_nodes", "asm_sd": "void ereb::k_asm_rows<0>",
              "asm_t": "void ereb::k_asm_rows<1>"},
}


def step_scalars(mat, k: int):
    """dt and heating rates of timestep k: dt varies within [0.6, 1] x 1e11 s
    as the reference's Courant and temperature-change limits make it vary,
    and timesum advances by dt."""
    dt = ereb.DT_LONGEST * (1.0 - 0.4 * ((k * 5) % 7) / 7.0)
    hrs, hrf = mat.radioactive_heating(START_TIME + k * ereb.DT_LONGEST)
    return dt, hrs, hrf


def log(*a):
    print(*a, file=sys.stderr, flush=True)


INPUT_GRIDS = ["vx", "vy", "vxf", "vyf", "wyx", "ETAP", "ETAPHI", "gx", "gy", "DT", "tk2",
               "pr0", "pf0", "DMP", "HA", "HS", "DHP", "DSXX", "DSXY"]


def make_inputs(g, mat, mk):
    """The synthetic grids that the solves and grid-only operations would
    produce, derived once from a Float64 scatter + node pass (ETA and PHI for
    the bulk viscosity, tk1 for tk2), so that every form runs on identical
    inputs."""
    vx, vy = synth.stream_velocity(g)
    hrs, hrf = mat.radioactive_heating(START_TIME)
    with ereb.Context(g, mat, ereb.FP64) as ctx:
        ctx.set_markers(mk)
        ctx.step_scatter(hrs, hrf, mat.marker_property_mode)
        ctx.step_node_properties()
        ETA, PHI, tk1 = ctx.get_grid("ETA"), ctx.get_grid("PHI"), ctx.get_grid("tk1")
    grids = {"vx": vx, "vy": vy, "vxf": vx, "vyf": vy, "wyx": synth.rotation_rate(g, vx, vy)}
    grids["ETAP"], grids["ETAPHI"] = synth.bulk_viscosity(ETA, np.maximum(PHI, 1e-4))
    grids["gx"], grids["gy"] = synth.disk_gravity(g)
    S = g.stag_shape
    grids["DT"] = synth.smooth_field(g, S, 0.05, 2, 1)
    grids["tk2"] = np.asfortranarray(tk1 + grids["DT"])
    grids["pr0"] = np.full(S, mat.psurface, order="F") + synth.smooth_field(g, S, 1e5, 1, 1)
    grids["pf0"] = np.full(S, mat.psurface, order="F") + synth.smooth_field(g, S, 9e4, 1, 1)
    for n in ("DMP", "HA", "HS", "DHP"):
        grids[n] = np.zeros(S, order="F")
    grids["DSXX"] = synth.smooth_field(g, S, 1e3, 3, 2)
    grids["DSXY"] = synth.smooth_field(g, g.basic_shape, 1e3, 2, 3)
    return grids, hrs, hrf


def setup(ctx: ereb.Context, mk, grids):
    ctx.set_markers(mk)
    for n in INPUT_GRIDS:
        ctx.set_grid(n, grids[n])
    ctx.sync()


def step(ctx, mat, mode, timestep):
    """One transfer step: scatter, node properties, Stokes-Darcy and thermal
    assembly, stress and temperature gathers, RK4 (ereb_step_transfer, on
    the context's two streams; everything on one stream while profiling),
    with this timestep's dt and heating rates."""
    dt, hrs, hrf = step_scalars(mat, timestep)
    ctx.step_transfer(hrs, hrf, mode, dt, timestep)


def cpu_baseline(g, mat, mk, grids, hrs, hrf, steps: int, budget_s: float = 25.0):
    """The oracle (C restatement of the reference CPU path) built at -O3 with
    OpenMP in the gathers, where the reference has @threads, timed inside C
    over whole timesteps of the same workload in the same step order (no
    Python in the timed loop; ereb_cpu_step.c)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc_mod

    mode = mat.marker_property_mode
    mk = {k: np.array(v, copy=True) for k, v in mk.items()}
    # the marker properties of the first step, untimed (the marker vectors
    # they are written to must exist as Float64 arrays)
    props, _ = orc_mod.Oracle(g, mat).marker_to_nodes(mk, hrs, hrf, mode)
    mk.update(props)
    for f in orc_mod.BENCH_MARKERS:
        mk[f] = np.ascontiguousarray(mk[f], dtype=np.float64)
    done, el, threads = orc_mod.bench_transfer(g, mat, mk, grids, hrs, hrf, mode,
                                               ereb.DT_LONGEST, steps, budget_s)
    M = mk["xm"].size
    return {"value": M * done / el, "unit": "marker-steps/s", "cores": threads, "kind": "port",
            "sample": f"{done} full timestep(s) of the {M:,}-marker workload, same step "
                      f"order, C restatement of the reference CPU path (-O3, no FMA, OpenMP "
                      f"with {threads} threads in the two gathers where the reference has "
                      f"@threads; scatter, assemblies and RK4 single-threaded as in the "
                      f"reference), timed inside C, {el:.1f} s"}


def run_form(g, mat, mk, grids, hrs, hrf, precision, steps, warmup):
    """Warmup + timed steps of one precision form, then the same number of
    steps with HIP events around every kernel class (kept out of the timed
    steps, whose launches they would separate); returns the live context
    (state after warmup + 2 x steps), marker-steps/s and ms per step of the
    timed steps, and the mean duration of each kernel class."""
    mode = mat.marker_property_mode
    ctx = ereb.Context(g, mat, precision)
    setup(ctx, mk, grids)
    ts = 2
    for _ in range(warmup):
        step(ctx, mat, mode, ts)
        ts += 1
    ctx.sync()
    cap0, _ = ctx.graph_stats()
    t0 = time.perf_counter()
    for _ in range(steps):
        step(ctx, mat, mode, ts)
        ts += 1
    ctx.sync()
    el = time.perf_counter() - t0
    recaptures = ctx.graph_stats()[0] - cap0
    ctx.profile(True)
    ctx.profile_reset()
    for _ in range(steps):
        step(ctx, mat, mode, ts)
        ts += 1
    ctx.sync()
    kern = {}
    for k in ereb.lib.KERNEL_CLASSES:
        ms, n = ctx.profile_get(k)
        kern[k] = round(ms / max(n, 1), 5)
    ctx.profile(False)
    M = mk["xm"].size
    return ctx, M * steps / el, el * 1e3 / steps, kern, recaptures


def copy_bandwidth(nbytes: int = 1 << 30, reps: int = 5) -> float:
    """Device-to-device copy rate of this GPU in GB/s (bytes read plus
    bytes written per second), HIP events around hipMemcpy on the null
    stream: the practical ceiling next to the 8 TB/s vendor peak."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so.7")  # the runtime the library links
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    e0, e1 = ctypes.c_void_p(), ctypes.c_void_p()
    ok = hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    ok = ok and hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    ok = ok and hip.hipEventCreate(ctypes.byref(e0)) == 0 and hip.hipEventCreate(ctypes.byref(e1)) == 0
    gbps = float("nan")
    if ok:
        hip.hipMemset(a, 1, ctypes.c_size_t(nbytes))
        D2D = 3  # hipMemcpyDeviceToDevice
        hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), D2D)
        hip.hipEventRecord(e0, None)
        for _ in range(reps):
            hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), D2D)
        hip.hipEventRecord(e1, None)
        hip.hipEventSynchronize(e1)
        ms = ctypes.c_float()
        if hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1) == 0 and ms.value > 0:
            gbps = 2.0 * nbytes * reps / (ms.value * 1e-3) / 1e9
    for e in (e0, e1):
        if e.value:
            hip.hipEventDestroy(e)
    for p in (a, b):
        if p.value:
            hip.hipFree(p)
    return gbps


def _finite(x: float, nd: int = 1):
    return round(x, nd) if x == x else None


def _traffic(workload: str, form: str, kclass: str):
    """HBM bytes per launch of a class's main kernel from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (each scaled by what the counter
    reads for a plain copy of known size), when one exists."""
    pre = TRAFFIC_KERNEL.get(form, {}).get(kclass)
    pmc = os.path.join(ROOT, "profiles", f"hbm_traffic_{workload}.json")
    if not pre or not os.path.exists(pmc):
        return None
    for k, v in json.load(open(pmc)).get("kernels", {}).items():
        if k.startswith(pre):
            return v["bytes_per_launch"]
    return None


def class_bytes(form, kclass, M, N, nnz):
    bm, bn = ALG_BYTES[form][kclass]
    extra = {"asm_sd": 8 * nnz[0], "asm_t": 8 * nnz[1]}.get(kclass, 0)
    return M * bm + N * bn + extra


def rooflines(g, M, kern, form, workload, nnz, ms_per_step):
    """Per-class rooflines (algorithmic bytes over the class's mean device
    time per step), the dominant class first, and the whole step's."""
    N = g.Nx1 * g.Ny1
    out = {}
    for k in STEP_CLASSES:
        ms = kern.get(k, 0.0)
        if ms <= 0:
            continue
        nb = class_bytes(form, k, M, N, nnz)
        if nb == 0:
            continue
        ach = nb / (ms * 1e-3) / 1e9
        out[k] = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                  "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                  "traffic": _traffic(workload, form, k), "ms": ms,
                  "algorithmic_bytes": nb}
    dom = max(STEP_CLASSES, key=lambda k: kern.get(k, 0.0))
    tot = sum(class_bytes(form, k, M, N, nnz) for k in STEP_CLASSES)
    step_gbs = tot / (ms_per_step * 1e-3) / 1e9
    step = {"algorithmic_bytes": tot, "ms_per_step": ms_per_step,
            "achieved": round(step_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(step_gbs / HBM_PEAK_GBS, 4), "target_frac": 0.60,
            "kernel_ms_sum": round(sum(kern.get(k, 0.0) for k in STEP_CLASSES), 4)}
    return dom, out, step


# algorithmic bytes per marker of the transfers outside the timed step
# (every Julia-visible marker vector read or written once; tm as 4 bytes)
NEXT_ROW_BYTES = {
    # xm ym tm tkm etatotalm (r) + etatotalm etavpm (w)
    "viscosity": 2 * 8 + 4 + 8 + 8 + 2 * 8,
    # xm ym tm phim (r) + phim (w)
    "porosity": 2 * 8 + 4 + 8 + 8,
    # xm ym XWsolidm0 (r) and the sort order (4)
    "melt": 3 * 8 + 4,
    # xm ym tm XWsolidm0 phim pfm0 (r) + pfm0 XWsolidm phinewm XWsolidm0 phim (w), perm
    "reaction": 2 * 8 + 4 + 3 * 8 + 5 * 8 + 4,
    # xm ym read by the two lattice passes
    "replenish": 2 * 2 * 8,
}


def next_rows_leg(ctx, g, mat, hrs, hrf, reps: int = 5):
    """Mean device time of each remaining marker<->grid transfer on the
    resident Float64 state (HIP events on the library stream), called in
    the reference's order after a scatter (so the P-grid scatters reuse its
    sort, as at E:6629 and E:7019), with synthetic YNY_inv_ETA, APHI, pf,
    pr and ps grids; the pressure back-trace is per node."""
    rng = np.random.default_rng(5)
    B, S = g.basic_shape, g.stag_shape
    ctx.set_grid("YNY_inv_ETA", 1e-19 * (1 + rng.random(B)))
    ctx.set_yny((rng.random(B) < 0.3).astype(np.uint8))
    ctx.set_grid("APHI", 1e-12 * (rng.random(S) - 0.5))
    for n in ("pf", "pr", "ps"):
        ctx.set_grid(n, mat.psurface + 1e5 * rng.random(S))
    hyd = ereb.Hydration()
    dt = ereb.DT_LONGEST
    M0 = ctx.marknum
    out = {}
    ctx.step_scatter(hrs, hrf, mat.marker_property_mode)

    def calls():
        ctx.step_marker_viscosity()
        ctx.step_thermochemical_reaction(dt, 2, 3, hyd=hyd)
        ctx.step_marker_porosity(dt, hyd=hyd)
        ctx.step_melt_composition()
        ctx.step_backtrace_pressures(dt)
        ctx.step_replenish_markers()

    calls()  # first launches load the code objects; replenishment may grow the arrays
    M = ctx.marknum
    ctx.sync()
    ctx.profile(True)
    ctx.profile_reset()
    for _ in range(reps):
        calls()
    ctx.sync()
    for k in ("viscosity", "reaction", "porosity", "melt", "backtrace", "replenish"):
        ms, n = ctx.profile_get(k)
        ms /= max(n, 1)
        if k == "backtrace":
            nbytes = 10 * 8 * g.Ny1 * g.Nx1
        else:
            nbytes = M * NEXT_ROW_BYTES[k] + (3 * 8 * g.Ny1 * g.Nx1 if k in ("melt", "reaction")
                                             else 0)
        out[k] = {"ms": round(ms, 4), "GB_per_s": round(nbytes / (ms * 1e-3) / 1e9, 1)}
    ctx.profile(False)
    out["added_markers"] = ctx.marknum - M0
    return out


def handoff_leg(ctx, reps: int = 5):
    """Solver hand-off of the last assembled Stokes-Darcy system (host wall
    time, device-to-host copies included): the pattern-change query, a full
    fetch (1-based Int64 colptr/rowval, nzval, R) into pageable arrays, and
    the values-only fetch (nzval, R) into pageable and page-locked arrays."""
    which = ereb.SYS_STOKES_DARCY
    n, nnz = ctx.csc_size(which)
    full_b = 8 * (n + 1) + 16 * nnz + 8 * n
    vals_b = 8 * nnz + 8 * n
    page = (np.empty(n + 1, np.int64), np.empty(nnz, np.int64), np.empty(nnz),
            np.empty(n))
    pinned = (None, None, ctx.host_array(nnz, np.float64), ctx.host_array(n, np.float64))

    def timed(f):
        f()
        t0 = time.perf_counter()
        for _ in range(reps):
            f()
        return (time.perf_counter() - t0) / reps * 1e3

    q = timed(lambda: ctx.csc_pattern_changed(which))
    full = timed(lambda: ctx.csc(which, out=page))
    vp = timed(lambda: ctx.csc(which, pattern=False, out=(None, None) + page[2:]))
    vq = timed(lambda: ctx.csc(which, pattern=False, out=pinned))
    return {"system": f"Stokes-Darcy n={n:,} nnz={nnz:,}",
            "pattern_query_ms": round(q, 4),
            "full_fetch_pageable": {"ms": round(full, 3), "GB_per_s": round(full_b / full / 1e6, 2)},
            "values_pageable": {"ms": round(vp, 3), "GB_per_s": round(vals_b / vp / 1e6, 2)},
            "values_pinned": {"ms": round(vq, 3), "GB_per_s": round(vals_b / vq / 1e6, 2)}}


def full_timestep_leg(ctx, mat, hrs, hrf, reps: int = 5):
    """Every GPU row of one reference timestep in the reference's order
    (E:6419-7065, host solves excluded): scatter, thermochemical reaction,
    node properties, both assemblies, viscosity, stress and temperature
    gathers, porosity, melt composition, RK4, pressure back-trace,
    replenishment; eager launches, host wall time per timestep (the
    replenishment reads the new marker count back, which synchronises).
    Run after next_rows_leg, which sets the grids these rows read."""
    hyd = ereb.Hydration()
    dt = ereb.DT_LONGEST
    mode = mat.marker_property_mode

    def timestep(ts):
        ctx.step_scatter(hrs, hrf, mode)
        ctx.step_thermochemical_reaction(dt, ts, 3, hyd=hyd)
        ctx.step_node_properties()
        ctx.step_assemble_stokes_darcy(dt)
        ctx.step_assemble_thermal(dt)
        ctx.step_marker_viscosity()
        ctx.step_gather_stress()
        ctx.step_gather_temperature(ts)
        ctx.step_marker_porosity(dt, hyd=hyd)
        ctx.step_melt_composition()
        ctx.step_advect_rk4(dt, mode)
        ctx.step_backtrace_pressures(dt)
        ctx.step_replenish_markers(maps=False)  # mdis, mnum stay on the device

    timestep(2)
    ctx.sync()
    M = ctx.marknum
    t0 = time.perf_counter()
    for k in range(reps):
        timestep(3 + k)
    ctx.sync()
    ms = (time.perf_counter() - t0) / reps * 1e3
    return {"ms_per_timestep": round(ms, 4), "marker_steps_per_s": M / (ms * 1e-3),
            "markers": M, "order": "scatter, reaction, node properties, assemblies, "
            "viscosity, gathers, porosity, melt composition, RK4, back-trace, replenishment"}


BREAKS = (263.0, 268.0, 273.0, 278.0, 410.0)  # kinks and steps of T-dependent properties


def large_leg(steps: int = 5, warmup: int = 2):
    """The 20,070,400-marker workloads (561x561 basic nodes, 8x8 markers per
    cell, planetesimal R = 40 and 65 km): the transfer step of both forms
    (marker-steps/s, the dominant class's roofline and the step's HBM
    fraction) and one checkpointed adjoint step of the Float64 form against
    its forward step."""
    out = {}
    for name in ("large40", "large65"):
        g, mat = synth.workload_geometry(name, mode=1)
        mode = mat.marker_property_mode
        mk = synth.make_markers(g, mat, seed=0)
        M = mk["xm"].size
        grids, hrs, hrf = make_inputs(g, mat, mk)
        res = {"markers": M, "grid": [g.Ny, g.Nx], "planet_radius_km": g.rplanet / 1e3}
        for prec, form in ((ereb.FP64, "fp64"), (ereb.MIXED, "mixed")):
            ctx, v, ms, kern, recap = run_form(g, mat, mk, grids, hrs, hrf, prec, steps, warmup)
            nnz = (ctx.csc_size(ereb.SYS_STOKES_DARCY)[1], ctx.csc_size(ereb.SYS_THERMAL)[1])
            dom, roofs, step_roof = rooflines(g, M, kern, form, name, nnz, ms)
            res[form] = {"value": v, "ms_per_step": ms,
                         "roofline": dict(roofs[dom], kernel_class=dom),
                         "step_roofline": step_roof, "kernel_ms": kern,
                         "launches": STEP_LAUNCHES}
            ctx.close()
            log(f"[bench] {name} {form} {ms:.3f} ms/step")
        with ereb.Context(g, mat, ereb.FP64) as ctx:
            vx, vy = synth.stream_velocity(g)
            ctx.set_markers(mk)
            for n, a in (("vx", vx), ("vy", vy), ("vxf", vx), ("vyf", vy),
                         ("wyx", np.zeros(g.basic_shape, order="F"))):
                ctx.set_grid(n, a)
            dt = ereb.DT_LONGEST
            ctx.adjoint_tk1_loss(1, 1, dt, mode, hrs, hrf)  # warm
            reps = 3
            t0 = time.perf_counter()
            for _ in range(reps):
                ctx.adjoint_tk1_loss(1, 1, dt, mode, hrs, hrf, grad=False)
            t_fwd = (time.perf_counter() - t0) / reps
            t0 = time.perf_counter()
            for _ in range(reps):
                ctx.adjoint_tk1_loss(1, 1, dt, mode, hrs, hrf)
            t_adj = (time.perf_counter() - t0) / reps
        res["adjoint_step"] = {"forward_ms": t_fwd * 1e3, "adjoint_ms": t_adj * 1e3,
                               "adjoint_forward_ratio": t_adj / t_fwd}
        out[name] = res
    return out


# relative steps of the fourth-order central differences of the adjoint
# checks (adjoint_leg)
FD_STEP = {"tkm": 1e-5, "phim": 1e-3}


def adjoint_leg(nsteps: int = 20, every: int = 5, nfd: int = 64, workload: str = "adj"):
    """Gradient of L = sum(tk1^2) after nsteps steps with respect to the
    initial marker temperature and porosity: time of the forward map, time
    of the checkpointed adjoint (forward with checkpoints + reverse sweep),
    and fourth-order central differences, (8 (L(x+h) - L(x-h)) - (L(x+2h) -
    L(x-2h))) / 12h, on nfd random planet markers 2 K or more from the
    temperature breakpoints of the property laws.  Relative steps 1e-5 for
    tkm (2h = 6e-3 K) and 1e-3 for phim: one marker moves tk1 through node
    sums of ~100 markers, so the quotient's rounding noise (~1.5 u / h, u
    the relative rounding of a loss difference) is ~1e-7 at these steps,
    and the truncation error (O(h^4)) below 1e-12."""
    g, mat = synth.workload_geometry(workload, mode=1)
    mode = mat.marker_property_mode
    mk = synth.make_markers(g, mat, seed=0)
    M = mk["xm"].size
    hrs, hrf = mat.radioactive_heating(START_TIME)
    vx, vy = synth.stream_velocity(g)
    dt = ereb.DT_LONGEST
    with ereb.Context(g, mat, ereb.FP64) as ctx:
        ctx.set_markers(mk)
        for n, a in (("vx", vx), ("vy", vy), ("vxf", vx), ("vyf", vy),
                     ("wyx", np.zeros(g.basic_shape, order="F"))):
            ctx.set_grid(n, a)
        # the gradient goes to page-locked host arrays, as the solver hand-off's
        # fetches do (a pageable destination adds ~1 ms of staged copying)
        gout = (ctx.host_array(M, np.float64), ctx.host_array(M, np.float64))
        ctx.adjoint_tk1_loss(nsteps, every, dt, mode, hrs, hrf, grad=True, out=gout)  # warm
        reps = 3
        t0 = time.perf_counter()
        for _ in range(reps):
            L, _, _, _ = ctx.adjoint_tk1_loss(nsteps, every, dt, mode, hrs, hrf, grad=False)
        t_fwd = (time.perf_counter() - t0) / reps
        t0 = time.perf_counter()
        for _ in range(reps):
            L2, gT, gP, _ = ctx.adjoint_tk1_loss(nsteps, every, dt, mode, hrs, hrf, out=gout)
        t_adj = (time.perf_counter() - t0) / reps
        gT, gP = np.array(gT), np.array(gP)
        rng = np.random.default_rng(12345)
        ok = np.flatnonzero((mk["tm"] < 3) & np.all(
            np.abs(mk["tkm"][:, None] - np.array(BREAKS)[None, :]) >= 2.0, axis=1))
        idx = rng.choice(ok, nfd, replace=False)
        fd_err = {}
        for name, grad in (("tkm", gT), ("phim", gP)):
            fd = np.empty(nfd)
            for q, m in enumerate(idx):
                h = FD_STEP[name] * abs(mk[name][m])
                tk = {}
                for k in (1, -1, 2, -2):
                    a = mk[name].copy()
                    a[m] += k * h
                    ctx.set_marker(name, a)
                    tk[k] = ctx.adjoint_tk1_loss(nsteps, every, dt, mode, hrs, hrf,
                                                 grad=False, want_tk1=True)[3]
                dL = lambda u, v: float(np.sum((tk[u] - tk[v]) * (tk[u] + tk[v])))  # noqa: E731
                fd[q] = (8.0 * dL(1, -1) - dL(2, -2)) / (12.0 * h)
            ctx.set_marker(name, mk[name])
            ad = grad[idx]
            fd_err[name] = {
                "max_rel_err": float(np.max(np.abs(fd - ad) / np.abs(ad))),
                "normwise_rel_err": float(np.max(np.abs(fd - ad)) / np.max(np.abs(ad))),
            }
    return {
        "workload": f"{workload}: {g.Nx}x{g.Ny} basic nodes, {g.Nxmc}x{g.Nymc} jittered "
                    f"markers/cell ({M:,} markers), prescribed stream-function velocity",
        "loss": "sum over P nodes of tk1^2 after the last step (tk2 := tk1)",
        "gradient_destination": "page-locked host arrays (ereb_host_alloc)",
        "steps": nsteps, "checkpoint_every": every, "markers": M,
        "forward_ms": t_fwd * 1e3, "adjoint_ms": t_adj * 1e3,
        "adjoint_forward_ratio": t_adj / t_fwd,
        "forward_marker_steps_per_s": M * nsteps / t_fwd,
        "fd_check": {"markers": nfd, "h_rel": FD_STEP, **fd_err},
    }


def host_lu_leg(systems):
    """Host factorisation and solve of the last assembled systems, as the
    reference solves them on the host (UMFPACK through LinearSolve, E:6719-6749;
    LT \\ RT, E:6925).  UMFPACK is not installable here: scipy's SuperLU
    (splu, COLAMD ordering) stands in, labelled as such."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla

    out = {"solver": "SuperLU (scipy.sparse.linalg.splu, COLAMD), not UMFPACK",
           "threads": 1}
    for name, (cp, rv, nz, R) in systems.items():
        n = R.size
        A = sp.csc_matrix((nz, rv - 1, cp - 1), shape=(n, n))
        t0 = time.perf_counter()
        lu = sla.splu(A, permc_spec="COLAMD")
        t1 = time.perf_counter()
        x = lu.solve(R)
        t2 = time.perf_counter()
        res = float(np.linalg.norm(A @ x - R) / max(np.linalg.norm(R), 1e-300))
        out[name] = {"n": n, "nnz": int(A.nnz), "factor_s": round(t1 - t0, 3),
                     "solve_s": round(t2 - t1, 4), "rel_residual": float(f"{res:.3e}")}
        log(f"[bench] host LU {name}: n={n:,} factor {t1 - t0:.2f} s, solve {t2 - t1:.3f} s")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="fp64", choices=sorted(synth.WORKLOADS))
    ap.add_argument("--cpu-steps", type=int, default=12)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-next-rows", action="store_true")
    ap.add_argument("--no-host-lu", action="store_true")
    ap.add_argument("--no-large", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path "
              f"(--gpus {a.gpus} requested)", file=sys.stderr)
        sys.exit(2)

    g, mat = synth.workload_geometry(a.workload, mode=1)
    mode = mat.marker_property_mode
    log(f"[bench] workload {a.workload}: {g.Nx}x{g.Ny} nodes, {g.Nxmc}x{g.Nymc} markers/cell")
    mk = synth.make_markers(g, mat, seed=0)
    M = mk["xm"].size
    grids, hrs, hrf = make_inputs(g, mat, mk)
    log(f"[bench] {M:,} markers; FP64 form: warmup {a.warmup}, {a.steps} timed steps")
    ctx64, value, ms, kern, recap = run_form(g, mat, mk, grids, hrs, hrf, ereb.FP64, a.steps,
                                             a.warmup)
    log(f"[bench] FP64 {ms:.3f} ms/step")
    nnz = (ctx64.csc_size(ereb.SYS_STOKES_DARCY)[1], ctx64.csc_size(ereb.SYS_THERMAL)[1])
    dom, roofs, step_roof = rooflines(g, M, kern, "fp64", a.workload, nnz, ms)
    res = {
        "metric": METRIC,
        "value": value,
        "unit": "marker-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{a.workload}: {g.Nx}x{g.Ny} basic nodes, "
                               f"{g.Nxmc}x{g.Nymc} jittered markers/cell ({M:,} markers), "
                               f"planetesimal R={g.rplanet / 1e3:.0f} km in 140 km box, "
                               f"stream-function velocity Courant 0.5, dt 0.6-1e11 s varying "
                               f"per step, mode {mode}",
                   "markers": M, "grid": [g.Ny, g.Nx], "form": "fp64",
                   "launches": STEP_LAUNCHES},
        "roofline": dict(roofs[dom], kernel_class=dom,
                         measured_copy_GBps=_finite(copy_bandwidth())),
        "rooflines": roofs,
        "step_roofline": step_roof,
        "kernel_ms": kern,
    }
    if not a.no_mixed:
        log("[bench] mixed-precision form")
        ctxm, vm, msm, kernm, recapm = run_form(g, mat, mk, grids, hrs, hrf, ereb.MIXED,
                                                a.steps, a.warmup)
        errs = precision.field_errors(ctxm, ctx64)
        wk, wv = precision.worst(errs)
        nodal = set(ereb.SUM_FIELDS + precision.NODE_FIELDS)
        wn, wnv = precision.worst(errs, nodal)
        nnzm = (ctxm.csc_size(ereb.SYS_STOKES_DARCY)[1], ctxm.csc_size(ereb.SYS_THERMAL)[1])
        domm, roofsm, steproofm = rooflines(g, M, kernm, "mixed", a.workload, nnzm, msm)
        res["mixed"] = {
            "value": vm, "ms_per_step": msm, "speedup_vs_fp64": vm / value,
            "dtype": "f32 properties/weights/interpolation, f64 state/normalisation/coefficients",
            "roofline": dict(roofsm[domm], kernel_class=domm), "rooflines": roofsm,
            "step_roofline": steproofm, "launches": STEP_LAUNCHES,
            "kernel_ms": kernm,
            "pattern_difference": {
                "stokes_darcy": precision.pattern_difference(ctxm.csc(0), ctx64.csc(0)),
                "thermal": precision.pattern_difference(ctxm.csc(1), ctx64.csc(1))},
            "errors_vs_fp64": {
                "after_steps": a.warmup + 2 * a.steps,
                "metric": "per field [max|a-ref|/max|ref|, rms(a-ref)/rms(ref)]",
                "worst_field": wk, "worst_max_rel": wv,
                "worst_nodal_field": wn, "worst_nodal_max_rel": wnv,
                "fields": {k: [float(f"{e[0]:.3e}"), float(f"{e[1]:.3e}")]
                           for k, e in errs.items()},
            },
        }
        log(f"[bench] mixed {msm:.3f} ms/step, worst nodal field {wn} {wnv:.2e}")
        ctxm.close()
    log("[bench] solver hand-off")
    res["handoff"] = handoff_leg(ctx64)
    systems = {"thermal": ctx64.csc(ereb.SYS_THERMAL),
               "stokes_darcy": ctx64.csc(ereb.SYS_STOKES_DARCY)}
    if not a.no_next_rows:
        log("[bench] remaining transfer rows")
        res["next_rows"] = next_rows_leg(ctx64, g, mat, hrs, hrf)
        log("[bench] full reference timestep on the GPU")
        res["full_timestep"] = full_timestep_leg(ctx64, mat, hrs, hrf)
        log(f"[bench] full timestep {res['full_timestep']['ms_per_timestep']:.3f} ms")
    ctx64.close()
    if not a.no_adjoint:
        log("[bench] adjoint leg")
        res["adjoint"] = adjoint_leg()
        log(f"[bench] adjoint/forward {res['adjoint']['adjoint_forward_ratio']:.2f}")
    if not a.no_large:
        log("[bench] 20M-marker workloads (R = 40 and 65 km)")
        res["large"] = large_leg()
    if not a.no_cpu_baseline:
        log("[bench] CPU baseline (oracle at -O3, OpenMP in the gathers)")
        res["cpu_baseline"] = cpu_baseline(g, mat, mk, grids, hrs, hrf, a.cpu_steps)
    if not a.no_host_lu:
        log("[bench] host LU of the assembled systems (SuperLU)")
        res["host_lu"] = host_lu_leg(systems)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
