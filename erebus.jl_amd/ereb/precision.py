"""Per-field error of one resident state against another of the same
workload, e.g. the mixed-precision form against the Float64 form after the
same steps on the same inputs.

For every field a: max|a - ref| / max|ref| (normwise relative maximum) and
rms(a - ref) / rms(ref) (relative RMS).  Fields: the 30 SUM grids of the
scatter, the 27 node fields of the normalisation, the marker state and the
7 marker property vectors, and both assembled systems (matrix values on the
union of the two sparsity patterns, and right-hand side).
"""
from __future__ import annotations

import numpy as np

from .lib import SUM_FIELDS, Context

NODE_FIELDS = [
    "ETA0", "ETA", "GGG", "SXY0", "COH", "TEN", "FRI",          # basic E:1970
    "RHOX", "RHOFX", "KX", "PHIX", "RX",                        # Vx E:2033
    "RHOY", "RHOFY", "KY", "PHIY", "RY",                        # Vy E:2086
    "RHO", "RHOCP", "ALPHA", "ALPHAF", "HR", "GGGP", "SXX0", "tk1", "PHI", "BETAPHI",  # P
]
MARKER_STATE = ["xm", "ym", "tkm", "sxxm", "sxym"]
MARKER_PROPS = ["rhototalm", "rhocptotalm", "etatotalm", "hrtotalm", "ktotalm",
                "tkm_rhocptotalm", "etafluidcur_inv_kphim"]


def rel_errors(a: np.ndarray, ref: np.ndarray) -> tuple[float, float]:
    """(max|a-ref|/max|ref|, rms(a-ref)/rms(ref)); 0 when both are zero."""
    a = np.asarray(a, dtype=np.float64).ravel()
    ref = np.asarray(ref, dtype=np.float64).ravel()
    d = np.abs(a - ref)
    den_max = np.abs(ref).max() if ref.size else 0.0
    den_rms = np.sqrt(np.mean(ref * ref)) if ref.size else 0.0
    emax = float(d.max() / den_max) if den_max > 0 else float(d.max() if d.size else 0.0)
    erms = float(np.sqrt(np.mean(d * d)) / den_rms) if den_rms > 0 else emax
    return emax, erms


def field_errors(ctx: Context, ref: Context) -> dict:
    """Per-field errors of ctx against ref (same geometry and marker count)."""
    out: dict[str, tuple[float, float]] = {}
    for n in SUM_FIELDS + NODE_FIELDS:
        out[n] = rel_errors(ctx.get_grid(n), ref.get_grid(n))
    for n in MARKER_STATE + MARKER_PROPS:
        out[n] = rel_errors(ctx.get_marker(n), ref.get_marker(n))
    for which, name in ((0, "stokes_darcy"), (1, "thermal")):
        a, b = _csc_dense_pairs(ctx.csc(which), ref.csc(which))
        out[f"{name}.nzval"] = rel_errors(a, b)
        out[f"{name}.rhs"] = rel_errors(ctx.csc(which)[3], ref.csc(which)[3])
    return out


def _csc_dense_pairs(A, B):
    """Values of two CSC matrices on the union of their patterns (an entry
    absent from one side counts as zero there: the zero-skip of the
    assembly drops coefficients that round to exactly zero)."""
    def keyed(m):
        cp, rv, nz, _ = m
        cols = np.repeat(np.arange(cp.size - 1, dtype=np.int64), np.diff(cp))
        return (cols * (1 << 32) + (rv - 1)), nz
    ka, va = keyed(A)
    kb, vb = keyed(B)
    keys = np.union1d(ka, kb)
    a = np.zeros(keys.size)
    b = np.zeros(keys.size)
    a[np.searchsorted(keys, ka)] = va
    b[np.searchsorted(keys, kb)] = vb
    return a, b


def pattern_difference(A, B) -> int:
    """Number of entries present in one CSC pattern and not the other."""
    def keys(m):
        cp, rv = m[0], m[1]
        cols = np.repeat(np.arange(cp.size - 1, dtype=np.int64), np.diff(cp))
        return cols * (1 << 32) + (rv - 1)
    return int(np.setxor1d(keys(A), keys(B)).size)


def worst(errs: dict, prefix_filter=None) -> tuple[str, float]:
    """Field with the largest normwise maximum error."""
    items = [(k, v[0]) for k, v in errs.items()
             if prefix_filter is None or k in prefix_filter]
    k, v = max(items, key=lambda kv: kv[1])
    return k, v
