"""Shared helpers of the test suite: run the HIP path and compare with the oracle."""
from __future__ import annotations

import numpy as np

from monc_amd import tvd


def interior(a: np.ndarray) -> np.ndarray:
    return a[2:-2, 2:-2, :]


def rel_err(a: np.ndarray, b: np.ndarray) -> float:
    """max|a-b| / max|b| over the interior columns (all levels)."""
    ai, bi = interior(a), interior(b)
    den = float(np.max(np.abs(bi)))
    num = float(np.max(np.abs(ai - bi)))
    return num / den if den > 0 else num


def gpu_tendencies(state, dtm, fs, ms, ss, sources=None, precision=tvd.FP64, advect=7, tiling=None, active=None):
    """Library timestep over a State; returns the s-fields (FP64 numpy).
    tiling = (kernel, rows, strip) for monc_tvd_set_tiling, or None; active:
    the th and q fields' active flags by field id (None: all active)."""
    t = tvd.TvdAdvection.initialise(state, precision=precision, advect=advect)
    try:
        if tiling is not None:
            t.set_tiling(*tiling)
        if active is not None:
            for f in range(3, 4 + state.n_q):
                t.set_field_active(f, bool(active[f]))
        n = len(state.fields)
        for f in range(n):
            t.set_field(f, tvd.SOURCE, np.zeros(t.field_shape()) if sources is None else sources[f])
        t.halo_fill()
        t.tendencies(dtm, fs, ms, ss)
        return [t.get_field(f, tvd.SOURCE) for f in range(n)]
    finally:
        t.close()


def gpu_forward(state, dtm, nsteps, precision=tvd.FP64, advect=7, tiling=None):
    t = tvd.TvdAdvection.initialise(state, precision=precision, advect=advect)
    try:
        if tiling is not None:
            t.set_tiling(*tiling)
        t.halo_fill()
        t.forward_steps(dtm, nsteps)
        return [t.get_field(f, tvd.PROG) for f in range(len(state.fields))]
    finally:
        t.close()
