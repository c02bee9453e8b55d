"""Benchmark of the TVD advection forward step on one MI355X.

Workload (the plain FP64 form at real shape): a 512x512x256 local domain
(dx = dy = 25 m, z tanh-stretched to 5 km), fields u, v, w, th, q with the
seeded distributions of the synthetic set-up, advection-only forward-Euler
steps at CFL 0.3.  One step = the TVD tendencies of all five fields from the
pre-step state plus the update field <- field + dt*s and the periodic halo of
the new state (monc_tvd_forward_steps): four launches of the march
kernel, one each for u, v, w and the (th, q) pair, with tile shapes picked by
the library's autotuner before the untimed warm-up.  The same step in FP32 is
reported beside it.

Output: one JSON line.  value = interior cell-updates per second (all five
fields) of the FP64 form, wall clock over the timed steps; roofline = the
dominant kernel's algorithmic bytes per launch over its average duration
(HIP events recorded by the library around every launch on its stream), with
`traffic` the HBM bytes rocprofv3 measured for that kernel at that tile shape
(profiles/march_kernel_traffic_512x512x256.json; null for another shape);
cpu_baseline = the C restatement (oracle) timed on a bounded sample on this
host's CPU.  Beside it: the FP32 form, the FP32 adjoint step on the same
domain, the MONC leapfrog timestep (FP64), the sibling components
th_advection and pw_advection on the same FP64 state (kernel time and HBM
fraction), and the batched ensemble (48 FP32 members of seeded sizes
64..256 x 64..256 x 64..128, one persistent launch per step, graph-replayed).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "Gcell-updates/s & %HBM peak: FP32/FP64 fwd + adjoint at stated error vs FP64"


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


# Algorithmic HBM bytes per interior cell of one launch of each step kernel
# (element size es): u, v and w kernels read u, v, w and write their field;
# the scalar kernel reads u, v, w and its two scalars and writes both.
KIND_FIELDS = {"u": 4, "v": 4, "w": 4, "scalars": 7, "cell": 10}
KERNEL_NAMES = {0: "march_kernel", 1: "cell_kernel"}


def run_form(state, precision: int, steps: int, warmup: int, dt: float, kernel: int = 0, rows: int = 0,
             tune: bool = True):
    import torch

    from monc_amd import tvd

    t = tvd.TvdAdvection.initialise(state, precision=precision)
    try:
        stream = torch.cuda.current_stream()
        t.set_stream(stream.cuda_stream)
        t.set_tiling(kernel, rows, 0)
        t.halo_fill()
        tuned = t.autotune(dt, reps=2) if tune else None
        for _ in range(warmup):
            t.forward_steps(dt, 1)
        t.synchronize()
        t.set_profiling(True)
        t.kernel_times(reset=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        t.forward_steps(dt, steps)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        kt = t.kernel_times(reset=True)
        t.set_profiling(False)
        return wall, kt, tuned
    finally:
        t.close()


def roofline(kt: dict, cells: int, es: int):
    """Dominant kernel of the step: the kind with the largest total time."""
    kind = max((k for k in kt if k in KIND_FIELDS), key=lambda k: kt[k][0])
    ms_total, n = kt[kind]
    avg_ms = ms_total / n
    bytes_launch = KIND_FIELDS[kind] * es * cells
    achieved = bytes_launch / (avg_ms * 1e-3) / 1e9
    return kind, avg_ms, bytes_launch, achieved


def default_shape(precision: str, kind: str):
    """The library's default march tile (rows, x columns) for this workload
    (256 levels over 512 x 512 columns)."""
    return [2, 64] if precision == "fp64" and kind == "scalars" else [3, 32]


def measured_traffic(precision: str, kind: str, nz: int, shape):
    """HBM bytes per launch that rocprofv3 measured for this kernel at the
    tile shape it was profiled in, or None for another shape."""
    p = os.path.join(ROOT, "profiles", "march_kernel_traffic_512x512x256.json")
    if nz != 256 or not os.path.exists(p):
        return None
    with open(p) as fh:
        t = json.load(fh)
    prof = t.get("shapes", {}).get(precision, {}).get(kind)
    if shape is None or prof is None or list(shape) != list(prof):
        return None
    return t["bytes_per_launch"].get(precision, {}).get(kind)


def adjoint_form(state, dt: float, steps: int):
    """FP32 adjoint step (monc_tvd_adjoint_step) on the bench domain: ms per
    step against the FP32 forward step on the same context."""
    import numpy as np
    import torch

    from monc_amd import tvd

    t = tvd.TvdAdvection.initialise(state, precision=tvd.FP32)
    try:
        t.set_stream(torch.cuda.current_stream().cuda_stream)
        t.halo_fill()
        rng = np.random.default_rng(1)
        for f in range(len(state.fields)):
            lam = np.zeros(t.field_shape())
            lam[2:-2, 2:-2, :] = rng.standard_normal((state.nx, state.ny, state.nz))
            t.set_field(f, tvd.ADJOINT, lam)
        t.adjoint_step(dt)
        t.forward_steps(dt, 1)
        t.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            t.adjoint_step(dt)
        t.synchronize()
        t1 = time.perf_counter()
        t.forward_steps(dt, steps)
        t.synchronize()
        t2 = time.perf_counter()
        cells = state.nx * state.ny * state.nz
        return {"ms_per_step": round((t1 - t0) / steps * 1e3, 4),
                "value": round(cells * steps / (t1 - t0) / 1e9, 4), "unit": "Gcell-updates/s",
                "over_forward_step": round((t1 - t0) / (t2 - t1), 3),
                "bytes_per_cell_update": 80}
    finally:
        t.close()


def leapfrog_form(state, dt: float, steps: int):
    """FP64 MONC timestep (monc_tvd_leapfrog_steps: TVD + stepfields +
    swap_smooth, Robert filter tsmth = 0.06 as MONC's default)."""
    from monc_amd import tvd

    t = tvd.TvdAdvection.initialise(state, precision=tvd.FP64)
    try:
        t.halo_fill()
        t.leapfrog_steps(dt, 0.06, 2)
        t.synchronize()
        t0 = time.perf_counter()
        t.leapfrog_steps(dt, 0.06, steps, first_forward=False)
        t.synchronize()
        el = time.perf_counter() - t0
        cells = state.nx * state.ny * state.nz
        return {"ms_per_step": round(el / steps * 1e3, 4), "value": round(cells * steps / el / 1e9, 4),
                "unit": "Gcell-updates/s", "bytes_per_cell_update": 160, "dtype": "f64"}
    finally:
        t.close()


def thadv_form(state, steps: int):
    """th_advection's source terms (monc_thadv_tendencies: reference-state
    vertical advection and mean baroclinicity, every column) on the FP64
    state: one launch reads sth, u, v, w and writes sth, 40 B per cell
    (halo columns included), priced against HBM with HIP events."""
    import numpy as np

    from monc_amd import tvd

    t = tvd.TvdAdvection.initialise(state, precision=tvd.FP64)
    try:
        nz = state.nz
        t.set_field(tvd.TH, tvd.SOURCE, np.zeros(t.field_shape()))
        thref = 290.0 + 0.004 * state.grid.zn
        dth = np.zeros(nz)
        dth[:-1] = thref[1:] - thref[:-1]
        kw = dict(dthref=dth, thref=thref, baroclinicity=True, thref0=300.0, fcoriol=1.0e-4, g=9.81,
                  rate_x=2.0e-3, rate_y=-1.0e-3)
        t.thadv_tendencies(True, **kw)
        t.synchronize()
        t.set_profiling(True)
        t.kernel_times(reset=True)
        for _ in range(steps):
            t.thadv_tendencies(True, **kw)
        t.synchronize()
        # timed under the library's whole-field kind (halo fill, th_advection)
        ms, n = t.kernel_times(reset=True).get("halo", (0.0, 0))
        t.set_profiling(False)
        per = ms / max(n, 1)
        nbytes = 40.0 * (state.nx + 4) * (state.ny + 4) * (nz - 1)
        return {"kernel": "thadv_kernel<double>", "launches": n, "kernel_ms": round(per, 4),
                "bytes_per_launch": nbytes, "achieved_GBs": round(nbytes / per / 1e6, 2),
                "frac_hbm": round(nbytes / per / 1e6 / 8000.0, 4), "dtype": "f64"}
    finally:
        t.close()


def pw_form(state, steps: int):
    """pw_advection's centred tendencies (monc_pw_tendencies, flow, th and q)
    on the FP64 state: one launch reads u, v, w, th, q and sq and writes su,
    sv, sw, sth and sq, 88 B per interior cell, priced against HBM with HIP
    events."""
    import numpy as np

    from monc_amd import tvd

    t = tvd.TvdAdvection.initialise(state, precision=tvd.FP64)
    try:
        for n in range(len(state.fields)):
            t.set_field(n, tvd.SOURCE, np.zeros(t.field_shape()))
        tc = dict(tcx=0.25 * state.cx, tcy=0.25 * state.cy)
        t.pw_tendencies(**tc)
        t.synchronize()
        t.set_profiling(True)
        t.kernel_times(reset=True)
        for _ in range(steps):
            t.pw_tendencies(**tc)
        t.synchronize()
        # timed under the library's whole-field kind
        ms, n = t.kernel_times(reset=True).get("halo", (0.0, 0))
        t.set_profiling(False)
        per = ms / max(n, 1)
        nbytes = 88.0 * state.nx * state.ny * (state.nz - 1)
        return {"kernel": "pw_kernel<double>", "launches": n, "kernel_ms": round(per, 4),
                "bytes_per_launch": nbytes, "achieved_GBs": round(nbytes / per / 1e6, 2),
                "frac_hbm": round(nbytes / per / 1e6 / 8000.0, 4),
                "gcell_per_s": round(state.nx * state.ny * state.nz / per / 1e6, 3), "dtype": "f64"}
    finally:
        t.close()


def ensemble_form(steps: int):
    """48 FP32 members of seeded uneven sizes: forward steps by one persistent
    launch per step over all members' tiles, with and without graph replay,
    tile shape from the library's autotuner."""
    import torch

    from monc_amd import setup_state as ss
    from monc_amd import tvd
    from monc_amd.ensemble import TvdEnsemble

    sts = ss.ensemble_states(48, seed=0)
    dts = [ss.cfl_dt(st, 0.3) for st in sts]
    cells = sum(st.nx * st.ny * st.nz for st in sts)
    e = TvdEnsemble.from_states(sts, dts, precision=tvd.FP32)
    del sts
    try:
        e.set_stream(torch.cuda.current_stream().cuda_stream)
        e.halo_fill()
        shape = e.autotune(reps=2)
        out = {"members": 48, "cells": cells, "shape": {"waves": shape[0], "strip": shape[1]}}
        for graph in (True, False):
            e.set_tiling(shape[0], shape[1], graph)
            e.forward_steps(2)
            e.synchronize()
            t0 = time.perf_counter()
            e.forward_steps(steps)
            e.synchronize()
            el = time.perf_counter() - t0
            out["graph" if graph else "direct"] = {"ms_per_step": round(el / steps * 1e3, 4),
                                                   "value": round(cells * steps / el / 1e9, 4)}
        out["unit"] = "Gcell-updates/s"
        out["bytes_per_cell_update"] = 40
        return out
    finally:
        e.close()


def cpu_baseline(nz: int, seconds_hint: float = 12.0):
    """The oracle (single-threaded C restatement) on a bounded sample of the
    same workload: one forward step of a 288x288 column block at full depth."""
    import numpy as np

    import oracle
    from monc_amd import setup_state as ss

    n = 288
    st = ss.make_state(nz, n, n, 25.0, 5000.0, seed=0, n_q=1, with_zfields=False)
    dt = ss.cfl_dt(st, 0.3)
    t0 = time.perf_counter()
    srcs, _ = oracle.tendencies(nz, n, n, st.cx, st.cy, st.grid.metric_vectors(), dt, 1, 1, 1, st.fields)
    _ = [oracle.forward_step(f, s, dt) for f, s in zip(st.fields, srcs)]
    el = time.perf_counter() - t0
    cells = n * n * nz
    return {"value": cells / el / 1e9, "unit": "Gcell-updates/s", "cores": 1, "kind": "port",
            "sample": f"one forward step (5 fields) of a {n}x{n}x{nz} block of the workload, "
                      f"{el:.1f} s on 1 core"}


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nx", type=int, default=512)
    ap.add_argument("--ny", type=int, default=512)
    ap.add_argument("--nz", type=int, default=256)
    ap.add_argument("--kernel", type=int, default=0, help="0 march (default), 1 cell")
    ap.add_argument("--rows", type=int, default=0, help="march kernel cell rows per block (0: default)")
    ap.add_argument("--no-tune", action="store_true", help="skip the tile-shape autotuner (untimed)")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the adjoint, leapfrog and ensemble measurements")
    ap.add_argument("--only", choices=["adjoint", "leapfrog", "ensemble", "thadv", "pw"], default=None,
                    help="run only this side measurement (for profiling) and print its JSON object")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="after the timed work, run three 512 MiB device-to-device copies: the known-size "
                         "reference that scales rocprofv3 FETCH_SIZE/WRITE_SIZE to bytes")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        sys.exit(2)

    import torch

    from monc_amd import setup_state as ss
    from monc_amd import tvd

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    tvd.load_library()
    nx, ny, nz = args.nx, args.ny, args.nz
    log(f"bench: building {nx}x{ny}x{nz} state")
    st = ss.make_state(nz, ny, nx, 25.0, 5000.0, seed=0, n_q=1, with_zfields=False)
    dt = ss.cfl_dt(st, 0.3)
    cells = nx * ny * nz

    if args.only:
        log(f"bench: {args.only} only")
        side = {"adjoint": lambda: adjoint_form(st, dt, max(2, args.steps // 2)),
                "leapfrog": lambda: leapfrog_form(st, dt, max(2, args.steps // 2)),
                "ensemble": lambda: ensemble_form(max(4, args.steps // 2)),
                "thadv": lambda: thadv_form(st, max(4, args.steps)),
                "pw": lambda: pw_form(st, max(4, args.steps))}[args.only]()
        print(json.dumps({args.only: side}), flush=True)
        return

    log("bench: FP64 form")
    tune = not args.no_tune and args.kernel == 0 and args.rows == 0
    wall64, kt64, tuned64 = run_form(st, tvd.FP64, args.steps, args.warmup, dt, args.kernel, args.rows, tune)
    ms_step = wall64 / args.steps * 1e3
    value = cells * args.steps / wall64 / 1e9
    kind, kavg, kbytes, achieved = roofline(kt64, cells, 8)
    traffic = measured_traffic("fp64", kind, nz, (tuned64 or {}).get(kind) if tuned64 else
                               (default_shape("fp64", kind) if args.kernel == 0 and args.rows == 0 else None))
    step_kernel_ms = sum(v[0] for k, v in kt64.items() if k in KIND_FIELDS) / args.steps
    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded u=5+N(0,1), v=N(0,1), w=N(0,0.5), th=300+0.003z+U(-0.1,0.1), "
                "q=0.01exp(-z/2000)+U(0,1e-4))",
        "config": {"workload": "plain FP64 TVD advection forward-Euler step, 512x512x256, dx=dy=25 m, "
                               "z tanh-stretched to 5 km, fields u,v,w,th,q, CFL 0.3",
                   "nx": nx, "ny": ny, "nz": nz, "fields": 5, "dt_s": dt},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": f"{KERNEL_NAMES[args.kernel]} ({kind})", "kernel_ms": round(kavg, 4),
                     "algorithmic_bytes_per_launch": kbytes,
                     "step": {"kernel_ms": round(step_kernel_ms, 4), "bytes_per_cell_update": 80,
                              "achieved_GBs": round(80 * cells / (step_kernel_ms * 1e-3) / 1e9, 2)},
                     "kernels_ms_per_launch": {k: round(v[0] / v[1], 4) for k, v in kt64.items()}},
        "tile_shapes": {"fp64": tuned64},
    }
    if not args.no_fp32:
        log("bench: FP32 form")
        wall32, kt32, tuned32 = run_form(st, tvd.FP32, args.steps, args.warmup, dt, args.kernel, args.rows, tune)
        v32 = cells * args.steps / wall32 / 1e9
        kind32, kavg32, kbytes32, a32 = roofline(kt32, cells, 4)
        step32 = sum(v[0] for k, v in kt32.items() if k in KIND_FIELDS) / args.steps
        out["fp32"] = {"value": round(v32, 4), "unit": "Gcell-updates/s",
                       "ms_per_step": round(wall32 / args.steps * 1e3, 4),
                       "dominant_kernel": kind32, "kernel_ms": round(kavg32, 4), "achieved_GBs": round(a32, 2),
                       "frac": round(a32 / HBM_PEAK_GBS, 4),
                       "step_achieved_GBs": round(40 * cells / (step32 * 1e-3) / 1e9, 2),
                       "speedup_vs_fp64": round(v32 / value, 3)}
        out["tile_shapes"]["fp32"] = tuned32
    if not args.no_extra:
        log("bench: FP32 adjoint step")
        out["adjoint_fp32"] = adjoint_form(st, dt, max(2, args.steps // 2))
        log("bench: FP64 leapfrog timestep")
        out["leapfrog_fp64"] = leapfrog_form(st, dt, max(2, args.steps // 2))
        log("bench: th_advection")
        out["thadv_fp64"] = thadv_form(st, max(4, args.steps))
        log("bench: pw_advection")
        out["pw_fp64"] = pw_form(st, max(4, args.steps))
        log("bench: batched ensemble (48 FP32 members)")
        out["ensemble_fp32"] = ensemble_form(max(4, args.steps // 2))
    if args.calibrate_copy:
        a = torch.ones(64 * 1024 * 1024, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    if not args.no_cpu:
        log("bench: CPU baseline (oracle) on a bounded sample")
        out["cpu_baseline"] = cpu_baseline(nz)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
