"""Benchmark of the TVD advection forward step on one MI355X.

Workload (the plain FP64 form at real shape): a 512x512x256 local domain
(dx = dy = 25 m, z tanh-stretched to 5 km), fields u, v, w, th, q with the
seeded distributions of the synthetic set-up, advection-only forward-Euler
steps at CFL 0.3.  One step = the TVD tendencies of all five fields from the
pre-step state plus the update field <- field + dt*s and the periodic halo of
the new state (monc_tvd_forward_steps): one launch per field kind (u, v, w
and the (th, q) pair), each in the library's default kernel and tile shape
(FP64 u, v, w: the LDS-staged kernel; the rest: the march kernel; --tune runs
the library's autotuner first).  The same step in FP32 is reported beside it.

Output: one JSON line.  value = interior cell-updates per second (all five
fields) of the FP64 form, wall clock over the timed steps; roofline = the
dominant kernel's algorithmic bytes per launch over its average duration
(HIP events recorded by the library around every launch on its stream),
with the HBM traffic, VALU instruction count and vector-memory busy that
rocprofv3 measured for that kernel at that shape
(profiles/step_counters_512x512x256.json; marked stale for another kernel or
shape); cpu_baseline = the C restatement (oracle) of the path on the host's
cores, one process per core on x slabs of the same workload, run before the
GPU is touched.  Beside it: the FP32 form, the FP32 adjoint step on the same
domain, the MONC leapfrog timestep (FP64), the drop-in tendency call itself
(64^3 and the bench domain, forward and centred; 120 / 160 B per
cell-update), the sibling components
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
# VALU issue peak: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz (FP64 and unpacked
# FP32 alike), lane-instructions per second
VALU_PEAK = 256 * 4 * 16 * 2.4e9
PROFILE = os.path.join(ROOT, "profiles", "step_counters_512x512x256.json")
METRIC = "Gcell-updates/s & %HBM peak: FP32/FP64 fwd + adjoint at stated error vs FP64"


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


# Algorithmic HBM bytes per interior cell of one launch of each step kernel
# (element size es): u, v and w kernels read u, v, w and write their field;
# the scalar kernel reads u, v, w and its two scalars and writes both.
KIND_FIELDS = {"u": 4, "v": 4, "w": 4, "scalars": 7, "cell": 10}
KERNEL_NAMES = {0: "march_kernel", 1: "cell_kernel", 2: "stage_kernel"}


def run_form(state, precision: int, steps: int, warmup: int, dt: float, kernel: int = 0, rows: int = 0,
             tune: bool = True):
    import torch

    from monc_amd import tvd

    t = tvd.TvdAdvection.initialise(state, precision=precision)
    try:
        stream = torch.cuda.current_stream()
        t.set_stream(stream.cuda_stream)
        if kernel or rows:
            t.set_tiling(kernel, rows, 0)
        t.halo_fill()
        if tune:
            t.autotune(dt, reps=2)
        tuned = {k: list(v) for k, v in t.tiling().items()}
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


def profiled(precision: str, kind: str, nz: int, tiling):
    """The rocprofv3 counters of this kernel kind at the kernel and tile shape
    it ran in (kernel, rows, strip), or None when the profile holds another
    kernel or shape (stale) or another column height."""
    if nz != 256 or not os.path.exists(PROFILE) or tiling is None:
        return None
    with open(PROFILE) as fh:
        prof = json.load(fh)
    entry = prof.get(precision, {}).get(kind)
    if entry is None or list(entry["tiling"]) != list(tiling):
        return None
    return entry


def cells_of(nz: int) -> int:
    """Interior cells of the profiled workload (512 x 512 columns)."""
    return 512 * 512 * nz


def roofline_entry(precision: str, kind: str, kavg_ms: float, alg_bytes: float, nz: int, tiling):
    """HBM fraction at algorithmic bytes, and, from the profile at the same
    kernel and shape, the measured HBM traffic, the VALU fraction
    (SQ_INSTS_VALU x 64 / time / VALU_PEAK) and the vector-memory busy (TA);
    `limiter` is the largest of the three fractions."""
    t = kavg_ms * 1e-3
    achieved = alg_bytes / t / 1e9
    out = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel_ms": round(kavg_ms, 4),
           "algorithmic_bytes_per_launch": alg_bytes, "tiling": tiling}
    prof = profiled(precision, kind, nz, tiling)
    if prof is None:
        out["counters"] = "stale: no profile of this kernel at this shape"
        return out
    traffic = prof["fetch_bytes"] + prof["write_bytes"]
    valu = prof["valu_insts"] * 64 / t / VALU_PEAK
    fracs = {"hbm": traffic / t / (HBM_PEAK_GBS * 1e9), "valu": valu, "vmem": prof["ta_busy"]}
    out.update({"traffic": traffic, "traffic_over_algorithmic": round(traffic / alg_bytes, 3),
                "traffic_frac": round(fracs["hbm"], 4),
                "valu": {"insts_per_launch": prof["valu_insts"], "achieved": round(prof["valu_insts"] * 64 / t, 0),
                         "peak": VALU_PEAK, "unit": "lane-instructions/s", "frac": round(valu, 4)},
                "vmem_busy": prof["ta_busy"], "limiter": max(fracs, key=fracs.get),
                # vector-memory read instructions (each 64 lanes) per 64-cell column of the launch
                "vmem_wave_loads_per_column": round(prof["vmem_rd"] / (cells_of(nz) / 64.0), 2),
                "profiled_kernel_ms": prof["ms"],
                "counters": f"profile {os.path.relpath(PROFILE, ROOT)} ({prof['kernel']})"})
    return out


_KIND_OF = {0: "u", 1: "v", 2: "w", 3: "scalars"}


def _step_kernel(name: str):
    """(precision, kind) of a forward-step kernel named in a rocprofv3 record,
    or None: march_kernel<T, KIND, NS, MODE=1, ...> and stage_kernel<T, KIND,
    ...>."""
    import re

    m = re.search(r"(march|stage)_kernel<(float|double), (\d+), (\d+)(?:, (\d+))?", name)
    if m is None or (m.group(1) == "march" and m.group(5) != "1"):
        return None
    return ("fp64" if m.group(2) == "double" else "fp32"), _KIND_OF[int(m.group(3))]


def collate(d: str) -> None:
    """Reduce the rocprofv3 runs of the step under directory d to the counters
    bench.py prices the dominant kernel with (written to the profile file).
    Expected: d/kt/kt_kernel_stats.csv (--kernel-trace --stats), d/bench.json
    (the JSON line of that run, for the tile shapes), and the counter passes
    d/f (FETCH_SIZE), d/w (WRITE_SIZE), d/sq (SQ_INSTS_VALU, SQ_INSTS_VMEM_RD,
    GRBM_GUI_ACTIVE), d/ta (TA_TA_BUSY_sum, TD_TD_BUSY_sum, GRBM_GUI_ACTIVE),
    each d/<p>/<p>_counter_collection.csv.  The f and w runs end with three
    512 MiB device-to-device copies (--calibrate-copy): the last three copy
    dispatches give the factors from FETCH_SIZE and WRITE_SIZE (KiB) to bytes.
    Busy fractions: VALU = SQ_INSTS_VALU x 4 cycles / (GRBM_GUI_ACTIVE / 8 XCDs
    x 1024 SIMDs); TA and TD = *_BUSY_sum / (GRBM_GUI_ACTIVE / 8 x 256 CUs)."""
    import csv
    from collections import defaultdict

    def counters(p):
        per = defaultdict(lambda: defaultdict(list))   # kernel -> counter -> values
        copies = defaultdict(dict)                     # dispatch -> counter -> value
        with open(os.path.join(d, p, f"{p}_counter_collection.csv")) as fh:
            for r in csv.DictReader(fh):
                v = float(r["Counter_Value"])
                if "copyBuffer" in r["Kernel_Name"]:
                    copies[int(r["Dispatch_Id"])][r["Counter_Name"]] = v
                key = _step_kernel(r["Kernel_Name"])
                if key is not None:
                    per[key][r["Counter_Name"]].append(v)
        return ({k: {c: sum(v) / len(v) for c, v in cs.items()} for k, cs in per.items()},
                [copies[i] for i in sorted(copies)[-3:]])

    def factor(copy_list, counter):
        vals = sorted(c[counter] for c in copy_list)
        return 512 * 1024 / vals[1]

    stats, names = {}, {}
    with open(os.path.join(d, "kt", "kt_kernel_stats.csv")) as fh:
        for r in csv.DictReader(fh):
            key = _step_kernel(r["Name"])
            if key is not None:
                stats[key] = float(r["AverageNs"]) * 1e-6
                names[key] = r["Name"].split("(")[0].replace("void ", "")
    with open(os.path.join(d, "bench.json")) as fh:
        shapes = json.loads(fh.read().strip().splitlines()[-1])["tile_shapes"]
    f, fcp = counters("f")
    w, wcp = counters("w")
    sq, _ = counters("sq")
    ta, _ = counters("ta")
    ff, wf = factor(fcp, "FETCH_SIZE"), factor(wcp, "WRITE_SIZE")
    prof = {"workload": "bench.py default FP64 and FP32 step, 512x512x256, library default shapes",
            "calibration": {"fetch_bytes_per_KiB_count": round(ff * 1024, 1),
                            "write_bytes_per_KiB_count": round(wf * 1024, 1)}}
    for key in sorted(stats):
        prec, kind = key
        cyc_xcd = sq[key]["GRBM_GUI_ACTIVE"] / 8
        prof.setdefault(prec, {})[kind] = {
            "kernel": names[key], "tiling": shapes[prec][kind], "ms": round(stats[key], 4),
            "fetch_bytes": round(f[key]["FETCH_SIZE"] * 1024 * ff),
            "write_bytes": round(w[key]["WRITE_SIZE"] * 1024 * wf),
            "valu_insts": round(sq[key]["SQ_INSTS_VALU"]), "vmem_rd": round(sq[key]["SQ_INSTS_VMEM_RD"]),
            "valu_busy": round(sq[key]["SQ_INSTS_VALU"] * 4 / (cyc_xcd * 1024), 3),
            "ta_busy": round(ta[key]["TA_TA_BUSY_sum"] / (ta[key]["GRBM_GUI_ACTIVE"] / 8 * 256), 3),
            "td_busy": round(ta[key]["TD_TD_BUSY_sum"] / (ta[key]["GRBM_GUI_ACTIVE"] / 8 * 256), 3)}
    with open(PROFILE, "w") as fh:
        json.dump(prof, fh, indent=1)
        fh.write("\n")
    cells = 512 * 512 * 256
    print("| precision | kind | kernel | tiling | avg ms | fetch GB | write GB | traffic / alg. | "
          "alg. GB/s | VALU instr. (M) | VALU busy | VMEM reads (M) | TA busy | TD busy |")
    print("|---|---|---|---|---|---|---|---|---|---|---|---|---|---|")
    for prec in ("fp64", "fp32"):
        es = 8 if prec == "fp64" else 4
        for kind, e in prof.get(prec, {}).items():
            alg = KIND_FIELDS[kind] * es * cells
            print(f"| {prec} | {kind} | `{e['kernel']}` | {e['tiling']} | {e['ms']:.3f} | "
                  f"{e['fetch_bytes'] / 1e9:.2f} | {e['write_bytes'] / 1e9:.2f} | "
                  f"{(e['fetch_bytes'] + e['write_bytes']) / alg:.2f} | {alg / e['ms'] / 1e6:.0f} | "
                  f"{e['valu_insts'] / 1e6:.0f} | {e['valu_busy']:.2f} | {e['vmem_rd'] / 1e6:.1f} | "
                  f"{e['ta_busy']:.2f} | {e['td_busy']:.2f} |")
    print(f"calibration: FETCH_SIZE x {ff:.3f}, WRITE_SIZE x {wf:.3f}")


def adjoint_form(state, dt: float, steps: int):
    """FP32 adjoint step (monc_tvd_adjoint_step) on the bench domain: ms per
    step against the FP32 forward step (monc_tvd_forward_step_into), with the limiter
    arms recorded by the forward step (monc_tvd_forward_step_record, the
    rollout's recompute) and with them re-derived from the state; and the
    recording step's cost against the plain step."""
    import numpy as np
    import torch

    from monc_amd import setup_state as ss
    from monc_amd import tvd

    t = tvd.TvdAdvection.initialise(state, precision=tvd.FP32)
    try:
        t.set_stream(torch.cuda.current_stream().cuda_stream)
        dev = torch.device("cuda", 0)
        x = [torch.from_numpy(ss.halo_fill(f.copy())).to(dev, torch.float32).contiguous() for f in state.fields]
        for f, a in enumerate(x):
            t.bind_field(f, tvd.PROG, a)
        y = [torch.empty_like(a) for a in x]
        rec = [torch.empty(t.field_shape(), dtype=torch.uint8, device=dev) for _ in x]
        rng = np.random.default_rng(1)
        lam0 = []
        for f in range(len(state.fields)):
            lam = np.zeros(t.field_shape())
            lam[2:-2, 2:-2, :] = rng.standard_normal((state.nx, state.ny, state.nz))
            lam0.append(lam)
            t.set_field(f, tvd.ADJOINT, lam)

        def timed(fn, n):
            t.synchronize()
            t0 = time.perf_counter()
            for _ in range(n):
                fn()
            t.synchronize()
            return (time.perf_counter() - t0) / n

        t.adjoint_step(dt)  # warm-up, arms re-derived
        derived = timed(lambda: t.adjoint_step(dt), steps)
        t.forward_step_into(dt, x, y, rec)
        for f in range(len(x)):
            t.bind_record(f, rec[f])
        t.adjoint_step(dt)
        recorded = timed(lambda: t.adjoint_step(dt), steps)
        for f in range(len(x)):
            t.bind_record(f, None)
        plain_into = timed(lambda: t.forward_step_into(dt, x, y), steps)
        rec_into = timed(lambda: t.forward_step_into(dt, x, y, rec), steps)
        fwd = plain_into  # the forward step between buffers, without the copy-back of bound fields
        cells = state.nx * state.ny * state.nz
        return {"ms_per_step": round(recorded * 1e3, 4), "value": round(cells / recorded / 1e9, 4),
                "unit": "Gcell-updates/s", "over_forward_step": round(recorded / fwd, 3),
                "arms": "recorded by the forward step",
                "rederived_ms_per_step": round(derived * 1e3, 4),
                "rederived_over_forward_step": round(derived / fwd, 3),
                "record_step_ms": round(rec_into * 1e3, 4), "plain_step_into_ms": round(plain_into * 1e3, 4),
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


def tendency_form(state, dtm: float, stepping: int, calls: int):
    """The drop-in call itself (monc_tvd_tendencies, the timestep callback over
    every column) on resident FP64 fields: forward stepping reads the five
    prognostics and read-modify-writes the five source fields (120 B per
    cell-update), centred stepping also reads the five z-fields (160 B).
    Kernel time from the library's HIP events, summed over the launches of
    one call."""
    import numpy as np

    from monc_amd import tvd

    t = tvd.TvdAdvection.initialise(state, precision=tvd.FP64)
    try:
        if stepping != tvd.FORWARD_STEPPING and not len(state.zfields):
            # a state built without z-fields: the fields stand in for them
            for n, f in enumerate(state.fields):
                t.set_field(n, tvd.ZFIELD, f)
        t.zero_sources(range(len(state.fields)))
        t.halo_fill()
        t.tendencies(dtm, stepping, stepping, stepping)
        t.synchronize()
        t.set_profiling(True)
        t.kernel_times(reset=True)
        t0 = time.perf_counter()
        for _ in range(calls):
            t.tendencies(dtm, stepping, stepping, stepping)
        t.synchronize()
        wall = (time.perf_counter() - t0) / calls
        kt = t.kernel_times(reset=True)
        t.set_profiling(False)
        ms = sum(v[0] for v in kt.values()) / calls
        cells = state.nx * state.ny * state.nz
        nbytes = (120.0 if stepping == tvd.FORWARD_STEPPING else 160.0) * cells
        return {"cells": cells, "stepping": "forward" if stepping == tvd.FORWARD_STEPPING else "centred",
                "kernel_ms": round(ms, 4), "wall_ms": round(wall * 1e3, 4),
                "value": round(cells / (ms * 1e-3) / 1e9, 3), "unit": "Gcell-updates/s",
                "bytes_per_cell_update": nbytes / cells, "achieved_GBs": round(nbytes / ms / 1e6, 1),
                "frac_hbm": round(nbytes / ms / 1e6 / HBM_PEAK_GBS, 4), "dtype": "f64",
                "kernels_ms_per_call": {k: round(v[0] / calls, 4) for k, v in kt.items()}}
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


_CPU_STATE = None


def _cpu_slab(args):
    """One process's share of the CPU baseline: the TVD tendencies and the
    forward step of an x slab (its x halo taken from the neighbouring
    columns, y periodic), timed around the oracle calls only."""
    import numpy as np

    import oracle
    i0, i1 = args
    st, dt = _CPU_STATE
    fields = [np.ascontiguousarray(f[i0 - 2:i1 + 2]) for f in st.fields]
    t0 = time.perf_counter()
    srcs, _ = oracle.tendencies(st.nz, st.ny, i1 - i0, st.cx, st.cy, st.grid.metric_vectors(), dt, 1, 1, 1, fields)
    _ = [oracle.forward_step(f, s, dt) for f, s in zip(fields, srcs)]
    return time.perf_counter() - t0


def cpu_baseline(nz: int, ny: int, nx: int):
    """The oracle (the C restatement of the path, -O2, no OpenMP of its own)
    on the host's cores: one forward step of the whole workload, split into
    x slabs, one process per core (as MONC's MPI ranks split the domain; the
    halo exchange is the slabs' copy of neighbouring columns, outside the
    timed region).  Runs before the GPU is initialised (the workers are
    forked)."""
    import multiprocessing as mp

    from monc_amd import setup_state as ss
    global _CPU_STATE
    cores = max(1, min(os.cpu_count() or 1, 16))
    st = ss.make_state(nz, ny, nx, 25.0, 5000.0, seed=0, n_q=1, with_zfields=False)
    dt = ss.cfl_dt(st, 0.3)
    _CPU_STATE = (st, dt)
    edges = [2 + (nx * p) // cores for p in range(cores + 1)]
    slabs = [(edges[p], edges[p + 1]) for p in range(cores)]
    t0 = time.perf_counter()
    with mp.get_context("fork").Pool(cores) as pool:
        times = pool.map(_cpu_slab, slabs)
    wall = time.perf_counter() - t0
    _CPU_STATE = None
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    cells = nx * ny * nz
    return {"value": round(cells / max(times) / 1e9, 5), "unit": "Gcell-updates/s", "cores": cores,
            "kind": "port", "nproc": os.cpu_count(), "cpu_model": model,
            "sample": f"one forward step (5 fields) of the whole {nx}x{ny}x{nz} workload by the C restatement "
                      f"(not the reference Fortran) in {cores} processes on x slabs; slowest slab "
                      f"{max(times):.2f} s, pool wall {wall:.2f} s"}


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nx", type=int, default=512)
    ap.add_argument("--ny", type=int, default=512)
    ap.add_argument("--nz", type=int, default=256)
    ap.add_argument("--kernel", type=int, default=0,
                    help="0 the library's default per kind, 1 cell, 2 stage for every kind")
    ap.add_argument("--rows", type=int, default=0, help="cell rows per block (0: default)")
    ap.add_argument("--tune", action="store_true", help="run the library's autotuner first (untimed)")
    ap.add_argument("--no-tune", action="store_true", help="(the default) keep the library's default shapes")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the adjoint, leapfrog and ensemble measurements")
    ap.add_argument("--only", choices=["adjoint", "leapfrog", "ensemble", "thadv", "pw", "tendency"], default=None,
                    help="run only this side measurement (for profiling) and print its JSON object")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="after the timed work, run three 512 MiB device-to-device copies: the known-size "
                         "reference that scales rocprofv3 FETCH_SIZE/WRITE_SIZE to bytes")
    ap.add_argument("--collate", metavar="DIR", default=None,
                    help="no GPU work: reduce the rocprofv3 runs of the step under DIR to the profile file")
    args = ap.parse_args()
    if args.collate:
        collate(args.collate)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        sys.exit(2)

    cpu = None
    if not args.no_cpu and not args.only:
        # before anything initialises the GPU: the baseline forks its workers
        log("bench: CPU baseline (the C restatement on the host's cores)")
        cpu = cpu_baseline(args.nz, args.ny, args.nx)

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
                "tendency": lambda: {s: tendency_form(st, dt, m, max(4, args.steps // 2))
                                     for s, m in (("forward", tvd.FORWARD_STEPPING), ("centred", tvd.CENTRED_STEPPING))},
                "pw": lambda: pw_form(st, max(4, args.steps))}[args.only]()
        print(json.dumps({args.only: side}), flush=True)
        return

    log("bench: FP64 form")
    tune = args.tune and args.kernel == 0 and args.rows == 0
    wall64, kt64, tuned64 = run_form(st, tvd.FP64, args.steps, args.warmup, dt, args.kernel, args.rows, tune)
    ms_step = wall64 / args.steps * 1e3
    value = cells * args.steps / wall64 / 1e9
    kind, kavg, kbytes, achieved = roofline(kt64, cells, 8)
    step_kernel_ms = sum(v[0] for k, v in kt64.items() if k in KIND_FIELDS) / args.steps
    roof = roofline_entry("fp64", kind, kavg, kbytes, nz, tuned64.get(kind))
    roof["kernel"] = f"{KERNEL_NAMES[tuned64[kind][0]]} ({kind})"
    step_traffic = [profiled("fp64", k, nz, tuned64.get(k)) for k in ("u", "v", "w", "scalars")]
    roof["step"] = {"kernel_ms": round(step_kernel_ms, 4), "bytes_per_cell_update": 80,
                    "achieved_GBs": round(80 * cells / (step_kernel_ms * 1e-3) / 1e9, 2),
                    "frac": round(80 * cells / (step_kernel_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                    "traffic_over_algorithmic": (round(sum(p["fetch_bytes"] + p["write_bytes"] for p in step_traffic)
                                                       / (80.0 * cells), 3)
                                                 if all(step_traffic) else None)}
    roof["kernels_ms_per_launch"] = {k: round(v[0] / v[1], 4) for k, v in kt64.items()}
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
        "roofline": roof,
        "tile_shapes": {"fp64": tuned64},
    }
    if not args.no_fp32:
        log("bench: FP32 form")
        wall32, kt32, tuned32 = run_form(st, tvd.FP32, args.steps, args.warmup, dt, args.kernel, args.rows, tune)
        v32 = cells * args.steps / wall32 / 1e9
        kind32, kavg32, kbytes32, a32 = roofline(kt32, cells, 4)
        step32 = sum(v[0] for k, v in kt32.items() if k in KIND_FIELDS) / args.steps
        r32 = roofline_entry("fp32", kind32, kavg32, kbytes32, nz, tuned32.get(kind32))
        st32 = [profiled("fp32", k, nz, tuned32.get(k)) for k in ("u", "v", "w", "scalars")]
        out["fp32"] = {"value": round(v32, 4), "unit": "Gcell-updates/s",
                       "ms_per_step": round(wall32 / args.steps * 1e3, 4),
                       "dominant_kernel": f"{KERNEL_NAMES[tuned32[kind32][0]]} ({kind32})",
                       "roofline": r32,
                       "step_achieved_GBs": round(40 * cells / (step32 * 1e-3) / 1e9, 2),
                       "step_frac": round(40 * cells / (step32 * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                       "step_traffic_over_algorithmic": (round(sum(p["fetch_bytes"] + p["write_bytes"] for p in st32)
                                                               / (40.0 * cells), 3) if all(st32) else None),
                       "kernels_ms_per_launch": {k: round(v[0] / v[1], 4) for k, v in kt32.items()},
                       "speedup_vs_fp64": round(v32 / value, 3)}
        out["tile_shapes"]["fp32"] = tuned32
    if not args.no_extra:
        log("bench: FP32 adjoint step")
        out["adjoint_fp32"] = adjoint_form(st, dt, max(2, args.steps // 2))
        log("bench: FP64 leapfrog timestep")
        out["leapfrog_fp64"] = leapfrog_form(st, dt, max(2, args.steps // 2))
        log("bench: drop-in tendency calls (64^3 and the bench domain, FP64)")
        c1 = ss.make_state(64, 64, 64, 50.0, 2000.0, seed=0, n_q=1)
        out["tendency_fp64"] = {
            "c1_64cubed_forward": tendency_form(c1, 1.0, tvd.FORWARD_STEPPING, 50),
            "c1_64cubed_centred": tendency_form(c1, 1.0, tvd.CENTRED_STEPPING, 50),
            "512x512x256_forward": tendency_form(st, dt, tvd.FORWARD_STEPPING, max(4, args.steps // 2)),
            "512x512x256_centred": tendency_form(st, dt, tvd.CENTRED_STEPPING, max(4, args.steps // 2))}
        del c1
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
    if cpu is not None:
        out["cpu_baseline"] = cpu
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
