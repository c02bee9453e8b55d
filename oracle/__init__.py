"""ctypes front end of the C restatement in tvd_oracle.c -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may import
this package, and only as the checker.  The product library never uses it.

Parity status per oracle function (the checks live in tests/test_oracle.py):

* oracle_tvd_timestep -- pinned by closed forms (constant field, zero flow,
  linear field under uniform flow, 1-D ULTIMATE-QUICKEST of Leonard 1991 with
  the universal limiter), invariants (mass conservation, translation
  F+c -> s+c*D, periodic shift equivariance, x/y mirror and x<->y
  transposition symmetry), and the y-flux wrap equivalence.  The reference
  Fortran is not buildable here (it needs a Fortran `mpi` module compiled for
  amdflang; the image only has a gfortran-format mpi.mod and no gfortran), so
  no reference-executed vectors exist.  What remains unpinned by any check:
  the result at exact ties (signed zeros, U == 0 exactly) follows the source
  text only.
* oracle_halo_fill -- pinned by the periodic-image identity on random data.
* oracle_vertical_metrics -- pinned by the telescoping identities
  4*rhon(k)*dz(k)*tzc1(k) = rho(k-1) and their tzd counterparts.
* oracle_forward_step -- pinned by direct evaluation of field + dtm*s.
* oracle_profile_total -- pinned by exact integer sums (every partial sum
  exactly representable) and by comparison with a pairwise sum on random data.
* oracle_step_field, oracle_swap_classic, oracle_swap_robert and their
  composition leapfrog_run (the MONC step of an advection-only run) -- pinned
  by the textbook leapfrog with the Robert-Asselin filter
  x(t) + nu*(x(t+1) - 2x(t) + x(t-1)) evaluated independently, the forward
  first step f + dtm*s, the galilean identities (new field + gal, new z-field
  + (1 - tsmth)*gal for u and v), exact swaps at s = 0, and scalar mass
  conservation through both filters.
* oracle_thadv (th_advection, the first sibling plugin) -- pinned by the
  centred vertical advection of a reference profile on a uniform grid
  (-w (thref(k+1) - thref(k-1)) / (2 dz), exact for a quadratic thref), the
  mass-weighted split of each interface's transport on an anelastic grid
  (each of the two cells takes rho(k) w(k) dthref(k) / 2 per unit mass-depth),
  the thermal-wind form of the mean baroclinicity term (-u.grad(thbar) with
  grad(thbar) = (thref0 f / g) (dVg/dz, -dUg/dz)), galilean pairing, and the
  untouched level 1 and halo-inclusive columns.
* oracle_pw (pw_advection, the second sibling plugin) -- pinned against the
  TVD oracle: for constant advected fields (scalars, and u, v advected by
  themselves) the centred scheme's tendency is half of TVD's, whose constant
  field case is pinned above (both reduce to c times the discrete flow
  divergence); by exact centred advection of linear fields under uniform
  flow (-U dq/dx, -V dq/dy, -W dq/dz, -V du/dy, -U dv/dx, -U dw/dx); and by
  the reference's overwrite of su, sv, sw, sth against the accumulation into
  sq, untouched level 1, sw left alone at the top and halo columns skipped.
"""
from __future__ import annotations

import contextlib
import ctypes
import subprocess
from pathlib import Path
from typing import Optional, Sequence

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle_tvd.so"

_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def build_variant(plant: int, outdir: Path) -> Path:
    """Build the restatement with the deliberate mistake PLANT = plant into
    outdir (for the pinning test only: it shows that each planted mistake
    fails at least one check)."""
    out = Path(outdir) / f"liboracle_plant{plant}.so"
    subprocess.run(["gcc", "-O2", "-fPIC", "-std=c11", "-ffp-contract=off", "-fno-fast-math", f"-DPLANT={plant}",
                    "-shared", "-o", str(out), str(HERE / "tvd_oracle.c"), "-lm"], check=True)
    return out


@contextlib.contextmanager
def use_library(path):
    """Temporarily route every call of this module through another build of
    the restatement (a planted-mistake variant)."""
    global _lib
    saved = _lib
    _lib = _load(Path(path))
    try:
        yield
    finally:
        _lib = saved


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = _load(LIB)
    return _lib


def _load(path: Path) -> ctypes.CDLL:
    L = ctypes.CDLL(str(path))
    dp = ctypes.POINTER(ctypes.c_double)
    pp = ctypes.POINTER(dp)
    ci, cd = ctypes.c_int, ctypes.c_double
    ip = ctypes.POINTER(ctypes.c_int)
    L.oracle_tvd_timestep.argtypes = [ci, ci, ci, cd, cd, dp, dp, dp, dp, dp, dp, dp, dp, cd, ci, ci, ci, ci, ci,
                                      ci, ip, pp, pp, pp, ci, dp, dp, dp, dp, dp, dp]
    L.oracle_tvd_timestep.restype = ci
    L.oracle_halo_fill.argtypes = [dp, ci, ci, ci]
    L.oracle_vertical_metrics.argtypes = [ci, dp, dp, dp, dp, dp, dp, dp, dp, dp, dp, dp, dp]
    L.oracle_horizontal_c.argtypes = [cd]
    L.oracle_horizontal_c.restype = cd
    L.oracle_forward_step.argtypes = [dp, dp, dp, ci, ci, ci, cd]
    L.oracle_profile_total.argtypes = [dp, ci, ci, ci, dp]
    L.oracle_step_field.argtypes = [dp, dp, dp, dp, ci, ci, ci, ci, ci, cd, cd, ci, cd, cd]
    L.oracle_swap_classic.argtypes = [dp, dp, ci, ci, ci]
    L.oracle_swap_robert.argtypes = [dp, dp, dp, ci, cd, cd, ci, ci, ci]
    L.oracle_thadv.argtypes = [dp, dp, dp, dp, dp, dp, dp, dp, ci, ci, ci, ci, ci] + [cd] * 7
    pp = ctypes.POINTER(dp)
    L.oracle_pw.argtypes = [dp, dp, dp, dp, pp, dp, dp, dp, dp, pp, ci, dp, dp, dp, dp, cd, cd, cd, cd, ci, ci,
                            ci, ci, ci, ci]
    return L


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _pp(arrs: Sequence[np.ndarray]):
    dp = ctypes.POINTER(ctypes.c_double)
    return (dp * len(arrs))(*[_p(a) for a in arrs])


def tendencies(nz: int, ny: int, nx: int, cx: float, cy: float, metrics: dict, dtm: float, field_stepping: int,
               momentum_stepping: int, scalar_stepping: int, fields: Sequence[np.ndarray],
               zfields: Optional[Sequence[np.ndarray]] = None, sources: Optional[Sequence[np.ndarray]] = None,
               advect: int = 7, rec_field: int = -1, n_tracers: int = 0, active: Optional[Sequence[int]] = None,
               arms: bool = False):
    """One timestep_callback sweep.  Returns (sources, faces) where sources are
    the accumulated s-fields (copies of `sources` or zeros) and faces is
    (fx, fy, fz) for rec_field (or None); with `arms`, (fx, fy, fz, ax, ay, az),
    a* the limiter arm of each face (0 fc, 1 fd, 2 QUICKEST, 3 fu + (fc-fu)/S;
    NaN where no face was computed).  fields: u, v, w, th, the q fields,
    then n_tracers tracers; advect bits 1 flow, 2 theta, 4 q, 8 tracers;
    active: per-field flags (None: all active)."""
    n = len(fields)
    n_q = n - 4 - n_tracers
    if n_q < 0:
        raise ValueError("fewer fields than u, v, w, th and the tracers")
    act = None
    if active is not None:
        if len(active) != n:
            raise ValueError("one active flag per field")
        act = (ctypes.c_int * n)(*[1 if a else 0 for a in active])
    shape = (nx + 4, ny + 4, nz)
    fields = [np.ascontiguousarray(f, dtype=np.float64) for f in fields]
    zfields = [np.ascontiguousarray(f, dtype=np.float64) for f in (zfields if zfields else fields)]
    srcs = [np.zeros(shape) if sources is None else np.array(sources[i], dtype=np.float64, copy=True)
            for i in range(n)]
    m = {k: np.ascontiguousarray(v, dtype=np.float64) for k, v in metrics.items()}
    faces = None
    rec = [None] * 6
    if rec_field >= 0:
        faces = tuple(np.full(shape, np.nan) for _ in range(6 if arms else 3))
        rec = [_p(a) for a in faces] + [None] * (6 - len(faces))
    rc = lib().oracle_tvd_timestep(nz, ny, nx, cx, cy, _p(m["rdz"]), _p(m["rdzn"]), _p(m["dz"]), _p(m["dzn"]),
                                   _p(m["tzc1"]), _p(m["tzc2"]), _p(m["tzd1"]), _p(m["tzd2"]), dtm,
                                   field_stepping, momentum_stepping, scalar_stepping, advect, n_q, n_tracers,
                                   act, _pp(fields), _pp(zfields), _pp(srcs), rec_field, *rec)
    if rc != 0:
        raise RuntimeError(f"oracle_tvd_timestep failed with {rc}")
    return srcs, faces


def state_tendencies(state, dtm: float, field_stepping: int = 1, momentum_stepping: int = 1,
                     scalar_stepping: int = 1, sources=None, advect: int = 7, rec_field: int = -1, active=None,
                     arms: bool = False):
    return tendencies(state.nz, state.ny, state.nx, state.cx, state.cy, state.grid.metric_vectors(), dtm,
                      field_stepping, momentum_stepping, scalar_stepping, state.fields, state.zfields or None,
                      sources, advect, rec_field, getattr(state, "n_tracers", 0), active, arms)


def halo_fill(f: np.ndarray) -> np.ndarray:
    f = np.ascontiguousarray(f, dtype=np.float64)
    nx, ny, nz = f.shape[0] - 4, f.shape[1] - 4, f.shape[2]
    lib().oracle_halo_fill(_p(f), nz, ny, nx)
    return f


def vertical_metrics(z, zn, rho, rhon) -> dict:
    nz = len(z)
    out = {k: np.zeros(nz) for k in ("dz", "dzn", "rdz", "rdzn", "tzc1", "tzc2", "tzd1", "tzd2")}
    a = [np.ascontiguousarray(x, dtype=np.float64) for x in (z, zn, rho, rhon)]
    lib().oracle_vertical_metrics(nz, *[_p(x) for x in a], *[_p(out[k]) for k in
                                  ("dz", "dzn", "rdz", "rdzn", "tzc1", "tzc2", "tzd1", "tzd2")])
    return out


def horizontal_c(dxx: float) -> float:
    return lib().oracle_horizontal_c(dxx)


def forward_step(f: np.ndarray, s: np.ndarray, dtm: float) -> np.ndarray:
    """stepfields forward step over the interior columns; halo left as in f."""
    f = np.ascontiguousarray(f, dtype=np.float64)
    s = np.array(s, dtype=np.float64, copy=True)
    out = f.copy()
    nx, ny, nz = f.shape[0] - 4, f.shape[1] - 4, f.shape[2]
    lib().oracle_forward_step(_p(out), _p(f), _p(s), nz, ny, nx, dtm)
    return out


def forward_run(state, dtm: float, nsteps: int, advect: int = 7, active=None):
    """nsteps of: halo fill, FORWARD tendencies from zero, forward step, for all
    advected fields (the resident stepping the library fuses)."""
    fields = [halo_fill(f.copy()) for f in state.fields]
    n_tr = getattr(state, "n_tracers", 0)
    for _ in range(nsteps):
        srcs, _ = tendencies(state.nz, state.ny, state.nx, state.cx, state.cy, state.grid.metric_vectors(), dtm,
                             1, 1, 1, fields, None, None, advect, -1, n_tr, active)
        adv = advected(len(fields), advect, n_tr, active)
        fields = [halo_fill(forward_step(f, srcs[n], dtm)) if n in adv else f for n, f in enumerate(fields)]
    return fields


def advected(n: int, advect: int, n_tracers: int = 0, active=None):
    """Field ids a sweep advects (tvdadvection.F90:591-594, :661, :706)."""
    act = (lambda f: True) if active is None else (lambda f: bool(active[f]))
    n_q = n - 4 - n_tracers
    s = set()
    if advect & 1:
        s |= {0, 1, 2}
    if advect & 2 and act(3):
        s.add(3)
    if advect & 4:
        s |= {f for f in range(4, 4 + n_q) if act(f)}
    if advect & 8:
        s |= set(range(4 + n_q, n))
    return s


def profile_total(t3d: np.ndarray) -> np.ndarray:
    """Per-level total of a (nx, ny, nz) tendency over its columns, in the
    reference's column order."""
    t = np.ascontiguousarray(t3d, dtype=np.float64)
    nx, ny, nz = t.shape
    out = np.zeros(nz)
    lib().oracle_profile_total(_p(t), nz, ny, nx, _p(out))
    return out


def tendency_diagnostics(state, dtm: float, field_stepping: int, momentum_stepping: int, scalar_stepping: int,
                         sources, rprefrcp=None, advect: int = 7):
    """save_precomponent_tendencies / compute_component_tendencies around one
    sweep: per field tend_3d = s_after - s_before over the interior columns,
    its profile total, and (with rprefrcp) the tabs tendency of th.  Returns
    (after, tend3d, profiles, tabs3d, tabs_profile)."""
    after, _ = state_tendencies(state, dtm, field_stepping, momentum_stepping, scalar_stepping, sources, advect)
    t3d = [a[2:-2, 2:-2, :] - b[2:-2, 2:-2, :] for a, b in zip(after, sources)]
    prof = [profile_total(t) for t in t3d]
    tabs = tabs_p = None
    if rprefrcp is not None:
        r = np.asarray(rprefrcp, dtype=np.float64)
        tabs = after[3][2:-2, 2:-2, :] * r - sources[3][2:-2, 2:-2, :] * r
        tabs_p = profile_total(tabs)
    return after, t3d, prof, tabs, tabs_p


CENTRED, FORWARD = 0, 1  # state.F90:15


def step_field(f, zf, s, dtm: float, direction: int, flow_field: bool = False, gal: float = 0.0,
               timesmooth: bool = False, tsmth: float = 0.0, with_sav: bool = False):
    """stepfields.F90:553-627 on copies: returns (field, zfield, sav or None);
    s(1) is zeroed in a copy of s as the reference does in place."""
    f = np.array(f, dtype=np.float64, copy=True)
    zf = np.array(zf, dtype=np.float64, copy=True)
    s = np.array(s, dtype=np.float64, copy=True)
    sav = np.zeros_like(f) if with_sav else None
    nx, ny, nz = f.shape[0] - 4, f.shape[1] - 4, f.shape[2]
    c1, c2 = 1.0 - 2.0 * tsmth, tsmth
    lib().oracle_step_field(_p(f), _p(zf), _p(s), _p(sav) if sav is not None else None, nz, ny, nx,
                            1 if flow_field else 0, direction, dtm, gal, 1 if timesmooth else 0, c1, c2)
    return f, zf, sav


def swap_classic(f, zf):
    f = np.array(f, dtype=np.float64, copy=True)
    zf = np.array(zf, dtype=np.float64, copy=True)
    nx, ny, nz = f.shape[0] - 4, f.shape[1] - 4, f.shape[2]
    lib().oracle_swap_classic(_p(f), _p(zf), nz, ny, nx)
    return f, zf


def swap_robert(f, zf, sav, kind: int, tsmth: float, gal: float = 0.0):
    """kind 0: u or v (galilean shift gal), 1: w, 2: scalar."""
    f = np.array(f, dtype=np.float64, copy=True)
    zf = np.array(zf, dtype=np.float64, copy=True)
    nx, ny, nz = f.shape[0] - 4, f.shape[1] - 4, f.shape[2]
    sv = np.ascontiguousarray(sav, dtype=np.float64) if sav is not None else None
    lib().oracle_swap_robert(_p(f), _p(zf), _p(sv) if sv is not None else None, kind, tsmth, gal, nz, ny, nx)
    return f, zf


def thadv(sth, u, v, w, tzc1, tzc2, dthref=None, thref=None, anelastic: bool = True,
          baroclinicity: bool = False, thref0: float = 0.0, fcoriol: float = 0.0, g: float = 9.81,
          rate_x: float = 0.0, rate_y: float = 0.0, ugal: float = 0.0, vgal: float = 0.0):
    """th_advection (thadvection.F90:124-211) on a copy of the th source
    field sth, from the prognostics u, v, w; returns the new source field."""
    out = np.array(sth, dtype=np.float64, copy=True)
    nx, ny, nz = out.shape[0] - 4, out.shape[1] - 4, out.shape[2]

    def vec(a):
        return np.ascontiguousarray(a if a is not None else np.zeros(nz), dtype=np.float64)

    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (u, v, w)]
    vs = [vec(a) for a in (tzc1, tzc2, dthref, thref)]
    lib().oracle_thadv(_p(out), *[_p(a) for a in arrs], *[_p(a) for a in vs], nz, ny, nx, int(bool(anelastic)),
                       int(bool(baroclinicity)), thref0, fcoriol, g, rate_x, rate_y, ugal, vgal)
    return out


def pw(fields, sources, grid, cx: float, cy: float, tcx: float, tcy: float, advect_flow: bool = True,
       advect_th: bool = True, advect_q: bool = True):
    """pw_advection (pwadvection.F90:258-610) on copies of the source fields
    [su, sv, sw, sth, sq(1..)] from the prognostics [u, v, w, th, q(1..)];
    returns the new source fields."""
    f = [np.ascontiguousarray(a, dtype=np.float64) for a in fields]
    s = [np.array(a, dtype=np.float64, copy=True) for a in sources]
    nx, ny, nz = f[0].shape[0] - 4, f[0].shape[1] - 4, f[0].shape[2]
    n_q = len(f) - 4
    dp = ctypes.POINTER(ctypes.c_double)
    qs = (dp * max(n_q, 1))(*[_p(a) for a in f[4:]])
    sqs = (dp * max(n_q, 1))(*[_p(a) for a in s[4:]])
    m = [np.ascontiguousarray(getattr(grid, n), dtype=np.float64) for n in ("tzc1", "tzc2", "tzd1", "tzd2")]
    lib().oracle_pw(_p(s[0]), _p(s[1]), _p(s[2]), _p(s[3]), sqs, _p(f[0]), _p(f[1]), _p(f[2]), _p(f[3]), qs, n_q,
                    *[_p(a) for a in m], cx, cy, tcx, tcy, nz, ny, nx, int(bool(advect_flow)), int(bool(advect_th)),
                    int(bool(advect_q)))
    return s


def leapfrog_run(state, dtm: float, tsmth: float, nsteps: int, ugal: float = 0.0, vgal: float = 0.0,
                 first_forward: bool = True, zfields=None):
    """nsteps MONC timesteps of an advection-only run, column semantics of
    timestepper.F90:65-122: halo_swapper fills fields and z-fields, the
    sources are cleared (clearsourceterms), tvd_advection adds the tendencies
    with momentum and scalar stepping CENTRED (steppingdirection.F90:27-32) and
    field stepping FORWARD on the first step when first_forward, CENTRED
    otherwise (:41-48), stepfields steps every field (stepfields.F90:314-351:
    u, v flow fields with ugal/vgal and sav, w with sav, th and q time-smoothed
    when centred), and swap_smooth swaps (classic when forward, Robert filter
    when centred; swapsmooth.F90:44-57).  Returns (fields, zfields)."""
    fields = [halo_fill(np.array(f, dtype=np.float64, copy=True)) for f in state.fields]
    zsrc = zfields if zfields is not None else (state.zfields or state.fields)
    zfs = [halo_fill(np.array(f, dtype=np.float64, copy=True)) for f in zsrc]
    for step in range(nsteps):
        direction = FORWARD if (first_forward and step == 0) else CENTRED
        fields = [halo_fill(f) for f in fields]
        zfs = [halo_fill(z) for z in zfs]
        n_tr = getattr(state, "n_tracers", 0)
        srcs, _ = tendencies(state.nz, state.ny, state.nx, state.cx, state.cy, state.grid.metric_vectors(), dtm,
                             direction, CENTRED, CENTRED, fields, zfs, None, 15 if n_tr else 7, -1, n_tr)
        new_f, new_z = [], []
        for n, (f, z, s) in enumerate(zip(fields, zfs, srcs)):
            flow = n in (0, 1)
            gal = ugal if n == 0 else (vgal if n == 1 else 0.0)
            scalar = n >= 3
            fs, zs, sav = step_field(f, z, s, dtm, direction, flow_field=flow, gal=gal,
                                     timesmooth=scalar and direction == CENTRED, tsmth=tsmth,
                                     with_sav=not scalar)
            if direction == FORWARD:
                fs, zs = swap_classic(fs, zs)
            else:
                fs, zs = swap_robert(fs, zs, sav, 2 if scalar else (1 if n == 2 else 0), tsmth, gal)
            new_f.append(fs)
            new_z.append(zs)
        fields, zfs = new_f, new_z
    return [halo_fill(f) for f in fields], [halo_fill(z) for z in zfs]
