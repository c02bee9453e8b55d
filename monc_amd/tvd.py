"""Python mirror of MONC's tvd_advection component over libmonc_tvd.so.

The reference component exposes init / timestep / finalise callbacks on the
model state (components/tvdadvection/src/tvdadvection.F90:62-106, :286-481,
:528-599, :485-524).  `TvdAdvection` keeps that surface for the test harness and
the benchmark: `initialise` copies the grid metrics to the device, `timestep`
adds the advection tendencies to the source fields, `finalise` frees the
device state.  All numerics run in the HIP library; if the library is not
built, constructing a `TvdAdvection` raises.
"""
from __future__ import annotations

import ctypes
from pathlib import Path
from typing import Optional, Sequence

import numpy as np

from .setup_state import State

LIB_PATH = Path(__file__).resolve().parent / "libmonc_tvd.so"

OK, ERR_ARG, ERR_HIP, ERR_STATE = 0, 1, 2, 3
CENTRED_STEPPING, FORWARD_STEPPING = 0, 1  # state.F90:15
FP64, FP32 = 0, 1
ADVECT_FLOW, ADVECT_THETA, ADVECT_Q, ADVECT_TRACER = 1, 2, 4, 8
U, V, W, TH = 0, 1, 2, 3
PROG, ZFIELD, SOURCE, ADJOINT, TANGENT = 0, 1, 2, 3, 4

EXPORTS = (
    "monc_tvd_create", "monc_tvd_destroy", "monc_tvd_set_field", "monc_tvd_get_field",
    "monc_tvd_bind_field", "monc_tvd_field_device_ptr", "monc_tvd_set_stream",
    "monc_tvd_synchronize", "monc_tvd_halo_fill", "monc_tvd_tendencies",
    "monc_tvd_forward_steps", "monc_tvd_last_error", "monc_tvd_set_tiling",
    "monc_tvd_set_profiling", "monc_tvd_kernel_times", "monc_tvd_weighted_sum",
    "monc_tvd_adjoint_step", "monc_tvd_adjoint_fold", "monc_tvd_tangent_steps", "monc_tvd_march_shapes",
    "monc_tvd_set_diagnostics", "monc_tvd_get_diagnostic", "monc_tvd_autotune",
    "monc_tvd_leapfrog_steps", "monc_tvd_forward_step_into", "monc_tvd_forward_step_record",
    "monc_tvd_bind_record", "monc_thadv_tendencies",
    "monc_pw_tendencies", "monc_tvd_set_field_active", "monc_tvd_test_fail_allocations",
    "monc_tvd_weighted_sum_noalloc", "monc_tvd_get_tiling", "monc_tvd_get_component_diagnostic",
    "monc_tvd_ensemble_create", "monc_tvd_ensemble_destroy", "monc_tvd_ensemble_layout",
    "monc_tvd_ensemble_set_field", "monc_tvd_ensemble_get_field", "monc_tvd_ensemble_copy_pool",
    "monc_tvd_ensemble_set_dt", "monc_tvd_ensemble_set_stream", "monc_tvd_ensemble_synchronize",
    "monc_tvd_ensemble_halo_fill", "monc_tvd_ensemble_forward_steps", "monc_tvd_ensemble_adjoint_step",
    "monc_tvd_ensemble_adjoint_fold", "monc_tvd_ensemble_set_tiling", "monc_tvd_ensemble_autotune",
)
DIAG_TENDENCY, DIAG_ZFACE = 1, 2
DIAG_3D, DIAG_PROFILE, DIAG_TABS_3D, DIAG_TABS_PROFILE, DIAG_ZFACES = 0, 1, 2, 3, 4
COMPONENT_TVD, COMPONENT_PW, COMPONENT_THADV = 0, 1, 2
KERNEL_MARCH, KERNEL_CELL, KERNEL_STAGE = 0, 1, 2
KERNEL_KINDS = ("u", "v", "w", "scalars", "halo", "cell", "unused")


def Q(n: int) -> int:
    return 4 + n


def TRACER(n_q: int, n: int) -> int:
    return 4 + n_q + n


class TvdError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"monc_tvd error {code}: {msg}")
        self.code = code


_lib: Optional[ctypes.CDLL] = None


def load_library(path: Path = LIB_PATH) -> ctypes.CDLL:
    """Load libmonc_tvd.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not path.exists():
        raise RuntimeError(f"{path} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = ctypes.CDLL(str(path))
    c_int, c_double, c_void_p = ctypes.c_int, ctypes.c_double, ctypes.c_void_p
    dp = ctypes.POINTER(ctypes.c_double)
    lib.monc_tvd_create.argtypes = [c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_double, c_double,
                                    dp, dp, dp, dp, dp, dp, dp, dp, c_int, ctypes.POINTER(c_void_p)]
    lib.monc_tvd_set_field_active.argtypes = [c_void_p, c_int, c_int]
    lib.monc_tvd_test_fail_allocations.argtypes = [c_int]
    lib.monc_tvd_destroy.argtypes = [c_void_p]
    lib.monc_tvd_set_field.argtypes = [c_void_p, c_int, c_int, dp]
    lib.monc_tvd_get_field.argtypes = [c_void_p, c_int, c_int, dp]
    lib.monc_tvd_bind_field.argtypes = [c_void_p, c_int, c_int, c_void_p]
    lib.monc_tvd_field_device_ptr.argtypes = [c_void_p, c_int, c_int, ctypes.POINTER(c_void_p)]
    lib.monc_tvd_set_stream.argtypes = [c_void_p, c_void_p]
    lib.monc_tvd_set_tiling.argtypes = [c_void_p, c_int, c_int, c_int]
    lib.monc_tvd_march_shapes.argtypes = [c_void_p, ctypes.POINTER(c_int), c_int]
    lib.monc_tvd_set_diagnostics.argtypes = [c_void_p, c_int, dp]
    lib.monc_tvd_autotune.argtypes = [c_void_p, c_double, c_int, ctypes.POINTER(c_int)]
    lib.monc_tvd_get_tiling.argtypes = [c_void_p, ctypes.POINTER(c_int)]
    lib.monc_tvd_get_diagnostic.argtypes = [c_void_p, c_int, c_int, dp]
    lib.monc_tvd_get_component_diagnostic.argtypes = [c_void_p, c_int, c_int, c_int, dp]
    lib.monc_tvd_set_profiling.argtypes = [c_void_p, c_int]
    lib.monc_tvd_weighted_sum.argtypes = [c_void_p, c_int, c_int, dp, dp]
    lib.monc_tvd_weighted_sum_noalloc.argtypes = [c_void_p, c_int, c_int, dp, dp]
    lib.monc_tvd_kernel_times.argtypes = [c_void_p, ctypes.POINTER(c_double), ctypes.POINTER(ctypes.c_longlong),
                                          c_int, c_int]
    lib.monc_tvd_synchronize.argtypes = [c_void_p]
    lib.monc_tvd_halo_fill.argtypes = [c_void_p]
    lib.monc_tvd_tendencies.argtypes = [c_void_p, c_double, c_int, c_int, c_int]
    lib.monc_tvd_forward_steps.argtypes = [c_void_p, c_double, c_int]
    lib.monc_tvd_forward_step_into.argtypes = [c_void_p, c_double, ctypes.POINTER(c_void_p),
                                               ctypes.POINTER(c_void_p)]
    lib.monc_tvd_forward_step_record.argtypes = [c_void_p, c_double, ctypes.POINTER(c_void_p),
                                                 ctypes.POINTER(c_void_p), ctypes.POINTER(c_void_p)]
    lib.monc_tvd_bind_record.argtypes = [c_void_p, c_int, c_void_p]
    lib.monc_tvd_leapfrog_steps.argtypes = [c_void_p, c_double, c_double, c_double, c_double, c_int, c_int]
    lib.monc_tvd_adjoint_step.argtypes = [c_void_p, c_double]
    lib.monc_tvd_adjoint_fold.argtypes = [c_void_p]
    dp = ctypes.POINTER(c_double)
    lib.monc_thadv_tendencies.argtypes = [c_void_p, c_int, dp, c_int, dp] + [c_double] * 7
    lib.monc_pw_tendencies.argtypes = [c_void_p, c_int, c_int, c_int, c_double, c_double]
    lib.monc_tvd_tangent_steps.argtypes = [c_void_p, c_double, c_int]
    ip = ctypes.POINTER(c_int)
    llp = ctypes.POINTER(ctypes.c_longlong)
    lib.monc_tvd_ensemble_create.argtypes = [c_int, ip, ip, ip, dp, dp, dp, c_int, ctypes.POINTER(c_void_p)]
    lib.monc_tvd_ensemble_destroy.argtypes = [c_void_p]
    lib.monc_tvd_ensemble_layout.argtypes = [c_void_p, c_int, llp, llp]
    lib.monc_tvd_ensemble_set_field.argtypes = [c_void_p, c_int, c_int, c_int, dp]
    lib.monc_tvd_ensemble_get_field.argtypes = [c_void_p, c_int, c_int, c_int, dp]
    lib.monc_tvd_ensemble_copy_pool.argtypes = [c_void_p, c_int, c_void_p, c_void_p]
    lib.monc_tvd_ensemble_set_dt.argtypes = [c_void_p, dp]
    lib.monc_tvd_ensemble_set_stream.argtypes = [c_void_p, c_void_p]
    lib.monc_tvd_ensemble_synchronize.argtypes = [c_void_p]
    lib.monc_tvd_ensemble_halo_fill.argtypes = [c_void_p]
    lib.monc_tvd_ensemble_forward_steps.argtypes = [c_void_p, c_int]
    lib.monc_tvd_ensemble_adjoint_step.argtypes = [c_void_p, c_void_p]
    lib.monc_tvd_ensemble_adjoint_fold.argtypes = [c_void_p]
    lib.monc_tvd_ensemble_set_tiling.argtypes = [c_void_p, c_int, c_int, c_int]
    lib.monc_tvd_ensemble_autotune.argtypes = [c_void_p, c_int, ip]
    lib.monc_tvd_last_error.argtypes = []
    lib.monc_tvd_last_error.restype = ctypes.c_char_p
    for name in EXPORTS:
        if name != "monc_tvd_last_error":
            getattr(lib, name).restype = c_int
    _lib = lib
    return lib


def _check(rc: int) -> None:
    if rc != OK:
        raise TvdError(rc, load_library().monc_tvd_last_error().decode())


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class TvdAdvection:
    """One device-resident local domain of the tvd_advection component."""

    def __init__(self, nz: int, ny: int, nx: int, cx: float, cy: float, metrics: dict, n_q: int = 1,
                 advect: int = ADVECT_FLOW | ADVECT_THETA | ADVECT_Q, precision: int = FP64, halo: int = 2,
                 n_tracers: int = 0):
        self._lib = load_library()
        self.nz, self.ny, self.nx, self.n_q, self.n_tracers = nz, ny, nx, n_q, n_tracers
        self.precision = precision
        self._keep = [np.ascontiguousarray(metrics[k], dtype=np.float64) for k in
                      ("rdz", "rdzn", "dz", "dzn", "tzc1", "tzc2", "tzd1", "tzd2")]
        for v in self._keep:
            if v.shape != (nz,):
                raise ValueError("metric vectors must have length nz")
        h = ctypes.c_void_p()
        _check(self._lib.monc_tvd_create(nz, ny, nx, halo, n_q, n_tracers, advect, cx, cy,
                                         *[_dptr(v) for v in self._keep], precision, ctypes.byref(h)))
        self._h = h

    # the component surface -------------------------------------------------
    @classmethod
    def initialise(cls, state: State, precision: int = FP64,
                   advect: int = ADVECT_FLOW | ADVECT_THETA | ADVECT_Q) -> "TvdAdvection":
        """initialisation_callback: metrics to the device, fields staged.  With
        tracers in the state, ADVECT_THETA also sets ADVECT_TRACER, as the
        reference derives advect_tracer from the theta option
        (tvdadvection.F90:352)."""
        n_tr = getattr(state, "n_tracers", 0)
        if n_tr and advect & ADVECT_THETA:
            advect |= ADVECT_TRACER
        t = cls(state.nz, state.ny, state.nx, state.cx, state.cy, state.grid.metric_vectors(), n_q=state.n_q,
                advect=advect, precision=precision, n_tracers=n_tr)
        for n, f in enumerate(state.fields):
            t.set_field(n, PROG, f)
        for n, f in enumerate(state.zfields):
            t.set_field(n, ZFIELD, f)
        return t

    def timestep(self, dtm: float, field_stepping: int = FORWARD_STEPPING, momentum_stepping: int = CENTRED_STEPPING,
                 scalar_stepping: int = CENTRED_STEPPING) -> None:
        """timestep_callback over every column: s += advection tendencies."""
        self.tendencies(dtm, field_stepping, momentum_stepping, scalar_stepping)

    def finalise(self) -> None:
        self.close()

    # entry points ------------------------------------------------------------
    def field_shape(self):
        return (self.nx + 4, self.ny + 4, self.nz)

    def set_field(self, field: int, variant: int, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.shape != self.field_shape():
            raise ValueError(f"field shape {a.shape} != {self.field_shape()}")
        _check(self._lib.monc_tvd_set_field(self._h, field, variant, _dptr(a)))

    def get_field(self, field: int, variant: int) -> np.ndarray:
        out = np.empty(self.field_shape(), dtype=np.float64)
        _check(self._lib.monc_tvd_get_field(self._h, field, variant, _dptr(out)))
        return out

    def zero_sources(self, fields: Sequence[int]) -> None:
        z = np.zeros(self.field_shape())
        for f in fields:
            self.set_field(f, SOURCE, z)

    def bind_field(self, field: int, variant: int, tensor) -> None:
        """Bind a torch CUDA tensor of shape field_shape() and the context dtype."""
        import torch
        want = torch.float64 if self.precision == FP64 else torch.float32
        if not tensor.is_cuda or tensor.dtype != want or tuple(tensor.shape) != self.field_shape() \
                or not tensor.is_contiguous():
            raise ValueError("bind_field needs a contiguous CUDA tensor of the context dtype and field shape")
        _check(self._lib.monc_tvd_bind_field(self._h, field, variant, ctypes.c_void_p(tensor.data_ptr())))

    def set_field_active(self, field: int, active: bool) -> None:
        """th%active / q(n)%active: an inactive th or q field is not advected."""
        _check(self._lib.monc_tvd_set_field_active(self._h, field, 1 if active else 0))

    @property
    def nfields(self) -> int:
        return 4 + self.n_q + self.n_tracers

    def device_ptr(self, field: int, variant: int) -> int:
        p = ctypes.c_void_p()
        _check(self._lib.monc_tvd_field_device_ptr(self._h, field, variant, ctypes.byref(p)))
        return int(p.value or 0)

    def set_tiling(self, kernel: int = KERNEL_MARCH, rows: int = 0, strip: int = 0) -> None:
        """kernel KERNEL_MARCH (default; rows = cell rows per block, 0 the
        default shape), KERNEL_CELL or KERNEL_STAGE (each kind staged in
        LDS); strip = x columns per block (0: the default)."""
        _check(self._lib.monc_tvd_set_tiling(self._h, kernel, rows, strip))

    def autotune(self, dtm: float, reps: int = 3) -> dict:
        """Time every compiled march and stage shape per field kind on this
        context's fields and keep the fastest; returns {kind: (kernel, rows,
        strip)} with kernel KERNEL_MARCH or KERNEL_STAGE."""
        out = (ctypes.c_int * 12)()
        _check(self._lib.monc_tvd_autotune(self._h, dtm, reps, out))
        return {k: (out[3 * i], out[3 * i + 1], out[3 * i + 2]) for i, k in enumerate(("u", "v", "w", "scalars"))}

    def tiling(self) -> dict:
        """{kind: (kernel, rows, strip)} the forward step uses, defaults resolved."""
        out = (ctypes.c_int * 12)()
        _check(self._lib.monc_tvd_get_tiling(self._h, out))
        return {k: (out[3 * i], out[3 * i + 1], out[3 * i + 2]) for i, k in enumerate(("u", "v", "w", "scalars"))}

    def march_shapes(self) -> list:
        """Cell rows per block of the march shapes compiled for this column height, default first."""
        buf = (ctypes.c_int * 16)()
        n = self._lib.monc_tvd_march_shapes(self._h, buf, 16)
        if n < 0:
            _check(n)
        return [buf[i] for i in range(min(n, 16))]

    def set_diagnostics(self, flags: int, rprefrcp: Optional[np.ndarray] = None) -> None:
        """DIAG_TENDENCY and/or DIAG_ZFACE recording on every tendencies() call
        (DIAG_TENDENCY also on every pw_tendencies() and thadv_tendencies()
        call, kept per component); rprefrcp (length nz) enables the tabs
        tendency."""
        r = None
        if rprefrcp is not None:
            r = np.ascontiguousarray(rprefrcp, dtype=np.float64)
            if r.shape != (self.nz,):
                raise ValueError("rprefrcp must have length nz")
        _check(self._lib.monc_tvd_set_diagnostics(self._h, flags, _dptr(r) if r is not None else None))

    def get_diagnostic(self, field: int, which: int, component: int = COMPONENT_TVD) -> np.ndarray:
        """A recorded diagnostic of a component (COMPONENT_TVD, COMPONENT_PW,
        COMPONENT_THADV): DIAG_3D / DIAG_TABS_3D as (nx, ny, nz) arrays
        (x slowest, the (z,y,x) column-major layout), DIAG_PROFILE /
        DIAG_TABS_PROFILE as (nz,), DIAG_ZFACES (TVD only) in the field shape."""
        if which in (DIAG_3D, DIAG_TABS_3D):
            out = np.zeros((self.nx, self.ny, self.nz))
        elif which in (DIAG_PROFILE, DIAG_TABS_PROFILE):
            out = np.zeros(self.nz)
        else:
            out = np.zeros(self.field_shape())
        if component == COMPONENT_TVD:
            _check(self._lib.monc_tvd_get_diagnostic(self._h, field, which, _dptr(out)))
        else:
            _check(self._lib.monc_tvd_get_component_diagnostic(self._h, component, field, which, _dptr(out)))
        return out

    def weighted_sum(self, field: int, variant: int, weights: np.ndarray) -> float:
        """sum_k w(k) sum_{j,i} f(k,j,i) over the interior, accumulated in FP64 on the device."""
        w = np.ascontiguousarray(weights, dtype=np.float64)
        if w.shape != (self.nz,):
            raise ValueError("weights must have length nz")
        out = np.zeros(1)
        _check(self._lib.monc_tvd_weighted_sum(self._h, field, variant, _dptr(w), _dptr(out)))
        return float(out[0])

    def weighted_sum_noalloc(self, field: int, variant: int, weights: np.ndarray) -> float:
        """weighted_sum through the entry point that allocates no device
        memory (fixed per-GPU scratch, nz <= 4096)."""
        w = np.ascontiguousarray(weights, dtype=np.float64)
        if w.shape != (self.nz,):
            raise ValueError("weights must have length nz")
        out = np.zeros(1)
        _check(self._lib.monc_tvd_weighted_sum_noalloc(self._h, field, variant, _dptr(w), _dptr(out)))
        return float(out[0])

    def set_profiling(self, on: bool) -> None:
        _check(self._lib.monc_tvd_set_profiling(self._h, 1 if on else 0))

    def kernel_times(self, reset: bool = True) -> dict:
        """{kind: (total ms, launches)} from the HIP events of profiled launches."""
        n = len(KERNEL_KINDS)
        ms = (ctypes.c_double * n)()
        cnt = (ctypes.c_longlong * n)()
        _check(self._lib.monc_tvd_kernel_times(self._h, ms, cnt, n, 1 if reset else 0))
        return {KERNEL_KINDS[k]: (ms[k], cnt[k]) for k in range(n) if cnt[k]}

    def set_stream(self, stream_handle: int) -> None:
        _check(self._lib.monc_tvd_set_stream(self._h, ctypes.c_void_p(stream_handle)))

    def synchronize(self) -> None:
        _check(self._lib.monc_tvd_synchronize(self._h))

    def halo_fill(self) -> None:
        _check(self._lib.monc_tvd_halo_fill(self._h))

    def tendencies(self, dtm: float, field_stepping: int, momentum_stepping: int, scalar_stepping: int) -> None:
        _check(self._lib.monc_tvd_tendencies(self._h, dtm, field_stepping, momentum_stepping, scalar_stepping))

    def forward_steps(self, dtm: float, nsteps: int) -> None:
        _check(self._lib.monc_tvd_forward_steps(self._h, dtm, nsteps))

    def _check_tensor(self, t, what: str) -> None:
        import torch
        want = torch.float64 if self.precision == FP64 else torch.float32
        if not t.is_cuda or t.dtype != want or tuple(t.shape) != self.field_shape() or not t.is_contiguous():
            raise ValueError(f"{what} must be a contiguous CUDA tensor of the context dtype and field shape")

    def forward_step_into(self, dtm: float, src, dst, rec=None) -> None:
        """One forward step from the torch tensors `src` (one per field) into
        `dst`, without touching the context's own fields.  Every tensor must be
        a contiguous CUDA tensor of the context's dtype and field shape, and no
        destination may share memory with a source or another destination.
        With `rec` (one uint8 CUDA tensor of the field shape per field, or None
        for a field the step leaves alone) the step also records every face's
        limiter arm there (monc_tvd_forward_step_record)."""
        n = len(src)
        if n != len(dst) or n != self.nfields:
            raise ValueError("one source and one destination tensor per field")
        for i, t in enumerate(src):
            self._check_tensor(t, f"source {i}")
        for i, t in enumerate(dst):
            self._check_tensor(t, f"destination {i}")
        nbytes = int(np.prod(self.field_shape())) * (8 if self.precision == FP64 else 4)
        spans = [(t.data_ptr(), t.data_ptr() + nbytes) for t in src]
        outs = [(t.data_ptr(), t.data_ptr() + nbytes) for t in dst]
        for i, (a0, a1) in enumerate(outs):
            for j, (b0, b1) in enumerate(spans + [o for m, o in enumerate(outs) if m != i]):
                if a0 < b1 and b0 < a1:
                    raise ValueError(f"destination {i} overlaps another source or destination")
        a = (ctypes.c_void_p * n)(*[t.data_ptr() for t in src])
        b = (ctypes.c_void_p * n)(*[t.data_ptr() for t in dst])
        if rec is None:
            _check(self._lib.monc_tvd_forward_step_into(self._h, dtm, a, b))
            return
        if len(rec) != n:
            raise ValueError("one record tensor (or None) per field")
        for i, t in enumerate(rec):
            if t is not None:
                self._check_record(t, f"record {i}")
        r = (ctypes.c_void_p * n)(*[None if t is None else t.data_ptr() for t in rec])
        _check(self._lib.monc_tvd_forward_step_record(self._h, dtm, a, b, r))

    def _check_record(self, t, what: str) -> None:
        import torch
        if not t.is_cuda or t.dtype != torch.uint8 or tuple(t.shape) != self.field_shape() or not t.is_contiguous():
            raise ValueError(f"{what} must be a contiguous CUDA uint8 tensor of the field shape")

    def bind_record(self, field: int, tensor) -> None:
        """The limiter arms adjoint_step reads for `field` (a uint8 tensor a
        recording step wrote), or None to re-derive them from the state."""
        if tensor is not None:
            self._check_record(tensor, f"record of field {field}")
        _check(self._lib.monc_tvd_bind_record(self._h, field, None if tensor is None else tensor.data_ptr()))

    def leapfrog_steps(self, dtm: float, tsmth: float, nsteps: int, ugal: float = 0.0, vgal: float = 0.0,
                       first_forward: bool = True) -> None:
        """MONC timesteps of an advection-only run (clearsourceterms, TVD,
        stepfields, swap_smooth), resident: the first one forward if asked,
        then leapfrog with the Robert filter."""
        _check(self._lib.monc_tvd_leapfrog_steps(self._h, dtm, tsmth, ugal, vgal, nsteps, 1 if first_forward else 0))

    def adjoint_step(self, dtm: float) -> None:
        _check(self._lib.monc_tvd_adjoint_step(self._h, dtm))

    def adjoint_fold(self) -> None:
        _check(self._lib.monc_tvd_adjoint_fold(self._h))

    def thadv_tendencies(self, anelastic: bool, dthref=None, baroclinicity: bool = False, thref=None,
                         thref0: float = 0.0, fcoriol: float = 0.0, g: float = 9.81, rate_x: float = 0.0,
                         rate_y: float = 0.0, ugal: float = 0.0, vgal: float = 0.0) -> None:
        """th_advection's source terms added to the th source field
        (thadvection.F90:124-211); dthref, thref: length-nz FP64 vectors."""
        def vec(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.float64)
            if a.shape != (self.nz,):
                raise ValueError(f"expected a vector of length nz={self.nz}, got {a.shape}")
            return a
        dth, thr = vec(dthref), vec(thref)
        dp = ctypes.POINTER(ctypes.c_double)
        _check(self._lib.monc_thadv_tendencies(
            self._h, int(bool(anelastic)), dth.ctypes.data_as(dp) if dth is not None else None,
            int(bool(baroclinicity)), thr.ctypes.data_as(dp) if thr is not None else None,
            thref0, fcoriol, g, rate_x, rate_y, ugal, vgal))

    def pw_tendencies(self, advect_flow: bool = True, advect_th: bool = True, advect_q: bool = True, *,
                      tcx: float, tcy: float) -> None:
        """pw_advection's centred tendencies (pwadvection.F90:258-610): su, sv,
        sw and sth overwritten, sq added to, interior columns."""
        _check(self._lib.monc_pw_tendencies(self._h, int(bool(advect_flow)), int(bool(advect_th)),
                                            int(bool(advect_q)), tcx, tcy))

    def tangent_steps(self, dtm: float, nsteps: int) -> None:
        _check(self._lib.monc_tvd_tangent_steps(self._h, dtm, nsteps))

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.monc_tvd_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
