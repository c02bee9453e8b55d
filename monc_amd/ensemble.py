"""Batched ensembles of uneven-size TVD domains on one device.

`TvdEnsemble` wraps the library's monc_tvd_ensemble_* entry points: many
independent local domains (members) of their own sizes, spacings, metric
vectors and timesteps, each advecting u, v, w, th and q, stepped together by
ONE persistent launch per forward-Euler step (pairs of steps replayed from a
captured hipGraph), with the adjoint step, halo fill and fold as persistent
grid-stride launches over all members.

All members share one device pool per variant; `pool_view` gives the
(nx+4, ny+4, nz) view of one member's field inside a pool tensor, the layout
of the single-domain entry points.

`EnsembleRollout` is the differentiable n-step rollout of a whole ensemble
(a `torch.autograd.Function` over pool tensors) with checkpoints every k
steps and recompute of every segment but the last (whose states the forward
pass keeps), as `adjoint.TvdRollout` does for one domain.
"""
from __future__ import annotations

import ctypes
from typing import List, Optional, Sequence

import numpy as np

from . import tvd
from .setup_state import State

FIELDS = 5  # u, v, w, th, q


class TvdEnsemble:
    """Device-resident ensemble of uneven domains (u, v, w, th, q each)."""

    def __init__(self, states: Sequence[State], precision: int = tvd.FP32):
        self._lib = tvd.load_library()
        self.n = len(states)
        self.precision = precision
        self.sizes = [(s.nz, s.ny, s.nx) for s in states]
        c_int = ctypes.c_int
        nz = (c_int * self.n)(*[s.nz for s in states])
        ny = (c_int * self.n)(*[s.ny for s in states])
        nx = (c_int * self.n)(*[s.nx for s in states])
        cx = np.array([s.cx for s in states], dtype=np.float64)
        cy = np.array([s.cy for s in states], dtype=np.float64)
        met = np.concatenate([np.concatenate([s.grid.metric_vectors()[k] for k in
                                              ("rdz", "rdzn", "dz", "dzn", "tzc1", "tzc2", "tzd1", "tzd2")])
                              for s in states]).astype(np.float64)
        h = ctypes.c_void_p()
        tvd._check(self._lib.monc_tvd_ensemble_create(self.n, nz, ny, nx, tvd._dptr(cx), tvd._dptr(cy),
                                                      tvd._dptr(met), precision, ctypes.byref(h)))
        self._h = h
        off, st = ctypes.c_longlong(), ctypes.c_longlong()
        tvd._check(self._lib.monc_tvd_ensemble_layout(self._h, -1, ctypes.byref(off), ctypes.byref(st)))
        self.pool_elems = int(off.value)
        self.offsets, self.strides = [], []
        for m in range(self.n):
            tvd._check(self._lib.monc_tvd_ensemble_layout(self._h, m, ctypes.byref(off), ctypes.byref(st)))
            self.offsets.append(int(off.value))
            self.strides.append(int(st.value))

    @classmethod
    def from_states(cls, states: Sequence[State], dts: Sequence[float], precision: int = tvd.FP32):
        e = cls(states, precision)
        for m, s in enumerate(states):
            for f in range(FIELDS):
                e.set_field(m, f, tvd.PROG, s.fields[f])
        e.set_dt(dts)
        return e

    # staging ------------------------------------------------------------------
    def field_shape(self, m: int):
        nz, ny, nx = self.sizes[m]
        return (nx + 4, ny + 4, nz)

    def set_field(self, m: int, f: int, variant: int, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.shape != self.field_shape(m):
            raise ValueError(f"member {m}: field shape {a.shape} != {self.field_shape(m)}")
        tvd._check(self._lib.monc_tvd_ensemble_set_field(self._h, m, f, variant, tvd._dptr(a)))

    def get_field(self, m: int, f: int, variant: int = tvd.PROG) -> np.ndarray:
        out = np.empty(self.field_shape(m), dtype=np.float64)
        tvd._check(self._lib.monc_tvd_ensemble_get_field(self._h, m, f, variant, tvd._dptr(out)))
        return out

    def set_dt(self, dts: Sequence[float]) -> None:
        d = np.ascontiguousarray(dts, dtype=np.float64)
        if d.shape != (self.n,):
            raise ValueError("one dtm per member")
        tvd._check(self._lib.monc_tvd_ensemble_set_dt(self._h, tvd._dptr(d)))

    # device pools ---------------------------------------------------------------
    def torch_dtype(self):
        import torch
        return torch.float64 if self.precision == tvd.FP64 else torch.float32

    def new_pool(self):
        """A zeroed device tensor in pool layout (checkpoints, gradients)."""
        import torch
        return torch.zeros(self.pool_elems, dtype=self.torch_dtype(), device="cuda")

    def _pool_arg(self, t):
        if t is None:
            return None
        if not t.is_cuda or t.dtype != self.torch_dtype() or t.numel() != self.pool_elems or not t.is_contiguous():
            raise ValueError("a pool buffer is a contiguous CUDA tensor of pool_elems elements of the ensemble dtype")
        return ctypes.c_void_p(t.data_ptr())

    def save_pool(self, variant: int, dst) -> None:
        tvd._check(self._lib.monc_tvd_ensemble_copy_pool(self._h, variant, self._pool_arg(dst), None))

    def load_pool(self, variant: int, src) -> None:
        tvd._check(self._lib.monc_tvd_ensemble_copy_pool(self._h, variant, None, self._pool_arg(src)))

    def pool_view(self, pool, m: int, f: int):
        """(nx+4, ny+4, nz) view of member m's field f inside a pool tensor."""
        nz, ny, nx = self.sizes[m]
        o = self.offsets[m] + f * self.strides[m]
        return pool[o:o + nz * (ny + 4) * (nx + 4)].view(nx + 4, ny + 4, nz)

    # stepping -------------------------------------------------------------------
    def set_stream(self, stream_handle: int) -> None:
        tvd._check(self._lib.monc_tvd_ensemble_set_stream(self._h, ctypes.c_void_p(stream_handle)))

    def synchronize(self) -> None:
        tvd._check(self._lib.monc_tvd_ensemble_synchronize(self._h))

    def halo_fill(self) -> None:
        tvd._check(self._lib.monc_tvd_ensemble_halo_fill(self._h))

    def forward_steps(self, nsteps: int) -> None:
        tvd._check(self._lib.monc_tvd_ensemble_forward_steps(self._h, nsteps))

    def adjoint_step(self, state=None) -> None:
        tvd._check(self._lib.monc_tvd_ensemble_adjoint_step(self._h, self._pool_arg(state)))

    def adjoint_fold(self) -> None:
        tvd._check(self._lib.monc_tvd_ensemble_adjoint_fold(self._h))

    def set_tiling(self, waves: int, strip: int, graph: bool = True) -> None:
        tvd._check(self._lib.monc_tvd_ensemble_set_tiling(self._h, waves, strip, 1 if graph else 0))

    def autotune(self, reps: int = 2):
        out = (ctypes.c_int * 2)()
        tvd._check(self._lib.monc_tvd_ensemble_autotune(self._h, reps, out))
        return int(out[0]), int(out[1])

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.monc_tvd_ensemble_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class EnsembleRollout:
    """Differentiable n-step rollout of a whole ensemble over pool tensors."""

    def __init__(self, ens: TvdEnsemble, checkpoint_every: int = 8):
        import torch
        if checkpoint_every < 1:
            raise ValueError("checkpoint_every must be >= 1")
        if not torch.cuda.is_available():
            raise RuntimeError("EnsembleRollout needs a GPU: the rollout and its adjoint run only in the HIP library")
        self.ens = ens
        self.k = checkpoint_every
        self._seg: List = []  # segment states, reused across segments and rollouts
        self._seg_tag = None  # the forward pass whose last segment they hold

    def _sync_stream(self) -> None:
        import torch
        self.ens.set_stream(torch.cuda.current_stream().cuda_stream)

    def _seg_buffers(self, n: int):
        while len(self._seg) < n:
            self._seg.append(self.ens.new_pool())
        return self._seg

    def rollout(self, x0, nsteps: int):
        """X_n (pool tensor), differentiable w.r.t. x0 (pool tensor)."""
        return _EnsembleRolloutFn.apply(self, int(nsteps), x0)

    def close(self) -> None:
        self._seg = []


def _rollout_backward_segments(eng: EnsembleRollout, ckpts, nsteps: int, tag) -> None:
    e = eng.ens
    seg = eng._seg_buffers(eng.k)
    for start, ck in reversed(ckpts):
        n = min(eng.k, nsteps - start)
        if ck is None and eng._seg_tag is not tag:
            raise RuntimeError("another rollout of this engine reused its segment buffers before this backward")
        if ck is not None:  # recompute the segment's states from its checkpoint
            e.load_pool(tvd.PROG, ck)
            for m in range(n):
                if m:
                    e.forward_steps(1)
                e.save_pool(tvd.PROG, seg[m])
        for m in reversed(range(n)):
            e.adjoint_step(seg[m])
    eng._seg_tag = None


try:
    import torch

    class _EnsembleRolloutFn(torch.autograd.Function):
        @staticmethod
        def forward(ctx, eng: EnsembleRollout, nsteps: int, x0):
            e = eng.ens
            eng._sync_stream()
            e.load_pool(tvd.PROG, x0.detach().contiguous())
            e.halo_fill()
            ckpts = []
            done = 0
            tag = object()
            while done < nsteps:
                n = min(eng.k, nsteps - done)
                if done + n < nsteps:
                    ck = e.new_pool()
                    e.save_pool(tvd.PROG, ck)
                    ckpts.append((done, ck))
                    e.forward_steps(n)
                else:  # last segment: keep every state the backward pass needs
                    seg = eng._seg_buffers(n)
                    for m in range(n):
                        e.save_pool(tvd.PROG, seg[m])
                        e.forward_steps(1)
                    ckpts.append((done, None))
                    eng._seg_tag = tag
                done += n
            out = e.new_pool()
            e.save_pool(tvd.PROG, out)
            ctx.eng, ctx.nsteps, ctx.ckpts, ctx.tag = eng, nsteps, ckpts, tag
            return out

        @staticmethod
        def backward(ctx, g):
            eng = ctx.eng
            e = eng.ens
            eng._sync_stream()
            if g is None:
                g = e.new_pool()
            e.load_pool(tvd.ADJOINT, g.contiguous())
            e.adjoint_fold()
            _rollout_backward_segments(eng, ctx.ckpts, ctx.nsteps, ctx.tag)
            ctx.ckpts = None
            grad = e.new_pool()
            e.save_pool(tvd.ADJOINT, grad)
            return None, None, grad
except ImportError:  # pragma: no cover - torch is part of the harness image
    pass
