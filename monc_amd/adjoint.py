"""Differentiable forward-Euler rollouts of the TVD advection step.

`TvdRollout` owns one device-resident domain (FP32 or FP64) and exposes an
n-step rollout X_0 -> X_n (X = u, v, w, th, q on the (nx+4, ny+4, nz) arrays
with their periodic halo) as a `torch.autograd.Function`:

* forward: the state is advanced by the library's fused step
  (monc_tvd_forward_steps) and a copy of it is kept every `checkpoint_every`
  steps;
* backward: segment by segment from the last, the states inside the segment
  are recomputed from its checkpoint, then the hand-written adjoint step
  (monc_tvd_adjoint_step) carries lambda back one step at a time.  The
  recompute records the limiter arm of every face
  (monc_tvd_forward_step_record) and each adjoint step reads the arms of its
  state (monc_tvd_bind_record); every recomputed state is bitwise the state
  the forward pass stepped from, so the arms are those the forward pass took.
  With `record_arms=False`, or columns of more than 256 levels, the adjoint
  re-derives them from the state instead.

The states of the last segment are kept as the forward pass computes them
(in the segment buffers the backward pass uses), so the backward pass
recomputes every segment but the last.

Memory: ceil(n/k) checkpoints plus k + 1 segment states, each one state (all
fields with halo) in the context precision, and k records of one byte per
point and advected field.

The incoming gradient may be non-zero on halo entries (a loss over whole
arrays): it is folded onto the periodic sources first (monc_tvd_adjoint_fold).
The gradient returned for the initial state is zero on the halo, because the
rollout overwrites the input halo with the periodic image of its interior.
"""
from __future__ import annotations

from typing import List, Sequence

import torch

from . import tvd
from .setup_state import State


class TvdRollout:
    """Device domain + checkpoint buffers for differentiable rollouts."""

    def __init__(self, state: State, precision: int = tvd.FP32, checkpoint_every: int = 8,
                 advect: int = tvd.ADVECT_FLOW | tvd.ADVECT_THETA | tvd.ADVECT_Q, record_arms: bool = True):
        if checkpoint_every < 1:
            raise ValueError("checkpoint_every must be >= 1")
        if not torch.cuda.is_available():
            raise RuntimeError("TvdRollout needs a GPU: the rollout and its adjoint run only in the HIP library")
        self.nfields = len(state.fields)
        self.precision = precision
        self.dtype = torch.float64 if precision == tvd.FP64 else torch.float32
        self.k = checkpoint_every
        n_tr = getattr(state, "n_tracers", 0)
        if n_tr and advect & tvd.ADVECT_THETA:
            advect |= tvd.ADVECT_TRACER  # the reference advects tracers under the theta option
        self.t = tvd.TvdAdvection(state.nz, state.ny, state.nx, state.cx, state.cy, state.grid.metric_vectors(),
                                  n_q=state.n_q, advect=advect, precision=precision, n_tracers=n_tr)
        self.shape = self.t.field_shape()
        dev = torch.device("cuda", torch.cuda.current_device())
        self.prog = [torch.zeros(self.shape, dtype=self.dtype, device=dev) for _ in range(self.nfields)]
        self.adj = [torch.zeros(self.shape, dtype=self.dtype, device=dev) for _ in range(self.nfields)]
        self._bind_prog(self.prog)
        for f in range(self.nfields):
            self.t.bind_field(f, tvd.ADJOINT, self.adj[f])
        adv = set()
        if advect & tvd.ADVECT_FLOW:
            adv |= {0, 1, 2}
        if advect & tvd.ADVECT_THETA:
            adv.add(3)
        if advect & tvd.ADVECT_Q:
            adv |= set(range(4, 4 + state.n_q))
        if advect & tvd.ADVECT_TRACER:
            adv |= set(range(4 + state.n_q, self.nfields))
        self._fixed = [f for f in range(self.nfields) if f not in adv]  # fields the step leaves as they are
        # the recording step runs the march kernel: columns of at most 256 levels
        self.record = bool(record_arms) and state.nz <= 256
        self._seg: List[List[torch.Tensor]] = []  # recompute buffers, reused across segments
        self._rec: List[List] = []  # limiter arms of the step from each segment state
        self._seg_tag = None  # the forward pass whose last segment the buffers hold

    # helpers -------------------------------------------------------------------
    def _bind_prog(self, fields: Sequence[torch.Tensor]) -> None:
        for f, a in enumerate(fields):
            self.t.bind_field(f, tvd.PROG, a)

    def _sync_stream(self) -> None:
        self.t.set_stream(torch.cuda.current_stream().cuda_stream)

    def _load(self, fields: Sequence[torch.Tensor]) -> None:
        for p, f in zip(self.prog, fields):
            p.copy_(f)

    def _step_into(self, dt: float, src: Sequence[torch.Tensor], dst: Sequence[torch.Tensor], rec=None) -> None:
        """dst = one forward step of src (monc_tvd_forward_step_into), recording
        the limiter arms into `rec` when given."""
        self.t.forward_step_into(dt, src, dst, rec)
        for f in self._fixed:
            dst[f].copy_(src[f])

    def _rec_buffers(self, n: int) -> List[List]:
        while len(self._rec) < n:
            self._rec.append([None if f in self._fixed else
                              torch.empty(self.shape, dtype=torch.uint8, device=self.prog[0].device)
                              for f in range(self.nfields)])
        return self._rec

    def _bind_rec(self, rec) -> None:
        for f in range(self.nfields):
            if f not in self._fixed:
                self.t.bind_record(f, None if rec is None else rec[f])

    def _snapshot(self) -> List[torch.Tensor]:
        return [p.clone() for p in self.prog]

    def _seg_buffers(self, n: int) -> List[List[torch.Tensor]]:
        while len(self._seg) < n:
            self._seg.append([torch.empty_like(p) for p in self.prog])
        return self._seg

    # public --------------------------------------------------------------------
    def rollout(self, fields: Sequence[torch.Tensor], dt: float, nsteps: int):
        """X_n as a tuple of tensors, differentiable w.r.t. `fields` (X_0)."""
        return RolloutFn.apply(self, float(dt), int(nsteps), *fields)

    @torch.no_grad()
    def forward_only(self, fields: Sequence[torch.Tensor], dt: float, nsteps: int) -> List[torch.Tensor]:
        self._sync_stream()
        self._load(fields)
        self.t.halo_fill()
        self.t.forward_steps(dt, nsteps)
        return self._snapshot()

    def close(self) -> None:
        self.t.close()
        self._seg = []
        self._rec = []


class RolloutFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, eng: TvdRollout, dt: float, nsteps: int, *fields):
        eng._sync_stream()
        eng._load(fields)
        eng.t.halo_fill()
        ckpts = []
        done = 0
        tag = object()
        while done < nsteps:
            n = min(eng.k, nsteps - done)
            if done + n < nsteps:
                ckpts.append((done, eng._snapshot()))
                eng.t.forward_steps(dt, n)
            else:  # last segment: keep every state (and arms) the backward pass needs
                seg = eng._seg_buffers(n)
                rec = eng._rec_buffers(n) if eng.record else [None] * n
                for b, p in zip(seg[0], eng.prog):
                    b.copy_(p)
                for m in range(1, n):
                    eng._step_into(dt, seg[m - 1], seg[m], rec[m - 1])
                eng._step_into(dt, seg[n - 1], eng.prog, rec[n - 1])
                ckpts.append((done, None))
                eng._seg_tag = tag
            done += n
        ctx.eng, ctx.dt, ctx.nsteps, ctx.ckpts, ctx.tag = eng, dt, nsteps, ckpts, tag
        return tuple(eng._snapshot())

    @staticmethod
    def backward(ctx, *grads):
        eng, dt = ctx.eng, ctx.dt
        eng._sync_stream()
        for a, g in zip(eng.adj, grads):
            if g is None:
                a.zero_()
            else:
                a.copy_(g)
        eng.t.adjoint_fold()
        seg = eng._seg_buffers(eng.k + 1)
        rec = eng._rec_buffers(eng.k) if eng.record else [None] * eng.k
        for start, ck in reversed(ctx.ckpts):
            n = min(eng.k, ctx.nsteps - start)
            if ck is None and eng._seg_tag is not ctx.tag:
                raise RuntimeError("another rollout of this engine reused its segment buffers before this backward")
            if ck is not None:
                # recompute X_start .. X_{start+n-1} in the working buffers, and
                # the arms of the step from each (the last step's state is the
                # next checkpoint, written to a spare buffer)
                for s, c in zip(seg[0], ck):
                    s.copy_(c)
                for m in range(1, n):
                    eng._step_into(dt, seg[m - 1], seg[m], rec[m - 1])
                if eng.record:
                    eng._step_into(dt, seg[n - 1], seg[n], rec[n - 1])
            # adjoint sweep over the segment, newest state first
            for m in reversed(range(n)):
                eng._bind_prog(seg[m])
                if eng.record:
                    eng._bind_rec(rec[m])
                eng.t.adjoint_step(dt)
        eng._bind_prog(eng.prog)
        if eng.record:
            eng._bind_rec(None)
        eng._seg_tag = None
        ctx.ckpts = None
        return (None, None, None) + tuple(a.clone() for a in eng.adj)
