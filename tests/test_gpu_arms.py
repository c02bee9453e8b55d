"""Limiter arms recorded by the forward step for the adjoint.

* monc_tvd_forward_step_record writes, per face owner and field, the arm of
  the limiter each of its three faces took; they match the arms the oracle
  records (which tests/test_oracle.py pins against Leonard's bounds), on
  ragged column heights, for every field kind, FP64 and FP32, with the march
  kernel and with the stage kernel (row tiles and strips partial and exact),
  and the step itself matches monc_tvd_forward_step_into in the same tiling
  bit for bit;
* the adjoint step reading the recorded arms (monc_tvd_bind_record) agrees
  with the adjoint that re-derives them from the state.
"""
import numpy as np
import pytest
import torch

import oracle
from monc_amd import setup_state as ss
from monc_amd import tvd

pytestmark = pytest.mark.gpu

GRIDS = [(33, 10, 12, 1), (70, 9, 8, 2), (130, 6, 7, 1)]  # nz, ny, nx, n_q


def _record(st, precision, tiling=None):
    t = tvd.TvdAdvection.initialise(st, precision=precision)
    dt = torch.float64 if precision == tvd.FP64 else torch.float32
    dev = torch.device("cuda", 0)
    try:
        if tiling is not None:
            t.set_tiling(*tiling)
        t.set_stream(torch.cuda.current_stream().cuda_stream)
        src = [torch.from_numpy(ss.halo_fill(f.copy())).to(dev, dt).contiguous() for f in st.fields]
        dst = [torch.empty_like(s) for s in src]
        ref = [torch.empty_like(s) for s in src]
        rec = [torch.full(t.field_shape(), 255, dtype=torch.uint8, device=dev) for _ in src]
        dtm = ss.cfl_dt(st, 0.3)
        t.forward_step_into(dtm, src, dst, rec)
        t.forward_step_into(dtm, src, ref)
        torch.cuda.synchronize()
        return dtm, [r.cpu().numpy() for r in rec], [d.cpu().numpy() for d in dst], [r.cpu().numpy() for r in ref]
    finally:
        t.close()


# the march kernel (the default at these sizes), and the stage kernel in
# 4-row tiles (ny = 10, 9, 6: three, three and two tiles, the last partial)
# with strips of 5 columns, and in 11-row tiles (the scalars in 4) with
# strips of 16
TILINGS = [None, (tvd.KERNEL_STAGE, 4, 5), (tvd.KERNEL_STAGE, 11, 16)]


@pytest.mark.parametrize("tiling", TILINGS, ids=["march", "stage4", "stage11"])
@pytest.mark.parametrize("precision", [tvd.FP64, tvd.FP32], ids=["fp64", "fp32"])
@pytest.mark.parametrize("grid", GRIDS, ids=[f"{g[2]}x{g[1]}x{g[0]}" for g in GRIDS])
def test_recorded_arms_match_the_oracle(grid, precision, tiling):
    nz, ny, nx, n_q = grid
    st = ss.make_state(nz, ny, nx, 50.0, 25.0 * (nz - 1), seed=41, n_q=n_q, with_zfields=False)
    dtm, rec, out, ref = _record(st, precision, tiling)
    for n in range(len(st.fields)):  # recording does not change the step
        assert np.array_equal(out[n], ref[n]), f"field {n}: the recording step differs from the plain step"
    seen = set()
    for n in range(len(st.fields)):
        _, faces = oracle.state_tendencies(st, dtm, 1, 1, 1, rec_field=n, arms=True)
        ax, ay, az = faces[3], faces[4], faces[5]
        r = rec[n].astype(np.int64)
        kmax = nz - 1 if n == 2 else nz
        # x faces: owners j = 3..ny+2, i = 2..nx+2, k = 2..kmax
        K = slice(1, kmax)
        gx, ox = r[1:nx + 2, 2:ny + 2, K] & 3, ax[1:nx + 2, 2:ny + 2, K]
        # y faces: owners j = 2..ny+1 (the oracle takes row ny+2 from the wrap), i = 3..nx+2
        gy, oy = (r[2:nx + 2, 1:ny + 1, K] >> 2) & 3, ay[2:nx + 2, 1:ny + 1, K]
        # z faces: owner k's top face is the oracle's flux_z(k+1); k = kmin..nz-1
        kmin = 1 if n == 2 else 2
        gz, oz = (r[2:nx + 2, 2:ny + 2, kmin - 1:nz - 1] >> 4) & 3, az[2:nx + 2, 2:ny + 2, kmin:nz]
        for what, g, o in (("x", gx, ox), ("y", gy, oy), ("z", gz, oz)):
            assert not np.isnan(o).any(), f"field {n} {what}: the oracle records no arm for a compared face"
            bad = int(np.count_nonzero(g != o.astype(np.int64)))
            # FP32 faces sit within rounding of a limiter switch at a few points
            # (th, stored as 300 K, the most: its differences keep few digits)
            allowed = 0 if precision == tvd.FP64 else max(2, g.size // 400)
            assert bad <= allowed, f"field {n} {what}: {bad} of {g.size} arms differ from the oracle's"
            seen |= set(np.unique(g).tolist())
    assert seen == {0, 1, 2, 3}, f"arms seen: {sorted(seen)}"


def _adjoint(st, precision, dtm, lam, record, tiling=None):
    t = tvd.TvdAdvection.initialise(st, precision=precision)
    dt = torch.float64 if precision == tvd.FP64 else torch.float32
    dev = torch.device("cuda", 0)
    try:
        if tiling is not None:
            t.set_tiling(*tiling)
        t.set_stream(torch.cuda.current_stream().cuda_stream)
        t.halo_fill()
        state = [torch.from_numpy(ss.halo_fill(f.copy())).to(dev, dt).contiguous() for f in st.fields]
        for f, s in enumerate(state):
            t.bind_field(f, tvd.PROG, s)
        if record:
            scratch = [torch.empty_like(s) for s in state]
            rec = [torch.empty(t.field_shape(), dtype=torch.uint8, device=dev) for _ in state]
            t.forward_step_into(dtm, state, scratch, rec)
            for f in range(len(state)):
                t.bind_record(f, rec[f])
        for f in range(len(st.fields)):
            t.set_field(f, tvd.ADJOINT, lam[f])
        t.adjoint_step(dtm)
        out = [t.get_field(f, tvd.ADJOINT) for f in range(len(st.fields))]
        torch.cuda.synchronize()
        return out
    finally:
        t.close()


@pytest.mark.parametrize("tiling", [None, (tvd.KERNEL_STAGE, 11, 16)], ids=["march", "stage11"])
@pytest.mark.parametrize("precision", [tvd.FP64, tvd.FP32], ids=["fp64", "fp32"])
def test_adjoint_with_recorded_arms_matches_rederived(precision, tiling):
    st = ss.make_state(70, 20, 24, 50.0, 1700.0, seed=42, n_q=2, with_zfields=False)
    dtm = ss.cfl_dt(st, 0.3)
    rng = np.random.default_rng(3)
    lam = []
    for f in st.fields:
        a = np.zeros(f.shape)
        a[2:-2, 2:-2, :] = rng.standard_normal(a[2:-2, 2:-2, :].shape)
        lam.append(a)
    a = _adjoint(st, precision, dtm, lam, record=False)
    b = _adjoint(st, precision, dtm, lam, record=True, tiling=tiling)
    tol = 1e-12 if precision == tvd.FP64 else 1e-5
    for n, (x, y) in enumerate(zip(a, b)):
        err = float(np.linalg.norm(x - y) / np.linalg.norm(x))
        assert err <= tol, f"field {n}: recorded-arm adjoint differs from the re-derived one by {err:.2e}"


@pytest.mark.parametrize("precision", [tvd.FP64, tvd.FP32], ids=["fp64", "fp32"])
def test_recording_step_matches_the_step_at_scale(precision):
    """512x512x256, default tiling (the stage kernel for the flows, and in
    FP64 for the scalar pair): the recording step writes the plain step's
    values bit for bit, and an arm for every face of every interior owner."""
    st = ss.make_state(256, 512, 512, 25.0, 5000.0, seed=0, n_q=1, with_zfields=False)
    t = tvd.TvdAdvection.initialise(st, precision=precision)
    dt = torch.float64 if precision == tvd.FP64 else torch.float32
    dev = torch.device("cuda", 0)
    try:
        assert all(t.tiling()[k][0] == tvd.KERNEL_STAGE for k in ("u", "v", "w"))
        t.set_stream(torch.cuda.current_stream().cuda_stream)
        src = [torch.from_numpy(ss.halo_fill(f.copy())).to(dev, dt).contiguous() for f in st.fields]
        dst = [torch.empty_like(s) for s in src]
        ref = [torch.empty_like(s) for s in src]
        rec = [torch.full(t.field_shape(), 255, dtype=torch.uint8, device=dev) for _ in src]
        dtm = ss.cfl_dt(st, 0.3)
        t.forward_step_into(dtm, src, dst, rec)
        t.forward_step_into(dtm, src, ref)
        torch.cuda.synchronize()
        for n in range(len(src)):
            assert torch.equal(dst[n], ref[n]), f"field {n}: the recording step differs from the plain step"
            # owners k = 1..nz, j = 3..ny+2, i = 3..nx+2 (stored x, y, z-fastest)
            inner = rec[n][2:-2, 2:-2, :]
            assert int((inner == 255).sum()) == 0, f"field {n}: owners without a recorded arm"
            assert int((inner >= 64).sum()) == 0, f"field {n}: codes outside the six arm bits"
    finally:
        t.close()
