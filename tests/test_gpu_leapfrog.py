"""Resident MONC timesteps of an advection-only run (monc_tvd_leapfrog_steps):
clearsourceterms + TVD tendencies + stepfields + swap_smooth in one pass per
field kind, checked through the C ABI against the oracle's composition of
the reference's step (oracle.leapfrog_run) at 1e-12 in FP64, for a forward
first step followed by leapfrog steps with the Robert filter, galilean
shifts, ragged and anelastic grids, a start from existing z-fields and
caller-bound buffers; FP32 against FP64."""
import numpy as np
import pytest

import oracle
from monc_amd import setup_state as ss
from monc_amd import tvd
from tests.helpers import interior

pytestmark = pytest.mark.gpu

# (nz, ny, nx, seed, anelastic scale)
GRIDS = [(16, 12, 12, 0, None), (70, 8, 9, 1, 1000.0), (130, 6, 7, 2, None), (33, 10, 6, 3, 2000.0),
         (256, 5, 6, 4, None)]


def _state(case):
    nz, ny, nx, seed, an = case
    st = ss.make_state(nz, ny, nx, 50.0, 25.0 * (nz - 1), seed=seed, n_q=1, anelastic_scale=an)
    return st, ss.cfl_dt(st, 0.15)


def _rel(a, b):
    ai, bi = interior(a), interior(b)
    return float(np.max(np.abs(ai - bi)) / max(np.max(np.abs(bi)), 1e-300))


def _gpu_leapfrog(st, dt, tsmth, nsteps, ugal=0.0, vgal=0.0, first_forward=True, precision=tvd.FP64, tiling=None):
    t = tvd.TvdAdvection.initialise(st, precision=precision)
    try:
        if tiling is not None:
            t.set_tiling(*tiling)
        t.halo_fill()
        t.leapfrog_steps(dt, tsmth, nsteps, ugal, vgal, first_forward)
        return ([t.get_field(n, tvd.PROG) for n in range(len(st.fields))],
                [t.get_field(n, tvd.ZFIELD) for n in range(len(st.fields))])
    finally:
        t.close()


# the march kernel (the default at these sizes), and the stage kernel in
# 4-row tiles with strips of 5 columns and in 11-row tiles with strips of 16
# (the scalars in every step, the flows in the forward first step)
@pytest.mark.parametrize("tiling", [None, (tvd.KERNEL_STAGE, 4, 5), (tvd.KERNEL_STAGE, 11, 16)],
                         ids=["march", "stage4", "stage11"])
@pytest.mark.parametrize("case", GRIDS, ids=[f"{c[2]}x{c[1]}x{c[0]}" for c in GRIDS])
def test_leapfrog_matches_oracle(gpu, case, tiling):
    st, dt = _state(case)
    f, z = _gpu_leapfrog(st, dt, 0.06, 4, ugal=1.5, vgal=-0.7, tiling=tiling)
    rf, rz = oracle.leapfrog_run(st, dt, 0.06, 4, ugal=1.5, vgal=-0.7)
    for n in range(5):
        assert _rel(f[n], rf[n]) <= 1e-12, ("field", n, _rel(f[n], rf[n]))
        assert _rel(z[n], rz[n]) <= 1e-12, ("zfield", n, _rel(z[n], rz[n]))
        # the kernels write the periodic halo of both time levels
        assert np.array_equal(f[n], ss.halo_fill(f[n].copy())), n
        assert np.array_equal(z[n], ss.halo_fill(z[n].copy())), n


def test_leapfrog_from_existing_zfields(gpu):
    st, dt = _state(GRIDS[0])
    f, z = _gpu_leapfrog(st, dt, 0.1, 3, first_forward=False)
    rf, rz = oracle.leapfrog_run(st, dt, 0.1, 3, first_forward=False)
    for n in range(5):
        assert _rel(f[n], rf[n]) <= 1e-12, n
        assert _rel(z[n], rz[n]) <= 1e-12, n


def test_leapfrog_bound_buffers_odd_steps(gpu):
    import torch
    st, dt = _state(GRIDS[3])
    t = tvd.TvdAdvection(st.nz, st.ny, st.nx, st.cx, st.cy, st.grid.metric_vectors(), n_q=1)
    try:
        prog = [torch.tensor(ss.halo_fill(f.copy()), device=gpu) for f in st.fields]
        zf = [torch.tensor(ss.halo_fill(f.copy()), device=gpu) for f in st.zfields]
        for n in range(5):
            t.bind_field(n, tvd.PROG, prog[n])
            t.bind_field(n, tvd.ZFIELD, zf[n])
        t.leapfrog_steps(dt, 0.06, 3)
        t.synchronize()
        rf, rz = oracle.leapfrog_run(st, dt, 0.06, 3)
        for n in range(5):
            assert _rel(prog[n].cpu().numpy(), rf[n]) <= 1e-12, n
            assert _rel(zf[n].cpu().numpy(), rz[n]) <= 1e-12, n
    finally:
        t.close()


# the FP32 step marching, and staged in 11-row tiles (momentum advecting its
# z-fields as a fourth staged array)
@pytest.mark.parametrize("tiling", [None, (tvd.KERNEL_STAGE, 11, 16)], ids=["march", "stage11"])
def test_leapfrog_fp32_vs_fp64(gpu, tiling):
    st, dt = _state(GRIDS[1])
    f64, z64 = _gpu_leapfrog(st, dt, 0.06, 20)
    f32, z32 = _gpu_leapfrog(st, dt, 0.06, 20, precision=tvd.FP32, tiling=tiling)
    for n in range(5):
        ai, bi = interior(f32[n]), interior(f64[n])
        l2 = float(np.linalg.norm(ai - bi) / np.linalg.norm(bi))
        assert l2 <= 1e-5, (n, l2)
