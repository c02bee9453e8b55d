"""pw_advection (pwadvection.F90:258-610), the second sibling plugin on the
TVD boundary.  The oracle's restatement is pinned against the (separately
pinned) TVD oracle and by exact centred advection of linear fields (CPU
tests); the HIP kernel behind monc_pw_tendencies is checked against the
oracle through the C ABI (GPU tests, 1e-12 in FP64)."""
import numpy as np
import pytest

import oracle
from monc_amd import setup_state as ss

FS, CS = 1, 0


def _state(nz=12, ny=8, nx=10, seed=0, anel=None, n_q=2, dz=None):
    st = ss.make_state(nz, ny, nx, 50.0, 25.0 * (nz - 1), seed=seed, n_q=n_q, anelastic_scale=anel)
    return st


def _tc(st):
    """tcx = 0.25/dx, tcy = 0.25/dy (gridmanager.F90:2638-2641)."""
    return 0.25 / (1.0 / st.cx), 0.25 / (1.0 / st.cy)


def _pw(st, sources=None, **kw):
    tcx, tcy = _tc(st)
    src = sources if sources is not None else [np.zeros_like(f) for f in st.fields]
    return oracle.pw(st.fields, src, st.grid, st.cx, st.cy, tcx, tcy, **kw)


def _tvd(st, fs=FS):
    s, _ = oracle.state_tendencies(st, 1.0, fs, FS, FS)
    return s


def _inner(a, top=True):
    return a[2:-2, 2:-2, 1:] if top else a[2:-2, 2:-2, 1:-1]


@pytest.mark.parametrize("anel", [None, 6000.0])
def test_constant_scalars_give_half_the_tvd_tendency(anel):
    """For a constant advected scalar both schemes reduce to c times a
    discrete flow divergence; the centred form carries half of TVD's."""
    st = _state(seed=2, anel=anel)
    st.fields[3][:] = 301.25
    st.fields[4][:] = -2.5
    st.fields[5][:] = 0.75
    pw = _pw(st)
    tvd = _tvd(st)
    for f in (3, 4, 5):
        ref = 0.5 * _inner(tvd[f])
        np.testing.assert_allclose(_inner(pw[f]), ref, rtol=0, atol=1e-12 * np.abs(ref).max())


@pytest.mark.parametrize("anel", [None, 6000.0])
@pytest.mark.parametrize("f", [0, 1])
def test_constant_momentum_gives_half_the_tvd_tendency(anel, f):
    """u (or v) constant, as field and as its own flow, under random other
    flows: the self-advection term vanishes and the transverse terms are half
    of TVD's advection of the constant field on the dual grid."""
    st = _state(seed=3 + f, anel=anel)
    st.fields[f][:] = [4.5, -1.25][f]
    st.zfields[f][:] = st.fields[f]
    pw = _pw(st, advect_th=False, advect_q=False)
    tvd = _tvd(st, fs=CS)
    ref = 0.5 * _inner(tvd[f])
    np.testing.assert_allclose(_inner(pw[f]), ref, rtol=0, atol=1e-12 * np.abs(ref).max())


def _uniform_grid_state(nz, ny, nx, dz, dxy):
    st = _state(nz, ny, nx, seed=5)
    st.grid = ss.vertical_grid(np.arange(nz) * dz)
    st.cx = st.cy = 1.0 / dxy
    return st


def test_linear_scalars_under_uniform_flow_are_advected_exactly():
    """q = a + bx x + by y + bz zn under uniform (U, V, W), with the halos
    holding the linear extension: the centred scheme gives -(U bx + V by +
    W bz) exactly at levels 2..nz-1 (the top level is one-sided)."""
    nz, ny, nx, dz, dxy = 10, 6, 7, 20.0, 50.0
    st = _uniform_grid_state(nz, ny, nx, dz, dxy)
    U, V, W = 3.0, -1.5, 0.4
    bx, by, bz = 2.0e-3, -1.0e-3, 5.0e-3
    I, J, K = np.meshgrid(np.arange(nx + 4), np.arange(ny + 4), np.arange(nz), indexing="ij")
    zn = st.grid.zn
    q = 1.0 + bx * I * dxy + by * J * dxy + bz * zn[K]
    st.fields[0][:], st.fields[1][:], st.fields[2][:] = U, V, W
    for f in (3, 4, 5):
        st.fields[f][:] = q
    pw = _pw(st, advect_flow=False)
    expect = -(U * bx + V * by + W * bz)
    for f in (3, 4, 5):
        np.testing.assert_allclose(_inner(pw[f], top=False), expect, rtol=1e-10)


def test_linear_momentum_under_uniform_flow_is_advected_exactly():
    """u linear in y under uniform v: su = -V du/dy; v linear in x under
    uniform u: sv = -U dv/dx; w linear in x under uniform u: sw = -U dw/dx
    (w = 0 in the first two cases; rho = 1, uniform dz)."""
    nz, ny, nx, dz, dxy = 9, 6, 7, 20.0, 50.0
    I, J, K = np.meshgrid(np.arange(nx + 4), np.arange(ny + 4), np.arange(nz), indexing="ij")
    a = 1.0e-3
    st = _uniform_grid_state(nz, ny, nx, dz, dxy)
    st.fields[0][:] = 2.0 + a * J * dxy
    st.fields[1][:] = -1.5
    st.fields[2][:] = 0.0
    pw = _pw(st, advect_th=False, advect_q=False)
    np.testing.assert_allclose(_inner(pw[0]), 1.5 * a, rtol=1e-10)
    st.fields[0][:] = 3.0
    st.fields[1][:] = 0.5 + a * I * dxy
    pw = _pw(st, advect_th=False, advect_q=False)
    np.testing.assert_allclose(_inner(pw[1]), -3.0 * a, rtol=1e-10)
    st.fields[1][:] = 0.0
    st.fields[2][:] = 0.2 + a * I * dxy
    pw = _pw(st, advect_th=False, advect_q=False)
    np.testing.assert_allclose(_inner(pw[2], top=False), -3.0 * a, rtol=1e-10)


def test_overwritten_and_accumulated_sources():
    """su, sv, sw and sth are overwritten, sq(n) accumulates; level 1, sw at
    the top and the halo columns are left as they were."""
    st = _state(seed=7, anel=3000.0)
    rng = np.random.default_rng(8)
    src1 = [rng.standard_normal(f.shape) for f in st.fields]
    src2 = [rng.standard_normal(f.shape) for f in st.fields]
    a, b = _pw(st, src1), _pw(st, src2)
    for f in range(6):
        top = f != 2
        if f < 4:
            np.testing.assert_array_equal(_inner(a[f], top), _inner(b[f], top))
        else:
            np.testing.assert_allclose(_inner(a[f]) - _inner(src1[f]), _inner(b[f]) - _inner(src2[f]),
                                       rtol=0, atol=1e-9)
        np.testing.assert_array_equal(a[f][:, :, 0], src1[f][:, :, 0])
        for sl in (np.s_[:2], np.s_[-2:], np.s_[:, :2], np.s_[:, -2:]):
            np.testing.assert_array_equal(a[f][sl], src1[f][sl])
    np.testing.assert_array_equal(a[2][:, :, -1], src1[2][:, :, -1])


# ---------------------------------------------------------------------------
# The HIP kernel through the C ABI

# (nz, ny, nx, seed, anelastic scale or None, n_q)
# nx = 37 and 32 make the kernel march three strips of 16 columns with a
# partial last strip, and two exact strips (seeding of later strips)
GRIDS = [(16, 12, 12, 0, None, 1), (33, 10, 6, 1, 2000.0, 2), (70, 8, 9, 2, None, 1), (130, 5, 7, 3, 1500.0, 3),
         (24, 6, 37, 4, 3000.0, 1), (20, 7, 32, 5, None, 2)]


def _gpu_case(case, flags, precision=None):
    from monc_amd import tvd
    nz, ny, nx, seed, an, n_q = case
    st = ss.make_state(nz, ny, nx, 50.0, 25.0 * (nz - 1), seed=seed, n_q=n_q, anelastic_scale=an)
    rng = np.random.default_rng(seed + 200)
    src = [rng.standard_normal(f.shape) * 1e-2 for f in st.fields]
    tcx, tcy = _tc(st)
    ref = oracle.pw(st.fields, src, st.grid, st.cx, st.cy, tcx, tcy, *flags)
    t = tvd.TvdAdvection(nz, ny, nx, st.cx, st.cy, st.grid.metric_vectors(), n_q=n_q,
                         precision=tvd.FP64 if precision is None else precision)
    try:
        for n in range(len(st.fields)):
            t.set_field(n, tvd.PROG, st.fields[n])
            t.set_field(n, tvd.SOURCE, src[n])
        t.pw_tendencies(*flags, tcx=tcx, tcy=tcy)
        t.synchronize()
        return [t.get_field(n, tvd.SOURCE) for n in range(len(st.fields))], ref, src
    finally:
        t.close()


@pytest.mark.gpu
@pytest.mark.parametrize("case", GRIDS, ids=[f"{c[2]}x{c[1]}x{c[0]}" for c in GRIDS])
@pytest.mark.parametrize("flags", [(True, True, True), (True, False, False), (False, True, True)],
                         ids=["all", "flow", "scalars"])
def test_pw_matches_oracle(gpu, case, flags):
    got, ref, before = _gpu_case(case, flags)
    for n in range(len(got)):
        inc, rinc = got[n] - before[n], ref[n] - before[n]
        scale = max(float(np.max(np.abs(ref[n]))), 1e-300)
        err = float(np.max(np.abs(got[n] - ref[n]))) / scale
        assert err <= 1e-12, (n, err)
        assert np.array_equal(got[n][:, :, 0], before[n][:, :, 0]), n


@pytest.mark.gpu
@pytest.mark.parametrize("case", [GRIDS[1], GRIDS[4]], ids=["6x10x33", "37x6x24"])
def test_pw_fp32_vs_fp64(gpu, case):
    from monc_amd import tvd
    got, ref, before = _gpu_case(case, (True, True, True), precision=tvd.FP32)
    for n in range(len(got)):
        l2 = float(np.linalg.norm(got[n] - ref[n]) / np.linalg.norm(ref[n]))
        assert l2 <= 1e-5, (n, l2)


@pytest.mark.gpu
def test_pw_skips_tracers_and_inactive_theta(gpu):
    """pw_advection's q loop covers the q fields only (pwadvection.F90:333)
    and th only when active (:403): tracers and an inactive th keep their
    sources."""
    from monc_amd import tvd
    st = ss.make_state(20, 8, 10, 50.0, 475.0, seed=9, n_q=1, n_tracers=1)
    rng = np.random.default_rng(9)
    src = [rng.standard_normal(f.shape) * 1e-2 for f in st.fields]
    tcx, tcy = _tc(st)
    ref = oracle.pw(st.fields[:5], src[:5], st.grid, st.cx, st.cy, tcx, tcy, True, False, True)
    t = tvd.TvdAdvection.initialise(st)
    try:
        t.set_field_active(tvd.TH, False)
        for n in range(len(st.fields)):
            t.set_field(n, tvd.SOURCE, src[n])
        t.pw_tendencies(True, True, True, tcx=tcx, tcy=tcy)
        got = [t.get_field(n, tvd.SOURCE) for n in range(len(st.fields))]
    finally:
        t.close()
    np.testing.assert_array_equal(got[3], src[3])
    np.testing.assert_array_equal(got[5], src[5])
    for n in (0, 1, 2, 4):
        err = float(np.max(np.abs(got[n] - ref[n]))) / float(np.max(np.abs(ref[n])))
        assert err <= 1e-12, (n, err)
