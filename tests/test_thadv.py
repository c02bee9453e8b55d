"""th_advection (thadvection.F90:124-211), the first sibling plugin on the TVD
boundary.  The oracle's restatement is pinned by the physics it discretises
(CPU tests below); the HIP kernel behind monc_thadv_tendencies is checked
against the oracle through the C ABI (GPU tests, 1e-12 in FP64)."""
import numpy as np
import pytest

import oracle
from monc_amd import setup_state as ss


def _dthref(thref):
    """gridmanager.F90:2418-2420: dthref(k) = thref(k+1) - thref(k), 0 at the top."""
    d = np.zeros_like(thref)
    d[:-1] = thref[1:] - thref[:-1]
    return d


def _zeros(nz, ny, nx):
    return np.zeros((nx + 4, ny + 4, nz))


def test_reference_state_term_is_centred_vertical_advection():
    """Uniform dz, rho = 1, uniform w = c: the term is -c times the centred
    difference of thref, so for a quadratic thref(zn) it is exactly
    -c * dthref/dz at zn(k) (k = 2..nz-1); the top level sees the one-sided
    half difference (dthref(nz) = 0); level 1 is untouched; every column,
    halos included, gets it."""
    nz, ny, nx, dz, c = 14, 3, 4, 20.0, 0.7
    g = ss.vertical_grid(np.arange(nz) * dz)
    a, b, q = 300.0, 0.004, 2.0e-6
    thref = a + b * g.zn + q * g.zn ** 2
    w = np.full((nx + 4, ny + 4, nz), c)
    out = oracle.thadv(_zeros(nz, ny, nx), _zeros(nz, ny, nx), _zeros(nz, ny, nx), w, g.tzc1, g.tzc2,
                       dthref=_dthref(thref), anelastic=True)
    expect = -c * (b + 2.0 * q * g.zn)
    for col in ((0, 0), (2, 3), (nx + 3, ny + 3)):
        col_out = out[col[0], col[1], :]
        assert col_out[0] == 0.0
        np.testing.assert_allclose(col_out[1:nz - 1], expect[1:nz - 1], rtol=1e-12, atol=1e-15)
        np.testing.assert_allclose(col_out[nz - 1], -c * (thref[nz - 1] - thref[nz - 2]) / (2.0 * dz), rtol=1e-12)


@pytest.mark.parametrize("m", [2, 7, 18])
def test_each_interface_transport_is_split_equally_by_mass(m):
    """Anelastic stretched grid, w = 1 at one interface level m of one column
    (1-based), zero elsewhere: only levels m and m+1 of that column change,
    and each takes -rho(m) dthref(m) / 2 per unit mass-depth rhon(k) dz(k)."""
    nz, ny, nx = 24, 3, 3
    g = ss.vertical_grid(ss.tanh_stretched_z(nz, 25.0 * (nz - 1)), anelastic_scale=1500.0)
    rng = np.random.default_rng(m)
    thref = 290.0 + np.cumsum(rng.uniform(0.05, 0.3, nz))
    dth = _dthref(thref)
    w = _zeros(nz, ny, nx)
    w[2, 3, m - 1] = 1.0
    out = oracle.thadv(_zeros(nz, ny, nx), _zeros(nz, ny, nx), _zeros(nz, ny, nx), w, g.tzc1, g.tzc2,
                       dthref=dth, anelastic=True)
    changed = np.argwhere(out != 0.0)
    assert {tuple(x) for x in changed} == {(2, 3, m - 1), (2, 3, m)}
    half = -g.rho[m - 1] * dth[m - 1] / 2.0
    for k in (m, m + 1):  # 1-based levels
        np.testing.assert_allclose(g.rhon[k - 1] * g.dz[k - 1] * out[2, 3, k - 1], half, rtol=1e-12)


def test_boussinesq_has_no_reference_state_term():
    nz, ny, nx = 10, 3, 3
    g = ss.vertical_grid(ss.tanh_stretched_z(nz, 25.0 * (nz - 1)))
    rng = np.random.default_rng(1)
    sth = rng.standard_normal((nx + 4, ny + 4, nz))
    w = rng.standard_normal(sth.shape)
    out = oracle.thadv(sth, _zeros(nz, ny, nx), _zeros(nz, ny, nx), w, g.tzc1, g.tzc2,
                       dthref=np.linspace(0.1, 0.3, nz), anelastic=False)
    assert np.array_equal(out, sth)


@pytest.mark.parametrize("anel", [False, True])
def test_baroclinicity_term_is_advection_of_the_thermal_wind_theta(anel):
    """With geostrophic shears dUg/dz = rate_x and dVg/dz = rate_y, thermal
    wind balance gives the mean potential temperature the horizontal
    gradient (thref f / g) * (rate_y, -rate_x); the term must be its
    advection -(u+ugal) dth/dx - (v+vgal) dth/dy, with the gradient taken
    here by finite differences of that linear field (thref0 in the
    Boussinesq form, thref(k) in the anelastic one)."""
    nz, ny, nx, dxy = 9, 5, 6, 50.0
    g = ss.vertical_grid(ss.tanh_stretched_z(nz, 25.0 * (nz - 1)), anelastic_scale=1200.0 if anel else None)
    rng = np.random.default_rng(2)
    thref = 295.0 + 0.003 * g.zn
    thref0, f, grav, rx, ry, ugal, vgal = 300.0, 1.0e-4, 9.81, 2.0e-3, -1.5e-3, 1.7, -0.4
    u = rng.standard_normal((nx + 4, ny + 4, nz)) + 5.0
    v = rng.standard_normal(u.shape)
    out = oracle.thadv(_zeros(nz, ny, nx), u, v, _zeros(nz, ny, nx), g.tzc1, g.tzc2, thref=thref,
                       anelastic=anel, baroclinicity=True, thref0=thref0, fcoriol=f, g=grav, rate_x=rx,
                       rate_y=ry, ugal=ugal, vgal=vgal)
    x = np.arange(3) * dxy
    theta0 = thref if anel else np.full(nz, thref0)
    for k in range(1, nz):
        coef = theta0[k] * f / grav
        tx = coef * ry * x  # theta along x at fixed y, and along y at fixed x
        ty = -coef * rx * x
        dtdx = (tx[2] - tx[0]) / (2 * dxy)
        dtdy = (ty[2] - ty[0]) / (2 * dxy)
        expect = -(u[:, :, k] + ugal) * dtdx - (v[:, :, k] + vgal) * dtdy
        np.testing.assert_allclose(out[:, :, k], expect, rtol=1e-10, atol=1e-18)
    assert np.all(out[:, :, 0] == 0.0)


def test_galilean_pairing():
    nz, ny, nx = 8, 3, 4
    g = ss.vertical_grid(ss.tanh_stretched_z(nz, 25.0 * (nz - 1)))
    rng = np.random.default_rng(3)
    u = rng.standard_normal((nx + 4, ny + 4, nz))
    v = rng.standard_normal(u.shape)
    kw = dict(anelastic=False, baroclinicity=True, thref0=300.0, fcoriol=1e-4, g=9.81, rate_x=1e-3, rate_y=2e-3)
    z = _zeros(nz, ny, nx)
    a = oracle.thadv(z, u, v, z, g.tzc1, g.tzc2, ugal=0.8, vgal=-0.3, **kw)
    b = oracle.thadv(z, u + 0.8, v - 0.3, z, g.tzc1, g.tzc2, **kw)
    np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-18)
    still = oracle.thadv(z, np.full(u.shape, -0.8), np.full(u.shape, 0.3), z, g.tzc1, g.tzc2, ugal=0.8, vgal=-0.3,
                         **kw)
    assert np.all(still == 0.0)


# ---------------------------------------------------------------------------
# The HIP kernel through the C ABI

# (nz, ny, nx, seed, anelastic scale or None)
GRIDS = [(16, 12, 12, 0, 1000.0), (33, 10, 6, 1, None), (70, 8, 9, 2, 2000.0), (130, 5, 7, 3, 1500.0)]


def _gpu_case(case, anelastic, baroclinicity, precision=None):
    from monc_amd import tvd
    nz, ny, nx, seed, an = case
    st = ss.make_state(nz, ny, nx, 50.0, 25.0 * (nz - 1), seed=seed, n_q=1, anelastic_scale=an)
    rng = np.random.default_rng(seed + 100)
    sth = rng.standard_normal(st.fields[3].shape) * 1e-3
    thref = 290.0 + 0.004 * st.grid.zn
    kw = dict(dthref=_dthref(thref), thref=thref, anelastic=anelastic, baroclinicity=baroclinicity,
              thref0=300.0, fcoriol=1.1e-4, g=9.81, rate_x=2.5e-3, rate_y=-1.2e-3, ugal=1.3, vgal=-0.6)
    ref = oracle.thadv(sth, st.fields[0], st.fields[1], st.fields[2], st.grid.tzc1, st.grid.tzc2, **kw)
    t = tvd.TvdAdvection(nz, ny, nx, st.cx, st.cy, st.grid.metric_vectors(), n_q=1,
                         precision=tvd.FP64 if precision is None else precision)
    try:
        for n in range(3):
            t.set_field(n, tvd.PROG, st.fields[n])
        t.set_field(tvd.TH, tvd.SOURCE, sth)
        t.thadv_tendencies(**kw)
        t.synchronize()
        return t.get_field(tvd.TH, tvd.SOURCE), ref, sth
    finally:
        t.close()


@pytest.mark.gpu
@pytest.mark.parametrize("case", GRIDS, ids=[f"{c[2]}x{c[1]}x{c[0]}" for c in GRIDS])
@pytest.mark.parametrize("flags", [(True, False), (False, True), (True, True)],
                         ids=["refstate", "baroclinicity", "both"])
def test_thadv_matches_oracle(gpu, case, flags):
    got, ref, before = _gpu_case(case, *flags)
    inc, rinc = got - before, ref - before
    err = float(np.max(np.abs(inc - rinc)) / max(float(np.max(np.abs(rinc))), 1e-300))
    assert err <= 1e-12, err
    assert np.array_equal(got[:, :, 0], before[:, :, 0])  # level 1 untouched


@pytest.mark.gpu
def test_thadv_fp32_vs_fp64(gpu):
    from monc_amd import tvd
    got, ref, before = _gpu_case(GRIDS[2], True, True, precision=tvd.FP32)
    inc, rinc = got - before, ref - before
    assert float(np.linalg.norm(inc - rinc) / np.linalg.norm(rinc)) <= 1e-5


@pytest.mark.gpu
def test_thadv_argument_errors(gpu):
    from monc_amd import tvd
    st = ss.make_state(12, 6, 6, 50.0, 275.0, seed=0, n_q=1)
    t = tvd.TvdAdvection(st.nz, st.ny, st.nx, st.cx, st.cy, st.grid.metric_vectors(), n_q=1)
    try:
        with pytest.raises(tvd.TvdError, match="th source|never set"):
            t.thadv_tendencies(True, dthref=np.zeros(st.nz))
        t.set_field(tvd.TH, tvd.SOURCE, np.zeros(st.fields[3].shape))
        with pytest.raises(tvd.TvdError, match="dthref"):
            t.thadv_tendencies(True)
        with pytest.raises(ValueError):
            t.thadv_tendencies(True, dthref=np.zeros(st.nz + 1))
    finally:
        t.close()
