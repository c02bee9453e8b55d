"""Parity of the HIP path with the oracle (FP64 drop-in form), through the C ABI.

Bar: max|gpu - oracle| / max|oracle| <= 1e-12 per field over the interior,
for every source field (su, sv, sw, sth, sq).  Full-size (512x512x256) parity
is checked on windows cut from the full domain, plus conservation.
"""
import numpy as np
import pytest

import oracle
from monc_amd import setup_state as ss
from monc_amd import tvd
from tests.helpers import gpu_forward, gpu_tendencies, interior, rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-12
NAMES = ["su", "sv", "sw", "sth", "sq", "sq2"]


def _compare(gpu_s, ref_s, tol=TOL, rows=None):
    errs = {}
    for n, (a, b) in enumerate(zip(gpu_s, ref_s)):
        if rows is not None:
            a, b = a[:, rows, :], b[:, rows, :]
            e = float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
        else:
            e = rel_err(a, b)
        errs[NAMES[n]] = e
    assert all(e <= tol for e in errs.values()), errs
    return errs


@pytest.mark.parametrize("fs,ms,sc", [(1, 1, 1), (0, 0, 0), (1, 0, 0), (0, 1, 0)])
@pytest.mark.parametrize("anelastic", [None, 8000.0])
@pytest.mark.parametrize("tiling", [None, (1, 0, 0), (0, 4, 5), (0, 8, 16), (0, 4, 1)])
def test_tiny_grid_stepping_modes(gpu, fs, ms, sc, anelastic, tiling):
    st = ss.make_state(16, 10, 12, 50.0, 500.0, seed=3, n_q=2, anelastic_scale=anelastic)
    ref, _ = oracle.state_tendencies(st, 1.0, fs, ms, sc)
    got = gpu_tendencies(st, 1.0, fs, ms, sc, tiling=tiling)
    _compare(got, ref)


# every compiled march shape (cell rows per block) for column heights that
# fill 1..4 sub-columns of 64 levels partially and exactly, with strips that
# do and do not divide nx
MARCH_CASES = [(nz, rows) for nz in (4, 5, 63, 64) for rows in (12, 8, 4)]
MARCH_CASES += [(nz, rows) for nz in (65, 70, 127, 128) for rows in (6, 4, 2)]
MARCH_CASES += [(nz, rows) for nz in (129, 150, 191, 192) for rows in (4, 2)]
MARCH_CASES += [(nz, rows) for nz in (193, 200, 256) for rows in (4, 2, 3, 1)]


@pytest.mark.parametrize("nz,rows", MARCH_CASES)
@pytest.mark.parametrize("strip", [5, 16])
def test_march_shapes(gpu, nz, rows, strip):
    st = ss.make_state(nz, 9, 11, 50.0, 25.0 * (nz - 1), seed=nz + rows, n_q=1)
    for fs in (1, 0):
        ref, _ = oracle.state_tendencies(st, 1.0, fs, fs, fs)
        got = gpu_tendencies(st, 1.0, fs, fs, fs, tiling=(tvd.KERNEL_MARCH, rows, strip))
        _compare(got, ref)


@pytest.mark.parametrize("fs", [1, 0])
def test_config_c1_64cubed(gpu, fs):
    """64^3, dx=dy=50 m, z tanh-stretched to 2 km, dt = 1 s, seed 0."""
    st = ss.make_state(64, 64, 64, 50.0, 2000.0, seed=0, n_q=1)
    ms = sc = fs
    ref, _ = oracle.state_tendencies(st, 1.0, fs, ms, sc)
    got = gpu_tendencies(st, 1.0, fs, ms, sc)
    errs = _compare(got, ref)
    print("C1 errors", errs)


def test_accumulates_and_overwrites_top(gpu):
    """s is accumulated below the top and overwritten at k = nz (u, v, scalars)."""
    st = ss.make_state(12, 8, 9, 50.0, 300.0, seed=5, n_q=1)
    rng = np.random.default_rng(11)
    src = [rng.standard_normal(st.shape()) for _ in st.fields]
    ref, _ = oracle.state_tendencies(st, 0.7, 1, 1, 1, sources=src)
    got = gpu_tendencies(st, 0.7, 1, 1, 1, sources=src)
    _compare(got, ref)
    # levels the reference never writes keep their input
    for n in range(len(src)):
        np.testing.assert_array_equal(interior(got[n])[:, :, 0], interior(src[n])[:, :, 0])
    np.testing.assert_array_equal(interior(got[2])[:, :, -1], interior(src[2])[:, :, -1])


def test_exact_zeros_and_ties(gpu):
    """Exact zero and negative-zero flows, flat patches (|curvature| = |delta|)."""
    st = ss.make_state(16, 8, 8, 50.0, 500.0, seed=9, n_q=1)
    rng = np.random.default_rng(2)
    for f in (0, 1, 2):
        a = st.fields[f]
        m = rng.random(a.shape) < 0.15
        a[m] = 0.0
        m = rng.random(a.shape) < 0.1
        a[m] = -0.0
        ss.halo_fill(a)
    th = st.fields[3]
    th[4:7, 3:6, 2:9] = 300.0
    ss.halo_fill(th)
    st.fields[2][:, :, 0] = 0.0
    st.fields[2][:, :, -1] = 0.0
    for fs in (1, 0):
        ref, _ = oracle.state_tendencies(st, 2.0, fs, fs, fs)
        got = gpu_tendencies(st, 2.0, fs, fs, fs)
        _compare(got, ref)


def test_halo_fill_matches_periodic_copy(gpu):
    st = ss.make_state(8, 6, 7, 50.0, 200.0, seed=1, n_q=1)
    rng = np.random.default_rng(4)
    garbage = [f + rng.standard_normal(f.shape) * 5 for f in st.fields]
    for g in garbage:  # interior untouched, halo scrambled
        g[2:-2, 2:-2, :] = rng.standard_normal(g[2:-2, 2:-2, :].shape)
    t = tvd.TvdAdvection(st.nz, st.ny, st.nx, st.cx, st.cy, st.grid.metric_vectors(), n_q=st.n_q)
    try:
        for n, g in enumerate(garbage):
            t.set_field(n, tvd.PROG, g)
            t.set_field(n, tvd.ZFIELD, g * 2)
        t.halo_fill()
        for n, g in enumerate(garbage):
            np.testing.assert_array_equal(t.get_field(n, tvd.PROG), oracle.halo_fill(g.copy()))
            np.testing.assert_array_equal(t.get_field(n, tvd.ZFIELD), oracle.halo_fill(g * 2))
    finally:
        t.close()


@pytest.mark.parametrize("tiling", [None, (1, 0, 0), (0, 8, 7), (0, 4, 3)])
def test_forward_steps_match_oracle_run(gpu, tiling):
    st = ss.make_state(20, 24, 20, 50.0, 600.0, seed=12, n_q=1, with_zfields=False)
    dt = ss.cfl_dt(st, 0.3)
    ref = oracle.forward_run(st, dt, 5)
    got = gpu_forward(st, dt, 5, tiling=tiling)
    for n, (a, b) in enumerate(zip(got, ref)):
        e = float(np.max(np.abs(a - b)) / np.max(np.abs(b)))  # halos included
        assert e <= TOL, (n, e)


def _window(f, i0, j0, wx, wy):
    """Sub-block with real neighbour data as its halo: interior i in [i0, i0+wx)."""
    return np.ascontiguousarray(f[i0 - 2:i0 + wx + 2, j0 - 2:j0 + wy + 2, :])


def test_full_size_windows_and_conservation(gpu):
    """512x512x256 (dx=dy=25 m, z to 5 km), one FORWARD tendency call on the
    GPU; eight 12x12 windows compared with the oracle (its last row excluded:
    there the oracle's y-flux wrap assumes periodicity), and conservation of
    the scalar tendencies over the whole domain."""
    nz, ny, nx = 256, 512, 512
    st = ss.make_state(nz, ny, nx, 25.0, 5000.0, seed=0, n_q=1, with_zfields=False)
    got = gpu_tendencies(st, 1.0, 1, 1, 1)
    rng = np.random.default_rng(0)
    corners = [(2, 2), (nx - 10, ny - 10), (2, ny - 10), (nx - 10, 2)]
    corners += [tuple(int(v) for v in rng.integers(2, 500, 2)) for _ in range(4)]
    wx = wy = 12
    for (i0, j0) in corners:
        fields = [_window(f, i0, j0, wx, wy) for f in st.fields]
        ref, _ = oracle.tendencies(nz, wy, wx, st.cx, st.cy, st.grid.metric_vectors(), 1.0, 1, 1, 1, fields)
        mine = [_window(g, i0, j0, wx, wy) for g in got]
        _compare([m[2:-2] for m in mine], [r[2:-2] for r in ref], rows=slice(2, wy + 1))
    g = st.grid
    for n in (3, 4):  # sum_k rhon dz sum_ji s = 0 up to round-off
        s = interior(got[n])
        tot = float((s.sum(axis=(0, 1))[1:] * g.rhon[1:] * g.dz[1:]).sum())
        mag = float((np.abs(s).sum(axis=(0, 1))[1:] * g.rhon[1:] * g.dz[1:]).sum())
        assert abs(tot) <= 1e-12 * mag, (n, tot, mag)


@pytest.mark.parametrize("precision", [tvd.FP64, tvd.FP32])
def test_autotune_keeps_state_and_parity(gpu, precision):
    """The autotuner times every shape without advancing the state, picks a
    compiled shape per kernel, and the tuned step still matches the oracle
    (FP64 at 1e-12; FP32 to the FP32 bar of the precision tests)."""
    st = ss.make_state(70, 24, 20, 50.0, 1500.0, seed=31, n_q=1, with_zfields=False)
    dt = ss.cfl_dt(st, 0.3)
    t = tvd.TvdAdvection.initialise(st, precision=precision)
    try:
        t.halo_fill()
        before = [t.get_field(n, tvd.PROG) for n in range(5)]
        chosen = t.autotune(dt, reps=1)
        shapes = t.march_shapes()
        assert set(chosen) == {"u", "v", "w", "scalars"}
        for kernel, rows, strip in chosen.values():
            if kernel == tvd.KERNEL_STAGE:
                assert rows in (4, 6, 8, 11, 15) and strip in (32, 64, 128)
            else:
                assert kernel == tvd.KERNEL_MARCH and rows in shapes and strip in (8, 16, 32, 64)
        for n in range(5):
            np.testing.assert_array_equal(t.get_field(n, tvd.PROG), before[n])
        t.forward_steps(dt, 3)
        got = [t.get_field(n, tvd.PROG) for n in range(5)]
    finally:
        t.close()
    ref = oracle.forward_run(st, dt, 3)
    tol = TOL if precision == tvd.FP64 else 2e-6
    for n, (a, b) in enumerate(zip(got, ref)):
        e = float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
        assert e <= tol, (n, e)


# ---------------------------------------------------------------------------
# Field selection at the boundary: tracers advected under the theta selector
# (tvdadvection.F90:352, :594), th and q fields gated by their active flags
# (:706, :661), all at 1e-12 against the oracle.
SELECTION_CASES = [
    # (advect bits for the library (initialise adds tracers to theta), oracle bits, active th, q1, q2)
    (7, 15, 1, 1, 1),    # everything
    (3, 11, 1, 1, 1),    # th = tvd, q = another scheme: th and tracers, no q
    (5, 5, 1, 1, 1),     # q = tvd, th = another scheme: no th, no tracers
    (7, 15, 0, 1, 0),    # th and q(2) inactive
    (6, 14, 1, 0, 1),    # scalars only, q(1) inactive
]


@pytest.mark.parametrize("lib_bits,oracle_bits,th_on,q1_on,q2_on", SELECTION_CASES)
@pytest.mark.parametrize("fs", [1, 0])
def test_tracers_and_active_fields(gpu, lib_bits, oracle_bits, th_on, q1_on, q2_on, fs):
    st = ss.make_state(20, 10, 12, 50.0, 500.0, seed=41, n_q=2, n_tracers=2)
    assert st.n_q == 2 and len(st.fields) == 8
    active = [1, 1, 1, th_on, q1_on, q2_on, 1, 1]
    rng = np.random.default_rng(5)
    src = [rng.standard_normal(st.shape()) for _ in st.fields]  # untouched fields keep these
    ref, _ = oracle.state_tendencies(st, 1.0, fs, fs, fs, sources=src, advect=oracle_bits, active=active)
    got = gpu_tendencies(st, 1.0, fs, fs, fs, sources=src, advect=lib_bits, active=active)
    adv = oracle.advected(8, oracle_bits, 2, active)
    for n, (a, b) in enumerate(zip(got, ref)):
        if n in adv:
            assert rel_err(a, b) <= TOL, (n, rel_err(a, b))
        else:
            np.testing.assert_array_equal(a, src[n])  # not advected: the source field is untouched


def test_tracer_is_advected_as_a_scalar(gpu):
    """A tracer holding the same data as a q field receives the same tendency
    (advect_tracer_fields and advect_q_fields both call advect_scalar_field,
    tvdadvection.F90:661-694).  The q field runs in the (th, q) pair kernel and
    the tracer in the single-field one, whose roundings may differ in the last
    bits, so the bar is the parity tolerance."""
    st = ss.make_state(24, 9, 11, 50.0, 600.0, seed=42, n_q=1, n_tracers=1)
    st.fields[5] = st.fields[4].copy()
    st.zfields[5] = st.zfields[4].copy()
    got = gpu_tendencies(st, 1.0, 1, 1, 1)
    assert rel_err(got[5], got[4]) <= TOL


def test_forward_steps_with_tracers_and_inactive_q(gpu):
    st = ss.make_state(20, 12, 10, 50.0, 600.0, seed=43, n_q=2, n_tracers=1, with_zfields=False)
    dt = ss.cfl_dt(st, 0.3)
    active = [1, 1, 1, 1, 1, 0, 1]
    ref = oracle.forward_run(st, dt, 4, advect=15, active=active)
    t = tvd.TvdAdvection.initialise(st)
    try:
        t.set_field_active(tvd.Q(1), False)
        t.halo_fill()
        t.forward_steps(dt, 4)
        got = [t.get_field(n, tvd.PROG) for n in range(len(st.fields))]
    finally:
        t.close()
    for n, (a, b) in enumerate(zip(got, ref)):
        e = float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
        assert e <= TOL, (n, e)
    np.testing.assert_array_equal(got[5], oracle.halo_fill(st.fields[5].copy()))  # inactive q(2) unchanged


# ---------------------------------------------------------------------------
# The stage kernel (per field kind, fields staged in LDS): column heights that
# fill one to four 64-level bands partially and exactly, row counts that do
# and do not divide ny, strips that do and do not divide nx, one and two
# scalars, against the oracle's forward run.
STAGE_CASES = [(nz, rows) for nz in (4, 5, 63, 64, 65, 70, 128, 150, 193, 256) for rows in (4, 6, 8, 11, 15)]


@pytest.mark.parametrize("nz,rows", STAGE_CASES)
@pytest.mark.parametrize("strip,n_q", [(5, 1), (16, 2)])
def test_stage_kernel_matches_oracle(gpu, nz, rows, strip, n_q):
    """rows = 15 is compiled for u, v and w only: the scalars then take 4 rows."""
    st = ss.make_state(nz, 11, 13, 50.0, 25.0 * (nz - 1), seed=nz + rows + strip, n_q=n_q, with_zfields=False)
    dt = ss.cfl_dt(st, 0.3)
    ref = oracle.forward_run(st, dt, 3)
    got = gpu_forward(st, dt, 3, tiling=(tvd.KERNEL_STAGE, rows, strip))
    for n, (a, b) in enumerate(zip(got, ref)):
        e = float(np.max(np.abs(a - b)) / np.max(np.abs(b)))  # halos included
        assert e <= TOL, (n, e)


STAGE_ACCUM_CASES = [(nz, rows) for nz in (5, 64, 70, 150, 256) for rows in (4, 6, 11, 15)]


@pytest.mark.parametrize("nz,rows", STAGE_ACCUM_CASES)
@pytest.mark.parametrize("fs,ms,sc", [(1, 1, 1), (1, 0, 0), (0, 0, 0)])
def test_stage_kernel_drop_in_call(gpu, nz, rows, fs, ms, sc):
    """The drop-in call (s += tendency, overwritten at k = nz) through the
    stage kernel, into non-zero sources; with centred field stepping the
    momentum fields advect their z-fields, staged as a fourth array beside
    the flows (at 15 rows in 4-row tiles)."""
    st = ss.make_state(nz, 11, 13, 50.0, 25.0 * (nz - 1), seed=nz + 7 * rows + fs, n_q=2,
                       anelastic_scale=6000.0 if rows == 6 else None)
    rng = np.random.default_rng(nz + rows)
    src = [ss.halo_fill(0.01 * rng.standard_normal(f.shape)) for f in st.fields]
    ref, _ = oracle.state_tendencies(st, 1.0, fs, ms, sc, src)
    got = gpu_tendencies(st, 1.0, fs, ms, sc, sources=src, tiling=(tvd.KERNEL_STAGE, rows, 5))
    _compare(got, ref)


def test_default_drop_in_call_stages_at_scale(gpu):
    """512x512x256 FP64: the default drop-in call (u, v, w staged) against the
    march kernel forced for every kind, both into the same sources."""
    st = ss.make_state(256, 512, 512, 25.0, 5000.0, seed=0, n_q=1, with_zfields=False)
    dt = ss.cfl_dt(st, 0.3)
    rng = np.random.default_rng(5)
    src = [ss.halo_fill(0.01 * rng.standard_normal(st.fields[0].shape)) for _ in range(2)]
    src += [np.zeros(st.fields[0].shape)] * 3
    t = tvd.TvdAdvection.initialise(st, precision=tvd.FP64)
    try:
        assert [t.tiling()[k][0] for k in ("u", "v", "w")] == [tvd.KERNEL_STAGE] * 3  # the default
        out = []
        for tiling in (None, (tvd.KERNEL_MARCH, 3, 32)):
            if tiling is not None:
                t.set_tiling(*tiling)
            for f in range(5):
                t.set_field(f, tvd.SOURCE, src[f])
            t.halo_fill()
            t.tendencies(dt, 1, 1, 1)
            out.append([t.get_field(f, tvd.SOURCE) for f in range(5)])
        assert [k[0] for k in list(t.tiling().values())[:3]] == [tvd.KERNEL_MARCH] * 3  # forced
    finally:
        t.close()
    for n in range(5):
        e = rel_err(out[0][n], out[1][n])
        assert e <= TOL, (n, e)


@pytest.mark.parametrize("rows", [4, 6, 8, 11, 15])
def test_stage_kernel_fp32_matches_march(gpu, rows):
    st = ss.make_state(130, 20, 18, 50.0, 2500.0, seed=77, n_q=1, with_zfields=False)
    dt = ss.cfl_dt(st, 0.3)
    a = gpu_forward(st, dt, 4, precision=tvd.FP32, tiling=(tvd.KERNEL_STAGE, rows, 16))
    b = gpu_forward(st, dt, 4, precision=tvd.FP32)
    for n, (x, y) in enumerate(zip(a, b)):
        e = float(np.max(np.abs(x - y)) / np.max(np.abs(y)))
        assert e <= 1e-6, (n, e)
