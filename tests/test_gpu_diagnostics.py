"""Tendency diagnostics of the component through the C ABI
(tvdadvection.F90:1064-1188 and the tvd_dgs_terms copies of :615-717):
tend_3d = s_after - s_before per advected field, its per-level total over the
interior columns, the tabs tendency (s*rprefrcp after minus before) and the
per-column flux_z, against the oracle.

Bars: tend_3d and flux_z to 1e-12 relative to the field's largest magnitude
(the same arithmetic as the tendencies); profile totals to 1e-12 relative to
the per-level sum of |tend| (the device sums in another order)."""
import numpy as np
import pytest

import oracle
from monc_amd import setup_state as ss
from monc_amd import tvd

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _run(st, dtm, fs, ms, sc, sources, flags, rprefrcp=None, tiling=None):
    t = tvd.TvdAdvection.initialise(st, precision=tvd.FP64)
    try:
        if tiling is not None:
            t.set_tiling(*tiling)
        for f, s in enumerate(sources):
            t.set_field(f, tvd.SOURCE, s)
        t.set_diagnostics(flags, rprefrcp)
        t.halo_fill()
        t.tendencies(dtm, fs, ms, sc)
        n = len(st.fields)
        out = {"s": [t.get_field(f, tvd.SOURCE) for f in range(n)]}
        if flags & tvd.DIAG_TENDENCY:
            out["t3d"] = [t.get_diagnostic(f, tvd.DIAG_3D) for f in range(n)]
            out["prof"] = [t.get_diagnostic(f, tvd.DIAG_PROFILE) for f in range(n)]
            if rprefrcp is not None:
                out["tabs"] = t.get_diagnostic(0, tvd.DIAG_TABS_3D)
                out["tabs_p"] = t.get_diagnostic(0, tvd.DIAG_TABS_PROFILE)
        if flags & tvd.DIAG_ZFACE:
            out["zf"] = [t.get_diagnostic(f, tvd.DIAG_ZFACES) for f in range(n)]
        return out
    finally:
        t.close()


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _prof_ok(got, ref, t3d):
    scale = np.abs(t3d).sum(axis=(0, 1))
    assert np.all(np.abs(got - ref) <= TOL * np.maximum(scale, 1e-300)), (got, ref)


@pytest.mark.parametrize("fs", [1, 0])
@pytest.mark.parametrize("tiling", [None, (tvd.KERNEL_CELL, 0, 0)])
def test_tendency_diagnostics(gpu, fs, tiling):
    st = ss.make_state(24, 12, 14, 50.0, 700.0, seed=21, n_q=2, anelastic_scale=8000.0)
    rng = np.random.default_rng(3)
    src = [ss.halo_fill(0.01 * rng.standard_normal(f.shape)) for f in st.fields]
    r = 1.0 + 0.05 * rng.random(st.nz)
    got = _run(st, 0.8, fs, fs, fs, src, tvd.DIAG_TENDENCY, r, tiling)
    after, t3d, prof, tabs, tabs_p = oracle.tendency_diagnostics(st, 0.8, fs, fs, fs, src, r)
    for n in range(len(st.fields)):
        assert _rel(got["t3d"][n], t3d[n]) <= TOL, n
        _prof_ok(got["prof"][n], prof[n], t3d[n])
    assert _rel(got["tabs"], tabs) <= TOL
    _prof_ok(got["tabs_p"], tabs_p, tabs)


@pytest.mark.parametrize("tiling", [None, (tvd.KERNEL_MARCH, 4, 5), (tvd.KERNEL_CELL, 0, 0)])
@pytest.mark.parametrize("nz", [16, 70])
def test_zface_diagnostics(gpu, tiling, nz):
    """flux_z(k) per interior column for every field, 0 at k = 1; w carries
    its own bottom face at k = 2 and the top face at k = nz."""
    st = ss.make_state(nz, 10, 9, 50.0, 30.0 * nz, seed=nz, n_q=1)
    src = [np.zeros(f.shape) for f in st.fields]
    got = _run(st, 1.0, 1, 1, 1, src, tvd.DIAG_ZFACE, None, tiling)
    for n in range(len(st.fields)):
        _, faces = oracle.state_tendencies(st, 1.0, 1, 1, 1, rec_field=n)
        ref = faces[2][2:-2, 2:-2, :]
        mine = got["zf"][n][2:-2, 2:-2, :]
        assert _rel(mine, ref) <= TOL, n
        assert np.all(mine[:, :, 0] == 0.0)
        assert np.all(got["zf"][n][:2] == 0.0) and np.all(got["zf"][n][:, :2] == 0.0)  # halo columns untouched


def test_diagnostics_need_recording(gpu):
    st = ss.make_state(8, 6, 6, 50.0, 200.0, seed=1)
    t = tvd.TvdAdvection.initialise(st, precision=tvd.FP64)
    try:
        with pytest.raises(tvd.TvdError):
            t.get_diagnostic(3, tvd.DIAG_3D)
        with pytest.raises(tvd.TvdError):
            t.set_diagnostics(8)
        with pytest.raises(tvd.TvdError):  # no kernel 3
            t.set_tiling(3, 0, 0)
        with pytest.raises(tvd.TvdError):  # no 5-row stage shape
            t.set_tiling(tvd.KERNEL_STAGE, 5, 0)
    finally:
        t.close()


# ---------------------------------------------------------------------------
# The sibling components' tendency diagnostics on the same context
# (pwadvection.F90:619-743, thadvection.F90:228-270): tend_3d = s_after -
# s_before of each field the component changes, over the interior columns,
# its per-level total and the tabs tendency, recorded per component.

def _diag_ref(after, before, fields, r):
    t3d = {f: after[f][2:-2, 2:-2, :] - before[f][2:-2, 2:-2, :] for f in fields}
    prof = {f: oracle.profile_total(t3d[f]) for f in fields}
    tabs = None
    if tvd.TH in fields:
        tabs = after[tvd.TH][2:-2, 2:-2, :] * r - before[tvd.TH][2:-2, 2:-2, :] * r
    return t3d, prof, tabs


def _state_error(fn):
    with pytest.raises(tvd.TvdError) as e:
        fn()
    assert e.value.code == tvd.ERR_STATE


@pytest.mark.parametrize("flags", [(True, True, True), (True, False, False), (False, True, True)],
                         ids=["all", "flow", "scalars"])
def test_pw_tendency_diagnostics(gpu, flags):
    st = ss.make_state(24, 12, 37, 50.0, 575.0, seed=31, n_q=2, anelastic_scale=5000.0)
    rng = np.random.default_rng(31)
    src = [ss.halo_fill(0.01 * rng.standard_normal(f.shape)) for f in st.fields]
    r = 1.0 + 0.05 * rng.random(st.nz)
    tcx, tcy = 0.25 * st.cx, 0.25 * st.cy
    ref = oracle.pw(st.fields, src, st.grid, st.cx, st.cy, tcx, tcy, *flags)
    fields = ([0, 1, 2] if flags[0] else []) + ([3] if flags[1] else []) + ([4, 5] if flags[2] else [])
    t3d, prof, tabs = _diag_ref(ref, src, fields, r)
    t = tvd.TvdAdvection.initialise(st, precision=tvd.FP64)
    try:
        for f, s in enumerate(src):
            t.set_field(f, tvd.SOURCE, s)
        t.set_diagnostics(tvd.DIAG_TENDENCY, r)
        t.pw_tendencies(*flags, tcx=tcx, tcy=tcy)
        for f in range(len(st.fields)):
            if f in fields:
                got = t.get_diagnostic(f, tvd.DIAG_3D, tvd.COMPONENT_PW)
                assert _rel(got, t3d[f]) <= TOL, f
                _prof_ok(t.get_diagnostic(f, tvd.DIAG_PROFILE, tvd.COMPONENT_PW), prof[f], t3d[f])
            else:
                _state_error(lambda: t.get_diagnostic(f, tvd.DIAG_3D, tvd.COMPONENT_PW))
        if tabs is not None:
            got = t.get_diagnostic(0, tvd.DIAG_TABS_3D, tvd.COMPONENT_PW)
            assert _rel(got, tabs) <= TOL
            _prof_ok(t.get_diagnostic(0, tvd.DIAG_TABS_PROFILE, tvd.COMPONENT_PW), oracle.profile_total(tabs), tabs)
        else:
            _state_error(lambda: t.get_diagnostic(0, tvd.DIAG_TABS_3D, tvd.COMPONENT_PW))
        # each component keeps its own record: no TVD call was made
        _state_error(lambda: t.get_diagnostic(0, tvd.DIAG_3D))
        with pytest.raises(tvd.TvdError) as e:
            t.get_diagnostic(0, tvd.DIAG_ZFACES, tvd.COMPONENT_PW)
        assert e.value.code == tvd.ERR_ARG
    finally:
        t.close()


def test_component_records_are_separate(gpu):
    """pw_advection, then TVD, then th_advection on one context: each
    component's tend_3d is its own increment of the sources."""
    st = ss.make_state(20, 9, 11, 50.0, 475.0, seed=33, n_q=1, anelastic_scale=4000.0)
    rng = np.random.default_rng(33)
    s0 = [ss.halo_fill(0.01 * rng.standard_normal(f.shape)) for f in st.fields]
    tcx, tcy = 0.25 * st.cx, 0.25 * st.cy
    s1 = oracle.pw(st.fields, s0, st.grid, st.cx, st.cy, tcx, tcy, True, True, True)
    s2, _ = oracle.state_tendencies(st, 0.7, 1, 1, 1, s1)
    thref = 290.0 + 0.004 * st.grid.zn
    dthref = np.zeros(st.nz)
    dthref[:-1] = thref[1:] - thref[:-1]
    kw = dict(dthref=dthref, thref=thref, anelastic=True, baroclinicity=True, thref0=300.0, fcoriol=1.0e-4,
              g=9.81, rate_x=2.0e-3, rate_y=-1.0e-3, ugal=0.4, vgal=-0.2)
    s3 = oracle.thadv(s2[3], st.fields[0], st.fields[1], st.fields[2], st.grid.tzc1, st.grid.tzc2, **kw)
    t = tvd.TvdAdvection.initialise(st, precision=tvd.FP64)
    try:
        for f, s in enumerate(s0):
            t.set_field(f, tvd.SOURCE, s)
        t.set_diagnostics(tvd.DIAG_TENDENCY)
        t.halo_fill()
        t.pw_tendencies(True, True, True, tcx=tcx, tcy=tcy)
        t.tendencies(0.7, 1, 1, 1)
        t.thadv_tendencies(**kw)
        for f in range(len(st.fields)):
            want = {tvd.COMPONENT_PW: s1[f] - s0[f], tvd.COMPONENT_TVD: s2[f] - s1[f]}
            for comp, d in want.items():
                assert _rel(t.get_diagnostic(f, tvd.DIAG_3D, comp), d[2:-2, 2:-2, :]) <= TOL, (f, comp)
        got = t.get_diagnostic(tvd.TH, tvd.DIAG_3D, tvd.COMPONENT_THADV)
        assert _rel(got, (s3 - s2[3])[2:-2, 2:-2, :]) <= TOL
        _state_error(lambda: t.get_diagnostic(tvd.U, tvd.DIAG_3D, tvd.COMPONENT_THADV))
    finally:
        t.close()


@pytest.mark.parametrize("terms", [(True, False), (False, True), (False, False)],
                         ids=["refstate", "baroclinicity", "none"])
def test_thadv_tendency_diagnostics(gpu, terms):
    """th's increment and tabs over the interior columns (the component
    updates halo columns too; its diagnostics skip them), zero when no term
    is switched on, and nothing while th is inactive."""
    st = ss.make_state(30, 8, 10, 50.0, 725.0, seed=35, n_q=1, anelastic_scale=3000.0)
    rng = np.random.default_rng(35)
    sth = ss.halo_fill(1e-3 * rng.standard_normal(st.fields[3].shape))
    r = 1.0 + 0.05 * rng.random(st.nz)
    thref = 290.0 + 0.004 * st.grid.zn
    dthref = np.zeros(st.nz)
    dthref[:-1] = thref[1:] - thref[:-1]
    kw = dict(dthref=dthref, thref=thref, anelastic=terms[0], baroclinicity=terms[1], thref0=300.0,
              fcoriol=1.1e-4, g=9.81, rate_x=2.5e-3, rate_y=-1.2e-3, ugal=1.3, vgal=-0.6)
    ref = oracle.thadv(sth, st.fields[0], st.fields[1], st.fields[2], st.grid.tzc1, st.grid.tzc2, **kw)
    t3d = (ref - sth)[2:-2, 2:-2, :]
    tabs = ref[2:-2, 2:-2, :] * r - sth[2:-2, 2:-2, :] * r
    for active in (True, False):
        t = tvd.TvdAdvection(st.nz, st.ny, st.nx, st.cx, st.cy, st.grid.metric_vectors(), n_q=1)
        try:
            for n in range(3):
                t.set_field(n, tvd.PROG, st.fields[n])
            t.set_field(tvd.TH, tvd.SOURCE, sth)
            t.set_field_active(tvd.TH, active)
            t.set_diagnostics(tvd.DIAG_TENDENCY, r)
            t.thadv_tendencies(**kw)
            if not active:
                _state_error(lambda: t.get_diagnostic(tvd.TH, tvd.DIAG_3D, tvd.COMPONENT_THADV))
                continue
            got = t.get_diagnostic(tvd.TH, tvd.DIAG_3D, tvd.COMPONENT_THADV)
            gtabs = t.get_diagnostic(0, tvd.DIAG_TABS_3D, tvd.COMPONENT_THADV)
            if terms == (False, False):
                assert not np.any(got) and not np.any(gtabs)
            else:
                assert _rel(got, t3d) <= TOL
                assert _rel(gtabs, tabs) <= TOL
            _prof_ok(t.get_diagnostic(tvd.TH, tvd.DIAG_PROFILE, tvd.COMPONENT_THADV), oracle.profile_total(t3d), t3d)
            _prof_ok(t.get_diagnostic(0, tvd.DIAG_TABS_PROFILE, tvd.COMPONENT_THADV), oracle.profile_total(tabs),
                     tabs)
        finally:
            t.close()
