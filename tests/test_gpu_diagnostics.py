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
