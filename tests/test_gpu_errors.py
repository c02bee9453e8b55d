"""Error paths on the GPU: a non-finite value in a field stays within its
stencil and leaves every other cell's result as the oracle's, a failed device
allocation surfaces as MONC_TVD_ERR_HIP and leaves the context usable, and
non-finite scalars are refused."""
import numpy as np
import pytest

import oracle
from monc_amd import setup_state as ss
from monc_amd import tvd
from tests.helpers import gpu_tendencies

pytestmark = pytest.mark.gpu


def _footprint(shape, nx, ny, i, j, k, r=3):
    """Interior cells (0-based interior indices) within r cells of array
    point (i, j, k), with the periodic wrap of the interior."""
    m = np.zeros(shape, dtype=bool)
    for di in range(-r, r + 1):
        for dj in range(-r, r + 1):
            ii, jj = (i - 2 + di) % nx, (j - 2 + dj) % ny
            m[ii, jj, max(k - r, 0):k + r + 1] = True
    return m


def test_non_finite_values_stay_in_their_stencil(gpu):
    """A NaN in th and an Inf in q: every cell farther than the stencil
    radius (3) from the bad value matches the oracle at 1e-12, the bad
    value's own cell is non-finite in both, and the GPU's non-finite cells lie
    inside the footprint.  Inside it the reference spreads non-finite values
    further than the GPU does: its upwind weights multiply the unselected
    neighbour by 0 (0 * NaN = NaN, ultimateflux.F90:616-676), where the
    kernels select the operand."""
    st = ss.make_state(20, 9, 10, 50.0, 500.0, seed=51, n_q=1)
    bad = {3: (5, 6, 7, np.nan), 4: (3, 3, 12, np.inf)}
    for f, (i, j, k, v) in bad.items():
        st.fields[f][i, j, k] = v
        ss.halo_fill(st.fields[f])
    ref, _ = oracle.state_tendencies(st, 1.0, 1, 1, 1)
    got = gpu_tendencies(st, 1.0, 1, 1, 1)
    for n, (a, b) in enumerate(zip(got, ref)):
        ai, bi = a[2:-2, 2:-2], b[2:-2, 2:-2]
        if n in bad:
            i, j, k, _ = bad[n]
            fp = _footprint(ai.shape, st.nx, st.ny, i, j, k)
            assert not np.isfinite(ai[i - 2, j - 2, k]) and not np.isfinite(bi[i - 2, j - 2, k]), n
            assert np.all(np.isfinite(ai[~fp])), n
        else:
            fp = np.zeros(ai.shape, dtype=bool)
            assert np.all(np.isfinite(ai)), n  # th's and q's bad values stay in th and q
        scale = float(np.max(np.abs(bi[~fp])))
        assert float(np.max(np.abs(ai[~fp] - bi[~fp]))) <= 1e-12 * scale, n


def test_allocation_failure_returns_hip_error_and_context_survives(gpu):
    st = ss.make_state(12, 6, 6, 50.0, 275.0, seed=52, n_q=1)
    lib = tvd.load_library()
    t = tvd.TvdAdvection.initialise(st)
    try:
        lib.monc_tvd_test_fail_allocations(1)
        with pytest.raises(tvd.TvdError) as e:  # the first source field is allocated here
            t.set_field(tvd.U, tvd.SOURCE, np.zeros(t.field_shape()))
        assert e.value.code == tvd.ERR_HIP and "out of memory" in str(e.value).lower()
        assert lib.monc_tvd_test_fail_allocations(0) == 0
        with pytest.raises(tvd.TvdError) as e:  # never set: a state error, not a crash
            t.tendencies(1.0, 1, 1, 1)
        assert e.value.code == tvd.ERR_STATE
        t.zero_sources(range(len(st.fields)))
        t.halo_fill()
        t.tendencies(1.0, 1, 1, 1)
        got = [t.get_field(n, tvd.SOURCE) for n in range(len(st.fields))]
    finally:
        lib.monc_tvd_test_fail_allocations(0)
        t.close()
    ref, _ = oracle.state_tendencies(st, 1.0, 1, 1, 1)
    for a, b in zip(got, ref):
        assert float(np.max(np.abs(a - b))) <= 1e-12 * float(np.max(np.abs(b)))


@pytest.mark.parametrize("dtm", [float("nan"), float("inf")])
def test_non_finite_timestep_is_refused(gpu, dtm):
    st = ss.make_state(12, 6, 6, 50.0, 275.0, seed=53, n_q=1)
    t = tvd.TvdAdvection.initialise(st)
    try:
        for call in (lambda: t.tendencies(dtm, 1, 1, 1), lambda: t.forward_steps(dtm, 1),
                     lambda: t.leapfrog_steps(dtm, 0.06, 1)):
            with pytest.raises(tvd.TvdError) as e:
                call()
            assert e.value.code == tvd.ERR_ARG and "finite" in str(e.value)
    finally:
        t.close()
