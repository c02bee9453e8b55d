"""The no-allocation entry point monc_tvd_weighted_sum_noalloc against
monc_tvd_weighted_sum, through the C ABI: identical results on two calls in a
row (the second on a smaller domain, so it reads the scratch the first one
left), no library allocation during the calls, and a clear refusal beyond
the scratch's level limit."""
import numpy as np
import pytest

from monc_amd import setup_state as ss
from monc_amd import tvd

pytestmark = pytest.mark.gpu


def _ctx(nz, ny, nx, seed, precision):
    st = ss.make_state(nz, ny, nx, 50.0, 25.0 * (nz - 1), seed=seed, n_q=1, with_zfields=False)
    t = tvd.TvdAdvection.initialise(st, precision=precision)
    w = st.grid.rhon * st.grid.dz
    w[0] = 0.0
    return t, w


@pytest.mark.parametrize("precision", [tvd.FP64, tvd.FP32])
def test_noalloc_sum_matches_the_allocating_entry_point(gpu, precision):
    lib = tvd.load_library()
    cases = [(256, 96, 80, 61), (40, 10, 12, 62)]  # the second call is smaller
    for nz, ny, nx, seed in cases:
        t, w = _ctx(nz, ny, nx, seed, precision)
        try:
            ref = [t.weighted_sum(f, tvd.PROG, w) for f in (tvd.TH, tvd.Q(0))]
            lib.monc_tvd_test_fail_allocations(1 << 20)  # any library allocation now fails
            try:
                got = [t.weighted_sum_noalloc(f, tvd.PROG, w) for f in (tvd.TH, tvd.Q(0))]
            finally:
                assert lib.monc_tvd_test_fail_allocations(0) == 1 << 20  # none was attempted
            assert got == ref, (nz, got, ref)
            assert abs(ref[0]) > 0
        finally:
            t.close()


def test_noalloc_sum_refuses_columns_beyond_its_scratch(gpu):
    nz = 4097
    st = ss.make_state(nz, 4, 4, 50.0, 25.0 * (nz - 1), seed=63, n_q=1, with_zfields=False)
    t = tvd.TvdAdvection.initialise(st)
    try:
        with pytest.raises(tvd.TvdError) as e:
            t.weighted_sum_noalloc(tvd.TH, tvd.PROG, np.ones(nz))
        assert e.value.code == tvd.ERR_ARG and "4096" in str(e.value)
    finally:
        t.close()
