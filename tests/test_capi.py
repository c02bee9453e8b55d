"""The C ABI library loads, exports every entry point include/monc_tvd.h
declares, and rejects invalid arguments before touching a device (CPU only)."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from monc_amd import tvd

HEADER = Path(__file__).resolve().parents[1] / "include" / "monc_tvd.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(monc_(?:tvd|thadv|pw)_[a-z_]+)\s*\(", text)))


def test_header_declares_the_binding_set():
    assert set(declared_functions()) == set(tvd.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = tvd.load_library()
    for name in declared_functions():
        assert hasattr(lib, name), name


def _create(nz=8, ny=6, nx=6, halo=2, n_q=1, flags=7, cx=0.02, cy=0.02, prec=0, null_metric=False, n_tracers=0):
    lib = tvd.load_library()
    v = np.ones(nz)
    dp = v.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    mets = [dp] * 8
    if null_metric:
        mets[3] = ctypes.POINTER(ctypes.c_double)()
    h = ctypes.c_void_p()
    rc = lib.monc_tvd_create(nz, ny, nx, halo, n_q, n_tracers, flags, cx, cy, *mets, prec, ctypes.byref(h))
    return rc, lib.monc_tvd_last_error().decode(), h


@pytest.mark.parametrize("kw,needle", [
    (dict(halo=3), "halo"),
    (dict(halo=1), "halo"),
    (dict(nz=3), "too small"),
    (dict(ny=3), "too small"),
    (dict(prec=7), "precision"),
    (dict(flags=16), "advect flags"),
    (dict(n_tracers=-1), "n_tracers"),
    (dict(n_tracers=65), "n_tracers"),
    (dict(cx=0.0), "positive"),
    (dict(n_q=-1), "n_q"),
    (dict(null_metric=True), "metric vector"),
])
def test_create_rejects_invalid_arguments(kw, needle):
    rc, msg, h = _create(**kw)
    assert rc == tvd.ERR_ARG
    assert needle in msg
    assert not h.value


def test_null_context_is_an_argument_error():
    lib = tvd.load_library()
    assert lib.monc_tvd_halo_fill(None) == tvd.ERR_ARG
    assert "null context" in lib.monc_tvd_last_error().decode()
    assert lib.monc_tvd_tendencies(None, 1.0, 1, 1, 1) == tvd.ERR_ARG
    assert lib.monc_thadv_tendencies(None, 1, None, 0, None, 300.0, 1e-4, 9.81, 0.0, 0.0, 0.0, 0.0) == tvd.ERR_ARG
    assert "null context" in lib.monc_tvd_last_error().decode()
    assert lib.monc_pw_tendencies(None, 1, 1, 1, 0.005, 0.005) == tvd.ERR_ARG
    assert "null context" in lib.monc_tvd_last_error().decode()
    out = np.zeros(4)
    assert lib.monc_tvd_get_component_diagnostic(None, tvd.COMPONENT_PW, 0, tvd.DIAG_PROFILE,
                                                 out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == tvd.ERR_ARG
    assert "null context" in lib.monc_tvd_last_error().decode()
    assert lib.monc_tvd_destroy(None) == tvd.OK


def _ens_create(sizes, prec=1, cx=0.02):
    lib = tvd.load_library()
    n = len(sizes)
    nz = (ctypes.c_int * n)(*[s[0] for s in sizes])
    ny = (ctypes.c_int * n)(*[s[1] for s in sizes])
    nx = (ctypes.c_int * n)(*[s[2] for s in sizes])
    c = np.full(max(n, 1), cx)
    met = np.ones(8 * max(sum(s[0] for s in sizes), 1))
    dp = ctypes.POINTER(ctypes.c_double)
    h = ctypes.c_void_p()
    rc = lib.monc_tvd_ensemble_create(n, nz, ny, nx, c.ctypes.data_as(dp), c.ctypes.data_as(dp),
                                      met.ctypes.data_as(dp), prec, ctypes.byref(h))
    return rc, lib.monc_tvd_last_error().decode(), h


@pytest.mark.parametrize("kw,needle", [
    (dict(sizes=[]), "n_members"),
    (dict(sizes=[(16, 8, 8), (3, 8, 8)]), "too small"),
    (dict(sizes=[(300, 8, 8)]), "above the ensemble"),
    (dict(sizes=[(16, 8, 8)], prec=5), "precision"),
    (dict(sizes=[(16, 8, 8)], cx=-1.0), "positive"),
])
def test_ensemble_create_rejects_invalid_arguments(kw, needle):
    rc, msg, h = _ens_create(**kw)
    assert rc == tvd.ERR_ARG
    assert needle in msg
    assert not h.value


def test_null_ensemble_is_an_argument_error():
    lib = tvd.load_library()
    assert lib.monc_tvd_ensemble_forward_steps(None, 2) == tvd.ERR_ARG
    assert "null ensemble" in lib.monc_tvd_last_error().decode()
    assert lib.monc_tvd_ensemble_destroy(None) == tvd.OK


def test_ensemble_workload_sizes():
    from monc_amd import setup_state as ss
    sts = ss.ensemble_states(6, seed=0, nxy=(8, 12), nzs=(16, 20), ztop=400.0)
    assert len(sts) == 6
    for st in sts:
        assert st.nx in (8, 12) and st.ny in (8, 12) and st.nz in (16, 20)
        assert len(st.fields) == 5


@pytest.mark.parametrize("kw", [dict(cx=float("inf")), dict(cy=float("nan"))])
def test_create_rejects_non_finite_spacings(kw):
    rc, msg, h = _create(**kw)
    assert rc == tvd.ERR_ARG and "finite" in msg and not h.value


def test_failed_allocation_is_a_hip_error():
    """The allocation hook makes create's first device allocation (the metric
    upload) fail before reaching the runtime: the call returns
    MONC_TVD_ERR_HIP with the runtime's message and no context, and the hook
    is consumed."""
    lib = tvd.load_library()
    assert lib.monc_tvd_test_fail_allocations(1) == 0
    rc, msg, h = _create()
    assert rc == tvd.ERR_HIP and "out of memory" in msg.lower() and not h.value
    assert lib.monc_tvd_test_fail_allocations(0) == 0


def test_forward_step_into_checks_its_tensors():
    """The Python layer refuses buffers the kernels would overrun or alias,
    before any call reaches the library (CPU tensors are refused first)."""
    import torch
    t = tvd.TvdAdvection.__new__(tvd.TvdAdvection)
    t.nz, t.ny, t.nx, t.n_q, t.n_tracers, t.precision = 8, 6, 6, 1, 0, tvd.FP64
    good = [torch.zeros(t.field_shape(), dtype=torch.float64) for _ in range(5)]
    with pytest.raises(ValueError, match="one source"):
        t.forward_step_into(1.0, good[:4], good)
    with pytest.raises(ValueError, match="CUDA tensor"):
        t.forward_step_into(1.0, good, [g.clone() for g in good])
