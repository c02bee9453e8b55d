"""Raw-file exchange with the Fortran drop-in driver (monc_amd/fortran)."""
import numpy as np


def write_input(path, state, fs, ms, ss, dtm, precision=0, advect=(1, 1, 1), th_active=1, q_active=None):
    """Header: local_grid%size in MONC's (Z_INDEX, Y_INDEX, X_INDEX) order,
    number_q_fields, n_tracers, the stepping flags, precision, the three
    advection selectors, th%active, then q(n)%active per q field."""
    nq = state.n_q
    qa = [1] * nq if q_active is None else [int(bool(a)) for a in q_active]
    with open(path, "wb") as f:
        np.array([state.nz, state.ny, state.nx, nq, state.n_tracers, fs, ms, ss, precision, *advect,
                  int(bool(th_active))], dtype=np.int32).tofile(f)
        np.array(qa, dtype=np.int32).tofile(f)
        np.array([state.cx, state.cy, dtm], dtype=np.float64).tofile(f)
        for k in ("rdz", "rdzn", "dz", "dzn", "tzc1", "tzc2", "tzd1", "tzd2"):
            np.asarray(getattr(state.grid, k), dtype=np.float64).tofile(f)
        for a in state.fields:
            np.ascontiguousarray(a, dtype=np.float64).tofile(f)
        for a in (state.zfields or state.fields):
            np.ascontiguousarray(a, dtype=np.float64).tofile(f)


def read_output(path, state):
    """The source fields, then per non-tracer field (u, v, w, th, q(:)) the
    tend_3d diagnostic as an (nx, ny, nz) array and its profile total."""
    data = np.fromfile(path, dtype=np.float64)
    shape = state.shape()
    n = int(np.prod(shape))
    nf = len(state.fields)
    sources = [data[i * n:(i + 1) * n].reshape(shape) for i in range(nf)]
    m = state.nx * state.ny * state.nz
    off = nf * n
    diags = []
    for _ in range(4 + state.n_q):
        t3d = data[off:off + m].reshape(state.nx, state.ny, state.nz)
        prof = data[off + m:off + m + state.nz]
        diags.append((t3d, prof))
        off += m + state.nz
    assert off == data.size
    return sources, diags
