"""Synthetic MONC states for the TVD advection path: grid, metrics and fields.

This is the harness-side stand-in for what MONC's set-up components produce
before the first timestep:

* the vertical grid and its metric vectors, as
  components/gridmanager/src/gridmanager.F90:2449-2456 (dz, dzn, rdz, rdzn),
  :2492-2494 (zn(1) = -zn(2)), :2048-2067 (tzc1/2, tzd1/2) and :2624-2633
  (cx = 1/real(dx), the spacing rounded to default real first);
* the prognostic fields with the staggering of
  components/simplesetup/src/simplesetup.F90:70-101 (u, v and scalars at zn,
  w at z), filled with the seeded distributions of the benchmark configs;
* the single-rank periodic halo fill of
  model_core/src/communication/halocommunication.F90:1058-1168.

Arrays use the reference's memory layout: Fortran data(k,j,i) of shape
(nz, ny+4, nx+4) is a C-contiguous numpy array of shape (nx+4, ny+4, nz),
indexed [i-1, j-1, k-1].
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

HALO = 2


@dataclass
class VerticalGrid:
    """Vertical coordinate and the FP64 metric vectors TVD reads (length nz)."""
    z: np.ndarray
    zn: np.ndarray
    rho: np.ndarray
    rhon: np.ndarray
    dz: np.ndarray
    dzn: np.ndarray
    rdz: np.ndarray
    rdzn: np.ndarray
    tzc1: np.ndarray
    tzc2: np.ndarray
    tzd1: np.ndarray
    tzd2: np.ndarray

    @property
    def nz(self) -> int:
        return int(self.z.shape[0])

    def metric_vectors(self) -> Dict[str, np.ndarray]:
        return {n: getattr(self, n) for n in ("rdz", "rdzn", "dz", "dzn", "tzc1", "tzc2", "tzd1", "tzd2")}


def tanh_stretched_z(nz: int, ztop: float, dzmin: float = 10.0, dzmax: float = 40.0) -> np.ndarray:
    """Level heights z(1..nz): z(1)=0, dz(k)=dzmin+(dzmax-dzmin)*(1+tanh((k-kc)/kw))/2
    for k=2..nz with kw=nz/8 and kc chosen so that z(nz)=ztop (bisection)."""
    if not (dzmin * (nz - 1) < ztop < dzmax * (nz - 1)):
        raise ValueError(f"ztop={ztop} unreachable with dz in [{dzmin},{dzmax}] over {nz - 1} layers")
    k = np.arange(2, nz + 1, dtype=np.float64)
    kw = nz / 8.0

    def top(kc: float) -> float:
        return float(np.sum(dzmin + (dzmax - dzmin) * 0.5 * (1.0 + np.tanh((k - kc) / kw))))

    lo, hi = -10.0 * nz, 10.0 * nz  # top(lo) ~ dzmax*(nz-1) > ztop > top(hi)
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if top(mid) > ztop:
            lo = mid
        else:
            hi = mid
    kc = 0.5 * (lo + hi)
    dz = dzmin + (dzmax - dzmin) * 0.5 * (1.0 + np.tanh((k - kc) / kw))
    z = np.zeros(nz)
    z[1:] = np.cumsum(dz)
    z[-1] = ztop
    return z


def vertical_grid(z: np.ndarray, anelastic_scale: Optional[float] = None) -> VerticalGrid:
    """Metrics of gridmanager.F90 for level heights z.  rho = rhon = 1
    (Boussinesq, gridmanager.F90:2402-2421) unless anelastic_scale is given, in
    which case rho = exp(-z/H) at w levels and rhon = exp(-zn/H)."""
    z = np.asarray(z, dtype=np.float64)
    nz = z.shape[0]
    zn = np.zeros(nz)
    zn[1:] = 0.5 * (z[:-1] + z[1:])
    zn[0] = -zn[1]
    if anelastic_scale is None:
        rho = np.ones(nz)
        rhon = np.ones(nz)
    else:
        rho = np.exp(-z / anelastic_scale)
        rhon = np.exp(-zn / anelastic_scale)
    dz = np.zeros(nz)
    dzn = np.zeros(nz)
    rdz = np.zeros(nz)
    rdzn = np.zeros(nz)
    dz[1:] = z[1:] - z[:-1]
    dzn[1:] = zn[1:] - zn[:-1]
    rdz[1:] = 1.0 / dz[1:]
    rdzn[1:] = 1.0 / dzn[1:]
    tzc1 = np.zeros(nz)
    tzc2 = np.zeros(nz)
    tzd1 = np.zeros(nz)
    tzd2 = np.zeros(nz)
    # 0-based index m = k-1; k = 2..nz for tzc (k = nz included), 2..nz-1 for tzd
    m = np.arange(1, nz)
    tzc1[m] = 0.25 * rdz[m] * rho[m - 1] / rhon[m]
    tzc2[m] = 0.25 * rdz[m] * rho[m] / rhon[m]
    m = np.arange(1, nz - 1)
    tzd1[m] = 0.25 * rdzn[m + 1] * rhon[m] / rho[m]
    tzd2[m] = 0.25 * rdzn[m + 1] * rhon[m + 1] / rho[m]
    return VerticalGrid(z=z, zn=zn, rho=rho, rhon=rhon, dz=dz, dzn=dzn, rdz=rdz, rdzn=rdzn, tzc1=tzc1,
                        tzc2=tzc2, tzd1=tzd1, tzd2=tzd2)


def horizontal_c(dxx: float) -> float:
    """cx = 1./dx with dx = real(dxx) (single precision), gridmanager.F90:2627-2633."""
    return 1.0 / float(np.float32(dxx))


def halo_fill(f: np.ndarray) -> np.ndarray:
    """Periodic depth-2 fill of an (nx+4, ny+4, nz) array in place (corners included)."""
    nx = f.shape[0] - 4
    ny = f.shape[1] - 4
    f[:, 0:2, :] = f[:, ny:ny + 2, :]
    f[:, ny + 2:ny + 4, :] = f[:, 2:4, :]
    f[0:2, :, :] = f[nx:nx + 2, :, :]
    f[nx + 2:nx + 4, :, :] = f[2:4, :, :]
    return f


def empty_field(nz: int, ny: int, nx: int, dtype=np.float64) -> np.ndarray:
    return np.zeros((nx + 4, ny + 4, nz), dtype=dtype)


@dataclass
class State:
    """Fields of one local domain: ids 0..2 = u, v, w; 3 = th; then the n_q
    q fields and, after them, the n_tracers tracers (state.F90:86-92)."""
    nz: int
    ny: int
    nx: int
    cx: float
    cy: float
    grid: VerticalGrid
    fields: List[np.ndarray]
    zfields: List[np.ndarray] = field(default_factory=list)
    n_tracers: int = 0

    @property
    def n_q(self) -> int:
        return len(self.fields) - 4 - self.n_tracers

    def shape(self):
        return (self.nx + 4, self.ny + 4, self.nz)


def nondivergent_flow(nz: int, ny: int, nx: int, dxy: float, z: np.ndarray, rng: np.random.Generator,
                      u_mean: float = 5.0, uamp: float = 1.5, wamp: float = 0.3, modes: int = 6):
    """Interior (nx, ny, nz) u, v, w whose discrete divergence (the scalar flux
    divergence of the advection scheme for rho = 1: (u(i-1)-u(i))/dx +
    (v(j-1)-v(j))/dy + (w(k-1)-w(k))/dz(k)) vanishes at every cell k >= 2.
    Horizontal part from a streamfunction psi at cell corners (u = -d psi/dy,
    v = d psi/dx); vertical part from chi at (u face, w level) with chi = 0 at
    levels 1 and nz (u -= d chi/dz, w = d chi/dx), so w = 0 there.  Fourier
    modes with seeded wavenumbers and phases, periodic in x and y."""
    x = np.arange(nx)[:, None, None] / nx
    y = np.arange(ny)[None, :, None] / ny
    zf = ((z - z[0]) / (z[-1] - z[0]))[None, None, :]
    psi = np.zeros((nx, ny, nz))
    chi = np.zeros((nx, ny, nz))
    for _ in range(modes):
        kx, ky = rng.integers(1, 4, size=2)
        ph, ph2, kz = rng.uniform(0, 2 * np.pi), rng.uniform(0, 2 * np.pi), rng.uniform(0.5, 2.0)
        a = uamp * ny * dxy / (2 * np.pi * ky * np.sqrt(modes))
        psi += a * np.sin(2 * np.pi * (kx * x + ky * y) + ph) * (1.0 + 0.5 * np.sin(2 * np.pi * kz * zf + ph2))
        b = wamp * nx * dxy / (2 * np.pi * kx * np.sqrt(modes))
        chi += b * np.sin(2 * np.pi * (kx * x + ky * y) + ph2) * np.sin(np.pi * zf)
    chi[:, :, 0] = 0.0
    chi[:, :, -1] = 0.0
    rdz = np.zeros(nz)
    rdz[1:] = 1.0 / (z[1:] - z[:-1])
    u = u_mean - (psi - np.roll(psi, 1, axis=1)) / dxy
    u[:, :, 1:] -= (chi[:, :, 1:] - chi[:, :, :-1]) * rdz[None, None, 1:]
    v = (psi - np.roll(psi, 1, axis=0)) / dxy
    w = (chi - np.roll(chi, 1, axis=0)) / dxy
    return u, v, w


def make_state(nz: int, ny: int, nx: int, dxy: float, ztop: float, seed: int, n_q: int = 1,
               anelastic_scale: Optional[float] = None, u_mean: float = 5.0, lapse: float = 0.003,
               with_zfields: bool = True, flow: str = "noise", n_tracers: int = 0) -> State:
    """Seeded synthetic state: u = u_mean + N(0,1), v = N(0,1), w = N(0,0.5) with
    w = 0 at k = 1 and nz, th = 300 + lapse*zn + U(-0.1,0.1),
    q = 0.01*exp(-zn/2000) + U(0,1e-4) (extra q fields reuse the q recipe).
    Draw order u, v, w, th noise, q noise, each as the (nz, ny, nx) interior in
    Fortran order, then the periodic halo fill.  z-fields (the t-1 level for
    centred stepping) are the same fields plus a second seeded perturbation.
    flow = "nondivergent" replaces u, v, w (after the same draws) by the
    discretely divergence-free flow of `nondivergent_flow`, drawn next from
    the same generator (dx = dy = dxy).  n_tracers tracers follow the q fields:
    tracer n = exp(-((zn - z_n)/500)^2) + U(0, 0.01) (a blob level z_n spread
    over the column), drawn after everything above."""
    z = tanh_stretched_z(nz, ztop)
    g = vertical_grid(z, anelastic_scale)
    rng = np.random.default_rng(seed)
    zn = g.zn[None, None, :]

    def interior(a: np.ndarray) -> np.ndarray:
        f = empty_field(nz, ny, nx)
        f[2:nx + 2, 2:ny + 2, :] = a
        return halo_fill(f)

    shp = (nx, ny, nz)
    u = u_mean + rng.standard_normal(shp)
    v = rng.standard_normal(shp)
    w = 0.5 * rng.standard_normal(shp)
    w[:, :, 0] = 0.0
    w[:, :, -1] = 0.0
    th = 300.0 + lapse * zn + rng.uniform(-0.1, 0.1, shp)
    qs = [0.01 * np.exp(-zn / 2000.0) + rng.uniform(0.0, 1e-4, shp) for _ in range(max(n_q, 0))]
    if flow == "nondivergent":
        u, v, w = nondivergent_flow(nz, ny, nx, dxy, z, rng, u_mean=u_mean)
    elif flow != "noise":
        raise ValueError(f"unknown flow kind {flow!r}")
    trs = [np.exp(-((zn - (n + 1) * ztop / (n_tracers + 1)) / 500.0) ** 2) + rng.uniform(0.0, 0.01, shp)
           for n in range(max(n_tracers, 0))]
    fields = [interior(a) for a in [u, v, w, th, *qs, *trs]]
    zfields: List[np.ndarray] = []
    if with_zfields:
        rng2 = np.random.default_rng(seed + 7919)
        for n, f in enumerate(fields):
            scale = [0.05, 0.05, 0.02, 0.01, 1e-6][min(n, 4)]
            zf = f.copy()
            zf[2:nx + 2, 2:ny + 2, :] += scale * rng2.standard_normal(shp)
            if n == 2:
                zf[:, :, 0] = 0.0
                zf[:, :, -1] = 0.0
            zfields.append(halo_fill(zf))
    c = horizontal_c(dxy)
    return State(nz=nz, ny=ny, nx=nx, cx=c, cy=c, grid=g, fields=fields, zfields=zfields,
                 n_tracers=max(n_tracers, 0))


def cfl_dt(state: State, cfl: float = 0.3) -> float:
    """dt = cfl / (max|u| cx + max|v| cy + max_k max|w| rdzn(k+1)), the CFL
    measure of components/cfltest/src/cfltest.F90:85-88, from the interior."""
    nx, ny = state.nx, state.ny
    sl = (slice(2, nx + 2), slice(2, ny + 2))
    u = np.abs(state.fields[0][sl]).max()
    v = np.abs(state.fields[1][sl]).max()
    wmax_k = np.abs(state.fields[2][sl]).max(axis=(0, 1))
    rdzn_next = np.zeros(state.nz)
    rdzn_next[:-1] = state.grid.rdzn[1:]
    wterm = float(np.max(wmax_k * rdzn_next))
    return cfl / (u * state.cx + v * state.cy + wterm)


ENSEMBLE_NXY = (64, 96, 128, 192, 256)
ENSEMBLE_NZ = (64, 96, 128)


def ensemble_states(n_members: int = 48, seed: int = 0, dxy: float = 50.0, ztop: float = 2000.0,
                    nxy=ENSEMBLE_NXY, nzs=ENSEMBLE_NZ) -> List[State]:
    """Members of the batched-ensemble workload: sizes nx, ny drawn from `nxy`
    and nz from `nzs`, mean u ~ U(0, 10) m/s and lapse rate ~ U(0.001, 0.006)
    K/m, drawn in that order per member from default_rng(seed); member m's
    fields are make_state(..., seed=seed+m) (dx = dy = dxy, z tanh-stretched to
    ztop, one q field)."""
    rng = np.random.default_rng(seed)
    out = []
    for m in range(n_members):
        nx = int(rng.choice(nxy))
        ny = int(rng.choice(nxy))
        nz = int(rng.choice(nzs))
        u_mean = float(rng.uniform(0.0, 10.0))
        lapse = float(rng.uniform(0.001, 0.006))
        out.append(make_state(nz, ny, nx, dxy, ztop, seed=seed + m, n_q=1, u_mean=u_mean, lapse=lapse,
                              with_zfields=False))
    return out
