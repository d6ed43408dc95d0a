"""Seeded synthetic domains for the wind solve (inputs, not part of the solve).

The vertical grid and terrain-following metrics restate HICAR's own generators
so that the solver sees the kind of operator HICAR builds:

* ``auto_dz`` case 1 (ICON cubic half-levels)   src/objects/domain_obj.F90:1504-1632
* ``smooth_array_2d`` (moving box, truncated)    src/utilities/array_utilities.F90:412-460
* ``array_offset_x/y_2d``                        src/utilities/array_utilities.F90:158-197
* ``split_topography`` (h1 = smooth(h), h2=h-h1)  src/objects/domain_obj.F90:2279-2354
* SLEVE jacobians, dzdx/dzdy, dzdx_u/dzdy_v      src/objects/domain_obj.F90:1647-1924
* ``advection_dz = dz_levels(k)``                src/objects/domain_obj.F90:2092-2094

Builder choices where the reference leaves a value open (stretch_fac 0.7, SLEVE
decay 2/6, n = 1.35, smoothing 5 x 100, z0 = 0.1 m, power-law exponent 1/7) are
the defaults listed for these workloads. On one tile HICAR never writes the
outermost faces dzdx_u(ims), dzdx_u(ime+1), dzdy_v(jms), dzdy_v(jme+1) on the
SLEVE path, so they keep their allocation value 0 here too.

Array layout (numpy C order == Fortran (i,k,j)):
  mass (ny, nz, nx), u (ny, nz, nx+1), v (ny+1, nz, nx), 2-D (ny, nx).
All fields are float64.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


# --------------------------------------------------------------- vertical grid
def auto_dz(nz: int, lowest: float, top: float, stretch_fac: float = 0.7) -> np.ndarray:
    """ICON-style cubic half-levels (domain_obj.F90:1532-1544); returns dz(1:nz)."""
    x1 = (2.0 * stretch_fac - 1.0) * lowest
    b = (top - (x1 / 6.0) * nz ** 3 - (lowest - x1 / 6.0) * nz) / (nz ** 2 - 1.0 / 3.0 * nz ** 3 - 2.0 / 3.0 * nz)
    a = (x1 - 2.0 * b) / 6.0
    c = lowest - (a + b)
    jk = np.arange(nz + 1, dtype=np.float64)
    vct = a * jk ** 3 + b * jk ** 2 + c * jk
    dz = np.diff(vct)
    if abs(vct[0]) > 1e-6 or np.any(dz <= 0):
        raise ValueError("auto_dz: half-levels are not monotonically increasing")
    return dz


# ------------------------------------------------------------------ smoothing
def _box_sum_1d(a: np.ndarray, w: int, axis: int) -> tuple[np.ndarray, np.ndarray]:
    """Sum over the truncated window [i-w, i+w] along axis, and the window length."""
    n = a.shape[axis]
    cs = np.cumsum(a, axis=axis)
    zero_shape = list(a.shape)
    zero_shape[axis] = 1
    cs = np.concatenate([np.zeros(zero_shape), cs], axis=axis)
    idx = np.arange(n)
    hi = np.minimum(n, idx + w + 1)
    lo = np.maximum(0, idx - w)
    s = np.take(cs, hi, axis=axis) - np.take(cs, lo, axis=axis)
    return s, (hi - lo).astype(np.float64)


def smooth_array_2d(a: np.ndarray, windowsize: int, nsmooths: int = 1) -> np.ndarray:
    """Moving (2w+1)^2 box mean, truncated at the edges (array_utilities.F90:412-460).
    ``a`` is indexed [j, i]."""
    out = np.array(a, dtype=np.float64)
    for _ in range(nsmooths):
        s, ni = _box_sum_1d(out, windowsize, axis=1)
        s, nj = _box_sum_1d(s, windowsize, axis=0)
        out = s / (nj[:, None] * ni[None, :])
    return out


def array_offset_x_2d(a: np.ndarray) -> np.ndarray:
    """Cell -> u-face interpolation with linear extrapolation (array_utilities.F90:158-176)."""
    ny, nx = a.shape
    o = np.empty((ny, nx + 1))
    o[:, 0] = 1.5 * a[:, 0] - 0.5 * a[:, 1]
    o[:, 1:nx] = (a[:, : nx - 1] + a[:, 1:]) * 0.5
    o[:, nx] = 1.5 * a[:, nx - 1] - 0.5 * a[:, nx - 2]
    return o


def array_offset_y_2d(a: np.ndarray) -> np.ndarray:
    """Cell -> v-face interpolation (array_utilities.F90:178-197)."""
    return array_offset_x_2d(a.T).T


# -------------------------------------------------------------------- terrain
def gaussian_hill(nx: int, ny: int, dx: float, height: float, sigma: float) -> np.ndarray:
    x = (np.arange(nx) - (nx - 1) / 2.0) * dx
    y = (np.arange(ny) - (ny - 1) / 2.0) * dx
    r2 = x[None, :] ** 2 + y[:, None] ** 2
    return height * np.exp(-r2 / (2.0 * sigma ** 2))


def fractal_terrain(nx: int, ny: int, relief: float, rng: np.random.Generator) -> np.ndarray:
    """Random-phase Fourier synthesis, amplitude ~ |k|^-1 (power ~ k^-2), rescaled to
    [0, relief]."""
    ky = np.fft.fftfreq(ny)[:, None]
    kx = np.fft.rfftfreq(nx)[None, :]
    kk = np.sqrt(kx ** 2 + ky ** 2)
    amp = np.zeros_like(kk)
    amp[kk > 0] = 1.0 / kk[kk > 0]
    phase = rng.uniform(0.0, 2.0 * np.pi, size=kk.shape)
    spec = amp * np.exp(1j * phase)
    h = np.fft.irfft2(spec, s=(ny, nx))
    h -= h.min()
    h *= relief / max(h.max(), 1e-30)
    return h


# ---------------------------------------------------------------------- SLEVE
@dataclass
class Sleve:
    decay_L: float = 2.0
    decay_S: float = 6.0
    n: float = 1.35
    smooth_window: int = 5
    smooth_cycles: int = 100


def split_topography(terrain: np.ndarray, sl: Sleve):
    """h1 = smooth(h), h2 = h - h1 on mass, u and v grids (domain_obj.F90:2279-2354)."""
    h1 = smooth_array_2d(terrain, sl.smooth_window, sl.smooth_cycles)
    h2 = terrain - h1
    tu = array_offset_x_2d(terrain)
    h1_u = smooth_array_2d(tu, sl.smooth_window, sl.smooth_cycles)
    h2_u = tu - h1_u
    tv = array_offset_y_2d(terrain)
    h1_v = smooth_array_2d(tv, sl.smooth_window, sl.smooth_cycles)
    h2_v = tv - h1_v
    return h1, h2, h1_u, h2_u, h1_v, h2_v


def sleve_metrics(terrain: np.ndarray, dz: np.ndarray, dx: float, sl: Sleve | None = None) -> dict:
    """SLEVE metric fields as setup_sleve builds them on one tile
    (domain_obj.F90:1647-1924, max_level = nz)."""
    sl = sl or Sleve()
    ny, nx = terrain.shape
    nz = dz.size
    h1, h2, h1_u, h2_u, h1_v, h2_v = split_topography(terrain, sl)
    H = float(dz.sum())
    s1, s2, n = H / sl.decay_L, H / sl.decay_S, sl.n

    def bfun(z, s):
        return math.sinh((H / s) ** n - (z / s) ** n) / math.sinh((H / s) ** n)

    def dbfun(z, s):
        return -n / (s ** n) * z ** (n - 1) * math.cosh((H / s) ** n - (z / s) ** n) / math.sinh((H / s) ** n)

    jac = np.empty((ny, nz, nx))
    jac_u = np.empty((ny, nz, nx + 1))
    jac_v = np.empty((ny + 1, nz, nx))
    jac_w = np.empty((ny, nz, nx))
    dzdx = np.empty((ny, nz, nx))
    dzdy = np.empty((ny, nz, nx))
    zmass = np.empty((ny, nz, nx))
    z_agl_u = np.empty((ny, nz, nx + 1))
    z_agl_v = np.empty((ny + 1, nz, nx))
    zc = np.concatenate([[0.0], np.cumsum(dz)])
    gamma_terms = []
    for k in range(nz):
        z_i = zc[k + 1]
        z_m = zc[k] + dz[k] / 2.0
        b1i, b2i = bfun(z_i, s1), bfun(z_i, s2)
        b1m, b2m = bfun(z_m, s1), bfun(z_m, s2)
        db1i, db2i = dbfun(z_i, s1), dbfun(z_i, s2)
        db1m, db2m = dbfun(z_m, s1), dbfun(z_m, s2)
        gamma_terms.append(1.0 + h1.max() * db1i + h2.max() * db2i)
        zm = z_m + h1 * b1m + h2 * b2m
        zmass[:, k, :] = zm
        z_agl_u[:, k, :] = z_m + h1_u * (b1m - 1.0) + h2_u * (b2m - 1.0)
        z_agl_v[:, k, :] = z_m + h1_v * (b1m - 1.0) + h2_v * (b2m - 1.0)
        jac[:, k, :] = 1.0 + h1 * db1m + h2 * db2m
        jac_u[:, k, :] = 1.0 + h1_u * db1m + h2_u * db2m
        jac_v[:, k, :] = 1.0 + h1_v * db1m + h2_v * db2m
        jac_w[:, k, :] = 1.0 + h1 * db1i + h2 * db2i
        dzdx[:, k, 1:nx - 1] = (b1m * (h1[:, 2:] - h1[:, :-2]) + b2m * (h2[:, 2:] - h2[:, :-2])) / (2 * dx)
        dzdx[:, k, 0] = (-zm[:, 2] + 4 * zm[:, 1] - 3 * zm[:, 0]) / (2 * dx)
        dzdx[:, k, nx - 1] = (zm[:, nx - 3] - 4 * zm[:, nx - 2] + 3 * zm[:, nx - 1]) / (2 * dx)
        dzdy[1:ny - 1, k, :] = (b1m * (h1[2:, :] - h1[:-2, :]) + b2m * (h2[2:, :] - h2[:-2, :])) / (2 * dx)
        dzdy[0, k, :] = (-zm[2, :] + 4 * zm[1, :] - 3 * zm[0, :]) / (2 * dx)
        dzdy[ny - 1, k, :] = (zm[ny - 3, :] - 4 * zm[ny - 2, :] + 3 * zm[ny - 1, :]) / (2 * dx)
    jac_w[:, nz - 1, :] = 1.0
    dzdx_u = np.zeros((ny, nz, nx + 1))
    dzdx_u[:, :, 1:nx] = (dzdx[:, :, : nx - 1] + dzdx[:, :, 1:]) * 0.5
    dzdy_v = np.zeros((ny + 1, nz, nx))
    dzdy_v[1:ny, :, :] = (dzdy[: ny - 1] + dzdy[1:]) * 0.5
    adv_dz = np.broadcast_to(dz[None, :, None], (ny, nz, nx)).copy()
    if min(jac.min(), jac_u.min(), jac_v.min(), jac_w.min()) <= 0:
        raise ValueError("SLEVE transform is not invertible for this terrain (jacobian <= 0)")
    return dict(dzdx=dzdx, dzdy=dzdy, jacobian=jac, jacobian_u=jac_u, jacobian_v=jac_v,
                jacobian_w=jac_w, dzdx_u=dzdx_u, dzdy_v=dzdy_v, advection_dz=adv_dz,
                terrain=np.array(terrain, dtype=np.float64), z=zmass, z_agl_u=z_agl_u,
                z_agl_v=z_agl_v, gamma_min=float(min(gamma_terms)), dz_levels=dz)


def flat_metrics(nx: int, ny: int, dz: np.ndarray) -> dict:
    """Flat terrain: every SLEVE factor reduces to the identity."""
    return sleve_metrics(np.zeros((ny, nx)), dz, 1.0)


# ---------------------------------------------------------------------- winds
def log_profile_winds(m: dict, speed_100m: float, z0: float, noise: float,
                      rng: np.random.Generator):
    zu = np.maximum(m["z_agl_u"], z0 * 1.0001)
    u = speed_100m * np.log(zu / z0) / math.log(100.0 / z0)
    v = np.zeros_like(m["z_agl_v"])
    u = u + rng.normal(0.0, noise, size=u.shape)
    v = v + rng.normal(0.0, noise, size=v.shape)
    w = np.zeros_like(m["jacobian"])
    return u, v, w


def power_law_winds(m: dict, speed_10m: float, shear_deg: float, noise: float,
                    rng: np.random.Generator, exponent: float = 1.0 / 7.0,
                    direction_deg: float = 0.0):
    nz = m["jacobian"].shape[1]
    theta = np.deg2rad(direction_deg + shear_deg * np.arange(nz) / max(nz - 1, 1))
    zu = np.maximum(m["z_agl_u"], 0.1)
    zv = np.maximum(m["z_agl_v"], 0.1)
    su = speed_10m * (zu / 10.0) ** exponent
    sv = speed_10m * (zv / 10.0) ** exponent
    u = su * np.cos(theta)[None, :, None] + rng.normal(0.0, noise, size=su.shape)
    v = sv * np.sin(theta)[None, :, None] + rng.normal(0.0, noise, size=sv.shape)
    w = np.zeros_like(m["jacobian"])
    return u, v, w


def lapse_rate_alpha(shape, dx: float, rng: np.random.Generator, smooth_m: float = 2000.0):
    """alpha = 0.3 + 0.7 (G - 4)/(9.8 - 4) from a seeded lapse-rate field G in
    [4, 9.8] K/km, box-smoothed over `smooth_m` horizontally (and 3 levels
    vertically), then stretched back to span [4, 9.8]."""
    ny, nz, nx = shape
    g = rng.uniform(4.0, 9.8, size=shape)
    w = max(1, int(round(smooth_m / dx / 2.0)))
    for k in range(nz):
        g[:, k, :] = smooth_array_2d(g[:, k, :], w, 1)
    s, nk = _box_sum_1d(g, 1, axis=1)
    g = s / nk[None, :, None]
    g = 4.0 + (g - g.min()) * (9.8 - 4.0) / max(g.max() - g.min(), 1e-30)
    return 0.3 + 0.7 * (g - 4.0) / (9.8 - 4.0)


# ---------------------------------------------------------------------- cases
@dataclass
class Case:
    name: str
    nx: int
    ny: int
    nz: int
    dx: float
    metrics: dict
    u: np.ndarray
    v: np.ndarray
    w: np.ndarray
    alpha: np.ndarray
    rho: np.ndarray
    info: dict = field(default_factory=dict)

    @property
    def cells(self) -> int:
        return self.nx * self.ny * self.nz


def gaussian_hill_case(nx=64, ny=64, nz=20, dx=250.0, seed=0) -> Case:
    """64x64x20, dx = 250 m, 1000 m Gaussian hill (sigma 2 km), 20 levels 20 m -> 10 km,
    log-profile westerly 10 m/s at 100 m + N(0, 0.5) noise, alpha = 1."""
    rng = np.random.default_rng(seed)
    dz = auto_dz(nz, 20.0, 10000.0)
    m = sleve_metrics(gaussian_hill(nx, ny, dx, 1000.0, 2000.0), dz, dx)
    u, v, w = log_profile_winds(m, 10.0, 0.1, 0.5, rng)
    alpha = np.ones((ny, nz, nx))
    return Case("gaussian_hill", nx, ny, nz, dx, m, u, v, w, alpha, np.ones((ny, nz, nx)))


def fractal_case(nx=512, ny=512, nz=40, dx=50.0, relief=2500.0, speed=8.0, seed=1,
                 varying_alpha=False, lowest=10.0, top=12000.0, noise=1.0,
                 direction_deg=0.0) -> Case:
    """Fractal terrain (k^-2 spectrum) rescaled to [0, relief], stretched levels
    lowest -> top, power-law wind `speed` at 10 m with 25 deg direction shear +
    N(0, noise) on u and v; alpha = 1 or the lapse-rate field."""
    rng = np.random.default_rng(seed)
    dz = auto_dz(nz, lowest, top)
    m = sleve_metrics(fractal_terrain(nx, ny, relief, rng), dz, dx)
    u, v, w = power_law_winds(m, speed, 25.0, noise, rng, direction_deg=direction_deg)
    if varying_alpha:
        alpha = lapse_rate_alpha((ny, nz, nx), dx, rng)
    else:
        alpha = np.ones((ny, nz, nx))
    return Case("fractal", nx, ny, nz, dx, m, u, v, w, alpha, np.ones((ny, nz, nx)),
                info=dict(relief=relief, speed=speed, seed=seed))


def flat_plane_case(nx=32, ny=32, nz=20, dx=250.0) -> Case:
    """The reference unit test's set-up (tests/unit/test_wind_iterative.F90:68-87,
    179-207): flat 250 m plane, 20 SLEVE levels, alpha = 1, rho = 1,
    u = 8 cos(2 pi (i - ids)/Lx), v = 6 sin(2 pi (j - jds)/Ly), w = 0."""
    dz = auto_dz(nz, 20.0, 8000.0)
    m = sleve_metrics(np.zeros((ny, nx)), dz, dx)
    lx = float(max(nx - 1, 1))
    ly = float(max(ny - 1, 1))
    iu = np.arange(nx + 1, dtype=np.float64)  # ii - ids with ids = 1, ii = 1..nx+1
    jv = np.arange(ny + 1, dtype=np.float64)
    u = np.broadcast_to(8.0 * np.cos(2.0 * np.pi * iu / lx)[None, None, :], (ny, nz, nx + 1)).copy()
    v = np.broadcast_to(6.0 * np.sin(2.0 * np.pi * jv / ly)[:, None, None], (ny + 1, nz, nx)).copy()
    w = np.zeros((ny, nz, nx))
    return Case("flat_plane", nx, ny, nz, dx, m, u, v, w, np.ones((ny, nz, nx)),
                np.ones((ny, nz, nx)))
