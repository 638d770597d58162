"""Seeded synthetic initial conditions for the doubly periodic f-plane.

The reference reads its initial state from initial.nc (src/mod_input.f90:
21-252); this module builds the same variables directly:
dp, pt, u (ud), v (vd) on (nlay, ny, nx), gzs and ts on (ny, nx), the hybrid
coefficients ak/bk on nlay+1 levels and the xc/yc axes.

Recipe (one per grid size, fully determined by the seed):
  * hybrid sigma-p levels evenly spaced in ln p from 1000 hPa (k=1, surface)
    to p_top = 10 hPa (k=nlay+1), with Ak = p_top and
    Bk = (p_ref - p_top) / (1e5 - p_top), so Ak(1) = p_top, Bk(1) = 1,
    Bk(nlay+1) = 0 (the remap's target grid, src/mod_remap.f90:57-59);
  * surface pressure 1000 hPa + a 5 hPa Gaussian bump (sigma 500 km) at the
    domain centre; plev(k) = Ak + Bk (ps - p_top);
  * temperature 280 K at the surface, 6.5 K/km lapse rate up to 11 km,
    isothermal 208.5 K above; the interface geopotential is the analytic
    hydrostatic height of that profile and pt is taken from the model's own
    hydrostatic relation gz(k+1) - gz(k) = cp pt (pkap(k) - pkap(k+1))
    (src/mod_sw_dyn.f90:621-626), so pt = T / p^kappa with no p0 (Q15);
  * u, v = 10 m/s + an independent k^-3-spectrum noise field per level and
    component (std 1 m/s, random phases from the seed);
  * gzs = 0, ts = 280 K; optional k^-3 temperature noise (ensembles).
"""
from __future__ import annotations

import numpy as np

R_DRY = float(np.float32(287.052874))  # src/mod_constants.f90:11 (binary32 literal)
CP_DRY = 1005.0
KAPPA = R_DRY / CP_DRY


def spectral_noise(rng: np.random.Generator, ny: int, nx: int, slope: float = -3.0) -> np.ndarray:
    """Periodic 2-D field with isotropic power spectrum ~ k^slope, zero mean,
    unit standard deviation, random phases."""
    ky = np.fft.fftfreq(ny) * ny
    kx = np.fft.rfftfreq(nx) * nx
    kk = np.sqrt(ky[:, None] ** 2 + kx[None, :] ** 2)
    amp = np.zeros_like(kk)
    nz = kk > 0
    amp[nz] = kk[nz] ** (slope / 2.0)
    phase = rng.uniform(0.0, 2.0 * np.pi, size=kk.shape)
    spec = amp * np.exp(1j * phase)
    field = np.fft.irfft2(spec, s=(ny, nx))
    field -= field.mean()
    sd = field.std()
    return field / sd if sd > 0 else field


def height_of_pressure(p, ps, T0=280.0, lapse=6.5e-3, T_strat=208.5, g=9.807):
    """Height (m) of pressure p over a surface at ps for the lapse-rate /
    isothermal profile."""
    expo = R_DRY * lapse / g
    z_t = (T0 - T_strat) / lapse
    p_t = ps * (T_strat / T0) ** (1.0 / expo)
    z_low = (T0 / lapse) * (1.0 - (p / ps) ** expo)
    z_high = z_t + (R_DRY * T_strat / g) * np.log(p_t / p)
    return np.where(p >= p_t, z_low, z_high)


def make_initial_state(nx: int, ny: int, nlay: int, dx: float, dy: float | None = None, *,
                       seed: int = 20261203, p_top: float = 1000.0, ps_ref: float = 1e5,
                       bump: float = 500.0, sigma: float = 500e3, T0: float = 280.0,
                       lapse: float = 6.5e-3, T_strat: float = 208.5, u0: float = 10.0,
                       v0: float = 10.0, wind_noise: float = 1.0, t_noise: float = 0.0,
                       ts: float = 280.0, gravity: float = 9.807) -> dict:
    dy = dx if dy is None else dy
    rng = np.random.default_rng(seed)
    Lx, Ly = nx * dx, ny * dy
    xc = np.linspace((dx - Lx) / 2, (Lx - dx) / 2, nx)
    yc = np.linspace((dy - Ly) / 2, (Ly - dy) / 2, ny)

    # hybrid coefficients (k = 1 .. nlay+1 -> index 0 .. nlay)
    p_ref = np.exp(np.linspace(np.log(ps_ref), np.log(p_top), nlay + 1))
    p_ref[0], p_ref[-1] = ps_ref, p_top
    ak = np.full(nlay + 1, p_top)
    bk = (p_ref - p_top) / (ps_ref - p_top)
    bk[0], bk[-1] = 1.0, 0.0

    r2 = xc[None, :] ** 2 + yc[:, None] ** 2
    ps = ps_ref + bump * np.exp(-r2 / (2.0 * sigma ** 2))
    plev = ak[:, None, None] + bk[:, None, None] * (ps[None] - p_top)
    dp = plev[:-1] - plev[1:]

    z = height_of_pressure(plev, ps[None], T0, lapse, T_strat, gravity)
    gz = gravity * z
    pkap = plev ** KAPPA
    pt = (gz[1:] - gz[:-1]) / (CP_DRY * (pkap[:-1] - pkap[1:]))
    if t_noise > 0:
        # temperature perturbation -> pt perturbation at fixed layer p^kappa
        playkap = (pkap[1:] - pkap[:-1]) / np.log(pkap[1:] / pkap[:-1])
        for k in range(nlay):
            pt[k] += t_noise * spectral_noise(rng, ny, nx) / playkap[k]

    u = np.empty((nlay, ny, nx))
    v = np.empty((nlay, ny, nx))
    for k in range(nlay):
        u[k] = u0 + wind_noise * spectral_noise(rng, ny, nx)
        v[k] = v0 + wind_noise * spectral_noise(rng, ny, nx)

    return dict(dp=np.ascontiguousarray(dp), pt=np.ascontiguousarray(pt), u=u, v=v,
                gzs=np.zeros((ny, nx)), ts=np.full((ny, nx), ts), ak=ak, bk=bk,
                xc=xc, yc=yc, dx=dx, dy=dy)


def radius_of(ic: dict) -> np.ndarray:
    """Distance from the domain centre (src/mod_input.f90:233-243)."""
    xc, yc = ic["xc"], ic["yc"]
    return np.sqrt(xc[None, :] ** 2 + yc[:, None] ** 2)
