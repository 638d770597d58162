"""Shared helpers for the GPU parity tests: build the same initial state in
the HIP library and in the CPU oracle and compare fields."""
import numpy as np

import oracle as O
from atmos import driver
from atmos.config import Config
from atmos.ic import make_initial_state
from atmos.model import Model

PARITY = 1e-12  # norm-relative FP64 tolerance (max|gpu-oracle| / max|oracle|)


def make_pair(nx, ny, nlay, dx=100e3, seed=20261203, **ickw):
    ic = make_initial_state(nx, ny, nlay, dx, seed=seed, **ickw)
    cfg = Config(dt_max=60.0, t_final=1e12, dt_output=1e12, dt_radiation=1e12,
                 dt_physics=1e12, dt_remap=300.0, radiation_on=False, physics_on=False)
    gpu = Model(nx, ny, nlay, dx, dx)
    orc = O.Oracle(nx, ny, nlay, dx, dx)
    sg = driver.init_model(gpu, ic, Config(**cfg.__dict__))
    so = driver.init_model(orc, ic, Config(**cfg.__dict__))
    return gpu, orc, cfg, sg, so, ic


ORC_ID = {"dp": O.DP, "pt": O.PT, "ud": O.UD, "vd": O.VD, "net_flux": O.NET_FLUX,
          "dpc": O.DPC, "ptc": O.PTC, "uc": O.UC, "vc": O.VC, "gz": O.GZ, "pkap": O.PKAP,
          "plev": O.PLEV, "ts": O.TS}


def compare(gpu, orc, name, region=None):
    """norm-relative difference of a field over the interior (default) or
    over a region (a0, a1) of the halo'd index space (0-based, inclusive)."""
    g = gpu.get_full(name)
    o = orc.field(ORC_ID[name])
    if o.ndim == 2:
        o = o[None]
    if region is None:
        h = 7
        sl = (slice(None), slice(h, h + gpu.ny), slice(h, h + gpu.nx))
    else:
        a0, a1 = region
        sl = (slice(None), slice(a0, g.shape[1] - a1), slice(a0, g.shape[2] - a1))
    a, b = g[sl], o[sl]
    den = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / den) if den > 0 else float(np.max(np.abs(a - b)))
