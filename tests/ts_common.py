"""Shared column sets for the two-stream tests."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "atmosphere-model_amd"))

from atmos.twostream import synthetic_columns  # noqa: E402

GUARD = 1e-4


def conditioned_columns(ncol, nlay=64, n_sw=32, n_lw=32, seed=0, cloud_fraction=0.2):
    """The synthetic recipe with the optical depths kept away from zero:
    k in [0.02, 1] m^2/kg with no pressure dependence and water vapour
    1e-3 kg/kg throughout, so every layer's gas optical depth lies between
    ~1e-4 and ~1 (clouds add theirs).  There the reference's formulas lose
    no more than ~1e-13 of a column's flux scale to rounding, and FP64
    parity can be held to 1e-12."""
    cols = synthetic_columns(ncol, nlay=nlay, n_sw=n_sw, n_lw=n_lw, seed=seed, k_range=(0.02, 1.0),
                             cloud_fraction=cloud_fraction)
    t = cols["optics"]
    for i in range(n_sw + n_lw):
        t[5 * i + 1] = t[5 * i + 2] = 0.0
    cols["h2o"][:] = 1e-3
    return cols


def column_scale(ref):
    """Per-column flux scale: the largest |value| of the column."""
    return np.maximum(np.max(np.abs(ref), axis=-1, keepdims=True), 1e-300)


def subset(cols, idx):
    """The columns idx of a column dict (host arrays)."""
    out = dict(cols)
    for n in ("play", "tlay", "h2o", "cloud_tau", "cloud_ssa", "cloud_asy", "plev", "tlev", "mu0", "sfc_alb",
              "tsfc", "irradiance"):
        out[n] = np.ascontiguousarray(np.asarray(cols[n])[idx])
    out["ncol"] = len(idx)
    return out


def rel_to_column(a, ref):
    """max over levels of |a - ref| / the column's scale, per column."""
    return np.max(np.abs(a - ref), axis=-1) / column_scale(ref)[..., 0]
