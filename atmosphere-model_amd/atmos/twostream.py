"""General two-stream correlated-k columns on the GPU (atm_twostream).

The column solver of src/mod_radiation.f90 (solve_shortwave_column 158-268,
solve_longwave_column 270-347, calculate_fluxes 349-381,
delta_eddington_rescaling 383-398) over inputs the reference fixes: the
asymmetry, solar zenith angle, surface albedo and scattering come per layer
and per column, and the gas optics are an analytic k-distribution
(ln k = c0 + c1 x + c2 x^2 + c3 y + c4 y^2, x = ln(p / p_ref), y = T - t_ref,
per g-point, k in m^2 per kg of water vapour).  Arrays are [column][level]
with the level index contiguous; level 0 is the surface, layer k lies
between levels k and k + 1.

`synthetic_columns` builds the seeded workload: levels evenly spaced in
log p from the surface pressure to 10 hPa, a 6.5 K/km lapse rate below
11 km and isothermal above, water vapour exponential in height (1e-2 kg/kg at
the surface, 2 km scale height), cloud in 20 % of layers (single-scattering
albedo U[0, 0.99], asymmetry 0.85), cos(zenith) U[0.1, 1], surface albedo
U[0.05, 0.3], absorption coefficients log-uniform in [1e-6, 1e2] m^2/kg at
(p_ref, t_ref) with quadratic ln-p and T dependence, and the Planck source
from the layer temperatures.

`twostream` runs the library on device arrays; there is no CPU path.
torch brings its own HIP runtime: in a process that uses both, torch must
bring up the device (torch.cuda.init()) before the library's first call.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

GRAVITY = 9.807
P_REF = 1e5
T_REF = 250.0
GUARD_FP64 = 1e-4
LAYER_FIELDS = ("play", "tlay", "h2o", "cloud_tau", "cloud_ssa", "cloud_asy")
LEVEL_FIELDS = ("plev", "tlev")
COLUMN_FIELDS = ("mu0", "sfc_alb", "tsfc", "irradiance")
STATS = ("guarded_layers", "bad_columns")


def optics_size(n_sw: int, n_lw: int) -> int:
    """Length of the packed optics table: n_sw x 5 shortwave coefficients,
    n_lw x 5 longwave, n_sw Rayleigh coefficients (m^2 per kg of air),
    n_sw solar fractions, n_lw Planck fractions."""
    return 6 * n_sw + 6 * n_lw + n_sw


def synthetic_optics(n_sw: int, n_lw: int, rng: np.random.Generator, k_range=(1e-6, 1e2)):
    lo, hi = np.log(k_range[0]), np.log(k_range[1])

    def coeffs(n):
        c = np.empty((n, 5))
        c[:, 0] = rng.uniform(lo, hi, n)          # ln k at (p_ref, t_ref)
        c[:, 1] = rng.uniform(0.0, 1.0, n)        # pressure broadening, per ln p
        c[:, 2] = rng.uniform(-0.1, 0.1, n)       # per (ln p)^2
        c[:, 3] = rng.uniform(-0.01, 0.01, n)     # per K
        c[:, 4] = rng.uniform(-5e-5, 5e-5, n)     # per K^2
        return c

    sw, lw = coeffs(n_sw), coeffs(n_lw)
    rayleigh = rng.uniform(0.0, 2e-5, n_sw)
    solar = rng.uniform(0.5, 1.5, n_sw)
    solar /= solar.sum()
    planck = rng.uniform(0.5, 1.5, n_lw)
    planck /= planck.sum()
    return np.concatenate([sw.ravel(), lw.ravel(), rayleigh, solar, planck])


def synthetic_columns(ncol: int, nlay: int = 64, n_sw: int = 32, n_lw: int = 32, seed: int = 0,
                      k_range=(1e-6, 1e2), h2o_floor: float = 0.0, cloud_fraction: float = 0.2) -> dict:
    """The seeded column workload (host float64 arrays).  k_range and
    h2o_floor move the optical depths away from zero (the parity tests'
    well-conditioned case); the defaults are the workload itself."""
    rng = np.random.default_rng(seed)
    ps = rng.uniform(0.98e5, 1.02e5, ncol)
    frac = np.linspace(0.0, 1.0, nlay + 1)
    plev = ps[:, None] * (1e3 / ps[:, None]) ** frac[None, :]
    play = 0.5 * (plev[:, :-1] + plev[:, 1:])
    tsfc = rng.uniform(275.0, 305.0, ncol)

    def height(p):
        return -7.5e3 * np.log(p / ps[:, None])

    def temperature(z):
        return tsfc[:, None] - 6.5e-3 * np.minimum(z, 11e3)

    zl = height(play)
    tlay = temperature(zl) + rng.normal(0.0, 0.5, (ncol, nlay))
    tlev = np.empty((ncol, nlay + 1))
    if nlay > 1:
        tlev[:, 1:-1] = 0.5 * (tlay[:, :-1] + tlay[:, 1:])
        tlev[:, 0] = 1.5 * tlay[:, 0] - 0.5 * tlay[:, 1]
        tlev[:, -1] = 1.5 * tlay[:, -1] - 0.5 * tlay[:, -2]
    else:
        tlev[:, 0] = tlay[:, 0] + 3.0
        tlev[:, 1] = tlay[:, 0] - 3.0
    h2o = np.maximum(1e-2 * np.exp(-zl / 2e3), h2o_floor)
    cloud = rng.uniform(size=(ncol, nlay)) < cloud_fraction
    cloud_tau = np.where(cloud, np.exp(rng.uniform(np.log(0.1), np.log(50.0), (ncol, nlay))), 0.0)
    cloud_ssa = np.where(cloud, rng.uniform(0.0, 0.99, (ncol, nlay)), 0.0)
    cloud_asy = np.where(cloud, 0.85, 0.0)
    cols = {"play": play, "plev": plev, "tlay": tlay, "tlev": tlev, "h2o": h2o, "cloud_tau": cloud_tau,
            "cloud_ssa": cloud_ssa, "cloud_asy": cloud_asy, "mu0": rng.uniform(0.1, 1.0, ncol),
            "sfc_alb": rng.uniform(0.05, 0.3, ncol), "tsfc": tsfc, "irradiance": np.full(ncol, 1361.0),
            "optics": synthetic_optics(n_sw, n_lw, rng, k_range)}
    cols.update(ncol=ncol, nlay=nlay, n_sw=n_sw, n_lw=n_lw, p_ref=P_REF, t_ref=T_REF, gravity=GRAVITY)
    return {k: (np.ascontiguousarray(v, dtype=np.float64) if isinstance(v, np.ndarray) else v)
            for k, v in cols.items()}


def select_gpoints(cols: dict, sw=(), lw=()) -> dict:
    """The same columns with the optics table cut down to the listed
    shortwave and longwave g-points (in that order)."""
    n_sw, n_lw = cols["n_sw"], cols["n_lw"]
    t = np.asarray(cols["optics"])
    csw = t[:5 * n_sw].reshape(n_sw, 5)
    clw = t[5 * n_sw:5 * (n_sw + n_lw)].reshape(n_lw, 5)
    rest = t[5 * (n_sw + n_lw):]
    ray, sol, plk = rest[:n_sw], rest[n_sw:2 * n_sw], rest[2 * n_sw:]
    sw, lw = list(sw), list(lw)
    out = dict(cols)
    out["optics"] = np.concatenate([csw[sw].ravel(), clw[lw].ravel(), ray[sw], sol[sw], plk[lw]])
    out["n_sw"], out["n_lw"] = len(sw), len(lw)
    return out


class Columns(C.Structure):
    """atm_ts_columns (include/atm.h)."""
    _fields_ = [("ncol", C.c_int64), ("nlay", C.c_int32), ("n_sw", C.c_int32), ("n_lw", C.c_int32),
                ("p_ref", C.c_double), ("t_ref", C.c_double), ("gravity", C.c_double), ("guard", C.c_double),
                ("optics", C.c_void_p)] + [(n, C.c_void_p) for n in LAYER_FIELDS + LEVEL_FIELDS + COLUMN_FIELDS]


class Stats(C.Structure):
    _fields_ = [(n, C.c_int64) for n in STATS]


def _lib():
    from .model import load_library
    L = load_library()
    if not getattr(L, "_ts_declared", False):
        for name in ("atm_twostream", "atm_twostream_noalloc"):
            if hasattr(L, name):
                f = getattr(L, name)
                f.argtypes = [C.POINTER(Columns), C.c_int, C.c_void_p, C.c_void_p, C.POINTER(Stats), C.c_void_p]
                f.restype = C.c_int
        L.atm_twostream_last_error.restype = C.c_char_p
        L.atm_stream_synchronize.argtypes = [C.c_void_p]
        L.atm_stream_timer_begin.argtypes = [C.c_void_p]
        L.atm_stream_timer_end.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
        L._ts_declared = True
    return L


_STREAM = None


def stream():
    """The library-runtime stream the columns run on (created once)."""
    global _STREAM
    if _STREAM is None:
        from .model import stream_create
        _STREAM = stream_create()
    return _STREAM


def synchronize():
    _lib().atm_stream_synchronize(C.c_void_p(stream()))


def timer_begin():
    if _lib().atm_stream_timer_begin(C.c_void_p(stream())) != 0:
        raise RuntimeError("atm_stream_timer_begin failed")


def timer_end() -> float:
    ms = C.c_double(0.0)
    if _lib().atm_stream_timer_end(C.c_void_p(stream()), C.byref(ms)) != 0:
        raise RuntimeError("atm_stream_timer_end failed")
    return ms.value


def to_device(cols: dict, device="cuda"):
    """The column arrays as contiguous float64 tensors on the device."""
    import torch
    out = dict(cols)
    for n in LAYER_FIELDS + LEVEL_FIELDS + COLUMN_FIELDS + ("optics",):
        out[n] = torch.as_tensor(np.ascontiguousarray(cols[n]), dtype=torch.float64).to(device).contiguous()
    return out


def _struct(cols: dict, guard: float) -> Columns:
    import torch
    s = Columns()
    s.ncol, s.nlay, s.n_sw, s.n_lw = int(cols["ncol"]), int(cols["nlay"]), int(cols["n_sw"]), int(cols["n_lw"])
    s.p_ref, s.t_ref, s.gravity, s.guard = cols["p_ref"], cols["t_ref"], cols["gravity"], guard
    shapes = {**{n: (s.ncol, s.nlay) for n in LAYER_FIELDS}, **{n: (s.ncol, s.nlay + 1) for n in LEVEL_FIELDS},
              **{n: (s.ncol,) for n in COLUMN_FIELDS}, "optics": (optics_size(s.n_sw, s.n_lw),)}
    for n, shp in shapes.items():
        t = cols[n]
        if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()
                and tuple(t.shape) == shp):
            raise ValueError(f"{n}: expected a contiguous float64 CUDA tensor of shape {shp}")
        setattr(s, n, t.data_ptr())
    return s


def twostream(cols: dict, precision: str = "fp64", guard: float = GUARD_FP64, fluxes: bool = False,
              noalloc: bool = False, stats: bool = True, out=None, wait: bool = True):
    """Net flux [column][level] (and, with fluxes, the (4, ncol, nlay + 1)
    shortwave up, shortwave down, longwave up, longwave down) of device
    columns (to_device); returns (net, fluxes or None, stats dict or None).
    The library runs on its own stream (stream()); torch's work is waited
    for first, and the library's after the call unless wait is False (then
    call synchronize() before torch reads the outputs)."""
    import torch
    L = _lib()
    s = _struct(cols, guard)
    shape = (s.ncol, s.nlay + 1)
    net = out if out is not None else torch.empty(shape, dtype=torch.float64, device="cuda")
    fl = torch.empty((4,) + shape, dtype=torch.float64, device="cuda") if fluxes else None
    st = Stats()
    fn = L.atm_twostream_noalloc if noalloc else L.atm_twostream
    torch.cuda.synchronize()
    rc = fn(C.byref(s), {"fp64": 0, "mixed": 1}[precision], net.data_ptr(), fl.data_ptr() if fluxes else None,
            C.byref(st) if stats else None, C.c_void_p(stream()))
    if rc != 0:
        from .model import AtmError
        raise AtmError(rc, L.atm_twostream_last_error().decode())
    if wait:
        synchronize()
    return net, fl, ({n: int(getattr(st, n)) for n in STATS} if stats else None)
