"""CPU oracle for the atmosphere hot path -- TEST INFRASTRUCTURE ONLY.

ctypes binding of liboracle.so (oracle/atm_oracle.c), a plain-C FP64
restatement of the reference Fortran model's hot path.  Only tests/,
__graft_entry__.smoke() and the CPU-baseline leg of bench.py may import this
package, and only as the checker; the HIP product never calls it.

`Oracle` mirrors the reference module procedures (init_sw_dyn,
cgrid_dynamics_step, ..., perform_vertical_remapping,
calculate_radiative_heating) over arrays in the reference's halo'd Fortran
layout, exposed as numpy views shaped (k, j, i) (i fastest in memory).
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

DP, PT, UD, VD, TS, PLAY, PLAYKAP, PLEV, PKAP, GZ, NET_FLUX, AK, BK, RADIUS, DPC, PTC, UC, VC = range(18)
PCM, PPM_UNCONSTRAINED, PPM_CONSTRAINED, PPM_NONNEGATIVE = 0, 1, 2, 3
HALO = 7


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE, "liboracle.so"])


def lib():
    global _LIB
    if _LIB is None:
        # ORACLE_LIB names another build (the sanitizer build of the host tests)
        path = os.environ.get("ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        dp = C.POINTER(C.c_double)
        vp = C.c_void_p
        L.orc_create.restype = vp
        L.orc_create.argtypes = [C.c_int, C.c_int, C.c_int] + [C.c_double] * 6
        L.orc_destroy.argtypes = [vp]
        L.orc_field.restype = dp
        L.orc_field.argtypes = [vp, C.c_int]
        L.orc_field_size.restype = C.c_long
        L.orc_field_size.argtypes = [vp, C.c_int]
        for name in ["orc_init_sw_dyn", "orc_cgrid_halo", "orc_dgrid_halo", "orc_allocate_remapping",
                     "orc_remap", "orc_allocate_radiation", "orc_radiation_halo", "orc_allocate_physics",
                     "orc_physics_tendencies", "orc_physics_halo"]:
            getattr(L, name).argtypes = [vp]
        L.orc_apply_physics.argtypes = [vp, C.c_double]
        L.orc_output_init.argtypes = [vp]
        L.orc_accumulate_output.argtypes = [vp, C.c_double]
        L.orc_output_mean.argtypes = [vp, dp]
        L.orc_physics_column_tend.argtypes = [vp, C.c_int, C.c_int, dp, dp, dp]
        L.orc_cgrid_step.argtypes = [vp, C.c_double]
        L.orc_dgrid_step.argtypes = [vp, C.c_double]
        L.orc_radiation.argtypes = [vp, C.c_double]
        L.orc_radiation.restype = C.c_int
        L.orc_halo.argtypes = [vp, dp, C.c_int]
        for name in ["orc_sw_abs", "orc_lw_abs"]:
            getattr(L, name).restype = C.c_double
            getattr(L, name).argtypes = [C.c_int, C.c_double, C.c_double, C.POINTER(C.c_int)]
        L.orc_planck.restype = C.c_double
        L.orc_planck.argtypes = [C.c_int, C.c_double, C.POINTER(C.c_int)]
        for name in ["orc_sw_sca", "orc_solar_fraction"]:
            getattr(L, name).restype = C.c_double
            getattr(L, name).argtypes = [C.c_int]
        L.orc_solar_irradiance.restype = C.c_double
        L.orc_solar_irradiance.argtypes = [C.c_double]
        L.orc_constants.argtypes = [dp]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_set_threads.restype = C.c_int
        L.orc_flux_pencil.argtypes = [dp, dp, dp, C.c_int, C.c_int]
        L.orc_vertical_ppm.argtypes = [C.c_int, dp, dp, dp, dp, dp, vp]
        L.orc_remap_column.argtypes = [C.c_int, dp, dp, dp, dp]
        L.orc_calculate_fluxes.argtypes = [C.c_int, dp, dp, dp, dp, C.c_double, C.c_double, C.c_double, dp, dp]
        L.orc_delta_eddington.argtypes = [dp, dp, dp]
        L.orc_interpolate_to_corners.argtypes = [dp, dp, C.c_int, C.c_int]
        L.orc_sw_column.restype = C.c_int
        L.orc_sw_column.argtypes = [C.c_int, C.c_int, dp, dp, dp, C.c_double, dp, dp]
        L.orc_lw_column.restype = C.c_int
        L.orc_lw_column.argtypes = [C.c_int, C.c_int, dp, dp, dp, dp, C.c_double, dp, dp]
        L.orc_ts_set_jitter.argtypes = [C.c_int]
        for name in ("orc_ts_column_g", "orc_ts_column_g_ld", "orc_ts_column_g_q"):
            getattr(L, name).argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                         dp] + [dp] * 8 + [C.c_double] * 6 + [dp, dp]
        for name in ("orc_twostream", "orc_twostream_ld", "orc_twostream_q"):
            getattr(L, name).argtypes = [C.c_long, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, dp] + \
                [dp] * 8 + [dp] * 4 + [C.c_double, C.c_double, dp, dp]
        _LIB = L
    return _LIB


def set_threads(n: int = 0) -> int:
    """Threads of the oracle's parallel loops (n <= 0: leave as is);
    returns the count in use.  Results do not depend on it."""
    return lib().orc_set_threads(int(n))


def _p(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Oracle:
    """One model instance on one image (halo width 7), reference semantics."""

    def __init__(self, nx, ny, nlay, dx, dy, gravity=9.807, coriolis=1e-5,
                 divergence_damping=0.15, vorticity_damping=0.005):
        self.nx, self.ny, self.nlay = nx, ny, nlay
        self._L = lib()
        self._h = self._L.orc_create(nx, ny, nlay, dx, dy, gravity, coriolis,
                                     divergence_damping, vorticity_damping)
        ni, nj = nx + 2 * HALO, ny + 2 * HALO
        self._views = {}
        for fid in range(18):
            n = self._L.orc_field_size(self._h, fid)
            ptr = self._L.orc_field(self._h, fid)
            arr = np.ctypeslib.as_array(ptr, shape=(n,))
            if fid in (AK, BK):
                self._views[fid] = arr
            elif fid in (TS, RADIUS):
                self._views[fid] = arr.reshape(nj, ni)
            else:
                self._views[fid] = arr.reshape(n // (ni * nj), nj, ni)

    def __del__(self):
        if getattr(self, "_h", None):
            self._L.orc_destroy(self._h)
            self._h = None

    def field(self, fid):
        return self._views[fid]

    def interior(self, fid):
        a = self._views[fid]
        return a[..., HALO:HALO + self.ny, HALO:HALO + self.nx]

    # initial state (src/mod_input.f90:151-252)
    def put(self, name, arr):
        arr = np.asarray(arr, dtype=np.float64)
        if name in ("ak", "bk"):
            self._views[AK if name == "ak" else BK][:] = arr
        elif name == "gzs":
            self.interior(GZ)[0] = arr
        else:
            fid = {"dp": DP, "pt": PT, "ud": UD, "vd": VD, "ts": TS, "radius": RADIUS,
                   "net_flux": NET_FLUX}[name]
            self.interior(fid)[...] = arr

    def get(self, name):
        fid = {"dp": DP, "pt": PT, "ud": UD, "vd": VD, "ts": TS, "net_flux": NET_FLUX,
               "gz": GZ, "plev": PLEV, "pkap": PKAP}[name]
        return self.interior(fid).copy()

    def init_prognostic_fields(self):
        top = self._views[AK][0]
        self._views[PLEV][-1] = top
        self._views[PKAP][-1] = top ** constants()["kappa"]

    def initial_halo_exchange(self):
        for fid in (DP, PT, UD, VD):
            self.halo_exchange(fid)
        self._L.orc_halo(self._h, _p(self._views[GZ]), self.nlay)
        self.halo_exchange(TS)

    def check_stability(self):
        return all(np.isfinite(self.interior(f)).all() for f in (DP, PT, UD, VD))

    # reference module procedures
    def init_sw_dyn(self):
        self._L.orc_init_sw_dyn(self._h)

    def cgrid_dynamics_step(self, dt):
        self._L.orc_cgrid_step(self._h, dt)

    def cgrid_halo_exchange(self):
        self._L.orc_cgrid_halo(self._h)

    def dgrid_dynamics_step(self, dt):
        self._L.orc_dgrid_step(self._h, dt)

    def dgrid_halo_exchange(self):
        self._L.orc_dgrid_halo(self._h)

    def allocate_remapping_arrays(self):
        self._L.orc_allocate_remapping(self._h)

    def perform_vertical_remapping(self):
        self._L.orc_remap(self._h)

    def allocate_radiation_arrays(self):
        self._L.orc_allocate_radiation(self._h)

    def calculate_radiative_heating(self, time):
        if self._L.orc_radiation(self._h, time):
            raise ValueError("gas-optics lookup outside the tabulated range")

    def radiation_halo_exchange(self):
        self._L.orc_radiation_halo(self._h)

    def allocate_physics_arrays(self):
        self._L.orc_allocate_physics(self._h)

    def calculate_physics_tendencies(self):
        self._L.orc_physics_tendencies(self._h)

    def apply_physics_tendencies(self, dt):
        self._L.orc_apply_physics(self._h, dt)

    def physics_halo_exchange(self):
        self._L.orc_physics_halo(self._h)

    def physics_column_tendencies(self, i, j):
        """(u_tend, v_tend, pt_tend) of column (i, j), 1-based reference
        indices (i = 1 is the first interior point)."""
        u, v, t = (np.zeros(self.nlay) for _ in range(3))
        self._L.orc_physics_column_tend(self._h, i, j, _p(u), _p(v), _p(t))
        return u, v, t

    def init_output(self):
        self._L.orc_output_init(self._h)

    def accumulate_output(self, dt):
        self._L.orc_accumulate_output(self._h, dt)

    def output_mean(self):
        """write_output's time means as (6, nlay+1, ny, nx): dp, pt, u, v, ts, nflx."""
        out = np.zeros((6, self.nlay + 1, self.ny, self.nx))
        self._L.orc_output_mean(self._h, _p(out))
        return out

    def halo_exchange(self, fid):
        a = self._views[fid]
        nk = 0 if a.ndim == 2 else a.shape[0]
        self._L.orc_halo(self._h, _p(a), nk)


# ---- unit-level helpers ----------------------------------------------------
def sw_abs(g, p, T):
    e = C.c_int(0)
    v = lib().orc_sw_abs(g, p, T, C.byref(e))
    return v, e.value


def lw_abs(g, p, T):
    e = C.c_int(0)
    v = lib().orc_lw_abs(g, p, T, C.byref(e))
    return v, e.value


def planck(g, T):
    e = C.c_int(0)
    v = lib().orc_planck(g, T, C.byref(e))
    return v, e.value


def sw_sca(g):
    return lib().orc_sw_sca(g)


def solar_fraction(g):
    return lib().orc_solar_fraction(g)


def solar_irradiance(r):
    return lib().orc_solar_irradiance(r)


def constants():
    out = np.zeros(5)
    lib().orc_constants(_p(out))
    return dict(zip(["pi", "R_dry", "cp_dry", "cv_dry", "kappa"], out))


def flux_pencil(q, c, variant, f0=None):
    q = np.ascontiguousarray(q, dtype=np.float64)
    c = np.ascontiguousarray(c, dtype=np.float64)
    f = np.full_like(q, np.nan) if f0 is None else np.array(f0, dtype=np.float64)
    lib().orc_flux_pencil(_p(f), _p(q), _p(c), len(q), variant)
    return f


def vertical_ppm(var, coord):
    var = np.ascontiguousarray(var, dtype=np.float64)
    coord = np.ascontiguousarray(coord, dtype=np.float64)
    L = len(var)
    top, btm, curv, raw = np.zeros(L), np.zeros(L), np.zeros(L), np.zeros(L + 1)
    lib().orc_vertical_ppm(L, _p(var), _p(coord), _p(top), _p(btm), _p(curv), raw.ctypes.data)
    return top, btm, curv, raw


def remap_column(src_var, src_coord, dst_coord):
    src_var = np.ascontiguousarray(src_var, dtype=np.float64)
    src_coord = np.ascontiguousarray(src_coord, dtype=np.float64)
    dst_coord = np.ascontiguousarray(dst_coord, dtype=np.float64)
    out = np.zeros(len(src_var))
    lib().orc_remap_column(len(src_var), _p(src_var), _p(out), _p(src_coord), _p(dst_coord))
    return out


def calculate_fluxes(R, T, Sup, Sdn, toa_dn, sfc_up, sfc_alb):
    R, T, Sup, Sdn = (np.ascontiguousarray(a, dtype=np.float64) for a in (R, T, Sup, Sdn))
    L = len(R)
    fu, fd = np.zeros(L + 1), np.zeros(L + 1)
    lib().orc_calculate_fluxes(L, _p(R), _p(T), _p(Sup), _p(Sdn), toa_dn, sfc_up, sfc_alb, _p(fu), _p(fd))
    return fu, fd


def delta_eddington(tau, ssa, g):
    a, b, c = (np.array([v], dtype=np.float64) for v in (tau, ssa, g))
    lib().orc_delta_eddington(_p(a), _p(b), _p(c))
    return a[0], b[0], c[0]


def interpolate_to_corners(q):
    q = np.ascontiguousarray(q, dtype=np.float64)
    nj, ni = q.shape
    out = np.zeros_like(q)
    lib().orc_interpolate_to_corners(_p(out), _p(q), ni, nj)
    return out


def sw_column(g, play, dm, T, irradiance):
    play, dm, T = (np.ascontiguousarray(a, dtype=np.float64) for a in (play, dm, T))
    L = len(play)
    fu, fd = np.zeros(L + 1), np.zeros(L + 1)
    err = lib().orc_sw_column(L, g, _p(play), _p(dm), _p(T), irradiance, _p(fu), _p(fd))
    return fu, fd, err


def lw_column(g, play, dm, T, Tlev, ts):
    play, dm, T, Tlev = (np.ascontiguousarray(a, dtype=np.float64) for a in (play, dm, T, Tlev))
    L = len(play)
    fu, fd = np.zeros(L + 1), np.zeros(L + 1)
    err = lib().orc_lw_column(L, g, _p(play), _p(dm), _p(T), _p(Tlev), ts, _p(fu), _p(fd))
    return fu, fd, err


def reference_gas_optics(queries):
    """Run the compiled reference harness (oracle/_ref) on (kind, g, p, T)
    rows; returns (results, constants) or None where it is not built."""
    exe = os.path.join(_HERE, "_ref", "ref_gas_optics")
    if not os.path.exists(exe):
        return None
    import tempfile
    q = np.ascontiguousarray(queries, dtype="<f8")
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as fh:
            fh.write(np.int32(len(q)).tobytes())
            fh.write(q.tobytes())
        subprocess.check_call([exe, fin, fout])
        res = np.fromfile(fout, dtype="<f8")
    return res[:len(q)], res[len(q):]


# ---- general two-stream columns (arrays [column][level], level 0 = surface) ----
_TS_LAYER = ("play", "tlay", "h2o", "cloud_tau", "cloud_ssa", "cloud_asy")


def _ts_arrays(cols):
    a = {n: np.ascontiguousarray(cols[n], dtype=np.float64) for n in
         _TS_LAYER + ("plev", "tlev", "mu0", "sfc_alb", "tsfc", "irradiance", "optics")}
    return a


def twostream(cols, guard, jitter=False, extended=None, fluxes=False):
    """Net flux [column][level] (and, with fluxes, the (4, ncol, nlay + 1)
    sw up, sw down, lw up, lw down): orc_twostream over the columns of a
    dict as built by atmos.twostream.synthetic_columns.  jitter: every
    exponential moved by one ulp (orc_ts_set_jitter), to measure the
    formulas' own sensitivity to the rounding of exp.  extended: the same
    statements in "ld" (long double, orc_twostream_ld) or "q" (binary128,
    orc_twostream_q), accuracy references."""
    a = _ts_arrays(cols)
    ncol, L = int(cols["ncol"]), int(cols["nlay"])
    net = np.zeros((ncol, L + 1))
    fl = np.zeros((4, ncol, L + 1)) if fluxes else None
    fn = {None: lib().orc_twostream, "ld": lib().orc_twostream_ld, "q": lib().orc_twostream_q}[extended]
    lib().orc_ts_set_jitter(1 if jitter else 0)
    try:
        fn(ncol, L, int(cols["n_sw"]), int(cols["n_lw"]), float(cols["p_ref"]), float(cols["t_ref"]),
           _p(a["optics"]), _p(a["play"]), _p(a["plev"]), _p(a["tlay"]), _p(a["tlev"]), _p(a["h2o"]),
           _p(a["cloud_tau"]), _p(a["cloud_ssa"]), _p(a["cloud_asy"]), _p(a["mu0"]), _p(a["sfc_alb"]),
           _p(a["tsfc"]), _p(a["irradiance"]), float(cols["gravity"]), float(guard), _p(net),
           _p(fl) if fluxes else None)
    finally:
        lib().orc_ts_set_jitter(0)
    return (net, fl) if fluxes else net


def twostream_column_g(cols, c, kind, g, guard, jitter=False, extended=None):
    """(flux up, flux down) of column c for one g-point (kind 0 shortwave,
    1 longwave; the shortwave down flux includes the direct beam)."""
    a = _ts_arrays(cols)
    L = int(cols["nlay"])
    row = {n: np.ascontiguousarray(a[n][c]) for n in _TS_LAYER + ("plev", "tlev")}
    fu, fd = np.zeros(L + 1), np.zeros(L + 1)
    fn = {None: lib().orc_ts_column_g, "ld": lib().orc_ts_column_g_ld, "q": lib().orc_ts_column_g_q}[extended]
    lib().orc_ts_set_jitter(1 if jitter else 0)
    try:
        fn(L, kind, g, int(cols["n_sw"]), int(cols["n_lw"]), float(cols["p_ref"]), float(cols["t_ref"]),
           _p(a["optics"]), _p(row["play"]), _p(row["plev"]), _p(row["tlay"]), _p(row["tlev"]), _p(row["h2o"]),
           _p(row["cloud_tau"]), _p(row["cloud_ssa"]), _p(row["cloud_asy"]), float(a["mu0"][c]),
           float(a["sfc_alb"][c]), float(a["tsfc"][c]), float(a["irradiance"][c]), float(cols["gravity"]),
           float(guard), _p(fu), _p(fd))
    finally:
        lib().orc_ts_set_jitter(0)
    return fu, fd
