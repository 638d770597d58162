"""ctypes binding of libatmos_hip.so (include/atm.h) with the reference's
module-procedure names, so the same host driver runs the GPU library and
the CPU oracle.  Field arrays cross the boundary in the reference's halo'd
Fortran layout (numpy shape (nk, ny+14, nx+14), i fastest).

There is no fallback: if the HIP library is missing or no GPU is visible,
constructing a Model raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HALO = 7
_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libatmos_hip.so")

FIELDS = {"dp": 0, "pt": 1, "ud": 2, "vd": 3, "ts": 4, "gz": 5, "net_flux": 6,
          "plev": 7, "pkap": 8, "radius": 9, "dpc": 10, "ptc": 11, "uc": 12, "vc": 13,
          "tl_dp": 14, "tl_pt": 15, "tl_ud": 16, "tl_vd": 17,
          "ad_dp": 18, "ad_pt": 19, "ad_ud": 20, "ad_vd": 21}
STATUS = {0: "ok", 1: "argument", 2: "hip", 3: "table range", 4: "non-finite", 5: "state"}
EXPORTS = ["atm_create", "atm_destroy", "atm_last_error", "atm_set_stream", "atm_set_coordinate",
           "atm_put_field", "atm_get_field", "atm_put_field_device", "atm_get_field_device",
           "atm_init_prognostic_fields", "atm_initial_halo_exchange", "atm_init_sw_dyn",
           "atm_cgrid_step", "atm_cgrid_halo", "atm_dgrid_step", "atm_dgrid_halo",
           "atm_remap_init", "atm_remap", "atm_radiation_init", "atm_radiation",
           "atm_radiation_halo", "atm_halo_exchange", "atm_check_stability", "atm_sync",
           "atm_build_info", "atm_set_path", "atm_get_path", "atm_timer_enable", "atm_timer_select", "atm_timer_read",
           "atm_physics_init", "atm_physics_tendencies", "atm_apply_physics", "atm_physics_halo",
           "atm_output_init", "atm_accumulate_output", "atm_output_mean", "atm_stream_create",
           "atm_stream_destroy", "atm_graph_begin", "atm_graph_end", "atm_graph_launch", "atm_graph_free",
           "atm_autotune", "atm_get_tuning", "atm_set_tuning", "atm_get_tiles", "atm_set_tiles", "atm_lin_init", "atm_tl_step", "atm_ad_step",
           "atm_twostream", "atm_twostream_noalloc", "atm_twostream_last_error", "atm_stream_synchronize", "atm_stream_timer_begin",
           "atm_stream_timer_end", "atm_poison", "atm_check_stability_begin", "atm_check_stability_end",
           "atm_ensemble_create", "atm_ensemble_destroy", "atm_ensemble_cgrid_step", "atm_ensemble_cgrid_halo",
           "atm_ensemble_dgrid_step", "atm_ensemble_dgrid_halo", "atm_ensemble_remap", "atm_ensemble_radiation", "atm_ensemble_graph_begin",
           "atm_ensemble_graph_end", "atm_ensemble_graph_launch"]
PATHS = {"fused": 0, "reference": 1}
PRECISIONS = {"fp64": 0, "mixed": 1}
TIMERS = {"c1": 0, "c2": 1, "d1": 2, "d2": 3, "scan": 4, "halo": 5, "remap": 6, "radiation": 7,
          "physics": 8}


class AtmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"atm status {code} ({STATUS.get(code, '?')}): {msg}")
        self.code = code


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nlay", C.c_int32),
                ("dx", C.c_double), ("dy", C.c_double)]


class Params(C.Structure):
    _fields_ = [("gravity", C.c_double), ("coriolis_parameter", C.c_double),
                ("divergence_damping_coefficient", C.c_double),
                ("vorticity_damping_coefficient", C.c_double)]


_LIB = None


def load_library(path: str = LIB_PATH):
    """Load libatmos_hip.so and declare its signatures (no GPU needed)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise RuntimeError(f"HIP library not built: {path} (run __graft_entry__.build())")
    L = C.CDLL(path)
    vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.c_int
    L.atm_create.argtypes = [C.POINTER(Grid), C.POINTER(Params), ip, C.POINTER(vp)]
    L.atm_last_error.restype = C.c_char_p
    L.atm_last_error.argtypes = [vp]
    L.atm_build_info.restype = C.c_char_p
    for n in ["atm_destroy", "atm_init_prognostic_fields", "atm_initial_halo_exchange",
              "atm_init_sw_dyn", "atm_cgrid_halo", "atm_dgrid_halo", "atm_remap_init", "atm_remap",
              "atm_radiation_init", "atm_radiation_halo", "atm_sync", "atm_physics_init",
              "atm_physics_tendencies", "atm_physics_halo", "atm_output_init"]:
        getattr(L, n).argtypes = [vp]
    L.atm_set_stream.argtypes = [vp, vp]
    L.atm_stream_create.argtypes = [C.POINTER(vp)]
    L.atm_stream_destroy.argtypes = [vp]
    for n in ["atm_graph_begin", "atm_graph_end", "atm_graph_free"]:
        getattr(L, n).argtypes = [vp]
    L.atm_graph_launch.argtypes = [vp, ip]
    L.atm_set_coordinate.argtypes = [vp, dp, dp]
    for n in ["atm_put_field", "atm_get_field"]:
        getattr(L, n).argtypes = [vp, ip, dp]
    for n in ["atm_put_field_device", "atm_get_field_device"]:
        getattr(L, n).argtypes = [vp, ip, vp]
    for n in ["atm_cgrid_step", "atm_dgrid_step", "atm_radiation", "atm_apply_physics", "atm_accumulate_output"]:
        getattr(L, n).argtypes = [vp, C.c_double]
    L.atm_halo_exchange.argtypes = [vp, ip]
    L.atm_output_mean.argtypes = [vp, dp]
    L.atm_check_stability.argtypes = [vp, C.POINTER(C.c_int)]
    L.atm_check_stability_begin.argtypes = [vp]
    L.atm_check_stability_end.argtypes = [vp, C.POINTER(C.c_int)]
    L.atm_set_path.argtypes = [vp, ip]
    L.atm_get_path.argtypes = [vp, C.POINTER(C.c_int)]
    L.atm_timer_enable.argtypes = [vp, ip]
    L.atm_timer_select.argtypes = [vp, C.c_uint]
    L.atm_poison.argtypes = [vp, ip, C.c_double]
    L.atm_timer_read.argtypes = [vp, ip, C.POINTER(C.c_int64), C.POINTER(C.c_double)]
    L.atm_lin_init.argtypes = [vp]
    for n in ["atm_tl_step", "atm_ad_step"]:
        getattr(L, n).argtypes = [vp, C.c_double, ip]
    L.atm_autotune.argtypes = [vp, ip]
    L.atm_get_tuning.argtypes = [vp, C.POINTER(C.c_int)]
    L.atm_set_tuning.argtypes = [vp, ip]
    L.atm_get_tiles.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.atm_set_tiles.argtypes = [vp, ip, ip]
    L.atm_ensemble_create.argtypes = [C.POINTER(vp), ip, C.POINTER(vp)]
    L.atm_ensemble_graph_launch.argtypes = [vp, ip]
    L.atm_ensemble_radiation.argtypes = [vp, C.c_double]
    for n in ["atm_ensemble_destroy", "atm_ensemble_cgrid_halo", "atm_ensemble_dgrid_halo", "atm_ensemble_remap",
              "atm_ensemble_graph_begin", "atm_ensemble_graph_end"]:
        getattr(L, n).argtypes = [vp]
    for n in ["atm_ensemble_cgrid_step", "atm_ensemble_dgrid_step"]:
        getattr(L, n).argtypes = [vp, C.c_double]
    _LIB = L
    return L


def stream_create() -> int:
    """A HIP stream of default priority from the library's runtime (handle as int)."""
    L = load_library()
    s = C.c_void_p()
    rc = L.atm_stream_create(C.byref(s))
    if rc:
        raise AtmError(rc, "atm_stream_create")
    return s.value


def stream_destroy(handle: int) -> None:
    rc = load_library().atm_stream_destroy(C.c_void_p(handle))
    if rc:
        raise AtmError(rc, "atm_stream_destroy")


class BatchedSteps:
    """The batched ensemble steps (atm_ensemble_*) over Models that run the
    fused path in one precision: each call (the dynamics steps, their halo
    fills, the remap, radiation) does for every member what the Model's own
    call does,
    with one launch per stage for all of them on the first member's stream
    (joined to the others' streams by events).  Create it after the members'
    initialisation and close it before the members."""

    def __init__(self, models):
        self._L = load_library()
        self.models = list(models)
        arr = (C.c_void_p * len(self.models))(*[m._h for m in self.models])
        h = C.c_void_p()
        rc = self._L.atm_ensemble_create(arr, len(self.models), C.byref(h))
        if rc:
            self._raise(rc)
        self._h = h

    def _raise(self, rc):
        msgs = [m._L.atm_last_error(m._h).decode() for m in self.models if m._h]
        raise AtmError(rc, "; ".join(x for x in msgs if x) or "atm_ensemble")

    def _ck(self, rc):
        if rc:
            self._raise(rc)

    def cgrid_dynamics_step(self, dt):
        self._ck(self._L.atm_ensemble_cgrid_step(self._h, dt))

    def cgrid_halo_exchange(self):
        self._ck(self._L.atm_ensemble_cgrid_halo(self._h))

    def dgrid_dynamics_step(self, dt):
        self._ck(self._L.atm_ensemble_dgrid_step(self._h, dt))

    def dgrid_halo_exchange(self):
        self._ck(self._L.atm_ensemble_dgrid_halo(self._h))

    def perform_vertical_remapping(self):
        self._ck(self._L.atm_ensemble_remap(self._h))

    def calculate_radiative_heating(self, time):
        self._ck(self._L.atm_ensemble_radiation(self._h, float(time)))

    def graph_begin(self):
        """Record the following batched and member step calls into one HIP
        graph instead of running them."""
        self._ck(self._L.atm_ensemble_graph_begin(self._h))

    def graph_end(self):
        self._ck(self._L.atm_ensemble_graph_end(self._h))

    def graph_launch(self, times=1):
        self._ck(self._L.atm_ensemble_graph_launch(self._h, int(times)))

    def close(self):
        if getattr(self, "_h", None):
            self._L.atm_ensemble_destroy(self._h)
            self._h = None


def _dptr(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Model:
    """One model instance on the current GPU (FP64 drop-in path)."""

    def __init__(self, nx, ny, nlay, dx, dy=None, gravity=9.807, coriolis=1e-5,
                 divergence_damping=0.15, vorticity_damping=0.005, precision="fp64"):
        self.nx, self.ny, self.nlay = nx, ny, nlay
        self.precision = precision
        self.NI, self.NJ = nx + 2 * HALO, ny + 2 * HALO
        self._L = load_library()
        h = C.c_void_p()
        grid = Grid(nx, ny, nlay, dx, dx if dy is None else dy)
        prm = Params(gravity, coriolis, divergence_damping, vorticity_damping)
        rc = self._L.atm_create(C.byref(grid), C.byref(prm), PRECISIONS[precision], C.byref(h))
        if rc != 0:
            raise AtmError(rc, f"atm_create failed (no GPU, bad sizes, or {precision} not available at this size)")
        self._h = h
        self._ak = self._bk = None
        self._gzs = None

    def close(self):
        if getattr(self, "_h", None):
            self._L.atm_destroy(self._h)
            self._h = None

    __del__ = close

    def _ck(self, rc):
        if rc != 0:
            raise AtmError(rc, self._L.atm_last_error(self._h).decode())

    def nk(self, name):
        return {"ts": 1, "radius": 1, "gz": self.nlay + 1, "net_flux": self.nlay + 1,
                "plev": self.nlay + 1, "pkap": self.nlay + 1}.get(name, self.nlay)

    # ---- field transfer (halo'd layout) ----
    def put_full(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(self.nk(name), self.NJ, self.NI)
        self._ck(self._L.atm_put_field(self._h, FIELDS[name], _dptr(a)))

    def get_full(self, name):
        a = np.empty((self.nk(name), self.NJ, self.NI))
        self._ck(self._L.atm_get_field(self._h, FIELDS[name], _dptr(a)))
        return a

    def put(self, name, arr):
        """Interior values (nk, ny, nx) or (ny, nx); the halo is zeroed."""
        arr = np.asarray(arr, dtype=np.float64)
        if name in ("ak", "bk"):
            if name == "ak":
                self._ak = np.ascontiguousarray(arr)
            else:
                self._bk = np.ascontiguousarray(arr)
            if self._ak is not None and self._bk is not None:
                self._ck(self._L.atm_set_coordinate(self._h, _dptr(self._ak), _dptr(self._bk)))
            return
        if name == "gzs":
            full = np.zeros((self.nlay + 1, self.NJ, self.NI))
            full[0, HALO:HALO + self.ny, HALO:HALO + self.nx] = arr
            self.put_full("gz", full)
            return
        full = np.zeros((self.nk(name), self.NJ, self.NI))
        full[..., HALO:HALO + self.ny, HALO:HALO + self.nx] = arr
        self.put_full(name, full)

    def get(self, name):
        a = self.get_full(name)[..., HALO:HALO + self.ny, HALO:HALO + self.nx]
        return np.ascontiguousarray(a[0] if self.nk(name) == 1 else a)

    # ---- reference module procedures ----
    def init_prognostic_fields(self):
        self._ck(self._L.atm_init_prognostic_fields(self._h))

    def initial_halo_exchange(self):
        self._ck(self._L.atm_initial_halo_exchange(self._h))

    def init_sw_dyn(self):
        self._ck(self._L.atm_init_sw_dyn(self._h))

    def allocate_radiation_arrays(self):
        self._ck(self._L.atm_radiation_init(self._h))

    def allocate_remapping_arrays(self):
        self._ck(self._L.atm_remap_init(self._h))

    def allocate_physics_arrays(self):
        self._ck(self._L.atm_physics_init(self._h))

    def calculate_physics_tendencies(self):
        self._ck(self._L.atm_physics_tendencies(self._h))

    def apply_physics_tendencies(self, dt):
        self._ck(self._L.atm_apply_physics(self._h, dt))

    def physics_halo_exchange(self):
        self._ck(self._L.atm_physics_halo(self._h))

    def init_output(self):
        self._ck(self._L.atm_output_init(self._h))

    def accumulate_output(self, dt):
        self._ck(self._L.atm_accumulate_output(self._h, dt))

    def output_mean(self):
        """write_output's time means as (6, nlay+1, ny, nx): dp, pt, u, v, ts, nflx."""
        out = np.zeros((6, self.nlay + 1, self.ny, self.nx))
        self._ck(self._L.atm_output_mean(self._h, _dptr(out)))
        return out

    def calculate_radiative_heating(self, time):
        self._ck(self._L.atm_radiation(self._h, time))

    def radiation_halo_exchange(self):
        self._ck(self._L.atm_radiation_halo(self._h))

    def cgrid_dynamics_step(self, dt):
        self._ck(self._L.atm_cgrid_step(self._h, dt))

    def cgrid_halo_exchange(self):
        self._ck(self._L.atm_cgrid_halo(self._h))

    def dgrid_dynamics_step(self, dt):
        self._ck(self._L.atm_dgrid_step(self._h, dt))

    def dgrid_halo_exchange(self):
        self._ck(self._L.atm_dgrid_halo(self._h))

    def perform_vertical_remapping(self):
        self._ck(self._L.atm_remap(self._h))

    def halo_exchange(self, name):
        self._ck(self._L.atm_halo_exchange(self._h, FIELDS[name]))

    def check_stability(self):
        s = C.c_int(0)
        self._ck(self._L.atm_check_stability(self._h, C.byref(s)))
        return bool(s.value)

    def check_stability_begin(self):
        """Queue a stability check behind the calls issued so far (returns at once)."""
        self._ck(self._L.atm_check_stability_begin(self._h))

    def check_stability_end(self):
        """The result of the check queued by check_stability_begin (waits for it only)."""
        s = C.c_int(0)
        self._ck(self._L.atm_check_stability_end(self._h, C.byref(s)))
        return bool(s.value)

    def sync(self):
        self._ck(self._L.atm_sync(self._h))

    @property
    def path(self):
        p = C.c_int(0)
        self._ck(self._L.atm_get_path(self._h, C.byref(p)))
        return {v: k for k, v in PATHS.items()}[p.value]

    def set_path(self, name):
        self._ck(self._L.atm_set_path(self._h, PATHS[name]))

    def autotune(self, reps=5):
        """Time the fused path's candidate launch shapes inside trial steps
        on this grid and keep the fastest (results do not depend on the
        choice); returns the chosen corner-kernel levels per workgroup (the
        chosen tile rows: `tiles`)."""
        self._ck(self._L.atm_autotune(self._h, int(reps)))
        return self.tuning

    @property
    def tuning(self):
        v = C.c_int(0)
        self._ck(self._L.atm_get_tuning(self._h, C.byref(v)))
        return v.value

    def set_tuning(self, corner_levels):
        self._ck(self._L.atm_set_tuning(self._h, int(corner_levels)))

    @property
    def tiles(self):
        """Rows of the dp/pt tiles and of the wind tiles (32 cells wide)."""
        a, b = C.c_int(0), C.c_int(0)
        self._ck(self._L.atm_get_tiles(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def set_tiles(self, pair_rows, wind_rows):
        self._ck(self._L.atm_set_tiles(self._h, int(pair_rows), int(wind_rows)))

    def timer_enable(self, on=True):
        self._ck(self._L.atm_timer_enable(self._h, int(bool(on))))

    def timer_select(self, names=None):
        """Time only these kernel groups (keys of TIMERS; None: all)."""
        mask = (1 << len(TIMERS)) - 1 if names is None else sum(1 << TIMERS[n] for n in set(names))
        self._ck(self._L.atm_timer_select(self._h, C.c_uint(mask)))

    def poison(self, padding=True, outer_halo=True, value=float("nan")):
        """Write `value` into memory no kernel may read for a result inside
        the domain: the row padding and buffer tails, and/or the halo cells
        farther than 4 from the interior (atm_poison)."""
        self._ck(self._L.atm_poison(self._h, (1 if padding else 0) | (2 if outer_halo else 0), value))

    def timer_read(self, name):
        """(launches, total milliseconds) of one kernel group since enabling."""
        n, ms = C.c_int64(0), C.c_double(0)
        self._ck(self._L.atm_timer_read(self._h, TIMERS[name], C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def graph_begin(self):
        """Record the following step calls into a HIP graph instead of running them."""
        self._ck(self._L.atm_graph_begin(self._h))

    def graph_end(self):
        self._ck(self._L.atm_graph_end(self._h))

    def graph_launch(self, times=1):
        self._ck(self._L.atm_graph_launch(self._h, int(times)))

    def graph_free(self):
        self._ck(self._L.atm_graph_free(self._h))

    def set_stream(self, stream_handle):
        self._ck(self._L.atm_set_stream(self._h, C.c_void_p(stream_handle)))

    # ---- device-resident transfers (halo'd layout, stream-ordered) ----
    def put_device(self, name, ptr):
        """Copy a device array (nk, ny+14, nx+14) at address ptr into a field."""
        self._ck(self._L.atm_put_field_device(self._h, FIELDS[name], C.c_void_p(ptr)))

    def get_device(self, name, ptr):
        """Copy a field into the device array at address ptr."""
        self._ck(self._L.atm_get_field_device(self._h, FIELDS[name], C.c_void_p(ptr)))

    # ---- tangent-linear and adjoint of one step (no reference counterpart) ----
    def lin_init(self):
        self._ck(self._L.atm_lin_init(self._h))

    def tl_step(self, dt, remap=False):
        """Advance the state one step and the tangent fields (tl_*) alongside."""
        self._ck(self._L.atm_tl_step(self._h, dt, int(bool(remap))))

    def ad_step(self, dt, remap=False):
        """Replace the adjoint fields (ad_*) of the step's end state by those
        of its start state, linearised about the current state."""
        self._ck(self._L.atm_ad_step(self._h, dt, int(bool(remap))))
