"""The model's files in the netCDF classic formats, through libatmos_io.so
(include/atm_nc.h): the initial file, the time-mean output and the restart.

The reference reads and writes these with netCDF-Fortran
(src/mod_netcdf.f90, src/mod_input.f90:21-252, src/mod_output.f90:37-261).
This module keeps its file contents and checks:

* `read_initial(path)`: read_initial_file + init_prognostic_fields
  (mod_input.f90:21-148, 150-246): the axes xc, yc, xf, yf, lay, lev and t
  must be present with consistent sizes, the centre spacing uniform to
  1e-3; dx = xc(2) - xc(1), Lx = max(xc) - min(xc) + dx, t_initial = t(1);
  dp, pt, u, v, gzs, ts, ak, bk read (at the last record) in double.
  Returns the dict of `atmos.ic.make_initial_state` plus t_initial, Lx, Ly.
* `write_initial(path, ic)`: the file read_initial_file expects, with the
  record dimension t first as the classic formats require.
* `OutputFile`: output/output.nc of init_output / write_output
  (mod_output.f90:37-92, 144-177): float32 means of dp, pt, u, v, ts, nflx
  on (t, [lay|lev,] y, x), one record per write.
* `write_restart(path, ...)`: write_restart_file (mod_output.f90:179-230),
  plus the ak, bk and lev it lacks (Q13), so read_initial accepts it; in
  double by default so a restart continues bit for bit.

Errors from the library raise NcError with its message; read_initial's
checks raise ValueError with the reference's wording.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# ATM_IO_LIB names another build of the library (the sanitizer build of the
# host tests)
LIB_PATH = os.environ.get("ATM_IO_LIB") or os.path.join(os.path.dirname(_HERE), "lib", "libatmos_io.so")

NOWRITE, WRITE, CLOBBER, NOCLOBBER = 0, 1, 0, 4
CDF1, CDF2, CDF5 = 0, 0x0200, 0x0020
UNLIMITED, GLOBAL = 0, -1
BYTE, CHAR, SHORT, INT, FLOAT, DOUBLE = 1, 2, 3, 4, 5, 6
UBYTE, USHORT, UINT, INT64, UINT64 = 7, 8, 9, 10, 11
MAX_NAME = 256
EXPORTS = ["atm_nc_open", "atm_nc_create", "atm_nc_enddef", "atm_nc_redef", "atm_nc_sync",
           "atm_nc_close", "atm_nc_inq", "atm_nc_inq_format", "atm_nc_inq_dim", "atm_nc_inq_dimid",
           "atm_nc_inq_var", "atm_nc_inq_varid", "atm_nc_def_dim", "atm_nc_def_var",
           "atm_nc_inq_att", "atm_nc_inq_attname", "atm_nc_get_att_double", "atm_nc_get_att_text",
           "atm_nc_put_att_double", "atm_nc_put_att_text", "atm_nc_get_vara_double",
           "atm_nc_get_varm_double", "atm_nc_put_vara_double", "atm_nc_put_varm_double",
           "atm_nc_strerror"]


class NcError(RuntimeError):
    def __init__(self, code: int, msg: str, where: str = ""):
        super().__init__(f"netCDF status {code} in {where}: {msg}")
        self.code = code


_LIB = None


def load_library(path: str = LIB_PATH):
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise RuntimeError(f"I/O library not built: {path} (run __graft_entry__.build())")
    L = C.CDLL(path)
    ip, ipp, sz, szp = C.c_int, C.POINTER(C.c_int), C.c_size_t, C.POINTER(C.c_size_t)
    dp, pdp = C.POINTER(C.c_double), C.POINTER(C.c_ssize_t)
    L.atm_nc_open.argtypes = [C.c_char_p, ip, ipp]
    L.atm_nc_create.argtypes = [C.c_char_p, ip, ipp]
    for n in ["atm_nc_enddef", "atm_nc_redef", "atm_nc_sync", "atm_nc_close"]:
        getattr(L, n).argtypes = [ip]
    L.atm_nc_inq.argtypes = [ip, ipp, ipp, ipp, ipp]
    L.atm_nc_inq_format.argtypes = [ip, ipp]
    L.atm_nc_inq_dim.argtypes = [ip, ip, C.c_char_p, szp]
    L.atm_nc_inq_dimid.argtypes = [ip, C.c_char_p, ipp]
    L.atm_nc_inq_var.argtypes = [ip, ip, C.c_char_p, ipp, ipp, ipp, ipp]
    L.atm_nc_inq_varid.argtypes = [ip, C.c_char_p, ipp]
    L.atm_nc_def_dim.argtypes = [ip, C.c_char_p, sz, ipp]
    L.atm_nc_def_var.argtypes = [ip, C.c_char_p, ip, ip, ipp, ipp]
    L.atm_nc_inq_att.argtypes = [ip, ip, C.c_char_p, ipp, szp]
    L.atm_nc_inq_attname.argtypes = [ip, ip, ip, C.c_char_p]
    L.atm_nc_get_att_double.argtypes = [ip, ip, C.c_char_p, dp]
    L.atm_nc_get_att_text.argtypes = [ip, ip, C.c_char_p, C.c_char_p]
    L.atm_nc_put_att_double.argtypes = [ip, ip, C.c_char_p, ip, sz, dp]
    L.atm_nc_put_att_text.argtypes = [ip, ip, C.c_char_p, sz, C.c_char_p]
    L.atm_nc_get_vara_double.argtypes = [ip, ip, szp, szp, dp]
    L.atm_nc_put_vara_double.argtypes = [ip, ip, szp, szp, dp]
    L.atm_nc_get_varm_double.argtypes = [ip, ip, szp, szp, pdp, dp]
    L.atm_nc_put_varm_double.argtypes = [ip, ip, szp, szp, pdp, dp]
    L.atm_nc_strerror.argtypes = [ip]
    L.atm_nc_strerror.restype = C.c_char_p
    _LIB = L
    return L


def _sizes(vals):
    return (C.c_size_t * max(1, len(vals)))(*[int(v) for v in vals])


class NcFile:
    """One open file: a thin wrapper of the C calls (C dimension order)."""

    def __init__(self, path: str, mode: str = "r", fmt: int = CDF2):
        self.L = load_library()
        self.path = path
        ncid = C.c_int()
        if mode == "r":
            self._check(self.L.atm_nc_open(path.encode(), NOWRITE, C.byref(ncid)), "open")
        elif mode == "a":
            self._check(self.L.atm_nc_open(path.encode(), WRITE, C.byref(ncid)), "open")
        elif mode == "w":
            self._check(self.L.atm_nc_create(path.encode(), CLOBBER | fmt, C.byref(ncid)), "create")
        else:
            raise ValueError(f"mode {mode!r}")
        self.ncid = ncid.value

    def _check(self, rc: int, where: str):
        if rc != 0:
            raise NcError(rc, self.L.atm_nc_strerror(rc).decode(), f"{where} ({self.path})")

    # -- lifetime ---------------------------------------------------------
    def close(self):
        if self.ncid is not None:
            rc = self.L.atm_nc_close(self.ncid)
            self.ncid = None
            self._check(rc, "close")

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def enddef(self):
        self._check(self.L.atm_nc_enddef(self.ncid), "enddef")

    def redef(self):
        self._check(self.L.atm_nc_redef(self.ncid), "redef")

    def sync(self):
        self._check(self.L.atm_nc_sync(self.ncid), "sync")

    # -- inquiry ----------------------------------------------------------
    @property
    def format(self) -> int:
        f = C.c_int()
        self._check(self.L.atm_nc_inq_format(self.ncid, C.byref(f)), "inq_format")
        return f.value

    def inq(self):
        nd, nv, na, ul = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        self._check(self.L.atm_nc_inq(self.ncid, C.byref(nd), C.byref(nv), C.byref(na), C.byref(ul)), "inq")
        return nd.value, nv.value, na.value, ul.value

    @property
    def dimensions(self) -> dict:
        nd, _, _, _ = self.inq()
        out = {}
        for i in range(nd):
            name = C.create_string_buffer(MAX_NAME + 1)
            n = C.c_size_t()
            self._check(self.L.atm_nc_inq_dim(self.ncid, i, name, C.byref(n)), "inq_dim")
            out[name.value.decode()] = n.value
        return out

    @property
    def unlimited(self) -> str | None:
        ul = self.inq()[3]
        if ul < 0:
            return None
        return list(self.dimensions)[ul]

    def var_info(self, name: str):
        vid = self.varid(name)
        nm = C.create_string_buffer(MAX_NAME + 1)
        xt, nd, na = C.c_int(), C.c_int(), C.c_int()
        dims = (C.c_int * 1024)()
        self._check(self.L.atm_nc_inq_var(self.ncid, vid, nm, C.byref(xt), C.byref(nd), dims, C.byref(na)),
                    "inq_var")
        dnames = list(self.dimensions)
        return dict(varid=vid, type=xt.value, dims=[dnames[dims[i]] for i in range(nd.value)], natts=na.value)

    @property
    def variables(self) -> list:
        _, nv, _, _ = self.inq()
        out = []
        for i in range(nv):
            nm = C.create_string_buffer(MAX_NAME + 1)
            self._check(self.L.atm_nc_inq_var(self.ncid, i, nm, None, None, None, None), "inq_var")
            out.append(nm.value.decode())
        return out

    def varid(self, name: str) -> int:
        v = C.c_int()
        self._check(self.L.atm_nc_inq_varid(self.ncid, name.encode(), C.byref(v)), f"inq_varid {name}")
        return v.value

    def has_var(self, name: str) -> bool:
        return name in self.variables

    def shape(self, name: str) -> tuple:
        d = self.dimensions
        return tuple(d[n] for n in self.var_info(name)["dims"])

    def attributes(self, name: str | None = None) -> dict:
        vid = GLOBAL if name is None else self.varid(name)
        n = self.inq()[2] if name is None else self.var_info(name)["natts"]
        out = {}
        for i in range(n):
            an = C.create_string_buffer(MAX_NAME + 1)
            self._check(self.L.atm_nc_inq_attname(self.ncid, vid, i, an), "inq_attname")
            xt, ln = C.c_int(), C.c_size_t()
            self._check(self.L.atm_nc_inq_att(self.ncid, vid, an.value, C.byref(xt), C.byref(ln)), "inq_att")
            if xt.value == CHAR:
                buf = C.create_string_buffer(ln.value + 1)
                self._check(self.L.atm_nc_get_att_text(self.ncid, vid, an.value, buf), "get_att_text")
                out[an.value.decode()] = buf.raw[: ln.value].decode("latin-1")
            else:
                vals = np.zeros(max(1, ln.value))
                self._check(self.L.atm_nc_get_att_double(self.ncid, vid, an.value,
                                                         vals.ctypes.data_as(C.POINTER(C.c_double))),
                            "get_att_double")
                out[an.value.decode()] = vals[: ln.value]
        return out

    # -- definitions -------------------------------------------------------
    def def_dim(self, name: str, size: int) -> int:
        d = C.c_int()
        self._check(self.L.atm_nc_def_dim(self.ncid, name.encode(), int(size), C.byref(d)), f"def_dim {name}")
        return d.value

    def def_var(self, name: str, xtype: int, dims: tuple = ()) -> int:
        names = list(self.dimensions)
        ids = (C.c_int * max(1, len(dims)))(*[names.index(d) for d in dims])
        v = C.c_int()
        self._check(self.L.atm_nc_def_var(self.ncid, name.encode(), xtype, len(dims), ids, C.byref(v)),
                    f"def_var {name}")
        return v.value

    def put_att(self, name: str | None, att: str, value, xtype: int = DOUBLE):
        vid = GLOBAL if name is None else self.varid(name)
        if isinstance(value, str):
            b = value.encode()
            self._check(self.L.atm_nc_put_att_text(self.ncid, vid, att.encode(), len(b), b), "put_att_text")
        else:
            a = np.ascontiguousarray(np.atleast_1d(value), dtype=np.float64)
            self._check(self.L.atm_nc_put_att_double(self.ncid, vid, att.encode(), xtype, a.size,
                                                     a.ctypes.data_as(C.POINTER(C.c_double))), "put_att")

    # -- values ------------------------------------------------------------
    def read(self, name: str, start=None, count=None) -> np.ndarray:
        vid = self.varid(name)
        shp = self.shape(name)
        start = [0] * len(shp) if start is None else list(start)
        count = [s - o for s, o in zip(shp, start)] if count is None else list(count)
        out = np.empty(count if count else (), dtype=np.float64)
        self._check(self.L.atm_nc_get_vara_double(self.ncid, vid, _sizes(start), _sizes(count),
                                                  out.ctypes.data_as(C.POINTER(C.c_double))), f"get {name}")
        return out

    def write(self, name: str, values, start=None):
        vid = self.varid(name)
        a = np.ascontiguousarray(values, dtype=np.float64)
        nd = len(self.var_info(name)["dims"])
        count = list(a.shape)
        if len(count) < nd:  # leading record / singleton dimensions
            count = [1] * (nd - len(count)) + count
        start = [0] * nd if start is None else list(start)
        self._check(self.L.atm_nc_put_vara_double(self.ncid, vid, _sizes(start), _sizes(count),
                                                  a.ctypes.data_as(C.POINTER(C.c_double))), f"put {name}")


# ---------------------------------------------------------------------------
# the model's files
# ---------------------------------------------------------------------------

def _uniform(points: np.ndarray, d: float, tol: float = 1e-3) -> bool:
    # mod_input.f90:118-119, 131-132
    return bool(np.all(np.abs(points[1:] - points[:-1]) - abs(d) < tol))


def read_initial(path: str) -> dict:
    """read_initial_file + init_prognostic_fields (mod_input.f90:21-246)."""
    with NcFile(path) as f:
        dims = f.dimensions
        t = f.unlimited
        sizes = {"x": None, "y": None, "z": None}
        for n in ["xc", "yc", "xf", "yf", "lay", "lev"]:  # 42-101
            if n not in dims or n == t:
                raise ValueError(f"Could not find axis `{n}` in initial condition netCDF.")
            key, size = {"xc": ("x", dims[n]), "xf": ("x", dims[n]), "yc": ("y", dims[n]),
                         "yf": ("y", dims[n]), "lay": ("z", dims[n]), "lev": ("z", dims[n] - 1)}[n]
            if sizes[key] is None:
                sizes[key] = size
            elif sizes[key] != size:
                what = {"x": "x axis", "y": "y axis", "z": "vertical axis"}[key]
                raise ValueError(f"Inconsistent {what} sizes in initial condition netCDF.")
        nx, ny, nlay = sizes["x"], sizes["y"], sizes["z"]
        xc, yc = f.read("xc"), f.read("yc")
        dx, dy = xc[1] - xc[0], yc[1] - yc[0]
        for pts, d, ax in ((xc, dx, "x"), (yc, dy, "y")):
            if not _uniform(pts, d):
                raise ValueError(f"Spacing of {ax}-coordinate points in initial condition netCDF is not uniform.")
        if t is None:
            raise ValueError("Could not find axis `t` in initial condition netCDF.")
        t_initial = float(f.read(t)[0])
        nrec = dims[t]

        def field(name, nk):
            # read_3d_variable / read_2d_variable: the last record, any order
            # of the x/y/lay dimensions (mod_netcdf.f90:520-624)
            if not f.has_var(name):
                raise ValueError(f"Could not find variable `{name}` in initial condition netCDF.")
            info = f.var_info(name)
            start = [nrec - 1 if d == t else 0 for d in info["dims"]]
            count = [1 if d == t else dims[d] for d in info["dims"]]
            a = f.read(name, start, count)
            keep = [d for d in info["dims"] if d != t]
            a = a.reshape([dims[d] for d in keep])
            # target axis of each file dimension: (lay, y, x) or (y, x)
            role = {"x": 2, "xc": 2, "xf": 2, "y": 1, "yc": 1, "yf": 1, "lay": 0}
            if nk == 1:
                role = {k: v - 1 for k, v in role.items() if k != "lay"}
            if any(d not in role for d in keep) or len(keep) != (2 if nk == 1 else 3):
                raise ValueError(f"Variable `{name}` has dimensions {info['dims']}, not x, y"
                                 + ("" if nk == 1 else ", lay") + " (and t).")
            order = [role[d] for d in keep]
            return np.ascontiguousarray(np.transpose(a, np.argsort(order)))

        ic = dict(dp=field("dp", nlay), pt=field("pt", nlay), u=field("u", nlay), v=field("v", nlay),
                  gzs=field("gzs", 1), ts=field("ts", 1), ak=f.read("ak"), bk=f.read("bk"),
                  xc=xc, yc=yc, dx=float(dx), dy=float(dy), t_initial=t_initial,
                  Lx=float(xc.max() - xc.min() + dx), Ly=float(yc.max() - yc.min() + dy),
                  nx=nx, ny=ny, nlay=nlay)
        return ic


def write_initial(path: str, ic: dict, t_initial: float = 0.0, fmt: int = CDF2, xtype: int = DOUBLE):
    """An initial file of `ic` (atmos.ic.make_initial_state's dict) that
    read_initial_file accepts; fields on (t, lay, y, x)."""
    nlay, ny, nx = ic["dp"].shape
    xc, yc, dx, dy = ic["xc"], ic["yc"], ic["dx"], ic["dy"]
    with NcFile(path, "w", fmt) as f:
        f.def_dim("t", UNLIMITED)
        for n, size in (("xc", nx), ("yc", ny), ("xf", nx), ("yf", ny), ("lay", nlay), ("lev", nlay + 1)):
            f.def_dim(n, size)
        f.def_var("t", DOUBLE, ("t",))
        for n in ("xc", "yc", "xf", "yf"):
            f.def_var(n, DOUBLE, (n,))
        f.def_var("lay", INT, ("lay",))
        f.def_var("lev", INT, ("lev",))
        for n, dims in (("dp", ("t", "lay", "yc", "xc")), ("pt", ("t", "lay", "yc", "xc")),
                        ("u", ("t", "lay", "yf", "xc")), ("v", ("t", "lay", "yc", "xf")),
                        ("gzs", ("t", "yc", "xc")), ("ts", ("t", "yc", "xc"))):
            f.def_var(n, xtype, dims)
        f.def_var("ak", DOUBLE, ("lev",))
        f.def_var("bk", DOUBLE, ("lev",))
        f.enddef()
        f.write("t", [t_initial], start=[0])
        f.write("xc", xc)
        f.write("yc", yc)
        f.write("xf", xc - dx / 2)
        f.write("yf", yc - dy / 2)
        f.write("lay", np.arange(1, nlay + 1))
        f.write("lev", np.arange(1, nlay + 2))
        for n, src in (("dp", "dp"), ("pt", "pt"), ("u", "u"), ("v", "v"), ("gzs", "gzs"), ("ts", "ts")):
            f.write(n, ic[src][None])
        f.write("ak", ic["ak"])
        f.write("bk", ic["bk"])


def centre_axis(n: int, L: float, d: float) -> np.ndarray:
    # -.5*L + (i + .5)*d, i = 0 .. n-1 (mod_output.f90:68-72, 193-203)
    return np.array([-.5 * L + (i + .5) * d for i in range(n)])


def face_axis(n: int, L: float, d: float) -> np.ndarray:
    return np.array([-.5 * L + i * d for i in range(n)])


class OutputFile:
    """output/output.nc of init_output / write_output (mod_output.f90:37-92,
    144-177): axes t (record), x, y, lay, lev; float32 variables dp, pt, u,
    v (t, lay, y, x), ts (t, y, x), nflx (t, lev, y, x)."""

    NAMES = ("dp", "pt", "u", "v", "ts", "nflx")

    def __init__(self, path: str, nx: int, ny: int, nlay: int, dx: float, dy: float,
                 Lx: float | None = None, Ly: float | None = None, fmt: int = CDF1):
        Lx = nx * dx if Lx is None else Lx
        Ly = ny * dy if Ly is None else Ly
        self.nlay = nlay
        self.f = NcFile(path, "w", fmt)
        f = self.f
        f.def_dim("t", UNLIMITED)
        f.def_var("t", FLOAT, ("t",))
        for n, size in (("x", nx), ("y", ny), ("lay", nlay), ("lev", nlay + 1)):
            f.def_dim(n, size)
            f.def_var(n, FLOAT if n in ("x", "y") else INT, (n,))
        for n in ("dp", "pt", "u", "v"):
            f.def_var(n, FLOAT, ("t", "lay", "y", "x"))
        f.def_var("ts", FLOAT, ("t", "y", "x"))
        f.def_var("nflx", FLOAT, ("t", "lev", "y", "x"))
        f.enddef()
        f.write("x", centre_axis(nx, Lx, dx))
        f.write("y", centre_axis(ny, Ly, dy))
        f.write("lay", np.arange(1, nlay + 1))
        f.write("lev", np.arange(1, nlay + 2))
        self.nrec = 0

    def write(self, time: float, means: np.ndarray):
        """means: (6, nlay+1, ny, nx) time means in the order dp, pt, u, v,
        ts, nflx (atm_output_mean); ts is level 1."""
        L = self.nlay
        r = self.nrec
        self.f.write("t", [time], start=[r])
        for i, n in enumerate(("dp", "pt", "u", "v")):
            self.f.write(n, means[i, :L][None], start=[r, 0, 0, 0])
        self.f.write("ts", means[4, 0][None], start=[r, 0, 0])
        self.f.write("nflx", means[5][None], start=[r, 0, 0, 0])
        self.nrec += 1

    def close(self):
        self.f.close()


def write_restart(path: str, time: float, dp, pt, ud, vd, gzs, ts, ak, bk, dx: float, dy: float,
                  Lx: float | None = None, Ly: float | None = None, xtype: int = DOUBLE, fmt: int = CDF2):
    """write_restart_file (mod_output.f90:179-230) of interior fields
    ((nlay, ny, nx) and (ny, nx)), plus ak, bk on a lev axis."""
    nlay, ny, nx = dp.shape
    Lx = nx * dx if Lx is None else Lx
    Ly = ny * dy if Ly is None else Ly
    with NcFile(path, "w", fmt) as f:
        f.def_dim("t", UNLIMITED)
        f.def_var("t", xtype, ("t",))
        for n, size in (("xc", nx), ("xf", nx), ("yc", ny), ("yf", ny), ("lay", nlay), ("lev", nlay + 1)):
            f.def_dim(n, size)
            f.def_var(n, INT if n in ("lay", "lev") else xtype, (n,))
        for n, dims in (("dp", ("t", "lay", "yc", "xc")), ("pt", ("t", "lay", "yc", "xc")),
                        ("u", ("t", "lay", "yf", "xc")), ("v", ("t", "lay", "yc", "xf")),
                        ("gzs", ("t", "yc", "xc")), ("ts", ("t", "yc", "xc"))):
            f.def_var(n, xtype, dims)
        f.def_var("ak", DOUBLE, ("lev",))
        f.def_var("bk", DOUBLE, ("lev",))
        f.enddef()
        f.write("t", [time], start=[0])
        f.write("xc", centre_axis(nx, Lx, dx))
        f.write("xf", face_axis(nx, Lx, dx))
        f.write("yc", centre_axis(ny, Ly, dy))
        f.write("yf", face_axis(ny, Ly, dy))
        f.write("lay", np.arange(1, nlay + 1))
        f.write("lev", np.arange(1, nlay + 2))
        for n, a in (("dp", dp), ("pt", pt), ("u", ud), ("v", vd), ("gzs", gzs), ("ts", ts)):
            f.write(n, np.asarray(a)[None])
        f.write("ak", ak)
        f.write("bk", bk)
