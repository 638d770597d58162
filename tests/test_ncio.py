"""The netCDF classic-format I/O (include/atm_nc.h, atmos.ncio) on the CPU.

Pins, none of which restates the library's own code:
  * scipy.io.netcdf_file, an independent implementation of the classic
    formats, reads what libatmos_io writes and libatmos_io reads what scipy
    writes, for every classic type, record and non-record variables,
    attributes and padding;
  * two files assembled byte by byte from the published format
    specification (CDF-1 with a non-record variable, CDF-2 with a lone
    record variable of 2-byte values, whose records are not padded);
  * the reference's initial-file semantics (src/mod_input.f90:21-246) on a
    file laid out the way scripts/ic_generator.py orders its dimensions;
  * the restart round trip that the reference's own restart cannot make
    (Q13): written, read back as an initial file, identical;
  * the module `netcdf` (atmosphere-model_amd/fortran/netcdf.f90), driven
    by tests/fortran/nc_host.f90 through the nf90 call sequences of the
    reference's mod_netcdf.
"""
import os
import struct
import subprocess

import numpy as np
import pytest
from scipy.io import netcdf_file

from atmos import ncio
from atmos.ic import make_initial_state

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NC_HOST = os.path.join(ROOT, "tests", "fortran", "_build", "nc_host")


def test_library_exports():
    L = ncio.load_library()
    for name in ncio.EXPORTS:
        assert hasattr(L, name), name


# ---------------------------------------------------------------------------
# byte-level pins from the format specification
# ---------------------------------------------------------------------------

def _name(s):
    b = s.encode()
    return struct.pack(">i", len(b)) + b + b"\0" * (-len(b) % 4)


def test_cdf1_bytes_match_specification(tmp_path):
    p = str(tmp_path / "a.nc")
    with ncio.NcFile(p, "w", ncio.CDF1) as f:
        f.def_dim("x", 2)
        f.def_var("v", ncio.INT, ("x",))
        f.enddef()
        f.write("v", [1, -2])
    absent = b"\0" * 8
    header = (b"CDF\x01" + struct.pack(">i", 0)
              + struct.pack(">ii", 0x0A, 1) + _name("x") + struct.pack(">i", 2)
              + absent
              + struct.pack(">ii", 0x0B, 1) + _name("v") + struct.pack(">ii", 1, 0) + absent
              + struct.pack(">ii", 4, 8))
    begin = len(header) + 4
    expected = header + struct.pack(">i", begin) + struct.pack(">ii", 1, -2)
    assert open(p, "rb").read() == expected


def test_cdf2_lone_record_variable_bytes(tmp_path):
    p = str(tmp_path / "r.nc")
    with ncio.NcFile(p, "w", ncio.CDF2) as f:
        f.def_dim("t", ncio.UNLIMITED)
        f.def_dim("x", 3)
        f.def_var("r", ncio.SHORT, ("t", "x"))
        f.enddef()
        f.write("r", [[1, 2, 3], [-4, -5, -6]], start=[0, 0])
    absent = b"\0" * 8
    header = (b"CDF\x02" + struct.pack(">i", 2)
              + struct.pack(">ii", 0x0A, 2) + _name("t") + struct.pack(">i", 0) + _name("x")
              + struct.pack(">i", 3)
              + absent
              + struct.pack(">ii", 0x0B, 1) + _name("r") + struct.pack(">iii", 2, 0, 1) + absent
              + struct.pack(">ii", 3, 8))
    begin = len(header) + 8
    # a lone record variable's records are 6 bytes apart, not 8
    expected = header + struct.pack(">q", begin) + struct.pack(">6h", 1, 2, 3, -4, -5, -6)
    assert open(p, "rb").read() == expected


# ---------------------------------------------------------------------------
# against scipy's implementation
# ---------------------------------------------------------------------------

TYPES = [("b", ncio.BYTE), ("h", ncio.SHORT), ("i", ncio.INT), ("f", ncio.FLOAT), ("d", ncio.DOUBLE)]


@pytest.mark.parametrize("version", [1, 2])
def test_reads_scipy_files(tmp_path, version):
    rng = np.random.default_rng(7)
    p = str(tmp_path / f"s{version}.nc")
    vals = {}
    with netcdf_file(p, "w", version=version) as f:
        f.createDimension("t", None)
        f.createDimension("y", 3)
        f.createDimension("x", 5)
        f.history = "written by scipy"
        f.scale = np.float32(0.5)
        for code, _ in TYPES:
            a = (rng.integers(-100, 100, (3, 5)) if code in "bhi" else rng.normal(size=(3, 5))).astype(code)
            v = f.createVariable("s_" + code, code, ("y", "x"))
            v[:] = a
            v.units = "m s-1"
            v.valid_range = np.array([-1, 1], dtype=code)
            vals["s_" + code] = a
            r = (rng.integers(-100, 100, (4, 3)) if code in "bhi" else rng.normal(size=(4, 3))).astype(code)
            w = f.createVariable("r_" + code, code, ("t", "y"))
            w[:] = r
            vals["r_" + code] = r
        c = f.createVariable("c", "c", ("x",))
        c[:] = np.array(list("hello"), dtype="S1")
    with ncio.NcFile(p) as g:
        assert g.format == version
        assert g.dimensions == {"t": 4, "y": 3, "x": 5}
        assert g.unlimited == "t"
        for name, a in vals.items():
            got = g.read(name)
            assert got.shape == a.shape
            assert np.array_equal(got, a.astype(np.float64)), name
        assert g.attributes()["history"] == "written by scipy"
        assert g.attributes()["scale"][0] == 0.5
        at = g.attributes("s_h")
        assert at["units"] == "m s-1" and list(at["valid_range"]) == [-1, 1]
        # hyperslabs of a record variable and of a fixed one
        assert np.array_equal(g.read("r_d", [1, 1], [2, 2]), vals["r_d"][1:3, 1:3])
        assert np.array_equal(g.read("s_i", [2, 1], [1, 3]), vals["s_i"][2:3, 1:4].astype(float))
        with pytest.raises(ncio.NcError) as e:
            g.read("c")
        assert e.value.code == -56  # numeric read of a char variable


@pytest.mark.parametrize("fmt,version", [(ncio.CDF1, 1), (ncio.CDF2, 2)])
def test_scipy_reads_our_files(tmp_path, fmt, version):
    rng = np.random.default_rng(11)
    p = str(tmp_path / "o.nc")
    vals = {}
    with ncio.NcFile(p, "w", fmt) as f:
        f.def_dim("t", ncio.UNLIMITED)
        f.def_dim("y", 4)
        f.def_dim("x", 3)
        for code, xt in TYPES:
            f.def_var("s_" + code, xt, ("y", "x"))
            f.def_var("r_" + code, xt, ("t", "x"))
        f.put_att(None, "title", "atmos test")
        f.put_att("s_d", "scale_factor", [2.0, 3.0])
        f.enddef()
        for code, _ in TYPES:
            a = rng.integers(-100, 100, (4, 3)) if code in "bhi" else rng.normal(size=(4, 3))
            f.write("s_" + code, a)
            r = rng.integers(-100, 100, (5, 3)) if code in "bhi" else rng.normal(size=(5, 3))
            f.write("r_" + code, r, start=[0, 0])
            vals["s_" + code], vals["r_" + code] = a, r
    with netcdf_file(p, "r", mmap=False) as s:
        assert s.version_byte == version
        assert s.dimensions["y"] == 4 and s.dimensions["t"] is None
        assert s.title == b"atmos test"
        assert list(s.variables["s_d"].scale_factor) == [2.0, 3.0]
        for code, _ in TYPES:
            for kind in ("s_", "r_"):
                got = s.variables[kind + code][:]
                ref = vals[kind + code]
                if code == "f":
                    ref = ref.astype(np.float32)
                assert np.array_equal(got, ref.astype(got.dtype)), kind + code


def test_redef_moves_written_values(tmp_path):
    """Definitions added after values are written (the reference's
    create_*_axis / create_variable pattern) move those values."""
    p = str(tmp_path / "m.nc")
    with ncio.NcFile(p, "w", ncio.CDF1) as f:
        f.def_dim("t", ncio.UNLIMITED)
        f.def_var("t", ncio.FLOAT, ("t",))
        f.enddef()
        f.write("t", [12.5], start=[0])
        for n, size in (("x", 6), ("lay", 2)):
            f.redef()
            f.def_dim(n, size)
            f.def_var(n, ncio.DOUBLE, (n,))
            f.enddef()
            f.write(n, np.arange(size) * 1.5)
        f.redef()
        f.def_var("q", ncio.DOUBLE, ("t", "lay", "x"))
        f.put_att("q", "units", "K")
        f.enddef()
        f.write("q", np.arange(12.0).reshape(1, 2, 6), start=[0, 0, 0])
    with netcdf_file(p, "r", mmap=False) as s:
        assert s.variables["t"][:].tolist() == [12.5]
        assert s.variables["x"][:].tolist() == list(np.arange(6) * 1.5)
        assert s.variables["lay"][:].tolist() == [0.0, 1.5]
        assert np.array_equal(s.variables["q"][:], np.arange(12.0).reshape(1, 2, 6))
        assert s.variables["q"].units == b"K"


def test_redef_lone_record_variable_gains_a_neighbour(tmp_path):
    """A lone record variable's unpadded records are re-laid with padding
    when a second record variable is defined."""
    p = str(tmp_path / "l.nc")
    with ncio.NcFile(p, "w") as f:
        f.def_dim("t", ncio.UNLIMITED)
        f.def_dim("x", 3)
        f.def_var("a", ncio.SHORT, ("t", "x"))
        f.enddef()
        f.write("a", [[1, 2, 3], [4, 5, 6]], start=[0, 0])
        f.redef()
        f.def_var("b", ncio.SHORT, ("t", "x"))
        f.enddef()
        f.write("b", [[-1, -2, -3]], start=[1, 0])
    with netcdf_file(p, "r", mmap=False) as s:
        assert s.variables["a"][:].tolist() == [[1, 2, 3], [4, 5, 6]]
        assert s.variables["b"][:].tolist() == [[-32767] * 3, [-1, -2, -3]]


def test_records_added_out_of_order_are_filled(tmp_path):
    p = str(tmp_path / "f.nc")
    with ncio.NcFile(p, "w") as f:
        f.def_dim("t", ncio.UNLIMITED)
        f.def_dim("x", 2)
        f.def_var("a", ncio.DOUBLE, ("t", "x"))
        f.def_var("b", ncio.INT, ("t", "x"))
        f.def_var("c", ncio.FLOAT, ("t",))
        f.enddef()
        f.write("a", [[1.0, 2.0]], start=[2, 0])
    with ncio.NcFile(p) as g:
        assert g.dimensions["t"] == 3
        a, b, c = g.read("a"), g.read("b"), g.read("c")
    fill_d, fill_i, fill_f = 9.9692099683868690e36, -2147483647, float(np.float32(9.9692099683868690e36))
    assert np.array_equal(a[:2], np.full((2, 2), fill_d)) and a[2].tolist() == [1.0, 2.0]
    assert np.array_equal(b, np.full((3, 2), fill_i))
    assert np.array_equal(c, np.full(3, fill_f))
    with netcdf_file(p, "r", mmap=False) as s:  # scipy sees the same records
        assert s.variables["a"][:].shape == (3, 2)


def test_fill_value_attribute(tmp_path):
    p = str(tmp_path / "fv.nc")
    with ncio.NcFile(p, "w") as f:
        f.def_dim("x", 4)
        f.def_var("a", ncio.SHORT, ("x",))
        f.put_att("a", "_FillValue", [-99], ncio.SHORT)
        f.enddef()
        f.write("a", [7], start=[2])
    with ncio.NcFile(p) as g:
        assert g.read("a").tolist() == [-99, -99, 7, -99]


def test_mapped_transfer(tmp_path):
    """varm: element (i0, i1) of the block at values[i0*imap0 + i1*imap1]."""
    import ctypes as C
    p = str(tmp_path / "v.nc")
    a = np.arange(20.0).reshape(4, 5)
    with ncio.NcFile(p, "w") as f:
        f.def_dim("y", 4)
        f.def_dim("x", 5)
        f.def_var("a", ncio.DOUBLE, ("y", "x"))
        f.enddef()
        f.write("a", a)
        # read the 3x4 block at (1, 1) transposed into a (4, 3) buffer
        out = np.zeros((4, 3))
        start, count = ncio._sizes([1, 1]), ncio._sizes([3, 4])
        imap = (C.c_ssize_t * 2)(1, 3)
        rc = f.L.atm_nc_get_varm_double(f.ncid, f.varid("a"), start, count, imap,
                                        out.ctypes.data_as(C.POINTER(C.c_double)))
        assert rc == 0
        assert np.array_equal(out, a[1:4, 1:5].T)
        # write it back transposed into a fresh variable region
        rc = f.L.atm_nc_put_varm_double(f.ncid, f.varid("a"), ncio._sizes([0, 0]), count, imap,
                                        (out + 100).ctypes.data_as(C.POINTER(C.c_double)))
        assert rc == 0
        assert np.array_equal(f.read("a", [0, 0], [3, 4]), a[1:4, 1:5] + 100)


def test_cdf5_round_trip(tmp_path):
    p = str(tmp_path / "c5.nc")
    with ncio.NcFile(p, "w", ncio.CDF5) as f:
        f.def_dim("t", ncio.UNLIMITED)
        f.def_dim("x", 3)
        for n, xt in (("u8", ncio.UBYTE), ("u16", ncio.USHORT), ("u32", ncio.UINT), ("i64", ncio.INT64),
                      ("u64", ncio.UINT64), ("d", ncio.DOUBLE)):
            f.def_var(n, xt, ("t", "x"))
        f.enddef()
        vals = {"u8": [0, 200, 255], "u16": [0, 60000, 65535], "u32": [0, 4e9, 4294967295],
                "i64": [-2.0 ** 62, 0, 2.0 ** 62], "u64": [0, 1, 2.0 ** 63], "d": [np.pi, -np.e, 1e300]}
        for n, v in vals.items():
            f.write(n, [v], start=[0, 0])
    raw = open(p, "rb").read(12)
    assert raw[:4] == b"CDF\x05" and struct.unpack(">q", raw[4:12])[0] == 1  # 64-bit record count
    with ncio.NcFile(p) as g:
        assert g.format == 5
        for n, v in vals.items():
            assert g.read(n)[0].tolist() == [float(x) for x in v], n


def test_errors(tmp_path):
    p = str(tmp_path / "e.nc")
    with ncio.NcFile(p, "w") as f:
        f.def_dim("t", ncio.UNLIMITED)
        f.def_dim("x", 2)
        with pytest.raises(ncio.NcError) as e:
            f.def_dim("x", 3)
        assert e.value.code == -42  # name in use
        with pytest.raises(ncio.NcError) as e:
            f.def_dim("t2", ncio.UNLIMITED)
        assert e.value.code == -54  # second record dimension
        with pytest.raises(ncio.NcError) as e:
            f.def_var("bad", ncio.DOUBLE, ("x", "t"))
        assert e.value.code == -47  # record dimension not first
        f.def_var("f", ncio.FLOAT, ("x",))
        f.def_var("s", ncio.SHORT, ("x",))
        with pytest.raises(ncio.NcError) as e:
            f.write("f", [1.0, 2.0])
        assert e.value.code == -39  # in define mode
        f.enddef()
        with pytest.raises(ncio.NcError) as e:
            f.def_dim("y", 2)
        assert e.value.code == -38  # not in define mode
        with pytest.raises(ncio.NcError) as e:
            f.write("f", [1.0, 2.0, 3.0])
        assert e.value.code == -57  # start + count past the end
        with pytest.raises(ncio.NcError) as e:
            f.write("f", [1.0], start=[3])
        assert e.value.code == -40  # start past the end
        with pytest.raises(ncio.NcError) as e:
            f.write("f", [1e39, 1.0])
        assert e.value.code == -60  # out of float range, still written
        with pytest.raises(ncio.NcError) as e:
            f.write("s", [40000, 1])
        assert e.value.code == -60
        assert f.read("s")[1] == 1
        with pytest.raises(ncio.NcError) as e:
            f.read("nope")
        assert e.value.code == -49
    with ncio.NcFile(p) as g:
        assert g.read("f")[0] == np.inf
        with pytest.raises(ncio.NcError) as e:
            g.write("f", [1.0, 2.0])
        assert e.value.code == -37  # read-only
    h = str(tmp_path / "h5.nc")
    with open(h, "wb") as fh:
        fh.write(b"\x89HDF\r\n\x1a\n" + b"\0" * 64)
    with pytest.raises(ncio.NcError) as e:
        ncio.NcFile(h)
    assert e.value.code == -51 and "HDF5" in str(e.value)
    t = str(tmp_path / "trunc.nc")
    open(t, "wb").write(open(p, "rb").read()[:20])
    with pytest.raises(ncio.NcError) as e:
        ncio.NcFile(t)
    assert e.value.code == -51


# ---------------------------------------------------------------------------
# the model's files
# ---------------------------------------------------------------------------

IC_KEYS = ("dp", "pt", "u", "v", "gzs", "ts", "ak", "bk", "xc", "yc")


def test_initial_file_round_trip(tmp_path):
    ic = make_initial_state(24, 16, 6, 50e3)
    p = str(tmp_path / "initial.nc")
    ncio.write_initial(p, ic, t_initial=7200.0)
    r = ncio.read_initial(p)
    for k in IC_KEYS:
        assert np.array_equal(r[k], ic[k]), k
    assert (r["nx"], r["ny"], r["nlay"]) == (24, 16, 6)
    assert r["t_initial"] == 7200.0
    assert r["dx"] == ic["xc"][1] - ic["xc"][0]  # mod_input.f90:115
    assert r["Lx"] == ic["xc"].max() - ic["xc"].min() + r["dx"]  # 127


def _ic_generator_layout(path, ic, t=0.0, order=("yc", "xc", "lay")):
    """A classic file with the variables scripts/ic_generator.py writes
    (initial.nc), in its dimension order after the record dimension."""
    nlay, ny, nx = ic["dp"].shape
    with netcdf_file(path, "w", version=2) as f:
        f.createDimension("t", None)
        for n, s in (("xc", nx), ("yc", ny), ("xf", nx), ("yf", ny), ("lay", nlay), ("lev", nlay + 1)):
            f.createDimension(n, s)
        f.createVariable("t", "d", ("t",))[:] = [t]
        f.createVariable("xc", "d", ("xc",))[:] = ic["xc"]
        f.createVariable("yc", "d", ("yc",))[:] = ic["yc"]
        f.createVariable("xf", "d", ("xf",))[:] = ic["xc"] - ic["dx"] / 2
        f.createVariable("yf", "d", ("yf",))[:] = ic["yc"] - ic["dy"] / 2
        f.createVariable("lay", "i", ("lay",))[:] = np.arange(1, nlay + 1)
        f.createVariable("lev", "i", ("lev",))[:] = np.arange(1, nlay + 2)
        to = {"lay": 0, "yc": 1, "yf": 1, "xc": 2, "xf": 2}
        for n, src, dims in (("dp", "dp", order), ("pt", "pt", order),
                             ("u", "u", tuple("xf" if d == "xc" else d for d in order)),
                             ("v", "v", tuple("yf" if d == "yc" else d for d in order))):
            a = np.transpose(ic[src], [to[d] for d in dims])
            f.createVariable(n, "d", ("t",) + dims)[:] = a[None]
        for n in ("gzs", "ts"):
            f.createVariable(n, "d", ("t", "yc", "xc"))[:] = ic[n][None]
        f.createVariable("ak", "d", ("lev",))[:] = ic["ak"]
        f.createVariable("bk", "d", ("lev",))[:] = ic["bk"]


@pytest.mark.parametrize("order", [("yc", "xc", "lay"), ("lay", "yc", "xc"), ("xc", "lay", "yc")])
def test_read_initial_any_dimension_order(tmp_path, order):
    ic = make_initial_state(12, 8, 4, 100e3)
    p = str(tmp_path / "ig.nc")
    _ic_generator_layout(p, ic, 60.0, order)
    r = ncio.read_initial(p)
    for k in IC_KEYS:
        assert np.array_equal(r[k], ic[k]), k
    assert r["t_initial"] == 60.0


def test_read_initial_checks(tmp_path):
    ic = make_initial_state(12, 8, 4, 100e3)
    p = str(tmp_path / "bad.nc")
    # a missing axis
    with netcdf_file(p, "w", version=2) as f:
        f.createDimension("t", None)
        f.createDimension("xc", 12)
    with pytest.raises(ValueError, match="Could not find axis `yc`"):
        ncio.read_initial(p)
    # inconsistent x sizes
    q = str(tmp_path / "bad2.nc")
    with netcdf_file(q, "w", version=2) as f:
        f.createDimension("t", None)
        for n, s in (("xc", 12), ("yc", 8), ("xf", 11), ("yf", 8), ("lay", 4), ("lev", 5)):
            f.createDimension(n, s)
    with pytest.raises(ValueError, match="Inconsistent x axis sizes"):
        ncio.read_initial(q)
    # non-uniform spacing (beyond the 1e-3 tolerance)
    ic2 = dict(ic)
    ic2["xc"] = ic["xc"].copy()
    ic2["xc"][5] += 1.0
    _ic_generator_layout(p, ic2)
    with pytest.raises(ValueError, match="x-coordinate points .* not uniform"):
        ncio.read_initial(p)
    ic2["xc"] = ic["xc"].copy()
    ic2["xc"][5] += 1e-4  # within tolerance
    _ic_generator_layout(p, ic2)
    assert ncio.read_initial(p)["nx"] == 12


def test_restart_is_readable_as_initial_file(tmp_path):
    """write_restart adds ak, bk and lev (missing from the reference's
    restart, Q13); in double it restores the state exactly, in float it is
    the float32 rounding of it."""
    ic = make_initial_state(16, 8, 5, 100e3)
    for xtype, cast in ((ncio.DOUBLE, np.float64), (ncio.FLOAT, np.float32)):
        p = str(tmp_path / f"restart{xtype}.nc")
        ncio.write_restart(p, 4321.0, ic["dp"], ic["pt"], ic["u"], ic["v"], ic["gzs"], ic["ts"], ic["ak"],
                           ic["bk"], ic["dx"], ic["dy"], xtype=xtype)
        r = ncio.read_initial(p)
        for k in ("dp", "pt", "u", "v", "gzs", "ts"):
            assert np.array_equal(r[k], ic[k].astype(cast).astype(np.float64)), k
        assert np.array_equal(r["ak"], ic["ak"]) and np.array_equal(r["bk"], ic["bk"])
        assert r["t_initial"] == 4321.0
        # centre axis as write_restart_file builds it (mod_output.f90:193-195)
        assert np.array_equal(r["xc"], ncio.centre_axis(16, 16 * 100e3, 100e3).astype(cast))


def test_restart_continues_oracle_run_bit_for_bit(tmp_path):
    """10 steps of the oracle equal 5 steps, a restart file, and 5 steps
    from that file as the initial state (remaps at 300 s and 600 s)."""
    import oracle as O
    from atmos import driver
    from atmos.config import Config

    def cfg(t0=0.0):
        return Config(dt_max=60.0, t_initial=t0, t_final=1e12, dt_output=1e12, dt_radiation=1e12,
                      dt_physics=1e12, dt_remap=300.0, radiation_on=False, physics_on=False)

    def run(b, c, st, n):
        t = c.t_initial
        for _ in range(n):
            t += driver.step(b, c, st, t)
        return t

    nx, ny, nlay = 32, 16, 6
    p = str(tmp_path / "initial.nc")
    ncio.write_initial(p, make_initial_state(nx, ny, nlay, 100e3))
    ic = ncio.read_initial(p)
    a = O.Oracle(nx, ny, nlay, ic["dx"], ic["dy"])
    c = cfg()
    run(a, c, driver.init_model(a, ic, c), 10)
    b = O.Oracle(nx, ny, nlay, ic["dx"], ic["dy"])
    c = cfg()
    t = run(b, c, driver.init_model(b, ic, c), 5)
    r = str(tmp_path / "restart.nc")
    ncio.write_restart(r, t, b.interior(O.DP), b.interior(O.PT), b.interior(O.UD), b.interior(O.VD),
                       b.interior(O.GZ)[0], b.interior(O.TS), ic["ak"], ic["bk"], ic["dx"], ic["dy"])
    ic2 = ncio.read_initial(r)
    d = O.Oracle(nx, ny, nlay, ic2["dx"], ic2["dy"])
    c = cfg(ic2["t_initial"])
    run(d, c, driver.init_model(d, ic2, c), 5)
    for f in (O.DP, O.PT, O.UD, O.VD):
        assert np.array_equal(d.interior(f), a.interior(f))


def test_output_file_layout(tmp_path):
    nx, ny, nlay = 8, 4, 3
    p = str(tmp_path / "output.nc")
    rng = np.random.default_rng(3)
    means = [rng.normal(size=(6, nlay + 1, ny, nx)) * 1e3 for _ in range(2)]
    out = ncio.OutputFile(p, nx, ny, nlay, 1e3, 2e3)
    out.write(100.0, means[0])
    out.write(200.0, means[1])
    out.close()
    with netcdf_file(p, "r", mmap=False) as s:
        assert s.version_byte == 1
        assert s.variables["t"][:].tolist() == [100.0, 200.0]
        assert s.variables["x"][:].tolist() == [-4e3 + (i + .5) * 1e3 for i in range(nx)]
        assert s.variables["lev"][:].tolist() == [1, 2, 3, 4]
        for r in range(2):
            for i, n in enumerate(("dp", "pt", "u", "v")):
                assert s.variables[n].dimensions == ("t", "lay", "y", "x")
                assert np.array_equal(s.variables[n][r], means[r][i, :nlay].astype(np.float32))
            assert np.array_equal(s.variables["ts"][r], means[r][4, 0].astype(np.float32))
            assert np.array_equal(s.variables["nflx"][r], means[r][5].astype(np.float32))


# ---------------------------------------------------------------------------
# the Fortran module `netcdf`
# ---------------------------------------------------------------------------

needs_nc_host = pytest.mark.skipif(not os.path.exists(NC_HOST), reason="nc_host not built (needs flang)")


@needs_nc_host
def test_fortran_reads_initial_file(tmp_path):
    nx, ny, nlay = 12, 8, 4
    ic = make_initial_state(nx, ny, nlay, 100e3)
    for k, writer in enumerate((lambda p: ncio.write_initial(p, ic, t_initial=90.0),
                                lambda p: _ic_generator_layout(p, ic, 90.0))):
        p, b = str(tmp_path / f"i{k}.nc"), str(tmp_path / f"i{k}.bin")
        writer(p)
        subprocess.run([NC_HOST, "read", p, str(nx), str(ny), str(nlay), b], check=True, timeout=60)
        raw = np.fromfile(b, dtype="<f8")
        o, n3, n2 = 0, nx * ny * nlay, nx * ny
        for name in ("dp", "pt", "u", "v"):
            assert np.array_equal(raw[o:o + n3].reshape(nlay, ny, nx), ic[name]), name
            o += n3
        for name in ("gzs", "ts"):
            assert np.array_equal(raw[o:o + n2].reshape(ny, nx), ic[name]), name
            o += n2
        assert np.array_equal(raw[o:o + nlay + 1], ic["ak"])
        assert np.array_equal(raw[o + nlay + 1:o + 2 * nlay + 2], ic["bk"])
        assert np.array_equal(raw[o + 2 * nlay + 2:o + 2 * nlay + 2 + nx], ic["xc"])
        assert raw[-1] == 90.0


@needs_nc_host
def test_fortran_writes_like_mod_netcdf(tmp_path):
    nx, ny, nlay = 5, 4, 3
    p = str(tmp_path / "w.nc")
    subprocess.run([NC_HOST, "write", p, str(nx), str(ny), str(nlay)], check=True, timeout=60)
    with netcdf_file(p, "r", mmap=False) as s:
        assert s.version_byte == 1 and s.dimensions["t"] is None
        assert s.variables["t"][:].tolist() == [1800.5, 3600.25]
        assert s.variables["x"][:].tolist() == [-.5 * nx + i + .5 for i in range(nx)]
        assert s.variables["lay"][:].tolist() == [1, 2, 3]
        assert s.variables["dp"].dimensions == ("t", "lay", "y", "x")
        i, j, k = np.meshgrid(np.arange(1, nx + 1), np.arange(1, ny + 1), np.arange(1, nlay + 1), indexing="ij")
        for r in (1, 2):
            v = (i + 100 * j + 10000 * k + 1e6 * r).transpose(2, 1, 0)
            assert np.array_equal(s.variables["dp"][r - 1], v.astype(np.float32))
            assert np.array_equal(s.variables["ts"][r - 1], v[0].astype(np.float32))
        assert np.array_equal(s.variables["gzs"][:], v[0] + 0.125)  # the last record's level 1
