"""The model's files around the GPU path: a restart written from the device
state and read back as an initial file continues the run bit for bit (the
FP64 restart with ak, bk and lev that the reference's float32 restart lacks,
Q13), and the time means written to output.nc match the oracle's."""
import os

import numpy as np
import pytest
from scipy.io import netcdf_file

import oracle as O
from atmos import driver, ncio
from atmos.config import Config
from atmos.ic import make_initial_state
from atmos.model import Model

pytestmark = pytest.mark.gpu


def _cfg(t0=0.0):
    return Config(dt_max=60.0, t_initial=t0, t_final=1e12, dt_output=1e12, dt_radiation=1e12,
                  dt_physics=1e12, dt_remap=300.0, radiation_on=False, physics_on=False)


def _run(backend, cfg, st, nsteps):
    t = cfg.t_initial
    for _ in range(nsteps):
        t += driver.step(backend, cfg, st, t)
    return t


def test_restart_continues_bit_for_bit(tmp_path):
    nx, ny, nlay = 64, 32, 10
    print(f"restart round trip {nx}x{ny}x{nlay}", flush=True)
    ic = make_initial_state(nx, ny, nlay, 100e3)
    p = str(tmp_path / "initial.nc")
    ncio.write_initial(p, ic)
    ic_file = ncio.read_initial(p)

    # uninterrupted: 10 steps (remaps at 300 s and 600 s)
    a = Model(nx, ny, nlay, ic_file["dx"], ic_file["dy"])
    cfg = _cfg()
    st = driver.init_model(a, ic_file, cfg)
    _run(a, cfg, st, 10)

    # 5 steps, restart file, 5 more from it
    b = Model(nx, ny, nlay, ic_file["dx"], ic_file["dy"])
    cfg = _cfg()
    st = driver.init_model(b, ic_file, cfg)
    t = _run(b, cfg, st, 5)
    assert t == 300.0
    r = str(tmp_path / "restart.nc")
    ncio.write_restart(r, t, b.get("dp"), b.get("pt"), b.get("ud"), b.get("vd"), b.get("gz")[0], b.get("ts"),
                       ic_file["ak"], ic_file["bk"], ic_file["dx"], ic_file["dy"])
    del b
    ic2 = ncio.read_initial(r)
    assert ic2["t_initial"] == 300.0
    c = Model(nx, ny, nlay, ic2["dx"], ic2["dy"])
    cfg2 = _cfg(ic2["t_initial"])
    st2 = driver.init_model(c, ic2, cfg2)
    _run(c, cfg2, st2, 5)
    for name in ("dp", "pt", "ud", "vd"):
        assert np.array_equal(c.get(name), a.get(name)), name


def test_output_file_of_device_means(tmp_path):
    nx, ny, nlay = 32, 16, 8
    print(f"output.nc of device means {nx}x{ny}x{nlay}", flush=True)
    ic = make_initial_state(nx, ny, nlay, 100e3)
    gpu, orc = Model(nx, ny, nlay, 100e3, 100e3), O.Oracle(nx, ny, nlay, 100e3, 100e3)
    p = str(tmp_path / "output.nc")
    out = ncio.OutputFile(p, nx, ny, nlay, 100e3, 100e3)
    means = []
    for b in (gpu, orc):
        cfg = _cfg()
        st = driver.init_model(b, ic, cfg)
        b.init_output()
        t = 0.0
        rec = []
        for n in range(6):
            dt = driver.step(b, cfg, st, t)
            b.accumulate_output(dt)
            t += dt
            if n % 3 == 2:
                m = b.output_mean()
                rec.append((t, m))
                if b is gpu:
                    out.write(t, m)
        means.append(rec)
    out.close()
    with netcdf_file(p, "r", mmap=False) as s:
        assert s.variables["t"][:].tolist() == [180.0, 360.0]
        for r, (t, m) in enumerate(means[1]):
            for i, n in enumerate(("dp", "pt", "u", "v")):
                got = s.variables[n][r].astype(np.float64)
                want = m[i, :nlay]
                assert np.max(np.abs(got - want)) <= 1e-6 * np.max(np.abs(want)), n
            got = s.variables["ts"][r].astype(np.float64)
            assert np.max(np.abs(got - m[4, 0])) <= 1e-6 * np.max(np.abs(m[4, 0]))
