"""CPU tests of the host side: the C-ABI library loads and exports every
symbol include/atm.h declares; config/duration parsing and the time-step
logic of src/mod_model.f90; the initial-condition recipe."""
import os
import re

import numpy as np
import pytest

from atmos import driver
from atmos.config import Config, ConfigError, parse_duration, read_config_file
from atmos.ic import CP_DRY, KAPPA, make_initial_state

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "atm.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(atm_\w+)\s*\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    import ctypes
    from atmos import model
    syms = declared_symbols()
    assert len(syms) >= 25
    lib = ctypes.CDLL(model.LIB_PATH)
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    assert sorted(model.EXPORTS) == syms
    model.load_library()  # signatures declare without a GPU


def test_no_gpu_create_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    from atmos.model import AtmError, Model
    with pytest.raises(AtmError):
        Model(16, 16, 8, 1e5)


def test_no_gpu_ensemble_fails_loudly():
    """Without a device the ensemble's first stream already fails; nothing is
    left half built."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    from atmos.ensemble import Ensemble
    from atmos.model import AtmError
    with pytest.raises(AtmError):
        Ensemble([(2, 32, 32, 8)], 100e3, 60.0)


@pytest.mark.parametrize("text,sec", [("PT20S", 20), ("PT2M", 120), ("PT3H", 10800), ("P5D", 432000),
                                      ("PT100S", 100), ("P1Y2M3DT4H5M6S", 31536000 + 2 * 2592000 + 3 * 86400 + 4 * 3600 + 5 * 60 + 6),
                                      ("P1W", 604800), ("PT1.5H", 5400)])
def test_parse_duration(text, sec):
    assert parse_duration(text) == sec


@pytest.mark.parametrize("bad", ["T5M", "PTXS", "P-1D"])
def test_parse_duration_errors(bad):
    with pytest.raises(ConfigError):
        parse_duration(bad)


def test_read_reference_style_namelist(tmp_path):
    p = tmp_path / "config.nml"
    p.write_text("""&io_control
    initial_filename = 'initial.nc'
    save_restart_file = T
    restart_filename = 'output/restart.nc'
/
&time_control
    max_timestep = "PT20S"
    run_duration = "PT2000S"
    output_interval = "PT1H" ! comment
    radiation_interval = "PT3H"
    physics_interval = "PT5M"
    remap_interval = "PT100S"
    radiation_on = F
/
&physics_parameters
    gravity = 9.807
    coriolis_parameter = 1e-5
    divergence_damping_coefficient = 0.15
    vorticity_damping_coefficient = 0.005
/
""")
    c = read_config_file(str(p))
    assert (c.dt_max, c.t_final, c.dt_remap, c.dt_output) == (20, 2000, 100, 3600)
    assert c.radiation_on is False and c.physics_on is True and c.save_restart_file is True
    assert c.gravity == 9.807 and c.coriolis_parameter == 1e-5


def test_calculate_dt_remap_every_five_steps():
    cfg = Config(dt_max=20.0, t_final=2000.0, dt_output=1e9, dt_radiation=1e9, dt_physics=1e9,
                 dt_remap=100.0, radiation_on=False, physics_on=False)
    st = driver.DriverState(0.0, 0.0, 0.0, 0.0)
    t, remaps = 0.0, []
    for n in range(1, 101):
        dt = driver.calculate_dt(t, cfg, st)
        assert dt == 20.0
        if st.remap_this_step:
            remaps.append(n)
            st.remap_this_step = False
        t += dt
    assert remaps == list(range(5, 101, 5))
    assert t == 2000.0


def test_calculate_dt_shortens_to_hit_events():
    cfg = Config(dt_max=120.0, t_final=1000.0, dt_output=250.0, dt_radiation=1e9, dt_physics=1e9,
                 dt_remap=1e9)
    st = driver.DriverState(0.0, 0.0, 0.0, 0.0)
    dts, t = [], 0.0
    while t < cfg.t_final:
        dt = driver.calculate_dt(t, cfg, st)
        dts.append(dt)
        t += dt
    assert dts[:3] == [120.0, 120.0, 10.0] and abs(sum(dts) - 1000.0) < 1e-9


def test_initial_condition_recipe():
    ic = make_initial_state(32, 24, 20, 100e3)
    ak, bk = ic["ak"], ic["bk"]
    assert ak[0] == 1000.0 and bk[0] == 1.0 and bk[-1] == 0.0  # Q16
    plev = ak[:, None, None] + bk[:, None, None] * (ic["dp"].sum(0)[None] + 1000.0 - 1000.0)
    assert np.all(ic["dp"] > 0)
    # surface pressure 1000 hPa + 5 hPa bump at the centre
    ps = ic["dp"].sum(0) + 1000.0
    assert abs(ps.max() - 1e5 - 500.0) < 30 and ps.min() > 1e5
    # layer temperature pt * p^kappa is within the lapse-rate/isothermal profile
    pl = 1000.0 + np.cumsum(ic["dp"][::-1], axis=0)[::-1]
    pu = np.concatenate([pl[1:], np.full((1,) + pl.shape[1:], 1000.0)])
    pk, pk1 = pl ** KAPPA, pu ** KAPPA
    T = ic["pt"] * (pk1 - pk) / np.log(pk1 / pk)
    assert T.max() < 281.0 and T.min() > 208.0
    # winds: 10 m/s + unit-std noise; deterministic per seed
    assert abs(ic["u"].mean() - 10.0) < 1e-9 and abs(ic["u"][0].std() - 1.0) < 1e-9
    ic2 = make_initial_state(32, 24, 20, 100e3)
    assert np.array_equal(ic["u"], ic2["u"]) and np.array_equal(ic["pt"], ic2["pt"])
