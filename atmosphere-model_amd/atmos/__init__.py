"""atmos -- host side of the MI355X (gfx950) atmosphere hot-path library.

* `model.Model`: ctypes binding of libatmos_hip.so with the reference's
  module-procedure names (cgrid_dynamics_step, perform_vertical_remapping,
  calculate_radiative_heating, ...).
* `driver`: init_model / run_model / calculate_dt of src/mod_model.f90.
* `config`: config.nml namelists and ISO-8601 durations.
* `ic`: seeded synthetic initial conditions.
* `ncio`: initial, output and restart files in the netCDF classic formats
  (libatmos_io.so).
* `ensemble`: independent members of uneven grid sizes stepped together.
"""
from . import config, driver, ensemble, ic, model, ncio  # noqa: F401

__all__ = ["config", "driver", "ensemble", "ic", "model", "ncio"]
