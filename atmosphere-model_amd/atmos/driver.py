"""Host driver with the reference's call order and time-step logic.

Mirrors src/mod_model.f90: `init_model` (36-89), `run_model` (91-199) and
`calculate_dt` (201-237), driving any backend object that exposes the
reference module procedures by their reference names (the HIP `Model`, or
the CPU oracle in the tests):

    put(name, interior_array), init_prognostic_fields(), initial_halo_exchange(),
    init_sw_dyn(), allocate_radiation_arrays(), allocate_physics_arrays(),
    allocate_remapping_arrays(), calculate_radiative_heating(time),
    radiation_halo_exchange(), calculate_physics_tendencies(),
    cgrid_dynamics_step(dt), cgrid_halo_exchange(), dgrid_dynamics_step(dt),
    apply_physics_tendencies(dt), perform_vertical_remapping(),
    dgrid_halo_exchange(), physics_halo_exchange(), check_stability()

radiation_on / physics_on are harness switches (the reference has none,
Q11): off means the procedures are not called and net_flux stays zero.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from .config import Config

TIME_EPS = float(np.float32(1e-2))  # mod_model.f90:205, default-real literal


@dataclasses.dataclass
class DriverState:
    previous_write_time: float
    previous_radiation_time: float
    previous_physics_time: float
    previous_remap_time: float
    write_this_step: bool = False
    calculate_radiation_this_step: bool = True
    calculate_physics_this_step: bool = True
    remap_this_step: bool = False
    stable: bool = True


def init_model(backend, ic: dict, config: Config) -> DriverState:
    """init_model (mod_model.f90:36-89) after the initial file is read."""
    config.nlay = ic["dp"].shape[0]
    config.ny, config.nx = ic["dp"].shape[1:]
    config.dx, config.dy = ic["dx"], ic["dy"]
    for name, key in (("dp", "dp"), ("pt", "pt"), ("ud", "u"), ("vd", "v"),
                      ("gzs", "gzs"), ("ts", "ts"), ("ak", "ak"), ("bk", "bk")):
        backend.put(name, ic[key])
    xc, yc = ic["xc"], ic["yc"]
    backend.put("radius", np.sqrt(xc[None, :] ** 2 + yc[:, None] ** 2))
    backend.init_prognostic_fields()
    backend.initial_halo_exchange()
    backend.init_sw_dyn()
    backend.allocate_radiation_arrays()
    backend.allocate_physics_arrays()
    backend.allocate_remapping_arrays()
    t0 = config.t_initial
    return DriverState(t0, t0, t0, t0, calculate_radiation_this_step=True,
                       calculate_physics_this_step=True, remap_this_step=False,
                       write_this_step=False)


def calculate_dt(time: float, config: Config, st: DriverState) -> float:
    """calculate_dt (mod_model.f90:201-237)."""
    cand = [config.dt_max,
            config.t_final - time,
            st.previous_write_time + config.dt_output - time,
            st.previous_radiation_time + config.dt_radiation - time,
            st.previous_physics_time + config.dt_physics - time,
            st.previous_remap_time + config.dt_remap - time]
    pos = [c for c in cand if c > 0]
    dt = min(pos)  # minloc(mask > 0): first minimum; the value is what matters
    if abs(dt - cand[2]) < TIME_EPS:
        st.write_this_step = True
        st.previous_write_time = time + dt
    if abs(dt - cand[3]) < TIME_EPS:
        st.calculate_radiation_this_step = True
        st.previous_radiation_time = time + dt
    if abs(dt - cand[4]) < TIME_EPS:
        st.calculate_physics_this_step = True
        st.previous_physics_time = time + dt
    if abs(dt - cand[5]) < TIME_EPS:
        st.remap_this_step = True
        st.previous_remap_time = time + dt
    return dt


def step(backend, config: Config, st: DriverState, time: float, hooks=None) -> float:
    """One pass of the main loop body (mod_model.f90:117-183); returns dt, or
    raises FloatingPointError when check_stability fails."""
    if st.calculate_radiation_this_step:
        if config.radiation_on:
            backend.calculate_radiative_heating(time)
            backend.radiation_halo_exchange()
        st.calculate_radiation_this_step = False
    dt = calculate_dt(time, config, st)
    if config.physics_on and st.calculate_physics_this_step:
        backend.calculate_physics_tendencies()
    st.calculate_physics_this_step = False
    backend.cgrid_dynamics_step(dt / 2.0)
    backend.cgrid_halo_exchange()
    backend.dgrid_dynamics_step(dt)
    if config.physics_on:
        backend.apply_physics_tendencies(dt)
    if st.remap_this_step:
        backend.perform_vertical_remapping()
        st.remap_this_step = False
    backend.dgrid_halo_exchange()
    if config.physics_on:
        backend.physics_halo_exchange()
    if hooks and "after_step" in hooks:
        hooks["after_step"](dt)
    return dt


def run_model(backend, config: Config, st: DriverState, max_steps: int | None = None,
              check_every: int = 1) -> tuple:
    """run_model (mod_model.f90:91-199) without output; returns
    (time, steps, stable).  check_every > 1 checks stability less often
    (the reference checks every step)."""
    time = config.t_initial
    n = 0
    while time < config.t_final:
        if max_steps is not None and n >= max_steps:
            break
        dt = step(backend, config, st, time)
        n += 1
        if n % check_every == 0 and not backend.check_stability():
            st.stable = False
            break
        time = time + dt
    return time, n, st.stable
