"""Run configuration with the reference's namelist semantics.

`parse_duration` follows src/mod_util.f90:33-159 (ISO-8601 durations; a year
is 365 days, a month 30 days; 'M' before 'T' is months, after it minutes).
`read_config_file` reads the three namelists of src/mod_config.f90:31-93
(io_control, time_control, physics_parameters) from a config.nml.  The
additive optional keys `radiation_on` and `physics_on` (time_control) switch
those stages off; they default to the reference behaviour (on).
"""
from __future__ import annotations

import dataclasses
import re


class ConfigError(ValueError):
    pass


def _read_real(s: str) -> float:
    # mod_util.f90:161-172: Fortran F edit; anything unreadable is -1
    try:
        return float(s.strip()) if s.strip() else -1.0
    except ValueError:
        return -1.0


def parse_duration(text: str) -> float:
    """ISO-8601 duration -> seconds, as src/mod_util.f90:33-159."""
    s = text.strip()
    if not s.startswith("P"):
        raise ConfigError(f"Could not read ISO8601 duration `{text}`")
    i_time = s.find("T")
    if i_time >= 0:
        date, tpart = s[:i_time], s[i_time + 1:]
    else:
        date, tpart = s, ""
    fields = {}
    prev = 0  # index of 'P'
    for key in "YMWD":
        i = date.find(key)
        if i >= 0:
            v = _read_real(date[prev + 1:i])
            if v < 0:
                raise ConfigError(f"Could not read ISO8601 duration `{text}`")
            fields[key] = v
            prev = i
    if i_time >= 0:
        prev = -1
        for key in "HMS":
            i = tpart.find(key)
            if i >= 0:
                v = _read_real(tpart[prev + 1:i])
                if v < 0:
                    raise ConfigError(f"Could not read ISO8601 duration `{text}`")
                fields["t" + key] = v
                prev = i
    # default-real literal multipliers are exact in binary32
    return (31536000.0 * fields.get("Y", 0.0) + 2592000.0 * fields.get("M", 0.0)
            + 604800.0 * fields.get("W", 0.0) + 86400.0 * fields.get("D", 0.0)
            + 3600.0 * fields.get("tH", 0.0) + 60.0 * fields.get("tM", 0.0)
            + 1.0 * fields.get("tS", 0.0))


@dataclasses.dataclass
class Config:
    # io_control
    initial_filename: str = "initial.nc"
    save_restart_file: bool = False
    restart_filename: str = "output/restart.nc"
    # time_control (seconds)
    dt_max: float = 120.0
    t_final: float = 5 * 86400.0
    dt_output: float = 3600.0
    dt_radiation: float = 3 * 3600.0
    dt_physics: float = 300.0
    dt_remap: float = 3 * 3600.0
    t_initial: float = 0.0
    radiation_on: bool = True
    physics_on: bool = True
    # physics_parameters (FP64, read from the namelist)
    gravity: float = 9.807
    coriolis_parameter: float = 1e-5
    divergence_damping_coefficient: float = 0.15
    vorticity_damping_coefficient: float = 0.005
    # domain (from the initial condition)
    nx: int = -1
    ny: int = -1
    nlay: int = -1
    dx: float = -1.0
    dy: float = -1.0


_NML_GROUP = re.compile(r"&(\w+)(.*?)/", re.S)
_NML_ITEM = re.compile(r"(\w+)\s*=\s*('[^']*'|\"[^\"]*\"|[^\s,!]+)")


def _value(tok: str):
    t = tok.strip()
    if t[:1] in "'\"":
        return t[1:-1]
    if t.upper() in ("T", ".TRUE.", "TRUE"):
        return True
    if t.upper() in ("F", ".FALSE.", "FALSE"):
        return False
    return float(t.replace("d", "e").replace("D", "e"))


def read_config_file(path: str) -> Config:
    """Read config.nml (the three namelists of src/mod_config.f90:38-50)."""
    with open(path) as fh:
        text = "\n".join(line.split("!", 1)[0] for line in fh)
    groups = {}
    for name, body in _NML_GROUP.findall(text):
        groups[name.lower()] = {k.lower(): _value(v) for k, v in _NML_ITEM.findall(body)}
    for need in ("io_control", "time_control", "physics_parameters"):
        if need not in groups:
            raise ConfigError(f"Could not read `{need}` namelist")
    io, tc, pp = groups["io_control"], groups["time_control"], groups["physics_parameters"]
    c = Config()
    c.initial_filename = io.get("initial_filename", c.initial_filename)
    c.save_restart_file = bool(io.get("save_restart_file", c.save_restart_file))
    c.restart_filename = io.get("restart_filename", c.restart_filename)
    c.t_final = parse_duration(tc["run_duration"])
    c.dt_max = parse_duration(tc["max_timestep"])
    c.dt_output = parse_duration(tc["output_interval"])
    c.dt_radiation = parse_duration(tc["radiation_interval"])
    c.dt_physics = parse_duration(tc["physics_interval"])
    c.dt_remap = parse_duration(tc["remap_interval"])
    c.radiation_on = bool(tc.get("radiation_on", True))
    c.physics_on = bool(tc.get("physics_on", True))
    c.gravity = pp["gravity"]
    c.coriolis_parameter = pp["coriolis_parameter"]
    c.divergence_damping_coefficient = pp["divergence_damping_coefficient"]
    c.vorticity_damping_coefficient = pp["vorticity_damping_coefficient"]
    return c
