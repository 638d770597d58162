"""An ensemble of independent models of uneven grid sizes on one GPU.

Each member is a `Model` with its own device state and its own seeded
initial state (temperature noise with a k^-3 spectrum on top of the common
recipe).  The members share one schedule: the time step, the remap
interval and the radiation interval.  A step issues every member's
launches in the reference's per-step call order (`driver.step`) on that
member's HIP stream, then checks every member's stability; members are
dealt round-robin to a few streams of default priority, so the launches of
different members overlap on the GPU while each member's own launches stay
in order.
"""
from __future__ import annotations

import dataclasses

from . import driver
from .config import Config
from .ic import make_initial_state
from .model import Model, stream_create, stream_destroy


@dataclasses.dataclass
class Member:
    model: Model
    state: driver.DriverState
    config: Config
    time: float
    cells: int


class Ensemble:
    """specs: (count, nx, ny, nlay) groups.  Member i is seeded with seed + i.
    autotune: each member times the fused path's launch shapes on its own
    grid at construction (results do not depend on the choice)."""

    def __init__(self, specs, dx: float, dt: float, remap_every: int = 5, radiation_every: int = 10,
                 precision: str = "fp64", nstreams: int = 4, seed: int = 20261205, t_noise: float = 0.5,
                 autotune: bool = False):
        if nstreams < 1:
            raise ValueError("nstreams must be >= 1")
        self.streams: list[int] = []
        self.members: list[Member] = []
        try:
            self.streams = [stream_create() for _ in range(nstreams)]
            i = 0
            for count, nx, ny, nlay in specs:
                for _ in range(count):
                    ic = make_initial_state(nx, ny, nlay, dx, seed=seed + i, t_noise=t_noise)
                    m = Model(nx, ny, nlay, dx, precision=precision)
                    self.members.append(Member(m, None, None, 0.0, nx * ny * nlay))
                    m.set_stream(self.streams[i % nstreams])
                    cfg = Config(dt_max=dt, t_final=1e12, dt_output=1e12, dt_radiation=radiation_every * dt,
                                 dt_physics=1e12, dt_remap=remap_every * dt,
                                 radiation_on=radiation_every > 0, physics_on=False)
                    st = driver.init_model(m, ic, cfg)
                    if autotune:  # launch shapes for this member's own grid
                        m.autotune()
                    self.members[-1] = Member(m, st, cfg, cfg.t_initial, nx * ny * nlay)
                    i += 1
        except BaseException:
            self.close()
            raise

    @property
    def cells(self) -> int:
        return sum(mb.cells for mb in self.members)

    def step(self) -> bool:
        """One step of every member; False when a member is no longer stable
        (check_stability, mod_fields.f90:69-106)."""
        dts = [driver.step(mb.model, mb.config, mb.state, mb.time) for mb in self.members]
        ok = True
        for mb, dt in zip(self.members, dts):
            ok = mb.model.check_stability() and ok
            mb.time = mb.time + dt
        return ok

    def sync(self):
        for mb in self.members:
            mb.model.sync()

    def close(self):
        """Release the members' contexts, then the streams they ran on
        (closing twice is harmless)."""
        for mb in self.members:
            mb.model.close()  # waits for the member's stream
        self.members.clear()
        for s in self.streams:
            stream_destroy(s)
        self.streams = []
