"""An ensemble of independent models of uneven grid sizes on one GPU.

Each member is a `Model` with its own device state and its own seeded
initial state (temperature noise with a k^-3 spectrum on top of the common
recipe).  The members share one schedule: the time step, the remap
interval and the radiation interval.  Members are dealt round-robin to a
few HIP streams of default priority, so the launches of different members
overlap on the GPU while each member's own launches stay in order.

A step issues every member's calls in the reference's per-step order
(`driver.step`, mod_model.f90:117-183) and checks every member's stability
(check_stability, mod_fields.f90:69-106).  With deferred checks (the
default) a member's check is queued behind its step and read after the
next step has been issued, so the streams never drain between steps: the
result `step()` returns is that of the previous step, and `finish()` reads
the last one.  `checks="sync"` waits for every member's check each step.

Graph mode (`graph_period=P`): each member records P steps of dynamics and
remaps as one HIP graph (atm_graph_begin/end) during the first period and
replays it for every later one, with radiation issued eagerly at each
period's start and one deferred check per period.  The period must hold
whole remap intervals and an even number of buffer switches (each D step
and each remap switches), and radiation may fall due only at period
starts.  The recorded time steps are those of the first period; the remap
pattern of every later period is checked against it.

Batched mode (`launch="batched"`): the stages of a step (C step, its halo
fill, D step, the remap, the D halo fill) and radiation run as one launch
per stage for all members together (atm_ensemble_*: a member table
indexed by workgroup), on the first member's stream, joined to the other
members' streams by events.  The members share the time step and the
remap and radiation schedules.  Every member's values are
those of its own launches, bit for bit.  With a graph period as well, the
batched launches of a whole period are recorded as one graph
(atm_ensemble_graph_*) and replayed for every later period, under
the same rules as the members' own graphs; the recorded time steps and
remap pattern are checked against each later period's.
"""
from __future__ import annotations

import dataclasses

from . import driver
from .config import Config
from .ic import make_initial_state
from .model import BatchedSteps, Model, stream_create, stream_destroy


@dataclasses.dataclass
class Member:
    model: Model
    state: driver.DriverState
    config: Config
    time: float
    cells: int
    pending: bool = False  # a deferred check is queued
    pattern: list | None = None  # graph mode: remap flags of the recorded period


class Ensemble:
    """specs: (count, nx, ny, nlay) groups.  Member i is seeded with seed + i.
    autotune: each member times the fused path's launch shapes on its own
    grid at construction (results do not depend on the choice).
    launch: "streams" (each member's own launches, members dealt to nstreams
    streams) or "batched" (one launch per dynamics, remap and radiation
    stage for all members)."""

    def __init__(self, specs, dx: float, dt: float, remap_every: int = 5, radiation_every: int = 10,
                 precision: str = "fp64", nstreams: int = 4, seed: int = 20261205, t_noise: float = 0.5,
                 autotune: bool = False, checks: str = "deferred", graph_period: int = 0,
                 launch: str = "streams"):
        if nstreams < 1:
            raise ValueError("nstreams must be >= 1")
        if checks not in ("deferred", "sync"):
            raise ValueError("checks: 'deferred' or 'sync'")
        if launch not in ("streams", "batched"):
            raise ValueError("launch: 'streams' or 'batched'")
        self.launch = launch
        self.batch = None
        self.batch_pattern = None  # batched graph mode: (dt, remap flags) of the recorded period
        self.remap_every, self.radiation_every = remap_every, radiation_every
        self.checks = checks
        self.graph_period = 0
        if graph_period:
            self._validate_period(graph_period)
            self.graph_period = graph_period
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
            if launch == "batched":
                self.batch = BatchedSteps([mb.model for mb in self.members])
        except BaseException:
            self.close()
            raise

    @property
    def cells(self) -> int:
        return sum(mb.cells for mb in self.members)

    def step(self) -> bool:
        """One step of every member; False when a member is no longer stable
        (with deferred checks: was not stable after the previous step)."""
        if self.graph_period:
            raise RuntimeError("graph mode steps whole periods: period()")
        if self.batch is not None:
            self._batched_step()
            return self._check()
        dts = [driver.step(mb.model, mb.config, mb.state, mb.time) for mb in self.members]
        for mb, dt in zip(self.members, dts):
            mb.time = mb.time + dt
        return self._check()

    def _batched_radiation(self):
        """calculate_radiative_heating of the members whose radiation is due:
        one batched call when all of them are (they share the schedule)."""
        due = [mb.state.calculate_radiation_this_step and mb.config.radiation_on for mb in self.members]
        if all(due):
            self.batch.calculate_radiative_heating(self.members[0].time)
        for mb, d in zip(self.members, due):
            if d:
                if not all(due):
                    mb.model.calculate_radiative_heating(mb.time)
                mb.model.radiation_halo_exchange()
            mb.state.calculate_radiation_this_step = False

    def _batched_step(self):
        """driver.step (mod_model.f90:117-183) for every member, its stages
        batched over the members."""
        dts = []
        if any(mb.state.calculate_radiation_this_step for mb in self.members):
            self._batched_radiation()
        for mb in self.members:
            m, cfg, st = mb.model, mb.config, mb.state
            dts.append(driver.calculate_dt(mb.time, cfg, st))
            if cfg.physics_on and st.calculate_physics_this_step:
                m.calculate_physics_tendencies()
            st.calculate_physics_this_step = False
        dt = dts[0]
        if any(d != dt for d in dts):
            raise RuntimeError("batched members must share the time step")
        b = self.batch
        b.cgrid_dynamics_step(dt / 2.0)
        b.cgrid_halo_exchange()
        b.dgrid_dynamics_step(dt)
        for mb in self.members:
            if mb.config.physics_on:
                mb.model.apply_physics_tendencies(dt)
        if self._remap_due([mb.state.remap_this_step for mb in self.members]):
            b.perform_vertical_remapping()
        for mb in self.members:
            mb.state.remap_this_step = False
        b.dgrid_halo_exchange()
        for mb in self.members:
            if mb.config.physics_on:
                mb.model.physics_halo_exchange()
            mb.time = mb.time + dt

    @staticmethod
    def _remap_due(flags) -> bool:
        if any(flags) and not all(flags):
            raise RuntimeError("batched members must remap together")
        return all(flags)

    def _validate_period(self, P: int):
        if self.remap_every < 1 or P % self.remap_every or (P + P // self.remap_every) % 2:
            raise ValueError("graph_period must hold whole remap intervals and an even number of "
                             "buffer switches (D steps + remaps)")
        if self.radiation_every > 0 and self.radiation_every % P:
            raise ValueError("radiation must fall due at period starts (radiation_every % graph_period == 0)")

    def use_graphs(self, period: int):
        """Switch to graph mode from here on (at a period start of the
        schedule: radiation due now or not within the period)."""
        self._validate_period(period)
        self.graph_period = period
        self.batch_pattern = None  # a batched ensemble records afresh

    def period(self) -> bool:
        """Graph mode: graph_period steps of every member (the first call
        records each member's graph); False when a member was not stable
        after the previous period."""
        if not self.graph_period:
            raise RuntimeError("period() needs graph_period")
        if self.batch is not None:
            self._batched_period()
        else:
            for mb in self.members:
                self._member_period(mb)
        return self._check()

    def finish(self) -> bool:
        """Reads the deferred checks still queued; False when a member was
        not stable after its last step."""
        ok = True
        for mb in self.members:
            if mb.pending:
                mb.pending = False
                ok = mb.model.check_stability_end() and ok
        return ok

    def _check(self) -> bool:
        ok = True
        for mb in self.members:
            if self.checks == "sync":
                ok = mb.model.check_stability() and ok
                continue
            if mb.pending:
                mb.pending = False
                ok = mb.model.check_stability_end() and ok
            mb.model.check_stability_begin()
            mb.pending = True
        return ok

    def _member_period(self, mb: Member):
        m, st, cfg = mb.model, mb.state, mb.config
        if st.calculate_radiation_this_step:  # due at the period start
            if cfg.radiation_on:
                m.calculate_radiative_heating(mb.time)
                m.radiation_halo_exchange()
            st.calculate_radiation_this_step = False
        record = mb.pattern is None
        if record:
            if mb.pending:  # a graph recording starts with no check queued
                mb.pending = False
                if not m.check_stability_end():
                    raise FloatingPointError("ensemble member unstable before its graph was recorded")
            m.graph_begin()
        t, pattern = mb.time, []
        for k in range(self.graph_period):
            dt = driver.calculate_dt(t, cfg, st)
            remap = st.remap_this_step
            st.remap_this_step = False
            if record:
                m.cgrid_dynamics_step(dt / 2.0)
                m.cgrid_halo_exchange()
                m.dgrid_dynamics_step(dt)
                if remap:
                    m.perform_vertical_remapping()
                m.dgrid_halo_exchange()
            pattern.append(remap)
            t = t + dt
            if st.calculate_radiation_this_step and k + 1 < self.graph_period:
                raise RuntimeError("radiation fell due inside a graph period")
        if record:
            m.graph_end()
            mb.pattern = pattern
        elif pattern != mb.pattern:
            raise RuntimeError("the remap pattern of this period differs from the recorded one")
        m.graph_launch(1)
        mb.time = t

    def _batched_period(self):
        """graph_period steps of every member as one recording of batched
        and member calls (recorded by the first call, replayed after), with
        radiation issued eagerly at the period start."""
        if any(mb.state.calculate_radiation_this_step for mb in self.members):  # due at the period start
            self._batched_radiation()
        record = self.batch_pattern is None
        b = self.batch
        if record:
            for mb in self.members:  # a recording starts with no check queued
                if mb.pending:
                    mb.pending = False
                    if not mb.model.check_stability_end():
                        raise FloatingPointError("ensemble member unstable before the graph was recorded")
            b.graph_begin()
        times, pattern = [mb.time for mb in self.members], []
        for k in range(self.graph_period):
            dts, remaps = [], []
            for i, mb in enumerate(self.members):
                dts.append(driver.calculate_dt(times[i], mb.config, mb.state))
                remaps.append(mb.state.remap_this_step)
                mb.state.remap_this_step = False
            dt = dts[0]
            if any(d != dt for d in dts):
                raise RuntimeError("batched members must share the time step")
            remap = self._remap_due(remaps)
            if record:
                b.cgrid_dynamics_step(dt / 2.0)
                b.cgrid_halo_exchange()
                b.dgrid_dynamics_step(dt)
                if remap:
                    b.perform_vertical_remapping()
                b.dgrid_halo_exchange()
            pattern.append((dt, tuple(remaps)))
            times = [t + dt for t in times]
            if k + 1 < self.graph_period and any(mb.state.calculate_radiation_this_step for mb in self.members):
                raise RuntimeError("radiation fell due inside a graph period")
        if record:
            b.graph_end()
            self.batch_pattern = pattern
        elif pattern != self.batch_pattern:
            raise RuntimeError("the time steps or remap pattern of this period differ from the recorded ones")
        b.graph_launch(1)
        for mb, t in zip(self.members, times):
            mb.time = t

    def sync(self):
        for mb in self.members:
            mb.model.sync()

    def close(self):
        """Release the batched steps, the members' contexts, then the
        streams they ran on (closing twice is harmless)."""
        if self.batch is not None:
            self.batch.close()
            self.batch = None
        for mb in self.members:
            mb.model.close()  # waits for the member's stream
        self.members.clear()
        for s in self.streams:
            stream_destroy(s)
        self.streams = []
