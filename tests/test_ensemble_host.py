"""CPU tests of the batched ensemble's host logic (atmos/ensemble.py): which
calls a batched step and a batched graph period issue, in which order, for
members that share the schedule, and the refusals when they do not.  The
models and the batched steps are recording stand-ins, so no GPU is needed;
the GPU tests check the values the real calls give."""
import pytest

from atmos import driver
from atmos.config import Config
from atmos.ensemble import Ensemble, Member


class Log(list):
    pass


class FakeModel:
    def __init__(self, log, i):
        self.log, self.i = log, i

    def __getattr__(self, name):  # calculate_radiative_heating, radiation_halo_exchange, ...
        return lambda *a: self.log.append((name, self.i))

    def check_stability_begin(self):
        self.log.append(("check_begin", self.i))

    def check_stability_end(self):
        self.log.append(("check_end", self.i))
        return True


class FakeBatch:
    def __init__(self, log):
        self.log = log

    def __getattr__(self, name):  # cgrid_dynamics_step, perform_vertical_remapping, graph_begin, ...
        return lambda *a: self.log.append((name, "batch"))


def make(n=3, dt=60.0, remap_every=5, radiation_every=10, dt_remap=None, graph_period=0):
    log = Log()
    ens = object.__new__(Ensemble)
    ens.launch = "batched"
    ens.batch = FakeBatch(log)
    ens.batch_pattern = None
    ens.remap_every, ens.radiation_every = remap_every, radiation_every
    ens.checks = "deferred"
    ens.graph_period = graph_period
    ens.streams = []
    ens.members = []
    for i in range(n):
        cfg = Config(dt_max=dt, t_final=1e12, dt_output=1e12, dt_radiation=radiation_every * dt, dt_physics=1e12,
                     dt_remap=(dt_remap[i] if dt_remap else remap_every) * dt, radiation_on=True,
                     physics_on=False)
        st = driver.DriverState(previous_write_time=0., previous_radiation_time=0., previous_physics_time=0.,
                                previous_remap_time=0.)
        ens.members.append(Member(FakeModel(log, i), st, cfg, 0.0, 1))
    return ens, log


def batch_calls(log):
    return [name for name, who in log if who == "batch"]


def test_batched_step_issues_one_call_per_stage():
    ens, log = make()
    ens._batched_step()  # step 1: radiation due at t = 0, no remap
    assert batch_calls(log) == ["calculate_radiative_heating", "cgrid_dynamics_step", "cgrid_halo_exchange",
                                "dgrid_dynamics_step", "dgrid_halo_exchange"]
    # the net flux's halo fill stays per member, after the batched radiation
    assert [who for name, who in log if name == "radiation_halo_exchange"] == [0, 1, 2]
    assert all(name != "calculate_radiative_heating" or who == "batch" for name, who in log)
    log.clear()
    for _ in range(4):  # steps 2..5: the remap falls due at the fifth
        ens._batched_step()
    assert batch_calls(log).count("perform_vertical_remapping") == 1
    i = batch_calls(log).index("perform_vertical_remapping")
    assert batch_calls(log)[i - 1] == "dgrid_dynamics_step" and batch_calls(log)[i + 1] == "dgrid_halo_exchange"
    assert all(mb.time == 5 * 60.0 for mb in ens.members)
    assert not any(mb.state.remap_this_step for mb in ens.members)


def test_members_must_share_the_schedule():
    ens, _ = make(dt_remap=[5, 5, 3])  # a member that remaps on its own schedule
    for _ in range(2):
        ens._batched_step()
    with pytest.raises(RuntimeError, match="remap together"):
        ens._batched_step()
    ens, _ = make()
    ens.members[1].config.dt_max = 30.0
    with pytest.raises(RuntimeError, match="share the time step"):
        ens._batched_step()


def test_radiation_due_for_some_members_runs_per_member():
    ens, log = make()
    ens.members[1].state.calculate_radiation_this_step = False
    ens._batched_radiation()
    assert "calculate_radiative_heating" not in batch_calls(log)
    assert [who for name, who in log if name == "calculate_radiative_heating"] == [0, 2]
    assert [who for name, who in log if name == "radiation_halo_exchange"] == [0, 2]
    assert not any(mb.state.calculate_radiation_this_step for mb in ens.members)


def test_batched_graph_period_records_once_then_replays():
    ens, log = make(graph_period=10)
    ens._batched_period()  # records the period: radiation eager, then the recording, then one launch
    calls = batch_calls(log)
    assert calls[0] == "calculate_radiative_heating"
    assert calls[1] == "graph_begin" and calls[-2] == "graph_end" and calls[-1] == "graph_launch"
    assert calls.count("dgrid_dynamics_step") == 10 and calls.count("perform_vertical_remapping") == 2
    assert ens.batch_pattern is not None and len(ens.batch_pattern) == 10
    log.clear()
    ens._batched_period()  # replays: radiation eager, one launch, nothing recorded
    assert batch_calls(log) == ["calculate_radiative_heating", "graph_launch"]
    assert all(mb.time == 20 * 60.0 for mb in ens.members)
    # a period whose time steps differ from the recorded ones is refused
    for mb in ens.members:
        mb.config.t_final = mb.time + 9.5 * 60.0
    with pytest.raises(RuntimeError, match="differ from the recorded"):
        ens._batched_period()
