"""Graphs of whole steps (atm_graph_begin / end / launch): every launch of a
recorded 10-step cycle (two remaps) leaves the state an eager run of the
same calls reaches, bit for bit, with eager work (checks, reads) between
the launches; calls that wait on the device are refused while recording,
and a recording that does not end where it started is refused."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DT = 60.0


def _model(path, precision):
    from atmos import driver
    from atmos.config import Config
    from atmos.ic import make_initial_state
    from atmos.model import Model
    nx, ny, nlay = 64, 32, 12
    ic = make_initial_state(nx, ny, nlay, 100e3, seed=21)
    m = Model(nx, ny, nlay, 100e3, precision=precision)
    if m.path != path:
        m.set_path(path)
    cfg = Config(dt_max=DT, t_final=1e12, dt_output=1e12, dt_radiation=1e12, dt_physics=1e12, dt_remap=1e12,
                 radiation_on=False, physics_on=False)
    driver.init_model(m, ic, cfg)
    return m


def _step(m, k):
    m.cgrid_dynamics_step(DT / 2)
    m.cgrid_halo_exchange()
    m.dgrid_dynamics_step(DT)
    if k % 5 == 4:
        m.perform_vertical_remapping()
    m.dgrid_halo_exchange()


@pytest.mark.parametrize("path,precision", [("fused", "fp64"), ("fused", "mixed"), ("reference", "fp64")])
def test_graph_launches_match_eager_steps(gpu_available, path, precision):
    from atmos.model import AtmError
    print(f"graph {path} {precision}", flush=True)
    eager, rec = _model(path, precision), _model(path, precision)
    rec.graph_begin()
    for k in range(10):
        _step(rec, k)
    with pytest.raises(AtmError) as e:
        rec.check_stability()
    assert e.value.code == 5
    rec.graph_end()
    # nothing ran yet: the recorded model still holds its initial state
    assert np.array_equal(rec.get("dp"), eager.get("dp"))
    for r in range(3):
        rec.graph_launch(1)
        for k in range(10):
            _step(eager, k)
        assert rec.check_stability() and eager.check_stability()
        for name in ("dp", "pt", "ud", "vd"):
            assert np.array_equal(rec.get(name), eager.get(name)), (r, name)
    # two launches in a row, then an eager step on the recorded model
    with pytest.raises(AtmError) as e:
        rec.graph_launch(-1)
    assert e.value.code == 1 and "times" in str(e.value)
    rec.graph_launch(0)  # nothing runs
    rec.graph_launch(2)
    for _ in range(2):
        for k in range(10):
            _step(eager, k)
    _step(rec, 0)
    _step(eager, 0)
    for name in ("dp", "pt", "ud", "vd"):
        assert np.array_equal(rec.get(name), eager.get(name)), name
    rec.graph_free()


def test_graph_must_end_where_it_started(gpu_available):
    from atmos.model import AtmError
    m = _model("fused", "fp64")
    m.graph_begin()
    _step(m, 0)  # one D step: the other copy is current
    with pytest.raises(AtmError) as e:
        m.graph_end()
    assert e.value.code == 5
    with pytest.raises(AtmError):
        m.graph_launch(1)  # nothing recorded
    # the refused recording had no effect: the context steps on from where
    # it was
    ref = _model("fused", "fp64")
    _step(m, 1)
    _step(ref, 1)
    assert m.check_stability()
    for name in ("dp", "pt", "ud", "vd"):
        assert np.array_equal(m.get(name), ref.get(name)), name


def _step_acc(m, k):
    _step(m, k)
    m.accumulate_output(DT)


def test_graph_advances_the_output_time_per_launch(gpu_available):
    """accumulate_output recorded in a graph: each launch adds the recorded
    dt values to the accumulation time, so the means equal those of eager
    steps; a mean before any launch and after a refused recording finds
    nothing accumulated."""
    from atmos.model import AtmError
    print("graph accumulation", flush=True)
    eager, rec = _model("fused", "fp64"), _model("fused", "fp64")
    for m in (eager, rec):
        m.init_output()
    rec.graph_begin()
    _step_acc(rec, 0)  # one D step: refused at the end
    with pytest.raises(AtmError):
        rec.graph_end()
    rec.graph_begin()
    for k in range(2):
        _step_acc(rec, k)
    rec.graph_end()
    with pytest.raises(AtmError) as e:
        rec.output_mean()
    assert e.value.code == 5 and "nothing accumulated" in str(e.value)
    rec.graph_launch(3)
    for _ in range(3):
        for k in range(2):
            _step_acc(eager, k)
    a, b = rec.output_mean(), eager.output_mean()
    assert np.array_equal(a, b)
    rec.graph_free()


def test_graph_recorded_after_a_d_step_forms_playkap(gpu_available):
    """A recording that starts after an eager D step (playkap valid then)
    and ends with a remap, with a nonzero net flux: every launch forms
    playkap in its first C step, so two launches match eager steps."""
    print("graph playkap", flush=True)
    eager, rec = _model("fused", "fp64"), _model("fused", "fp64")
    rng = np.random.default_rng(5)
    nf = rng.uniform(-50.0, 50.0, size=(eager.nlay + 1, eager.ny, eager.nx))
    for m in (eager, rec):
        m.put("net_flux", nf)
        m.halo_exchange("net_flux")
        _step(m, 0)
    rec.graph_begin()
    for k in range(10):
        _step(rec, k)
    rec.graph_end()
    rec.graph_launch(2)
    for _ in range(2):
        for k in range(10):
            _step(eager, k)
    for name in ("dp", "pt", "ud", "vd"):
        assert np.array_equal(rec.get(name), eager.get(name)), name
    rec.graph_free()


def test_stability_scanned_after_a_reference_path_launch(gpu_available):
    """The reference-path D step does not report into the stability word:
    after a launch of such a recording the check scans the fields and sees
    the non-finite values the launch produced."""
    print("graph stability tracking", flush=True)
    m = _model("reference", "fp64")
    ud = m.get("ud")
    ud[3, 10, 20] = 1e200  # finite; a step overflows from it
    m.put("ud", ud)
    m.halo_exchange("ud")
    assert m.check_stability()  # finite, and now tracked
    m.graph_begin()
    for k in range(2):
        _step(m, k)
    m.graph_end()
    assert m.check_stability()  # nothing ran
    m.graph_launch(1)
    assert not m.check_stability()
    m.graph_free()


def test_refused_radiation_init_stays_uninitialised(gpu_available):
    """nlay above the radiation kernel's limit is refused on every call, and
    radiation then refuses to run (no kernel is launched)."""
    from atmos.model import AtmError, Model
    print("radiation init refusal", flush=True)
    m = Model(8, 8, 257, 100e3)
    for _ in range(2):
        with pytest.raises(AtmError) as e:
            m.allocate_radiation_arrays()
        assert e.value.code == 1 and "nlay" in str(e.value)
    with pytest.raises(AtmError) as e:
        m.calculate_radiative_heating(0.0)
    assert e.value.code == 5
    m.close()
