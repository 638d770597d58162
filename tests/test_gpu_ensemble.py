"""The uneven ensemble: members of different grid sizes stepped together,
their launches on shared HIP streams, give each member exactly the state it
reaches when stepped alone on its own (default) stream."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_members_match_solo_runs(gpu_available, precision):
    from atmos import driver
    from atmos.config import Config
    from atmos.ensemble import Ensemble
    from atmos.ic import make_initial_state
    from atmos.model import Model
    specs = [(1, 64, 32, 12), (2, 32, 32, 10), (1, 32, 16, 6)]
    dx, dt, seed, steps = 100e3, 60.0, 77, 12
    print(f"ensemble {precision} {specs} x{steps}", flush=True)
    ens = Ensemble(specs, dx, dt, remap_every=5, radiation_every=10, precision=precision, nstreams=2,
                   seed=seed, t_noise=0.5)
    assert len(ens.members) == 4
    assert ens.cells == 64 * 32 * 12 + 2 * 32 * 32 * 10 + 32 * 16 * 6
    for _ in range(steps):
        assert ens.step()
    ens.sync()
    i = 0
    for count, nx, ny, nlay in specs:
        for _ in range(count):
            ic = make_initial_state(nx, ny, nlay, dx, seed=seed + i, t_noise=0.5)
            m = Model(nx, ny, nlay, dx, precision=precision)
            cfg = Config(dt_max=dt, t_final=1e12, dt_output=1e12, dt_radiation=10 * dt, dt_physics=1e12,
                         dt_remap=5 * dt, radiation_on=True, physics_on=False)
            st = driver.init_model(m, ic, cfg)
            t, n, stable = driver.run_model(m, cfg, st, max_steps=steps)
            assert stable and n == steps
            mb = ens.members[i]
            assert mb.time == t
            for name in ("dp", "pt", "ud", "vd", "net_flux"):
                assert np.array_equal(mb.model.get(name), m.get(name)), (i, name)
            i += 1
    ens.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batched_launches_match_member_launches(gpu_available, precision):
    """One launch per dynamics and remap stage for all members
    (atm_ensemble_*: a member table indexed by workgroup) gives every member
    the bits its own launches give: uneven grids, sizes that are not
    multiples of the tiles, levels of 6 to 12, three remaps and a radiation
    call in 12 steps."""
    from atmos.ensemble import Ensemble
    specs = [(1, 64, 32, 12), (2, 40, 24, 10), (1, 37, 23, 6), (1, 32, 16, 7)]
    dx, dt, seed, steps = 100e3, 60.0, 91, 12
    print(f"batched ensemble {precision} {specs} x{steps}", flush=True)
    a = Ensemble(specs, dx, dt, remap_every=4, radiation_every=10, precision=precision, nstreams=2, seed=seed)
    b = Ensemble(specs, dx, dt, remap_every=4, radiation_every=10, precision=precision, seed=seed,
                 launch="batched")
    for _ in range(steps):
        assert a.step()
        assert b.step()
    assert a.finish() and b.finish()
    a.sync()
    b.sync()
    for i, (ma, mb) in enumerate(zip(a.members, b.members)):
        assert ma.time == mb.time
        for name in ("dp", "pt", "ud", "vd", "net_flux", "gz", "pkap"):
            assert np.array_equal(ma.model.get(name), mb.model.get(name)), (i, name)
    a.close()
    b.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batched_graph_periods_match_batched_launches(gpu_available, precision):
    """Periods of batched steps recorded once as one graph (batched launches,
    the members' remaps, the stream joins) and replayed give every member
    the bits of the same steps issued eagerly; a replay after an odd number
    of eager buffer switches is refused."""
    from atmos.ensemble import Ensemble
    from atmos.model import AtmError
    specs = [(1, 64, 32, 12), (2, 40, 24, 10), (1, 37, 23, 6)]
    dx, dt, seed, period, periods = 100e3, 60.0, 93, 10, 3
    print(f"batched graph ensemble {precision} {specs} {periods}x{period}", flush=True)
    a = Ensemble(specs, dx, dt, remap_every=5, radiation_every=10, precision=precision, nstreams=2, seed=seed,
                 launch="batched")
    b = Ensemble(specs, dx, dt, remap_every=5, radiation_every=10, precision=precision, nstreams=2, seed=seed,
                 launch="batched", graph_period=period)
    for _ in range(period * periods):
        assert a.step()
    for _ in range(periods):
        assert b.period()
    assert a.finish() and b.finish()
    a.sync()
    b.sync()
    for i, (ma, mb) in enumerate(zip(a.members, b.members)):
        assert ma.time == mb.time
        for name in ("dp", "pt", "ud", "vd", "net_flux", "gz", "pkap"):
            assert np.array_equal(ma.model.get(name), mb.model.get(name)), (i, name)
    b._batched_step()  # one D step: the other copy of the fields is current
    with pytest.raises(AtmError, match="current copy"):
        b.batch.graph_launch(1)
    a.close()
    b.close()


def test_batched_launches_refuse_mixed_members(gpu_available):
    """The batched steps need every member on the fused path and in one
    precision; a member that is not is named in the error, at creation or at
    a later call, and the members' call order is checked as for one model
    (the remap after the D step)."""
    from atmos.config import Config
    from atmos import driver
    from atmos.ic import make_initial_state
    from atmos.model import AtmError, BatchedSteps, Model, stream_create, stream_destroy
    cfg = Config(dt_max=60.0, t_final=1e12, dt_output=1e12, dt_radiation=1e12, dt_physics=1e12, dt_remap=300.0,
                 radiation_on=False, physics_on=False)
    s = stream_create()

    def member(prec="fp64"):
        m = Model(32, 16, 6, 100e3, precision=prec)
        m.set_stream(s)
        driver.init_model(m, make_initial_state(32, 16, 6, 100e3), cfg)
        return m

    m0, m1 = member(), member("mixed")
    with pytest.raises(AtmError, match="one precision"):
        BatchedSteps([m0, m1])
    m1.close()
    m = member()
    m.set_path("reference")
    with pytest.raises(AtmError, match="fused path"):
        BatchedSteps([m0, m])
    m.set_path("fused")
    bs = BatchedSteps([m0, m])
    with pytest.raises(AtmError, match="follows"):  # the batched remap comes after the D step
        bs.perform_vertical_remapping()
    bs.cgrid_dynamics_step(30.0)
    bs.cgrid_halo_exchange()
    bs.dgrid_dynamics_step(60.0)
    with pytest.raises(AtmError, match="must follow"):  # the order of the calls, checked per member
        bs.cgrid_dynamics_step(30.0)
    bs.perform_vertical_remapping()
    bs.dgrid_halo_exchange()
    m.set_path("reference")  # a member that leaves the fused path later
    with pytest.raises(AtmError, match="fused path"):
        bs.cgrid_dynamics_step(30.0)
    bs.close()
    m.close()
    m0.close()
    stream_destroy(s)


@pytest.mark.parametrize("launch", ["streams", "batched"])
def test_ensemble_at_the_workload_sizes(gpu_available, launch):
    """The uneven ensemble as the workload states it: 24 members at
    192x192x48, 6 at 256x256x48, 2 at 384x384x64, mixed precision, radiation
    every 10 steps, remap every 5, seeded T noise per member; 10 steps.
    Every member stays stable and conserves its mass to FP64 rounding; one
    member of each size equals its solo run bit for bit."""
    from atmos import driver
    from atmos.config import Config
    from atmos.ensemble import Ensemble
    from atmos.ic import make_initial_state
    from atmos.model import Model
    specs = [(24, 192, 192, 48), (6, 256, 256, 48), (2, 384, 384, 64)]
    dx, dt, seed, steps = 50e3, 40.0, 500, 10
    print(f"ensemble at workload sizes {specs} x{steps}, {launch}", flush=True)
    ens = Ensemble(specs, dx, dt, remap_every=5, radiation_every=10, precision="mixed", nstreams=4, seed=seed,
                   t_noise=0.5, launch=launch)
    assert len(ens.members) == 32
    mass0 = [mb.model.get("dp").sum() for mb in ens.members]
    for _ in range(steps):
        assert ens.step()
    ens.sync()
    for mb, m0 in zip(ens.members, mass0):
        assert abs(mb.model.get("dp").sum() - m0) / m0 <= 1e-13
    firsts = {0: specs[0][1:], 24: specs[1][1:], 30: specs[2][1:]}
    for i, (nx, ny, nlay) in firsts.items():
        ic = make_initial_state(nx, ny, nlay, dx, seed=seed + i, t_noise=0.5)
        m = Model(nx, ny, nlay, dx, precision="mixed")
        cfg = Config(dt_max=dt, t_final=1e12, dt_output=1e12, dt_radiation=10 * dt, dt_physics=1e12,
                     dt_remap=5 * dt, radiation_on=True, physics_on=False)
        st = driver.init_model(m, ic, cfg)
        t, n, stable = driver.run_model(m, cfg, st, max_steps=steps)
        assert stable and n == steps
        for name in ("dp", "pt", "ud", "vd", "net_flux"):
            assert np.array_equal(ens.members[i].model.get(name), m.get(name)), (i, name)
        print(f"  member {i} {nx}x{ny}x{nlay} equals its solo run", flush=True)
    ens.close()


@pytest.mark.parametrize("mode", ["deferred", "graph"])
def test_deferred_checks_and_graph_periods_match_sync_steps(gpu_available, mode):
    """Deferred stability checks (read one step late, no drain) and 10-step
    periods replayed from each member's HIP graph give every member the
    same bits as steps with a waited check each (radiation every 10 steps,
    remap every 5; one member's grid is overhung by the tiles)."""
    from atmos.ensemble import Ensemble
    specs = [(1, 64, 32, 12), (2, 40, 24, 10), (1, 32, 16, 6)]
    dx, dt, seed = 100e3, 60.0, 91
    print(f"ensemble {mode} against sync checks {specs} x20", flush=True)
    kw = dict(remap_every=5, radiation_every=10, precision="mixed", nstreams=2, seed=seed, t_noise=0.5)
    ref = Ensemble(specs, dx, dt, checks="sync", **kw)
    for _ in range(20):
        assert ref.step()
    ens = Ensemble(specs, dx, dt, checks="deferred", graph_period=10 if mode == "graph" else 0, **kw)
    if mode == "graph":
        for _ in range(2):
            assert ens.period()
    else:
        for _ in range(20):
            assert ens.step()
    assert ens.finish()
    ref.sync()
    ens.sync()
    for i, (a, b) in enumerate(zip(ref.members, ens.members)):
        assert a.time == b.time
        for name in ("dp", "pt", "ud", "vd", "net_flux"):
            assert np.array_equal(a.model.get(name), b.model.get(name)), (i, name)
    ref.close()
    ens.close()


def test_deferred_check_reports_what_the_waited_check_reports(gpu_available):
    """The deferred check finds a NaN put into dp (a scan) and a NaN that a
    step carries into the state from net_flux (the D-step kernels' tracking
    word); it is refused when none is queued, when one is already queued,
    and the waited check and a graph recording are refused while one is."""
    from atmos import driver
    from atmos.config import Config
    from atmos.ic import make_initial_state
    from atmos.model import AtmError, Model
    nx, ny, nlay = 40, 24, 8
    print("deferred check detection and refusals", flush=True)
    ic = make_initial_state(nx, ny, nlay, 100e3)
    cfg = Config(dt_max=60.0, t_final=1e12, dt_output=1e12, dt_radiation=1e12, dt_physics=1e12, dt_remap=300.0,
                 radiation_on=False, physics_on=False)
    m = Model(nx, ny, nlay, 100e3)
    st = driver.init_model(m, ic, cfg)
    with pytest.raises(AtmError):
        m.check_stability_end()
    t = 0.0
    for _ in range(2):  # the first deferred check scans, the second reads the tracking word
        t += driver.step(m, cfg, st, t)
        m.check_stability_begin()
        with pytest.raises(AtmError):
            m.check_stability_begin()
        with pytest.raises(AtmError):
            m.check_stability()
        with pytest.raises(AtmError):
            m.graph_begin()
        assert m.check_stability_end()
    # a NaN through net_flux reaches pt in the next step's D kernels (tracked)
    a = m.get("net_flux")
    a[3, 10, 20] = np.nan
    m.put("net_flux", a)
    t += driver.step(m, cfg, st, t)
    m.check_stability_begin()
    assert not m.check_stability_end()
    # a NaN put into dp is found by the scan the put makes necessary
    m2 = Model(nx, ny, nlay, 100e3)
    st2 = driver.init_model(m2, ic, cfg)
    driver.step(m2, cfg, st2, 0.0)
    m2.check_stability_begin()
    assert m2.check_stability_end()
    dp = m2.get_full("dp")
    dp[2, 12, 15] = np.nan
    m2.put_full("dp", dp)
    m2.check_stability_begin()
    assert not m2.check_stability_end()
    assert not m2.check_stability()
    m.close()
    m2.close()
