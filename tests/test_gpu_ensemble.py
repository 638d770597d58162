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


def test_ensemble_at_the_workload_sizes(gpu_available):
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
    print(f"ensemble at workload sizes {specs} x{steps}", flush=True)
    ens = Ensemble(specs, dx, dt, remap_every=5, radiation_every=10, precision="mixed", nstreams=4, seed=seed,
                   t_noise=0.5)
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
