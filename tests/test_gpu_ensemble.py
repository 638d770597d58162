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
