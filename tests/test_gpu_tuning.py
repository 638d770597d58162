"""Launch-shape tuning (atm_autotune / atm_get_tuning / atm_set_tuning): every
candidate shape of the corner kernel (levels per workgroup 1, 2, 4, 8) gives
the same bits after steps with remaps, in both precisions and at a level
count that no group size divides; the autotuner's trial steps leave the run
unchanged; it picks one of the candidates; bad arguments and calls out of
order are refused."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DT = 60.0


def _run(precision, corner_levels, steps=7, tune_after=None, tiles=(8, 16), nx=64, ny=32):
    from atmos import driver
    from atmos.config import Config
    from atmos.ic import make_initial_state
    from atmos.model import Model
    nlay = 13
    ic = make_initial_state(nx, ny, nlay, 100e3, seed=5)
    m = Model(nx, ny, nlay, 100e3, precision=precision)
    assert m.path == "fused"
    m.set_tuning(corner_levels)
    assert m.tuning == corner_levels
    m.set_tiles(*tiles)
    assert m.tiles == tiles
    cfg = Config(dt_max=DT, t_final=1e12, dt_output=1e12, dt_radiation=1e12, dt_physics=1e12, dt_remap=3 * DT,
                 radiation_on=False, physics_on=False)
    st = driver.init_model(m, ic, cfg)
    if tune_after is None:
        _, n, stable = driver.run_model(m, cfg, st, max_steps=steps)
        assert stable and n == steps
    else:  # trial steps of the autotuner in the middle of the run
        t = cfg.t_initial
        for _ in range(tune_after):
            t += driver.step(m, cfg, st, t)
        m.autotune(reps=1)
        m.set_tuning(corner_levels)
        m.set_tiles(*tiles)
        for _ in range(steps - tune_after):
            t += driver.step(m, cfg, st, t)
        assert m.check_stability()
    return {name: m.get_full(name) for name in ("dp", "pt", "ud", "vd")}


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_corner_levels_do_not_change_results(gpu_available, precision):
    print(f"tuning {precision}", flush=True)
    ref = _run(precision, 4)
    for n in (1, 2, 8):
        got = _run(precision, n)
        for name in ref:
            assert np.array_equal(got[name], ref[name]), (n, name)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_tile_rows_do_not_change_results(gpu_available, precision):
    """dp/pt tiles of 32 x 16 and wind tiles of 32 x 8 (on 64 x 32 and on a
    grid the tiles overhang) give the bits of the default shapes."""
    for nx, ny in ((64, 32), (40, 24)):
        print(f"tile rows {precision} {nx}x{ny}", flush=True)
        ref = _run(precision, 4, nx=nx, ny=ny)
        for tiles in ((16, 16), (8, 8), (16, 8)):
            got = _run(precision, 4, tiles=tiles, nx=nx, ny=ny)
            for name in ref:
                assert np.array_equal(got[name], ref[name]), (nx, ny, tiles, name)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_autotune_trial_steps_leave_the_run_unchanged(gpu_available, precision):
    """The trial steps write the other copy of the fields and the step's
    work arrays; a remap right after the next step reads the other copy's
    wind ring, so the run goes through one (remap every 3 steps, tuned after
    step 2)."""
    print(f"autotune mid-run {precision}", flush=True)
    ref = _run(precision, 4)
    got = _run(precision, 4, tune_after=2)
    for name in ref:
        assert np.array_equal(got[name], ref[name]), name


def test_autotune_choice_and_arguments(gpu_available):
    from atmos import driver
    from atmos.config import Config
    from atmos.ic import make_initial_state
    from atmos.model import AtmError, Model
    m = Model(64, 32, 13, 100e3)
    assert m.tuning == 4  # default
    with pytest.raises(AtmError) as e:  # the trial steps need the dynamics initialised
        m.autotune(reps=2)
    assert e.value.code == 5
    cfg = Config(dt_max=DT, t_final=1e12, dt_output=1e12, dt_radiation=1e12, dt_physics=1e12, dt_remap=3 * DT,
                 radiation_on=False, physics_on=False)
    driver.init_model(m, make_initial_state(64, 32, 13, 100e3, seed=5), cfg)
    assert m.autotune(reps=2) in (1, 2, 4, 8)
    assert m.tiles[0] in (8, 16) and m.tiles[1] in (8, 16)
    for bad in (lambda: m.autotune(reps=0), lambda: m.set_tuning(3), lambda: m.set_tiles(4, 16),
                lambda: m.set_tiles(8, 32)):
        with pytest.raises(AtmError) as e:
            bad()
        assert e.value.code == 1
