"""GPU parity: the HIP library against the CPU oracle on the same seeded
inputs, stage by stage and over multi-step runs (FP64, 1e-12 norm-relative).
"""
import numpy as np
import pytest

from gpu_common import PARITY, compare, make_pair

pytestmark = pytest.mark.gpu


def _print(tag, **kw):
    print(tag, " ".join(f"{k}={v:.3e}" for k, v in kw.items()), flush=True)


def test_halo_fill_closed_form(gpu_available):
    from atmos.model import Model
    m = Model(19, 11, 5, 1e5)
    rng = np.random.default_rng(0)
    a = rng.standard_normal((5, 11 + 14, 19 + 14))
    m.put_full("dp", a)
    m.halo_exchange("dp")
    b = m.get_full("dp")
    ii = (np.arange(19 + 14) - 7) % 19 + 7
    jj = (np.arange(11 + 14) - 7) % 11 + 7
    expect = a[:, jj][:, :, ii]
    assert np.array_equal(b, expect)


@pytest.mark.parametrize("shape", [(32, 32, 20), (40, 24, 16)])
def test_cgrid_step_stage(gpu_available, shape):
    gpu, orc, cfg, sg, so, ic = make_pair(*shape)
    for b in (gpu, orc):
        b.cgrid_dynamics_step(30.0)
    # dpc/ptc valid on is+5..ie-5, uc/vc updated on the interior
    errs = {n: compare(gpu, orc, n, (5, 5)) for n in ("dpc", "ptc")}
    errs.update({n: compare(gpu, orc, n) for n in ("uc", "vc")})
    _print(f"cgrid {shape}", **errs)
    for n, e in errs.items():
        assert e <= PARITY, (n, e)


@pytest.mark.parametrize("shape", [(32, 32, 20), (40, 24, 16)])
def test_dgrid_step_stage(gpu_available, shape):
    gpu, orc, cfg, sg, so, ic = make_pair(*shape)
    for b in (gpu, orc):
        b.cgrid_dynamics_step(30.0)
        b.cgrid_halo_exchange()
        b.dgrid_dynamics_step(60.0)
    errs = {n: compare(gpu, orc, n, (4, 4)) for n in ("dp", "pt")}
    errs.update({n: compare(gpu, orc, n) for n in ("ud", "vd")})
    _print(f"dgrid {shape}", **errs)
    for n, e in errs.items():
        assert e <= PARITY, (n, e)


def test_remap_stage(gpu_available):
    gpu, orc, cfg, sg, so, ic = make_pair(32, 32, 20)
    for b in (gpu, orc):
        b.cgrid_dynamics_step(30.0)
        b.cgrid_halo_exchange()
        b.dgrid_dynamics_step(60.0)
        b.perform_vertical_remapping()
    errs = {n: compare(gpu, orc, n) for n in ("dp", "pt", "ud", "vd")}
    _print("remap", **errs)
    for n, e in errs.items():
        assert e <= PARITY, (n, e)


def test_radiation_stage(gpu_available):
    gpu, orc, cfg, sg, so, ic = make_pair(24, 20, 20)
    for b in (gpu, orc):
        b.calculate_radiative_heating(0.0)
        b.radiation_halo_exchange()
    e = compare(gpu, orc, "net_flux", (0, 0))
    _print("radiation", net_flux=e)
    assert e <= PARITY


@pytest.mark.parametrize("shape,steps,rad", [((32, 32, 20), 10, True), ((64, 64, 20), 100, False),
                                             ((32, 32, 72), 6, True), ((8, 8, 4), 6, True),
                                             ((37, 23, 10), 6, True), ((32, 16, 6), 6, True),
                                             ((32, 32, 13), 6, True)])
def test_driver_run(gpu_available, shape, steps, rad):
    from atmos import driver
    gpu, orc, cfg, sg, so, ic = make_pair(*shape)
    cfg.radiation_on = rad
    cfg.dt_radiation = 600.0 if rad else 1e12
    for b, st in ((gpu, sg), (orc, so)):
        t, n, stable = driver.run_model(b, cfg, st, max_steps=steps)
        assert stable and n == steps
    errs = {n: compare(gpu, orc, n) for n in ("dp", "pt", "ud", "vd")}
    if rad:
        errs["net_flux"] = compare(gpu, orc, "net_flux")
    _print(f"run {shape} x{steps} rad={rad}", **errs)
    for n, e in errs.items():
        assert e <= PARITY, (n, e)


@pytest.mark.parametrize("shape,steps", [((32, 32, 20), 11), ((64, 32, 16), 6), ((37, 23, 10), 11),
                                         ((200, 200, 20), 6), ((8, 8, 4), 11), ((33, 17, 5), 11)])
def test_fused_path_bitwise_equals_reference_path(gpu_available, shape, steps):
    """The fused LDS-tiled kernels and the reference-structured kernels
    produce identical bits, including remap and radiation steps; on grids
    that are not multiples of the tiles (37x23, 200x200, 8x8, 33x17) the
    last tiles overhang the domain and store only the cells inside."""
    from atmos import driver
    from atmos.config import Config
    from atmos.ic import make_initial_state
    from atmos.model import Model
    nx, ny, nlay = shape
    ic = make_initial_state(nx, ny, nlay, 100e3)
    cfg = dict(dt_max=60.0, t_final=1e12, dt_output=1e12, dt_radiation=300.0, dt_physics=1e12,
               dt_remap=180.0, radiation_on=True, physics_on=False)
    a, b = Model(nx, ny, nlay, 100e3), Model(nx, ny, nlay, 100e3)
    assert a.path == "fused"
    b.set_path("reference")
    for m in (a, b):
        st = driver.init_model(m, ic, Config(**cfg))
        driver.run_model(m, Config(**cfg), st, max_steps=steps)
    for name in ("dp", "pt", "ud", "vd", "net_flux"):
        x, y = a.get_full(name), b.get_full(name)
        assert np.array_equal(x, y), (name, float(np.max(np.abs(x - y))))
    print(f"fused == reference bitwise {shape} x{steps}", flush=True)


def _unstable(ic, rng, frac=0.3):
    """Perturb pt so that some interfaces are statically unstable and the
    dry convective adjustment mixes there."""
    pt = ic["pt"].copy()
    mask = rng.random(pt.shape) < frac
    pt[mask] *= 1.0 + rng.uniform(0.01, 0.05, mask.sum())
    ic = dict(ic)
    ic["pt"] = pt
    return ic


def test_physics_stage(gpu_available):
    """calculate_physics_tendencies + apply_physics_tendencies (dry
    convective adjustment, surface heating) on a partly unstable state."""
    import oracle as O
    from atmos import driver
    from atmos.config import Config
    from atmos.ic import make_initial_state
    from atmos.model import Model
    nx, ny, nlay = 32, 24, 16
    rng = np.random.default_rng(77)
    ic = _unstable(make_initial_state(nx, ny, nlay, 100e3), rng)
    cfg = Config(dt_max=60.0, t_final=1e12, dt_output=1e12, dt_radiation=1e12, dt_physics=300.0,
                 dt_remap=300.0, radiation_on=False, physics_on=True)
    gpu, orc = Model(nx, ny, nlay, 100e3), O.Oracle(nx, ny, nlay, 100e3, 100e3)
    nf = np.zeros((nlay + 1, ny, nx))
    nf[0] = rng.uniform(-80.0, 150.0, (ny, nx))
    for b in (gpu, orc):
        driver.init_model(b, ic, Config(**cfg.__dict__))
        b.put("net_flux", nf)
        b.calculate_physics_tendencies()
        b.apply_physics_tendencies(60.0)
    errs = {n: compare(gpu, orc, n) for n in ("dp", "pt", "ud", "vd", "ts")}
    errs.update({n: compare(gpu, orc, n, (1, 1)) for n in ("gz", "plev", "pkap")})
    _print("physics stage", **errs)
    moved = np.max(np.abs(gpu.get("pt") - ic["pt"]))
    assert moved > 0.0
    for n, e in errs.items():
        assert e <= PARITY, (n, e)


@pytest.mark.parametrize("path", ["fused", "reference"])
def test_driver_run_with_physics(gpu_available, path):
    """The CPU reference configuration's schedule (dry mixing on, physics and
    remap every 5 minutes, dt = 60 s) for 20 steps with radiation on."""
    import oracle as O
    from atmos import driver
    from atmos.config import Config
    from atmos.ic import make_initial_state
    from atmos.model import Model
    nx, ny, nlay = 64, 64, 20
    rng = np.random.default_rng(3)
    ic = _unstable(make_initial_state(nx, ny, nlay, 100e3), rng, 0.1)
    cfg = dict(dt_max=60.0, t_final=1e12, dt_output=1e12, dt_radiation=600.0, dt_physics=300.0,
               dt_remap=300.0, radiation_on=True, physics_on=True)
    gpu, orc = Model(nx, ny, nlay, 100e3), O.Oracle(nx, ny, nlay, 100e3, 100e3)
    if path != gpu.path:
        gpu.set_path(path)
    for b in (gpu, orc):
        st = driver.init_model(b, ic, Config(**cfg))
        t, n, stable = driver.run_model(b, Config(**cfg), st, max_steps=20)
        assert stable and n == 20
    errs = {n: compare(gpu, orc, n) for n in ("dp", "pt", "ud", "vd", "ts", "net_flux")}
    _print(f"driver physics {path} 64x64x20 x20", **errs)
    for n, e in errs.items():
        assert e <= PARITY, (n, e)


@pytest.mark.parametrize("path", ["fused", "reference"])
def test_unequal_spacing(gpu_available, path):
    """dx != dy, the one case where Q2 shows: the y Courant numbers use
    dt/dx (mod_sw_dyn.f90:133, 349, 448) while the denominators, the
    Laplacians and the pressure gradient use dy.  12 steps with two remaps
    and radiation at 100 km x 75 km."""
    import oracle as O
    from atmos import driver
    from atmos.config import Config
    from atmos.ic import make_initial_state
    from atmos.model import Model
    nx, ny, nlay, dx, dy = 64, 32, 16, 100e3, 75e3
    ic = make_initial_state(nx, ny, nlay, dx, dy, seed=5)
    cfg = dict(dt_max=60.0, t_final=1e12, dt_output=1e12, dt_radiation=600.0, dt_physics=1e12,
               dt_remap=300.0, radiation_on=True, physics_on=False)
    gpu, orc = Model(nx, ny, nlay, dx, dy), O.Oracle(nx, ny, nlay, dx, dy)
    if path != gpu.path:
        gpu.set_path(path)
    for b in (gpu, orc):
        st = driver.init_model(b, ic, Config(**cfg))
        t, n, stable = driver.run_model(b, Config(**cfg), st, max_steps=12)
        assert stable and n == 12
    errs = {n: compare(gpu, orc, n) for n in ("dp", "pt", "ud", "vd", "net_flux")}
    _print(f"dx != dy {path} {nx}x{ny}x{nlay} x12", **errs)
    for n, e in errs.items():
        assert e <= PARITY, (n, e)


def test_radiation_table_range_error(gpu_available):
    """A temperature outside the gas-optics table (the reference aborts,
    mod_gas_optics.f90:4562-4592) comes back as ATM_ERR_TABLE_RANGE."""
    from atmos import driver
    from atmos.config import Config
    from atmos.ic import make_initial_state
    from atmos.model import AtmError, Model
    nx, ny, nlay = 32, 16, 8
    ic = make_initial_state(nx, ny, nlay, 100e3)
    ic = dict(ic)
    ic["pt"] = ic["pt"] * 3.0  # ~900 K: beyond every temperature row
    m = Model(nx, ny, nlay, 100e3)
    cfg = Config(dt_max=60.0, t_final=1e12, dt_output=1e12, dt_radiation=1e12, dt_physics=1e12,
                 dt_remap=1e12, radiation_on=True, physics_on=False)
    driver.init_model(m, ic, cfg)
    with pytest.raises(AtmError) as e:
        m.calculate_radiative_heating(0.0)
    assert e.value.code == 3


def test_radiation_nan_is_range_error(gpu_available):
    """A NaN surface temperature fails the table's range test (the
    reference's test lets NaN through to a read before the table): the call
    reports ATM_ERR_TABLE_RANGE instead of reading outside the table."""
    from atmos import driver
    from atmos.config import Config
    from atmos.ic import make_initial_state
    from atmos.model import AtmError, Model
    nx, ny, nlay = 32, 16, 8
    ic = dict(make_initial_state(nx, ny, nlay, 100e3))
    ic["ts"] = ic["ts"].copy()
    ic["ts"].flat[37] = np.nan
    m = Model(nx, ny, nlay, 100e3)
    cfg = Config(dt_max=60.0, t_final=1e12, dt_output=1e12, dt_radiation=1e12, dt_physics=1e12,
                 dt_remap=1e12, radiation_on=True, physics_on=False)
    driver.init_model(m, ic, cfg)
    with pytest.raises(AtmError) as e:
        m.calculate_radiative_heating(0.0)
    assert e.value.code == 3


def test_radiation_on_table_points(gpu_available):
    """Surface temperatures on the Planck table's points, both ends among
    them, take the lookup's exact-hit branch (mod_gas_optics.f90:4576-4577)
    and the edges of the table search; fluxes as the oracle's."""
    import oracle as O
    from atmos import driver
    from atmos.config import Config
    from atmos.ic import make_initial_state
    from atmos.model import Model
    nx, ny, nlay = 24, 20, 20
    ic = dict(make_initial_state(nx, ny, nlay, 100e3, seed=11))
    ic["ts"] = ic["ts"].copy()
    pts = [120.0, 121.0, 200.0, 287.0, 349.0, 350.0, 120.5, 349.5, 300.0 + 2.0 ** -40]
    for c, v in enumerate(pts):
        ic["ts"].flat[c * 7] = v
    cfg = Config(dt_max=60.0, t_final=1e12, dt_output=1e12, dt_radiation=1e12, dt_physics=1e12,
                 dt_remap=1e12, radiation_on=False, physics_on=False)
    gpu, orc = Model(nx, ny, nlay, 100e3), O.Oracle(nx, ny, nlay, 100e3, 100e3)
    for b in (gpu, orc):
        driver.init_model(b, ic, Config(**cfg.__dict__))
        b.calculate_radiative_heating(0.0)
    e = compare(gpu, orc, "net_flux")
    _print("radiation on table points", net_flux=e)
    assert e <= PARITY


def test_call_order_errors(gpu_available):
    """The state machine of the boundary: a second D step before the halo
    fill, and physics before its allocation, are refused (status 5)."""
    from atmos import driver
    from atmos.config import Config
    from atmos.ic import make_initial_state
    from atmos.model import AtmError, Model
    nx, ny, nlay = 32, 16, 8
    m = Model(nx, ny, nlay, 100e3)
    assert m.path == "fused"
    with pytest.raises(AtmError) as e:
        m.calculate_physics_tendencies()
    assert e.value.code == 5
    ic = make_initial_state(nx, ny, nlay, 100e3)
    cfg = Config(dt_max=60.0, t_final=1e12, dt_output=1e12, dt_radiation=1e12, dt_physics=1e12,
                 dt_remap=1e12, radiation_on=False, physics_on=False)
    driver.init_model(m, ic, cfg)
    m.cgrid_dynamics_step(30.0)
    m.cgrid_halo_exchange()
    m.dgrid_dynamics_step(60.0)
    with pytest.raises(AtmError) as e:
        m.dgrid_dynamics_step(60.0)
    assert e.value.code == 5
    m.dgrid_halo_exchange()
    assert m.check_stability()


def test_output_accumulation(gpu_available):
    """accumulate_output every step and write_output's means, GPU vs oracle,
    with radiation on (net_flux) over 6 steps."""
    import oracle as O
    from atmos import driver
    from atmos.config import Config
    from atmos.ic import make_initial_state
    from atmos.model import Model
    nx, ny, nlay = 32, 16, 12
    ic = make_initial_state(nx, ny, nlay, 100e3)
    cfg = dict(dt_max=60.0, t_final=1e12, dt_output=1e12, dt_radiation=120.0, dt_physics=1e12,
               dt_remap=180.0, radiation_on=True, physics_on=False)
    gpu, orc = Model(nx, ny, nlay, 100e3), O.Oracle(nx, ny, nlay, 100e3, 100e3)
    means = []
    for b in (gpu, orc):
        c = Config(**cfg)
        st = driver.init_model(b, ic, c)
        b.init_output()
        t = 0.0
        for _ in range(6):
            dt = driver.step(b, c, st, t)
            assert b.check_stability()
            t += dt
            b.accumulate_output(dt)
        means.append(b.output_mean())
    a, o = means
    errs = {f: float(np.max(np.abs(a[i] - o[i])) / max(np.max(np.abs(o[i])), 1e-300))
            for i, f in enumerate(("dp", "pt", "u", "v", "ts", "nflx"))}
    _print("output means x6", **errs)
    for f, e in errs.items():
        assert e <= PARITY, (f, e)


@pytest.mark.parametrize("route", ["net_flux", "gz"])
def test_stability_check_sees_nonfinite_from_a_step(gpu_available, route):
    """A NaN entering through a field the check does not scan (net_flux:
    the C-step heating rate -> pt; surface gz: the pressure gradient -> ud,
    vd) reaches the prognostic state in the next step, and the check finds
    it while it answers from the D-step kernels' tracking word."""
    from atmos import driver
    gpu, orc, cfg, sg, so, ic = make_pair(64, 32, 8)
    assert gpu.path == "fused"
    t = 0.0
    for _ in range(2):  # the first check scans the fields, the second reads the word
        t += driver.step(gpu, cfg, sg, t)
        assert gpu.check_stability()
    a = gpu.get(route)
    a[0 if route == "gz" else 3, 10, 20] = np.nan
    gpu.put(route, a)
    t += driver.step(gpu, cfg, sg, t)
    bad = [n for n in ("dp", "pt", "ud", "vd") if not np.all(np.isfinite(gpu.get(n)))]
    print(f"{route}: non-finite after the step in {bad}", flush=True)
    assert bad
    assert not gpu.check_stability()
    assert not gpu.check_stability()


def _ps_T(m, ptop):
    kappa = float(np.float32(287.052874)) / 1005.0
    dp, pt = m.get("dp"), m.get("pt")
    plev = np.concatenate([np.cumsum(dp[::-1], axis=0)[::-1] + ptop, np.full((1,) + dp.shape[1:], ptop)])
    pk = plev ** kappa
    return plev[0], pt * ((pk[1:] - pk[:-1]) / np.log(pk[1:] / pk[:-1]))


@pytest.mark.parametrize("shape,dy", [((64, 64, 20), 100e3), ((64, 64, 20), 75e3), ((40, 24, 8), 100e3)])
def test_mixed_precision_error_and_mass(gpu_available, shape, dy):
    """ATM_PRECISION_MIXED against the FP64 path: 30 steps with remaps every
    5 at 64x64x20 from the same state, at dx = dy and dx != dy, and on a
    grid the tiles overhang (40x24x8).  Surface pressure and temperature
    agree to float-level error, and the total mass (sum of dp) changes only
    by FP64 rounding in both forms (the float fluxes are shared by the two
    cells of each face)."""
    from atmos import driver
    from atmos.config import Config
    from atmos.ic import make_initial_state
    from atmos.model import AtmError, Model
    nx, ny, nlay = shape
    ic = make_initial_state(nx, ny, nlay, 100e3, dy)
    kw = dict(dt_max=60.0, t_final=1e12, dt_output=1e12, dt_radiation=1e12, dt_physics=1e12, dt_remap=300.0,
              radiation_on=False, physics_on=False)
    out = {}
    for prec in ("fp64", "mixed"):
        m = Model(nx, ny, nlay, 100e3, dy, precision=prec)
        cfg = Config(**kw)
        st = driver.init_model(m, ic, cfg)
        mass0 = m.get("dp").sum()
        t = 0.0
        for _ in range(30):
            t += driver.step(m, cfg, st, t)
            assert m.check_stability()
        ps, T = _ps_T(m, ic["ak"][0])
        out[prec] = dict(ps=ps, T=T, ud=m.get("ud"), vd=m.get("vd"), drift=(m.get("dp").sum() - mass0) / mass0)
        if prec == "mixed":
            with pytest.raises(AtmError):
                m.set_path("reference")
    a, b = out["mixed"], out["fp64"]
    rr = {k: float(np.sqrt(np.mean((a[k] - b[k]) ** 2)) / np.sqrt(np.mean(b[k] ** 2))) for k in ("ps", "T", "ud", "vd")}
    _print(f"mixed vs fp64 x30 {shape} dy={dy:.0f} rel rms", **rr, drift_mixed=abs(a["drift"]), drift_fp64=abs(b["drift"]))
    assert rr["ps"] <= 1e-6 and rr["T"] <= 1e-6, rr
    assert rr["ud"] <= 1e-4 and rr["vd"] <= 1e-4, rr
    assert abs(a["drift"]) <= 1e-13 and abs(b["drift"]) <= 1e-13


def test_full_size_parity(gpu_available):
    """The benchmark workload itself: 512x512x64, dx = 25 km, dt = 20 s,
    remap every 5 steps, 6 steps (one remap) from the seeded state, GPU
    fused path against the (threaded) oracle, 1e-12 norm-relative on the
    prognostic fields, and the total mass change of both at FP64 rounding."""
    import oracle as O
    from atmos import driver
    from atmos.config import Config
    from atmos.ic import make_initial_state
    from atmos.model import Model
    nx, ny, nlay, dx = 512, 512, 64, 25e3
    print(f"full size {nx}x{ny}x{nlay}", flush=True)
    ic = make_initial_state(nx, ny, nlay, dx)
    kw = dict(dt_max=20.0, t_final=1e12, dt_output=1e12, dt_radiation=1e12, dt_physics=1e12, dt_remap=100.0,
              radiation_on=False, physics_on=False)
    gpu, orc = Model(nx, ny, nlay, dx), O.Oracle(nx, ny, nlay, dx, dx)
    assert gpu.path == "fused"
    mass = {}
    for b in (gpu, orc):
        cfg = Config(**kw)
        st = driver.init_model(b, ic, cfg)
        t = 0.0
        for _ in range(6):
            t += driver.step(b, cfg, st, t)
    errs = {n: compare(gpu, orc, n) for n in ("dp", "pt", "ud", "vd")}
    _print("full size x6", **errs)
    for n, e in errs.items():
        assert e <= PARITY, (n, e)
    m0 = ic["dp"].sum()
    assert abs(gpu.get("dp").sum() - m0) / m0 <= 1e-13


def test_mixed_precision_radiation(gpu_available):
    """Radiation of the mixed form (float g-point optics, double adding
    method and sums) against the FP64 path on the same state: net_flux to
    float-level error (1e-4 norm-relative; the two-stream coefficients lose
    digits where 1 - lambda^2 is small)."""
    from atmos import driver
    from atmos.config import Config
    from atmos.ic import make_initial_state
    from atmos.model import Model
    n, L = 64, 20
    ic = make_initial_state(n, n, L, 25e3)
    out = {}
    for prec in ("fp64", "mixed"):
        m = Model(n, n, L, 25e3, precision=prec)
        driver.init_model(m, ic, Config(dt_max=20.0, t_final=1e12, dt_output=1e12, dt_radiation=1e12,
                                        dt_physics=1e12, dt_remap=100.0, radiation_on=False, physics_on=False))
        m.calculate_radiative_heating(0.0)
        out[prec] = m.get("net_flux")
    a, b = out["mixed"], out["fp64"]
    e = float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
    _print("mixed radiation", net_flux=e)
    assert e <= 1e-4


def test_cpu_reference_configuration_100_steps(gpu_available):
    """The CPU reference configuration's exact schedule: 64x64x20, dx = dy =
    100 km, dt = 60 s, seeded recipe, radiation off, dry convective mixing
    on (physics and remap every 5 minutes), 100 steps; GPU fused path
    against the oracle at 1e-12 on every prognostic field and ts."""
    import oracle as O
    from atmos import driver
    from atmos.config import Config
    from atmos.ic import make_initial_state
    from atmos.model import Model
    nx, ny, nlay = 64, 64, 20
    print("cpu reference configuration 64x64x20 x100 physics on, radiation off", flush=True)
    ic = make_initial_state(nx, ny, nlay, 100e3)
    cfg = dict(dt_max=60.0, t_final=1e12, dt_output=1e12, dt_radiation=1e12, dt_physics=300.0,
               dt_remap=300.0, radiation_on=False, physics_on=True)
    gpu, orc = Model(nx, ny, nlay, 100e3), O.Oracle(nx, ny, nlay, 100e3, 100e3)
    for b in (gpu, orc):
        st = driver.init_model(b, ic, Config(**cfg))
        t, n, stable = driver.run_model(b, Config(**cfg), st, max_steps=100)
        assert stable and n == 100
    errs = {n: compare(gpu, orc, n) for n in ("dp", "pt", "ud", "vd", "ts")}
    _print("cpu configuration x100", **errs)
    for n, e in errs.items():
        assert e <= PARITY, (n, e)


def test_full_size_parity_100_steps(gpu_available):
    """The benchmark workload's whole run: 512x512x64, dx = 25 km, dt = 20 s,
    remap every 5 steps, 100 steps (20 remaps) from the seeded state, GPU
    fused path against the threaded oracle at 1e-12 on the prognostic
    fields; the total mass of both conserved to FP64 rounding."""
    import oracle as O
    from atmos import driver
    from atmos.config import Config
    from atmos.ic import make_initial_state
    from atmos.model import Model
    nx, ny, nlay, dx = 512, 512, 64, 25e3
    print(f"full size {nx}x{ny}x{nlay} x100", flush=True)
    ic = make_initial_state(nx, ny, nlay, dx)
    kw = dict(dt_max=20.0, t_final=1e12, dt_output=1e12, dt_radiation=1e12, dt_physics=1e12, dt_remap=100.0,
              radiation_on=False, physics_on=False)
    gpu, orc = Model(nx, ny, nlay, dx), O.Oracle(nx, ny, nlay, dx, dx)
    for b in (gpu, orc):
        cfg = Config(**kw)
        st = driver.init_model(b, ic, cfg)
        t = 0.0
        for k in range(100):
            t += driver.step(b, cfg, st, t)
            if b is orc and (k + 1) % 20 == 0:
                print(f"  oracle step {k + 1}", flush=True)
    errs = {n: compare(gpu, orc, n) for n in ("dp", "pt", "ud", "vd")}
    _print("full size x100", **errs)
    for n, e in errs.items():
        assert e <= PARITY, (n, e)
    m0 = ic["dp"].sum()
    for b in (gpu, orc):
        assert abs(b.get("dp").sum() - m0) / m0 <= 1e-13


def test_mixed_form_against_the_oracle_100_steps(gpu_available):
    """The mixed form's error measured against the oracle itself (not the
    FP64 GPU path): 64x64x20, dt = 60 s, remap every 5 steps, radiation every
    10 steps, 100 steps.  The north star's bound is 1e-5 relative RMS in
    surface pressure and temperature."""
    import oracle as O
    from atmos import driver
    from atmos.config import Config
    from atmos.ic import make_initial_state
    from atmos.model import Model
    nx, ny, nlay = 64, 64, 20
    print("mixed form vs oracle 64x64x20 x100, radiation every 10", flush=True)
    ic = make_initial_state(nx, ny, nlay, 100e3)
    kw = dict(dt_max=60.0, t_final=1e12, dt_output=1e12, dt_radiation=600.0, dt_physics=1e12, dt_remap=300.0,
              radiation_on=True, physics_on=False)
    gpu, orc = Model(nx, ny, nlay, 100e3, precision="mixed"), O.Oracle(nx, ny, nlay, 100e3, 100e3)
    for b in (gpu, orc):
        st = driver.init_model(b, ic, Config(**kw))
        t, n, stable = driver.run_model(b, Config(**kw), st, max_steps=100)
        assert stable and n == 100
    ps_g, T_g = _ps_T(gpu, ic["ak"][0])
    ps_o, T_o = _ps_T(orc, ic["ak"][0])
    rr = {"ps": float(np.sqrt(np.mean((ps_g - ps_o) ** 2)) / np.sqrt(np.mean(ps_o ** 2))),
          "T": float(np.sqrt(np.mean((T_g - T_o) ** 2)) / np.sqrt(np.mean(T_o ** 2)))}
    _print("mixed vs oracle x100 rel rms", **rr)
    assert rr["ps"] <= 1e-5 and rr["T"] <= 1e-5, rr
