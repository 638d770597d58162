"""Tangent-linear and adjoint kernels of one model step on the GPU
(atm_lin_init / atm_tl_step / atm_ad_step) and gradients through a window
of steps with binomial checkpointing behind torch.autograd (atmos.adjoint):

* the linearised kernels' forward values are the FP64 step's, bit for bit
  (fused path: eager C, D, remap, halo calls);
* tangents match central finite differences, the adjoint passes the
  dot-product test, and both match the host build of the same program
  (tests/test_lin_host.py checks that one against the oracle);
* the window's gradient of a surface-pressure cost matches finite
  differences of the cost, and does not depend on the number of snapshots.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STATE = ("dp", "pt", "ud", "vd")
DT = 60.0


def _model(nx=64, ny=32, nlay=10, seed=3, precision="fp64"):
    from atmos import driver
    from atmos.config import Config
    from atmos.ic import make_initial_state
    from atmos.model import Model
    from atmos.ic import spectral_noise
    ic = make_initial_state(nx, ny, nlay, 100e3, seed=seed)
    # Far from the pressure bump the surface pressure is uniform to ~1e-5
    # Pa, so neighbouring dp values tie to that precision and finite-
    # difference perturbations would flip the limiters' min/max selections
    # all over the far field; smooth noise of 1 Pa in dp (0.1 K in T)
    # removes those ties.
    rng = np.random.default_rng(seed + 1000)
    for k in range(nlay):
        ic["dp"][k] += spectral_noise(rng, ny, nx)
        ic["pt"][k] *= 1.0 + 3e-4 * spectral_noise(rng, ny, nx)
    m = Model(nx, ny, nlay, 100e3, precision=precision)
    cfg = Config(dt_max=DT, t_final=1e12, dt_output=1e12, dt_radiation=1e12, dt_physics=1e12, dt_remap=1e12,
                 radiation_on=False, physics_on=False)
    driver.init_model(m, ic, cfg)
    return m, ic


def _state(m, prefix=""):
    return {n: m.get(prefix + n) for n in STATE}


def _put(m, d, prefix=""):
    for n in STATE:
        m.put(prefix + n, d[n])
        if not prefix:
            m.halo_exchange(n)


def _direction(x0, seed):
    rng = np.random.default_rng(seed)
    scale = {"dp": 1.0, "pt": 1e-3, "ud": 0.1, "vd": 0.1}
    return {n: rng.standard_normal(x0[n].shape) * scale[n] for n in STATE}


@pytest.mark.parametrize("remap", [False, True])
def test_linearised_forward_is_the_fp64_step(gpu_available, remap):
    print(f"lin forward remap={remap}", flush=True)
    a, _ = _model()
    b, _ = _model()
    b.lin_init()
    for _ in range(3):
        a.cgrid_dynamics_step(DT / 2)
        a.cgrid_halo_exchange()
        a.dgrid_dynamics_step(DT)
        if remap:
            a.perform_vertical_remapping()
        a.dgrid_halo_exchange()
        b.tl_step(DT, remap)
    for n in STATE:
        assert np.array_equal(a.get(n), b.get(n)), n
    # the halo the linearised step leaves is the periodic fill
    full = b.get_full("ud")
    assert np.array_equal(full[:, :7, 7:-7], full[:, -14:-7, 7:-7])


@pytest.mark.parametrize("remap", [False, True])
def test_tangent_fd_and_adjoint_dot_product(gpu_available, remap):
    print(f"tangent/adjoint remap={remap}", flush=True)
    m, _ = _model()
    m.lin_init()
    x0 = _state(m)
    dx = _direction(x0, 1)

    def run(x, t=None):
        _put(m, x)
        _put(m, t if t is not None else {n: np.zeros_like(x[n]) for n in STATE}, "tl_")
        m.tl_step(DT, remap)
        return _state(m), _state(m, "tl_")

    _, tl = run(x0, dx)
    # Central differences straddle every limiter or upwind switch that lies
    # within eps of the state; over 20k columns a few do at any eps, and
    # there the difference quotient is not a derivative.  The field norm of
    # (FD - TL) stays small: it converges to the tangent wherever the
    # nearest switch is farther than eps (checked pointwise on small grids
    # by the host build's tests).
    eps = 1e-6
    xp, _ = run({n: x0[n] + eps * dx[n] for n in STATE})
    xm, _ = run({n: x0[n] - eps * dx[n] for n in STATE})
    for n in STATE:
        fd = (xp[n] - xm[n]) / (2 * eps)
        d = np.linalg.norm(fd - tl[n]) / np.linalg.norm(tl[n])
        assert d < 1e-4, (n, d)
    rng = np.random.default_rng(2)
    y = {n: rng.standard_normal(x0[n].shape) for n in STATE}
    _put(m, x0)
    _put(m, y, "ad_")
    m.ad_step(DT, remap)
    ad = _state(m, "ad_")
    for n in STATE:
        assert np.array_equal(m.get(n), x0[n])  # the state is left as it was
    lhs = sum(float(np.sum(tl[n] * y[n])) for n in STATE)
    rhs = sum(float(np.sum(dx[n] * ad[n])) for n in STATE)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs), (lhs, rhs)


def test_gpu_linearisation_matches_the_host_build(gpu_available):
    """Same program and bodies on the host: tangents to rounding, adjoints
    to the reordering of the device's atomic additions."""
    from lin_host import HostLin
    print("gpu vs host linearisation", flush=True)
    m, ic = _model(nx=32, ny=16, nlay=8, seed=9)
    m.lin_init()
    h = HostLin(ic)
    x0 = _state(m)
    dx = _direction(x0, 4)
    _put(m, dx, "tl_")
    h.set_tangent(dx)
    m.tl_step(DT, True)
    h.tl_step(DT, True)
    tg, th = _state(m, "tl_"), h.tangent()
    for n in STATE:
        assert np.max(np.abs(tg[n] - th[n])) <= 1e-13 * np.max(np.abs(th[n])), n
    y = {n: np.random.default_rng(5).standard_normal(x0[n].shape) for n in STATE}
    _put(m, x0)
    h.set_state(x0)
    _put(m, y, "ad_")
    h.set_adjoint(y)
    m.ad_step(DT, True)
    h.ad_step(DT, True)
    ag, ah = _state(m, "ad_"), h.adjoint()
    for n in STATE:
        assert np.max(np.abs(ag[n] - ah[n])) <= 1e-12 * np.max(np.abs(ah[n])), n


def _window_grad(m, x0, schedule, snaps, box, p_top):
    import torch

    from atmos.adjoint import Window, surface_pressure_cost, window_apply
    w = Window(m, schedule, snapshots=snaps)
    xs = [torch.tensor(x0[n], dtype=torch.float64, device="cuda", requires_grad=True) for n in STATE]
    out = window_apply(w, *xs)
    J = surface_pressure_cost(out[0], p_top, box)
    J.backward()
    return J.item(), [x.grad.cpu().numpy() for x in xs], w


def test_window_gradient_matches_finite_differences(gpu_available):
    import torch

    from atmos.adjoint import Window, recompute_steps, schedule_of, surface_pressure_cost, window_apply
    print("window gradient", flush=True)
    m, ic = _model(nx=64, ny=32, nlay=10, seed=12)
    x0 = _state(m)
    p_top = float(ic["ak"][0])
    sched = schedule_of(12, DT, 5)
    box = (8, 24, 16, 48)
    J0, g, w = _window_grad(m, x0, sched, 3, box, p_top)
    assert w.counts["adjoint_steps"] == 12
    assert w.counts["forward_steps"] == 12 + recompute_steps(12, 3) - sum(
        p - q for p, q in zip(w._spine(12), [0] + w._spine(12)[:-1]))
    assert w.counts["max_snapshots"] <= 3
    # the same gradient with every state kept (no recomputation)
    _put(m, x0)
    _, g_all, _ = _window_grad(m, x0, sched, 12, box, p_top)
    for a, b in zip(g, g_all):
        assert np.max(np.abs(a - b)) <= 1e-11 * np.max(np.abs(b))
    # finite differences of J along T (pt) and the winds
    rng = np.random.default_rng(8)
    d = {"dp": np.zeros_like(x0["dp"]), "pt": 1e-3 * rng.standard_normal(x0["pt"].shape),
         "ud": rng.standard_normal(x0["ud"].shape), "vd": rng.standard_normal(x0["vd"].shape)}
    gdot = sum(float(np.sum(gi * d[n])) for gi, n in zip(g, STATE))

    def J_at(eps):
        ww = Window(m, sched, snapshots=3)
        xs = [torch.tensor(x0[n] + eps * d[n], dtype=torch.float64, device="cuda") for n in STATE]
        out = window_apply(ww, *xs)
        return float(surface_pressure_cost(out[0], p_top, box))

    eps = 1e-7
    fd = (J_at(eps) - J_at(-eps)) / (2 * eps)
    print(f"J {J0:.6e} <grad, d> {gdot:.12e} fd {fd:.12e}", flush=True)
    assert abs(fd - gdot) <= 1e-6 * abs(gdot), (fd, gdot)


def test_window_gradient_is_the_host_adjoint_and_the_frozen_branch_derivative(gpu_available):
    """The GPU window's gradient (binomial checkpointing, 3 snapshots) is the
    host build's adjoint through stored states, and matches finite
    differences of the model with its branches frozen (the host build's
    decision tape, eps 1e-7) to 1e-6."""
    from lin_host import HostLin
    from test_lin_frozen_fd import frozen_J, window_forward, window_gradient

    from atmos.adjoint import schedule_of
    print("window gradient vs host adjoint and frozen-branch FD", flush=True)
    m, ic = _model(nx=64, ny=32, nlay=10, seed=12)
    x0 = _state(m)
    p_top = float(ic["ak"][0])
    sched = schedule_of(12, DT, 5)
    box = (8, 24, 16, 48)
    J0, g, _ = _window_grad(m, x0, sched, 3, box, p_top)
    h = HostLin(ic)
    remaps = [r for _, r in sched]
    states, final = window_forward(h, x0, remaps, record=True)
    gh = window_gradient(h, states, final, remaps, p_top, box)
    for a, n in zip(g, STATE):
        assert np.max(np.abs(a - gh[n])) <= 1e-10 * np.max(np.abs(gh[n])), n
    rng = np.random.default_rng(8)
    d = {"dp": np.zeros_like(x0["dp"]), "pt": 1e-3 * rng.standard_normal(x0["pt"].shape),
         "ud": rng.standard_normal(x0["ud"].shape), "vd": rng.standard_normal(x0["vd"].shape)}
    gdot = sum(float(np.sum(gi * d[n])) for gi, n in zip(g, STATE))
    eps = 1e-7
    jp, _ = frozen_J(h, {n: x0[n] + eps * d[n] for n in STATE}, remaps, p_top, box)
    jm, _ = frozen_J(h, {n: x0[n] - eps * d[n] for n in STATE}, remaps, p_top, box)
    fd = (jp - jm) / (2 * eps)
    print(f"<grad, d> {gdot:.12e} frozen FD {fd:.12e}", flush=True)
    assert abs(fd - gdot) <= 1e-6 * abs(gdot)
