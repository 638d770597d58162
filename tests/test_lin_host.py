"""Tangent-linear and adjoint of one model step, host build of the same
step program and per-point bodies the HIP kernels run (csrc/lin_host.cpp):

* its forward values are the reference step's: the oracle's C + D step and
  remap (oracle/atm_oracle.c) to rounding (pow_kappa against glibc pow);
* the tangent-linear model matches central finite differences of the
  forward model (perturbations small enough not to flip a limiter);
* the adjoint passes the dot-product test <M dx, y> = <dx, M^T y>;
* a wrong linearisation fails: a tangent taken about a different state
  (one step further on) misses the finite differences.
"""
import numpy as np
import pytest

from lin_host import STATE, HostLin

NX, NY, NLAY, DT = 12, 10, 6, 60.0


def _ic(seed=7):
    from atmos.ic import make_initial_state
    return make_initial_state(NX, NY, NLAY, 100e3, seed=seed)


def _oracle_steps(ic, nsteps, remap):
    import oracle as O
    from atmos import driver
    from atmos.config import Config
    orc = O.Oracle(NX, NY, NLAY, 100e3, 100e3)
    cfg = Config(dt_max=DT, t_final=1e12, dt_output=1e12, dt_radiation=1e12, dt_physics=1e12,
                 dt_remap=1e12, radiation_on=False, physics_on=False)
    driver.init_model(orc, ic, cfg)
    for _ in range(nsteps):
        orc.cgrid_dynamics_step(DT / 2)
        orc.cgrid_halo_exchange()
        orc.dgrid_dynamics_step(DT)
        if remap:
            orc.perform_vertical_remapping()
        orc.dgrid_halo_exchange()
    return {n: orc.interior(f).copy() for n, f in zip(STATE, (O.DP, O.PT, O.UD, O.VD))}


def _direction(x0, seed=1):
    rng = np.random.default_rng(seed)
    scale = {"dp": 1.0, "pt": 1e-4, "ud": 0.1, "vd": 0.1}
    return {n: rng.standard_normal(x0[n].shape) * scale[n] for n in STATE}


def _run(h, x, dt, remap, tangent=None):
    h.set_state(x)
    h.set_tangent(tangent if tangent is not None else {n: np.zeros_like(x[n]) for n in STATE})
    h.tl_step(dt, remap)
    return h.state(), h.tangent()


@pytest.mark.parametrize("remap", [False, True])
def test_forward_values_are_the_reference_step(remap):
    ic = _ic()
    h = HostLin(ic)
    for _ in range(2):
        h.tl_step(DT, remap)
    x = h.state()
    ref = _oracle_steps(ic, 2, remap)
    for n in STATE:
        d = np.max(np.abs(x[n] - ref[n])) / np.max(np.abs(ref[n]))
        assert d < 1e-13, (n, d)


@pytest.mark.parametrize("remap", [False, True])
def test_tangent_linear_matches_finite_differences(remap):
    ic = _ic()
    h = HostLin(ic)
    x0 = h.state()
    dx = _direction(x0)
    _, tl = _run(h, x0, DT, remap, dx)
    eps = 1e-5
    xp, _ = _run(h, {n: x0[n] + eps * dx[n] for n in STATE}, DT, remap)
    xm, _ = _run(h, {n: x0[n] - eps * dx[n] for n in STATE}, DT, remap)
    for n in STATE:
        fd = (xp[n] - xm[n]) / (2 * eps)
        d = np.max(np.abs(fd - tl[n])) / np.max(np.abs(tl[n]))
        assert d < 1e-5, (n, d)


def test_tangent_about_another_state_misses_finite_differences():
    """The check has teeth: the tangent linearised one step further on (a
    plausible bookkeeping slip of a trajectory) is far from the finite
    differences about the intended state."""
    ic = _ic()
    h = HostLin(ic)
    x0 = h.state()
    dx = _direction(x0)
    x1, _ = _run(h, x0, DT, True)
    _, tl_wrong = _run(h, x1, DT, True, dx)
    eps = 1e-5
    xp, _ = _run(h, {n: x0[n] + eps * dx[n] for n in STATE}, DT, True)
    xm, _ = _run(h, {n: x0[n] - eps * dx[n] for n in STATE}, DT, True)
    worst = max(np.max(np.abs((xp[n] - xm[n]) / (2 * eps) - tl_wrong[n])) / np.max(np.abs(tl_wrong[n]))
                for n in STATE)
    assert worst > 1e-3


@pytest.mark.parametrize("remap", [False, True])
def test_adjoint_dot_product(remap):
    ic = _ic(seed=11)
    h = HostLin(ic)
    x0 = h.state()
    dx = _direction(x0, seed=2)
    _, tl = _run(h, x0, DT, remap, dx)
    rng = np.random.default_rng(3)
    for field in STATE + ("all",):
        y = {n: (rng.standard_normal(x0[n].shape) if field in (n, "all") else np.zeros_like(x0[n]))
             for n in STATE}
        h.set_state(x0)
        h.set_adjoint(y)
        h.ad_step(DT, remap)
        ad = h.adjoint()
        assert np.array_equal(h.state()["dp"], x0["dp"])  # the state is left as it was
        lhs = sum(float(np.sum(tl[n] * y[n])) for n in STATE)
        rhs = sum(float(np.sum(dx[n] * ad[n])) for n in STATE)
        assert abs(lhs - rhs) <= 1e-12 * abs(lhs), (field, lhs, rhs)


def test_adjoint_of_a_two_step_window_is_the_composition():
    """Two adjoint steps in reverse order give the adjoint of two tangent
    steps (the dot-product test over the window)."""
    ic = _ic(seed=5)
    h = HostLin(ic)
    x0 = h.state()
    dx = _direction(x0, seed=4)
    x1, t1 = _run(h, x0, DT, False, dx)
    _, t2 = _run(h, x1, DT, True, t1)
    rng = np.random.default_rng(6)
    y = {n: rng.standard_normal(x0[n].shape) for n in STATE}
    h.set_state(x1)
    h.set_adjoint(y)
    h.ad_step(DT, True)
    h.set_state(x0)
    h.ad_step(DT, False)
    ad = h.adjoint()
    lhs = sum(float(np.sum(t2[n] * y[n])) for n in STATE)
    rhs = sum(float(np.sum(dx[n] * ad[n])) for n in STATE)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_binomial_schedule():
    from atmos.adjoint import binomial_reps, binomial_split, recompute_steps
    assert binomial_reps(90, 9) == 3  # 10 snapshots, the start's included: C(12, 3) = 220 >= 90 > C(11, 2)
    # cost of the optimal schedule equals a brute-force search of all splits
    def brute(n, s):
        if n <= 1:
            return 0
        if s == 0:
            return n * (n - 1) // 2
        return min(m + brute(n - m, s - 1) + brute(m, s) for m in range(1, n))
    for n in range(1, 12):
        for s in range(1, 4):
            assert recompute_steps(n, s) == brute(n, s - 1)
            m = binomial_split(n, s - 1)
            if n > 1 and s > 1:
                assert m + brute(n - m, s - 2) + brute(m, s - 1) == brute(n, s - 1)
