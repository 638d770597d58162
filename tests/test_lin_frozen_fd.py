"""The gradient of a surface-pressure cost through a window of steps against
finite differences of the model with its branches frozen (host build of
the linearised program, csrc/lin_host.cpp, in its frozen-branch mode).

The reference's constrained PPM (mod_sw_dyn.f90:802-829) switches between
a parabola and first-order upwind where 3|eL + eR| > |eL - eR| on both
cells: a discontinuous switch.  Over a window of many steps and many
columns some of those switches sit within ~1e-7 of their threshold, so the
cost J is piecewise smooth with jumps at a spacing far below any usable
epsilon, and central differences of the model itself measure the jumps.
The tangent-linear and adjoint models are the derivative of the model with
each switch held on the branch the trajectory took; with the host build's
decisions recorded on the base trajectory and replayed in the perturbed
runs, finite differences differentiate exactly that model, and the adjoint
gradient has to match them to the rounding of J over epsilon.
"""
import numpy as np
import pytest

from lin_host import STATE, HostLin, tape, tape_stats

DT = 60.0


def _ic(nx, ny, nlay, seed):
    from atmos.ic import make_initial_state, spectral_noise
    ic = make_initial_state(nx, ny, nlay, 100e3, seed=seed)
    rng = np.random.default_rng(seed + 1000)
    for k in range(nlay):
        ic["dp"][k] += spectral_noise(rng, ny, nx)
        ic["pt"][k] *= 1.0 + 3e-4 * spectral_noise(rng, ny, nx)
    return ic


def cost(dp, p_top, box):
    """mean over the box of (ps - mean ps)^2, ps = p_top + sum_k dp"""
    j0, j1, i0, i1 = box
    ps = p_top + dp.sum(axis=0)[j0:j1, i0:i1]
    return float(np.mean((ps - ps.mean()) ** 2))


def cost_grad(dp, p_top, box):
    j0, j1, i0, i1 = box
    ps = p_top + dp.sum(axis=0)
    g = np.zeros_like(dp)
    box_ps = ps[j0:j1, i0:i1]
    g[:, j0:j1, i0:i1] = 2.0 * (box_ps - box_ps.mean()) / box_ps.size
    return g


def schedule(n, every):
    return [(k + 1) % every == 0 for k in range(n)]


def window_forward(h, x0, sched, record=False):
    """States at the start of each step (and the final one); with record,
    the decisions of the run go to the tape."""
    h.set_state(x0)
    h.set_tangent({n: np.zeros_like(x0[n]) for n in STATE})
    states = []
    if record:
        tape("record")
    for remap in sched:
        states.append(h.state())
        h.tl_step(DT, remap)
    tape("off")
    return states, h.state()


def window_gradient(h, states, final, sched, p_top, box):
    """Adjoint of J through the window from the stored states."""
    h.set_adjoint({"dp": cost_grad(final["dp"], p_top, box), **{n: np.zeros_like(final[n]) for n in STATE[1:]}})
    for k in range(len(sched) - 1, -1, -1):
        h.set_state(states[k])
        h.ad_step(DT, sched[k])
    return h.adjoint()


def frozen_J(h, x, sched, p_top, box):
    """J of the window from x with every decision replayed from the tape."""
    h.set_state(x)
    tape("replay")
    for remap in sched:
        h.tl_step(DT, remap)
    st = tape_stats()
    tape("off")
    assert st["overrun"] == 0 and st["replayed"] == st["recorded"]
    return cost(h.state()["dp"], p_top, box), st["flips"]


def live_J(h, x, sched, p_top, box):
    h.set_state(x)
    for remap in sched:
        h.tl_step(DT, remap)
    return cost(h.state()["dp"], p_top, box)


def gradient_and_fd(nx, ny, nlay, nsteps, seed, eps_list, box):
    ic = _ic(nx, ny, nlay, seed)
    h = HostLin(ic)
    p_top = float(ic["ak"][0])
    x0 = h.state()
    sched = schedule(nsteps, 5)
    states, final = window_forward(h, x0, sched, record=True)
    g = window_gradient(h, states, final, sched, p_top, box)
    rng = np.random.default_rng(8)
    d = {"dp": np.zeros_like(x0["dp"]), "pt": 1e-3 * rng.standard_normal(x0["pt"].shape),
         "ud": rng.standard_normal(x0["ud"].shape), "vd": rng.standard_normal(x0["vd"].shape)}
    gd = sum(float(np.sum(g[n] * d[n])) for n in STATE)
    out = []
    for eps in eps_list:
        jp, fp = frozen_J(h, {n: x0[n] + eps * d[n] for n in STATE}, sched, p_top, box)
        jm, fm = frozen_J(h, {n: x0[n] - eps * d[n] for n in STATE}, sched, p_top, box)
        fd = (jp - jm) / (2 * eps)
        lp = live_J(h, {n: x0[n] + eps * d[n] for n in STATE}, sched, p_top, box)
        lm = live_J(h, {n: x0[n] - eps * d[n] for n in STATE}, sched, p_top, box)
        out.append({"eps": eps, "fd_frozen": fd, "fd_live": (lp - lm) / (2 * eps), "flips": fp + fm})
    return gd, out, cost(final["dp"], p_top, box), g, d, x0, sched, ic


def test_window_gradient_matches_frozen_branch_differences():
    gd, res, J, *_ = gradient_and_fd(32, 16, 8, 12, seed=12, eps_list=(1e-7, 1e-5), box=(4, 12, 8, 24))
    for r in res:
        print(f"J {J:.6e} eps {r['eps']:g}: <grad, d> {gd:.12e} frozen FD {r['fd_frozen']:.12e} "
              f"live FD {r['fd_live']:.12e} (decisions flipped in the live runs: {r['flips']})")
        assert abs(r["fd_frozen"] - gd) <= 1e-6 * abs(gd), r
    # the frozen differences converge: O(eps^2) truncation, O(1/eps) rounding
    assert abs(res[0]["fd_frozen"] - res[1]["fd_frozen"]) <= 1e-6 * abs(gd)


def test_replay_without_perturbation_flips_nothing():
    ic = _ic(16, 12, 6, seed=3)
    h = HostLin(ic)
    x0 = h.state()
    sched = schedule(6, 5)
    _, final = window_forward(h, x0, sched, record=True)
    p_top = float(ic["ak"][0])
    box = (2, 10, 2, 14)
    J, flips = frozen_J(h, x0, sched, p_top, box)
    assert flips == 0 and J == cost(final["dp"], p_top, box)
    # a perturbation large enough to move the state across switches flips
    # some, and the frozen run then differs from the live one
    rng = np.random.default_rng(1)
    xp = {n: x0[n] + (rng.standard_normal(x0[n].shape) if n in ("ud", "vd") else 0.0) for n in STATE}
    jf, flips = frozen_J(h, xp, sched, p_top, box)
    assert flips > 0 and jf != live_J(h, xp, sched, p_top, box)
