"""Gradients through a window of model steps on the GPU.

`Window` runs a fixed schedule of steps (dt, remap due) forward on a
`Model` and back through the library's adjoint of one step
(`Model.ad_step`: the hand-written adjoint kernels, linearised about the
stored trajectory with limiter and upwind switches frozen).  The trajectory
is not stored step by step: binomial checkpointing (Griewank's "revolve"
schedule) keeps at most `snapshots` states and recomputes the rest, so a
window of n steps needs at most r forward recomputations of any step, with
r the smallest integer for which C(snapshots + r, snapshots) >= n (90 steps,
10 snapshots: r = 3).

`WindowFunction` puts a window behind torch.autograd: its forward maps the
initial interior fields (dp, pt, ud, vd as float64 CUDA tensors of shape
(nlay, ny, nx)) to the final ones, its backward returns their gradients.
The forward trajectory runs on the model's own path and precision (the
mixed form when the model was created so); the linearised step forms its
forward values in FP64 from each stored state.

torch carries its own HIP runtime beside the one the library links; in a
process that uses both, torch must bring up the device (torch.cuda.init())
before the first Model is created.

The library runs on the model's own stream; every hand-over of a tensor
between torch and the library waits for the side that wrote it (torch's
device work before the library reads a tensor, the library's stream before
torch reads one or frees one the library copied from).
"""
from __future__ import annotations

from functools import lru_cache
from math import comb

import numpy as np

STATE = ("dp", "pt", "ud", "vd")
HALO = 7


def binomial_reps(n: int, snaps: int) -> int:
    """Smallest r with C(snaps + r, snaps) >= n: the most times revolve
    advances any one step (n - 1 without a snapshot)."""
    if snaps <= 0:
        return max(n - 1, 0)
    r = 0
    while comb(snaps + r, snaps) < n:
        r += 1
    return r


@lru_cache(maxsize=None)
def _reverse_cost(n: int, snaps: int) -> tuple:
    """(forward steps, first split) to reverse n steps from a held start
    state with `snaps` more snapshots free: advancing m steps, snapshotting,
    reversing the right part with one snapshot fewer and the left part with
    as many, T(n, s) = min_m [m + T(n - m, s - 1) + T(m, s)]; without a free
    snapshot every step is recomputed from the start, T(n, 0) = n(n-1)/2.
    The minimum is Griewank's binomial bound."""
    if n <= 1:
        return 0, 0
    if snaps <= 0:
        return n * (n - 1) // 2, 0
    best, arg = None, 1
    for m in range(1, n):
        c = m + _reverse_cost(n - m, snaps - 1)[0] + _reverse_cost(m, snaps)[0]
        if best is None or c < best:
            best, arg = c, m
    return best, arg


def binomial_split(n: int, snaps: int) -> int:
    """Steps to advance from the start of a range of n steps before taking
    the next snapshot, with `snaps` snapshots free (0: none to take)."""
    if n <= 1 or snaps <= 0:
        return 0
    return _reverse_cost(n, snaps)[1]


def recompute_steps(n: int, snaps: int) -> int:
    """Forward steps the reversal of n steps costs with `snaps` snapshots
    (the start state's included)."""
    return _reverse_cost(n, snaps - 1)[0]


class Window:
    """A schedule of steps with binomial checkpointing on a Model (which must
    have run init_model; remap steps need its remap arrays)."""

    def __init__(self, model, schedule, snapshots=10):
        import torch

        self.torch = torch
        self.m = model
        self.schedule = [(float(dt), bool(rm)) for dt, rm in schedule]
        self.snaps = int(snapshots)
        if self.snaps < 1:
            raise ValueError("at least one snapshot")
        self.shape_full = (model.nlay, model.ny + 2 * HALO, model.nx + 2 * HALO)
        self.m.lin_init()
        self.store = {}  # step index -> halo'd state tensors
        self.counts = {"forward_steps": 0, "adjoint_steps": 0, "max_snapshots": 0}

    # ---- state transfer ----
    def _take(self):
        t = self.torch
        snap = [t.empty(self.shape_full, dtype=t.float64, device="cuda") for _ in STATE]
        t.cuda.synchronize()
        for n, a in zip(STATE, snap):
            self.m.get_device(n, a.data_ptr())
        self.m.sync()
        return snap

    def _give(self, snap):
        self.torch.cuda.synchronize()
        for n, a in zip(STATE, snap):
            self.m.put_device(n, a.data_ptr())
        self.m.sync()

    def _interior(self, names):
        t = self.torch
        full = [t.empty(self.shape_full, dtype=t.float64, device="cuda") for _ in names]
        t.cuda.synchronize()
        for n, a in zip(names, full):
            self.m.get_device(n, a.data_ptr())
        self.m.sync()
        return [a[:, HALO:-HALO, HALO:-HALO].contiguous() for a in full]

    def _put_interior(self, names, tensors):
        t = self.torch
        full = []
        for x in tensors:
            a = t.zeros(self.shape_full, dtype=t.float64, device="cuda")
            a[:, HALO:-HALO, HALO:-HALO] = x
            full.append(a)
        t.cuda.synchronize()
        for n, a in zip(names, full):
            self.m.put_device(n, a.data_ptr())
        self.m.sync()

    def set_state(self, dp, pt, ud, vd):
        """Initial interior fields; the halos are filled periodically."""
        self._put_interior(STATE, (dp, pt, ud, vd))
        for n in STATE:
            self.m.halo_exchange(n)

    # ---- forward ----
    def _advance(self, k0, k1):
        m = self.m
        for k in range(k0, k1):
            dt, remap = self.schedule[k]
            m.cgrid_dynamics_step(dt / 2.0)
            m.cgrid_halo_exchange()
            m.dgrid_dynamics_step(dt)
            if remap:
                m.perform_vertical_remapping()
            m.dgrid_halo_exchange()
        self.counts["forward_steps"] += k1 - k0

    def _spine(self, n):
        """Snapshot positions the checkpointing takes on its first sweep."""
        pos, start, free = [], 0, self.snaps - 1  # the start state holds one snapshot
        while n - start > 1 and free > 0:
            start += binomial_split(n - start, free)
            pos.append(start)
            free -= 1
        return pos

    def forward(self):
        """Run the schedule from the current state, keeping the start state
        and the first sweep's snapshots; returns the final interior fields."""
        n = len(self.schedule)
        self.store = {0: self._take()}
        k = 0
        for p in self._spine(n):
            self._advance(k, p)
            self.store[p] = self._take()
            k = p
        self._advance(k, n)
        self.counts["max_snapshots"] = max(self.counts["max_snapshots"], len(self.store))
        return self._interior(STATE)

    # ---- adjoint ----
    def _ad(self, k):
        dt, remap = self.schedule[k]
        self.m.ad_step(dt, remap)
        self.counts["adjoint_steps"] += 1

    def _reverse(self, start, end, free):
        """Adjoint over steps [start, end): the snapshot at `start` is held and
        the adjoint fields hold the adjoint at `end`."""
        while end - start > 0:
            if end - start == 1:
                self._give(self.store[start])
                self._ad(start)
                return
            if free <= 0:
                # no snapshot left: recompute from `start` for each step
                for k in range(end - 1, start - 1, -1):
                    self._give(self.store[start])
                    self._advance(start, k)
                    self._ad(k)
                return
            mid = start + binomial_split(end - start, free)
            if mid not in self.store:
                self._give(self.store[start])
                self._advance(start, mid)
                self.store[mid] = self._take()
                self.counts["max_snapshots"] = max(self.counts["max_snapshots"], len(self.store))
            self._reverse(mid, end, free - 1)
            del self.store[mid]
            end = mid  # then [start, mid) with the same snapshots free

    def backward(self, gdp, gpt, gud, gvd):
        """Gradients of the initial interior fields from those of the final."""
        self._put_interior(tuple("ad_" + n for n in STATE), (gdp, gpt, gud, gvd))
        self._reverse(0, len(self.schedule), self.snaps - 1)
        grads = self._interior(tuple("ad_" + n for n in STATE))
        self._give(self.store[0])  # leave the model at the window's start
        return grads


def _function():
    import torch

    class WindowFunction(torch.autograd.Function):
        """(dp, pt, ud, vd) at the window's start -> at its end; backward
        through the library's adjoint kernels with binomial checkpointing."""

        @staticmethod
        def forward(ctx, dp, pt, ud, vd, window):
            window.set_state(dp, pt, ud, vd)
            out = window.forward()
            ctx.window = window
            return tuple(out)

        @staticmethod
        def backward(ctx, gdp, gpt, gud, gvd):
            z = [g if g is not None else torch.zeros_like(ctx.window._interior(("dp",))[0]) for g in
                 (gdp, gpt, gud, gvd)]
            grads = ctx.window.backward(*[g.contiguous().to(torch.float64) for g in z])
            return (*grads, None)

    return WindowFunction


WindowFunction = None


def window_apply(window, dp, pt, ud, vd):
    """WindowFunction.apply(dp, pt, ud, vd, window)."""
    global WindowFunction
    if WindowFunction is None:
        WindowFunction = _function()
    return WindowFunction.apply(dp, pt, ud, vd, window)


def surface_pressure_cost(dp, p_top, box):
    """J = mean over a box of (ps - mean ps)^2 with ps = p_top + sum_k dp
    (the surface pressure, src/mod_remap.f90:42-44); box = (j0, j1, i0, i1)."""
    j0, j1, i0, i1 = box
    ps = p_top + dp.sum(dim=0)[j0:j1, i0:i1]
    return ((ps - ps.mean()) ** 2).mean()


def schedule_of(nsteps, dt, remap_every):
    """(dt, remap due) for each step: remap after every remap_every-th step."""
    return [(dt, remap_every > 0 and (k + 1) % remap_every == 0) for k in range(nsteps)]


def numpy_dot(a, b):
    return float(sum(np.sum(np.asarray(x, dtype=float) * np.asarray(y, dtype=float)) for x, y in zip(a, b)))
