#!/usr/bin/env python3
"""Throughput of the FP64 dynamics + remap hot path on one MI355X.

Workload (the plain FP64 GPU drop-in configuration): 512 x 512 x 64 doubly
periodic grid, dx = dy = 25 km, dt = 20 s, vertical remap every 5 steps
(every 100 s), radiation and physics off, seeded synthetic initial state
(atmos.ic).  One step is the reference's main-loop body: C-grid half step,
C halo fill, D-grid step, the remap when due, D halo fill and the
stability check (src/mod_model.f90:117-177).

Prints one JSON line: value = grid-cell-steps per second
(nx * ny * nlay * steps / seconds) with all state resident in HBM, a
roofline entry for the dominant kernel (algorithmic bytes per launch over
its HIP-event-measured average duration, against 8 TB/s) and a CPU
baseline: the oracle (a C restatement of the reference algorithm, threaded
over the host cores with OpenMP) timed on a bounded sample of the same
workload, with a single-core figure beside it.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "atmosphere-model_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "cell-steps/s FP64 vs mixed at stated error; adjoint:forward cost ratio"
HBM_PEAK_GBS = 8000.0

# algorithmic HBM bytes per interior cell for one launch of each kernel group
# (compulsory reads + writes of the fields each kernel consumes/produces)
# radiation: FP64 operations per layer and g-point of the column solve
# (k_rad_columns: shortwave optics + two-stream coefficients ~145, longwave
# ~85, the two adding passes ~44; exp, sqrt and each division count as one)
RAD_FLOP_PER_LAYER_G = 280
RAD_G_POINTS = 16
FP64_VECTOR_PEAK_TFLOPS = 78.6

# Algorithmic bytes of a whole step per cell (FP64 fields): the C stage's
# compulsory reads and writes (9 x 8 B), the D stage's (13 x 8 B), two halo
# fills (~7 B) and a remap every 5 steps (8 x 8 B / 5): ~196 B.
WHOLE_STEP_BYTES = 72 + 104 + 7 + 12.8

ALG_BYTES_PER_CELL = {
    "c1": 8 * (6 + 4),   # read dp, pt, ud, vd, playkap, net_flux; write dpc, ptc, uc, vc
    "c2": 8 * (6 + 2),   # read uc, vc, ud, vd, pkap, gz; write uc, vc
    "d1": 8 * (5 + 2),   # read dp, pt, uc, vc, heat; write dp, pt
    "d2": 8 * (6 + 2),   # read ud, vd, uc, vc, pkap, gz; write ud, vd
    "remap": 8 * (4 + 4),  # read dp, pt, ud, vd; write dp, pt, ud, vd
}
# the kernel groups timed with HIP events in the timed steps: the roofline's candidates
ROOFLINE_GROUPS = tuple(ALG_BYTES_PER_CELL)


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--nx", type=int, default=512)
    ap.add_argument("--ny", type=int, default=512)
    ap.add_argument("--nlay", type=int, default=64)
    ap.add_argument("--dx", type=float, default=25e3)
    ap.add_argument("--dt", type=float, default=20.0)
    ap.add_argument("--remap-every", type=int, default=5)
    ap.add_argument("--path", choices=["fused", "reference"], default="fused")
    ap.add_argument("--cpu-baseline", choices=["on", "off"], default="on")
    ap.add_argument("--cpu-nx", type=int, default=512, help="bounded CPU sample: nx = ny")
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="threads of the CPU baseline (0: OMP_NUM_THREADS or all host cores)")
    ap.add_argument("--mixed", choices=["on", "off"], default="on",
                    help="also time the mixed-precision form and its error after --mixed-error-steps steps")
    ap.add_argument("--mixed-error-steps", type=int, default=100)
    ap.add_argument("--radiation-calls", type=int, default=3,
                    help="also time this many radiation calls on the same grid (0: skip)")
    ap.add_argument("--ensemble", choices=["on", "off"], default="on",
                    help="also time the uneven ensemble: 24 x 192x192x48 + 6 x 256x256x48 + 2 x 384x384x64, "
                         "mixed precision, radiation every 10 steps")
    ap.add_argument("--ensemble-steps", type=int, default=100)
    ap.add_argument("--ensemble-streams", type=int, default=4)
    ap.add_argument("--adjoint", choices=["on", "off", "steps"], default="on",
                    help="gradient through a checkpointed window (tangent-linear/adjoint workload); "
                         "steps: only single tangent-linear and adjoint steps at the same size, "
                         "through the C ABI without torch (for profiling the step kernels)")
    ap.add_argument("--adj-nx", type=int, default=256)
    ap.add_argument("--adj-nlay", type=int, default=48)
    ap.add_argument("--adj-steps", type=int, default=90)
    ap.add_argument("--adj-snapshots", type=int, default=10)
    ap.add_argument("--twostream", choices=["on", "off"], default="on",
                    help="general two-stream columns: 262144 x 64, 32 + 32 g-points, FP64 and mixed")
    ap.add_argument("--ts-ncol", type=int, default=262144)
    ap.add_argument("--ts-calls", type=int, default=5)
    ap.add_argument("--autotune", choices=["on", "off"], default="on",
                    help="time the fused path's tile and launch-shape candidates inside trial steps on each "
                         "grid first and keep the fastest (results unchanged)")
    ap.add_argument("--checks", choices=["deferred", "sync"], default="deferred",
                    help="the per-step stability check of the timed steps: queued behind each step and read "
                         "after the next step is issued (deferred), or waited for before the next step (sync)")
    ap.add_argument("--graph", choices=["on", "off"], default="on",
                    help="also time a recorded 10-step graph against the same steps issued one call at a "
                         "time, at 64x64x20 (the CPU configuration's grid) and at the main grid")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of this many MiB (read + write) a few times, "
                         "to scale HBM counters (FETCH_SIZE/WRITE_SIZE) against a known size")
    return ap.parse_args()


def calibrate_copy(mib):
    import torch
    n = mib * (1 << 20) // 8
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        torch.mul(x, 1.0, out=y)
    torch.cuda.synchronize()
    print(json.dumps({"calibrate_copy_mib": mib, "bytes_read_per_launch": n * 8, "bytes_written_per_launch": n * 8}))


ENSEMBLE_SPECS = [(24, 192, 192, 48), (6, 256, 256, 48), (2, 384, 384, 64)]


def whole_step_roofline(cell_steps_per_s):
    """The whole step's algorithmic bytes per cell-step times the rate,
    against the HBM peak (radiation and physics not counted)."""
    gbs = WHOLE_STEP_BYTES * cell_steps_per_s / 1e9
    return {"bytes_per_cell_step": WHOLE_STEP_BYTES, "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": gbs / HBM_PEAK_GBS}


def ensemble_timing(args):
    """The uneven ensemble: 32 members of three grid sizes, each from its own
    seeded state (k^-3 temperature noise of 0.5 K), mixed precision, dx and
    dt of the main workload, remap every 5 steps and radiation every 10.
    Whole steps of all members (dynamics, remap, radiation, the stability
    check of every member after every step), timed after 10 warm-up steps:
    first with each check waited for before the next step is issued (the
    streams drain every step), then with deferred checks (the value: each
    member's check is queued behind its step and read after the next step is
    issued), then with each member's 10-step period replayed from a HIP graph
    (radiation eager at the period start, one deferred check per period);
    then a second ensemble of the same members whose dynamics stages run as
    one launch per stage for all members (atm_ensemble_*, deferred checks),
    eagerly and then with each 10-step period replayed from one graph of
    the batched launches and the members' remaps."""
    from atmos.ensemble import Ensemble
    ens = Ensemble(ENSEMBLE_SPECS, args.dx, args.dt, remap_every=5, radiation_every=10, precision="mixed",
                   nstreams=args.ensemble_streams, autotune=args.autotune == "on", checks="sync")
    n_sync = max(10, args.ensemble_steps // 20 * 10)
    n_def = max(10, args.ensemble_steps // 10 * 10)
    for _ in range(10):
        if not ens.step():
            raise RuntimeError("ensemble member unstable in warm-up")
    ens.sync()
    cells = ens.cells

    def timed(run, n):
        t0 = time.perf_counter()
        for _ in range(n):
            if not run():
                raise RuntimeError("ensemble member unstable")
        if not ens.finish():
            raise RuntimeError("ensemble member unstable")
        ens.sync()
        return time.perf_counter() - t0

    sec_sync = timed(ens.step, n_sync)
    ens.checks = "deferred"
    sec = timed(ens.step, n_def)
    ens.use_graphs(10)
    if not ens.period() or not ens.finish():  # records each member's graph, runs the first period
        raise RuntimeError("ensemble member unstable")
    ens.sync()
    n_per = max(1, n_def // 10)
    sec_graph = timed(ens.period, n_per)
    out = {"members": " + ".join(f"{c} x {nx}x{ny}x{L}" for c, nx, ny, L in ENSEMBLE_SPECS), "cells": cells,
           "precision": "mixed", "steps": n_def, "remap_every": 5, "radiation_every": 10,
           "value": cells * n_def / sec, "unit": "cell-steps/s",
           "ms_per_step": sec / n_def * 1e3,
           "launch": f"each member's own launches, members dealt to {args.ensemble_streams} HIP streams; "
                     "each member's stability check deferred by one step (no drain between steps)",
           "whole_step_roofline": whole_step_roofline(cells * n_def / sec),
           "sync_checks": {"steps": n_sync, "value": cells * n_sync / sec_sync,
                           "ms_per_step": sec_sync / n_sync * 1e3},
           "graph": {"period_steps": 10, "periods": n_per, "value": cells * 10 * n_per / sec_graph,
                     "ms_per_step": sec_graph / (10 * n_per) * 1e3,
                     "checks": "one deferred check per member per period"},
           "corner_levels_per_workgroup": {f"{nx}x{ny}x{L}": sorted({mb.model.tuning for mb in ens.members
                                                                     if mb.cells == nx * ny * L})
                                           for _, nx, ny, L in ENSEMBLE_SPECS},
           "tile_rows_dp_pt_wind": {f"{nx}x{ny}x{L}": sorted({mb.model.tiles for mb in ens.members
                                                              if mb.cells == nx * ny * L})
                                    for _, nx, ny, L in ENSEMBLE_SPECS}}
    ens.close()
    ens = Ensemble(ENSEMBLE_SPECS, args.dx, args.dt, remap_every=5, radiation_every=10, precision="mixed",
                   launch="batched")
    for _ in range(10):
        if not ens.step():
            raise RuntimeError("ensemble member unstable in warm-up")
    if not ens.finish():
        raise RuntimeError("ensemble member unstable in warm-up")
    ens.sync()
    sec_b = timed(ens.step, n_def)
    out["batched"] = {"steps": n_def, "value": cells * n_def / sec_b, "ms_per_step": sec_b / n_def * 1e3,
                      "launch": "one launch per stage for all 32 members (radiation profiles, zero fill, "
                                "columns; C1, halo, scans, C2, C halo, D1, halo, scans, corners, D2, remap "
                                "grid, winds, energy inputs, energy, D halo) on the first member's stream, "
                                "joined by events to the members' 4 streams; deferred checks"}
    ens.use_graphs(10)
    if not ens.period() or not ens.finish():  # records the period's graph, runs the first period
        raise RuntimeError("ensemble member unstable")
    ens.sync()
    sec_bg = timed(ens.period, n_per)
    out["batched"]["graph"] = {"period_steps": 10, "periods": n_per, "value": cells * 10 * n_per / sec_bg,
                               "ms_per_step": sec_bg / (10 * n_per) * 1e3,
                               "launch": "the batched launches of a 10-step period as one graph on the "
                                         "first member's stream (radiation eager at the period start, one "
                                         "deferred check per member per period)"}
    ens.close()
    return out


def graph_timing(args):
    """Steps recorded as one graph (atm_graph_begin / end, 10 steps with two
    remaps) against the same calls issued one at a time, FP64, fused path,
    no per-step stability check in either (one check per 10 steps)."""
    from atmos import driver
    from atmos.config import Config
    from atmos.ic import make_initial_state
    from atmos.model import Model
    dt = args.dt

    def steps(m):
        for k in range(10):
            m.cgrid_dynamics_step(dt / 2)
            m.cgrid_halo_exchange()
            m.dgrid_dynamics_step(dt)
            if k % 5 == 4:
                m.perform_vertical_remapping()
            m.dgrid_halo_exchange()

    res = {}
    for nx, ny, L, cycles in ((64, 64, 20, 100), (args.nx, args.ny, args.nlay, 10)):
        ic = make_initial_state(nx, ny, L, args.dx)
        out = {}
        for mode in ("eager", "graph"):
            m = Model(nx, ny, L, args.dx)
            cfg = Config(dt_max=dt, t_final=1e12, dt_output=1e12, dt_radiation=1e12, dt_physics=1e12,
                         dt_remap=1e12, radiation_on=False, physics_on=False)
            driver.init_model(m, ic, cfg)
            if mode == "graph":
                m.graph_begin()
                steps(m)
                m.graph_end()
                run = lambda: m.graph_launch(1)  # noqa: E731
            else:
                run = lambda: steps(m)  # noqa: E731
            run()
            if not m.check_stability():
                raise RuntimeError("graph timing: unstable")
            t0 = time.perf_counter()
            for _ in range(cycles):
                run()
                if not m.check_stability():
                    raise RuntimeError("graph timing: unstable")
            sec = time.perf_counter() - t0
            out[f"{mode}_ms_per_step"] = sec / (10 * cycles) * 1e3
            m.close()
        res[f"{nx}x{ny}x{L}"] = out
    return res


def cpu_baseline(args):
    """The oracle (C restatement of the reference) on a bounded sample of the
    same workload, threaded over the host's cores (OpenMP: OMP_NUM_THREADS or
    all cores), and on one core for reference."""
    import oracle as O
    from atmos import driver
    from atmos.config import Config
    from atmos.ic import make_initial_state

    def run(n, steps):
        L = args.nlay
        ic = make_initial_state(n, n, L, args.dx)
        cfg = Config(dt_max=args.dt, t_final=1e12, dt_output=1e12, dt_radiation=1e12, dt_physics=1e12,
                     dt_remap=args.remap_every * args.dt, radiation_on=False, physics_on=False)
        o = O.Oracle(n, n, L, args.dx, args.dx)
        st = driver.init_model(o, ic, cfg)
        t0 = time.perf_counter()
        _, done, _ = driver.run_model(o, cfg, st, max_steps=steps)
        sec = time.perf_counter() - t0
        return n * n * L * done / sec, done, sec

    default = O.set_threads(0)
    cores = O.set_threads(args.cpu_threads) if args.cpu_threads > 0 else default
    n, L = args.cpu_nx, args.nlay
    v, steps, sec = run(n, args.cpu_steps)
    O.set_threads(1)
    n1 = args.cpu_nx // 2
    v1, steps1, sec1 = run(n1, max(1, args.cpu_steps // 2))
    O.set_threads(default)
    return {"value": v, "unit": "cell-steps/s", "cores": cores, "kind": "port",
            "sample": f"{n}x{n}x{L}, {steps} steps incl. {steps // args.remap_every} remap, {sec:.1f} s on "
                      f"{cores} host threads (oracle/atm_oracle.c, -O2, OpenMP over levels/columns, glibc libm)",
            "single_core": {"value": v1, "cores": 1,
                            "sample": f"{n1}x{n1}x{L}, {steps1} steps, {sec1:.1f} s on 1 host core"}}


def surface_pressure_and_temperature(m, ptop):
    """ps = p_top + sum(dp) and layer temperature T = pt * playkap (the
    dynamics' playkap form, mod_sw_dyn.f90:99-101) of a model's state."""
    import numpy as np
    kappa = float(np.float32(287.052874)) / 1005.0
    dp, pt = m.get("dp"), m.get("pt")
    plev = np.concatenate([np.cumsum(dp[::-1], axis=0)[::-1] + ptop, np.full((1,) + dp.shape[1:], ptop)])
    pk = plev ** kappa
    return plev[0], pt * ((pk[1:] - pk[:-1]) / np.log(pk[1:] / pk[:-1]))


def mixed_timing(args, ic, cfg_kw):
    """The mixed-precision form (ATM_PRECISION_MIXED) on the same workload:
    throughput over the same steps, and after --mixed-error-steps steps from
    the initial state its relative RMS and max error against the FP64 path
    in surface pressure, temperature and the prognostic fields, with the
    drift of the total mass."""
    import numpy as np

    from atmos import driver
    from atmos.config import Config
    from atmos.model import Model
    nx, ny, L = args.nx, args.ny, args.nlay
    m = Model(nx, ny, L, args.dx, precision="mixed")
    st = driver.init_model(m, ic, Config(**cfg_kw))
    if args.autotune == "on":
        m.autotune()
    cfg = Config(**cfg_kw)
    t = 0.0
    for _ in range(args.warmup):
        t += driver.step(m, cfg, st, t)
        assert m.check_stability()
    m.sync()
    # timed as the FP64 steps are: events around the roofline's candidates
    # only, each step's check read once the next step is issued
    m.timer_select(ROOFLINE_GROUPS)
    m.timer_enable(True)
    t0 = time.perf_counter()
    pending = False
    for _ in range(args.steps):
        t += driver.step(m, cfg, st, t)
        stable = m.check_stability_end() if pending else True
        m.check_stability_begin()
        pending = True
        if not stable:
            raise SystemExit("mixed: unstable")
    if pending and not m.check_stability_end():
        raise SystemExit("mixed: unstable")
    m.sync()
    sec = time.perf_counter() - t0
    # every group over a short pass after the timed steps
    n_groups = max(10, args.steps // 5 // args.remap_every * args.remap_every)
    m.timer_select(None)
    m.timer_enable(True)
    for _ in range(n_groups):
        t += driver.step(m, cfg, st, t)
        assert m.check_stability()
    kernels = {}
    for name in ("c1", "c2", "d1", "d2", "scan", "halo", "remap"):
        n, ms = m.timer_read(name)
        if n:
            kernels[name] = ms / n_groups
    m.timer_enable(False)
    rad = radiation_timing(m, args, t) if args.radiation_calls > 0 else None
    del m
    # error after N steps from the same initial state, dynamics + remap +
    # radiation every 10 steps (each form with its own radiation)
    cfg_err = dict(cfg_kw, radiation_on=True, dt_radiation=10 * args.dt)
    res = {}
    for prec in ("fp64", "mixed"):
        mm = Model(nx, ny, L, args.dx, precision=prec)
        c = Config(**cfg_err)
        s2 = driver.init_model(mm, ic, c)
        mass0 = float(mm.get("dp").sum())
        tt = 0.0
        for _ in range(args.mixed_error_steps):
            tt += driver.step(mm, c, s2, tt)
        ps, T = surface_pressure_and_temperature(mm, float(ic["ak"][0]))
        res[prec] = {"ps": ps, "T": T, "mass_drift": (float(mm.get("dp").sum()) - mass0) / mass0,
                     **{k: mm.get(k) for k in ("dp", "pt", "ud", "vd")}}
        del mm

    def rel(a, b):
        return {"rel_rms": float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2))),
                "rel_max": float(np.max(np.abs(a - b)) / np.max(np.abs(b)))}
    a, b = res["mixed"], res["fp64"]
    err = {k: rel(a[k], b[k]) for k in ("ps", "T", "dp", "pt", "ud", "vd")}
    err["mass_drift"] = {"mixed": a["mass_drift"], "fp64": b["mass_drift"]}
    # radiation alone on the same (initial) state: net flux of the mixed
    # form (float g-point optics) against the FP64 solve
    nf = {}
    for prec in ("fp64", "mixed"):
        mm = Model(nx, ny, L, args.dx, precision=prec)
        driver.init_model(mm, ic, Config(**cfg_kw))
        mm.calculate_radiative_heating(0.0)
        nf[prec] = mm.get("net_flux")
        del mm
    rad_err = rel(nf["mixed"], nf["fp64"])
    value = nx * ny * L * args.steps / sec
    return {"value": value, "unit": "cell-steps/s", "ms_per_step": sec / args.steps * 1e3,
            "dtype": "f32 reconstruction, fluxes and heating rate; f64 updates, pressures, pkap, gz, "
                     "pressure gradient, remap",
            "kernels_ms_per_step": kernels,
            "error_vs_fp64": {"steps": args.mixed_error_steps, "radiation_every": 10, **err},
            "radiation_ms_per_call": rad["ms_per_call"] if rad else None,
            "radiation_net_flux_error_vs_fp64": rad_err}


def radiation_timing(m, args, time_s):
    """calculate_radiative_heating on the benchmark state (all interior
    columns, 16 g-points, shortwave and longwave), timed with HIP events."""
    m.calculate_radiative_heating(time_s)  # warm-up (first call also checks the tables)
    m.sync()
    m.timer_enable(True)
    for _ in range(args.radiation_calls):
        m.calculate_radiative_heating(time_s)
    m.sync()
    n, ms = m.timer_read("radiation")
    m.timer_enable(False)
    per = ms / n
    cols = args.nx * args.ny
    flops = RAD_FLOP_PER_LAYER_G * RAD_G_POINTS * args.nlay * cols
    achieved = flops / (per / 1e3) / 1e12
    return {"ms_per_call": per, "columns_per_s": cols / (per / 1e3), "calls": n,
            "achieved": achieved, "peak": FP64_VECTOR_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / FP64_VECTOR_PEAK_TFLOPS, "bound": "valu",
            "flop_per_call": flops}


# Algorithmic flops of the two-stream columns, counted from the
# reference's statements (src/mod_radiation.f90:158-398) with every +, -, *,
# /, sqrt and exp one flop: per layer and shortwave g-point 149 (gas,
# Rayleigh and cloud optical depths 21 with the k evaluation, delta-Eddington
# 9, coefficients, guard and direct-beam R/T 89, the beam and its sources 3,
# diffuse R/T 27), 19 for calculate_fluxes' two sweeps, 1 for adding the
# beam, 2 for the g-point sums; per layer and longwave g-point 72 (optics
# 40, R/T and the Planck sources 32) + 19 + 2; per layer and column 9 shared
# by the g-points (ln p, T - T_ref, dm, cloud scattering, T^4).
TS_FLOPS_SW = 149 + 19 + 1 + 2
TS_FLOPS_LW = 72 + 19 + 2
TS_FLOPS_LAYER = 9
FP64_VALU_PEAK_TFLOPS = 78.6


def twostream_timing(args):
    """The general two-stream columns at the workload's size: FP64 (the
    reference's formulas) and the mixed form, timed with HIP events on the
    library's stream over back-to-back calls; errors against each other on
    every column, against the binary128 evaluation on a sample, and per
    g-point; the oracle on a bounded sample of columns on the host."""
    import numpy as np
    import torch

    import oracle as O
    from atmos import twostream as TS

    ncol, L, nsw, nlw = args.ts_ncol, 64, 32, 32
    cols = TS.synthetic_columns(ncol, nlay=L, n_sw=nsw, n_lw=nlw, seed=20261203 + 300)
    dev = TS.to_device(cols)
    flops = ncol * L * (nsw * TS_FLOPS_SW + nlw * TS_FLOPS_LW + TS_FLOPS_LAYER)
    res = {"workload": f"{ncol} columns x {L} layers, {nsw} SW + {nlw} LW g-points, synthetic k-distribution, "
                       "clouds in 20% of layers (g 0.85), mu0 U[0.1,1], albedo U[0.05,0.3]",
           "flops_per_call": flops}
    nets = {}
    for prec in ("fp64", "mixed"):
        out = torch.empty((ncol, L + 1), dtype=torch.float64, device="cuda")
        _, _, st = TS.twostream(dev, prec, out=out)  # warm-up, with diagnostics
        TS.timer_begin()
        for _ in range(args.ts_calls):
            TS.twostream(dev, prec, out=out, stats=False, wait=False)
        ms = TS.timer_end() / args.ts_calls
        nets[prec] = out.cpu().numpy()
        tf = flops / (ms * 1e-3) / 1e12
        res[prec] = {"ms_per_call": ms, "columns_per_s": ncol / (ms * 1e-3), "tflops": tf,
                     "frac_fp64_valu_peak": tf / FP64_VALU_PEAK_TFLOPS, **st}
    res["speedup_mixed"] = res["fp64"]["ms_per_call"] / res["mixed"]["ms_per_call"]

    def rel(a, b):
        sc = np.maximum(np.max(np.abs(b), axis=-1), 1e-300)
        e = np.max(np.abs(a - b), axis=-1) / sc
        return {"max": float(e.max()), "rms": float(np.sqrt(np.mean(e ** 2)))}

    res["mixed_vs_fp64"] = rel(nets["mixed"], nets["fp64"])
    idx = np.arange(0, ncol, max(1, ncol // 256))
    sub = {**cols, **{n: cols[n][idx] for n in TS.LAYER_FIELDS + TS.LEVEL_FIELDS + TS.COLUMN_FIELDS},
           "ncol": len(idx)}
    q = O.twostream(sub, TS.GUARD_FP64, extended="q")
    res["vs_binary128"] = {"sample_columns": len(idx), "fp64": rel(nets["fp64"][idx], q),
                           "mixed": rel(nets["mixed"][idx], q)}
    # per g-point: the same columns with one g-point's optics
    per_g = []
    for kind, n in (("sw", nsw), ("lw", nlw)):
        for g in range(n):
            one = TS.select_gpoints(cols, sw=[g] if kind == "sw" else [], lw=[g] if kind == "lw" else [])
            d1 = {**dev, "optics": torch.tensor(one["optics"], dtype=torch.float64, device="cuda"),
                  "n_sw": one["n_sw"], "n_lw": one["n_lw"]}
            a, _, _ = TS.twostream(d1, "fp64", stats=False)
            b, _, _ = TS.twostream(d1, "mixed", stats=False)
            per_g.append((kind, g, rel(b.cpu().numpy(), a.cpu().numpy())))
    worst = max(per_g, key=lambda t: t[2]["max"])
    res["mixed_vs_fp64_per_g"] = {"max_over_g": worst[2]["max"], "worst_g": f"{worst[0]}{worst[1]}",
                                  "rms_over_g": float(np.sqrt(np.mean([t[2]["rms"] ** 2 for t in per_g])))}
    # the oracle (C restatement, FP64) on a bounded sample of columns
    nthreads = O.set_threads(0)
    ncpu = min(ncol, 8192)
    sub = {**cols, **{n: cols[n][:ncpu] for n in TS.LAYER_FIELDS + TS.LEVEL_FIELDS + TS.COLUMN_FIELDS},
           "ncol": ncpu}
    t0 = time.perf_counter()
    O.twostream(sub, TS.GUARD_FP64)
    sec = time.perf_counter() - t0
    res["cpu_oracle"] = {"columns_per_s": ncpu / sec, "cores": nthreads,
                         "sample": f"{ncpu} columns, {sec:.2f} s on {nthreads} host threads (oracle/ts_columns.inc)"}
    return res


def adjoint_timing(args):
    """Gradient of a surface-pressure cost through a window of steps: the
    tangent-linear/adjoint workload (256 x 256 x 48, dx = 50 km, dt = 40 s,
    a 90-step window with remaps every 5 steps, 10 snapshots of binomial
    checkpointing; J = mean-square deviation of the surface pressure over a
    32 x 32 box at the end).  The forward trajectory runs in the mixed form
    (or FP64), the adjoint kernels in FP64.  Ratio = wall time of the
    gradient (checkpoint recomputation + adjoint steps) / wall time of the
    forward window.  A central finite difference of J (eps 1e-7 along a
    random direction of pt, ud, vd) checks the gradient on an FP64 run."""
    import numpy as np
    import torch

    from atmos import driver
    from atmos.adjoint import (Window, binomial_reps, recompute_steps, schedule_of, surface_pressure_cost,
                               window_apply)
    from atmos.config import Config
    from atmos.ic import make_initial_state
    from atmos.model import Model

    nx, L, dx, dt, n, snaps = args.adj_nx, args.adj_nlay, 50e3, 40.0, args.adj_steps, args.adj_snapshots
    ic = make_initial_state(nx, nx, L, dx, seed=20261203 + 400)
    # Far from the pressure bump the recipe's surface pressure is uniform, so
    # neighbouring dp values tie exactly and every limiter min/max there sits
    # on its switch: J is not differentiable at that state and finite
    # differences measure the kinks.  Smooth noise of 1 Pa in dp and 0.1 K in
    # T removes the ties.
    from atmos.ic import spectral_noise
    rng = np.random.default_rng(20261203 + 401)
    for k in range(L):
        ic["dp"][k] += spectral_noise(rng, nx, nx)
        ic["pt"][k] *= 1.0 + 3e-4 * spectral_noise(rng, nx, nx)
    cfg = Config(dt_max=dt, t_final=1e12, dt_output=1e12, dt_radiation=1e12, dt_physics=1e12, dt_remap=5 * dt,
                 radiation_on=False, physics_on=False)
    sched = schedule_of(n, dt, 5)
    c0 = nx // 2 - 16
    box = (c0, c0 + 32, c0, c0 + 32)
    p_top = float(ic["ak"][0])
    res = {"grid": f"{nx}x{nx}x{L}", "dx_km": dx / 1e3, "dt": dt, "window_steps": n, "snapshots": snaps,
           "state": "seeded recipe + smooth noise (1 Pa dp, 0.1 K T)",
           "remap_every": 5, "repetitions": binomial_reps(n, snaps - 1),
           "recomputed_steps": recompute_steps(n, snaps)}

    def make(precision):
        m = Model(nx, nx, L, dx, precision=precision)
        driver.init_model(m, ic, cfg)
        return m

    def grad(m, timing):
        w = Window(m, sched, snapshots=snaps)
        x0 = [torch.tensor(m.get(name), dtype=torch.float64, device="cuda", requires_grad=True)
              for name in ("dp", "pt", "ud", "vd")]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = window_apply(w, *x0)
        J = surface_pressure_cost(out[0], p_top, box)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        J.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        if timing is not None:
            timing.update(forward_s=t1 - t0, gradient_s=t2 - t1, counts=dict(w.counts))
        return J.item(), [x.grad for x in x0], [x.detach() for x in x0], w

    for precision in ("mixed", "fp64"):
        m = make(precision)
        grad(m, None)  # warm-up (allocations, kernel loading)
        t = {}
        J, g, x0, w = grad(m, t)
        res[precision] = {"forward_ms": t["forward_s"] * 1e3, "gradient_ms": t["gradient_s"] * 1e3,
                          "ratio": t["gradient_s"] / t["forward_s"], "J": J, **t["counts"]}
        if precision == "fp64":
            # finite differences of J along pt, ud, vd (T and the winds)
            rng = np.random.default_rng(7)
            d = [torch.zeros_like(x0[0]), 1e-3 * torch.tensor(rng.standard_normal(x0[1].shape), device="cuda"),
                 torch.tensor(rng.standard_normal(x0[2].shape), device="cuda"),
                 torch.tensor(rng.standard_normal(x0[3].shape), device="cuda")]
            gd = float(sum((gi * di).sum() for gi, di in zip(g, d)))
            eps = 1e-7

            def J_at(e):
                ww = Window(m, sched, snapshots=snaps)
                out = window_apply(ww, *[xi + e * di for xi, di in zip(x0, d)])
                return float(surface_pressure_cost(out[0], p_top, box))

            fd = (J_at(eps) - J_at(-eps)) / (2 * eps)
            fd6 = (J_at(10 * eps) - J_at(-10 * eps)) / (20 * eps)
            res["fd_check"] = {
                "eps": eps, "grad_dot_d": gd, "fd": fd, "rel_err": abs(fd - gd) / abs(gd),
                "fd_eps_x10": fd6, "rel_err_eps_x10": abs(fd6 - gd) / abs(gd),
                "note": "differences of the model itself: the constrained PPM switch "
                        "(mod_sw_dyn.f90:802-829) is discontinuous and, over 90 steps of 3.1M points, "
                        "switches within ~1e-7 of their threshold flip, so J jumps between "
                        "eps and -eps; the frozen-branch differences the linearisation is the "
                        "derivative of are checked in tests/test_lin_frozen_fd.py and "
                        "tests/test_gpu_adjoint.py"}
            # the tangent-linear model through the same window along d (its
            # directional derivative of J) against the adjoint's <grad, d>
            xs = [x.cpu().numpy() for x in x0]
            for name, x in zip(("dp", "pt", "ud", "vd"), xs):
                m.put(name, x)
                m.halo_exchange(name)
            for name, di in zip(("dp", "pt", "ud", "vd"), d):
                m.put("tl_" + name, di.cpu().numpy())
            for dt_k, remap in sched:
                m.tl_step(dt_k, remap)
            dpN, tl = m.get("dp"), m.get("tl_dp")
            ps = p_top + dpN.sum(axis=0)
            j0, j1, i0, i1 = box
            bps = ps[j0:j1, i0:i1]
            dJ = float(np.sum((2.0 * (bps - bps.mean()) / bps.size)[None] * tl[:, j0:j1, i0:i1]))
            res["tl_check"] = {"tl_dJ": dJ, "grad_dot_d": gd, "rel_err": abs(dJ - gd) / abs(gd)}
        # one adjoint step alone and one linearised forward step alone
        m.timer_enable(False)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(5):
            m.ad_step(dt, False)
        torch.cuda.synchronize()
        res[precision]["ad_step_ms"] = (time.perf_counter() - t0) / 5 * 1e3
        del m, w
        torch.cuda.empty_cache()
    res["ratio"] = res["mixed"]["ratio"]
    return res


def adjoint_step_timing(args):
    """Single linearised steps of the tangent-linear/adjoint workload's grid
    (256 x 256 x 48, dx = 50 km, dt = 40 s), with and without a remap:
    wall time per adjoint step and per tangent-linear step, against a
    forward step of the fused FP64 path, with every adjoint and tangent
    field nonzero (seeded noise).  No torch: only the C ABI."""
    import numpy as np

    from atmos import driver
    from atmos.config import Config
    from atmos.ic import make_initial_state
    from atmos.model import Model

    nx, L, dx, dt = args.adj_nx, args.adj_nlay, 50e3, 40.0
    ic = make_initial_state(nx, nx, L, dx, seed=20261203 + 400)
    cfg = Config(dt_max=dt, t_final=1e12, dt_output=1e12, dt_radiation=1e12, dt_physics=1e12, dt_remap=5 * dt,
                 radiation_on=False, physics_on=False)
    m = Model(nx, nx, L, dx)
    st = driver.init_model(m, ic, cfg)
    m.lin_init()
    rng = np.random.default_rng(20261203 + 402)
    res = {"grid": f"{nx}x{nx}x{L}", "dt": dt}
    for kind in ("ad", "tl"):
        for remap in (False, True):
            for name in ("dp", "pt", "ud", "vd"):
                m.put(f"{kind}_{name}", rng.standard_normal((L, nx, nx)))
            step = m.ad_step if kind == "ad" else m.tl_step
            step(dt, remap)
            m.sync()
            n = 5
            t0 = time.perf_counter()
            for _ in range(n):
                step(dt, remap)
            m.sync()
            res[f"{kind}_step_ms{'_remap' if remap else ''}"] = (time.perf_counter() - t0) / n * 1e3
    t = 0.0
    for _ in range(3):
        t += driver.step(m, cfg, st, t)
    m.sync()
    t0 = time.perf_counter()
    n = 10
    for _ in range(n):
        t += driver.step(m, cfg, st, t)
        if not m.check_stability():
            raise SystemExit("adjoint steps: the forward run is unstable")
    m.sync()
    res["forward_step_ms"] = (time.perf_counter() - t0) / n * 1e3
    m.close()
    return res


def main():
    args = parse()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              f"(the path would shard by 2-D tiles with 7-wide halo exchanges)", file=sys.stderr)
        sys.exit(2)
    import numpy as np  # noqa: F401

    if args.adjoint == "on" or args.twostream == "on":
        # torch carries its own HIP runtime beside the library's; it has to
        # bring up the device first in a process that uses both
        import torch
        torch.cuda.init()
    from atmos import driver
    from atmos.config import Config
    from atmos.ic import make_initial_state
    from atmos.model import Model

    nx, ny, L = args.nx, args.ny, args.nlay
    ic = make_initial_state(nx, ny, L, args.dx)
    cfg_kw = dict(dt_max=args.dt, t_final=1e12, dt_output=1e12, dt_radiation=1e12, dt_physics=1e12,
                  dt_remap=args.remap_every * args.dt, radiation_on=False, physics_on=False)
    cfg = Config(**cfg_kw)
    m = Model(nx, ny, L, args.dx)
    if args.path != m.path:
        m.set_path(args.path)
    st = driver.init_model(m, ic, cfg)
    tuning = m.autotune() if args.autotune == "on" else m.tuning
    tile_rows = m.tiles
    time_s = 0.0
    for _ in range(args.warmup):
        time_s += driver.step(m, cfg, st, time_s)
        assert m.check_stability()
    m.sync()
    # HIP events in the timed region around the roofline's candidate kernels
    # only; the other groups' times come from a short pass after it
    m.timer_select(ROOFLINE_GROUPS)
    m.timer_enable(True)
    t0 = time.perf_counter()
    pending = False
    for _ in range(args.steps):
        time_s += driver.step(m, cfg, st, time_s)
        if args.checks == "sync":
            stable = m.check_stability()
        else:  # every step's check, read once the next step is issued
            stable = m.check_stability_end() if pending else True
            m.check_stability_begin()
            pending = True
        if not stable:
            raise SystemExit("unstable")
    if pending and not m.check_stability_end():
        raise SystemExit("unstable")
    m.sync()
    sec = time.perf_counter() - t0
    cells = nx * ny * L
    value = cells * args.steps / sec

    kernels = {}
    for name in ROOFLINE_GROUPS:
        n, ms = m.timer_read(name)
        if n:
            kernels[name] = (n, ms)
    # every group over a short pass (whole remap intervals) after the timed steps
    n_groups = max(10, args.steps // 5 // args.remap_every * args.remap_every)
    m.timer_select(None)
    m.timer_enable(True)
    for _ in range(n_groups):
        time_s += driver.step(m, cfg, st, time_s)
        assert m.check_stability()
    groups = {}
    for name in ("c1", "c2", "d1", "d2", "scan", "halo", "remap"):
        n, ms = m.timer_read(name)
        if n:
            groups[name] = ms / n_groups
    m.timer_enable(False)
    dom = max((k for k in kernels if k in ALG_BYTES_PER_CELL), key=lambda k: kernels[k][1])
    n, ms = kernels[dom]
    avg_s = ms / n / 1e3
    alg = ALG_BYTES_PER_CELL[dom] * cells
    achieved = alg / avg_s / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as fh:
            traffic = json.load(fh).get(f"{args.path}:{dom}:{nx}x{ny}x{L}")
    out = {
        "metric": METRIC, "value": value, "unit": "cell-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": sec / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"FP64 dyn+remap {nx}x{ny}x{L}, dx={args.dx / 1e3:g} km, dt={args.dt:g} s, "
                               f"remap every {args.remap_every} steps, radiation/physics off, seeded IC",
                   "nx": nx, "ny": ny, "nlay": L, "path": args.path, "precision": "fp64",
                   "stability_check": "every step, " + ("read after the next step is issued" if args.checks == "deferred"
                                                         else "waited for before the next step")},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": "profiles/traffic.json: FETCH_SIZE + WRITE_SIZE of this kernel "
                                       "from a committed rocprofv3 pass, scaled by the copy calibration",
                     "alg_bytes_per_launch": alg, "avg_launch_ms": avg_s * 1e3},
        "whole_step_roofline": whole_step_roofline(value),
        "kernels_ms_per_step": groups,
        "kernels_ms_per_step_source": f"HIP-event means over {n_groups} steps after the timed ones, every group "
                                      "timed; the timed steps time only the roofline's candidates "
                                      f"({', '.join(ROOFLINE_GROUPS)})",
        "adjoint_forward_cost_ratio": None,
        "tuning": {"corner_levels_per_workgroup": tuning, "tile_rows_dp_pt_wind": list(tile_rows),
                   "autotune": args.autotune}
    }
    if args.radiation_calls > 0:
        out["radiation"] = radiation_timing(m, args, time_s)
    if args.mixed == "on" and args.path == "fused":
        del m
        out["mixed"] = mixed_timing(args, ic, cfg_kw)
        if out.get("radiation") and out["mixed"].get("radiation_ms_per_call"):
            # derived from the measured times above: a step plus a tenth of a
            # radiation call (radiation every 10 steps)
            f64 = out["ms_per_step"] + out["radiation"]["ms_per_call"] / 10
            mix = out["mixed"]["ms_per_step"] + out["mixed"]["radiation_ms_per_call"] / 10
            out["mixed"]["with_radiation_every_10"] = {"fp64_ms_per_step": f64, "mixed_ms_per_step": mix,
                                                       "speedup": f64 / mix, "derived": True}
    if args.adjoint == "on":
        out["adjoint"] = adjoint_timing(args)
        out["adjoint_forward_cost_ratio"] = out["adjoint"]["ratio"]
    elif args.adjoint == "steps":
        out["adjoint_steps"] = adjoint_step_timing(args)
    if args.twostream == "on":
        out["twostream"] = twostream_timing(args)
    if args.ensemble == "on" and args.path == "fused":
        out["ensemble"] = ensemble_timing(args)
    if args.graph == "on" and args.path == "fused":
        out["graph"] = graph_timing(args)
    if args.cpu_baseline == "on":
        out["cpu_baseline"] = cpu_baseline(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
