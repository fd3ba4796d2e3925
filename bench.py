"""Benchmark of the MI355X time-advance hot path.

Workload (default): the single curvilinear block 512x256x128, FP64, as the
headline configuration states it: x uniform and periodic, slip walls on both
y-faces (type_wall_bc = 2 with the slip flag) with tanh clustering (ratio
1.05), a wavy lower wall y' = y + 0.05 sin(2 pi x/Lx)(1 - y/Ly), z periodic,
3-D GCL metrics, an isentropic vortex (strength 5, radius 0.5) advected at
Mach 0.5 in a uniform free stream, 4-stage low-storage RK, 5-pt viscous terms
in the last stage, shock-capturing filter with Ducros sensor (quiescent).

One "step" = one full RK time step (4 stages + filter) of that block, all
inputs resident in HBM.  Metric: Mcell-RK-stages/s = cells * nrk * steps / t.

Prints ONE JSON line.  Usage:
    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s

# HBM bytes per k_stage launch on the 512x256x128 curvilinear blocks (tuned
# tiles, k-chunk 128), averaged over the 4 stages (3 Euler launches + the
# viscous last stage): rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE per dispatch,
# FETCH_SIZE scaled by 2.000 and WRITE_SIZE by 1.000 (their readings for a
# 1 GiB device copy).  Summaries: profiles/step_kernels_c3_slip_channel_512x256x128_fp64.md
# (the channel with slip walls), profiles/step_kernels_c3_channel_512x256x128_fp64.md
# (no-slip walls), profiles/step_kernels_c3_512x256x128_fp64.md (the periodic block).
STAGE_TRAFFIC = {
    ("channel", 512, 256, 128): ((3 * 4.598e9 + 5.484e9) / 4,
                                 "profiles/step_kernels_c3_slip_channel_512x256x128_fp64.md"),
    ("channel_noslip", 512, 256, 128): ((3 * 4.597e9 + 5.510e9) / 4,
                                        "profiles/step_kernels_c3_channel_512x256x128_fp64.md"),
    ("periodic", 512, 256, 128): ((3 * 4.589e9 + 5.338e9) / 4, "profiles/step_kernels_c3_512x256x128_fp64.md"),
}


def stage_words_per_cell(case) -> float:
    """Algorithmic FP64 words per interior cell of one stage-kernel launch,
    averaged over the nrk stages: stage 1 reads U^n and writes U1 (10 W),
    later stages also read U^n (15 W); curvilinear metrics add 10 W per stage
    (9 metric terms and 1/J) on a 3-D curvilinear block, nothing per cell on a
    2.5-D one (its 2-D metrics are read once per column); the viscous last
    stage reads the 4 viscous increments (4 W)."""
    w = [10.0] + [15.0] * (case.nrk - 1)
    if int(case.curvilinear) == 1:
        w = [x + 10.0 for x in w]
    if case.iorder_visc:
        w[-1] += 4.0
    return sum(w) / len(w)


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(case, nsteps: int, single_steps: int = 1):
    """Time the CPU oracle on the host cores for a bounded sample: every
    thread OpenMP gives it (all host cores), then one thread."""
    from oracle.oracle import Oracle, num_threads, set_num_threads
    o = Oracle(case)
    U = case.copy_state()
    o.step(U, 1)  # first-touch of its work arrays
    threads = num_threads()

    def timed(n):
        U = case.copy_state()
        t0 = time.perf_counter()
        o.step(U, n)
        return time.perf_counter() - t0

    t = timed(nsteps)
    val = case.ncells * case.nrk * nsteps / t / 1e6
    set_num_threads(1)
    t1 = timed(single_steps)
    set_num_threads(threads)
    val1 = case.ncells * case.nrk * single_steps / t1 / 1e6
    return {"value": round(val, 3), "unit": "Mcell-RK-stages/s", "cores": threads,
            "kind": "port",
            "sample": f"{nsteps} full RK steps of the same {case.nx}x{case.ny}x{case.nz} block, "
                      f"oracle/musica_oracle.c, OpenMP over k on {threads} threads, {t:.2f} s; "
                      f"{single_steps} step on 1 thread, {t1:.2f} s",
            "cpu_model": _cpu_model(), "nproc": os.cpu_count(),
            "single_thread": {"value": round(val1, 3), "unit": "Mcell-RK-stages/s", "cores": 1,
                              "steps": single_steps, "seconds": round(t1, 2)}}


def copy_calibration(nbytes: int = 1 << 30, reps: int = 5) -> int:
    """A plain device-to-device copy of nbytes, reps times: under rocprofv3
    --pmc FETCH_SIZE / WRITE_SIZE it gives the counter's reading for a known
    number of bytes read and written per dispatch."""
    import torch
    x = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "dispatches": reps}), flush=True)
    return 0


def run_leg(case, precision: int, steps: int, warmup: int, tune: bool, total_steps: int = 0):
    """Advance `case` on the GPU: autotune, warm up, then time `steps` steps
    replayed from the captured graph (the production path); then the same
    number of steps launched one by one with per-kernel HIP events on the
    library's stream (the per-kernel split and the stage kernel's launch
    time); then, untimed, up to `total_steps` steps in all (the state the
    precision comparison reads)."""
    from musica2_newgrid_amd import _lib
    from musica2_newgrid_amd.solver import GpuSolver

    solver = GpuSolver(case, precision=precision)
    if tune:
        solver.autotune()
    tuned = {"euler_stages": solver.stage_config(0), "last_stage": solver.stage_config(1),
             "visc": solver.stage_config(2), "vdiv": solver.stage_config(3), "filter": solver.stage_config(4)}
    solver.step(warmup)
    solver.sync()
    t0 = time.perf_counter()
    solver.step(steps)
    solver.sync()
    t = time.perf_counter() - t0

    solver.profile(True)
    solver.sync()
    t0 = time.perf_counter()
    solver.step(steps)
    solver.sync()
    t_eager = time.perf_counter() - t0
    times = {_lib.K_NAMES[k]: solver.kernel_time(k) for k in _lib.K_NAMES}
    solver.profile(False)
    done = warmup + 2 * steps
    if total_steps > done:
        solver.step(total_steps - done)
        done = total_steps
    rho = solver.get_state()[0]
    solver.close()
    return {"t": t, "t_eager": t_eager, "times": times, "tuned": tuned, "rho": rho, "steps_done": done}


def multiblock_leg(steps: int, warmup: int):
    """The uneven 7-block workload (384x192x96, 2x256x128x96, 3x128x128x64,
    96^3): one batched launch per kernel and stage for all blocks, steps
    replayed from a captured HIP graph, autotuned tiles."""
    from musica2_newgrid_amd import _lib, multiblock as mb
    from musica2_newgrid_amd.solver import GpuSolver

    flags, blocks, bcs = mb.uneven_islands()
    cells = sum(bc.ncells for bc in bcs)
    with GpuSolver(flags, topology=blocks, block_cases=bcs) as s:
        s.autotune()
        s.step(warmup)
        s.sync()
        t0 = time.perf_counter()
        s.step(steps)
        s.sync()
        t = time.perf_counter() - t0
        # per-kernel split from a separate, eagerly launched and event-timed pass
        s.profile(True)
        s.step(2)
        times = {_lib.K_NAMES[k]: s.kernel_time(k) for k in _lib.K_NAMES}
        s.profile(False)
    return {"value": round(cells * flags.nrk * steps / t / 1e6, 2), "unit": "Mcell-RK-stages/s",
            "ms_per_step": round(t / steps * 1e3, 4), "steps": steps, "blocks": len(blocks), "cells": cells,
            "workload": "7 curvilinear blocks 384x192x96, 2x256x128x96, 3x128x128x64, 96^3 in 4 periodic "
                        "islands of 1-to-1 interfaces (all orientation kinds), M=0.3 + 1e-3 pressure pulse, "
                        "viscous, selective filter, RK4, FP64, graph replay",
            "kernel_ms_per_step_eager": {k: round(ms / 2, 4) for k, (n, ms) in times.items()}}


def tgv_leg(steps: int = 100):
    """The TGV 64^3 (Cartesian, viscous, selective filter, RK4, CFL 0.8, 100
    steps): graph replay, small enough to be launch-bound."""
    from musica2_newgrid_amd import cases
    from musica2_newgrid_amd.solver import GpuSolver
    case = cases.tgv(64)
    with GpuSolver(case) as s:
        s.autotune()
        s.step(5)
        s.sync()
        t0 = time.perf_counter()
        s.step(steps)
        s.sync()
        t = time.perf_counter() - t0
    return {"value": round(case.ncells * case.nrk * steps / t / 1e6, 2), "unit": "Mcell-RK-stages/s",
            "ms_per_step": round(t / steps * 1e3, 4), "steps": steps,
            "workload": "TGV 64^3, Cartesian periodic, M=0.1, Re=1600, viscous, selective filter chi=0.2, "
                        "RK4, CFL 0.8, FP64, graph replay"}


def channel_leg(steps: int, warmup: int):
    """A 512x256x128 Cartesian channel with no-slip walls on both j-faces
    (stretched y-grid), periodic x and z: the wall variants of the stage and
    filter kernels plus the wall-row kernels, autotuned, graph replay."""
    from musica2_newgrid_amd import _lib, cases
    from musica2_newgrid_amd.solver import GpuSolver
    case = cases.channel(512, 256, 128, lengths=(8.0, 2.0, 2.0))
    with GpuSolver(case) as s:
        s.autotune()
        s.step(warmup)
        s.sync()
        t0 = time.perf_counter()
        s.step(steps)
        s.sync()
        t = time.perf_counter() - t0
        s.profile(True)
        s.step(2)
        times = {_lib.K_NAMES[k]: s.kernel_time(k) for k in _lib.K_NAMES}
        s.profile(False)
    return {"value": round(case.ncells * case.nrk * steps / t / 1e6, 2), "unit": "Mcell-RK-stages/s",
            "ms_per_step": round(t / steps * 1e3, 4), "steps": steps,
            "workload": "Cartesian channel 512x256x128, no-slip adiabatic walls on both j-faces (tanh-stretched "
                        "y), periodic x and z, M=0.3 parabolic profile + perturbations + pressure pulse, "
                        "viscous, selective filter, RK4, FP64, graph replay",
            "kernel_ms_per_step_eager": {k: round(ms / 2, 4) for k, (n, ms) in times.items()}}


def c2_leg(nx: int, ny: int, nz: int, steps: int, warmup: int, tune: bool, c3_value: float, mixed: bool):
    """The same block and flow run as 2.5-D curvilinear (the reference's
    is_curv: 2-D metrics in (i, j), Cartesian z), which a z-extruded wavy grid
    allows: FP64 and mixed, each with the stage kernel's roofline."""
    import numpy as np

    from musica2_newgrid_amd import _lib, cases
    case = cases.vortex_wavy_c2(nx, ny, nz, filter=2)
    res = {"workload": f"2.5-D curvilinear block {nx}x{ny}x{nz} (x = xi + ax sin(2pi eta/Ly), "
                       "y = eta + ay sin(2pi xi/Lx), extruded in z), isentropic vortex M=0.5, viscous, "
                       "shock filter + Ducros, RK4"}
    legs = [("fp64", _lib.PREC_FP64)] + ([("mixed", _lib.PREC_MIXED)] if mixed else [])
    rho = {}
    for name, prec in legs:
        r = run_leg(case, prec, steps, warmup, tune)
        nst, ms_stage = r["times"]["stage"]
        avg = ms_stage / nst / 1e3
        byts = stage_words_per_cell(case) * 8.0 * case.ncells
        val = case.ncells * case.nrk * steps / r["t"] / 1e6
        res[name] = {"value": round(val, 2), "unit": "Mcell-RK-stages/s",
                     "ms_per_step": round(r["t"] / steps * 1e3, 4),
                     "stage_tiles": {k: {"tile": f"{v[0]}x{v[1]}", "kchunk": v[2]} for k, v in r["tuned"].items()},
                     "stage_roofline": {"achieved_GBs": round(byts / avg / 1e9, 1),
                                        "frac": round(byts / avg / 1e9 / HBM_PEAK_GBS, 4),
                                        "avg_launch_ms": round(avg * 1e3, 4),
                                        "algorithmic_bytes_per_launch": byts},
                     "kernel_ms": {k: {"launches": n, "ms": round(ms, 3)} for k, (n, ms) in r["times"].items()}}
        rho[name] = r["rho"][5:-5, 5:-5, 5:-5]
    res["fp64"]["speedup_vs_c3_fp64"] = round(res["fp64"]["value"] / c3_value, 3)
    if mixed:
        a, b = rho["mixed"], rho["fp64"]
        res["mixed"]["speedup_vs_fp64"] = round(res["mixed"]["value"] / res["fp64"]["value"], 3)
        res["mixed"]["rho_linf_vs_fp64"] = float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
    return res


def periodic_leg(nx: int, ny: int, nz: int, steps: int, warmup: int, tune: bool, filt: int):
    """The same vortex on a fully periodic wavy block (no walls), FP64: the
    curvilinear workload the headline used before the walled channel ran."""
    from musica2_newgrid_amd import _lib, cases
    case = cases.vortex_wavy(nx, ny, nz, filter=filt)
    r = run_leg(case, _lib.PREC_FP64, steps, warmup, tune)
    nst, ms_stage = r["times"]["stage"]
    avg = ms_stage / nst / 1e3
    byts = stage_words_per_cell(case) * 8.0 * case.ncells
    traffic = STAGE_TRAFFIC.get(("periodic", nx, ny, nz))
    return {"value": round(case.ncells * case.nrk * steps / r["t"] / 1e6, 2), "unit": "Mcell-RK-stages/s",
            "ms_per_step": round(r["t"] / steps * 1e3, 4),
            "workload": f"curvilinear block {nx}x{ny}x{nz} with 3-D GCL metrics of a fully periodic wavy map, "
                        f"isentropic vortex M=0.5, viscous, shock filter + Ducros, RK4, FP64, graph replay",
            "stage_roofline": {"achieved_GBs": round(byts / avg / 1e9, 1),
                               "frac": round(byts / avg / 1e9 / HBM_PEAK_GBS, 4), "avg_launch_ms": round(avg * 1e3, 4),
                               "algorithmic_bytes_per_launch": byts, "traffic": traffic[0] if traffic else None}}


def shock_bl_leg(nx: int, ny: int, nz: int, steps: int, warmup: int, tune: bool, error_steps: int, mixed: bool):
    """The shock/boundary-layer workload (Cartesian, Mach 2 oblique shock of a
    10 degree wedge entering at jmax onto a Pohlhausen boundary layer,
    supersonic inlet at imin, characteristic outflow at imax and jmax, no-slip
    adiabatic wall at jmin, z periodic, seeded velocity noise): FP64 and mixed
    throughput (graph replay), the stage kernel's roofline, and the density
    of the mixed run against the FP64 run after `error_steps` steps."""
    import numpy as np

    from musica2_newgrid_amd import _lib, cases
    case = cases.shock_bl(nx, ny, nz)
    info = case.info
    res = {"workload": f"Cartesian block {nx}x{ny}x{nz}, Mach {info['mach']} inflow (supersonic inlet at imin), "
                       f"oblique shock of a {info['wedge_deg']} deg wedge (beta {info['beta_deg']:.2f} deg) entering "
                       f"at jmax onto a Pohlhausen boundary layer (Re_delta {info['re_delta']:g}), characteristic "
                       f"outflow at imax and jmax, no-slip adiabatic wall at jmin, z periodic, velocity noise "
                       f"1e-4 U (seed {info['seed']}), viscous, shock filter + Ducros, RK4, graph replay",
           "cells": case.ncells}
    legs = [("fp64", _lib.PREC_FP64)] + ([("mixed", _lib.PREC_MIXED)] if mixed else [])
    rho = {}
    for name, prec in legs:
        print(f"bench: shock/boundary-layer {name}, {steps} timed steps, {error_steps} in all", file=sys.stderr,
              flush=True)
        r = run_leg(case, prec, steps, warmup, tune, error_steps)
        nst, ms_stage = r["times"]["stage"]
        avg = ms_stage / nst / 1e3
        byts = stage_words_per_cell(case) * 8.0 * case.ncells
        val = case.ncells * case.nrk * steps / r["t"] / 1e6
        res[name] = {"value": round(val, 2), "unit": "Mcell-RK-stages/s",
                     "ms_per_step": round(r["t"] / steps * 1e3, 4),
                     "stage_tiles": {k: {"tile": f"{v[0]}x{v[1]}", "kchunk": v[2]} for k, v in r["tuned"].items()},
                     "stage_roofline": {"achieved_GBs": round(byts / avg / 1e9, 1),
                                        "frac": round(byts / avg / 1e9 / HBM_PEAK_GBS, 4),
                                        "avg_launch_ms": round(avg * 1e3, 4),
                                        "algorithmic_bytes_per_launch": byts},
                     "kernel_ms": {k: {"launches": n, "ms": round(ms, 3)} for k, (n, ms) in r["times"].items()},
                     "steps_done": r["steps_done"]}
        rho[name] = r["rho"][5:-5, 5:-5, 5:-5]
        if not np.all(np.isfinite(rho[name])):
            res[name]["error"] = "non-finite density"
    if mixed:
        a, b = rho["mixed"], rho["fp64"]
        res["mixed"]["speedup_vs_fp64"] = round(res["mixed"]["value"] / res["fp64"]["value"], 3)
        res["mixed"]["rho_linf_vs_fp64"] = float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
        res["mixed"]["rho_l2_vs_fp64"] = float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2)))
        res["mixed"]["after_steps"] = res["mixed"]["steps_done"]
    return res


def serve_leg(nreq: int, nsteps: int, seed: int = 5):
    """Serving loop: a Poisson stream of TGV / isentropic-vortex requests on
    cubes of edge 64..256, Re log-uniform in [400, 4000], FP64 or mixed,
    admitted by an HBM budget and co-batched."""
    import numpy as np

    from musica2_newgrid_amd import serving
    edges = (64, 96, 128, 192, 256)
    mean_cells = sum(e ** 3 for e in edges) / len(edges)
    budget = 64 << 30
    # arrival rate for about 80 % of the device: the service rate measured on
    # a saturated mix of the same request kinds
    svc = serving.service_rate(edges, budget_bytes=budget)
    rate = 0.8 / (mean_cells * 4 * nsteps / svc)
    reqs = serving.make_requests(nreq, seed=seed, edges=edges, nsteps=nsteps, rate=rate)
    stats = {}
    lat, wall, work = serving.run_stream(reqs, budget_bytes=budget, quantum=10, stats=stats, progress=True)
    p50, p99 = np.percentile(lat, [50, 99])
    return {"requests": nreq, "steps_per_request": nsteps, "arrival_rate_per_s": round(rate, 3),
            "measured_service_rate": round(svc / 1e6, 2), "hbm_budget_GiB": budget >> 30, "wall_s": round(wall, 3),
            "throughput": round(work / wall / 1e6, 2), "unit": "Mcell-RK-stages/s",
            "p50_ms": round(p50 * 1e3, 2), "p99_ms": round(p99 * 1e3, 2), "seed": seed,
            "co_batching": {k: (round(v, 3) if isinstance(v, float) else v) for k, v in stats.items()}}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision leg")
    ap.add_argument("--no-tune", action="store_true", help="skip the tile autotuner")
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--error-steps", type=int, default=1000,
                    help="steps after which the mixed leg's density error against the FP64 leg is measured")
    ap.add_argument("--nx", type=int, default=512)
    ap.add_argument("--ny", type=int, default=256)
    ap.add_argument("--nz", type=int, default=128)
    ap.add_argument("--no-multiblock", action="store_true", help="skip the uneven 7-block leg")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving-loop leg")
    ap.add_argument("--no-tgv", action="store_true", help="skip the TGV 64^3 leg")
    ap.add_argument("--no-channel", action="store_true", help="skip the walled-channel leg")
    ap.add_argument("--no-c2", action="store_true", help="skip the 2.5-D curvilinear leg")
    ap.add_argument("--no-shock-bl", action="store_true", help="skip the shock/boundary-layer leg")
    ap.add_argument("--no-periodic", action="store_true", help="skip the fully periodic curvilinear leg")
    ap.add_argument("--headline", choices=["channel", "channel_noslip", "periodic"], default="channel",
                    help="the curvilinear headline block: the channel with slip walls as stated, the same "
                         "channel with no-slip walls (is_wall2 = F), or the fully periodic wavy block")
    ap.add_argument("--shock-bl-size", type=int, nargs=3, default=(1024, 512, 8), metavar=("NX", "NY", "NZ"))
    ap.add_argument("--shock-bl-error-steps", type=int, default=2000,
                    help="steps after which the shock/boundary-layer mixed density is compared with FP64")
    ap.add_argument("--dissipation", choices=["shock", "artvisc_shock"], default="shock",
                    help="dissipation of the main leg: the shock filter (filter_o10_shock) or the artificial "
                         "viscosity with shock capturing (artvisc_o9_shock)")
    ap.add_argument("--serve-requests", type=int, default=16)
    ap.add_argument("--serve-steps", type=int, default=100)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain device copy of known size (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration()
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py runs on one GPU only; --gpus must be 1"}))
        return 2

    import numpy as np

    from musica2_newgrid_amd import _lib, cases

    diss = {"shock": (2, "shock filter + Ducros"),
            "artvisc_shock": (4, "artificial viscosity with shock capturing + Ducros")}[args.dissipation]
    if args.headline == "channel":
        case = cases.vortex_channel(args.nx, args.ny, args.nz, filter=diss[0], slip=True)
        inf = case.info
        workload = (f"curvilinear channel {args.nx}x{args.ny}x{args.nz}: x uniform and periodic, slip walls "
                    f"(type_wall_bc 2, flag s, adiabatic) on both y-faces with tanh clustering (largest "
                    f"neighbouring-spacing ratio {inf['ratio']}), wavy lower wall y' = y + {inf['amp']} "
                    f"sin(2 pi x/Lx)(1 - y/Ly), z uniform and periodic, 3-D GCL metrics; isentropic vortex "
                    f"(strength 5, radius 0.5) at Mach 0.5 in a uniform free stream; viscous, {diss[1]} "
                    f"(quiescent), RK4, FP64, steps replayed from the captured HIP graph")
    elif args.headline == "channel_noslip":
        case = cases.vortex_channel(args.nx, args.ny, args.nz, filter=diss[0])
        inf = case.info
        workload = (f"curvilinear channel {args.nx}x{args.ny}x{args.nz}: x uniform and periodic, no-slip adiabatic "
                    f"walls on both y-faces with tanh clustering (largest neighbouring-spacing ratio "
                    f"{inf['ratio']}), wavy lower wall y' = y + {inf['amp']} sin(2 pi x/Lx)(1 - y/Ly), z uniform and "
                    f"periodic, 3-D GCL metrics; isentropic vortex (strength 5, radius 0.5) at Mach 0.5 in a "
                    f"free stream meeting the walls through a tanh layer of thickness {inf['delta']}; viscous, "
                    f"{diss[1]} (quiescent), RK4, FP64, steps replayed from the captured HIP graph")
    else:
        case = cases.vortex_wavy(args.nx, args.ny, args.nz, filter=diss[0])
        workload = (f"curvilinear block {args.nx}x{args.ny}x{args.nz} with 3-D GCL metrics of a fully "
                    f"periodic wavy map (x, y and z each displaced by a product of sines; periodic "
                    f"faces, no walls, no tanh stretching), isentropic vortex M=0.5, viscous, "
                    f"{diss[1]}, RK4, FP64, steps replayed from the captured HIP graph")
    total = 0 if args.no_mixed else args.error_steps
    print(f"bench: {args.nx}x{args.ny}x{args.nz} FP64 leg, {args.steps} steps", file=sys.stderr, flush=True)
    fp = run_leg(case, _lib.PREC_FP64, args.steps, args.warmup, not args.no_tune, total)
    t = fp["t"]
    nst, ms_stage = fp["times"]["stage"]

    value = case.ncells * case.nrk * args.steps / t / 1e6
    avg_stage_s = ms_stage / nst / 1e3
    bytes_launch = stage_words_per_cell(case) * 8.0 * case.ncells
    achieved = bytes_launch / avg_stage_s / 1e9
    traffic = STAGE_TRAFFIC.get((args.headline, args.nx, args.ny, args.nz))
    out = {
        "metric": "Mcell-RK-stages/s",
        "value": round(value, 2),
        "unit": "Mcell-RK-stages/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": workload,
                   "cells": case.ncells, "nrk": case.nrk, "precision": "fp64",
                   "stage_tiles": {k: {"tile": f"{v[0]}x{v[1]}", "kchunk": v[2]} for k, v in fp["tuned"].items()}},
        "timing": {"value": "graph replay of the timed steps",
                   "roofline": f"per-launch HIP events on the library's stream over {args.steps} more steps "
                               f"launched one by one ({fp['t_eager'] / args.steps * 1e3:.4f} ms per step)"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic[0] if traffic else None,
                     "traffic_source": ("rocprofv3 --pmc FETCH_SIZE/WRITE_SIZE, " + traffic[1]) if traffic else None,
                     "kernel": "k_stage", "avg_launch_ms": round(avg_stage_s * 1e3, 4),
                     "algorithmic_bytes_per_launch": bytes_launch},
        "kernel_ms": {k: {"launches": n, "ms": round(ms, 3)} for k, (n, ms) in fp["times"].items()},
    }
    if not args.no_mixed:
        print("bench: mixed-precision leg", file=sys.stderr, flush=True)
        mx = run_leg(case, _lib.PREC_MIXED, args.steps, args.warmup, not args.no_tune, total)
        assert mx["steps_done"] == fp["steps_done"]
        a = mx["rho"][5:-5, 5:-5, 5:-5]
        b = fp["rho"][5:-5, 5:-5, 5:-5]
        mval = case.ncells * case.nrk * args.steps / mx["t"] / 1e6
        out["mixed"] = {
            "value": round(mval, 2), "unit": "Mcell-RK-stages/s",
            "ms_per_step": round(mx["t"] / args.steps * 1e3, 4),
            "speedup_vs_fp64": round(mval / value, 3),
            "rho_linf_vs_fp64": float(np.max(np.abs(a - b)) / np.max(np.abs(b))),
            "rho_l2_vs_fp64": float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2))),
            "after_steps": mx["steps_done"],
            "kernel_ms": {k: {"launches": n, "ms": round(ms, 3)} for k, (n, ms) in mx["times"].items()},
        }
    if args.headline != "periodic" and not args.no_periodic:
        print("bench: periodic curvilinear leg", file=sys.stderr, flush=True)
        out["periodic_c3"] = periodic_leg(args.nx, args.ny, args.nz, args.steps, args.warmup, not args.no_tune, diss[0])
    if not args.no_c2:
        print("bench: 2.5-D curvilinear leg", file=sys.stderr, flush=True)
        out["c2"] = c2_leg(args.nx, args.ny, args.nz, args.steps, args.warmup, not args.no_tune, value,
                           not args.no_mixed)
    if not args.no_shock_bl:
        print("bench: shock/boundary-layer leg", file=sys.stderr, flush=True)
        out["shock_bl"] = shock_bl_leg(*args.shock_bl_size, args.steps, args.warmup, not args.no_tune,
                                       args.shock_bl_error_steps, not args.no_mixed)
    if not args.no_tgv:
        print("bench: TGV 64^3 leg", file=sys.stderr, flush=True)
        out["tgv64"] = tgv_leg()
    if not args.no_channel:
        print("bench: walled-channel leg", file=sys.stderr, flush=True)
        out["channel"] = channel_leg(args.steps, args.warmup)
    if not args.no_multiblock:
        print("bench: uneven 7-block leg", file=sys.stderr, flush=True)
        out["multiblock"] = multiblock_leg(args.steps, args.warmup)
    if not args.no_serve:
        print("bench: serving leg", file=sys.stderr, flush=True)
        out["serving"] = serve_leg(args.serve_requests, args.serve_steps)
    if not args.no_cpu:
        print("bench: CPU baseline leg", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(case, args.cpu_steps)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
