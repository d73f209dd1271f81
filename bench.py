#!/usr/bin/env python3
"""bench.py — particle-substeps/s of the MI355X HIP XPBD substep (plain form, one GPU).

Workload (default): gravity pour of 1,048,576 polydisperse spheres, radius ~ U[0.8, 1.2] mm,
rho = 2500 kg/m^3, 160 x 160 mm periodic in x and y with a floor, seeded random non-overlapping
start (random sequential addition in a 700 mm column), dt = 5e-6 s, mu = 0.5, e = 0.3, solver
iterations (pos 12, vel 8), FP64 state.  A "step" is one substep (one demStep); W untimed
substeps, then K timed substeps bracketed by device synchronisation.

The JSON line carries:
  roofline     — the dominant kernel (the one with the largest summed time: the PGS velocity
                 sweep k_vsweep or the position sweep k_psweep), achieved = algorithmic bytes it
                 moved / its summed launch time from HIP events on its own stream, against the
                 8 TB/s HBM peak.  Algorithmic bytes: 601 B per active manifold per velocity
                 sweep, 388 B per contact per position sweep.
  cpu_baseline — the CPU oracle (serial C restatement, 1 thread) timed on the same seeded input
                 for a few substeps, rank 0 only.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "peclet-dem_amd"))

HBM_PEAK_GBS = 8000.0
BYTES = {"vsweep": 601, "psweep": 388}
METRIC = "particle-substeps/s at 1M spheres: plain vs batched vs fwd+adjoint, %HBM roof"


def substep_bytes(st):
    """Algorithmic bytes of one substep from its counters (FP64 SoA, spheres)."""
    N = st["num_real"]
    G = st["num_particles"] - N
    P, C, M = st["num_pairs"], st["num_contacts"], st["num_manifolds"]
    b = 684 * N + 460 * G + 152 * P + 236 * C + 1661 * M
    b += 88 * st["uncolored_vel"] + 8 * N * st["rounds_vel"]
    b += 80 * st["uncolored_pos"] + 8 * (N + G) * st["rounds_pos"]
    b += 601 * (st["sweeps_vel"] + st["sweeps_stab"]) * st["active_manifolds"]
    b += 388 * st["sweeps_pos"] * C
    return b


def make_workload(kind, n, seed, column):
    from peclet.dem.workloads import lattice_bed, sphere_bed

    if kind == "lattice":
        return lattice_bed(n, 0.16, seed)
    return sphere_bed(n, 0.8e-3, 1.2e-3, (0.16, 0.16, 0.4), seed, column=column)


def run_cpu_baseline(bed, dt, substeps):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_binding import oracle_sim
    from peclet.dem.workloads import setup_gravity_bed

    sim = setup_gravity_bed(oracle_sim(len(bed["pos"])), bed, dt, 0.5, 0.3)
    t0 = time.perf_counter()
    for _ in range(substeps):
        sim.step(dt)
    t = time.perf_counter() - t0
    return len(bed["pos"]) * substeps / t, t


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=32, help="timed substeps")
    ap.add_argument("--warmup", type=int, default=8, help="untimed substeps")
    ap.add_argument("--n", type=int, default=1048576)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--column", type=float, default=0.7, help="seeding column height (m)")
    ap.add_argument("--cpu-substeps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", choices=["pour", "lattice"], default="pour",
                    help="pour: the 1M-sphere gravity pour (default); lattice: a contact-rich "
                         "bed of spheres on a slightly overlapping cubic lattice")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); the "
              "multi-GPU spatial decomposition is a design, not an implementation", file=sys.stderr)
        sys.exit(2)

    import torch

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    from peclet.dem import Simulation
    from peclet.dem.workloads import setup_gravity_bed

    dt = 5e-6
    t_gen = time.perf_counter()
    bed = make_workload(args.workload, args.n, args.seed, args.column)
    t_gen = time.perf_counter() - t_gen
    sim = setup_gravity_bed(Simulation(args.n), bed, dt, 0.5, 0.3)
    print(f"# workload ready: n={args.n} (generation {t_gen:.1f}s)", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        sim.step(dt)
    sim._b.synchronize(sim._h)
    sim._b.set_timing(sim._h, 2)
    stats = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sim.step(dt)
        stats.append(sim.get_stats())
    sim._b.synchronize(sim._h)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    ms = np.zeros(2)
    launches = np.zeros(2, np.int64)
    rows = np.zeros(2, np.int64)
    import ctypes as C

    sim._b.get_kernel_times(sim._h, ms.ctypes.data_as(C.POINTER(C.c_double)),
                            launches.ctypes.data_as(C.POINTER(C.c_long)),
                            rows.ctypes.data_as(C.POINTER(C.c_long)), 2)
    value = args.n * args.steps / t
    names = ["k_vsweep", "k_psweep"]
    dom = int(np.argmax(ms))
    per_row = BYTES["vsweep"] if dom == 0 else BYTES["psweep"]
    bytes_per_launch = per_row * rows[dom] / max(launches[dom], 1)
    avg_launch_s = ms[dom] / 1e3 / max(launches[dom], 1)
    achieved = bytes_per_launch / avg_launch_s / 1e9 if avg_launch_s > 0 else 0.0
    total_bytes = sum(substep_bytes(s) for s in stats)
    stage = np.zeros(32)
    nst = sim._b.get_stage_times(sim._h, stage.ctypes.data_as(C.POINTER(C.c_double)), 32)
    stage_ms = {sim._b.stage_name(k).decode(): float(stage[k]) / args.steps for k in range(nst)}
    mean = {k: float(np.mean([s[k] for s in stats])) for k in stats[0]}

    cpu = None
    if not args.no_cpu:
        cps, tc = run_cpu_baseline(bed, dt, args.cpu_substeps)
        cpu = {"value": cps, "unit": "particle-substeps/s", "cores": 1, "kind": "port",
               "sample": f"oracle (serial C restatement) on the same seeded {args.n}-sphere input, "
                         f"first {args.cpu_substeps} substeps, {tc:.1f} s"}

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"plain gravity pour, {args.n} polydisperse spheres r~U[0.8,1.2] mm, "
                         "160x160 mm periodic xy + floor, RSA start in a 700 mm column, "
                         "dt=5e-6 s, mu=0.5, e=0.3, iters pos 12 / vel 8"
                         if args.workload == "pour" else
                         f"lattice bed, {args.n} spheres r~U[0.99,1.01] mm on a 1.996 r cubic "
                         "lattice, 160x160 mm periodic xy + floor, dt=5e-6 s, mu=0.5, e=0.3, "
                         "iters pos 12 / vel 8"),
            "form": "plain",
            "substeps_timed": args.steps,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": names[dom],
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "bytes_per_launch": bytes_per_launch,
            "avg_launch_us": avg_launch_s * 1e6,
            "launches": int(launches[dom]),
        },
        "cpu_baseline": cpu,
        "substep_roofline": {
            "algorithmic_bytes_per_substep": total_bytes / args.steps,
            "achieved_GBs": total_bytes / t / 1e9,
            "frac": total_bytes / t / 1e9 / HBM_PEAK_GBS,
        },
        "kernel_ms": {names[k]: float(ms[k]) for k in range(2)},
        "stage_ms_per_substep": stage_ms,
        "counters_mean": mean,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
