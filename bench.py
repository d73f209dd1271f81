#!/usr/bin/env python3
"""bench.py — particle-substeps/s of the MI355X HIP XPBD substep (one GPU), plain, batched and
forward+adjoint forms.

A "step" is 8 substeps (8 demStep calls, as the reference's drivers loop them); ``--steps K``
times K steps = 8K substeps after ``--warmup W`` untimed steps.

Headline (``value``; ``--workload settled``, the default): the 1M-sphere bed at rest, where every
substep does full contact work: 1,048,576 spheres r ~ U[0.99, 1.01] mm on a 1.996 r cubic lattice
resting on the floor (neighbours overlap slightly: ~13.8M candidate pairs, ~3.2M contacts,
~2.1M active velocity rows, 8-10 colours), 160 x 160 mm periodic in x and y, dt = 5e-6 s,
mu = 0.5, e = 0.3, iterations (pos 12, vel 8), FP64.  ``--workload pour`` makes the headline the
gravity pour below; ``--workload cylinders`` the hollow-cylinder bed.

Further legs in the same JSON line (each can be switched off):
  pour          the 1M-sphere gravity pour as a whole job: 1,048,576 spheres r ~ U[0.8, 1.2] mm,
                rho 2500 kg/m^3, random non-overlapping start seeded across the periodic x/y
                seams (a 700 mm column: the 400 mm column of the specification is above the
                random-sequential-addition jamming density), floor, dt = 5e-6 s, mu 0.5, e 0.3,
                500 steps x 8 substeps timed (``--pour-steps``), whatever ``--steps`` is.
  cylinders     200,000 hollow cylinders (r_out 2 mm, r_in 1.2 mm, L 4 mm, 256-point seeded
                shell, random orientations), 200 x 200 mm periodic xy + floor, mu 0.4, e 0.2,
                dt 5e-6 s; ``--cyl-steps`` steps timed.
  hertz         the pour's spheres under the Hertz-Mindlin engine (step_hertz; E = 1e7 Pa,
                nu = 0.3), in a non-periodic box with an SDF floor (the Hertz path takes neither
                periodic axes nor analytic planes); ``--hertz-steps`` steps of 8 substeps.
  batched       256 independent periodic-xy sphere beds stepped as one Batch, against the same
                systems stepped one after another.
  fwd_adjoint   50,000 spheres: a recorded forward run plus the adjoint of the loss, against
                the plain forward run of the same substeps.
  cpu_baseline  the CPU oracle (OpenMP over the host cores, thread count stated) on the same
                seeded headline input for a bounded number of substeps.

Roofline (``roofline``): the kernel with the largest summed time over the headline's timed
substeps, timed with HIP events on the simulation's stream around each of its launches (every
``--timing-every``-th substep); achieved = its algorithmic bytes per launch / its average launch
time, against the 8 TB/s HBM peak.  Algorithmic bytes per unit (reads + writes once, FP64, int32
indices):
    k_vsweep, k_vsolve_lds, k_vsolve_block  601 per active velocity row per executed sweep
    k_psweep, k_psolve_block                388 per contact row per executed sweep
    k_predict                               240 per real particle
    k_reintegrate                           264 per slot
    k_bp_query                              44 per slot + 8 per candidate pair written
    k_np_pairs                              144 per pair processed + 104 per contact written
A launch made after its sweep loop's device-side stop flag was set returns at once: it counts
its time and no bytes.  ``traffic`` = measured HBM bytes per launch of that kernel (rocprofv3
--pmc FETCH_SIZE and WRITE_SIZE, scaled by the same counters on a plain copy of known size), read
from profiles/traffic_<workload>.json when present, else null.
``substep_roofline``: whole-substep algorithmic bytes from the per-substep counters.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "peclet-dem_amd"))

HBM_PEAK_GBS = 8000.0
METRIC = "particle-substeps/s at 1M spheres: plain vs batched vs fwd+adjoint, %HBM roof"


SUBSTEPS_PER_STEP = 8


def substep_bytes(st):
    """Algorithmic bytes of one substep from its counters (FP64 SoA, spheres)."""
    N = st["num_real"]
    G = st["num_particles"] - N
    P, Cn, M = st["num_pairs"], st["num_contacts"], st["num_manifolds"]
    b = 684 * N + 460 * G + 152 * P + 236 * Cn + 1661 * M
    # the colourings as the incidence-list rounds move them: 256 B per coloured edge and 40 B per
    # body, whichever round form ran (the contend / commit rounds move more)
    if st["active_manifolds"] > 0:
        b += 256 * st["active_manifolds"] + 40 * N
    if Cn > 0:
        b += 256 * Cn + 40 * (N + G)
    b += 601 * (st["sweeps_vel"] + st["sweeps_stab"]) * st["active_manifolds"]
    b += 388 * st["sweeps_pos"] * Cn
    return b


def kernel_bytes(name, units, log):
    """Algorithmic bytes of all timed launches of one kernel: units = the library's summed units
    for it (rows x executed sweeps, slots or pairs), log = the per-substep counters of the timed
    substeps."""
    if name in ("k_vsweep", "k_vsolve_lds"):  # active velocity rows x executed sweeps
        return 601 * units
    if name in ("k_psweep", "k_psolve_block"):  # contact rows x executed sweeps
        return 388 * units
    if name == "k_bp_query":  # slots queried, plus the candidate pairs they write
        return 44 * units + sum(8 * s["num_pairs"] for s in log)
    if name == "k_np_pairs":  # pairs processed, plus the contacts written
        return 144 * units + 104 * sum(s["num_contacts"] for s in log)
    if name == "k_predict":
        return 240 * units
    if name == "k_reintegrate":
        return 264 * units
    raise KeyError(name)


def make_workload(kind, n, seed, column):
    from peclet.dem.workloads import cylinder_bed, cylinder_pack, lattice_bed, sphere_bed

    if kind == "settled":
        return lattice_bed(n, 0.16, seed)
    if kind == "cylinders":
        return cylinder_bed(n, 0.2, seed)
    if kind == "cylinders_packed":
        return cylinder_pack(n, 0.2, seed)
    return sphere_bed(n, 0.8e-3, 1.2e-3, (0.16, 0.16, 0.4), seed, column=column)


def setup(sim, kind, bed, dt):
    from peclet.dem.workloads import setup_cylinder_bed, setup_gravity_bed

    if kind in ("cylinders", "cylinders_packed"):
        return setup_cylinder_bed(sim, bed, dt, 0.4, 0.2)
    return setup_gravity_bed(sim, bed, dt, 0.5, 0.3)


WORKLOAD_TEXT = {
    "settled": ("resting 1M-sphere bed: {n} spheres r~U[0.99,1.01] mm on a 1.996 r cubic lattice "
                "on the floor (contacts everywhere), 160x160 mm periodic xy, dt=5e-6 s, mu=0.5, "
                "e=0.3, iters pos 12 / vel 8"),
    "pour": ("gravity pour, {n} polydisperse spheres r~U[0.8,1.2] mm, 160x160 mm periodic xy + "
             "floor, random start seeded across the periodic seams in a {col} mm column, "
             "dt=5e-6 s, mu=0.5, e=0.3, iters pos 12 / vel 8"),
    "cylinders_packed": ("{n} hollow cylinders r_out 2 mm, r_in 1.2 mm, L 4 mm, 256-point seeded "
                         "shell, packed at rest on a 4.1 mm jittered lattice (the contact-rich state "
                         "of a settled bed), random orientations, 200x200 mm periodic xy + floor, "
                         "dt=5e-6 s, mu=0.4, e=0.2, iters pos 12 / vel 8"),
    "cylinders": ("{n} hollow cylinders r_out 2 mm, r_in 1.2 mm, L 4 mm, 256-point seeded shell, "
                  "random orientations, 200x200 mm periodic xy + floor, random start, dt=5e-6 s, "
                  "mu=0.4, e=0.2, iters pos 12 / vel 8"),
}


def oracle_threads():
    """Host threads for the CPU baseline: the box's CPU share (OMP_NUM_THREADS when set)."""
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    return max(1, min(16, os.cpu_count() or 1))


def run_cpu_baseline(kind, bed, dt, substeps, threads):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_binding import oracle_binding, oracle_sim

    oracle_binding().lib.pdref_set_threads(threads)
    sim = setup(oracle_sim(len(bed["pos"])), kind, bed, dt)
    t0 = time.perf_counter()
    for _ in range(substeps):
        sim.step(dt)
    t = time.perf_counter() - t0
    return len(bed["pos"]) * substeps / t, t


def kernel_times(sim):
    """Per-kernel HIP-event totals of the library: names, ms, launches, units."""
    b, h = sim._b, sim._h
    count = b.get_kernel_times(h, None, None, None, 0)
    names = [b.kernel_name(k).decode() for k in range(count)]
    ms = np.zeros(count)
    launches = np.zeros(count, np.int64)
    units = np.zeros(count, np.int64)
    b.get_kernel_times(h, ms.ctypes.data_as(C.POINTER(C.c_double)),
                       launches.ctypes.data_as(C.POINTER(C.c_long)),
                       units.ctypes.data_as(C.POINTER(C.c_long)), count)
    return names, ms, launches, units


def stats_log(sim):
    """The per-substep counters the library logged while timing at level 2."""
    from peclet.dem._binding import Stats

    n = sim._b.get_stats_log(sim._h, None, 0)
    arr = (Stats * max(n, 1))()
    sim._b.get_stats_log(sim._h, arr, n)
    return [arr[k].as_dict() for k in range(n)]


def load_traffic(workload, kernel):
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    k = d.get("kernels", {}).get(kernel)
    return None if k is None else float(k["hbm_bytes_per_launch"])


def copy_probe(mib):
    """A plain device-to-device copy of known size (hipMemcpy, run as the runtime's blit kernel),
    for scaling the HBM byte counters. Uses the HIP runtime the library itself links."""
    hip = C.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so"))
    nbytes = C.c_size_t(mib << 20)
    a, b = C.c_void_p(), C.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(C.byref(p), nbytes) != 0:
            raise RuntimeError("copy probe: hipMalloc failed")
    hip.hipMemset(a, 1, nbytes)
    hip.hipDeviceSynchronize()
    for _ in range(3):
        if hip.hipMemcpy(b, a, nbytes, 3) != 0:  # hipMemcpyDeviceToDevice
            raise RuntimeError("copy probe: hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print(f"# copy probe: 3 copies of {mib} MiB", file=sys.stderr, flush=True)


BATCH_ROUNDS = 3  # alternating sequential / batched timing rounds of the batched leg


def run_batched(num, seed, steps, seq_steps, workers, autotune=True):
    """Config-4 style ensemble: aggregate particle-substeps/s of one batch, against the same
    systems stepped one after another (a single worker)."""
    from peclet.dem import Batch
    from peclet.dem.workloads import ensemble, setup_gravity_bed

    beds = ensemble(num, seed)
    dt = 1e-5
    batch = Batch([len(b["pos"]) for b in beds], workers=workers)
    for sim, bed in zip(batch.systems, beds):
        setup_gravity_bed(sim, bed, dt, bed["mu"], bed["e"])
    total = sum(len(b["pos"]) for b in beds)
    batch.step(dt, 2)  # warm-up: buffers reach their working size
    if autotune:
        batch.autotune(dt, 2)  # the largest system of each size class is tuned
    # the one-worker (sequential) and the batched timings alternate over BATCH_ROUNDS rounds, so
    # a drift of the box's clocks or host load over the leg reaches both alike
    t_seq = t_b = 0.0
    per_round = max(1, steps // BATCH_ROUNDS)
    steps = per_round * BATCH_ROUNDS
    seq_per_round = max(1, seq_steps // BATCH_ROUNDS)
    seq_steps = seq_per_round * BATCH_ROUNDS
    for _ in range(BATCH_ROUNDS):
        batch.set_workers(1)
        t0 = time.perf_counter()
        batch.step(dt, seq_per_round)
        t_seq += time.perf_counter() - t0
        batch.set_workers(workers)
        t0 = time.perf_counter()
        batch.step(dt, per_round)
        t_b += time.perf_counter() - t0
    # algorithmic bytes of one batch substep from each system's last substep counters
    bps = sum(substep_bytes(sim.get_stats()) for sim in batch.systems)
    batch.close()
    v_b = total * steps / t_b
    v_s = total * seq_steps / t_seq
    return {"value": v_b, "unit": "particle-substeps/s", "systems": num, "particles": total,
            "substeps_timed": steps, "workers": workers, "ms_per_substep": 1e3 * t_b / steps,
            "sequential_value": v_s, "sequential_substeps_timed": seq_steps,
            "speedup_vs_sequential": v_b / v_s, "autotuned": autotune,
            "algorithmic_bytes_per_substep": bps,
            "hbm_frac": bps * steps / t_b / 1e9 / HBM_PEAK_GBS,
            "workload": f"{num} independent periodic-xy sphere beds, N~U[500,20000], "
                        "r~U[0.8,1.2] mm, mu~U[0.2,0.8], e~U[0.1,0.6], phi0=0.1, dt=1e-5 s"}


ADJOINT_REPEATS = 3  # timed runs per form; the line reports their medians


def run_adjoint(n, substeps, every, warm, dense):
    """Forward + adjoint form: particle-substeps/s of a recorded forward run plus its adjoint
    (gradient of the mean bed height w.r.t. initial velocities, friction and restitution),
    against the plain forward run of the same substeps from the same state (medians of
    ADJOINT_REPEATS timed runs, after one untimed recording and adjoint)."""
    from peclet.dem import Simulation
    from peclet.dem.workloads import lattice_bed, setup_gravity_bed, sphere_bed

    dt = 1e-5
    if dense:
        bed = lattice_bed(n, 0.06, 5)
        what = (f"{n} spheres r~U[0.99,1.01] mm on a cubic lattice resting on the floor "
                "(contacts everywhere), 60x60 mm periodic xy")
    else:
        bed = sphere_bed(n, 0.9e-3, 1.1e-3, (0.06, 0.06, 0.2), 5, column=0.2)
        what = (f"{n} spheres r~U[0.9,1.1] mm, RSA start in 60x60x200 mm periodic xy + floor, "
                "v~N(0,0.01 m/s)")

    def fresh():
        sim = setup_gravity_bed(Simulation(n), bed, dt, 0.5, 0.3)
        if not dense:
            sim.set_velocities(np.random.default_rng(5).normal(0.0, 0.01, (n, 3)))
        sim.step_n(dt, warm)
        sim.synchronize()
        return sim

    grad = np.zeros((n, 13))
    grad[:, 2] = 1.0 / n

    def recorded(sim):
        t0 = time.perf_counter()
        sim.tape_begin(every)
        sim.step_n(dt, substeps)
        sim.tape_end()
        sim.synchronize()
        t1 = time.perf_counter()
        g = sim.tape_adjoint(grad)
        return t1 - t0, time.perf_counter() - t1, g

    # one untimed recording and adjoint first (first launches of the adjoint kernels, workspace
    # allocation), then the median of `REPEATS` timed runs of each form, each from a fresh start
    sim = fresh()
    recorded(sim)
    sim.close()
    plain, fwd, adj = [], [], []
    for _ in range(ADJOINT_REPEATS):
        sim = fresh()
        t0 = time.perf_counter()
        sim.step_n(dt, substeps)
        sim.synchronize()
        plain.append(time.perf_counter() - t0)
        sim.close()
        sim = fresh()
        tf, ta, (g0, gp) = recorded(sim)
        fwd.append(tf)
        adj.append(ta)
        contacts = sim.num_contacts()
        fwd_bytes = substep_bytes(sim.get_stats()) * substeps
        sim.close()
    t_plain, t_fwd, t_adj = (float(np.median(x)) for x in (plain, fwd, adj))
    return {"value": n * substeps / (t_fwd + t_adj), "unit": "particle-substeps/s",
            "workload": what + f", dt=1e-5 s, mu=0.5, e=0.3, {warm} substeps before the "
                               f"recording, {substeps} recorded, checkpoint every {every}",
            "repeats": ADJOINT_REPEATS, "contacts_at_end": contacts, "forward_ms": 1e3 * t_plain,
            "recorded_forward_ms": 1e3 * t_fwd, "adjoint_ms": 1e3 * t_adj,
            "cost_vs_forward": (t_fwd + t_adj) / t_plain,
            # the forward substeps' algorithmic bytes (last substep's counters) over the whole
            # forward+adjoint time; the adjoint's own traffic is not modelled
            "hbm_frac_forward_bytes": fwd_bytes / (t_fwd + t_adj) / 1e9 / HBM_PEAK_GBS,
            "grad_norm_v0": float(np.linalg.norm(g0[:, 3:6])),
            "grad_mu": float(gp[0]), "grad_restitution": float(gp[1])}


def timed_run(sim, dt, substeps, timing_every, level, label):
    """Time `substeps` substeps of `sim` (device-synchronised on both sides) with the library's
    per-kernel events on every `timing_every`-th substep; returns throughput, the dominant
    kernel's roofline and the mean counters."""
    import torch

    n = sim.num_particles()
    sim._b.synchronize(sim._h)
    sim.set_tuning("timing_every", timing_every)
    sim._b.set_timing(sim._h, level)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    report = max(1, substeps // 8)
    for k in range(substeps):
        sim.step(dt)
        if (k + 1) % report == 0:
            print(f"# {label}: substep {k + 1}/{substeps}", file=sys.stderr, flush=True)
    sim._b.synchronize(sim._h)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    out = {"value": n * substeps / t, "unit": "particle-substeps/s", "substeps_timed": substeps,
           "ms_per_substep": 1e3 * t / substeps,
           "ms_per_step": 1e3 * t / substeps * SUBSTEPS_PER_STEP}
    if level < 2:
        sim._b.set_timing(sim._h, 0)
        return out
    names, ms, launches, units = kernel_times(sim)
    stats = stats_log(sim)
    assert len(stats) == substeps, (len(stats), substeps)
    sampled = stats[::timing_every]  # the substeps whose launches were timed
    kbytes = np.array([kernel_bytes(nm, int(units[q]), sampled) for q, nm in enumerate(names)],
                      dtype=float)
    dom = int(np.argmax(ms))
    nl = max(int(launches[dom]), 1)
    bytes_per_launch = kbytes[dom] / nl
    avg_launch_s = ms[dom] / 1e3 / nl
    achieved = bytes_per_launch / avg_launch_s / 1e9 if avg_launch_s > 0 else 0.0
    total_bytes = sum(substep_bytes(st) for st in stats)
    stage = np.zeros(32)
    nst = sim._b.get_stage_times(sim._h, stage.ctypes.data_as(C.POINTER(C.c_double)), 32)
    hstage = np.zeros(32)
    sim._b.get_stage_host_times(sim._h, hstage.ctypes.data_as(C.POINTER(C.c_double)), 32)
    out["roofline"] = {
        "bound": "hbm",
        "kernel": names[dom],
        "timed_every_nth_substep": timing_every,
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS,
        "traffic": None,
        "bytes_per_launch": bytes_per_launch,
        "avg_launch_us": avg_launch_s * 1e6,
        "launches": int(launches[dom]),
    }
    out["substep_roofline"] = {
        "algorithmic_bytes_per_substep": total_bytes / substeps,
        "achieved_GBs": total_bytes / t / 1e9,
        "frac": total_bytes / t / 1e9 / HBM_PEAK_GBS,
    }
    out["kernels"] = {names[q]: {"ms": float(ms[q]), "launches": int(launches[q]),
                                 "avg_us": float(ms[q]) * 1e3 / max(int(launches[q]), 1),
                                 "GBs": (kbytes[q] / (ms[q] / 1e3) / 1e9) if ms[q] > 0 else 0.0}
                      for q in range(len(names))}
    out["stage_ms_per_substep"] = {sim._b.stage_name(k).decode(): float(stage[k]) / len(sampled)
                                   for k in range(nst)}
    out["stage_host_ms_per_substep"] = {
        sim._b.stage_name(k).decode(): float(hstage[k]) / len(sampled) for k in range(nst)}
    out["counters_mean"] = {k: float(np.mean([st[k] for st in stats])) for k in stats[0]}
    sim._b.set_timing(sim._h, 0)
    return out


def plain_leg(kind, n, seed, column, dt, warmup, substeps, timing_every, level, tuning=(),
              autotune=False):
    """Build one plain workload, run `warmup` untimed substeps (and, with `autotune`, the
    autotuner's trial substeps: its knobs change no result) and time `substeps`."""
    from peclet.dem import Simulation

    t_gen = time.perf_counter()
    bed = make_workload(kind, n, seed, column)
    t_gen = time.perf_counter() - t_gen
    sim = setup(Simulation(n), kind, bed, dt)
    for name, value in tuning:
        sim.set_tuning(name, value)
    print(f"# {kind}: n={n} ready (generation {t_gen:.1f}s)", file=sys.stderr, flush=True)
    for _ in range(warmup):
        sim.step(dt)
    chosen = None
    if autotune:
        chosen, best = sim.autotune(dt, substeps=2, use_cache=False)
        print(f"# {kind}: autotuned {chosen} ({best:.3f} ms per substep in its trials)",
              file=sys.stderr, flush=True)
    res = timed_run(sim, dt, substeps, timing_every, level, kind)
    res["tuning"] = chosen
    res["workload"] = WORKLOAD_TEXT[kind].format(n=n, col=f"{column * 1e3:.0f}")
    if "roofline" in res:
        res["roofline"]["traffic"] = load_traffic(kind, res["roofline"]["kernel"])
    return sim, bed, res


def hertz_leg(n, seed, steps):
    """The pour's spheres under the Hertz-Mindlin engine in a non-periodic box with an SDF floor:
    particle-substeps/s over `steps` x 8 substeps of step_hertz."""
    import torch

    from peclet.dem import Simulation, build_wall_sdf
    from peclet.dem.workloads import sphere_bed

    dt = 5e-6
    bed = sphere_bed(n, 0.8e-3, 1.2e-3, (0.16, 0.16, 0.4), seed, column=0.7,
                     periodic=(False, False, False))
    sim = Simulation(n)
    sim.set_sphere_shape(bed["base_radius"])
    sim.set_domain(0.16, 0.16, 0.7, False, False, False)
    floor = build_wall_sdf(lambda p: p[:, 2], ((-0.01, -0.01, -0.01), (0.17, 0.17, 0.71)),
                           resolution=(4, 4, 4))
    floor.add_to(sim, restitution=0.3, friction=0.5)
    sim.set_positions(bed["pos"])
    sim.set_scales(bed["scale"])
    sim.set_inv_mass(bed["inv_mass"])
    sim.set_inv_inertia(bed["inv_inertia"])
    sim.set_gravity(0.0, 0.0, -9.81)
    sim.set_material_params(0.3, 0.0, 0.5)
    sim.set_hertz_material(0, 1e7, 0.3)
    sim.step_hertz(dt, SUBSTEPS_PER_STEP)  # warm-up: pair list and buffers
    sim.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sim.step_hertz(dt, steps * SUBSTEPS_PER_STEP)
    sim.synchronize()
    t = time.perf_counter() - t0
    sub = steps * SUBSTEPS_PER_STEP
    pairs = len(sim.debug_hertz_pairs()[0])
    sim.close()
    return {"value": n * sub / t, "unit": "particle-substeps/s", "substeps_timed": sub,
            "ms_per_substep": 1e3 * t / sub, "pairs_at_end": pairs,
            "workload": f"{n} spheres r~U[0.8,1.2] mm, random start in a 160x160x700 mm "
                        "non-periodic box, SDF floor, Hertz-Mindlin E=1e7 Pa nu=0.3, mu=0.5, "
                        "e=0.3, dt=5e-6 s, skin 0.3 r_min"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed steps (8 substeps each)")
    ap.add_argument("--warmup", type=int, default=2, help="untimed steps (8 substeps each)")
    ap.add_argument("--workload", choices=["settled", "pour", "cylinders"], default="settled",
                    help="the headline workload (default: the resting 1M-sphere bed)")
    ap.add_argument("--n", type=int, default=None,
                    help="particles of the headline (default 1,048,576 spheres; 200,000 "
                         "cylinders)")
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--column", type=float, default=0.7, help="pour seeding column height (m)")
    ap.add_argument("--pour-steps", type=int, default=500,
                    help="steps of the pour leg (the whole 500-step job by default)")
    ap.add_argument("--cyl-steps", type=int, default=2, help="steps of the cylinder leg")
    ap.add_argument("--packed-substeps", type=int, default=4,
                    help="timed substeps of the packed-cylinder leg")
    ap.add_argument("--hertz-steps", type=int, default=4, help="steps of the Hertz leg")
    ap.add_argument("--no-pour", action="store_true")
    ap.add_argument("--no-cylinders", action="store_true")
    ap.add_argument("--no-hertz", action="store_true")
    ap.add_argument("--cpu-substeps", type=int, default=2)
    ap.add_argument("--cpu-threads", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="also run a plain copy of MIB MiB (scales the HBM byte counters)")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--tuning", action="append", default=[], metavar="NAME=VALUE",
                    help="any library tuning knob (pd_set_tuning) for the headline, or for "
                         "the leg that --only runs")
    ap.add_argument("--no-autotune", action="store_true",
                    help="time the headline and the batch with the library's default knobs "
                         "instead of autotuning them in the untimed warm-up")
    ap.add_argument("--timing-every", type=int, default=8, metavar="N",
                    help="record timing events on every N-th timed substep (bounds their host "
                         "cost; averages come from the sampled substeps)")
    ap.add_argument("--timing-level", type=int, default=2, choices=[0, 1, 2],
                    help="library timing in the timed region: 2 = per-kernel events and counter "
                         "log (needed for roofline), 1 = stage events only, 0 = none")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="run only the forward+adjoint leg and print its JSON")
    ap.add_argument("--only", choices=["headline", "pour", "cylinders", "cylinders_packed", "hertz",
                                       "batched"],
                    default=None, help="run one leg and print its JSON (profiling)")
    ap.add_argument("--adjoint-n", type=int, default=50000)
    ap.add_argument("--adjoint-case", choices=["both", "loose_start", "dense_bed"], default="both",
                    help="with --adjoint-only: which start to run")
    ap.add_argument("--adjoint-substeps", type=int, default=16)
    ap.add_argument("--batch-systems", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=24)
    ap.add_argument("--batch-workers", type=int, default=16)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU sharding is described in the design notes, not implemented",
              file=sys.stderr)
        sys.exit(2)
    if args.adjoint_only:
        print(json.dumps({k: run_adjoint(args.adjoint_n, args.adjoint_substeps, 8, 8, d)
                          for k, d in (("loose_start", False), ("dense_bed", True))
                          if args.adjoint_case in ("both", k)}))
        return
    if args.n is None:
        args.n = 200000 if args.workload == "cylinders" else 1048576

    import torch

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)

    dt = 5e-6
    tuning = [(kv.split("=", 1)[0], int(kv.split("=", 1)[1])) for kv in args.tuning]
    if args.only == "pour":
        sim, _, r = plain_leg("pour", 1048576, args.seed, args.column, dt, 8,
                              args.pour_steps * SUBSTEPS_PER_STEP, args.timing_every,
                              args.timing_level, tuning)
        sim.close()
        if args.copy_probe:
            copy_probe(args.copy_probe)
        print(json.dumps(r))
        return
    if args.only == "cylinders":
        _, _, r = plain_leg("cylinders", 200000, 3, args.column, dt, SUBSTEPS_PER_STEP,
                            args.cyl_steps * SUBSTEPS_PER_STEP, 1, args.timing_level, tuning)
        print(json.dumps(r))
        return
    if args.only == "cylinders_packed":
        _, _, r = plain_leg("cylinders_packed", 200000, 3, args.column, dt, 1,
                            args.packed_substeps, 1, args.timing_level, tuning)
        print(json.dumps(r))
        return
    if args.only == "hertz":
        print(json.dumps(hertz_leg(1048576, args.seed, args.hertz_steps)))
        return
    if args.only == "batched":
        print(json.dumps(run_batched(args.batch_systems, 4, args.batch_steps,
                                     max(2, args.batch_steps // 2), args.batch_workers)))
        return

    sim, bed, head = plain_leg(args.workload, args.n, args.seed, args.column, dt,
                               args.warmup * SUBSTEPS_PER_STEP, args.steps * SUBSTEPS_PER_STEP,
                               args.timing_every, args.timing_level, tuning,
                               autotune=not args.no_autotune and not tuning)
    sim.close()
    del sim
    if args.copy_probe:
        copy_probe(args.copy_probe)
    if args.only == "headline":
        print(json.dumps(head))
        return

    legs = {}
    if not args.no_pour and args.workload != "pour":
        _, _, legs["pour"] = plain_leg("pour", 1048576, args.seed, args.column, dt, 8,
                                       args.pour_steps * SUBSTEPS_PER_STEP, args.timing_every,
                                       args.timing_level)
        print(f"# pour: {legs['pour']['value']:.3e} particle-substeps/s", file=sys.stderr,
              flush=True)
    if not args.no_cylinders and args.workload != "cylinders":
        _, _, legs["cylinders"] = plain_leg("cylinders", 200000, 3, args.column, dt,
                                            SUBSTEPS_PER_STEP,
                                            args.cyl_steps * SUBSTEPS_PER_STEP, 1,
                                            args.timing_level)
        print(f"# cylinders: {legs['cylinders']['value']:.3e} particle-substeps/s",
              file=sys.stderr, flush=True)
        # the same cylinders packed at rest: every contact colour up to the last class (62)
        _, _, legs["cylinders_packed"] = plain_leg("cylinders_packed", 200000, 3, args.column, dt,
                                                   1, args.packed_substeps, 1, args.timing_level)
        print(f"# cylinders packed: {legs['cylinders_packed']['value']:.3e} particle-substeps/s",
              file=sys.stderr, flush=True)
    if not args.no_hertz:
        legs["hertz"] = hertz_leg(1048576, args.seed, args.hertz_steps)
        print(f"# hertz: {legs['hertz']['value']:.3e} particle-substeps/s", file=sys.stderr,
              flush=True)

    batched = None
    if not args.no_batched:
        batched = run_batched(args.batch_systems, 4, args.batch_steps,
                              max(2, args.batch_steps // 2), args.batch_workers,
                              not args.no_autotune)
        print(f"# batched: {batched['value']:.3e} particle-substeps/s "
              f"({batched['speedup_vs_sequential']:.2f}x sequential)", file=sys.stderr, flush=True)

    fwd_adjoint = None
    if not args.no_adjoint:
        fwd_adjoint = {
            "loose_start": run_adjoint(args.adjoint_n, args.adjoint_substeps, 8, 8, False),
            "dense_bed": run_adjoint(args.adjoint_n, args.adjoint_substeps, 8, 8, True)}
        for k, v in fwd_adjoint.items():
            print(f"# fwd+adjoint {k}: {v['value']:.3e} particle-substeps/s, "
                  f"{v['cost_vs_forward']:.1f}x the forward time", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or oracle_threads()
        cps, tc = run_cpu_baseline(args.workload, bed, dt, args.cpu_substeps, threads)
        cpu = {"value": cps, "unit": "particle-substeps/s", "cores": threads, "kind": "port",
               "sample": f"oracle (C restatement of the reference algorithm, OpenMP, {threads} "
                         f"threads, {os.cpu_count()} logical CPUs visible) on the same seeded "
                         f"{args.n}-particle headline input from its start, first "
                         f"{args.cpu_substeps} substeps, {tc:.1f} s"}

    line = {
        "metric": METRIC,
        "value": head["value"],
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "ms_per_substep": head["ms_per_substep"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": head["workload"],
            "form": "plain",
            "tuning": head.get("tuning"),
            "substeps_per_step": SUBSTEPS_PER_STEP,
            "substeps_timed": head["substeps_timed"],
        },
        "roofline": head.get("roofline"),
        "cpu_baseline": cpu,
        "substep_roofline": head.get("substep_roofline"),
        "counters_mean": head.get("counters_mean"),
        "kernels": head.get("kernels"),
        "stage_ms_per_substep": head.get("stage_ms_per_substep"),
        "stage_host_ms_per_substep": head.get("stage_host_ms_per_substep"),
        "legs": legs,
        "batched": batched,
        "fwd_adjoint": fwd_adjoint,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
