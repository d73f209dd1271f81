"""GPU parity on the five benchmark workloads (HIP path through the C-ABI against the oracle).

* Settling of 2,048 monodisperse spheres (r 1 mm, rho 2500, 20 x 20 x 80 mm, periodic x/y,
  floor, v0 = 0, dt 1e-5 s, mu 0.5, e 0.3): 8,000 substeps (1,000 steps, through the contact-rich
  phase of the fill); pairs, contacts, manifolds and both colourings equal at every one of the
  first 240 substeps and every 40th after, every counter and the state (1e-12 relative) at every
  substep, and the bed statistics (height, packing fraction, coordination) equal at the end.
* 1,048,576 spheres: the gravity pour (r ~ U[0.8, 1.2] mm, seeded across the periodic seams) and
  a resting lattice bed (contacts everywhere), 8 substeps each, everything compared every
  substep.
* Hollow cylinders (r_out 2 mm, r_in 1.2 mm, L 4 mm) with a 256-point shell: 125 bodies against
  the oracle; 200,000 bodies stepped on the GPU alone with invariants checked (finite state,
  lambda >= 0, |lambda_T| <= mu * lambda on the ledger).
* 256 independent systems (N ~ U[500, 20000], mu ~ U[0.2, 0.8], e ~ U[0.1, 0.6], phi0 = 0.1):
  every system of one batch bit-identical to the same system stepped alone, four against the
  oracle.

The oracle runs with the host's cores (its results do not depend on the thread count).
"""
import os

import numpy as np
import pytest

from oracle_binding import oracle_binding, oracle_sim
from test_gpu_parity import assert_ledger, assert_state, assert_stats, assert_topology


def hip_sim(capacity):
    from peclet.dem import Simulation

    return Simulation(capacity)


@pytest.fixture
def oracle_threads():
    lib = oracle_binding().lib
    lib.pdref_set_threads(max(1, min(16, os.cpu_count() or 1)))
    yield
    lib.pdref_set_threads(1)


def compare(h, r, tag, topology=True, ledger=True):
    assert_stats(h, r, tag)
    assert_state(h, r, tag)
    if topology:
        assert_topology(h, r, tag)
    if ledger:
        assert_ledger(h, r, tag)
        assert np.array_equal(h.debug_grounded(), r.debug_grounded()), f"{tag}: grounded"


def bed_statistics(sim, radius, area):
    """Bed height (99th percentile of the sphere tops), packing fraction below half of it and
    the mean number of sphere-sphere contacts per sphere (manifolds with a penetrating point)."""
    p = sim.get_positions()
    top = np.quantile(p[:, 2] + radius, 0.99)
    low = p[:, 2] < 0.5 * top
    phi = (4.0 / 3.0 * np.pi * radius[low] ** 3).sum() / (area * 0.5 * top)
    ids, data = sim.debug_manifolds()
    return top, phi, len(ids)


@pytest.mark.gpu
@pytest.mark.timeout(900)
def test_settling_2048_spheres(hip_lib, oracle_threads):
    from peclet.dem.workloads import setup_gravity_bed, sphere_bed

    n, dt = 2048, 1e-5
    bed = sphere_bed(n, 1e-3, 1e-3, (0.02, 0.02, 0.08), 1)
    h = setup_gravity_bed(hip_sim(n), bed, dt, 0.5, 0.3)
    r = setup_gravity_bed(oracle_sim(n), bed, dt, 0.5, 0.3)
    active = stab = 0
    for k in range(8000):
        h.step(dt)
        r.step(dt)
        tag = f"settling 2048 substep {k}"
        full = k < 240 or k % 40 == 0
        if full:
            print(tag, flush=True)
        compare(h, r, tag, topology=full, ledger=full)
        st = h.get_stats()
        active = max(active, st["active_manifolds"])
        stab += st["stab_ran"]
    assert h.get_stats()["num_particles"] > n  # ghosts across the periodic seams
    assert active > 1000 and stab > 0  # the velocity solve and its stabilization did work
    sh = bed_statistics(h, bed["radius"], 0.02 * 0.02)
    sr = bed_statistics(r, bed["radius"], 0.02 * 0.02)
    print("bed height, packing fraction, manifolds:", sh, flush=True)
    assert sh == sr


def _compare_run(h, r, dt, substeps, label):
    for k in range(substeps):
        h.step(dt)
        r.step(dt)
        tag = f"{label} substep {k}"
        print(tag, flush=True)
        compare(h, r, tag)


@pytest.mark.gpu
@pytest.mark.timeout(900)
def test_pour_1m_spheres(hip_lib, oracle_threads):
    from peclet.dem.workloads import setup_gravity_bed, sphere_bed

    n, dt = 1048576, 5e-6
    bed = sphere_bed(n, 0.8e-3, 1.2e-3, (0.16, 0.16, 0.4), 2, column=0.7)
    h = setup_gravity_bed(hip_sim(n), bed, dt, 0.5, 0.3)
    r = setup_gravity_bed(oracle_sim(n), bed, dt, 0.5, 0.3)
    _compare_run(h, r, dt, 8, "pour 1M")
    st = h.get_stats()
    assert st["num_particles"] - n > 10000 and st["num_contacts"] > 100000


@pytest.mark.gpu
@pytest.mark.timeout(900)
def test_resting_bed_1m_spheres(hip_lib, oracle_threads):
    from peclet.dem.workloads import lattice_bed, setup_gravity_bed

    n, dt = 1048576, 5e-6
    bed = lattice_bed(n, 0.16, 2)
    h = setup_gravity_bed(hip_sim(n), bed, dt, 0.5, 0.3)
    r = setup_gravity_bed(oracle_sim(n), bed, dt, 0.5, 0.3)
    _compare_run(h, r, dt, 8, "resting bed 1M")
    st = h.get_stats()
    assert st["active_manifolds"] > 1000000 and st["num_particles"] > n


def cylinder_scene(n, seed, spacing, vel, side):
    """n config-3 hollow cylinders on a jittered lattice (pitch `spacing`, `side` x `side` per
    layer) over the floor of a periodic x/y box `side` pitches wide, random orientations,
    velocities ~ N(0, vel)."""
    from peclet.dem.workloads import cylinder_shell

    rng = np.random.default_rng(seed)
    r_out, r_in, length = 2e-3, 1.2e-3, 4e-3
    rb = np.sqrt(r_out ** 2 + (0.5 * length) ** 2)
    k = np.arange(n)
    iz, rem = np.divmod(k, side * side)
    iy, ix = np.divmod(rem, side)
    pos = np.stack([(ix + 0.5) * spacing, (iy + 0.5) * spacing, rb + iz * spacing], 1)
    pos += rng.uniform(-0.05, 0.05, pos.shape) * spacing
    pos[:, 2] = np.maximum(pos[:, 2], rb)
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    m = 2500.0 * np.pi * (r_out ** 2 - r_in ** 2) * length
    i_axis = 0.5 * m * (r_out ** 2 + r_in ** 2)
    i_diam = m * (3.0 * (r_out ** 2 + r_in ** 2) + length ** 2) / 12.0
    L = side * spacing
    return dict(shape=(2, r_out, length, r_out - r_in), shell=cylinder_shell(256, r_out, r_in,
                                                                             length),
                domain=(L, L, float(pos[:, 2].max() + 4 * rb)), pos=pos, quat=q,
                vel=rng.normal(0.0, vel, (n, 3)), inv_mass=np.full(n, 1.0 / m),
                inv_inertia=np.tile([1.0 / i_diam, 1.0 / i_axis, 1.0 / i_diam], (n, 1)))


def apply_cylinders(sim, sc, dt):
    sim.initialize_shape(*sc["shape"])
    sim.set_shell_points(sc["shell"])
    sim.set_domain(*sc["domain"], True, True, False)
    sim.add_plane(0.0, 0.0, 0.0, 0.0, 0.0, 1.0)
    sim.set_positions(sc["pos"])
    sim.set_quaternions(sc["quat"])
    sim.set_velocities(sc["vel"])
    sim.set_inv_mass(sc["inv_mass"])
    sim.set_inv_inertia(sc["inv_inertia"])
    sim.set_gravity(0.0, 0.0, -9.81)
    sim.set_material_params(0.2, 0.0, 0.4)
    sim.set_solver_iterations(12, 8)
    sim.set_dt(dt)
    return sim


@pytest.mark.gpu
@pytest.mark.timeout(600)
@pytest.mark.parametrize("tuning", [{}, {"block_solver_rows": 0, "serial_class_waves_min_rows": 1},
                                    {"block_solver_rows": 0, "serial_class_waves_min_rows": 1,
                                     "serial_class_wave_launch_rows": 8}],
                         ids=["default", "per-colour-waves", "per-colour-wave-launches"])
def test_hollow_cylinders_256_point_shell(hip_lib, oracle_threads, tuning):
    # per-colour-waves: the position sweeps run one launch per colour and the contact colour
    # 62 (the many contacts of one pair of shells share both bodies) runs in waves, in one
    # workgroup or (wave-launches) each wave of at least 8 rows in a launch of its own
    n, dt = 125, 5e-6
    sc = cylinder_scene(n, 3, 3.9e-3, 0.2, 5)
    h = apply_cylinders(hip_sim(n), sc, dt)
    for key, value in tuning.items():
        h.set_tuning(key, value)
    r = apply_cylinders(oracle_sim(n), sc, dt)
    assert np.array_equal(h.get_shell_points(), r.get_shell_points())
    assert len(h.get_shell_points()) == 256
    contacts = 0
    for k in range(40):
        h.step(dt)
        r.step(dt)
        tag = f"cylinders 125 substep {k}"
        print(tag, flush=True)
        compare(h, r, tag)
        contacts = max(contacts, h.get_stats()["num_contacts"])
    assert contacts > 200


def ledger_cone_violations(sim, prev, mu):
    """Rows of the ledger whose friction impulse leaves the Coulomb cone of
    src/solver_velocity.hpp:760-855: |lambda_T| <= mu (max(p, 0) |N| + max(q, 0)), with |N| the
    manifold's normal-sum length and q the position impulse the row was warm-started with (the
    first entry of its key in the previous substep's ledger `prev`; it is added only for resting
    contacts, so including it always gives an upper bound).  The ledger holds the manifolds
    sorted by (key, manifold index), keys (min << 32) | max of the real indices (a wall's other
    index 0xFFFFFFFF), ~0 for a manifold without points.  Returns (rows checked, violations)."""
    keys, lam, lamT, _ = sim.debug_ledger()
    ids, data = sim.debug_manifolds()
    a = ids[:, 3].astype(np.uint64)
    rb = ids[:, 4].astype(np.int64)
    b = np.where(rb >= 0, rb, np.int64(0xFFFFFFFF)).astype(np.uint64)
    mkey = (np.minimum(a, b) << np.uint64(32)) | np.maximum(a, b)
    mkey[ids[:, 2] <= 0] = np.uint64(0xFFFFFFFFFFFFFFFF)
    order = np.lexsort((np.arange(len(mkey)), mkey))
    assert np.array_equal(keys, mkey[order])  # the ledger is the manifolds in key order
    lenN = np.linalg.norm(data[order, 0:3], axis=1)
    q0 = {}
    for k, q in zip(prev[0].tolist(), prev[3].tolist()):
        q0.setdefault(k, q)
    live = keys != np.uint64(0xFFFFFFFFFFFFFFFF)
    bad = 0
    for k, p, t, ln in zip(keys[live].tolist(), lam[live], lamT[live], lenN[live]):
        bound = mu * (max(p, 0.0) * ln + max(q0.get(k, 0.0), 0.0))
        bad += np.linalg.norm(t) > bound * (1.0 + 1e-9) + 1e-300
    return int(live.sum()), bad


@pytest.mark.gpu
@pytest.mark.timeout(800)
def test_hollow_cylinders_200k_invariants(hip_lib):
    import time

    n, dt, mu = 200000, 5e-6, 0.4
    sc = cylinder_scene(n, 3, 4.6e-3, 0.5, 43)  # 43 x 4.6 mm: the 200 mm footprint
    h = apply_cylinders(hip_sim(n), sc, dt)
    prev = None
    for k in range(16):
        if k == 15:
            prev = h.debug_ledger()
        t0 = time.perf_counter()
        h.step(dt)
        st = h.get_stats()
        print(f"cylinders 200k substep {k}: {st['num_contacts']} contacts, "
              f"{st['num_colors_pos']} contact colours, {st['rounds_pos']} rounds, "
              f"{time.perf_counter() - t0:.2f} s", flush=True)
    st = h.get_stats()
    assert st["num_contacts"] > 1000 and st["num_particles"] > n
    for g in ("get_positions", "get_velocities", "get_angular_velocities", "get_quaternions"):
        assert np.all(np.isfinite(getattr(h, g)())), g
    keys, lam, _, _ = h.debug_ledger()
    live = keys != np.uint64(0xFFFFFFFFFFFFFFFF)
    assert np.all(lam[live] >= 0.0)
    rows, bad = ledger_cone_violations(h, prev, mu)
    print(f"cone checked on {rows} ledger rows: {bad} outside", flush=True)
    assert rows > 1000 and bad == 0


@pytest.mark.gpu
@pytest.mark.timeout(900)
def test_ensemble_256_systems(hip_lib, oracle_threads):
    from peclet.dem import Batch
    from peclet.dem.workloads import ensemble, setup_gravity_bed

    beds = ensemble(256, 4)
    dt, substeps = 1e-5, 16
    rng = np.random.default_rng(4)
    vels = [rng.normal(0.0, 0.2, (len(b["pos"]), 3)) for b in beds]

    def prepare(sim, k):
        setup_gravity_bed(sim, beds[k], dt, beds[k]["mu"], beds[k]["e"])
        sim.set_velocities(vels[k])
        return sim

    batch = Batch([len(b["pos"]) for b in beds], workers=16)
    for k, sim in enumerate(batch.systems):
        prepare(sim, k)
    batch.step(dt, substeps)
    sizes = np.array([len(b["pos"]) for b in beds])
    assert sizes.min() >= 500 and sizes.max() <= 20000
    contacts = 0
    for k in range(len(beds)):
        alone = prepare(hip_sim(len(beds[k]["pos"])), k)
        alone.step_n(dt, substeps)
        for g in ("get_positions", "get_velocities", "get_angular_velocities",
                  "get_quaternions"):
            assert np.array_equal(getattr(batch.systems[k], g)(), getattr(alone, g)()), (k, g)
        contacts += batch.systems[k].get_stats()["active_manifolds"]
        alone.close()
    assert contacts > 0
    for k in np.argsort(sizes)[[0, 1, 128, 255]]:
        print(f"system {k}: {sizes[k]} spheres against the oracle", flush=True)
        r = prepare(oracle_sim(int(sizes[k])), int(k))
        for _ in range(substeps):
            r.step(dt)
        assert_state(batch.systems[int(k)], r, f"system {k}")
        assert_stats(batch.systems[int(k)], r, f"system {k}")
    batch.close()
