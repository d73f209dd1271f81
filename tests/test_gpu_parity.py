"""GPU parity: the HIP substep against the CPU oracle on identical seeded scenes.

Bar (FP64): integer/index outputs (pair set, contact and manifold lists, colourings, counters)
bit-exact; floating state within 1e-12 relative (max-norm relative to the array's max-norm).
"""
import numpy as np
import pytest

from oracle_binding import oracle_sim
from scenes import apply_scene, make_scene

RTOL = 1e-12
INT_STATS = ["num_real", "num_particles", "num_pairs", "num_contacts", "num_manifolds",
             "active_manifolds", "num_colors_vel", "num_colors_pos", "rounds_vel", "rounds_pos",
             "sweeps_vel", "sweeps_stab", "sweeps_pos", "stab_ran", "uncolored_vel",
             "uncolored_pos", "ml_levels"]


def hip_sim(capacity):
    from peclet.dem import Simulation

    return Simulation(capacity)


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    if a.size == 0:
        return 0.0
    scale = max(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max() / scale)


def assert_state(h, r, tag):
    for name in ["get_positions", "get_velocities", "get_quaternions", "get_angular_velocities"]:
        a, b = getattr(h, name)(), getattr(r, name)()
        assert a.shape == b.shape, (tag, name)
        err = rel_err(a, b)
        assert err <= RTOL, f"{tag}: {name} rel err {err:.3e}"


def assert_stats(h, r, tag):
    sh, sr = h.get_stats(), r.get_stats()
    for k in INT_STATS:
        assert sh[k] == sr[k], f"{tag}: stat {k}: hip {sh[k]} oracle {sr[k]}"
    for k in ["max_overlap", "max_approach"]:
        assert abs(sh[k] - sr[k]) <= RTOL * max(abs(sr[k]), 1e-300), f"{tag}: {k}"


def assert_topology(h, r, tag, colours=True):
    assert np.array_equal(h.debug_pairs(), r.debug_pairs()), f"{tag}: pairs"
    hi, hd = h.debug_contacts()
    ri, rd = r.debug_contacts()
    assert np.array_equal(hi, ri), f"{tag}: contact bodies"
    assert rel_err(hd, rd) <= RTOL, f"{tag}: contact geometry"
    mi, md = h.debug_manifolds()
    qi, qd = r.debug_manifolds()
    assert np.array_equal(mi, qi), f"{tag}: manifold ids"
    assert rel_err(md, qd) <= RTOL, f"{tag}: manifold sums"
    if not colours:
        return
    assert np.array_equal(h.debug_manifold_colors(), r.debug_manifold_colors()), f"{tag}: mcolors"
    assert np.array_equal(h.debug_contact_colors(), r.debug_contact_colors()), f"{tag}: ccolors"


def assert_ledger(h, r, tag):
    hk, hl, ht, hp = h.debug_ledger()
    rk, rl, rt, rp = r.debug_ledger()
    assert np.array_equal(hk, rk), f"{tag}: ledger keys"
    live = rk != np.uint64(0xFFFFFFFFFFFFFFFF)
    assert rel_err(hl[live], rl[live]) <= RTOL, f"{tag}: ledger lambda"
    assert rel_err(ht[live], rt[live]) <= RTOL, f"{tag}: ledger lambdaT"
    # position-channel impulse: only the first entry of each key run is ever read
    head = live.copy()
    head[1:] &= rk[1:] != rk[:-1]
    assert rel_err(hp[head], rp[head]) <= RTOL, f"{tag}: ledger position impulse"


def run_pair(kind, n, steps, seed=0, check_topology_every=1, block_rows=None, tuning=None,
             log=None, **over):
    scene = make_scene(kind, n, seed)
    scene.update(over)
    h = apply_scene(hip_sim(n), scene)
    if block_rows is not None:
        h.set_tuning("block_solver_rows", block_rows)
    for key, value in (tuning or {}).items():
        h.set_tuning(key, value)
    r = apply_scene(oracle_sim(n), scene)
    for k in range(steps):
        tag = f"{kind} n={n} seed={seed} substep {k}"
        print(tag, flush=True)
        h.step(scene["dt"])
        r.step(scene["dt"])
        assert_stats(h, r, tag)
        if log is not None:
            log.append(r.get_stats())
        if check_topology_every and k % check_topology_every == 0:
            assert_topology(h, r, tag, colours=scene.get("use_gs", True))
        assert_state(h, r, tag)
        assert_ledger(h, r, tag)
        assert np.array_equal(h.debug_grounded(), r.debug_grounded()), f"{tag}: grounded"
    return h, r


@pytest.mark.gpu
def test_bed_first_substep(hip_lib):
    run_pair("bed", 300, 1)


@pytest.mark.gpu
@pytest.mark.parametrize("block_rows", [None, 0])
def test_bed_multistep(hip_lib, block_rows):
    # None: sweep loops inside one workgroup; 0: one launch per colour and sweep
    run_pair("bed", 300, 25, block_rows=block_rows)


@pytest.mark.gpu
@pytest.mark.parametrize("block_rows", [None, 0])
def test_bed_periodic_multistep(hip_lib, block_rows):
    h, r = run_pair("bed_periodic", 400, 25, block_rows=block_rows)
    assert h.get_stats()["num_particles"] > h.num_particles()  # ghosts were generated


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["bed", "bed_periodic", "gas"])
@pytest.mark.parametrize("tuning", [{"velocity_loop_lds": 0}, {"velocity_loop_lds_bodies": 24},
                                    {"velocity_loop_lds_keys": 16}],
                         ids=["global", "lds-and-global", "lds-fallback"])
def test_velocity_loop_body_storage(hip_lib, kind, tuning):
    # the one-workgroup velocity loops default to bodies staged in LDS; here they keep them in
    # global memory, or stage only 24 bodies and leave the rest in global memory, or give up
    # staging above 16 distinct bodies (the gated global-memory loop then runs): every variant
    # is the oracle's result
    run_pair(kind, 300, 15, tuning=tuning)


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["bed", "bed_periodic", "gas"])
@pytest.mark.parametrize("pack", [0, 1])
def test_per_colour_launches_body_records(hip_lib, kind, pack):
    # one launch per colour and sweep (the large-bed path), with each body's velocity and
    # constants packed into one record for the stage (1) or read from the separate arrays (0)
    run_pair(kind, 400, 15, tuning={"block_solver_rows": 0, "velocity_pack_bodies": pack})


@pytest.mark.gpu
@pytest.mark.parametrize("block_rows", [None, 0])
def test_contact_colouring_on_the_main_stream(hip_lib, block_rows):
    # the contact colouring runs on the second stream by default, beside the velocity stage;
    # here on the main stream behind it
    run_pair("bed_periodic", 400, 10, block_rows=block_rows,
             tuning={"side_stream_min_contacts": 1 << 30})


@pytest.mark.gpu
@pytest.mark.parametrize("tail", [0, 64, 100000])
def test_position_tail_split(hip_lib, tail):
    # per-colour position sweeps with no trailing colours in the one-workgroup tail launch (0),
    # the last small colours there (64) and every colour there (100000)
    run_pair("bed_periodic", 400, 12, tuning={"block_solver_rows": 0, "position_tail_rows": tail})


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["bed_periodic", "gas"])
@pytest.mark.parametrize("tail", [0, 64, 100000])
def test_velocity_tail_split(hip_lib, kind, tail):
    # per-colour velocity sweeps (PGS, warm start, stabilization; one-shot GS for the gas) with
    # none, the last small or every colour in the one-workgroup tail launch
    run_pair(kind, 400, 12, tuning={"block_solver_rows": 0, "velocity_tail_rows": tail})


@pytest.mark.gpu
@pytest.mark.parametrize("tuning", [{}, {"block_solver_rows": 0}, {"block_solver_rows": 0,
                                     "serial_class_waves_min_rows": 1}],
                         ids=["default", "per-colour", "per-colour-waves"])
def test_big_sphere_last_colour_class(hip_lib, tuning):
    # both colourings' class 62 holds rows sharing the big sphere (the first substep's manifolds,
    # every substep's contacts): one row after another in one workgroup's loop, one thread in a
    # per-colour launch, or in waves of body-disjoint rows
    log = []
    run_pair("big_among_small", 4000, 4, tuning=tuning, log=log)
    assert log[0]["num_colors_vel"] == 63 and log[0]["num_colors_pos"] == 63


@pytest.mark.gpu
@pytest.mark.parametrize("tuning", [{}, {"colouring_incidence_min_edges": 0}],
                         ids=["contend-commit", "incidence"])
def test_sphere_in_shell_thousands_of_edges_at_one_body(hip_lib, tuning):
    # ~5000 contacts and manifolds share one sphere: ~5000 colouring rounds, the last class
    # holding thousands of rows at that sphere, and (incidence rounds) a per-body list sorted by
    # one workgroup against the list staged through LDS in several chunks
    log = []
    run_pair("sphere_in_shell", 5001, 2, tuning=tuning, log=log)
    assert log[0]["num_contacts"] > 4096 and log[0]["num_colors_pos"] == 63
    assert log[0]["rounds_pos"] > 4096


@pytest.mark.gpu
@pytest.mark.parametrize("n,steps", [(2500, 4), (4000, 3)])
def test_velocity_stage_large_beds(hip_lib, n, steps):
    # 2500 spheres: ~4500-6700 rows over ~2900 bodies, so the staged stage keeps 2400 bodies in
    # LDS and the rest in global memory, with several rounds per colour and the stabilization
    # loop running; 4000 spheres: more than 8192 rows, so staging gives up and the gated
    # global-memory launches run the stage
    run_pair("bed_periodic", n, steps)


@pytest.mark.gpu
def test_bed_small_periodic_box_twins(hip_lib):
    # box of ~2 diameters: both periodic twins of a contact exist (duplicate ledger keys)
    run_pair("bed_small_box", 24, 20)


@pytest.mark.gpu
@pytest.mark.parametrize("kind,n,steps", [("bed", 300, 12), ("bed_periodic", 4000, 3),
                                          ("bed_small_box", 24, 10), ("gas", 300, 10),
                                          ("cylinders", 27, 6), ("big_among_small", 4000, 3)])
def test_colouring_by_incidence_lists(hip_lib, kind, n, steps):
    # every colouring (manifolds and contacts) runs its rounds over per-body edge lists sorted by
    # priority; bodies with more than 16 edges (the big sphere, cylinders touching at many shell
    # points) have their lists sorted by one workgroup each.  Colours, round counts and the
    # per-round uncoloured totals are compared with the oracle's every substep.
    log = []
    run_pair(kind, n, steps, tuning={"colouring_incidence_min_edges": 0}, log=log)
    if kind == "big_among_small":
        assert log[0]["num_colors_pos"] == 63


@pytest.mark.gpu
@pytest.mark.parametrize("block_rows", [None, 0])
def test_gas_one_shot_gs(hip_lib, block_rows):
    run_pair("gas", 300, 20, block_rows=block_rows)


@pytest.mark.gpu
def test_hollow_cylinders(hip_lib):
    run_pair("cylinders", 27, 8)


@pytest.mark.gpu
def test_boxes(hip_lib):
    run_pair("boxes", 27, 8)


@pytest.mark.gpu
def test_cylinders_one_launch_per_colour(hip_lib):
    run_pair("cylinders", 27, 8, block_rows=0)


@pytest.mark.gpu
def test_escalate_stabilization(hip_lib):
    run_pair("bed", 200, 6, stab="escalate")


@pytest.mark.gpu
@pytest.mark.parametrize("gates,kind,n,steps,levels", [
    (None, "bed", 200, 8, 1), (None, "bed_periodic", 400, 12, 2),
    (None, "bed_periodic", 4000, 5, 2), ("1", "bed", 200, 8, 2),
    ("1", "bed_periodic", 400, 12, 5), ("2", "bed_periodic", 400, 12, 1),
    ("7", "bed_periodic", 400, 12, 0)])
def test_multilevel_stabilization(hip_lib, monkeypatch, gates, kind, n, steps, levels):
    """Stabilization mode "multilevel": the aggregation levels, their colourings, the coarse
    sweeps and the quasi-static stop equal the oracle's every substep, under the default slip
    gate and the persistent (1), cone (2) and all (7) gate masks (PECLET_DEM_ML_GATES, read when
    a simulation is created).  `levels`: the deepest hierarchy the run builds, so the coarse
    path runs where it should."""
    if gates is None:
        monkeypatch.delenv("PECLET_DEM_ML_GATES", raising=False)
    else:
        monkeypatch.setenv("PECLET_DEM_ML_GATES", gates)
    log = []
    run_pair(kind, n, steps, stab="multilevel", log=log)
    assert max(st["ml_levels"] for st in log) == levels
    assert sum(st["stab_ran"] for st in log) > 0


@pytest.mark.gpu
@pytest.mark.parametrize("block_rows", [None, 0])
def test_ordered_stabilization(hip_lib, block_rows):
    h, r = run_pair("bed", 200, 6, stab="ordered", block_rows=block_rows)


@pytest.mark.gpu
def test_pair_materials(hip_lib):
    n = 200
    ids = (np.arange(n) % 3).astype(np.int32)
    run_pair("bed", n, 6, mat_ids=ids, pair_materials=[(0, 1, 0.7, 0.1), (2, 2, 0.05, 0.9)])


@pytest.mark.gpu
def test_empty_and_single(hip_lib):
    for n in (0, 1):
        scene = make_scene("bed", max(n, 1), 0)
        scene["pos"] = scene["pos"][:n]
        for key in ("scale", "inv_mass", "inv_inertia", "vel", "angvel"):
            scene[key] = scene[key][:n]
        h = apply_scene(hip_sim(4), scene)
        r = apply_scene(oracle_sim(4), scene)
        for _ in range(3):
            h.step(scene["dt"])
            r.step(scene["dt"])
        assert_stats(h, r, f"n={n}")
        assert_state(h, r, f"n={n}")


@pytest.mark.gpu
def test_compute_overlaps(hip_lib):
    scene = make_scene("bed_periodic", 300, 3)
    h = apply_scene(hip_sim(300), scene)
    r = apply_scene(oracle_sim(300), scene)
    a, b = h.compute_overlaps(), r.compute_overlaps()
    assert a == pytest.approx(b, rel=RTOL, abs=0)
    assert h.num_contacts() == r.num_contacts()


@pytest.mark.gpu
def test_thermostat(hip_lib):
    # KE is reduced in a different (fixed) order on the GPU: tolerance, not bit equality
    scene = make_scene("gas", 200, 5)
    scene["thermostat"] = (0.5, 0.1, 1.0)
    h = apply_scene(hip_sim(200), scene)
    r = apply_scene(oracle_sim(200), scene)
    for _ in range(5):
        h.step(scene["dt"])
        r.step(scene["dt"])
    assert rel_err(h.get_velocities(), r.get_velocities()) <= 1e-12


@pytest.mark.gpu
def test_bed_bit_exact_state(hip_lib):
    """With -ffp-contract=off on both sides the FP64 state agrees to the last bit."""
    h, r = run_pair("bed_periodic", 300, 12, seed=4, check_topology_every=0)
    for name in ["get_positions", "get_velocities", "get_quaternions", "get_angular_velocities"]:
        a, b = getattr(h, name)(), getattr(r, name)()
        assert np.array_equal(a, b), f"{name}: {int((a != b).sum())} of {a.size} differ"


@pytest.mark.gpu
def test_gas_legacy_friction(hip_lib):
    # zero gravity with friction: count-averaged Coulomb friction after the GS sweeps
    run_pair("gas", 300, 15, material=(0.8, 0.0, 0.4))


@pytest.mark.gpu
def test_bed_jacobi_velocity_with_friction(hip_lib):
    # velocity_use_gs=False: count-averaged Jacobi velocity and position solves, legacy friction
    run_pair("bed_periodic", 300, 10, use_gs=False)


@pytest.mark.gpu
def test_bed_jacobi_frictionless(hip_lib):
    run_pair("bed", 300, 10, use_gs=False, material=(0.3, 0.0, 0.0))


@pytest.mark.gpu
def test_bed_in_sdf_box_with_moving_floor(hip_lib):
    run_pair("bed_walls", 300, 15)


@pytest.mark.gpu
def test_cylinders_in_sdf_box(hip_lib):
    sc_walls = [dict(f=lambda p: np.minimum(p[:, 2], 20.0 - p[:, 0]), bounds=((-1, -1, -1),
                     (21, 21, 21)), res=(24, 24, 24), e=0.3, mu=0.4)]
    run_pair("cylinders", 27, 6, walls=sc_walls, planes=[])


@pytest.mark.gpu
def test_frictional_wall_drives_legacy_friction(hip_lib):
    # zero gravity, frictionless bodies, frictional walls: the count-averaged friction path
    sc_walls = [dict(f=lambda p: 5.4 - np.abs(p[:, 0] - 6.0), bounds=((-1, -1, -1), (13, 13, 13)),
                     res=(20, 20, 20), e=0.9, mu=0.5)]
    run_pair("gas", 300, 12, walls=sc_walls)


@pytest.mark.gpu
def test_grid_sdf_boxes(hip_lib):
    run_pair("gridboxes", 27, 20)


@pytest.mark.gpu
@pytest.mark.parametrize("kind,n", [("bed_periodic", 300), ("cylinders", 27), ("gridboxes", 27)])
def test_sdf_grid_export(hip_lib, kind, n, tmp_path):
    scene = make_scene(kind, n, 2)
    h = apply_scene(hip_sim(n), scene)
    r = apply_scene(oracle_sim(n), scene)
    for _ in range(2):
        h.step(scene["dt"])
        r.step(scene["dt"])
    res = (20, 18, 24)
    a, b = h.get_sdf_grid(res), r.get_sdf_grid(res)
    assert a.shape == res
    assert np.array_equal(a, b)
    f = tmp_path / "bed.vti"
    h.export_sdf(str(f), res)
    text = f.read_text()
    vals = text.split('format="ascii">')[1].split("</DataArray>")[0].split()
    assert len(vals) == res[0] * res[1] * res[2]
    np.testing.assert_allclose(np.array(vals, float), b.ravel().astype(np.float32), rtol=1e-5,
                               atol=1e-6)


@pytest.mark.gpu
def test_zero_copy_views(hip_lib):
    scene = make_scene("bed", 50, 1)
    h = apply_scene(hip_sim(50), scene)
    p = h.get_positions_view()
    v = h.get_velocities_view()
    im = h.get_inv_mass_view()
    f = h.get_external_forces_view()
    assert p.shape == (50, 3) and v.shape == (50, 3) and im.shape == (50,) and f.shape == (50, 3)
    assert p.__cuda_array_interface__["strides"] == (32, 8)
    np.testing.assert_array_equal(p.numpy(), h.get_positions())
    np.testing.assert_array_equal(v.numpy(), h.get_velocities())
    np.testing.assert_array_equal(im.numpy(), scene["inv_mass"])
    # writes through the view reach the simulation (external forces are read by predict)
    force = np.tile([0.0, 0.0, -5.0], (50, 1))
    f.assign(force)
    r = apply_scene(oracle_sim(50), scene)
    r.set_external_forces(force)
    h.step(scene["dt"])
    r.step(scene["dt"])
    assert np.array_equal(h.get_positions(), r.get_positions())
    h.step(scene["dt"])
    np.testing.assert_array_equal(p.numpy(), h.get_positions())  # still aliases the state


@pytest.mark.gpu
def test_timing_is_passive_and_logs_counters(hip_lib):
    # level-2 timing records events and logs counters; it must not change a single bit of state
    import ctypes as C

    from peclet.dem._binding import Stats

    scene = make_scene("bed_periodic", 400, 3)
    a = apply_scene(hip_sim(400), scene)
    b = apply_scene(hip_sim(400), scene)
    b._b.set_timing(b._h, 2)
    plain = []
    for _ in range(12):
        a.step(scene["dt"])
        b.step(scene["dt"])
        plain.append(a.get_stats())
    for name in ["get_positions", "get_velocities", "get_quaternions", "get_angular_velocities"]:
        assert np.array_equal(getattr(a, name)(), getattr(b, name)()), name
    n = b._b.get_stats_log(b._h, None, 0)
    assert n == 12
    arr = (Stats * n)()
    assert b._b.get_stats_log(b._h, arr, n) == n
    for k in range(n):
        assert arr[k].as_dict() == plain[k], k
    count = b._b.get_kernel_times(b._h, None, None, None, 0)
    ms = np.zeros(count)
    launches = np.zeros(count, np.int64)
    units = np.zeros(count, np.int64)
    b._b.get_kernel_times(b._h, ms.ctypes.data_as(C.POINTER(C.c_double)),
                          launches.ctypes.data_as(C.POINTER(C.c_long)),
                          units.ctypes.data_as(C.POINTER(C.c_long)), count)
    names = [b._b.kernel_name(k).decode() for k in range(count)]
    q = names.index("k_predict")
    assert launches[q] == 12 and units[q] == 12 * 400 and ms[q] > 0
    v = names.index("k_vsolve_lds")
    want = sum(s["active_manifolds"] * (s["sweeps_vel"] + s["sweeps_stab"]) for s in plain)
    assert units[v] == want
    stage = np.zeros(32)
    nst = b._b.get_stage_times(b._h, stage.ctypes.data_as(C.POINTER(C.c_double)), 32)
    assert nst > 0 and np.all(stage[:nst] >= 0) and stage[:nst].sum() > 0


@pytest.mark.gpu
def test_sampled_timing(hip_lib):
    # timing_every = 3: events on substeps 0, 3, 6, 9 only; the counter log still has every substep
    import ctypes as C

    scene = make_scene("bed_periodic", 400, 5)
    a = apply_scene(hip_sim(400), scene)
    b = apply_scene(hip_sim(400), scene)
    b.set_tuning("timing_every", 3)
    b._b.set_timing(b._h, 2)
    for _ in range(11):
        a.step(scene["dt"])
        b.step(scene["dt"])
    for name in ["get_positions", "get_velocities"]:
        assert np.array_equal(getattr(a, name)(), getattr(b, name)()), name
    assert b._b.get_stats_log(b._h, None, 0) == 11
    count = b._b.get_kernel_times(b._h, None, None, None, 0)
    ms = np.zeros(count)
    launches = np.zeros(count, np.int64)
    units = np.zeros(count, np.int64)
    b._b.get_kernel_times(b._h, ms.ctypes.data_as(C.POINTER(C.c_double)),
                          launches.ctypes.data_as(C.POINTER(C.c_long)),
                          units.ctypes.data_as(C.POINTER(C.c_long)), count)
    names = [b._b.kernel_name(k).decode() for k in range(count)]
    q = names.index("k_predict")
    assert launches[q] == 4 and units[q] == 4 * 400
    with pytest.raises(ValueError):
        b.set_tuning("timing_every", 0)


@pytest.mark.gpu
def test_ghosts_outgrow_the_slot_buffers(hip_lib):
    # 200 spheres with three periodic images each (600 ghosts) in slot buffers reserved for
    # ~310 slots: the first substep finds the ghosts do not fit, grows the buffers and runs the
    # ghosts and the broadphase again; every substep must still equal the oracle's
    scene = make_scene("column_small_box", 200, 4)
    h = hip_sim(200)
    h.set_tuning("slot_headroom", 0)
    h = apply_scene(h, scene)
    r = apply_scene(oracle_sim(200), scene)
    for k in range(6):
        tag = f"column_small_box substep {k}"
        print(tag, flush=True)
        h.step(scene["dt"])
        r.step(scene["dt"])
        if k == 0:
            assert h.get_stats()["num_particles"] - 200 == 600  # every sphere has three images
        assert_stats(h, r, tag)
        assert_topology(h, r, tag)
        assert_state(h, r, tag)
        assert_ledger(h, r, tag)


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["bed_periodic", "gas"])
def test_candidate_pairs_equal_a_fresh_broadphase(hip_lib, kind):
    # the candidate list kept across substeps (default) against a broadphase every substep
    # (skin 0): identical pair lists and identical state, substep after substep
    scene = make_scene(kind, 400, 7)
    a = apply_scene(hip_sim(400), scene)
    b = hip_sim(400)
    b.set_tuning("candidate_skin_permille", 0)
    b = apply_scene(b, scene)
    for k in range(30):
        a.step(scene["dt"])
        b.step(scene["dt"])
        assert np.array_equal(a.debug_pairs(), b.debug_pairs()), (kind, k)
    for name in ["get_positions", "get_velocities", "get_angular_velocities"]:
        assert np.array_equal(getattr(a, name)(), getattr(b, name)()), name


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["bed_periodic", "gas"])
def test_growth_ramp(hip_lib, kind):
    # particles grown from 0.9 of their target scales towards them (updateGrowthScalesKokkos,
    # src/integration.hpp:275-283; the factor ramp of src/sim.hpp:100-108): the scales, the
    # factor and every substep's contacts and state equal the oracle's while the ramp runs and
    # after it has reached 1
    n = 300
    scene = make_scene(kind, n, 5)
    h = apply_scene(hip_sim(n), scene)
    r = apply_scene(oracle_sim(n), scene)
    rate = np.log(1.0 / 0.9) / (15 * scene["dt"])  # the factor reaches 1 at substep 15
    for s in (h, r):
        s.set_growth_params(rate, 0.9)
    grew = False
    for k in range(30):
        tag = f"growth {kind} substep {k}"
        print(tag, flush=True)
        h.step(scene["dt"])
        r.step(scene["dt"])
        assert h.get_growth_factor() == r.get_growth_factor(), tag
        assert np.array_equal(h.get_scales(), r.get_scales()), tag
        grew |= 0.9 < h.get_growth_factor() < 1.0
        assert_stats(h, r, tag)
        assert_topology(h, r, tag, colours=scene.get("use_gs", True))
        assert_state(h, r, tag)
        assert_ledger(h, r, tag)
    assert grew and h.get_growth_factor() == 1.0
