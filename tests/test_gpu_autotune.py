"""The autotuner picks performance knobs by timing a simulation's own substeps; results do not
depend on them, so a tuned run equals an untuned run of the same substeps bit for bit."""
import numpy as np
import pytest

from scenes import apply_scene, make_scene


def state(sim):
    return np.concatenate([sim.get_positions(), sim.get_velocities(),
                           sim.get_angular_velocities(), sim.get_quaternions()], axis=1)


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["bed_periodic", "gas"])
def test_autotuned_run_equals_untuned(hip_lib, kind):
    from peclet.dem import Simulation

    n = 2000
    scene = make_scene(kind, n, 7)
    tuned = apply_scene(Simulation(n), scene)
    plain = apply_scene(Simulation(n), scene)
    cfg, ms = tuned.autotune(scene["dt"], substeps=2, use_cache=False)
    print(kind, cfg, ms, flush=True)
    assert cfg["candidate_skin_permille"] in (0, 50, 100, 200, 400)
    assert cfg["block_solver_rows"] in (0, 2048, 16384, 65536)
    assert cfg["velocity_loop_lds"] in (0, 1) and cfg["vsweep_waves"] in (3, 4) and ms > 0.0
    # the tuner stepped the simulation: 11 trials of 1 + 2 substeps at most, at least 3
    steps = 0
    while steps < 64 and not np.array_equal(state(plain), state(tuned)):
        plain.step(scene["dt"])
        steps += 1
    assert 3 <= steps <= 33
    for _ in range(10):
        tuned.step(scene["dt"])
        plain.step(scene["dt"])
    assert np.array_equal(state(tuned), state(plain))
    # a second system of the same size class takes the cached choice without stepping
    other = apply_scene(Simulation(n), scene)
    s0 = state(other)
    cfg2, ms2 = other.autotune(scene["dt"], substeps=2)
    assert ms2 is None and np.array_equal(state(other), s0)


@pytest.mark.gpu
def test_batch_autotune_keeps_each_system_equal_to_its_own_run(hip_lib):
    from peclet.dem import Batch, Simulation

    scenes = [make_scene("bed_periodic", 300, 1), make_scene("bed_periodic", 280, 2)]
    batch = Batch([300, 280], workers=2)
    for sim, sc in zip(batch.systems, scenes):
        apply_scene(sim, sc)
    dt = scenes[0]["dt"]
    batch.autotune(dt, 1)
    batch.step(dt, 6)
    for k, sc in enumerate(scenes):
        alone = apply_scene(Simulation(len(sc["pos"])), sc)
        target = batch.systems[k]
        steps = 0
        while steps < 64 and not np.array_equal(state(alone), state(target)):
            alone.step(dt)
            steps += 1
        assert steps < 64, f"system {k} never matched its own run"
    batch.close()
