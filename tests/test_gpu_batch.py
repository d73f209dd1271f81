"""Batched ensembles: every system of a batch ends bit-identical to the same system stepped
alone, and to the oracle."""
import numpy as np
import pytest

from oracle_binding import oracle_sim
from scenes import apply_scene, make_scene


@pytest.mark.gpu
@pytest.mark.parametrize("tuning", [{}, {"colouring_incidence_min_edges": 0}],
                         ids=["default", "incidence-colouring"])
def test_batch_systems_equal_independent_runs(hip_lib, tuning):
    from peclet.dem import Batch, Simulation

    specs = [("bed_periodic", 300, 1, dict()), ("bed", 150, 2, dict(material=(0.5, 0.0, 0.3))),
             ("gas", 200, 3, dict()), ("bed_periodic", 40, 4, dict(use_gs=False))]
    scenes = []
    for kind, n, seed, over in specs:
        sc = make_scene(kind, n, seed)
        sc.update(over)
        scenes.append(sc)
    batch = Batch([len(sc["pos"]) for sc in scenes], workers=4)
    alone = []
    for sim, sc in zip(batch.systems, scenes):
        apply_scene(sim, sc)
        alone.append(apply_scene(Simulation(len(sc["pos"])), sc))
        for key, value in tuning.items():
            sim.set_tuning(key, value)  # (the systems stepped alone keep the defaults)
    dt = scenes[0]["dt"]  # the batch steps every system with the same dt
    for _ in range(4):
        batch.step(dt, 2)
    for k in range(len(scenes)):
        for _ in range(8):
            alone[k].step(dt)
    ref = apply_scene(oracle_sim(len(scenes[0]["pos"])), scenes[0])
    for _ in range(8):
        ref.step(dt)
    for k in range(len(scenes)):
        for g in ("get_positions", "get_velocities", "get_angular_velocities", "get_quaternions"):
            a, b = getattr(batch.systems[k], g)(), getattr(alone[k], g)()
            assert np.array_equal(a, b), f"system {k}: {g}"
    for g in ("get_positions", "get_velocities"):
        assert np.array_equal(getattr(batch.systems[0], g)(), getattr(ref, g)()), g
    batch.close()


@pytest.mark.gpu
def test_batch_reports_a_failing_system(hip_lib):
    from peclet.dem import Batch

    batch = Batch([10, 10], workers=2)
    sc = make_scene("bed", 10, 0)
    apply_scene(batch.systems[0], sc)
    apply_scene(batch.systems[1], sc)
    # an open recording whose masses change fails the system's next substep
    batch.systems[1].tape_begin(4)
    batch.systems[1].set_inv_mass(np.full(10, 2.0))
    with pytest.raises(RuntimeError, match="system 1"):
        batch.step(1e-5, 1)
    batch.close()
