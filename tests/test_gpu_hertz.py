"""GPU parity of the Hertz–Mindlin engine and SDF walls: HIP against the oracle, bit for bit.

Forces are gathered per body in the serial pair order on the GPU, so positions, velocities,
angular velocities, quaternions and the cached pair list with its shear history must agree
exactly after thousands of steps with Verlet rebuilds.
"""
import numpy as np
import pytest

from oracle_binding import oracle_sim
from peclet.dem import build_wall_sdf


def hip_sim(n):
    from peclet.dem import Simulation

    return Simulation(n)


def box_wall(lo, hi, res):
    lx, ly = hi[0] - lo[0], hi[1] - lo[1]
    return build_wall_sdf(lambda p: np.minimum.reduce(
        [p[:, 0] - lo[0], lo[0] + lx - p[:, 0], p[:, 1] - lo[1], lo[1] + ly - p[:, 1],
         p[:, 2]]), (lo, hi), resolution=res)


def pile_scene(sim, n, seed, shape=None, mats=False, sp=1.12, row=5, sdf=None):
    rng = np.random.default_rng(seed)
    lo, hi = (0.0, 0.0, -1.0), (8.0, 8.0, 20.0)
    if sdf is not None:
        sim.set_sdf_shape(*sdf)
        shape = "grid"
    elif shape is None:
        sim.set_sphere_shape(0.5)
    else:
        sim.initialize(**shape)
    sim.set_domain(lo, hi)
    sim.enable_periodicity(False, False, False)
    sim.set_material_params(0.4, 0.0, 0.5)
    sim.set_hertz_material(0, 1e7, 0.25)
    sim.set_hertz_material(1, 3e7, 0.3)
    wid = box_wall(lo, hi, (24, 24, 48)).add_to(sim, restitution=0.3, friction=0.6)
    pts = np.array([(1 + sp * (k % row), 1 + sp * ((k // row) % row), 0.56 + sp * (k // row ** 2))
                    for k in range(n)]) + rng.uniform(-0.05, 0.05, (n, 3))
    sim.set_positions(pts)
    sim.set_scales(rng.uniform(0.9, 1.1, n))
    m = rng.uniform(0.8, 1.2, n)
    sim.set_inv_mass(1 / m)
    sim.set_inv_inertia(np.tile((1 / (0.4 * m * 0.25))[:, None], (1, 3)))
    sim.set_velocities(rng.normal(0, 1.0, (n, 3)))
    sim.set_angular_velocities(rng.normal(0, 0.5, (n, 3)))
    if shape is not None:
        q = rng.normal(size=(n, 4))
        sim.set_quaternions(q / np.linalg.norm(q, axis=1, keepdims=True))
    if mats:
        sim.set_material_ids((np.arange(n) % 2).astype(np.int32))
        sim.set_pair_material(0, 1, 0.2, 0.3)
        sim.set_pair_material(1, 1, 0.7, 0.9)
        sim.set_wall_material_id(wid, 1)
    sim.set_wall_velocity(wid, (0.01, 0.0, 0.0), (0.0, 0.0, 0.2), (4.0, 4.0, 0.0))
    sim.set_gravity(0, 0, -10.0)
    return sim


def assert_same(h, r, tag):
    for g in ("get_positions", "get_velocities", "get_angular_velocities", "get_quaternions"):
        a, b = getattr(h, g)(), getattr(r, g)()
        if not np.array_equal(a, b):
            err = np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)
            raise AssertionError(f"{tag}: {g} differs (max rel {err:.3e}, "
                                 f"{int((a != b).sum())} of {a.size})")
    hp, hx = h.debug_hertz_pairs()
    rp, rx = r.debug_hertz_pairs()
    assert np.array_equal(hp, rp), f"{tag}: hertz pairs"
    assert np.array_equal(hx, rx), f"{tag}: shear history"


def run(n, seed, calls, steps, skin=0.3, **kw):
    h = pile_scene(hip_sim(n), n, seed, **kw)
    r = pile_scene(oracle_sim(n), n, seed, **kw)
    for c in range(calls):
        tag = f"n={n} seed={seed} {kw} call {c}"
        print(tag, flush=True)
        h.step_hertz(1e-4, steps, skin)
        r.step_hertz(1e-4, steps, skin)
        assert_same(h, r, tag)
    return h, r


@pytest.mark.gpu
def test_hertz_binary_collision(hip_lib):
    sims = []
    for s in (hip_sim(2), oracle_sim(2)):
        s.set_sphere_shape(0.5)
        s.set_domain((0, 0, 0), (20, 20, 20))
        s.enable_periodicity(False, False, False)
        s.set_material_params(0.5, 0.0, 0.3)
        s.set_hertz_material(0, 1e7, 0.25)
        s.set_positions(np.array([[8, 10, 10], [12, 10.3, 10.0]]))
        s.set_velocities(np.array([[1, 0, 0], [-1, 0, 0.0]]))
        s.step_hertz(5e-5, 60000)
        sims.append(s)
    assert_same(sims[0], sims[1], "binary")


@pytest.mark.gpu
def test_hertz_pile_in_box(hip_lib):
    h, _ = run(60, 3, 4, 1000)
    assert h.get_positions()[:, 2].max() < 8.0


@pytest.mark.gpu
def test_hertz_pile_small_skin_many_rebuilds(hip_lib):
    run(60, 5, 3, 500, skin=0.02)


@pytest.mark.gpu
def test_hertz_pair_and_wall_materials(hip_lib):
    run(50, 7, 2, 800, mats=True)


@pytest.mark.gpu
def test_hertz_hollow_cylinders_with_walls(hip_lib):
    run(27, 11, 2, 600, shape=dict(shape_type=2, radius=0.5, height=0.6, thickness=0.2), sp=1.2,
        row=3)


@pytest.mark.gpu
def test_hertz_rejects_planes(hip_lib):
    s = hip_sim(2)
    s.set_positions(np.array([[1, 1, 1], [3, 3, 3.0]]))
    s.set_domain((0, 0, 0), (10, 10, 10))
    s.enable_periodicity(False, False, False)
    s.add_plane(0, 0, 0, 0, 0, 1)
    with pytest.raises(RuntimeError):
        s.step_hertz(1e-4)


@pytest.mark.gpu
def test_hertz_grid_sdf_boxes(hip_lib):
    from scenes import make_scene

    sdf = make_scene("gridboxes", 8, 0)["sdf_shape"]
    run(27, 13, 2, 300, sdf=sdf, sp=1.8, row=3)
