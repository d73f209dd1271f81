"""Pins the oracle's Hertz–Mindlin engine and SDF walls (CPU).

Closed forms and the reference's own acceptance battery (test_hertz.py): the static Hertz
force law, binary restitution through the damping formula, stick on an incline, kinetic slide
distance, slide-to-roll 5/7, and Verlet-list invariance (the skin must not change the result).
"""
import math

import numpy as np
import pytest

from oracle_binding import oracle_sim
from peclet.dem import build_wall_sdf

E0, NU0 = 1.0e7, 0.25
DT = 5e-5


def base(n, lo, hi, e, mu):
    s = oracle_sim(n)
    s.set_sphere_shape(0.5)
    s.set_domain(lo, hi)
    s.enable_periodicity(False, False, False)
    s.set_material_params(e, 0.0, mu)
    s.set_hertz_material(0, E0, NU0)
    return s


def floor_wall(s, lo, hi, res, e, mu):
    w = build_wall_sdf(lambda p: p[:, 2], (lo, hi), resolution=res)
    return w.add_to(s, restitution=e, friction=mu)


def test_static_hertz_force_law():
    """Two resting spheres overlapping by delta: one step gives dv = F m^-1 dt with
    F = 4/3 E* sqrt(R* delta) delta, E* = E / (2 (1 - nu^2)), R* = r / 2."""
    for delta in (1e-3, 4e-3, 1.6e-2):
        s = base(2, (0, 0, 0), (20, 20, 20), 0.5, 0.0)
        s.set_positions(np.array([[10 - 0.5 + delta / 2, 10, 10], [10 + 0.5 - delta / 2, 10, 10]]))
        s.set_inv_mass(np.array([1.0, 1.0]))
        s.set_gravity(0, 0, 0)
        dt = 1e-9
        s.step_hertz(dt, 1)
        v = s.get_velocities()
        es = E0 / (2 * (1 - NU0 ** 2))
        f = 4.0 / 3.0 * es * math.sqrt(0.25 * delta) * delta
        assert v[1, 0] / dt == pytest.approx(f, rel=1e-9)
        assert v[0, 0] / dt == pytest.approx(-f, rel=1e-9)
        assert np.all(v[:, 1:] == 0)


def binary(e):
    s = base(2, (0, 0, 0), (20, 20, 20), e, 0.0)
    s.set_positions(np.array([[8, 10, 10], [12, 10, 10.0]]))
    s.set_inv_mass(np.ones(2))
    s.set_inv_inertia(np.full((2, 3), 1.0))
    s.set_velocities(np.array([[1, 0, 0], [-1, 0, 0.0]]))
    s.set_gravity(0, 0, 0)
    s.step_hertz(DT, 60000)
    v = s.get_velocities()
    return (v[1, 0] - v[0, 0]) / 2.0


@pytest.mark.parametrize("e", [0.2, 0.5, 0.8])
def test_binary_restitution_damping_formula(e):
    """test_hertz.py:27-37,99-103"""
    assert abs(binary(e) - e) < 0.04


def slab_slide(mu, steps=200000):
    pts = np.array([(4 + 1.9 * i, 4 + 1.9 * j, 0.55 + 1.9 * k)
                    for i in range(4) for j in range(4) for k in range(2)])
    n = len(pts)
    s = base(n, (0, 0, -1.0), (60, 12, 8), 0.2, mu)
    floor_wall(s, (0, 0, -1.0), (60, 12, 8), (96, 24, 24), 0.2, mu)
    s.set_positions(pts)
    s.set_inv_mass(np.ones(n))
    s.set_inv_inertia(np.zeros((n, 3)))
    s.set_gravity(5.0, 0.0, -10.0)
    x0 = pts[:, 0].mean()
    s.step_hertz(DT, steps)
    return float(s.get_positions()[:, 0].mean() - x0)


def test_slab_stick_and_slip():
    """test_hertz.py:40-60,105-109 (rotation locked, tan = 0.5)"""
    assert abs(slab_slide(0.9)) < 1.0
    assert slab_slide(0.05) > 5.0


def single(mu, gx, v0, inv_i, steps):
    s = base(1, (0, 0, -1.0), (200, 8, 6), 0.0, mu)
    floor_wall(s, (0, 0, -1.0), (200, 8, 6), (128, 16, 16), 0.0, mu)
    s.set_positions(np.array([[5, 4, 0.5]]))
    s.set_inv_mass(np.ones(1))
    s.set_inv_inertia(np.full((1, 3), inv_i))
    s.set_velocities(np.array([[v0, 0, 0.0]]))
    s.set_gravity(gx, 0.0, -10.0)
    s.step_hertz(DT, steps)
    return s


@pytest.mark.parametrize("mu,acc,tol", [(0.3, 5.0, 0.1), (0.6, 2.0, 0.15)])
def test_kinetic_slide_distance(mu, acc, tol):
    """test_hertz.py:63-75,111-118: a = gx - mu gz with rotation locked"""
    s = single(mu, 8.0, 0.0, 0.0, 160000)
    t = 160000 * DT
    d = s.get_positions()[0, 0] - 5.0
    assert abs(d / (0.5 * acc * t * t) - 1) < tol


def test_slide_to_roll_five_sevenths():
    """test_hertz.py:78-93,120-123"""
    s = single(0.5, 0.0, 4.0, 1.0 / (0.4 * 0.25), 120000)
    v = s.get_velocities()[0]
    w = s.get_angular_velocities()[0]
    assert abs(v[0] / 4.0 - 5 / 7) < 0.03
    assert abs(w[1] * 0.5 / 4.0 - 5 / 7) < 0.05


def pile(skin_frac, steps=3000, n=60, seed=3):
    rng = np.random.default_rng(seed)
    lo, hi = (0, 0, -1.0), (8, 8, 20)
    s = base(n, lo, hi, 0.4, 0.5)
    box = build_wall_sdf(lambda p: np.minimum.reduce([p[:, 0], 8 - p[:, 0], p[:, 1], 8 - p[:, 1],
                                                     p[:, 2]]), (lo, hi), resolution=(24, 24, 48))
    box.add_to(s, restitution=0.4, friction=0.5)
    pts = np.array([(1 + 1.2 * (k % 5), 1 + 1.2 * ((k // 5) % 5), 0.6 + 1.1 * (k // 25))
                    for k in range(n)]) + rng.uniform(-0.05, 0.05, (n, 3))
    s.set_positions(pts)
    s.set_scales(rng.uniform(0.9, 1.1, n))
    m = rng.uniform(0.8, 1.2, n)
    s.set_inv_mass(1 / m)
    s.set_inv_inertia(np.tile((1 / (0.4 * m * 0.25))[:, None], (1, 3)))
    s.set_velocities(rng.normal(0, 0.3, (n, 3)))
    s.set_gravity(0, 0, -10.0)
    s.step_hertz(1e-4, steps, skin_frac)
    return s


def test_verlet_skin_does_not_change_the_trajectory():
    """Only touching pairs exert forces and shear history is carried by key, so a small skin
    (frequent rebuilds) and a large one (rare rebuilds) give the same bits."""
    a, b = pile(0.05), pile(1.5)
    assert len(a.debug_hertz_pairs()[0]) < len(b.debug_hertz_pairs()[0])
    for g in ("get_positions", "get_velocities", "get_angular_velocities"):
        np.testing.assert_array_equal(getattr(a, g)(), getattr(b, g)())
    p = a.get_positions()
    assert p[:, 2].max() < 6.0 and p[:, 2].min() > 0.3  # settled inside the box


def test_hertz_rejects_periodic_and_planes():
    s = base(2, (0, 0, 0), (10, 10, 10), 0.5, 0.0)
    s.set_positions(np.array([[2, 2, 2], [5, 5, 5.0]]))
    s.enable_periodicity(True, False, False)
    with pytest.raises(RuntimeError):
        s.step_hertz(1e-4)
    s.enable_periodicity(False, False, False)
    s.add_plane(0, 0, 0, 0, 0, 1)
    with pytest.raises(RuntimeError):
        s.step_hertz(1e-4)
    with pytest.raises(ValueError):
        s.set_hertz_material(8, 1e7, 0.3)


# --- grid-SDF shapes under the Hertz engine: per-point shell springs against another body or an
# SDF wall (hertzShapePairForces, src/solver_hertz.hpp:454-606; the shape branch of
# hertzWallForces, :148-338), pinned by the closed forms of the reference's test_hertz_shapes.py

def _grid(f, lo, hi, res):
    axes = [np.linspace(lo, hi, res)] * 3
    X, Y, Z = np.meshgrid(*axes, indexing="ij")
    p = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    return f(p).reshape(res, res, res).ravel(order="F"), (hi - lo) / (res - 1)


def _fibonacci_sphere(n, r):
    k = np.arange(n) + 0.5
    z = 1.0 - 2.0 * k / n
    th = math.pi * (3.0 - math.sqrt(5.0)) * np.arange(n)
    s = np.sqrt(1.0 - z * z)
    return r * np.stack([s * np.cos(th), s * np.sin(th), z], 1)


def _cube_shell(a, m):
    """Points on the faces of the cube [-a, a]^3: an m x m lattice per face (edges and corners
    included once)."""
    t = np.linspace(-a, a, m)
    pts = {(round(x, 12), round(y, 12), round(z, 12))
           for x in t for y in t for z in t
           if max(abs(x), abs(y), abs(z)) >= a - 1e-12}
    return np.array(sorted(pts))


def _sdf_sphere(s, r=0.5):
    g, h = _grid(lambda p: np.linalg.norm(p, axis=1) - r, -0.7, 0.7, 56)
    s.set_sdf_shape(g, 56, 56, 56, (-0.7,) * 3, (h,) * 3, _fibonacci_sphere(220, r),
                    (1.0, 1.0, 1.0), r)


def _sdf_cube(s, a=0.5):
    def f(p):
        q = np.abs(p) - a
        return np.linalg.norm(np.maximum(q, 0.0), axis=1) + np.minimum(np.max(q, axis=1), 0.0)

    g, h = _grid(f, -0.8, 0.8, 56)
    s.set_sdf_shape(g, 56, 56, 56, (-0.8,) * 3, (h,) * 3, _cube_shell(a, 8), (6.0, 6.0, 6.0),
                    a * math.sqrt(3.0))


@pytest.mark.parametrize("e", [0.3, 0.8])
def test_sdf_sphere_binary_restitution(e):
    """Two grid-SDF spheres (220-point shells) colliding head-on at relative speed 2: the
    effective restitution is the analytic path's e within the shell discretisation, 0.08
    (test_hertz_shapes.py:34-51, :82-86)."""
    s = oracle_sim(2)
    _sdf_sphere(s)
    s.set_domain((0, 0, 0), (20, 20, 20))
    s.enable_periodicity(False, False, False)
    s.set_positions(np.array([[9.45, 10.0, 10.0], [10.55, 10.0, 10.0]]))  # 0.1 apart
    s.set_scales_uniform(1.0)
    s.set_inv_mass(np.ones(2))
    s.set_inv_inertia(np.ones((2, 3)))
    s.set_velocities(np.array([[1.0, 0, 0], [-1.0, 0, 0]]))
    s.set_gravity(0, 0, 0)
    s.set_material_params(e, 0.0, 0.0)
    s.set_hertz_material(0, E0, NU0)
    s.step_hertz(2e-5, 6000)  # contact at t = 0.05 s, over well before 0.12 s
    v = s.get_velocities()
    e_eff = (v[1, 0] - v[0, 0]) / 2.0
    assert abs(e_eff - e) < 0.08, e_eff
    assert abs(v[0, 0] + v[1, 0]) < 1e-9  # momentum
    assert v[0, 0] < 0.0 < v[1, 0]  # separating again
    p = s.get_positions()
    assert p[1, 0] - p[0, 0] > 1.0  # out of contact (centres more than a diameter apart)


def _cube_on_incline(mu, steps):
    s = oracle_sim(1)
    _sdf_cube(s)
    s.set_domain((0, 0, -1.0), (60, 12, 8))
    s.enable_periodicity(False, False, False)
    w = build_wall_sdf(lambda p: p[:, 2], ((0, 0, -1.0), (60, 12, 8)), resolution=(96, 24, 24))
    w.add_to(s, restitution=0.1, friction=mu)
    s.set_positions(np.array([[4.0, 4.0, 0.505]]))
    s.set_scales_uniform(1.0)
    s.set_inv_mass(np.ones(1))
    s.set_inv_inertia(np.full((1, 3), 6.0))  # I = (2/3) m a^2 = 1/6 for a = 0.5
    s.set_velocities(np.zeros((1, 3)))
    s.set_gravity(5.0, 0.0, -10.0)  # a tan(theta) = 0.5 incline
    s.set_material_params(0.1, 0.0, mu)
    s.set_hertz_material(0, E0, NU0)
    s.step_hertz(2e-5, steps)
    p = s.get_positions()[0]
    return float(p[0] - 4.0), float(p[2])


def test_cube_sticks_and_slides_on_an_incline():
    """A cube resting flat on a tan(theta) = 0.5 incline with free rotation sticks at mu = 0.9
    (tan(theta) < mu: no motion beyond the spring settling) and slides at mu = 0.05 (kinetic
    acceleration g (sin - mu cos) = 5 - 0.05 x 10 = 4.5, so ~9 units in 2 s), staying flat on the
    floor (test_hertz_shapes.py:54-92)."""
    d_hi, z_hi = _cube_on_incline(0.9, 50000)  # 1 s
    d_lo, z_lo = _cube_on_incline(0.05, 100000)  # 2 s
    assert abs(d_hi) < 1.0, d_hi
    assert d_lo > 5.0, d_lo
    assert d_lo < 0.5 * 4.5 * 2.0 ** 2 * 1.05  # no faster than the kinetic closed form
    assert 0.3 < z_hi < 0.7 and 0.3 < z_lo < 0.7
