"""Pins the CPU oracle: reference test inputs, brute force, closed forms and independent algorithms.

Each check is chosen so a plausible mistake (dropped term, wrong sign or index, swapped operand)
fails it; none re-types the oracle's own formula.
"""
import math
import os

import numpy as np
import pytest

from oracle_binding import oracle_binding, oracle_sim

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def ptr(a, ct):
    import ctypes as C

    return a.ctypes.data_as(C.POINTER(ct))


def oracle_broadphase(pos, rad, n_real, margin, box_cap, brute=False):
    import ctypes as C

    b = oracle_binding().lib
    fn = b.pdref_broadphase_bruteforce if brute else b.pdref_broadphase
    fn.restype = C.c_int
    dp = C.POINTER(C.c_double)
    fn.argtypes = [dp, dp, C.c_int, C.c_int, C.c_double, C.c_double, C.POINTER(C.c_int), C.c_int]
    pos = np.ascontiguousarray(pos, np.float64)
    rad = np.ascontiguousarray(rad, np.float64)
    n = fn(ptr(pos, C.c_double), ptr(rad, C.c_double), len(pos), n_real, margin, box_cap, None, 0)
    out = np.zeros((max(n, 1), 2), np.int32)
    fn(ptr(pos, C.c_double), ptr(rad, C.c_double), len(pos), n_real, margin, box_cap,
       ptr(out, C.c_int), n)
    return out[:n]


# --------------------------------------------------------------------------- broadphase
def test_broadphase_reference_inputs_match_brute_force():
    """tests/arborx/test_broadphase_arborx.cpp:19-89: 4000 particles (3600 real), margin 0.05."""
    d = np.fromfile(os.path.join(GOLD, "broadphase.bin"), np.float32).reshape(-1, 4)
    pos, rad = d[:, :3].astype(np.float64), d[:, 3].astype(np.float64)
    grid = oracle_broadphase(pos, rad, 3600, 0.05, 0.0)
    brute = oracle_broadphase(pos, rad, 3600, 0.05, 0.0, brute=True)
    assert np.array_equal(grid, brute)
    # the reference test's own criterion (|dx| <= bi + bj per axis), evaluated in numpy
    b = rad + 0.05
    expect = []
    for i in range(3600):
        j = np.arange(i + 1, 4000)
        s = b[i] + b[j]
        ok = (np.abs(pos[i, 0] - pos[j, 0]) <= s) & (np.abs(pos[i, 1] - pos[j, 1]) <= s) & \
             (np.abs(pos[i, 2] - pos[j, 2]) <= s)
        expect += [(i, k) for k in j[ok]]
    assert len(expect) > 1000
    assert set(map(tuple, grid.tolist())) == set(expect)
    # canonical order: lexicographic, unique
    keys = grid[:, 0].astype(np.int64) * 10**6 + grid[:, 1]
    assert np.all(np.diff(keys) > 0)


def test_broadphase_parks_non_finite_and_oversized_boxes():
    rng = np.random.default_rng(3)
    pos = rng.uniform(0, 10, (500, 3))
    rad = rng.uniform(0.2, 0.5, 500)
    pos[7] = np.nan
    pos[11, 1] = np.nan
    rad[13] = 1e6          # larger than the box cap -> parked
    pos[17] = [1e12, 0, 0]  # far away, finite
    grid = oracle_broadphase(pos, rad, 450, 0.05, 10.0)
    brute = oracle_broadphase(pos, rad, 450, 0.05, 10.0, brute=True)
    assert np.array_equal(grid, brute)
    # the parked NaN boxes coincide at -1e30 with zero half-width, so they pair with each other
    assert (7, 11) in set(map(tuple, grid.tolist()))
    assert not any(13 in p and 7 not in p and 11 not in p for p in grid.tolist())


# --------------------------------------------------------------------------- ghosts
def test_ghost_counts_reference_inputs():
    """tests/kokkos/test_periodicity.cpp:20-61: 600 reals in [-5,5]^3, band 0.5, all periodic."""
    p = np.fromfile(os.path.join(GOLD, "periodicity.bin"), np.float32).reshape(-1, 3).astype(float)
    s = oracle_sim(600)
    s.set_sphere_shape(0.5)
    s.set_domain((-5, -5, -5), (5, 5, 5))
    s.enable_periodicity(True, True, True)
    s.set_positions(p)
    s.set_solver_iterations(0, 0)
    s.step(1e-9)
    near = (p < -5 + 0.5) | (p > 5 - 0.5)
    expected = int(sum(2 ** k - 1 for k in near.sum(1)))
    assert s.get_stats()["num_particles"] == 600 + expected
    assert expected > 100


# --------------------------------------------------------------------------- integration
def test_ballistic_trapezoid_is_exact_for_constant_gravity():
    s = oracle_sim(1)
    s.set_sphere_shape(0.1)
    s.set_domain(100, 100, 100, False, False, False)
    x0 = np.array([[50.0, 50.0, 80.0]])
    v0 = np.array([[0.3, -0.2, 1.5]])
    s.set_positions(x0)
    s.set_velocities(v0)
    g = np.array([0.0, 0.0, -9.81])
    s.set_gravity(*g)
    dt, n = 1e-3, 200
    for _ in range(n):
        s.step(dt)
    t = n * dt
    np.testing.assert_allclose(s.get_positions()[0], x0[0] + v0[0] * t + 0.5 * g * t * t,
                               rtol=0, atol=1e-12)
    np.testing.assert_allclose(s.get_velocities()[0], v0[0] + g * t, rtol=0, atol=1e-12)


def test_periodic_wrap():
    s = oracle_sim(1)
    s.set_sphere_shape(0.1)
    s.set_domain(10, 10, 10, True, True, False)
    s.set_positions(np.array([[9.95, 0.02, 5.0]]))
    s.set_velocities(np.array([[1.0, -1.0, 0.0]]))
    s.step(0.1)
    np.testing.assert_allclose(s.get_positions()[0], [0.05, 9.92, 5.0], atol=1e-12)


def test_growth_ramp():
    s = oracle_sim(3)
    s.set_positions(np.array([[0, 0, 0], [5, 0, 0], [10, 0, 0.0]]))
    s.set_scales(np.array([1.0, 2.0, 3.0]))
    s.set_growth_params(2.0, 0.25)
    np.testing.assert_allclose(s.get_scales(), [0.25, 0.5, 0.75])
    f = 0.25
    for _ in range(3):
        s.step(0.1)
        f = min(1.0, f * math.exp(2.0 * 0.1))
    assert s.get_growth_factor() == pytest.approx(f, rel=1e-15)
    np.testing.assert_allclose(s.get_scales(), np.array([1.0, 2.0, 3.0]) * f, rtol=1e-15)


def test_thermostat_berendsen_factor():
    rng = np.random.default_rng(1)
    n = 50
    s = oracle_sim(n)
    s.set_sphere_shape(0.01)
    s.set_domain(100, 100, 100, False, False, False)
    s.set_positions(rng.uniform(0, 100, (n, 3)))
    v = rng.normal(0, 1.0, (n, 3))
    s.set_velocities(v)
    s.set_thermostat(0.3, 0.05, 1.0)
    dt = 0.01
    s.step(dt)
    t_cur = (2.0 * 0.5 * (v ** 2).sum()) / (3.0 * n)
    lam = math.sqrt(1.0 + dt / 0.05 * (0.3 / t_cur - 1.0))
    np.testing.assert_allclose(s.get_velocities(), v * lam, rtol=1e-13)


# --------------------------------------------------------------------------- contacts
def two_spheres(gap, e=0.5, mu=0.0, g=(0, 0, 0), v=((0, 0, 0), (0, 0, 0)), iters=(1, 0)):
    s = oracle_sim(2)
    s.set_sphere_shape(0.5)
    s.set_domain((-10, -10, -10), (10, 10, 10))
    s.enable_periodicity(False, False, False)
    s.set_positions(np.array([[-0.5 - gap / 2, 0.1, 0.2], [0.5 + gap / 2, 0.1, 0.2]]))
    s.set_velocities(np.array(v, float))
    s.set_gravity(*g)
    s.set_material_params(e, 0.0, mu)
    s.set_solver_iterations(*iters)
    return s


def test_overlap_removal_splits_equally():
    delta = 0.02
    s = two_spheres(-delta, iters=(1, 0))
    s.step(1e-4)
    p = s.get_positions()
    np.testing.assert_allclose(p[0], [-0.5, 0.1, 0.2], atol=1e-12)
    np.testing.assert_allclose(p[1], [0.5, 0.1, 0.2], atol=1e-12)
    ids, data = s.debug_contacts()
    assert ids.tolist() == [[0, 1]]
    assert data[0, 9] == pytest.approx(-delta, abs=1e-12)           # signed distance
    np.testing.assert_allclose(data[0, :3], [-1, 0, 0], atol=1e-9)   # normal on B's surface


@pytest.mark.parametrize("e", [0.2, 0.5, 0.8, 0.95])
@pytest.mark.parametrize("g", [0.0, -1e-6])
def test_binary_restitution_exact(e, g):
    """test_colored_gs.py:42-57 (g = 0, one-shot GS) and test_cone_friction.py:102-119 (PGS)."""
    s = two_spheres(3.0, e=e, mu=0.5 if g else 0.0, g=(0, 0, g), v=((1, 0, 0), (-1, 0, 0)),
                    iters=(10, 12))
    for _ in range(400):
        s.step(0.01)
    v = s.get_velocities()
    assert (v[1, 0] - v[0, 0]) / 2.0 == pytest.approx(e, abs=1e-6)
    assert abs(v[:, 0].sum()) < 1e-12


def sphere_on_floor(v0=(0, 0, 0), w0=(0, 0, 0), mu=0.5, e=0.0, beta=0.0, g=(0, 0, -10.0),
                    inv_i=None, z=0.5):
    s = oracle_sim(1)
    s.set_sphere_shape(0.5)
    s.set_domain((0, 0, -1), (200, 8, 6))
    s.enable_periodicity(False, False, False)
    s.add_plane(0, 0, 0, 0, 0, 1)
    s.set_positions(np.array([[5.0, 4.0, z]]))
    s.set_inv_mass(np.ones(1))
    s.set_inv_inertia(np.full((1, 3), 1.0 / (0.4 * 0.25) if inv_i is None else inv_i))
    s.set_velocities(np.array([v0], float))
    s.set_angular_velocities(np.array([w0], float))
    s.set_gravity(*g)
    s.set_material_params(e, beta, mu)
    s.set_solver_iterations(12, 8)
    return s


def test_resting_sphere_stays_put():
    s = sphere_on_floor()
    for _ in range(300):
        s.step(0.005)
    assert abs(s.get_velocities()[0, 2]) < 1e-9
    assert s.get_positions()[0, 2] == pytest.approx(0.5, abs=1e-6)


def test_slide_to_roll_five_sevenths():
    """test_cone_friction.py:74-99,137-140: terminal v = 5/7 v0 and w r = v."""
    s = sphere_on_floor(v0=(4.0, 0, 0))
    for _ in range(1200):
        s.step(0.005)
    v = s.get_velocities()[0]
    w = s.get_angular_velocities()[0]
    assert v[0] / 4.0 == pytest.approx(5 / 7, abs=0.01)
    assert w[1] * 0.5 / 4.0 == pytest.approx(5 / 7, abs=0.01)


@pytest.mark.parametrize("mu,theory", [(0.3, 0.5 * 5 * 64), (0.6, 0.5 * 2 * 64)])
def test_kinetic_slide_distance(mu, theory):
    """test_cone_friction.py:49-71,129-135: locked rotation, a = gx - mu gz."""
    s = sphere_on_floor(mu=mu, g=(8.0, 0, -10.0), inv_i=0.0)
    for _ in range(800):
        s.step(0.01)
    d = s.get_positions()[0, 0] - 5.0
    assert abs(d / theory - 1) < 0.05


def test_walton_oblique_impact():
    """test_cone_friction.py:142-188 (stick and slide regimes of the (e, mu, beta) law)."""
    def oblique(vx, vz, e=0.6, mu=0.4, beta=0.5):
        s = sphere_on_floor(v0=(vx, 0, -vz), mu=mu, e=e, beta=beta, g=(0, 0, -0.02), z=0.62)
        for _ in range(400):
            s.step(0.002)
        v = s.get_velocities()[0]
        w = s.get_angular_velocities()[0]
        return v[0], v[2], v[0] - w[1] * 0.5

    vxp, vzp, utp = oblique(1.0, 2.0)
    assert abs(vzp - 1.2) < 0.03 and abs(utp + 0.5) < 0.04
    assert abs(vxp - (1 - (2 / 7) * 1.5)) < 0.04
    vxp, vzp, _ = oblique(6.0, 2.0)
    assert abs(vzp - 1.2) < 0.03 and abs(vxp - (6.0 - 0.4 * 1.6 * 2.0)) < 0.06
    vxp, _, utp = oblique(1.0, 2.0, beta=0.0)
    assert abs(utp) < 0.03 and abs(vxp - (1 - 2 / 7)) < 0.04


def test_free_cylinder_cloud_conserves_momentum():
    """test_colored_gs.py:60-106: g = 0 cloud of cylinders; total linear momentum is conserved."""
    n = 24
    rng = np.random.default_rng(0)
    s = oracle_sim(n)
    s.initialize(shape_type=2, radius=0.5, height=1.0, thickness=0.2)
    s.set_domain((-20, -20, -20), (20, 20, 20))
    s.enable_periodicity(False, False, False)
    s.set_material_params(0.7, 0.0, 0.0)
    s.set_solver_iterations(8, 10)
    s.set_positions(rng.uniform(-3, 3, (n, 3)))
    v = rng.normal(0, 2.0, (n, 3))
    s.set_velocities(v)
    p0 = v.sum(0)
    spun = 0.0
    for i in range(300):
        s.step(0.004)
        spun = max(spun, np.abs(s.get_angular_velocities()).max())
    assert spun > 0.1  # off-centre hits happened
    np.testing.assert_allclose(s.get_velocities().sum(0), p0, atol=1e-11)


# --------------------------------------------------------------------------- colouring
def splitmix_priority(idx):
    z = (idx + 0x9E3779B9) & 0xFFFFFFFF
    z = ((z ^ (z >> 16)) * 0x21F0AAAD) & 0xFFFFFFFF
    z = ((z ^ (z >> 15)) * 0x735A2D97) & 0xFFFFFFFF
    z ^= z >> 15
    return ((z & 0x7FFFFFFF) << 32) | idx


def greedy_in_priority_order(ends, active):
    """Sequential greedy colouring in decreasing priority: a different algorithm that must give the
    Jones-Plassmann result (an edge commits only after every higher-priority neighbour)."""
    colour = np.full(len(ends), -2, np.int64)
    masks = {}
    order = sorted((i for i in range(len(ends)) if active[i]), key=splitmix_priority, reverse=True)
    for e in order:
        a, b = ends[e]
        forbidden = masks.get(a, 0) | (masks.get(b, 0) if b >= 0 else 0)
        c = 0
        while c < 62 and forbidden & (1 << c):
            c += 1
        colour[e] = c
        masks[a] = masks.get(a, 0) | (1 << c)
        if b >= 0:
            masks[b] = masks.get(b, 0) | (1 << c)
    return colour


@pytest.fixture(scope="module")
def dense_bed():
    from scenes import apply_scene, make_scene

    sc = make_scene("bed_periodic", 300, 9)
    s = apply_scene(oracle_sim(300), sc)
    for _ in range(3):
        s.step(sc["dt"])
    return s


def test_contact_colouring_equals_priority_greedy(dense_bed):
    ids, _ = dense_bed.debug_contacts()
    col = dense_bed.debug_contact_colors()
    ref = greedy_in_priority_order([tuple(x) for x in ids.tolist()], [True] * len(ids))
    assert np.array_equal(col, ref)
    assert col.max() >= 4


def test_manifold_colouring_equals_priority_greedy(dense_bed):
    ids, _ = dense_bed.debug_manifolds()
    col = dense_bed.debug_manifold_colors()
    assert dense_bed.get_stats()["num_particles"] > dense_bed.num_particles()  # ghosts present
    ends_real = [(ra, rb) for _, _, _, ra, rb in ids.tolist()]
    active = (col != -2).tolist()
    # the oracle's own rule for -2: empty manifold or periodic duplicate (realA > realB)
    for (a, b, npts, ra, rb), act in zip(ids.tolist(), active):
        assert act == (npts > 0 and not (b >= 0 and ra > rb))
    ref = greedy_in_priority_order(ends_real, active)
    assert np.array_equal(col, ref)


def test_manifold_sums_regrouped_independently(dense_bed):
    """Group the contacts by canonical pair key in numpy and re-sum the manifold fields."""
    cids, cd = dense_bed.debug_contacts()
    mids, md = dense_bed.debug_manifolds()
    a, b = cids[:, 0].astype(np.int64), cids[:, 1].astype(np.int64)
    key = np.where(b < 0, (a << 32) | 0xFFFFFFFF, (np.minimum(a, b) << 32) | np.maximum(a, b))
    order = np.argsort(key, kind="stable")
    uk, start = np.unique(key[order], return_index=True)
    assert len(uk) == len(mids)
    n, ra, rb, dist = cd[:, 0:3], cd[:, 3:6], cd[:, 6:9], cd[:, 9]
    act = dist <= 0
    flip = (b >= 0) & (b < a)
    na = np.where(flip[:, None], n, -n)
    shift = n * dist[:, None] * 0.5
    rcan = np.where(flip[:, None], rb + shift, ra - shift)
    roth = np.where(flip[:, None], ra - shift, rb + shift)
    fields = [na, np.cross(rcan, na), np.cross(roth, -na), rcan, roth]
    bounds = list(start) + [len(order)]
    for m in range(len(uk)):
        sel = order[bounds[m]:bounds[m + 1]]
        sel = sel[act[sel]]
        assert mids[m, 2] == len(sel)
        for f, arr in enumerate(fields):
            np.testing.assert_allclose(md[m, 3 * f:3 * f + 3], arr[sel].sum(0), rtol=0,
                                       atol=1e-15 + 1e-12 * np.abs(arr[sel]).sum())


def test_grounded_levels_double_buffered_chain():
    """A vertical stack on the floor: each sweep lifts the support level one body further up."""
    n = 12
    s = oracle_sim(n)
    s.set_sphere_shape(0.5)
    s.set_domain(10, 10, 40, False, False, False)
    s.add_plane(0, 0, 0, 0, 0, 1)
    s.set_positions(np.array([[5.0, 5.0, 0.5 + 0.999 * k] for k in range(n)]))
    s.set_gravity(0, 0, -10.0)
    s.set_solver_iterations(1, 1)
    s.step(1e-4)
    lv = s.debug_grounded()
    expect = [256 - (k + 1) if k < 8 else 0 for k in range(n)]
    assert lv.tolist() == expect


def test_api_errors():
    s = oracle_sim(2)
    with pytest.raises(ValueError):
        s.set_stabilization_mode("sideways")
    with pytest.raises(ValueError):
        s.set_pair_material(0, 9, 0.1, 0.1)


# --------------------------------------------------------------------------- shapes
def cylinder_count(r, h, t):
    """Point count of the reference generator (src/shapes_portable.hpp:21-76, spacing rule
    src/sim.hpp:958-964) evaluated from its ceil formulas."""
    md = min(r, t) if min(r, t) >= 1e-4 else r
    sp = max(min(r * 0.3, md * 0.5), 1e-3)
    nv = math.ceil(h / sp)
    n = math.ceil(2 * math.pi * r / sp) * (nv + 1)
    n += math.ceil(2 * math.pi * (r - t) / sp) * (nv + 1) if t > 0 else 0
    nr = math.ceil(t / sp)
    for i in range(nr + 1):
        n += 2 * math.ceil(2 * math.pi * (r - t + t * i / nr) / sp)
    return n


@pytest.mark.parametrize("dims,count", [((2e-3, 4e-3, 0.8e-3), 147), ((2.0, 4.0, 0.8), 715)])
def test_hollow_cylinder_shell_on_surface(dims, count):
    r, h, t = dims
    s = oracle_sim(1)
    s.initialize(shape_type=2, radius=r, height=h, thickness=t)
    pts = s.get_shell_points()
    assert len(pts) == count == cylinder_count(r, h, t)
    rho = np.hypot(pts[:, 0], pts[:, 2])
    on_wall = (np.isclose(rho, r, rtol=1e-6) | np.isclose(rho, r - t, rtol=1e-6)) & \
        (np.abs(pts[:, 1]) <= h / 2 * (1 + 1e-9))
    on_cap = np.isclose(np.abs(pts[:, 1]), h / 2, rtol=1e-9) & (rho >= (r - t) * (1 - 1e-6)) & \
        (rho <= r * (1 + 1e-6))
    assert np.all(on_wall | on_cap)


def test_box_shell_on_surface_with_corners():
    s = oracle_sim(1)
    s.initialize(shape_type=3, radius=1.0)
    pts = s.get_shell_points()
    m = np.abs(pts).max(1)
    np.testing.assert_allclose(m, 1.0, rtol=1e-12)
    corners = {tuple(c) for c in np.sign(pts[np.all(np.isclose(np.abs(pts), 1.0), 1)]).tolist()}
    assert len(corners) == 8


def test_cylinder_standing_on_plane_contacts_bottom_ring():
    """Axis along the plane normal: exactly the shell points at y = -h/2 (bottom cap annuli and
    the j = 0 rows of both walls) lie within the margin, each at the same signed distance."""
    r, h, t = 2.0, 4.0, 0.8
    s = oracle_sim(1)
    s.initialize(shape_type=2, radius=r, height=h, thickness=t)
    s.set_domain((-10, -10, -10), (10, 10, 10))
    s.enable_periodicity(False, False, False)
    s.add_plane(0, 0, 0, 0, 1, 0)
    sink = 0.01
    s.set_positions(np.array([[0.0, h / 2 - sink, 0.0]]))
    s.set_solver_iterations(0, 0)
    s.step(1e-9)
    ids, data = s.debug_contacts()
    pts = s.get_shell_points()
    bottom = np.isclose(pts[:, 1], -h / 2)
    assert len(ids) == int(bottom.sum()) > 20
    assert np.all(ids[:, 1] == -1)
    np.testing.assert_allclose(data[:, 9], -sink, atol=1e-12)
    np.testing.assert_allclose(data[:, 0:3], np.tile([0, 1, 0], (len(ids), 1)), atol=1e-12)


def test_ledger_sorted_and_cone_bounded(dense_bed):
    keys, lam, lamT, pimp = dense_bed.debug_ledger()
    live = keys != np.uint64(0xFFFFFFFFFFFFFFFF)
    assert live.sum() > 50
    assert np.all(np.diff(keys.astype(np.float64)) >= 0)
    assert np.all(lam[live] >= 0)
    assert np.all(np.isfinite(lamT[live]))


# --------------------------------------------------------------------------- SDF walls (XPBD)
def sphere_on(floor, v0=(0, 0, 0), mu=0.5, steps=1200, dt=0.005, belt=None):
    s = oracle_sim(1)
    s.set_sphere_shape(0.5)
    lo, hi = (0, 0, -1), (200, 8, 6)
    s.set_domain(lo, hi)
    s.enable_periodicity(False, False, False)
    if floor == "plane":
        s.add_plane(0, 0, 0, 0, 0, 1)
    else:
        from peclet.dem import build_wall_sdf

        w = build_wall_sdf(lambda p: p[:, 2], (lo, hi), resolution=(64, 8, 16))
        wid = w.add_to(s, restitution=-1.0, friction=mu)
        if belt is not None:
            s.set_wall_velocity(wid, (belt, 0, 0))
    s.set_positions(np.array([[5.0, 4.0, 0.5]]))
    s.set_inv_mass(np.ones(1))
    s.set_inv_inertia(np.full((1, 3), 1.0 / (0.4 * 0.25)))
    s.set_velocities(np.array([v0], float))
    s.set_gravity(0, 0, -10.0)
    s.set_material_params(0.0, 0.0, mu)
    s.set_solver_iterations(12, 8)
    for _ in range(steps):
        s.step(dt)
    return s


def test_wall_floor_matches_plane_floor():
    """A linear SDF floor samples exactly like the plane z = 0 (normal (0,0,1), gap z - r)."""
    a = sphere_on("plane", v0=(4.0, 0, 0))
    b = sphere_on("wall", v0=(4.0, 0, 0))
    np.testing.assert_allclose(b.get_positions(), a.get_positions(), rtol=0, atol=1e-9)
    np.testing.assert_allclose(b.get_velocities(), a.get_velocities(), rtol=0, atol=1e-9)
    assert b.get_velocities()[0, 0] / 4.0 == pytest.approx(5 / 7, abs=0.01)


def test_moving_wall_belt_two_sevenths():
    """A sphere dropped on a belt moving at V ends rolling at 2/7 V (angular momentum about the
    contact point is conserved during the slip)."""
    s = sphere_on("wall", belt=2.0, steps=1200)
    v = s.get_velocities()[0]
    w = s.get_angular_velocities()[0]
    assert v[0] / 2.0 == pytest.approx(2 / 7, abs=0.01)
    assert (v[0] - w[1] * 0.5) == pytest.approx(2.0, abs=0.02)  # contact point moves with the belt


# --------------------------------------------------------------------------- grid-SDF shapes
def box_sdf(p, h=1.0):
    d = np.abs(p) - h
    outside = np.linalg.norm(np.maximum(d, 0.0), axis=1)
    inside = np.minimum(np.max(d, axis=1), 0.0)
    return outside + inside


def grid_of(f, lo, hi, n):
    axes = [np.linspace(lo, hi, n)] * 3
    X, Y, Z = np.meshgrid(*axes, indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    g = f(pts).reshape(n, n, n).ravel(order="F")  # x fastest
    return g, (lo, lo, lo), ((hi - lo) / (n - 1),) * 3


def two_bodies(shape_setup, sep, n=2):
    s = oracle_sim(n)
    shape_setup(s)
    s.set_domain((-10, -10, -10), (10, 10, 10))
    s.enable_periodicity(False, False, False)
    s.set_positions(np.array([[-sep / 2, 0.1, -0.05], [sep / 2, 0.1, -0.05]]))
    s.set_solver_iterations(0, 0)
    s.step(1e-9)
    return s


def test_grid_box_matches_analytic_box():
    """Face-to-face boxes: the trilinear grid of the (piecewise linear) box SDF reproduces the
    analytic box's contacts away from edges."""
    ref_shell = oracle_sim(1)
    ref_shell.initialize(shape_type=3, radius=1.0)
    shell = ref_shell.get_shell_points()
    g, origin, spacing = grid_of(box_sdf, -1.6, 1.6, 65)

    def grid_box(s):
        s.set_sdf_shape(g, 65, 65, 65, origin, spacing, shell, (1.5, 1.5, 1.5), 1.0)

    def analytic_box(s):
        s.initialize(shape_type=3, radius=1.0)

    a = two_bodies(grid_box, 1.96)
    b = two_bodies(analytic_box, 1.96)
    ia, da = a.debug_contacts()
    ib, db = b.debug_contacts()
    assert len(ia) == len(ib) > 8
    np.testing.assert_array_equal(ia, ib)
    face = np.abs(np.abs(db[:, 0]) - 1.0) < 1e-9  # contacts whose normal is the face normal
    assert face.sum() > 4
    np.testing.assert_allclose(da[face, 9], db[face, 9], atol=1e-9)
    np.testing.assert_allclose(da[face, 0:3], db[face, 0:3], atol=1e-6)


def test_grid_sdf_off_grid_residual():
    """A probe beyond the stored lattice reads the clamped sample plus its distance to the
    clamp point, so a unit-sphere grid ending at 1.05 still measures a 0.08 gap exactly."""
    sph = lambda p: np.linalg.norm(p, axis=1) - 1.0  # noqa: E731
    g, origin, spacing = grid_of(sph, -1.05, 1.05, 43)
    shell = np.array([[1.0, 0.0, 0.0], [-1.0, 0.0, 0.0]])

    def grid_sphere(s):
        s.set_sdf_shape(g, 43, 43, 43, origin, spacing, shell, (2.5, 2.5, 2.5), 1.0)

    s = oracle_sim(2)
    grid_sphere(s)
    s.set_domain((-10, -10, -10), (10, 10, 10))
    s.enable_periodicity(False, False, False)
    s.set_positions(np.array([[0.0, 0.0, 0.0], [2.08, 0.0, 0.0]]))
    s.set_solver_iterations(0, 0)
    s.step(1e-9)
    ids, d = s.debug_contacts()
    k = [q for q, (a, b) in enumerate(ids.tolist()) if a == 0 and b == 1]
    assert len(k) == 1
    assert d[k[0], 9] == pytest.approx(0.08, abs=1e-12)


def test_sdf_shape_errors():
    s = oracle_sim(1)
    with pytest.raises(RuntimeError):
        s.set_sdf_shape(np.zeros(8), 2, 2, 1, (0, 0, 0), (1, 1, 1), np.zeros((1, 3)), (1, 1, 1), 1.0)
    with pytest.raises(RuntimeError):
        s.set_sdf_shape(np.zeros(8), 2, 2, 2, (0, 0, 0), (1, 1, 1), np.zeros((0, 3)), (1, 1, 1), 1.0)


def test_legacy_friction_oblique_binary_collision():
    """Zero gravity with friction: count-averaged Coulomb friction after the GS restitution.
    Equal spheres (m = 1, r = 0.5, I^-1 = 10) meet head-on with approach 2 and tangential slip
    2: the friction bound is mu times the impulse that stops the approach (2 / w_n, w_n = 2),
    and since that is below the stick impulse 2 / 7 the contact slides: each body's tangential
    velocity changes by mu * 1 and its spin by r * mu * I^-1."""
    mu, e = 0.1, 0.5
    s = two_spheres(-1e-6, e=e, mu=mu, v=((1, 1, 0), (-1, -1, 0)), iters=(1, 4))
    s.set_inv_inertia(np.full((2, 3), 10.0))
    s.set_inv_mass(np.ones(2))
    s.step(1e-6)
    v = s.get_velocities()
    w = s.get_angular_velocities()
    # (the predicted positions tilt the normal by ~2e-6 during the 1e-6 s step)
    np.testing.assert_allclose(v[:, 0], [-e, e], atol=2e-5)           # GS restitution
    np.testing.assert_allclose(v[:, 1], [1 - mu, -1 + mu], atol=2e-5)  # sliding friction
    np.testing.assert_allclose(w[:, 2], [-0.5 * mu * 10, -0.5 * mu * 10], atol=2e-5)
    np.testing.assert_allclose(v.sum(0), [0, 0, 0], atol=1e-12)


# --------------------------------------------------------------------------- SDF export
def test_sdf_grid_single_sphere_is_exact_distance():
    """One sphere: inside its splat band every voxel holds the exact distance; the Eikonal
    far field grows with slope <= 1 (a distance function) and stays close to the exact value."""
    s = oracle_sim(1)
    s.set_sphere_shape(1.0)
    s.set_domain((0, 0, 0), (8, 8, 8))
    s.enable_periodicity(False, False, False)
    s.set_positions(np.array([[4.0, 4.0, 4.0]]))
    n = 32
    g = s.get_sdf_grid((n, n, n)).ravel()  # x-fastest samples
    idx = np.arange(n ** 3)
    x, y, z = idx % n, (idx // n) % n, idx // (n * n)
    c = (np.stack([x, y, z], 1) + 0.5) * (8.0 / n)
    exact = np.linalg.norm(c - 4.0, axis=1) - 1.0
    # splatted cells (inside the particle's 1.2 r bounding box) nearer than 1.2 r stay fixed
    lo, hi = np.floor((4.0 - 1.2) / (8.0 / n)), np.ceil((4.0 + 1.2) / (8.0 / n))
    ijk = np.stack([x, y, z], 1)
    band = (exact < 1.2 - 1e-9) & np.all((ijk >= lo) & (ijk <= hi), axis=1)
    np.testing.assert_allclose(g[band], exact[band], atol=1e-12)
    far = ~band
    assert np.all(g[far] >= exact[far] - 1e-9)          # Eikonal never undershoots the distance
    assert np.max(g[far] - exact[far]) < 0.35             # first-order scheme error
    assert np.all(g[far] < 100.0)                         # everything reached


def test_sdf_grid_periodic_wrap():
    s = oracle_sim(1)
    s.set_sphere_shape(1.0)
    s.set_domain((0, 0, 0), (8, 8, 8))
    s.enable_periodicity(True, False, False)
    s.set_positions(np.array([[0.2, 4.0, 4.0]]))
    n = 16
    g = s.get_sdf_grid((n, n, n)).ravel()
    # the voxel just left of x = 8 sees the sphere through the periodic boundary
    x_hi = (n - 1) + n * (n // 2) + n * n * (n // 2)
    c = np.array([(n - 0.5) * 0.5, (n // 2 + 0.5) * 0.5, (n // 2 + 0.5) * 0.5])
    assert g[x_hi] == pytest.approx(np.linalg.norm(c - [8.2, 4.0, 4.0]) - 1.0, abs=1e-12)


def test_threaded_oracle_is_bit_identical_to_one_thread():
    """The oracle's OpenMP split (element-wise loops, rows of one colour class) changes no bit:
    a resting lattice bed large enough for every split, stepped with 1 and 4 threads."""
    from peclet.dem.workloads import lattice_bed, setup_gravity_bed

    lib = oracle_binding().lib
    bed = lattice_bed(30000, 0.04, 4)
    out = []
    try:
        for threads in (1, 4):
            lib.pdref_set_threads(threads)
            sim = setup_gravity_bed(oracle_sim(30000), bed, 5e-6, 0.5, 0.3)
            stab = 0
            for _ in range(4):
                sim.step(5e-6)
                stab += sim.get_stats()["sweeps_stab"]
            st = sim.get_stats()
            out.append((np.concatenate([sim.get_positions(), sim.get_velocities(),
                                        sim.get_angular_velocities()], 1),
                        st["max_approach"], st["max_overlap"], stab))
    finally:
        lib.pdref_set_threads(1)
    assert out[0][3] > 0  # the stabilization loop ran (all three row kinds exercised)
    assert np.array_equal(out[0][0], out[1][0])
    assert out[0][1:] == out[1][1:]


def test_periodic_rsa_seeds_across_the_seams():
    """Periodic beds are seeded over the whole period: spheres straddle the x/y seams, no two
    spheres overlap through a periodic image, and the native and Python placements agree."""
    from peclet.dem.workloads import sphere_bed

    a = sphere_bed(2048, 1e-3, 1e-3, (0.02, 0.02, 0.08), 1, native=False)
    b = sphere_bed(2048, 1e-3, 1e-3, (0.02, 0.02, 0.08), 1, native=True)
    assert np.array_equal(a["pos"], b["pos"])
    p, r = a["pos"], a["radius"]
    L = np.array([0.02, 0.02, np.inf])
    d = p[:, None, :] - p[None, :, :]
    d[..., :2] -= L[:2] * np.round(d[..., :2] / L[:2])
    dist = np.sqrt((d ** 2).sum(-1))
    np.fill_diagonal(dist, np.inf)
    assert (dist - (r[:, None] + r[None, :])).min() >= 0.0
    straddle = ((p[:, 0] < r) | (p[:, 0] > 0.02 - r) | (p[:, 1] < r) | (p[:, 1] > 0.02 - r))
    assert straddle.sum() > 100
    assert (p[:, 2] - r).min() >= 0.0 and (p[:, 2] + r).max() <= 0.08


# --- the Jacobi (count-averaged) solves: set_velocity_use_gs(False) runs them over every
# manifold / contact (solveVelocityJacobi + applyVelocityDeltasAveraged,
# src/solver_velocity.hpp:38-255; positionSolveJacobi, src/solver_position.hpp:34-134 with
# applyUpdates, src/integration.hpp:168-184); the colouring leftovers run the same functions over
# the uncoloured edges only.

def _supported_ball(n_sup, angle, gap, vz, e, pos_iters=1):
    """A unit ball (radius 0.5, mass 1) above n_sup fixed unit spheres placed symmetrically at
    polar angle `angle` from the vertical, each at centre distance 1 + gap; zero gravity."""
    s = oracle_sim(n_sup + 1)
    s.set_sphere_shape(0.5)
    s.set_domain((-5, -5, -5), (5, 5, 5))
    s.enable_periodicity(False, False, False)
    d = 1.0 + gap
    pos = [[0.0, 0.0, 0.0]]
    for k in range(n_sup):
        phi = 2.0 * math.pi * k / n_sup
        pos.append([-d * math.sin(angle) * math.cos(phi), -d * math.sin(angle) * math.sin(phi),
                    -d * math.cos(angle)])
    s.set_positions(np.array(pos))
    s.set_inv_mass(np.array([1.0] + [0.0] * n_sup))
    s.set_inv_inertia(np.array([[10.0] * 3] + [[0.0] * 3] * n_sup))
    s.set_velocities(np.array([[0.0, 0.0, vz]] + [[0.0] * 3] * n_sup))
    s.set_gravity(0, 0, 0)
    s.set_material_params(e, 0.0, 0.0)
    s.set_solver_iterations(pos_iters, 4)
    s.set_velocity_use_gs(False)
    return s


@pytest.mark.parametrize("e", [0.0, 0.3, 0.9])
def test_jacobi_isolated_contact_is_binary_restitution(e):
    """One contact: count 1, weight min(1, 2/1) = 1, so the Jacobi impulse is the exact binary
    collision: the approach velocity comes back times -e."""
    s = _supported_ball(1, 0.0, -1e-3, -1.0, e)
    s.step(1e-4)
    v = s.get_velocities()
    assert v[0, 2] == pytest.approx(e, abs=1e-12)
    assert np.all(v[1] == 0.0) and abs(v[0, 0]) < 1e-12 and abs(v[0, 1]) < 1e-12


@pytest.mark.parametrize("n_sup", [2, 3, 4])
@pytest.mark.parametrize("e", [0.2, 0.6])
def test_jacobi_symmetric_supports_count_averaged(n_sup, e):
    """k symmetric fixed supports at 30 degrees: each contact alone would reverse its own normal
    approach, (1+e) V cos(a) along n_i; the k deltas are summed and scaled by min(1, 2/k), so the
    ball leaves with v_z = -V + min(1, 2/k) k (1+e) V cos^2(a) = -V + 2 (1+e) V cos^2(a) for every
    k >= 2 (the first sweep; afterwards every contact separates and is skipped)."""
    a, V, dt = math.radians(30.0), 1.0, 1e-4
    s = _supported_ball(n_sup, a, -1e-3, -V, e)
    s.step(dt)
    v = s.get_velocities()[0]
    # contact normals at the predicted positions (ball moved by -V dt); the narrowphase's
    # central-difference normal matches the exact direction to O(eps^2) = 1e-8
    d = 1.0 - 1e-3
    nz = (d * math.cos(a) - V * dt) / math.hypot(d * math.sin(a), d * math.cos(a) - V * dt)
    assert v[2] == pytest.approx(-V + 2.0 * (1.0 + e) * V * nz ** 2, abs=1e-7)
    # the nominal-angle value differs by 1e-4: the check resolves the normals
    assert abs(v[2] - (-V + 2.0 * (1.0 + e) * V * math.cos(a) ** 2)) > 1e-5
    assert abs(v[0]) < 1e-8 and abs(v[1]) < 1e-8  # symmetric (up to the supports' rounding)


def test_jacobi_position_averaging_converges_geometrically():
    """A ball overlapping two fixed spheres at 30 degrees by delta each: the position rows keep
    the narrowphase's normals and arms, so each constraint is linear in the ball's position.
    Every Jacobi sweep moves the ball by the mean of the two single-contact corrections,
    delta_k cos(a) upwards, and each overlap shrinks by the factor sin^2(a): after k sweeps the
    ball has risen by delta cos(a) (1 - sin^(2k) a) / (1 - sin^2 a) = delta (1 - 4^-k) / cos(a)."""
    a, delta = math.radians(30.0), 1e-3
    for k in (1, 2, 12):
        s = _supported_ball(2, a, -delta, 0.0, 0.0, pos_iters=k)
        s.step(1e-4)
        z = s.get_positions()[0, 2]
        assert z == pytest.approx(delta * (1.0 - 0.25 ** k) / math.cos(a), rel=1e-9), k
        assert s.get_stats()["sweeps_pos"] == k  # no early stop on the Jacobi path


# --- stabilization modes "escalate" (3) and "ordered" (4), src/sim.hpp:356-415: a resting
# column of spheres on the floor being compressed (zero restitution), against a one-dimensional
# model of the same solve.
# Every contact of the column is vertical, so the PGS row reduces to scalars: with approach
# speed a (lower body's v_z minus the upper's), inverse-mass sum w and accumulated impulse p,
# the row applies d = max(0, p + a / w) - p (restitution zero), the
# upper body gains d / m, the lower loses it, and the sweep residual is max |d| w.  The model
# orders the rows the way each mode specifies (colour classes; or height levels bottom-up, then
# top-down) and stops where the mode stops.

def _column(n, mode, dt=1e-3, g=10.0, iters=8):
    s = oracle_sim(n)
    s.set_sphere_shape(0.5)
    s.set_domain((-5, -5, -1), (5, 5, 2 * n + 5))
    s.enable_periodicity(False, False, False)
    s.add_plane(0, 0, 0, 0, 0, 1)
    ov = 1e-4
    s.set_positions(np.array([[0.0, 0.0, 0.5 - ov + k * (1.0 - ov)] for k in range(n)]))
    s.set_inv_mass(np.ones(n))
    s.set_inv_inertia(np.full((n, 3), 10.0))
    vel = np.zeros((n, 3))
    vel[:, 2] = -0.05 * (np.arange(n) + 1)  # the column is being compressed
    s.set_velocities(vel)
    s.set_gravity(0, 0, -g)
    s.set_material_params(0.0, 0.0, 0.0)
    s.set_solver_iterations(4, iters)
    s.set_stabilization_mode({0: "off", 1: "onesided", 3: "escalate", 4: "ordered"}[mode])
    return s


def _column_model(n, rows, colours, mode, dt, g, iters):
    """rows: (lower, upper) body pairs (lower -1 = the floor); returns (v_z, main sweeps,
    stabilization sweeps)."""
    v = -0.05 * (np.arange(n) + 1) - g * dt  # predicted velocities (first substep, no warm start)
    p = np.zeros(len(rows))
    vrest = 2.0 * dt * g

    def row(k):
        lo, up = rows[k]
        a = (v[lo] if lo >= 0 else 0.0) - v[up]
        w = 1.0 + (1.0 if lo >= 0 else 0.0)
        new = max(0.0, p[k] + a / w)
        d = new - p[k]
        p[k] = new
        v[up] += d
        if lo >= 0:
            v[lo] -= d
        return abs(d) * w

    by_colour = [k for c in sorted(set(colours)) for k in range(len(rows)) if colours[k] == c]
    sweeps = stab = 0
    m = 0.0
    for _ in range(iters):
        m = max(row(k) for k in by_colour)
        sweeps += 1
        if m <= 0.02 * vrest:
            break
    if mode != 0 and m > vrest:
        if mode == 3:
            for _ in range(256):
                m = max(row(k) for k in by_colour)
                stab += 1
                if m <= vrest:
                    break
        elif mode == 4:
            # level of a row = the lower body's height level (floor row: 0); key = level, colour
            level = [max(lo, 0) for lo, up in rows]
            order = sorted(range(len(rows)), key=lambda k: (level[k], colours[k], k))
            for _ in range(2 * iters):
                m = max(row(k) for k in order)
                m = max(m, max(row(k) for k in reversed(order)))
                stab += 1
                if m <= vrest:
                    break
    return v, sweeps, stab


@pytest.mark.parametrize("mode", [0, 3, 4])
@pytest.mark.parametrize("n", [8, 16])
def test_column_stabilization_matches_the_1d_model(mode, n):
    dt, g, iters = 1e-3, 10.0, 4
    s = _column(n, mode, dt, g, iters)
    s.step(dt)
    ids, _ = s.debug_manifolds()
    colours = s.debug_manifold_colors()
    z = s.get_positions()[:, 2]
    rows, cols = [], []
    for (a, b), c in zip(ids[:, :2], colours):
        if b < 0:
            rows.append((-1, int(a)))
        else:
            lo, up = (a, b) if z[a] < z[b] else (b, a)
            rows.append((int(lo), int(up)))
        cols.append(int(c))
    assert len(rows) == n  # the floor contact and n - 1 neighbour contacts
    v, sweeps, stab = _column_model(n, rows, cols, mode, dt, g, iters)
    st = s.get_stats()
    assert st["sweeps_vel"] == sweeps and st["sweeps_stab"] == stab, (st, sweeps, stab)
    vz = s.get_velocities()[:, 2]
    np.testing.assert_allclose(vz, v, rtol=0, atol=1e-12 * g * dt)
    if mode:
        assert stab > 0  # the column did not converge in the main loop
    if mode == 3:
        assert np.abs(vz).max() <= n * 2.0 * dt * g  # drained to the rest threshold per contact


def _column_substep(v, p_prev, rows, colours, persistent, grounded_lower, dt, g, iters, mode):
    """One substep of the compressed column under the PGS path in one dimension: predicted
    velocities, vn0, warm start from the previous impulses, the main loop, then the
    stabilization (mode 1: side flags hold the lower body of a persistent, slowly approaching
    row whose lower body is grounded and not rising; mode 3: plain sweeps to the threshold).
    Returns (velocities, impulses, main sweeps, stabilization sweeps)."""
    v = v - g * dt
    n = len(rows)
    vrest = 2.0 * dt * g

    def approach(k):
        lo, up = rows[k]
        return (v[lo] if lo >= 0 else 0.0) - v[up]

    vn0 = [approach(k) for k in range(n)]
    by_colour = [k for c in sorted(set(colours)) for k in range(n) if colours[k] == c]
    p = np.array(p_prev, float)
    for k in by_colour:  # warm start: the previous impulse pushes the pair apart
        lo, up = rows[k]
        v[up] += p[k]
        if lo >= 0:
            v[lo] -= p[k]
    held = [False] * n

    def row(k):
        lo, up = rows[k]
        a = approach(k)
        lower_moves = lo >= 0 and not held[k]
        w = 1.0 + (1.0 if lower_moves else 0.0)
        new = max(0.0, p[k] + a / w)
        d = new - p[k]
        p[k] = new
        v[up] += d
        if lower_moves:
            v[lo] -= d
        return abs(d) * w

    sweeps = stab = 0
    m = 0.0
    for _ in range(iters):
        m = max(row(k) for k in by_colour)
        sweeps += 1
        if m <= 0.02 * vrest:
            break
    if mode != 0 and m > vrest:
        if mode == 1:
            for k in range(n):
                lo, up = rows[k]
                # the lower body rises no faster than 8 dt g (its velocity along -g)
                held[k] = (lo >= 0 and persistent[k] and abs(vn0[k]) <= 8.0 * dt * g
                           and v[lo] <= 8.0 * dt * g and grounded_lower[k])
        for _ in range(2 * iters if mode == 1 else 256):
            m = max(row(k) for k in by_colour)
            stab += 1
            if m <= vrest:
                break
    return v, p, sweeps, stab


def _column_levels(n, rows, level):
    """Grounded levels after one substep's update: decay by 8, then 8 double-buffered sweeps
    in which the floor sets its sphere to 255 and a sphere resting on a grounded one takes the
    lower one's level minus 1."""
    g = np.maximum(level - 8, 0)
    for _ in range(8):
        gn = g.copy()
        for lo, up in rows:
            if lo < 0:
                gn[up] = 255
            elif g[lo] - 1 > 0 and g[lo] - 1 > gn[up]:
                gn[up] = g[lo] - 1
        g = gn
    return g


def _column_rows(s):
    ids, _ = s.debug_manifolds()
    colours = s.debug_manifold_colors()
    z = s.get_positions()[:, 2]
    rows = []
    for a, b in ids[:, :2]:
        if b < 0:
            rows.append((-1, int(a)))
        else:
            rows.append((int(a), int(b)) if z[a] < z[b] else (int(b), int(a)))
    return rows, [int(c) for c in colours]


@pytest.mark.parametrize("mode", [1, 3])
def test_column_two_substeps_warm_start_and_side_flags(mode):
    """Two substeps of the compressed column: the second is warm-started from the first's
    ledger (every row persistent), and under the one-sided mode its stabilization holds the
    lower body of each slowly approaching, grounded row (the grounded levels reach the 8 lowest
    spheres in the first substep's 8 sweeps); the one-dimensional model with the same rules
    gives the oracle's sweep counts and velocities."""
    n, dt, g, iters = 12, 1e-3, 10.0, 4
    s = _column(n, mode, dt, g, iters)
    s.step(dt)
    rows1, cols1 = _column_rows(s)
    v1, p1, sw1, st1 = _column_substep(-0.05 * (np.arange(n) + 1.0), np.zeros(len(rows1)),
                                       rows1, cols1, [False] * len(rows1), [False] * len(rows1),
                                       dt, g, iters, mode)
    st = s.get_stats()
    assert (st["sweeps_vel"], st["sweeps_stab"]) == (sw1, st1)
    np.testing.assert_allclose(s.get_velocities()[:, 2], v1, rtol=0, atol=1e-12 * g * dt)
    key1 = {r: p1[k] for k, r in enumerate(rows1)}
    s.step(dt)
    rows2, cols2 = _column_rows(s)
    prev = [key1.get(r, 0.0) for r in rows2]
    persistent = [r in key1 for r in rows2]
    level = _column_levels(n, rows2, _column_levels(n, rows1, np.zeros(n, int)))
    grounded = [lo >= 0 and level[lo] > 0 for lo, up in rows2]
    v2, p2, sw2, st2 = _column_substep(v1, prev, rows2, cols2, persistent, grounded, dt, g,
                                       iters, mode)
    st = s.get_stats()
    assert (st["sweeps_vel"], st["sweeps_stab"]) == (sw2, st2), (st, sw2, st2)
    np.testing.assert_allclose(s.get_velocities()[:, 2], v2, rtol=0, atol=1e-12 * g * dt)
    assert st2 > 0 and all(persistent) and any(prev)
    assert np.array_equal(s.debug_grounded(), level)


# ---- stabilization mode "multilevel" (src/solver_multilevel.hpp, src/sim.hpp:303-355) ----
def _ml_pair(layout, mode="multilevel", dt=1e-3, g=10.0, u=5e-2, u_upper=None):
    """Two unit-mass spheres (radius 0.5) touching each other, the lower one (or both, side by
    side) touching the floor, moving down at u (quasi-static: u + g dt <= 8 g dt), frictionless,
    e = 0, one velocity iteration (so the stabilization loop is capped at two)."""
    s = oracle_sim(2)
    s.set_sphere_shape(0.5)
    s.set_domain((-5, -5, -1), (5, 5, 10))
    s.enable_periodicity(False, False, False)
    s.add_plane(0, 0, 0, 0, 0, 1)
    if layout == "vertical":
        s.set_positions(np.array([[0.0, 0.0, 0.5], [0.0, 0.0, 1.5]]))
    else:
        s.set_positions(np.array([[0.0, 0.0, 0.5], [1.0, 0.0, 0.5]]))
    s.set_inv_mass(np.ones(2))
    s.set_inv_inertia(np.full((2, 3), 10.0))
    s.set_velocities(np.array([[0.0, 0.0, -u], [0.0, 0.0, -(u if u_upper is None else u_upper)]]))
    s.set_gravity(0, 0, -g)
    s.set_material_params(0.0, 0.0, 0.0)
    s.set_solver_iterations(4, 1)
    s.set_stabilization_mode(mode)
    return s, u + g * dt


def test_multilevel_two_sphere_column_closed_form(monkeypatch):
    """Stacked pair on the floor, approach a = u + g dt.  The main sweep leaves the pair at
    -a/2 (in either colour order); the gate fires; one level aggregates the pair (mass 2); the
    first loop iteration's fine sweep leaves it at -a/4 (momentum -a/2), and the coarse floor row
    stops the aggregate: floor impulse 2a in total, the pair row's 0.75a, both bodies at rest;
    the second iteration changes nothing.  Plain sweeps (escalate) and the one-sided pass end
    elsewhere."""
    monkeypatch.delenv("PECLET_DEM_ML_GATES", raising=False)
    s, a = _ml_pair("vertical")
    s.step(1e-3)
    st = s.get_stats()
    assert (st["stab_ran"], st["ml_levels"], st["sweeps_vel"], st["sweeps_stab"]) == (1, 1, 1, 2)
    ids, _ = s.debug_manifolds()
    keys, lam, _, _ = s.debug_ledger()
    floor = [k for k in range(len(ids)) if ids[k, 1] < 0]
    assert len(ids) == 2 and len(floor) == 1
    # the ledger is sorted by key; the wall row's key sorts after the pair's (body B = -1)
    lam_floor = lam[np.argmax(keys)]
    lam_pair = lam[np.argmin(keys)]
    assert abs(lam_floor - 2.0 * a) <= 1e-12 * a and abs(lam_pair - 0.75 * a) <= 1e-12 * a
    np.testing.assert_allclose(s.get_velocities(), 0.0, atol=1e-12 * a)
    for other in ("escalate", "onesided"):
        o, _ = _ml_pair("vertical", other)
        o.step(1e-3)
        assert o.get_stats()["ml_levels"] == 0
        assert np.abs(o.get_velocities()[:, 2]).max() > 0.1 * a


@pytest.mark.parametrize("gates,layout,levels", [
    (None, "vertical", 1), (None, "horizontal", 1),  # slip gate: vt0 = 0, both rows eligible
    ("1", "vertical", 0),                            # persistent: nothing persists on a first substep
    ("2", "vertical", 1), ("2", "horizontal", 0),    # cone: only a support-oriented pair merges
    ("4", "vertical", 1)])
def test_multilevel_eligibility_gates(monkeypatch, gates, layout, levels):
    if gates is None:
        monkeypatch.delenv("PECLET_DEM_ML_GATES", raising=False)
    else:
        monkeypatch.setenv("PECLET_DEM_ML_GATES", gates)
    s, _ = _ml_pair(layout)
    s.step(1e-3)
    st = s.get_stats()
    assert st["stab_ran"] == 1 and st["ml_levels"] == levels


def test_multilevel_ballistic_pair_never_merges(monkeypatch):
    """A pair approaching faster than 8 g dt is ballistic: it is not aggregated (the floor row
    of the resting lower sphere stays quasi-static)."""
    monkeypatch.delenv("PECLET_DEM_ML_GATES", raising=False)
    s, _ = _ml_pair("vertical", u=0.0, u_upper=0.2)  # the pair approaches at 20 g dt
    s.step(1e-3)
    st = s.get_stats()
    assert st["stab_ran"] == 1 and st["ml_levels"] == 0


@pytest.mark.parametrize("seed", [3, 4])
def test_multilevel_conserves_momentum_without_walls(monkeypatch, seed):
    """Every coarse impulse acts on two aggregates with opposite signs, the restriction is
    mass-weighted and the prolongation adds one velocity change to every member, so in a
    wall-free periodic cluster the substep changes the total momentum by exactly the gravity
    impulse, while the hierarchy is non-trivial."""
    monkeypatch.delenv("PECLET_DEM_ML_GATES", raising=False)
    rng = np.random.default_rng(seed)
    dt, g, side = 1e-3, 10.0, 4
    n = side ** 3
    s = oracle_sim(n)
    s.set_sphere_shape(0.5)
    box = side * 0.99
    s.set_domain((0, 0, 0), (box, box, box))
    s.enable_periodicity(True, True, True)
    idx = np.stack(np.meshgrid(*[np.arange(side)] * 3, indexing="ij"), -1).reshape(-1, 3)
    s.set_positions((idx + 0.5) * 0.99 + rng.uniform(-0.004, 0.004, (n, 3)))
    inv_m = rng.uniform(0.5, 2.0, n)
    s.set_inv_mass(inv_m)
    s.set_inv_inertia(np.full((n, 3), 10.0))
    v = rng.normal(0.0, 3.0 * dt * g, (n, 3))
    s.set_velocities(v)
    s.set_gravity(0, 0, -g)
    s.set_material_params(0.0, 0.0, 0.5)
    s.set_solver_iterations(4, 2)
    s.set_stabilization_mode("multilevel")
    m = 1.0 / inv_m
    p0 = (v * m[:, None]).sum(0)
    s.step(dt)
    st = s.get_stats()
    assert st["stab_ran"] == 1 and st["ml_levels"] >= 2
    p1 = (s.get_velocities() * m[:, None]).sum(0)
    drift = p1 - p0 - np.array([0.0, 0.0, -g * dt]) * m.sum()
    assert np.abs(drift).max() <= 1e-13 * np.abs(v * m[:, None]).sum()
