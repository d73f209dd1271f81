"""Seeded scenes applied identically to the HIP Simulation and the CPU oracle."""
from __future__ import annotations

import numpy as np

SPHERE, HOLLOW_CYLINDER, BOX = 1, 2, 3


def random_quats(rng, n):
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    return q


def lattice_bed(n, r, seed, spacing=1.95, jitter=0.03, box_xy=None, layers=None, vel=0.05):
    """Spheres on a jittered cubic lattice resting on z=0, slightly overlapping their neighbours."""
    rng = np.random.default_rng(seed)
    side = box_xy if box_xy is not None else int(np.ceil(np.sqrt(n / (layers or 4))))
    pts = []
    k = 0
    while len(pts) < n:
        iz, rem = divmod(k, side * side)
        iy, ix = divmod(rem, side)
        pts.append(((ix + 0.5) * spacing * r, (iy + 0.5) * spacing * r, r + iz * spacing * r))
        k += 1
    pos = np.array(pts, np.float64)
    pos += rng.uniform(-jitter * r, jitter * r, pos.shape)
    pos[:, 2] = np.maximum(pos[:, 2], 0.97 * r)
    v = rng.normal(0.0, vel, pos.shape)
    return pos, v, side * spacing * r


def apply_scene(sim, scene):
    """scene: dict produced by make_scene; calls the reference API on `sim`."""
    s = scene
    if "sdf_shape" in s:
        sim.set_sdf_shape(*s["sdf_shape"])
    elif s["shape"][0] == SPHERE:
        sim.set_sphere_shape(s["shape"][1])
    else:
        sim.initialize_shape(*s["shape"])
    if s.get("shell") is not None:
        sim.set_shell_points(s["shell"])
    if "domain_minmax" in s:
        sim.set_domain(*s["domain_minmax"])
        sim.enable_periodicity(*s["periodic"])
    else:
        sim.set_domain(*s["domain"], *s["periodic"])
    for pl in s.get("planes", []):
        sim.add_plane(*pl)
    for w in s.get("walls", []):
        from peclet.dem import build_wall_sdf

        wid = build_wall_sdf(w["f"], w["bounds"], resolution=w["res"]).add_to(
            sim, restitution=w["e"], friction=w["mu"])
        if "vel" in w:
            sim.set_wall_velocity(wid, *w["vel"])
        if "mat" in w:
            sim.set_wall_material_id(wid, w["mat"])
    sim.set_positions(s["pos"])
    if "scale" in s:
        sim.set_scales(s["scale"])
    if "inv_mass" in s:
        sim.set_inv_mass(s["inv_mass"])
    if "inv_inertia" in s:
        sim.set_inv_inertia(s["inv_inertia"])
    if "quat" in s:
        sim.set_quaternions(s["quat"])
    if "vel" in s:
        sim.set_velocities(s["vel"])
    if "angvel" in s:
        sim.set_angular_velocities(s["angvel"])
    sim.set_gravity(*s["gravity"])
    sim.set_material_params(*s["material"])
    sim.set_solver_iterations(*s["iters"])
    if "stab" in s:
        sim.set_stabilization_mode(s["stab"])
    if "use_gs" in s:
        sim.set_velocity_use_gs(s["use_gs"])
    if "pair_materials" in s:
        sim.set_material_ids(s["mat_ids"])
        for a, b, e, mu in s["pair_materials"]:
            sim.set_pair_material(a, b, e, mu)
    if "thermostat" in s:
        sim.set_thermostat(*s["thermostat"])
    sim.set_dt(s["dt"])
    return sim


def make_scene(kind, n=256, seed=0):
    rng = np.random.default_rng(seed)
    if kind in ("bed", "bed_periodic", "bed_small_box"):
        r = 1e-3
        pos, v, L = lattice_bed(n, r, seed, vel=0.02)
        scale = rng.uniform(0.97, 1.03, n)
        rho = 2500.0
        rad = r * scale
        m = rho * 4.0 / 3.0 * np.pi * rad ** 3
        per = (kind != "bed")
        if kind == "bed_small_box":
            L = 4.2 * r
        sc = dict(shape=(SPHERE, r), domain=(L, L, 40 * r), periodic=(per, per, False),
                  planes=[(0, 0, 0, 0, 0, 1)], pos=pos, scale=scale, inv_mass=1.0 / m,
                  inv_inertia=np.repeat((1.0 / (0.4 * m * rad ** 2))[:, None], 3, 1), vel=v,
                  angvel=rng.normal(0, 5.0, (n, 3)), gravity=(0, 0, -9.81),
                  material=(0.3, 0.2, 0.5), iters=(12, 8), dt=1e-5)
        if kind == "bed_small_box":
            sc["pos"][:, 0] %= L
            sc["pos"][:, 1] %= L
        return sc
    if kind == "big_among_small":
        # a sphere of ten radii in a bed of unit ones, the nearest of which overlap it by up to
        # half a radius: on the first substep it holds more than 62 active manifolds, so both
        # colourings' last class (62) holds rows that share the big sphere
        r = 1e-3
        pos, v, L = lattice_bed(n, r, seed, vel=0.02, layers=12)
        big = np.array([[L / 2, L / 2, 11.5 * r]])
        keep = np.linalg.norm(pos - big[0], axis=1) > 10.5 * r
        pos = np.concatenate([pos[keep], big])
        m_ = len(pos)
        scale = np.concatenate([rng.uniform(0.97, 1.03, m_ - 1), [10.0]])
        rad = r * scale
        m = 2500.0 * 4.0 / 3.0 * np.pi * rad ** 3
        return dict(shape=(SPHERE, r), domain=(L, L, 60 * r), periodic=(True, True, False),
                    planes=[(0, 0, 0, 0, 0, 1)], pos=pos, scale=scale, inv_mass=1.0 / m,
                    inv_inertia=np.repeat((1.0 / (0.4 * m * rad ** 2))[:, None], 3, 1),
                    vel=np.concatenate([v[keep], np.zeros((1, 3))]),
                    angvel=rng.normal(0, 5.0, (m_, 3)), gravity=(0, 0, -9.81),
                    material=(0.3, 0.2, 0.5), iters=(12, 8), dt=1e-5)
    if kind == "sphere_in_shell":
        # a sphere of forty radii wrapped in one layer of n - 1 unit spheres on a Fibonacci
        # lattice, each overlapping it by 0.2 radii: thousands of contacts and manifolds share the
        # big sphere (its edge lists span several LDS chunks of the per-body sort)
        r = 1e-3
        m_small = n - 1
        k = np.arange(m_small) + 0.5
        phi = np.arccos(1.0 - 2.0 * k / m_small)
        theta = np.pi * (1.0 + 5 ** 0.5) * k
        L = 0.12
        c = np.array([L / 2, L / 2, L / 2])
        R = (40.0 + 1.0 - 0.2) * r
        shell = c + R * np.stack([np.cos(theta) * np.sin(phi), np.sin(theta) * np.sin(phi),
                                  np.cos(phi)], 1)
        pos = np.concatenate([shell, c[None]])
        scale = np.concatenate([np.ones(m_small), [40.0]])
        rad = r * scale
        m = 2500.0 * 4.0 / 3.0 * np.pi * rad ** 3
        return dict(shape=(SPHERE, r), domain=(L, L, L), periodic=(False, False, False),
                    planes=[], pos=pos, scale=scale, inv_mass=1.0 / m,
                    inv_inertia=np.repeat((1.0 / (0.4 * m * rad ** 2))[:, None], 3, 1),
                    vel=rng.normal(0, 0.01, (n, 3)), angvel=rng.normal(0, 1.0, (n, 3)),
                    gravity=(0, 0, 0), material=(0.3, 0.2, 0.5), iters=(12, 8), dt=1e-5)
    if kind == "column_small_box":
        # a 2 x 2 column of spheres in a periodic box two diameters wide: every sphere has three
        # periodic images
        r = 1e-3
        pos, v, L = lattice_bed(n, r, seed, box_xy=2, vel=0.02)
        scale = rng.uniform(0.97, 1.03, n)
        rad = r * scale
        m = 2500.0 * 4.0 / 3.0 * np.pi * rad ** 3
        return dict(shape=(SPHERE, r), domain=(L, L, (n / 4 + 8) * 2 * r),
                    periodic=(True, True, False), planes=[(0, 0, 0, 0, 0, 1)], pos=pos,
                    scale=scale, inv_mass=1.0 / m,
                    inv_inertia=np.repeat((1.0 / (0.4 * m * rad ** 2))[:, None], 3, 1), vel=v,
                    angvel=rng.normal(0, 5.0, (n, 3)), gravity=(0, 0, -9.81),
                    material=(0.3, 0.2, 0.5), iters=(12, 8), dt=1e-5)
    if kind == "bed_walls":
        # non-periodic bed in an SDF box whose floor moves like a belt
        sc = make_scene("bed", n, seed)
        L = sc["domain"][0]
        lo, hi = (-0.1 * L, -0.1 * L, -2e-3), (1.1 * L, 1.1 * L, sc["domain"][2])
        sc["periodic"] = (False, False, False)
        sc["planes"] = []
        sc["walls"] = [dict(f=lambda p: np.minimum.reduce([p[:, 0], L - p[:, 0], p[:, 1],
                                                           L - p[:, 1], p[:, 2]]),
                            bounds=(lo, hi), res=(24, 24, 40), e=0.2, mu=0.6,
                            vel=((1e-3, 0.0, 0.0), (0.0, 0.0, 2.0), (L / 2, L / 2, 0.0)))]
        sc["pos"][:, :2] = np.clip(sc["pos"][:, :2], 1.05e-3, L - 1.05e-3)
        return sc
    if kind == "gridboxes":
        # the box scene with shape 0 imported as a grid SDF of the box and the box's shell
        sc = make_scene("boxes", n, seed)
        axes = [np.linspace(-0.8, 0.8, 33)] * 3
        X, Y, Z = np.meshgrid(*axes, indexing="ij")
        p = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
        d = np.abs(p) - 0.5
        f = np.linalg.norm(np.maximum(d, 0.0), axis=1) + np.minimum(np.max(d, axis=1), 0.0)
        g = f.reshape(33, 33, 33).ravel(order="F")
        k = np.linspace(-0.5, 0.5, 5)
        shell = np.array([(x, y, z) for x in k for y in k for z in k
                          if max(abs(x), abs(y), abs(z)) == 0.5])
        sc["sdf_shape"] = (g, 33, 33, 33, (-0.8, -0.8, -0.8), (0.05, 0.05, 0.05), shell,
                           (6.0, 6.0, 6.0), 0.5)
        return sc
    if kind == "gas":
        # granular gas in a closed box, g = 0: one-shot coloured GS path
        L = 12.0
        pos = rng.uniform(0.6, L - 0.6, (n, 3))
        return dict(shape=(SPHERE, 0.5), domain=(L, L, L), periodic=(True, True, True), pos=pos,
                    vel=rng.normal(0, 1.0, (n, 3)), gravity=(0, 0, 0), material=(0.8, 0.0, 0.0),
                    iters=(6, 8), dt=0.02)
    if kind in ("cylinders", "boxes"):
        shape = (HOLLOW_CYLINDER, 0.5, 1.0, 0.2) if kind == "cylinders" else (BOX, 0.5, 0.0, 0.0)
        side = int(np.ceil(n ** (1 / 3)))
        g = np.stack(np.meshgrid(*[np.arange(side)] * 3, indexing="ij"), -1).reshape(-1, 3)[:n]
        pos = (g + 0.5) * 1.05 + rng.uniform(-0.05, 0.05, (n, 3))
        pos[:, 2] += 0.3
        L = side * 1.05 + 1.0
        return dict(shape=shape, domain=(L, L, 3 * L), periodic=(True, True, False),
                    planes=[(0, 0, 0, 0, 0, 1)], pos=pos, quat=random_quats(rng, n),
                    vel=rng.normal(0, 0.2, (n, 3)), angvel=rng.normal(0, 1.0, (n, 3)),
                    gravity=(0, 0, -9.81), material=(0.4, 0.1, 0.4), iters=(10, 6), dt=2e-3)
    raise ValueError(kind)
