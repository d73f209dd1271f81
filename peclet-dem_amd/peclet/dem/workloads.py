"""Seeded synthetic workloads: non-overlapping random sphere packings in a box.

``rsa_positions`` is random sequential addition (rejection sampling): candidates are drawn from
``numpy.random.default_rng(seed)`` and accepted in draw order when they overlap no accepted sphere
and clear the floor.  ``order="cell"`` returns the accepted spheres sorted along the cells of a
uniform grid (z-major), an input ordering with spatial locality; ``order="draw"`` keeps draw order.
"""
from __future__ import annotations

import math

import numpy as np


def _draw(rng, m, lo, hi, rmax, per):
    """m candidate centres: over the whole period on periodic axes, else inside [lo+rmax,
    hi-rmax] (clear of the faces)."""
    u = rng.random((m, 3))
    a = np.where(per, lo, lo + rmax)
    w = np.where(per, hi - lo, hi - lo - 2.0 * rmax)
    return a + u * w


def _rsa_native(n, radii, lo, hi, rng, max_tries, per):
    """Same acceptance rule as the Python loop below, run by the library's pd_rsa_place."""
    import ctypes as C

    from . import _product_binding

    lib = _product_binding().lib
    fn = lib.pd_rsa_place
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int)
    fn.restype = C.c_long
    fn.argtypes = [dp, C.c_long, dp, C.c_int, C.c_int, dp, dp, ip, dp, ip]
    out = np.zeros((n, 3), np.float64)
    lo3 = np.ascontiguousarray(lo, np.float64)
    hi3 = np.ascontiguousarray(hi, np.float64)
    per3 = np.ascontiguousarray(per, np.int32)
    placed = C.c_int(0)
    rmax = float(radii.max())
    total = 0
    while placed.value < n:
        m = max(4 * (n - placed.value), n // 2, 4096)
        total += m
        if total > max_tries * max(n, 1):
            raise RuntimeError(f"rsa_positions: could only place {placed.value} of {n} spheres")
        cand = np.ascontiguousarray(_draw(rng, m, lo3, hi3, rmax, per3.astype(bool)))
        fn(cand.ctypes.data_as(dp), m, radii.ctypes.data_as(dp), n, placed.value,
           lo3.ctypes.data_as(dp), hi3.ctypes.data_as(dp), per3.ctypes.data_as(ip),
           out.ctypes.data_as(dp), C.byref(placed))
    return out


def rsa_positions(n, radii, lo, hi, seed, max_tries=200, order="cell", native=None,
                  periodic=(False, False, False)):
    """Place ``n`` spheres with the given radii (array of n) uniformly in the box [lo, hi].

    On a non-periodic axis candidates are drawn clear of both faces (inside [lo+rmax, hi-rmax],
    rmax the largest radius).  On a periodic axis ([lo, hi) is one period) they are drawn over the
    whole period and overlaps are tested against the nearest periodic image, so spheres straddle
    the seam like anywhere else in the bed.  ``native`` (default: when n > 20000) runs the
    acceptance loop in the C library.
    """
    radii = np.ascontiguousarray(radii, np.float64)
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    per = np.asarray(periodic, bool)
    L = hi - lo
    rng = np.random.default_rng(seed)
    rmax = float(radii.max()) if n else 0.0
    h = 2.0 * rmax
    dims = np.maximum(((hi - lo) / h).astype(np.int64), 1)
    if native is None:
        native = n > 20000
    if native and n:
        out = _rsa_native(n, radii, lo, hi, rng, max_tries, per)
        if order == "cell":
            k = np.minimum(((out - lo) / h).astype(np.int64), dims - 1)
            key = (k[:, 2] * dims[1] + k[:, 1]) * dims[0] + k[:, 0]
            perm = np.argsort(key, kind="stable")
            return out[perm], perm
        return out, np.arange(n)
    cells: dict = {}
    out = np.empty((n, 3), np.float64)
    placed = 0
    batch = 4096
    tries = 0
    while placed < n:
        if tries > max_tries * max(n, 1):
            raise RuntimeError(f"rsa_positions: could only place {placed} of {n} spheres")
        cand = _draw(rng, batch, lo, hi, rmax, per)
        for c in cand:
            tries += 1
            r = radii[placed]
            key = tuple(np.minimum(((c - lo) / h).astype(np.int64), dims - 1))
            # neighbour cells per axis, wrapped on periodic axes, each visited once
            near = []
            for ax in range(3):
                lst = []
                for dc in (-1, 0, 1):
                    x = key[ax] + dc
                    if per[ax]:
                        x %= dims[ax]
                    elif x < 0 or x >= dims[ax]:
                        continue
                    if x not in lst:
                        lst.append(x)
                near.append(lst)
            ok = True
            for cz in near[2]:
                for cy in near[1]:
                    for cx in near[0]:
                        lst = cells.get((cx, cy, cz))
                        if not lst:
                            continue
                        for j in lst:
                            d = out[j] - c
                            d = np.where(per, d - L * np.round(d / L), d)
                            if d[0] * d[0] + d[1] * d[1] + d[2] * d[2] < (r + radii[j]) ** 2:
                                ok = False
                                break
                        if not ok:
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if ok:
                out[placed] = c
                cells.setdefault(key, []).append(placed)
                placed += 1
                if placed == n:
                    break
    if order == "cell":
        k = np.minimum(((out - lo) / h).astype(np.int64), dims - 1)
        key = (k[:, 2] * dims[1] + k[:, 1]) * dims[0] + k[:, 0]
        perm = np.argsort(key, kind="stable")
        return out[perm], perm
    return out, np.arange(n)


def sphere_bed(n, r_lo, r_hi, box, seed, rho=2500.0, base_radius=1e-3, column=None,
               order="cell", native=None, periodic=(True, True, False)):
    """Polydisperse sphere bed in SI units.

    Returns dict(pos, scale, inv_mass, inv_inertia, radius, domain) for a bed of ``n`` spheres
    with radius ~ U[r_lo, r_hi] (metres) seeded in ``[0, box[0]] x [0, box[1]] x [0, column]``,
    across the seams of the ``periodic`` axes (default x and y, the beds' footprint).
    """
    rng = np.random.default_rng(seed + 7919)
    radius = rng.uniform(r_lo, r_hi, n) if r_hi > r_lo else np.full(n, float(r_lo))
    col = box[2] if column is None else column
    pos, perm = rsa_positions(n, radius, (0.0, 0.0, 0.0), (box[0], box[1], col), seed,
                              order=order, native=native, periodic=periodic)
    radius = radius[perm]
    mass = rho * 4.0 / 3.0 * math.pi * radius ** 3
    inv_mass = 1.0 / mass
    inv_i = 1.0 / (0.4 * mass * radius ** 2)
    return {
        "pos": pos,
        "radius": radius,
        "scale": radius / base_radius,
        "inv_mass": inv_mass,
        "inv_inertia": np.repeat(inv_i[:, None], 3, axis=1),
        "domain": (float(box[0]), float(box[1]), float(box[2])),
        "base_radius": base_radius,
    }


def setup_gravity_bed(sim, bed, dt, mu, e, iters=(12, 8), periodic=(True, True, False)):
    """Configure a Simulation for a gravity-settling bed (floor at z = 0)."""
    sim.set_sphere_shape(bed["base_radius"])
    lx, ly, lz = bed["domain"]
    sim.set_domain(lx, ly, lz, *periodic)
    sim.add_plane(0.0, 0.0, 0.0, 0.0, 0.0, 1.0)
    sim.set_positions(bed["pos"])
    sim.set_scales(bed["scale"])
    sim.set_inv_mass(bed["inv_mass"])
    sim.set_inv_inertia(bed["inv_inertia"])
    sim.set_gravity(0.0, 0.0, -9.81)
    sim.set_material_params(e, 0.0, mu)
    sim.set_solver_iterations(*iters)
    sim.set_dt(dt)
    return sim


def lattice_bed(n, box_xy, seed, r=1e-3, spacing=1.996, jitter=0.004, spread=0.01, rho=2500.0,
                base_radius=1e-3):
    """Contact-rich bed: spheres on a cubic lattice (pitch ``spacing * r``) over the floor, filling
    a ``box_xy`` x ``box_xy`` periodic footprint layer by layer; radii ~ r * U[1-spread, 1+spread]
    and centres jittered by ``jitter * r``, so most lattice neighbours overlap slightly."""
    rng = np.random.default_rng(seed)
    pitch = spacing * r
    side = max(1, int(box_xy / pitch))
    k = np.arange(n)
    iz, rem = np.divmod(k, side * side)
    iy, ix = np.divmod(rem, side)
    pos = np.stack([(ix + 0.5) * pitch, (iy + 0.5) * pitch, r + iz * pitch], 1).astype(np.float64)
    pos += rng.uniform(-jitter * r, jitter * r, pos.shape)
    radius = r * rng.uniform(1.0 - spread, 1.0 + spread, n)
    pos[:, 2] = np.maximum(pos[:, 2], radius * 0.999)
    mass = rho * 4.0 / 3.0 * math.pi * radius ** 3
    inv_i = 1.0 / (0.4 * mass * radius ** 2)
    L = side * pitch
    return {
        "pos": pos,
        "radius": radius,
        "scale": radius / base_radius,
        "inv_mass": 1.0 / mass,
        "inv_inertia": np.repeat(inv_i[:, None], 3, axis=1),
        "domain": (L, L, float(pos[:, 2].max() + 4 * r)),
        "base_radius": base_radius,
    }


def ensemble(num, seed, n_lo=500, n_hi=20000, r_lo=0.8e-3, r_hi=1.2e-3, mu=(0.2, 0.8),
             e=(0.1, 0.6), phi0=0.1):
    """Independent sphere packings for a batched run: per system (drawn from ``seed`` and the
    system index) a particle count ~ U[n_lo, n_hi], radii ~ U[r_lo, r_hi], friction ~ U[mu],
    restitution ~ U[e], in a cubic box of solid fraction ``phi0`` (periodic in x and y, floor at
    z = 0).  Returns a list of dicts with the bed (see ``sphere_bed``) plus ``mu`` and ``e``."""
    out = []
    for k in range(num):
        rng = np.random.default_rng([seed, k])
        n = int(rng.integers(n_lo, n_hi + 1))
        m_mu = float(rng.uniform(*mu))
        m_e = float(rng.uniform(*e))
        mean_r3 = (r_hi ** 4 - r_lo ** 4) / (4.0 * (r_hi - r_lo))
        side = (n * 4.0 / 3.0 * math.pi * mean_r3 / phi0) ** (1.0 / 3.0)
        bed = sphere_bed(n, r_lo, r_hi, (side, side, side), seed * 100003 + k, native=True)
        bed["mu"] = m_mu
        bed["e"] = m_e
        out.append(bed)
    return out


def cylinder_shell(n, r_out, r_in, length, seed=3):
    """``n`` surface points of a hollow cylinder (axis y, canonical frame) on a seeded lattice:
    points are split over the outer wall, inner wall and the two annular caps in proportion to
    their areas and laid on golden-angle lattices with a seeded phase."""
    rng = np.random.default_rng(seed)
    a_out = 2 * math.pi * r_out * length
    a_in = 2 * math.pi * r_in * length
    a_cap = math.pi * (r_out ** 2 - r_in ** 2)
    areas = np.array([a_out, a_in, a_cap, a_cap])
    counts = np.floor(n * areas / areas.sum()).astype(int)
    counts[0] += n - counts.sum()
    golden = math.pi * (3.0 - math.sqrt(5.0))
    pts = []
    for part, m in enumerate(counts):
        phase = rng.uniform(0, 2 * math.pi)
        k = np.arange(m) + 0.5
        th = phase + golden * np.arange(m)
        if part < 2:
            r = r_out if part == 0 else r_in
            y = -0.5 * length + length * k / m
            pts.append(np.stack([r * np.cos(th), y, r * np.sin(th)], 1))
        else:
            r = np.sqrt(r_in ** 2 + (r_out ** 2 - r_in ** 2) * k / m)
            y = np.full(m, 0.5 * length if part == 2 else -0.5 * length)
            pts.append(np.stack([r * np.cos(th), y, r * np.sin(th)], 1))
    return np.concatenate(pts, 0)


def cylinder_bed(n, box_xy, seed, r_out=2e-3, r_in=1.2e-3, length=4e-3, shell_points=256,
                 rho=2500.0, column=None):
    """Hollow cylinders seeded (RSA on the bounding sphere) in a ``box_xy`` x ``box_xy`` column
    with uniform random orientations; returns the bed dict plus ``shell`` and the shape tuple."""
    rb = math.sqrt(r_out ** 2 + (0.5 * length) ** 2)
    radii = np.full(n, rb)
    col = column if column is not None else n * (4.0 / 3.0 * math.pi * rb ** 3) / (0.28 * box_xy ** 2)
    pos, _ = rsa_positions(n, radii, (0.0, 0.0, 0.0), (box_xy, box_xy, col), seed, native=True,
                           periodic=(True, True, False))
    rng = np.random.default_rng(seed + 1)
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    vol = math.pi * (r_out ** 2 - r_in ** 2) * length
    m = rho * vol
    # principal moments of a hollow cylinder about its axis (y) and a diameter
    i_axis = 0.5 * m * (r_out ** 2 + r_in ** 2)
    i_diam = m * (3.0 * (r_out ** 2 + r_in ** 2) + length ** 2) / 12.0
    return {
        "pos": pos,
        "quat": q,
        "inv_mass": np.full(n, 1.0 / m),
        "inv_inertia": np.tile([1.0 / i_diam, 1.0 / i_axis, 1.0 / i_diam], (n, 1)),
        "shell": cylinder_shell(shell_points, r_out, r_in, length),
        "shape": (2, r_out, length, r_out - r_in),
        "domain": (float(box_xy), float(box_xy), float(col)),
        "base_radius": r_out,
    }


def cylinder_pack(n, box_xy, seed, pitch=4.1e-3, r_out=2e-3, r_in=1.2e-3, length=4e-3,
                  shell_points=256, rho=2500.0):
    """Hollow cylinders packed on a jittered cubic lattice of the given pitch (smaller than the
    bounding diameter, so neighbours interpenetrate their bounding spheres and most touch) over
    the floor of a periodic ``box_xy`` box, random orientations, at rest: the contact-rich state
    of a settled bed.  Same dict as :func:`cylinder_bed`."""
    rb = math.sqrt(r_out ** 2 + (0.5 * length) ** 2)
    side = max(1, int(box_xy / pitch))
    rng = np.random.default_rng(seed)
    k = np.arange(n)
    iz, rem = np.divmod(k, side * side)
    iy, ix = np.divmod(rem, side)
    pos = np.stack([(ix + 0.5) * pitch, (iy + 0.5) * pitch, rb + iz * pitch], 1)
    pos += rng.uniform(-0.05, 0.05, pos.shape) * pitch
    pos[:, 2] = np.maximum(pos[:, 2], rb)
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    vol = math.pi * (r_out ** 2 - r_in ** 2) * length
    m = rho * vol
    i_axis = 0.5 * m * (r_out ** 2 + r_in ** 2)
    i_diam = m * (3.0 * (r_out ** 2 + r_in ** 2) + length ** 2) / 12.0
    L = side * pitch
    return {
        "pos": pos,
        "quat": q,
        "inv_mass": np.full(n, 1.0 / m),
        "inv_inertia": np.tile([1.0 / i_diam, 1.0 / i_axis, 1.0 / i_diam], (n, 1)),
        "shell": cylinder_shell(shell_points, r_out, r_in, length),
        "shape": (2, r_out, length, r_out - r_in),
        "domain": (float(L), float(L), float(pos[:, 2].max() + 4 * rb)),
        "base_radius": r_out,
    }


def setup_cylinder_bed(sim, bed, dt, mu, e, iters=(12, 8)):
    sim.initialize_shape(*bed["shape"])
    sim.set_shell_points(bed["shell"])
    lx, ly, lz = bed["domain"]
    sim.set_domain(lx, ly, lz, True, True, False)
    sim.add_plane(0.0, 0.0, 0.0, 0.0, 0.0, 1.0)
    sim.set_positions(bed["pos"])
    sim.set_quaternions(bed["quat"])
    sim.set_inv_mass(bed["inv_mass"])
    sim.set_inv_inertia(bed["inv_inertia"])
    sim.set_gravity(0.0, 0.0, -9.81)
    sim.set_material_params(e, 0.0, mu)
    sim.set_solver_iterations(*iters)
    sim.set_dt(dt)
    return sim
