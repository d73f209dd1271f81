"""World-space SDF walls: sample a container distance function onto the lattice ``add_sdf_wall``
takes.  Same interface as the reference's ``build_wall_sdf`` / ``WallSDF``
(packaging/particle_builder.py:397-457): the callable is positive in the void where grains live
and negative inside the solid; the lattice is node-centred over ``bounds`` with ``resolution``
nodes per axis, stored x-fastest.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class WallSDF:
    grid: np.ndarray      # (nx, ny, nz) samples at origin + (i, j, k) * spacing
    origin: np.ndarray    # (3,)
    spacing: np.ndarray   # (3,)

    def add_to(self, sim, *, restitution: float = 0.0, friction: float = 0.0) -> int:
        nx, ny, nz = self.grid.shape
        flat = np.asarray(self.grid, dtype=np.float64).ravel(order="F")  # x fastest
        return sim.add_sdf_wall(flat, int(nx), int(ny), int(nz),
                                tuple(float(v) for v in self.origin),
                                tuple(float(v) for v in self.spacing),
                                float(restitution), float(friction))


def build_wall_sdf(f, bounds, resolution=64) -> WallSDF:
    if not callable(f):
        raise TypeError("f must be a callable f(points)->distance (positive in the void)")
    lo = np.asarray(bounds[0], float)
    hi = np.asarray(bounds[1], float)
    res = (int(resolution),) * 3 if np.isscalar(resolution) else tuple(int(r) for r in resolution)
    axes = [np.linspace(lo[d], hi[d], res[d]) for d in range(3)]
    X, Y, Z = np.meshgrid(*axes, indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    phi = np.asarray(f(pts), float).reshape(res)
    spacing = np.array([(hi[d] - lo[d]) / (res[d] - 1) for d in range(3)])
    return WallSDF(grid=phi, origin=lo.copy(), spacing=spacing)
