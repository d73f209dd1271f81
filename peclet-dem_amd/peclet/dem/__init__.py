"""peclet.dem — MI355X-native XPBD discrete-element substep behind the reference's Python API.

``Simulation`` keeps the method names, argument names and defaults of the reference nanobind class
(src/dem_bindings.cpp:72-387) and drives the C-ABI library ``libpecletdem_hip.so`` (declared in
``include/peclet_dem.h``), whose substep runs as hand-written HIP kernels on the GPU.  State is
FP64; array inputs may be float32 or float64 and getters return float64 NumPy arrays.

There is no CPU fallback: importing works without a GPU, but constructing a ``Simulation`` loads
the HIP library and fails loudly if it or the GPU is missing.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from ._binding import Binding, Stats
from .walls import WallSDF, build_wall_sdf

__version__ = "0.3.2"
__all__ = ["Simulation", "Batch", "execution_space", "finalize", "export_lammps", "library_path",
           "build_wall_sdf", "WallSDF"]

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_NAME = "libpecletdem_hip.so"
_binding_cache: Binding | None = None

# reference src/dem_portable.hpp:77
SHAPE_GRID_SDF, SPHERE, HOLLOW_CYLINDER, BOX = 0, 1, 2, 3
_STAB_MODES = {"off": 0, "onesided": 1, "multilevel": 2, "escalate": 3, "ordered": 4}


def library_path() -> str:
    """Location of the HIP C-ABI library (``PECLET_DEM_LIB`` overrides)."""
    return os.environ.get("PECLET_DEM_LIB", os.path.join(_HERE, _LIB_NAME))


def _product_binding() -> Binding:
    global _binding_cache
    if _binding_cache is None:
        path = library_path()
        if not os.path.exists(path):
            raise RuntimeError(
                f"peclet.dem: the HIP library {path} is not built; run __graft_entry__.build() "
                "(there is no CPU fallback)")
        _binding_cache = Binding(path, "pd_", product=True)
    return _binding_cache


def __getattr__(name):
    if name == "execution_space":
        try:
            return _product_binding().execution_space().decode()
        except RuntimeError:
            return "HIP (library not built)"
    raise AttributeError(name)


class _DeviceArray:
    """A strided FP64 device array owned by a Simulation.

    It carries ``__cuda_array_interface__`` (pointer, shape, byte strides) for consumers that
    share this library's HIP runtime, and copies to / from host NumPy arrays through the
    library itself (``numpy()``, ``assign()``), ordered on the simulation's stream.
    """

    def __init__(self, ptr, shape, strides, owner):
        self._owner = owner
        self.ptr = int(ptr)
        self.shape = tuple(shape)
        self.strides = tuple(strides)
        self.__cuda_array_interface__ = {"shape": self.shape, "strides": self.strides,
                                         "typestr": "<f8", "data": (self.ptr, False),
                                         "version": 2}

    def _block(self):
        n = self.shape[0]
        row = self.strides[0] // 8
        return n, row, np.empty(n * row, np.float64)

    def numpy(self):
        n, row, buf = self._block()
        if n:
            self._owner._chk(self._owner._b.read_device(
                self._owner._h, C.c_void_p(self.ptr), buf.ctypes.data_as(C.c_void_p), buf.nbytes))
        cols = self.shape[1] if len(self.shape) > 1 else 1
        out = buf.reshape(n, row)[:, :cols]
        return out.copy() if len(self.shape) > 1 else out[:, 0].copy()

    def assign(self, values):
        """Write ``values`` (same shape) into the aliased state."""
        n, row, buf = self._block()
        if not n:
            return
        self._owner._chk(self._owner._b.read_device(
            self._owner._h, C.c_void_p(self.ptr), buf.ctypes.data_as(C.c_void_p), buf.nbytes))
        cols = self.shape[1] if len(self.shape) > 1 else 1
        blk = buf.reshape(n, row)
        blk[:, :cols] = np.asarray(values, np.float64).reshape(n, cols)
        self._owner._chk(self._owner._b.write_device(
            self._owner._h, C.c_void_p(self.ptr), buf.ctypes.data_as(C.c_void_p), buf.nbytes))


class Batch:
    """Independent systems advanced together (no reference counterpart).

    ``systems[k]`` is a full :class:`Simulation` view of system ``k``; configure each through
    the usual API, then :meth:`step` advances every system by the same substeps.  Each system
    runs on its own HIP stream and a pool of host workers keeps many of them in flight, so a
    batch of small systems fills the GPU where one system cannot.  Every system's result is
    the same as stepping it alone.
    """

    _live: "set[Batch]" = set()

    def __init__(self, capacities, workers: int = 16):
        self._b = _product_binding()
        caps = np.ascontiguousarray(np.asarray(capacities, dtype=np.int32))
        h = self._b.batch_create(int(caps.size), caps.ctypes.data_as(C.POINTER(C.c_int)))
        if not h:
            raise RuntimeError(self._b.error())
        self._h = C.c_void_p(h)
        self.systems = [Simulation(0, _binding=self._b,
                                   _handle=self._b.batch_system(self._h, k))
                        for k in range(int(caps.size))]
        self.set_workers(workers)
        Batch._live.add(self)

    def __len__(self):
        return len(self.systems)

    def set_workers(self, workers: int):
        rc = self._b.batch_set_workers(self._h, int(workers))
        if rc:
            raise ValueError(self._b.error())

    def step(self, dt, substeps: int = 1):
        if self._b.batch_step(self._h, float(dt), int(substeps)) != 0:
            raise RuntimeError(self._b.error())

    def autotune(self, dt, substeps: int = 4):
        """Autotune every system (see :meth:`Simulation.autotune`): the largest system of each
        size class is tuned by stepping it, the others take the cached choice."""
        rc = self._b.batch_autotune(self._h, float(dt), int(substeps))
        if rc != 0:
            raise (ValueError if rc == -2 else RuntimeError)(self._b.error())

    def close(self):
        if getattr(self, "_h", None):
            for s in self.systems:
                s._h = None
            self._b.batch_destroy(self._h)
            self._h = None
        Batch._live.discard(self)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def finalize() -> None:
    """Release every live Simulation (reference module ``finalize``, src/dem_bindings.cpp:61-69)."""
    for sim in list(Simulation._live):
        sim.close()
    for batch in list(Batch._live):
        batch.close()


def _f64(a, cols=None) -> np.ndarray:
    arr = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    if cols is not None:
        arr = arr.reshape(-1, cols) if arr.ndim != 2 else arr
    return arr


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Simulation:
    """One particle system.  Mirrors ``peclet::dem::Simulation`` (src/sim.hpp:909-1670)."""

    _live: "set[Simulation]" = set()

    def __init__(self, capacity: int, _binding: Binding | None = None, _handle=None):
        self._b = _binding if _binding is not None else _product_binding()
        self._owned = _handle is None
        h = self._b.create(int(capacity)) if _handle is None else _handle
        if not h:
            raise RuntimeError(self._b.error())
        self._h = C.c_void_p(h)
        if self._owned:
            Simulation._live.add(self)

    # ------------------------------------------------------------------ plumbing
    def _chk(self, rc: int):
        if rc == -2:
            raise ValueError(self._b.error())
        if rc != 0:
            raise RuntimeError(self._b.error())
        return rc

    def close(self):
        if getattr(self, "_h", None):
            if getattr(self, "_owned", True):
                self._b.destroy(self._h)
            self._h = None
        Simulation._live.discard(self)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _get(self, fn, cols):
        n = self._b.num_particles(self._h)
        out = np.empty((n, cols) if cols > 1 else (n,), np.float64)
        fn(self._h, _ptr(out), n)
        return out

    # ------------------------------------------------------------------ shape
    def set_sphere_shape(self, radius):
        self._chk(self._b.initialize_shape(self._h, SPHERE, float(radius), 0.0, 0.0))

    def initialize_shape(self, shape_type, radius, height=0.0, thickness=0.0):
        self._chk(self._b.initialize_shape(self._h, int(shape_type), float(radius), float(height),
                                           float(thickness)))

    def initialize(self, shape_type, radius=0.5, height=2.0, thickness=0.0):
        self.initialize_shape(shape_type, radius, height, thickness)

    def set_shell_points(self, points):
        """Extension: replace the surface point shell of the current (non-sphere) shape."""
        pts = _f64(points, 3)
        self._chk(self._b.set_shell_points(self._h, _ptr(pts), pts.shape[0]))

    def get_shell_points(self):
        n = self._b.get_shell_points(self._h, None, 0)
        out = np.empty((n, 3), np.float64)
        self._b.get_shell_points(self._h, _ptr(out), n)
        return out

    def set_sdf_shape(self, grid, nx, ny, nz, origin, spacing, shell, inv_inertia,
                      bounding_radius):
        """Shape 0 becomes an imported grid SDF (canonical space, x-fastest samples, negative
        inside) probed by the (N,3) surface point shell."""
        g = np.ascontiguousarray(np.asarray(grid, dtype=np.float64).ravel())
        if g.size != int(nx) * int(ny) * int(nz):
            raise RuntimeError("setSdfShape: grid.size() must equal nx*ny*nz")
        sh = _f64(shell, 3)
        self._chk(self._b.set_sdf_shape(self._h, _ptr(g), int(nx), int(ny), int(nz),
                                        *map(float, origin), *map(float, spacing), _ptr(sh),
                                        int(sh.shape[0]), *map(float, inv_inertia),
                                        float(bounding_radius)))

    # ------------------------------------------------------------------ domain
    def set_domain(self, *args, **kw):
        """``set_domain(lx, ly, lz, px=True, py=True, pz=False)`` or ``set_domain(min, max)``."""
        if (len(args) == 2 and not kw) or ("min" in kw):
            mn = args[0] if args else kw["min"]
            mx = args[1] if len(args) > 1 else kw["max"]
            self._chk(self._b.set_domain_minmax(self._h, *map(float, mn), *map(float, mx)))
            return
        names = ["lx", "ly", "lz", "px", "py", "pz"]
        vals = dict(zip(names, args))
        vals.update(kw)
        self._chk(self._b.set_domain(self._h, float(vals["lx"]), float(vals["ly"]),
                                     float(vals["lz"]), int(bool(vals.get("px", True))),
                                     int(bool(vals.get("py", True))),
                                     int(bool(vals.get("pz", False)))))

    def enable_periodicity(self, x, y, z):
        self._chk(self._b.enable_periodicity(self._h, int(bool(x)), int(bool(y)), int(bool(z))))

    def _domain(self):
        mn = np.zeros(3)
        mx = np.zeros(3)
        self._b.get_domain(self._h, _ptr(mn), _ptr(mx))
        return mn, mx

    def get_domain_min(self):
        return tuple(float(v) for v in self._domain()[0])

    def get_domain_max(self):
        return tuple(float(v) for v in self._domain()[1])

    # ------------------------------------------------------------------ parameters
    def set_gravity(self, gx, gy, gz):
        self._chk(self._b.set_gravity(self._h, float(gx), float(gy), float(gz)))

    def set_thermostat(self, temperature, tau, kB=1.0):
        self._chk(self._b.set_thermostat(self._h, float(temperature), float(tau), float(kB)))

    def set_solver_iterations(self, pos, vel):
        self._chk(self._b.set_solver_iterations(self._h, int(pos), int(vel)))

    def set_velocity_use_gs(self, use_gs):
        self._chk(self._b.set_velocity_use_gs(self._h, int(bool(use_gs))))

    def set_stabilization(self, enabled):
        self._chk(self._b.set_stabilization_mode(self._h, 1 if enabled else 0))

    def set_stabilization_mode(self, mode):
        if mode not in _STAB_MODES:
            raise ValueError("set_stabilization_mode: expected 'off', 'onesided', 'multilevel', "
                             "'escalate' or 'ordered'")
        self._chk(self._b.set_stabilization_mode(self._h, _STAB_MODES[mode]))

    def set_global_scale(self, s):
        self._chk(self._b.set_global_scale(self._h, float(s)))

    def set_dt(self, dt):
        self._chk(self._b.set_dt(self._h, float(dt)))

    def set_material_params(self, restitution_normal, restitution_tangent=0.0, friction=0.0):
        self._chk(self._b.set_material_params(self._h, float(restitution_normal),
                                              float(restitution_tangent), float(friction)))

    def set_material_ids(self, ids):
        a = np.ascontiguousarray(np.asarray(ids, dtype=np.int32).ravel())
        self._chk(self._b.set_material_ids(self._h, a.ctypes.data_as(C.POINTER(C.c_int)), a.size))

    def set_pair_material(self, a, b, restitution, friction):
        self._chk(self._b.set_pair_material(self._h, int(a), int(b), float(restitution),
                                            float(friction)))

    def set_hertz_material(self, mat, youngs, poisson):
        """Young's modulus and Poisson ratio of material ``mat`` (defaults 1e9, 0.25)."""
        self._chk(self._b.set_hertz_material(self._h, int(mat), float(youngs), float(poisson)))

    def step_hertz(self, dt, substeps=1, skin_frac=0.3):
        """``substeps`` explicit Hertz–Mindlin steps (spheres or point-shell shapes, SDF walls,
        non-periodic, no planes) on a Verlet pair list with skin ``skin_frac`` x min radius."""
        self._chk(self._b.step_hertz(self._h, float(dt), int(substeps), float(skin_frac)))

    def add_plane(self, *args, **kw):
        """``add_plane(px, py, pz, nx, ny, nz)`` or ``add_plane(point, normal)``."""
        if len(args) == 2 or "point" in kw:
            p = args[0] if args else kw["point"]
            n = args[1] if len(args) > 1 else kw["normal"]
            vals = [*map(float, p), *map(float, n)]
        else:
            vals = [float(v) for v in args]
        self._chk(self._b.add_plane(self._h, *vals))

    def add_sdf_wall(self, grid, nx, ny, nz, origin, spacing, restitution=0.0, friction=0.0):
        """Static world-space SDF wall: flat ``nx*ny*nz`` samples, x fastest, positive in the
        void; returns the wall index."""
        g = np.ascontiguousarray(np.asarray(grid, dtype=np.float64).ravel())
        if g.size != int(nx) * int(ny) * int(nz):
            raise RuntimeError("addSdfWall: grid.size() must equal nx*ny*nz")
        rc = self._b.add_sdf_wall(self._h, _ptr(g), int(nx), int(ny), int(nz),
                                  *map(float, origin), *map(float, spacing), float(restitution),
                                  float(friction))
        if rc < 0:
            self._chk(rc)
        return rc

    def set_wall_velocity(self, wall_index, lin_vel=(0.0, 0.0, 0.0), ang_vel=(0.0, 0.0, 0.0),
                          center=(0.0, 0.0, 0.0)):
        self._chk(self._b.set_wall_velocity(self._h, int(wall_index), *map(float, lin_vel),
                                            *map(float, ang_vel), *map(float, center)))

    def set_wall_material_id(self, wid, mat):
        self._chk(self._b.set_wall_material_id(self._h, int(wid), int(mat)))

    # ------------------------------------------------------------------ particle arrays
    def set_positions(self, a):
        """(N,3) positions, or (N,4) with column 3 = inverse mass (0 -> 1), as the reference."""
        arr = np.asarray(a)
        if arr.ndim == 2 and arr.shape[1] in (3, 4):
            xyz = _f64(arr[:, :3])
            self._chk(self._b.set_positions(self._h, _ptr(xyz), xyz.shape[0]))
            if arr.shape[1] == 4:
                w = np.asarray(arr[:, 3], np.float64)
                im = np.ascontiguousarray(np.where(w == 0.0, 1.0, w))
                self._chk(self._b.set_inv_mass(self._h, _ptr(im), im.shape[0]))
        else:
            xyz = _f64(arr).ravel()
            self._chk(self._b.set_positions(self._h, _ptr(xyz), xyz.size // 3))

    def _set3(self, fn, a, cols=3):
        arr = _f64(a).ravel()
        self._chk(fn(self._h, _ptr(arr), arr.size // cols))

    def set_velocities(self, a):
        self._set3(self._b.set_velocities, a)

    def set_external_forces(self, a):
        self._set3(self._b.set_external_forces, a)

    def clear_external_forces(self):
        self._chk(self._b.clear_external_forces(self._h))

    def set_quaternions(self, a):
        self._set3(self._b.set_quaternions, a, 4)

    def set_angular_velocities(self, a):
        self._set3(self._b.set_angular_velocities, a)

    def set_inv_inertia(self, a):
        self._set3(self._b.set_inv_inertia, a)

    def set_inv_mass(self, a):
        self._set3(self._b.set_inv_mass, a, 1)

    def set_scales_uniform(self, s):
        self._chk(self._b.set_scales_uniform(self._h, float(s)))

    def set_scales(self, a):
        self._set3(self._b.set_scales, a, 1)

    def set_growth_params(self, rate, new_factor=-1.0):
        self._chk(self._b.set_growth_params(self._h, float(rate), float(new_factor)))

    def get_growth_factor(self):
        return float(self._b.get_growth_factor(self._h))

    def get_growth_rate(self):
        return float(self._b.get_growth_rate(self._h))

    def get_positions(self):
        return self._get(self._b.get_positions, 3)

    def get_velocities(self):
        return self._get(self._b.get_velocities, 3)

    def get_quaternions(self):
        return self._get(self._b.get_quaternions, 4)

    def get_angular_velocities(self):
        return self._get(self._b.get_angular_velocities, 3)

    def get_inv_inertia(self):
        return self._get(self._b.get_inv_inertia, 3)

    def get_inv_mass(self):
        return self._get(self._b.get_inv_mass, 1)

    def get_scales(self):
        return self._get(self._b.get_scales, 1)

    def get_masses(self):
        im = self.get_inv_mass()
        with np.errstate(divide="ignore"):
            return np.where(im > 0.0, 1.0 / np.where(im > 0.0, im, 1.0), 0.0)

    # ------------------------------------------------------------------ stepping
    def step(self, dt=0.0):
        """One XPBD substep; like the reference, ``dt`` becomes the time step (src/sim.hpp:1456)."""
        self._chk(self._b.step(self._h, float(dt)))

    def step_n(self, dt, n):
        """Extension: ``n`` substeps of size ``dt`` with no host round trip between them."""
        self._chk(self._b.step_n(self._h, float(dt), int(n)))

    # ------------------------------------------------------------------ differentiation
    # Extension (no counterpart in the reference): the state of a body is 13 numbers, position,
    # velocity, angular velocity and orientation (x, y, z, w); see ``peclet.dem.adjoint``.
    def tape_begin(self, checkpoint_every=16):
        """Record the following substeps for ``tape_tangent`` / ``tape_adjoint``."""
        self._chk(self._b.tape_begin(self._h, int(checkpoint_every)))

    def tape_end(self):
        self._chk(self._b.tape_end(self._h))

    def tape_length(self):
        return int(self._b.tape_length(self._h))

    def tape_generation(self):
        """Number of the current recording (one more at every ``tape_begin``; 0 before)."""
        return int(self._b.tape_generation(self._h))

    def _state_array(self, a, what):
        n = self.num_particles()
        a = _f64(a)
        if a.shape != (n, 13):
            raise ValueError(f"{what} must have shape ({n}, 13), got {a.shape}")
        return a

    def tape_replay(self):
        """End state (n, 13) of the recorded substeps replayed from the first checkpoint."""
        out = np.empty((self.num_particles(), 13), np.float64)
        self._chk(self._b.tape_replay(self._h, _ptr(out), out.size))
        return out

    def tape_tangent(self, dstate0=None, dmu=0.0, drest=0.0):
        """Forward mode: (end state, its derivative along dstate0 (n, 13), dmu, drest)."""
        n = self.num_particles()
        d0 = None if dstate0 is None else self._state_array(dstate0, "dstate0")
        val = np.empty((n, 13), np.float64)
        tan = np.empty((n, 13), np.float64)
        self._chk(self._b.tape_tangent(self._h, None if d0 is None else _ptr(d0), float(dmu),
                                       float(drest), _ptr(val), _ptr(tan), val.size))
        return val, tan

    def tape_adjoint(self, grad_state_end):
        """Reverse mode: (dL/d start state (n, 13), dL/d (friction, restitution)) for
        grad_state_end = dL/d end state (n, 13)."""
        g = self._state_array(grad_state_end, "grad_state_end")
        g0 = np.empty_like(g)
        gp = np.zeros(2, np.float64)
        self._chk(self._b.tape_adjoint(self._h, _ptr(g), _ptr(g0), _ptr(gp), g.size))
        return g0, gp

    def synchronize(self):
        """Wait for the simulation's queued GPU work."""
        self._chk(self._b.synchronize(self._h))

    def compute_overlaps(self):
        out = C.c_double(0.0)
        self._chk(self._b.compute_overlaps(self._h, C.byref(out)))
        return out.value

    def num_particles(self):
        return int(self._b.num_particles(self._h))

    def num_contacts(self):
        return int(self._b.num_contacts(self._h))

    def num_manifolds(self):
        return int(self._b.num_manifolds(self._h))

    def max_overlap(self):
        return float(self._b.max_overlap(self._h))

    get_num_contacts = num_contacts
    get_num_manifolds = num_manifolds
    get_max_overlap = max_overlap

    def get_profiling_info(self):
        return {"num_particles": self.num_particles(), "num_contacts": self.num_contacts(),
                "num_manifolds": self.num_manifolds(), "max_overlap": self.max_overlap()}

    def get_stats(self) -> dict:
        """Extension: counters of the last substep (pairs, contacts, colours, sweeps, ...)."""
        st = Stats()
        self._chk(self._b.get_stats(self._h, C.byref(st)))
        return st.as_dict()

    # ------------------------------------------------------------------ diagnostics (last substep)
    def debug_pairs(self):
        n = self._b.get_pairs(self._h, None, 0)
        out = np.empty((n, 2), np.int32)
        self._b.get_pairs(self._h, out.ctypes.data_as(C.POINTER(C.c_int)), n)
        return out

    def debug_contacts(self):
        n = self._b.get_contacts(self._h, None, None, 0)
        ids = np.empty((n, 2), np.int32)
        data = np.empty((n, 10), np.float64)
        self._b.get_contacts(self._h, ids.ctypes.data_as(C.POINTER(C.c_int)), _ptr(data), n)
        return ids, data

    def debug_contact_colors(self):
        n = self._b.get_contact_colors(self._h, None, 0)
        out = np.empty(n, np.int32)
        self._b.get_contact_colors(self._h, out.ctypes.data_as(C.POINTER(C.c_int)), n)
        return out

    def debug_manifolds(self):
        n = self._b.get_manifolds(self._h, None, None, 0)
        ids = np.empty((n, 5), np.int32)
        data = np.empty((n, 15), np.float64)
        self._b.get_manifolds(self._h, ids.ctypes.data_as(C.POINTER(C.c_int)), _ptr(data), n)
        return ids, data

    def debug_manifold_colors(self):
        n = self._b.get_manifold_colors(self._h, None, 0)
        out = np.empty(n, np.int32)
        self._b.get_manifold_colors(self._h, out.ctypes.data_as(C.POINTER(C.c_int)), n)
        return out

    def debug_ledger(self):
        n = self._b.get_ledger(self._h, None, None, None, None, 0)
        keys = np.empty(n, np.uint64)
        lam = np.empty(n, np.float64)
        lamT = np.empty((n, 3), np.float64)
        pimp = np.empty(n, np.float64)
        self._b.get_ledger(self._h, keys.ctypes.data_as(C.POINTER(C.c_uint64)), _ptr(lam),
                           _ptr(lamT), _ptr(pimp), n)
        return keys, lam, lamT, pimp

    def autotune(self, dt, substeps=4, use_cache=True):
        """Choose the performance knobs (candidate_skin_permille, block_solver_rows,
        velocity_loop_lds, vsweep_waves) by timing this simulation's own substeps; the simulation advances by
        the trial substeps (results do not depend on the knobs).  Returns the chosen setting and
        the best trial's ms per substep (None when taken from the size-class cache)."""
        chosen = (C.c_long * 4)()
        ms = C.c_double(0.0)
        self._chk(self._b.autotune(self._h, float(dt), int(substeps), 1 if use_cache else 0,
                                   chosen, C.byref(ms)))
        return ({"candidate_skin_permille": int(chosen[0]), "block_solver_rows": int(chosen[1]),
                 "velocity_loop_lds": int(chosen[2]), "vsweep_waves": int(chosen[3])},
                None if ms.value < 0 else ms.value)

    def set_tuning(self, name, value):
        """Launch-shape knob of the HIP library (extension); results do not depend on it."""
        self._chk(self._b.set_tuning(self._h, name.encode(), int(value)))

    def debug_hertz_pairs(self):
        n = self._b.get_hertz_pairs(self._h, None, None, 0)
        pairs = np.empty((n, 2), np.int32)
        xi = np.empty((n, 3), np.float64)
        self._b.get_hertz_pairs(self._h, pairs.ctypes.data_as(C.POINTER(C.c_int)), _ptr(xi), n)
        return pairs, xi

    def debug_grounded(self):
        n = self._b.num_particles(self._h)
        out = np.empty(n, np.int32)
        self._b.get_grounded(self._h, out.ctypes.data_as(C.POINTER(C.c_int)), n)
        return out

    # ------------------------------------------------------------------ export (host side)
    def export_lammps(self, filename, step):
        self._chk(self._b.export_lammps(self._h, os.fsencode(filename), int(step)))

    def write_vtp(self, filename):
        self._chk(self._b.write_vtp(self._h, os.fsencode(filename)))

    # ------------------------------------------------------------------ zero-copy device views
    # (src/dem_bindings.cpp:302-335).  Strided FP64 device arrays that alias the state (the
    # reference's are float32); they stay valid until the particle count grows.
    def _view(self, name, cols, stride_doubles):
        ptr = self._b.get_device_pointer(self._h, name)
        if not ptr:
            raise RuntimeError(self._b.error())
        self._b.synchronize(self._h)
        n = self.num_particles()
        shape = (n, cols) if cols > 1 else (n,)
        strides = (8 * stride_doubles, 8) if cols > 1 else (8 * stride_doubles,)
        return _DeviceArray(ptr, shape, strides, self)

    def get_positions_view(self):
        return self._view(b"pos", 3, 4)

    def get_velocities_view(self):
        return self._view(b"vel", 3, 8)

    def get_external_forces_view(self):
        return self._view(b"ext_force", 3, 4)

    def get_inv_mass_view(self):
        return self._view(b"inv_mass", 1, 8)

    def get_sdf_grid(self, resolution):
        """Packed-bed SDF on a (rx, ry, rz) voxel grid over the domain (negative inside).  As in
        the reference binding, the x-fastest samples are returned as a C-ordered (rx, ry, rz)
        array."""
        rx, ry, rz = (int(v) for v in resolution)
        out = np.empty(rx * ry * rz, np.float64)
        self._chk(self._b.get_sdf_grid(self._h, rx, ry, rz, _ptr(out)))
        return out.reshape(rx, ry, rz)

    def get_sdf_grid_noalloc(self, resolution, out):
        """``get_sdf_grid`` without allocation (an extension for callers that cannot allocate):
        the x-fastest samples are written into ``out``, a caller-owned float64 CUDA tensor of
        rx*ry*rz elements on the simulation's GPU; at most 2,097,152 voxels.  Calls that overlap
        on one GPU share the library's working array and race."""
        rx, ry, rz = (int(v) for v in resolution)
        if out.dtype != __import__("torch").float64 or not out.is_cuda or not out.is_contiguous():
            raise ValueError("out must be a contiguous float64 CUDA tensor")
        if out.numel() != rx * ry * rz:
            raise ValueError(f"out must hold {rx * ry * rz} elements, has {out.numel()}")
        self._chk(self._b.get_sdf_grid_noalloc(self._h, rx, ry, rz, out.data_ptr()))
        return out

    def export_sdf(self, filename, resolution):
        rx, ry, rz = (int(v) for v in resolution)
        self._chk(self._b.export_sdf(self._h, os.fsencode(filename), rx, ry, rz))


def export_lammps(filename, step, pos, vel, quats, radii, box_min=None, box_max=None,
                  pbc_enabled=False):
    """Module-level LAMMPS dump writer (reference src/dem_bindings.cpp:436-464)."""
    b = _product_binding()
    p = _f64(pos).ravel()
    v = _f64(vel).ravel()
    q = _f64(quats).ravel()
    r = _f64(radii).ravel()
    mn = _f64(box_min).ravel() if box_min is not None else None
    mx = _f64(box_max).ravel() if box_max is not None else None
    has_box = int(mn is not None and mx is not None)
    rc = b.write_lammps_dump(os.fsencode(filename), int(step), _ptr(p), _ptr(v), _ptr(q), _ptr(r),
                             p.size // 3, _ptr(mn) if has_box else None,
                             _ptr(mx) if has_box else None, int(bool(pbc_enabled)))
    if rc != 0:
        raise RuntimeError(b.error())
