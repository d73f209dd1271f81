"""ctypes binding to a peclet-dem C-ABI shared library.

The product library (``libpecletdem_hip.so``, prefix ``pd_``) exports the functions declared in
``include/peclet_dem.h``.  The binding is written against that symbol table; the prefix is a
constructor argument only so a test can bind a second library that exports the same table.
"""
from __future__ import annotations

import ctypes as C
import os

_i = C.c_int
_d = C.c_double
_p = C.c_void_p
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_u64p = C.POINTER(C.c_uint64)


class Stats(C.Structure):
    """Per-substep counters (mirror of ``pd_stats``)."""

    _fields_ = [
        ("num_real", C.c_int),
        ("num_particles", C.c_int),
        ("num_pairs", C.c_int),
        ("num_contacts", C.c_int),
        ("num_manifolds", C.c_int),
        ("active_manifolds", C.c_int),
        ("num_colors_vel", C.c_int),
        ("num_colors_pos", C.c_int),
        ("rounds_vel", C.c_int),
        ("rounds_pos", C.c_int),
        ("sweeps_vel", C.c_int),
        ("sweeps_stab", C.c_int),
        ("sweeps_pos", C.c_int),
        ("stab_ran", C.c_int),
        ("uncolored_vel", C.c_longlong),
        ("uncolored_pos", C.c_longlong),
        ("max_overlap", C.c_double),
        ("max_approach", C.c_double),
        ("ml_levels", C.c_int),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


# name -> (restype, argtypes); every function also exists with the library's prefix.
SIGNATURES = {
    "create": (_p, [_i]),
    "destroy": (None, [_p]),
    "last_error": (C.c_char_p, []),
    "initialize_shape": (_i, [_p, _i, _d, _d, _d]),
    "set_shell_points": (_i, [_p, _dp, _i]),
    "get_shell_points": (_i, [_p, _dp, _i]),
    "set_domain": (_i, [_p, _d, _d, _d, _i, _i, _i]),
    "set_domain_minmax": (_i, [_p, _d, _d, _d, _d, _d, _d]),
    "enable_periodicity": (_i, [_p, _i, _i, _i]),
    "get_domain": (_i, [_p, _dp, _dp]),
    "set_gravity": (_i, [_p, _d, _d, _d]),
    "set_thermostat": (_i, [_p, _d, _d, _d]),
    "set_solver_iterations": (_i, [_p, _i, _i]),
    "set_velocity_use_gs": (_i, [_p, _i]),
    "set_stabilization_mode": (_i, [_p, _i]),
    "set_global_scale": (_i, [_p, _d]),
    "set_dt": (_i, [_p, _d]),
    "set_material_params": (_i, [_p, _d, _d, _d]),
    "set_material_ids": (_i, [_p, _ip, _i]),
    "set_pair_material": (_i, [_p, _i, _i, _d, _d]),
    "add_plane": (_i, [_p, _d, _d, _d, _d, _d, _d]),
    "set_positions": (_i, [_p, _dp, _i]),
    "set_velocities": (_i, [_p, _dp, _i]),
    "set_external_forces": (_i, [_p, _dp, _i]),
    "clear_external_forces": (_i, [_p]),
    "set_quaternions": (_i, [_p, _dp, _i]),
    "set_angular_velocities": (_i, [_p, _dp, _i]),
    "set_inv_inertia": (_i, [_p, _dp, _i]),
    "set_inv_mass": (_i, [_p, _dp, _i]),
    "set_scales": (_i, [_p, _dp, _i]),
    "set_scales_uniform": (_i, [_p, _d]),
    "set_growth_params": (_i, [_p, _d, _d]),
    "get_growth_factor": (_d, [_p]),
    "get_growth_rate": (_d, [_p]),
    "num_particles": (_i, [_p]),
    "get_positions": (_i, [_p, _dp, _i]),
    "get_velocities": (_i, [_p, _dp, _i]),
    "get_quaternions": (_i, [_p, _dp, _i]),
    "get_angular_velocities": (_i, [_p, _dp, _i]),
    "get_inv_inertia": (_i, [_p, _dp, _i]),
    "get_inv_mass": (_i, [_p, _dp, _i]),
    "get_scales": (_i, [_p, _dp, _i]),
    "step": (_i, [_p, _d]),
    "compute_overlaps": (_i, [_p, _dp]),
    "num_contacts": (_i, [_p]),
    "num_manifolds": (_i, [_p]),
    "max_overlap": (_d, [_p]),
    "get_stats": (_i, [_p, C.POINTER(Stats)]),
    "get_pairs": (_i, [_p, _ip, _i]),
    "get_contacts": (_i, [_p, _ip, _dp, _i]),
    "get_contact_colors": (_i, [_p, _ip, _i]),
    "get_manifolds": (_i, [_p, _ip, _dp, _i]),
    "get_manifold_colors": (_i, [_p, _ip, _i]),
    "get_ledger": (_i, [_p, _u64p, _dp, _dp, _dp, _i]),
    "get_grounded": (_i, [_p, _ip, _i]),
    "add_sdf_wall": (_i, [_p, _dp, _i, _i, _i, _d, _d, _d, _d, _d, _d, _d, _d]),
    "set_wall_velocity": (_i, [_p, _i, _d, _d, _d, _d, _d, _d, _d, _d, _d]),
    "set_wall_material_id": (_i, [_p, _i, _i]),
    "num_walls": (_i, [_p]),
    "set_hertz_material": (_i, [_p, _i, _d, _d]),
    "step_hertz": (_i, [_p, _d, _i, _d]),
    "get_hertz_pairs": (_i, [_p, _ip, _dp, _i]),
    "set_sdf_shape": (_i, [_p, _dp, _i, _i, _i, _d, _d, _d, _d, _d, _d, _dp, _i, _d, _d, _d, _d]),
    "get_sdf_grid": (_i, [_p, _i, _i, _i, _dp]),
}

# Product-only entry points (no counterpart in a restatement library).
PRODUCT_SIGNATURES = {
    "autotune": (_i, [_p, _d, _i, _i, C.POINTER(C.c_long), _dp]),
    "batch_autotune": (_i, [_p, _d, _i]),
    "get_sdf_grid_noalloc": (_i, [_p, _i, _i, _i, _p]),
    "step_n": (_i, [_p, _d, _i]),
    "synchronize": (_i, [_p]),
    "execution_space": (C.c_char_p, []),
    "export_lammps": (_i, [_p, C.c_char_p, _i]),
    "write_vtp": (_i, [_p, C.c_char_p]),
    "write_lammps_dump": (_i, [C.c_char_p, _i, _dp, _dp, _dp, _dp, _i, _dp, _dp, _i]),
    "get_device_pointer": (_p, [_p, C.c_char_p]),
    "set_timing": (_i, [_p, _i]),
    "get_stage_times": (_i, [_p, _dp, _i]),
    "get_stage_host_times": (_i, [_p, _dp, _i]),
    "stage_name": (C.c_char_p, [_i]),
    "get_kernel_times": (_i, [_p, _dp, C.POINTER(C.c_long), C.POINTER(C.c_long), _i]),
    "get_stats_log": (_i, [_p, C.POINTER(Stats), _i]),
    "kernel_name": (C.c_char_p, [_i]),
    "set_tuning": (_i, [_p, C.c_char_p, C.c_long]),
    "export_sdf": (_i, [_p, C.c_char_p, _i, _i, _i]),
    "read_device": (_i, [_p, _p, _p, C.c_size_t]),
    "write_device": (_i, [_p, _p, _p, C.c_size_t]),
    "batch_create": (_p, [_i, _ip]),
    "batch_destroy": (None, [_p]),
    "batch_size": (_i, [_p]),
    "batch_system": (_p, [_p, _i]),
    "batch_set_workers": (_i, [_p, _i]),
    "batch_step": (_i, [_p, _d, _i]),
    "rsa_place": (C.c_long, [_dp, C.c_long, _dp, _i, _i, _dp, _dp, _ip, _dp, _ip]),
    "tape_begin": (_i, [_p, _i]),
    "tape_end": (_i, [_p]),
    "tape_length": (_i, [_p]),
    "tape_generation": (C.c_long, [_p]),
    "tape_replay": (_i, [_p, _dp, C.c_long]),
    "tape_tangent": (_i, [_p, _dp, _d, _d, _dp, _dp, C.c_long]),
    "tape_adjoint": (_i, [_p, _dp, _dp, _dp, C.c_long]),
}


class Binding:
    """Resolved function table of one C-ABI library."""

    def __init__(self, path: str, prefix: str, product: bool):
        if not os.path.exists(path):
            raise RuntimeError(f"peclet.dem: shared library not found: {path}")
        self.path = path
        self.lib = C.CDLL(path)
        self.prefix = prefix
        table = dict(SIGNATURES)
        if product:
            table.update(PRODUCT_SIGNATURES)
        for name, (res, args) in table.items():
            fn = getattr(self.lib, prefix + name)
            fn.restype = res
            fn.argtypes = args
            setattr(self, name, fn)

    def error(self) -> str:
        msg = self.last_error()
        return msg.decode() if msg else "unknown error"
