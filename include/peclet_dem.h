/*
 * peclet_dem.h — C-ABI of the MI355X-native peclet-dem substep library (libpecletdem_hip.so).
 *
 * Plain C types only (opaque handle, pointers, sizes).  Each entry point replaces one method of
 * the reference C++ facade peclet::dem::Simulation (src/sim.hpp) as bound to Python by
 * src/dem_bindings.cpp; the citation on each declaration names the reference interface it
 * replaces.  Arrays cross the boundary as host FP64 buffers, row-major (N,3) / (N,4) / (N,).
 *
 * Status codes: 0 = OK, -1 = runtime error, -2 = invalid argument (the reference throws
 * std::runtime_error / std::invalid_argument in those cases); pd_last_error() has the message.
 * Handles are not thread-safe; one handle owns one HIP stream on the current device.
 */
#ifndef PECLET_DEM_H
#define PECLET_DEM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pd_sim pd_sim;

/* Per-substep counters of the last pd_step (no reference counterpart: the reference exposes only
 * num_contacts / num_manifolds / max_overlap, src/dem_bindings.cpp:358-366). */
typedef struct pd_stats {
  int num_real;
  int num_particles;       /* real + periodic ghost slots */
  int num_pairs;           /* broadphase candidate pairs */
  int num_contacts;        /* narrowphase contacts */
  int num_manifolds;       /* contact manifolds */
  int active_manifolds;    /* manifolds solved by the velocity sweeps (colour >= 0) */
  int num_colors_vel;
  int num_colors_pos;
  int rounds_vel;          /* Jones-Plassmann rounds, manifold graph */
  int rounds_pos;          /* Jones-Plassmann rounds, contact graph */
  int sweeps_vel;          /* executed main velocity sweeps */
  int sweeps_stab;         /* executed stabilization sweeps */
  int sweeps_pos;          /* executed position sweeps */
  int stab_ran;
  long long uncolored_vel; /* sum over colouring rounds of uncoloured manifolds */
  long long uncolored_pos; /* sum over colouring rounds of uncoloured contacts */
  double max_overlap;
  double max_approach;
  int ml_levels;           /* multilevel stabilization: coarse levels built (0 otherwise) */
} pd_stats;

/* --- lifetime --------------------------------------------------------------------------- */
pd_sim* pd_create(int capacity);      /* Simulation(int capacity), src/sim.hpp:911 */
void pd_destroy(pd_sim* s);           /* ~Simulation / releaseViews, src/sim.hpp:917-936 */
const char* pd_last_error(void);      /* exception message of the failed call (thread-local) */
const char* pd_execution_space(void); /* module attr execution_space, src/dem_bindings.cpp:70 */
int pd_synchronize(pd_sim* s);        /* Kokkos::fence equivalent: wait for the handle's stream */

/* --- shape ------------------------------------------------------------------------------ */
/* initializeShape(shape_type, radius, height, thickness), src/sim.hpp:949-1021 */
int pd_initialize_shape(pd_sim* s, int shape_type, double radius, double height,
                        double thickness);
/* setSdfShape(grid, nx, ny, nz, origin, spacing, shell, invInertia, boundingRadius),
 * src/sim.hpp:1034-1084: shape 0 becomes an imported grid SDF (x-fastest samples, negative
 * inside) probed by the (nPts,3) point shell; boundingRadius is the canonical bound */
int pd_set_sdf_shape(pd_sim* s, const double* grid, int nx, int ny, int nz, double ox, double oy,
                     double oz, double sx, double sy, double sz, const double* shell, int nPts,
                     double iix, double iiy, double iiz, double boundingRadius);
/* Extension (no reference counterpart): replace the surface point shell of shape 0 (N,3). */
int pd_set_shell_points(pd_sim* s, const double* pts, int n);
int pd_get_shell_points(pd_sim* s, double* out, int cap);

/* --- domain and parameters -------------------------------------------------------------- */
int pd_set_domain(pd_sim* s, double lx, double ly, double lz, int px, int py,
                  int pz);                                     /* setDomain, src/sim.hpp:1086 */
int pd_set_domain_minmax(pd_sim* s, double mnx, double mny, double mnz, double mxx, double mxy,
                         double mxz);                          /* setDomainMinMax, :1091 */
int pd_enable_periodicity(pd_sim* s, int x, int y, int z);    /* enablePeriodicity, :1101 */
int pd_get_domain(pd_sim* s, double* mn3, double* mx3);       /* getDomainMin/Max, :1106-1111 */
int pd_set_gravity(pd_sim* s, double gx, double gy, double gz);          /* setGravity, :1112 */
int pd_set_thermostat(pd_sim* s, double temperature, double tau, double kB); /* :1113-1117 */
int pd_set_solver_iterations(pd_sim* s, int pos, int vel);   /* setSolverIterations, :1117 */
int pd_set_velocity_use_gs(pd_sim* s, int use_gs);           /* setVelocityUseGS, :1123 */
int pd_set_stabilization_mode(pd_sim* s, int mode);  /* setStabilizationMode, :1136-1152 */
int pd_set_global_scale(pd_sim* s, double gs);                /* setGlobalScale, :1171 */
int pd_set_dt(pd_sim* s, double dt);                          /* setDt, :1175 */
int pd_set_material_params(pd_sim* s, double e_n, double e_t, double mu); /* :1179-1183 */
int pd_set_material_ids(pd_sim* s, const int* ids, int n);    /* setMaterialIds, :1187 */
int pd_set_pair_material(pd_sim* s, int a, int b, double e, double mu); /* :1196-1215 */
int pd_add_plane(pd_sim* s, double px, double py, double pz, double nx, double ny,
                 double nz);                                  /* addPlane, :1226-1232 */

/* --- particle arrays (host FP64) -------------------------------------------------------- */
int pd_set_positions(pd_sim* s, const double* xyz, int n);    /* setPositions, :1290-1330 */
int pd_set_velocities(pd_sim* s, const double* v, int n);     /* setVelocities, :1348 */
int pd_set_external_forces(pd_sim* s, const double* f, int n); /* setExternalForces, :1358 */
int pd_clear_external_forces(pd_sim* s);                      /* clearExternalForces, :1367 */
int pd_set_quaternions(pd_sim* s, const double* q, int n);    /* setQuaternions, :1372 */
int pd_set_angular_velocities(pd_sim* s, const double* w, int n); /* :1383 */
int pd_set_inv_inertia(pd_sim* s, const double* ii, int n);   /* setInvInertia, :1392 */
int pd_set_inv_mass(pd_sim* s, const double* im, int n);      /* setInvMass, :1401 */
int pd_set_scales(pd_sim* s, const double* sc, int n);        /* setScales, :1340 */
int pd_set_scales_uniform(pd_sim* s, double sc);              /* setScalesUniform, :1332 */
int pd_set_growth_params(pd_sim* s, double rate, double new_factor); /* :1417-1425 */
double pd_get_growth_factor(pd_sim* s);                       /* growthFactor, :1426 */
double pd_get_growth_rate(pd_sim* s);                         /* getGrowthRate, :1427 */
int pd_num_particles(pd_sim* s);                              /* numParticles, :1542 */
int pd_get_positions(pd_sim* s, double* out, int n);          /* getPositions, :1440 */
int pd_get_velocities(pd_sim* s, double* out, int n);         /* getVelocities, :1444 */
int pd_get_quaternions(pd_sim* s, double* out, int n);        /* getQuaternions, :1448 */
int pd_get_angular_velocities(pd_sim* s, double* out, int n); /* :1407 */
int pd_get_inv_inertia(pd_sim* s, double* out, int n);        /* getInvInertia, :1411 */
int pd_get_inv_mass(pd_sim* s, double* out, int n);           /* (getMasses input), :1430 */
int pd_get_scales(pd_sim* s, double* out, int n);             /* getScales, :1452 */
/* Device pointer of a live FP64 state array, the zero-copy get_*_view exports of
 * src/dem_bindings.cpp:302-335.  Row strides in doubles: "pos" 4 (x, y, z, pad), "ext_force" 4,
 * "vel" 8 (the {v, w} record: v at 0, angular velocity at 4), "inv_mass" 8 (packed record). */
void* pd_get_device_pointer(pd_sim* s, const char* name);
/* Copies between caller host memory and device memory of this library's runtime, ordered on the
 * handle's stream (for consumers of pd_get_device_pointer that run no HIP code of their own). */
int pd_read_device(pd_sim* s, const void* src, void* host, size_t bytes);
int pd_write_device(pd_sim* s, void* dst, const void* host, size_t bytes);

/* --- the substep ------------------------------------------------------------------------ */
int pd_step(pd_sim* s, double dt);        /* Simulation::step -> demStep, src/sim.hpp:1456, :97 */
/* Extension: n substeps of size dt without returning to the caller in between. */
int pd_step_n(pd_sim* s, double dt, int n);
int pd_compute_overlaps(pd_sim* s, double* out); /* computeOverlapsKokkos, src/sim.hpp:494 */
int pd_num_contacts(pd_sim* s);                  /* numContacts, src/sim.hpp:1550 */
int pd_num_manifolds(pd_sim* s);                 /* numManifolds, src/sim.hpp:1551 */
double pd_max_overlap(pd_sim* s);                /* maxOverlap, src/sim.hpp:1552-1556 */
int pd_get_stats(pd_sim* s, pd_stats* out);

/* --- diagnostics of the last substep (count returned; NULL output = count only) --------- */
int pd_get_pairs(pd_sim* s, int* out, int cap);                      /* (i, j) pairs */
int pd_get_contacts(pd_sim* s, int* ids, double* data, int cap);     /* n, rA, rB, dist */
int pd_get_contact_colors(pd_sim* s, int* out, int cap);
int pd_get_manifolds(pd_sim* s, int* ids, double* data, int cap);    /* A,B,pts,realA,realB */
int pd_get_manifold_colors(pd_sim* s, int* out, int cap);
int pd_get_ledger(pd_sim* s, uint64_t* keys, double* lambda, double* lambdaT, double* posImp,
                  int cap);                                          /* warm-start store */
int pd_get_grounded(pd_sim* s, int* out, int n);

/* --- host-side export ------------------------------------------------------------------- */
int pd_export_lammps(pd_sim* s, const char* filename, int step); /* exportLammps, :1463 */
int pd_write_vtp(pd_sim* s, const char* filename);               /* writeVtp, :1560-1601 */
/* module-level export_lammps, src/dem_bindings.cpp:436-464 / writeLammpsDump, src/io.hpp:23 */
int pd_write_lammps_dump(const char* filename, int step, const double* pos, const double* vel,
                         const double* quat, const double* radii, int n, const double* box_min,
                         const double* box_max, int pbc);

/* --- workload utility (no reference counterpart) ------------------------------------- */
/* Random sequential addition of n spheres (radii[i]) from a caller-drawn candidate stream (3
 * doubles per candidate, inside the box); periodic[3] (may be null) marks axes on which [lo, hi)
 * is one period and overlaps are tested against the nearest image; see pd_sim.hip for the
 * contract. */
long pd_rsa_place(const double* cand, long ncand, const double* radii, int n, int start,
                  const double* lo, const double* hi, const int* periodic, double* out,
                  int* placed);

/* getSdfGrid(rx, ry, rz), src/sim.hpp:1538-1542 -> src/output_sdf.hpp: packed-bed SDF on an
 * (rx,ry,rz) voxel grid over the domain, x-fastest, negative inside (out: rx*ry*rz doubles) */
int pd_get_sdf_grid(pd_sim* s, int rx, int ry, int rz, double* out);
/* exportSdf(filename, rx, ry, rz), src/sim.hpp:1478-1483: the same grid as an ASCII VTI file */
int pd_export_sdf(pd_sim* s, const char* filename, int rx, int ry, int rz);
/* No counterpart in the reference: pd_get_sdf_grid for callers that cannot allocate (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc).  It allocates nothing: its
 * working arrays (17 bytes per voxel) live in a file-scope device array of the library, one per
 * GPU, held until the process ends, so rx*ry*rz may be at most 2,097,152 voxels (e.g. 128^3;
 * larger grids return -2 with an error naming the limit).  The grid (same values as
 * pd_get_sdf_grid) is written to out_device, caller-owned device memory of rx*ry*rz doubles on
 * the handle's GPU; the call returns once it is there.  Every call on one GPU shares that
 * array: calls that overlap on one GPU (from several threads or handles) race. */
int pd_get_sdf_grid_noalloc(pd_sim* s, int rx, int ry, int rz, double* out_device);

/* --- SDF walls and the Hertz–Mindlin engine --------------------------------------------- */
/* addSdfWall(grid, nx, ny, nz, origin, spacing, restitution, friction), src/sim.hpp:1242-1277:
 * flat x-fastest samples, positive in the void; returns the wall index (>= 0) or a status */
int pd_add_sdf_wall(pd_sim* s, const double* grid, int nx, int ny, int nz, double ox, double oy,
                    double oz, double sx, double sy, double sz, double restitution,
                    double friction);
/* setWallVelocity(wall, linVel, angVel, center), src/sim.hpp:1280-1287 */
int pd_set_wall_velocity(pd_sim* s, int wid, double lx, double ly, double lz, double ax,
                         double ay, double az, double cx, double cy, double cz);
int pd_set_wall_material_id(pd_sim* s, int wid, int mat); /* setWallMaterialId, :1217-1224 */
int pd_num_walls(pd_sim* s);
int pd_set_hertz_material(pd_sim* s, int mat, double youngs, double poisson); /* :1153-1165 */
/* stepHertz(dt, substeps, skin_frac) -> demStepHertz, src/sim.hpp:1167-1170, :628-750 */
int pd_step_hertz(pd_sim* s, double dt, int nsteps, double skin_frac);
/* diagnostic: the cached Hertz pair list (N,2) and shear histories (N,3); returns the count */
int pd_get_hertz_pairs(pd_sim* s, int* pairs, double* xi, int cap);

/* --- batched ensembles (no reference counterpart: the reference steps one system per
 * Simulation).  Independent systems, each a full pd_sim configured through the calls above,
 * advanced together; every system's result equals its own pd_step sequence. */
typedef struct pd_batch pd_batch;
pd_batch* pd_batch_create(int n, const int* capacities);
void pd_batch_destroy(pd_batch* b);
int pd_batch_size(pd_batch* b);
pd_sim* pd_batch_system(pd_batch* b, int k); /* owned by the batch */
int pd_batch_set_workers(pd_batch* b, int workers);
/* pd_autotune for every system, largest first, sharing the choice per size class */
int pd_batch_autotune(pd_batch* b, double dt, int substeps);
int pd_batch_step(pd_batch* b, double dt, int nsteps); /* nsteps substeps of every system */

/* --- differentiation of the substep (extension: the reference has no adjoint) -------------
 * State layout, per real body, 13 doubles: position (3), velocity (3), angular velocity (3),
 * orientation quaternion (x, y, z, w).  Supported while recording: spheres (no shell, no grid
 * SDF), planes, periodic axes, gravity, the coloured Gauss-Seidel velocity path with the
 * Coulomb cone, stabilization modes off / one-sided; no SDF walls, pair-material table, growth
 * or thermostat (pd_step returns an error naming the feature).  Masses, inertia, scales and
 * external forces are captured when the recording opens; setting any of them while a recording
 * is open makes the next substep fail, and the derivative calls use the captured values.
 *
 * pd_tape_begin opens a recording (dropping any earlier one): every following substep stores
 * its discrete structure (contacts, manifolds, colour orders, sweep counts, stabilization
 * flags, warm-start map), and every `checkpoint_every` substeps the body state and the
 * warm-start ledger.  pd_tape_end closes it.  Every recording gets a new generation number
 * (pd_tape_generation, 0 before the first), so a caller holding derivatives of an earlier
 * recording can tell that it was replaced.  The state arrays of the calls below hold `count`
 * doubles, which must be 13 x the recorded particle count (else -2, ArgError). */
int pd_tape_begin(pd_sim* s, int checkpoint_every);
int pd_tape_end(pd_sim* s);
int pd_tape_length(pd_sim* s); /* recorded substeps */
long pd_tape_generation(pd_sim* s);
/* replay of the recorded substeps from the first checkpoint (equals the stepped state) */
int pd_tape_replay(pd_sim* s, double* state_out, long count);
/* forward mode: the end state and its directional derivative along (dstate0, dmu, drest),
 * where mu and rest are the friction coefficient and normal restitution; dstate0 may be NULL */
int pd_tape_tangent(pd_sim* s, const double* dstate0, double dmu, double drest, double* state_out,
                    double* dstate_out, long count);
/* reverse mode (adjoint): given dL/d(end state), returns dL/d(start state) and
 * dL/d(mu, rest) in grad_params[2]; checkpointed segments are replayed, then reversed
 * substep by substep (sweeps in reverse colour order) */
int pd_tape_adjoint(pd_sim* s, const double* grad_state_end, double* grad_state0,
                    double* grad_params, long count);

/* --- measurement ------------------------------------------------------------------------ */
/* Timing with HIP events on the handle's stream (no reference counterpart; the reference's
 * get_profiling_info returns an empty dict, src/dem_bindings.cpp:385): level 0 off, 1 per-stage
 * times, 2 also one event pair around every launch of the main kernels (PGS velocity sweep,
 * position sweep, broadphase query, narrowphase, predict, re-integrate, and the one-workgroup
 * sweep loops).  Events are resolved lazily, so timing adds no host synchronisation to a
 * substep.  Values accumulate until pd_set_timing is called again; at level 2 each substep's
 * counters are also logged (pd_get_stats_log). */
int pd_set_timing(pd_sim* s, int level);
/* Tuning (no reference counterpart): "block_solver_rows" = the largest velocity / position
 * sweep loop (in rows) solved inside one workgroup instead of one launch per colour (default
 * 16384; 0 disables); "candidate_skin_permille" = the skin of the candidate pair list kept
 * across substeps, in thousandths of the largest radius (default 100; 0 runs the broadphase
 * every substep); "slot_headroom" = slots pd_set_positions reserves beyond n + n/4 for
 * periodic ghosts (default 4096; more are allocated when a substep needs them);
 * "timing_every" = N records timing events only on substeps 0, N, 2N, ...
 * after pd_set_timing (default 1), which bounds the events' host cost; "velocity_loop_lds" = 1
 * (default) lets a one-workgroup velocity loop of at most 8192 rows keep the velocities of up
 * to "velocity_loop_lds_bodies" of its bodies in LDS (default and maximum 2400; the others stay
 * in global memory) when its rows touch at most "velocity_loop_lds_keys" distinct bodies
 * (default and maximum 12288), 0 keeps them all in global memory; "vsweep_waves" = 3 (default)
 * or 4 selects the per-colour PGS sweep compiled for that many waves per SIMD (4: at most 128
 * VGPRs, with some register spilling); "velocity_pack_bodies" = 1 (default) lets the per-colour
 * velocity launches read each body's velocity and constants from one 128-byte record packed for
 * the stage, 0 from the separate arrays; "serial_class_waves_min_rows" (default 4096): a last
 * contact colour class (62, whose rows may share bodies) of at least that many rows runs in
 * waves of body-disjoint rows instead of one row after another, a wave of at least
 * "serial_class_wave_launch_rows" rows (default 512) in a launch of its own and runs of smaller
 * waves in one workgroup; "colouring_incidence_min_edges" (default 65536): a colouring whose
 * uncoloured edges number at least that many and at least a quarter of its bodies runs its
 * Jones-Plassmann rounds over per-body edge lists sorted by priority (one visit per edge)
 * instead of contend / commit passes over every uncoloured edge (0: every colouring does);
 * "position_tail_rows" (default 2048): in per-colour position sweeps, the trailing contact colours
 * holding at most that many rows together run in one workgroup launch that also ends the sweep;
 * "velocity_tail_rows" (default 0, off): the same for the per-colour velocity sweeps (without a
 * last manifold class 62);
 * "adjoint_keep_segment" = 1 (default) lets pd_tape_adjoint replay each substep of a checkpoint
 * segment once, into workspaces and tapes of its own, when a segment's worth fits in a quarter of
 * the free device memory; 0 replays the segment for its states and each substep again right
 * before its reverse; "side_stream_min_contacts" (default 0): contact lists of at least that
 * many contacts are coloured on the simulation's second stream, beside the velocity stage,
 * shorter ones on its main stream.  Results do not depend on any of them. */
int pd_set_tuning(pd_sim* s, const char* name, long value);
/* Autotuner (no reference counterpart): chooses candidate_skin_permille, block_solver_rows,
 * velocity_loop_lds and vsweep_waves for this simulation by timing its own substeps (coordinate
 * descent over {0, 50, 100, 200, 400}, {0, 2048, 16384, 65536}, {1, 0} and {3, 4}; each trial is
 * one untimed and `substeps` timed substeps, so the call advances the simulation by up to
 * 11 x (1 + substeps) substeps; results do not depend on the knobs).  The choice is cached per
 * size class (shape kind, log2 particle count, log2 contact density); with use_cache != 0 a
 * cached class is applied without stepping.  chosen[4] (may be NULL) receives (skin, rows, lds,
 * waves); best_ms (may be NULL) the best trial's ms per substep (-1 when taken from the cache). */
int pd_autotune(pd_sim* s, double dt, int substeps, int use_cache, long* chosen,
                double* best_ms);
int pd_get_stage_times(pd_sim* s, double* out_ms, int cap);
/* host wall time between the same stage marks (enqueueing plus waiting), summed like the stage
 * times; a stage whose host time exceeds its device time is bound by launch overhead */
int pd_get_stage_host_times(pd_sim* s, double* out_ms, int cap);
const char* pd_stage_name(int stage);
/* per timed kernel k: summed ms, launch count and units processed (k_vsweep: manifold rows,
 * k_psweep: contact rows, k_bp_query: slots queried, k_np_pairs: pairs, k_predict: particles,
 * k_reintegrate: slots, k_vsolve_lds: manifold rows x sweeps of the one-workgroup velocity
 * loops (staged in LDS, with the global-memory k_vsolve_block when they fall back),
 * k_psolve_block: contact rows x sweeps); returns the number of timed kernels */
int pd_get_kernel_times(pd_sim* s, double* ms, long* launches, long* rows, int cap);
/* the counters of every substep since pd_set_timing(s, 2), oldest first: copies at most cap
 * records into out (out may be NULL) and returns how many are logged */
int pd_get_stats_log(pd_sim* s, pd_stats* out, int cap);
const char* pd_kernel_name(int k);

#ifdef __cplusplus
}
#endif

#endif /* PECLET_DEM_H */
