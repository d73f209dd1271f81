/*
 * pdref — CPU oracle for the peclet-dem XPBD substep (demStep) in FP64.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a serial C restatement of the reference algorithm
 * (src/sim.hpp:97-488 and the kernel headers it calls) used as the checker for the HIP
 * implementation.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it; the product library never links or calls it.
 *
 * The API mirrors the product C-ABI (include/peclet_dem.h) one-to-one with the prefix `pdref_`
 * instead of `pd_`, so a test drives both with identical calls.
 */
#ifndef PDREF_H
#define PDREF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Per-substep counters (same layout as pd_stats in include/peclet_dem.h). */
typedef struct pdref_stats {
  int num_real;
  int num_particles;       /* real + periodic ghost slots */
  int num_pairs;           /* broadphase candidate pairs P */
  int num_contacts;        /* narrowphase contacts C */
  int num_manifolds;       /* manifolds M */
  int active_manifolds;    /* manifolds with a colour >= 0 (solved by the velocity sweeps) */
  int num_colors_vel;      /* manifold colours */
  int num_colors_pos;      /* contact colours */
  int rounds_vel;          /* Jones-Plassmann rounds, manifold graph */
  int rounds_pos;          /* Jones-Plassmann rounds, contact graph */
  int sweeps_vel;          /* executed main velocity sweeps S_v */
  int sweeps_stab;         /* executed stabilization sweeps */
  int sweeps_pos;          /* executed position sweeps S_p */
  int stab_ran;            /* 1 if the stabilization pass ran */
  long long uncolored_vel; /* sum over rounds of uncoloured manifolds entering the round */
  long long uncolored_pos; /* sum over rounds of uncoloured contacts entering the round */
  double max_overlap;
  double max_approach;
  int ml_levels;           /* multilevel stabilization: coarse levels built (0 otherwise) */
} pdref_stats;

void* pdref_create(int capacity);
void pdref_destroy(void* s);
/* host threads for the element-wise loops and the rows of one colour class (default 1); results
 * are bit-identical for every thread count */
int pdref_set_threads(int n);
const char* pdref_last_error(void);

int pdref_initialize_shape(void* s, int shape_type, double radius, double height,
                           double thickness);
int pdref_set_shell_points(void* s, const double* pts, int n);
int pdref_get_shell_points(void* s, double* out, int cap);
int pdref_set_domain(void* s, double lx, double ly, double lz, int px, int py, int pz);
int pdref_set_domain_minmax(void* s, double mnx, double mny, double mnz, double mxx, double mxy,
                            double mxz);
int pdref_enable_periodicity(void* s, int x, int y, int z);
int pdref_get_domain(void* s, double* mn3, double* mx3);
int pdref_set_gravity(void* s, double gx, double gy, double gz);
int pdref_set_thermostat(void* s, double temperature, double tau, double kB);
int pdref_set_solver_iterations(void* s, int pos, int vel);
int pdref_set_velocity_use_gs(void* s, int use_gs);
int pdref_set_stabilization_mode(void* s, int mode);
int pdref_set_global_scale(void* s, double gs);
int pdref_set_dt(void* s, double dt);
int pdref_set_material_params(void* s, double e_n, double e_t, double mu);
int pdref_set_material_ids(void* s, const int* ids, int n);
int pdref_set_pair_material(void* s, int a, int b, double e, double mu);
int pdref_add_plane(void* s, double px, double py, double pz, double nx, double ny, double nz);
int pdref_set_positions(void* s, const double* xyz, int n);
int pdref_set_velocities(void* s, const double* v, int n);
int pdref_set_external_forces(void* s, const double* f, int n);
int pdref_clear_external_forces(void* s);
int pdref_set_quaternions(void* s, const double* q, int n);
int pdref_set_angular_velocities(void* s, const double* w, int n);
int pdref_set_inv_inertia(void* s, const double* ii, int n);
int pdref_set_inv_mass(void* s, const double* im, int n);
int pdref_set_scales(void* s, const double* sc, int n);
int pdref_set_scales_uniform(void* s, double sc);
int pdref_set_growth_params(void* s, double rate, double new_factor);
double pdref_get_growth_factor(void* s);
double pdref_get_growth_rate(void* s);

int pdref_num_particles(void* s);
int pdref_get_positions(void* s, double* out, int n);
int pdref_get_velocities(void* s, double* out, int n);
int pdref_get_quaternions(void* s, double* out, int n);
int pdref_get_angular_velocities(void* s, double* out, int n);
int pdref_get_inv_inertia(void* s, double* out, int n);
int pdref_get_inv_mass(void* s, double* out, int n);
int pdref_get_scales(void* s, double* out, int n);

int pdref_step(void* s, double dt);
int pdref_set_sdf_shape(void* s, const double* grid, int nx, int ny, int nz, double ox,
                        double oy, double oz, double sx, double sy, double sz,
                        const double* shell, int nPts, double iix, double iiy, double iiz,
                        double boundingRadius);
/* packed-bed SDF on an (rx, ry, rz) voxel grid over the domain, x-fastest (src/output_sdf.hpp) */
int pdref_get_sdf_grid(void* s, int rx, int ry, int rz, double* out);
/* SDF walls and the Hertz–Mindlin engine (src/sim.hpp:1153-1287, :628-750) */
int pdref_add_sdf_wall(void* s, const double* grid, int nx, int ny, int nz, double ox, double oy,
                       double oz, double sx, double sy, double sz, double restitution,
                       double friction);
int pdref_set_wall_velocity(void* s, int wid, double lx, double ly, double lz, double ax,
                            double ay, double az, double cx, double cy, double cz);
int pdref_set_wall_material_id(void* s, int wid, int mat);
int pdref_num_walls(void* s);
int pdref_set_hertz_material(void* s, int mat, double youngs, double poisson);
int pdref_step_hertz(void* s, double dt, int nsteps, double skin_frac);
int pdref_get_hertz_pairs(void* s, int* pairs, double* xi, int cap);
int pdref_compute_overlaps(void* s, double* out);
int pdref_num_contacts(void* s);
int pdref_num_manifolds(void* s);
double pdref_max_overlap(void* s);
int pdref_get_stats(void* s, pdref_stats* out);

/* Diagnostics of the last substep (counts returned; NULL out = query the count). */
int pdref_get_pairs(void* s, int* out, int cap);
int pdref_get_contacts(void* s, int* ids, double* data, int cap);
int pdref_get_contact_colors(void* s, int* out, int cap);
int pdref_get_manifolds(void* s, int* ids, double* data, int cap);
int pdref_get_manifold_colors(void* s, int* out, int cap);
int pdref_get_ledger(void* s, uint64_t* keys, double* lambda, double* lambdaT, double* posImp,
                     int cap);
int pdref_get_grounded(void* s, int* out, int n);

/* Standalone broadphase over caller arrays (tests): sorted (i<j) pairs, returns the count. */
int pdref_broadphase(const double* pos, const double* rad, int numParticles, int numReal,
                     double margin, double boxCap, int* outPairs, int cap);
int pdref_broadphase_bruteforce(const double* pos, const double* rad, int numParticles,
                                int numReal, double margin, double boxCap, int* outPairs, int cap);

#ifdef __cplusplus
}
#endif

#endif /* PDREF_H */
