/*
 * pdref.c — CPU oracle (serial, FP64) for the peclet-dem XPBD substep.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP implementation.  It is never linked into,
 * loaded by, or called from the product library.
 *
 * Every function restates the reference algorithm and cites the reference file:line it follows.
 * The reference computes in FP32; this restatement computes the same operation sequence in FP64.
 * Where the multi-threaded reference has no defined order (atomic appends / atomic float adds /
 * racy sweeps) this file fixes one canonical order, which the HIP path reproduces:
 *   - ghost slots: real index ascending, then the (a,b,c) loop order of the emitter
 *   - broadphase pairs: lexicographic (i, j)
 *   - contacts: pair order x shell index, then planes (real i, plane, shell index)
 *   - manifold sums: contacts in stable key order
 *   - warm-start application: colour class by colour class, manifolds ascending within a class
 *   - grounded-level sweeps: double buffered (every sweep reads the previous sweep's levels)
 *   - position-impulse reduction: contacts ascending within a manifold
 *   - sorted ledger: stable (key, manifold index)
 * Build with -ffp-contract=off so no multiply-add is fused.
 *
 * Threads (pdref_set_threads, default 1): element-wise loops and the rows of one colour class
 * (which touch disjoint bodies; colour 62 stays serial) are split over OpenMP threads.  Each
 * element is computed by the same statements either way and the running maxima are combined
 * with max, so every result is bit-identical to the one-thread run.
 *
 * Pinning (tests/test_oracle.py, tests/test_oracle_hertz.py): reference test inputs, brute force,
 * closed forms and independent algorithms; the Jacobi solves by binary restitution, symmetric
 * supports and the geometric convergence of the position average; the escalate and ordered
 * loops, warm start, persistence, grounded levels and the one-sided side flags by a
 * one-dimensional column model over two substeps; the multilevel stabilization by a two-sphere
 * closed form, its eligibility gates and momentum conservation in a wall-free cluster; the Hertz
 * shape paths by the SDF-sphere restitution and cube stick/slide closed forms.  No built
 * function is left unpinned.
 */
#include "pdref.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static int g_threads = 1;
int pdref_set_threads(int n) {
  g_threads = n < 1 ? 1 : n;
  return 0;
}
/* threads for a loop of n independent elements (small loops stay serial) */
static int nthreads(long n) { return n >= 2048 ? g_threads : 1; }

typedef double real;
typedef struct { real x, y, z; } V3;
typedef struct { real x, y, z, w; } Q4;

/* ------------------------------------------------------------------------------------------ */
/* error reporting                                                                             */
/* ------------------------------------------------------------------------------------------ */
static char g_err[512];
static int fail(const char* msg) {
  snprintf(g_err, sizeof g_err, "%s", msg);
  return -1;
}
/* invalid arguments (the reference's std::invalid_argument, a ValueError in Python) */
static int badarg(const char* msg) {
  snprintf(g_err, sizeof g_err, "%s", msg);
  return -2;
}
const char* pdref_last_error(void) { return g_err; }

/* ------------------------------------------------------------------------------------------ */
/* vector / quaternion math — src/dem_portable.hpp:25-74                                       */
/* ------------------------------------------------------------------------------------------ */
static V3 mk3(real x, real y, real z) { V3 r; r.x = x; r.y = y; r.z = z; return r; }
static V3 add3(V3 a, V3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
static V3 sub3(V3 a, V3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
static V3 scale3(V3 a, real s) { return mk3(a.x * s, a.y * s, a.z * s); }
static real dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static real len3(V3 v) { return sqrt(dot3(v, v)); }
static V3 cross3(V3 a, V3 b) {
  return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static int nonzero3(V3 a) { return a.x != 0.0 || a.y != 0.0 || a.z != 0.0; }
/* v' = v + w t + q x t, t = 2 q x v  (src/dem_portable.hpp:59-63) */
static V3 rotate(Q4 q, V3 v) {
  const V3 qv = mk3(q.x, q.y, q.z);
  const V3 t = scale3(cross3(qv, v), 2.0);
  return add3(add3(v, scale3(t, q.w)), cross3(qv, t));
}
static V3 invRotate(Q4 q, V3 v) {
  Q4 c; c.x = -q.x; c.y = -q.y; c.z = -q.z; c.w = q.w;
  return rotate(c, v);
}
static Q4 quatInverse(Q4 q) { Q4 r; r.x = -q.x; r.y = -q.y; r.z = -q.z; r.w = q.w; return r; }
static Q4 quatMult(Q4 a, Q4 b) {
  Q4 r;
  r.x = a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y;
  r.y = a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x;
  r.z = a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w;
  r.w = a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z;
  return r;
}

/* ------------------------------------------------------------------------------------------ */
/* analytic SDFs in canonical space — src/dem_portable.hpp:77-113                              */
/* ------------------------------------------------------------------------------------------ */
enum { SHAPE_GRID_SDF = 0, SHAPE_SPHERE = 1, SHAPE_HOLLOW_CYLINDER = 2, SHAPE_BOX = 3 };

static real sdfSphere(V3 p, const real* prm) { return len3(p) - prm[0]; }
static real sdfHollowCylinder(V3 p, const real* prm) {
  const real r_outer = prm[0], h = prm[1], thick = prm[2];
  const real r = sqrt(p.x * p.x + p.z * p.z);
  const real r_mid = r_outer - thick * 0.5;
  const real dx = fabs(r - r_mid) - thick * 0.5;
  const real dy = fabs(p.y) - h * 0.5;
  const real ox = fmax(dx, 0.0), oy = fmax(dy, 0.0);
  const real outside = sqrt(ox * ox + oy * oy);
  const real inside = fmin(fmax(dx, dy), 0.0);
  return outside + inside;
}
static real sdfBox(V3 p, const real* prm) {
  const real dx = fabs(p.x) - prm[0];
  const real dy = fabs(p.y) - prm[1];
  const real dz = fabs(p.z) - prm[2];
  const real ox = fmax(dx, 0.0), oy = fmax(dy, 0.0), oz = fmax(dz, 0.0);
  const real outside = sqrt(ox * ox + oy * oy + oz * oz);
  const real inside = fmin(fmax(dx, fmax(dy, dz)), 0.0);
  return outside + inside;
}
static real sdfEval(V3 p, int type, const real* prm) {
  if (type == SHAPE_SPHERE) return sdfSphere(p, prm);
  if (type == SHAPE_HOLLOW_CYLINDER) return sdfHollowCylinder(p, prm);
  if (type == SHAPE_BOX) return sdfBox(p, prm);
  return 1e9;
}

/* ------------------------------------------------------------------------------------------ */
/* records — src/contact_preprocessing.hpp:31-63, src/narrowphase.hpp:29-44, integration:19   */
/* ------------------------------------------------------------------------------------------ */
typedef struct { V3 min, max, size; int px, py, pz; } Domain;
typedef struct { V3 point, normal; } Plane;

typedef struct {
  int bodyA, bodyB;
  V3 normal, rA, rB;
  real dist;
  real frictionLambdaN;
  V3 boundaryVel;
  real boundaryRestitution, boundaryFriction;
} Contact;

typedef struct {
  int bodyA, bodyB;
  V3 normalSum, tauA, tauB, rASum, rBSum;
  int numPoints;
  V3 wallVelSum;
  real restitutionSum, frictionSum;
} Manifold;

#define KMAXMAT 8
#define MAXPLANES 8
#define MAXWALLS 8
#define HZ_MAXWALLS 4 /* Particles::kHertzMaxWalls, src/particles.hpp:130 */

/* static world-space SDF wall, src/narrowphase.hpp:59-76 */
typedef struct {
  int nx, ny, nz;
  long off;
  V3 origin, invSpacing;
  V3 linVel, angVel, center;
  real restitution, friction;
  int materialId;
} Wall;

typedef struct Sim {
  int capacity;
  int numReal, numParticles;
  /* shape 0 (src/sim.hpp:949-1021) */
  int shapeType;
  real shapeParams[4];
  int numShell;
  V3* shell;
  real baseRadius;
  V3 defaultInvI;
  /* host parameters (src/particles.hpp:173-191) */
  Domain dom;
  V3 gravity;
  real dt, globalScale, growthRate, growthFactor;
  real thermoTau, thermoTemp, thermoKB;
  real frictionDynamic, restitutionNormal, restitutionTangent;
  int posIters, velIters, useGS, stabMode;
  Plane planes[MAXPLANES];
  int numPlanes;
  int hasPairTable;
  real pairTable[KMAXMAT * KMAXMAT * 2];
  /* per-slot state (reals then ghosts) */
  int slotCap;
  V3 *pos, *vel, *angVel, *invI, *extForce, *posPred, *velPred, *angVelPred;
  V3 *deltaVel, *deltaAngVel, *deltaPos;
  Q4 *quat, *quatPred;
  real *invMass, *scale, *targetScale, *rad;
  int *realIdx, *counts;
  unsigned char* materialId;
  int* grounded; /* per real body, persistent */
  int groundedCap;
  /* per-substep buffers */
  int *pairs; int np, pairCap;
  Contact* contacts; int nc, contactCap;
  Manifold* manifolds; int nm, manifoldCap;
  int *contactSlot, *mColor, *cColor;
  uint64_t* pairKeys;
  unsigned char *persistent, *sideFlags;
  real *lambdaAcc, *posImpulse, *vn0, *posLambda;
  V3 *lambdaT, *vt0;
  /* warm-start ledger (sorted by key) */
  uint64_t* prevKeys;
  real *prevLambda, *prevPosImpulse;
  V3* prevLambdaT;
  int prevCount, prevCap;
  real maxOverlap, maxApproach;
  real maxApproachQS; /* the multilevel loop's quasi-static residual */
  int mlGates;        /* multilevel eligibility gates (PECLET_DEM_ML_GATES, default slip) */
  int numColorsVel, numColorsPos;
  pdref_stats stats;
  /* grid-SDF particle shape (src/sim.hpp:1034-1084) */
  real* sdfGrid;
  int gnx, gny, gnz;
  V3 gOrigin, gInvSpacing;
  /* SDF walls (src/sim.hpp:1242-1287) */
  Wall walls[MAXWALLS];
  int numWalls;
  real* wallGrid;
  long wallGridN;
  /* Hertz–Mindlin engine state (src/particles.hpp:113-130) */
  real hertzE[KMAXMAT], hertzNu[KMAXMAT];
  real hertzContactRadiusFrac;
  int hNumPairs, hPrevCount, hPairCap, hSlotCap, hCandCap;
  int* hPairs;
  uint64_t *hKeys, *hPrevKeys;
  V3 *hXi, *hPrevXi;
  real* hSn;
  V3 *hRefPos, *hForce, *hTorque, *hXiWall;
  real* hSnWall;
  int* hWallCand;
  int hNumWallCand;
} Sim;

static void* xrealloc(void* p, size_t n) {
  void* r = realloc(p, n ? n : 1);
  if (!r) { fprintf(stderr, "pdref: out of memory\n"); abort(); }
  return r;
}

/* grow every per-slot array to hold n slots; new slots are zeroed */
static void ensureSlots(Sim* S, int n) {
  if (n <= S->slotCap) return;
  int nc = n + n / 4 + 64;
#define GROW(f, T)                                                         \
  S->f = (T*)xrealloc(S->f, sizeof(T) * (size_t)nc);                       \
  memset(S->f + S->slotCap, 0, sizeof(T) * (size_t)(nc - S->slotCap));
  GROW(pos, V3) GROW(vel, V3) GROW(angVel, V3) GROW(invI, V3) GROW(extForce, V3)
  GROW(posPred, V3) GROW(velPred, V3) GROW(angVelPred, V3) GROW(deltaVel, V3)
  GROW(deltaAngVel, V3) GROW(deltaPos, V3) GROW(quat, Q4) GROW(quatPred, Q4)
  GROW(invMass, real) GROW(scale, real) GROW(targetScale, real) GROW(rad, real)
  GROW(realIdx, int) GROW(counts, int) GROW(materialId, unsigned char)
#undef GROW
  S->slotCap = nc;
}

static void ensureContacts(Sim* S, int n) {
  if (n <= S->contactCap) return;
  int nc = n + n / 4 + 64;
  S->contacts = (Contact*)xrealloc(S->contacts, sizeof(Contact) * (size_t)nc);
  S->contactSlot = (int*)xrealloc(S->contactSlot, sizeof(int) * (size_t)nc);
  S->cColor = (int*)xrealloc(S->cColor, sizeof(int) * (size_t)nc);
  S->posLambda = (real*)xrealloc(S->posLambda, sizeof(real) * (size_t)nc);
  S->contactCap = nc;
}

static void ensureManifolds(Sim* S, int n) {
  if (n <= S->manifoldCap) return;
  int nc = n + n / 4 + 64;
  size_t oc = (size_t)S->manifoldCap;
  S->manifolds = (Manifold*)xrealloc(S->manifolds, sizeof(Manifold) * (size_t)nc);
  S->mColor = (int*)xrealloc(S->mColor, sizeof(int) * (size_t)nc);
  S->pairKeys = (uint64_t*)xrealloc(S->pairKeys, sizeof(uint64_t) * (size_t)nc);
  S->persistent = (unsigned char*)xrealloc(S->persistent, (size_t)nc);
  S->sideFlags = (unsigned char*)xrealloc(S->sideFlags, (size_t)nc);
  S->lambdaAcc = (real*)xrealloc(S->lambdaAcc, sizeof(real) * (size_t)nc);
  S->posImpulse = (real*)xrealloc(S->posImpulse, sizeof(real) * (size_t)nc);
  S->vn0 = (real*)xrealloc(S->vn0, sizeof(real) * (size_t)nc);
  S->lambdaT = (V3*)xrealloc(S->lambdaT, sizeof(V3) * (size_t)nc);
  S->vt0 = (V3*)xrealloc(S->vt0, sizeof(V3) * (size_t)nc);
  memset(S->vn0 + oc, 0, sizeof(real) * ((size_t)nc - oc));
  S->manifoldCap = nc;
}

static void ensureLedger(Sim* S, int n) {
  if (n <= S->prevCap) return;
  int nc = n + n / 4 + 64;
  S->prevKeys = (uint64_t*)xrealloc(S->prevKeys, sizeof(uint64_t) * (size_t)nc);
  S->prevLambda = (real*)xrealloc(S->prevLambda, sizeof(real) * (size_t)nc);
  S->prevPosImpulse = (real*)xrealloc(S->prevPosImpulse, sizeof(real) * (size_t)nc);
  S->prevLambdaT = (V3*)xrealloc(S->prevLambdaT, sizeof(V3) * (size_t)nc);
  memset(S->prevPosImpulse + S->prevCap, 0, sizeof(real) * (size_t)(nc - S->prevCap));
  S->prevCap = nc;
}

static void ensureGrounded(Sim* S, int n) {
  if (n <= S->groundedCap) return;
  S->grounded = (int*)xrealloc(S->grounded, sizeof(int) * (size_t)n);
  memset(S->grounded + S->groundedCap, 0, sizeof(int) * (size_t)(n - S->groundedCap));
  S->groundedCap = n;
}

static Sim* asSim(void* s) { return (Sim*)s; }

/* ------------------------------------------------------------------------------------------ */
/* shell generators — src/shapes_portable.hpp:21-111 (computed in FP64)                        */
/* ------------------------------------------------------------------------------------------ */
typedef struct { V3* p; int n, cap; } PtList;
static void ptPush(PtList* L, real x, real y, real z) {
  if (L->n == L->cap) {
    L->cap = L->cap ? 2 * L->cap : 256;
    L->p = (V3*)xrealloc(L->p, sizeof(V3) * (size_t)L->cap);
  }
  L->p[L->n++] = mk3(x, y, z);
}
static void genCylinderShell(PtList* L, real radius, real height, real thickness, real spacing) {
  const real r_outer = radius, r_inner = radius - thickness, h = height;
  real circumference = 2.0 * M_PI * r_outer;
  int n_angular = (int)ceil(circumference / spacing);
  const int n_vertical = (int)ceil(h / spacing);
  for (int i = 0; i < n_angular; ++i) {
    const real theta = 2.0 * M_PI * i / n_angular;
    for (int j = 0; j <= n_vertical; ++j) {
      const real y = -0.5 * h + h * j / n_vertical;
      ptPush(L, r_outer * cos(theta), y, r_outer * sin(theta));
    }
  }
  if (thickness > 0.0) {
    circumference = 2.0 * M_PI * r_inner;
    n_angular = (int)ceil(circumference / spacing);
    for (int i = 0; i < n_angular; ++i) {
      const real theta = 2.0 * M_PI * i / n_angular;
      for (int j = 0; j <= n_vertical; ++j) {
        const real y = -0.5 * h + h * j / n_vertical;
        ptPush(L, r_inner * cos(theta), y, r_inner * sin(theta));
      }
    }
  }
  const real dr = spacing;
  const int n_radial = (int)ceil(thickness / dr);
  for (int i = 0; i <= n_radial; ++i) {
    const real r = r_inner + (r_outer - r_inner) * i / n_radial;
    circumference = 2.0 * M_PI * r;
    n_angular = (int)ceil(circumference / spacing);
    for (int j = 0; j < n_angular; ++j) {
      const real theta = 2.0 * M_PI * j / n_angular;
      const real x = r * cos(theta), z = r * sin(theta);
      ptPush(L, x, 0.5 * h, z);
      ptPush(L, x, -0.5 * h, z);
    }
  }
}
static int nseg(real half, real spacing) {
  int n = (int)ceil(2.0 * half / spacing);
  return n < 1 ? 1 : n;
}
static real lin(real half, int i, int n) { return -half + 2.0 * half * (real)i / (real)n; }
static void genBoxShell(PtList* L, real hx, real hy, real hz, real spacing) {
  const int nx = nseg(hx, spacing), ny = nseg(hy, spacing), nz = nseg(hz, spacing);
  for (int j = 0; j <= ny; ++j)
    for (int k = 0; k <= nz; ++k) {
      const real y = lin(hy, j, ny), z = lin(hz, k, nz);
      ptPush(L, hx, y, z);
      ptPush(L, -hx, y, z);
    }
  for (int i = 0; i <= nx; ++i)
    for (int k = 0; k <= nz; ++k) {
      const real x = lin(hx, i, nx), z = lin(hz, k, nz);
      ptPush(L, x, hy, z);
      ptPush(L, x, -hy, z);
    }
  for (int i = 0; i <= nx; ++i)
    for (int j = 0; j <= ny; ++j) {
      const real x = lin(hx, i, nx), y = lin(hy, j, ny);
      ptPush(L, x, y, hz);
      ptPush(L, x, y, -hz);
    }
}

/* ------------------------------------------------------------------------------------------ */
/* facade setup — src/sim.hpp:911-1330                                                         */
/* ------------------------------------------------------------------------------------------ */
int pdref_initialize_shape(void* s, int shape_type, double radius, double height,
                           double thickness) {
  /* src/sim.hpp:949-1021 */
  Sim* S = asSim(s);
  S->baseRadius = radius;
  real prm[4] = {radius, 0, 0, 0};
  PtList L = {0, 0, 0};
  if (shape_type == SHAPE_HOLLOW_CYLINDER) {
    prm[1] = height;
    prm[2] = thickness;
    real min_dim = radius < thickness ? radius : thickness;
    if (min_dim < 1e-4) min_dim = radius;
    real a = radius * 0.3, b = min_dim * 0.5;
    real spacing = a < b ? a : b;
    if (spacing < 1e-3) spacing = 1e-3;
    genCylinderShell(&L, radius, height, thickness, spacing);
  } else if (shape_type == SHAPE_BOX) {
    prm[1] = radius;
    prm[2] = radius;
    real spacing = radius * 0.5 > 1e-3 ? radius * 0.5 : 1e-3;
    genBoxShell(&L, radius, radius, radius, spacing);
  } else {
    shape_type = SHAPE_SPHERE;
  }
  S->shapeType = shape_type;
  memcpy(S->shapeParams, prm, sizeof prm);
  free(S->shell);
  S->shell = L.p;
  S->numShell = L.n;
  real ix = 1.0, iy = 1.0, iz = 1.0;
  if (shape_type == SHAPE_SPHERE) {
    if (S->baseRadius > 0.0) {
      real v = 2.5 / (S->baseRadius * S->baseRadius);
      ix = iy = iz = v;
    }
  } else if (shape_type == SHAPE_HOLLOW_CYLINDER) {
    real r_out = S->baseRadius, r_in = S->baseRadius - thickness;
    if (r_in < 0) r_in = 0;
    real term_r = r_out * r_out + r_in * r_in;
    real I_zz = 0.5 * term_r;
    real I_xx = (1.0 / 12.0) * (3.0 * term_r + height * height);
    if (I_xx > 1e-6) { ix = 1.0 / I_xx; iy = 1.0 / I_xx; }
    if (I_zz > 1e-6) iz = 1.0 / I_zz;
  } else if (shape_type == SHAPE_BOX) {
    real Lb = 2.0 * S->baseRadius;
    real I = (1.0 / 6.0) * Lb * Lb;
    if (I > 1e-6) ix = iy = iz = 1.0 / I;
  }
  S->defaultInvI = mk3(ix, iy, iz);
  return 0;
}

/* Extension (no reference counterpart): replace shape 0's surface point shell. */
int pdref_set_shell_points(void* s, const double* pts, int n) {
  Sim* S = asSim(s);
  if (n < 0) return badarg("set_shell_points: negative count");
  if (S->shapeType == SHAPE_SPHERE && n > 0)
    return badarg("set_shell_points: the analytic sphere uses a single centre probe");
  free(S->shell);
  S->shell = (V3*)xrealloc(NULL, sizeof(V3) * (size_t)(n > 0 ? n : 1));
  for (int i = 0; i < n; ++i) S->shell[i] = mk3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
  S->numShell = n;
  return 0;
}

int pdref_get_shell_points(void* s, double* out, int cap) {
  Sim* S = asSim(s);
  if (out)
    for (int i = 0; i < S->numShell && i < cap; ++i) {
      out[3 * i] = S->shell[i].x;
      out[3 * i + 1] = S->shell[i].y;
      out[3 * i + 2] = S->shell[i].z;
    }
  return S->numShell;
}

void* pdref_create(int capacity) {
  /* src/sim.hpp:911-916 + src/particles.hpp:173-191 defaults */
  if (capacity < 0) { fail("capacity must be >= 0"); return NULL; }
  Sim* S = (Sim*)calloc(1, sizeof(Sim));
  if (!S) return NULL;
  S->capacity = capacity;
  S->dt = 1e-3;
  S->globalScale = 1.0;
  S->growthFactor = -1.0;
  S->baseRadius = 1.0;
  S->thermoKB = 1.0;
  S->posIters = 10;
  S->velIters = 0;
  S->useGS = 1;
  S->stabMode = 1;
  {
    /* src/sim.hpp:318-321: the gate mask is read from the environment, default kGateSlip */
    const char* e = getenv("PECLET_DEM_ML_GATES");
    S->mlGates = e ? atoi(e) : 4;
  }
  S->defaultInvI = mk3(2.5, 2.5, 2.5);
  for (int m = 0; m < KMAXMAT; ++m) { S->hertzE[m] = 1.0e9; S->hertzNu[m] = 0.25; }
  S->hertzContactRadiusFrac = 0.5;
  S->hNumPairs = -1;
  ensureSlots(S, capacity > 0 ? capacity : 1);
  pdref_initialize_shape(S, SHAPE_SPHERE, 1.0, 0.0, 0.0);
  return S;
}

void pdref_destroy(void* s) {
  Sim* S = asSim(s);
  if (!S) return;
  free(S->shell);
  free(S->pos); free(S->vel); free(S->angVel); free(S->invI); free(S->extForce);
  free(S->posPred); free(S->velPred); free(S->angVelPred); free(S->deltaVel);
  free(S->deltaAngVel); free(S->deltaPos); free(S->quat); free(S->quatPred);
  free(S->invMass); free(S->scale); free(S->targetScale); free(S->rad);
  free(S->realIdx); free(S->counts); free(S->materialId); free(S->grounded);
  free(S->pairs); free(S->contacts); free(S->manifolds); free(S->contactSlot);
  free(S->mColor); free(S->cColor); free(S->pairKeys); free(S->persistent);
  free(S->sideFlags); free(S->lambdaAcc); free(S->posImpulse); free(S->vn0);
  free(S->posLambda); free(S->lambdaT); free(S->vt0); free(S->prevKeys);
  free(S->prevLambda); free(S->prevPosImpulse); free(S->prevLambdaT);
  free(S->wallGrid); free(S->sdfGrid); free(S->hPairs); free(S->hKeys); free(S->hPrevKeys); free(S->hXi);
  free(S->hPrevXi); free(S->hSn); free(S->hRefPos); free(S->hForce); free(S->hTorque);
  free(S->hXiWall); free(S->hSnWall); free(S->hWallCand);
  free(S);
}

int pdref_set_domain(void* s, double lx, double ly, double lz, int px, int py, int pz) {
  Sim* S = asSim(s); /* src/sim.hpp:1086-1089 */
  S->dom.min = mk3(0, 0, 0);
  S->dom.max = mk3(lx, ly, lz);
  S->dom.size = mk3(lx, ly, lz);
  S->dom.px = px != 0; S->dom.py = py != 0; S->dom.pz = pz != 0;
  return 0;
}
int pdref_set_domain_minmax(void* s, double mnx, double mny, double mnz, double mxx, double mxy,
                            double mxz) {
  Sim* S = asSim(s); /* src/sim.hpp:1091-1100 */
  S->dom.min = mk3(mnx, mny, mnz);
  S->dom.max = mk3(mxx, mxy, mxz);
  S->dom.size = mk3(mxx - mnx, mxy - mny, mxz - mnz);
  return 0;
}
int pdref_enable_periodicity(void* s, int x, int y, int z) {
  Sim* S = asSim(s);
  S->dom.px = x != 0; S->dom.py = y != 0; S->dom.pz = z != 0;
  return 0;
}
int pdref_get_domain(void* s, double* mn, double* mx) {
  Sim* S = asSim(s);
  mn[0] = S->dom.min.x; mn[1] = S->dom.min.y; mn[2] = S->dom.min.z;
  mx[0] = S->dom.max.x; mx[1] = S->dom.max.y; mx[2] = S->dom.max.z;
  return 0;
}
int pdref_set_gravity(void* s, double gx, double gy, double gz) {
  asSim(s)->gravity = mk3(gx, gy, gz);
  return 0;
}
int pdref_set_thermostat(void* s, double temperature, double tau, double kB) {
  Sim* S = asSim(s);
  S->thermoTemp = temperature; S->thermoTau = tau; S->thermoKB = kB;
  return 0;
}
int pdref_set_solver_iterations(void* s, int pos, int vel) {
  Sim* S = asSim(s);
  S->posIters = pos; S->velIters = vel;
  return 0;
}
int pdref_set_velocity_use_gs(void* s, int use_gs) { asSim(s)->useGS = use_gs != 0; return 0; }
int pdref_set_stabilization_mode(void* s, int mode) {
  if (mode < 0 || mode > 4) return badarg("stabilization mode must be 0..4");
  asSim(s)->stabMode = mode;
  return 0;
}
int pdref_set_global_scale(void* s, double gs) { asSim(s)->globalScale = gs; return 0; }
int pdref_set_dt(void* s, double dt) { asSim(s)->dt = dt; return 0; }
int pdref_set_material_params(void* s, double e_n, double e_t, double mu) {
  Sim* S = asSim(s); /* src/sim.hpp:1179-1183 */
  S->restitutionNormal = e_n; S->frictionDynamic = mu; S->restitutionTangent = e_t;
  return 0;
}
int pdref_set_material_ids(void* s, const int* ids, int n) {
  Sim* S = asSim(s); /* src/sim.hpp:1187-1193 */
  for (int i = 0; i < S->numReal && i < n; ++i) S->materialId[i] = (unsigned char)ids[i];
  return 0;
}
int pdref_set_pair_material(void* s, int a, int b, double e, double mu) {
  Sim* S = asSim(s); /* src/sim.hpp:1196-1215 */
  if (a < 0 || b < 0 || a >= KMAXMAT || b >= KMAXMAT) return badarg("material id out of range");
  if (!S->hasPairTable) {
    for (int i = 0; i < KMAXMAT * KMAXMAT; ++i) {
      S->pairTable[2 * i] = S->restitutionNormal;
      S->pairTable[2 * i + 1] = S->frictionDynamic;
    }
    S->hasPairTable = 1;
  }
  S->pairTable[(a * KMAXMAT + b) * 2] = e; S->pairTable[(a * KMAXMAT + b) * 2 + 1] = mu;
  S->pairTable[(b * KMAXMAT + a) * 2] = e; S->pairTable[(b * KMAXMAT + a) * 2 + 1] = mu;
  return 0;
}
int pdref_add_plane(void* s, double px, double py, double pz, double nx, double ny, double nz) {
  Sim* S = asSim(s); /* src/sim.hpp:1226-1232: silently ignored beyond capacity 8 */
  if (S->numPlanes < MAXPLANES) {
    S->planes[S->numPlanes].point = mk3(px, py, pz);
    S->planes[S->numPlanes].normal = mk3(nx, ny, nz);
    S->numPlanes++;
  }
  return 0;
}
int pdref_set_positions(void* s, const double* xyz, int n) {
  Sim* S = asSim(s); /* src/sim.hpp:1290-1330 */
  if (n < 0) return badarg("negative particle count");
  ensureSlots(S, n);
  ensureGrounded(S, n);
  S->numReal = n;
  S->numParticles = n;
  S->hNumPairs = -1; /* particle indices changed: the Hertz pair cache is stale, :1294-1295 */
  S->hPrevCount = 0;
  for (int i = 0; i < n; ++i) {
    S->pos[i] = mk3(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
    S->quat[i].x = 0; S->quat[i].y = 0; S->quat[i].z = 0; S->quat[i].w = 1;
    S->invMass[i] = 1.0;
    S->scale[i] = 1.0;
    S->invI[i] = S->defaultInvI;
    S->vel[i] = mk3(0, 0, 0);
    S->angVel[i] = mk3(0, 0, 0);
    S->targetScale[i] = 1.0;
  }
  return 0;
}
#define SETTER3(name, field)                                             \
  int pdref_##name(void* s, const double* v, int n) {                    \
    Sim* S = asSim(s);                                                   \
    for (int i = 0; i < S->numReal && i < n; ++i)                        \
      S->field[i] = mk3(v[3 * i], v[3 * i + 1], v[3 * i + 2]);           \
    return 0;                                                            \
  }
SETTER3(set_velocities, vel)
SETTER3(set_external_forces, extForce)
SETTER3(set_angular_velocities, angVel)
SETTER3(set_inv_inertia, invI)
#undef SETTER3
int pdref_clear_external_forces(void* s) {
  Sim* S = asSim(s);
  memset(S->extForce, 0, sizeof(V3) * (size_t)S->slotCap);
  return 0;
}
int pdref_set_quaternions(void* s, const double* q, int n) {
  Sim* S = asSim(s);
  for (int i = 0; i < S->numReal && i < n; ++i) {
    S->quat[i].x = q[4 * i]; S->quat[i].y = q[4 * i + 1];
    S->quat[i].z = q[4 * i + 2]; S->quat[i].w = q[4 * i + 3];
  }
  return 0;
}
int pdref_set_inv_mass(void* s, const double* im, int n) {
  Sim* S = asSim(s);
  for (int i = 0; i < S->numReal && i < n; ++i) S->invMass[i] = im[i];
  return 0;
}
int pdref_set_scales(void* s, const double* sc, int n) {
  Sim* S = asSim(s); /* src/sim.hpp:1340-1346 */
  for (int i = 0; i < S->numReal && i < n; ++i) S->targetScale[i] = sc[i];
  for (int i = 0; i < S->slotCap; ++i) S->scale[i] = S->targetScale[i];
  return 0;
}
int pdref_set_scales_uniform(void* s, double sc) {
  Sim* S = asSim(s); /* src/sim.hpp:1332-1338 */
  for (int i = 0; i < S->numReal; ++i) S->scale[i] = sc;
  for (int i = 0; i < S->slotCap; ++i) S->targetScale[i] = S->scale[i];
  return 0;
}
int pdref_set_growth_params(void* s, double rate, double new_factor) {
  Sim* S = asSim(s); /* src/sim.hpp:1417-1425 */
  if (S->growthFactor == -1.0)
    S->growthFactor = (new_factor > 0.0) ? new_factor : 0.01;
  else if (new_factor > 0.0)
    S->growthFactor = new_factor;
  S->growthRate = rate;
  if (S->growthFactor > 0.0)
    for (int i = 0; i < S->numReal; ++i) S->scale[i] = S->targetScale[i] * S->growthFactor;
  return 0;
}
double pdref_get_growth_factor(void* s) { return asSim(s)->growthFactor; }
double pdref_get_growth_rate(void* s) { return asSim(s)->growthRate; }

int pdref_num_particles(void* s) { return asSim(s)->numReal; }
#define GETTER3(name, field)                                       \
  int pdref_##name(void* s, double* out, int n) {                  \
    Sim* S = asSim(s);                                             \
    for (int i = 0; i < S->numReal && i < n; ++i) {                \
      out[3 * i] = S->field[i].x;                                  \
      out[3 * i + 1] = S->field[i].y;                              \
      out[3 * i + 2] = S->field[i].z;                              \
    }                                                              \
    return S->numReal;                                             \
  }
GETTER3(get_positions, pos)
GETTER3(get_velocities, vel)
GETTER3(get_angular_velocities, angVel)
GETTER3(get_inv_inertia, invI)
#undef GETTER3
int pdref_get_quaternions(void* s, double* out, int n) {
  Sim* S = asSim(s);
  for (int i = 0; i < S->numReal && i < n; ++i) {
    out[4 * i] = S->quat[i].x; out[4 * i + 1] = S->quat[i].y;
    out[4 * i + 2] = S->quat[i].z; out[4 * i + 3] = S->quat[i].w;
  }
  return S->numReal;
}
int pdref_get_inv_mass(void* s, double* out, int n) {
  Sim* S = asSim(s);
  for (int i = 0; i < S->numReal && i < n; ++i) out[i] = S->invMass[i];
  return S->numReal;
}
int pdref_get_scales(void* s, double* out, int n) {
  Sim* S = asSim(s);
  for (int i = 0; i < S->numReal && i < n; ++i) out[i] = S->scale[i];
  return S->numReal;
}

/* ------------------------------------------------------------------------------------------ */
/* integration — src/integration.hpp                                                           */
/* ------------------------------------------------------------------------------------------ */
/* src/sim.hpp:61-70 */
static real maxOwnedRadius(const Sim* S) {
  if (S->numReal <= 0) return S->globalScale * S->baseRadius;
  real mx = -DBL_MAX;
  for (int i = 0; i < S->numReal; ++i) mx = S->scale[i] > mx ? S->scale[i] : mx;
  return mx * S->globalScale * S->baseRadius;
}

/* src/integration.hpp:44-93 */
static void predictVelocity(Sim* S) {
  const real dt = S->dt;
#pragma omp parallel for num_threads(nthreads(S->numReal)) schedule(static)
  for (int i = 0; i < S->numReal; ++i) {
    const real invM = S->invMass[i];
    V3 v = S->vel[i];
    if (invM > 0.0) {
      v = add3(v, scale3(S->gravity, dt));
      v = add3(v, scale3(S->extForce[i], invM * dt));
    }
    S->velPred[i] = v;
    V3 wpred = S->angVel[i];
    if (invM > 0.0) {
      const V3 invI = S->invI[i];
      if (invI.x > 0.0 || invI.y > 0.0 || invI.z > 0.0) {
        const Q4 q = S->quat[i];
        V3 wb = invRotate(q, S->angVel[i]);
        const V3 Ib = mk3((invI.x > 1e-9) ? 1.0 / invI.x : 0.0, (invI.y > 1e-9) ? 1.0 / invI.y : 0.0,
                          (invI.z > 1e-9) ? 1.0 / invI.z : 0.0);
        const V3 Lb = mk3(Ib.x * wb.x, Ib.y * wb.y, Ib.z * wb.z);
        const V3 wxL = cross3(wb, Lb);
        const V3 alpha = mk3(-invI.x * wxL.x, -invI.y * wxL.y, -invI.z * wxL.z);
        wb = add3(wb, scale3(alpha, dt));
        wpred = rotate(q, wb);
      }
    }
    S->angVelPred[i] = wpred;
    V3 x = S->pos[i];
    if (invM > 0.0) x = add3(x, scale3(v, dt));
    S->posPred[i] = x;
    S->quatPred[i] = S->quat[i];
    S->deltaVel[i] = mk3(0, 0, 0);
    S->deltaAngVel[i] = mk3(0, 0, 0);
    S->deltaPos[i] = mk3(0, 0, 0);
    S->counts[i] = 0;
  }
}

/* src/periodicity.hpp:43-94, canonical slot order (real ascending, then the a,b,c loops) */
static void generateGhosts(Sim* S, real skin) {
  const Domain d = S->dom;
  int top = S->numReal;
  /* first pass: exact ghost count so no ghost is ever dropped */
  int need = S->numReal;
  for (int i = 0; i < S->numReal; ++i) {
    const V3 p = S->pos[i];
    const int sx = d.px ? ((p.x < d.min.x + skin) ? 1 : ((p.x > d.max.x - skin) ? -1 : 0)) : 0;
    const int sy = d.py ? ((p.y < d.min.y + skin) ? 1 : ((p.y > d.max.y - skin) ? -1 : 0)) : 0;
    const int sz = d.pz ? ((p.z < d.min.z + skin) ? 1 : ((p.z > d.max.z - skin) ? -1 : 0)) : 0;
    need += (sx ? 2 : 1) * (sy ? 2 : 1) * (sz ? 2 : 1) - 1;
  }
  ensureSlots(S, need);
  for (int i = 0; i < S->numReal; ++i) {
    const V3 p = S->pos[i];
    const int sx = d.px ? ((p.x < d.min.x + skin) ? 1 : ((p.x > d.max.x - skin) ? -1 : 0)) : 0;
    const int sy = d.py ? ((p.y < d.min.y + skin) ? 1 : ((p.y > d.max.y - skin) ? -1 : 0)) : 0;
    const int sz = d.pz ? ((p.z < d.min.z + skin) ? 1 : ((p.z > d.max.z - skin) ? -1 : 0)) : 0;
    const int ax[2] = {0, sx}, ay[2] = {0, sy}, az[2] = {0, sz};
    const int nx = sx == 0 ? 1 : 2, ny = sy == 0 ? 1 : 2, nz = sz == 0 ? 1 : 2;
    const V3 pPred = S->posPred[i];
    for (int a = 0; a < nx; ++a)
      for (int b = 0; b < ny; ++b)
        for (int c = 0; c < nz; ++c) {
          const int ix = ax[a], iy = ay[b], iz = az[c];
          if (ix == 0 && iy == 0 && iz == 0) continue;
          const int slot = top++;
          const V3 shift = mk3(ix * d.size.x, iy * d.size.y, iz * d.size.z);
          S->pos[slot] = add3(p, shift);
          S->posPred[slot] = add3(pPred, shift);
          S->invMass[slot] = S->invMass[i];
          S->vel[slot] = S->vel[i];
          S->velPred[slot] = S->velPred[i];
          S->quat[slot] = S->quat[i];
          S->quatPred[slot] = S->quatPred[i];
          S->angVel[slot] = S->angVel[i];
          S->angVelPred[slot] = S->angVelPred[i];
          S->scale[slot] = S->scale[i];
          S->materialId[slot] = S->materialId[i];
          S->invI[slot] = mk3(0, 0, 0);
          S->realIdx[slot] = i;
        }
  }
  S->numParticles = top;
}

/* src/integration.hpp:112-148 */
static void applyVelocityAndPredictPosition(Sim* S) {
  const real dt = S->dt;
#pragma omp parallel for num_threads(nthreads(S->numParticles)) schedule(static)
  for (int i = 0; i < S->numParticles; ++i) {
    const V3 vStart = S->vel[i];
    const V3 vFinal = S->velPred[i];
    S->velPred[i] = vFinal;
    S->vel[i] = vFinal;
    const real invM = S->invMass[i];
    V3 x = S->pos[i];
    if (invM > 0.0) x = add3(x, scale3(add3(vStart, vFinal), 0.5 * dt));
    S->posPred[i] = x;
    const Q4 q = S->quat[i];
    const V3 omega = S->angVelPred[i];
    S->angVel[i] = omega;
    Q4 qp = q;
    if (invM > 0.0) {
      qp.x += 0.5 * dt * (omega.x * q.w + omega.y * q.z - omega.z * q.y);
      qp.y += 0.5 * dt * (omega.y * q.w + omega.z * q.x - omega.x * q.z);
      qp.z += 0.5 * dt * (omega.z * q.w + omega.x * q.y - omega.y * q.x);
      qp.w += 0.5 * dt * (-omega.x * q.x - omega.y * q.y - omega.z * q.z);
      const real len = sqrt(qp.x * qp.x + qp.y * qp.y + qp.z * qp.z + qp.w * qp.w);
      if (len > 1e-9) {
        const real sc = 1.0 / len;
        qp.x = qp.x * sc; qp.y = qp.y * sc; qp.z = qp.z * sc; qp.w = qp.w * sc;
      }
    }
    S->quatPred[i] = qp;
  }
}

/* src/integration.hpp:187-215 */
static void finalCommit(Sim* S) {
  const Domain d = S->dom;
#pragma omp parallel for num_threads(nthreads(S->numReal)) schedule(static)
  for (int i = 0; i < S->numReal; ++i) {
    V3 p = S->posPred[i];
    if (d.px) {
      if (p.x < d.min.x) p.x += d.size.x;
      else if (p.x >= d.max.x) p.x -= d.size.x;
    }
    if (d.py) {
      if (p.y < d.min.y) p.y += d.size.y;
      else if (p.y >= d.max.y) p.y -= d.size.y;
    }
    if (d.pz) {
      if (p.z < d.min.z) p.z += d.size.z;
      else if (p.z >= d.max.z) p.z -= d.size.z;
    }
    S->pos[i] = p;
    S->quat[i] = S->quatPred[i];
  }
}

/* src/integration.hpp:229-271 (KE summed serially in real-index order) */
static void applyThermostat(Sim* S) {
  real kt = 0.0, kr = 0.0;
  for (int i = 0; i < S->numReal; ++i) {
    const real invM = S->invMass[i];
    const V3 v = S->vel[i];
    if (invM > 0.0) kt += 0.5 * (1.0 / invM) * (v.x * v.x + v.y * v.y + v.z * v.z);
    const V3 invI = S->invI[i];
    if (invI.x > 0.0 || invI.y > 0.0 || invI.z > 0.0) {
      const V3 wb = invRotate(S->quat[i], S->angVel[i]);
      const real Ix = (invI.x > 0.0) ? 1.0 / invI.x : 0.0;
      const real Iy = (invI.y > 0.0) ? 1.0 / invI.y : 0.0;
      const real Iz = (invI.z > 0.0) ? 1.0 / invI.z : 0.0;
      kr += 0.5 * (Ix * wb.x * wb.x + Iy * wb.y * wb.y + Iz * wb.z * wb.z);
    }
  }
  const real ndof = 3.0 * S->numReal;
  real lam[2];
  const real ke[2] = {kt, kr};
  for (int k = 0; k < 2; ++k) {
    lam[k] = 1.0;
    if (ndof <= 0 || S->thermoKB <= 0) continue;
    const real Tcur = (2.0 * ke[k]) / (ndof * S->thermoKB);
    if (Tcur < 1e-6) continue;
    real f = 1.0 + (S->dt / S->thermoTau) * (S->thermoTemp / Tcur - 1.0);
    if (f < 0.0) f = 0.0;
    lam[k] = sqrt(f);
  }
  for (int i = 0; i < S->numReal; ++i) {
    S->vel[i] = scale3(S->vel[i], lam[0]);
    S->angVel[i] = scale3(S->angVel[i], lam[1]);
  }
}

/* ------------------------------------------------------------------------------------------ */
/* broadphase — src/broadphase_arborx.hpp:36-107 (closed-interval AABB overlap, i<j)          */
/* ------------------------------------------------------------------------------------------ */
typedef struct { real x, y, z, b; } Box;

/* PECLET_DEM_SANE_BOX, src/broadphase_arborx.hpp:41-52 */
static Box saneBox(const double* pos, const double* rad, int i, real margin, real boxCap) {
  Box B;
  B.x = pos[3 * i]; B.y = pos[3 * i + 1]; B.z = pos[3 * i + 2];
  B.b = rad[i] + margin;
  if (!(B.b == B.b) || !(B.x == B.x) || !(B.y == B.y) || !(B.z == B.z) ||
      (boxCap > 0.0 && B.b > boxCap)) {
    B.x = B.y = B.z = -1.0e30;
    B.b = 0.0;
  }
  return B;
}
static int boxOverlap(const Box* a, const Box* c) {
  /* ArborX Box::intersects: disjoint iff min_a > max_c or max_a < min_c on some axis */
  if (a->x - a->b > c->x + c->b || a->x + a->b < c->x - c->b) return 0;
  if (a->y - a->b > c->y + c->b || a->y + a->b < c->y - c->b) return 0;
  if (a->z - a->b > c->z + c->b || a->z + a->b < c->z - c->b) return 0;
  return 1;
}

typedef struct { int* v; int n, cap; } IntList;
static void ilPush(IntList* L, int x) {
  if (L->n == L->cap) {
    L->cap = L->cap ? 2 * L->cap : 1024;
    L->v = (int*)xrealloc(L->v, sizeof(int) * (size_t)L->cap);
  }
  L->v[L->n++] = x;
}
static void isortSmall(int* a, int n) {
  for (int i = 1; i < n; ++i) {
    int x = a[i], j = i - 1;
    while (j >= 0 && a[j] > x) { a[j + 1] = a[j]; --j; }
    a[j + 1] = x;
  }
}

int pdref_broadphase_bruteforce(const double* pos, const double* rad, int numParticles,
                                int numReal, double margin, double boxCap, int* outPairs,
                                int cap) {
  int count = 0;
  for (int i = 0; i < numReal; ++i) {
    Box bi = saneBox(pos, rad, i, margin, boxCap);
    for (int j = i + 1; j < numParticles; ++j) {
      Box bj = saneBox(pos, rad, j, margin, boxCap);
      if (boxOverlap(&bi, &bj)) {
        if (outPairs && count < cap) { outPairs[2 * count] = i; outPairs[2 * count + 1] = j; }
        ++count;
      }
    }
  }
  return count;
}

/* Uniform-grid restatement of the same criterion, pairs emitted in lexicographic order.  The cell
 * edge is >= twice the largest half-width so overlapping boxes always sit in adjacent cells. */
static int broadphaseGrid(const double* pos, const double* rad, int numParticles, int numReal,
                          real margin, real boxCap, IntList* out) {
  out->n = 0;
  if (numReal <= 0 || numParticles <= 0) return 0;
  Box* B = (Box*)xrealloc(NULL, sizeof(Box) * (size_t)numParticles);
  real lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
  real maxB = 0.0;
  for (int i = 0; i < numParticles; ++i) {
    B[i] = saneBox(pos, rad, i, margin, boxCap);
    const real c[3] = {B[i].x, B[i].y, B[i].z};
    if (B[i].b > maxB && isfinite(B[i].b)) maxB = B[i].b;
    if (B[i].x == -1.0e30) continue; /* parked */
    for (int d = 0; d < 3; ++d)
      if (isfinite(c[d])) {
        if (c[d] < lo[d]) lo[d] = c[d];
        if (c[d] > hi[d]) hi[d] = c[d];
      }
  }
  for (int d = 0; d < 3; ++d)
    if (lo[d] > hi[d]) { lo[d] = 0; hi[d] = 0; }
  real h = 2.0 * maxB * 1.0001 + 1e-300;
  int dim[3];
  for (;;) {
    double cells = 1.0;
    for (int d = 0; d < 3; ++d) {
      double nd = floor((hi[d] - lo[d]) / h) + 1.0;
      if (!(nd >= 1.0)) nd = 1.0;
      if (nd > 1e9) nd = 1e9;
      dim[d] = (int)nd;
      cells *= nd;
    }
    if (cells <= 4.0 * numParticles + 64.0) break;
    h *= 1.5;
  }
  const long ncell = (long)dim[0] * dim[1] * dim[2];
  int* cellStart = (int*)calloc((size_t)ncell + 1, sizeof(int));
  int* cellOf = (int*)xrealloc(NULL, sizeof(int) * (size_t)numParticles);
  int* order = (int*)xrealloc(NULL, sizeof(int) * (size_t)numParticles);
  int(*cc)[3] = (int(*)[3])xrealloc(NULL, sizeof(int[3]) * (size_t)numParticles);
  for (int i = 0; i < numParticles; ++i) {
    const real c[3] = {B[i].x, B[i].y, B[i].z};
    for (int d = 0; d < 3; ++d) {
      const double f = (c[d] - lo[d]) / h;
      int k;
      if (!(f >= 0.0)) k = 0;
      else if (f >= (double)(dim[d] - 1)) k = dim[d] - 1;
      else k = (int)f;
      cc[i][d] = k;
    }
    cellOf[i] = (int)(((long)cc[i][2] * dim[1] + cc[i][1]) * dim[0] + cc[i][0]);
    cellStart[cellOf[i] + 1]++;
  }
  for (long c = 0; c < ncell; ++c) cellStart[c + 1] += cellStart[c];
  int* fill = (int*)xrealloc(NULL, sizeof(int) * (size_t)ncell);
  memcpy(fill, cellStart, sizeof(int) * (size_t)ncell);
  for (int i = 0; i < numParticles; ++i) order[fill[cellOf[i]]++] = i;
  free(fill);
  IntList tmp = {0, 0, 0};
  for (int i = 0; i < numReal; ++i) {
    tmp.n = 0;
    for (int dz = -1; dz <= 1; ++dz) {
      const int z = cc[i][2] + dz;
      if (z < 0 || z >= dim[2]) continue;
      for (int dy = -1; dy <= 1; ++dy) {
        const int y = cc[i][1] + dy;
        if (y < 0 || y >= dim[1]) continue;
        for (int dx = -1; dx <= 1; ++dx) {
          const int x = cc[i][0] + dx;
          if (x < 0 || x >= dim[0]) continue;
          const long c = ((long)z * dim[1] + y) * dim[0] + x;
          for (int k = cellStart[c]; k < cellStart[c + 1]; ++k) {
            const int j = order[k];
            if (j > i && boxOverlap(&B[i], &B[j])) ilPush(&tmp, j);
          }
        }
      }
    }
    isortSmall(tmp.v, tmp.n);
    for (int k = 0; k < tmp.n; ++k) { ilPush(out, i); ilPush(out, tmp.v[k]); }
  }
  free(tmp.v); free(B); free(cellStart); free(cellOf); free(order); free(cc);
  return out->n / 2;
}

int pdref_broadphase(const double* pos, const double* rad, int numParticles, int numReal,
                     double margin, double boxCap, int* outPairs, int cap) {
  IntList L = {0, 0, 0};
  int n = broadphaseGrid(pos, rad, numParticles, numReal, margin, boxCap, &L);
  if (outPairs)
    for (int k = 0; k < n && k < cap; ++k) {
      outPairs[2 * k] = L.v[2 * k];
      outPairs[2 * k + 1] = L.v[2 * k + 1];
    }
  free(L.v);
  return n;
}

/* src/sim.hpp:82-94 (pairs are sized exactly; nothing to grow) */
static void findCollisions(Sim* S, real margin) {
  const real boxCap = fmax(fmax(S->dom.size.x, S->dom.size.y), S->dom.size.z);
  IntList L = {S->pairs, 0, S->pairCap * 2};
  S->np = broadphaseGrid((const double*)S->posPred, S->rad, S->numParticles, S->numReal, margin,
                         boxCap, &L);
  S->pairs = L.v;
  S->pairCap = L.cap / 2;
}

/* ------------------------------------------------------------------------------------------ */
/* narrowphase — src/narrowphase.hpp:179-268 (pairs) and :371-449 (planes)                     */
/* ------------------------------------------------------------------------------------------ */
static void pushContact(Sim* S, const Contact* c) {
  ensureContacts(S, S->nc + 1);
  S->contacts[S->nc++] = *c;
}
/* trilinear grid SDF with the off-grid residual added (object convention),
 * src/narrowphase.hpp:97-135 */
static real sampleGridSdf(const Sim* S, V3 p) {
  const real fx = (p.x - S->gOrigin.x) * S->gInvSpacing.x;
  const real fy = (p.y - S->gOrigin.y) * S->gInvSpacing.y;
  const real fz = (p.z - S->gOrigin.z) * S->gInvSpacing.z;
  const real cx = fmin(fmax(fx, 0.0), (real)(S->gnx - 1));
  const real cy = fmin(fmax(fy, 0.0), (real)(S->gny - 1));
  const real cz = fmin(fmax(fz, 0.0), (real)(S->gnz - 1));
  const int ix = (int)cx, iy = (int)cy, iz = (int)cz;
  const int ix1 = ix < S->gnx - 1 ? ix + 1 : ix;
  const int iy1 = iy < S->gny - 1 ? iy + 1 : iy;
  const int iz1 = iz < S->gnz - 1 ? iz + 1 : iz;
  const real tx = cx - ix, ty = cy - iy, tz = cz - iz;
  const long nxny = (long)S->gnx * S->gny;
  const real* g = S->sdfGrid;
#define AT(x, y, z) g[(long)(z) * nxny + (long)(y) * S->gnx + (x)]
  const real c00 = AT(ix, iy, iz) * (1.0 - tx) + AT(ix1, iy, iz) * tx;
  const real c10 = AT(ix, iy1, iz) * (1.0 - tx) + AT(ix1, iy1, iz) * tx;
  const real c01 = AT(ix, iy, iz1) * (1.0 - tx) + AT(ix1, iy, iz1) * tx;
  const real c11 = AT(ix, iy1, iz1) * (1.0 - tx) + AT(ix1, iy1, iz1) * tx;
#undef AT
  const real c0 = c00 * (1.0 - ty) + c10 * ty;
  const real c1 = c01 * (1.0 - ty) + c11 * ty;
  const real val = c0 * (1.0 - tz) + c1 * tz;
  const real rx = (S->gInvSpacing.x > 0.0) ? (fx - cx) / S->gInvSpacing.x : 0.0;
  const real ry = (S->gInvSpacing.y > 0.0) ? (fy - cy) / S->gInvSpacing.y : 0.0;
  const real rz = (S->gInvSpacing.z > 0.0) ? (fz - cz) / S->gInvSpacing.z : 0.0;
  return val + sqrt(rx * rx + ry * ry + rz * rz);
}
/* sdfEvalShape, src/narrowphase.hpp:169-173 */
static real sdfShape(const Sim* S, V3 p) {
  if (S->shapeType == SHAPE_GRID_SDF) return sampleGridSdf(S, p);
  return sdfEval(p, S->shapeType, S->shapeParams);
}

static void detectContacts(Sim* S, real margin) {
  const real gs = S->globalScale;
  const int countA = S->numShell;
  const int sphereA = (S->shapeType == SHAPE_SPHERE);
  const int iter = countA > 0 ? countA : 1;
  for (int idx = 0; idx < S->np; ++idx) {
    const int idA = S->pairs[2 * idx], idB = S->pairs[2 * idx + 1];
    const V3 posA = S->posPred[idA], posB = S->posPred[idB];
    const Q4 qA = S->quatPred[idA], qB = S->quatPred[idB];
    const real effScaleA = S->scale[idA] * gs, effScaleB = S->scale[idB] * gs;
    for (int k = 0; k < iter; ++k) {
      V3 pLocalA = mk3(0, 0, 0);
      real pointRadius = 0.0;
      if (countA > 0) pLocalA = S->shell[k];
      else if (sphereA) pointRadius = S->shapeParams[0] * effScaleA;
      const V3 pWorld = add3(posA, rotate(qA, scale3(pLocalA, effScaleA)));
      const V3 pLocalB = invRotate(qB, sub3(pWorld, posB));
      const V3 pCanB = scale3(pLocalB, 1.0 / effScaleB);
      const real dist = sdfShape(S, pCanB) * effScaleB;
      const real effDist = dist - pointRadius;
      if (effDist >= margin) continue;
      if (effDist < 0.0 && -effDist > S->maxOverlap) S->maxOverlap = -effDist;
      const real eps = 1e-4;
      V3 nLoc = mk3(sdfShape(S, mk3(pCanB.x + eps, pCanB.y, pCanB.z)) -
                        sdfShape(S, mk3(pCanB.x - eps, pCanB.y, pCanB.z)),
                    sdfShape(S, mk3(pCanB.x, pCanB.y + eps, pCanB.z)) -
                        sdfShape(S, mk3(pCanB.x, pCanB.y - eps, pCanB.z)),
                    sdfShape(S, mk3(pCanB.x, pCanB.y, pCanB.z + eps)) -
                        sdfShape(S, mk3(pCanB.x, pCanB.y, pCanB.z - eps)));
      const real len = len3(nLoc);
      nLoc = (len > 1e-9) ? scale3(nLoc, 1.0 / len) : mk3(0, 1, 0);
      const V3 nWorld = rotate(qB, nLoc);
      const V3 pSurfA = sub3(pWorld, scale3(nWorld, pointRadius));
      const V3 rA = sub3(pSurfA, posA);
      const V3 rB = sub3(sub3(pSurfA, scale3(nWorld, effDist)), posB);
      Contact c;
      memset(&c, 0, sizeof c);
      c.bodyA = idA; c.bodyB = idB;
      c.normal = nWorld; c.rA = rA; c.rB = rB;
      c.dist = effDist;
      c.boundaryRestitution = -1.0; c.boundaryFriction = -1.0;
      if (S->hasPairTable) {
        const int t = ((int)S->materialId[idA] * KMAXMAT + (int)S->materialId[idB]) * 2;
        c.boundaryRestitution = S->pairTable[t];
        c.boundaryFriction = S->pairTable[t + 1];
      }
      pushContact(S, &c);
    }
  }
}

static void detectBoundary(Sim* S, real margin) {
  for (int i = 0; i < S->numReal; ++i) {
    const V3 posA = S->posPred[i];
    const real s = S->scale[i] * S->globalScale;
    real baseR = S->shapeParams[0];
    if (baseR == 0.0) baseR = 1.0;
    const real radius = baseR * s;
    const int numPts = S->numShell;
    const Q4 qA = S->quatPred[i];
    for (int pi = 0; pi < S->numPlanes; ++pi) {
      const Plane pl = S->planes[pi];
      if (numPts > 0) {
        for (int k = 0; k < numPts; ++k) {
          const V3 rA = rotate(qA, scale3(S->shell[k], s));
          const V3 pwk = add3(posA, rA);
          const real dist = dot3(sub3(pwk, pl.point), pl.normal);
          if (dist >= margin) continue;
          if (dist < 0.0 && -dist > S->maxOverlap) S->maxOverlap = -dist;
          Contact c;
          memset(&c, 0, sizeof c);
          c.bodyA = i; c.bodyB = -1;
          c.normal = pl.normal; c.rA = rA; c.rB = pl.point; c.dist = dist;
          c.boundaryRestitution = -1.0; c.boundaryFriction = -1.0;
          pushContact(S, &c);
        }
      } else {
        const real dist = dot3(sub3(posA, pl.point), pl.normal) - radius;
        if (dist >= margin) continue;
        if (dist < 0.0 && -dist > S->maxOverlap) S->maxOverlap = -dist;
        Contact c;
        memset(&c, 0, sizeof c);
        c.bodyA = i; c.bodyB = -1;
        c.normal = pl.normal; c.rA = scale3(pl.normal, -radius); c.rB = pl.point; c.dist = dist;
        c.boundaryRestitution = -1.0; c.boundaryFriction = -1.0;
        pushContact(S, &c);
      }
    }
  }
}

/* ------------------------------------------------------------------------------------------ */
/* contact -> manifold reduction — src/contact_preprocessing.hpp:320-471                       */
/* ------------------------------------------------------------------------------------------ */
static uint64_t pairKey(const Contact* c) {
  const int idA = c->bodyA, idB = c->bodyB;
  if (idB < 0) return ((uint64_t)(unsigned)idA << 32) | 0xFFFFFFFFu;
  const unsigned u = (unsigned)(idA < idB ? idA : idB);
  const unsigned v = (unsigned)(idA < idB ? idB : idA);
  return ((uint64_t)u << 32) | v;
}

static Manifold transformContact(const Contact* c) {
  Manifold m;
  memset(&m, 0, sizeof m);
  const int idA = c->bodyA, idB = c->bodyB;
  if (c->dist > 0.0) { m.numPoints = 0; return m; }
  const int flip = (idB >= 0 && idB < idA);
  m.numPoints = 1;
  const V3 n_aligned = flip ? c->normal : mk3(-c->normal.x, -c->normal.y, -c->normal.z);
  const V3 shift = mk3(c->normal.x * c->dist * 0.5, c->normal.y * c->dist * 0.5,
                       c->normal.z * c->dist * 0.5);
  const V3 rA_mid = sub3(c->rA, shift);
  const V3 rB_mid = add3(c->rB, shift);
  const V3 r_can = flip ? rB_mid : rA_mid;
  const V3 r_other = flip ? rA_mid : rB_mid;
  m.normalSum = n_aligned;
  m.tauA = cross3(r_can, n_aligned);
  m.tauB = cross3(r_other, mk3(-n_aligned.x, -n_aligned.y, -n_aligned.z));
  m.rASum = r_can;
  m.rBSum = r_other;
  m.wallVelSum = c->boundaryVel;
  m.restitutionSum = c->boundaryRestitution;
  m.frictionSum = c->boundaryFriction;
  return m;
}

typedef struct { uint64_t key; int idx; } KeyIdx;
static int cmpKeyIdx(const void* a, const void* b) {
  const KeyIdx* x = (const KeyIdx*)a;
  const KeyIdx* y = (const KeyIdx*)b;
  if (x->key != y->key) return x->key < y->key ? -1 : 1;
  return x->idx < y->idx ? -1 : (x->idx > y->idx ? 1 : 0);
}

static void reduceToManifolds(Sim* S) {
  const int n = S->nc;
  S->nm = 0;
  if (n == 0) return;
  KeyIdx* ki = (KeyIdx*)xrealloc(NULL, sizeof(KeyIdx) * (size_t)n);
  for (int i = 0; i < n; ++i) { ki[i].key = pairKey(&S->contacts[i]); ki[i].idx = i; }
  qsort(ki, (size_t)n, sizeof(KeyIdx), cmpKeyIdx); /* stable by construction (idx tie-break) */
  int nseg = 0;
  for (int p = 0; p < n; ++p)
    if (p == 0 || ki[p].key != ki[p - 1].key) ++nseg;
  ensureManifolds(S, nseg);
  int seg = -1;
  for (int p = 0; p < n; ++p) {
    if (p == 0 || ki[p].key != ki[p - 1].key) {
      ++seg;
      Manifold* m = &S->manifolds[seg];
      memset(m, 0, sizeof *m);
      const uint64_t key = ki[p].key;
      const unsigned v = (unsigned)(key & 0xFFFFFFFFu);
      m->bodyA = (int)(key >> 32);
      m->bodyB = (v == 0xFFFFFFFFu) ? -1 : (int)v;
    }
    const Manifold t = transformContact(&S->contacts[ki[p].idx]);
    Manifold* m = &S->manifolds[seg];
    m->numPoints += t.numPoints;
    m->normalSum = add3(m->normalSum, t.normalSum);
    m->tauA = add3(m->tauA, t.tauA);
    m->tauB = add3(m->tauB, t.tauB);
    m->rASum = add3(m->rASum, t.rASum);
    m->rBSum = add3(m->rBSum, t.rBSum);
    m->wallVelSum = add3(m->wallVelSum, t.wallVelSum);
    m->restitutionSum += t.restitutionSum;
    m->frictionSum += t.frictionSum;
    S->contactSlot[ki[p].idx] = seg;
  }
  S->nm = nseg;
  free(ki);
}

/* ------------------------------------------------------------------------------------------ */
/* Jones-Plassmann colouring — src/solver_velocity.hpp:277-374, src/solver_position.hpp:155-241 */
/* ------------------------------------------------------------------------------------------ */
/* splitmix32 priority, src/contact_preprocessing.hpp:310-317 */
static long long colorKey(int idx) {
  unsigned z = (unsigned)idx + 0x9e3779b9u;
  z = (z ^ (z >> 16)) * 0x21f0aaadu;
  z = (z ^ (z >> 15)) * 0x735a2d97u;
  z ^= (z >> 15);
  return ((long long)(z & 0x7fffffffu) << 32) | (long long)(unsigned)idx;
}

/* Generic JP over edges (ea[e], eb[e]) on `numBodies` vertices; colour[e] == -1 enters, -2 skips.
 * Returns the colour count (max colour + 1, the reference's reduction), fills rounds and the
 * per-round uncoloured totals. */
static int jpColor(int numEdges, const int* ea, const int* eb, int numBodies, int* colour,
                   int* roundsOut, long long* uncolOut, int* leftoverOut) {
  long long* winner = (long long*)xrealloc(NULL, sizeof(long long) * (size_t)(numBodies + 1));
  uint64_t* mask = (uint64_t*)calloc((size_t)numBodies + 1, sizeof(uint64_t));
  int remaining = 1, prevRemaining = -1, rounds = 0;
  long long uncol = 0;
  const int maxRounds = numBodies + 2;
  for (int round = 0; round < maxRounds && remaining > 0; ++round) {
    for (int b = 0; b < numBodies; ++b) winner[b] = -1;
    for (int e = 0; e < numEdges; ++e) {
      if (colour[e] != -1) continue;
      ++uncol;
      const long long key = colorKey(e);
      if (key > winner[ea[e]]) winner[ea[e]] = key;
      if (eb[e] >= 0 && key > winner[eb[e]]) winner[eb[e]] = key;
    }
    int rem = 0;
    for (int e = 0; e < numEdges; ++e) {
      if (colour[e] != -1) continue;
      const long long key = colorKey(e);
      const int a = ea[e], b = eb[e];
      if (winner[a] != key || (b >= 0 && winner[b] != key)) { ++rem; continue; }
      uint64_t forbidden = mask[a];
      if (b >= 0) forbidden |= mask[b];
      int c = 0;
      while (c < 62 && (forbidden & ((uint64_t)1 << c))) ++c;
      colour[e] = c;
      mask[a] |= (uint64_t)1 << c;
      if (b >= 0) mask[b] |= (uint64_t)1 << c;
    }
    ++rounds;
    if (rem == prevRemaining) break;
    prevRemaining = rem;
    remaining = rem;
  }
  int maxc = -1, left = 0;
  for (int e = 0; e < numEdges; ++e) {
    if (colour[e] > maxc) maxc = colour[e];
    if (colour[e] == -1) ++left;
  }
  free(winner);
  free(mask);
  *roundsOut = rounds;
  *uncolOut = uncol;
  *leftoverOut = left;
  return maxc + 1;
}

static int colorManifolds(Sim* S, int* leftover) {
  *leftover = 0;
  S->stats.rounds_vel = 0;
  S->stats.uncolored_vel = 0;
  if (S->nm <= 0 || S->numReal <= 0) return 0;
  int* ea = (int*)xrealloc(NULL, sizeof(int) * (size_t)S->nm);
  int* eb = (int*)xrealloc(NULL, sizeof(int) * (size_t)S->nm);
  for (int idx = 0; idx < S->nm; ++idx) {
    const Manifold* m = &S->manifolds[idx];
    ea[idx] = S->realIdx[m->bodyA];
    eb[idx] = m->bodyB >= 0 ? S->realIdx[m->bodyB] : -1;
    if (m->numPoints <= 0) S->mColor[idx] = -2;
    else if (m->bodyB >= 0 && ea[idx] > eb[idx]) S->mColor[idx] = -2;
    else S->mColor[idx] = -1;
  }
  int nc = jpColor(S->nm, ea, eb, S->numReal, S->mColor, &S->stats.rounds_vel,
                   &S->stats.uncolored_vel, leftover);
  free(ea);
  free(eb);
  return nc;
}

static int colorContacts(Sim* S, int* leftover) {
  *leftover = 0;
  S->stats.rounds_pos = 0;
  S->stats.uncolored_pos = 0;
  if (S->nc <= 0 || S->numParticles <= 0) return 0;
  int* ea = (int*)xrealloc(NULL, sizeof(int) * (size_t)S->nc);
  int* eb = (int*)xrealloc(NULL, sizeof(int) * (size_t)S->nc);
  for (int idx = 0; idx < S->nc; ++idx) {
    ea[idx] = S->contacts[idx].bodyA;
    eb[idx] = S->contacts[idx].bodyB;
    S->cColor[idx] = -1;
  }
  int nc = jpColor(S->nc, ea, eb, S->numParticles, S->cColor, &S->stats.rounds_pos,
                   &S->stats.uncolored_pos, leftover);
  free(ea);
  free(eb);
  return nc;
}

/* edges grouped by colour class, ascending index within a class */
typedef struct { int* start; int* list; int ncol; } ColourClasses;
static void buildClasses(ColourClasses* C, const int* colour, int n, int ncol) {
  if (ncol < 0) ncol = 0;
  C->ncol = ncol;
  C->start = (int*)calloc((size_t)ncol + 1, sizeof(int));
  C->list = (int*)xrealloc(NULL, sizeof(int) * (size_t)(n > 0 ? n : 1));
  for (int e = 0; e < n; ++e)
    if (colour[e] >= 0 && colour[e] < ncol) C->start[colour[e] + 1]++;
  for (int c = 0; c < ncol; ++c) C->start[c + 1] += C->start[c];
  int* fill = (int*)xrealloc(NULL, sizeof(int) * (size_t)(ncol + 1));
  memcpy(fill, C->start, sizeof(int) * (size_t)(ncol + 1));
  for (int e = 0; e < n; ++e)
    if (colour[e] >= 0 && colour[e] < ncol) C->list[fill[colour[e]]++] = e;
  free(fill);
}
static void freeClasses(ColourClasses* C) { free(C->start); free(C->list); }

/* ------------------------------------------------------------------------------------------ */
/* warm start / persistence / grounded levels — src/contact_preprocessing.hpp:70-253           */
/* ------------------------------------------------------------------------------------------ */
static uint64_t pairKeyOf(const Sim* S, const Manifold* m) {
  const unsigned a = (unsigned)S->realIdx[m->bodyA];
  const unsigned b = (m->bodyB >= 0) ? (unsigned)S->realIdx[m->bodyB] : 0xFFFFFFFFu;
  const unsigned hi = a < b ? a : b, lo = a < b ? b : a;
  return ((uint64_t)hi << 32) | lo;
}
static int lowerBound(const uint64_t* keys, int n, uint64_t k) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (keys[mid] < k) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

/* src/contact_preprocessing.hpp:113-156 */
static void gatherWarmLambda(Sim* S) {
#pragma omp parallel for num_threads(nthreads(S->nm)) schedule(static)
  for (int idx = 0; idx < S->nm; ++idx) {
    const Manifold* m = &S->manifolds[idx];
    int dup = 0;
    if (m->numPoints > 0 && m->bodyB >= 0 && S->realIdx[m->bodyA] > S->realIdx[m->bodyB]) dup = 1;
    if (m->numPoints <= 0 || dup) {
      S->pairKeys[idx] = ~(uint64_t)0;
      S->lambdaAcc[idx] = 0.0;
      S->lambdaT[idx] = mk3(0, 0, 0);
      S->posImpulse[idx] = 0.0;
      continue;
    }
    const uint64_t k = pairKeyOf(S, m);
    S->pairKeys[idx] = k;
    const int lo = lowerBound(S->prevKeys, S->prevCount, k);
    const int hit = (lo < S->prevCount && S->prevKeys[lo] == k);
    S->lambdaAcc[idx] = hit ? S->prevLambda[lo] : 0.0;
    S->lambdaT[idx] = hit ? S->prevLambdaT[lo] : mk3(0, 0, 0);
    S->posImpulse[idx] = hit ? S->prevPosImpulse[lo] : 0.0;
  }
}

/* src/contact_preprocessing.hpp:78-109 */
static void markPersistent(Sim* S) {
  for (int idx = 0; idx < S->nm; ++idx) {
    const Manifold* m = &S->manifolds[idx];
    if (m->numPoints <= 0) {
      S->pairKeys[idx] = ~(uint64_t)0;
      S->persistent[idx] = 0;
      continue;
    }
    const uint64_t k = pairKeyOf(S, m);
    S->pairKeys[idx] = k;
    const int lo = lowerBound(S->prevKeys, S->prevCount, k);
    S->persistent[idx] = (lo < S->prevCount && S->prevKeys[lo] == k) ? 1 : 0;
  }
}

/* src/contact_preprocessing.hpp:213-253 with double-buffered sweeps */
static void updateGroundedLevels(Sim* S, V3 gHat, int sweeps, int decay) {
  int* g = S->grounded;
  for (int i = 0; i < S->numReal; ++i) {
    const int v = g[i] - decay;
    g[i] = v > 0 ? v : 0;
  }
  int* gn = (int*)xrealloc(NULL, sizeof(int) * (size_t)(S->numReal > 0 ? S->numReal : 1));
  for (int sweep = 0; sweep < sweeps; ++sweep) {
    memcpy(gn, g, sizeof(int) * (size_t)S->numReal);
    for (int idx = 0; idx < S->nm; ++idx) {
      const Manifold* m = &S->manifolds[idx];
      if (m->numPoints <= 0) continue;
      const int realA = S->realIdx[m->bodyA];
      if (m->bodyB < 0) {
        gn[realA] = 255;
        continue;
      }
      const int realB = S->realIdx[m->bodyB];
      const V3 dx = sub3(S->posPred[m->bodyA], S->posPred[m->bodyB]);
      const real up = -(dx.x * gHat.x + dx.y * gHat.y + dx.z * gHat.z);
      const real thr = 0.3 * sqrt(dot3(dx, dx));
      if (up > thr) {
        const int lvl = g[realB] - 1;
        if (lvl > 0 && lvl > gn[realA]) gn[realA] = lvl;
      } else if (up < -thr) {
        const int lvl = g[realA] - 1;
        if (lvl > 0 && lvl > gn[realB]) gn[realB] = lvl;
      }
    }
    memcpy(g, gn, sizeof(int) * (size_t)S->numReal);
  }
  free(gn);
}

/* ------------------------------------------------------------------------------------------ */
/* velocity solve — src/solver_velocity.hpp                                                    */
/* ------------------------------------------------------------------------------------------ */
/* src/solver_velocity.hpp:28-34 */
static real genInvMass(V3 tau, V3 invIlocal, Q4 q) {
  if (invIlocal.x == invIlocal.y && invIlocal.y == invIlocal.z) return dot3(tau, tau) * invIlocal.x;
  const V3 t = invRotate(q, tau);
  return t.x * t.x * invIlocal.x + t.y * t.y * invIlocal.y + t.z * t.z * invIlocal.z;
}
/* dw_world = R (invI_local * (R^T J)) */
static V3 angDelta(Q4 q, V3 J, V3 invI) {
  const V3 Jl = invRotate(q, J);
  const V3 dwl = mk3(Jl.x * invI.x, Jl.y * invI.y, Jl.z * invI.z);
  return rotate(q, dwl);
}
static void addTo(V3* v, V3 d) { v->x += d.x; v->y += d.y; v->z += d.z; }
static void addScaled(V3* v, V3 J, real s) {
  v->x += J.x * s; v->y += J.y * s; v->z += J.z * s;
}
static void addNegScaled(V3* v, V3 J, real s) {
  v->x += -J.x * s; v->y += -J.y * s; v->z += -J.z * s;
}

/* src/solver_velocity.hpp:390-447 */
static void computeVn0(Sim* S) {
#pragma omp parallel for num_threads(nthreads(S->nm)) schedule(static)
  for (int idx = 0; idx < S->nm; ++idx) {
    const Manifold* m = &S->manifolds[idx];
    if (m->numPoints <= 0) continue;
    const int idA = m->bodyA, idB = m->bodyB;
    const int realA = S->realIdx[idA];
    if (idB >= 0 && realA > S->realIdx[idB]) continue;
    const real invN = 1.0 / (real)m->numPoints;
    const V3 Nsum = m->normalSum, TauA = m->tauA, TauB = m->tauB;
    const V3 vA = S->velPred[realA], wA = S->angVelPred[realA];
    V3 vB = mk3(0, 0, 0), wB = mk3(0, 0, 0);
    if (idB >= 0) {
      vB = S->velPred[S->realIdx[idB]];
      wB = S->angVelPred[S->realIdx[idB]];
    } else {
      vB = scale3(m->wallVelSum, invN);
    }
    const V3 rAavg = scale3(m->rASum, invN);
    const V3 rBavg = scale3(m->rBSum, invN);
    const V3 diffCenters = (idB < 0) ? rAavg : sub3(rAavg, rBavg);
    const V3 vGrowth = scale3(diffCenters, S->growthRate);
    real vn = dot3(vA, Nsum) + dot3(wA, TauA) + dot3(vB, mk3(-Nsum.x, -Nsum.y, -Nsum.z)) +
              dot3(wB, TauB);
    vn += dot3(vGrowth, Nsum);
    S->vn0[idx] = vn;
    const real lenN = sqrt(dot3(Nsum, Nsum));
    V3 vt = mk3(0, 0, 0);
    if (lenN > 1e-9) {
      const V3 nhat = scale3(Nsum, 1.0 / lenN);
      const int realB2 = (idB >= 0) ? S->realIdx[idB] : -1;
      const V3 wAv = S->angVelPred[realA];
      const V3 wBv = (realB2 >= 0) ? S->angVelPred[realB2] : mk3(0, 0, 0);
      const V3 vrel = sub3(add3(vA, cross3(wAv, rAavg)), add3(vB, cross3(wBv, rBavg)));
      vt = sub3(vrel, scale3(nhat, dot3(vrel, nhat)));
    }
    S->vt0[idx] = vt;
  }
}

/* src/solver_velocity.hpp:498-605 (one manifold) */
static void warmStartApplyOne(Sim* S, int idx) {
  const real p = S->lambdaAcc[idx];
  V3 lt = S->lambdaT[idx];
  const int hasT = nonzero3(lt);
  if (p == 0.0 && !hasT) return;
  const Manifold* m = &S->manifolds[idx];
  if (m->numPoints <= 0) return;
  const int idA = m->bodyA, idB = m->bodyB;
  const int realA = S->realIdx[idA];
  const int realB = (idB >= 0) ? S->realIdx[idB] : idB;
  if (idB >= 0 && realA > realB) return;
  const real invN = 1.0 / (real)m->numPoints;
  const V3 Nsum = m->normalSum, TauA = m->tauA, TauB = m->tauB;
  const V3 rAavg = scale3(m->rASum, invN);
  const V3 rBavg = scale3(m->rBSum, invN);
  const V3 diffCenters = (idB < 0) ? rAavg : sub3(rAavg, rBavg);
  const real alignment = dot3(Nsum, diffCenters);
  const real sgn = (alignment > 0.0) ? 1.0 : -1.0;
  const real lambda = -sgn * p;
  const V3 Jlin = scale3(Nsum, lambda);
  const real invMassA = S->invMass[realA];
  const real invMassB = (idB >= 0) ? S->invMass[realB] : 0.0;
  addScaled(&S->velPred[realA], Jlin, invMassA);
  addTo(&S->angVelPred[realA], angDelta(S->quat[realA], scale3(TauA, lambda), S->invI[realA]));
  if (idB >= 0) {
    addNegScaled(&S->velPred[realB], Jlin, invMassB);
    addTo(&S->angVelPred[realB], angDelta(S->quat[realB], scale3(TauB, lambda), S->invI[realB]));
  }
  if (hasT) {
    const real lenN2 = sqrt(dot3(Nsum, Nsum));
    if (lenN2 > 1e-9) {
      const V3 nhat = scale3(Nsum, 1.0 / lenN2);
      lt = sub3(lt, scale3(nhat, dot3(lt, nhat)));
    }
    S->lambdaT[idx] = lt;
    addScaled(&S->velPred[realA], lt, invMassA);
    addTo(&S->angVelPred[realA], angDelta(S->quat[realA], cross3(rAavg, lt), S->invI[realA]));
    if (idB >= 0) {
      addNegScaled(&S->velPred[realB], lt, invMassB);
      addTo(&S->angVelPred[realB],
            angDelta(S->quat[realB], cross3(rBavg, scale3(lt, -1.0)), S->invI[realB]));
    }
  }
}

/* src/solver_velocity.hpp:456-495 */
static void computeSideFlags(Sim* S, V3 gHat, real riseThr, real approachThr) {
  for (int idx = 0; idx < S->nm; ++idx) {
    const unsigned char wasPersistent = S->persistent[idx];
    S->sideFlags[idx] = 0;
    const Manifold* m = &S->manifolds[idx];
    if (m->numPoints <= 0 || m->bodyB < 0 || wasPersistent == 0) continue;
    if (fabs(S->vn0[idx]) > approachThr) continue;
    const int realA = S->realIdx[m->bodyA], realB = S->realIdx[m->bodyB];
    if (realA > realB) continue;
    const V3 dx = sub3(S->posPred[m->bodyA], S->posPred[m->bodyB]);
    const real up = -(dx.x * gHat.x + dx.y * gHat.y + dx.z * gHat.z);
    const real thr = 0.3 * sqrt(dot3(dx, dx));
    if (up > thr && -dot3(S->velPred[realB], gHat) <= riseThr && S->grounded[realB] > 0)
      S->sideFlags[idx] = 1;
    else if (up < -thr && -dot3(S->velPred[realA], gHat) <= riseThr && S->grounded[realA] > 0)
      S->sideFlags[idx] = 2;
  }
}

/* PGSManifoldSweep::solveOne, src/solver_velocity.hpp:641-858 */
static void pgsSolveOne(Sim* S, int idx, real restVelThreshold) {
  const Manifold* m = &S->manifolds[idx];
  const int idA = m->bodyA, idB = m->bodyB;
  const int realA = S->realIdx[idA];
  const int realB = (idB >= 0) ? S->realIdx[idB] : idB;
  const real invMassA = S->invMass[realA];
  const real invMassB = (idB >= 0) ? S->invMass[realB] : 0.0;
  const V3 invIA = S->invI[realA];
  const V3 invIB = (idB >= 0) ? S->invI[realB] : mk3(0, 0, 0);
  Q4 qB; qB.x = 0; qB.y = 0; qB.z = 0; qB.w = 1;
  const Q4 qA = S->quat[realA];
  if (idB >= 0) qB = S->quat[realB];
  const V3 vA = S->velPred[realA], wA = S->angVelPred[realA];
  V3 vB = mk3(0, 0, 0), wB = mk3(0, 0, 0);
  const real invN = 1.0 / (real)m->numPoints;
  real restitution = S->restitutionNormal;
  if (idB >= 0) {
    vB = S->velPred[realB];
    wB = S->angVelPred[realB];
  } else {
    vB = scale3(m->wallVelSum, invN);
  }
  {
    const real ra = m->restitutionSum * invN;
    if (ra >= 0.0) restitution = ra;
  }
  const V3 Nsum = m->normalSum, TauA = m->tauA, TauB = m->tauB;
  const V3 rAavg = scale3(m->rASum, invN);
  const V3 rBavg = scale3(m->rBSum, invN);
  const real lenN = sqrt(dot3(Nsum, Nsum));
  if (lenN < 1e-9) return;
  const V3 diffCenters = (idB < 0) ? rAavg : sub3(rAavg, rBavg);
  const V3 vGrowth = scale3(diffCenters, S->growthRate);
  real vn = dot3(vA, Nsum) + dot3(wA, TauA) + dot3(vB, mk3(-Nsum.x, -Nsum.y, -Nsum.z)) +
            dot3(wB, TauB);
  vn += dot3(vGrowth, Nsum);
  const real alignment = dot3(Nsum, diffCenters);
  const real sgn = (alignment > 0.0) ? 1.0 : -1.0;
  const real Nsq = dot3(Nsum, Nsum);
  real wA_n = Nsq * invMassA + genInvMass(TauA, invIA, qA);
  real wB_n = Nsq * invMassB + genInvMass(TauB, invIB, qB);
  int applyA = 1, applyB = 1;
  const unsigned char sf = S->sideFlags[idx];
  if (sf == 1) { wB_n = 0.0; applyB = 0; restitution = 0.0; }
  else if (sf == 2) { wA_n = 0.0; applyA = 0; restitution = 0.0; }
  const real wTotal = wA_n + wB_n;
  if (wTotal <= 0.0) return;
  const real vn0 = S->vn0[idx];
  const real v0til = sgn * vn0;
  if (fabs(vn0) < restVelThreshold * lenN) restitution = 0.0;
  const real target = (v0til > 0.0) ? -restitution * v0til : 0.0;
  const real vtil = sgn * vn;
  const real dp = (vtil - target) / wTotal;
  const real pOld = S->lambdaAcc[idx];
  real pNew = pOld + dp;
  if (pNew < 0.0) pNew = 0.0;
  const real dApplied = pNew - pOld;
  if (dApplied != 0.0) {
    S->lambdaAcc[idx] = pNew;
    const real appr = fabs(dApplied) * wTotal / lenN;
    if (appr > S->maxApproach) S->maxApproach = appr;
    const real lambda = -sgn * dApplied;
    const V3 Jlin = scale3(Nsum, lambda);
    const V3 JangA = scale3(TauA, lambda);
    const V3 JangB = scale3(TauB, lambda);
    if (applyA) {
      addScaled(&S->velPred[realA], Jlin, invMassA);
      addTo(&S->angVelPred[realA], angDelta(qA, JangA, invIA));
    }
    if (idB >= 0 && applyB) {
      addNegScaled(&S->velPred[realB], Jlin, invMassB);
      addTo(&S->angVelPred[realB], angDelta(qB, JangB, invIB));
    }
  }
  /* friction cone, src/solver_velocity.hpp:760-855 */
  real mu = S->frictionDynamic;
  const real fa = m->frictionSum * invN;
  if (fa >= 0.0) mu = fa;
  const V3 ltOld = S->lambdaT[idx];
  const int haveOld = nonzero3(ltOld);
  if (!(mu > 0.0 || haveOld)) return;
  const V3 nhat = scale3(Nsum, 1.0 / lenN);
  const V3 vA2 = S->velPred[realA], wA2 = S->angVelPred[realA];
  V3 vB2 = mk3(0, 0, 0), wB2 = mk3(0, 0, 0);
  if (idB >= 0) {
    vB2 = S->velPred[realB];
    wB2 = S->angVelPred[realB];
  } else {
    vB2 = scale3(m->wallVelSum, invN);
  }
  const V3 vrel = sub3(add3(vA2, cross3(wA2, rAavg)), add3(vB2, cross3(wB2, rBavg)));
  const V3 vt = sub3(vrel, scale3(nhat, dot3(vrel, nhat)));
  real beta = S->restitutionTangent;
  if (beta != 0.0 && (sf != 0 || fabs(vn0) < restVelThreshold * lenN)) beta = 0.0;
  V3 vtErr = vt;
  if (beta != 0.0) {
    V3 v0 = S->vt0[idx];
    v0 = sub3(v0, scale3(nhat, dot3(v0, nhat)));
    vtErr = add3(vt, scale3(v0, beta));
  }
  const real vtLen = sqrt(dot3(vtErr, vtErr));
  V3 ltNew = ltOld;
  real wT = invMassA + invMassB;
  if (vtLen > 1e-9) {
    const V3 that = scale3(vtErr, 1.0 / vtLen);
    const real wAt = (sf == 2) ? 0.0 : invMassA + genInvMass(cross3(rAavg, that), invIA, qA);
    const real wBt =
        (sf == 1 || idB < 0) ? 0.0 : invMassB + genInvMass(cross3(rBavg, that), invIB, qB);
    wT = wAt + wBt;
    if (wT > 1e-9) ltNew = add3(ltOld, scale3(that, -vtLen / wT));
  }
  ltNew = sub3(ltNew, scale3(nhat, dot3(ltNew, nhat)));
  real nTot = fmax(S->lambdaAcc[idx], 0.0) * lenN;
  if (fabs(vn0) < restVelThreshold * lenN) nTot += fmax(S->posImpulse[idx], 0.0);
  const real bound = mu * nTot;
  const real ltLen = sqrt(dot3(ltNew, ltNew));
  if (ltLen > bound) ltNew = (bound > 0.0) ? scale3(ltNew, bound / ltLen) : mk3(0, 0, 0);
  const V3 dApp = sub3(ltNew, ltOld);
  if (nonzero3(dApp)) {
    S->lambdaT[idx] = ltNew;
    const real dLen = sqrt(dot3(dApp, dApp));
    if (wT > 1e-9 && dLen * wT > S->maxApproach) S->maxApproach = dLen * wT;
    if (applyA) {
      addScaled(&S->velPred[realA], dApp, invMassA);
      addTo(&S->angVelPred[realA], angDelta(qA, cross3(rAavg, dApp), invIA));
    }
    if (idB >= 0 && applyB) {
      addNegScaled(&S->velPred[realB], dApp, invMassB);
      addTo(&S->angVelPred[realB], angDelta(qB, cross3(rBavg, scale3(dApp, -1.0)), invIB));
    }
  }
}

/* src/solver_velocity.hpp:989-1133 (one manifold of the one-shot coloured GS, g = 0 path) */
static void gsSolveOne(Sim* S, int idx, real restVelThreshold) {
  const Manifold* m = &S->manifolds[idx];
  const int idA = m->bodyA, idB = m->bodyB;
  const int realA = S->realIdx[idA];
  const int realB = (idB >= 0) ? S->realIdx[idB] : idB;
  const real invMassA = S->invMass[realA];
  const real invMassB = (idB >= 0) ? S->invMass[realB] : 0.0;
  const V3 invIA = S->invI[realA];
  const V3 invIB = (idB >= 0) ? S->invI[realB] : mk3(0, 0, 0);
  Q4 qB; qB.x = 0; qB.y = 0; qB.z = 0; qB.w = 1;
  const Q4 qA = S->quat[realA];
  if (idB >= 0) qB = S->quat[realB];
  const V3 vA = S->velPred[realA], wA = S->angVelPred[realA];
  V3 vB = mk3(0, 0, 0), wB = mk3(0, 0, 0);
  if (idB >= 0) {
    vB = S->velPred[realB];
    wB = S->angVelPred[realB];
  }
  const V3 Nsum = m->normalSum, TauA = m->tauA, TauB = m->tauB;
  const real invN = 1.0 / (real)m->numPoints;
  real restitution = S->restitutionNormal;
  if (idB < 0) vB = scale3(m->wallVelSum, invN);
  {
    const real ra = m->restitutionSum * invN;
    if (ra >= 0.0) restitution = ra;
  }
  const V3 rAavg = scale3(m->rASum, invN);
  const V3 rBavg = scale3(m->rBSum, invN);
  const real lenN = sqrt(dot3(Nsum, Nsum));
  if (lenN < 1e-9) return;
  const V3 diffCenters = (idB < 0) ? rAavg : sub3(rAavg, rBavg);
  const V3 vGrowth = scale3(diffCenters, S->growthRate);
  real vn = dot3(vA, Nsum) + dot3(wA, TauA) + dot3(vB, mk3(-Nsum.x, -Nsum.y, -Nsum.z)) +
            dot3(wB, TauB);
  vn += dot3(vGrowth, Nsum);
  const real alignment = dot3(Nsum, diffCenters);
  if (alignment > 0.0) {
    if (vn < 0.0) return;
  } else {
    if (vn > 0.0) return;
  }
  const real Nsq = dot3(Nsum, Nsum);
  const real wA_n = Nsq * invMassA + genInvMass(TauA, invIA, qA);
  const real wB_n = Nsq * invMassB + genInvMass(TauB, invIB, qB);
  const real wTotal = wA_n + wB_n;
  if (wTotal <= 0.0) return;
  const real appr = fabs(vn) / lenN;
  if (appr > S->maxApproach) S->maxApproach = appr;
  if (fabs(vn) < restVelThreshold * lenN) restitution = 0.0;
  const real lambda = (-restitution * vn - vn) / wTotal;
  const V3 Jlin = scale3(Nsum, lambda);
  addScaled(&S->velPred[realA], Jlin, invMassA);
  addTo(&S->angVelPred[realA], angDelta(qA, scale3(TauA, lambda), invIA));
  if (idB >= 0) {
    addNegScaled(&S->velPred[realB], Jlin, invMassB);
    addTo(&S->angVelPred[realB], angDelta(qB, scale3(TauB, lambda), invIB));
  }
}

/* Jacobi restitution, src/solver_velocity.hpp:38-224 (serial accumulation in manifold order);
 * onlyUncoloured = the colour-saturation fallback pass (colour == -1 only), src/sim.hpp:245-256 */
static void solveVelocityJacobi(Sim* S, real restVelThreshold, int onlyUncoloured) {
  for (int idx = 0; idx < S->nm; ++idx) {
    if (onlyUncoloured && S->mColor[idx] != -1) continue;
    const Manifold* m = &S->manifolds[idx];
    if (m->numPoints <= 0) continue;
    const int idA = m->bodyA, idB = m->bodyB;
    const int realA = S->realIdx[idA];
    int realB = idB;
    if (idB >= 0) {
      realB = S->realIdx[idB];
      if (realA > realB) continue;
    }
    const real invMassA = S->invMass[realA];
    const real invMassB = (idB >= 0) ? S->invMass[realB] : 0.0;
    const V3 invIA = S->invI[realA];
    const V3 invIB = (idB >= 0) ? S->invI[realB] : mk3(0, 0, 0);
    Q4 qB; qB.x = 0; qB.y = 0; qB.z = 0; qB.w = 1;
    const Q4 qA = S->quat[realA];
    if (idB >= 0) qB = S->quat[realB];
    const V3 vA = S->velPred[realA], wA = S->angVelPred[realA];
    V3 vB = mk3(0, 0, 0), wB = mk3(0, 0, 0);
    if (idB >= 0) { vB = S->velPred[realB]; wB = S->angVelPred[realB]; }
    const V3 Nsum = m->normalSum, TauA = m->tauA, TauB = m->tauB;
    const real invN = 1.0 / (real)m->numPoints;
    real restitution = S->restitutionNormal;
    if (idB < 0) vB = scale3(m->wallVelSum, invN);
    {
      const real ra = m->restitutionSum * invN;
      if (ra >= 0.0) restitution = ra;
    }
    const V3 rAavg = scale3(m->rASum, invN);
    const V3 rBavg = scale3(m->rBSum, invN);
    const real lenN = sqrt(dot3(Nsum, Nsum));
    if (lenN < 1e-9) continue;
    const V3 diffCenters = (idB < 0) ? rAavg : sub3(rAavg, rBavg);
    const V3 vGrowth = scale3(diffCenters, S->growthRate);
    real vn = dot3(vA, Nsum) + dot3(wA, TauA) + dot3(vB, mk3(-Nsum.x, -Nsum.y, -Nsum.z)) +
              dot3(wB, TauB);
    vn += dot3(vGrowth, Nsum);
    const real alignment = dot3(Nsum, diffCenters);
    if (alignment > 0.0) {
      if (vn < 0.0) continue;
    } else {
      if (vn > 0.0) continue;
    }
    const real Nsq = dot3(Nsum, Nsum);
    const real wTotal = (Nsq * invMassA + genInvMass(TauA, invIA, qA)) +
                        (Nsq * invMassB + genInvMass(TauB, invIB, qB));
    if (wTotal <= 0.0) continue;
    if (fabs(vn) < restVelThreshold * lenN) restitution = 0.0;
    const real lambda = (-restitution * vn - vn) / wTotal;
    const V3 Jlin = scale3(Nsum, lambda);
    addScaled(&S->deltaVel[realA], Jlin, invMassA);
    addTo(&S->deltaAngVel[realA], angDelta(qA, scale3(TauA, lambda), invIA));
    if (idB >= 0) {
      addNegScaled(&S->deltaVel[realB], Jlin, invMassB);
      addTo(&S->deltaAngVel[realB], angDelta(qB, scale3(TauB, lambda), invIB));
      S->counts[realB]++;
    }
    S->counts[realA]++;
  }
}
/* src/solver_velocity.hpp:233-255 */
static void applyVelocityDeltasAveraged(Sim* S) {
  for (int i = 0; i < S->numParticles; ++i) {
    const int count = S->counts[i];
    if (count <= 0) continue;
    const real f = fmin(1.0, 2.0 / (real)count);
    S->velPred[i].x += S->deltaVel[i].x * f;
    S->angVelPred[i].x += S->deltaAngVel[i].x * f;
    S->velPred[i].y += S->deltaVel[i].y * f;
    S->angVelPred[i].y += S->deltaAngVel[i].y * f;
    S->velPred[i].z += S->deltaVel[i].z * f;
    S->angVelPred[i].z += S->deltaAngVel[i].z * f;
    S->deltaVel[i] = mk3(0, 0, 0);
    S->deltaAngVel[i] = mk3(0, 0, 0);
    S->counts[i] = 0;
  }
}

/* src/contact_preprocessing.hpp:160-192 (stable sort) */
static void commitPairKeysLambda(Sim* S) {
  const int n = S->nm;
  if (n <= 0) return;
  ensureLedger(S, n);
  KeyIdx* ki = (KeyIdx*)xrealloc(NULL, sizeof(KeyIdx) * (size_t)n);
  for (int i = 0; i < n; ++i) { ki[i].key = S->pairKeys[i]; ki[i].idx = i; }
  qsort(ki, (size_t)n, sizeof(KeyIdx), cmpKeyIdx);
  for (int i = 0; i < n; ++i) {
    S->prevKeys[i] = ki[i].key;
    S->prevLambda[i] = S->lambdaAcc[ki[i].idx];
    S->prevLambdaT[i] = S->lambdaT[ki[i].idx];
  }
  free(ki);
}

/* ------------------------------------------------------------------------------------------ */
/* legacy friction cluster — src/solver_friction.hpp (serial accumulation in contact order)    */
/* ------------------------------------------------------------------------------------------ */
/* src/solver_position.hpp:21-24 */
static real computeW(V3 r, V3 dir, real invM, V3 invI) {
  const V3 rn = cross3(r, dir);
  return invM + rn.x * rn.x * invI.x + rn.y * rn.y * invI.y + rn.z * rn.z * invI.z;
}
/* src/solver_friction.hpp:87-120; planeFriction kept per slot (load, count) */
static void computePlaneLoad(Sim* S, real* pfLoad, real* pfCount) {
  for (int i = 0; i < S->numParticles; ++i) { pfLoad[i] = 0.0; pfCount[i] = 0.0; }
  for (int idx = 0; idx < S->nc; ++idx) {
    Contact* c = &S->contacts[idx];
    if (c->bodyB >= 0) continue;
    const int idA = c->bodyA;
    const real invMA = S->invMass[idA];
    if (invMA <= 0.0) continue;
    const V3 vAc = add3(S->velPred[idA], cross3(S->angVelPred[idA], c->rA));
    const real approach = -dot3(sub3(vAc, c->boundaryVel), c->normal);
    if (approach <= 0.0) continue;
    const real w_n = computeW(c->rA, c->normal, invMA, S->invI[idA]);
    c->frictionLambdaN = (w_n > 1e-6) ? approach / w_n : 0.0;
    if (approach / invMA > pfLoad[idA]) pfLoad[idA] = approach / invMA;
  }
}
/* src/solver_friction.hpp:32-85 */
static void accumulateNormalImpulse(Sim* S) {
  for (int idx = 0; idx < S->nc; ++idx) {
    Contact* c = &S->contacts[idx];
    if (c->bodyB < 0) {
      const int idA = c->bodyA;
      const real invMA = S->invMass[idA];
      if (invMA <= 0.0) continue;
      const V3 vAc = add3(S->velPred[idA], cross3(S->angVelPred[idA], c->rA));
      const real approach = -dot3(sub3(vAc, c->boundaryVel), c->normal);
      if (approach <= 0.0) continue;
      const real w_n = computeW(c->rA, c->normal, invMA, S->invI[idA]);
      if (w_n > 1e-6) c->frictionLambdaN += approach / w_n;
      continue;
    }
    const int realA = S->realIdx[c->bodyA], realB = S->realIdx[c->bodyB];
    if (realA > realB) continue;
    const real invMA = S->invMass[realA], invMB = S->invMass[realB];
    const V3 vAc = add3(S->velPred[realA], cross3(S->angVelPred[realA], c->rA));
    const V3 vBc = add3(S->velPred[realB], cross3(S->angVelPred[realB], c->rB));
    const V3 vrel = add3(sub3(vAc, vBc), scale3(sub3(c->rA, c->rB), S->growthRate));
    const real approach = -dot3(vrel, c->normal);
    if (approach <= 0.0) continue;
    const real w_n = computeW(c->rA, c->normal, invMA, S->invI[realA]) +
                     computeW(c->rB, c->normal, invMB, S->invI[realB]);
    if (w_n > 1e-6) c->frictionLambdaN += approach / w_n;
  }
}
/* src/solver_friction.hpp:123-140 */
static void countFrictionContacts(Sim* S, real* pfCount) {
  for (int idx = 0; idx < S->nc; ++idx) {
    const Contact* c = &S->contacts[idx];
    if (c->frictionLambdaN <= 0.0) continue;
    pfCount[S->realIdx[c->bodyA]] += 1.0;
    if (c->bodyB >= 0) pfCount[S->realIdx[c->bodyB]] += 1.0;
  }
}
/* src/solver_friction.hpp:142-220 + applyVelocityDeltas src/integration.hpp:97-109 */
static void solveContactFriction(Sim* S, const real* pfCount) {
  for (int i = 0; i < S->numParticles; ++i) {
    S->deltaVel[i] = mk3(0, 0, 0);
    S->deltaAngVel[i] = mk3(0, 0, 0);
  }
  for (int idx = 0; idx < S->nc; ++idx) {
    const Contact* c = &S->contacts[idx];
    const real lambda_n = c->frictionLambdaN;
    if (lambda_n <= 0.0) continue;
    const int idA = c->bodyA, idB = c->bodyB;
    const int realA = S->realIdx[idA];
    const int realB = (idB >= 0) ? S->realIdx[idB] : -1;
    const real invMA = S->invMass[realA];
    const real invMB = (idB >= 0) ? S->invMass[realB] : 0.0;
    const V3 invIA = S->invI[realA];
    const V3 invIB = (idB >= 0) ? S->invI[realB] : mk3(0, 0, 0);
    const V3 rA = c->rA, rB = c->rB, n = c->normal;
    const V3 vAc = add3(S->velPred[realA], cross3(S->angVelPred[realA], rA));
    V3 vBc;
    if (idB >= 0) vBc = add3(S->velPred[realB], cross3(S->angVelPred[realB], rB));
    else vBc = c->boundaryVel;
    const V3 vrel = sub3(vAc, vBc);
    const real vn = dot3(vrel, n);
    const V3 vt = sub3(vrel, scale3(n, vn));
    const real vt_len = sqrt(dot3(vt, vt));
    if (vt_len < 1e-7) continue;
    const V3 t = scale3(vt, 1.0 / vt_len);
    const V3 rnA = cross3(rA, t), rnB = cross3(rB, t);
    const real w_t = invMA + invMB + rnA.x * rnA.x * invIA.x + rnA.y * rnA.y * invIA.y +
                     rnA.z * rnA.z * invIA.z + rnB.x * rnB.x * invIB.x +
                     rnB.y * rnB.y * invIB.y + rnB.z * rnB.z * invIB.z;
    if (w_t < 1e-6) continue;
    const real bound = lambda_n;
    const real nA = pfCount[realA];
    const real nB = (idB >= 0) ? pfCount[realB] : 0.0;
    const real inv_n = 1.0 / fmax(fmax(nA, nB), 1.0);
    real lt = -vt_len / w_t;
    const real mu = (c->boundaryFriction >= 0.0) ? c->boundaryFriction : S->frictionDynamic;
    const real maxf = mu * bound;
    if (lt < -maxf) lt = -maxf;
    lt *= inv_n;
    S->deltaVel[realA].x += t.x * lt * invMA;
    S->deltaVel[realA].y += t.y * lt * invMA;
    S->deltaVel[realA].z += t.z * lt * invMA;
    S->deltaAngVel[realA].x += rnA.x * invIA.x * lt;
    S->deltaAngVel[realA].y += rnA.y * invIA.y * lt;
    S->deltaAngVel[realA].z += rnA.z * invIA.z * lt;
    if (idB >= 0) {
      S->deltaVel[realB].x += -t.x * lt * invMB;
      S->deltaVel[realB].y += -t.y * lt * invMB;
      S->deltaVel[realB].z += -t.z * lt * invMB;
      S->deltaAngVel[realB].x += -rnB.x * invIB.x * lt;
      S->deltaAngVel[realB].y += -rnB.y * invIB.y * lt;
      S->deltaAngVel[realB].z += -rnB.z * invIB.z * lt;
    }
  }
  for (int i = 0; i < S->numParticles; ++i) {
    S->velPred[i] = add3(S->velPred[i], S->deltaVel[i]);
    S->angVelPred[i] = add3(S->angVelPred[i], S->deltaAngVel[i]);
    S->deltaVel[i] = mk3(0, 0, 0);
    S->deltaAngVel[i] = mk3(0, 0, 0);
  }
}

/* ------------------------------------------------------------------------------------------ */
/* position solve — src/solver_position.hpp                                                    */
/* ------------------------------------------------------------------------------------------ */
/* linearised constraint shared by the GS and Jacobi forms, src/solver_position.hpp:276-312 */
static int positionConstraint(const Sim* S, const Contact* c, V3* nOut, V3* rAOut, V3* rBOut,
                              real* COut) {
  const int idA = c->bodyA, idB = c->bodyB;
  const V3 pA = S->posPred[idA];
  const Q4 qA = S->quatPred[idA];
  V3 pB = mk3(0, 0, 0);
  Q4 qB; qB.x = 0; qB.y = 0; qB.z = 0; qB.w = 1;
  if (idB >= 0) { pB = S->posPred[idB]; qB = S->quatPred[idB]; }
  const Q4 qAdelta = quatMult(qA, quatInverse(S->quat[idA]));
  V3 rA = rotate(qAdelta, c->rA);
  V3 rB = c->rB;
  V3 n = c->normal;
  if (idB >= 0) {
    const Q4 qBdelta = quatMult(qB, quatInverse(S->quat[idB]));
    rB = rotate(qBdelta, rB);
    n = rotate(qBdelta, n);
  }
  real C;
  if (idB < 0) {
    n = c->normal;
    C = dot3(sub3(add3(pA, rA), c->rB), n);
  } else {
    C = dot3(sub3(add3(pA, rA), add3(pB, rB)), n);
  }
  *nOut = n; *rAOut = rA; *rBOut = rB; *COut = C;
  return 1;
}

/* src/solver_position.hpp:259-326 (one contact) */
static void positionSolveOne(Sim* S, int idx) {
  const Contact* c = &S->contacts[idx];
  const int idA = c->bodyA, idB = c->bodyB;
  const real invMassA = S->invMass[idA];
  const real invMassB = (idB >= 0) ? S->invMass[idB] : 0.0;
  V3 n, rA, rB;
  real C;
  positionConstraint(S, c, &n, &rA, &rB, &C);
  if (C >= 0.0) return;
  const V3 invIA = S->invI[idA];
  const V3 invIB = (idB >= 0) ? S->invI[idB] : mk3(0, 0, 0);
  const real wTotal = computeW(rA, n, invMassA, invIA) + computeW(rB, n, invMassB, invIB);
  if (wTotal < 1e-6) return;
  const real dLambda = -C / wTotal;
  S->posLambda[idx] += dLambda;
  S->posPred[idA].x += n.x * dLambda * invMassA;
  S->posPred[idA].y += n.y * dLambda * invMassA;
  S->posPred[idA].z += n.z * dLambda * invMassA;
  if (idB >= 0) {
    S->posPred[idB].x += -n.x * dLambda * invMassB;
    S->posPred[idB].y += -n.y * dLambda * invMassB;
    S->posPred[idB].z += -n.z * dLambda * invMassB;
  }
  if (-C > S->maxOverlap) S->maxOverlap = -C;
}

/* Jacobi XPBD, src/solver_position.hpp:34-132 + applyUpdates src/integration.hpp:168-184;
 * onlyUncoloured = the colour-saturation fallback pass (colour == -1 only), src/sim.hpp:454-462 */
static void positionSolveJacobi(Sim* S, int onlyUncoloured) {
  for (int idx = 0; idx < S->nc; ++idx) {
    if (onlyUncoloured && S->cColor[idx] != -1) continue;
    const Contact* c = &S->contacts[idx];
    const int idA = c->bodyA, idB = c->bodyB;
    const real invMassA = S->invMass[idA];
    const real invMassB = (idB >= 0) ? S->invMass[idB] : 0.0;
    V3 n, rA, rB;
    real C;
    positionConstraint(S, c, &n, &rA, &rB, &C);
    if (C >= 0.0) continue;
    const V3 invIA = S->invI[idA];
    const V3 invIB = (idB >= 0) ? S->invI[idB] : mk3(0, 0, 0);
    const real wTotal = computeW(rA, n, invMassA, invIA) + computeW(rB, n, invMassB, invIB);
    if (wTotal < 1e-6) continue;
    const real dLambda = -C / wTotal;
    S->deltaPos[idA].x += n.x * dLambda * invMassA;
    S->deltaPos[idA].y += n.y * dLambda * invMassA;
    S->deltaPos[idA].z += n.z * dLambda * invMassA;
    if (idB >= 0) {
      S->deltaPos[idB].x += -n.x * dLambda * invMassB;
      S->deltaPos[idB].y += -n.y * dLambda * invMassB;
      S->deltaPos[idB].z += -n.z * dLambda * invMassB;
      S->counts[idB]++;
    }
    S->counts[idA]++;
    if (-C > S->maxOverlap) S->maxOverlap = -C;
  }
  for (int i = 0; i < S->numParticles; ++i) {
    const int count = S->counts[i];
    if (count <= 0) continue;
    const real f = 1.0 / (real)count;
    S->posPred[i] = add3(S->posPred[i], scale3(S->deltaPos[i], f));
    S->velPred[i] = add3(S->velPred[i], scale3(S->deltaVel[i], f));
    S->deltaPos[i] = mk3(0, 0, 0);
    S->deltaVel[i] = mk3(0, 0, 0);
    S->counts[i] = 0;
  }
}

/* src/contact_preprocessing.hpp:501-538 (per-manifold sums in contact order; duplicate keys:
 * the highest manifold index writes last) */
static void commitPosImpulse(Sim* S) {
  for (int m = 0; m < S->nm; ++m) S->vn0[m] = 0.0;
  for (int i = 0; i < S->nc; ++i) {
    const real l = S->posLambda[i];
    if (l != 0.0) S->vn0[S->contactSlot[i]] += l / S->dt;
  }
  for (int idx = 0; idx < S->nm; ++idx) {
    const uint64_t k = S->pairKeys[idx];
    if (k == ~(uint64_t)0) continue;
    const int lo = lowerBound(S->prevKeys, S->prevCount, k);
    if (lo < S->prevCount && S->prevKeys[lo] == k) S->prevPosImpulse[lo] = S->vn0[idx];
  }
}

/* ------------------------------------------------------------------------------------------ */
/* the substep — src/sim.hpp:97-488                                                            */
/* ------------------------------------------------------------------------------------------ */
/* One colour class of rows: the rows of a class (below colour 62) touch disjoint bodies, so they
 * are split over the threads; each thread keeps its own running maxima (a shallow copy of the
 * Sim: the arrays are shared) and the maxima are combined with max afterwards. */
typedef void (*RowFn)(Sim* S, int idx, real thr);
static void runClass(Sim* S, const int* list, int n, int colour, RowFn fn, real thr) {
  const int nt = colour == 62 ? 1 : nthreads(n);
  if (nt <= 1) {
    for (int k = 0; k < n; ++k) fn(S, list[k], thr);
    return;
  }
  real ma = S->maxApproach, mo = S->maxOverlap, mq = S->maxApproachQS;
#pragma omp parallel num_threads(nt)
  {
    Sim L = *S;
#pragma omp for schedule(static)
    for (int k = 0; k < n; ++k) fn(&L, list[k], thr);
#pragma omp critical(pdref_class_max)
    {
      if (L.maxApproach > ma) ma = L.maxApproach;
      if (L.maxOverlap > mo) mo = L.maxOverlap;
      if (L.maxApproachQS > mq) mq = L.maxApproachQS;
    }
  }
  S->maxApproach = ma;
  S->maxOverlap = mo;
  S->maxApproachQS = mq;
}
static void warmRow(Sim* S, int idx, real thr) { (void)thr; warmStartApplyOne(S, idx); }
static void posRow(Sim* S, int idx, real thr) { (void)thr; positionSolveOne(S, idx); }
static void runClasses(Sim* S, const ColourClasses* C, RowFn fn, real thr) {
  for (int c = 0; c < C->ncol; ++c)
    runClass(S, C->list + C->start[c], C->start[c + 1] - C->start[c], c, fn, thr);
}

/* ------------------------------------------------------------------------------------------ */
/* stabilization mode "multilevel", src/solver_multilevel.hpp (GraphMG-style aggregation) and  */
/* its driver src/sim.hpp:303-355.                                                             */
/* Canonical orders where the reference sums with atomics: a group's mass and momentum are     */
/* summed over the aggregation tree (a level-l group = its leader's level-(l-1) sum plus its   */
/* mate's), and the rows of the coarse colour 62, which may share a group, run in ascending    */
/* manifold order.                                                                             */
/* ------------------------------------------------------------------------------------------ */
#define ML_MAX_LEVELS 10 /* kMlMaxLevels, src/solver_multilevel.hpp:44 */
#define ML_SLOT_SKIP 63  /* kMlSlotSkip, :45 */
#define ML_GATE_PERSISTENT 1 /* kGatePersistent, :82 */
#define ML_GATE_CONE 2
#define ML_GATE_SLIP 4
typedef struct {
  int numLevels;
  int groupOff[ML_MAX_LEVELS], numGroups[ML_MAX_LEVELS], parentOff[ML_MAX_LEVELS];
  int numColors[ML_MAX_LEVELS];
  unsigned char* slot; /* [level][nm]: the manifold's colour at that level, ML_SLOT_SKIP = none */
  int* parent;         /* parent maps (finer-level id -> group id), pooled by level */
  int* child;          /* [2 per group]: its leader and its mate (-1 for a singleton) */
  real *massG, *invMassG;
  V3 *velG, *velG0;
  V3* mom;             /* momentum tree: n bodies, then the pooled groups */
  int* grp;            /* real body -> group of the current level */
} MlHier;
/* mldetail::effMass, src/solver_multilevel.hpp:68-70 */
static real mlEffMass(real invMass) { return 1.0 / fmax(invMass, 1e-30); }
/* mldetail::eligible, src/solver_multilevel.hpp:88-119 */
static int mlEligible(const Sim* S, int idx, real qsThr, int gateMask, V3 gHat) {
  const Manifold* m = &S->manifolds[idx];
  if (m->numPoints <= 0 || S->mColor[idx] < 0) return 0;
  if ((gateMask & ML_GATE_PERSISTENT) && S->persistent[idx] == 0) return 0;
  if (fabs(S->vn0[idx]) > qsThr) return 0;
  if (gateMask & ML_GATE_SLIP) {
    const V3 vt = S->vt0[idx];
    if (dot3(vt, vt) > qsThr * qsThr) return 0;
  }
  if (m->bodyB < 0) return 1;
  if (gateMask & ML_GATE_CONE) {
    const V3 dx = sub3(S->posPred[m->bodyA], S->posPred[m->bodyB]);
    const real up = -(dx.x * gHat.x + dx.y * gHat.y + dx.z * gHat.z);
    return fabs(up) > 0.3 * sqrt(dot3(dx, dx));
  }
  return 1;
}
static void mlFree(MlHier* H) {
  free(H->slot); free(H->parent); free(H->child); free(H->massG); free(H->invMassG);
  free(H->velG); free(H->velG0); free(H->mom); free(H->grp);
  memset(H, 0, sizeof *H);
}
/* buildContactHierarchyKokkos, src/solver_multilevel.hpp:126-347: per level one round of
 * random-priority matching (a manifold joins its two groups when its key is the largest at
 * both), leader-ordered compaction, the group masses, and the colouring of the level's crossing
 * manifolds by rounds of the same arbitration (at most 64 rounds; a round that colours nothing
 * more ends it).  The pools are sized by the capacity as the reference's (5 and 4 x capacity,
 * src/particles.hpp:244-248). */
static void mlBuild(Sim* S, MlHier* H, real qsThr, int gateMask, V3 gHat) {
  memset(H, 0, sizeof *H);
  const int nm = S->nm, n = S->numReal;
  if (nm <= 0 || n <= 0) return;
  const int parentCap = 5 * S->capacity, groupCap = 4 * S->capacity;
  H->slot = (unsigned char*)xrealloc(NULL, (size_t)nm * ML_MAX_LEVELS);
  memset(H->slot, ML_SLOT_SKIP, (size_t)nm * ML_MAX_LEVELS);
  H->parent = (int*)xrealloc(NULL, sizeof(int) * (size_t)(parentCap > 0 ? parentCap : 1));
  H->child = (int*)xrealloc(NULL, sizeof(int) * 2 * (size_t)(groupCap > 0 ? groupCap : 1));
  H->massG = (real*)xrealloc(NULL, sizeof(real) * (size_t)(groupCap > 0 ? groupCap : 1));
  H->invMassG = (real*)xrealloc(NULL, sizeof(real) * (size_t)(groupCap > 0 ? groupCap : 1));
  H->velG = (V3*)xrealloc(NULL, sizeof(V3) * (size_t)(groupCap > 0 ? groupCap : 1));
  H->velG0 = (V3*)xrealloc(NULL, sizeof(V3) * (size_t)(groupCap > 0 ? groupCap : 1));
  H->mom = (V3*)xrealloc(NULL, sizeof(V3) * ((size_t)n + (size_t)(groupCap > 0 ? groupCap : 1)));
  H->grp = (int*)xrealloc(NULL, sizeof(int) * (size_t)n);
  long long* winner = (long long*)xrealloc(NULL, sizeof(long long) * (size_t)n);
  int* mate = (int*)xrealloc(NULL, sizeof(int) * (size_t)n);
  uint64_t* mask = (uint64_t*)xrealloc(NULL, sizeof(uint64_t) * (size_t)n);
  unsigned char* elig = (unsigned char*)xrealloc(NULL, (size_t)nm);
  for (int idx = 0; idx < nm; ++idx)
    elig[idx] = (unsigned char)mlEligible(S, idx, qsThr, gateMask, gHat);
  for (int i = 0; i < n; ++i) H->grp[i] = i;
  int ngPrev = n, parentOff = 0, groupOff = 0;
  for (int lvl = 1; lvl <= ML_MAX_LEVELS; ++lvl) {
    if (parentOff + ngPrev > parentCap) break;
    /* matching, :169-207 */
    for (int g = 0; g < ngPrev; ++g) winner[g] = -1;
    for (int idx = 0; idx < nm; ++idx) {
      const Manifold* m = &S->manifolds[idx];
      if (m->bodyB < 0 || !elig[idx]) continue;
      const int gA = H->grp[S->realIdx[m->bodyA]], gB = H->grp[S->realIdx[m->bodyB]];
      if (gA == gB) continue;
      const long long key = colorKey(idx);
      if (key > winner[gA]) winner[gA] = key;
      if (key > winner[gB]) winner[gB] = key;
    }
    for (int g = 0; g < ngPrev; ++g) mate[g] = g;
    for (int idx = 0; idx < nm; ++idx) {
      const Manifold* m = &S->manifolds[idx];
      if (m->bodyB < 0 || !elig[idx]) continue;
      const int gA = H->grp[S->realIdx[m->bodyA]], gB = H->grp[S->realIdx[m->bodyB]];
      if (gA == gB) continue;
      const long long key = colorKey(idx);
      if (winner[gA] == key && winner[gB] == key) {
        mate[gA] = gB;
        mate[gB] = gA;
      }
    }
    /* compaction, :209-232: a matched pair's leader (the smaller id) and singletons, in id order */
    int ngNew = 0;
    int* parent = H->parent + parentOff;
    for (int g = 0; g < ngPrev; ++g)
      if (mate[g] >= g) parent[g] = ngNew++;
    for (int g = 0; g < ngPrev; ++g)
      if (mate[g] < g) parent[g] = parent[mate[g]];
    if (ngNew >= ngPrev || ngNew > (9 * ngPrev) / 10) break;
    if (groupOff + ngNew > groupCap) break;
    for (int g = 0; g < ngPrev; ++g)
      if (mate[g] >= g) {
        int* c = H->child + 2 * (size_t)(groupOff + parent[g]);
        c[0] = g;
        c[1] = mate[g] != g ? mate[g] : -1;
      }
    for (int i = 0; i < n; ++i) H->grp[i] = parent[H->grp[i]];
    /* group masses, :243-256 (summed over the tree) */
    for (int g = 0; g < ngNew; ++g) {
      const int* c = H->child + 2 * (size_t)(groupOff + g);
      const int prevOff = lvl == 1 ? -1 : H->groupOff[lvl - 2];
      real mass = prevOff < 0 ? mlEffMass(S->invMass[c[0]]) : H->massG[prevOff + c[0]];
      if (c[1] >= 0) mass += prevOff < 0 ? mlEffMass(S->invMass[c[1]]) : H->massG[prevOff + c[1]];
      H->massG[groupOff + g] = mass;
      H->invMassG[groupOff + g] = 1.0 / mass;
    }
    /* colouring of the crossing manifolds, :257-327 */
    unsigned char* slot = H->slot + (size_t)(lvl - 1) * nm;
    for (int g = 0; g < ngNew; ++g) mask[g] = 0;
    int remaining = 1, prevRemaining = -1, maxc = -1;
    for (int round = 0; round < 64 && remaining > 0; ++round) {
      for (int g = 0; g < ngNew; ++g) winner[g] = -1;
      for (int idx = 0; idx < nm; ++idx) {
        if (!elig[idx] || slot[idx] != ML_SLOT_SKIP) continue;
        const Manifold* m = &S->manifolds[idx];
        const int gA = H->grp[S->realIdx[m->bodyA]];
        const int gB = (m->bodyB >= 0) ? H->grp[S->realIdx[m->bodyB]] : -1;
        if (gB == gA) continue;
        const long long key = colorKey(idx);
        if (key > winner[gA]) winner[gA] = key;
        if (gB >= 0 && key > winner[gB]) winner[gB] = key;
      }
      int rem = 0;
      for (int idx = 0; idx < nm; ++idx) {
        if (!elig[idx] || slot[idx] != ML_SLOT_SKIP) continue;
        const Manifold* m = &S->manifolds[idx];
        const int gA = H->grp[S->realIdx[m->bodyA]];
        const int gB = (m->bodyB >= 0) ? H->grp[S->realIdx[m->bodyB]] : -1;
        if (gB == gA) continue;
        const long long key = colorKey(idx);
        if (winner[gA] != key || (gB >= 0 && winner[gB] != key)) {
          ++rem;
          continue;
        }
        uint64_t forbidden = mask[gA];
        if (gB >= 0) forbidden |= mask[gB];
        int c = 0;
        while (c < ML_SLOT_SKIP - 1 && (forbidden & ((uint64_t)1 << c))) ++c;
        slot[idx] = (unsigned char)c;
        mask[gA] |= (uint64_t)1 << c;
        if (gB >= 0) mask[gB] |= (uint64_t)1 << c;
      }
      if (rem == prevRemaining) break;
      prevRemaining = rem;
      remaining = rem;
    }
    for (int idx = 0; idx < nm; ++idx)
      if (slot[idx] != ML_SLOT_SKIP && slot[idx] > maxc) maxc = slot[idx];
    H->numColors[lvl - 1] = maxc + 1;
    H->groupOff[lvl - 1] = groupOff;
    H->numGroups[lvl - 1] = ngNew;
    H->parentOff[lvl - 1] = parentOff;
    H->numLevels = lvl;
    parentOff += ngPrev;
    groupOff += ngNew;
    ngPrev = ngNew;
    if (ngNew <= 32) break;
  }
  free(winner); free(mate); free(mask); free(elig);
}
/* One coarse row, the body of ml_coarse_pgs (src/solver_multilevel.hpp:405-466): translation
 * only, e = 0, on the shared impulse accumulator. */
static void mlCoarseRow(Sim* S, MlHier* H, int lvl, int idx) {
  const int off = H->groupOff[lvl - 1];
  const Manifold* m = &S->manifolds[idx];
  const int gA = H->grp[S->realIdx[m->bodyA]];
  const int gB = (m->bodyB >= 0) ? H->grp[S->realIdx[m->bodyB]] : -1;
  const real invN = 1.0 / (real)m->numPoints;
  const V3 Nsum = m->normalSum;
  const real lenN = sqrt(dot3(Nsum, Nsum));
  if (lenN < 1e-9) return;
  const V3 rAavg = scale3(m->rASum, invN);
  const V3 rBavg = scale3(m->rBSum, invN);
  const V3 diffCenters = (gB < 0) ? rAavg : sub3(rAavg, rBavg);
  V3* velG = H->velG + off;
  const V3 vA = velG[gA];
  const V3 vB = (gB >= 0) ? velG[gB] : scale3(m->wallVelSum, invN);
  const real vn = dot3(sub3(vA, vB), Nsum);
  const real alignment = dot3(Nsum, diffCenters);
  const real sgn = (alignment > 0.0) ? 1.0 : -1.0;
  const real invMA = H->invMassG[off + gA];
  const real invMB = (gB >= 0) ? H->invMassG[off + gB] : 0.0;
  const real w = dot3(Nsum, Nsum) * (invMA + invMB);
  if (w <= 0.0) return;
  const real vtil = sgn * vn;
  const real dp = vtil / w;
  const real pOld = S->lambdaAcc[idx];
  real pNew = pOld + dp;
  if (pNew < 0.0) pNew = 0.0;
  const real d = pNew - pOld;
  if (d == 0.0) return;
  S->lambdaAcc[idx] = pNew;
  const real appr = fabs(d) * w / lenN;
  if (appr > S->maxApproachQS) S->maxApproachQS = appr;
  const real lambda = -sgn * d;
  const V3 J = scale3(Nsum, lambda);
  velG[gA].x += J.x * invMA;
  velG[gA].y += J.y * invMA;
  velG[gA].z += J.z * invMA;
  if (gB >= 0) {
    velG[gB].x -= J.x * invMB;
    velG[gB].y -= J.y * invMB;
    velG[gB].z -= J.z * invMB;
  }
}
/* multilevelCoarseCycleKokkos, src/solver_multilevel.hpp:354-483: per level, compose the body
 * map, restrict (V = sum m v / sum m, summed over the tree), `coarseSweeps` coloured coarse
 * sweeps, prolongate (every member takes its group's velocity change). */
static void mlCoarseCycle(Sim* S, MlHier* H, int coarseSweeps) {
  const int n = S->numReal, nm = S->nm;
  for (int i = 0; i < n; ++i) H->grp[i] = i;
  for (int lvl = 1; lvl <= H->numLevels; ++lvl) {
    const int off = H->groupOff[lvl - 1], ng = H->numGroups[lvl - 1];
    const int* parent = H->parent + H->parentOff[lvl - 1];
    for (int i = 0; i < n; ++i) H->grp[i] = parent[H->grp[i]];
    /* restrict: momenta of the bodies, then up the tree to this level */
    for (int i = 0; i < n; ++i) {
      const real m = mlEffMass(S->invMass[i]);
      H->mom[i] = mk3(m * S->velPred[i].x, m * S->velPred[i].y, m * S->velPred[i].z);
    }
    for (int l = 1; l <= lvl; ++l) {
      const int lo = H->groupOff[l - 1];
      const V3* below = l == 1 ? H->mom : H->mom + n + H->groupOff[l - 2];
      for (int g = 0; g < H->numGroups[l - 1]; ++g) {
        const int* c = H->child + 2 * (size_t)(lo + g);
        V3 p = below[c[0]];
        if (c[1] >= 0) p = add3(p, below[c[1]]);
        H->mom[n + lo + g] = p;
      }
    }
    V3* velG = H->velG + off;
    V3* velG0 = H->velG0 + off;
    for (int g = 0; g < ng; ++g) {
      const real invM = 1.0 / H->massG[off + g];
      const V3 p = H->mom[n + off + g];
      velG[g] = mk3(p.x * invM, p.y * invM, p.z * invM);
      velG0[g] = velG[g];
    }
    const unsigned char* slot = H->slot + (size_t)(lvl - 1) * nm;
    for (int s = 0; s < coarseSweeps; ++s)
      for (int color = 0; color < H->numColors[lvl - 1]; ++color)
        for (int idx = 0; idx < nm; ++idx)
          if (slot[idx] == color) mlCoarseRow(S, H, lvl, idx);
    for (int i = 0; i < n; ++i) {
      const int g = H->grp[i];
      S->velPred[i].x += velG[g].x - velG0[g].x;
      S->velPred[i].y += velG[g].y - velG0[g].y;
      S->velPred[i].z += velG[g].z - velG0[g].z;
    }
  }
}
/* the fine smoothing sweep of the multilevel loop tracks, besides the full residual, the
 * quasi-static one: corrections of rows with |vn0| <= 4 vRest |N| (src/solver_velocity.hpp:
 * 721-723, 827-831) */
static void pgsSolveOneQS(Sim* S, int idx, real restVelThreshold) {
  const real before = S->maxApproach;
  S->maxApproach = 0.0;
  pgsSolveOne(S, idx, restVelThreshold);
  const real a = S->maxApproach;
  S->maxApproach = a > before ? a : before;
  const V3 Nsum = S->manifolds[idx].normalSum;
  if (fabs(S->vn0[idx]) <= 4.0 * restVelThreshold * sqrt(dot3(Nsum, Nsum)) && a > S->maxApproachQS)
    S->maxApproachQS = a;
}

static void fillRad(Sim* S, int n) {
  for (int i = 0; i < n; ++i) S->rad[i] = S->scale[i] * S->globalScale * S->baseRadius;
}

static void detectWallSdf(Sim* S, real margin);
static real wallFrictionMax(const Sim* S);
static void orderedStabilization(Sim* S, V3 gHat, real vRestS);

static int demStep(Sim* S) {
  memset(&S->stats, 0, sizeof S->stats);
  /* growth ramp, src/sim.hpp:100-108 */
  if (S->growthFactor != -1.0 && S->growthRate != 0.0) {
    S->growthFactor *= exp(S->growthRate * S->dt);
    if (S->growthFactor > 1.0) S->growthFactor = 1.0;
  }
  if (S->growthFactor > 0.0)
    for (int i = 0; i < S->numReal; ++i) S->scale[i] = S->targetScale[i] * S->growthFactor;

  const real maxRad = maxOwnedRadius(S);
  const real margin = 0.1 * maxRad;
  predictVelocity(S);
  for (int i = 0; i < S->numReal; ++i) S->realIdx[i] = i;
  generateGhosts(S, maxRad);
  ensureGrounded(S, S->numReal);
  fillRad(S, S->numParticles);
  findCollisions(S, margin);

  S->nc = 0;
  S->maxOverlap = 0.0;
  detectContacts(S, margin);
  detectBoundary(S, margin);
  detectWallSdf(S, margin);
  ensureContacts(S, S->nc);
  reduceToManifolds(S);
  ensureManifolds(S, S->nm);
  const int nm = S->nm, nc = S->nc;

  /* a frictional wall drives friction even for frictionless bodies, src/sim.hpp:169-170 */
  const int friction = (S->frictionDynamic > 0.0 || wallFrictionMax(S) > 0.0);
  const int usePersist = (S->gravity.x != 0.0 || S->gravity.y != 0.0 || S->gravity.z != 0.0);
  const int legacyFriction = friction && !(usePersist && S->useGS);
  real* pfLoad = NULL;
  real* pfCount = NULL;
  if (legacyFriction) {
    pfLoad = (real*)xrealloc(NULL, sizeof(real) * (size_t)(S->numParticles + 1));
    pfCount = (real*)xrealloc(NULL, sizeof(real) * (size_t)(S->numParticles + 1));
    computePlaneLoad(S, pfLoad, pfCount);
  }

  int velLeftover = 0;
  const int numColors = S->useGS ? colorManifolds(S, &velLeftover) : 0;
  S->numColorsVel = numColors;
  ColourClasses VC;
  buildClasses(&VC, S->mColor, nm, S->useGS ? numColors : 0);

  const int usePGS = usePersist && S->useGS;
  const real gMagP = sqrt(S->gravity.x * S->gravity.x + S->gravity.y * S->gravity.y +
                          S->gravity.z * S->gravity.z);
  const V3 gHat = usePersist ? mk3(S->gravity.x / gMagP, S->gravity.y / gMagP, S->gravity.z / gMagP)
                             : mk3(0, 0, 0);
  if (usePGS) {
    gatherWarmLambda(S);
    markPersistent(S);
    updateGroundedLevels(S, gHat, 8, 8);
    memset(S->sideFlags, 0, (size_t)nm);
    computeVn0(S);
    runClasses(S, &VC, warmRow, 0.0);
  }
  for (int it = 0; it < S->velIters; ++it) {
    if (legacyFriction) accumulateNormalImpulse(S);
    const real vRest = 2.0 * S->dt *
                       sqrt(S->gravity.x * S->gravity.x + S->gravity.y * S->gravity.y +
                            S->gravity.z * S->gravity.z);
    if (S->useGS) {
      S->maxApproach = 0.0;
      runClasses(S, &VC, usePGS ? pgsSolveOne : gsSolveOne, vRest);
      if (velLeftover > 0) {
        solveVelocityJacobi(S, vRest, 1);
        applyVelocityDeltasAveraged(S);
      }
      S->stats.sweeps_vel++;
      if (S->maxApproach <= (usePGS ? 0.02 * vRest : vRest)) break;
    } else {
      solveVelocityJacobi(S, vRest, 0);
      applyVelocityDeltasAveraged(S);
      S->stats.sweeps_vel++;
    }
  }
  if (usePGS) {
    const real vRestS = 2.0 * S->dt * gMagP;
    const int smode = S->stabMode;
    if (smode != 0 && S->maxApproach > vRestS) {
      S->stats.stab_ran = 1;
      if (smode == 1) {
        computeSideFlags(S, gHat, 8.0 * S->dt * gMagP, 8.0 * S->dt * gMagP);
        for (int it = 0; it < 2 * S->velIters; ++it) {
          S->maxApproach = 0.0;
          runClasses(S, &VC, pgsSolveOne, vRestS);
          S->stats.sweeps_stab++;
          if (S->maxApproach <= vRestS) break;
        }
      } else if (smode == 2) {
        /* src/sim.hpp:303-355: hierarchy once, then fine sweep + coarse cycle until the
         * quasi-static residual drains */
        MlHier H;
        mlBuild(S, &H, 8.0 * S->dt * gMagP, S->mlGates, gHat);
        S->stats.ml_levels = H.numLevels;
        for (int it = 0; it < 2 * S->velIters; ++it) {
          S->maxApproachQS = 0.0;
          runClasses(S, &VC, pgsSolveOneQS, vRestS);
          S->stats.sweeps_stab++;
          if (H.numLevels > 0) mlCoarseCycle(S, &H, 2);
          if (S->maxApproachQS <= vRestS) break;
        }
        mlFree(&H);
      } else if (smode == 3) {
        for (int it = 0; it < 256; ++it) {
          S->maxApproach = 0.0;
          runClasses(S, &VC, pgsSolveOne, vRestS);
          S->stats.sweeps_stab++;
          if (S->maxApproach <= vRestS) break;
        }
      } else if (smode == 4) {
        orderedStabilization(S, gHat, vRestS);
      }
    }
  }
  if (usePGS) {
    commitPairKeysLambda(S);
    S->prevCount = nm;
  }
  if (legacyFriction) {
    countFrictionContacts(S, pfCount);
    solveContactFriction(S, pfCount);
  }
  free(pfLoad);
  free(pfCount);
  freeClasses(&VC);

  applyVelocityAndPredictPosition(S);

  int posLeftover = 0;
  const int numPosColors = S->useGS ? colorContacts(S, &posLeftover) : 0;
  S->numColorsPos = numPosColors;
  ColourClasses PC;
  buildClasses(&PC, S->cColor, nc, S->useGS ? numPosColors : 0);
  const real posTol = 1e-4 * S->baseRadius * S->globalScale;
  for (int i = 0; i < nc; ++i) S->posLambda[i] = 0.0;
  for (int it = 0; it < S->posIters; ++it) {
    if (S->useGS) {
      S->maxOverlap = 0.0;
      runClasses(S, &PC, posRow, 0.0);
      if (posLeftover > 0) positionSolveJacobi(S, 1);
      S->stats.sweeps_pos++;
      if (S->maxOverlap < posTol) break;
    } else {
      positionSolveJacobi(S, 0);
      S->stats.sweeps_pos++;
    }
  }
  freeClasses(&PC);
  if (usePGS && nm > 0) commitPosImpulse(S);
  finalCommit(S);
  if (S->thermoTau > 0.0 && S->dt > 0.0) applyThermostat(S);

  S->stats.num_real = S->numReal;
  S->stats.num_particles = S->numParticles;
  S->stats.num_pairs = S->np;
  S->stats.num_contacts = S->nc;
  S->stats.num_manifolds = S->nm;
  int act = 0;
  for (int i = 0; i < nm; ++i) act += S->mColor[i] >= 0;
  S->stats.active_manifolds = S->useGS ? act : 0;
  S->stats.num_colors_vel = numColors > 0 ? numColors : 0;
  S->stats.num_colors_pos = numPosColors > 0 ? numPosColors : 0;
  S->stats.max_overlap = S->maxOverlap;
  S->stats.max_approach = S->maxApproach;
  return 0;
}

int pdref_step(void* s, double dt) {
  Sim* S = asSim(s); /* src/sim.hpp:1456-1459 */
  S->dt = dt;
  return demStep(S);
}

/* src/sim.hpp:494-541 */
int pdref_compute_overlaps(void* s, double* out) {
  Sim* S = asSim(s);
  const real maxRad = maxOwnedRadius(S);
  const real margin = 0.1 * maxRad;
  S->numParticles = S->numReal;
  for (int i = 0; i < S->slotCap; ++i) {
    S->posPred[i] = S->pos[i];
    S->quatPred[i] = S->quat[i];
  }
  for (int i = 0; i < S->numReal; ++i) S->realIdx[i] = i;
  generateGhosts(S, maxRad);
  fillRad(S, S->numParticles);
  findCollisions(S, margin);
  S->nc = 0;
  S->maxOverlap = 0.0;
  detectContacts(S, margin);
  detectBoundary(S, margin);
  detectWallSdf(S, margin);
  S->numParticles = S->numReal;
  if (out) *out = S->maxOverlap;
  return 0;
}

int pdref_num_contacts(void* s) { return asSim(s)->nc; }
int pdref_num_manifolds(void* s) { return asSim(s)->nm; }
double pdref_max_overlap(void* s) { return asSim(s)->maxOverlap; }
int pdref_get_stats(void* s, pdref_stats* out) {
  *out = asSim(s)->stats;
  return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* diagnostics                                                                                 */
/* ------------------------------------------------------------------------------------------ */
int pdref_get_pairs(void* s, int* out, int cap) {
  Sim* S = asSim(s);
  if (out)
    for (int k = 0; k < S->np && k < cap; ++k) {
      out[2 * k] = S->pairs[2 * k];
      out[2 * k + 1] = S->pairs[2 * k + 1];
    }
  return S->np;
}
/* ids: (bodyA, bodyB); data: normal(3), rA(3), rB(3), dist */
int pdref_get_contacts(void* s, int* ids, double* data, int cap) {
  Sim* S = asSim(s);
  for (int k = 0; k < S->nc && k < cap; ++k) {
    const Contact* c = &S->contacts[k];
    if (ids) { ids[2 * k] = c->bodyA; ids[2 * k + 1] = c->bodyB; }
    if (data) {
      double* d = data + 10 * k;
      d[0] = c->normal.x; d[1] = c->normal.y; d[2] = c->normal.z;
      d[3] = c->rA.x; d[4] = c->rA.y; d[5] = c->rA.z;
      d[6] = c->rB.x; d[7] = c->rB.y; d[8] = c->rB.z;
      d[9] = c->dist;
    }
  }
  return S->nc;
}
int pdref_get_contact_colors(void* s, int* out, int cap) {
  Sim* S = asSim(s);
  if (out)
    for (int k = 0; k < S->nc && k < cap; ++k) out[k] = S->cColor[k];
  return S->nc;
}
/* ids: (bodyA, bodyB, numPoints, realA, realB); data: normalSum, tauA, tauB, rASum, rBSum (15) */
int pdref_get_manifolds(void* s, int* ids, double* data, int cap) {
  Sim* S = asSim(s);
  for (int k = 0; k < S->nm && k < cap; ++k) {
    const Manifold* m = &S->manifolds[k];
    if (ids) {
      ids[5 * k] = m->bodyA;
      ids[5 * k + 1] = m->bodyB;
      ids[5 * k + 2] = m->numPoints;
      ids[5 * k + 3] = S->realIdx[m->bodyA];
      ids[5 * k + 4] = m->bodyB >= 0 ? S->realIdx[m->bodyB] : -1;
    }
    if (data) {
      double* d = data + 15 * k;
      const V3 f[5] = {m->normalSum, m->tauA, m->tauB, m->rASum, m->rBSum};
      for (int j = 0; j < 5; ++j) { d[3 * j] = f[j].x; d[3 * j + 1] = f[j].y; d[3 * j + 2] = f[j].z; }
    }
  }
  return S->nm;
}
int pdref_get_manifold_colors(void* s, int* out, int cap) {
  Sim* S = asSim(s);
  if (out)
    for (int k = 0; k < S->nm && k < cap; ++k) out[k] = S->mColor[k];
  return S->nm;
}
int pdref_get_ledger(void* s, uint64_t* keys, double* lambda, double* lambdaT, double* posImp,
                     int cap) {
  Sim* S = asSim(s);
  for (int k = 0; k < S->prevCount && k < cap; ++k) {
    if (keys) keys[k] = S->prevKeys[k];
    if (lambda) lambda[k] = S->prevLambda[k];
    if (lambdaT) {
      lambdaT[3 * k] = S->prevLambdaT[k].x;
      lambdaT[3 * k + 1] = S->prevLambdaT[k].y;
      lambdaT[3 * k + 2] = S->prevLambdaT[k].z;
    }
    if (posImp) posImp[k] = S->prevPosImpulse[k];
  }
  return S->prevCount;
}
int pdref_get_grounded(void* s, int* out, int n) {
  Sim* S = asSim(s);
  for (int i = 0; i < S->numReal && i < n && i < S->groundedCap; ++i) out[i] = S->grounded[i];
  return S->numReal;
}

/* ------------------------------------------------------------------------------------------ */
/* SDF walls — src/sim.hpp:1216-1287; trilinear sampling src/narrowphase.hpp:137-165            */
/* ------------------------------------------------------------------------------------------ */
int pdref_add_sdf_wall(void* s, const double* grid, int nx, int ny, int nz, double ox, double oy,
                       double oz, double sx, double sy, double sz, double restitution,
                       double friction) {
  Sim* S = asSim(s);
  if (nx < 2 || ny < 2 || nz < 2) return fail("addSdfWall: grid dims must be >= 2 on each axis");
  if (S->numWalls >= MAXWALLS) return fail("addSdfWall: at most 8 walls");
  const long cells = (long)nx * ny * nz;
  Wall* w = &S->walls[S->numWalls];
  memset(w, 0, sizeof *w);
  w->nx = nx; w->ny = ny; w->nz = nz;
  w->off = S->wallGridN;
  w->origin = mk3(ox, oy, oz);
  w->invSpacing = mk3(sx > 0 ? 1.0 / sx : 0.0, sy > 0 ? 1.0 / sy : 0.0, sz > 0 ? 1.0 / sz : 0.0);
  w->restitution = restitution;
  w->friction = friction;
  w->materialId = -1;
  S->wallGrid = (real*)xrealloc(S->wallGrid, sizeof(real) * (size_t)(S->wallGridN + cells));
  memcpy(S->wallGrid + S->wallGridN, grid, sizeof(real) * (size_t)cells);
  S->wallGridN += cells;
  return S->numWalls++;
}
int pdref_set_wall_velocity(void* s, int wid, double lx, double ly, double lz, double ax,
                            double ay, double az, double cx, double cy, double cz) {
  Sim* S = asSim(s); /* src/sim.hpp:1280-1287 */
  if (wid < 0 || wid >= S->numWalls) return fail("setWallVelocity: wall index out of range");
  S->walls[wid].linVel = mk3(lx, ly, lz);
  S->walls[wid].angVel = mk3(ax, ay, az);
  S->walls[wid].center = mk3(cx, cy, cz);
  return 0;
}
int pdref_set_wall_material_id(void* s, int wid, int mat) {
  Sim* S = asSim(s); /* src/sim.hpp:1217-1224 */
  if (wid < 0 || wid >= S->numWalls) return badarg("wall index out of range");
  S->walls[wid].materialId = mat;
  return 0;
}
int pdref_num_walls(void* s) { return asSim(s)->numWalls; }

static real sampleWallSdf(V3 p, const Wall* w, const real* grid) {
  const real fx = (p.x - w->origin.x) * w->invSpacing.x;
  const real fy = (p.y - w->origin.y) * w->invSpacing.y;
  const real fz = (p.z - w->origin.z) * w->invSpacing.z;
  const real cx = fmin(fmax(fx, 0.0), (real)(w->nx - 1));
  const real cy = fmin(fmax(fy, 0.0), (real)(w->ny - 1));
  const real cz = fmin(fmax(fz, 0.0), (real)(w->nz - 1));
  const int ix = (int)cx, iy = (int)cy, iz = (int)cz;
  const int ix1 = ix < w->nx - 1 ? ix + 1 : ix;
  const int iy1 = iy < w->ny - 1 ? iy + 1 : iy;
  const int iz1 = iz < w->nz - 1 ? iz + 1 : iz;
  const real tx = cx - ix, ty = cy - iy, tz = cz - iz;
  const long nxny = (long)w->nx * w->ny;
  const real* g = grid + w->off;
#define AT(x, y, z) g[(long)(z) * nxny + (long)(y) * w->nx + (x)]
  const real c00 = AT(ix, iy, iz) * (1.0 - tx) + AT(ix1, iy, iz) * tx;
  const real c10 = AT(ix, iy1, iz) * (1.0 - tx) + AT(ix1, iy1, iz) * tx;
  const real c01 = AT(ix, iy, iz1) * (1.0 - tx) + AT(ix1, iy, iz1) * tx;
  const real c11 = AT(ix, iy1, iz1) * (1.0 - tx) + AT(ix1, iy1, iz1) * tx;
#undef AT
  const real c0 = c00 * (1.0 - ty) + c10 * ty;
  const real c1 = c01 * (1.0 - ty) + c11 * ty;
  const real val = c0 * (1.0 - tz) + c1 * tz;
  const real rx = (w->invSpacing.x > 0.0) ? (fx - cx) / w->invSpacing.x : 0.0;
  const real ry = (w->invSpacing.y > 0.0) ? (fy - cy) / w->invSpacing.y : 0.0;
  const real rz = (w->invSpacing.z > 0.0) ? (fz - cz) / w->invSpacing.z : 0.0;
  return val - sqrt(rx * rx + ry * ry + rz * rz);
}
static V3 wallGradient(V3 p, const Wall* w, const real* grid) {
  const real hx = 0.5 / w->invSpacing.x;
  return mk3(sampleWallSdf(mk3(p.x + hx, p.y, p.z), w, grid) -
                 sampleWallSdf(mk3(p.x - hx, p.y, p.z), w, grid),
             sampleWallSdf(mk3(p.x, p.y + hx, p.z), w, grid) -
                 sampleWallSdf(mk3(p.x, p.y - hx, p.z), w, grid),
             sampleWallSdf(mk3(p.x, p.y, p.z + hx), w, grid) -
                 sampleWallSdf(mk3(p.x, p.y, p.z - hx), w, grid));
}

/* ------------------------------------------------------------------------------------------ */
/* Hertz–Mindlin soft-sphere engine — src/solver_hertz.hpp, driver src/sim.hpp:543-750          */
/* FP64 constants: sqrt(5/6) and pi^2 in full precision (the reference writes them as floats). */
/* ------------------------------------------------------------------------------------------ */
#define HZ_SQRT56 0.91287092917527690
#define HZ_PI2 9.8696044010893586

int pdref_set_hertz_material(void* s, int mat, double youngs, double poisson) {
  Sim* S = asSim(s); /* src/sim.hpp:1153-1165 */
  if (mat < 0 || mat >= KMAXMAT) return badarg("material id out of range");
  S->hertzE[mat] = youngs;
  S->hertzNu[mat] = poisson;
  return 0;
}

/* src/solver_hertz.hpp:27-37 */
static void hertzPairEMu(const Sim* S, int ma, int mb, real* e, real* mu) {
  if (S->hasPairTable) {
    const int t = (ma * KMAXMAT + mb) * 2;
    *e = S->pairTable[t];
    *mu = S->pairTable[t + 1];
  } else {
    *e = S->restitutionNormal;
    *mu = S->frictionDynamic;
  }
}
/* src/solver_hertz.hpp:41-48 */
static real hertzBetaD(real e) {
  if (e <= 0.0) return -1.0;
  if (e >= 1.0) return 0.0;
  const real le = log(e);
  return le / sqrt(le * le + HZ_PI2);
}
static void hertzStars(const Sim* S, int ma, int mb, real* eStar, real* gStar) {
  const real Ei = S->hertzE[ma], Ej = S->hertzE[mb];
  const real ni = S->hertzNu[ma], nj = S->hertzNu[mb];
  *eStar = 1.0 / ((1.0 - ni * ni) / Ei + (1.0 - nj * nj) / Ej);
  *gStar = 1.0 / (2.0 * (2.0 - ni) * (1.0 + ni) / Ei + 2.0 * (2.0 - nj) * (1.0 + nj) / Ej);
}
/* src/solver_hertz.hpp:53-84: force on A at the contact point; updates the shear history */
static V3 hertzForce(real delta, V3 nhat, V3 vrel, real rStar, real mStar, real eStar,
                     real gStar, real bd, real mu, real dt, V3* xi) {
  const real sq = sqrt(rStar * delta);
  const real sn = 2.0 * eStar * sq;
  const real st = 8.0 * gStar * sq;
  const real kn = (4.0 / 3.0) * eStar * sq;
  const real etaN = -2.0 * HZ_SQRT56 * bd * sqrt(sn * mStar);
  const real etaT = -2.0 * HZ_SQRT56 * bd * sqrt(st * mStar);
  const real vn = dot3(vrel, nhat);
  const real fnSpring = kn * delta;
  const real fn = fnSpring - etaN * vn;
  const V3 vt = sub3(vrel, scale3(nhat, vn));
  V3 x = add3(*xi, scale3(vt, dt));
  x = sub3(x, scale3(nhat, dot3(x, nhat)));
  V3 ft = add3(scale3(x, -st), scale3(vt, -etaT));
  const real ftLen = sqrt(dot3(ft, ft));
  const real ftMax = mu * fabs(fnSpring);
  if (ftLen > ftMax) {
    const real sc = (ftLen > 1e-20) ? ftMax / ftLen : 0.0;
    ft = scale3(ft, sc);
    if (st > 1e-20) x = scale3(add3(ft, scale3(vt, etaT)), -1.0 / st);
  }
  *xi = x;
  return add3(scale3(nhat, fn), ft);
}
static void hzAccum(Sim* S, int i, V3 f, V3 t) {
  S->hForce[i] = add3(S->hForce[i], f);
  S->hTorque[i] = add3(S->hTorque[i], t);
}
static V3 neg3(V3 a) { return mk3(-a.x, -a.y, -a.z); }

/* spheres: src/solver_hertz.hpp:88-144 (accumulation in pair order) */
static void hertzPairForces(Sim* S, real dt) {
  for (int idx = 0; idx < S->hNumPairs; ++idx) {
    const int i = S->hPairs[2 * idx], j = S->hPairs[2 * idx + 1];
    const V3 dx = sub3(S->pos[i], S->pos[j]);
    const real d = sqrt(dot3(dx, dx));
    const real ri = S->rad[i], rj = S->rad[j];
    const real delta = ri + rj - d;
    if (delta <= 0.0 || d < 1e-12) {
      S->hXi[idx] = mk3(0, 0, 0);
      continue;
    }
    const V3 nhat = scale3(dx, 1.0 / d);
    const V3 rA = scale3(nhat, -(ri - 0.5 * delta));
    const V3 rB = scale3(nhat, (rj - 0.5 * delta));
    const V3 vrel = sub3(add3(S->vel[i], cross3(S->angVel[i], rA)),
                         add3(S->vel[j], cross3(S->angVel[j], rB)));
    const int ma = S->materialId[i], mb = S->materialId[j];
    real e, mu, eStar, gStar;
    hertzPairEMu(S, ma, mb, &e, &mu);
    hertzStars(S, ma, mb, &eStar, &gStar);
    const real rStar = ri * rj / (ri + rj);
    const real imSum = S->invMass[i] + S->invMass[j];
    const real mStar = (imSum > 0.0) ? 1.0 / imSum : 0.0;
    const V3 f = hertzForce(delta, nhat, vrel, rStar, mStar, eStar, gStar, hertzBetaD(e), mu, dt,
                            &S->hXi[idx]);
    hzAccum(S, i, f, cross3(rA, f));
    hzAccum(S, j, neg3(f), cross3(rB, scale3(f, -1.0)));
  }
}

/* point-shell shapes: src/solver_hertz.hpp:454-606 (one Mindlin patch history per pair) */
static void hertzShapePairForces(Sim* S, real dt) {
  const real frac = S->hertzContactRadiusFrac;
  const int sphere = S->shapeType == SHAPE_SPHERE;
  for (int idx = 0; idx < S->hNumPairs; ++idx) {
    const int idA = S->hPairs[2 * idx], idB = S->hPairs[2 * idx + 1];
    const V3 posA = S->pos[idA], posB = S->pos[idB];
    const Q4 qA = S->quat[idA], qB = S->quat[idB];
    const real effScaleA = S->scale[idA] * S->globalScale, effScaleB = S->scale[idB] * S->globalScale;
    const int countA = S->numShell;
    const int iter = countA > 0 ? countA : 1;
    const int ma = S->materialId[idA], mb = S->materialId[idB];
    real e, mu, eStar, gStar;
    hertzPairEMu(S, ma, mb, &e, &mu);
    hertzStars(S, ma, mb, &eStar, &gStar);
    const real rcA = (sphere ? 1.0 : frac) * S->shapeParams[0] * effScaleA;
    const real rcB = (sphere ? 1.0 : frac) * S->shapeParams[0] * effScaleB;
    const real rStar = rcA * rcB / (rcA + rcB);
    const real imSum = S->invMass[idA] + S->invMass[idB];
    const real mStar = (imSum > 0.0) ? 1.0 / imSum : 0.0;
    const real bd = hertzBetaD(e);
    const V3 vA = S->vel[idA], wA = S->angVel[idA], vB = S->vel[idB], wB = S->angVel[idB];
    real fnSum = 0.0, snSum = 0.0, dMax = 0.0;
    V3 cSum = mk3(0, 0, 0), nSum = mk3(0, 0, 0), fAcc = mk3(0, 0, 0);
    V3 tAccA = mk3(0, 0, 0), tAccB = mk3(0, 0, 0);
    const real snTotPrev = S->hSn[idx];
    const real dampC = (mStar > 0.0) ? -2.0 * HZ_SQRT56 * bd * sqrt(mStar) : 0.0;
    for (int k = 0; k < iter; ++k) {
      V3 pLocalA = mk3(0, 0, 0);
      real pointRadius = 0.0;
      if (countA > 0) pLocalA = S->shell[k];
      else if (sphere) pointRadius = S->shapeParams[0] * effScaleA;
      const V3 pWorld = add3(posA, rotate(qA, scale3(pLocalA, effScaleA)));
      const V3 pCanB = scale3(invRotate(qB, sub3(pWorld, posB)), 1.0 / effScaleB);
      const real dist = sdfShape(S, pCanB) * effScaleB;
      const real effDist = dist - pointRadius;
      if (effDist >= 0.0) continue;
      const real delta = -effDist;
      const real eps = 1e-4;
      V3 nLoc = mk3(sdfShape(S, mk3(pCanB.x + eps, pCanB.y, pCanB.z)) -
                        sdfShape(S, mk3(pCanB.x - eps, pCanB.y, pCanB.z)),
                    sdfShape(S, mk3(pCanB.x, pCanB.y + eps, pCanB.z)) -
                        sdfShape(S, mk3(pCanB.x, pCanB.y - eps, pCanB.z)),
                    sdfShape(S, mk3(pCanB.x, pCanB.y, pCanB.z + eps)) -
                        sdfShape(S, mk3(pCanB.x, pCanB.y, pCanB.z - eps)));
      const real len = len3(nLoc);
      nLoc = (len > 1e-9) ? scale3(nLoc, 1.0 / len) : mk3(0, 1, 0);
      const V3 nW = rotate(qB, nLoc);
      const V3 pSurfA = sub3(pWorld, scale3(nW, pointRadius));
      const V3 rA = sub3(pSurfA, posA);
      const V3 rB = sub3(add3(pSurfA, scale3(nW, delta)), posB);
      const V3 vrelP = sub3(add3(vA, cross3(wA, rA)), add3(vB, cross3(wB, rB)));
      const real vnP = dot3(vrelP, nW);
      const real sq = sqrt(rStar * delta);
      const real fnSpring = (4.0 / 3.0) * eStar * sq * delta;
      const real snI = 2.0 * eStar * sq;
      const real snTot = (snTotPrev > snI) ? snTotPrev : snI;
      const real etaI = dampC * snI / sqrt(snTot);
      const V3 f = scale3(nW, fnSpring - etaI * vnP);
      fAcc = add3(fAcc, f);
      tAccA = add3(tAccA, cross3(rA, f));
      tAccB = add3(tAccB, cross3(rB, scale3(f, -1.0)));
      fnSum += fnSpring;
      snSum += 2.0 * eStar * sq;
      cSum = add3(cSum, scale3(pSurfA, fnSpring));
      nSum = add3(nSum, scale3(nW, fnSpring));
      if (delta > dMax) dMax = delta;
    }
    if (fnSum <= 0.0) {
      S->hXi[idx] = mk3(0, 0, 0);
      S->hSn[idx] = 0.0;
      continue;
    }
    const V3 cpt = scale3(cSum, 1.0 / fnSum);
    V3 nhat = nSum;
    const real nl = len3(nhat);
    if (nl > 1e-9) nhat = scale3(nhat, 1.0 / nl);
    const V3 rA = sub3(cpt, posA), rB = sub3(cpt, posB);
    const V3 vrel = sub3(add3(vA, cross3(wA, rA)), add3(vB, cross3(wB, rB)));
    S->hSn[idx] = snSum;
    const V3 vt = sub3(vrel, scale3(nhat, dot3(vrel, nhat)));
    const real st = 8.0 * gStar * sqrt(rStar * dMax);
    const real etaT = -2.0 * HZ_SQRT56 * bd * sqrt(st * mStar);
    V3 x = S->hXi[idx];
    x = add3(x, scale3(vt, dt));
    x = sub3(x, scale3(nhat, dot3(x, nhat)));
    V3 ft = add3(scale3(x, -st), scale3(vt, -etaT));
    const real ftLen = len3(ft);
    const real ftMax = mu * fnSum;
    if (ftLen > ftMax) {
      const real sc = (ftLen > 1e-20) ? ftMax / ftLen : 0.0;
      ft = scale3(ft, sc);
      if (st > 1e-20) x = scale3(add3(ft, scale3(vt, etaT)), -1.0 / st);
    }
    S->hXi[idx] = x;
    fAcc = add3(fAcc, ft);
    tAccA = add3(tAccA, cross3(rA, ft));
    tAccB = add3(tAccB, cross3(rB, scale3(ft, -1.0)));
    hzAccum(S, idA, fAcc, tAccA);
    hzAccum(S, idB, neg3(fAcc), tAccB);
  }
}

/* wall (e, mu, E*, G*) of particle material ma against wall w, src/solver_hertz.hpp:187-199 */
static void hertzWallMaterial(const Sim* S, int ma, const Wall* w, real* e, real* mu,
                              real* eStar, real* gStar) {
  if (w->materialId >= 0 && S->hasPairTable) {
    hertzPairEMu(S, ma, w->materialId, e, mu);
  } else {
    *e = w->restitution;
    *mu = w->friction;
  }
  const int mw = (w->materialId >= 0) ? w->materialId : ma;
  hertzStars(S, ma, mw, eStar, gStar);
}

/* src/solver_hertz.hpp:148-338, candidates in ascending (particle, wall) order */
static void hertzWallForces(Sim* S, real dt, int hasShapes) {
  const real frac = S->hertzContactRadiusFrac;
  for (int ci = 0; ci < S->hNumWallCand; ++ci) {
    const int enc = S->hWallCand[ci];
    const int i = enc / 8, wi = enc % 8;
    const V3 p = S->pos[i];
    const real r = S->rad[i];
    const Wall* w = &S->walls[wi];
    const int slot = i * HZ_MAXWALLS + wi;
    const int ma = S->materialId[i];
    real e, mu, eStar, gStar;
    hertzWallMaterial(S, ma, w, &e, &mu, &eStar, &gStar);
    const real mStar = (S->invMass[i] > 0.0) ? 1.0 / S->invMass[i] : 0.0;
    if (hasShapes) {
      const Q4 qP = S->quat[i];
      const real effScale = S->scale[i] * S->globalScale;
      const real rc = frac * S->shapeParams[0] * effScale;
      const real bd = hertzBetaD(e);
      const V3 vP = S->vel[i], wP = S->angVel[i];
      const real snTotPrev = S->hSnWall[slot];
      const real dampC = (mStar > 0.0) ? -2.0 * HZ_SQRT56 * bd * sqrt(mStar) : 0.0;
      real fnSum = 0.0, snSum = 0.0, dMax = 0.0;
      V3 cSum = mk3(0, 0, 0), nSum = mk3(0, 0, 0), fAcc = mk3(0, 0, 0), tAcc = mk3(0, 0, 0);
      for (int k = 0; k < S->numShell; ++k) {
        const V3 pW = add3(p, rotate(qP, scale3(S->shell[k], effScale)));
        const real sd = sampleWallSdf(pW, w, S->wallGrid);
        if (sd >= 0.0) continue;
        const real delta = -sd;
        const V3 g = wallGradient(pW, w, S->wallGrid);
        const real gl = len3(g);
        if (gl < 1e-12) continue;
        const V3 nhat = scale3(g, 1.0 / gl);
        const V3 rA = sub3(pW, p);
        const V3 vWall = add3(w->linVel, cross3(w->angVel, sub3(pW, w->center)));
        const V3 vrelP = sub3(add3(vP, cross3(wP, rA)), vWall);
        const real vnP = dot3(vrelP, nhat);
        const real sq = sqrt(rc * delta);
        const real fnSpring = (4.0 / 3.0) * eStar * sq * delta;
        const real snI = 2.0 * eStar * sq;
        const real snTot = (snTotPrev > snI) ? snTotPrev : snI;
        const real etaI = dampC * snI / sqrt(snTot);
        const V3 f = scale3(nhat, fnSpring - etaI * vnP);
        fAcc = add3(fAcc, f);
        tAcc = add3(tAcc, cross3(rA, f));
        fnSum += fnSpring;
        snSum += 2.0 * eStar * sq;
        cSum = add3(cSum, scale3(pW, fnSpring));
        nSum = add3(nSum, scale3(nhat, fnSpring));
        if (delta > dMax) dMax = delta;
      }
      if (fnSum <= 0.0) {
        S->hXiWall[slot] = mk3(0, 0, 0);
        S->hSnWall[slot] = 0.0;
        continue;
      }
      const V3 cpt = scale3(cSum, 1.0 / fnSum);
      V3 nhat = nSum;
      const real nl = len3(nhat);
      if (nl > 1e-9) nhat = scale3(nhat, 1.0 / nl);
      const V3 rA = sub3(cpt, p);
      const V3 vWall = add3(w->linVel, cross3(w->angVel, sub3(cpt, w->center)));
      const V3 vrel = sub3(add3(vP, cross3(wP, rA)), vWall);
      S->hSnWall[slot] = snSum;
      const V3 vt = sub3(vrel, scale3(nhat, dot3(vrel, nhat)));
      const real st = 8.0 * gStar * sqrt(rc * dMax);
      const real etaT = -2.0 * HZ_SQRT56 * bd * sqrt(st * mStar);
      V3 x = S->hXiWall[slot];
      x = add3(x, scale3(vt, dt));
      x = sub3(x, scale3(nhat, dot3(x, nhat)));
      V3 ft = add3(scale3(x, -st), scale3(vt, -etaT));
      const real ftLen = len3(ft);
      const real ftMax = mu * fnSum;
      if (ftLen > ftMax) {
        const real sc = (ftLen > 1e-20) ? ftMax / ftLen : 0.0;
        ft = scale3(ft, sc);
        if (st > 1e-20) x = scale3(add3(ft, scale3(vt, etaT)), -1.0 / st);
      }
      S->hXiWall[slot] = x;
      fAcc = add3(fAcc, ft);
      tAcc = add3(tAcc, cross3(rA, ft));
      hzAccum(S, i, fAcc, tAcc);
      continue;
    }
    const real sd = sampleWallSdf(p, w, S->wallGrid);
    const real delta = r - sd;
    if (delta <= 0.0) {
      S->hXiWall[slot] = mk3(0, 0, 0);
      continue;
    }
    const V3 g = wallGradient(p, w, S->wallGrid);
    const real gl = sqrt(dot3(g, g));
    if (gl < 1e-12) continue;
    const V3 nhat = scale3(g, 1.0 / gl);
    const V3 rA = scale3(nhat, -(r - 0.5 * delta));
    const V3 cpt = add3(p, rA);
    const V3 vWall = add3(w->linVel, cross3(w->angVel, sub3(cpt, w->center)));
    const V3 vrel = sub3(add3(S->vel[i], cross3(S->angVel[i], rA)), vWall);
    const V3 f = hertzForce(delta, nhat, vrel, r, mStar, eStar, gStar, hertzBetaD(e), mu, dt,
                            &S->hXiWall[slot]);
    hzAccum(S, i, f, cross3(rA, f));
  }
}

/* src/solver_hertz.hpp:374-422: symplectic Euler, forces then cleared */
static void hertzIntegrate(Sim* S, real dt, int integrateOrientation) {
  const V3 g = S->gravity;
  for (int i = 0; i < S->numReal; ++i) {
    const real im = S->invMass[i];
    if (im > 0.0) {
      const V3 iI = S->invI[i];
      const V3 T = S->hTorque[i];
      V3 dw;
      if (iI.x == iI.y && iI.y == iI.z) {
        dw = mk3(T.x * iI.x, T.y * iI.x, T.z * iI.x);
      } else {
        const Q4 q = S->quat[i];
        const V3 tl = invRotate(q, T);
        dw = rotate(q, mk3(tl.x * iI.x, tl.y * iI.y, tl.z * iI.z));
      }
      const V3 F = S->hForce[i];
      S->vel[i].x += (F.x * im + g.x) * dt;
      S->vel[i].y += (F.y * im + g.y) * dt;
      S->vel[i].z += (F.z * im + g.z) * dt;
      S->angVel[i].x += dw.x * dt;
      S->angVel[i].y += dw.y * dt;
      S->angVel[i].z += dw.z * dt;
      S->pos[i].x += S->vel[i].x * dt;
      S->pos[i].y += S->vel[i].y * dt;
      S->pos[i].z += S->vel[i].z * dt;
      if (integrateOrientation) {
        Q4 q = S->quat[i];
        const V3 w = S->angVel[i];
        const real h = 0.5 * dt;
        Q4 dq;
        dq.x = h * (w.x * q.w + w.y * q.z - w.z * q.y);
        dq.y = h * (w.y * q.w + w.z * q.x - w.x * q.z);
        dq.z = h * (w.z * q.w + w.x * q.y - w.y * q.x);
        dq.w = h * (-w.x * q.x - w.y * q.y - w.z * q.z);
        q.x = q.x + dq.x; q.y = q.y + dq.y; q.z = q.z + dq.z; q.w = q.w + dq.w;
        const real qn = sqrt(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
        if (qn > 1e-12) {
          S->quat[i].x = q.x / qn;
          S->quat[i].y = q.y / qn;
          S->quat[i].z = q.z / qn;
          S->quat[i].w = q.w / qn;
        }
      }
    }
    S->hForce[i] = mk3(0, 0, 0);
    S->hTorque[i] = mk3(0, 0, 0);
  }
}

static void hzEnsureSlots(Sim* S, int n) {
  if (n <= S->hSlotCap) return;
  const int nc = n + n / 4 + 64;
  S->hRefPos = (V3*)xrealloc(S->hRefPos, sizeof(V3) * (size_t)nc);
  S->hForce = (V3*)xrealloc(S->hForce, sizeof(V3) * (size_t)nc);
  S->hTorque = (V3*)xrealloc(S->hTorque, sizeof(V3) * (size_t)nc);
  S->hXiWall = (V3*)xrealloc(S->hXiWall, sizeof(V3) * (size_t)nc * HZ_MAXWALLS);
  S->hSnWall = (real*)xrealloc(S->hSnWall, sizeof(real) * (size_t)nc * HZ_MAXWALLS);
  S->hWallCand = (int*)xrealloc(S->hWallCand, sizeof(int) * (size_t)nc * HZ_MAXWALLS);
  const size_t old = (size_t)S->hSlotCap;
  memset(S->hRefPos + old, 0, sizeof(V3) * (nc - old));
  memset(S->hForce + old, 0, sizeof(V3) * (nc - old));
  memset(S->hTorque + old, 0, sizeof(V3) * (nc - old));
  memset(S->hXiWall + old * HZ_MAXWALLS, 0, sizeof(V3) * (nc - old) * HZ_MAXWALLS);
  memset(S->hSnWall + old * HZ_MAXWALLS, 0, sizeof(real) * (nc - old) * HZ_MAXWALLS);
  S->hSlotCap = nc;
}

/* the key of the Hertz pair cache is (hi << 32) | lo, src/sim.hpp:571-573 */
static uint64_t hzKey(int a, int b) {
  const uint64_t lo = (uint64_t)(unsigned)(a < b ? a : b), hi = (uint64_t)(unsigned)(a < b ? b : a);
  return (hi << 32) | lo;
}
typedef struct { uint64_t key; int idx; } HzKI;
static int cmpHzKI(const void* a, const void* b) {
  const HzKI* x = (const HzKI*)a;
  const HzKI* y = (const HzKI*)b;
  return x->key < y->key ? -1 : (x->key > y->key ? 1 : 0);
}
/* src/sim.hpp:596-624: save the current list sorted by key for the next rebuild */
static void hertzCommitHistory(Sim* S) {
  const int n = S->hNumPairs;
  if (n <= 0) {
    S->hPrevCount = 0;
    return;
  }
  HzKI* v = (HzKI*)xrealloc(NULL, sizeof(HzKI) * (size_t)n);
  for (int i = 0; i < n; ++i) { v[i].key = S->hKeys[i]; v[i].idx = i; }
  qsort(v, (size_t)n, sizeof(HzKI), cmpHzKI); /* keys are unique */
  S->hPrevKeys = (uint64_t*)xrealloc(S->hPrevKeys, sizeof(uint64_t) * (size_t)n);
  S->hPrevXi = (V3*)xrealloc(S->hPrevXi, sizeof(V3) * (size_t)n);
  for (int i = 0; i < n; ++i) {
    S->hPrevKeys[i] = v[i].key;
    S->hPrevXi[i] = S->hXi[v[i].idx];
  }
  free(v);
  S->hPrevCount = n;
}
/* src/sim.hpp:545-594: Verlet list with margin = skin, history carried by key */
static void hertzRebuildPairs(Sim* S, real skin) {
  for (int i = 0; i < S->numReal; ++i) S->posPred[i] = S->pos[i];
  S->numParticles = S->numReal;
  findCollisions(S, skin);
  const int np = S->np;
  if (np > S->hPairCap) {
    const int nc = np + np / 4 + 64;
    S->hPairs = (int*)xrealloc(S->hPairs, sizeof(int) * 2 * (size_t)nc);
    S->hKeys = (uint64_t*)xrealloc(S->hKeys, sizeof(uint64_t) * (size_t)nc);
    S->hXi = (V3*)xrealloc(S->hXi, sizeof(V3) * (size_t)nc);
    S->hSn = (real*)xrealloc(S->hSn, sizeof(real) * (size_t)nc);
    S->hPairCap = nc;
  }
  for (int idx = 0; idx < np; ++idx) {
    const int a = S->pairs[2 * idx], b = S->pairs[2 * idx + 1];
    S->hPairs[2 * idx] = a;
    S->hPairs[2 * idx + 1] = b;
    const uint64_t k = hzKey(a, b);
    S->hKeys[idx] = k;
    const int l = lowerBound(S->hPrevKeys, S->hPrevCount, k);
    const int hit = (l < S->hPrevCount && S->hPrevKeys[l] == k);
    S->hXi[idx] = hit ? S->hPrevXi[l] : mk3(0, 0, 0);
    S->hSn[idx] = 0.0;
  }
  S->hNumPairs = np;
  for (int i = 0; i < S->numReal; ++i) S->hRefPos[i] = S->pos[i];
}
/* src/solver_hertz.hpp:343-370 */
static void hertzBuildWallCandidates(Sim* S, real skin) {
  int n = 0;
  for (int i = 0; i < S->numReal; ++i)
    for (int wi = 0; wi < S->numWalls; ++wi) {
      const real sd = sampleWallSdf(S->pos[i], &S->walls[wi], S->wallGrid);
      if (sd < S->rad[i] + skin) S->hWallCand[n++] = i * 8 + wi;
    }
  S->hNumWallCand = n;
}
static real hertzMaxDisp2(const Sim* S) {
  real m = 0.0;
  for (int i = 0; i < S->numReal; ++i) {
    const V3 d = sub3(S->pos[i], S->hRefPos[i]);
    const real dd = dot3(d, d);
    if (dd > m) m = dd;
  }
  return m;
}

/* demStepHertz, src/sim.hpp:628-750 */
int pdref_step_hertz(void* s, double dt, int nsteps, double skinFrac) {
  Sim* S = asSim(s);
  S->dt = dt;
  if (S->dom.px || S->dom.py || S->dom.pz)
    return fail("step_hertz: periodic domains not supported");
  if (S->numPlanes > 0)
    return fail("step_hertz: analytic planes not supported (use an SDF wall)");
  if (S->numWalls > HZ_MAXWALLS) return fail("step_hertz: too many SDF walls");
  const int n = S->numReal;
  hzEnsureSlots(S, n > 0 ? n : 1);
  fillRad(S, n);
  real minRad = 3.4e38;
  for (int i = 0; i < n; ++i)
    if (S->rad[i] < minRad) minRad = S->rad[i];
  const real skin = skinFrac * minRad;
  const real rebuildAt = 0.25 * skin * skin;
  int rebuild = (S->hNumPairs < 0);
  const int hasShapes = S->numShell > 0;
  for (int k = 0; k < nsteps; ++k) {
    if (rebuild) {
      hertzCommitHistory(S);
      hertzRebuildPairs(S, skin);
      if (S->numWalls > 0) hertzBuildWallCandidates(S, skin);
      rebuild = 0;
    }
    if (hasShapes) hertzShapePairForces(S, dt);
    else hertzPairForces(S, dt);
    if (S->numWalls > 0 && S->hNumWallCand > 0) hertzWallForces(S, dt, hasShapes);
    hertzIntegrate(S, dt, hasShapes);
    if ((k & 15) == 15 || k == nsteps - 1)
      if (hertzMaxDisp2(S) > rebuildAt) rebuild = 1;
  }
  for (int i = 0; i < n; ++i) {
    S->posPred[i] = S->pos[i];
    S->velPred[i] = S->vel[i];
  }
  return 0;
}
/* debug: the cached Hertz pair list with its shear history */
int pdref_get_hertz_pairs(void* s, int* pairs, double* xi, int cap) {
  Sim* S = asSim(s);
  const int n = S->hNumPairs > 0 ? S->hNumPairs : 0;
  for (int k = 0; k < n && k < cap; ++k) {
    if (pairs) { pairs[2 * k] = S->hPairs[2 * k]; pairs[2 * k + 1] = S->hPairs[2 * k + 1]; }
    if (xi) { xi[3 * k] = S->hXi[k].x; xi[3 * k + 1] = S->hXi[k].y; xi[3 * k + 2] = S->hXi[k].z; }
  }
  return n;
}

/* wall contacts of the XPBD substep, src/narrowphase.hpp:278-367: per real, per wall, per probe */
static real wallFrictionMax(const Sim* S) {
  real m = 0.0;
  for (int w = 0; w < S->numWalls; ++w)
    if (S->walls[w].friction > m) m = S->walls[w].friction;
  return m;
}
static void detectWallSdf(Sim* S, real margin) {
  const real eps = 1e-4;
  for (int i = 0; i < S->numReal; ++i) {
    const V3 posA = S->posPred[i];
    const real s = S->scale[i] * S->globalScale;
    real baseR = S->shapeParams[0];
    if (baseR == 0.0) baseR = 1.0;
    const real radius = baseR * s;
    const int numPts = S->numShell;
    const Q4 qA = S->quatPred[i];
    for (int wi = 0; wi < S->numWalls; ++wi) {
      const Wall* w = &S->walls[wi];
      const int iter = numPts > 0 ? numPts : 1;
      for (int k = 0; k < iter; ++k) {
        V3 rA = mk3(0, 0, 0);
        if (numPts > 0) rA = rotate(qA, scale3(S->shell[k], s));
        const V3 pw = add3(posA, rA);
        const real sdf = sampleWallSdf(pw, w, S->wallGrid);
        V3 n = mk3(sampleWallSdf(mk3(pw.x + eps, pw.y, pw.z), w, S->wallGrid) -
                       sampleWallSdf(mk3(pw.x - eps, pw.y, pw.z), w, S->wallGrid),
                   sampleWallSdf(mk3(pw.x, pw.y + eps, pw.z), w, S->wallGrid) -
                       sampleWallSdf(mk3(pw.x, pw.y - eps, pw.z), w, S->wallGrid),
                   sampleWallSdf(mk3(pw.x, pw.y, pw.z + eps), w, S->wallGrid) -
                       sampleWallSdf(mk3(pw.x, pw.y, pw.z - eps), w, S->wallGrid));
        const real ln = len3(n);
        n = (ln > 1e-9) ? scale3(n, 1.0 / ln) : mk3(0, 1, 0);
        const real dist = (numPts > 0) ? sdf : sdf - radius;
        if (dist >= margin) continue;
        const V3 rAeff = (numPts > 0) ? rA : scale3(n, -radius);
        const V3 pSurfA = add3(posA, rAeff);
        const V3 pWall = sub3(pSurfA, scale3(n, dist));
        if (dist < 0.0 && -dist > S->maxOverlap) S->maxOverlap = -dist;
        Contact c;
        memset(&c, 0, sizeof c);
        c.bodyA = i;
        c.bodyB = -1;
        c.normal = n;
        c.rA = rAeff;
        c.rB = pWall;
        c.dist = dist;
        c.boundaryVel = add3(w->linVel, cross3(w->angVel, sub3(pWall, w->center)));
        if (w->materialId >= 0 && S->hasPairTable) {
          const int t = ((int)S->materialId[i] * KMAXMAT + w->materialId) * 2;
          c.boundaryRestitution = S->pairTable[t];
          c.boundaryFriction = S->pairTable[t + 1];
        } else {
          c.boundaryRestitution = w->restitution;
          c.boundaryFriction = w->friction;
        }
        pushContact(S, &c);
      }
    }
  }
}

/* setSdfShape(grid, nx, ny, nz, origin, spacing, shell, invInertia, boundingRadius),
 * src/sim.hpp:1034-1084: shape 0 becomes a grid SDF probed by the given point shell */
int pdref_set_sdf_shape(void* s, const double* grid, int nx, int ny, int nz, double ox,
                        double oy, double oz, double sx, double sy, double sz,
                        const double* shell, int nPts, double iix, double iiy, double iiz,
                        double boundingRadius) {
  Sim* S = asSim(s);
  if (nx < 2 || ny < 2 || nz < 2) return fail("setSdfShape: grid dims must be >= 2 on each axis");
  if (nPts <= 0) return fail("setSdfShape: empty surface point shell");
  const long cells = (long)nx * ny * nz;
  S->sdfGrid = (real*)xrealloc(S->sdfGrid, sizeof(real) * (size_t)cells);
  memcpy(S->sdfGrid, grid, sizeof(real) * (size_t)cells);
  S->gnx = nx; S->gny = ny; S->gnz = nz;
  S->gOrigin = mk3(ox, oy, oz);
  S->gInvSpacing = mk3(sx > 0 ? 1.0 / sx : 0.0, sy > 0 ? 1.0 / sy : 0.0, sz > 0 ? 1.0 / sz : 0.0);
  free(S->shell);
  S->shell = (V3*)xrealloc(NULL, sizeof(V3) * (size_t)nPts);
  for (int i = 0; i < nPts; ++i) S->shell[i] = mk3(shell[3 * i], shell[3 * i + 1], shell[3 * i + 2]);
  S->numShell = nPts;
  S->shapeType = SHAPE_GRID_SDF;
  S->shapeParams[0] = boundingRadius;
  S->shapeParams[1] = S->shapeParams[2] = S->shapeParams[3] = 0.0;
  S->baseRadius = boundingRadius;
  S->defaultInvI = mk3(iix, iiy, iiz);
  return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* stabilization mode 4 ("ordered"), src/sim.hpp:356-397: fresh height-from-floor levels, the  */
/* coloured manifolds bucketed by (level, colour), then bottom-up + top-down PGS passes         */
/* ------------------------------------------------------------------------------------------ */
#define LEVEL_INF (1 << 28)
/* src/contact_preprocessing.hpp:255-297: the fixed point of the min-propagation is the exact
 * support distance, whatever order the relaxation visits the manifolds in */
static void computeHeightLevels(const Sim* S, V3 gHat, int* h) {
  for (int i = 0; i < S->numReal; ++i) h[i] = LEVEL_INF;
  for (int s = 0; s < 1024; ++s) {
    int changed = 0;
    for (int idx = 0; idx < S->nm; ++idx) {
      const Manifold* m = &S->manifolds[idx];
      if (m->numPoints <= 0) continue;
      const int realA = S->realIdx[m->bodyA];
      if (m->bodyB < 0) {
        if (h[realA] > 0) { h[realA] = 0; changed++; }
        continue;
      }
      const int realB = S->realIdx[m->bodyB];
      const V3 dx = sub3(S->posPred[m->bodyA], S->posPred[m->bodyB]);
      const real up = -(dx.x * gHat.x + dx.y * gHat.y + dx.z * gHat.z);
      const real thr = 0.3 * sqrt(dot3(dx, dx));
      if (up > thr) {
        const int cand = h[realB];
        if (cand < LEVEL_INF && h[realA] > cand + 1) { h[realA] = cand + 1; changed++; }
      } else if (up < -thr) {
        const int cand = h[realA];
        if (cand < LEVEL_INF && h[realB] > cand + 1) { h[realB] = cand + 1; changed++; }
      }
    }
    if (!changed) break;
  }
}
typedef struct { int key, idx; } LvlKey;
static int cmpLvl(const void* a, const void* b) {
  const LvlKey* x = (const LvlKey*)a;
  const LvlKey* y = (const LvlKey*)b;
  if (x->key != y->key) return x->key < y->key ? -1 : 1;
  return x->idx < y->idx ? -1 : (x->idx > y->idx);
}
static void orderedStabilization(Sim* S, V3 gHat, real vRestS) {
  int* h = (int*)xrealloc(NULL, sizeof(int) * (size_t)(S->numReal + 1));
  computeHeightLevels(S, gHat, h);
  /* src/solver_velocity.hpp:914-965: key = min level (clamped 1023) * 64 + colour */
  LvlKey* v = (LvlKey*)xrealloc(NULL, sizeof(LvlKey) * (size_t)(S->nm + 1));
  int n = 0;
  for (int idx = 0; idx < S->nm; ++idx) {
    const Manifold* m = &S->manifolds[idx];
    const int c = S->mColor[idx];
    if (m->numPoints <= 0 || c < 0) continue;
    int lv = h[S->realIdx[m->bodyA]];
    if (m->bodyB >= 0) {
      const int hb = h[S->realIdx[m->bodyB]];
      if (hb < lv) lv = hb;
    }
    if (lv > 1023) lv = 1023;
    v[n].key = lv * 64 + c;
    v[n].idx = idx;
    n++;
  }
  qsort(v, (size_t)n, sizeof(LvlKey), cmpLvl);
  /* buckets of equal key; within one the manifolds share a colour, so order is free */
  for (int it = 0; it < 2 * S->velIters; ++it) {
    S->maxApproach = 0.0;
    for (int k = 0; k < n; ++k) pgsSolveOne(S, v[k].idx, vRestS);
    int b = n;
    while (b > 0) {  /* top-down: buckets in descending key order */
      int a = b - 1;
      while (a > 0 && v[a - 1].key == v[b - 1].key) --a;
      for (int k = a; k < b; ++k) pgsSolveOne(S, v[k].idx, vRestS);
      b = a;
    }
    S->stats.sweeps_stab++;
    if (S->maxApproach <= vRestS) break;
  }
  free(v);
  free(h);
}

/* ------------------------------------------------------------------------------------------ */
/* packed-bed SDF reconstruction (get_sdf_grid / export_sdf), src/output_sdf.hpp:25-197:        */
/* splat each particle's exact analytic distance over its bounding band (minimum), then a       */
/* Jacobi fast-iterative Eikonal solve for the far field, periodic-wrapped; x-fastest output    */
/* ------------------------------------------------------------------------------------------ */
int pdref_get_sdf_grid(void* s, int rx, int ry, int rz, double* out) {
  Sim* S = asSim(s);
  if (rx < 1 || ry < 1 || rz < 1) return badarg("resolution must be >= 1 on each axis");
  const long total = (long)rx * ry * rz;
  const V3 o = S->dom.min;
  const V3 vox = mk3((S->dom.max.x - S->dom.min.x) / rx, (S->dom.max.y - S->dom.min.y) / ry,
                     (S->dom.max.z - S->dom.min.z) / rz);
  real* g = (real*)xrealloc(NULL, sizeof(real) * (size_t)total);
  real* g2 = (real*)xrealloc(NULL, sizeof(real) * (size_t)total);
  unsigned char* fixedCell = (unsigned char*)calloc((size_t)total, 1);
  for (long i = 0; i < total; ++i) g[i] = 100.0;
  const int px = S->dom.px, py = S->dom.py, pz = S->dom.pz;
  for (int i = 0; i < S->numReal; ++i) {
    const V3 pw = S->pos[i];
    const Q4 q = S->quat[i];
    const real sc = S->scale[i];
    real rbound = 1.0;
    if (S->shapeType == SHAPE_HOLLOW_CYLINDER) {
      const real r = S->shapeParams[0], h = S->shapeParams[1];
      rbound = sqrt(r * r + (h * 0.5) * (h * 0.5));
    } else if (S->shapeType == SHAPE_SPHERE || S->shapeType == SHAPE_GRID_SDF) {
      rbound = S->shapeParams[0];
    }
    rbound *= sc;
    rbound *= 1.2;
    int minx = (int)floor((pw.x - rbound - o.x) / vox.x), maxx = (int)ceil((pw.x + rbound - o.x) / vox.x);
    int miny = (int)floor((pw.y - rbound - o.y) / vox.y), maxy = (int)ceil((pw.y + rbound - o.y) / vox.y);
    int minz = (int)floor((pw.z - rbound - o.z) / vox.z), maxz = (int)ceil((pw.z + rbound - o.z) / vox.z);
    if (!px) { if (minx < 0) minx = 0; if (maxx > rx - 1) maxx = rx - 1; }
    if (!py) { if (miny < 0) miny = 0; if (maxy > ry - 1) maxy = ry - 1; }
    if (!pz) { if (minz < 0) minz = 0; if (maxz > rz - 1) maxz = rz - 1; }
    for (int z = minz; z <= maxz; ++z)
      for (int y = miny; y <= maxy; ++y)
        for (int x = minx; x <= maxx; ++x) {
          const V3 vp = mk3(o.x + (x + 0.5) * vox.x, o.y + (y + 0.5) * vox.y, o.z + (z + 0.5) * vox.z);
          const V3 plocal = scale3(invRotate(q, sub3(vp, pw)), 1.0 / sc);
          real distc = 100.0;
          if (S->shapeType == SHAPE_HOLLOW_CYLINDER || S->shapeType == SHAPE_SPHERE ||
              S->shapeType == SHAPE_GRID_SDF)
            distc = sdfShape(S, plocal);
          const real dist = distc * sc;
          const int wx = (x % rx + rx) % rx, wy = (y % ry + ry) % ry, wz = (z % rz + rz) % rz;
          const long vi = (long)wz * rx * ry + (long)wy * rx + wx;
          if (dist < g[vi]) g[vi] = dist;
          if (dist < rbound) fixedCell[vi] = 1;
        }
  }
  int iters = rx > ry ? rx : ry;
  if (rz > iters) iters = rz;
  iters *= 4;
  const real h = vox.x;
  real* in = g;
  real* o2 = g2;
  const long sy = rx, sz = (long)rx * ry;
  for (int it = 0; it < iters; ++it) {
    for (long idx = 0; idx < total; ++idx) {
      if (fixedCell[idx]) { o2[idx] = in[idx]; continue; }
      const int z = (int)(idx / sz);
      const long rem = idx % sz;
      const int y = (int)(rem / sy), x = (int)(rem % sy);
      real xm = 100.0, xp = 100.0, ym = 100.0, yp = 100.0, zm = 100.0, zp = 100.0;
      if (x > 0) xm = in[idx - 1]; else if (px) xm = in[idx - 1 + rx];
      if (x < rx - 1) xp = in[idx + 1]; else if (px) xp = in[idx + 1 - rx];
      if (y > 0) ym = in[idx - sy]; else if (py) ym = in[(long)z * sz + (long)(ry - 1) * sy + x];
      if (y < ry - 1) yp = in[idx + sy]; else if (py) yp = in[(long)z * sz + x];
      if (z > 0) zm = in[idx - sz]; else if (pz) zm = in[(long)(rz - 1) * sz + (long)y * sy + x];
      if (z < rz - 1) zp = in[idx + sz]; else if (pz) zp = in[(long)y * sy + x];
      real a = fmin(xm, xp), b = fmin(ym, yp), c = fmin(zm, zp), t;
      if (a > b) { t = a; a = b; b = t; }
      if (b > c) { t = b; b = c; c = t; }
      if (a > b) { t = a; a = b; b = t; }
      real u = 100.0;
      const real u1 = a + h;
      if (u1 <= b) {
        u = u1;
      } else {
        const real d2 = 2.0 * h * h - (a - b) * (a - b);
        const real u2 = (d2 < 0.0) ? 100.0 : (a + b + sqrt(d2)) * 0.5;
        if (u2 <= c) {
          u = u2;
        } else {
          const real s3 = a + b + c, qd = s3 * s3 - 3.0 * (a * a + b * b + c * c - h * h);
          u = (qd < 0.0) ? 100.0 : (s3 + sqrt(qd)) / 3.0;
        }
      }
      o2[idx] = u;
    }
    real* tmp = in;
    in = o2;
    o2 = tmp;
  }
  if (out) memcpy(out, in, sizeof(real) * (size_t)total);
  free(g);
  free(g2);
  free(fixedCell);
  return 0;
}
