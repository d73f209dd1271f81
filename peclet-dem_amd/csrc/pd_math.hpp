// pd_math.hpp — FP64 vector/quaternion math and analytic SDFs for the substep kernels.
//
// The operation order of every helper follows the reference math (src/dem_portable.hpp:25-113,
// src/solver_velocity.hpp:28-34, src/solver_position.hpp:21-31) so that, compiled with
// -ffp-contract=off, each result is the correctly rounded value of the same expression tree.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define PD_HD __host__ __device__ __forceinline__

namespace pd {

struct d3 {
  double x, y, z;
};
struct q4 {
  double x, y, z, w;
};

PD_HD d3 mk3(double x, double y, double z) { return d3{x, y, z}; }
PD_HD d3 xyz(const double4& a) { return d3{a.x, a.y, a.z}; }
PD_HD q4 toq(const double4& a) { return q4{a.x, a.y, a.z, a.w}; }
PD_HD double4 tod4(d3 a, double w) { return make_double4(a.x, a.y, a.z, w); }
PD_HD double4 tod4(q4 a) { return make_double4(a.x, a.y, a.z, a.w); }
PD_HD d3 add3(d3 a, d3 b) { return d3{a.x + b.x, a.y + b.y, a.z + b.z}; }
PD_HD d3 sub3(d3 a, d3 b) { return d3{a.x - b.x, a.y - b.y, a.z - b.z}; }
PD_HD d3 scale3(d3 a, double s) { return d3{a.x * s, a.y * s, a.z * s}; }
PD_HD d3 neg3(d3 a) { return d3{-a.x, -a.y, -a.z}; }
PD_HD double dot3(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
PD_HD double len3(d3 v) { return sqrt(dot3(v, v)); }
PD_HD d3 cross3(d3 a, d3 b) {
  return d3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
PD_HD bool nonzero3(d3 a) { return a.x != 0.0 || a.y != 0.0 || a.z != 0.0; }

// v' = v + w t + q x t with t = 2 q x v  (src/dem_portable.hpp:59-63)
PD_HD d3 rotate(q4 q, d3 v) {
  const d3 qv{q.x, q.y, q.z};
  const d3 t = scale3(cross3(qv, v), 2.0);
  return add3(add3(v, scale3(t, q.w)), cross3(qv, t));
}
PD_HD d3 invRotate(q4 q, d3 v) { return rotate(q4{-q.x, -q.y, -q.z, q.w}, v); }
PD_HD q4 quatInverse(q4 q) { return q4{-q.x, -q.y, -q.z, q.w}; }
PD_HD q4 quatMult(q4 a, q4 b) {
  return q4{a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
            a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
            a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w,
            a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}

// ---- analytic SDFs in canonical space (src/dem_portable.hpp:77-113) ----
enum ShapeKind { SHAPE_GRID_SDF = 0, SHAPE_SPHERE = 1, SHAPE_HOLLOW_CYLINDER = 2, SHAPE_BOX = 3 };

PD_HD double sdfSphere(d3 p, double r) { return len3(p) - r; }
PD_HD double sdfHollowCylinder(d3 p, double r_outer, double h, double thick) {
  const double r = sqrt(p.x * p.x + p.z * p.z);
  const double r_mid = r_outer - thick * 0.5;
  const double dx = fabs(r - r_mid) - thick * 0.5;
  const double dy = fabs(p.y) - h * 0.5;
  const double ox = fmax(dx, 0.0), oy = fmax(dy, 0.0);
  const double outside = sqrt(ox * ox + oy * oy);
  const double inside = fmin(fmax(dx, dy), 0.0);
  return outside + inside;
}
PD_HD double sdfBox(d3 p, double hx, double hy, double hz) {
  const double dx = fabs(p.x) - hx;
  const double dy = fabs(p.y) - hy;
  const double dz = fabs(p.z) - hz;
  const double ox = fmax(dx, 0.0), oy = fmax(dy, 0.0), oz = fmax(dz, 0.0);
  const double outside = sqrt(ox * ox + oy * oy + oz * oz);
  const double inside = fmin(fmax(dx, fmax(dy, dz)), 0.0);
  return outside + inside;
}

struct ShapeP {
  int type;
  int numShell;
  double p0, p1, p2, p3;
  // imported grid SDF (type SHAPE_GRID_SDF): samples x fastest at origin + (x,y,z) / invSpacing
  const double* grid;
  int nx, ny, nz;
  d3 origin, invSpacing;
};

// trilinear grid SDF with the off-grid residual added (object convention),
// src/narrowphase.hpp:97-135
PD_HD double sampleGridSdf(const ShapeP& s, d3 p) {
  const double fx = (p.x - s.origin.x) * s.invSpacing.x;
  const double fy = (p.y - s.origin.y) * s.invSpacing.y;
  const double fz = (p.z - s.origin.z) * s.invSpacing.z;
  const double cx = fmin(fmax(fx, 0.0), (double)(s.nx - 1));
  const double cy = fmin(fmax(fy, 0.0), (double)(s.ny - 1));
  const double cz = fmin(fmax(fz, 0.0), (double)(s.nz - 1));
  const int ix = (int)cx, iy = (int)cy, iz = (int)cz;
  const int ix1 = ix < s.nx - 1 ? ix + 1 : ix;
  const int iy1 = iy < s.ny - 1 ? iy + 1 : iy;
  const int iz1 = iz < s.nz - 1 ? iz + 1 : iz;
  const double tx = cx - ix, ty = cy - iy, tz = cz - iz;
  const long nxny = (long)s.nx * s.ny;
  const double* g = s.grid;
  auto at = [&](int x, int y, int z) { return g[(long)z * nxny + (long)y * s.nx + x]; };
  const double c00 = at(ix, iy, iz) * (1.0 - tx) + at(ix1, iy, iz) * tx;
  const double c10 = at(ix, iy1, iz) * (1.0 - tx) + at(ix1, iy1, iz) * tx;
  const double c01 = at(ix, iy, iz1) * (1.0 - tx) + at(ix1, iy, iz1) * tx;
  const double c11 = at(ix, iy1, iz1) * (1.0 - tx) + at(ix1, iy1, iz1) * tx;
  const double c0 = c00 * (1.0 - ty) + c10 * ty;
  const double c1 = c01 * (1.0 - ty) + c11 * ty;
  const double val = c0 * (1.0 - tz) + c1 * tz;
  const double rx = (s.invSpacing.x > 0.0) ? (fx - cx) / s.invSpacing.x : 0.0;
  const double ry = (s.invSpacing.y > 0.0) ? (fy - cy) / s.invSpacing.y : 0.0;
  const double rz = (s.invSpacing.z > 0.0) ? (fz - cz) / s.invSpacing.z : 0.0;
  return val + sqrt(rx * rx + ry * ry + rz * rz);
}

// sdfEvalShape, src/narrowphase.hpp:169-173
PD_HD double sdfShape(const ShapeP& s, d3 p) {
  if (s.type == SHAPE_SPHERE) return sdfSphere(p, s.p0);
  if (s.type == SHAPE_HOLLOW_CYLINDER) return sdfHollowCylinder(p, s.p0, s.p1, s.p2);
  if (s.type == SHAPE_BOX) return sdfBox(p, s.p0, s.p1, s.p2);
  if (s.type == SHAPE_GRID_SDF && s.grid) return sampleGridSdf(s, p);
  return 1e9;
}

// v^T I_world^-1 v, isotropic shortcut as the reference (src/solver_velocity.hpp:28-34)
PD_HD double genInvMass(d3 tau, d3 invI, q4 q) {
  if (invI.x == invI.y && invI.y == invI.z) return dot3(tau, tau) * invI.x;
  const d3 t = invRotate(q, tau);
  return t.x * t.x * invI.x + t.y * t.y * invI.y + t.z * t.z * invI.z;
}
// dw_world = R (invI_local * (R^T J))
PD_HD d3 angDelta(q4 q, d3 J, d3 invI) {
  const d3 Jl = invRotate(q, J);
  const d3 dwl{Jl.x * invI.x, Jl.y * invI.y, Jl.z * invI.z};
  return rotate(q, dwl);
}
// src/solver_position.hpp:21-24
PD_HD double computeW(d3 r, d3 dir, double invM, d3 invI) {
  const d3 rn = cross3(r, dir);
  return invM + rn.x * rn.x * invI.x + rn.y * rn.y * invI.y + rn.z * rn.z * invI.z;
}

// splitmix32 priority (src/contact_preprocessing.hpp:310-317), stored +1 so 0 means "no winner"
PD_HD unsigned long long colorKey1(int idx) {
  unsigned z = (unsigned)idx + 0x9e3779b9u;
  z = (z ^ (z >> 16)) * 0x21f0aaadu;
  z = (z ^ (z >> 15)) * 0x735a2d97u;
  z ^= (z >> 15);
  return ((((unsigned long long)(z & 0x7fffffffu)) << 32) | (unsigned long long)(unsigned)idx) + 1ull;
}

// order-preserving map double -> uint64 (for atomic min/max of doubles of any sign)
union DU {
  double d;
  unsigned long long u;
};
PD_HD unsigned long long dbits(double v) {
  DU x;
  x.d = v;
  return x.u;
}
PD_HD double bitsd(unsigned long long u) {
  DU x;
  x.u = u;
  return x.d;
}
PD_HD unsigned long long ordKey(double v) {
  const unsigned long long b = dbits(v);
  return (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
}
PD_HD double ordVal(unsigned long long k) {
  return bitsd((k & 0x8000000000000000ull) ? (k & 0x7fffffffffffffffull) : ~k);
}

}  // namespace pd
