// pd_hertz.hpp — Hertz–Mindlin soft-sphere engine and SDF walls (gfx950, FP64).
//
// src/solver_hertz.hpp (forces, integration) and src/sim.hpp:543-750 (Verlet driver).
// Forces are computed per pair / per wall candidate into per-contact records, then every body
// gathers its contributions in ascending pair order (pairs where it is the second body, which
// all precede, then its own range, then its walls in wall order) and integrates.  That is the
// order in which a serial run adds them, so no atomics are needed and the result is
// reproducible bit for bit.  Per-material-pair constants that need a logarithm (the damping
// ratio) are precomputed on the host.
#pragma once

#include "pd_kernels.hpp"

namespace pd {

constexpr int HZ_MAXWALLS = 4;  // Particles::kHertzMaxWalls, src/particles.hpp:130
#define HZ_SQRT56 0.91287092917527690
#define HZ_PI2 9.8696044010893586

// static world-space SDF wall, src/narrowphase.hpp:59-76
struct WallP {
  int nx, ny, nz, materialId;
  long off;
  d3 origin, invSpacing, linVel, angVel, center;
  double restitution, friction;
};

struct HzParams {
  double dt, gs, baseR, frac;
  d3 g;
  int numReal, numWalls, hasShapes, numPairs;
  ShapeP shape;
  WallP walls[HZ_MAXWALLS];
};

// per material pair (ma * 8 + mb) and per (material, wall) (ma * 8 + wall):
// [0] damping ratio beta_d, [1] mu, [2] E*, [3] G*
struct HzTab {
  double pair[64][4];
  double wall[64][4];
};

// src/solver_hertz.hpp:41-48 (host only: log is not correctly rounded on every library)
inline double hzBetaD(double e) {
  if (e <= 0.0) return -1.0;
  if (e >= 1.0) return 0.0;
  const double le = std::log(e);
  return le / std::sqrt(le * le + HZ_PI2);
}
PD_HD void hzStars(double Ei, double Ej, double ni, double nj, double* eStar, double* gStar) {
  *eStar = 1.0 / ((1.0 - ni * ni) / Ei + (1.0 - nj * nj) / Ej);
  *gStar = 1.0 / (2.0 * (2.0 - ni) * (1.0 + ni) / Ei + 2.0 * (2.0 - nj) * (1.0 + nj) / Ej);
}

struct HzDev {
  int2* pairs;        // cached Verlet pairs (lexicographic)
  int* iStart;        // [numReal + 1] own-range offsets into pairs
  int* jStart;        // [numReal + 1] offsets into jList
  int* jList;         // pair indices where the body is the second body, ascending
  double4* xi;        // shear history (.w = last patch stiffness for shapes)
  double4* f;         // per pair: force on the first body (.w = 1 when in contact)
  double4* tA;        // per pair: torque on the first body
  double4* tB;        // per pair: torque on the second body
  unsigned long long* prevKeys;
  double4* prevXi;
  double4* refPos;
  int* wStart;        // [numReal + 1] wall candidates per body (ascending wall)
  int* wCand;         // encoded i * 8 + wall
  double4* xiWall;    // [numReal * 4] (.w = last patch stiffness)
  double4* wf;        // per candidate force (.w = 1 when in contact)
  double4* wt;        // per candidate torque
  double* wallGrid;
  const HzTab* tab;
};

// trilinear wall SDF with the off-grid residual subtracted, src/narrowphase.hpp:137-165
__device__ __forceinline__ double sampleWallSdf(d3 p, const WallP& w, const double* grid) {
  const double fx = (p.x - w.origin.x) * w.invSpacing.x;
  const double fy = (p.y - w.origin.y) * w.invSpacing.y;
  const double fz = (p.z - w.origin.z) * w.invSpacing.z;
  const double cx = fmin(fmax(fx, 0.0), (double)(w.nx - 1));
  const double cy = fmin(fmax(fy, 0.0), (double)(w.ny - 1));
  const double cz = fmin(fmax(fz, 0.0), (double)(w.nz - 1));
  const int ix = (int)cx, iy = (int)cy, iz = (int)cz;
  const int ix1 = ix < w.nx - 1 ? ix + 1 : ix;
  const int iy1 = iy < w.ny - 1 ? iy + 1 : iy;
  const int iz1 = iz < w.nz - 1 ? iz + 1 : iz;
  const double tx = cx - ix, ty = cy - iy, tz = cz - iz;
  const long nxny = (long)w.nx * w.ny;
  const double* g = grid + w.off;
  auto at = [&](int x, int y, int z) { return g[(long)z * nxny + (long)y * w.nx + x]; };
  const double c00 = at(ix, iy, iz) * (1.0 - tx) + at(ix1, iy, iz) * tx;
  const double c10 = at(ix, iy1, iz) * (1.0 - tx) + at(ix1, iy1, iz) * tx;
  const double c01 = at(ix, iy, iz1) * (1.0 - tx) + at(ix1, iy, iz1) * tx;
  const double c11 = at(ix, iy1, iz1) * (1.0 - tx) + at(ix1, iy1, iz1) * tx;
  const double c0 = c00 * (1.0 - ty) + c10 * ty;
  const double c1 = c01 * (1.0 - ty) + c11 * ty;
  const double val = c0 * (1.0 - tz) + c1 * tz;
  const double rx = (w.invSpacing.x > 0.0) ? (fx - cx) / w.invSpacing.x : 0.0;
  const double ry = (w.invSpacing.y > 0.0) ? (fy - cy) / w.invSpacing.y : 0.0;
  const double rz = (w.invSpacing.z > 0.0) ? (fz - cz) / w.invSpacing.z : 0.0;
  return val - sqrt(rx * rx + ry * ry + rz * rz);
}
__device__ __forceinline__ d3 wallGradient(d3 p, const WallP& w, const double* grid) {
  const double hx = 0.5 / w.invSpacing.x;
  return d3{sampleWallSdf(d3{p.x + hx, p.y, p.z}, w, grid) -
                sampleWallSdf(d3{p.x - hx, p.y, p.z}, w, grid),
            sampleWallSdf(d3{p.x, p.y + hx, p.z}, w, grid) -
                sampleWallSdf(d3{p.x, p.y - hx, p.z}, w, grid),
            sampleWallSdf(d3{p.x, p.y, p.z + hx}, w, grid) -
                sampleWallSdf(d3{p.x, p.y, p.z - hx}, w, grid)};
}

// src/solver_hertz.hpp:53-84: force on A at the contact point; updates the shear history
__device__ __forceinline__ d3 hertzForce(double delta, d3 nhat, d3 vrel, double rStar,
                                         double mStar, double eStar, double gStar, double bd,
                                         double mu, double dt, d3& xi) {
  const double sq = sqrt(rStar * delta);
  const double sn = 2.0 * eStar * sq;
  const double st = 8.0 * gStar * sq;
  const double kn = (4.0 / 3.0) * eStar * sq;
  const double etaN = -2.0 * HZ_SQRT56 * bd * sqrt(sn * mStar);
  const double etaT = -2.0 * HZ_SQRT56 * bd * sqrt(st * mStar);
  const double vn = dot3(vrel, nhat);
  const double fnSpring = kn * delta;
  const double fn = fnSpring - etaN * vn;
  const d3 vt = sub3(vrel, scale3(nhat, vn));
  d3 x = add3(xi, scale3(vt, dt));
  x = sub3(x, scale3(nhat, dot3(x, nhat)));
  d3 ft = add3(scale3(x, -st), scale3(vt, -etaT));
  const double ftLen = sqrt(dot3(ft, ft));
  const double ftMax = mu * fabs(fnSpring);
  if (ftLen > ftMax) {
    const double sc = (ftLen > 1e-20) ? ftMax / ftLen : 0.0;
    ft = scale3(ft, sc);
    if (st > 1e-20) x = scale3(add3(ft, scale3(vt, etaT)), -1.0 / st);
  }
  xi = x;
  return add3(scale3(nhat, fn), ft);
}

// radii into posPred.w for the broadphase; minimum radius as an ordered key
__global__ __launch_bounds__(TPB) void k_hz_prep(const Dev* __restrict__ Dp, HzParams H, unsigned long long* minKey) {
  const Dev& D = *Dp;
  __shared__ unsigned long long sh[WAVES];
  const int i = gtid();
  unsigned long long k = ~0ull;
  if (i < H.numReal) {
    const double rad = D.scale[i] * H.gs * H.baseR;
    const double4 p = D.pos[i];
    D.posPred[i] = make_double4(p.x, p.y, p.z, rad);
    k = ordKey(rad);
  }
  k = waveMinU64(k);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = k;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long m = sh[0];
    for (int w = 1; w < WAVES; ++w) m = sh[w] < m ? sh[w] : m;
    if (m != ~0ull) atomicMin(minKey, m);
  }
}

// history of the outgoing list in key order: the jList order is (second body, first body)
// ascending, which is ascending (hi << 32 | lo)  (src/sim.hpp:596-624)
__global__ void k_hz_commit(HzDev Z, int n) {
  const int q = gtid();
  if (q >= n) return;
  const int p = Z.jList[q];
  const int2 ab = Z.pairs[p];
  Z.prevKeys[q] = ((unsigned long long)(unsigned)ab.y << 32) | (unsigned)ab.x;
  Z.prevXi[q] = Z.xi[p];
}
// new list: copy, carry history by key, clear the patch stiffness (src/sim.hpp:559-590)
__global__ void k_hz_carry(const Dev* __restrict__ Dp, HzDev Z, int np, int prevCount) {
  const Dev& D = *Dp;
  const int p = gtid();
  if (p >= np) return;
  const int2 ab = D.pairs[p];
  Z.pairs[p] = ab;
  const unsigned long long k = ((unsigned long long)(unsigned)ab.y << 32) | (unsigned)ab.x;
  int l = 0, h = prevCount;
  while (l < h) {
    const int m = (l + h) >> 1;
    if (Z.prevKeys[m] < k) l = m + 1;
    else h = m;
  }
  const bool hit = (l < prevCount && Z.prevKeys[l] == k);
  const double4 x = hit ? Z.prevXi[l] : make_double4(0, 0, 0, 0);
  Z.xi[p] = make_double4(x.x, x.y, x.z, 0.0);
}
__global__ void k_hz_jcount(HzDev Z, int np, int* cnt) {
  const int p = gtid();
  if (p >= np) return;
  atomicAdd(&cnt[Z.pairs[p].y], 1);
}
// cnt is reused as the fill cursor (zeroed before)
__global__ void k_hz_jfill(HzDev Z, int np, int* cnt) {
  const int p = gtid();
  if (p >= np) return;
  const int j = Z.pairs[p].y;
  Z.jList[Z.jStart[j] + atomicAdd(&cnt[j], 1)] = p;
}
__global__ void k_hz_jsort(HzDev Z, HzParams H) {
  const int b = gtid();
  if (b >= H.numReal) return;
  int* seg = Z.jList + Z.jStart[b];
  const int n = Z.jStart[b + 1] - Z.jStart[b];
  for (int a = 1; a < n; ++a) {
    const int x = seg[a];
    int c = a - 1;
    while (c >= 0 && seg[c] > x) {
      seg[c + 1] = seg[c];
      --c;
    }
    seg[c + 1] = x;
  }
}
__global__ void k_hz_ref(const Dev* __restrict__ Dp, HzDev Z, HzParams H) {
  const Dev& D = *Dp;
  const int i = gtid();
  if (i >= H.numReal) return;
  Z.refPos[i] = D.pos[i];
}
// wall candidates: bodies within radius + skin of a wall's zero level (src/solver_hertz.hpp:343-370)
template <bool WRITE>
__global__ void k_hz_wallcand(const Dev* __restrict__ Dp, HzDev Z, HzParams H, double skin, int* cnt) {
  const Dev& D = *Dp;
  const int i = gtid();
  if (i >= H.numReal) return;
  const double4 pp = D.posPred[i];
  const d3 p{pp.x, pp.y, pp.z};
  int n = 0;
  const int base = WRITE ? Z.wStart[i] : 0;
  for (int wi = 0; wi < H.numWalls; ++wi) {
    const double s = sampleWallSdf(p, H.walls[wi], Z.wallGrid);
    if (s < pp.w + skin) {
      if (WRITE) Z.wCand[base + n] = i * 8 + wi;
      ++n;
    }
  }
  if (!WRITE) cnt[i] = n;
}

// spheres, src/solver_hertz.hpp:88-144
__global__ __launch_bounds__(TPB) void k_hz_pairs(const Dev* __restrict__ Dp, HzDev Z, HzParams H) {
  const Dev& D = *Dp;
  const int idx = gtid();
  if (idx >= H.numPairs) return;
  const int2 ab = Z.pairs[idx];
  const int i = ab.x, j = ab.y;
  const d3 pi = xyz(D.pos[i]), pj = xyz(D.pos[j]);
  const d3 dx = sub3(pi, pj);
  const double d = sqrt(dot3(dx, dx));
  const double ri = D.posPred[i].w, rj = D.posPred[j].w;
  const double delta = ri + rj - d;
  if (delta <= 0.0 || d < 1e-12) {
    Z.xi[idx] = make_double4(0, 0, 0, 0);
    Z.f[idx] = make_double4(0, 0, 0, 0);
    return;
  }
  const d3 nhat = scale3(dx, 1.0 / d);
  const d3 rA = scale3(nhat, -(ri - 0.5 * delta));
  const d3 rB = scale3(nhat, (rj - 0.5 * delta));
  const Body64 bi = D.vw0[i], bj = D.vw0[j];
  const d3 vrel = sub3(add3(xyz(bi.v), cross3(xyz(bi.w), rA)),
                       add3(xyz(bj.v), cross3(xyz(bj.w), rB)));
  const int t = (int)D.matId[i] * 8 + (int)D.matId[j];
  const double* c = Z.tab->pair[t];
  const double rStar = ri * rj / (ri + rj);
  const double imSum = D.bc[i].iim.w + D.bc[j].iim.w;
  const double mStar = (imSum > 0.0) ? 1.0 / imSum : 0.0;
  const double4 x4 = Z.xi[idx];
  d3 x{x4.x, x4.y, x4.z};
  const d3 f = hertzForce(delta, nhat, vrel, rStar, mStar, c[2], c[3], c[0], c[1], H.dt, x);
  Z.xi[idx] = make_double4(x.x, x.y, x.z, 0.0);
  Z.f[idx] = tod4(f, 1.0);
  Z.tA[idx] = tod4(cross3(rA, f), 0.0);
  Z.tB[idx] = tod4(cross3(rB, scale3(f, -1.0)), 0.0);
}

// point-shell shapes, src/solver_hertz.hpp:454-606
__global__ __launch_bounds__(TPB) void k_hz_shape_pairs(const Dev* __restrict__ Dp, HzDev Z, HzParams H) {
  const Dev& D = *Dp;
  const int idx = gtid();
  if (idx >= H.numPairs) return;
  const int2 ab = Z.pairs[idx];
  const int idA = ab.x, idB = ab.y;
  const d3 posA = xyz(D.pos[idA]), posB = xyz(D.pos[idB]);
  const BodyC cA = D.bc[idA], cB = D.bc[idB];
  const q4 qA = toq(cA.q), qB = toq(cB.q);
  const double effScaleA = D.scale[idA] * H.gs, effScaleB = D.scale[idB] * H.gs;
  const int countA = H.shape.numShell;
  const bool sphere = (H.shape.type == SHAPE_SPHERE);
  const int iter = countA > 0 ? countA : 1;
  const double* c = Z.tab->pair[(int)D.matId[idA] * 8 + (int)D.matId[idB]];
  const double eStar = c[2], gStar = c[3], bd = c[0], mu = c[1];
  const double rcA = (sphere ? 1.0 : H.frac) * H.shape.p0 * effScaleA;
  const double rcB = (sphere ? 1.0 : H.frac) * H.shape.p0 * effScaleB;
  const double rStar = rcA * rcB / (rcA + rcB);
  const double imSum = cA.iim.w + cB.iim.w;
  const double mStar = (imSum > 0.0) ? 1.0 / imSum : 0.0;
  const Body64 bA = D.vw0[idA], bB = D.vw0[idB];
  const d3 vA = xyz(bA.v), wA = xyz(bA.w), vB = xyz(bB.v), wB = xyz(bB.w);
  double fnSum = 0.0, snSum = 0.0, dMax = 0.0;
  d3 cSum{0, 0, 0}, nSum{0, 0, 0}, fAcc{0, 0, 0}, tAccA{0, 0, 0}, tAccB{0, 0, 0};
  const double4 x4 = Z.xi[idx];
  const double snTotPrev = x4.w;
  const double dampC = (mStar > 0.0) ? -2.0 * HZ_SQRT56 * bd * sqrt(mStar) : 0.0;
  for (int k = 0; k < iter; ++k) {
    d3 pLocalA{0.0, 0.0, 0.0};
    double pointRadius = 0.0;
    if (countA > 0) pLocalA = xyz(D.shell[k]);
    else if (sphere) pointRadius = H.shape.p0 * effScaleA;
    const d3 pWorld = add3(posA, rotate(qA, scale3(pLocalA, effScaleA)));
    const d3 pCanB = scale3(invRotate(qB, sub3(pWorld, posB)), 1.0 / effScaleB);
    const double dist = sdfShape(H.shape, pCanB) * effScaleB;
    const double effDist = dist - pointRadius;
    if (effDist >= 0.0) continue;
    const double delta = -effDist;
    const double eps = 1e-4;
    d3 nLoc{sdfShape(H.shape, d3{pCanB.x + eps, pCanB.y, pCanB.z}) -
                sdfShape(H.shape, d3{pCanB.x - eps, pCanB.y, pCanB.z}),
            sdfShape(H.shape, d3{pCanB.x, pCanB.y + eps, pCanB.z}) -
                sdfShape(H.shape, d3{pCanB.x, pCanB.y - eps, pCanB.z}),
            sdfShape(H.shape, d3{pCanB.x, pCanB.y, pCanB.z + eps}) -
                sdfShape(H.shape, d3{pCanB.x, pCanB.y, pCanB.z - eps})};
    const double len = len3(nLoc);
    nLoc = (len > 1e-9) ? scale3(nLoc, 1.0 / len) : d3{0.0, 1.0, 0.0};
    const d3 nW = rotate(qB, nLoc);
    const d3 pSurfA = sub3(pWorld, scale3(nW, pointRadius));
    const d3 rA = sub3(pSurfA, posA);
    const d3 rB = sub3(add3(pSurfA, scale3(nW, delta)), posB);
    const d3 vrelP = sub3(add3(vA, cross3(wA, rA)), add3(vB, cross3(wB, rB)));
    const double vnP = dot3(vrelP, nW);
    const double sq = sqrt(rStar * delta);
    const double fnSpring = (4.0 / 3.0) * eStar * sq * delta;
    const double snI = 2.0 * eStar * sq;
    const double snTot = (snTotPrev > snI) ? snTotPrev : snI;
    const double etaI = dampC * snI / sqrt(snTot);
    const d3 f = scale3(nW, fnSpring - etaI * vnP);
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
    Z.xi[idx] = make_double4(0, 0, 0, 0);
    Z.f[idx] = make_double4(0, 0, 0, 0);
    return;
  }
  const d3 cpt = scale3(cSum, 1.0 / fnSum);
  d3 nhat = nSum;
  const double nl = len3(nhat);
  if (nl > 1e-9) nhat = scale3(nhat, 1.0 / nl);
  const d3 rA = sub3(cpt, posA), rB = sub3(cpt, posB);
  const d3 vrel = sub3(add3(vA, cross3(wA, rA)), add3(vB, cross3(wB, rB)));
  const d3 vt = sub3(vrel, scale3(nhat, dot3(vrel, nhat)));
  const double st = 8.0 * gStar * sqrt(rStar * dMax);
  const double etaT = -2.0 * HZ_SQRT56 * bd * sqrt(st * mStar);
  d3 x{x4.x, x4.y, x4.z};
  x = add3(x, scale3(vt, H.dt));
  x = sub3(x, scale3(nhat, dot3(x, nhat)));
  d3 ft = add3(scale3(x, -st), scale3(vt, -etaT));
  const double ftLen = len3(ft);
  const double ftMax = mu * fnSum;
  if (ftLen > ftMax) {
    const double sc = (ftLen > 1e-20) ? ftMax / ftLen : 0.0;
    ft = scale3(ft, sc);
    if (st > 1e-20) x = scale3(add3(ft, scale3(vt, etaT)), -1.0 / st);
  }
  Z.xi[idx] = make_double4(x.x, x.y, x.z, snSum);
  fAcc = add3(fAcc, ft);
  tAccA = add3(tAccA, cross3(rA, ft));
  tAccB = add3(tAccB, cross3(rB, scale3(ft, -1.0)));
  Z.f[idx] = tod4(fAcc, 1.0);
  Z.tA[idx] = tod4(tAccA, 0.0);
  Z.tB[idx] = tod4(tAccB, 0.0);
}

// wall forces per candidate, src/solver_hertz.hpp:148-338
__global__ __launch_bounds__(TPB) void k_hz_walls(const Dev* __restrict__ Dp, HzDev Z, HzParams H, int numCand) {
  const Dev& D = *Dp;
  const int ci = gtid();
  if (ci >= numCand) return;
  const int enc = Z.wCand[ci];
  const int i = enc / 8, wi = enc % 8;
  const d3 p = xyz(D.pos[i]);
  const double r = D.posPred[i].w;
  const WallP& w = H.walls[wi];
  const int slot = i * HZ_MAXWALLS + wi;
  const double* c = Z.tab->wall[(int)D.matId[i] * 8 + wi];
  const double bd = c[0], mu = c[1], eStar = c[2], gStar = c[3];
  const BodyC bc = D.bc[i];
  const double mStar = (bc.iim.w > 0.0) ? 1.0 / bc.iim.w : 0.0;
  const Body64 b = D.vw0[i];
  const d3 vP = xyz(b.v), wP = xyz(b.w);
  if (H.hasShapes) {
    const q4 qP = toq(bc.q);
    const double effScale = D.scale[i] * H.gs;
    const double rc = H.frac * H.shape.p0 * effScale;
    const double4 x4 = Z.xiWall[slot];
    const double snTotPrev = x4.w;
    const double dampC = (mStar > 0.0) ? -2.0 * HZ_SQRT56 * bd * sqrt(mStar) : 0.0;
    double fnSum = 0.0, snSum = 0.0, dMax = 0.0;
    d3 cSum{0, 0, 0}, nSum{0, 0, 0}, fAcc{0, 0, 0}, tAcc{0, 0, 0};
    for (int k = 0; k < H.shape.numShell; ++k) {
      const d3 pW = add3(p, rotate(qP, scale3(xyz(D.shell[k]), effScale)));
      const double sd = sampleWallSdf(pW, w, Z.wallGrid);
      if (sd >= 0.0) continue;
      const double delta = -sd;
      const d3 g = wallGradient(pW, w, Z.wallGrid);
      const double gl = len3(g);
      if (gl < 1e-12) continue;
      const d3 nhat = scale3(g, 1.0 / gl);
      const d3 rA = sub3(pW, p);
      const d3 vWall = add3(w.linVel, cross3(w.angVel, sub3(pW, w.center)));
      const d3 vrelP = sub3(add3(vP, cross3(wP, rA)), vWall);
      const double vnP = dot3(vrelP, nhat);
      const double sq = sqrt(rc * delta);
      const double fnSpring = (4.0 / 3.0) * eStar * sq * delta;
      const double snI = 2.0 * eStar * sq;
      const double snTot = (snTotPrev > snI) ? snTotPrev : snI;
      const double etaI = dampC * snI / sqrt(snTot);
      const d3 f = scale3(nhat, fnSpring - etaI * vnP);
      fAcc = add3(fAcc, f);
      tAcc = add3(tAcc, cross3(rA, f));
      fnSum += fnSpring;
      snSum += 2.0 * eStar * sq;
      cSum = add3(cSum, scale3(pW, fnSpring));
      nSum = add3(nSum, scale3(nhat, fnSpring));
      if (delta > dMax) dMax = delta;
    }
    if (fnSum <= 0.0) {
      Z.xiWall[slot] = make_double4(0, 0, 0, 0);
      Z.wf[ci] = make_double4(0, 0, 0, 0);
      return;
    }
    const d3 cpt = scale3(cSum, 1.0 / fnSum);
    d3 nhat = nSum;
    const double nl = len3(nhat);
    if (nl > 1e-9) nhat = scale3(nhat, 1.0 / nl);
    const d3 rA = sub3(cpt, p);
    const d3 vWall = add3(w.linVel, cross3(w.angVel, sub3(cpt, w.center)));
    const d3 vrel = sub3(add3(vP, cross3(wP, rA)), vWall);
    const d3 vt = sub3(vrel, scale3(nhat, dot3(vrel, nhat)));
    const double st = 8.0 * gStar * sqrt(rc * dMax);
    const double etaT = -2.0 * HZ_SQRT56 * bd * sqrt(st * mStar);
    d3 x{x4.x, x4.y, x4.z};
    x = add3(x, scale3(vt, H.dt));
    x = sub3(x, scale3(nhat, dot3(x, nhat)));
    d3 ft = add3(scale3(x, -st), scale3(vt, -etaT));
    const double ftLen = len3(ft);
    const double ftMax = mu * fnSum;
    if (ftLen > ftMax) {
      const double sc = (ftLen > 1e-20) ? ftMax / ftLen : 0.0;
      ft = scale3(ft, sc);
      if (st > 1e-20) x = scale3(add3(ft, scale3(vt, etaT)), -1.0 / st);
    }
    Z.xiWall[slot] = make_double4(x.x, x.y, x.z, snSum);
    fAcc = add3(fAcc, ft);
    tAcc = add3(tAcc, cross3(rA, ft));
    Z.wf[ci] = tod4(fAcc, 1.0);
    Z.wt[ci] = tod4(tAcc, 0.0);
    return;
  }
  const double sd = sampleWallSdf(p, w, Z.wallGrid);
  const double delta = r - sd;
  if (delta <= 0.0) {
    Z.xiWall[slot] = make_double4(0, 0, 0, 0);
    Z.wf[ci] = make_double4(0, 0, 0, 0);
    return;
  }
  Z.wf[ci] = make_double4(0, 0, 0, 0);
  const d3 g = wallGradient(p, w, Z.wallGrid);
  const double gl = sqrt(dot3(g, g));
  if (gl < 1e-12) return;
  const d3 nhat = scale3(g, 1.0 / gl);
  const d3 rA = scale3(nhat, -(r - 0.5 * delta));
  const d3 cpt = add3(p, rA);
  const d3 vWall = add3(w.linVel, cross3(w.angVel, sub3(cpt, w.center)));
  const d3 vrel = sub3(add3(vP, cross3(wP, rA)), vWall);
  const double4 x4 = Z.xiWall[slot];
  d3 x{x4.x, x4.y, x4.z};
  const d3 f = hertzForce(delta, nhat, vrel, r, mStar, eStar, gStar, bd, mu, H.dt, x);
  Z.xiWall[slot] = make_double4(x.x, x.y, x.z, x4.w);
  Z.wf[ci] = tod4(f, 1.0);
  Z.wt[ci] = tod4(cross3(rA, f), 0.0);
}

// gather in serial order, then symplectic Euler (src/solver_hertz.hpp:374-422)
__global__ __launch_bounds__(TPB) void k_hz_integrate(const Dev* __restrict__ Dp, HzDev Z, HzParams H) {
  const Dev& D = *Dp;
  const int i = gtid();
  if (i >= H.numReal) return;
  d3 F{0.0, 0.0, 0.0}, T{0.0, 0.0, 0.0};
  for (int q = Z.jStart[i], e = Z.jStart[i + 1]; q < e; ++q) {
    const int p = Z.jList[q];
    const double4 f = Z.f[p];
    if (f.w == 0.0) continue;
    F = add3(F, neg3(xyz(f)));
    T = add3(T, xyz(Z.tB[p]));
  }
  for (int p = Z.iStart[i], e = Z.iStart[i + 1]; p < e; ++p) {
    const double4 f = Z.f[p];
    if (f.w == 0.0) continue;
    F = add3(F, xyz(f));
    T = add3(T, xyz(Z.tA[p]));
  }
  if (H.numWalls > 0)
    for (int c = Z.wStart[i], e = Z.wStart[i + 1]; c < e; ++c) {
      const double4 f = Z.wf[c];
      if (f.w == 0.0) continue;
      F = add3(F, xyz(f));
      T = add3(T, xyz(Z.wt[c]));
    }
  const BodyC bc = D.bc[i];
  const double im = bc.iim.w;
  if (!(im > 0.0)) return;
  const d3 iI{bc.iim.x, bc.iim.y, bc.iim.z};
  d3 dw;
  if (iI.x == iI.y && iI.y == iI.z) {
    dw = d3{T.x * iI.x, T.y * iI.x, T.z * iI.x};
  } else {
    const q4 q = toq(bc.q);
    const d3 tl = invRotate(q, T);
    dw = rotate(q, d3{tl.x * iI.x, tl.y * iI.y, tl.z * iI.z});
  }
  Body64 b = D.vw0[i];
  const double dt = H.dt;
  b.v.x = b.v.x + (F.x * im + H.g.x) * dt;
  b.v.y = b.v.y + (F.y * im + H.g.y) * dt;
  b.v.z = b.v.z + (F.z * im + H.g.z) * dt;
  b.w.x = b.w.x + dw.x * dt;
  b.w.y = b.w.y + dw.y * dt;
  b.w.z = b.w.z + dw.z * dt;
  double4 p = D.pos[i];
  p.x = p.x + b.v.x * dt;
  p.y = p.y + b.v.y * dt;
  p.z = p.z + b.v.z * dt;
  D.vw0[i] = b;
  D.pos[i] = p;
  if (H.hasShapes) {
    q4 q = toq(bc.q);
    const d3 w{b.w.x, b.w.y, b.w.z};
    const double h = 0.5 * dt;
    const q4 dq{h * (w.x * q.w + w.y * q.z - w.z * q.y), h * (w.y * q.w + w.z * q.x - w.x * q.z),
                h * (w.z * q.w + w.x * q.y - w.y * q.x), h * (-w.x * q.x - w.y * q.y - w.z * q.z)};
    q = q4{q.x + dq.x, q.y + dq.y, q.z + dq.z, q.w + dq.w};
    const double qn = sqrt(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
    if (qn > 1e-12) D.bc[i].q = make_double4(q.x / qn, q.y / qn, q.z / qn, q.w / qn);
  }
}

// max |pos - ref|^2 since the last build (ordered bits of a non-negative double)
__global__ __launch_bounds__(TPB) void k_hz_disp(const Dev* __restrict__ Dp, HzDev Z, HzParams H,
                                                 unsigned long long* out) {
  const Dev& D = *Dp;
  __shared__ unsigned long long sh[WAVES];
  unsigned long long m = 0;
  for (int i = gtid(); i < H.numReal; i += gridDim.x * blockDim.x) {
    const d3 d = sub3(xyz(D.pos[i]), xyz(Z.refPos[i]));
    const unsigned long long k = dbits(dot3(d, d));
    m = k > m ? k : m;
  }
  blockAtomicMaxU64(out, m, sh);
}

// ============================================================================================
// wall contacts of the XPBD substep, src/narrowphase.hpp:278-367: per real body, per wall, per
// surface probe (shell point, or the centre for an analytic sphere); appended after the planes
// ============================================================================================
struct WallSet {
  int n;
  WallP w[8];
  const double* grid;
};
template <bool WRITE>
__global__ __launch_bounds__(TPB) void k_np_walls(const Dev* __restrict__ Dp, Params P, WallSet W, int slotBase) {
  const Dev& D = *Dp;
  __shared__ unsigned long long sh[WAVES];
  const int i = gtid();
  double ovl = 0.0;
  if (i < P.numReal) {
    const d3 posA = xyz(D.posPred[i]);
    const double s = D.scale[i] * P.gs;
    double baseR = P.shape.p0;
    if (baseR == 0.0) baseR = 1.0;
    const double radius = baseR * s;
    const int numPts = P.shape.numShell;
    const q4 qA = toq(D.quatPred[i]);
    const double eps = 1e-4;
    int n = 0;
    const int base = WRITE ? D.noff[slotBase + i] : 0;
    const int ma = (int)D.matId[i];
    for (int wi = 0; wi < W.n; ++wi) {
      const WallP& w = W.w[wi];
      const int iter = numPts > 0 ? numPts : 1;
      for (int k = 0; k < iter; ++k) {
        d3 rA{0.0, 0.0, 0.0};
        if (numPts > 0) rA = rotate(qA, scale3(xyz(D.shell[k]), s));
        const d3 pw = add3(posA, rA);
        const double sdf = sampleWallSdf(pw, w, W.grid);
        d3 nrm{sampleWallSdf(d3{pw.x + eps, pw.y, pw.z}, w, W.grid) -
                   sampleWallSdf(d3{pw.x - eps, pw.y, pw.z}, w, W.grid),
               sampleWallSdf(d3{pw.x, pw.y + eps, pw.z}, w, W.grid) -
                   sampleWallSdf(d3{pw.x, pw.y - eps, pw.z}, w, W.grid),
               sampleWallSdf(d3{pw.x, pw.y, pw.z + eps}, w, W.grid) -
                   sampleWallSdf(d3{pw.x, pw.y, pw.z - eps}, w, W.grid)};
        const double ln = len3(nrm);
        nrm = (ln > 1e-9) ? scale3(nrm, 1.0 / ln) : d3{0.0, 1.0, 0.0};
        const double dist = (numPts > 0) ? sdf : sdf - radius;
        if (dist >= P.margin) continue;
        if (!WRITE) {
          ++n;
          continue;
        }
        const d3 rAeff = (numPts > 0) ? rA : scale3(nrm, -radius);
        const d3 pSurfA = add3(posA, rAeff);
        const d3 pWall = sub3(pSurfA, scale3(nrm, dist));
        if (dist < 0.0 && -dist > ovl) ovl = -dist;
        const d3 vWall = add3(w.linVel, cross3(w.angVel, sub3(pWall, w.center)));
        double eb = w.restitution, mub = w.friction;
        if (w.materialId >= 0 && P.hasPairTable) {
          const int t = (ma * 8 + w.materialId) * 2;
          eb = D.pairTable[t];
          mub = D.pairTable[t + 1];
        }
        writeContact(D, base + n, i, -1, nrm, rAeff, pWall, dist, eb, mub, vWall);
        ++n;
      }
    }
    if (!WRITE) D.ncnt[slotBase + i] = n;
  }
  if (WRITE) blockAtomicMaxU64(D.maxOverlapBits, dbits(ovl), sh);
}

}  // namespace pd
