// pd_adjoint.hpp — differentiable replay of the XPBD substep (spheres, planes, periodic ghosts)
// and its hand-written adjoint.
//
// The forward substep (pd_kernels.hpp) records, per substep, only its discrete structure: the
// contact list, the manifold ranges, the velocity and position colour orders, the sweep counts,
// the stabilization flags and the warm-start map into the previous ledger.  The replay below
// re-evaluates every continuous value of the substep from the body state with the discrete
// structure held fixed, in the same expression trees as the forward kernels, so for T = double
// it reproduces the forward state bit for bit.  Each per-element function (predict, contact,
// manifold term, row set-up, warm start, PGS row with the Coulomb cone, re-integration, position
// row set-up, XPBD row) is written once, generic over the scalar type:
//   * T = double: the primal replay (it also records, per sweep, the inputs of every row);
//   * T = DN<K>: forward mode with K tangents — the tangent run of the whole substep, and, inside
//     the reverse kernels, the local Jacobian of one element (the vector-Jacobian product of an
//     element is assembled from forward-mode columns, with the branches — clamp, cone, skip —
//     taken where the primal took them).
// The reverse pass walks the substep backwards: final commit, position sweeps in reverse sweep
// and colour order, position-row set-up, re-integration, ledger, velocity sweeps (stabilization,
// main loop, warm start) in reverse order, row set-up, manifold reduction, contacts, ghosts,
// predict.  Element orderings follow src/sim.hpp:97-488.
#pragma once

#include "pd_kernels.hpp"

namespace pd {
namespace ad {

// ---------------------------------------------------------------- forward-mode numbers
template <int K>
struct DN {
  double v;
  double d[K];
  PD_HD DN() : v(0.0) {
#pragma unroll
    for (int i = 0; i < K; ++i) d[i] = 0.0;
  }
  PD_HD DN(double x) : v(x) {  // NOLINT: a constant is a number with zero tangents
#pragma unroll
    for (int i = 0; i < K; ++i) d[i] = 0.0;
  }
  friend PD_HD DN operator+(const DN& a, const DN& b) {
    DN r;
    r.v = a.v + b.v;
#pragma unroll
    for (int i = 0; i < K; ++i) r.d[i] = a.d[i] + b.d[i];
    return r;
  }
  friend PD_HD DN operator-(const DN& a, const DN& b) {
    DN r;
    r.v = a.v - b.v;
#pragma unroll
    for (int i = 0; i < K; ++i) r.d[i] = a.d[i] - b.d[i];
    return r;
  }
  friend PD_HD DN operator-(const DN& a) {
    DN r;
    r.v = -a.v;
#pragma unroll
    for (int i = 0; i < K; ++i) r.d[i] = -a.d[i];
    return r;
  }
  friend PD_HD DN operator*(const DN& a, const DN& b) {
    DN r;
    r.v = a.v * b.v;
#pragma unroll
    for (int i = 0; i < K; ++i) r.d[i] = a.v * b.d[i] + a.d[i] * b.v;
    return r;
  }
  friend PD_HD DN operator/(const DN& a, const DN& b) {
    DN r;
    r.v = a.v / b.v;
#pragma unroll
    for (int i = 0; i < K; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) / b.v;
    return r;
  }
  PD_HD DN& operator+=(const DN& b) { return *this = *this + b; }
};

PD_HD double val(double x) { return x; }
template <int K>
PD_HD double val(const DN<K>& x) { return x.v; }
PD_HD double tsqrt(double x) { return sqrt(x); }
template <int K>
PD_HD DN<K> tsqrt(const DN<K>& a) {
  DN<K> r;
  r.v = sqrt(a.v);
  const double h = r.v > 0.0 ? 0.5 / r.v : 0.0;
#pragma unroll
  for (int i = 0; i < K; ++i) r.d[i] = a.d[i] * h;
  return r;
}
// tangent-carrying types: the derivative may follow a branch the primal does not take
template <class T>
struct IsDual {
  static constexpr bool value = false;
};
template <int K>
struct IsDual<DN<K>> {
  static constexpr bool value = true;
};
// set the value of a, keep its tangents
PD_HD void setVal(double& a, double v) { a = v; }
template <int K>
PD_HD void setVal(DN<K>& a, double v) {
  a.v = v;
}
template <class T>
PD_HD T tabs(const T& a) {
  return val(a) < 0.0 ? -a : a;
}
// fmax(a, b) for a constant b
template <class T>
PD_HD T tmax(const T& a, double b) {
  return val(a) >= b ? a : T(b);
}

// ---------------------------------------------------------------- vectors of T
template <class T>
struct V3 {
  T x, y, z;
};
template <class T>
struct Q4 {
  T x, y, z, w;
};
template <class T>
PD_HD V3<T> v3(const d3& a) {
  return V3<T>{T(a.x), T(a.y), T(a.z)};
}
template <class T>
PD_HD Q4<T> q4t(const q4& a) {
  return Q4<T>{T(a.x), T(a.y), T(a.z), T(a.w)};
}
template <class T>
PD_HD V3<T> add3(V3<T> a, V3<T> b) {
  return V3<T>{a.x + b.x, a.y + b.y, a.z + b.z};
}
template <class T>
PD_HD V3<T> sub3(V3<T> a, V3<T> b) {
  return V3<T>{a.x - b.x, a.y - b.y, a.z - b.z};
}
template <class T, class S>
PD_HD V3<T> scale3(V3<T> a, S s) {
  const T t = T(s);
  return V3<T>{a.x * t, a.y * t, a.z * t};
}
template <class T>
PD_HD V3<T> neg3(V3<T> a) {
  return V3<T>{-a.x, -a.y, -a.z};
}
template <class T>
PD_HD T dot3(V3<T> a, V3<T> b) {
  return a.x * b.x + a.y * b.y + a.z * b.z;
}
template <class T>
PD_HD V3<T> cross3(V3<T> a, V3<T> b) {
  return V3<T>{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T>
PD_HD bool nonzero3(V3<T> a) {
  return val(a.x) != 0.0 || val(a.y) != 0.0 || val(a.z) != 0.0;
}
template <class T>
PD_HD V3<T> rotate(Q4<T> q, V3<T> v) {
  const V3<T> qv{q.x, q.y, q.z};
  const V3<T> t = scale3(cross3(qv, v), 2.0);
  return add3(add3(v, scale3(t, q.w)), cross3(qv, t));
}
template <class T>
PD_HD V3<T> invRotate(Q4<T> q, V3<T> v) {
  return rotate(Q4<T>{-q.x, -q.y, -q.z, q.w}, v);
}
template <class T>
PD_HD Q4<T> quatInverse(Q4<T> q) {
  return Q4<T>{-q.x, -q.y, -q.z, q.w};
}
template <class T>
PD_HD Q4<T> quatMult(Q4<T> a, Q4<T> b) {
  return Q4<T>{a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
               a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
               a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w,
               a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}
template <class T>
PD_HD T genInvMass(V3<T> tau, d3 invI, Q4<T> q) {
  if (invI.x == invI.y && invI.y == invI.z) return dot3(tau, tau) * T(invI.x);
  const V3<T> t = invRotate(q, tau);
  return t.x * t.x * T(invI.x) + t.y * t.y * T(invI.y) + t.z * t.z * T(invI.z);
}
template <class T>
PD_HD V3<T> angDelta(Q4<T> q, V3<T> J, d3 invI) {
  const V3<T> Jl = invRotate(q, J);
  const V3<T> dwl{Jl.x * T(invI.x), Jl.y * T(invI.y), Jl.z * T(invI.z)};
  return rotate(q, dwl);
}
template <class T>
PD_HD T computeW(V3<T> r, V3<T> dir, double invM, d3 invI) {
  const V3<T> rn = cross3(r, dir);
  return T(invM) + rn.x * rn.x * T(invI.x) + rn.y * rn.y * T(invI.y) + rn.z * rn.z * T(invI.z);
}
template <class T>
PD_HD T len3(V3<T> v) {
  return tsqrt(dot3(v, v));
}
template <class T>
PD_HD V3<T> ld3(const T* a) {
  return V3<T>{a[0], a[1], a[2]};
}
template <class T>
PD_HD void st3(T* a, V3<T> v) {
  a[0] = v.x;
  a[1] = v.y;
  a[2] = v.z;
}
template <class T>
PD_HD Q4<T> ld4(const T* a) {
  return Q4<T>{a[0], a[1], a[2], a[3]};
}
template <class T>
PD_HD void st4(T* a, Q4<T> v) {
  a[0] = v.x;
  a[1] = v.y;
  a[2] = v.z;
  a[3] = v.w;
}

// ---------------------------------------------------------------- element functions
// Each element maps NI inputs to NO outputs; constants are members.

// predictVelocityKokkos, src/integration.hpp:44-93: in x v w q -> out vh wh xt
struct ElPredict {
  static constexpr int NI = 13, NO = 9;
  double invM, dt;
  d3 invI, g, F;
  template <class T>
  PD_HD void operator()(const T (&in)[NI], T (&out)[NO]) const {
    V3<T> x = ld3(in), v = ld3(in + 3);
    const V3<T> w0 = ld3(in + 6);
    const Q4<T> q = ld4(in + 9);
    if (invM > 0.0) {
      v = add3(v, v3<T>(scale3(g, dt)));
      v = add3(v, v3<T>(scale3(F, invM * dt)));
    }
    V3<T> wpred = w0;
    if (invM > 0.0) {
      if (invI.x > 0.0 || invI.y > 0.0 || invI.z > 0.0) {
        V3<T> wb = invRotate(q, w0);
        const d3 Ib{(invI.x > 1e-9) ? 1.0 / invI.x : 0.0, (invI.y > 1e-9) ? 1.0 / invI.y : 0.0,
                    (invI.z > 1e-9) ? 1.0 / invI.z : 0.0};
        const V3<T> Lb{T(Ib.x) * wb.x, T(Ib.y) * wb.y, T(Ib.z) * wb.z};
        const V3<T> wxL = cross3(wb, Lb);
        const V3<T> alpha{T(-invI.x) * wxL.x, T(-invI.y) * wxL.y, T(-invI.z) * wxL.z};
        wb = add3(wb, scale3(alpha, dt));
        wpred = rotate(q, wb);
      }
    }
    if (invM > 0.0) x = add3(x, scale3(v, dt));
    st3(out, v);
    st3(out + 3, wpred);
    st3(out + 6, x);
  }
};

// sphere-vs-sphere contact (npPair, src/narrowphase.hpp:179-268, a sphere has no shell):
// in xA xB qA qB -> out n rA rB dist.  The orientation enters through the normal's central
// difference, taken along body B's axes with a step that is not small against the separation.
struct ElPair {
  static constexpr int NI = 14, NO = 10;
  double sA, sB, p0;
  template <class T>
  PD_HD static T sdf(V3<T> p, double r) {
    return len3(p) - T(r);
  }
  template <class T>
  PD_HD void operator()(const T (&in)[NI], T (&out)[NO]) const {
    const V3<T> posA = ld3(in), posB = ld3(in + 3);
    const Q4<T> qa = ld4(in + 6), qb = ld4(in + 10);
    const double pointRadius = p0 * sA;
    const V3<T> pLocalA{T(0.0), T(0.0), T(0.0)};
    const V3<T> pWorld = add3(posA, rotate(qa, scale3(pLocalA, sA)));
    const V3<T> pLocalB = invRotate(qb, sub3(pWorld, posB));
    const V3<T> pCanB = scale3(pLocalB, 1.0 / sB);
    const T dist = sdf(pCanB, p0) * T(sB);
    const T effDist = dist - T(pointRadius);
    const double eps = 1e-4;
    V3<T> nLoc{sdf(V3<T>{pCanB.x + T(eps), pCanB.y, pCanB.z}, p0) -
                   sdf(V3<T>{pCanB.x - T(eps), pCanB.y, pCanB.z}, p0),
               sdf(V3<T>{pCanB.x, pCanB.y + T(eps), pCanB.z}, p0) -
                   sdf(V3<T>{pCanB.x, pCanB.y - T(eps), pCanB.z}, p0),
               sdf(V3<T>{pCanB.x, pCanB.y, pCanB.z + T(eps)}, p0) -
                   sdf(V3<T>{pCanB.x, pCanB.y, pCanB.z - T(eps)}, p0)};
    const T len = len3(nLoc);
    nLoc = (val(len) > 1e-9) ? scale3(nLoc, T(1.0) / len) : V3<T>{T(0.0), T(1.0), T(0.0)};
    const V3<T> nWorld = rotate(qb, nLoc);
    const V3<T> pSurfA = sub3(pWorld, scale3(nWorld, pointRadius));
    const V3<T> rA = sub3(pSurfA, posA);
    const V3<T> rB = sub3(sub3(pSurfA, scale3(nWorld, effDist)), posB);
    st3(out, nWorld);
    st3(out + 3, rA);
    st3(out + 6, rB);
    out[9] = effDist;
  }
};

// sphere-vs-plane contact (npPlanes, src/narrowphase.hpp:371-449): in xA -> out n rA rB dist
struct ElPlane {
  static constexpr int NI = 3, NO = 10;
  d3 pp, pn;
  double radius;
  template <class T>
  PD_HD void operator()(const T (&in)[NI], T (&out)[NO]) const {
    const V3<T> posA = ld3(in);
    const T dist = dot3(sub3(posA, v3<T>(pp)), v3<T>(pn)) - T(radius);
    st3(out, v3<T>(pn));
    st3(out + 3, v3<T>(scale3(pn, -radius)));
    st3(out + 6, v3<T>(pp));
    out[9] = dist;
  }
};

// transformContact, src/contact_preprocessing.hpp:339-374: in n rA rB dist -> the contact's
// terms of the manifold sums (n, tauA, tauB, rA, rB)
struct ElTerm {
  static constexpr int NI = 10, NO = 15;
  bool flip;
  template <class T>
  PD_HD void operator()(const T (&in)[NI], T (&out)[NO]) const {
    const T z(0.0);
    V3<T> na{z, z, z}, tca{z, z, z}, tcb{z, z, z}, rcan{z, z, z}, roth{z, z, z};
    const T w = in[9];
    if (!(val(w) > 0.0)) {
      const V3<T> nvec = ld3(in);
      na = flip ? nvec : neg3(nvec);
      const V3<T> shift{nvec.x * w * T(0.5), nvec.y * w * T(0.5), nvec.z * w * T(0.5)};
      const V3<T> rAmid = sub3(ld3(in + 3), shift);
      const V3<T> rBmid = add3(ld3(in + 6), shift);
      rcan = flip ? rBmid : rAmid;
      roth = flip ? rAmid : rBmid;
      tca = cross3(rcan, na);
      tcb = cross3(roth, neg3(na));
    }
    st3(out, na);
    st3(out + 3, tca);
    st3(out + 6, tcb);
    st3(out + 9, rcan);
    st3(out + 12, roth);
  }
};

// constants of one velocity row's bodies
struct RowBodies {
  double invMassA, invMassB;
  d3 invIA, invIB;
  q4 qA, qB;  // qB = identity for a boundary row
  d3 vWall;   // boundary rows: the body-B velocity (wall velocity sum / points)
  bool hasB;
};

// velocity row set-up (k_vrows: gatherWarmLambdaKokkos / computeVn0Kokkos,
// src/contact_preprocessing.hpp:113, src/solver_velocity.hpp:390-447):
// in manifold sums (ns ta tb ra rb) and vA wA vB wB -> out Nsum TauA TauB rA rB vn0 vt0
struct ElRow {
  static constexpr int NI = 27, NO = 19;
  double invN, growthRate;
  bool hasB;
  d3 vWall;
  template <class T>
  PD_HD void operator()(const T (&in)[NI], T (&out)[NO]) const {
    const V3<T> Nsum = ld3(in), TauA = ld3(in + 3), TauB = ld3(in + 6);
    const V3<T> rAavg = scale3(ld3(in + 9), invN), rBavg = scale3(ld3(in + 12), invN);
    const T lenN = tsqrt(dot3(Nsum, Nsum));
    const V3<T> diffCenters = hasB ? sub3(rAavg, rBavg) : rAavg;
    const V3<T> vA = ld3(in + 15), wA = ld3(in + 18);
    const T z(0.0);
    V3<T> vB{z, z, z}, wB{z, z, z};
    if (hasB) {
      vB = ld3(in + 21);
      wB = ld3(in + 24);
    } else {
      vB = v3<T>(vWall);
    }
    const V3<T> vGrowth = scale3(diffCenters, growthRate);
    T vn = dot3(vA, Nsum) + dot3(wA, TauA) + dot3(vB, neg3(Nsum)) + dot3(wB, TauB);
    vn = vn + dot3(vGrowth, Nsum);
    V3<T> vt{z, z, z};
    if (val(lenN) > 1e-9) {
      const V3<T> nhat = scale3(Nsum, T(1.0) / lenN);
      const V3<T> wBv = hasB ? wB : V3<T>{z, z, z};
      const V3<T> vrel = sub3(add3(vA, cross3(wA, rAavg)), add3(vB, cross3(wBv, rBavg)));
      vt = sub3(vrel, scale3(nhat, dot3(vrel, nhat)));
    }
    st3(out, Nsum);
    st3(out + 3, TauA);
    st3(out + 6, TauB);
    st3(out + 9, rAavg);
    st3(out + 12, rBavg);
    out[15] = vn;
    st3(out + 16, vt);
  }
};

// Velocity-row input layout shared by the warm-start and PGS elements:
//   [0..12)  vA wA vB wB          [12] lambda   [13..16) lambdaT
//   [16..31) Nsum TauA TauB rA rB [31] vn0      [32..35) vt0   [35] posImpulse
//   [36] mu  [37] restitution
// outputs: [0..12) bodies  [12] lambda  [13..16) lambdaT
constexpr int VR_NI = 38, VR_NO = 16;

// row constants that k_vrows evaluates once: sign, |N|, 1/|N|, wA_n, wB_n, growth term
template <class T>
struct RowDerived {
  T lenN, invLenN, wAn, wBn, growth;
  double sgn;
};
template <class T>
PD_HD RowDerived<T> rowDerived(const T (&in)[VR_NI], const RowBodies& B, double growthRate) {
  RowDerived<T> r;
  const V3<T> Nsum = ld3(in + 16), TauA = ld3(in + 19), TauB = ld3(in + 22);
  const V3<T> rAavg = ld3(in + 25), rBavg = ld3(in + 28);
  r.lenN = tsqrt(dot3(Nsum, Nsum));
  const V3<T> diffCenters = B.hasB ? sub3(rAavg, rBavg) : rAavg;
  const T alignment = dot3(Nsum, diffCenters);
  r.sgn = (val(alignment) > 0.0) ? 1.0 : -1.0;
  const T Nsq = dot3(Nsum, Nsum);
  r.wAn = Nsq * T(B.invMassA) + genInvMass(TauA, B.invIA, q4t<T>(B.qA));
  r.wBn = Nsq * T(B.invMassB) + genInvMass(TauB, B.invIB, q4t<T>(B.qB));
  r.growth = dot3(scale3(diffCenters, growthRate), Nsum);
  r.invLenN = val(r.lenN) > 0.0 ? T(1.0) / r.lenN : T(0.0);
  return r;
}

// warmStartApplyKokkos, src/solver_velocity.hpp:498-605 (one row)
struct ElWarm {
  static constexpr int NI = VR_NI, NO = VR_NO;
  RowBodies B;
  double growthRate;
  template <class T>
  PD_HD void operator()(const T (&in)[NI], T (&out)[NO]) const {
#pragma unroll
    for (int i = 0; i < NO; ++i) out[i] = in[i];
    const T p = in[12];
    V3<T> lt = ld3(in + 13);
    // a zero warm start changes no value, but its tangents do (see ElPgs)
    const bool hasT = nonzero3(lt) || IsDual<T>::value;
    if (val(p) == 0.0 && !hasT) return;
    const RowDerived<T> R = rowDerived(in, B, growthRate);
    const V3<T> Nsum = ld3(in + 16), TauA = ld3(in + 19), TauB = ld3(in + 22);
    const V3<T> rAavg = ld3(in + 25), rBavg = ld3(in + 28);
    const Q4<T> qA = q4t<T>(B.qA), qB = q4t<T>(B.qB);
    V3<T> vA = ld3(in), wA = ld3(in + 3), vB = ld3(in + 6), wB = ld3(in + 9);
    const T lambda = T(-R.sgn) * p;
    const V3<T> Jlin = scale3(Nsum, lambda);
    vA = add3(vA, scale3(Jlin, B.invMassA));
    wA = add3(wA, angDelta(qA, scale3(TauA, lambda), B.invIA));
    if (B.hasB) {
      vB = add3(vB, scale3(neg3(Jlin), B.invMassB));
      wB = add3(wB, angDelta(qB, scale3(TauB, lambda), B.invIB));
    }
    if (hasT) {
      const T lenN2 = tsqrt(dot3(Nsum, Nsum));
      if (val(lenN2) > 1e-9) {
        const V3<T> nhat = scale3(Nsum, T(1.0) / lenN2);
        lt = sub3(lt, scale3(nhat, dot3(lt, nhat)));
      }
      st3(out + 13, lt);
      vA = add3(vA, scale3(lt, B.invMassA));
      wA = add3(wA, angDelta(qA, cross3(rAavg, lt), B.invIA));
      if (B.hasB) {
        vB = add3(vB, scale3(neg3(lt), B.invMassB));
        wB = add3(wB, angDelta(qB, cross3(rBavg, scale3(lt, -1.0)), B.invIB));
      }
    }
    st3(out, vA);
    st3(out + 3, wA);
    if (B.hasB) {
      st3(out + 6, vB);
      st3(out + 9, wB);
    }
  }
};

// PGSManifoldSweep::solveOne, src/solver_velocity.hpp:641-858 (one row, normal + friction cone)
struct ElPgs {
  static constexpr int NI = VR_NI, NO = VR_NO;
  RowBodies B;
  double growthRate, restVelThreshold, restT;
  int sf;  // stabilization side: 0 none, 1 B held, 2 A held
  template <class T>
  PD_HD void operator()(const T (&in)[NI], T (&out)[NO]) const {
#pragma unroll
    for (int i = 0; i < NO; ++i) out[i] = in[i];
    const RowDerived<T> R = rowDerived(in, B, growthRate);
    const T lenN = R.lenN;
    if (val(lenN) < 1e-9) return;
    const V3<T> Nsum = ld3(in + 16), TauA = ld3(in + 19), TauB = ld3(in + 22);
    const V3<T> rAavg = ld3(in + 25), rBavg = ld3(in + 28);
    const Q4<T> qA = q4t<T>(B.qA), qB = q4t<T>(B.qB);
    V3<T> vA = ld3(in), wA = ld3(in + 3);
    const T z(0.0);
    V3<T> vB{z, z, z}, wB{z, z, z};
    if (B.hasB) {
      vB = ld3(in + 6);
      wB = ld3(in + 9);
    } else {
      vB = v3<T>(B.vWall);
    }
    T restitution = in[37];
    T vn = dot3(vA, Nsum) + dot3(wA, TauA) + dot3(vB, neg3(Nsum)) + dot3(wB, TauB);
    vn = vn + R.growth;
    const double sgn = R.sgn;
    T wA_n = R.wAn, wB_n = R.wBn;
    bool applyA = true, applyB = true;
    if (sf == 1) {
      wB_n = T(0.0);
      applyB = false;
      restitution = T(0.0);
    } else if (sf == 2) {
      wA_n = T(0.0);
      applyA = false;
      restitution = T(0.0);
    }
    const T wTotal = wA_n + wB_n;
    if (!(val(wTotal) > 0.0)) return;
    const T vn0 = in[31];
    const T v0til = T(sgn) * vn0;
    if (val(tabs(vn0)) < restVelThreshold * val(lenN)) restitution = T(0.0);
    const T target = (val(v0til) > 0.0) ? -restitution * v0til : T(0.0);
    const T vtil = T(sgn) * vn;
    const T dp = (vtil - target) / wTotal;
    const T pOld = in[12];
    T pNew = pOld + dp;
    const bool clamped = val(pNew) < 0.0;
    if (clamped) pNew = T(0.0);
    const T dApplied = pNew - pOld;
    // The forward kernel skips an update that changes nothing (a converged row); its derivative
    // is not zero, so the tangent path applies it (adding exact zeros to the values).
    const bool applied = val(dApplied) != 0.0 || (IsDual<T>::value && !clamped);
    const T lamNow = applied ? pNew : pOld;
    if (applied) {
      out[12] = pNew;
      const T lambda = T(-sgn) * dApplied;
      const V3<T> Jlin = scale3(Nsum, lambda);
      if (applyA) {
        vA = add3(vA, scale3(Jlin, B.invMassA));
        wA = add3(wA, angDelta(qA, scale3(TauA, lambda), B.invIA));
      }
      if (B.hasB && applyB) {
        vB = add3(vB, scale3(neg3(Jlin), B.invMassB));
        wB = add3(wB, angDelta(qB, scale3(TauB, lambda), B.invIB));
      }
    }
    // friction cone, src/solver_velocity.hpp:760-855
    const T mu = in[36];
    const V3<T> ltOld = ld3(in + 13);
    const bool haveOld = nonzero3(ltOld);
    if (val(mu) > 0.0 || haveOld) {
      const V3<T> nhat = scale3(Nsum, R.invLenN);
      const V3<T> wB2 = B.hasB ? wB : V3<T>{z, z, z};
      const V3<T> vB2 = B.hasB ? vB : v3<T>(B.vWall);
      const V3<T> vrel = sub3(add3(vA, cross3(wA, rAavg)), add3(vB2, cross3(wB2, rBavg)));
      const V3<T> vt = sub3(vrel, scale3(nhat, dot3(vrel, nhat)));
      double beta = restT;
      if (beta != 0.0 && (sf != 0 || val(tabs(vn0)) < restVelThreshold * val(lenN))) beta = 0.0;
      V3<T> vtErr = vt;
      const V3<T> vt0 = ld3(in + 32);
      if (beta != 0.0) {
        const V3<T> v0 = sub3(vt0, scale3(nhat, dot3(vt0, nhat)));
        vtErr = add3(vt, scale3(v0, beta));
      }
      const T vtLen = tsqrt(dot3(vtErr, vtErr));
      V3<T> ltNew = ltOld;
      bool stickTangent = false;
      if (val(vtLen) > 1e-9) {
        const V3<T> that = scale3(vtErr, T(1.0) / vtLen);
        const T wAt = (sf == 2) ? T(0.0)
                                : T(B.invMassA) + genInvMass(cross3(rAavg, that), B.invIA, qA);
        const T wBt = (sf == 1 || !B.hasB)
                          ? T(0.0)
                          : T(B.invMassB) + genInvMass(cross3(rBavg, that), B.invIB, qB);
        const T wT = wAt + wBt;
        if (val(wT) > 1e-9) ltNew = add3(ltOld, scale3(that, -vtLen / wT));
      } else if (IsDual<T>::value) {
        // A sticking contact (|vt| below the 1e-9 guard) keeps its impulse, but any perturbation
        // above the guard is cancelled by the update lt - vtErr / wT.  For spheres the lever
        // arms are normal to the contact plane, so wT is the same for every tangent direction
        // and that update is linear: its tangents are taken here, the value is left unchanged.
        const d3 nh{val(nhat.x), val(nhat.y), val(nhat.z)};
        const d3 ax = fabs(nh.x) < 0.9 ? d3{1.0, 0.0, 0.0} : d3{0.0, 1.0, 0.0};
        d3 th = cross3(nh, ax);
        th = scale3(th, 1.0 / len3(th));
        const d3 ra{val(rAavg.x), val(rAavg.y), val(rAavg.z)};
        const d3 rb{val(rBavg.x), val(rBavg.y), val(rBavg.z)};
        const double wAt = (sf == 2) ? 0.0 : B.invMassA + pd::genInvMass(pd::cross3(ra, th), B.invIA, B.qA);
        const double wBt = (sf == 1 || !B.hasB)
                               ? 0.0
                               : B.invMassB + pd::genInvMass(pd::cross3(rb, th), B.invIB, B.qB);
        const double wT = wAt + wBt;
        if (wT > 1e-9) {
          const V3<T> moved = sub3(ltOld, scale3(vtErr, 1.0 / wT));
          ltNew = moved;
          setVal(ltNew.x, val(ltOld.x));
          setVal(ltNew.y, val(ltOld.y));
          setVal(ltNew.z, val(ltOld.z));
          stickTangent = true;
        }
      }
      ltNew = sub3(ltNew, scale3(nhat, dot3(ltNew, nhat)));
      T nTot = tmax(lamNow, 0.0) * lenN;
      if (val(tabs(vn0)) < restVelThreshold * val(lenN)) nTot = nTot + tmax(in[35], 0.0);
      const T bound = mu * nTot;
      const T ltLen = tsqrt(dot3(ltNew, ltNew));
      if (val(ltLen) > val(bound))
        ltNew = (val(bound) > 0.0) ? scale3(ltNew, bound / ltLen) : V3<T>{z, z, z};
      const V3<T> dApp = sub3(ltNew, ltOld);
      if (nonzero3(dApp) || stickTangent || IsDual<T>::value) {
        st3(out + 13, ltNew);
        if (applyA) {
          vA = add3(vA, scale3(dApp, B.invMassA));
          wA = add3(wA, angDelta(qA, cross3(rAavg, dApp), B.invIA));
        }
        if (B.hasB && applyB) {
          vB = add3(vB, scale3(neg3(dApp), B.invMassB));
          wB = add3(wB, angDelta(qB, cross3(rBavg, scale3(dApp, -1.0)), B.invIB));
        }
      }
    }
    st3(out, vA);
    st3(out + 3, wA);
    if (B.hasB) {
      st3(out + 6, vB);
      st3(out + 9, wB);
    }
  }
  // Reverse of operator() for T = double: xb = J^T yb with J the Jacobian the forward-mode
  // evaluation defines (same branches, the same always-applied updates and the same sticking
  // tangent), assembled in one backward pass instead of NI forward-mode columns.  M_A, M_B are
  // the bodies' world inverse inertia (angDelta is M J, genInvMass tau^T M tau).
  PD_HD void vjpRev(const double (&x)[NI], const double (&yb)[NO], double (&xb)[NI]) const {
    using pd::add3;
    using pd::cross3;
    using pd::dot3;
    using pd::scale3;
    using pd::sub3;
    for (int i = 0; i < NI; ++i) xb[i] = 0.0;
    for (int i = 0; i < NO; ++i) xb[i] = yb[i];  // the identity part (unchanged outputs)
    const d3 N{x[16], x[17], x[18]}, TA{x[19], x[20], x[21]}, TB{x[22], x[23], x[24]};
    const d3 rA{x[25], x[26], x[27]}, rB{x[28], x[29], x[30]};
    const double Nsq = dot3(N, N);
    const double lenN = sqrt(Nsq);
    if (lenN < 1e-9) return;
    const double invLenN = lenN > 0.0 ? 1.0 / lenN : 0.0;
    const d3 dc = B.hasB ? sub3(rA, rB) : rA;
    const double sgn = dot3(N, dc) > 0.0 ? 1.0 : -1.0;
    const d3 vA{x[0], x[1], x[2]}, wA{x[3], x[4], x[5]};
    const d3 vB = B.hasB ? d3{x[6], x[7], x[8]} : B.vWall;
    const d3 wB = B.hasB ? d3{x[9], x[10], x[11]} : d3{0.0, 0.0, 0.0};
    const double wAn = Nsq * B.invMassA + pd::genInvMass(TA, B.invIA, B.qA);
    const double wBn = Nsq * B.invMassB + pd::genInvMass(TB, B.invIB, B.qB);
    const bool applyA = sf != 2, applyB = sf != 1;
    const double wTotal = (applyA ? wAn : 0.0) + (applyB ? wBn : 0.0);
    const double growth = dot3(scale3(dc, growthRate), N);
    const double vn = dot3(vA, N) + dot3(wA, TA) + dot3(vB, pd::neg3(N)) + dot3(wB, TB) + growth;
    if (!(wTotal > 0.0)) return;
    const double vn0 = x[31];
    const bool slowN = fabs(vn0) < restVelThreshold * lenN;
    const bool useE = sf == 0 && !slowN;
    const double e = x[37];
    const double v0til = sgn * vn0;
    const double target = v0til > 0.0 ? -(useE ? e : 0.0) * v0til : 0.0;
    const double dp = (sgn * vn - target) / wTotal;
    const double pOld = x[12];
    const double pNew0 = pOld + dp;
    const bool clamped = pNew0 < 0.0;
    const double pNew = clamped ? 0.0 : pNew0;
    const double dApplied = pNew - pOld;
    const bool applied = dApplied != 0.0 || !clamped;
    const double lamNow = applied ? pNew : pOld;
    const double lambda = -sgn * dApplied;
    d3 vA1 = vA, wA1 = wA, vB1 = vB, wB1 = wB;
    if (applied) {
      const d3 Jlin = scale3(N, lambda);
      if (applyA) {
        vA1 = add3(vA, scale3(Jlin, B.invMassA));
        wA1 = add3(wA, pd::angDelta(B.qA, scale3(TA, lambda), B.invIA));
      }
      if (B.hasB && applyB) {
        vB1 = add3(vB, scale3(pd::neg3(Jlin), B.invMassB));
        wB1 = add3(wB, pd::angDelta(B.qB, scale3(TB, lambda), B.invIB));
      }
    }
    // adjoints of the stage values
    d3 gvA1{yb[0], yb[1], yb[2]}, gwA1{yb[3], yb[4], yb[5]};
    d3 gvB1{0, 0, 0}, gwB1{0, 0, 0};
    if (B.hasB) {
      gvB1 = d3{yb[6], yb[7], yb[8]};
      gwB1 = d3{yb[9], yb[10], yb[11]};
    }
    d3 gN{0, 0, 0}, gTA{0, 0, 0}, gTB{0, 0, 0}, grA{0, 0, 0}, grB{0, 0, 0};
    d3 gltOld{yb[13], yb[14], yb[15]}, gvt0{0, 0, 0};
    double glamNow = 0.0, glenN = 0.0, ginvLenN = 0.0, gmu = 0.0, gpimp = 0.0, gvn0 = 0.0, ge = 0.0;
    const double mu = x[36];
    const d3 ltOld{x[13], x[14], x[15]};
    const bool haveOld = ltOld.x != 0.0 || ltOld.y != 0.0 || ltOld.z != 0.0;
    if (mu > 0.0 || haveOld) {
      // ---- forward of the cone part
      const d3 nhat = scale3(N, invLenN);
      const d3 vrel = sub3(add3(vA1, cross3(wA1, rA)), add3(vB1, cross3(wB1, rB)));
      const double vrn = dot3(vrel, nhat);
      const d3 vt = sub3(vrel, scale3(nhat, vrn));
      const bool betaOn = restT != 0.0 && sf == 0 && !slowN;
      const d3 vt0{x[32], x[33], x[34]};
      const double v0n = dot3(vt0, nhat);
      d3 vtErr = vt;
      if (betaOn) vtErr = add3(vt, scale3(sub3(vt0, scale3(nhat, v0n)), restT));
      const double vtLen = sqrt(dot3(vtErr, vtErr));
      // which update moved the impulse: 1 the sliding update, 2 the sticking tangent, 0 none
      int upd = 0;
      d3 that{0, 0, 0}, cA{0, 0, 0}, cB{0, 0, 0};
      double wT = 0.0, c = 0.0, invLen = 0.0, wTc = 0.0;
      d3 ltN0 = ltOld;  // value
      if (vtLen > 1e-9) {
        invLen = 1.0 / vtLen;
        that = scale3(vtErr, invLen);
        cA = cross3(rA, that);
        cB = cross3(rB, that);
        const double wAt = sf == 2 ? 0.0 : B.invMassA + pd::genInvMass(cA, B.invIA, B.qA);
        const double wBt = (sf == 1 || !B.hasB) ? 0.0 : B.invMassB + pd::genInvMass(cB, B.invIB, B.qB);
        wT = wAt + wBt;
        if (wT > 1e-9) {
          c = -vtLen / wT;
          ltN0 = add3(ltOld, scale3(that, c));
          upd = 1;
        }
      } else {
        const d3 ax = fabs(nhat.x) < 0.9 ? d3{1.0, 0.0, 0.0} : d3{0.0, 1.0, 0.0};
        d3 th = cross3(nhat, ax);
        th = scale3(th, 1.0 / sqrt(dot3(th, th)));
        const double wAt = sf == 2 ? 0.0 : B.invMassA + pd::genInvMass(cross3(rA, th), B.invIA, B.qA);
        const double wBt =
            (sf == 1 || !B.hasB) ? 0.0 : B.invMassB + pd::genInvMass(cross3(rB, th), B.invIB, B.qB);
        wTc = wAt + wBt;
        if (wTc > 1e-9) upd = 2;  // value unchanged, tangents of lt - vtErr / wT
      }
      const double l0n = dot3(ltN0, nhat);
      const d3 ltN1 = sub3(ltN0, scale3(nhat, l0n));
      const bool lamPos = lamNow >= 0.0;
      const double tm = lamPos ? lamNow : 0.0;
      double nTot = tm * lenN;
      const bool pimpOn = slowN && x[35] >= 0.0;
      if (slowN) nTot = nTot + (pimpOn ? x[35] : 0.0);
      const double bound = mu * nTot;
      const double ltLen = sqrt(dot3(ltN1, ltN1));
      // cone: 0 inside, 1 scaled onto it, 2 cleared (zero bound)
      const int cone = ltLen > bound ? (bound > 0.0 ? 1 : 2) : 0;
      const double s = cone == 1 ? bound / ltLen : 1.0;
      const d3 ltN2 = cone == 1 ? scale3(ltN1, s) : (cone == 2 ? d3{0, 0, 0} : ltN1);
      const d3 dApp = sub3(ltN2, ltOld);
      // ---- reverse of the cone part
      d3 gdApp{0, 0, 0};
      d3 gltN2{yb[13], yb[14], yb[15]};
      gltOld = d3{0, 0, 0};
      if (applyA) {
        gdApp = add3(gdApp, scale3(gvA1, B.invMassA));
        const d3 u = pd::angDelta(B.qA, gwA1, B.invIA);  // M_A is symmetric
        grA = add3(grA, cross3(dApp, u));
        gdApp = add3(gdApp, cross3(u, rA));
      }
      if (B.hasB && applyB) {
        gdApp = sub3(gdApp, scale3(gvB1, B.invMassB));
        const d3 u = pd::angDelta(B.qB, gwB1, B.invIB);
        const d3 ndApp = pd::neg3(dApp);
        grB = add3(grB, cross3(ndApp, u));
        gdApp = sub3(gdApp, cross3(u, rB));
      }
      gltN2 = add3(gltN2, gdApp);
      gltOld = sub3(gltOld, gdApp);
      d3 gltN1{0, 0, 0};
      double gbound = 0.0;
      if (cone == 1) {
        gltN1 = scale3(gltN2, s);
        const double gs = dot3(gltN2, ltN1);
        gbound += gs / ltLen;
        const double gltLen = -gs * s / ltLen;
        gltN1 = add3(gltN1, scale3(ltN1, gltLen / ltLen));
      } else if (cone == 0) {
        gltN1 = gltN2;
      }
      gmu += gbound * nTot;
      const double gnTot = gbound * mu;
      if (pimpOn) gpimp += gnTot;
      glenN += gnTot * tm;
      if (lamPos) glamNow += gnTot * lenN;
      // ltN1 = ltN0 - nhat (ltN0 . nhat)
      d3 gnhat{0, 0, 0};
      const double g1n = dot3(gltN1, nhat);
      d3 gltN0 = sub3(gltN1, scale3(nhat, g1n));
      gnhat = sub3(gnhat, add3(scale3(gltN1, l0n), scale3(ltN0, g1n)));
      d3 gvtErr{0, 0, 0};
      double gvtLen = 0.0;
      gltOld = add3(gltOld, gltN0);
      if (upd == 1) {
        // ltN0 = ltOld + that c, c = -vtLen / wT
        d3 gthat = scale3(gltN0, c);
        const double gc = dot3(gltN0, that);
        gvtLen += -gc / wT;
        const double gwT = -gc * c / wT;
        if (sf != 2) {
          const d3 gcA = scale3(pd::angDelta(B.qA, cA, B.invIA), 2.0 * gwT);
          grA = add3(grA, cross3(that, gcA));
          gthat = add3(gthat, cross3(gcA, rA));
        }
        if (!(sf == 1 || !B.hasB)) {
          const d3 gcB = scale3(pd::angDelta(B.qB, cB, B.invIB), 2.0 * gwT);
          grB = add3(grB, cross3(that, gcB));
          gthat = add3(gthat, cross3(gcB, rB));
        }
        // that = vtErr / vtLen (as vtErr * (1 / vtLen))
        gvtErr = add3(gvtErr, scale3(gthat, invLen));
        const double ginv = dot3(gthat, vtErr);
        gvtLen += -ginv * invLen * invLen;
      } else if (upd == 2) {
        gvtErr = sub3(gvtErr, scale3(gltN0, 1.0 / wTc));
      }
      if (vtLen > 0.0) gvtErr = add3(gvtErr, scale3(vtErr, gvtLen / vtLen));
      // vtErr = vt + (vt0 - nhat (vt0 . nhat)) beta
      d3 gvt = gvtErr;
      if (betaOn) {
        const d3 gv0p = scale3(gvtErr, restT);
        const double gpn = dot3(gv0p, nhat);
        gvt0 = add3(gvt0, sub3(gv0p, scale3(nhat, gpn)));
        gnhat = sub3(gnhat, add3(scale3(gv0p, v0n), scale3(vt0, gpn)));
      }
      // vt = vrel - nhat (vrel . nhat)
      const double gtn = dot3(gvt, nhat);
      const d3 gvrel = sub3(gvt, scale3(nhat, gtn));
      gnhat = sub3(gnhat, add3(scale3(gvt, vrn), scale3(vrel, gtn)));
      // vrel = vA1 + wA1 x rA - vB1 - wB1 x rB
      gvA1 = add3(gvA1, gvrel);
      gwA1 = add3(gwA1, cross3(rA, gvrel));
      grA = add3(grA, cross3(gvrel, wA1));
      if (B.hasB) {
        gvB1 = sub3(gvB1, gvrel);
        const d3 ng = pd::neg3(gvrel);
        gwB1 = add3(gwB1, cross3(rB, ng));
        grB = add3(grB, cross3(ng, wB1));
      } else {
        grB = add3(grB, cross3(pd::neg3(gvrel), wB1));
      }
      // nhat = N invLenN
      gN = add3(gN, scale3(gnhat, invLenN));
      ginvLenN += dot3(gnhat, N);
    }
    // ---- reverse of the normal part
    double gpNew = 0.0, gpOld = 0.0;
    // out[12] and lamNow: pNew if applied, else pOld
    if (applied) gpNew += yb[12] + glamNow;
    else gpOld += yb[12] + glamNow;
    d3 gvA = gvA1, gwA = gwA1, gvB = gvB1, gwB = gwB1;
    if (applied) {
      d3 gJ{0, 0, 0};
      double glambda = 0.0;
      if (applyA) {
        gJ = add3(gJ, scale3(gvA1, B.invMassA));
        const d3 t = pd::angDelta(B.qA, gwA1, B.invIA);
        gTA = add3(gTA, scale3(t, lambda));
        glambda += dot3(t, TA);
      }
      if (B.hasB && applyB) {
        gJ = sub3(gJ, scale3(gvB1, B.invMassB));
        const d3 t = pd::angDelta(B.qB, gwB1, B.invIB);
        gTB = add3(gTB, scale3(t, lambda));
        glambda += dot3(t, TB);
      }
      gN = add3(gN, scale3(gJ, lambda));
      glambda += dot3(gJ, N);
      const double gdApplied = -sgn * glambda;
      gpNew += gdApplied;
      gpOld -= gdApplied;
    }
    double gvn = 0.0, gwTotal = 0.0;
    if (!clamped) {
      gpOld += gpNew;
      const double gdp = gpNew;
      gvn = sgn * gdp / wTotal;
      const double gtarget = -gdp / wTotal;
      gwTotal = -gdp * dp / wTotal;
      if (v0til > 0.0 && useE) {
        ge += -gtarget * v0til;
        gvn0 += sgn * (-e * gtarget);
      }
    }
    // wTotal
    double gwAn = applyA ? gwTotal : 0.0, gwBn = applyB ? gwTotal : 0.0;
    double gNsq = gwAn * B.invMassA + gwBn * B.invMassB;
    gTA = add3(gTA, scale3(pd::angDelta(B.qA, TA, B.invIA), 2.0 * gwAn));
    gTB = add3(gTB, scale3(pd::angDelta(B.qB, TB, B.invIB), 2.0 * gwBn));
    // vn = vA.N + wA.TA - vB.N + wB.TB + growth
    gvA = add3(gvA, scale3(N, gvn));
    gwA = add3(gwA, scale3(TA, gvn));
    gN = add3(gN, scale3(sub3(vA, vB), gvn));
    gTA = add3(gTA, scale3(wA, gvn));
    gTB = add3(gTB, scale3(wB, gvn));
    if (B.hasB) {
      gvB = sub3(gvB, scale3(N, gvn));
      gwB = add3(gwB, scale3(TB, gvn));
    }
    // growth = (dc growthRate) . N
    gN = add3(gN, scale3(dc, growthRate * gvn));
    const d3 gdc = scale3(N, growthRate * gvn);
    grA = add3(grA, gdc);
    if (B.hasB) grB = sub3(grB, gdc);
    // lenN, invLenN, Nsq
    glenN += -ginvLenN * invLenN * invLenN;
    gN = add3(gN, scale3(N, 2.0 * gNsq + glenN / lenN));
    // scatter
    xb[0] = gvA.x; xb[1] = gvA.y; xb[2] = gvA.z;
    xb[3] = gwA.x; xb[4] = gwA.y; xb[5] = gwA.z;
    if (B.hasB) {
      xb[6] = gvB.x; xb[7] = gvB.y; xb[8] = gvB.z;
      xb[9] = gwB.x; xb[10] = gwB.y; xb[11] = gwB.z;
    }
    xb[12] = gpOld;
    xb[13] = gltOld.x; xb[14] = gltOld.y; xb[15] = gltOld.z;
    xb[16] = gN.x; xb[17] = gN.y; xb[18] = gN.z;
    xb[19] = gTA.x; xb[20] = gTA.y; xb[21] = gTA.z;
    xb[22] = gTB.x; xb[23] = gTB.y; xb[24] = gTB.z;
    xb[25] = grA.x; xb[26] = grA.y; xb[27] = grA.z;
    xb[28] = grB.x; xb[29] = grB.y; xb[30] = grB.z;
    xb[31] = gvn0;
    xb[32] = gvt0.x; xb[33] = gvt0.y; xb[34] = gvt0.z;
    xb[35] = gpimp;
    xb[36] = gmu;
    xb[37] = ge;
  }
};

// applyVelocityAndPredictPositionKokkos, src/integration.hpp:112-148:
// in x v0 vf wf q -> out x~ q~
struct ElReint {
  static constexpr int NI = 16, NO = 7;
  double invM, dt;
  template <class T>
  PD_HD void operator()(const T (&in)[NI], T (&out)[NO]) const {
    V3<T> x = ld3(in);
    const V3<T> vStart = ld3(in + 3), vFinal = ld3(in + 6), omega = ld3(in + 9);
    const Q4<T> q = ld4(in + 12);
    if (invM > 0.0) x = add3(x, scale3(add3(vStart, vFinal), 0.5 * dt));
    Q4<T> qp = q;
    if (invM > 0.0) {
      const T h(0.5 * dt);
      qp.x = qp.x + h * (omega.x * q.w + omega.y * q.z - omega.z * q.y);
      qp.y = qp.y + h * (omega.y * q.w + omega.z * q.x - omega.x * q.z);
      qp.z = qp.z + h * (omega.z * q.w + omega.x * q.y - omega.y * q.x);
      qp.w = qp.w + h * (-omega.x * q.x - omega.y * q.y - omega.z * q.z);
      const T len = tsqrt(qp.x * qp.x + qp.y * qp.y + qp.z * qp.z + qp.w * qp.w);
      if (val(len) > 1e-9) {
        const T s = T(1.0) / len;
        qp = Q4<T>{qp.x * s, qp.y * s, qp.z * s, qp.w * s};
      }
    }
    st3(out, x);
    st4(out + 3, qp);
  }
};

// position row set-up (k_prows, src/solver_position.hpp:276-312):
// in n rA rB (contact) q~A qA q~B qB -> out n rA rB wTotal
struct ElProw {
  static constexpr int NI = 25, NO = 10;
  double invMassA, invMassB;
  d3 invIA, invIB;
  bool hasB;
  template <class T>
  PD_HD void operator()(const T (&in)[NI], T (&out)[NO]) const {
    const V3<T> n0 = ld3(in), ra0 = ld3(in + 3), rb0 = ld3(in + 6);
    const Q4<T> qA = ld4(in + 9), qA0 = ld4(in + 13);
    const Q4<T> qAdelta = quatMult(qA, quatInverse(qA0));
    const V3<T> rA = rotate(qAdelta, ra0);
    V3<T> rB = rb0, n = n0;
    if (hasB) {
      const Q4<T> qB = ld4(in + 17), qB0 = ld4(in + 21);
      const Q4<T> qBdelta = quatMult(qB, quatInverse(qB0));
      rB = rotate(qBdelta, rB);
      n = rotate(qBdelta, n);
    }
    const T wTotal = computeW(rA, n, invMassA, invIA) + computeW(rB, n, invMassB, invIB);
    st3(out, n);
    st3(out + 3, rA);
    st3(out + 6, rB);
    out[9] = wTotal;
  }
};

// solvePositionColoredGSKokkos, one contact (src/solver_position.hpp:259-326):
// in pA pB pLam n rA rB wTotal -> out pA pB pLam
struct ElPos {
  static constexpr int NI = 17, NO = 7;
  double invMassA, invMassB;
  bool hasB;
  template <class T>
  PD_HD void operator()(const T (&in)[NI], T (&out)[NO]) const {
#pragma unroll
    for (int i = 0; i < NO; ++i) out[i] = in[i];
    const V3<T> pA = ld3(in), pB = ld3(in + 3);
    const V3<T> n = ld3(in + 7), rA = ld3(in + 10), rB = ld3(in + 13);
    const T wTotal = in[16];
    const T C = hasB ? dot3(sub3(add3(pA, rA), add3(pB, rB)), n) : dot3(sub3(add3(pA, rA), rB), n);
    if (val(C) >= 0.0) return;
    if (val(wTotal) < 1e-6) return;
    const T dLambda = -C / wTotal;
    out[6] = in[6] + dLambda;
    out[0] = pA.x + n.x * dLambda * T(invMassA);
    out[1] = pA.y + n.y * dLambda * T(invMassA);
    out[2] = pA.z + n.z * dLambda * T(invMassA);
    if (hasB) {
      out[3] = pB.x + -n.x * dLambda * T(invMassB);
      out[4] = pB.y + -n.y * dLambda * T(invMassB);
      out[5] = pB.z + -n.z * dLambda * T(invMassB);
    }
  }
};

// tangents per forward-mode evaluation in the velocity-row products (38 inputs)
constexpr int VJP_K = 2;

// vector-Jacobian product of one element: xb = J^T yb, J assembled from K forward-mode columns
// per evaluation (the element's branches follow the primal input x)
template <int K, class F>
__device__ void vjp(const F& f, const double (&x)[F::NI], const double (&yb)[F::NO],
                    double (&xb)[F::NI]) {
#pragma unroll 1
  for (int j0 = 0; j0 < F::NI; j0 += K) {
    DN<K> in[F::NI];
#pragma unroll
    for (int i = 0; i < F::NI; ++i) {
      in[i].v = x[i];
#pragma unroll
      for (int t = 0; t < K; ++t) in[i].d[t] = (i == j0 + t) ? 1.0 : 0.0;
    }
    DN<K> out[F::NO];
    f(in, out);
#pragma unroll
    for (int t = 0; t < K; ++t) {
      double s = 0.0;
#pragma unroll
      for (int o = 0; o < F::NO; ++o) s += yb[o] * out[o].d[t];
#pragma unroll
      for (int i = 0; i < F::NI; ++i)
        if (i == j0 + t) xb[i] = s;
    }
  }
}

// the columns of part `part` of `nparts`: evaluations j0 = part*K, (part+nparts)*K, ...
template <int K>
PD_HD bool ownsColumn(int i, int part, int nparts) {
  return (i / K) % nparts == part;
}
template <int K, class F>
__device__ void vjpPart(const F& f, const double (&x)[F::NI], const double (&yb)[F::NO],
                        double (&xb)[F::NI], int part, int nparts) {
#pragma unroll 1
  for (int j0 = part * K; j0 < F::NI; j0 += nparts * K) {
    DN<K> in[F::NI];
#pragma unroll
    for (int i = 0; i < F::NI; ++i) {
      in[i].v = x[i];
#pragma unroll
      for (int t = 0; t < K; ++t) in[i].d[t] = (i == j0 + t) ? 1.0 : 0.0;
    }
    DN<K> out[F::NO];
    f(in, out);
#pragma unroll
    for (int t = 0; t < K; ++t) {
      double s = 0.0;
#pragma unroll
      for (int o = 0; o < F::NO; ++o) s += yb[o] * out[o].d[t];
#pragma unroll
      for (int i = 0; i < F::NI; ++i)
        if (i == j0 + t) xb[i] = s;
    }
  }
}

// ---------------------------------------------------------------- recorded structure
// One substep's discrete structure, captured after the forward substep (device pointers).
struct Topo {
  const int2* cAB;        // contact slots (b = -1: plane)
  const int* cPl;         // plane index of a plane contact
  const int2* mRange;     // manifold -> pair/plane contact range
  const int* vPerm;       // velocity row -> manifold
  const int2* vAB;        // velocity row -> real bodies
  const unsigned char* vFlag;  // row flags after the stage (side bits of the stabilization)
  const int* warmEntry;   // row -> entry of the previous ledger (or -1)
  const int* srcRowL;     // ledger entry -> velocity row whose impulse it stores (or -1)
  const int* pimpMan;     // ledger entry (first of its key) -> manifold writing its impulse
  const int* pPerm;       // position row -> contact
  const int* pRowOf;      // contact -> position row (or -1)
  int nSlots, nc, nm, nRows, nPRows;
};

// per-slot constants of the replay
struct SlotK {
  double4 iim;  // invI.xyz, invMass (ghost: zero inertia, the real's mass)
  double scale;
};

// replay buffers of one substep (T = double or DN<K>)
template <class T>
struct RB {
  T* X;           // [numReal][13] state at the substep start: x v w q
  T* Xn;          // [numReal][13] state at the end
  T* xs;          // [slot][3] committed position
  T* xt;          // [slot][3] predicted position
  T* v0;          // [slot][3] start velocity
  T* vh;          // [slot][6] predicted velocity and angular velocity
  T* q0;          // [slot][4] committed orientation
  T* vc;          // [real][6] velocities during the velocity solve
  T* xr;          // [slot][3] re-integrated position, then the position solve in place
  T* qr;          // [slot][4] re-integrated orientation
  T* cg;          // [contact][10] n rA rB dist
  T* ms;          // [manifold][15] sums
  T* rc;          // [row][20] Nsum TauA TauB rA rB vn0 vt0 posImpulse
  T* lam;         // [row][4] lambda, lambdaT
  T* pc;          // [prow][10] n rA rB wTotal
  T* pl;          // [prow] position lambda
  T* L;           // [entry][5] this substep's ledger: lambda, lambdaT, position impulse
  const T* Lp;    // [entry][5] previous ledger
  const T* par;   // [2] mu, restitution
  SlotK* sk;      // [slot]
  int* realOf;    // [slot]
  int* gcnt;      // [real + 1] ghost count, then (scanned) ghost offset
  int* goff;
  int* cMan;      // [contact] manifold
  double* vtape;  // [pass][row][16] velocity-row inputs (T = double, recording)
  double* ptape;  // [sweep][prow][7] position-row inputs
};

struct RK {  // replay constants
  Params P;
  double vRest;  // restitution velocity threshold 2 dt |g|
  const BodyC* bc;        // real bodies' masses and inertia (orientation from the replay)
  const double* scale;    // real scales
  const double4* extForce;
};

template <class T>
__device__ __forceinline__ void ldT(T* dst, const T* src, int n) {
  for (int i = 0; i < n; ++i) dst[i] = src[i];
}

// ---- slots: copy the start state, count ghosts (src/periodicity.hpp:25-38)
template <class T>
__global__ __launch_bounds__(TPB) void a_start(RB<T> B, RK K) {
  const int i = gtid();
  if (i >= K.P.numReal) return;
  const T* X = B.X + 13 * i;
  for (int c = 0; c < 3; ++c) {
    B.xs[3 * i + c] = X[c];
    B.v0[3 * i + c] = X[3 + c];
  }
  for (int c = 0; c < 4; ++c) B.q0[4 * i + c] = X[9 + c];
  int sx, sy, sz;
  ghostShifts(K.P, d3{val(X[0]), val(X[1]), val(X[2])}, sx, sy, sz);
  B.gcnt[i] = (sx ? 2 : 1) * (sy ? 2 : 1) * (sz ? 2 : 1) - 1;
  const BodyC b = K.bc[i];
  B.sk[i].iim = b.iim;
  B.sk[i].scale = K.scale[i];
  B.realOf[i] = i;
}

// ---- predict + ghosts (src/integration.hpp:44-93, src/periodicity.hpp:43-94)
template <class T>
__global__ __launch_bounds__(TPB) void a_predict(RB<T> B, RK K, int nSlots) {
  const int i = gtid();
  if (i >= K.P.numReal) return;
  const BodyC b = K.bc[i];
  ElPredict f;
  f.invM = b.iim.w;
  f.dt = K.P.dt;
  f.invI = xyz(b.iim);
  f.g = K.P.g;
  f.F = xyz(K.extForce[i]);
  T in[13], out[9];
  ldT(in, B.X + 13 * i, 13);
  f(in, out);
  for (int c = 0; c < 6; ++c) {
    B.vh[6 * i + c] = out[c];
    B.vc[6 * i + c] = out[c];
  }
  for (int c = 0; c < 3; ++c) B.xt[3 * i + c] = out[6 + c];
  // ghosts of this real, in the reference's (a, b, c) loop order
  int sx, sy, sz;
  ghostShifts(K.P, d3{val(in[0]), val(in[1]), val(in[2])}, sx, sy, sz);
  if (!sx && !sy && !sz) return;
  const int ax[2] = {0, sx}, ay[2] = {0, sy}, az[2] = {0, sz};
  const int nx = sx == 0 ? 1 : 2, ny = sy == 0 ? 1 : 2, nz = sz == 0 ? 1 : 2;
  int slot = K.P.numReal + B.goff[i];
  for (int a = 0; a < nx; ++a)
    for (int bb = 0; bb < ny; ++bb)
      for (int c = 0; c < nz; ++c) {
        const int ix = ax[a], iy = ay[bb], iz = az[c];
        if (ix == 0 && iy == 0 && iz == 0) continue;
        if (slot >= nSlots) return;  // the recorded slot count bounds the ghosts
        const double sh[3] = {ix * K.P.dsize.x, iy * K.P.dsize.y, iz * K.P.dsize.z};
        for (int k = 0; k < 3; ++k) {
          B.xs[3 * slot + k] = in[k] + T(sh[k]);
          B.xt[3 * slot + k] = out[6 + k] + T(sh[k]);
          B.v0[3 * slot + k] = in[3 + k];
        }
        for (int k = 0; k < 6; ++k) B.vh[6 * slot + k] = out[k];
        for (int k = 0; k < 4; ++k) B.q0[4 * slot + k] = in[9 + k];
        B.sk[slot].iim = make_double4(0.0, 0.0, 0.0, b.iim.w);
        B.sk[slot].scale = K.scale[i];
        B.realOf[slot] = i;
        ++slot;
      }
}

__device__ __forceinline__ double planeRadius(const RK& K, double sc) {
  double baseR = K.P.shape.p0;
  if (baseR == 0.0) baseR = 1.0;
  return baseR * (sc * K.P.gs);
}

// ---- contacts
template <class T>
__global__ __launch_bounds__(TPB) void a_contacts(RB<T> B, RK K, Topo t) {
  const int c = gtid();
  if (c >= t.nc) return;
  B.cMan[c] = -1;
  const int2 ab = t.cAB[c];
  T out[10];
  if (ab.y >= 0) {
    ElPair f;
    f.sA = B.sk[ab.x].scale * K.P.gs;
    f.sB = B.sk[ab.y].scale * K.P.gs;
    f.p0 = K.P.shape.p0;
    T in[14];
    ldT(in, B.xt + 3 * ab.x, 3);
    ldT(in + 3, B.xt + 3 * ab.y, 3);
    ldT(in + 6, B.q0 + 4 * ab.x, 4);
    ldT(in + 10, B.q0 + 4 * ab.y, 4);
    f(in, out);
  } else {
    const int pi = t.cPl[c];
    ElPlane f;
    f.pp = xyz(K.P.planeP[pi]);
    f.pn = xyz(K.P.planeN[pi]);
    f.radius = planeRadius(K, B.sk[ab.x].scale);
    T in[3];
    ldT(in, B.xt + 3 * ab.x, 3);
    f(in, out);
  }
  for (int k = 0; k < 10; ++k) B.cg[10 * c + k] = out[k];
}

// ---- manifold sums, contact order within the manifold (src/contact_preprocessing.hpp:378-471)
template <class T>
__global__ __launch_bounds__(TPB) void a_manifolds(RB<T> B, Topo t) {
  const int m = gtid();
  if (m >= t.nm) return;
  const int2 rg = t.mRange[m];
  T s[15];
  for (int k = 0; k < 15; ++k) s[k] = T(0.0);
  for (int q = 0; q < rg.y; ++q) {
    const int c = rg.x + q;
    B.cMan[c] = m;
    const int2 cab = t.cAB[c];
    ElTerm f;
    f.flip = (cab.y >= 0 && cab.y < cab.x);
    T in[10], out[15];
    ldT(in, B.cg + 10 * c, 10);
    f(in, out);
    for (int k = 0; k < 15; ++k) s[k] = s[k] + out[k];
  }
  for (int k = 0; k < 15; ++k) B.ms[15 * m + k] = s[k];
}

__device__ __forceinline__ int manifoldPoints(const Topo& t, const double* cgDist, int m) {
  // active contacts (dist <= 0) of manifold m
  const int2 rg = t.mRange[m];
  int np = 0;
  for (int q = 0; q < rg.y; ++q) np += !(cgDist[10 * (rg.x + q)] > 0.0);
  return np;
}

template <class T>
__device__ __forceinline__ RowBodies rowBodies(const RB<T>& B, const RK& K, int2 r) {
  RowBodies o;
  const BodyC ca = K.bc[r.x];
  o.invMassA = ca.iim.w;
  o.invIA = xyz(ca.iim);
  o.qA = q4{val(B.q0[4 * r.x]), val(B.q0[4 * r.x + 1]), val(B.q0[4 * r.x + 2]),
            val(B.q0[4 * r.x + 3])};
  o.hasB = r.y >= 0;
  o.vWall = d3{0.0, 0.0, 0.0};
  if (o.hasB) {
    const BodyC cb = K.bc[r.y];
    o.invMassB = cb.iim.w;
    o.invIB = xyz(cb.iim);
    o.qB = q4{val(B.q0[4 * r.y]), val(B.q0[4 * r.y + 1]), val(B.q0[4 * r.y + 2]),
              val(B.q0[4 * r.y + 3])};
  } else {
    o.invMassB = 0.0;
    o.invIB = d3{0.0, 0.0, 0.0};
    o.qB = q4{0.0, 0.0, 0.0, 1.0};
  }
  return o;
}

// number of active contacts of a row's manifold from the primal contact distances
template <class T>
__device__ __forceinline__ int rowPoints(const RB<T>& B, const Topo& t, int m) {
  const int2 rg = t.mRange[m];
  int np = 0;
  for (int q = 0; q < rg.y; ++q) np += !(val(B.cg[10 * (rg.x + q) + 9]) > 0.0);
  return np;
}

// ---- velocity rows: row constants, warm start from the previous ledger
template <class T>
__global__ __launch_bounds__(TPB) void a_rows(RB<T> B, RK K, Topo t) {
  const int k = gtid();
  if (k >= t.nRows) return;
  const int m = t.vPerm[k];
  const int2 r = t.vAB[k];
  ElRow f;
  f.invN = 1.0 / (double)rowPoints(B, t, m);
  f.growthRate = K.P.growthRate;
  f.hasB = r.y >= 0;
  f.vWall = d3{0.0 * f.invN, 0.0 * f.invN, 0.0 * f.invN};
  T in[27], out[19];
  ldT(in, B.ms + 15 * m, 15);
  ldT(in + 15, B.vh + 6 * r.x, 6);
  if (r.y >= 0) ldT(in + 21, B.vh + 6 * r.y, 6);
  else
    for (int c = 21; c < 27; ++c) in[c] = T(0.0);
  f(in, out);
  T* rc = B.rc + 20 * k;
  for (int c = 0; c < 19; ++c) rc[c] = out[c];
  const int e = t.warmEntry[k];
  T* lam = B.lam + 4 * k;
  if (e >= 0) {
    const T* L = B.Lp + 5 * e;
    for (int c = 0; c < 4; ++c) lam[c] = L[c];
    rc[19] = L[4];
  } else {
    for (int c = 0; c < 4; ++c) lam[c] = T(0.0);
    rc[19] = T(0.0);
  }
}

// gather one row's 38 inputs
template <class T>
__device__ __forceinline__ void vrowIn(const RB<T>& B, int k, int2 r, T (&in)[VR_NI]) {
  ldT(in, B.vc + 6 * r.x, 6);
  if (r.y >= 0) ldT(in + 6, B.vc + 6 * r.y, 6);
  else
    for (int c = 6; c < 12; ++c) in[c] = T(0.0);
  ldT(in + 12, B.lam + 4 * k, 4);
  ldT(in + 16, B.rc + 20 * k, 20);
  in[36] = B.par[0];
  in[37] = B.par[1];
}

// one colour of one velocity pass: mode 0 warm start, 1 PGS, 2 PGS with the stabilization side
template <class T>
__global__ __launch_bounds__(TPB) void a_vpass(RB<T> B, RK K, Topo t, int mode, int r0, int r1, double* tape) {
  const int k = r0 + gtid();
  if (k >= r1) return;
  const int2 r = t.vAB[k];
  T in[VR_NI], out[VR_NO];
  vrowIn(B, k, r, in);
  if (tape)
    for (int c = 0; c < 16; ++c) tape[16 * (size_t)k + c] = val(in[c]);
  const RowBodies RBo = rowBodies(B, K, r);
  if (mode == 0) {
    ElWarm f;
    f.B = RBo;
    f.growthRate = K.P.growthRate;
    f(in, out);
  } else {
    ElPgs f;
    f.B = RBo;
    f.growthRate = K.P.growthRate;
    f.restVelThreshold = K.vRest;
    f.restT = K.P.restT;
    const unsigned char fl = t.vFlag[k];
    f.sf = mode == 2 ? ((fl & F_SIDE1) ? 1 : ((fl & F_SIDE2) ? 2 : 0)) : 0;
    f(in, out);
  }
  for (int c = 0; c < 6; ++c) B.vc[6 * r.x + c] = out[c];
  if (r.y >= 0)
    for (int c = 0; c < 6; ++c) B.vc[6 * r.y + c] = out[6 + c];
  for (int c = 0; c < 4; ++c) B.lam[4 * k + c] = out[12 + c];
}

// ---- ledger of this substep: impulses of the velocity rows (commitPairKeysLambdaKokkos)
template <class T>
__global__ __launch_bounds__(TPB) void a_ledger(RB<T> B, Topo t) {
  const int e = gtid();
  if (e >= t.nm) return;
  const int k = t.srcRowL[e];
  T* L = B.L + 5 * e;
  for (int c = 0; c < 4; ++c) L[c] = k >= 0 ? B.lam[4 * k + c] : T(0.0);
  L[4] = T(0.0);
}

// ---- re-integration (every slot; a ghost keeps its predicted velocity)
template <class T>
__global__ __launch_bounds__(TPB) void a_reint(RB<T> B, RK K, int nSlots) {
  const int i = gtid();
  if (i >= nSlots) return;
  ElReint f;
  f.invM = B.sk[i].iim.w;
  f.dt = K.P.dt;
  T in[16], out[7];
  ldT(in, B.xs + 3 * i, 3);
  ldT(in + 3, B.v0 + 3 * i, 3);
  if (i < K.P.numReal) ldT(in + 6, B.vc + 6 * i, 6);
  else ldT(in + 6, B.vh + 6 * i, 6);
  ldT(in + 12, B.q0 + 4 * i, 4);
  f(in, out);
  for (int c = 0; c < 3; ++c) B.xr[3 * i + c] = out[c];
  for (int c = 0; c < 4; ++c) B.qr[4 * i + c] = out[3 + c];
}

// ---- position rows
template <class T>
__global__ __launch_bounds__(TPB) void a_prows(RB<T> B, Topo t) {
  const int k = gtid();
  if (k >= t.nPRows) return;
  const int c = t.pPerm[k];
  const int2 ab = t.cAB[c];
  ElProw f;
  f.invMassA = B.sk[ab.x].iim.w;
  f.invIA = xyz(B.sk[ab.x].iim);
  f.hasB = ab.y >= 0;
  f.invMassB = f.hasB ? B.sk[ab.y].iim.w : 0.0;
  f.invIB = f.hasB ? xyz(B.sk[ab.y].iim) : d3{0.0, 0.0, 0.0};
  T in[25], out[10];
  ldT(in, B.cg + 10 * c, 9);
  ldT(in + 9, B.qr + 4 * ab.x, 4);
  ldT(in + 13, B.q0 + 4 * ab.x, 4);
  if (f.hasB) {
    ldT(in + 17, B.qr + 4 * ab.y, 4);
    ldT(in + 21, B.q0 + 4 * ab.y, 4);
  } else {
    for (int q = 17; q < 25; ++q) in[q] = T(0.0);
  }
  f(in, out);
  for (int q = 0; q < 10; ++q) B.pc[10 * k + q] = out[q];
  B.pl[k] = T(0.0);
}

template <class T>
__global__ __launch_bounds__(TPB) void a_ppass(RB<T> B, Topo t, int r0, int r1, double* tape) {
  const int k = r0 + gtid();
  if (k >= r1) return;
  const int c = t.pPerm[k];
  const int2 ab = t.cAB[c];
  ElPos f;
  f.hasB = ab.y >= 0;
  f.invMassA = B.sk[ab.x].iim.w;
  f.invMassB = f.hasB ? B.sk[ab.y].iim.w : 0.0;
  T in[17], out[7];
  ldT(in, B.xr + 3 * ab.x, 3);
  if (f.hasB) ldT(in + 3, B.xr + 3 * ab.y, 3);
  else
    for (int q = 3; q < 6; ++q) in[q] = T(0.0);
  in[6] = B.pl[k];
  ldT(in + 7, B.pc + 10 * k, 10);
  if (tape)
    for (int q = 0; q < 7; ++q) tape[7 * (size_t)k + q] = val(in[q]);
  f(in, out);
  for (int q = 0; q < 3; ++q) B.xr[3 * ab.x + q] = out[q];
  if (f.hasB)
    for (int q = 0; q < 3; ++q) B.xr[3 * ab.y + q] = out[3 + q];
  B.pl[k] = out[6];
}

// ---- position impulse of the ledger entries (commitPosImpulseKokkos,
// src/contact_preprocessing.hpp:501-538): the last manifold of a key writes its sum / dt at the
// key's first entry
template <class T>
__global__ __launch_bounds__(TPB) void a_pimp(RB<T> B, RK K, Topo t) {
  const int e = gtid();
  if (e >= t.nm) return;
  const int m = t.pimpMan[e];
  if (m < 0) return;
  const int2 rg = t.mRange[m];
  T s(0.0);
  for (int q = 0; q < rg.y; ++q) {
    const int row = t.pRowOf[rg.x + q];
    const T l = row >= 0 ? B.pl[row] : T(0.0);
    if (val(l) != 0.0) s = s + l / T(K.P.dt);
  }
  B.L[5 * e + 4] = s;
}

// ---- final commit: periodic wrap of the position, orientation and velocities persist
template <class T>
__global__ __launch_bounds__(TPB) void a_final(RB<T> B, RK K) {
  const int i = gtid();
  if (i >= K.P.numReal) return;
  T p[3] = {B.xr[3 * i], B.xr[3 * i + 1], B.xr[3 * i + 2]};
  const int per[3] = {K.P.px, K.P.py, K.P.pz};
  const double lo[3] = {K.P.dmin.x, K.P.dmin.y, K.P.dmin.z};
  const double hi[3] = {K.P.dmax.x, K.P.dmax.y, K.P.dmax.z};
  const double L[3] = {K.P.dsize.x, K.P.dsize.y, K.P.dsize.z};
  for (int c = 0; c < 3; ++c) {
    if (!per[c]) continue;
    if (val(p[c]) < lo[c]) p[c] = p[c] + T(L[c]);
    else if (val(p[c]) >= hi[c]) p[c] = p[c] - T(L[c]);
  }
  T* Xn = B.Xn + 13 * i;
  for (int c = 0; c < 3; ++c) Xn[c] = p[c];
  for (int c = 0; c < 6; ++c) Xn[3 + c] = B.vc[6 * i + c];
  for (int c = 0; c < 4; ++c) Xn[9 + c] = B.qr[4 * i + c];
}

// ================================================================ reverse pass (double)
struct GB {  // adjoint buffers, same layout as RB<double>
  double* X;   // [real][13] in: adjoint of the end state; out: adjoint of the start state
  double* xs;  // [slot][3]
  double* xt;  // [slot][3]
  double* v0;  // [slot][3]
  double* vh;  // [slot][6]
  double* q0;  // [slot][4]
  double* vc;  // [real][6]
  double* xr;  // [slot][3]
  double* qr;  // [slot][4]
  double* cg;  // [contact][10]
  double* ms;  // [manifold][15]
  double* rc;  // [row][22]  (20 row constants, mu, restitution)
  double* lam; // [row][4]
  double* pc;  // [prow][10]
  double* pl;  // [prow]
  double* L;   // [entry][5] adjoint of this substep's ledger (in)
  double* Lp;  // [entry][5] adjoint of the previous ledger (out, accumulated)
};

__device__ __forceinline__ void aadd(double* a, double v) {
  if (v != 0.0) atomicAdd(a, v);
}

// ledger adjoint -> the rows and position rows that produced it
__global__ __launch_bounds__(TPB) void b_ledger(RB<double> B, GB G, RK K, Topo t) {
  const int e = gtid();
  if (e >= t.nm) return;
  const double* gl = G.L + 5 * e;
  const int k = t.srcRowL[e];
  if (k >= 0)
    for (int c = 0; c < 4; ++c) aadd(G.lam + 4 * k + c, gl[c]);
  const int m = t.pimpMan[e];
  if (m >= 0 && gl[4] != 0.0) {
    const int2 rg = t.mRange[m];
    for (int q = 0; q < rg.y; ++q) {
      const int row = t.pRowOf[rg.x + q];
      if (row >= 0 && B.pl[row] != 0.0) aadd(G.pl + row, gl[4] / K.P.dt);
    }
  }
}

// final commit: the wrap is a shift
__global__ __launch_bounds__(TPB) void b_final(GB G, RK K) {
  const int i = gtid();
  if (i >= K.P.numReal) return;
  const double* gX = G.X + 13 * i;
  for (int c = 0; c < 3; ++c) G.xr[3 * i + c] += gX[c];
  for (int c = 0; c < 6; ++c) G.vc[6 * i + c] += gX[3 + c];
  for (int c = 0; c < 4; ++c) G.qr[4 * i + c] += gX[9 + c];
}

__global__ __launch_bounds__(TPB) void b_ppass(RB<double> B, GB G, Topo t, int r0, int r1, const double* tape) {
  const int k = r0 + gtid();
  if (k >= r1) return;
  const int c = t.pPerm[k];
  const int2 ab = t.cAB[c];
  ElPos f;
  f.hasB = ab.y >= 0;
  f.invMassA = B.sk[ab.x].iim.w;
  f.invMassB = f.hasB ? B.sk[ab.y].iim.w : 0.0;
  double x[17], yb[7], xb[17];
  for (int q = 0; q < 7; ++q) x[q] = tape[7 * (size_t)k + q];
  for (int q = 0; q < 10; ++q) x[7 + q] = B.pc[10 * k + q];
  for (int q = 0; q < 3; ++q) yb[q] = G.xr[3 * ab.x + q];
  for (int q = 0; q < 3; ++q) yb[3 + q] = f.hasB ? G.xr[3 * ab.y + q] : 0.0;
  yb[6] = G.pl[k];
  vjp<1>(f, x, yb, xb);
  for (int q = 0; q < 3; ++q) G.xr[3 * ab.x + q] = xb[q];
  if (f.hasB)
    for (int q = 0; q < 3; ++q) G.xr[3 * ab.y + q] = xb[3 + q];
  G.pl[k] = xb[6];
  for (int q = 0; q < 10; ++q) G.pc[10 * k + q] += xb[7 + q];
}

__global__ __launch_bounds__(TPB) void b_prows(RB<double> B, GB G, Topo t) {
  const int k = gtid();
  if (k >= t.nPRows) return;
  const int c = t.pPerm[k];
  const int2 ab = t.cAB[c];
  ElProw f;
  f.invMassA = B.sk[ab.x].iim.w;
  f.invIA = xyz(B.sk[ab.x].iim);
  f.hasB = ab.y >= 0;
  f.invMassB = f.hasB ? B.sk[ab.y].iim.w : 0.0;
  f.invIB = f.hasB ? xyz(B.sk[ab.y].iim) : d3{0.0, 0.0, 0.0};
  double x[25], yb[10], xb[25];
  for (int q = 0; q < 9; ++q) x[q] = B.cg[10 * c + q];
  for (int q = 0; q < 4; ++q) {
    x[9 + q] = B.qr[4 * ab.x + q];
    x[13 + q] = B.q0[4 * ab.x + q];
    x[17 + q] = f.hasB ? B.qr[4 * ab.y + q] : 0.0;
    x[21 + q] = f.hasB ? B.q0[4 * ab.y + q] : 0.0;
  }
  for (int q = 0; q < 10; ++q) yb[q] = G.pc[10 * k + q];
  vjp<1>(f, x, yb, xb);
  for (int q = 0; q < 9; ++q) G.cg[10 * c + q] += xb[q];
  for (int q = 0; q < 4; ++q) {
    aadd(G.qr + 4 * ab.x + q, xb[9 + q]);
    aadd(G.q0 + 4 * ab.x + q, xb[13 + q]);
    if (f.hasB) {
      aadd(G.qr + 4 * ab.y + q, xb[17 + q]);
      aadd(G.q0 + 4 * ab.y + q, xb[21 + q]);
    }
  }
}

__global__ __launch_bounds__(TPB) void b_reint(RB<double> B, GB G, RK K, int nSlots) {
  const int i = gtid();
  if (i >= nSlots) return;
  ElReint f;
  f.invM = B.sk[i].iim.w;
  f.dt = K.P.dt;
  const bool real = i < K.P.numReal;
  double x[16], yb[7], xb[16];
  for (int c = 0; c < 3; ++c) {
    x[c] = B.xs[3 * i + c];
    x[3 + c] = B.v0[3 * i + c];
  }
  for (int c = 0; c < 6; ++c) x[6 + c] = real ? B.vc[6 * i + c] : B.vh[6 * i + c];
  for (int c = 0; c < 4; ++c) x[12 + c] = B.q0[4 * i + c];
  for (int c = 0; c < 3; ++c) yb[c] = G.xr[3 * i + c];
  for (int c = 0; c < 4; ++c) yb[3 + c] = G.qr[4 * i + c];
  vjp<1>(f, x, yb, xb);
  for (int c = 0; c < 3; ++c) {
    G.xs[3 * i + c] += xb[c];
    G.v0[3 * i + c] += xb[3 + c];
  }
  for (int c = 0; c < 6; ++c) {
    if (real) G.vc[6 * i + c] += xb[6 + c];
    else G.vh[6 * i + c] += xb[6 + c];
  }
  for (int c = 0; c < 4; ++c) aadd(G.q0 + 4 * i + c, xb[12 + c]);
}

// threads per velocity row in the reverse sweeps: each evaluates a disjoint set of the row's
// forward-mode columns (a colour holds too few rows to fill the chip one row per thread)
constexpr int VP_SPLIT = 8;
__global__ __launch_bounds__(TPB) void b_vpass(RB<double> B, GB G, RK K, Topo t, int mode, int r0,
                                               int r1, const double* tape) {
  const int tid = gtid();
  const int k = r0 + tid / VP_SPLIT;
  const int part = tid % VP_SPLIT;
  const bool live = k < r1;
  int2 r = make_int2(0, -1);
  double x[VR_NI], yb[VR_NO], xb[VR_NI];
  if (live) {
    r = t.vAB[k];
    for (int c = 0; c < 16; ++c) x[c] = tape[16 * (size_t)k + c];
    for (int c = 0; c < 20; ++c) x[16 + c] = B.rc[20 * k + c];
    x[36] = B.par[0];
    x[37] = B.par[1];
    for (int c = 0; c < 6; ++c) yb[c] = G.vc[6 * r.x + c];
    for (int c = 0; c < 6; ++c) yb[6 + c] = r.y >= 0 ? G.vc[6 * r.y + c] : 0.0;
    for (int c = 0; c < 4; ++c) yb[12 + c] = G.lam[4 * k + c];
  }
  // every thread of the row has read the adjoints before any overwrites them
  __syncthreads();
  if (!live) return;
  const RowBodies RBo = rowBodies(B, K, r);
  if (mode == 0) {
    ElWarm f;
    f.B = RBo;
    f.growthRate = K.P.growthRate;
    vjpPart<VJP_K>(f, x, yb, xb, part, VP_SPLIT);
  } else {
    ElPgs f;
    f.B = RBo;
    f.growthRate = K.P.growthRate;
    f.restVelThreshold = K.vRest;
    f.restT = K.P.restT;
    const unsigned char fl = t.vFlag[k];
    f.sf = mode == 2 ? ((fl & F_SIDE1) ? 1 : ((fl & F_SIDE2) ? 2 : 0)) : 0;
    vjpPart<VJP_K>(f, x, yb, xb, part, VP_SPLIT);
  }
#pragma unroll
  for (int c = 0; c < VR_NI; ++c) {
    if (!ownsColumn<VJP_K>(c, part, VP_SPLIT)) continue;
    if (c < 6) G.vc[6 * r.x + c] = xb[c];
    else if (c < 12) {
      if (r.y >= 0) G.vc[6 * r.y + c - 6] = xb[c];
    } else if (c < 16) G.lam[4 * k + c - 12] = xb[c];
    else G.rc[22 * k + c - 16] += xb[c];
  }
}

// The reverse PGS sweeps (mode 1 main loop, 2 stabilization): one thread per row, the row's
// vector-Jacobian product in one backward pass (ElPgs::vjpRev) instead of forward-mode columns.
__global__ __launch_bounds__(TPB) void b_vpass_pgs(RB<double> B, GB G, RK K, Topo t, int mode, int r0,
                                                   int r1, const double* tape) {
  const int k = r0 + gtid();
  if (k >= r1) return;
  const int2 r = t.vAB[k];
  double x[VR_NI], yb[VR_NO], xb[VR_NI];
  for (int c = 0; c < 16; ++c) x[c] = tape[16 * (size_t)k + c];
  for (int c = 0; c < 20; ++c) x[16 + c] = B.rc[20 * k + c];
  x[36] = B.par[0];
  x[37] = B.par[1];
  for (int c = 0; c < 6; ++c) yb[c] = G.vc[6 * r.x + c];
  for (int c = 0; c < 6; ++c) yb[6 + c] = r.y >= 0 ? G.vc[6 * r.y + c] : 0.0;
  for (int c = 0; c < 4; ++c) yb[12 + c] = G.lam[4 * k + c];
  ElPgs f;
  f.B = rowBodies(B, K, r);
  f.growthRate = K.P.growthRate;
  f.restVelThreshold = K.vRest;
  f.restT = K.P.restT;
  const unsigned char fl = t.vFlag[k];
  f.sf = mode == 2 ? ((fl & F_SIDE1) ? 1 : ((fl & F_SIDE2) ? 2 : 0)) : 0;
  f.vjpRev(x, yb, xb);
  for (int c = 0; c < 6; ++c) G.vc[6 * r.x + c] = xb[c];
  if (r.y >= 0)
    for (int c = 0; c < 6; ++c) G.vc[6 * r.y + c] = xb[6 + c];
  for (int c = 0; c < 4; ++c) G.lam[4 * k + c] = xb[12 + c];
  for (int c = 16; c < VR_NI; ++c) G.rc[22 * k + c - 16] += xb[c];
}

// warm-start values -> the previous ledger; the velocity-stage start -> the predicted velocities
__global__ __launch_bounds__(TPB) void b_warm(GB G, RK K, Topo t) {
  const int k = gtid();
  if (k < t.nRows) {
    const int e = t.warmEntry[k];
    if (e >= 0) {
      for (int c = 0; c < 4; ++c) aadd(G.Lp + 5 * e + c, G.lam[4 * k + c]);
      aadd(G.Lp + 5 * e + 4, G.rc[22 * k + 19]);
    }
  }
  if (k < K.P.numReal)
    for (int c = 0; c < 6; ++c) aadd(G.vh + 6 * k + c, G.vc[6 * k + c]);
}

__global__ __launch_bounds__(TPB) void b_rows(RB<double> B, GB G, RK K, Topo t, double* gpar) {
  const int k = gtid();
  if (k >= t.nRows) return;
  const int m = t.vPerm[k];
  const int2 r = t.vAB[k];
  ElRow f;
  f.invN = 1.0 / (double)rowPoints(B, t, m);
  f.growthRate = K.P.growthRate;
  f.hasB = r.y >= 0;
  f.vWall = d3{0.0 * f.invN, 0.0 * f.invN, 0.0 * f.invN};
  double x[27], yb[19], xb[27];
  for (int c = 0; c < 15; ++c) x[c] = B.ms[15 * m + c];
  for (int c = 0; c < 6; ++c) x[15 + c] = B.vh[6 * r.x + c];
  for (int c = 0; c < 6; ++c) x[21 + c] = r.y >= 0 ? B.vh[6 * r.y + c] : 0.0;
  for (int c = 0; c < 19; ++c) yb[c] = G.rc[22 * k + c];
  vjp<1>(f, x, yb, xb);
  for (int c = 0; c < 15; ++c) G.ms[15 * m + c] += xb[c];
  for (int c = 0; c < 6; ++c) aadd(G.vh + 6 * r.x + c, xb[15 + c]);
  if (r.y >= 0)
    for (int c = 0; c < 6; ++c) aadd(G.vh + 6 * r.y + c, xb[21 + c]);
  aadd(gpar, G.rc[22 * k + 20]);
  aadd(gpar + 1, G.rc[22 * k + 21]);
}

__global__ __launch_bounds__(TPB) void b_manifolds(RB<double> B, GB G, Topo t) {
  const int c = gtid();
  if (c >= t.nc) return;
  const int m = B.cMan[c];
  if (m < 0) return;
  const int2 cab = t.cAB[c];
  ElTerm f;
  f.flip = (cab.y >= 0 && cab.y < cab.x);
  double x[10], yb[15], xb[10];
  for (int q = 0; q < 10; ++q) x[q] = B.cg[10 * c + q];
  for (int q = 0; q < 15; ++q) yb[q] = G.ms[15 * m + q];
  vjp<1>(f, x, yb, xb);
  for (int q = 0; q < 10; ++q) G.cg[10 * c + q] += xb[q];
}

__global__ __launch_bounds__(TPB) void b_contacts(RB<double> B, GB G, RK K, Topo t) {
  const int c = gtid();
  if (c >= t.nc) return;
  const int2 ab = t.cAB[c];
  double yb[10];
  for (int q = 0; q < 10; ++q) yb[q] = G.cg[10 * c + q];
  if (ab.y >= 0) {
    ElPair f;
    f.sA = B.sk[ab.x].scale * K.P.gs;
    f.sB = B.sk[ab.y].scale * K.P.gs;
    f.p0 = K.P.shape.p0;
    double x[14], xb[14];
    for (int q = 0; q < 3; ++q) {
      x[q] = B.xt[3 * ab.x + q];
      x[3 + q] = B.xt[3 * ab.y + q];
    }
    for (int q = 0; q < 4; ++q) {
      x[6 + q] = B.q0[4 * ab.x + q];
      x[10 + q] = B.q0[4 * ab.y + q];
    }
    vjp<1>(f, x, yb, xb);
    for (int q = 0; q < 3; ++q) {
      aadd(G.xt + 3 * ab.x + q, xb[q]);
      aadd(G.xt + 3 * ab.y + q, xb[3 + q]);
    }
    for (int q = 0; q < 4; ++q) {
      aadd(G.q0 + 4 * ab.x + q, xb[6 + q]);
      aadd(G.q0 + 4 * ab.y + q, xb[10 + q]);
    }
  } else {
    const int pi = t.cPl[c];
    ElPlane f;
    f.pp = xyz(K.P.planeP[pi]);
    f.pn = xyz(K.P.planeN[pi]);
    f.radius = planeRadius(K, B.sk[ab.x].scale);
    double x[3], xb[3];
    for (int q = 0; q < 3; ++q) x[q] = B.xt[3 * ab.x + q];
    vjp<1>(f, x, yb, xb);
    for (int q = 0; q < 3; ++q) aadd(G.xt + 3 * ab.x + q, xb[q]);
  }
}

// ghosts -> their reals (a shifted copy), then predict; writes the start-state adjoint
__global__ __launch_bounds__(TPB) void b_predict(RB<double> B, GB G, RK K, int nSlots) {
  const int i = gtid();
  if (i >= K.P.numReal) return;
  double gxs[3], gxt[3], gv0[3], gvh[6], gq0[4];
  for (int c = 0; c < 3; ++c) {
    gxs[c] = G.xs[3 * i + c];
    gxt[c] = G.xt[3 * i + c];
    gv0[c] = G.v0[3 * i + c];
  }
  for (int c = 0; c < 6; ++c) gvh[c] = G.vh[6 * i + c];
  for (int c = 0; c < 4; ++c) gq0[c] = G.q0[4 * i + c];
  const int g0 = K.P.numReal + B.goff[i], g1 = g0 + B.gcnt[i];
  for (int s = g0; s < g1 && s < nSlots; ++s) {
    for (int c = 0; c < 3; ++c) {
      gxs[c] += G.xs[3 * s + c];
      gxt[c] += G.xt[3 * s + c];
      gv0[c] += G.v0[3 * s + c];
    }
    for (int c = 0; c < 6; ++c) gvh[c] += G.vh[6 * s + c];
    for (int c = 0; c < 4; ++c) gq0[c] += G.q0[4 * s + c];
  }
  const BodyC b = K.bc[i];
  ElPredict f;
  f.invM = b.iim.w;
  f.dt = K.P.dt;
  f.invI = xyz(b.iim);
  f.g = K.P.g;
  f.F = xyz(K.extForce[i]);
  double x[13], yb[9], xb[13];
  for (int c = 0; c < 13; ++c) x[c] = B.X[13 * i + c];
  for (int c = 0; c < 6; ++c) yb[c] = gvh[c];
  for (int c = 0; c < 3; ++c) yb[6 + c] = gxt[c];
  vjp<1>(f, x, yb, xb);
  double* gX = G.X + 13 * i;
  for (int c = 0; c < 3; ++c) {
    gX[c] = xb[c] + gxs[c];
    gX[3 + c] = xb[3 + c] + gv0[c];
    gX[6 + c] = xb[6 + c];
  }
  for (int c = 0; c < 4; ++c) gX[9 + c] = xb[9 + c] + gq0[c];
}

// ================================================================ tape capture (product side)
// plane index of the plane contacts, and the recorded structure's copies
__global__ void k_tape_contacts(const Dev* __restrict__ Dp, Params P, int nc, int2* cAB, int* cPl,
                                int* pRowOf) {
  const Dev& D = *Dp;
  const int c = gtid();
  if (c >= nc) return;
  const int2 ab = D.cAB[c];
  cAB[c] = ab;
  pRowOf[c] = D.pRowOf[c];
  int pl = -1;
  if (ab.y < 0) {
    const double4 n = D.cN[c], rb = D.cRB[c];
    for (int k = 0; k < P.numPlanes && pl < 0; ++k)
      if (n.x == P.planeN[k].x && n.y == P.planeN[k].y && n.z == P.planeN[k].z &&
          rb.x == P.planeP[k].x && rb.y == P.planeP[k].y && rb.z == P.planeP[k].z)
        pl = k;
  }
  cPl[c] = pl;
}
__global__ void k_tape_manifolds(const Dev* __restrict__ Dp, int nm, int2* mRange,
                                 unsigned long long* lKey, int* srcRowL, int* pimpMan) {
  const Dev& D = *Dp;
  const int e = gtid();
  if (e >= nm) return;
  mRange[e] = D.mRange[e];
  const unsigned long long key = D.lKey[e];
  lKey[e] = key;
  const int m = D.lIdx[e];
  srcRowL[e] = D.vRowOf[m];
  int w = -1;
  if (key != ~0ull && (e == 0 || D.lKey[e - 1] != key)) {
    int hi = e + 1;
    while (hi < nm && D.lKey[hi] == key) ++hi;
    w = D.lIdx[hi - 1];
  }
  pimpMan[e] = w;
}
__global__ void k_tape_rows(const Dev* __restrict__ Dp, int nRows, int nPRows, int* vPerm,
                            int2* vAB, unsigned char* vFlag, int* warmEntry,
                            const unsigned long long* prevKeys, int nPrev, int* pPerm) {
  const Dev& D = *Dp;
  const int k = gtid();
  if (k < nPRows) pPerm[k] = D.pPerm[k];
  if (k >= nRows) return;
  vPerm[k] = D.vPerm[k];
  const int2 r = D.vAB[k];
  vAB[k] = r;
  vFlag[k] = D.vFlag[k];
  const unsigned long long key = realKey(r.x, r.y);
  const int lo = lowerBound(prevKeys, nPrev, key);
  warmEntry[k] = (lo < nPrev && prevKeys[lo] == key) ? lo : -1;
}
// substep-start checkpoint: body state (x v w q) and the ledger values
__global__ void k_tape_state(const Dev* __restrict__ Dp, int numReal, double* X) {
  const Dev& D = *Dp;
  const int i = gtid();
  if (i >= numReal) return;
  const double4 p = D.pos[i];
  const Body64 v = D.vw0[i];
  const double4 q = D.bc[i].q;
  double* o = X + 13 * i;
  o[0] = p.x; o[1] = p.y; o[2] = p.z;
  o[3] = v.v.x; o[4] = v.v.y; o[5] = v.v.z;
  o[6] = v.w.x; o[7] = v.w.y; o[8] = v.w.z;
  o[9] = q.x; o[10] = q.y; o[11] = q.z; o[12] = q.w;
}
__global__ void k_tape_ledger(const Dev* __restrict__ Dp, int n, double* L,
                              unsigned long long* keys) {
  const Dev& D = *Dp;
  const int e = gtid();
  if (e >= n) return;
  const double4 lt = D.lLamT[e];
  double* o = L + 5 * e;
  o[0] = D.lLam[e];
  o[1] = lt.x; o[2] = lt.y; o[3] = lt.z; o[4] = lt.w;
  if (keys) keys[e] = D.lKey[e];
}

// ---- conversions between the double state and the tangent state
__global__ void k_seed_dn(const double* x, const double* dx, DN<1>* out, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  DN<1> r(x[i]);
  r.d[0] = dx ? dx[i] : 0.0;
  out[i] = r;
}
__global__ void k_split_dn(const DN<1>* in, double* x, double* dx, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (x) x[i] = in[i].v;
  if (dx) dx[i] = in[i].d[0];
}
// sum of one array into gpar (adjoint of the parameters)
__global__ void k_fill_d(double* a, double v, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

}  // namespace ad
}  // namespace pd
