// Checks the hand-written reverse of the PGS row (ElPgs::vjpRev) against the Jacobian its
// forward-mode evaluation defines: random rows over every branch (wall rows, held sides,
// restitution on and off, clamped and unapplied impulses, the Coulomb cone inside, on and
// cleared, sticking contacts, tangential restitution), J^T yb from 38 one-tangent columns
// against vjpRev, relative to the largest component.  Prints the worst error; exit 1 above 1e-10.
#include <cstdio>
#include <random>
#include "pd_adjoint.hpp"
using namespace pd;
using namespace pd::ad;
int main() {
  std::mt19937_64 rng(7);
  std::uniform_real_distribution<double> U(-1.0, 1.0);
  double worst = 0.0;
  long cnt[8] = {0};
  int trials = 0, bad = 0;
  for (int t = 0; t < 60000; ++t) {
    ElPgs f;
    const int kind = t % 8;
    f.B.hasB = (t % 3) != 0;
    f.B.invMassA = 0.5 + 0.5 * (U(rng) + 1.0);
    f.B.invMassB = f.B.hasB ? 0.5 + 0.5 * (U(rng) + 1.0) : 0.0;
    const bool aniso = (t % 5) < 2;
    f.B.invIA = aniso ? d3{1.0 + U(rng) * 0.5, 1.0 + U(rng) * 0.5, 1.0 + U(rng) * 0.5} : d3{2.0, 2.0, 2.0};
    f.B.invIB = f.B.hasB ? (aniso ? d3{1.0 + U(rng) * 0.5, 1.5, 0.8} : d3{3.0, 3.0, 3.0}) : d3{0, 0, 0};
    q4 qa{U(rng), U(rng), U(rng), U(rng)};
    double ql = sqrt(qa.x * qa.x + qa.y * qa.y + qa.z * qa.z + qa.w * qa.w);
    f.B.qA = q4{qa.x / ql, qa.y / ql, qa.z / ql, qa.w / ql};
    q4 qb{U(rng), U(rng), U(rng), U(rng)};
    ql = sqrt(qb.x * qb.x + qb.y * qb.y + qb.z * qb.z + qb.w * qb.w);
    f.B.qB = f.B.hasB ? q4{qb.x / ql, qb.y / ql, qb.z / ql, qb.w / ql} : q4{0, 0, 0, 1};
    f.B.vWall = f.B.hasB ? d3{0, 0, 0} : d3{U(rng) * 0.1, U(rng) * 0.1, 0.0};
    f.growthRate = (t % 7 == 0) ? 0.3 : 0.0;
    f.restVelThreshold = 0.05;
    f.restT = (t % 4 == 0) ? 0.4 : 0.0;
    f.sf = (t % 6 == 0) ? 1 : ((t % 6 == 1) ? 2 : 0);
    double x[38], yb[16];
    for (int i = 0; i < 38; ++i) x[i] = U(rng);
    // normal and lever arms of a sphere-like contact
    d3 n{U(rng), U(rng), U(rng)};
    const double nl = sqrt(dot3(n, n));
    n = scale3(n, 1.0 / nl);
    const double np = 1.0 + (t % 4);
    const d3 N = scale3(n, np * (0.8 + 0.2 * (U(rng) + 1.0)));
    const d3 rA = scale3(n, -0.5), rB = scale3(n, 0.5);
    const d3 rAp = add3(rA, d3{U(rng) * 0.1, U(rng) * 0.1, U(rng) * 0.1});
    const d3 TA = cross3(rAp, N), TB = cross3(rB, neg3(N));
    x[16] = N.x; x[17] = N.y; x[18] = N.z;
    x[19] = TA.x; x[20] = TA.y; x[21] = TA.z;
    x[22] = TB.x; x[23] = TB.y; x[24] = TB.z;
    x[25] = rAp.x; x[26] = rAp.y; x[27] = rAp.z;
    x[28] = rB.x; x[29] = rB.y; x[30] = rB.z;
    x[36] = kind == 5 ? 0.0 : 0.5 * (U(rng) + 1.0);  // mu
    x[37] = 0.5 * (U(rng) + 1.0);                    // e
    x[12] = kind == 1 ? 0.0 : 0.3 * (U(rng) + 1.0);  // pOld
    if (kind == 2) x[12] = 5.0;                      // clamped with an old impulse
    x[31] = (kind == 3) ? 0.01 : U(rng);             // vn0 (slow or not)
    x[35] = U(rng);                                  // pimp
    if (kind == 4) {                                 // a resting, sticking row
      for (int i = 0; i < 12; ++i) x[i] = 0.0;
      x[12] = 0.0;
      x[31] = 0.0;
      x[13] = x[14] = x[15] = 0.0;
      f.growthRate = 0.0;
      f.B.vWall = d3{0, 0, 0};
    }
    if (kind == 6) {  // a large old tangent impulse: the cone scales it
      x[13] = 3.0 * U(rng); x[14] = 3.0 * U(rng); x[15] = 3.0 * U(rng);
    }
    if (kind == 5) {  // no friction at all: the cone branch is skipped
      x[13] = x[14] = x[15] = 0.0;
    }
    if (kind == 7) {  // zero old impulses
      x[13] = x[14] = x[15] = 0.0;
    }
    for (int o = 0; o < 16; ++o) yb[o] = U(rng);
    double ref[38];
    for (int j = 0; j < 38; ++j) {
      DN<1> in[38];
      for (int i = 0; i < 38; ++i) {
        in[i].v = x[i];
        in[i].d[0] = i == j ? 1.0 : 0.0;
      }
      DN<1> out[16];
      f(in, out);
      double sum = 0.0;
      for (int o = 0; o < 16; ++o) sum += yb[o] * out[o].d[0];
      ref[j] = sum;
    }
    double xb[38];
    f.vjpRev(x, yb, xb);
    {  // branch census (recomputed here)
      const d3 Nn{x[16], x[17], x[18]};
      const double lenN = sqrt(dot3(Nn, Nn));
      const bool slowN = fabs(x[31]) < f.restVelThreshold * lenN;
      double o[16];
      {
        double xi[38];
        for (int i = 0; i < 38; ++i) xi[i] = x[i];
        f(xi, o);
      }
      const bool clamped0 = o[12] == 0.0 && x[12] != 0.0;
      const bool notApplied = o[12] == x[12];
      const bool lt0 = x[13] == 0 && x[14] == 0 && x[15] == 0;
      const bool fric = x[36] > 0.0 || !lt0;
      const double ltl = sqrt(o[13] * o[13] + o[14] * o[14] + o[15] * o[15]);
      cnt[0] += clamped0; cnt[1] += notApplied; cnt[2] += fric; cnt[3] += slowN;
      cnt[4] += (fric && f.restT != 0.0 && f.sf == 0 && !slowN);
      cnt[5] += (fric && ltl == 0.0);
      cnt[6] += (kind == 4 && fric);
      cnt[7] += (kind == 6);
    }
    double scale = 1e-30, err = 0.0;
    int at = -1;
    for (int j = 0; j < 38; ++j) scale = fmax(scale, fabs(ref[j]));
    for (int j = 0; j < 38; ++j) {
      const double e = fabs(xb[j] - ref[j]);
      if (e > err) { err = e; at = j; }
    }
    const double rel = err / scale;
    ++trials;
    if (rel > worst) worst = rel;
    if (rel > 1e-10) {
      if (bad < 12)
        printf("trial %d kind %d hasB %d sf %d restT %.1f aniso %d: rel %.3e at input %d (ref %.6e got %.6e)\n",
               t, kind, (int)f.B.hasB, f.sf, f.restT, (int)aniso, rel, at, ref[at], xb[at]);
      ++bad;
    }
  }
  printf("%d trials, %d above 1e-10, worst %.3e\n", trials, bad, worst);
  printf("census: clamped %ld notApplied %ld friction %ld slowN %ld beta %ld ltZero %ld stick %ld bigOld %ld\n",
         cnt[0], cnt[1], cnt[2], cnt[3], cnt[4], cnt[5], cnt[6], cnt[7]);
  return bad ? 1 : 0;
}
