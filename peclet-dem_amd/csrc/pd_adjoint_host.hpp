// pd_adjoint_host.hpp — recording of substeps, replay, forward-mode (tangent) runs and the adjoint
// (reverse) pass over the recorded substeps.  Included by pd_sim.hip after the simulation type.
//
// Recording: while a tape is open, every demStep checkpoints the body state every `every`
// substeps (with the warm-start ledger it starts from) and, after the substep, copies its
// discrete structure (pd_adjoint.hpp, ad::Topo) into a chunked device arena.
// Adjoint: segments of `every` substeps are processed last to first.  A segment is replayed
// forward from its checkpoint storing each substep's start state; then each substep, last to
// first, is replayed once more recording its sweep inputs and is reversed.  Memory: the tape
// (about 60 B per contact and substep), the checkpoints, and one segment of states.
#pragma once

namespace {

struct AdjArena {
  std::vector<void*> chunks;
  size_t used = 0, cap = 0, total = 0;
  void* take(size_t bytes) {
    bytes = (bytes + 255) & ~(size_t)255;
    if (bytes == 0) bytes = 256;
    if (used + bytes > cap) {
      const size_t sz = std::max(bytes, (size_t)64 << 20);
      void* p = nullptr;
      HIPCHECK(hipMalloc(&p, sz));
      chunks.push_back(p);
      used = 0;
      cap = sz;
      total += sz;
    }
    void* r = static_cast<char*>(chunks.back()) + used;
    used += bytes;
    return r;
  }
  template <class T>
  T* takeN(size_t n) {
    return static_cast<T*>(take(n * sizeof(T)));
  }
  void release() {
    for (void* p : chunks) (void)hipFree(p);
    chunks.clear();
    used = cap = total = 0;
  }
};

struct TapeStep {
  ad::Topo t{};
  Params P{};
  double vRest = 0.0;
  std::vector<int> vs, ps;
  int nColV = 0, nColP = 0, sweepsV = 0, stabRan = 0, sweepsS = 0, sweepsP = 0;
  const unsigned long long* lKey = nullptr;  // this substep's ledger keys (t.nm of them)
};
struct TapeCkpt {
  double* X = nullptr;  // [numReal][13]
  double* L = nullptr;  // [nL][5] the ledger the substep starts from
  int nL = 0;
};

struct AdjTape {
  bool on = false;
  int every = 16;
  int numReal = 0;
  long generation = 0;     // number of the recording (one more at every pd_tape_begin)
  long paramsVersion = 0;  // the simulation's parameter version when the recording opened
  // masses / inertia (bc), scales and external forces captured when the recording opened
  const BodyC* bc = nullptr;
  const double* scale = nullptr;
  const double4* extForce = nullptr;
  AdjArena arena;
  std::vector<TapeStep> steps;
  std::vector<TapeCkpt> ckpts;
  const unsigned long long* key0 = nullptr;  // keys of the ledger before the first substep
  int nKey0 = 0;
  void clear() {
    arena.release();
    bc = nullptr;
    scale = nullptr;
    extForce = nullptr;
    steps.clear();
    ckpts.clear();
    key0 = nullptr;
    nKey0 = 0;
  }
};
void adjFree(AdjTape* t) {
  if (!t) return;
  t->clear();
  delete t;
}

}  // namespace

// what the recorded substep supports; throws with the reason otherwise
void pd_sim::tapeCheck(bool usePGS, bool legacyFriction) const {
  auto no = [](const char* why) {
    throw ArgError(std::string("adjoint tape: ") + why + " is not supported by the adjoint");
  };
  if (shapeType != SHAPE_SPHERE || !shellHost.empty()) no("a non-spherical particle shape");
  if (!wallsHost.empty()) no("an SDF wall");
  if (hasPairTable) no("a pair-material table");
  if (growthFactor > 0.0 && growthRate != 0.0) no("particle growth");
  if (growthRate != 0.0) no("a growth rate");
  if (thermoTau > 0.0) no("the thermostat");
  if (!usePGS) no("zero gravity or the Jacobi velocity path");
  if (legacyFriction) no("the count-averaged friction path");
  if (stabMode != 0 && stabMode != 1) no("this stabilization mode");
}

// a new recording: its generation number and the captured per-body constants
void pd_sim::tapeCapture() {
  AdjTape& T = *tape;
  ++T.generation;
  T.paramsVersion = paramsVersion;
  T.numReal = numReal;
  const size_t n = (size_t)std::max(1, numReal);
  BodyC* b = T.arena.takeN<BodyC>(n);
  double* sc = T.arena.takeN<double>(n);
  double4* ef = T.arena.takeN<double4>(n);
  if (numReal > 0) {
    HIPCHECK(hipMemcpyAsync(b, bc.p, numReal * sizeof(BodyC), hipMemcpyDeviceToDevice, st));
    HIPCHECK(hipMemcpyAsync(sc, scale.p, numReal * sizeof(double), hipMemcpyDeviceToDevice, st));
    HIPCHECK(hipMemcpyAsync(ef, extForce.p, numReal * sizeof(double4), hipMemcpyDeviceToDevice,
                            st));
  }
  T.bc = b;
  T.scale = sc;
  T.extForce = ef;
}
// a state array of `count` doubles must hold the recorded bodies' 13 numbers
void pd_sim::tapeCount(long count) const {
  const long want = tape ? 13L * tape->numReal : 0L;
  if (count != want)
    throw ArgError("adjoint tape: a state array of " + std::to_string(count) +
                   " doubles was passed where the recording needs " + std::to_string(want) +
                   " (13 per body)");
}

void pd_sim::tapeSubstepStart(bool usePGS, bool legacyFriction) {
  AdjTape& T = *tape;
  tapeCheck(usePGS, legacyFriction);
  if (paramsVersion != T.paramsVersion)
    throw ArgError("adjoint tape: the particle state, masses, inertia, scales or external forces "
                   "were set while recording (open a new recording after setting them)");
  if (numReal != T.numReal) throw ArgError("adjoint tape: the particle count changed while recording");
  const int s = (int)T.steps.size();
  if (s % T.every != 0) return;
  TapeCkpt c;
  c.X = T.arena.takeN<double>((size_t)13 * std::max(1, numReal));
  if (numReal > 0) ad::k_tape_state<<<nblk(numReal), TPB, 0, st>>>(dD, numReal, c.X);
  c.nL = prevCount;
  c.L = T.arena.takeN<double>((size_t)5 * std::max(1, prevCount));
  unsigned long long* keys = nullptr;
  if (s == 0) {
    keys = T.arena.takeN<unsigned long long>(std::max(1, prevCount));
    T.key0 = keys;
    T.nKey0 = prevCount;
  }
  if (prevCount > 0) ad::k_tape_ledger<<<nblk(prevCount), TPB, 0, st>>>(dD, prevCount, c.L, keys);
  HIPCHECK(hipGetLastError());
  T.ckpts.push_back(c);
}

void pd_sim::tapeSubstepEnd(const Params& P, const std::vector<int>& vs, int numColors,
                            const std::vector<int>& ps, int numPosColors, int nRowsP,
                            int velLeftover, int posLeftover) {
  AdjTape& T = *tape;
  if (velLeftover > 0 || posLeftover > 0 || numColors > 62 || numPosColors > 62) {
    T.clear();
    T.on = false;
    taping = false;
    throw std::runtime_error(velLeftover > 0 || posLeftover > 0
                                 ? "adjoint tape: a colouring did not complete (Jacobi leftovers)"
                                 : "adjoint tape: the last colour (62) holds conflicting rows");
  }
  TapeStep S;
  S.P = P;
  S.P.numParticles = numParticles;
  S.vRest = 2.0 * dt * std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
  S.vs.assign(vs.begin(), vs.begin() + numColors + 1);
  S.ps.assign(ps.begin(), ps.begin() + numPosColors + 1);
  S.nColV = numColors;
  S.nColP = numPosColors;
  S.sweepsV = stats.sweeps_vel;
  S.stabRan = stats.stab_ran;
  S.sweepsS = stats.sweeps_stab;
  S.sweepsP = stats.sweeps_pos;
  ad::Topo& t = S.t;
  t.nSlots = numParticles;
  t.nc = nc;
  t.nm = nm;
  t.nRows = nRowsV;
  t.nPRows = nRowsP;
  int2* cAB_ = T.arena.takeN<int2>(std::max(1, nc));
  int* cPl_ = T.arena.takeN<int>(std::max(1, nc));
  int* pRowOf_ = T.arena.takeN<int>(std::max(1, nc));
  int2* mRange_ = T.arena.takeN<int2>(std::max(1, nm));
  unsigned long long* lKey_ = T.arena.takeN<unsigned long long>(std::max(1, nm));
  int* srcRowL_ = T.arena.takeN<int>(std::max(1, nm));
  int* pimpMan_ = T.arena.takeN<int>(std::max(1, nm));
  int* vPerm_ = T.arena.takeN<int>(std::max(1, nRowsV));
  int2* vAB_ = T.arena.takeN<int2>(std::max(1, nRowsV));
  unsigned char* vFlag_ = T.arena.takeN<unsigned char>(std::max(1, nRowsV));
  int* warm_ = T.arena.takeN<int>(std::max(1, nRowsV));
  int* pPerm_ = T.arena.takeN<int>(std::max(1, nRowsP));
  const unsigned long long* prevKeys = T.steps.empty() ? T.key0 : T.steps.back().lKey;
  const int nPrev = T.steps.empty() ? T.nKey0 : T.steps.back().t.nm;
  if (nc > 0) ad::k_tape_contacts<<<nblk(nc), TPB, 0, st>>>(dD, P, nc, cAB_, cPl_, pRowOf_);
  if (nm > 0)
    ad::k_tape_manifolds<<<nblk(nm), TPB, 0, st>>>(dD, nm, mRange_, lKey_, srcRowL_, pimpMan_);
  const int nr = std::max(nRowsV, nRowsP);
  if (nr > 0)
    ad::k_tape_rows<<<nblk(nr), TPB, 0, st>>>(dD, nRowsV, nRowsP, vPerm_, vAB_, vFlag_, warm_,
                                               prevKeys, nPrev, pPerm_);
  HIPCHECK(hipGetLastError());
  t.cAB = cAB_;
  t.cPl = cPl_;
  t.pRowOf = pRowOf_;
  t.mRange = mRange_;
  t.srcRowL = srcRowL_;
  t.pimpMan = pimpMan_;
  t.vPerm = vPerm_;
  t.vAB = vAB_;
  t.vFlag = vFlag_;
  t.warmEntry = warm_;
  t.pPerm = pPerm_;
  S.lKey = lKey_;
  T.steps.push_back(std::move(S));
}

namespace {
namespace adjrun {

inline int nb(long n) { return n > 0 ? (int)((n + TPB - 1) / TPB) : 1; }

// the next n elements of type E from a carving cursor, 256-byte aligned (a null cursor measures)
template <class E>
E* carvePiece(char*& cur, size_t n) {
  const size_t at = ((size_t)cur + 255) / 256 * 256;
  cur = (char*)(at + n * sizeof(E));
  return (E*)at;
}
template <class T>
struct WS {  // replay workspace
  DBuf<T> xs, xt, v0, vh, q0, vc, xr, qr, cg, ms, rc, lam, pc, pl, par;
  DBuf<ad::SlotK> sk;
  DBuf<int> realOf, gcnt, goff, cMan;
  struct Sizes {
    size_t slots = 1, ncm = 1, nmm = 1, rows = 1, prows = 1;
  };
  static Sizes sizes(const AdjTape& A) {
    Sizes z;
    for (const TapeStep& S : A.steps) {
      z.slots = std::max(z.slots, (size_t)S.t.nSlots);
      z.ncm = std::max(z.ncm, (size_t)S.t.nc);
      z.nmm = std::max(z.nmm, (size_t)S.t.nm);
      z.rows = std::max(z.rows, (size_t)S.t.nRows);
      z.prows = std::max(z.prows, (size_t)S.t.nPRows);
    }
    return z;
  }
  void ensure(const AdjTape& A, hipStream_t st) {
    const Sizes z = sizes(A);
    const size_t slots = z.slots, ncm = z.ncm, nmm = z.nmm, rows = z.rows, prows = z.prows;
    xs.ensure(3 * slots, st); xt.ensure(3 * slots, st); v0.ensure(3 * slots, st);
    vh.ensure(6 * slots, st); q0.ensure(4 * slots, st); vc.ensure(6 * slots, st);
    xr.ensure(3 * slots, st); qr.ensure(4 * slots, st); cg.ensure(10 * ncm, st);
    ms.ensure(15 * nmm, st); rc.ensure(20 * rows, st); lam.ensure(4 * rows, st);
    pc.ensure(10 * prows, st); pl.ensure(prows, st); par.ensure(2, st);
    sk.ensure(slots, st); realOf.ensure(slots, st); gcnt.ensure((size_t)A.numReal + 1, st);
    goff.ensure((size_t)A.numReal + 1, st); cMan.ensure(ncm, st);
  }
  void release() {
    xs.release(); xt.release(); v0.release(); vh.release(); q0.release(); vc.release();
    xr.release(); qr.release(); cg.release(); ms.release(); rc.release(); lam.release();
    pc.release(); pl.release(); par.release(); sk.release(); realOf.release(); gcnt.release();
    goff.release(); cMan.release();
  }
  // the same buffers carved from caller-owned memory (`cur` advances past them); such a
  // workspace owns nothing and is not released
  void carve(const AdjTape& A, char*& cur) {
    const Sizes z = sizes(A);
    auto put = [&](auto& b, size_t n) {
      using E = std::remove_reference_t<decltype(*b.p)>;
      b.p = carvePiece<E>(cur, n);
      b.cap = n;
    };
    put(xs, 3 * z.slots); put(xt, 3 * z.slots); put(v0, 3 * z.slots); put(vh, 6 * z.slots);
    put(q0, 4 * z.slots); put(vc, 6 * z.slots); put(xr, 3 * z.slots); put(qr, 4 * z.slots);
    put(cg, 10 * z.ncm); put(ms, 15 * z.nmm); put(rc, 20 * z.rows); put(lam, 4 * z.rows);
    put(pc, 10 * z.prows); put(pl, z.prows); put(par, 2); put(sk, z.slots);
    put(realOf, z.slots); put(gcnt, (size_t)A.numReal + 1); put(goff, (size_t)A.numReal + 1);
    put(cMan, z.ncm);
  }
  // bytes of one carved workspace and its tape pair of vt + pt doubles, a multiple of 256 (so
  // consecutive ones carved from a 256-byte aligned base carve exactly as measured)
  static size_t carveBytes(const AdjTape& A, size_t vt, size_t pt) {
    WS<T> w;
    char* cur = nullptr;
    w.carve(A, cur);
    (void)carvePiece<double>(cur, vt);
    (void)carvePiece<double>(cur, pt);
    return ((size_t)cur + 255) / 256 * 256;
  }
  ad::RB<T> rb(T* X, T* Xn, T* L, const T* Lp) {
    ad::RB<T> B{};
    B.X = X; B.Xn = Xn; B.xs = xs.p; B.xt = xt.p; B.v0 = v0.p; B.vh = vh.p; B.q0 = q0.p;
    B.vc = vc.p; B.xr = xr.p; B.qr = qr.p; B.cg = cg.p; B.ms = ms.p; B.rc = rc.p;
    B.lam = lam.p; B.pc = pc.p; B.pl = pl.p; B.L = L; B.Lp = Lp; B.par = par.p; B.sk = sk.p;
    B.realOf = realOf.p; B.gcnt = gcnt.p; B.goff = goff.p; B.cMan = cMan.p;
    return B;
  }
};

// the recording's captured per-body constants (not the simulation's live arrays)
inline ad::RK constants(const AdjTape& A, const TapeStep& S) {
  ad::RK K;
  K.P = S.P;
  K.vRest = S.vRest;
  K.bc = A.bc;
  K.scale = A.scale;
  K.extForce = A.extForce;
  return K;
}

// one substep of the replay; vtape / ptape (T = double) record the sweep inputs
template <class T>
void replayStep(pd_sim& s, const TapeStep& S, ad::RB<T> B, double* vtape, double* ptape) {
  hipStream_t st = s.st;
  const ad::RK K = constants(*s.tape, S);
  const ad::Topo& t = S.t;
  const int nR = K.P.numReal;
  ad::a_start<T><<<nb(nR), TPB, 0, st>>>(B, K);
  s.scanQuiet(B.gcnt, B.goff, nR);
  ad::a_predict<T><<<nb(nR), TPB, 0, st>>>(B, K, t.nSlots);
  if (t.nc > 0) ad::a_contacts<T><<<nb(t.nc), TPB, 0, st>>>(B, K, t);
  if (t.nm > 0) ad::a_manifolds<T><<<nb(t.nm), TPB, 0, st>>>(B, t);
  if (t.nRows > 0) {
    ad::a_rows<T><<<nb(t.nRows), TPB, 0, st>>>(B, K, t);
    const size_t stride = (size_t)16 * t.nRows;
    int pass = 0;
    auto run = [&](int mode) {
      double* tp = vtape ? vtape + stride * pass : nullptr;
      for (int c = 0; c < S.nColV; ++c) {
        const int r0 = S.vs[c], r1 = S.vs[c + 1];
        if (r1 > r0) ad::a_vpass<T><<<nb(r1 - r0), TPB, 0, st>>>(B, K, t, mode, r0, r1, tp);
      }
      ++pass;
    };
    run(0);
    for (int it = 0; it < S.sweepsV; ++it) run(1);
    if (S.stabRan)
      for (int it = 0; it < S.sweepsS; ++it) run(2);
  }
  if (t.nm > 0) ad::a_ledger<T><<<nb(t.nm), TPB, 0, st>>>(B, t);
  ad::a_reint<T><<<nb(t.nSlots), TPB, 0, st>>>(B, K, t.nSlots);
  if (t.nPRows > 0) {
    ad::a_prows<T><<<nb(t.nPRows), TPB, 0, st>>>(B, t);
    const size_t stride = (size_t)7 * t.nPRows;
    for (int it = 0; it < S.sweepsP; ++it) {
      double* tp = ptape ? ptape + stride * it : nullptr;
      for (int c = 0; c < S.nColP; ++c) {
        const int r0 = S.ps[c], r1 = S.ps[c + 1];
        if (r1 > r0) ad::a_ppass<T><<<nb(r1 - r0), TPB, 0, st>>>(B, t, r0, r1, tp);
      }
    }
  }
  if (t.nm > 0) ad::a_pimp<T><<<nb(t.nm), TPB, 0, st>>>(B, K, t);
  ad::a_final<T><<<nb(nR), TPB, 0, st>>>(B, K);
  HIPCHECK(hipGetLastError());
}

// the per-substep adjoint buffers, carved from one allocation so that one fill clears them all
struct GCarve {
  double *xs, *xt, *v0, *vh, *q0, *vc, *xr, *qr, *cg, *ms, *rc, *lam, *pc, *pl;
  size_t total;  // doubles used
};
inline GCarve carveG(double* base, size_t slots, size_t nc, size_t nm, size_t rows, size_t prows) {
  GCarve c{};
  size_t o = 0;
  // each piece starts on a 64-byte boundary
  auto take = [&](size_t n) {
    double* p = base ? base + o : nullptr;
    o += (n + 7) / 8 * 8;
    return p;
  };
  c.xs = take(3 * slots); c.xt = take(3 * slots); c.v0 = take(3 * slots); c.vh = take(6 * slots);
  c.q0 = take(4 * slots); c.vc = take(6 * slots); c.xr = take(3 * slots); c.qr = take(4 * slots);
  c.cg = take(10 * nc); c.ms = take(15 * nm); c.rc = take(22 * rows); c.lam = take(4 * rows);
  c.pc = take(10 * prows); c.pl = take(prows);
  c.total = o;
  return c;
}
struct GWS {  // adjoint workspace
  DBuf<double> X, Z, La, Lb, par;
  DBuf<double> vtape, ptape;
  size_t nL = 1, vtN = 1, ptN = 1;  // (tape lengths, doubles)
  void ensure(const AdjTape& A, hipStream_t st) {
    size_t nmm = 1, vt = 1, pt = 1, z = 1;
    for (const TapeStep& S : A.steps) {
      z = std::max(z, carveG(nullptr, S.t.nSlots, std::max(1, S.t.nc), std::max(1, S.t.nm),
                             std::max(1, S.t.nRows), std::max(1, S.t.nPRows)).total);
      nmm = std::max(nmm, (size_t)S.t.nm);
      const size_t passes = 1 + S.sweepsV + (S.stabRan ? S.sweepsS : 0);
      vt = std::max(vt, passes * 16 * (size_t)S.t.nRows);
      pt = std::max(pt, (size_t)S.sweepsP * 7 * (size_t)S.t.nPRows);
    }
    for (const TapeCkpt& c : A.ckpts) nmm = std::max(nmm, (size_t)c.nL);
    nL = nmm;
    vtN = vt;
    ptN = pt;
    X.ensure(13 * (size_t)std::max(1, A.numReal), st);
    Z.ensure(z, st);
    La.ensure(5 * nmm, st);
    Lb.ensure(5 * nmm, st); par.ensure(2, st); vtape.ensure(vt, st); ptape.ensure(pt, st);
  }
  void release() {
    X.release(); Z.release(); La.release(); Lb.release(); par.release();
    vtape.release(); ptape.release();
  }
};

// reverse of one substep: G.X holds the adjoint of the end state on entry and of the start state
// on exit; gL the adjoint of this substep's ledger, gLp (zeroed here) receives the previous one's
void reverseStep(pd_sim& s, const TapeStep& S, ad::RB<double> B, GWS& W, const double* vtape,
                 const double* ptape, double* gL, double* gLp, int nLp) {
  hipStream_t st = s.st;
  const ad::RK K = constants(*s.tape, S);
  const ad::Topo& t = S.t;
  const int nR = K.P.numReal;
  const GCarve C = carveG(W.Z.p, (size_t)t.nSlots, std::max(1, t.nc), std::max(1, t.nm),
                         std::max(1, t.nRows), std::max(1, t.nPRows));
  s.zero(W.Z.p, C.total);
  s.zero(gLp, 5 * (size_t)std::max(1, nLp));
  ad::GB G{};
  G.X = W.X.p; G.xs = C.xs; G.xt = C.xt; G.v0 = C.v0; G.vh = C.vh; G.q0 = C.q0;
  G.vc = C.vc; G.xr = C.xr; G.qr = C.qr; G.cg = C.cg; G.ms = C.ms; G.rc = C.rc;
  G.lam = C.lam; G.pc = C.pc; G.pl = C.pl; G.L = gL; G.Lp = gLp;
  if (t.nm > 0) ad::b_ledger<<<nb(t.nm), TPB, 0, st>>>(B, G, K, t);
  ad::b_final<<<nb(nR), TPB, 0, st>>>(G, K);
  if (t.nPRows > 0) {
    const size_t stride = (size_t)7 * t.nPRows;
    for (int it = S.sweepsP - 1; it >= 0; --it)
      for (int c = S.nColP - 1; c >= 0; --c) {
        const int r0 = S.ps[c], r1 = S.ps[c + 1];
        if (r1 > r0)
          ad::b_ppass<<<nb(r1 - r0), TPB, 0, st>>>(B, G, t, r0, r1, ptape + stride * it);
      }
    ad::b_prows<<<nb(t.nPRows), TPB, 0, st>>>(B, G, t);
  }
  ad::b_reint<<<nb(t.nSlots), TPB, 0, st>>>(B, G, K, t.nSlots);
  if (t.nRows > 0) {
    const size_t stride = (size_t)16 * t.nRows;
    const int passes = 1 + S.sweepsV + (S.stabRan ? S.sweepsS : 0);
    for (int p = passes - 1; p >= 0; --p) {
      const int mode = p == 0 ? 0 : (p <= S.sweepsV ? 1 : 2);
      for (int c = S.nColV - 1; c >= 0; --c) {
        const int r0 = S.vs[c], r1 = S.vs[c + 1];
        if (r1 <= r0) continue;
        if (mode == 0)
          ad::b_vpass<<<nb((long)(r1 - r0) * ad::VP_SPLIT), TPB, 0, st>>>(
              B, G, K, t, mode, r0, r1, vtape + stride * p);
        else
          ad::b_vpass_pgs<<<nb(r1 - r0), TPB, 0, st>>>(B, G, K, t, mode, r0, r1,
                                                       vtape + stride * p);
      }
    }
  }
  ad::b_warm<<<nb(std::max(t.nRows, nR)), TPB, 0, st>>>(G, K, t);
  if (t.nRows > 0) ad::b_rows<<<nb(t.nRows), TPB, 0, st>>>(B, G, K, t, W.par.p);
  if (t.nc > 0) {
    ad::b_manifolds<<<nb(t.nc), TPB, 0, st>>>(B, G, t);
    ad::b_contacts<<<nb(t.nc), TPB, 0, st>>>(B, G, K, t);
  }
  ad::b_predict<<<nb(nR), TPB, 0, st>>>(B, G, K, (int)t.nSlots);
  HIPCHECK(hipGetLastError());
}

void checkTape(const pd_sim& s, const AdjTape* A) {
  if (!A || A->steps.empty()) throw ArgError("no recorded substeps (pd_tape_begin, then step)");
  if (A->on) throw ArgError("the tape is still recording (pd_tape_end first)");
  if (A->ckpts.size() != (A->steps.size() + A->every - 1) / A->every)
    throw ArgError("the recording is incomplete (a substep failed while recording)");
  if (s.numReal != A->numReal) throw ArgError("the particle count changed since the recording");
  const Params& P0 = A->steps[0].P;
  for (const TapeStep& S : A->steps)
    if (S.P.mu != P0.mu || S.P.restN != P0.restN)
      throw ArgError("friction or restitution changed during the recording");
}

// T = double primal replay from the first checkpoint; returns the end state [numReal][13]
void replayAll(pd_sim& s, AdjTape* A, std::vector<double>& out) {
  checkTape(s, A);
  const int nR = A->numReal;
  WS<double> W;
  W.ensure(*A, s.st);
  size_t nL = 1;
  for (const TapeStep& S : A->steps) nL = std::max(nL, (size_t)S.t.nm);
  for (const TapeCkpt& c : A->ckpts) nL = std::max(nL, (size_t)c.nL);
  DBuf<double> Xa, Xb, La, Lb;
  Xa.ensure(13 * (size_t)std::max(1, nR), s.st);
  Xb.ensure(13 * (size_t)std::max(1, nR), s.st);
  La.ensure(5 * nL, s.st);
  Lb.ensure(5 * nL, s.st);
  HIPCHECK(hipMemcpyAsync(Xa.p, A->ckpts[0].X, sizeof(double) * 13 * nR, hipMemcpyDeviceToDevice, s.st));
  if (A->ckpts[0].nL > 0)
    HIPCHECK(hipMemcpyAsync(La.p, A->ckpts[0].L, sizeof(double) * 5 * A->ckpts[0].nL,
                            hipMemcpyDeviceToDevice, s.st));
  const double par[2] = {A->steps[0].P.mu, A->steps[0].P.restN};
  s.writeN(W.par.p, par, 2);
  double *X = Xa.p, *Xn = Xb.p, *Lp = La.p, *L = Lb.p;
  for (const TapeStep& S : A->steps) {
    replayStep<double>(s, S, W.rb(X, Xn, L, Lp), nullptr, nullptr);
    std::swap(X, Xn);
    std::swap(L, Lp);
  }
  out.resize((size_t)13 * nR);
  s.readN(out.data(), X, out.size());
  W.release();
  Xa.release(); Xb.release(); La.release(); Lb.release();
}

// forward mode: tangent of the end state along (dX0, dmu, de); returns value and tangent
void tangentAll(pd_sim& s, AdjTape* A, const std::vector<double>& dX0, double dmu, double de,
                std::vector<double>& val, std::vector<double>& tan) {
  using DT = ad::DN<1>;
  checkTape(s, A);
  const int nR = A->numReal;
  const size_t nX = (size_t)13 * nR;
  WS<DT> W;
  W.ensure(*A, s.st);
  size_t nL = 1;
  for (const TapeStep& S : A->steps) nL = std::max(nL, (size_t)S.t.nm);
  for (const TapeCkpt& c : A->ckpts) nL = std::max(nL, (size_t)c.nL);
  DBuf<DT> Xa, Xb, La, Lb;
  Xa.ensure(std::max<size_t>(1, nX), s.st);
  Xb.ensure(std::max<size_t>(1, nX), s.st);
  La.ensure(5 * nL, s.st);
  Lb.ensure(5 * nL, s.st);
  DBuf<double> tmp;
  tmp.ensure(std::max<size_t>(1, nX), s.st);
  s.writeN(tmp.p, dX0.data(), nX);
  ad::k_seed_dn<<<nb((long)nX), TPB, 0, s.st>>>(A->ckpts[0].X, tmp.p, Xa.p, (long)nX);
  if (A->ckpts[0].nL > 0)
    ad::k_seed_dn<<<nb(5L * A->ckpts[0].nL), TPB, 0, s.st>>>(A->ckpts[0].L, nullptr, La.p,
                                                             5L * A->ckpts[0].nL);
  DT par[2];
  par[0] = DT(A->steps[0].P.mu);
  par[0].d[0] = dmu;
  par[1] = DT(A->steps[0].P.restN);
  par[1].d[0] = de;
  s.writeN(W.par.p, par, 2);
  DT *X = Xa.p, *Xn = Xb.p, *Lp = La.p, *L = Lb.p;
  for (const TapeStep& S : A->steps) {
    replayStep<DT>(s, S, W.rb(X, Xn, L, Lp), nullptr, nullptr);
    std::swap(X, Xn);
    std::swap(L, Lp);
  }
  DBuf<double> vd, td;
  vd.ensure(std::max<size_t>(1, nX), s.st);
  td.ensure(std::max<size_t>(1, nX), s.st);
  ad::k_split_dn<<<nb((long)nX), TPB, 0, s.st>>>(X, vd.p, td.p, (long)nX);
  val.resize(nX);
  tan.resize(nX);
  s.readN(val.data(), vd.p, nX);
  s.readN(tan.data(), td.p, nX);
  W.release();
  Xa.release(); Xb.release(); La.release(); Lb.release(); tmp.release(); vd.release(); td.release();
}

// reverse mode: adjoint of the start state and of (mu, restitution) for the end-state adjoint gX
void adjointAll(pd_sim& s, AdjTape* A, const std::vector<double>& gXend, std::vector<double>& gX0,
                double* gpar) {
  checkTape(s, A);
  const int nR = A->numReal;
  const size_t nX = (size_t)13 * nR;
  const int nS = (int)A->steps.size();
  const int every = A->every;
  WS<double> W;
  W.ensure(*A, s.st);
  GWS G;
  G.ensure(*A, s.st);
  const double par[2] = {A->steps[0].P.mu, A->steps[0].P.restN};
  s.writeN(W.par.p, par, 2);
  s.zero(G.par.p, 2);
  s.writeN(G.X.p, gXend.data(), nX);
  // states of one segment: X_s for s in [s0, s1], ledgers L_{s0-1} .. L_{s1-1}
  DBuf<double> Xs, Ls;
  Xs.ensure((size_t)(every + 1) * std::max<size_t>(1, nX), s.st);
  Ls.ensure((size_t)(every + 1) * 5 * G.nL, s.st);
  double* gL = G.La.p;   // adjoint of the current substep's ledger
  double* gLp = G.Lb.p;  // adjoint of the previous one
  s.zero(gL, 5 * G.nL);
  // A segment's substeps are replayed once each into workspaces and tapes of their own when
  // those fit in a quarter of the free device memory; otherwise the segment is replayed for its
  // states and each substep again, with its tapes, right before its reverse.
  const int segMax = std::min(every, nS);
  bool keep = false;
  if (segMax > 1 && s.adjKeepSegment) {
    size_t freeB = 0, totalB = 0;
    HIPCHECK(hipMemGetInfo(&freeB, &totalB));
    keep = (size_t)segMax * W.carveBytes(*A, G.vtN, G.ptN) <= freeB / 4;
  }
  // one allocation for the segment's workspaces and tapes (one per substep)
  DBuf<char> segArena;
  std::vector<WS<double>> Wk;
  std::vector<double*> vtk, ptk;
  if (keep) {
    const size_t per = W.carveBytes(*A, G.vtN, G.ptN);
    segArena.ensure(per * segMax, s.st);
    Wk.resize(segMax);
    vtk.resize(segMax);
    ptk.resize(segMax);
    for (int i = 0; i < segMax; ++i) {
      char* cur = segArena.p + per * i;
      Wk[i].carve(*A, cur);
      vtk[i] = carvePiece<double>(cur, G.vtN);
      ptk[i] = carvePiece<double>(cur, G.ptN);
      s.writeN(Wk[i].par.p, par, 2);
    }
  }
  const int nSeg = (nS + every - 1) / every;
  for (int seg = nSeg - 1; seg >= 0; --seg) {
    const int s0 = seg * every, s1 = std::min(nS, s0 + every);
    const TapeCkpt& c = A->ckpts[seg];
    HIPCHECK(hipMemcpyAsync(Xs.p, c.X, sizeof(double) * nX, hipMemcpyDeviceToDevice, s.st));
    if (c.nL > 0)
      HIPCHECK(hipMemcpyAsync(Ls.p, c.L, sizeof(double) * 5 * c.nL, hipMemcpyDeviceToDevice, s.st));
    auto Xat = [&](int k) { return Xs.p + (size_t)(k - s0) * nX; };
    auto Lat = [&](int k) { return Ls.p + (size_t)(k - s0 + 1) * 5 * G.nL; };  // L_k, k >= s0-1
    if (keep) {
      // one replay per substep, each into its own workspace and tapes, then the reverses
      for (int k = s0; k < s1; ++k)
        replayStep<double>(s, A->steps[k], Wk[k - s0].rb(Xat(k), Xat(k + 1), Lat(k), Lat(k - 1)),
                           vtk[k - s0], ptk[k - s0]);
      for (int k = s1 - 1; k >= s0; --k) {
        const int nLp = k == 0 ? A->nKey0 : A->steps[k - 1].t.nm;
        reverseStep(s, A->steps[k], Wk[k - s0].rb(Xat(k), Xat(k + 1), Lat(k), Lat(k - 1)), G,
                    vtk[k - s0], ptk[k - s0], gL, gLp, nLp);
        std::swap(gL, gLp);
      }
      continue;
    }
    for (int k = s0; k < s1 - 1; ++k)
      replayStep<double>(s, A->steps[k], W.rb(Xat(k), Xat(k + 1), Lat(k), Lat(k - 1)), nullptr,
                         nullptr);
    for (int k = s1 - 1; k >= s0; --k) {
      const TapeStep& S = A->steps[k];
      replayStep<double>(s, S, W.rb(Xat(k), Xat(k + 1), Lat(k), Lat(k - 1)), G.vtape.p,
                         G.ptape.p);
      const int nLp = k == 0 ? A->nKey0 : A->steps[k - 1].t.nm;
      reverseStep(s, S, W.rb(Xat(k), Xat(k + 1), Lat(k), Lat(k - 1)), G, G.vtape.p, G.ptape.p,
                  gL, gLp, nLp);
      std::swap(gL, gLp);
    }
  }
  gX0.resize(nX);
  s.readN(gX0.data(), G.X.p, nX);
  s.readN(gpar, G.par.p, 2);
  segArena.release();  // (the carved workspaces own nothing)
  W.release();
  G.release();
  Xs.release();
  Ls.release();
}

}  // namespace adjrun
}  // namespace
