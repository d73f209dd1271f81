// pd_sim.hip — host orchestration of the XPBD substep and the C-ABI of libpecletdem_hip.so.
//
// One pd_sim owns one HIP stream and every device buffer of one particle system.  pd_step runs
// the substep of the reference demStep (src/sim.hpp:97-488) as a sequence of kernels from
// pd_kernels.hpp on that stream; buffer counts that size the next launch are read back to the
// host at the same points where the reference reads its device counters.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <limits>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "pd_adjoint.hpp"
#include "pd_hertz.hpp"
#include "pd_kernels.hpp"
#include "peclet_dem.h"

using namespace pd;

namespace {

thread_local std::string g_error;

struct ArgError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define HIPCHECK(x)                                                                 \
  do {                                                                              \
    hipError_t e_ = (x);                                                            \
    if (e_ != hipSuccess)                                                           \
      throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e_) + \
                               " at " #x);                                          \
  } while (0)

// every kernel bounds-checks its index, so an empty range still launches one block
inline int nblk(long n) { return n > 0 ? (int)((n + TPB - 1) / TPB) : 1; }
// grid of a launch mapped with xcdTid: a multiple of the 8 XCDs
inline int xcdGrid(long n) { return (nblk(n) + NUM_XCD - 1) / NUM_XCD * NUM_XCD; }

// Device allocation is serialised process-wide, and a buffer replaced by growth is not freed at
// once: it is retired and freed at a point where no stream can still read it (after a device
// synchronisation).  Systems of a batch grow their buffers from several host threads while the
// other systems' streams run; freeing there would race their in-flight work (a system's second
// stream, or kernels enqueued before its pointer table was refreshed) and interleave the
// runtime's allocator calls across threads.
std::mutex g_allocMutex;
std::vector<void*> g_retired;       // replaced buffers waiting for a safe free (under the mutex)
size_t g_retiredBytes = 0;
std::atomic<int> g_batchDepth{0};   // > 0 while a batch step drives systems from several threads
constexpr size_t RETIRED_LIMIT = size_t(4) << 30;

// free every retired buffer; the caller holds g_allocMutex
void drainRetiredLocked() {
  if (g_retired.empty()) return;
  HIPCHECK(hipDeviceSynchronize());
  for (void* p : g_retired) (void)hipFree(p);
  g_retired.clear();
  g_retiredBytes = 0;
}
// free the retired buffers at a point where the calling thread is the only one driving systems
void drainRetired() {
  if (g_batchDepth.load() > 0) return;
  std::lock_guard<std::mutex> lk(g_allocMutex);
  drainRetiredLocked();
}

// grow-only device buffer; growth keeps the first `keep` elements
template <class T>
struct DBuf {
  T* p = nullptr;
  size_t cap = 0;
  void release() {
    if (p) {
      std::lock_guard<std::mutex> lk(g_allocMutex);
      (void)hipFree(p);
    }
    p = nullptr;
    cap = 0;
  }
  void ensure(size_t n, hipStream_t st, size_t keep = 0) {
    if (n <= cap) return;
    size_t nc = n + n / 4 + 64;
    T* q = nullptr;
    std::lock_guard<std::mutex> lk(g_allocMutex);
    // many retired bytes: free them first (a device synchronisation; nothing still reads them)
    if (g_retiredBytes > RETIRED_LIMIT) drainRetiredLocked();
    HIPCHECK(hipMalloc(&q, nc * sizeof(T)));
    HIPCHECK(hipMemsetAsync(q, 0, nc * sizeof(T), st));
    if (p && keep) HIPCHECK(hipMemcpyAsync(q, p, std::min(keep, cap) * sizeof(T),
                                           hipMemcpyDeviceToDevice, st));
    HIPCHECK(hipStreamSynchronize(st));
    if (p) {
      g_retired.push_back(p);
      g_retiredBytes += cap * sizeof(T);
    }
    p = q;
    cap = nc;
  }
};

enum Stage {
  ST_PREDICT,
  ST_BROAD,
  ST_NARROW,
  ST_MANIFOLD,
  ST_COLOR_V,
  ST_PREP,
  ST_VSWEEP,
  ST_STAB,
  ST_LEDGER,
  ST_REINT,
  ST_COLOR_P,
  ST_PSWEEP,
  ST_COMMIT,
  ST_COUNT
};
const char* kStageNames[ST_COUNT] = {"predict_ghosts", "broadphase",   "narrowphase",
                                     "manifolds",      "color_vel",    "vel_prepare_warm",
                                     "vel_sweeps",     "stabilization", "ledger",
                                     "reintegrate",    "color_pos",    "pos_sweeps",
                                     "commit"};

struct AdjTape;
void adjFree(AdjTape* t);

}  // namespace

struct pd_sim {
  hipStream_t st = nullptr;
  int capacity = 0;
  int device = 0;  // the device current when the handle was created (its streams' device)
  int numReal = 0, numParticles = 0;
  // shape 0 (src/sim.hpp:949-1021)
  int shapeType = SHAPE_SPHERE;
  double shapeParams[4] = {1.0, 0, 0, 0};
  std::vector<double> shellHost;  // 3 per point
  double baseRadius = 1.0;
  double defaultInvI[3] = {2.5, 2.5, 2.5};
  // host parameters (src/particles.hpp:173-191)
  double dmin[3] = {0, 0, 0}, dmax[3] = {0, 0, 0}, dsize[3] = {0, 0, 0};
  int per[3] = {0, 0, 0};
  double g[3] = {0, 0, 0};
  double dt = 1e-3, globalScale = 1.0, growthRate = 0.0, growthFactor = -1.0;
  double thermoTau = 0.0, thermoTemp = 0.0, thermoKB = 1.0;
  double frictionDynamic = 0.0, restitutionNormal = 0.0, restitutionTangent = 0.0;
  int posIters = 10, velIters = 0, useGS = 1, stabMode = 1;
  double planeP[8][3], planeN[8][3];
  int numPlanes = 0;
  int hasPairTable = 0;
  double pairTable[128];
  // results
  int np = 0, nc = 0, nm = 0, prevCount = 0, nRowsV = 0;
  int ledgerBuckets = 0;  // real-index buckets of the sorted ledger (bucketStart has +1 entries)
  double maxOverlap = 0.0, maxApproach = 0.0;
  pd_stats stats{};
  // imported grid-SDF particle shape (src/sim.hpp:1034-1084)
  DBuf<double> sdfGridD;
  int gridDims[3] = {0, 0, 0};
  double gridOrigin[3] = {0, 0, 0}, gridInvSpacing[3] = {0, 0, 0};
  // SDF walls (src/sim.hpp:1242-1287) and the Hertz–Mindlin engine (src/particles.hpp:113-130)
  std::vector<WallP> wallsHost;
  std::vector<double> wallGridHost;
  DBuf<double> wallGridD;
  double hertzE[8] = {1e9, 1e9, 1e9, 1e9, 1e9, 1e9, 1e9, 1e9};
  double hertzNu[8] = {0.25, 0.25, 0.25, 0.25, 0.25, 0.25, 0.25, 0.25};
  double hertzFrac = 0.5;
  int hNumPairs = -1, hPrevCount = 0, hNumWallCand = 0;
  DBuf<int2> hPairs;
  DBuf<int> hIStart, hJStart, hJList, hJCnt, hWStart, hWCand, hWCnt;
  DBuf<double4> hXi, hF, hTA, hTB, hPrevXi, hRefPos, hXiWall, hWF, hWT;
  DBuf<unsigned long long> hPrevKeys;
  DBuf<HzTab> hTab;
  HzDev Z{};
  // sweep loops over at most this many rows run in one workgroup (k_vsolve_block/k_psolve_block)
  int blockRows = BLOCK_SOLVER_ROWS;
  bool ldsLoop = true;  // one-workgroup velocity loops stage their bodies in LDS when they fit
  bool vsweepWaves4 = false;  // per-colour PGS sweeps compiled for 4 waves per SIMD
  int ldsBodies = VL_BODIES;  // ... up to this many bodies (the rest stay in global memory)
  int ldsKeys = VL_KEYS;      // more distinct bodies than this: the global-memory loop
  // extra slots reserved by setPositions for periodic ghosts (more are added when needed)
  long slotHeadroom = 4096;
  // the largest scale, cached while no scale is written (every writer clears scaleKnown)
  bool scaleKnown = false;
  unsigned long long maxScaleKey = 0;
  // timing: stage events (level >= 1) and per-launch events of the main kernels (level >= 2).
  // Events come from a pool and are resolved lazily (at a query, or every few hundred substeps),
  // so timing adds no host synchronisation to a substep.
  int timing = 0;
  // events are recorded on every timingEvery-th substep after pd_set_timing (substeps 0, N, 2N..)
  int timingEvery = 1;
  long timingSubstep = 0;
  bool timedNow = false;
  double stageMs[ST_COUNT] = {};
  // host wall time between the same marks (enqueue work plus waits), at timing >= 1
  double stageHostMs[ST_COUNT] = {};
  std::chrono::steady_clock::time_point stageHostT[ST_COUNT + 1];
  // kernels timed per launch at level 2; units = what each launch processed
  enum { K_VSWEEP, K_PSWEEP, K_BPQUERY, K_NPPAIRS, K_PREDICT, K_REINTEGRATE, K_VBLOCK, K_PBLOCK,
         K_COUNT };
  double kernMs[K_COUNT] = {};
  long kernLaunches[K_COUNT] = {};
  long kernRows[K_COUNT] = {};
  struct KernRec {
    hipEvent_t a, b;
    int kid;
    long units;
  };
  std::vector<hipEvent_t> evFree;
  std::vector<KernRec> kernPending;
  std::vector<hipEvent_t> stagePending;  // ST_COUNT + 1 events per substep
  hipEvent_t stageCur[ST_COUNT + 1] = {};
  int stageMarked = -1;  // last stage marked in the current substep
  hipEvent_t kernOpen = nullptr;
  // per-substep counters, logged while timing >= 2
  std::vector<pd_stats> statsLog;
  hipEvent_t takeEvent() {
    if (evFree.empty()) {
      hipEvent_t e;
      // timestamps only: no system-scope fence when recorded (it would cost the timed work)
      HIPCHECK(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
      return e;
    }
    hipEvent_t e = evFree.back();
    evFree.pop_back();
    return e;
  }
  void kernBegin() {
    if (timing < 2 || !timedNow) return;
    kernOpen = takeEvent();
    HIPCHECK(hipEventRecord(kernOpen, st));
  }
  // returns the record's index (valid until the next resolve) or -1
  long kernEnd(int kid, long units) {
    if (timing < 2 || !timedNow) return -1;
    hipEvent_t b = takeEvent();
    HIPCHECK(hipEventRecord(b, st));
    kernPending.push_back({kernOpen, b, kid, units});
    return (long)kernPending.size() - 1;
  }
  // wait for every recorded event and fold the elapsed times into the totals
  void resolveTiming() {
    if (kernPending.empty() && stagePending.empty()) return;
    sync();
    for (const KernRec& r : kernPending) {
      float ms = 0.f;
      HIPCHECK(hipEventElapsedTime(&ms, r.a, r.b));
      kernMs[r.kid] += ms;
      kernLaunches[r.kid] += 1;
      kernRows[r.kid] += r.units;
      evFree.push_back(r.a);
      evFree.push_back(r.b);
    }
    kernPending.clear();
    for (size_t base = 0; base + ST_COUNT < stagePending.size(); base += ST_COUNT + 1) {
      for (int k = 0; k < ST_COUNT; ++k) {
        float ms = 0.f;
        HIPCHECK(hipEventElapsedTime(&ms, stagePending[base + k], stagePending[base + k + 1]));
        stageMs[k] += ms;
      }
      for (int k = 0; k <= ST_COUNT; ++k) evFree.push_back(stagePending[base + k]);
    }
    stagePending.clear();
  }
  void resetTiming(int level) {
    resolveTiming();
    timing = level;
    for (double& x : stageMs) x = 0.0;
    for (double& x : stageHostMs) x = 0.0;
    for (int k = 0; k < K_COUNT; ++k) {
      kernMs[k] = 0.0;
      kernLaunches[k] = 0;
      kernRows[k] = 0;
    }
    statsLog.clear();
    timingSubstep = 0;
  }
  // device buffers
  DBuf<double4> pos, posPred, quatPred, extForce, shell, sbox;
  DBuf<Body64> vw0, vw;
  DBuf<BodyC> bc;
  DBuf<double> scale, targetScale, pairTableD, thermo;
  DBuf<double4> dVel, dAng, dPos;
  DBuf<int> dCnt;
  DBuf<double> cLamN;
  DBuf<int> frCnt;
  DBuf<int> realIdx, ground, ground2, cellOf, cellRank, cellCnt, cellStart, order, pairCnt,
      pairStart;
  DBuf<unsigned char> matId, mCode, vFlag;
  DBuf<int> pairStage;  // broadphase partners staged slot-major
  DBuf<int> pairOfContact;  // point shapes: the pair of each pair contact
  DBuf<int2> pairs, cAB, mAB, mReal, mRange, mRange2, vAB, pAB;
  DBuf<int> ncnt, noff, cSlot, mCnt, mStart, mNp, mColor, cColor, vPerm, vRowOf, pPerm, pRowOf;
  DBuf<double4> cN, cRA, cRB, cBV, mNs, mTA, mTB, mRA, mRB, mWV, vN, vTA, vTB, vRA, vRB, vVt0,
      vLamT, vWV, pN,
      pRA, pRB, lLamT;
  DBuf<double> vVn0, vLam, pLam, lLam;
  DBuf<unsigned long long> mKey, lKey, uKey, red, bits, redPart;
  DBuf<int> supList, supCnt;  // support manifolds of the grounded-level sweeps
  DBuf<int> ctl;  // device loop control (CTL_* in pd_kernels.hpp)
  // order-exact parallel Jacobi / friction accumulation: per-element contributions and flags,
  // per-body incidence lists (count, start, list)
  DBuf<JContrib> jC;
  DBuf<int> jF, jIncCnt, jIncStart, jIncList;
  DBuf<int> lIdx, bucketCnt, bucketStart, counters, scanTmp, ghostCnt,
      ghostOff;
  // pinned host scratch
  static constexpr size_t HOST_SCRATCH_BYTES = 4096;
  void* hostScratch = nullptr;
  void* hostPost = nullptr;     // pinned area of the main stream's batched read-backs
  void* hostPostDev = nullptr;  // its device view

  Dev D{};

  // recording of substeps for the adjoint (pd_adjoint_host.hpp); null when never opened
  AdjTape* tape = nullptr;
  bool taping = false;  // a tape is open: every substep is recorded
  void tapeCheck(bool usePGS, bool legacyFriction) const;
  void tapeCapture();
  void tapeCount(long count) const;
  // bumped by every setter of masses, inertia, scales or external forces (a recording fails if
  // they change while it is open)
  long paramsVersion = 0;
  void tapeSubstepStart(bool usePGS, bool legacyFriction);
  void tapeSubstepEnd(const Params& P, const std::vector<int>& vs, int numColors,
                      const std::vector<int>& ps, int numPosColors, int nRowsP, int velLeftover,
                      int posLeftover);

  hipStream_t sideStream = nullptr;  // default priority, like st
  hipEvent_t evSide = nullptr;      // side stream done -> main stream
  hipEvent_t evContacts = nullptr;  // contact list built -> side stream

  pd_sim(int cap) : capacity(cap) {
    HIPCHECK(hipGetDevice(&device));
    HIPCHECK(hipStreamCreate(&st));
    HIPCHECK(hipStreamCreate(&sideStream));
    // the two streams exchange device memory only: no system-scope fence (cache writeback and
    // invalidation) when these are recorded
    HIPCHECK(hipEventCreateWithFlags(&evSide, hipEventDisableTiming | hipEventDisableSystemFence));
    HIPCHECK(hipEventCreateWithFlags(&evContacts,
                                     hipEventDisableTiming | hipEventDisableSystemFence));
    HIPCHECK(hipHostMalloc(&hostScratch, HOST_SCRATCH_BYTES, hipHostMallocDefault));
    HIPCHECK(hipMalloc((void**)&devD, sizeof(Dev)));
    dD = devD;
    HIPCHECK(hipHostMalloc((void**)&devRing, DEV_RING * sizeof(Dev), hipHostMallocDefault));
    for (int k = 0; k < DEV_RING; ++k) {
      HIPCHECK(hipEventCreateWithFlags(&devRingEv[k], hipEventDisableTiming));
      HIPCHECK(hipEventRecord(devRingEv[k], st));
    }
    HIPCHECK(hipHostMalloc(&jobV.pinned, 1024, hipHostMallocDefault));
    HIPCHECK(hipHostMalloc(&jobP.pinned, 1024, hipHostMallocDefault));
    // the read-back areas' sequence words start at 0, a number never posted
    std::memset(jobV.pinned, 0, 1024);
    std::memset(jobP.pinned, 0, 1024);
    HIPCHECK(hipHostMalloc(&hostPost, 1024, hipHostMallocDefault));
    std::memset(hostPost, 0, 1024);
    HIPCHECK(hipHostGetDevicePointer(&hostPostDev, hostPost, 0));
    HIPCHECK(hipHostGetDevicePointer(&jobV.pinnedDev, jobV.pinned, 0));
    HIPCHECK(hipHostGetDevicePointer(&jobP.pinnedDev, jobP.pinned, 0));
    jobV.stream = st;
    jobP.stream = sideStream;
    ensureSlots(cap > 0 ? cap : 1, 0);
    // src/sim.hpp:318-321: the multilevel gate mask comes from the environment, default slip
    if (const char* e = std::getenv("PECLET_DEM_ML_GATES")) mlGates = std::atoi(e);
    red.ensure(8, st);
    bits.ensure(4, st);
    counters.ensure(16, st);
    ctl.ensure(CTL_COUNT + 1, st);
    candFlag.ensure(1, st);
    supCnt.ensure(1, st);
    pairTableD.ensure(128, st);
    shell.ensure(1, st);
    initializeShape(SHAPE_SPHERE, 1.0, 0.0, 0.0);
    refreshDev();
  }
  ~pd_sim() {
    (void)hipStreamSynchronize(st);
    (void)hipStreamSynchronize(sideStream);
    adjFree(tape);
    DBufReleaseAll();
    for (const KernRec& r : kernPending) {
      (void)hipEventDestroy(r.a);
      (void)hipEventDestroy(r.b);
    }
    for (hipEvent_t e : stagePending) (void)hipEventDestroy(e);
    if (stageMarked >= 0)
      for (int k = 0; k <= stageMarked; ++k) (void)hipEventDestroy(stageCur[k]);
    for (hipEvent_t e : evFree) (void)hipEventDestroy(e);
    if (hostScratch) (void)hipHostFree(hostScratch);
    if (hostPost) (void)hipHostFree(hostPost);
    for (int k = 0; k < DEV_RING; ++k) (void)hipEventDestroy(devRingEv[k]);
    if (devRing) (void)hipHostFree(devRing);
    if (devD) (void)hipFree(devD);
    if (jobV.pinned) (void)hipHostFree(jobV.pinned);
    if (jobP.pinned) (void)hipHostFree(jobP.pinned);
    (void)hipEventDestroy(evSide);
    (void)hipEventDestroy(evContacts);
    (void)hipStreamDestroy(sideStream);
    (void)hipStreamDestroy(st);
  }
  void DBufReleaseAll() {
    pos.release(); posPred.release(); quatPred.release(); extForce.release(); shell.release();
    sbox.release(); vw0.release(); vw.release(); bc.release(); scale.release();
    targetScale.release(); pairTableD.release(); thermo.release(); realIdx.release();
    ground.release(); ground2.release(); cellOf.release(); cellRank.release(); cellCnt.release();
    cellStart.release(); gridD.release(); ncellD.release(); slotInfo.release();
    cellCntDirty = true;
    cand.release(); candStart.release(); candFlag.release(); candKeep.release();
    lRank.release(); ledgerCntDirty = true;
    candCode.release(); candRef.release(); candValid = false;
    order.release(); pairCnt.release(); pairStart.release(); pairStage.release();
    pairOfContact.release(); matId.release(); mCode.release();
    vFlag.release(); pairs.release(); cAB.release(); mAB.release(); mReal.release();
    mRange.release(); mRange2.release(); cBV.release(); mWV.release(); vWV.release();
    vAB.release(); pAB.release(); ncnt.release(); noff.release();
    cSlot.release(); mCnt.release(); mStart.release(); mNp.release(); mColor.release();
    cColor.release(); vPerm.release(); vRowOf.release(); pPerm.release(); pRowOf.release();
    cN.release(); cRA.release(); cRB.release(); mNs.release(); mTA.release(); mTB.release();
    mRA.release(); mRB.release(); vN.release(); vTA.release(); vTB.release(); vRA.release();
    vRB.release(); vVt0.release(); vLamT.release(); pN.release(); pRA.release(); pRB.release();
    lLamT.release(); vVn0.release(); vLam.release(); pLam.release(); lLam.release();
    mKey.release(); lKey.release(); uKey.release();
    red.release(); bits.release(); lIdx.release(); bucketCnt.release(); bucketStart.release();
    counters.release();
    scanTmp.release(); ghostCnt.release(); ghostOff.release();
    dVel.release(); dAng.release(); dPos.release(); dCnt.release(); cLamN.release();
    frCnt.release();
    jC.release(); jF.release(); jIncCnt.release(); jIncStart.release(); jIncList.release();
    wallGridD.release(); sdfGridD.release(); hPairs.release(); hIStart.release(); hJStart.release();
    hJList.release(); hJCnt.release(); hWStart.release(); hWCand.release(); hWCnt.release();
    hXi.release(); hF.release(); hTA.release(); hTB.release(); hPrevXi.release();
    hRefPos.release(); hXiWall.release(); hWF.release(); hWT.release(); hPrevKeys.release();
    hTab.release();
    jobV.release(); jobP.release(); redPart.release();
    ctl.release(); supList.release(); supCnt.release(); hLevel.release(); lvlKey.release(); lvlCnt.release();
    lvlStart.release(); lvlPerm.release(); lvlChanged.release(); lvlBuckets.release();
  }

  // ---------------------------------------------------------------- small helpers
  void sync() { HIPCHECK(hipStreamSynchronize(st)); }
  template <class T>
  T read1(const T* dptr) {
    HIPCHECK(hipMemcpyAsync(hostScratch, dptr, sizeof(T), hipMemcpyDeviceToHost, st));
    sync();
    T v;
    std::memcpy(&v, hostScratch, sizeof(T));
    return v;
  }
  template <class T>
  void readN(T* host, const T* dptr, size_t n) {
    if (!n) return;
    if (n * sizeof(T) <= HOST_SCRATCH_BYTES) {
      // small reads go through the pinned scratch (a direct copy, no staging)
      HIPCHECK(hipMemcpyAsync(hostScratch, dptr, n * sizeof(T), hipMemcpyDeviceToHost, st));
      sync();
      std::memcpy(host, hostScratch, n * sizeof(T));
      return;
    }
    HIPCHECK(hipMemcpyAsync(host, dptr, n * sizeof(T), hipMemcpyDeviceToHost, st));
    sync();
  }
  // Ranges read back together through one k_post launch into pinned memory `area` (host view;
  // areaDev is the device view): add() returns where each range will be, post() launches.
  // Wait until the posted word *seqp equals seq: the k_post that writes it wrote its values
  // first.  Polling the pinned word wakes the host as soon as the values land; the stream is
  // queried now and then so that a failed stream raises instead of spinning.
  // Polls the posted word; the stream is queried (to raise on a failed stream) only every
  // ~2 ms of waiting, since the runtime call takes locks that the other threads of a batch need
  // for their launches.
  static void waitPosted(const volatile int* seqp, int seq, hipStream_t s) {
    auto last = std::chrono::steady_clock::now();
    for (long spin = 1;; ++spin) {
      if (*seqp == seq) break;
      __builtin_ia32_pause();
      if ((spin & 1023) == 0) {
        const auto now = std::chrono::steady_clock::now();
        if (now - last < std::chrono::milliseconds(2)) continue;
        last = now;
        const hipError_t e = hipStreamQuery(s);
        if (e == hipSuccess) {
          if (*seqp == seq) break;
          throw std::runtime_error("read-back: the stream finished without posting");
        }
        if (e != hipErrorNotReady) HIPCHECK(e);
      }
    }
    std::atomic_thread_fence(std::memory_order_acquire);
  }
  int postSeq = 0;  // the last read-back number handed out (0 is never posted)
  struct PostBatch {
    PostSpec spec{};
    int used = 0;
    int* area = nullptr;
    int* areaDev = nullptr;
    int cap = 0;
    int seq = 0;
    int* seqHost = nullptr;
    int* seqDev = nullptr;
    const int* add(const int* src, int len) {
      if (spec.n >= 4 || used + len > cap) throw std::runtime_error("read-back batch too large");
      spec.src[spec.n] = src;
      spec.len[spec.n] = len;
      ++spec.n;
      const int* h = area + used;
      used += len;
      return h;
    }
    void post(hipStream_t s) {
      k_post<<<1, 128, 0, s>>>(spec, areaDev, seqDev, seq);
      HIPCHECK(hipGetLastError());
    }
    // the values are in host memory after it
    void wait(hipStream_t s) const { waitPosted(seqHost, seq, s); }
  };
  // the area's last int is the posted-sequence word
  PostBatch postBatch(void* pinnedArea, void* pinnedAreaDev, int capInts) {
    PostBatch b;
    b.area = static_cast<int*>(pinnedArea);
    b.areaDev = static_cast<int*>(pinnedAreaDev);
    b.cap = capInts - 1;
    b.seqHost = b.area + capInts - 1;
    b.seqDev = b.areaDev + capInts - 1;
    postSeq = postSeq == 0x7fffffff ? 1 : postSeq + 1;
    b.seq = postSeq;
    return b;
  }
  // several small reads with one synchronisation: enqueue each into the pinned scratch at
  // `off` (bytes, advanced), sync() once, then read the returned host pointers
  template <class T>
  const T* enqueueRead(const T* dptr, size_t n, size_t& off) {
    off = (off + 15) & ~(size_t)15;
    if (off + n * sizeof(T) > HOST_SCRATCH_BYTES) throw std::runtime_error("read-back batch too large");
    char* h = static_cast<char*>(hostScratch) + off;
    if (n) HIPCHECK(hipMemcpyAsync(h, dptr, n * sizeof(T), hipMemcpyDeviceToHost, st));
    off += n * sizeof(T);
    return reinterpret_cast<const T*>(h);
  }
  template <class T>
  void writeN(T* dptr, const T* host, size_t n) {
    if (!n) return;
    HIPCHECK(hipMemcpyAsync(dptr, host, n * sizeof(T), hipMemcpyHostToDevice, st));
    sync();
  }
  template <class T>
  void zero(T* dptr, size_t n) {
    if (n) HIPCHECK(hipMemsetAsync(dptr, 0, n * sizeof(T), st));
  }
  void mark(int stage) {
    if (!timing || !timedNow) return;
    if (stage == ST_PREDICT && stageMarked >= 0) {
      // a substep that ended early (an error) left its events open: recycle them
      for (int k = 0; k <= stageMarked; ++k) evFree.push_back(stageCur[k]);
      stageMarked = -1;
    }
    if (stage != stageMarked + 1) return;
    stageCur[stage] = takeEvent();
    HIPCHECK(hipEventRecord(stageCur[stage], st));
    stageHostT[stage] = std::chrono::steady_clock::now();
    if (stage > 0)
      stageHostMs[stage - 1] +=
          std::chrono::duration<double, std::milli>(stageHostT[stage] - stageHostT[stage - 1])
              .count();
    stageMarked = stage;
    if (stage == ST_COUNT) {
      stagePending.insert(stagePending.end(), stageCur, stageCur + ST_COUNT + 1);
      stageMarked = -1;
    }
  }

  // exclusive scan of n ints into out[0..n), total into out[n]; returns the total
  int scan(const int* in, int* out, int n) {
    scanQuiet(in, out, n);
    return n > 0 ? read1(out + n) : 0;
  }
  // the same scan without reading the total back (no host synchronisation)
  void scanQuiet(const int* in, int* out, int n) {
    if (n <= 0) {
      zero(out, 1);
      return;
    }
    const int nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (nb > 1 && n <= SCAN1_MAX) {
      k_scan_single<<<1, SCAN1_THREADS, 0, st>>>(in, n, out, nullptr);
      HIPCHECK(hipGetLastError());
      return;
    }
    scanTmp.ensure(nb + 1, st);
    if (nb > 1) k_scan_reduce<<<nb, TPB, 0, st>>>(in, n, scanTmp.p, nullptr);
    k_scan_down2<<<nb, TPB, 0, st>>>(in, n, scanTmp.p, out, nullptr);
    HIPCHECK(hipGetLastError());
  }
  // the same scan over a count that lives on the device (*nDev <= nMax)
  void scanDev(const int* in, int* out, const int* nDev, int nMax) {
    const int nb = std::max(1, (nMax + SCAN_TILE - 1) / SCAN_TILE);
    if (nb > 1 && nMax <= SCAN1_MAX) {
      k_scan_single<<<1, SCAN1_THREADS, 0, st>>>(in, nMax, out, nDev);
      HIPCHECK(hipGetLastError());
      return;
    }
    scanTmp.ensure(nb + 1, st);
    if (nb > 1) k_scan_reduce<<<nb, TPB, 0, st>>>(in, nMax, scanTmp.p, nDev);
    k_scan_down2<<<nb, TPB, 0, st>>>(in, nMax, scanTmp.p, out, nDev);
    HIPCHECK(hipGetLastError());
  }

  // ---------------------------------------------------------------- capacity
  void ensureSlots(int n, int keep) {
    pos.ensure(n, st, keep);
    posPred.ensure(n, st, keep);
    quatPred.ensure(n, st, keep);
    extForce.ensure(n, st, keep);
    vw0.ensure(n, st, keep);
    vw.ensure(n, st, keep);
    bc.ensure(n, st, keep);
    scale.ensure(n, st, keep);
    targetScale.ensure(n, st, keep);
    realIdx.ensure(n, st, keep);
    matId.ensure(n, st, keep);
    ground.ensure(n, st, keep);
    ground2.ensure(n, st, keep);
  }

  // a table member is a global-address-space pointer in the device pass (same bits)
  template <class T>
  static void gset(PD_G T*& dst, T* src) {
    dst = (PD_G T*)src;
  }
  void refreshDev() {
    Dev& d = D;
    gset(d.pos, pos.p); gset(d.vw0, vw0.p); gset(d.bc, bc.p); gset(d.posPred, posPred.p); gset(d.quatPred, quatPred.p);
    gset(d.vw, vw.p); gset(d.scale, scale.p); gset(d.targetScale, targetScale.p); gset(d.extForce, extForce.p);
    gset(d.realIdx, realIdx.p); gset(d.matId, matId.p); gset(d.ground, ground.p); gset(d.ground2, ground2.p);
    gset(d.shell, shell.p); gset(d.pairTable, pairTableD.p);
    gset(d.cellOf, cellOf.p); gset(d.cellRank, cellRank.p); gset(d.cellCnt, cellCnt.p); gset(d.cellStart, cellStart.p); gset(d.order, order.p);
    gset(d.sbox, sbox.p); gset(d.pairCnt, pairCnt.p); gset(d.pairStart, pairStart.p); gset(d.pairs, pairs.p);
    gset(d.ncnt, ncnt.p); gset(d.noff, noff.p); gset(d.cAB, cAB.p); gset(d.cN, cN.p); gset(d.cRA, cRA.p); gset(d.cRB, cRB.p);
    gset(d.cSlot, cSlot.p); gset(d.mCnt, mCnt.p); gset(d.mStart, mStart.p); gset(d.mAB, mAB.p); gset(d.mReal, mReal.p);
    gset(d.mRange, mRange.p); gset(d.mRange2, mRange2.p); gset(d.cBV, cBV.p); gset(d.mWV, mWV.p); gset(d.vWV, vWV.p);
    gset(d.mNs, mNs.p); gset(d.mTA, mTA.p); gset(d.mTB, mTB.p); gset(d.mRA, mRA.p);
    gset(d.mRB, mRB.p); gset(d.mNp, mNp.p); gset(d.mKey, mKey.p); gset(d.mCode, mCode.p); gset(d.mColor, mColor.p);
    gset(d.cColor, cColor.p); gset(d.vPerm, vPerm.p);
    gset(d.vRowOf, vRowOf.p); gset(d.pPerm, pPerm.p); gset(d.pRowOf, pRowOf.p); gset(d.vAB, vAB.p); gset(d.vN, vN.p);
    gset(d.vTA, vTA.p); gset(d.vTB, vTB.p); gset(d.vRA, vRA.p); gset(d.vRB, vRB.p); gset(d.vVt0, vVt0.p);
    gset(d.vLamT, vLamT.p); gset(d.vVn0, vVn0.p); gset(d.vLam, vLam.p); gset(d.vFlag, vFlag.p); gset(d.pAB, pAB.p);
    gset(d.pN, pN.p); gset(d.pRA, pRA.p); gset(d.pRB, pRB.p); gset(d.pLam, pLam.p); gset(d.lKey, lKey.p);
    gset(d.lLam, lLam.p); gset(d.lLamT, lLamT.p); gset(d.lIdx, lIdx.p); gset(d.uKey, uKey.p);
    gset(d.bucketCnt, bucketCnt.p); gset(d.bucketStart, bucketStart.p); gset(d.red, red.p);
    gset(d.counters, counters.p);
    gset(d.maxOverlapBits, bits.p); gset(d.maxApproachBits, bits.p + 1); gset(d.thermo, thermo.p);
    gset(d.dVel, dVel.p); gset(d.dAng, dAng.p); gset(d.dPos, dPos.p); gset(d.dCnt, dCnt.p);
    gset(d.cLamN, cLamN.p); gset(d.frCnt, frCnt.p);
    uploadDev();
  }
  // The kernels read the pointer set from device memory (a few bytes of launch arguments instead
  // of ~800): a changed set is copied, in stream order, from a ring of pinned host slots.
  static constexpr int DEV_RING = 16;
  Dev* devD = nullptr;       // device copy read by the kernels
  const Dev* dD = nullptr;   // = devD, the kernels' argument
  Dev* devRing = nullptr;    // pinned host slots
  hipEvent_t devRingEv[DEV_RING] = {};
  int devRingNext = 0;
  Dev devLast{};
  bool devUploaded = false;
  void uploadDev() {
    if (devUploaded && std::memcmp(&devLast, &D, sizeof(Dev)) == 0) return;
    const int k = devRingNext;
    devRingNext = (devRingNext + 1) % DEV_RING;
    HIPCHECK(hipEventSynchronize(devRingEv[k]));  // the slot's previous copy has run
    devRing[k] = D;
    HIPCHECK(hipMemcpyAsync(devD, &devRing[k], sizeof(Dev), hipMemcpyHostToDevice, st));
    HIPCHECK(hipEventRecord(devRingEv[k], st));
    devLast = D;
    devUploaded = true;
  }
  // zeroed Jacobi accumulators for the colour-saturation fallback
  void prepareFallback() {
    dVel.ensure(numParticles + 1, st);
    dAng.ensure(numParticles + 1, st);
    dPos.ensure(numParticles + 1, st);
    dCnt.ensure(numParticles + 1, st);
    refreshDev();
    zero(dVel.p, numParticles);
    zero(dAng.p, numParticles);
    zero(dPos.p, numParticles);
    zero(dCnt.p, numParticles);
  }

  // incidence lists of the elements (manifolds for MODE 0, contacts otherwise) that may add to
  // a body's accumulators: ascending element order per body (slots 0..numParticles)
  template <int MODE>
  void buildIncidence(int numElements, int onlyUncoloured) {
    const int nb = numParticles;
    jIncCnt.ensure(nb + 1, st);
    jIncStart.ensure(nb + 1, st);
    jF.ensure(numElements + 1, st);
    jC.ensure(numElements + 1, st);
    refreshDev();
    zero(jIncCnt.p, nb + 1);
    if (numElements > 0)
      k_jinc_count<MODE><<<nblk(numElements), TPB, 0, st>>>(dD, numElements, onlyUncoloured,
                                                            jIncCnt.p);
    const int total = scan(jIncCnt.p, jIncStart.p, nb);
    jIncList.ensure(total + 1, st);
    zero(jIncCnt.p, nb + 1);
    if (numElements > 0)
      k_jinc_fill<MODE><<<nblk(numElements), TPB, 0, st>>>(dD, numElements, onlyUncoloured,
                                                           jIncStart.p, jIncCnt.p, jIncList.p);
    k_jinc_sort<<<nblk(nb), TPB, 0, st>>>(nb, jIncStart.p, jIncList.p);
    HIPCHECK(hipGetLastError());
  }
  // one Jacobi velocity pass (src/solver_velocity.hpp:38-255) over the incidence lists built
  // for it: contributions, ordered per-body sums, averaged apply
  void jacobiVelocityPass(const Params& PF, double thr, int onlyUncoloured) {
    prepareFallback();
    if (nm > 0)
      k_jvel_contrib<<<nblk(nm), TPB, 0, st>>>(dD, PF, nm, thr, onlyUncoloured, jC.p, jF.p);
    k_jgather<0><<<nblk(numParticles), TPB, 0, st>>>(dD, numParticles, jIncStart.p, jIncList.p,
                                                     jC.p, jF.p);
    k_apply_vel_avg<<<nblk(numParticles), TPB, 0, st>>>(dD, numParticles);
    HIPCHECK(hipGetLastError());
  }
  // one Jacobi position pass (src/solver_position.hpp:34-134, src/integration.hpp:168-184)
  void jacobiPositionPass(int onlyUncoloured) {
    prepareFallback();
    if (nc > 0)
      k_jpos_contrib<<<nblk(nc), TPB, 0, st>>>(dD, nc, onlyUncoloured, jC.p, jF.p);
    k_jgather<1><<<nblk(numParticles), TPB, 0, st>>>(dD, numParticles, jIncStart.p, jIncList.p,
                                                     jC.p, jF.p);
    k_apply_pos_avg<<<nblk(numParticles), TPB, 0, st>>>(dD, numParticles);
    HIPCHECK(hipGetLastError());
  }

  Params params(double margin, double band) const {
    Params P{};
    P.dt = dt;
    P.gs = globalScale;
    P.baseRadius = baseRadius;
    P.margin = margin;
    P.band = band;
    P.boxCap = std::max(std::max(dsize[0], dsize[1]), dsize[2]);
    P.growthRate = growthRate;
    P.restN = restitutionNormal;
    P.restT = restitutionTangent;
    P.mu = frictionDynamic;
    P.g = d3{g[0], g[1], g[2]};
    P.dmin = d3{dmin[0], dmin[1], dmin[2]};
    P.dmax = d3{dmax[0], dmax[1], dmax[2]};
    P.dsize = d3{dsize[0], dsize[1], dsize[2]};
    P.px = per[0];
    P.py = per[1];
    P.pz = per[2];
    P.numReal = numReal;
    P.numParticles = numParticles;
    P.numPlanes = numPlanes;
    P.hasPairTable = hasPairTable;
    P.shape.type = shapeType;
    P.shape.numShell = (int)(shellHost.size() / 3);
    P.shape.p0 = shapeParams[0];
    P.shape.p1 = shapeParams[1];
    P.shape.p2 = shapeParams[2];
    P.shape.p3 = shapeParams[3];
    if (shapeType == SHAPE_GRID_SDF) {
      P.shape.grid = sdfGridD.p;
      P.shape.nx = gridDims[0];
      P.shape.ny = gridDims[1];
      P.shape.nz = gridDims[2];
      P.shape.origin = d3{gridOrigin[0], gridOrigin[1], gridOrigin[2]};
      P.shape.invSpacing = d3{gridInvSpacing[0], gridInvSpacing[1], gridInvSpacing[2]};
    }
    for (int k = 0; k < numPlanes; ++k) {
      P.planeP[k] = make_double4(planeP[k][0], planeP[k][1], planeP[k][2], 0.0);
      P.planeN[k] = make_double4(planeN[k][0], planeN[k][1], planeN[k][2], 0.0);
    }
    return P;
  }

  // ---------------------------------------------------------------- shape
  void uploadShell() {
    const size_t n = shellHost.size() / 3;
    shell.ensure(n > 0 ? n : 1, st);
    std::vector<double4> h(n);
    for (size_t i = 0; i < n; ++i)
      h[i] = make_double4(shellHost[3 * i], shellHost[3 * i + 1], shellHost[3 * i + 2], 0.0);
    writeN(shell.p, h.data(), n);
  }
  // src/sim.hpp:949-1021 + shell generators src/shapes_portable.hpp:21-111 (FP64)
  void initializeShape(int type, double radius, double height, double thickness) {
    baseRadius = radius;
    double prm[4] = {radius, 0, 0, 0};
    std::vector<double> pts;
    auto push = [&](double x, double y, double z) {
      pts.push_back(x);
      pts.push_back(y);
      pts.push_back(z);
    };
    if (type == SHAPE_HOLLOW_CYLINDER) {
      prm[1] = height;
      prm[2] = thickness;
      double min_dim = std::min(radius, thickness);
      if (min_dim < 1e-4) min_dim = radius;
      double spacing = std::min(radius * 0.3, min_dim * 0.5);
      if (spacing < 1e-3) spacing = 1e-3;
      const double r_outer = radius, r_inner = radius - thickness, h = height;
      double circumference = 2.0 * M_PI * r_outer;
      int n_angular = (int)std::ceil(circumference / spacing);
      const int n_vertical = (int)std::ceil(h / spacing);
      for (int i = 0; i < n_angular; ++i) {
        const double theta = 2.0 * M_PI * i / n_angular;
        for (int j = 0; j <= n_vertical; ++j)
          push(r_outer * std::cos(theta), -0.5 * h + h * j / n_vertical, r_outer * std::sin(theta));
      }
      if (thickness > 0.0) {
        circumference = 2.0 * M_PI * r_inner;
        n_angular = (int)std::ceil(circumference / spacing);
        for (int i = 0; i < n_angular; ++i) {
          const double theta = 2.0 * M_PI * i / n_angular;
          for (int j = 0; j <= n_vertical; ++j)
            push(r_inner * std::cos(theta), -0.5 * h + h * j / n_vertical,
                 r_inner * std::sin(theta));
        }
      }
      const int n_radial = (int)std::ceil(thickness / spacing);
      for (int i = 0; i <= n_radial; ++i) {
        const double r = r_inner + (r_outer - r_inner) * i / n_radial;
        circumference = 2.0 * M_PI * r;
        n_angular = (int)std::ceil(circumference / spacing);
        for (int j = 0; j < n_angular; ++j) {
          const double theta = 2.0 * M_PI * j / n_angular;
          const double x = r * std::cos(theta), z = r * std::sin(theta);
          push(x, 0.5 * h, z);
          push(x, -0.5 * h, z);
        }
      }
    } else if (type == SHAPE_BOX) {
      prm[1] = radius;
      prm[2] = radius;
      const double spacing = std::max(radius * 0.5, 1e-3);
      auto nseg = [&](double half) {
        int n = (int)std::ceil(2.0 * half / spacing);
        return n < 1 ? 1 : n;
      };
      auto lin = [](double half, int i, int n) { return -half + 2.0 * half * (double)i / (double)n; };
      const double hx = radius, hy = radius, hz = radius;
      const int nx = nseg(hx), ny = nseg(hy), nz = nseg(hz);
      for (int j = 0; j <= ny; ++j)
        for (int k = 0; k <= nz; ++k) {
          push(hx, lin(hy, j, ny), lin(hz, k, nz));
          push(-hx, lin(hy, j, ny), lin(hz, k, nz));
        }
      for (int i = 0; i <= nx; ++i)
        for (int k = 0; k <= nz; ++k) {
          push(lin(hx, i, nx), hy, lin(hz, k, nz));
          push(lin(hx, i, nx), -hy, lin(hz, k, nz));
        }
      for (int i = 0; i <= nx; ++i)
        for (int j = 0; j <= ny; ++j) {
          push(lin(hx, i, nx), lin(hy, j, ny), hz);
          push(lin(hx, i, nx), lin(hy, j, ny), -hz);
        }
    } else {
      type = SHAPE_SPHERE;
    }
    shapeType = type;
    std::memcpy(shapeParams, prm, sizeof prm);
    shellHost = pts;
    uploadShell();
    double ix = 1.0, iy = 1.0, iz = 1.0;
    if (type == SHAPE_SPHERE) {
      if (baseRadius > 0.0) ix = iy = iz = 2.5 / (baseRadius * baseRadius);
    } else if (type == SHAPE_HOLLOW_CYLINDER) {
      const double r_out = baseRadius;
      double r_in = baseRadius - thickness;
      if (r_in < 0) r_in = 0;
      const double term_r = r_out * r_out + r_in * r_in;
      const double I_zz = 0.5 * term_r;
      const double I_xx = (1.0 / 12.0) * (3.0 * term_r + height * height);
      if (I_xx > 1e-6) ix = iy = 1.0 / I_xx;
      if (I_zz > 1e-6) iz = 1.0 / I_zz;
    } else if (type == SHAPE_BOX) {
      const double L = 2.0 * baseRadius;
      const double I = (1.0 / 6.0) * L * L;
      if (I > 1e-6) ix = iy = iz = 1.0 / I;
    }
    defaultInvI[0] = ix;
    defaultInvI[1] = iy;
    defaultInvI[2] = iz;
  }

  // ---------------------------------------------------------------- packed-field access
  // field ids: 0 pos, 1 vel, 2 angVel, 3 quat, 4 invI, 5 invMass, 6 extForce
  void fieldLayout(int f, double*& base, int& stride, int& off, int& comps) {
    switch (f) {
      case 0: base = (double*)pos.p; stride = 4; off = 0; comps = 3; break;
      case 1: base = (double*)vw0.p; stride = 8; off = 0; comps = 3; break;
      case 2: base = (double*)vw0.p; stride = 8; off = 4; comps = 3; break;
      case 3: base = (double*)bc.p; stride = 8; off = 0; comps = 4; break;
      case 4: base = (double*)bc.p; stride = 8; off = 4; comps = 3; break;
      case 5: base = (double*)bc.p; stride = 8; off = 7; comps = 1; break;
      default: base = (double*)extForce.p; stride = 4; off = 0; comps = 3; break;
    }
  }
  void setField(int f, const double* src, int n) {
    ++paramsVersion;
    n = std::min(n, numReal);
    if (n <= 0) return;
    double* base;
    int stride, off, comps;
    fieldLayout(f, base, stride, off, comps);
    std::vector<double> h((size_t)numReal * stride);
    readN(h.data(), base, h.size());
    for (int i = 0; i < n; ++i)
      for (int c = 0; c < comps; ++c) h[(size_t)i * stride + off + c] = src[(size_t)i * comps + c];
    writeN(base, h.data(), h.size());
  }
  void getField(int f, double* out, int n) {
    n = std::min(n, numReal);
    if (n <= 0) return;
    double* base;
    int stride, off, comps;
    fieldLayout(f, base, stride, off, comps);
    std::vector<double> h((size_t)n * stride);
    readN(h.data(), base, h.size());
    for (int i = 0; i < n; ++i)
      for (int c = 0; c < comps; ++c) out[(size_t)i * comps + c] = h[(size_t)i * stride + off + c];
  }

  // src/sim.hpp:1290-1330
  void setPositions(const double* xyz, int n) {
    ++paramsVersion;
    if (n < 0) throw ArgError("negative particle count");
    // headroom for the periodic ghosts (src/periodicity.hpp:25-38); grown on demand
    const long est = (long)n + (long)n / 4 + slotHeadroom;
    ensureSlots((int)std::max<long>(est, capacity), numReal);
    numReal = n;
    numParticles = n;
    hNumPairs = -1;  // particle indices changed: the Hertz pair cache is stale, :1294-1295
    candValid = false;
    hPrevCount = 0;
    std::vector<double4> P((size_t)n);
    std::vector<BodyC> B((size_t)n);
    std::vector<Body64> V((size_t)n);
    std::vector<double> one((size_t)n, 1.0);
    for (int i = 0; i < n; ++i) {
      P[i] = make_double4(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], 0.0);
      B[i].q = make_double4(0, 0, 0, 1);
      B[i].iim = make_double4(defaultInvI[0], defaultInvI[1], defaultInvI[2], 1.0);
      V[i].v = make_double4(0, 0, 0, 0);
      V[i].w = make_double4(0, 0, 0, 0);
    }
    writeN(pos.p, P.data(), n);
    writeN(bc.p, B.data(), n);
    writeN(vw0.p, V.data(), n);
    writeN(scale.p, one.data(), n);
    scaleKnown = false;
    writeN(targetScale.p, one.data(), n);
  }

  // ---------------------------------------------------------------- the substep
  void broadAndNarrow(double margin) {
    Params P = params(margin, 0.0);
    broadphase(P);
    mark(ST_NARROW);
    narrowphase(P);
  }

  // uniform-grid broadphase over posPred (.w = radius) of numParticles slots with P.margin;
  // fills pairs / pairStart / np
  // Broadphase up to the pair scan, without a read-back: the slot count lives on the device
  // (slotInfo[0]), gridSlots sizes the launches (the slot loops are grid-strided, so any grid is
  // correct) and capSlots bounds the slot count.  The pair total is at pairStart[numReal].
  void broadphaseQuery(const Params& P, const int* nSlotsDev, int gridSlots, int capSlots) {
    const int nbb = std::min(nblk(gridSlots), RED_BLOCKS);
    const long cellCap = 4L * capSlots + 64 + 1;  // the grid keeps cells <= 4 x slots + 64
    redPart.ensure(8 * RED_BLOCKS, st);
    gridD.ensure(1, st);
    ncellD.ensure(1, st);
    cellCnt.ensure(cellCap, st);
    cellStart.ensure(cellCap + 1, st);
    cellOf.ensure(capSlots, st);
    cellRank.ensure(capSlots, st);
    order.ensure(capSlots, st);
    sbox.ensure(capSlots, st);
    pairCnt.ensure(numReal + 1, st);
    pairStart.ensure(numReal + 1, st);
    refreshDev();
    // cellCnt is all zero between broadphases (k_cell_fill clears what k_cell_count set); a pass
    // cut short by an error leaves it dirty
    if (cellCntDirty) zero(cellCnt.p, cellCnt.cap);
    cellCntDirty = true;
    k_bp_bounds<<<nbb, TPB, 0, st>>>(dD, P, nSlotsDev, redPart.p);
    k_bp_bounds_final<<<1, TPB, 0, st>>>(redPart.p, nbb, red.p, nSlotsDev, gridD.p, ncellD.p);
    k_cell_count<<<nblk(gridSlots), TPB, 0, st>>>(dD, P, nSlotsDev, gridD.p);
    scanDev(cellCnt.p, cellStart.p, ncellD.p, (int)cellCap);
    k_cell_fill<<<nblk(gridSlots), TPB, 0, st>>>(dD, P, nSlotsDev);
    HIPCHECK(hipGetLastError());
    cellCntDirty = false;
    pairStage.ensure((size_t)numReal * BP_CAP + 1, st);
    zero(pairCnt.p, numReal);
    kernBegin();
    k_bp_query<<<nblk(gridSlots), TPB, 0, st>>>(dD, P, nSlotsDev, gridD.p, pairStage.p);
    kernEnd(K_BPQUERY, gridSlots);
    scanQuiet(pairCnt.p, pairStart.p, numReal);
  }
  // the pair list from the staged partners, once the pair total np is known
  void broadphaseCompact(const Params& P) {
    pairs.ensure(np + 1, st);
    refreshDev();
    k_bp_compact<<<nblk(numReal), TPB, 0, st>>>(dD, P, gridD.p, pairStage.p);
    HIPCHECK(hipGetLastError());
  }
  // the whole broadphase over the numParticles slots the host already knows; fills pairs / np
  void broadphase(const Params& P) {
    slotInfo.ensure(4, st);
    k_set_int<<<1, 1, 0, st>>>(slotInfo.p, numParticles);
    broadphaseQuery(P, slotInfo.p, numParticles, numParticles);
    np = numReal > 0 ? read1(pairStart.p + numReal) : 0;
    broadphaseCompact(P);
  }
  // per-slot capacity: every array the ghost emitter writes
  int slotCapacity() const {
    size_t c = pos.cap;
    for (size_t x : {posPred.cap, quatPred.cap, bc.cap, vw0.cap, vw.cap, scale.cap, matId.cap,
                     realIdx.cap})
      c = std::min(c, x);
    return (int)std::min<size_t>(c, (size_t)INT32_MAX);
  }
  DBuf<int> slotInfo;  // [0] slot count of the substep, [1] ghosts did not fit
  DBuf<int> lRank;     // ledger: each manifold's place in its bucket
  bool ledgerCntDirty = true;
  // candidate pairs of the last rebuild (boxes grown by candSkin), sorted with per-real offsets,
  // the slots' plain boxes and the reals' ghost images at that rebuild
  DBuf<int2> cand;
  DBuf<int> candStart, candFlag;
  DBuf<unsigned char> candKeep, candCode;
  DBuf<double4> candRef;
  bool candValid = false;
  int candSlots = 0, candNumReal = -1;
  double candSkin = 0.0, candMargin = 0.0;
  double candSkinFrac = 0.1;  // skin = this x the largest radius (0: rebuild every substep)
  long candRebuilds = 0, candReuses = 0;
  int lastGhosts = 0;
  DBuf<Grid> gridD;
  DBuf<int> ncellD;
  bool cellCntDirty = true;

  // narrowphase: counts, scan, write (pairs first, then planes per real)
  WallSet wallSet() const {
    WallSet W{};
    W.n = (int)wallsHost.size();
    for (int k = 0; k < W.n && k < 8; ++k) W.w[k] = wallsHost[k];
    W.grid = wallGridD.p;
    return W;
  }
  // contact slots: pairs, then planes per real, then walls per real (the reference's order)
  // deferNc (spheres without walls): the contact total is not read here but with the manifold
  // total in buildManifolds (one host synchronisation for both); the contact buffers are sized
  // for the bound np + numReal x planes (a sphere touches each plane once, each pair at most
  // once), so the write pass never needs the exact count
  bool ncDeferred = false;
  long npRecNc = -1;  // timing record of the write pass, its units patched once nc is known
  void narrowphase(const Params& P, bool zeroOverlap = true, bool deferNc = false) {
    const int nslots = np + 2 * numReal;
    const WallSet W = wallSet();
    ncnt.ensure(nslots + 1, st);
    noff.ensure(nslots + 1, st);
    refreshDev();
    const bool waveNp = P.shape.numShell > 0;  // shell shapes: one wavefront per pair
    const int npBlocks = (int)std::max<long>(1, ((long)np + NP_WAVES - 1) / NP_WAVES);
    if (waveNp) k_np_pairs_wave<false><<<npBlocks, TPB, 0, st>>>(dD, P, np);
    else k_np_pairs<false><<<nblk(np), TPB, 0, st>>>(dD, P, np);
    // without walls every wall slot holds 0 contacts (written by the planes pass)
    k_np_planes<false><<<nblk(numReal), TPB, 0, st>>>(dD, P, np, W.n > 0 ? -1 : np + numReal);
    if (W.n > 0) k_np_walls<false><<<nblk(numReal), TPB, 0, st>>>(dD, P, W, np + numReal);
    if (waveNp) {
      nc = scan(ncnt.p, noff.p, nslots);
    } else {
      // point shapes: the scan also lists the pairs in contact (their counts are 0 or 1)
      pairOfContact.ensure(np + 1, st);
      const int nb = std::max(1, (nslots + SCAN_TILE - 1) / SCAN_TILE);
      scanTmp.ensure(nb + 1, st);
      if (nb > 1) k_scan_reduce<<<nb, TPB, 0, st>>>(ncnt.p, nslots, scanTmp.p, nullptr);
      k_scan_down2<<<nb, TPB, 0, st>>>(ncnt.p, nslots, scanTmp.p, noff.p, nullptr,
                                       pairOfContact.p, np);
      HIPCHECK(hipGetLastError());
      ncDeferred = deferNc && W.n == 0;
      if (!ncDeferred) nc = read1(noff.p + nslots);
    }
    const size_t ncCap =
        ncDeferred ? (size_t)np + (size_t)numReal * (size_t)std::max(P.numPlanes, 0) : (size_t)nc;
    cAB.ensure(ncCap + 1, st);
    cBV.ensure(ncCap + 1, st);
    cN.ensure(ncCap + 1, st);
    cRA.ensure(ncCap + 1, st);
    cRB.ensure(ncCap + 1, st);
    cSlot.ensure(ncCap + 1, st);
    refreshDev();
    if (zeroOverlap) zero(bits.p, 1);  // a substep clears it in its predict launch
    kernBegin();
    if (waveNp) {
      k_np_pairs_wave<true><<<npBlocks, TPB, 0, st>>>(dD, P, np);
    } else if (ncDeferred || nc > 0) {
      const int grid = std::max(1, std::min(nblk(ncDeferred ? np : nc), 4096));
      k_np_pairs_sparse<<<grid, TPB, 0, st>>>(dD, P, np, pairOfContact.p);
    }
    npRecNc = kernEnd(K_NPPAIRS, waveNp ? np : nc);
    k_np_planes<true><<<nblk(numReal), TPB, 0, st>>>(dD, P, np);
    if (W.n > 0) k_np_walls<true><<<nblk(numReal), TPB, 0, st>>>(dD, P, W, np + numReal);
    HIPCHECK(hipGetLastError());
  }

  void buildManifolds() {
    Params P = params(0.0, 0.0);
    mCnt.ensure(numReal + 1, st);
    mStart.ensure(numReal + 1, st);
    refreshDev();
    k_mcount<<<nblk(numReal), TPB, 0, st>>>(dD, P, np);
    if (ncDeferred) {
      // the manifold total and the deferred contact total in one read
      scanQuiet(mCnt.p, mStart.p, numReal);
      PostBatch rb = postBatch(hostPost, hostPostDev, 64);
      const int* hm = rb.add(mStart.p + numReal, 1);
      const int* hc = rb.add(noff.p + (np + 2 * numReal), 1);
      rb.post(st);
      rb.wait(st);
      nm = numReal > 0 ? *hm : 0;
      nc = *hc;
      ncDeferred = false;
      if (npRecNc >= 0 && npRecNc < (long)kernPending.size()) kernPending[npRecNc].units = nc;
    } else {
      nm = scan(mCnt.p, mStart.p, numReal);
    }
    const size_t M = nm + 1;
    mAB.ensure(M, st); mReal.ensure(M, st); mRange.ensure(M, st); mRange2.ensure(M, st);
    mWV.ensure(M, st); mNs.ensure(M, st);
    mTA.ensure(M, st); mTB.ensure(M, st); mRA.ensure(M, st); mRB.ensure(M, st);
    mNp.ensure(M, st); mKey.ensure(M, st); mCode.ensure(M, st); mColor.ensure(M, st);
    vRowOf.ensure(M, st);
    refreshDev();
    k_mbuild<<<nblk(numReal), TPB, 0, st>>>(dD, P, np);
    k_mreduce<<<nblk(nm), TPB, 0, st>>>(dD, nm);
    HIPCHECK(hipGetLastError());
  }

  static constexpr int JP_MAX_BATCH = 32;
  // incidence rounds expected to list at most this many edges run in one workgroup (k_jpi_tail)
  static constexpr long JPI_TAIL_EDGES = 4096;
  static constexpr long JP_GRID_CAP = 2048;
  // large colourings run their rounds over ordered incidence lists (colourBegin has the rule)
  long incMinEdges = 1L << 16;
  // One colouring in progress: its stream, its own work lists, counters, per-body winner / mask
  // slots, class buffers and pinned read-back area, so the manifold and contact colourings can
  // run on different streams at once.
  struct ColourJob {
    hipStream_t stream = nullptr;
    // rounds expected to hold at most this many edges run in one workgroup out of LDS
    long smallRows = JPS_EDGES * 3 / 4;
    bool noSmall = false;  // this colouring's list did not fit: launch-per-round batches only
    bool globalDirty = true;  // winner / mask hold another colouring's values
    int smallLoad = 1;        // whether the batch's LDS kernel loaded the global slots
    int numBodies = 0;
    int smallAt = -1;      // round of the current batch where the one-workgroup kernel starts
    void* pinned = nullptr;     // read-back area (pinned host memory)
    void* pinnedDev = nullptr;  // its device view
    DBuf<int> work[2], cnt, blockCnt, colCnt, colStart, clsCnt;
    DBuf<unsigned long long> winner, mask;
    std::vector<int> hist;  // list size entering each round, last colouring of this job
    // the colouring being run
    const int2* ends = nullptr;
    int* colour = nullptr;
    int* perm = nullptr;
    int* rowOf = nullptr;
    int numEdges = 0;
    std::vector<int> seen;
    int remaining = 0, rounds = 0, prevRemaining = -1, cur = 0, batch = 0, B = 0;
    long long uncol = 0;
    long maxRounds = 0, entering = 0;
    const int* enteringDev = nullptr;
    bool first = true;
    const int* hostCnt = nullptr;
    const int* hostStarts = nullptr;
    const int* postSeqHost = nullptr;  // the read-back's sequence word and number
    int postSeq = 0;
    // rounds over ordered incidence lists (k_jpi_*) instead of contend / commit rounds
    bool inc = false;
    DBuf<int> iDeg, iStart, iList, iArrived, iLong, iScan;
    DBuf<int2> iSorted, iSlot;
    DBuf<JpBody> iBody;
    DBuf<int4> iWork[2];
    const int* hostEntering = nullptr;  // the entering count in the first read-back (device count)
    long listSize = 0;                  // uncoloured edges entering the next round
    long frontier = 0;                  // edges listed for the next batch's first round
    std::vector<int> fhist, fseen;      // edges coloured per round, last / current colouring
    void release() {
      work[0].release(); work[1].release(); cnt.release(); blockCnt.release(); colCnt.release();
      colStart.release(); winner.release(); mask.release(); clsCnt.release();
      iDeg.release(); iStart.release(); iList.release(); iSorted.release(); iBody.release();
      iArrived.release(); iLong.release(); iScan.release(); iWork[0].release();
      iWork[1].release(); iSlot.release();
    }
  };
  ColourJob jobV, jobP;  // manifold colouring (main stream), contact colouring (side stream)

  // Jones-Plassmann colouring then the colour classes, with one host read per batch of rounds.
  // The uncoloured edges are in job.work[0], their count in *enteringDev (device) or
  // enteringHost.  Round j of a batch reads its list size from cnt[j] and appends the losers to
  // the other list, counting into cnt[j + 1]; every round colours at least the globally
  // highest-priority edge, so rounds after the list empties are no-ops.  The class kernels are
  // enqueued behind every batch, so the batch that finishes the colouring also brings back the
  // class starts.  The batch and grid sizes come from the job's previous colouring (the round
  // kernels loop over their lists, so any grid is correct); the host replays the reference loop
  // (src/solver_velocity.hpp:305-355) on the counts.
  void colourBegin(ColourJob& J, const int2* ends, int* colour, int numEdges, int numBodies,
                   const int* enteringDev, int enteringHost, int* perm, int* rowOf) {
    J.winner.ensure(numBodies + 1, st);
    J.mask.ensure(numBodies + 1, st);
    J.cnt.ensure(JP_MAX_BATCH + 2, st);  // (the last slot: the incidence path's long-list count)
    J.blockCnt.ensure(64 * CP_BLOCKS, st);
    J.colCnt.ensure(64, st);
    J.colStart.ensure(65, st);
    J.numBodies = numBodies;
    J.globalDirty = true;  // cleared lazily: a colouring run wholly in LDS never reads them
    J.ends = ends;
    J.colour = colour;
    J.perm = perm;
    J.rowOf = rowOf;
    J.numEdges = numEdges;
    J.seen.clear();
    J.remaining = 1;  // the reference starts with remaining = 1: one round always runs
    J.rounds = 0;
    J.uncol = 0;
    J.prevRemaining = -1;
    J.cur = 0;
    J.maxRounds = (long)numBodies + 2;
    J.batch = std::min<int>(JP_MAX_BATCH, std::max<int>(4, (int)J.hist.size() + 2));
    J.entering = enteringDev ? -1 : enteringHost;
    J.enteringDev = enteringDev;
    J.first = true;
    J.noSmall = false;
    // the incidence path pays passes over every body: it is taken for colourings whose entering
    // edges (this one's for a host count, else the job's last colouring's) number at least
    // incMinEdges and a quarter of the bodies; incMinEdges = 0 takes it for every colouring
    const long estEntering =
        enteringDev ? (J.hist.empty() ? (long)numEdges : (long)J.hist[0]) : (long)enteringHost;
    J.inc = incMinEdges == 0 || (estEntering >= incMinEdges && 4 * estEntering >= numBodies);
    J.fseen.clear();
    if (J.inc) {
      J.globalDirty = true;  // the masks are left as this colouring's, the winner slots untouched
      J.hostEntering = nullptr;
      J.batch = std::min<int>(JP_MAX_BATCH, std::max<int>(4, (int)J.fhist.size() + 2));
      colourLaunchInc(J);
    } else {
      colourLaunch(J);
    }
  }
  // Enqueue one batch of incidence-list rounds (k_jpi_*), the class kernels and the read-back of
  // the per-round counts and class starts.  The first batch builds the lists: incidence counts,
  // their scan, the unordered lists, then the per-body sorts, which also list round 0's edges.
  // cnt[j] is the size of round j's list; the host replays the reference loop on the counts.
  void colourLaunchInc(ColourJob& J) {
    hipStream_t s = J.stream;
    const int B = (int)std::min<long>(J.batch, J.maxRounds - J.rounds);
    J.B = B;
    int* longCnt = J.cnt.p + JP_MAX_BATCH + 1;
    auto jblk = [](long n) { return n > 0 ? (int)((n + JPI_THREADS - 1) / JPI_THREADS) : 1; };
    if (J.first) {
      const int nb = J.numBodies;
      const long ne = J.numEdges;
      J.iDeg.ensure(nb + 1, s);
      J.iStart.ensure(nb + 1, s);
      J.iBody.ensure(nb + 1, s);
      J.iLong.ensure(nb + 1, s);
      J.iList.ensure(2 * ne + 1, s);
      J.iSorted.ensure(2 * ne + 1, s);
      J.iSlot.ensure(ne + 1, s);
      J.iArrived.ensure(ne + 1, s);
      J.iWork[0].ensure(ne + 1, s);
      J.iWork[1].ensure(ne + 1, s);
      const int nbs = std::max(1, (nb + SCAN_TILE - 1) / SCAN_TILE);
      J.iScan.ensure(nbs + 1, s);
      const int gridE = (int)std::min<long>(nblk(ne), JP_GRID_CAP);
      const int* work0 = J.work[0].p;
      k_fill_i32<<<nblk(nb), TPB, 0, s>>>(J.iDeg.p, 0, nb);
      k_jpi_count<<<gridE, TPB, 0, s>>>(J.ends, work0, J.enteringDev, (int)J.entering, J.iDeg.p,
                                        J.iSlot.p, J.iArrived.p);
      if (nbs > 1) k_scan_reduce<<<nbs, TPB, 0, s>>>(J.iDeg.p, nb, J.iScan.p, nullptr);
      k_scan_down2<<<nbs, TPB, 0, s>>>(J.iDeg.p, nb, J.iScan.p, J.iStart.p, nullptr);
      k_jpi_fill<<<gridE, TPB, 0, s>>>(J.ends, work0, J.enteringDev, (int)J.entering,
                                       J.iStart.p, J.iSlot.p, J.iList.p);
      k_jp_counts<<<1, 64, 0, s>>>(J.cnt.p, nullptr, 0, JP_MAX_BATCH + 2);
      J.cur = 0;
      k_jpi_sort<<<jblk(nb), JPI_THREADS, 0, s>>>(J.ends, J.iStart.p, J.iList.p, J.iSorted.p, nb,
                                                  J.iBody.p, J.iArrived.p, J.iWork[0].p, J.cnt.p,
                                                  J.iLong.p, longCnt);
      k_jpi_sort_long<<<JPI_LONG_GRID, TPB, 0, s>>>(J.ends, J.iStart.p, J.iList.p, J.iSorted.p,
                                                    J.iLong.p, longCnt, J.iArrived.p,
                                                    J.iWork[0].p, J.cnt.p);
    } else {
      k_jp_counts<<<1, 64, 0, s>>>(J.cnt.p, nullptr, (int)J.frontier, B + 1);
    }
    for (int j = 0; j < B; ++j) {
      const int r = J.rounds + j;
      long est = std::max<long>(1, J.numEdges / 8);
      if (r < (int)J.fhist.size()) est = J.fhist[r];
      else if (j == 0 && !J.first) est = J.frontier;
      if (est <= JPI_TAIL_EDGES && r < (int)J.fhist.size()) {
        // the rest of the batch is expected to hold short lists: one workgroup runs them
        k_jpi_tail<<<1, JPI_THREADS, 0, s>>>(J.iWork[(J.cur + j) & 1].p,
                                             J.iWork[(J.cur + j + 1) & 1].p, J.cnt.p, j, B,
                                             J.colour, J.iSorted.p, J.iBody.p, J.iArrived.p);
        break;
      }
      const int grid = (int)std::min<long>(jblk(est + est / 4 + 64), JP_GRID_CAP / 4);
      k_jpi_round<<<grid, JPI_THREADS, 0, s>>>(J.iWork[(J.cur + j) & 1].p, J.cnt.p + j, J.colour,
                                               J.iSorted.p, J.iBody.p, J.iArrived.p,
                                               J.iWork[(J.cur + j + 1) & 1].p, J.cnt.p + j + 1);
    }
    const int nbc = std::min(nblk(J.numEdges), CP_BLOCKS);
    k_color_count<<<nbc, TPB, 0, s>>>(J.colour, J.numEdges, J.blockCnt.p);
    k_color_offsets<<<64, TPB, 0, s>>>(J.blockCnt.p, nbc, J.colCnt.p);
    k_color_starts<<<1, 64, 0, s>>>(J.colCnt.p, J.colStart.p);
    k_color_place<<<nbc, TPB, 0, s>>>(J.colour, J.numEdges, J.blockCnt.p, J.colStart.p, J.perm,
                                      J.rowOf);
    HIPCHECK(hipGetLastError());
    PostBatch rb = postBatch(J.pinned, J.pinnedDev, 256);
    J.hostCnt = rb.add(J.cnt.p, B + 1);
    J.hostStarts = rb.add(J.colStart.p, 65);
    if (J.first && J.enteringDev) J.hostEntering = rb.add(J.enteringDev, 1);
    rb.post(s);
    J.postSeqHost = rb.seqHost;
    J.postSeq = rb.seq;
    if (J.first) J.listSize = J.enteringDev ? -1 : J.entering;
    J.first = false;
  }
  // enqueue one batch of rounds, the class kernels and the read-back of counts and class starts
  void colourLaunch(ColourJob& J) {
    hipStream_t s = J.stream;
    int B = (int)std::min<long>(J.batch, J.maxRounds - J.rounds);
    // the list size expected at round j of this batch (from the last colouring of this job)
    auto estimate = [&](int j) {
      const int r = J.rounds + j;
      long est = J.numEdges;
      if (r < (int)J.hist.size()) est = J.hist[r];
      else if (j == 0 && J.entering >= 0) est = J.entering;
      else if (!J.hist.empty()) est = J.hist.back();
      return est;
    };
    J.smallAt = -1;
    if (!J.noSmall)
      for (int j = 0; j < B; ++j)
        if (estimate(j) <= J.smallRows) {
          // rounds from j on are expected to be small: one workgroup runs them out of LDS, so the
          // batch may as well run to the cap
          J.smallAt = j;
          B = (int)std::min<long>(JP_MAX_BATCH, J.maxRounds - J.rounds);
          break;
        }
    J.B = B;
    const bool launchRounds = J.smallAt != 0;  // some rounds of this batch are launched one by one
    // a colouring run whole by the one-workgroup kernel also sorts its classes there (if it does
    // not finish, the next batch runs the class kernels over every edge)
    const bool place = J.first && J.smallAt == 0;
    const int loadGlobal = J.globalDirty ? 0 : 1;
    if (launchRounds && J.globalDirty) {
      k_zero2_u64<<<nblk(J.numBodies), TPB, 0, s>>>(J.mask.p, J.winner.p, J.numBodies);
    }
    // after this batch the global slots are valid for every body a later round can touch: zeroed
    // above, or written back by the LDS kernel for the bodies of the edges it leaves
    J.globalDirty = false;
    // (a colouring that starts in the one-workgroup kernel sets its counters there)
    if (!(place && J.enteringDev))
      k_jp_counts<<<1, 64, 0, s>>>(J.cnt.p, J.first ? J.enteringDev : nullptr, (int)J.entering,
                                   B + 1);
    for (int j = 0; j < B; ++j) {
      const long est = estimate(j);
      int* work = J.work[(J.cur + j) & 1].p;
      int* next = J.work[(J.cur + j + 1) & 1].p;
      if (j == J.smallAt) {
        J.smallLoad = j == 0 ? loadGlobal : 1;
        k_jp_small<<<1, JPS_THREADS, 0, s>>>(J.ends, J.colour, work, next, J.cnt.p, j, B,
                                             J.winner.p, J.mask.p, J.smallLoad, place ? 1 : 0,
                                             J.perm, J.rowOf, J.colStart.p,
                                             place ? J.enteringDev : nullptr);
        break;
      }
      const int grid = (int)std::min<long>(nblk(est + est / 4 + 64), JP_GRID_CAP);
      k_jp_contend<<<grid, TPB, 0, s>>>(J.ends, work, J.cnt.p + j, J.winner.p);
      k_jp_commit<<<grid, TPB, 0, s>>>(J.ends, J.colour, work, J.cnt.p + j, J.winner.p, J.mask.p,
                                       next, J.cnt.p + j + 1);
    }
    if (!place) {
      const int nbc = std::min(nblk(J.numEdges), CP_BLOCKS);
      k_color_count<<<nbc, TPB, 0, s>>>(J.colour, J.numEdges, J.blockCnt.p);
      k_color_offsets<<<64, TPB, 0, s>>>(J.blockCnt.p, nbc, J.colCnt.p);
      k_color_starts<<<1, 64, 0, s>>>(J.colCnt.p, J.colStart.p);
      k_color_place<<<nbc, TPB, 0, s>>>(J.colour, J.numEdges, J.blockCnt.p, J.colStart.p, J.perm,
                                        J.rowOf);
    }
    HIPCHECK(hipGetLastError());
    PostBatch rb = postBatch(J.pinned, J.pinnedDev, 256);
    J.hostCnt = rb.add(J.cnt.p, B + 1);
    J.hostStarts = rb.add(J.colStart.p, 65);
    rb.post(s);
    J.postSeqHost = rb.seqHost;
    J.postSeq = rb.seq;
    J.first = false;
  }
  // wait for the job's read-back and replay its rounds; enqueue further batches until done.
  // Returns the uncoloured leftovers after the round cap.
  int colourFinish(ColourJob& J, std::vector<int>& starts, int& numColors, int& rounds,
                   long long& uncol) {
    while (J.inc) {
      waitPosted(J.postSeqHost, J.postSeq, J.stream);
      if (J.listSize < 0) J.listSize = *J.hostEntering;
      for (int j = 0; j < J.B && J.remaining > 0; ++j) {
        const long size = J.listSize;
        J.uncol += size;
        J.seen.push_back((int)size);
        J.fseen.push_back(J.hostCnt[j]);
        ++J.rounds;
        const long rem = size - J.hostCnt[j];
        if (rem == J.prevRemaining && rem > 0)
          throw std::runtime_error("colouring stalled: a round coloured no edge");
        J.prevRemaining = (int)rem;
        J.remaining = (int)rem;
        J.listSize = rem;
      }
      J.frontier = J.hostCnt[J.B];
      J.entering = J.remaining;
      J.cur = (J.cur + J.B) & 1;
      if (J.remaining <= 0 || J.rounds >= J.maxRounds) {
        starts.assign(J.hostStarts, J.hostStarts + 65);
        J.fhist.swap(J.fseen);
        break;
      }
      J.batch = std::min(J.batch * 2, JP_MAX_BATCH);
      colourLaunchInc(J);
    }
    for (; !J.inc;) {
      waitPosted(J.postSeqHost, J.postSeq, J.stream);
      // a small-colouring kernel that could not take its list changed nothing: replay the
      // rounds before it and continue with launch-per-round batches
      const bool failed = J.smallAt >= 0 && J.hostCnt[J.smallAt + 1] < 0;
      const int B = failed ? J.smallAt : J.B;
      if (failed) {
        // it changed nothing: global slots it would have started from zero are still stale
        if (J.smallAt == 0 && J.smallLoad == 0) J.globalDirty = true;
        J.noSmall = true;
        J.smallRows = std::max<long>(256, J.smallRows / 2);  // lists of this job are too wide
      }
      for (int j = 0; j < B && J.remaining > 0; ++j) {
        J.uncol += J.hostCnt[j];
        J.seen.push_back(J.hostCnt[j]);
        ++J.rounds;
        const int rem = J.hostCnt[j + 1];
        if (rem == J.prevRemaining && rem > 0)
          throw std::runtime_error("colouring stalled: a round coloured no edge");
        J.prevRemaining = rem;
        J.remaining = rem;
      }
      J.entering = J.hostCnt[B];
      J.cur = (J.cur + B) & 1;
      if (!failed && (J.remaining <= 0 || J.rounds >= J.maxRounds)) {
        starts.assign(J.hostStarts, J.hostStarts + 65);
        break;
      }
      if (!failed) J.batch = std::min(J.batch * 2, JP_MAX_BATCH);
      colourLaunch(J);
    }
    J.hist.swap(J.seen);
    // the last class may hold rows sharing a body, which run one after another: its rows go in
    // edge order
    if (starts[63] > starts[62]) {
      J.clsCnt.ensure(CLS_BLOCKS, J.stream);
      const int nb = std::max(1, std::min(CLS_BLOCKS, nblk(J.numEdges)));
      k_class_chunk_count<<<nb, 1024, 0, J.stream>>>(J.colour, J.numEdges, 62, J.clsCnt.p);
      k_class_chunk_place<<<nb, 1024, 0, J.stream>>>(J.colour, J.numEdges, 62, starts[62],
                                                     J.clsCnt.p, J.perm, J.rowOf);
      HIPCHECK(hipGetLastError());
    }
    int maxc = -1;
    for (int c = 0; c < 64; ++c)
      if (starts[c + 1] > starts[c]) maxc = c;
    numColors = maxc + 1;
    rounds = J.rounds;
    uncol = J.uncol;
    return J.remaining > 0 ? (int)J.entering : 0;
  }

  // Enqueue up to `iters` device-gated sweeps in geometric batches (1, 2, 4, ...), reading the
  // stop flag once per batch so a loop that converges early does not pay for launching the rest.
  template <class F>
  void sweepLoop(int iters, const int* stop, F&& one) {
    int done = 0, batch = 1;
    while (done < iters) {
      const int b = std::min(batch, iters - done);
      for (int j = 0; j < b; ++j) one();
      done += b;
      if (done < iters && read1(stop)) break;
      batch *= 2;
    }
  }
  // stabilization mode 4, src/sim.hpp:356-397: levels, (level, colour) buckets, ordered passes
  DBuf<int> hLevel, lvlKey, lvlCnt, lvlStart, lvlPerm, lvlChanged;
  DBuf<int4> lvlBuckets;
  void orderedStabilization(d3 gHat, double vRestS) {
    const int n = numReal, rows = nRowsV;
    hLevel.ensure(n + 1, st);
    lvlChanged.ensure(8, st);
    k_fill_i32<<<nblk(n), TPB, 0, st>>>(hLevel.p, LEVEL_INF, n);
    for (int s = 0; s < 1024; s += 8) {
      zero(lvlChanged.p, 8);
      for (int j = 0; j < 8; ++j)
        k_height_sweep<<<nblk(nm), TPB, 0, st>>>(dD, nm, gHat, hLevel.p, lvlChanged.p + j);
      int ch[8];
      readN(ch, lvlChanged.p, 8);
      bool settled = false;
      for (int j = 0; j < 8; ++j) settled |= (ch[j] == 0);
      if (settled) break;
    }
    const int NK = 1024 * 64;
    lvlKey.ensure(rows + 1, st);
    lvlPerm.ensure(rows + 1, st);
    lvlCnt.ensure(NK + 1, st);
    lvlStart.ensure(NK + 1, st);
    zero(lvlCnt.p, NK + 1);
    if (rows > 0) k_level_keys<<<nblk(rows), TPB, 0, st>>>(dD, rows, hLevel.p, lvlKey.p, lvlCnt.p);
    scanQuiet(lvlCnt.p, lvlStart.p, NK);
    zero(lvlCnt.p, NK + 1);
    if (rows > 0)
      k_level_place<<<nblk(rows), TPB, 0, st>>>(rows, lvlKey.p, lvlStart.p, lvlCnt.p, lvlPerm.p);
    std::vector<int> starts(NK + 1);
    readN(starts.data(), lvlStart.p, NK + 1);
    std::vector<int4> bk;
    for (int k = 0; k < NK; ++k)
      if (starts[k + 1] > starts[k]) bk.push_back(make_int4(starts[k], starts[k + 1], k & 63, 0));
    const int nb = (int)bk.size();
    Params P = params(0.0, 0.0);
    if (rows <= blockRows) {
      lvlBuckets.ensure(nb + 1, st);
      if (nb) writeN(lvlBuckets.p, bk.data(), nb);
      zero(ctl.p + CTL_SWEEPS_S, 1);
      k_ordered_block<<<1, VBLOCK_THREADS, 0, st>>>(dD, P, lvlPerm.p, lvlBuckets.p, nb,
                                                          2 * velIters, vRestS,
                                                          ctl.p + CTL_SWEEPS_S);
      stats.sweeps_stab = read1(ctl.p + CTL_SWEEPS_S);
      maxApproach = readMaxApproach();
      return;
    }
    for (int it = 0; it < 2 * velIters; ++it) {
      zero(bits.p + 1, 1);
      for (int pass = 0; pass < 2; ++pass)
        for (int q = 0; q < nb; ++q) {
          const int4 b = bk[pass == 0 ? q : nb - 1 - q];
          k_vsweep_perm<<<nblk(b.y - b.x), TPB, 0, st>>>(dD, P, lvlPerm.p, b.x, b.y, vRestS,
                                                          b.z == 62 ? 1 : 0);
        }
      stats.sweeps_stab++;
      maxApproach = readMaxApproach();
      if (maxApproach <= vRestS) break;
    }
  }
  // stabilization mode 2 ("multilevel"), src/sim.hpp:303-355 with src/solver_multilevel.hpp:
  // the hierarchy is built once, then each iteration runs one fine sweep (tracking the
  // quasi-static residual) and the coarse cycle, until that residual is at most vRestS.
  // packed body records of the velocity stage (large beds), active between pack and unpack
  DBuf<VBody> vbody;
  VBody* vbActive = nullptr;
  bool packBodies = true;
  bool vbUsedLast = false;     // the previous substep's velocity stage used the records
  bool vbFromPredict = false;  // this substep's predict launch wrote them
  DBuf<unsigned char> mlElig, mlSlot;
  DBuf<int> mlGrp, mlMate, mlLead, mlScan, mlParent, mlCol, mlPerm, mlRowOfTmp, mlCnt;
  DBuf<int> mlBlockCnt, mlColCnt, mlColStart, mlClsCnt;
  DBuf<int2> mlChild;
  DBuf<double> mlMassG, mlInvMassG;
  DBuf<double4> mlVelG, mlVelG0, mlMom;
  DBuf<unsigned long long> mlWinner, mlMask;
  int mlGates = ML_GATE_SLIP;
  struct MlLevel {
    int groupOff, numGroups, parentOff;
    std::vector<int> starts;  // colour classes of the level's crossing rows (65 entries)
  };
  MlDev mlDev(int level) {
    MlDev m{};
    m.elig = mlElig.p;
    m.slot = mlSlot.p + (size_t)level * nRowsV;
    m.grp = mlGrp.p;
    m.winner = mlWinner.p;
    m.mate = mlMate.p;
    m.lead = mlLead.p;
    m.scan = mlScan.p;
    m.parent = mlParent.p;
    m.child = mlChild.p;
    m.massG = mlMassG.p;
    m.invMassG = mlInvMassG.p;
    m.velG = mlVelG.p;
    m.velG0 = mlVelG0.p;
    m.mom = mlMom.p;
    m.mask = mlMask.p;
    m.col = mlCol.p;
    return m;
  }
  // buildContactHierarchyKokkos, src/solver_multilevel.hpp:126-347
  std::vector<MlLevel> mlBuild(double qsThr, d3 gHat) {
    std::vector<MlLevel> L;
    const int n = numReal, R = nRowsV;
    if (nm <= 0 || n <= 0 || R <= 0) return L;
    // the pools are sized by the capacity, as the reference's (src/particles.hpp:244-248); a
    // level needs room for the finer level's parents and its own groups
    const long parentCap = 5L * capacity, groupCap = 4L * capacity;
    const size_t parentNeed = std::min<long>(parentCap, 10L * n) + 1;
    const size_t groupNeed = std::min<long>(groupCap, 10L * n) + 1;
    mlElig.ensure(R + 1, st);
    mlSlot.ensure((size_t)ML_MAX_LEVELS * R + 1, st);
    mlGrp.ensure(n + 1, st);
    mlMate.ensure(n + 1, st);
    mlLead.ensure(n + 1, st);
    mlScan.ensure(n + 1, st);
    mlWinner.ensure(n + 1, st);
    mlMask.ensure(n + 1, st);
    mlParent.ensure(parentNeed, st);
    mlChild.ensure(groupNeed, st);
    mlMassG.ensure(groupNeed, st);
    mlInvMassG.ensure(groupNeed, st);
    mlVelG.ensure(groupNeed, st);
    mlVelG0.ensure(groupNeed, st);
    mlMom.ensure(n + groupNeed, st);
    mlCol.ensure(nm + 1, st);
    mlRowOfTmp.ensure(nm + 1, st);
    mlCnt.ensure(64, st);
    mlBlockCnt.ensure(64 * CP_BLOCKS, st);
    mlColCnt.ensure(64, st);
    mlColStart.ensure(65, st);
    mlClsCnt.ensure(CLS_BLOCKS, st);
    k_ml_elig<<<nblk(R), TPB, 0, st>>>(dD, mlDev(0), R, qsThr, mlGates, gHat);
    k_fill_u8<<<nblk((long)ML_MAX_LEVELS * R), TPB, 0, st>>>(mlSlot.p, ML_SKIP,
                                                             (long)ML_MAX_LEVELS * R);
    k_ml_iota<<<nblk(n), TPB, 0, st>>>(mlGrp.p, n);
    int ngPrev = n, parentOff = 0, groupOff = 0;
    for (int lvl = 1; lvl <= ML_MAX_LEVELS; ++lvl) {
      if ((long)parentOff + ngPrev > parentCap) break;
      const MlDev m = mlDev(lvl - 1);
      // matching
      k_fill_u64<<<nblk(ngPrev), TPB, 0, st>>>(mlWinner.p, 0ull, ngPrev);
      k_ml_iota<<<nblk(ngPrev), TPB, 0, st>>>(mlMate.p, ngPrev);
      k_ml_match_contend<<<nblk(R), TPB, 0, st>>>(dD, m, R);
      k_ml_match_commit<<<nblk(R), TPB, 0, st>>>(dD, m, R);
      k_ml_lead<<<nblk(ngPrev), TPB, 0, st>>>(m, ngPrev);
      scanQuiet(mlLead.p, mlScan.p, ngPrev);
      const int ngNew = read1(mlScan.p + ngPrev);
      if (ngNew >= ngPrev || ngNew > (9 * ngPrev) / 10) break;
      if ((long)groupOff + ngNew > groupCap) break;
      k_ml_parent<<<nblk(ngPrev), TPB, 0, st>>>(m, ngPrev, parentOff, groupOff);
      k_ml_compose<<<nblk(n), TPB, 0, st>>>(m, n, parentOff);
      k_ml_mass<<<nblk(ngNew), TPB, 0, st>>>(dD, m, ngNew, groupOff,
                                             lvl == 1 ? -1 : L.back().groupOff);
      // colouring of the crossing rows: at most 64 rounds, run in batches whose loser counts
      // are read once per batch (a round after the last contender changes nothing)
      k_fill_u64<<<nblk(ngNew), TPB, 0, st>>>(mlMask.p, 0ull, ngNew);
      int round = 0, remaining = 1, prevRemaining = -1, batch = 1;
      while (round < 64 && remaining > 0) {
        const int B = std::min(batch, 64 - round);
        zero(mlCnt.p, B);
        for (int j = 0; j < B; ++j) {
          k_fill_u64<<<nblk(ngNew), TPB, 0, st>>>(mlWinner.p, 0ull, ngNew);
          k_ml_col_contend<<<nblk(R), TPB, 0, st>>>(dD, m, R);
          k_ml_col_commit<<<nblk(R), TPB, 0, st>>>(dD, m, R, mlCnt.p + j);
        }
        int cnt[64];
        readN(cnt, mlCnt.p, B);
        bool stop = false;
        for (int j = 0; j < B && !stop; ++j) {
          ++round;
          // a round that left as many losers as the last one ends the colouring (the
          // reference's stall rule); that cannot happen while a contender is left, since the
          // largest key always wins
          if (cnt[j] == prevRemaining) stop = true;
          prevRemaining = cnt[j];
          remaining = cnt[j];
          if (remaining <= 0) stop = true;
        }
        if (stop) break;
        batch *= 2;
      }
      // the level's rows by colour (colour 62, whose rows may share a group, in manifold order)
      MlLevel lv{groupOff, ngNew, parentOff, {}};
      mlPerm.ensure((size_t)lvl * nm + 1, st, (size_t)(lvl - 1) * nm);
      int* perm = mlPerm.p + (size_t)(lvl - 1) * nm;
      k_ml_col_of<<<nblk(nm), TPB, 0, st>>>(dD, m, nm);
      const int nbc = std::min(nblk(nm), CP_BLOCKS);
      k_color_count<<<nbc, TPB, 0, st>>>(mlCol.p, nm, mlBlockCnt.p);
      k_color_offsets<<<64, TPB, 0, st>>>(mlBlockCnt.p, nbc, mlColCnt.p);
      k_color_starts<<<1, 64, 0, st>>>(mlColCnt.p, mlColStart.p);
      k_color_place<<<nbc, TPB, 0, st>>>(mlCol.p, nm, mlBlockCnt.p, mlColStart.p, perm,
                                         mlRowOfTmp.p);
      lv.starts.resize(65);
      readN(lv.starts.data(), mlColStart.p, 65);
      if (lv.starts[63] > lv.starts[62]) {
        const int nb = std::max(1, std::min(CLS_BLOCKS, nblk(nm)));
        k_class_chunk_count<<<nb, 1024, 0, st>>>(mlCol.p, nm, 62, mlClsCnt.p);
        k_class_chunk_place<<<nb, 1024, 0, st>>>(mlCol.p, nm, 62, lv.starts[62], mlClsCnt.p,
                                                 perm, mlRowOfTmp.p);
      }
      HIPCHECK(hipGetLastError());
      L.push_back(std::move(lv));
      parentOff += ngPrev;
      groupOff += ngNew;
      ngPrev = ngNew;
      if (ngNew <= 32) break;
    }
    return L;
  }
  // multilevelCoarseCycleKokkos, src/solver_multilevel.hpp:354-483 (2 coarse sweeps per level)
  void mlCoarseCycle(const std::vector<MlLevel>& L, unsigned long long* qsBits) {
    const int n = numReal;
    k_ml_iota<<<nblk(n), TPB, 0, st>>>(mlGrp.p, n);
    for (int lvl = 1; lvl <= (int)L.size(); ++lvl) {
      const MlLevel& lv = L[lvl - 1];
      const MlDev m = mlDev(lvl - 1);
      const int off = lv.groupOff, ng = lv.numGroups;
      k_ml_compose<<<nblk(n), TPB, 0, st>>>(m, n, lv.parentOff);
      k_ml_mom0<<<nblk(n), TPB, 0, st>>>(dD, m, n);
      for (int l = 1; l <= lvl; ++l)
        k_ml_tree<<<nblk(L[l - 1].numGroups), TPB, 0, st>>>(
            m, L[l - 1].numGroups, L[l - 1].groupOff, l == 1 ? 0 : n + L[l - 2].groupOff, n);
      k_ml_restrict<<<nblk(ng), TPB, 0, st>>>(m, ng, off, n);
      const int* perm = mlPerm.p + (size_t)(lvl - 1) * nm;
      for (int s = 0; s < 2; ++s)
        for (int c = 0; c < 63; ++c) {
          const int b = lv.starts[c], e = lv.starts[c + 1];
          if (e <= b) continue;
          if (c == 62)
            k_ml_coarse_serial<<<1, 64, 0, st>>>(dD, m, perm, b, e, off, qsBits);
          else
            k_ml_coarse<<<nblk(e - b), TPB, 0, st>>>(dD, m, perm, b, e, off, qsBits);
        }
      k_ml_prolong<<<nblk(n), TPB, 0, st>>>(dD, m, n, off);
    }
    HIPCHECK(hipGetLastError());
  }
  void multilevelStabilization(const std::vector<int>& vs, int numColors, d3 gHat,
                               double vRestS, double qsThr) {
    const std::vector<MlLevel> L = mlBuild(qsThr, gHat);
    stats.ml_levels = (int)L.size();
    unsigned long long* qsBits = bits.p + 3;
    Params P = params(0.0, 0.0);
    for (int it = 0; it < 2 * velIters; ++it) {
      zero(qsBits, 1);
      for (int c = 0; c < numColors; ++c) {
        const int b = vs[c], e = vs[c + 1];
        if (e <= b) continue;
        if (c == 62)
          k_vsweep_qs_serial<<<1, 64, 0, st>>>(dD, P, b, e, vRestS, qsBits);
        else
          k_vsweep_qs<<<nblk(e - b), TPB, 0, st>>>(dD, P, b, e, vRestS, qsBits);
      }
      stats.sweeps_stab++;
      if (!L.empty()) mlCoarseCycle(L, qsBits);
      HIPCHECK(hipGetLastError());
      if (bitsd(read1(qsBits)) <= vRestS) break;
    }
  }
  static ClassStarts classStarts(const std::vector<int>& s) {
    ClassStarts c{};
    for (int k = 0; k < 65 && k < (int)s.size(); ++k) c.s[k] = s[k];
    return c;
  }
  double readMaxApproach() {
    return bitsd(read1(bits.p + 1));
  }
  void velocitySweep(const std::vector<int>& vs, int numColors, int mode, double thr,
                     const int* stop) {
    Params P = params(0.0, 0.0);
    // trailing colours holding at most velTailRows rows together run in one workgroup launch
    // (with no last class 62, which runs after every other class)
    int cTail = numColors;
    if (numColors <= 62)
      while (cTail > 0 && vs[numColors] - vs[cTail - 1] <= velTailRows) --cTail;
    for (int c = 0; c < cTail; ++c) {
      const int b = vs[c], e = vs[c + 1];
      if (e <= b) continue;
      VBody* vb = vbActive;
      if (c == 62 && vWaves62 > 0) {
        if (vb)
          k_vsweep_waves<true><<<1, VBLOCK_THREADS, 0, st>>>(dD, P, vWavePerm.p, vWaveStart.p,
                                                             vWaves62, mode, thr, stop, vb);
        else
          k_vsweep_waves<false><<<1, VBLOCK_THREADS, 0, st>>>(dD, P, vWavePerm.p, vWaveStart.p,
                                                              vWaves62, mode, thr, stop, nullptr);
      } else if (c == 62) {
        k_vsweep_serial<<<1, 1, 0, st>>>(dD, P, b, e, mode, thr, stop, vb);
      } else {
        if (mode == 0) {
          if (vb)
            k_vsweep<0, true><<<xcdGrid(e - b), TPB, 0, st>>>(dD, P, b, e, thr, stop, vb);
          else
            k_vsweep<0><<<xcdGrid(e - b), TPB, 0, st>>>(dD, P, b, e, thr, stop);
        } else if (mode == 1) {
          kernBegin();
          if (vsweepWaves4 && vb)
            k_vsweep_o4<true><<<xcdGrid(e - b), TPB, 0, st>>>(dD, P, b, e, thr, stop, vb);
          else if (vsweepWaves4)
            k_vsweep_o4<<<xcdGrid(e - b), TPB, 0, st>>>(dD, P, b, e, thr, stop);
          else if (vb)
            k_vsweep<1, true><<<xcdGrid(e - b), TPB, 0, st>>>(dD, P, b, e, thr, stop, vb);
          else
            k_vsweep<1><<<xcdGrid(e - b), TPB, 0, st>>>(dD, P, b, e, thr, stop);
          tagSweepRec(kernEnd(K_VSWEEP, e - b));
        } else {
          if (vb)
            k_vsweep<2, true><<<xcdGrid(e - b), TPB, 0, st>>>(dD, P, b, e, thr, stop, vb);
          else
            k_vsweep<2><<<xcdGrid(e - b), TPB, 0, st>>>(dD, P, b, e, thr, stop);
        }
      }
    }
    if (cTail < numColors && vs[numColors] > vs[cTail]) {
      const ClassStarts cs = classStarts(vs);
      if (vbActive)
        k_vsweep_tail<true><<<1, VBLOCK_THREADS, 0, st>>>(dD, P, cs, cTail, numColors, mode, thr,
                                                          stop, vbActive);
      else
        k_vsweep_tail<false><<<1, VBLOCK_THREADS, 0, st>>>(dD, P, cs, cTail, numColors, mode,
                                                           thr, stop, nullptr);
    }
    HIPCHECK(hipGetLastError());
  }

  // One-workgroup velocity sweep loop in global memory (mode 1 PGS, 2 one-shot GS, or 0 the
  // warm-start pass) over the colour classes vs; `need` (optional) gates it on a device flag.
  void vLoop(int mode, const Params& P, const std::vector<int>& vs, int numColors, int iters,
             double thr, double stopThr, int* stop, int* count, const int* run,
             const int* need = nullptr) {
    const ClassStarts cs = classStarts(vs);
    if (mode == 0)
      k_vsolve_block<0><<<1, VBLOCK_THREADS, 0, st>>>(dD, P, cs, numColors, nm, iters, thr,
                                                       stopThr, stop, count, run, need);
    else if (mode == 1)
      k_vsolve_block<1><<<1, VBLOCK_THREADS, 0, st>>>(dD, P, cs, numColors, nm, iters, thr,
                                                       stopThr, stop, count, run, need);
    else
      k_vsolve_block<2><<<1, VBLOCK_THREADS, 0, st>>>(dD, P, cs, numColors, nm, iters, thr,
                                                       stopThr, stop, count, run, need);
  }
  bool vFusedRec = false;  // the timed K_VBLOCK record covers both loops of the stage
  long vBlockRec[2] = {-1, -1}, pBlockRec = -1;
  // per-colour sweep launches of the device-gated loops: (timing record, loop * 4096 + sweep);
  // a launch after its loop's stop flag was set returns at once, so its units become 0 once the
  // executed sweep counts are known
  std::vector<std::pair<long, int>> sweepRecs;
  int curSweepTag = -1;
  void tagSweepRec(long rec) {
    if (rec >= 0 && curSweepTag >= 0) sweepRecs.push_back({rec, curSweepTag});
  }
  int posSweepsPrev = 1;  // sweeps of the last batched position loop (first batch size)
  // contact lists of at least this many contacts start their colouring before the velocity stage
  static constexpr int EARLY_CONTACT_COLOURING = 1 << 18;
  // The contact colouring's last class (62) may hold rows sharing a body (shell shapes: the many
  // contacts of one pair), which run in row order.  From waveMinRows rows on they are scheduled
  // in waves on the host (one pass over the class) and run by one workgroup, wave after wave.
  long waveMinRows = 4096;
  DBuf<int> pWavePerm, pWaveStart, vWavePerm, vWaveStart;
  // the contact class's waves as launches: runs of small waves in one workgroup (x = first
  // wave, y = end wave, z = 0), a wave of at least waveLaunchRows rows over the grid (z = 1)
  long waveLaunchRows = 512;
  std::vector<int3> pWaveSegs;
  std::vector<int> pWaveStartHost;
  int vWaves62 = 0;  // waves of the manifold colouring's class 62 this substep (0: row order)
  std::vector<int> waveLast, waveStamp, waveOf, waveCnt, wavePermHost;
  int waveGen = 0;
  // rows [b, b + count) with bodies ends[]: waves in row order into perm / starts; returns the
  // wave count
  int serialClassWaves(const int2* ends, int count, int numIds, int rowBegin, DBuf<int>& perm,
                       DBuf<int>& starts) {
    if ((int)waveLast.size() < numIds) {
      waveLast.assign(numIds, 0);
      waveStamp.assign(numIds, 0);
      waveGen = 0;
    }
    if (++waveGen == 0x7fffffff) {
      std::fill(waveStamp.begin(), waveStamp.end(), 0);
      waveGen = 1;
    }
    waveOf.resize(count);
    int depth = 0;
    auto last = [&](int id) { return waveStamp[id] == waveGen ? waveLast[id] : 0; };
    for (int r = 0; r < count; ++r) {
      const int a = ends[r].x, b = ends[r].y;
      int w = last(a);
      if (b >= 0) w = std::max(w, last(b));
      ++w;
      waveOf[r] = w - 1;
      waveLast[a] = w;
      waveStamp[a] = waveGen;
      if (b >= 0) {
        waveLast[b] = w;
        waveStamp[b] = waveGen;
      }
      depth = std::max(depth, w);
    }
    waveCnt.assign(depth + 1, 0);
    for (int r = 0; r < count; ++r) waveCnt[waveOf[r] + 1]++;
    for (int w = 0; w < depth; ++w) waveCnt[w + 1] += waveCnt[w];
    wavePermHost.resize(count);
    std::vector<int> fill(waveCnt.begin(), waveCnt.end() - 1);
    for (int r = 0; r < count; ++r) wavePermHost[fill[waveOf[r]]++] = rowBegin + r;
    perm.ensure(count + 1, st);
    starts.ensure(depth + 2, st);
    writeN(perm.p, wavePermHost.data(), count);
    writeN(starts.p, waveCnt.data(), depth + 1);
    return depth;
  }
  // trailing contact colours holding at most this many rows together run in one workgroup
  long posTailRows = 2048;
  // trailing velocity colours holding at most this many rows together run in one workgroup
  // (default off: on the resting 1M bed it shortened the sweeps' own time but the substep's
  // wall time grew by ~0.1 ms)
  long velTailRows = 0;
  // contact lists of at least this many contacts are coloured on the side stream
  long sideStreamMinContacts = 0;
  // the adjoint replays each substep of a checkpoint segment once, into workspaces of its own
  bool adjKeepSegment = true;
  void demStep() {
    std::memset(&stats, 0, sizeof stats);
    vbActive = nullptr;  // a substep cut short by an error may have left the stage's records
    vBlockRec[0] = vBlockRec[1] = pBlockRec = -1;
    sweepRecs.clear();
    curSweepTag = -1;
    vFusedRec = false;
    timedNow = timing > 0 && timingSubstep % timingEvery == 0;
    ++timingSubstep;
    // a frictional wall drives friction even for frictionless bodies (src/sim.hpp:169-170)
    double wallFrictionMax = 0.0;
    for (const WallP& w : wallsHost) wallFrictionMax = std::max(wallFrictionMax, w.friction);
    const bool frictionOn = frictionDynamic > 0.0 || wallFrictionMax > 0.0;
    const bool usePersist = (g[0] != 0.0 || g[1] != 0.0 || g[2] != 0.0);
    // count-averaged Coulomb friction when the PGS cone is off (src/sim.hpp:173-176)
    const bool legacyFriction = frictionOn && !(usePersist && useGS);
    mark(ST_PREDICT);
    refreshDev();
    if (taping) tapeSubstepStart(usePersist && useGS, legacyFriction);
    // growth ramp, src/sim.hpp:100-108
    if (growthFactor != -1.0 && growthRate != 0.0) {
      growthFactor *= std::exp(growthRate * dt);
      if (growthFactor > 1.0) growthFactor = 1.0;
    }
    if (growthFactor > 0.0 && numReal > 0) {
      k_growth<<<nblk(numReal), TPB, 0, st>>>(scale.p, targetScale.p, growthFactor, numReal);
      scaleKnown = false;
    }
    // maxOwnedRadius, src/sim.hpp:61-70 (read back only when a scale changed)
    double maxRad = globalScale * baseRadius;
    if (numReal > 0) {
      if (!scaleKnown) {
        const int nbs = std::min(nblk(numReal), RED_BLOCKS);
        redPart.ensure(8 * RED_BLOCKS, st);
        k_max_scale<<<nbs, TPB, 0, st>>>(scale.p, numReal, redPart.p);
        k_max_final<<<1, TPB, 0, st>>>(redPart.p, nbs, red.p);
        maxScaleKey = read1(red.p);
        scaleKnown = true;
      }
      const unsigned long long k = maxScaleKey;
      const double mx = k ? ordVal(k) : -std::numeric_limits<double>::max();
      maxRad = mx * globalScale * baseRadius;
    }
    const double margin = 0.1 * maxRad;
    numParticles = numReal;
    Params P = params(margin, maxRad);
    // periodic ghosts, src/periodicity.hpp:43-94: counted in the predict launch.  The ghost total
    // is read back together with the pair total: the broadphase runs on the device-side slot
    // count, and if the ghosts did not fit the slot buffers they grow and the ghosts and the
    // broadphase run again.
    ghostCnt.ensure(numReal + 1, st);
    ghostOff.ensure(numReal + 1, st);
    slotInfo.ensure(4, st);
    SubstepStart S0{candFlag.p, counters.p, supCnt.p, ctl.p, CTL_COUNT, bits.p, nullptr, 8,
                    ghostCnt.p, nullptr};
    if (usePersist && useGS) S0.ground = ground.p;  // the grounded levels are updated this substep
    // the velocity stage's packed records are written by the predict launch when the previous
    // substep's stage used them (nothing changes the reals' velocities in between)
    vbFromPredict = packBodies && vbUsedLast && numReal > 0;
    if (vbFromPredict) {
      vbody.ensure(numReal, st);
      S0.vb = vbody.p;
    }
    kernBegin();
    k_predict<<<std::max(1, nblk(numReal)), TPB, 0, st>>>(dD, P, S0);
    kernEnd(K_PREDICT, numReal);
    scanQuiet(ghostCnt.p, ghostOff.p, numReal);
    const Params PB = params(margin, 0.0);
    // candidate pairs kept across substeps (grown boxes): used while they provably hold every
    // pair, rebuilt otherwise; the pair list is the same either way
    const double skin = candSkinFrac * maxRad;
    bool useCand = candValid && skin > 0.0 && candNumReal == numReal && candSkin == skin &&
                   candMargin == margin;
    int G = 0;
    for (;;) {
      const int capSlots = slotCapacity();
      refreshDev();
      k_ghost_emit<<<std::max(1, nblk(numReal)), TPB, 0, st>>>(dD, P, ghostOff.p, capSlots,
                                                               slotInfo.p);
      HIPCHECK(hipGetLastError());
      mark(ST_BROAD);
      const long gridSlots = std::min<long>(capSlots, (long)numReal + lastGhosts + 1024);
      if (useCand) {
        // speculative: filter the candidates and check, in the same read-back, that they hold
        pairCnt.ensure(numReal + 1, st);
        pairStart.ensure(numReal + 1, st);
        refreshDev();
        // (the same launch checks the ghost slots)
        k_cand_count<<<std::max(1, nblk(numReal)), TPB, 0, st>>>(
            dD, P, cand.p, candStart.p, candKeep.p, candRef.p, candCode.p, 0.25 * skin,
            candFlag.p, slotInfo.p, candSlots);
        scanQuiet(pairCnt.p, pairStart.p, numReal);
      } else {
        Params PS = PB;
        PS.skin = skin;
        broadphaseQuery(PS, slotInfo.p, (int)gridSlots, capSlots);
      }
      PostBatch rb = postBatch(hostPost, hostPostDev, 64);
      const int* npH = rb.add(pairStart.p + numReal, 1);
      const int* gH = rb.add(ghostOff.p + numReal, 1);
      const int* fH = useCand ? rb.add(candFlag.p, 1) : nullptr;
      rb.post(st);
      rb.wait(st);
      np = *npH;
      G = *gH;
      const bool stale = fH && *fH;
      if ((long)numReal + G > capSlots) {
        ensureSlots(numReal + G, numReal);
        continue;
      }
      if (stale) {
        useCand = false;  // a box moved too far or the images changed: rebuild
        continue;
      }
      break;
    }
    lastGhosts = G;
    numParticles = numReal + G;
    if (useCand) {
      ++candReuses;
      pairs.ensure(np + 1, st);
      refreshDev();
      k_cand_write<<<nblk(numReal), TPB, 0, st>>>(dD, numReal, cand.p, candStart.p, candKeep.p);
      HIPCHECK(hipGetLastError());
    } else if (skin > 0.0) {
      // rebuild: the grown broadphase's pairs become the candidates, then the exact filter
      Params PS = PB;
      PS.skin = skin;
      broadphaseCompact(PS);
      std::swap(cand, pairs);
      std::swap(candStart, pairStart);
      candKeep.ensure(np + 1, st);
      candRef.ensure(numParticles + 1, st);
      candCode.ensure(numReal + 1, st);
      candFlag.ensure(1, st);
      pairCnt.ensure(numReal + 1, st);
      pairStart.ensure(numReal + 1, st);
      refreshDev();
      k_cand_ref<<<nblk(numParticles), TPB, 0, st>>>(dD, P, numParticles, candRef.p, candCode.p);
      k_cand_count<<<nblk(numReal), TPB, 0, st>>>(dD, P, cand.p, candStart.p, candKeep.p, nullptr,
                                                  nullptr, 0.0, nullptr);
      np = scan(pairCnt.p, pairStart.p, numReal);
      pairs.ensure(np + 1, st);
      refreshDev();
      k_cand_write<<<nblk(numReal), TPB, 0, st>>>(dD, numReal, cand.p, candStart.p, candKeep.p);
      HIPCHECK(hipGetLastError());
      candValid = true;
      candSlots = numParticles;
      candNumReal = numReal;
      candSkin = skin;
      candMargin = margin;
      ++candRebuilds;
    } else {
      broadphaseCompact(PB);
    }
    mark(ST_NARROW);
    narrowphase(PB, false, true);
    mark(ST_MANIFOLD);
    buildManifolds();
    if (legacyFriction) {
      cLamN.ensure(nc + 1, st);
      frCnt.ensure(numParticles + 1, st);
      refreshDev();
      if (nc > 0) k_fr_plane_load<<<nblk(nc), TPB, 0, st>>>(dD, nc);
    }
    mark(ST_COLOR_V);
    // The contact colouring (src/solver_position.hpp:155-241) needs only the contact list: it
    // runs on the side stream.  A large contact list starts it once the manifold colouring is
    // done (the two colourings would slow each other), beside the whole velocity stage; a small
    // one after the velocity stage's launches are enqueued, so that enqueueing it does not hold
    // back a short stage.  Lists shorter than sideStreamMinContacts are coloured on the main
    // stream instead (measured: a batch of small systems ran 7 % slower that way; the cross-
    // stream waits cost less than the overlap gains).
    const bool colourP = useGS && nc > 0 && numParticles > 0;
    bool colourPStarted = false;
    jobP.stream = nc >= sideStreamMinContacts ? sideStream : st;
    const bool sideP = jobP.stream != st;
    if (sideP) HIPCHECK(hipEventRecord(evContacts, st));
    auto startContactColouring = [&]() {
      cColor.ensure(nc + 1, st);
      pPerm.ensure(nc + 1, st);
      pRowOf.ensure(nc + 1, st);
      jobP.work[0].ensure(nc + 1, st);
      jobP.work[1].ensure(nc + 1, st);
      refreshDev();
      // ordered after the contact list only (evContacts was recorded behind the manifolds)
      if (sideP) HIPCHECK(hipStreamWaitEvent(jobP.stream, evContacts, 0));
      k_jp_init_contacts<<<nblk(nc), TPB, 0, jobP.stream>>>(cColor.p, nc, jobP.work[0].p);
      colourBegin(jobP, cAB.p, cColor.p, nc, numParticles, nullptr, nc, pPerm.p, pRowOf.p);
      colourPStarted = true;
    };

    // manifold colouring, src/solver_velocity.hpp:277-374
    vPerm.ensure(nm + 1, st);
    jobV.work[0].ensure(nm + 1, st);
    jobV.work[1].ensure(nm + 1, st);
    refreshDev();
    // (the same pass writes each manifold's ledger key and support code and the support list)
    supList.ensure(nm + 1, st);
    supCnt.ensure(1, st);
    refreshDev();
    const double gMag0 = std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
    const d3 gHat0 = (usePersist && gMag0 > 0.0) ? d3{g[0] / gMag0, g[1] / gMag0, g[2] / gMag0}
                                                 : d3{0, 0, 0};
    k_jp_init_manifolds<<<std::max(1, std::min(nblk(nm), CHUNK_GRID)), TPB, 0, st>>>(
        dD, nm, jobV.work[0].p, counters.p + 1, vRowOf.p, gHat0, supList.p, supCnt.p);
    std::vector<int> vs;
    int numColors = 0;
    int velLeftover = 0;
    if (useGS && nm > 0 && numReal > 0) {
      colourBegin(jobV, mReal.p, mColor.p, nm, numReal, counters.p + 1, 0, vPerm.p, vRowOf.p);
      velLeftover = colourFinish(jobV, vs, numColors, stats.rounds_vel, stats.uncolored_vel);
    } else {
      vs.assign(65, 0);
    }
    if (colourP && nc >= EARLY_CONTACT_COLOURING) startContactColouring();
    nRowsV = vs[numColors];
    mark(ST_PREP);

    const bool usePGS = usePersist && useGS;
    const double gMagP = std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
    const d3 gHat = usePersist ? d3{g[0] / gMagP, g[1] / gMagP, g[2] / gMagP} : d3{0, 0, 0};
    const double vRest = 2.0 * dt * std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
    const double vRestS = 2.0 * dt * gMagP;
    const bool hostLoop = velLeftover > 0 || (usePGS && stabMode >= 2) || !useGS ||
                          legacyFriction;
    const bool smallV = nRowsV <= blockRows;
    // the whole velocity stage in one launch on LDS-staged bodies
    const bool fusedV = !hostLoop && smallV && velIters > 0 && ldsLoop && nRowsV <= VL_ROWS;
    const size_t R = nRowsV + 1;
    vAB.ensure(R, st); vN.ensure(R, st); vTA.ensure(R, st); vTB.ensure(R, st); vRA.ensure(R, st);
    vRB.ensure(R, st); vVt0.ensure(R, st); vLamT.ensure(R, st); vVn0.ensure(R, st);
    vLam.ensure(R, st); vFlag.ensure(R, st); vWV.ensure(R, st);
    refreshDev();

    if (nRowsV > 0)
      k_vrows<<<nblk(nm), TPB, 0, st>>>(dD, P, nm, prevCount > 0 ? ledgerBuckets : 0,
                                        usePGS ? 1 : 0);
    // a large manifold class 62 (rows sharing bodies) runs in waves in the per-colour sweeps
    vWaves62 = 0;
    if (numColors == 63 && vs[63] - vs[62] >= std::max(1L, waveMinRows)) {
      const int cnt62 = vs[63] - vs[62];
      std::vector<int2> ends(cnt62);
      readN(ends.data(), vAB.p + vs[62], cnt62);
      vWaves62 = serialClassWaves(ends.data(), cnt62, numReal, vs[62], vWavePerm, vWaveStart);
    }
    // the per-colour launches of a large bed read each row's bodies from one 128-byte record
    // per body, written by the predict launch or packed here, and unpacked when the stage ends
    const bool usePack = packBodies && !hostLoop && !fusedV && !smallV && nRowsV > 0 &&
                         numReal > 0 && (usePGS || velIters > 0);
    if (usePack) {
      vbody.ensure(numReal, st);
      if (!vbFromPredict) k_vb_pack<<<nblk(numReal), TPB, 0, st>>>(dD, vbody.p, numReal);
      vbActive = vbody.p;
    }
    vbUsedLast = usePack;
    if (usePGS) {
      // grounded levels, double-buffered (src/contact_preprocessing.hpp:213-253); the decay ran
      // in the predict launch
      int* ga = ground.p;
      int* gb = ground2.p;
      if (nm > 0 && nm <= GROUND_BLOCK_MAX) {
        // a short support list: the seed and the eight sweeps in one workgroup
        k_ground_block<<<1, GROUND_BLOCK_THREADS, 0, st>>>(dD, supList.p, supCnt.p, ga, gb);
      } else if (nm > 0) {
        const int gg = std::min(nblk(nm), 1024);
        k_ground_seed<<<gg, TPB, 0, st>>>(dD, supList.p, supCnt.p, ga, gb);
        for (int s = 0; s < 8; ++s) {
          k_ground_sweep_list<<<gg, TPB, 0, st>>>(dD, supList.p, supCnt.p, ga, gb);
          std::swap(ga, gb);
        }
      }
      // 8 swaps: the result is back in `ground`
      if (!fusedV) {
        if (nRowsV <= blockRows)
          vLoop(0, P, vs, numColors, 1, 0.0, 0.0, ctl.p + CTL_NONE, ctl.p + CTL_NONE, nullptr);
        else
          velocitySweep(vs, numColors, 0, 0.0, ctl.p + CTL_NONE);
      }
    }
    HIPCHECK(hipGetLastError());
    mark(ST_VSWEEP);
    if (fusedV) {
      // warm start, sweep loop, stabilization gate, side flags and stabilization loop in one
      // launch on bodies staged in LDS; above ldsKeys / 2 rows staging may give up, and the
      // same sequence of global-memory launches follows, gated on its flag
      const double stopThr = usePGS ? 0.02 * vRest : vRest;
      VStage S{};
      S.warm = usePGS ? 1 : 0;
      S.iters = velIters;
      S.thr = vRest;
      S.stopThr = stopThr;
      S.stop = ctl.p + CTL_STOP_V;
      S.count = ctl.p + CTL_SWEEPS_V;
      S.gate = usePGS ? 1 : 0;
      S.stabMode = stabMode;
      S.stabIters = 2 * velIters;
      S.stabThr = vRestS;
      S.stabRun = ctl.p + CTL_STAB_RUN;
      S.stabStop = ctl.p + CTL_STOP_S;
      S.stabCount = ctl.p + CTL_SWEEPS_S;
      S.gHat = gHat;
      S.riseThr = 8.0 * dt * gMagP;
      S.approachThr = 8.0 * dt * gMagP;
      int* fb = ctl.p + CTL_VFALLBACK;
      kernBegin();
      if (usePGS)
        k_vsolve_lds<1><<<1, VBLOCK_THREADS, 0, st>>>(dD, P, classStarts(vs), numColors, nm, S, fb,
                                                       ldsBodies, ldsKeys);
      else
        k_vsolve_lds<2><<<1, VBLOCK_THREADS, 0, st>>>(dD, P, classStarts(vs), numColors, nm, S, fb,
                                                       ldsBodies, ldsKeys);
      if (2 * (long)nRowsV > ldsKeys) {
        if (usePGS)
          vLoop(0, P, vs, numColors, 1, 0.0, 0.0, ctl.p + CTL_NONE, ctl.p + CTL_NONE, nullptr, fb);
        vLoop(usePGS ? 1 : 2, P, vs, numColors, velIters, vRest, stopThr, ctl.p + CTL_STOP_V,
              ctl.p + CTL_SWEEPS_V, nullptr, fb);
        if (usePGS) {
          k_stab_gate<<<1, 1, 0, st>>>(bits.p + 1, ctl.p, stabMode, vRestS, fb);
          if (stabMode == 1) {
            if (nRowsV > 0)
              k_side_flags<<<nblk(nRowsV), TPB, 0, st>>>(dD, nRowsV, gHat, S.riseThr,
                                                         S.approachThr, ctl.p + CTL_STAB_RUN, fb);
            vLoop(1, P, vs, numColors, 2 * velIters, vRestS, vRestS, ctl.p + CTL_STOP_S,
                  ctl.p + CTL_SWEEPS_S, ctl.p + CTL_STAB_RUN, fb);
          }
        }
      }
      vBlockRec[0] = kernEnd(K_VBLOCK, nRowsV);
      vFusedRec = true;
      HIPCHECK(hipGetLastError());
      mark(ST_STAB);
    } else if (!hostLoop) {
      // device-controlled adaptive stops: no host round trip per sweep
      const double stopThr = usePGS ? 0.02 * vRest : vRest;
      if (smallV && velIters > 0) {
        kernBegin();
        vLoop(usePGS ? 1 : 2, P, vs, numColors, velIters, vRest, stopThr, ctl.p + CTL_STOP_V,
              ctl.p + CTL_SWEEPS_V, nullptr);
        vBlockRec[0] = kernEnd(K_VBLOCK, nRowsV);
      } else if (!smallV) sweepLoop(velIters, ctl.p + CTL_STOP_V, [&, sv = 0]() mutable {
        // (each check clears the accumulator for the sweep after it)
        if (sv == 0) k_sweep_begin<<<1, 1, 0, st>>>(bits.p + 1, ctl.p + CTL_STOP_V);
        curSweepTag = sv++;
        velocitySweep(vs, numColors, usePGS ? 1 : 2, vRest, ctl.p + CTL_STOP_V);
        curSweepTag = -1;
        k_sweep_check<<<1, 1, 0, st>>>(bits.p + 1, ctl.p + CTL_STOP_V, ctl.p + CTL_SWEEPS_V,
                                       stopThr, 0, sv < velIters ? 1 : 0);
      });
      mark(ST_STAB);
      if (usePGS) {
        k_stab_gate<<<1, 1, 0, st>>>(bits.p + 1, ctl.p, stabMode, vRestS, nullptr);
        if (smallV) {
          // the gate stays on the device: the flags kernel and the loop read it
          if (stabMode == 1 && velIters > 0) {
            if (nRowsV > 0)
              k_side_flags<<<nblk(nRowsV), TPB, 0, st>>>(dD, nRowsV, gHat, 8.0 * dt * gMagP,
                                                         8.0 * dt * gMagP, ctl.p + CTL_STAB_RUN,
                                                         nullptr);
            kernBegin();
            vLoop(1, P, vs, numColors, 2 * velIters, vRestS, vRestS, ctl.p + CTL_STOP_S,
                  ctl.p + CTL_SWEEPS_S, ctl.p + CTL_STAB_RUN);
            vBlockRec[1] = kernEnd(K_VBLOCK, nRowsV);
          }
        } else if (stabMode == 1 && velIters > 0 && read1(ctl.p + CTL_STAB_RUN)) {
          if (nRowsV > 0)
            k_side_flags<<<nblk(nRowsV), TPB, 0, st>>>(dD, nRowsV, gHat, 8.0 * dt * gMagP,
                                                       8.0 * dt * gMagP, ctl.p + CTL_STAB_RUN,
                                                       nullptr, vbActive);
          sweepLoop(2 * velIters, ctl.p + CTL_STOP_S, [&, ss = 0]() mutable {
            if (ss == 0) k_sweep_begin<<<1, 1, 0, st>>>(bits.p + 1, ctl.p + CTL_STOP_S);
            curSweepTag = 4096 + ss++;
            velocitySweep(vs, numColors, 1, vRestS, ctl.p + CTL_STOP_S);
            curSweepTag = -1;
            k_sweep_check<<<1, 1, 0, st>>>(bits.p + 1, ctl.p + CTL_STOP_S, ctl.p + CTL_SWEEPS_S,
                                           vRestS, 0, ss < 2 * velIters ? 1 : 0);
          });
        }
      }
      if (vbActive) {
        k_vb_unpack<<<nblk(numReal), TPB, 0, st>>>(dD, vbody.p, numReal);
        vbActive = nullptr;
      }
      HIPCHECK(hipGetLastError());
    } else {
      Params PF = params(margin, maxRad);
      for (int it = 0; it < velIters; ++it) {
        if (legacyFriction && nc > 0)
          k_fr_accum_normal<<<nblk(nc), TPB, 0, st>>>(dD, nc, growthRate);
        if (!useGS) {
          // count-averaged Jacobi over every manifold (src/solver_velocity.hpp:38-255)
          if (it == 0) buildIncidence<0>(nm, 0);
          jacobiVelocityPass(PF, vRest, 0);
          stats.sweeps_vel++;
          continue;
        }
        zero(bits.p + 1, 1);
        velocitySweep(vs, numColors, usePGS ? 1 : 2, vRest, ctl.p + CTL_NONE);
        if (velLeftover > 0) {
          if (it == 0) buildIncidence<0>(nm, 1);
          jacobiVelocityPass(PF, vRest, 1);
        }
        stats.sweeps_vel++;
        maxApproach = readMaxApproach();
        if (maxApproach <= (usePGS ? 0.02 * vRest : vRest)) break;
      }
      mark(ST_STAB);
      if (usePGS) {
        maxApproach = readMaxApproach();
        if (stabMode != 0 && maxApproach > vRestS) {
          stats.stab_ran = 1;
          const int one = 1;
          writeN(ctl.p + CTL_STAB_RUN, &one, 1);
          if (stabMode == 1 && nRowsV > 0)
            k_side_flags<<<nblk(nRowsV), TPB, 0, st>>>(dD, nRowsV, gHat, 8.0 * dt * gMagP,
                                                       8.0 * dt * gMagP, ctl.p + CTL_STAB_RUN,
                                                       nullptr);
          if (stabMode == 4) {
            orderedStabilization(gHat, vRestS);
          } else if (stabMode == 2) {
            multilevelStabilization(vs, numColors, gHat, vRestS, 8.0 * dt * gMagP);
            maxApproach = readMaxApproach();
          } else {
            const int cap = stabMode == 1 ? 2 * velIters : 256;
            for (int it = 0; it < cap; ++it) {
              zero(bits.p + 1, 1);
              velocitySweep(vs, numColors, 1, vRestS, ctl.p + CTL_NONE);
              stats.sweeps_stab++;
              maxApproach = readMaxApproach();
              if (maxApproach <= vRestS) break;
            }
          }
        }
      }
    }
    if (colourP && !colourPStarted) startContactColouring();
    mark(ST_LEDGER);
    if (usePGS && nm > 0) {
      // commitPairKeysLambdaKokkos, src/contact_preprocessing.hpp:160-192
      bucketCnt.ensure(numReal + 2, st);
      bucketStart.ensure(numReal + 2, st);
      lKey.ensure(nm + 1, st);
      lIdx.ensure(nm + 1, st);
      lLam.ensure(nm + 1, st);
      lLamT.ensure(nm + 1, st);
      refreshDev();
      lRank.ensure(nm + 1, st);
      // bucketCnt is all zero between commits (the scatter clears what the count set); a commit
      // cut short by an error leaves it dirty
      if (ledgerCntDirty) zero(bucketCnt.p, bucketCnt.cap);
      ledgerCntDirty = true;
      k_ledger_count<<<std::min(nblk(nm), CHUNK_GRID), TPB, 0, st>>>(dD, nm, numReal, lRank.p);
      scanQuiet(bucketCnt.p, bucketStart.p, numReal + 1);
      k_ledger_scatter<<<nblk(nm), TPB, 0, st>>>(dD, nm, numReal, lRank.p);
      ledgerCntDirty = false;
      k_ledger_sort<<<nblk(numReal), TPB, 0, st>>>(dD, numReal);
      k_ledger_values<<<nblk(nm), TPB, 0, st>>>(dD, nm);
      HIPCHECK(hipGetLastError());
      prevCount = nm;
      ledgerBuckets = numReal;
    } else if (usePGS) {
      prevCount = 0;
    }
    if (legacyFriction && nc > 0) {
      // src/solver_friction.hpp:123-220 then applyVelocityDeltas (src/integration.hpp:97-109)
      zero(frCnt.p, numParticles);
      k_fr_count<<<nblk(nc), TPB, 0, st>>>(dD, nc);
      buildIncidence<2>(nc, 0);
      prepareFallback();
      k_jfr_contrib<<<nblk(nc), TPB, 0, st>>>(dD, nc, frictionDynamic, jC.p, jF.p);
      k_jgather<2><<<nblk(numParticles), TPB, 0, st>>>(dD, numParticles, jIncStart.p,
                                                       jIncList.p, jC.p, jF.p);
      k_fr_apply<<<nblk(numParticles), TPB, 0, st>>>(dD, numParticles);
      HIPCHECK(hipGetLastError());
    }
    mark(ST_REINT);
    P = params(margin, maxRad);
    kernBegin();
    k_reintegrate<<<nblk(numParticles), TPB, 0, st>>>(dD, P);
    kernEnd(K_REINTEGRATE, numParticles);
    mark(ST_COLOR_P);

    // contact colouring over raw slots (started on the side stream), src/solver_position.hpp:155-241
    cColor.ensure(nc + 1, st);
    pPerm.ensure(nc + 1, st);
    pRowOf.ensure(nc + 1, st);
    refreshDev();
    std::vector<int> ps;
    int numPosColors = 0;
    int posLeftover = 0;
    if (colourP) {
      posLeftover = colourFinish(jobP, ps, numPosColors, stats.rounds_pos, stats.uncolored_pos);
      if (sideP) {
        HIPCHECK(hipEventRecord(evSide, jobP.stream));
        HIPCHECK(hipStreamWaitEvent(st, evSide, 0));
      }
    } else {
      ps.assign(65, 0);
    }
    const int nRowsP = ps[numPosColors];
    pAB.ensure(nRowsP + 1, st);
    pN.ensure(nRowsP + 1, st);
    pRA.ensure(nRowsP + 1, st);
    pRB.ensure(nRowsP + 1, st);
    pLam.ensure(nRowsP + 1, st);
    refreshDev();
    // with the batched position loop ahead, the rows launch also clears its residual
    const bool posBatches = useGS && posLeftover == 0 && nRowsP > blockRows;
    if (nRowsP > 0)
      k_prows<<<nblk(nc), TPB, 0, st>>>(dD, nc, posBatches ? bits.p : nullptr);
    mark(ST_PSWEEP);
    const double posTol = 1e-4 * baseRadius * globalScale;
    const int* finalCtl = nullptr;
    const unsigned long long* finalBits = nullptr;
    if (!useGS) {
      // Jacobi XPBD over every contact, no adaptive stop (src/solver_position.hpp:34-134)
      if (posIters > 0) buildIncidence<1>(nc, 0);
      for (int it = 0; it < posIters; ++it) {
        jacobiPositionPass(0);
        stats.sweeps_pos++;
      }
      HIPCHECK(hipGetLastError());
    } else if (posLeftover == 0 && nRowsP <= blockRows) {
      if (posIters > 0) {
        kernBegin();
        k_psolve_block<<<1, BLOCK_SOLVER_THREADS, 0, st>>>(dD, classStarts(ps), numPosColors, nc,
                                                           posIters, posTol, ctl.p + CTL_STOP_P,
                                                           ctl.p + CTL_SWEEPS_P);
        pBlockRec = kernEnd(K_PBLOCK, nRowsP);
      }
      HIPCHECK(hipGetLastError());
    } else if (posLeftover == 0) {
      // batches of device-gated sweeps, the first as long as the previous substep's loop; each
      // batch ends with one read of the loop counters and residuals, which is also the substep's
      // final read-back once the loop has stopped
      // the trailing colours holding at most posTailRows rows together run in one workgroup
      // launch that also ends the sweep (stop check, accumulator cleared for the next sweep)
      int cFrom = numPosColors;
      while (cFrom > 0 && ps[numPosColors] - ps[cFrom - 1] <= posTailRows) --cFrom;
      const ClassStarts pcs = classStarts(ps);
      int waves62 = 0;
      if (numPosColors == 63 && ps[63] - ps[62] >= std::max(1L, waveMinRows)) {
        cFrom = 63;  // colour 62 leaves the tail launch: it runs in waves
        const int cnt62 = ps[63] - ps[62];
        std::vector<int2> ends(cnt62);
        readN(ends.data(), pAB.p + ps[62], cnt62);
        waves62 = serialClassWaves(ends.data(), cnt62, numParticles, ps[62], pWavePerm,
                                   pWaveStart);
        pWaveStartHost.assign(waveCnt.begin(), waveCnt.begin() + waves62 + 1);
        pWaveSegs.clear();
        for (int w = 0; w < waves62; ++w) {
          const bool big = pWaveStartHost[w + 1] - pWaveStartHost[w] >= waveLaunchRows;
          if (!big && !pWaveSegs.empty() && pWaveSegs.back().z == 0 && pWaveSegs.back().y == w)
            pWaveSegs.back().y = w + 1;
          else
            pWaveSegs.push_back(make_int3(w, w + 1, big ? 1 : 0));
        }
      }
      int done = 0, batch = std::max(1, std::min(posIters, posSweepsPrev));
      // (the residual was cleared by the rows launch)
      while (done < posIters) {
        const int b = std::min(batch, posIters - done);
        for (int j = 0; j < b; ++j) {
          for (int c = 0; c < cFrom; ++c) {
            const int cb = ps[c], ce = ps[c + 1];
            if (ce <= cb) continue;
            if (c == 62 && waves62 > 0) {
              for (const int3 g : pWaveSegs) {
                if (g.z) {
                  const int wb = pWaveStartHost[g.x], we = pWaveStartHost[g.y];
                  k_psweep_perm<<<nblk(we - wb), TPB, 0, st>>>(dD, pWavePerm.p, wb, we,
                                                               ctl.p + CTL_STOP_P);
                } else {
                  k_psweep_waves<<<1, BLOCK_SOLVER_THREADS, 0, st>>>(
                      dD, pWavePerm.p, pWaveStart.p + g.x, g.y - g.x, ctl.p + CTL_STOP_P);
                }
              }
            } else if (c == 62) {
              k_psweep_serial<<<1, 1, 0, st>>>(dD, cb, ce, ctl.p + CTL_STOP_P);
            } else {
              kernBegin();
              k_psweep<<<xcdGrid(ce - cb), TPB, 0, st>>>(dD, cb, ce, ctl.p + CTL_STOP_P);
              if (timing >= 2) {
                curSweepTag = 2 * 4096 + done + j;
                tagSweepRec(kernEnd(K_PSWEEP, ce - cb));
                curSweepTag = -1;
              }
            }
          }
          k_psweep_tail<<<1, BLOCK_SOLVER_THREADS, 0, st>>>(dD, pcs, cFrom, numPosColors, nc, posTol,
                                                            ctl.p + CTL_STOP_P,
                                                            ctl.p + CTL_SWEEPS_P, posIters);
        }
        HIPCHECK(hipGetLastError());
        done += b;
        PostBatch rb = postBatch(hostPost, hostPostDev, 64);
        finalCtl = rb.add(ctl.p, CTL_COUNT);
        finalBits = reinterpret_cast<const unsigned long long*>(
            rb.add(reinterpret_cast<const int*>(bits.p), 4));
        rb.post(st);
        rb.wait(st);
        if (finalCtl[CTL_STOP_P]) break;
        batch *= 2;
      }
      posSweepsPrev = finalCtl[CTL_SWEEPS_P];
    } else {
      // colouring leftovers: the coloured sweep, then Jacobi over the uncoloured contacts
      if (posIters > 0) buildIncidence<1>(nc, 1);
      for (int it = 0; it < posIters; ++it) {
        zero(bits.p, 1);
        for (int c = 0; c < numPosColors; ++c) {
          const int b = ps[c], e = ps[c + 1];
          if (e <= b) continue;
          if (c == 62)
            k_psweep_serial<<<1, 1, 0, st>>>(dD, b, e, ctl.p + CTL_NONE);
          else
            k_psweep<<<xcdGrid(e - b), TPB, 0, st>>>(dD, b, e, ctl.p + CTL_NONE);
        }
        jacobiPositionPass(1);
        stats.sweeps_pos++;
        if (bitsd(read1(bits.p)) < posTol) break;
      }
    }
    {
      // one read-back of the loop counters and the residuals (the position loop's last one
      // when it ran in batches)
      if (!finalCtl) {
        PostBatch rb = postBatch(hostPost, hostPostDev, 64);
        finalCtl = rb.add(ctl.p, CTL_COUNT);
        finalBits = reinterpret_cast<const unsigned long long*>(
            rb.add(reinterpret_cast<const int*>(bits.p), 4));
        rb.post(st);
        rb.wait(st);
      }
      const int* c = finalCtl;
      const unsigned long long* b2 = finalBits;
      if (!hostLoop) {
        stats.sweeps_vel = c[CTL_SWEEPS_V];
        stats.sweeps_stab = c[CTL_SWEEPS_S];
        stats.stab_ran = usePGS ? c[CTL_STAB_RUN] : 0;
      }
      if (useGS && posLeftover == 0) stats.sweeps_pos = c[CTL_SWEEPS_P];
      maxOverlap = bitsd(b2[0]);
      maxApproach = bitsd(b2[1]);
    }
    mark(ST_COMMIT);
    if (usePGS && nm > 0) k_commit_pos_impulse<<<nblk(nm), TPB, 0, st>>>(dD, nm, numReal, dt);
    k_final_commit<<<nblk(numReal), TPB, 0, st>>>(dD, P);
    if (thermoTau > 0.0 && dt > 0.0) applyThermostat();
    HIPCHECK(hipGetLastError());
    if (taping)
      tapeSubstepEnd(P, vs, numColors, ps, numPosColors, nRowsP, velLeftover, posLeftover);
    mark(ST_COUNT);
    stats.num_real = numReal;
    stats.num_particles = numParticles;
    stats.num_pairs = np;
    stats.num_contacts = nc;
    stats.num_manifolds = nm;
    stats.active_manifolds = nRowsV;
    stats.num_colors_vel = numColors;
    stats.num_colors_pos = numPosColors;
    stats.max_overlap = maxOverlap;
    stats.max_approach = maxApproach;
    if (timing >= 2) {
      // the one-workgroup loops processed rows x sweeps, known only now
      if (vBlockRec[0] >= 0)
        kernPending[vBlockRec[0]].units =
            (long)nRowsV * (stats.sweeps_vel + (vFusedRec ? stats.sweeps_stab : 0));
      if (vBlockRec[1] >= 0) kernPending[vBlockRec[1]].units = (long)nRowsV * stats.sweeps_stab;
      if (pBlockRec >= 0) kernPending[pBlockRec].units = (long)nRowsP * stats.sweeps_pos;
      for (const auto& rt : sweepRecs) {
        const int loop = rt.second / 4096, sweep = rt.second % 4096;
        const long ran = loop == 0 ? stats.sweeps_vel : loop == 1 ? stats.sweeps_stab
                                                                  : stats.sweeps_pos;
        if (sweep >= ran && rt.first < (long)kernPending.size()) kernPending[rt.first].units = 0;
      }
      statsLog.push_back(stats);
    }
    if (kernPending.size() > 8192 || stagePending.size() > 256 * (ST_COUNT + 1)) resolveTiming();
  }

  // applyThermostatKokkos, src/integration.hpp:229-271 (block partials summed in block order)
  // ---------------------------------------------------------------- Hertz–Mindlin engine
  void refreshHz() {
    Z.pairs = hPairs.p; Z.iStart = hIStart.p; Z.jStart = hJStart.p; Z.jList = hJList.p;
    Z.xi = hXi.p; Z.f = hF.p; Z.tA = hTA.p; Z.tB = hTB.p; Z.prevKeys = hPrevKeys.p;
    Z.prevXi = hPrevXi.p; Z.refPos = hRefPos.p; Z.wStart = hWStart.p; Z.wCand = hWCand.p;
    Z.xiWall = hXiWall.p; Z.wf = hWF.p; Z.wt = hWT.p; Z.wallGrid = wallGridD.p; Z.tab = hTab.p;
  }
  HzParams hzParams(double dtH) const {
    HzParams H{};
    H.dt = dtH;
    H.gs = globalScale;
    H.baseR = baseRadius;
    H.frac = hertzFrac;
    H.g = d3{g[0], g[1], g[2]};
    H.numReal = numReal;
    H.numWalls = (int)wallsHost.size();
    H.hasShapes = shellHost.size() > 0;
    H.numPairs = hNumPairs > 0 ? hNumPairs : 0;
    H.shape = params(0.0, 0.0).shape;
    for (int w = 0; w < H.numWalls && w < HZ_MAXWALLS; ++w) H.walls[w] = wallsHost[w];
    return H;
  }
  // per material pair / (material, wall): beta_d, mu, E*, G* (src/solver_hertz.hpp:27-48,
  // :187-199); computed on the host so the logarithm is the same one the oracle uses
  void uploadHertzTable() {
    HzTab T{};
    for (int a = 0; a < 8; ++a)
      for (int b = 0; b < 8; ++b) {
        double e = restitutionNormal, mu = frictionDynamic;
        if (hasPairTable) {
          e = pairTable[(a * 8 + b) * 2];
          mu = pairTable[(a * 8 + b) * 2 + 1];
        }
        double* c = T.pair[a * 8 + b];
        c[0] = hzBetaD(e);
        c[1] = mu;
        hzStars(hertzE[a], hertzE[b], hertzNu[a], hertzNu[b], &c[2], &c[3]);
      }
    for (int a = 0; a < 8; ++a)
      for (int w = 0; w < (int)wallsHost.size() && w < 8; ++w) {
        const WallP& W = wallsHost[w];
        double e = W.restitution, mu = W.friction;
        if (W.materialId >= 0 && hasPairTable) {
          e = pairTable[(a * 8 + W.materialId) * 2];
          mu = pairTable[(a * 8 + W.materialId) * 2 + 1];
        }
        const int mw = W.materialId >= 0 ? W.materialId : a;
        double* c = T.wall[a * 8 + w];
        c[0] = hzBetaD(e);
        c[1] = mu;
        hzStars(hertzE[a], hertzE[mw], hertzNu[a], hertzNu[mw], &c[2], &c[3]);
      }
    hTab.ensure(1, st);
    writeN(hTab.p, &T, 1);
  }
  // src/sim.hpp:596-624 then :545-594, and the wall candidates (src/solver_hertz.hpp:343-370)
  void hertzRebuild(double skin) {
    const int n = numReal;
    if (hNumPairs > 0) {
      hPrevKeys.ensure(hNumPairs, st);
      hPrevXi.ensure(hNumPairs, st);
      refreshHz();
      k_hz_commit<<<nblk(hNumPairs), TPB, 0, st>>>(Z, hNumPairs);
      hPrevCount = hNumPairs;
    } else {
      hPrevCount = 0;
    }
    numParticles = n;
    {  // the broadphase reads posPred: current positions, radius in .w
      HzParams H0 = hzParams(dt);
      k_hz_prep<<<nblk(n), TPB, 0, st>>>(dD, H0, red.p + 1);
    }
    Params P = params(skin, 0.0);
    broadphase(P);
    hPairs.ensure(np + 1, st);
    hXi.ensure(np + 1, st);
    hF.ensure(np + 1, st);
    hTA.ensure(np + 1, st);
    hTB.ensure(np + 1, st);
    hJList.ensure(np + 1, st);
    hIStart.ensure(n + 1, st);
    hJStart.ensure(n + 1, st);
    hJCnt.ensure(n + 1, st);
    refreshDev();
    refreshHz();
    k_hz_carry<<<nblk(np), TPB, 0, st>>>(dD, Z, np, hPrevCount);
    HIPCHECK(hipMemcpyAsync(hIStart.p, pairStart.p, (n + 1) * sizeof(int),
                            hipMemcpyDeviceToDevice, st));
    zero(hJCnt.p, n + 1);
    k_hz_jcount<<<nblk(np), TPB, 0, st>>>(Z, np, hJCnt.p);
    scanQuiet(hJCnt.p, hJStart.p, n);
    zero(hJCnt.p, n + 1);
    k_hz_jfill<<<nblk(np), TPB, 0, st>>>(Z, np, hJCnt.p);
    hNumPairs = np;
    HzParams H = hzParams(dt);
    k_hz_jsort<<<nblk(n), TPB, 0, st>>>(Z, H);
    k_hz_ref<<<nblk(n), TPB, 0, st>>>(dD, Z, H);
    if (!wallsHost.empty()) {
      hWStart.ensure(n + 1, st);
      hWCnt.ensure(n + 1, st);
      refreshHz();
      k_hz_wallcand<false><<<nblk(n), TPB, 0, st>>>(dD, Z, H, skin, hWCnt.p);
      hNumWallCand = scan(hWCnt.p, hWStart.p, n);
      hWCand.ensure(hNumWallCand + 1, st);
      hWF.ensure(hNumWallCand + 1, st);
      hWT.ensure(hNumWallCand + 1, st);
      refreshHz();
      k_hz_wallcand<true><<<nblk(n), TPB, 0, st>>>(dD, Z, H, skin, hWCnt.p);
    } else {
      hNumWallCand = 0;
    }
    HIPCHECK(hipGetLastError());
  }
  // demStepHertz, src/sim.hpp:628-750
  void stepHertz(double dtH, int nsteps, double skinFrac) {
    dt = dtH;
    if (per[0] || per[1] || per[2])
      throw std::runtime_error("step_hertz: periodic domains not supported");
    if (numPlanes > 0)
      throw std::runtime_error("step_hertz: analytic planes not supported (use an SDF wall)");
    if ((int)wallsHost.size() > HZ_MAXWALLS)
      throw std::runtime_error("step_hertz: too many SDF walls");
    const int n = numReal;
    hXiWall.ensure((size_t)(n > 0 ? n : 1) * HZ_MAXWALLS, st, hXiWall.cap);
    hRefPos.ensure(n + 1, st);
    uploadHertzTable();
    refreshDev();
    refreshHz();
    HzParams H = hzParams(dtH);
    HIPCHECK(hipMemsetAsync(red.p, 0xff, sizeof(unsigned long long), st));
    k_hz_prep<<<nblk(n), TPB, 0, st>>>(dD, H, red.p);
    const unsigned long long mk = read1(red.p);
    const double minRad = (mk == ~0ull) ? 3.4e38 : ordVal(mk);
    const double skin = skinFrac * minRad;
    const double rebuildAt = 0.25 * skin * skin;
    bool rebuild = hNumPairs < 0;
    for (int k = 0; k < nsteps; ++k) {
      if (rebuild) {
        hertzRebuild(skin);
        rebuild = false;
      }
      H = hzParams(dtH);
      if (H.numPairs > 0) {
        if (H.hasShapes) k_hz_shape_pairs<<<nblk(H.numPairs), TPB, 0, st>>>(dD, Z, H);
        else k_hz_pairs<<<nblk(H.numPairs), TPB, 0, st>>>(dD, Z, H);
      }
      if (H.numWalls > 0 && hNumWallCand > 0)
        k_hz_walls<<<nblk(hNumWallCand), TPB, 0, st>>>(dD, Z, H, hNumWallCand);
      if (H.numWalls == 0 || hNumWallCand == 0) H.numWalls = 0;
      k_hz_integrate<<<nblk(n), TPB, 0, st>>>(dD, Z, H);
      if ((k & 15) == 15 || k == nsteps - 1) {
        zero(bits.p + 2, 1);
        k_hz_disp<<<std::min(nblk(n), RED_BLOCKS), TPB, 0, st>>>(dD, Z, H, bits.p + 2);
        if (bitsd(read1(bits.p + 2)) > rebuildAt) rebuild = true;
      }
    }
    HIPCHECK(hipGetLastError());
    sync();
  }

  void applyThermostat() {
    const int nb = nblk(numReal);
    thermo.ensure(2 * nb + 2, st);
    refreshDev();
    Params P = params(0.0, 0.0);
    k_thermo_partial<<<nb, TPB, 0, st>>>(dD, P, thermo.p);
    std::vector<double> part(2 * nb);
    readN(part.data(), thermo.p, part.size());
    double ke[2] = {0.0, 0.0};
    for (int b = 0; b < nb; ++b) {
      ke[0] += part[2 * b];
      ke[1] += part[2 * b + 1];
    }
    const double ndof = 3.0 * numReal;
    double lam[2];
    for (int k = 0; k < 2; ++k) {
      lam[k] = 1.0;
      if (ndof <= 0 || thermoKB <= 0) continue;
      const double Tcur = (2.0 * ke[k]) / (ndof * thermoKB);
      if (Tcur < 1e-6) continue;
      double f = 1.0 + (dt / thermoTau) * (thermoTemp / Tcur - 1.0);
      if (f < 0.0) f = 0.0;
      lam[k] = std::sqrt(f);
    }
    k_thermo_scale<<<nb, TPB, 0, st>>>(dD, P, lam[0], lam[1]);
  }

  // computeOverlapsKokkos, src/sim.hpp:494-541
  double computeOverlaps() {
    refreshDev();
    double maxRad = globalScale * baseRadius;
    if (numReal > 0) {
      const int nbs = std::min(nblk(numReal), RED_BLOCKS);
      redPart.ensure(8 * RED_BLOCKS, st);
      k_max_scale<<<nbs, TPB, 0, st>>>(scale.p, numReal, redPart.p);
      k_max_final<<<1, TPB, 0, st>>>(redPart.p, nbs, red.p);
      const unsigned long long k = read1(red.p);
      maxRad = (k ? ordVal(k) : -std::numeric_limits<double>::max()) * globalScale * baseRadius;
    }
    const double margin = 0.1 * maxRad;
    numParticles = numReal;
    Params P = params(margin, maxRad);
    k_overlap_prep<<<nblk(numReal), TPB, 0, st>>>(dD, P);
    ghostCnt.ensure(numReal + 1, st);
    ghostOff.ensure(numReal + 1, st);
    k_ghost_count<<<nblk(numReal), TPB, 0, st>>>(dD, P, ghostCnt.p);
    const int G = scan(ghostCnt.p, ghostOff.p, numReal);
    if ((size_t)(numReal + G) > pos.cap) ensureSlots(numReal + G, numReal);
    refreshDev();
    k_ghost_emit<<<nblk(numReal), TPB, 0, st>>>(dD, P, ghostOff.p, slotCapacity());
    numParticles = numReal + G;
    broadAndNarrow(margin);
    numParticles = numReal;
    maxOverlap = bitsd(read1(bits.p));
    return maxOverlap;
  }
};

// ================================================================================================
// C-ABI
// ================================================================================================

#include "pd_adjoint_host.hpp"

namespace {
template <class F>
int guard(F&& f) {
  try {
    f();
    return 0;
  } catch (const ArgError& e) {
    g_error = e.what();
    return -2;
  } catch (const std::exception& e) {
    g_error = e.what();
    return -1;
  }
}
void writeLammps(const std::string& filename, int step, const double* pos, const double* vel,
                 const double* quat, const double* radii, int n, const double* bmn,
                 const double* bmx, bool pbc) {
  // writeLammpsDump, src/io.hpp:23-82
  double bmin[3], bmax[3];
  if (bmn && bmx) {
    for (int d = 0; d < 3; ++d) {
      bmin[d] = bmn[d];
      bmax[d] = bmx[d];
    }
  } else if (n == 0) {
    for (int d = 0; d < 3; ++d) bmin[d] = bmax[d] = 0.0;
  } else {
    for (int d = 0; d < 3; ++d) {
      bmin[d] = std::numeric_limits<double>::max();
      bmax[d] = -std::numeric_limits<double>::max();
    }
    for (int i = 0; i < n; ++i) {
      const double r = radii ? radii[i] : 0.0;
      for (int d = 0; d < 3; ++d) {
        bmin[d] = std::min(bmin[d], pos[3 * i + d] - r);
        bmax[d] = std::max(bmax[d], pos[3 * i + d] + r);
      }
    }
    for (int d = 0; d < 3; ++d) {
      const double m = 0.05 * (bmax[d] - bmin[d]);
      bmin[d] -= m;
      bmax[d] += m;
    }
  }
  std::ofstream file(filename);
  if (!file) throw std::runtime_error("Could not open file for writing: " + filename);
  file << "ITEM: TIMESTEP\n" << step << "\n";
  file << "ITEM: NUMBER OF ATOMS\n" << n << "\n";
  file << "ITEM: BOX BOUNDS " << (pbc ? "pp pp pp" : "ff ff ff") << "\n";
  for (int d = 0; d < 3; ++d) file << bmin[d] << " " << bmax[d] << "\n";
  file << "ITEM: ATOMS id type x y z radius vx vy vz qw qx qy qz\n";
  file << std::fixed << std::setprecision(6);
  for (int i = 0; i < n; ++i) {
    const double r = radii ? radii[i] : 1.0;
    const double vx = vel ? vel[3 * i] : 0.0, vy = vel ? vel[3 * i + 1] : 0.0,
                 vz = vel ? vel[3 * i + 2] : 0.0;
    double qx = 0, qy = 0, qz = 0, qw = 1;
    if (quat) {
      qx = quat[4 * i];
      qy = quat[4 * i + 1];
      qz = quat[4 * i + 2];
      qw = quat[4 * i + 3];
    }
    file << (i + 1) << " 1 " << pos[3 * i] << " " << pos[3 * i + 1] << " " << pos[3 * i + 2]
         << " " << r << " " << vx << " " << vy << " " << vz << " " << qw << " " << qx << " " << qy
         << " " << qz << "\n";
  }
}
}  // namespace

// ---- autotuner (no reference counterpart) -------------------------------------------------
// Chooses the performance knobs of a simulation by timing its own substeps: the candidate-list
// skin (candidate_skin_permille), the largest sweep loop run inside one workgroup
// (block_solver_rows), LDS staging of the one-workgroup velocity loop (velocity_loop_lds) and the
// register budget of the per-colour PGS sweep (vsweep_waves: 3 or 4 waves per SIMD).
// None of them changes a result, so the trials step the simulation itself.  Coordinate descent
// from the current setting: each knob's values are tried in turn (one untimed substep after a
// change, which may rebuild lists, then `substeps` timed ones) and a value is kept when it is
// at least 2 % faster.  The choice is cached per size class (shape kind, log2 of the particle
// count, log2 of 1 + 4 x contacts per particle) for later calls on systems of the same class.
struct TuneCfg {
  long skin, rows, lds, waves;
};
std::mutex g_tuneMutex;
std::map<std::tuple<int, int, int>, TuneCfg> g_tuneCache;
static std::tuple<int, int, int> tuneClass(const pd_sim* s) {
  const int n = std::max(1, s->numReal);
  int ln = 0, lc = 0;
  while ((2L << ln) <= n) ++ln;
  const long dens = 1 + (4L * s->nc) / n;
  while ((2L << lc) <= dens) ++lc;
  return {s->shapeType + (s->shellHost.empty() ? 0 : 16), ln, lc};
}
static TuneCfg tuneCurrent(const pd_sim* s) {
  return {(long)std::lround(s->candSkinFrac * 1000.0), (long)s->blockRows, s->ldsLoop ? 1L : 0L,
          s->vsweepWaves4 ? 4L : 3L};
}
static void tuneApply(pd_sim* s, const TuneCfg& c) {
  if ((long)std::lround(s->candSkinFrac * 1000.0) != c.skin) {
    s->candSkinFrac = c.skin / 1000.0;
    s->candValid = false;
  }
  s->blockRows = (int)c.rows;
  s->ldsLoop = c.lds != 0;
  s->vsweepWaves4 = c.waves == 4;
}
static double tuneTrial(pd_sim* s, double dt, int substeps) {
  s->dt = dt;
  s->demStep();
  s->sync();
  const auto t0 = std::chrono::steady_clock::now();
  for (int k = 0; k < substeps; ++k) {
    s->dt = dt;
    s->demStep();
  }
  s->sync();
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() /
         std::max(1, substeps);
}
static TuneCfg tuneRun(pd_sim* s, double dt, int substeps, double* bestMs) {
  TuneCfg cur = tuneCurrent(s);
  double best = tuneTrial(s, dt, substeps);
  const long skins[] = {0, 50, 100, 200, 400};
  const long rows[] = {0, 2048, 16384, 65536};
  const long lds[] = {1, 0};
  const long waves[] = {3, 4};
  auto sweep = [&](const long* vals, int nv, long TuneCfg::*field) {
    const long start = cur.*field;
    for (int q = 0; q < nv; ++q) {
      if (vals[q] == start) continue;
      TuneCfg c = cur;
      c.*field = vals[q];
      tuneApply(s, c);
      const double t = tuneTrial(s, dt, substeps);
      if (t < 0.98 * best) {
        best = t;
        cur = c;
      }
    }
    tuneApply(s, cur);
  };
  sweep(skins, 5, &TuneCfg::skin);
  sweep(rows, 4, &TuneCfg::rows);
  sweep(lds, 2, &TuneCfg::lds);
  sweep(waves, 2, &TuneCfg::waves);
  if (bestMs) *bestMs = best;
  return cur;
}

extern "C" {
int pd_autotune(pd_sim* s, double dt, int substeps, int use_cache, long* chosen,
                double* best_ms) {
  return guard([&] {
    if (substeps < 1) throw ArgError("autotune: substeps must be >= 1");
    const auto key = tuneClass(s);
    TuneCfg c{};
    bool hit = false;
    if (use_cache) {
      std::lock_guard<std::mutex> lk(g_tuneMutex);
      auto it = g_tuneCache.find(key);
      if (it != g_tuneCache.end()) {
        c = it->second;
        hit = true;
      }
    }
    double ms = -1.0;
    if (hit) {
      tuneApply(s, c);
    } else {
      c = tuneRun(s, dt, substeps, &ms);
      std::lock_guard<std::mutex> lk(g_tuneMutex);
      g_tuneCache[key] = c;
    }
    if (chosen) {
      chosen[0] = c.skin;
      chosen[1] = c.rows;
      chosen[2] = c.lds;
      chosen[3] = c.waves;
    }
    if (best_ms) *best_ms = ms;
    drainRetired();
  });
}
}  // extern "C"

extern "C" {

const char* pd_last_error(void) { return g_error.c_str(); }

const char* pd_execution_space(void) {
  static std::string name;
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
    name = std::string("HIP(") + prop.gcnArchName + ")";
  else
    name = "HIP(no device)";
  return name.c_str();
}

pd_sim* pd_create(int capacity) {
  if (capacity < 0) {
    g_error = "capacity must be >= 0";
    return nullptr;
  }
  try {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
      throw std::runtime_error("no HIP device available (the substep runs only on the GPU)");
    return new pd_sim(capacity);
  } catch (const std::exception& e) {
    g_error = e.what();
    return nullptr;
  }
}
void pd_destroy(pd_sim* s) {
  delete s;
  (void)guard([&] { drainRetired(); });
}
int pd_synchronize(pd_sim* s) {
  return guard([&] { s->sync(); });
}

int pd_initialize_shape(pd_sim* s, int t, double r, double h, double th) {
  return guard([&] { s->initializeShape(t, r, h, th); });
}
int pd_set_shell_points(pd_sim* s, const double* pts, int n) {
  return guard([&] {
    if (n < 0) throw ArgError("set_shell_points: negative count");
    if (s->shapeType == SHAPE_SPHERE && n > 0)
      throw ArgError("set_shell_points: the analytic sphere uses a single centre probe");
    s->shellHost.assign(pts, pts + 3 * (size_t)n);
    s->uploadShell();
  });
}
int pd_get_shell_points(pd_sim* s, double* out, int cap) {
  const int n = (int)(s->shellHost.size() / 3);
  if (out)
    for (int i = 0; i < n && i < cap; ++i)
      for (int c = 0; c < 3; ++c) out[3 * i + c] = s->shellHost[3 * i + c];
  return n;
}
int pd_set_domain(pd_sim* s, double lx, double ly, double lz, int px, int py, int pz) {
  s->dmin[0] = s->dmin[1] = s->dmin[2] = 0.0;
  s->dmax[0] = lx; s->dmax[1] = ly; s->dmax[2] = lz;
  s->dsize[0] = lx; s->dsize[1] = ly; s->dsize[2] = lz;
  s->per[0] = px != 0; s->per[1] = py != 0; s->per[2] = pz != 0;
  return 0;
}
int pd_set_domain_minmax(pd_sim* s, double a, double b, double c, double d, double e, double f) {
  s->dmin[0] = a; s->dmin[1] = b; s->dmin[2] = c;
  s->dmax[0] = d; s->dmax[1] = e; s->dmax[2] = f;
  s->dsize[0] = d - a; s->dsize[1] = e - b; s->dsize[2] = f - c;
  return 0;
}
int pd_enable_periodicity(pd_sim* s, int x, int y, int z) {
  s->per[0] = x != 0; s->per[1] = y != 0; s->per[2] = z != 0;
  return 0;
}
int pd_get_domain(pd_sim* s, double* mn, double* mx) {
  for (int d = 0; d < 3; ++d) {
    mn[d] = s->dmin[d];
    mx[d] = s->dmax[d];
  }
  return 0;
}
int pd_set_gravity(pd_sim* s, double gx, double gy, double gz) {
  s->g[0] = gx; s->g[1] = gy; s->g[2] = gz;
  return 0;
}
int pd_set_thermostat(pd_sim* s, double T, double tau, double kB) {
  s->thermoTemp = T; s->thermoTau = tau; s->thermoKB = kB;
  return 0;
}
int pd_set_solver_iterations(pd_sim* s, int pos, int vel) {
  s->posIters = pos; s->velIters = vel;
  return 0;
}
int pd_set_velocity_use_gs(pd_sim* s, int use_gs) {
  s->useGS = use_gs != 0;
  return 0;
}
int pd_set_stabilization_mode(pd_sim* s, int mode) {
  return guard([&] {
    if (mode < 0 || mode > 4) throw ArgError("stabilization mode must be 0..4");
    s->stabMode = mode;
  });
}
int pd_set_global_scale(pd_sim* s, double v) { s->globalScale = v; return 0; }
int pd_set_dt(pd_sim* s, double v) { s->dt = v; return 0; }
int pd_set_material_params(pd_sim* s, double en, double et, double mu) {
  s->restitutionNormal = en; s->frictionDynamic = mu; s->restitutionTangent = et;
  return 0;
}
int pd_set_material_ids(pd_sim* s, const int* ids, int n) {
  return guard([&] {
    const int m = std::min(n, s->numReal);
    if (m <= 0) return;
    std::vector<unsigned char> h(m);
    for (int i = 0; i < m; ++i) h[i] = (unsigned char)ids[i];
    s->writeN(s->matId.p, h.data(), m);
  });
}
int pd_set_pair_material(pd_sim* s, int a, int b, double e, double mu) {
  return guard([&] {
    if (a < 0 || b < 0 || a >= 8 || b >= 8) throw ArgError("material id out of range");
    if (!s->hasPairTable) {
      for (int i = 0; i < 64; ++i) {
        s->pairTable[2 * i] = s->restitutionNormal;
        s->pairTable[2 * i + 1] = s->frictionDynamic;
      }
      s->hasPairTable = 1;
    }
    s->pairTable[(a * 8 + b) * 2] = e; s->pairTable[(a * 8 + b) * 2 + 1] = mu;
    s->pairTable[(b * 8 + a) * 2] = e; s->pairTable[(b * 8 + a) * 2 + 1] = mu;
    s->writeN(s->pairTableD.p, s->pairTable, 128);
  });
}
int pd_add_plane(pd_sim* s, double px, double py, double pz, double nx, double ny, double nz) {
  if (s->numPlanes < 8) {
    s->planeP[s->numPlanes][0] = px; s->planeP[s->numPlanes][1] = py;
    s->planeP[s->numPlanes][2] = pz; s->planeN[s->numPlanes][0] = nx;
    s->planeN[s->numPlanes][1] = ny; s->planeN[s->numPlanes][2] = nz;
    s->numPlanes++;
  }
  return 0;
}
int pd_set_positions(pd_sim* s, const double* xyz, int n) {
  return guard([&] { s->setPositions(xyz, n); });
}
int pd_set_velocities(pd_sim* s, const double* v, int n) {
  return guard([&] { s->setField(1, v, n); });
}
int pd_set_external_forces(pd_sim* s, const double* f, int n) {
  return guard([&] { s->setField(6, f, n); });
}
int pd_clear_external_forces(pd_sim* s) {
  return guard([&] {
    ++s->paramsVersion;
    s->zero(s->extForce.p, s->extForce.cap);
    s->sync();
  });
}
int pd_set_quaternions(pd_sim* s, const double* q, int n) {
  return guard([&] { s->setField(3, q, n); });
}
int pd_set_angular_velocities(pd_sim* s, const double* w, int n) {
  return guard([&] { s->setField(2, w, n); });
}
int pd_set_inv_inertia(pd_sim* s, const double* ii, int n) {
  return guard([&] { s->setField(4, ii, n); });
}
int pd_set_inv_mass(pd_sim* s, const double* im, int n) {
  return guard([&] { s->setField(5, im, n); });
}
int pd_set_scales(pd_sim* s, const double* sc, int n) {
  // src/sim.hpp:1340-1346: target := given, scale := target
  return guard([&] {
    ++s->paramsVersion;
    const int m = std::min(n, s->numReal);
    if (m > 0) s->writeN(s->targetScale.p, sc, m);
    HIPCHECK(hipMemcpyAsync(s->scale.p, s->targetScale.p, s->scale.cap * sizeof(double),
                            hipMemcpyDeviceToDevice, s->st));
    s->sync();
    s->scaleKnown = false;
  });
}
int pd_set_scales_uniform(pd_sim* s, double v) {
  // src/sim.hpp:1332-1338
  return guard([&] {
    ++s->paramsVersion;
    std::vector<double> h(s->numReal, v);
    if (s->numReal > 0) s->writeN(s->scale.p, h.data(), s->numReal);
    HIPCHECK(hipMemcpyAsync(s->targetScale.p, s->scale.p, s->scale.cap * sizeof(double),
                            hipMemcpyDeviceToDevice, s->st));
    s->sync();
    s->scaleKnown = false;
  });
}
int pd_set_growth_params(pd_sim* s, double rate, double nf) {
  // src/sim.hpp:1417-1425
  return guard([&] {
    ++s->paramsVersion;
    if (s->growthFactor == -1.0) s->growthFactor = (nf > 0.0) ? nf : 0.01;
    else if (nf > 0.0) s->growthFactor = nf;
    s->growthRate = rate;
    if (s->growthFactor > 0.0 && s->numReal > 0) {
      k_growth<<<nblk(s->numReal), TPB, 0, s->st>>>(s->scale.p, s->targetScale.p,
                                                    s->growthFactor, s->numReal);
      s->sync();
      s->scaleKnown = false;
    }
  });
}
double pd_get_growth_factor(pd_sim* s) { return s->growthFactor; }
double pd_get_growth_rate(pd_sim* s) { return s->growthRate; }
int pd_num_particles(pd_sim* s) { return s->numReal; }
int pd_get_positions(pd_sim* s, double* out, int n) {
  guard([&] { s->getField(0, out, n); });
  return s->numReal;
}
int pd_get_velocities(pd_sim* s, double* out, int n) {
  guard([&] { s->getField(1, out, n); });
  return s->numReal;
}
int pd_get_quaternions(pd_sim* s, double* out, int n) {
  guard([&] { s->getField(3, out, n); });
  return s->numReal;
}
int pd_get_angular_velocities(pd_sim* s, double* out, int n) {
  guard([&] { s->getField(2, out, n); });
  return s->numReal;
}
int pd_get_inv_inertia(pd_sim* s, double* out, int n) {
  guard([&] { s->getField(4, out, n); });
  return s->numReal;
}
int pd_get_inv_mass(pd_sim* s, double* out, int n) {
  guard([&] { s->getField(5, out, n); });
  return s->numReal;
}
int pd_get_scales(pd_sim* s, double* out, int n) {
  guard([&] { s->readN(out, s->scale.p, std::min(n, s->numReal)); });
  return s->numReal;
}
void* pd_get_device_pointer(pd_sim* s, const char* name) {
  const std::string n(name ? name : "");
  if (n == "pos") return s->pos.p;
  if (n == "vel") return s->vw0.p;
  if (n == "ext_force") return s->extForce.p;
  if (n == "inv_mass") return ((double*)s->bc.p) + 7;
  g_error = "unknown device array: " + n;
  return nullptr;
}

int pd_read_device(pd_sim* s, const void* src, void* host, size_t bytes) {
  return guard([&] {
    HIPCHECK(hipMemcpyAsync(host, src, bytes, hipMemcpyDeviceToHost, s->st));
    s->sync();
  });
}
int pd_write_device(pd_sim* s, void* dst, const void* host, size_t bytes) {
  return guard([&] {
    HIPCHECK(hipMemcpyAsync(dst, host, bytes, hipMemcpyHostToDevice, s->st));
    s->sync();
  });
}
int pd_step(pd_sim* s, double dt) {
  return guard([&] {
    s->dt = dt;
    s->demStep();
    drainRetired();
  });
}
int pd_set_sdf_shape(pd_sim* s, const double* grid, int nx, int ny, int nz, double ox, double oy,
                     double oz, double sx, double sy, double sz, const double* shell, int nPts,
                     double iix, double iiy, double iiz, double boundingRadius) {
  return guard([&] {
    if (nx < 2 || ny < 2 || nz < 2)
      throw std::runtime_error("setSdfShape: grid dims must be >= 2 on each axis");
    if (nPts <= 0) throw std::runtime_error("setSdfShape: empty surface point shell");
    const size_t cells = (size_t)nx * ny * nz;
    s->sdfGridD.ensure(cells, s->st);
    s->writeN(s->sdfGridD.p, grid, cells);
    s->gridDims[0] = nx;
    s->gridDims[1] = ny;
    s->gridDims[2] = nz;
    s->gridOrigin[0] = ox;
    s->gridOrigin[1] = oy;
    s->gridOrigin[2] = oz;
    s->gridInvSpacing[0] = sx > 0 ? 1.0 / sx : 0.0;
    s->gridInvSpacing[1] = sy > 0 ? 1.0 / sy : 0.0;
    s->gridInvSpacing[2] = sz > 0 ? 1.0 / sz : 0.0;
    s->shellHost.assign(shell, shell + 3 * (size_t)nPts);
    s->uploadShell();
    s->shapeType = SHAPE_GRID_SDF;
    s->shapeParams[0] = boundingRadius;
    s->shapeParams[1] = s->shapeParams[2] = s->shapeParams[3] = 0.0;
    s->baseRadius = boundingRadius;
    s->defaultInvI[0] = iix;
    s->defaultInvI[1] = iiy;
    s->defaultInvI[2] = iiz;
  });
}
// getSdfGrid / exportSdf, src/sim.hpp:1478-1483,1538-1542 (x-fastest, over the domain box)
static SdfGridP sdfGridParams(const pd_sim* s, int rx, int ry, int rz) {
  if (rx < 1 || ry < 1 || rz < 1) throw ArgError("resolution must be >= 1 on each axis");
  SdfGridP G{};
  G.rx = rx; G.ry = ry; G.rz = rz;
  G.px = s->per[0]; G.py = s->per[1]; G.pz = s->per[2];
  G.origin = d3{s->dmin[0], s->dmin[1], s->dmin[2]};
  G.vox = d3{(s->dmax[0] - s->dmin[0]) / rx, (s->dmax[1] - s->dmin[1]) / ry,
             (s->dmax[2] - s->dmin[2]) / rz};
  return G;
}
// The SDF grid on the simulation's stream, in caller-given working arrays of total voxels each:
// key (u64), g1, g2 (double), fixedCell (u8).  key may share g2's memory: the keys are dead
// once g1 holds the splatted distances.  Returns the array holding the result (g1 or g2).
static double* sdfGridCompute(pd_sim* s, const SdfGridP& G, unsigned long long* key, double* g1,
                              double* g2, unsigned char* fixedCell) {
  const long total = (long)G.rx * G.ry * G.rz;
  const hipStream_t st = s->st;
  const int nb = (int)((total + TPB - 1) / TPB);
  k_fill_u64<<<nb, TPB, 0, st>>>(key, ordKey(100.0), (int)total);
  k_fill_u8<<<nb, TPB, 0, st>>>(fixedCell, 0, total);
  s->refreshDev();
  Params P = s->params(0.0, 0.0);
  if (s->numReal > 0) k_sdf_splat<<<nblk(s->numReal), TPB, 0, st>>>(s->dD, P, G, key, fixedCell);
  k_sdf_from_keys<<<nb, TPB, 0, st>>>(key, g1, total);
  const int iters = std::max(std::max(G.rx, G.ry), G.rz) * 4;
  double* in = g1;
  double* o2 = g2;
  for (int it = 0; it < iters; ++it) {
    k_sdf_eikonal<<<nb, TPB, 0, st>>>(G, in, o2, fixedCell, G.vox.x);
    std::swap(in, o2);
  }
  HIPCHECK(hipGetLastError());
  return in;
}
static void sdfGrid(pd_sim* s, int rx, int ry, int rz, std::vector<double>& out) {
  const SdfGridP G = sdfGridParams(s, rx, ry, rz);
  const long total = (long)rx * ry * rz;
  unsigned long long* key = nullptr;
  double *g1 = nullptr, *g2 = nullptr;
  unsigned char* fixedCell = nullptr;
  HIPCHECK(hipMalloc(&key, total * sizeof(unsigned long long)));
  HIPCHECK(hipMalloc(&g1, total * sizeof(double)));
  HIPCHECK(hipMalloc(&g2, total * sizeof(double)));
  HIPCHECK(hipMalloc(&fixedCell, total));
  try {
    const double* res = sdfGridCompute(s, G, key, g1, g2, fixedCell);
    out.resize(total);
    s->readN(out.data(), res, total);
  } catch (...) {
    (void)hipFree(key); (void)hipFree(g1); (void)hipFree(g2); (void)hipFree(fixedCell);
    throw;
  }
  HIPCHECK(hipFree(key));
  HIPCHECK(hipFree(g1));
  HIPCHECK(hipFree(g2));
  HIPCHECK(hipFree(fixedCell));
}

// The SDF export without allocation: its working arrays live in `scratch`, one file-scope device
// array per GPU (held until the process ends): g1 and g2 (8 bytes per voxel each; the keys
// share g2) and the fixed-cell flags (1 byte per voxel), for at most SDF_SCRATCH_VOXELS voxels.
constexpr long SDF_SCRATCH_VOXELS = 1L << 21;  // 2,097,152 voxels, e.g. 128 x 128 x 128
constexpr size_t SDF_SCRATCH_BYTES = (size_t)SDF_SCRATCH_VOXELS * 17;
static __device__ __align__(256) unsigned char scratch[SDF_SCRATCH_BYTES];
// `scratch` names a host-side stand-in in host code; its device address, per device, comes from
// hipGetSymbolAddress on the first call made on that device and is kept here
std::mutex g_scratchMutex;
constexpr int SCRATCH_MAX_DEVICES = 64;
void* g_scratchAddr[SCRATCH_MAX_DEVICES] = {};
static void* scratchAddress(int device) {
  if (device < 0 || device >= SCRATCH_MAX_DEVICES)
    throw std::runtime_error("scratch: device index out of range");
  std::lock_guard<std::mutex> lk(g_scratchMutex);
  if (!g_scratchAddr[device]) {
    int cur = 0;
    HIPCHECK(hipGetDevice(&cur));
    if (cur != device) HIPCHECK(hipSetDevice(device));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) HIPCHECK(hipSetDevice(cur));
    if (e != hipSuccess || !p)
      throw std::runtime_error(std::string("scratch: hipGetSymbolAddress failed: ") +
                               hipGetErrorString(e));
    g_scratchAddr[device] = p;
  }
  return g_scratchAddr[device];
}

int pd_get_sdf_grid(pd_sim* s, int rx, int ry, int rz, double* out) {
  return guard([&] {
    std::vector<double> g;
    sdfGrid(s, rx, ry, rz, g);
    if (out) std::memcpy(out, g.data(), g.size() * sizeof(double));
  });
}
int pd_get_sdf_grid_noalloc(pd_sim* s, int rx, int ry, int rz, double* out_device) {
  return guard([&] {
    const SdfGridP G = sdfGridParams(s, rx, ry, rz);
    const long total = (long)rx * ry * rz;
    if (total > SDF_SCRATCH_VOXELS)
      throw ArgError("get_sdf_grid_noalloc: " + std::to_string(rx) + " x " + std::to_string(ry) +
                     " x " + std::to_string(rz) + " = " + std::to_string(total) +
                     " voxels exceed the scratch capacity of " +
                     std::to_string(SDF_SCRATCH_VOXELS) + " voxels");
    if (!out_device) throw ArgError("get_sdf_grid_noalloc: out_device is required");
    unsigned char* base = static_cast<unsigned char*>(scratchAddress(s->device));
    double* g1 = reinterpret_cast<double*>(base);
    double* g2 = reinterpret_cast<double*>(base + SDF_SCRATCH_VOXELS * sizeof(double));
    unsigned long long* key = reinterpret_cast<unsigned long long*>(g2);
    unsigned char* fixedCell = base + 2 * SDF_SCRATCH_VOXELS * sizeof(double);
    const double* res = sdfGridCompute(s, G, key, g1, g2, fixedCell);
    HIPCHECK(hipMemcpyAsync(out_device, res, total * sizeof(double), hipMemcpyDeviceToDevice,
                            s->st));
    s->sync();
  });
}
// writeSdfVti, src/io.hpp:86-110 (ASCII ImageData, Float32 samples)
int pd_export_sdf(pd_sim* s, const char* filename, int rx, int ry, int rz) {
  return guard([&] {
    std::vector<double> g;
    sdfGrid(s, rx, ry, rz, g);
    const float mn[3] = {(float)s->dmin[0], (float)s->dmin[1], (float)s->dmin[2]};
    const float mx[3] = {(float)s->dmax[0], (float)s->dmax[1], (float)s->dmax[2]};
    const double sx = rx > 1 ? (double(mx[0]) - mn[0]) / (rx - 1) : 1.0;
    const double sy = ry > 1 ? (double(mx[1]) - mn[1]) / (ry - 1) : 1.0;
    const double sz = rz > 1 ? (double(mx[2]) - mn[2]) / (rz - 1) : 1.0;
    std::ofstream f(filename);
    if (!f) throw std::runtime_error(std::string("Could not open file for writing: ") + filename);
    f << "<?xml version=\"1.0\"?>\n";
    f << "<VTKFile type=\"ImageData\" version=\"0.1\" byte_order=\"LittleEndian\">\n";
    f << "  <ImageData WholeExtent=\"0 " << (rx - 1) << " 0 " << (ry - 1) << " 0 " << (rz - 1)
      << "\" Origin=\"" << mn[0] << " " << mn[1] << " " << mn[2] << "\" Spacing=\"" << sx << " "
      << sy << " " << sz << "\">\n";
    f << "    <Piece Extent=\"0 " << (rx - 1) << " 0 " << (ry - 1) << " 0 " << (rz - 1) << "\">\n";
    f << "      <PointData Scalars=\"sdf\">\n";
    f << "        <DataArray type=\"Float32\" Name=\"sdf\" format=\"ascii\">\n";
    for (double v : g) f << (float)v << " ";
    f << "\n        </DataArray>\n";
    f << "      </PointData>\n";
    f << "    </Piece>\n";
    f << "  </ImageData>\n";
    f << "</VTKFile>\n";
  });
}
int pd_set_hertz_material(pd_sim* s, int mat, double youngs, double poisson) {
  return guard([&] {
    if (mat < 0 || mat >= 8) throw ArgError("material id out of range");
    s->hertzE[mat] = youngs;
    s->hertzNu[mat] = poisson;
  });
}
int pd_step_hertz(pd_sim* s, double dt, int nsteps, double skin_frac) {
  return guard([&] { s->stepHertz(dt, nsteps, skin_frac); });
}
int pd_add_sdf_wall(pd_sim* s, const double* grid, int nx, int ny, int nz, double ox, double oy,
                    double oz, double sx, double sy, double sz, double restitution,
                    double friction) {
  int idx = -1;
  const int rc = guard([&] {
    if (nx < 2 || ny < 2 || nz < 2)
      throw std::runtime_error("addSdfWall: grid dims must be >= 2 on each axis");
    if (s->wallsHost.size() >= 8) throw std::runtime_error("addSdfWall: at most 8 walls");
    const long cells = (long)nx * ny * nz;
    WallP w{};
    w.nx = nx;
    w.ny = ny;
    w.nz = nz;
    w.materialId = -1;
    w.off = (long)s->wallGridHost.size();
    w.origin = d3{ox, oy, oz};
    w.invSpacing = d3{sx > 0 ? 1.0 / sx : 0.0, sy > 0 ? 1.0 / sy : 0.0, sz > 0 ? 1.0 / sz : 0.0};
    w.restitution = restitution;
    w.friction = friction;
    s->wallGridHost.insert(s->wallGridHost.end(), grid, grid + cells);
    s->wallGridD.ensure(s->wallGridHost.size(), s->st);
    s->writeN(s->wallGridD.p, s->wallGridHost.data(), s->wallGridHost.size());
    idx = (int)s->wallsHost.size();
    s->wallsHost.push_back(w);
  });
  return rc == 0 ? idx : rc;
}
int pd_set_wall_velocity(pd_sim* s, int wid, double lx, double ly, double lz, double ax,
                         double ay, double az, double cx, double cy, double cz) {
  return guard([&] {
    if (wid < 0 || wid >= (int)s->wallsHost.size())
      throw std::runtime_error("setWallVelocity: wall index out of range");
    WallP& w = s->wallsHost[wid];
    w.linVel = d3{lx, ly, lz};
    w.angVel = d3{ax, ay, az};
    w.center = d3{cx, cy, cz};
  });
}
int pd_set_wall_material_id(pd_sim* s, int wid, int mat) {
  return guard([&] {
    if (wid < 0 || wid >= (int)s->wallsHost.size()) throw ArgError("wall index out of range");
    s->wallsHost[wid].materialId = mat;
  });
}
int pd_num_walls(pd_sim* s) { return (int)s->wallsHost.size(); }
int pd_get_hertz_pairs(pd_sim* s, int* pairs, double* xi, int cap) {
  const int n = s->hNumPairs > 0 ? s->hNumPairs : 0;
  const int m = std::min(n, cap);
  if (m <= 0) return n;
  guard([&] {
    if (pairs) s->readN((int2*)pairs, s->hPairs.p, m);
    if (xi) {
      std::vector<double4> h(m);
      s->readN(h.data(), s->hXi.p, m);
      for (int k = 0; k < m; ++k) {
        xi[3 * k] = h[k].x;
        xi[3 * k + 1] = h[k].y;
        xi[3 * k + 2] = h[k].z;
      }
    }
  });
  return n;
}
int pd_step_n(pd_sim* s, double dt, int n) {
  return guard([&] {
    for (int k = 0; k < n; ++k) {
      s->dt = dt;
      s->demStep();
    }
    drainRetired();
  });
}
int pd_tape_begin(pd_sim* s, int every) {
  return guard([&] {
    if (every < 1) throw ArgError("checkpoint_every must be >= 1");
    s->sync();
    if (!s->tape) s->tape = new AdjTape();
    s->tape->clear();
    s->tape->every = every;
    s->tapeCapture();
    s->tape->on = true;
    s->taping = true;
  });
}
long pd_tape_generation(pd_sim* s) { return s->tape ? s->tape->generation : 0; }
int pd_tape_end(pd_sim* s) {
  return guard([&] {
    if (!s->tape || !s->tape->on) throw ArgError("no open tape (pd_tape_begin)");
    s->sync();
    s->tape->on = false;
    s->taping = false;
  });
}
int pd_tape_length(pd_sim* s) { return s->tape ? (int)s->tape->steps.size() : 0; }
int pd_tape_replay(pd_sim* s, double* state_out, long count) {
  return guard([&] {
    s->tapeCount(count);
    std::vector<double> out;
    adjrun::replayAll(*s, s->tape, out);
    if (state_out) std::memcpy(state_out, out.data(), out.size() * sizeof(double));
  });
}
int pd_tape_tangent(pd_sim* s, const double* dstate0, double dmu, double drest, double* state_out,
                    double* dstate_out, long count) {
  return guard([&] {
    s->tapeCount(count);
    const size_t n = s->tape ? (size_t)13 * s->tape->numReal : 0;
    std::vector<double> d0(n, 0.0), v, t;
    if (dstate0) std::memcpy(d0.data(), dstate0, n * sizeof(double));
    adjrun::tangentAll(*s, s->tape, d0, dmu, drest, v, t);
    if (state_out) std::memcpy(state_out, v.data(), n * sizeof(double));
    if (dstate_out) std::memcpy(dstate_out, t.data(), n * sizeof(double));
  });
}
int pd_tape_adjoint(pd_sim* s, const double* grad_state_end, double* grad_state0,
                    double* grad_params, long count) {
  return guard([&] {
    s->tapeCount(count);
    if (!grad_state_end) throw ArgError("grad_state_end is required");
    const size_t n = s->tape ? (size_t)13 * s->tape->numReal : 0;
    std::vector<double> gE(grad_state_end, grad_state_end + n), g0;
    double gp[2] = {0.0, 0.0};
    adjrun::adjointAll(*s, s->tape, gE, g0, gp);
    if (grad_state0) std::memcpy(grad_state0, g0.data(), n * sizeof(double));
    if (grad_params) {
      grad_params[0] = gp[0];
      grad_params[1] = gp[1];
    }
  });
}
int pd_compute_overlaps(pd_sim* s, double* out) {
  return guard([&] {
    const double v = s->computeOverlaps();
    if (out) *out = v;
  });
}
int pd_num_contacts(pd_sim* s) { return s->nc; }
int pd_num_manifolds(pd_sim* s) { return s->nm; }
double pd_max_overlap(pd_sim* s) { return s->maxOverlap; }
int pd_get_stats(pd_sim* s, pd_stats* out) {
  *out = s->stats;
  return 0;
}

int pd_get_pairs(pd_sim* s, int* out, int cap) {
  if (out) {
    const int n = std::min(cap, s->np);
    std::vector<int2> h(n);
    guard([&] { s->readN(h.data(), s->pairs.p, n); });
    for (int k = 0; k < n; ++k) {
      out[2 * k] = h[k].x;
      out[2 * k + 1] = h[k].y;
    }
  }
  return s->np;
}
int pd_get_contacts(pd_sim* s, int* ids, double* data, int cap) {
  const int n = std::min(cap, s->nc);
  if (n > 0 && (ids || data)) {
    std::vector<int2> ab(n);
    std::vector<double4> cn(n), ra(n), rb(n);
    guard([&] {
      s->readN(ab.data(), s->cAB.p, n);
      s->readN(cn.data(), s->cN.p, n);
      s->readN(ra.data(), s->cRA.p, n);
      s->readN(rb.data(), s->cRB.p, n);
    });
    for (int k = 0; k < n; ++k) {
      if (ids) {
        ids[2 * k] = ab[k].x;
        ids[2 * k + 1] = ab[k].y;
      }
      if (data) {
        double* d = data + 10 * k;
        d[0] = cn[k].x; d[1] = cn[k].y; d[2] = cn[k].z;
        d[3] = ra[k].x; d[4] = ra[k].y; d[5] = ra[k].z;
        d[6] = rb[k].x; d[7] = rb[k].y; d[8] = rb[k].z;
        d[9] = cn[k].w;
      }
    }
  }
  return s->nc;
}
int pd_get_contact_colors(pd_sim* s, int* out, int cap) {
  if (out) guard([&] { s->readN(out, s->cColor.p, std::min(cap, s->nc)); });
  return s->nc;
}
int pd_get_manifolds(pd_sim* s, int* ids, double* data, int cap) {
  const int n = std::min(cap, s->nm);
  if (n > 0 && (ids || data)) {
    std::vector<int2> ab(n), rr(n);
    std::vector<int> npt(n);
    std::vector<double4> f[5];
    for (auto& v : f) v.resize(n);
    guard([&] {
      s->readN(ab.data(), s->mAB.p, n);
      s->readN(rr.data(), s->mReal.p, n);
      s->readN(npt.data(), s->mNp.p, n);
      s->readN(f[0].data(), s->mNs.p, n);
      s->readN(f[1].data(), s->mTA.p, n);
      s->readN(f[2].data(), s->mTB.p, n);
      s->readN(f[3].data(), s->mRA.p, n);
      s->readN(f[4].data(), s->mRB.p, n);
    });
    for (int k = 0; k < n; ++k) {
      if (ids) {
        ids[5 * k] = ab[k].x;
        ids[5 * k + 1] = ab[k].y;
        ids[5 * k + 2] = npt[k];
        ids[5 * k + 3] = rr[k].x;
        ids[5 * k + 4] = rr[k].y;
      }
      if (data)
        for (int j = 0; j < 5; ++j) {
          data[15 * k + 3 * j] = f[j][k].x;
          data[15 * k + 3 * j + 1] = f[j][k].y;
          data[15 * k + 3 * j + 2] = f[j][k].z;
        }
    }
  }
  return s->nm;
}
int pd_get_manifold_colors(pd_sim* s, int* out, int cap) {
  if (out) guard([&] { s->readN(out, s->mColor.p, std::min(cap, s->nm)); });
  return s->nm;
}
int pd_get_ledger(pd_sim* s, uint64_t* keys, double* lambda, double* lambdaT, double* posImp,
                  int cap) {
  const int n = std::min(cap, s->prevCount);
  if (n > 0 && (keys || lambda || lambdaT || posImp)) {
    std::vector<unsigned long long> k(n);
    std::vector<double> l(n);
    std::vector<double4> lt(n);
    guard([&] {
      s->readN(k.data(), s->lKey.p, n);
      s->readN(l.data(), s->lLam.p, n);
      s->readN(lt.data(), s->lLamT.p, n);
    });
    for (int i = 0; i < n; ++i) {
      if (keys) keys[i] = k[i];
      if (lambda) lambda[i] = l[i];
      if (lambdaT) {
        lambdaT[3 * i] = lt[i].x;
        lambdaT[3 * i + 1] = lt[i].y;
        lambdaT[3 * i + 2] = lt[i].z;
      }
      if (posImp) posImp[i] = lt[i].w;
    }
  }
  return s->prevCount;
}
int pd_get_grounded(pd_sim* s, int* out, int n) {
  if (out) guard([&] { s->readN(out, s->ground.p, std::min(n, s->numReal)); });
  return s->numReal;
}

int pd_export_lammps(pd_sim* s, const char* filename, int step) {
  // exportLammps, src/sim.hpp:1463-1474
  return guard([&] {
    const int n = s->numReal;
    std::vector<double> p(3 * (size_t)n), v(3 * (size_t)n), q(4 * (size_t)n), sc(n), r(n);
    s->getField(0, p.data(), n);
    s->getField(1, v.data(), n);
    s->getField(3, q.data(), n);
    s->readN(sc.data(), s->scale.p, n);
    for (int i = 0; i < n; ++i) r[i] = sc[i] * s->globalScale * s->baseRadius;
    const bool pbc = s->per[0] || s->per[1] || s->per[2];
    writeLammps(filename, step, p.data(), v.data(), q.data(), r.data(), n, nullptr, nullptr, pbc);
  });
}
int pd_write_lammps_dump(const char* filename, int step, const double* pos, const double* vel,
                         const double* quat, const double* radii, int n, const double* bmn,
                         const double* bmx, int pbc) {
  return guard([&] { writeLammps(filename, step, pos, vel, quat, radii, n, bmn, bmx, pbc != 0); });
}
int pd_write_vtp(pd_sim* s, const char* filename) {
  // writeVtp, src/sim.hpp:1560-1601
  return guard([&] {
    const int n = s->numReal;
    std::vector<double> p(3 * (size_t)n), v(3 * (size_t)n), sc(n);
    s->getField(0, p.data(), n);
    s->getField(1, v.data(), n);
    s->readN(sc.data(), s->scale.p, n);
    std::ofstream out(filename);
    if (!out) throw std::runtime_error(std::string("Could not open file for writing: ") + filename);
    out << "<?xml version=\"1.0\"?>\n";
    out << "<VTKFile type=\"PolyData\" version=\"0.1\" byte_order=\"LittleEndian\">\n";
    out << "  <PolyData>\n";
    out << "    <Piece NumberOfPoints=\"" << n << "\" NumberOfVerts=\"0\" "
        << "NumberOfLines=\"0\" NumberOfStrips=\"0\" NumberOfPolys=\"0\">\n";
    out << "      <Points>\n";
    out << "        <DataArray type=\"Float32\" Name=\"Position\" NumberOfComponents=\"3\" "
           "format=\"ascii\">\n";
    for (int i = 0; i < n; ++i) out << p[3 * i] << " " << p[3 * i + 1] << " " << p[3 * i + 2] << " ";
    out << "\n        </DataArray>\n";
    out << "      </Points>\n";
    out << "      <PointData Scalars=\"Radius\">\n";
    out << "        <DataArray type=\"Float32\" Name=\"Radius\" NumberOfComponents=\"1\" "
           "format=\"ascii\">\n";
    for (int i = 0; i < n; ++i) out << sc[i] * s->globalScale * s->baseRadius << " ";
    out << "\n        </DataArray>\n";
    out << "        <DataArray type=\"Float32\" Name=\"Velocity\" NumberOfComponents=\"3\" "
           "format=\"ascii\">\n";
    for (int i = 0; i < n; ++i) out << v[3 * i] << " " << v[3 * i + 1] << " " << v[3 * i + 2] << " ";
    out << "\n        </DataArray>\n";
    out << "      </PointData>\n";
    out << "    </Piece>\n";
    out << "  </PolyData>\n";
    out << "</VTKFile>\n";
    std::printf("Exported VTP: %s\n", filename);
  });
}

// Host utility for seeded workloads (no reference counterpart): random sequential addition.
// Spheres i = 0..n-1 (radius radii[i]) take candidates from `cand` (3 per candidate, in draw
// order) until one overlaps no placed sphere; returns the number of candidates consumed, or -1 if
// the stream ran out (then *placed tells how many were placed; call again with more candidates
// and `start` = *placed to continue).  On an axis with periodic[d] != 0 (periodic may be null)
// the box [lo, hi) is one period: distances use the nearest periodic image, so spheres may
// straddle that face.
long pd_rsa_place(const double* cand, long ncand, const double* radii, int n, int start,
                  const double* lo, const double* hi, const int* periodic, double* out,
                  int* placed) {
  static thread_local std::vector<int> head, next;
  double rmax = 0.0;
  for (int i = 0; i < n; ++i) rmax = std::max(rmax, radii[i]);
  const double h = 2.0 * rmax > 0.0 ? 2.0 * rmax : 1.0;
  int dim[3], per[3];
  double L[3];
  for (int d = 0; d < 3; ++d) {
    dim[d] = std::max(1, (int)((hi[d] - lo[d]) / h));
    L[d] = hi[d] - lo[d];
    per[d] = periodic && periodic[d] && L[d] > 0.0;
  }
  const long ncell = (long)dim[0] * dim[1] * dim[2];
  head.assign(ncell, -1);
  next.assign(n, -1);
  auto cellOf = [&](const double* p, int c[3]) {
    for (int d = 0; d < 3; ++d) c[d] = std::min(dim[d] - 1, std::max(0, (int)((p[d] - lo[d]) / h)));
  };
  for (int i = 0; i < start; ++i) {
    int c[3];
    cellOf(out + 3 * i, c);
    const long k = ((long)c[2] * dim[1] + c[1]) * dim[0] + c[0];
    next[i] = head[k];
    head[k] = i;
  }
  // neighbour cell offsets per axis: -1..1, wrapped on periodic axes (each cell visited once)
  auto cells = [&](int d, int c, int* o) {
    int m = 0;
    for (int dc = -1; dc <= 1; ++dc) {
      int x = c + dc;
      if (per[d]) x = ((x % dim[d]) + dim[d]) % dim[d];
      else if (x < 0 || x >= dim[d]) continue;
      bool seen = false;
      for (int q = 0; q < m; ++q) seen |= o[q] == x;
      if (!seen) o[m++] = x;
    }
    return m;
  };
  long used = 0;
  int i = start;
  while (i < n) {
    if (used >= ncand) {
      *placed = i;
      return -1;
    }
    const double* p = cand + 3 * used;
    ++used;
    const double r = radii[i];
    int c[3], ox[3], oy[3], oz[3];
    cellOf(p, c);
    const int mx = cells(0, c[0], ox), my = cells(1, c[1], oy), mz = cells(2, c[2], oz);
    bool ok = true;
    for (int a = 0; a < mz && ok; ++a)
      for (int b = 0; b < my && ok; ++b)
        for (int e = 0; e < mx && ok; ++e) {
          for (int j = head[((long)oz[a] * dim[1] + oy[b]) * dim[0] + ox[e]]; j >= 0; j = next[j]) {
            double dv[3];
            for (int d = 0; d < 3; ++d) {
              dv[d] = out[3 * j + d] - p[d];
              if (per[d]) dv[d] -= L[d] * std::nearbyint(dv[d] / L[d]);
            }
            const double rr = r + radii[j];
            if (dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2] < rr * rr) {
              ok = false;
              break;
            }
          }
        }
    if (!ok) continue;
    out[3 * i] = p[0];
    out[3 * i + 1] = p[1];
    out[3 * i + 2] = p[2];
    const long k = ((long)c[2] * dim[1] + c[1]) * dim[0] + c[0];
    next[i] = head[k];
    head[k] = i;
    ++i;
  }
  *placed = n;
  return used;
}

int pd_set_tuning(pd_sim* s, const char* name, long value) {
  return guard([&] {
    const std::string k = name ? name : "";
    if (k == "block_solver_rows") {
      if (value < 0) throw ArgError("block_solver_rows must be >= 0");
      s->blockRows = (int)std::min<long>(value, 1L << 30);
    } else if (k == "velocity_loop_lds") {
      if (value != 0 && value != 1) throw ArgError("velocity_loop_lds must be 0 or 1");
      s->ldsLoop = value != 0;
    } else if (k == "velocity_loop_lds_bodies") {
      if (value < 1 || value > VL_BODIES)
        throw ArgError("velocity_loop_lds_bodies must be in [1, " + std::to_string(VL_BODIES) + "]");
      s->ldsBodies = (int)value;
    } else if (k == "velocity_loop_lds_keys") {
      if (value < 1 || value > VL_KEYS)
        throw ArgError("velocity_loop_lds_keys must be in [1, " + std::to_string(VL_KEYS) + "]");
      s->ldsKeys = (int)value;
    } else if (k == "candidate_skin_permille") {
      if (value < 0) throw ArgError("candidate_skin_permille must be >= 0");
      s->candSkinFrac = std::min<long>(value, 100000) / 1000.0;
      s->candValid = false;
    } else if (k == "slot_headroom") {
      if (value < 0) throw ArgError("slot_headroom must be >= 0");
      s->slotHeadroom = std::min<long>(value, 1L << 30);
    } else if (k == "vsweep_waves") {
      if (value != 3 && value != 4) throw ArgError("vsweep_waves must be 3 or 4");
      s->vsweepWaves4 = value == 4;
    } else if (k == "serial_class_waves_min_rows") {
      if (value < 0) throw ArgError("serial_class_waves_min_rows must be >= 0");
      s->waveMinRows = value;
    } else if (k == "serial_class_wave_launch_rows") {
      if (value < 1) throw ArgError("serial_class_wave_launch_rows must be >= 1");
      s->waveLaunchRows = value;
    } else if (k == "colouring_incidence_min_edges") {
      if (value < 0) throw ArgError("colouring_incidence_min_edges must be >= 0");
      s->incMinEdges = value;
    } else if (k == "adjoint_keep_segment") {
      if (value != 0 && value != 1) throw ArgError("adjoint_keep_segment must be 0 or 1");
      s->adjKeepSegment = value == 1;
    } else if (k == "side_stream_min_contacts") {
      if (value < 0) throw ArgError("side_stream_min_contacts must be >= 0");
      s->sideStreamMinContacts = value;
    } else if (k == "velocity_tail_rows") {
      if (value < 0) throw ArgError("velocity_tail_rows must be >= 0");
      s->velTailRows = value;
    } else if (k == "position_tail_rows") {
      if (value < 0) throw ArgError("position_tail_rows must be >= 0");
      s->posTailRows = value;
    } else if (k == "velocity_pack_bodies") {
      if (value != 0 && value != 1) throw ArgError("velocity_pack_bodies must be 0 or 1");
      s->packBodies = value == 1;
    } else if (k == "timing_every") {
      if (value < 1) throw ArgError("timing_every must be >= 1");
      s->timingEvery = (int)std::min<long>(value, 1L << 30);
    } else {
      throw ArgError("unknown tuning parameter: " + k);
    }
  });
}
int pd_set_timing(pd_sim* s, int level) {
  return guard([&] { s->resetTiming(level); });
}
int pd_get_kernel_times(pd_sim* s, double* ms, long* launches, long* rows, int cap) {
  const int rc = guard([&] { s->resolveTiming(); });
  if (rc) return rc;
  for (int k = 0; k < pd_sim::K_COUNT && k < cap; ++k) {
    ms[k] = s->kernMs[k];
    launches[k] = s->kernLaunches[k];
    rows[k] = s->kernRows[k];
  }
  return pd_sim::K_COUNT;
}
int pd_get_stage_times(pd_sim* s, double* out, int cap) {
  const int rc = guard([&] { s->resolveTiming(); });
  if (rc) return rc;
  for (int k = 0; k < ST_COUNT && k < cap; ++k) out[k] = s->stageMs[k];
  return ST_COUNT;
}
int pd_get_stage_host_times(pd_sim* s, double* out, int cap) {
  for (int k = 0; k < ST_COUNT && k < cap; ++k) out[k] = s->stageHostMs[k];
  return ST_COUNT;
}
int pd_get_stats_log(pd_sim* s, pd_stats* out, int cap) {
  const int n = (int)s->statsLog.size();
  if (out)
    for (int k = 0; k < n && k < cap; ++k) out[k] = s->statsLog[k];
  return n;
}
const char* pd_kernel_name(int k) {
  static const char* names[pd_sim::K_COUNT] = {
      "k_vsweep",  "k_psweep",      "k_bp_query",     "k_np_pairs",
      "k_predict", "k_reintegrate", "k_vsolve_lds", "k_psolve_block"};
  return (k >= 0 && k < pd_sim::K_COUNT) ? names[k] : "";
}
const char* pd_stage_name(int stage) {
  return (stage >= 0 && stage < ST_COUNT) ? kStageNames[stage] : "";
}

}  // extern "C"

// ============================================================================================
// batched ensembles: independent systems, each with its own stream and buffers, advanced
// concurrently by a pool of host workers (each worker drives one system at a time, so the
// GPU overlaps the kernels and read-backs of many small systems)
// ============================================================================================
struct pd_batch {
  std::vector<pd_sim*> sims;
  int workers = 16;
};

pd_batch* pd_batch_create(int n, const int* capacities) {
  if (n < 0) {
    g_error = "batch size must be >= 0";
    return nullptr;
  }
  pd_batch* b = new pd_batch();
  for (int k = 0; k < n; ++k) {
    pd_sim* s = pd_create(capacities ? capacities[k] : 0);
    if (!s) {
      for (pd_sim* q : b->sims) delete q;
      delete b;
      return nullptr;
    }
    b->sims.push_back(s);
  }
  return b;
}
void pd_batch_destroy(pd_batch* b) {
  if (!b) return;
  for (pd_sim* s : b->sims) delete s;
  delete b;
}
int pd_batch_size(pd_batch* b) { return (int)b->sims.size(); }
pd_sim* pd_batch_system(pd_batch* b, int k) {
  if (k < 0 || k >= (int)b->sims.size()) {
    g_error = "system index out of range";
    return nullptr;
  }
  return b->sims[k];
}
// every system takes the autotuned setting of its size class; the largest system of a class
// is tuned (stepping it), the others reuse the cached choice
int pd_batch_autotune(pd_batch* b, double dt, int substeps) {
  std::vector<int> order(b->sims.size());
  for (size_t k = 0; k < order.size(); ++k) order[k] = (int)k;
  std::stable_sort(order.begin(), order.end(),
                   [&](int x, int y) { return b->sims[x]->numReal > b->sims[y]->numReal; });
  for (int k : order) {
    const int rc = pd_autotune(b->sims[k], dt, substeps, 1, nullptr, nullptr);
    if (rc != 0) {
      g_error = "system " + std::to_string(k) + ": " + g_error;
      return rc;
    }
  }
  return 0;
}
int pd_batch_set_workers(pd_batch* b, int workers) {
  if (workers < 1) {
    g_error = "workers must be >= 1";
    return -2;
  }
  b->workers = workers;
  return 0;
}
int pd_batch_step(pd_batch* b, double dt, int nsteps) {
  const int S = (int)b->sims.size();
  std::atomic<int> next{0};
  std::mutex mtx;
  std::string err;
  auto work = [&] {
    for (;;) {
      const int k = next.fetch_add(1);
      if (k >= S) break;
      pd_sim* s = b->sims[k];
      try {
        for (int t = 0; t < nsteps; ++t) {
          s->dt = dt;
          s->demStep();
        }
      } catch (const std::exception& e) {
        std::lock_guard<std::mutex> lk(mtx);
        if (err.empty()) err = "system " + std::to_string(k) + ": " + e.what();
      }
    }
  };
  const int nt = std::max(1, std::min(b->workers, S));
  // buffers the workers replace are freed after they have all finished (drainRetired)
  ++g_batchDepth;
  std::vector<std::thread> pool;
  for (int w = 1; w < nt; ++w) pool.emplace_back(work);
  work();
  for (auto& th : pool) th.join();
  --g_batchDepth;
  if (err.empty()) {
    try {
      drainRetired();
    } catch (const std::exception& e) {
      err = e.what();
    }
  }
  if (!err.empty()) {
    g_error = err;
    return -1;
  }
  return 0;
}
