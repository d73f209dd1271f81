// pd_kernels.hpp — HIP kernels of the XPBD substep (gfx950, FP64, wave64).
//
// Data layout in HBM (all FP64, 32-byte double4 rows):
//   per slot (reals then periodic ghosts):
//     pos[s]      committed position            vw0[s]  committed {v, w} (64 B record)
//     bc[s]       {q, (invI.xyz, invMass)}      vw[s]   predicted {v, w} (64 B record)
//     posPred[s]  predicted position, .w = broadphase radius       quatPred[s]
//   contacts (narrowphase order): cAB int2, cN (n, dist), cRA (rA, e_b), cRB (rB, mu_b)
//   manifolds (key order): mAB, mReal int2, mRange int2, mNs (Nsum, e_sum), mTA (tauA, mu_sum),
//                          mTB, mRA, mRB double4, mNp int
//   velocity rows (colour-sorted): vAB int2, vN (Nsum, e), vRA (rA, mu), vRB (rB, sign),
//                                  vTA, vTB (tauB, 1/points) (several-point manifolds only),
//                                  vVt0, vLamT double4, vVn0, vLam double, vFlag uchar
//   position rows (colour-sorted): pAB int2, pN, pRA, pRB double4, pLam double
// Every kernel cites the reference function it implements; the canonical orders are those of the
// CPU oracle (ghost slots by real index, pairs lexicographic, contacts pair-major then planes,
// manifolds by key, warm start by colour class, double-buffered grounded sweeps).
#pragma once

#include "pd_math.hpp"

namespace pd {

constexpr int TPB = 256;

struct Body64 {
  double4 v;
  double4 w;
};
struct BodyC {
  double4 q;
  double4 iim;  // (invI.x, invI.y, invI.z, invMass)
};

struct Params {
  double dt, gs, baseRadius, margin, band, boxCap, growthRate;
  double skin;  // broadphase boxes grown by this much (candidate-list rebuilds), else 0
  double restN, restT, mu;
  d3 g;
  d3 dmin, dmax, dsize;
  int px, py, pz;
  int numReal, numParticles;
  int numPlanes;
  int hasPairTable;
  ShapeP shape;
  double4 planeP[8];
  double4 planeN[8];
};

struct Grid {
  d3 lo;
  double h;
  int dim[3];
};

// The device pointer table below lives in device memory and kernels read it through a pointer,
// so a pointer member would otherwise be a generic (flat) address: flat accesses also count
// against the LDS / scalar-load wait counter and take a 64-bit VGPR address.  In the device pass
// the members are declared in the global address space, so accesses through them compile to
// global_* instructions; the host pass sees plain pointers of the same size and layout.
#if defined(__HIP_DEVICE_COMPILE__)
#define PD_G __attribute__((address_space(1)))
#else
#define PD_G
#endif

// device pointers of one simulation
struct Dev {
  // per slot
  PD_G double4* pos;
  PD_G Body64* vw0;
  PD_G BodyC* bc;
  PD_G double4* posPred;
  PD_G double4* quatPred;
  PD_G Body64* vw;
  PD_G double* scale;
  PD_G double* targetScale;
  PD_G double4* extForce;
  PD_G int* realIdx;
  PD_G unsigned char* matId;
  PD_G int* ground;
  PD_G int* ground2;
  PD_G double4* shell;
  PD_G double* pairTable;
  // broadphase
  PD_G int* cellOf;
  PD_G int* cellRank;
  PD_G int* cellCnt;
  PD_G int* cellStart;
  PD_G int* order;
  PD_G double4* sbox;
  PD_G int* pairCnt;
  PD_G int* pairStart;
  PD_G int2* pairs;
  // narrowphase
  PD_G int* ncnt;  // [numPairs + numReal] contact counts (pairs, then planes per real)
  PD_G int* noff;  // exclusive scan of ncnt
  PD_G int2* cAB;
  PD_G double4* cN;
  PD_G double4* cRA;
  PD_G double4* cRB;
  PD_G double4* cBV;   // boundary (wall) surface velocity at the contact, zero otherwise
  PD_G int* cSlot;
  // manifolds
  PD_G int* mCnt;
  PD_G int* mStart;
  PD_G int2* mAB;
  PD_G int2* mReal;
  PD_G int2* mRange;
  PD_G int2* mRange2;  // boundary manifolds: the wall-contact range (planes are mRange)
  PD_G double4* mNs;
  PD_G double4* mTA;
  PD_G double4* mTB;
  PD_G double4* mRA;
  PD_G double4* mRB;
  PD_G double4* mWV;   // sum of the boundary velocities of the active contacts
  PD_G int* mNp;
  PD_G unsigned long long* mKey;
  PD_G unsigned char* mCode;
  // colouring
  PD_G int* mColor;
  PD_G int* cColor;
  PD_G int* vPerm;
  PD_G int* vRowOf;
  PD_G int* pPerm;
  PD_G int* pRowOf;
  // velocity rows
  PD_G int2* vAB;
  PD_G double4* vN;
  PD_G double4* vTA;
  PD_G double4* vTB;
  PD_G double4* vRA;
  PD_G double4* vRB;
  PD_G double4* vVt0;
  PD_G double4* vLamT;
  PD_G double4* vWV;   // boundary rows: wall velocity sum (the body-B velocity is vWV / points)
  PD_G double* vVn0;
  PD_G double* vLam;
  PD_G unsigned char* vFlag;
  // position rows
  PD_G int2* pAB;
  PD_G double4* pN;
  PD_G double4* pRA;
  PD_G double4* pRB;
  PD_G double* pLam;
  // warm-start ledger (sorted by key)
  PD_G unsigned long long* lKey;
  PD_G double* lLam;
  PD_G double4* lLamT;  // (lambdaT.xyz, position-channel impulse)
  PD_G int* lIdx;
  PD_G unsigned long long* uKey;
  PD_G int* bucketCnt;
  PD_G int* bucketStart;
  // scalars
  PD_G unsigned long long* red;  // reduction scratch (ordered keys)
  PD_G int* counters;            // [0] remaining JP edges, [1] uncoloured entering, [2..] misc
  PD_G unsigned long long* maxOverlapBits;
  PD_G unsigned long long* maxApproachBits;
  PD_G double* thermo;           // partial sums
  // Jacobi fallback accumulators (per slot)
  PD_G double4* dVel;
  PD_G double4* dAng;
  PD_G double4* dPos;
  PD_G int* dCnt;
  // legacy count-averaged friction (per contact normal impulse, per body contact count)
  PD_G double* cLamN;
  PD_G int* frCnt;
};

// velocity-row flag bits
constexpr unsigned char F_PERSIST = 1, F_SIDE1 = 2, F_SIDE2 = 4, F_UP = 8, F_DOWN = 16;
// a one-point manifold's row: its tangent lever arms are recomputed instead of read (rowTaus)
constexpr unsigned char F_ONE = 32;

__device__ __forceinline__ int gtid() { return blockIdx.x * blockDim.x + threadIdx.x; }
// Workgroups are dispatched to the 8 XCDs in turn (workgroup i to XCD i mod 8), each XCD with
// its own L2.  Launched over a grid that is a multiple of 8 (xcdGrid), this maps the workgroups
// so that XCD x takes the x-th contiguous eighth of the range: rows placed in edge order then
// keep their bodies in one XCD's L2 across the colour launches of a sweep.
constexpr int NUM_XCD = 8;
__device__ __forceinline__ int xcdTid() {
  const int per = gridDim.x / NUM_XCD;
  return ((blockIdx.x % NUM_XCD) * per + blockIdx.x / NUM_XCD) * blockDim.x + threadIdx.x;
}

// ---- wave64 / block reductions: one global atomic per block instead of one per thread ----
// block-aggregated append of `item` (when `take`) to list[*count]; every thread must call it
__device__ __forceinline__ void blockAppend(bool take, int item, int* list, int* count, int* shCnt,
                                            int* shBase) {
  if (threadIdx.x == 0) *shCnt = 0;
  __syncthreads();
  const int local = take ? atomicAdd(shCnt, 1) : 0;
  __syncthreads();
  if (threadIdx.x == 0 && *shCnt) *shBase = atomicAdd(count, *shCnt);
  __syncthreads();
  if (take) list[*shBase + local] = item;
}
// Chunked appends: a workgroup walks one contiguous chunk of the index range (chunkOf), counts
// what it appends, takes its base with one atomic and places its items in a second walk by their
// rank within the workgroup (blockRank).  Thousands of workgroups each taking one atomic per
// 256 items on the same counter serialise across the XCDs; a few hundred taking one each do not.
__device__ __forceinline__ int2 chunkOf(int n) {
  const int per = ((n + (int)gridDim.x - 1) / (int)gridDim.x + (int)blockDim.x - 1) /
                  (int)blockDim.x * (int)blockDim.x;
  const int c0 = (int)blockIdx.x * per;
  return make_int2(c0, min(c0 + per, n));
}
// exclusive rank of `take` among the workgroup's threads and the workgroup total; every thread
// of the workgroup must call it
__device__ __forceinline__ int blockRank(bool take, int* shWave, int& total) {
  const unsigned long long m = __ballot(take);
  const int lane = (int)(threadIdx.x & 63), wave = (int)(threadIdx.x >> 6);
  if (lane == 0) shWave[wave] = __popcll(m);
  __syncthreads();
  int before = 0;
  total = 0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
    const int c = shWave[w];
    if (w < wave) before += c;
    total += c;
  }
  __syncthreads();
  return before + __popcll(m & ((1ull << lane) - 1ull));
}
// block sum of v (every thread); the result is valid in every thread
__device__ __forceinline__ int blockSumI(int v, int* shWave) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int lane = (int)(threadIdx.x & 63), wave = (int)(threadIdx.x >> 6);
  if (lane == 0) shWave[wave] = v;
  __syncthreads();
  int t = 0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += shWave[w];
  __syncthreads();
  return t;
}
constexpr int CHUNK_GRID = 1024;  // workgroups of the chunked-append kernels
constexpr int WAVES = TPB / 64;
__device__ __forceinline__ unsigned long long waveMaxU64(unsigned long long v) {
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(v, off, 64);
    v = o > v ? o : v;
  }
  return v;
}
__device__ __forceinline__ int waveSumI(int v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
// every thread of the block must call these (no early return before them)
__device__ __forceinline__ void blockAtomicMaxU64(unsigned long long* addr, unsigned long long v,
                                                  unsigned long long* sh) {
  v = waveMaxU64(v);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long m = sh[0];
    for (int w = 1; w < WAVES; ++w) m = sh[w] > m ? sh[w] : m;
    // the maximum only grows during a launch: a block below the value already there skips its
    // atomic (a stale, lower read only costs an atomic that changes nothing)
    if (m && m > __hip_atomic_load(addr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      atomicMax(addr, m);
  }
}
__device__ __forceinline__ void blockAtomicAddI(int* addr, int v, int* sh) {
  v = waveSumI(v);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    int m = 0;
    for (int w = 0; w < WAVES; ++w) m += sh[w];
    if (m) atomicAdd(addr, m);
  }
}
__device__ __forceinline__ void atomicMaxD(unsigned long long* a, double v) {
  // non-negative doubles order like their bit patterns
  atomicMax(a, dbits(v));
}

// ============================================================================================
// generic helpers: fills, exclusive scan
// ============================================================================================
__global__ void k_fill_i32(int* a, int v, int n) {
  const int i = gtid();
  if (i < n) a[i] = v;
}
__global__ void k_fill_u64(unsigned long long* a, unsigned long long v, int n) {
  const int i = gtid();
  if (i < n) a[i] = v;
}
// zero two arrays of n elements in one launch
__global__ void k_zero2_u64(unsigned long long* a, unsigned long long* b, int n) {
  const int i = gtid();
  if (i < n) {
    a[i] = 0ull;
    b[i] = 0ull;
  }
}
__global__ void k_fill_u8(unsigned char* a, unsigned char v, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = TPB * SCAN_ITEMS;

// the SCAN_ITEMS values at base (0 past n); in and out of every scan are 16-byte aligned buffer
// starts, so full groups are two int4 accesses
static_assert(SCAN_ITEMS == 8, "scanLoad reads two int4");
__device__ __forceinline__ void scanLoad(const int* in, int n, int base, int* vals) {
  if (base + SCAN_ITEMS <= n) {
    const int4* p = reinterpret_cast<const int4*>(in + base);
    const int4 a = p[0], b = p[1];
    vals[0] = a.x; vals[1] = a.y; vals[2] = a.z; vals[3] = a.w;
    vals[4] = b.x; vals[5] = b.y; vals[6] = b.z; vals[7] = b.w;
  } else {
    for (int k = 0; k < SCAN_ITEMS; ++k) vals[k] = (base + k < n) ? in[base + k] : 0;
  }
}
__device__ int blockExclusiveScan(int v, int* sh, int* total) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int off = 1; off < TPB; off <<= 1) {
    const int x = (t >= off) ? sh[t - off] : 0;
    __syncthreads();
    sh[t] += x;
    __syncthreads();
  }
  const int incl = sh[t];
  *total = sh[TPB - 1];
  __syncthreads();
  return incl - v;
}

// nDev (optional): the element count on the device; n is then only the launch bound
__global__ void k_scan_reduce(const int* in, int n, int* blockSums, const int* nDev) {
  __shared__ int sh[TPB];
  if (nDev) n = *nDev;
  const int base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
  int vals[SCAN_ITEMS];
  scanLoad(in, n, base, vals);
  int s = 0;
  for (int k = 0; k < SCAN_ITEMS; ++k) s += vals[k];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int off = TPB / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) blockSums[blockIdx.x] = sh[0];
}
// scan of one tile; the tile's offset is the block-reduced sum of the earlier tiles' totals
// (written by k_scan_reduce), so no separate top-level pass is needed. The last tile writes
// the grand total to out[n].
// inv (optional): for the elements below invLimit, which hold 0 or 1, inv[out[k]] = k for every
// element holding 1 (the compaction of those elements)
__global__ void k_scan_down2(const int* in, int n, const int* blockSums, int* out,
                             const int* nDev, int* inv = nullptr, int invLimit = 0) {
  __shared__ int sh[TPB];
  if (nDev) n = *nDev;
  const int last = n > 0 ? (n - 1) / SCAN_TILE : 0;
  if ((int)blockIdx.x > last) return;
  int pre = 0;
  for (int i = threadIdx.x; i < (int)blockIdx.x; i += TPB) pre += blockSums[i];
  sh[threadIdx.x] = pre;
  __syncthreads();
  for (int off = TPB / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  const int prefix = sh[0];
  __syncthreads();
  const int base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
  int vals[SCAN_ITEMS];
  scanLoad(in, n, base, vals);
  int s = 0;
  for (int k = 0; k < SCAN_ITEMS; ++k) s += vals[k];
  int tot;
  int ex = blockExclusiveScan(s, sh, &tot) + prefix;
  int res[SCAN_ITEMS];
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    res[k] = ex;
    if (inv && base + k < invLimit && vals[k]) inv[ex] = base + k;
    ex += vals[k];
  }
  if (base + SCAN_ITEMS <= n) {
    int4* o = reinterpret_cast<int4*>(out + base);
    o[0] = make_int4(res[0], res[1], res[2], res[3]);
    o[1] = make_int4(res[4], res[5], res[6], res[7]);
  } else {
    for (int k = 0; k < SCAN_ITEMS; ++k)
      if (base + k < n) out[base + k] = res[k];
  }
  if ((int)blockIdx.x == last && threadIdx.x == 0) out[n] = prefix + tot;
}
// The same exclusive scan (total into out[n]) in one workgroup of SCAN1_THREADS, for lists of up
// to SCAN1_MAX elements (`in` and `out` distinct): each thread sums a contiguous chunk, one
// workgroup scan of the chunk sums, then each thread writes its chunk's prefixes.  One launch
// and one barrier instead of the two-kernel scan.
constexpr int SCAN1_THREADS = 1024;
constexpr int SCAN1_MAX = 32768;
__global__ __launch_bounds__(SCAN1_THREADS) void k_scan_single(const int* in, int n, int* out,
                                                               const int* nDev) {
  __shared__ int waveTot[SCAN1_THREADS / 64];
  if (nDev) n = *nDev;
  const int lane = (int)(threadIdx.x & 63), wave = (int)(threadIdx.x >> 6);
  const int per = (n + SCAN1_THREADS - 1) / SCAN1_THREADS;
  const int b0 = (int)threadIdx.x * per;
  const int e0 = min(b0 + per, n);
  int own = 0;
  for (int i = b0; i < e0; ++i) own += in[i];
  // inclusive scan across the wave, then across the waves' totals
  int inc = own;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int x = __shfl_up(inc, off, 64);
    if (lane >= off) inc += x;
  }
  if (lane == 63) waveTot[wave] = inc;
  __syncthreads();
  int before = 0, tot = 0;
  for (int w = 0; w < SCAN1_THREADS / 64; ++w) {
    const int t = waveTot[w];
    if (w < wave) before += t;
    tot += t;
  }
  int ex = before + inc - own;
  for (int i = b0; i < e0; ++i) {
    const int v = in[i];
    out[i] = ex;
    ex += v;
  }
  if (threadIdx.x == 0) out[n] = tot;
}

// ============================================================================================
// integration — src/integration.hpp
// ============================================================================================
// maxOwnedRadius reduction, src/sim.hpp:61-70 (ordered-key max, NaN skipped): per-block partials
// then k_max_final (no contended atomics)
__global__ __launch_bounds__(TPB) void k_max_scale(const double* scale, int n,
                                                   unsigned long long* part) {
  __shared__ unsigned long long sh[WAVES];
  unsigned long long k = 0;
  for (int i = gtid(); i < n; i += gridDim.x * blockDim.x) {
    const double s = scale[i];
    if (s == s) {
      const unsigned long long o = ordKey(s);
      k = o > k ? o : k;
    }
  }
  k = waveMaxU64(k);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = k;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long m = sh[0];
    for (int w = 1; w < WAVES; ++w) m = sh[w] > m ? sh[w] : m;
    part[blockIdx.x] = m;
  }
}
__global__ void k_max_final(const unsigned long long* part, int nb, unsigned long long* out) {
  __shared__ unsigned long long sh[WAVES];
  unsigned long long k = 0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) k = part[b] > k ? part[b] : k;
  k = waveMaxU64(k);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = k;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long m = sh[0];
    for (int w = 1; w < WAVES; ++w) m = sh[w] > m ? sh[w] : m;
    *out = m;
  }
}

// updateGrowthScalesKokkos, src/integration.hpp:275-283
__global__ void k_growth(double* scale, const double* target, double factor, int n) {
  const int i = gtid();
  if (i < n) scale[i] = target[i] * factor;
}

__device__ __forceinline__ void ghostShifts(const Params& P, d3 p, int& sx, int& sy, int& sz) {
  const double skin = P.band;
  sx = P.px ? ((p.x < P.dmin.x + skin) ? 1 : ((p.x > P.dmax.x - skin) ? -1 : 0)) : 0;
  sy = P.py ? ((p.y < P.dmin.y + skin) ? 1 : ((p.y > P.dmax.y - skin) ? -1 : 0)) : 0;
  sz = P.pz ? ((p.z < P.dmin.z + skin) ? 1 : ((p.z > P.dmax.z - skin) ? -1 : 0)) : 0;
}
// The per-substep control words the predict launch clears first (launches fewer): the
// candidate-check flag, the colouring counters, the support-list count, the sweep-loop flags and
// counters, the overlap accumulator; the grounded-level decay of the substep
// (src/contact_preprocessing.hpp:213-253, first step) when `ground` is given; and each real's
// ghost count (as k_ghost_count).
// One 128-byte record per body for the velocity stage of large beds: velocity, angular
// velocity, orientation and inverse mass / inertia in one cache line, so a row's body gather
// fetches two whole lines instead of four half-used ones (the two 64-byte records of the
// separate arrays).  Written by the predict launch (or packed from vw / bc before the stage) and
// unpacked after the stage.
struct alignas(128) VBody {
  double4 v, w, q, iim;
};
struct SubstepStart {
  int* candFlag;
  int* counters;
  int* supCnt;
  int* ctl;
  int nCtl;
  unsigned long long* overlapBits;
  int* ground;  // nullptr: no grounded levels this substep
  int decay;
  int* ghostCnt;  // ghost count per real (src/periodicity.hpp:43-94, pass 1), from pos
  VBody* vb;      // non-null: also write each real's packed velocity-stage record
};
// predictVelocityKokkos, src/integration.hpp:44-93 (+ realIndices and broadphase radius fill)
__global__ void k_predict(const Dev* __restrict__ Dp, Params P, SubstepStart S) {
  const Dev& D = *Dp;
  if (blockIdx.x == 0) {
    const int t = threadIdx.x;
    if (t < S.nCtl) S.ctl[t] = 0;
    if (t == 0) {
      *S.candFlag = 0;
      S.counters[0] = 0;
      S.counters[1] = 0;
      *S.supCnt = 0;
      *S.overlapBits = 0ull;
    }
  }
  const int i = gtid();
  if (i >= P.numReal) return;
  if (S.ground) {
    const int v = S.ground[i] - S.decay;
    S.ground[i] = v > 0 ? v : 0;
  }
  {
    int sx, sy, sz;
    ghostShifts(P, xyz(D.pos[i]), sx, sy, sz);
    S.ghostCnt[i] = (sx ? 2 : 1) * (sy ? 2 : 1) * (sz ? 2 : 1) - 1;
  }
  const BodyC b = D.bc[i];
  const Body64 v0 = D.vw0[i];
  const double invM = b.iim.w;
  const double dt = P.dt;
  d3 v = xyz(v0.v);
  if (invM > 0.0) {
    v = add3(v, scale3(P.g, dt));
    v = add3(v, scale3(xyz(D.extForce[i]), invM * dt));
  }
  d3 wpred = xyz(v0.w);
  if (invM > 0.0) {
    const d3 invI = xyz(b.iim);
    if (invI.x > 0.0 || invI.y > 0.0 || invI.z > 0.0) {
      const q4 q = toq(b.q);
      d3 wb = invRotate(q, xyz(v0.w));
      const d3 Ib{(invI.x > 1e-9) ? 1.0 / invI.x : 0.0, (invI.y > 1e-9) ? 1.0 / invI.y : 0.0,
                  (invI.z > 1e-9) ? 1.0 / invI.z : 0.0};
      const d3 Lb{Ib.x * wb.x, Ib.y * wb.y, Ib.z * wb.z};
      const d3 wxL = cross3(wb, Lb);
      const d3 alpha{-invI.x * wxL.x, -invI.y * wxL.y, -invI.z * wxL.z};
      wb = add3(wb, scale3(alpha, dt));
      wpred = rotate(q, wb);
    }
  }
  Body64 out;
  out.v = tod4(v, 0.0);
  out.w = tod4(wpred, 0.0);
  D.vw[i] = out;
  if (S.vb) {  // (field by field, as k_vb_pack)
    S.vb[i].v = out.v;
    S.vb[i].w = out.w;
    S.vb[i].q = b.q;
    S.vb[i].iim = b.iim;
  }
  d3 x = xyz(D.pos[i]);
  if (invM > 0.0) x = add3(x, scale3(v, dt));
  const double rad = D.scale[i] * P.gs * P.baseRadius;
  D.posPred[i] = tod4(x, rad);
  D.quatPred[i] = b.q;
  D.realIdx[i] = i;
}

// computeOverlapsKokkos preparation, src/sim.hpp:494-510: predicted := committed
__global__ void k_overlap_prep(const Dev* __restrict__ Dp, Params P) {
  const Dev& D = *Dp;
  const int i = gtid();
  if (i >= P.numReal) return;
  const double rad = D.scale[i] * P.gs * P.baseRadius;
  D.posPred[i] = tod4(xyz(D.pos[i]), rad);
  D.quatPred[i] = D.bc[i].q;
  D.vw[i] = D.vw0[i];
  D.realIdx[i] = i;
}

// generateGhostsKokkos, src/periodicity.hpp:43-94 — pass 1: ghost count per real
__global__ void k_ghost_count(const Dev* __restrict__ Dp, Params P, int* cnt) {
  const Dev& D = *Dp;
  const int i = gtid();
  if (i >= P.numReal) return;
  int sx, sy, sz;
  ghostShifts(P, xyz(D.pos[i]), sx, sy, sz);
  cnt[i] = (sx ? 2 : 1) * (sy ? 2 : 1) * (sz ? 2 : 1) - 1;
}
// pass 2: emit ghosts at numReal + off[i] in the (a,b,c) loop order
// slots at or past capSlots are not written (the host grows the buffers and emits again)
// slotInfo (optional): info[0] = reals + the ghosts that fit in capSlots, info[1] = 1 when some
// did not fit (from the ghost total off[numReal])
__global__ void k_ghost_emit(const Dev* __restrict__ Dp, Params P, const int* off, int capSlots,
                             int* slotInfo = nullptr) {
  const Dev& D = *Dp;
  const int i = gtid();
  if (slotInfo && i == 0) {
    const long n = (long)P.numReal + off[P.numReal];
    slotInfo[0] = (int)(n < capSlots ? n : capSlots);
    slotInfo[1] = n > capSlots ? 1 : 0;
  }
  if (i >= P.numReal) return;
  const d3 p = xyz(D.pos[i]);
  int sx, sy, sz;
  ghostShifts(P, p, sx, sy, sz);
  if (!sx && !sy && !sz) return;
  const int nx = sx == 0 ? 1 : 2, ny = sy == 0 ? 1 : 2, nz = sz == 0 ? 1 : 2;
  const double4 pp = D.posPred[i];
  const d3 pPred = xyz(pp);
  // (records copied member by member: whole local records would sit in scratch)
  const double4 bq = D.bc[i].q;
  const double4 giim = make_double4(0.0, 0.0, 0.0, D.bc[i].iim.w);
  const double4 v0v = D.vw0[i].v, v0w = D.vw0[i].w, v1v = D.vw[i].v, v1w = D.vw[i].w;
  const double4 qp = D.quatPred[i];
  const double sc = D.scale[i];
  const unsigned char mat = D.matId[i];
  int slot = P.numReal + off[i];
  for (int a = 0; a < nx; ++a)
    for (int bb = 0; bb < ny; ++bb)
      for (int c = 0; c < nz; ++c) {
        // (selects, not a small array indexed by the loop counters: that would sit in scratch)
        const int ix = a ? sx : 0, iy = bb ? sy : 0, iz = c ? sz : 0;
        if (ix == 0 && iy == 0 && iz == 0) continue;
        if (slot >= capSlots) return;
        const d3 shift{ix * P.dsize.x, iy * P.dsize.y, iz * P.dsize.z};
        D.pos[slot] = tod4(add3(p, shift), 0.0);
        D.posPred[slot] = tod4(add3(pPred, shift), pp.w);
        D.bc[slot].q = bq;
        D.bc[slot].iim = giim;
        D.vw0[slot].v = v0v;
        D.vw0[slot].w = v0w;
        D.vw[slot].v = v1v;
        D.vw[slot].w = v1w;
        D.quatPred[slot] = qp;
        D.scale[slot] = sc;
        D.matId[slot] = mat;
        D.realIdx[slot] = i;
        ++slot;
      }
}

// a few device int ranges copied in one launch to host-visible pinned memory (the host reads
// them after synchronising the stream): one small kernel instead of a runtime copy per range
struct PostSpec {
  const int* src[4];
  int len[4];
  int n;
};
// (after the values, one thread writes `seq` to *seqDst: the host polls that word instead of
// synchronising the stream)
__global__ void k_post(PostSpec S, int* dst, int* seqDst, int seq) {
  int off = 0;
  for (int r = 0; r < S.n; ++r) {
    for (int i = threadIdx.x; i < S.len[r]; i += blockDim.x) dst[off + i] = S.src[r][i];
    off += S.len[r];
  }
  __threadfence_system();  // this thread's values reach host memory
  __syncthreads();
  if (threadIdx.x == 0) *(volatile int*)seqDst = seq;
}
__global__ void k_set_int(int* dst, int v) { *dst = v; }

// applyVelocityAndPredictPositionKokkos, src/integration.hpp:112-148 (reals and ghosts)
__global__ void k_reintegrate(const Dev* __restrict__ Dp, Params P) {
  const Dev& D = *Dp;
  const int i = gtid();
  if (i >= P.numParticles) return;
  const double dt = P.dt;
  const Body64 vs = D.vw0[i];
  const Body64 vf = D.vw[i];
  const d3 vStart = xyz(vs.v);
  const d3 vFinal = xyz(vf.v);
  D.vw0[i] = vf;  // persist v_{n+1} and omega
  const BodyC b = D.bc[i];
  const double invM = b.iim.w;
  d3 x = xyz(D.pos[i]);
  if (invM > 0.0) x = add3(x, scale3(add3(vStart, vFinal), 0.5 * dt));
  D.posPred[i].x = x.x;
  D.posPred[i].y = x.y;
  D.posPred[i].z = x.z;
  const q4 q = toq(b.q);
  const d3 omega = xyz(vf.w);
  q4 qp = q;
  if (invM > 0.0) {
    qp.x += 0.5 * dt * (omega.x * q.w + omega.y * q.z - omega.z * q.y);
    qp.y += 0.5 * dt * (omega.y * q.w + omega.z * q.x - omega.x * q.z);
    qp.z += 0.5 * dt * (omega.z * q.w + omega.x * q.y - omega.y * q.x);
    qp.w += 0.5 * dt * (-omega.x * q.x - omega.y * q.y - omega.z * q.z);
    const double len = sqrt(qp.x * qp.x + qp.y * qp.y + qp.z * qp.z + qp.w * qp.w);
    if (len > 1e-9) {
      const double s = 1.0 / len;
      qp = q4{qp.x * s, qp.y * s, qp.z * s, qp.w * s};
    }
  }
  D.quatPred[i] = tod4(qp);
}

// finalCommitKokkos, src/integration.hpp:187-215
__global__ void k_final_commit(const Dev* __restrict__ Dp, Params P) {
  const Dev& D = *Dp;
  const int i = gtid();
  if (i >= P.numReal) return;
  d3 p = xyz(D.posPred[i]);
  if (P.px) {
    if (p.x < P.dmin.x) p.x += P.dsize.x;
    else if (p.x >= P.dmax.x) p.x -= P.dsize.x;
  }
  if (P.py) {
    if (p.y < P.dmin.y) p.y += P.dsize.y;
    else if (p.y >= P.dmax.y) p.y -= P.dsize.y;
  }
  if (P.pz) {
    if (p.z < P.dmin.z) p.z += P.dsize.z;
    else if (p.z >= P.dmax.z) p.z -= P.dsize.z;
  }
  D.pos[i] = tod4(p, 0.0);
  D.bc[i].q = D.quatPred[i];
}

// applyThermostatKokkos, src/integration.hpp:229-271 — per-block partial energies (fixed order)
__global__ void k_thermo_partial(const Dev* __restrict__ Dp, Params P, double* part) {
  const Dev& D = *Dp;
  __shared__ double st[TPB], sr[TPB];
  const int i = gtid();
  double kt = 0.0, kr = 0.0;
  if (i < P.numReal) {
    const BodyC b = D.bc[i];
    const Body64 v0 = D.vw0[i];
    const double invM = b.iim.w;
    const d3 v = xyz(v0.v);
    if (invM > 0.0) kt = 0.5 * (1.0 / invM) * (v.x * v.x + v.y * v.y + v.z * v.z);
    const d3 invI = xyz(b.iim);
    if (invI.x > 0.0 || invI.y > 0.0 || invI.z > 0.0) {
      const d3 wb = invRotate(toq(b.q), xyz(v0.w));
      const double Ix = (invI.x > 0.0) ? 1.0 / invI.x : 0.0;
      const double Iy = (invI.y > 0.0) ? 1.0 / invI.y : 0.0;
      const double Iz = (invI.z > 0.0) ? 1.0 / invI.z : 0.0;
      kr = 0.5 * (Ix * wb.x * wb.x + Iy * wb.y * wb.y + Iz * wb.z * wb.z);
    }
  }
  st[threadIdx.x] = kt;
  sr[threadIdx.x] = kr;
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, c = 0.0;
    for (int t = 0; t < TPB; ++t) {
      a += st[t];
      c += sr[t];
    }
    part[2 * blockIdx.x] = a;
    part[2 * blockIdx.x + 1] = c;
  }
}
__global__ void k_thermo_scale(const Dev* __restrict__ Dp, Params P, double lt, double lr) {
  const Dev& D = *Dp;
  const int i = gtid();
  if (i >= P.numReal) return;
  Body64 v = D.vw0[i];
  v.v = tod4(scale3(xyz(v.v), lt), 0.0);
  v.w = tod4(scale3(xyz(v.w), lr), 0.0);
  D.vw0[i] = v;
}

// ============================================================================================
// broadphase — src/broadphase_arborx.hpp:36-107, uniform grid, pairs (i<j) lexicographic
// ============================================================================================
// PECLET_DEM_SANE_BOX, src/broadphase_arborx.hpp:41-52
__device__ __forceinline__ double4 saneBox(double4 pp, double margin, double boxCap) {
  double x = pp.x, y = pp.y, z = pp.z, b = pp.w + margin;
  if (!(b == b) || !(x == x) || !(y == y) || !(z == z) || (boxCap > 0.0 && b > boxCap)) {
    x = y = z = -1.0e30;
    b = 0.0;
  }
  return make_double4(x, y, z, b);
}
// broadphase box: the sane box grown by P.skin unless parked (parking is the plain box's)
__device__ __forceinline__ double4 bpBox(double4 pp, const Params& P) {
  double4 B = saneBox(pp, P.margin, P.boxCap);
  if (P.skin > 0.0 && B.x != -1.0e30) B.w += P.skin;
  return B;
}
__device__ __forceinline__ bool boxOverlap(const double4& a, const double4& c) {
  if (a.x - a.w > c.x + c.w || a.x + a.w < c.x - c.w) return false;
  if (a.y - a.w > c.y + c.w || a.y + a.w < c.y - c.w) return false;
  if (a.z - a.w > c.z + c.w || a.z + a.w < c.z - c.w) return false;
  return true;
}

// bounds of the finite, non-parked box centres and the largest half-width
// red[0..2] = ordered min, red[3..5] = ordered max, red[6] = ordered max half-width
__device__ __forceinline__ unsigned long long waveMinU64(unsigned long long v) {
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(v, off, 64);
    v = o < v ? o : v;
  }
  return v;
}
constexpr int RED_BLOCKS = 256;
// part[b * 8 + 0..2] = ordered min, [3..5] = ordered max, [6] = ordered max half-width
// nSlots: the slot count on the device (the host may not know it yet)
__global__ __launch_bounds__(TPB) void k_bp_bounds(const Dev* __restrict__ Dp, Params P, const int* nSlots,
                                                   unsigned long long* part) {
  const Dev& D = *Dp;
  __shared__ unsigned long long sh[7][WAVES];
  unsigned long long v[7] = {~0ull, ~0ull, ~0ull, 0, 0, 0, 0};
  const int n = *nSlots;
  for (int s = gtid(); s < n; s += gridDim.x * blockDim.x) {
    const double4 B = bpBox(D.posPred[s], P);
    if (isfinite(B.w)) {
      const unsigned long long k = ordKey(B.w);
      v[6] = k > v[6] ? k : v[6];
    }
    if (B.x != -1.0e30) {
      const double c[3] = {B.x, B.y, B.z};
      for (int d = 0; d < 3; ++d)
        if (isfinite(c[d])) {
          const unsigned long long k = ordKey(c[d]);
          v[d] = k < v[d] ? k : v[d];
          v[3 + d] = k > v[3 + d] ? k : v[3 + d];
        }
    }
  }
  for (int d = 0; d < 3; ++d) v[d] = waveMinU64(v[d]);
  for (int d = 3; d < 7; ++d) v[d] = waveMaxU64(v[d]);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int d = 0; d < 7; ++d) sh[d][w] = v[d];
  __syncthreads();
  if (threadIdx.x < 7) {
    const int d = threadIdx.x;
    unsigned long long r = sh[d][0];
    for (int k = 1; k < WAVES; ++k) r = (d < 3) ? (sh[d][k] < r ? sh[d][k] : r) : (sh[d][k] > r ? sh[d][k] : r);
    part[blockIdx.x * 8 + d] = r;
  }
}
// Final bounds reduction, then the grid on the device (thread 0): cell edge 2 x the largest
// half-width (x 1.0001), grown by 1.5x until there are at most 4 x slots + 64 cells (the cell
// buffers' capacity).  Any grid with cells at least that wide finds every overlapping pair.
__global__ void k_bp_bounds_final(const unsigned long long* part, int nb, unsigned long long* red,
                                  const int* nSlots, Grid* grid, int* numCells) {
  __shared__ unsigned long long sh[7][WAVES];
  __shared__ unsigned long long fin[7];
  unsigned long long v[7] = {~0ull, ~0ull, ~0ull, 0, 0, 0, 0};
  for (int b = threadIdx.x; b < nb; b += blockDim.x)
    for (int d = 0; d < 7; ++d) {
      const unsigned long long x = part[b * 8 + d];
      v[d] = (d < 3) ? (x < v[d] ? x : v[d]) : (x > v[d] ? x : v[d]);
    }
  for (int d = 0; d < 3; ++d) v[d] = waveMinU64(v[d]);
  for (int d = 3; d < 7; ++d) v[d] = waveMaxU64(v[d]);
  if ((threadIdx.x & 63) == 0)
    for (int d = 0; d < 7; ++d) sh[d][threadIdx.x >> 6] = v[d];
  __syncthreads();
  if (threadIdx.x < 7) {
    const int d = threadIdx.x;
    unsigned long long r = sh[d][0];
    for (int k = 1; k < WAVES; ++k) r = (d < 3) ? (sh[d][k] < r ? sh[d][k] : r) : (sh[d][k] > r ? sh[d][k] : r);
    red[d] = r;
    fin[d] = r;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double lo[3], hi[3];
  for (int d = 0; d < 3; ++d) {
    lo[d] = fin[d] == ~0ull ? 0.0 : ordVal(fin[d]);
    hi[d] = fin[3 + d] == 0 ? 0.0 : ordVal(fin[3 + d]);
    if (lo[d] > hi[d]) lo[d] = hi[d] = 0.0;
  }
  const double maxB = fin[6] ? ordVal(fin[6]) : 0.0;
  Grid G;
  G.lo = d3{lo[0], lo[1], lo[2]};
  G.h = 2.0 * maxB * 1.0001 + 1e-300;
  for (;;) {
    double cells = 1.0;
    for (int d = 0; d < 3; ++d) {
      double nd = floor((hi[d] - lo[d]) / G.h) + 1.0;
      if (!(nd >= 1.0)) nd = 1.0;
      if (nd > 1e9) nd = 1e9;
      G.dim[d] = (int)nd;
      cells *= nd;
    }
    if (cells <= 4.0 * (*nSlots) + 64.0) break;
    G.h *= 1.5;
  }
  *grid = G;
  *numCells = G.dim[0] * G.dim[1] * G.dim[2];
}

__device__ __forceinline__ int cellCoord(double c, double lo, double h, int dim) {
  const double f = (c - lo) / h;
  if (!(f >= 0.0)) return 0;
  if (f >= (double)(dim - 1)) return dim - 1;
  return (int)f;
}
// cell of every slot and its rank inside the cell (cellCnt is all zero on entry)
__global__ void k_cell_count(const Dev* __restrict__ Dp, Params P, const int* nSlots, const Grid* grid) {
  const Dev& D = *Dp;
  const int n = *nSlots;
  const Grid G = *grid;
  for (int s = gtid(); s < n; s += gridDim.x * blockDim.x) {
    const double4 B = bpBox(D.posPred[s], P);
    const int cx = cellCoord(B.x, G.lo.x, G.h, G.dim[0]);
    const int cy = cellCoord(B.y, G.lo.y, G.h, G.dim[1]);
    const int cz = cellCoord(B.z, G.lo.z, G.h, G.dim[2]);
    const int c = (cz * G.dim[1] + cy) * G.dim[0] + cx;
    D.cellOf[s] = c;
    D.cellRank[s] = atomicAdd(&D.cellCnt[c], 1);
  }
}
// slots in cell order; every occupied cell's count is cleared again for the next broadphase
__global__ void k_cell_fill(const Dev* __restrict__ Dp, Params P, const int* nSlots) {
  const Dev& D = *Dp;
  const int n = *nSlots;
  for (int s = gtid(); s < n; s += gridDim.x * blockDim.x) {
    const int c = D.cellOf[s];
    const int k = D.cellStart[c] + D.cellRank[s];
    D.order[k] = s;
    D.sbox[k] = bpBox(D.posPred[s], P);
    D.cellCnt[c] = 0;
  }
}
constexpr int BP_CAP = 32;  // staged partners per real; a real with more is queried again
// Full-stencil query of one real i: every partner j > i in the 27 neighbouring cells, written as
// pairs (i, j) to out[0..n) (cell order; k_bp_compact sorts the segment)
__device__ __forceinline__ int bpQueryOne(const Dev& D, const Grid& G, int i, const double4 Bi,
                                          int2* out) {
  const int cx = cellCoord(Bi.x, G.lo.x, G.h, G.dim[0]);
  const int cy = cellCoord(Bi.y, G.lo.y, G.h, G.dim[1]);
  const int cz = cellCoord(Bi.z, G.lo.z, G.h, G.dim[2]);
  const int x0 = cx > 0 ? cx - 1 : 0;
  const int x1 = cx + 1 < G.dim[0] ? cx + 1 : G.dim[0] - 1;
  int n = 0;
  for (int z = cz - 1; z <= cz + 1; ++z) {
    if (z < 0 || z >= G.dim[2]) continue;
    for (int y = cy - 1; y <= cy + 1; ++y) {
      if (y < 0 || y >= G.dim[1]) continue;
      const int row = (z * G.dim[1] + y) * G.dim[0];
      const int k0 = D.cellStart[row + x0], k1 = D.cellStart[row + x1 + 1];
      for (int k = k0; k < k1; ++k) {
        const int j = D.order[k];
        if (j > i && boxOverlap(Bi, D.sbox[k])) out[n++] = make_int2(i, j);
      }
    }
  }
  return n;
}
// Half-stencil query: every slot (real or ghost) tests only its forward neighbourhood -- the rest
// of its own cell and the next cell in x (one contiguous range in cell order), the row y+1 at the
// same z, and the three rows at z+1 -- so each neighbouring pair of slots is tested once (box
// overlap is symmetric).  The pair belongs to its lower slot, which must be real; the partner is
// appended to that real's staged row with an atomic count.  pairCnt is zero on entry.  A real
// with more than BP_CAP partners is queried again in full by k_bp_compact.
// Hits are buffered four at a time so their count atomics are in flight together (each returns
// the partner's place in its row, and the store waits for it).
__device__ __forceinline__ void bpQuerySlot(const Dev& D, const Params& P, const Grid& G, int k,
                                            int* stage) {
  const int s = D.order[k];
  const double4 Bs = D.sbox[k];
  const int cx = cellCoord(Bs.x, G.lo.x, G.h, G.dim[0]);
  const int cy = cellCoord(Bs.y, G.lo.y, G.h, G.dim[1]);
  const int cz = cellCoord(Bs.z, G.lo.z, G.h, G.dim[2]);
  const int x0 = cx > 0 ? cx - 1 : 0;
  const int x1 = cx + 1 < G.dim[0] ? cx + 1 : G.dim[0] - 1;
  const int row = (cz * G.dim[1] + cy) * G.dim[0];
  const int numReal = P.numReal;
  int lo0 = 0, hi0 = 0, lo1 = 0, hi1 = 0, lo2 = 0, hi2 = 0, lo3 = 0, hi3 = 0, nh = 0;
  auto flush = [&]() __attribute__((always_inline)) {
    int s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    if (nh > 0) s0 = atomicAdd(&D.pairCnt[lo0], 1);
    if (nh > 1) s1 = atomicAdd(&D.pairCnt[lo1], 1);
    if (nh > 2) s2 = atomicAdd(&D.pairCnt[lo2], 1);
    if (nh > 3) s3 = atomicAdd(&D.pairCnt[lo3], 1);
    if (nh > 0 && s0 < BP_CAP) stage[(size_t)s0 * numReal + lo0] = hi0;
    if (nh > 1 && s1 < BP_CAP) stage[(size_t)s1 * numReal + lo1] = hi1;
    if (nh > 2 && s2 < BP_CAP) stage[(size_t)s2 * numReal + lo2] = hi2;
    if (nh > 3 && s3 < BP_CAP) stage[(size_t)s3 * numReal + lo3] = hi3;
    nh = 0;
  };
  auto hit = [&](int j) __attribute__((always_inline)) {
    const int lo = s < j ? s : j, hi = s < j ? j : s;
    if (lo >= numReal) return;  // two ghosts
    // shift in (a register queue: no indexed store, which would go to scratch memory)
    lo3 = lo2;
    hi3 = hi2;
    lo2 = lo1;
    hi2 = hi1;
    lo1 = lo0;
    hi1 = hi0;
    lo0 = lo;
    hi0 = hi;
    if (++nh == 4) flush();
  };
  // own cell after this slot, then cell x+1
  {
    const int e = D.cellStart[row + x1 + 1];
    for (int q = k + 1; q < e; ++q)
      if (boxOverlap(Bs, D.sbox[q])) hit(D.order[q]);
  }
  for (int w = 0; w < 4; ++w) {
    const int y = w == 0 ? cy + 1 : cy + w - 2;
    const int z = w == 0 ? cz : cz + 1;
    if (y < 0 || y >= G.dim[1] || z >= G.dim[2]) continue;
    const int r = (z * G.dim[1] + y) * G.dim[0];
    const int b = D.cellStart[r + x0], e = D.cellStart[r + x1 + 1];
    for (int q = b; q < e; ++q)
      if (boxOverlap(Bs, D.sbox[q])) hit(D.order[q]);
  }
  flush();
}
__global__ __launch_bounds__(TPB) void k_bp_query(const Dev* __restrict__ Dp, Params P, const int* nSlots,
                                                  const Grid* grid, int* stage) {
  const Dev& D = *Dp;
  const int n = *nSlots;
  const Grid G = *grid;
  for (int k = gtid(); k < n; k += gridDim.x * blockDim.x) bpQuerySlot(D, P, G, k, stage);
}
__global__ __launch_bounds__(TPB) void k_bp_compact(const Dev* __restrict__ Dp, Params P, const Grid* grid,
                                                    const int* stage) {
  const Dev& D = *Dp;
  const int i = gtid();
  if (i >= P.numReal) return;
  const Grid G = *grid;
  const int n = D.pairCnt[i];
  if (n == 0) return;
  int2* seg = D.pairs + D.pairStart[i];
  // rows of up to 16 partners (a dense bed's half-stencil rows hold ~12) are sorted in
  // registers; longer ones in place in global memory
  constexpr int R = 16;
  if (n <= R) {
    int v[R];
#pragma unroll
    for (int a = 0; a < R; ++a) v[a] = a < n ? stage[(size_t)a * P.numReal + i] : 0x7fffffff;
    // odd-even transposition sort of R values (fixed indices: stays in registers)
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int a = r & 1; a + 1 < R; a += 2) {
        const int lo = v[a] < v[a + 1] ? v[a] : v[a + 1];
        const int hi = v[a] < v[a + 1] ? v[a + 1] : v[a];
        v[a] = lo;
        v[a + 1] = hi;
      }
#pragma unroll
    for (int a = 0; a < R; ++a)
      if (a < n) seg[a] = make_int2(i, v[a]);
    return;
  }
  if (n <= BP_CAP) {
    for (int a = 0; a < n; ++a) seg[a] = make_int2(i, stage[(size_t)a * P.numReal + i]);
  } else {
    bpQueryOne(D, G, i, bpBox(D.posPred[i], P), seg);
  }
  // ascending partner order
  for (int a = 1; a < n; ++a) {
    const int2 x = seg[a];
    int b = a - 1;
    while (b >= 0 && seg[b].y > x.y) {
      seg[b + 1] = seg[b];
      --b;
    }
    seg[b + 1] = x;
  }
}

// ---- candidate pairs kept across substeps -------------------------------------------------
// A rebuild runs the broadphase with every box grown by the skin; its pair list (sorted, per-real
// offsets) is the candidate list, and each slot's plain box and each real's ghost images are
// recorded.  While no box centre has moved by more than skin / 4 on any axis, no box has grown
// by more than skin / 4 and every real has the same images, any two boxes that overlap now
// overlapped grown by the skin at the rebuild (|c0i - c0j| <= |ci - cj| + skin / 2 <= wi + wj +
// skin / 2 <= w0i + w0j + skin), so filtering the candidates with the exact test gives the exact
// pair list, in the same order.
__device__ __forceinline__ int ghostCode(const Params& P, d3 p) {
  int sx, sy, sz;
  ghostShifts(P, p, sx, sy, sz);
  return (sx + 1) + 3 * (sy + 1) + 9 * (sz + 1);
}
__global__ void k_cand_ref(const Dev* __restrict__ Dp, Params P, int numSlots, double4* ref, unsigned char* code) {
  const Dev& D = *Dp;
  const int s = gtid();
  if (s >= numSlots) return;
  ref[s] = saneBox(D.posPred[s], P.margin, P.boxCap);
  if (s < P.numReal) code[s] = (unsigned char)ghostCode(P, xyz(D.pos[s]));
}
// whether slot s's plain box B has left its rebuild box R by more than lim (written so that a
// NaN counts as moved)
__device__ __forceinline__ bool candMoved(const double4& B, const double4& R, double lim) {
  return !(fabs(B.x - R.x) <= lim && fabs(B.y - R.y) <= lim && fabs(B.z - R.z) <= lim &&
           B.w <= R.w + lim);
}
// the exact test on every candidate of real i (the broadphase's own box test); keep[c] marks
// the pairs, pairCnt[i] counts them.  With a flag, real i's own box and images are checked too,
// and (nSlots given) flag = 1 also when a ghost slot moved or the slot count changed.
__global__ __launch_bounds__(TPB) void k_cand_count(const Dev* __restrict__ Dp, Params P, const int2* cand,
                                                    const int* candStart, unsigned char* keep,
                                                    const double4* ref, const unsigned char* code,
                                                    double lim, int* flag,
                                                    const int* nSlots = nullptr,
                                                    int refSlots = 0) {
  const Dev& D = *Dp;
  const int i = gtid();
  if (nSlots) {
    const int ns = *nSlots;
    bool bad = (ns != refSlots);
    for (int s = P.numReal + i; s < ns && !bad; s += gridDim.x * blockDim.x)
      bad = candMoved(saneBox(D.posPred[s], P.margin, P.boxCap), ref[s], lim);
    if (bad) atomicOr(flag, 1);
  }
  if (i >= P.numReal) return;
  const int b = candStart[i], e = candStart[i + 1];
  int n = 0;
  const double4 Bi = saneBox(D.posPred[i], P.margin, P.boxCap);
  if (flag && (candMoved(Bi, ref[i], lim) ||
               code[i] != (unsigned char)ghostCode(P, xyz(D.pos[i]))))
    atomicOr(flag, 1);
  if (e > b) {
    for (int c = b; c < e; ++c) {
      const bool k = boxOverlap(Bi, saneBox(D.posPred[cand[c].y], P.margin, P.boxCap));
      keep[c] = k ? 1 : 0;
      n += k;
    }
  }
  D.pairCnt[i] = n;
}
__global__ __launch_bounds__(TPB) void k_cand_write(const Dev* __restrict__ Dp, int numReal, const int2* cand,
                                                    const int* candStart,
                                                    const unsigned char* keep) {
  const Dev& D = *Dp;
  const int i = gtid();
  if (i >= numReal) return;
  int2* out = D.pairs + D.pairStart[i];
  int n = 0;
  for (int c = candStart[i], e = candStart[i + 1]; c < e; ++c)
    if (keep[c]) out[n++] = cand[c];
}

// ============================================================================================
// narrowphase — src/narrowphase.hpp:179-268 (pairs) and :371-449 (planes)
// ============================================================================================
__device__ __forceinline__ void writeContact(const Dev& D, int slot, int a, int b, d3 n, d3 rA, d3 rB,
                                             double dist, double eb, double mub,
                                             d3 bv = d3{0.0, 0.0, 0.0}) {
  D.cAB[slot] = make_int2(a, b);
  D.cN[slot] = tod4(n, dist);
  D.cRA[slot] = tod4(rA, eb);
  D.cRB[slot] = tod4(rB, mub);
  D.cBV[slot] = tod4(bv, 0.0);
}

template <bool WRITE>
__device__ __forceinline__ double npPair(const Dev& D, const Params& P, int p) {
  const int2 ab = D.pairs[p];
  const int idA = ab.x, idB = ab.y;
  const d3 posA = xyz(D.posPred[idA]), posB = xyz(D.posPred[idB]);
  const q4 qA = toq(D.quatPred[idA]), qB = toq(D.quatPred[idB]);
  const double effScaleA = D.scale[idA] * P.gs, effScaleB = D.scale[idB] * P.gs;
  const int countA = P.shape.numShell;
  const bool sphereA = (P.shape.type == SHAPE_SPHERE);
  const int iter = countA > 0 ? countA : 1;
  double eb = -1.0, mub = -1.0;
  if (WRITE && P.hasPairTable) {
    const int t = ((int)D.matId[idA] * 8 + (int)D.matId[idB]) * 2;
    eb = D.pairTable[t];
    mub = D.pairTable[t + 1];
  }
  int n = 0;
  double ovl = 0.0;
  const int base = WRITE ? D.noff[p] : 0;
  for (int k = 0; k < iter; ++k) {
    d3 pLocalA{0.0, 0.0, 0.0};
    double pointRadius = 0.0;
    if (countA > 0) pLocalA = xyz(D.shell[k]);
    else if (sphereA) pointRadius = P.shape.p0 * effScaleA;
    const d3 pWorld = add3(posA, rotate(qA, scale3(pLocalA, effScaleA)));
    const d3 pLocalB = invRotate(qB, sub3(pWorld, posB));
    const d3 pCanB = scale3(pLocalB, 1.0 / effScaleB);
    const double dist = sdfShape(P.shape, pCanB) * effScaleB;
    const double effDist = dist - pointRadius;
    if (effDist >= P.margin) continue;
    if (!WRITE) {
      ++n;
      continue;
    }
    if (effDist < 0.0 && -effDist > ovl) ovl = -effDist;
    const double eps = 1e-4;
    d3 nLoc{sdfShape(P.shape, d3{pCanB.x + eps, pCanB.y, pCanB.z}) -
                sdfShape(P.shape, d3{pCanB.x - eps, pCanB.y, pCanB.z}),
            sdfShape(P.shape, d3{pCanB.x, pCanB.y + eps, pCanB.z}) -
                sdfShape(P.shape, d3{pCanB.x, pCanB.y - eps, pCanB.z}),
            sdfShape(P.shape, d3{pCanB.x, pCanB.y, pCanB.z + eps}) -
                sdfShape(P.shape, d3{pCanB.x, pCanB.y, pCanB.z - eps})};
    const double len = len3(nLoc);
    nLoc = (len > 1e-9) ? scale3(nLoc, 1.0 / len) : d3{0.0, 1.0, 0.0};
    const d3 nWorld = rotate(qB, nLoc);
    const d3 pSurfA = sub3(pWorld, scale3(nWorld, pointRadius));
    const d3 rA = sub3(pSurfA, posA);
    const d3 rB = sub3(sub3(pSurfA, scale3(nWorld, effDist)), posB);
    writeContact(D, base + n, idA, idB, nWorld, rA, rB, effDist, eb, mub);
    ++n;
  }
  if (!WRITE) D.ncnt[p] = n;
  return ovl;
}
template <bool WRITE>
__global__ __launch_bounds__(TPB) void k_np_pairs(const Dev* __restrict__ Dp, Params P, int numPairs) {
  const Dev& D = *Dp;
  __shared__ unsigned long long sh[WAVES];
  const int p = gtid();
  const double ovl = (p < numPairs) ? npPair<WRITE>(D, P, p) : 0.0;
  if (WRITE) blockAtomicMaxU64(D.maxOverlapBits, dbits(ovl), sh);
}
// Write pass of point shapes (at most one contact per pair) over the contacts only: contact c
// comes from pair pairOf[c] (the compaction written by the slot scan) and lands at its offset
// noff[pairOf[c]] = c, as in the full pass.  The pair contacts are the first noff[numPairs].
// (grid-stride over the pair contacts, whose count noff[numPairs] is read on the device)
__global__ __launch_bounds__(TPB) void k_np_pairs_sparse(const Dev* __restrict__ Dp, Params P, int numPairs,
                                                         const int* pairOf) {
  const Dev& D = *Dp;
  __shared__ unsigned long long sh[WAVES];
  const int n = D.noff[numPairs];
  double ovl = 0.0;
  for (int c = gtid(); c < n; c += gridDim.x * blockDim.x) {
    const double o = npPair<true>(D, P, pairOf[c]);
    ovl = o > ovl ? o : ovl;
  }
  blockAtomicMaxU64(D.maxOverlapBits, dbits(ovl), sh);
}

// Shell shapes: one wavefront per candidate pair.  The block stages the point shell in LDS; the
// 64 lanes take points k = lane, lane + 64, ...; a ballot prefix gives every hit its place in
// shell order, so the contacts (and their count) are the serial loop's.
constexpr int NP_WAVES = TPB / 64;
constexpr int NP_LDS_POINTS = 1024;  // larger shells read the shell from global memory
template <bool WRITE>
__global__ __launch_bounds__(TPB) void k_np_pairs_wave(const Dev* __restrict__ Dp, Params P, int numPairs) {
  const Dev& D = *Dp;
  __shared__ double4 shShell[NP_LDS_POINTS];
  __shared__ unsigned long long shMax[NP_WAVES];
  const int countA = P.shape.numShell;
  const bool inLds = countA <= NP_LDS_POINTS;
  if (inLds)
    for (int k = threadIdx.x; k < countA; k += blockDim.x) shShell[k] = D.shell[k];
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int p = blockIdx.x * NP_WAVES + wave;
  double ovl = 0.0;
  if (p < numPairs) {
    const int2 ab = D.pairs[p];
    const int idA = ab.x, idB = ab.y;
    const d3 posA = xyz(D.posPred[idA]), posB = xyz(D.posPred[idB]);
    const q4 qA = toq(D.quatPred[idA]), qB = toq(D.quatPred[idB]);
    const double effScaleA = D.scale[idA] * P.gs, effScaleB = D.scale[idB] * P.gs;
    double eb = -1.0, mub = -1.0;
    if (WRITE && P.hasPairTable) {
      const int tt = ((int)D.matId[idA] * 8 + (int)D.matId[idB]) * 2;
      eb = D.pairTable[tt];
      mub = D.pairTable[tt + 1];
    }
    const int base = WRITE ? D.noff[p] : 0;
    int n = 0;
    for (int k0 = 0; k0 < countA; k0 += 64) {
      const int k = k0 + lane;
      bool hit = false;
      d3 pWorld{0, 0, 0}, pCanB{0, 0, 0};
      double effDist = 0.0;
      if (k < countA) {
        const d3 pLocalA = xyz(inLds ? shShell[k] : D.shell[k]);
        pWorld = add3(posA, rotate(qA, scale3(pLocalA, effScaleA)));
        const d3 pLocalB = invRotate(qB, sub3(pWorld, posB));
        pCanB = scale3(pLocalB, 1.0 / effScaleB);
        effDist = sdfShape(P.shape, pCanB) * effScaleB;  // point radius 0 for shell points
        hit = effDist < P.margin;
      }
      const unsigned long long mask = __ballot(hit);
      if (WRITE && hit) {
        if (effDist < 0.0 && -effDist > ovl) ovl = -effDist;
        const double eps = 1e-4;
        d3 nLoc{sdfShape(P.shape, d3{pCanB.x + eps, pCanB.y, pCanB.z}) -
                    sdfShape(P.shape, d3{pCanB.x - eps, pCanB.y, pCanB.z}),
                sdfShape(P.shape, d3{pCanB.x, pCanB.y + eps, pCanB.z}) -
                    sdfShape(P.shape, d3{pCanB.x, pCanB.y - eps, pCanB.z}),
                sdfShape(P.shape, d3{pCanB.x, pCanB.y, pCanB.z + eps}) -
                    sdfShape(P.shape, d3{pCanB.x, pCanB.y, pCanB.z - eps})};
        const double len = len3(nLoc);
        nLoc = (len > 1e-9) ? scale3(nLoc, 1.0 / len) : d3{0.0, 1.0, 0.0};
        const d3 nWorld = rotate(qB, nLoc);
        const d3 pSurfA = sub3(pWorld, scale3(nWorld, 0.0));
        const d3 rA = sub3(pSurfA, posA);
        const d3 rB = sub3(sub3(pSurfA, scale3(nWorld, effDist)), posB);
        const int rank = __popcll(mask & ((1ull << lane) - 1ull));
        writeContact(D, base + n + rank, idA, idB, nWorld, rA, rB, effDist, eb, mub);
      }
      n += __popcll(mask);
    }
    if (!WRITE && lane == 0) D.ncnt[p] = n;
  }
  if (WRITE) {
    unsigned long long m = waveMaxU64(dbits(ovl));
    if (lane == 0) shMax[wave] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned long long b = shMax[0];
      for (int w = 1; w < NP_WAVES; ++w) b = shMax[w] > b ? shMax[w] : b;
      if (b) atomicMax(D.maxOverlapBits, b);
    }
  }
}

template <bool WRITE>
__device__ __forceinline__ double npPlanes(const Dev& D, const Params& P, int numPairs, int i) {
  const d3 posA = xyz(D.posPred[i]);
  const double s = D.scale[i] * P.gs;
  double baseR = P.shape.p0;
  if (baseR == 0.0) baseR = 1.0;
  const double radius = baseR * s;
  const int numPts = P.shape.numShell;
  const q4 qA = toq(D.quatPred[i]);
  int n = 0;
  double ovl = 0.0;
  const int base = WRITE ? D.noff[numPairs + i] : 0;
  for (int pi = 0; pi < P.numPlanes; ++pi) {
    const d3 pp = xyz(P.planeP[pi]), pn = xyz(P.planeN[pi]);
    if (numPts > 0) {
      for (int k = 0; k < numPts; ++k) {
        const d3 rA = rotate(qA, scale3(xyz(D.shell[k]), s));
        const d3 pwk = add3(posA, rA);
        const double dist = dot3(sub3(pwk, pp), pn);
        if (dist >= P.margin) continue;
        if (WRITE) {
          if (dist < 0.0 && -dist > ovl) ovl = -dist;
          writeContact(D, base + n, i, -1, pn, rA, pp, dist, -1.0, -1.0);
        }
        ++n;
      }
    } else {
      const double dist = dot3(sub3(posA, pp), pn) - radius;
      if (dist >= P.margin) continue;
      if (WRITE) {
        if (dist < 0.0 && -dist > ovl) ovl = -dist;
        writeContact(D, base + n, i, -1, pn, scale3(pn, -radius), pp, dist, -1.0, -1.0);
      }
      ++n;
    }
  }
  if (!WRITE) D.ncnt[numPairs + i] = n;
  return ovl;
}
// zeroSlots >= 0: also write a zero count at slot zeroSlots + i (the wall slots, without walls)
template <bool WRITE>
__global__ void k_np_planes(const Dev* __restrict__ Dp, Params P, int numPairs,
                            int zeroSlots = -1) {
  const Dev& D = *Dp;
  __shared__ unsigned long long sh[WAVES];
  const int i = gtid();
  if (!WRITE && zeroSlots >= 0 && i < P.numReal) D.ncnt[zeroSlots + i] = 0;
  const double ovl = (i < P.numReal) ? npPlanes<WRITE>(D, P, numPairs, i) : 0.0;
  if (WRITE) blockAtomicMaxU64(D.maxOverlapBits, dbits(ovl), sh);
}

// ============================================================================================
// contact -> manifold reduction — src/contact_preprocessing.hpp:320-471
// ============================================================================================
// manifolds per real body: its pairs with >= 1 contact, then one boundary manifold
__global__ void k_mcount(const Dev* __restrict__ Dp, Params P, int numPairs) {
  const Dev& D = *Dp;
  const int i = gtid();
  if (i >= P.numReal) return;
  int m = 0;
  for (int p = D.pairStart[i]; p < D.pairStart[i + 1]; ++p) m += D.ncnt[p] > 0;
  m += (D.ncnt[numPairs + i] + D.ncnt[numPairs + P.numReal + i]) > 0;
  D.mCnt[i] = m;
}
__global__ void k_mbuild(const Dev* __restrict__ Dp, Params P, int numPairs) {
  const Dev& D = *Dp;
  const int i = gtid();
  if (i >= P.numReal) return;
  int idx = D.mStart[i];
  for (int p = D.pairStart[i]; p < D.pairStart[i + 1]; ++p) {
    const int c = D.ncnt[p];
    if (c > 0) {
      D.mAB[idx] = D.pairs[p];
      D.mRange[idx] = make_int2(D.noff[p], c);
      D.mRange2[idx] = make_int2(0, 0);
      ++idx;
    }
  }
  // boundary manifold: plane contacts, then wall contacts (contact order)
  const int cb = D.ncnt[numPairs + i], cw = D.ncnt[numPairs + P.numReal + i];
  if (cb + cw > 0) {
    D.mAB[idx] = make_int2(i, -1);
    D.mRange[idx] = make_int2(D.noff[numPairs + i], cb);
    D.mRange2[idx] = make_int2(D.noff[numPairs + P.numReal + i], cw);
  }
}
// transformContact + ordered segmented sum, src/contact_preprocessing.hpp:339-466
__global__ __launch_bounds__(TPB) void k_mreduce(const Dev* __restrict__ Dp, int numManifolds) {
  const Dev& D = *Dp;
  const int m = gtid();
  if (m >= numManifolds) return;
  const int2 ab = D.mAB[m];
  const int2 rg = D.mRange[m], rg2 = D.mRange2[m];
  d3 ns{0, 0, 0}, ta{0, 0, 0}, tb{0, 0, 0}, ra{0, 0, 0}, rb{0, 0, 0}, wv{0, 0, 0};
  double es = 0.0, fs = 0.0;
  int np = 0;
  const int nTot = rg.y + rg2.y;
  for (int q = 0; q < nTot; ++q) {
    const int c = q < rg.y ? rg.x + q : rg2.x + (q - rg.y);
    D.cSlot[c] = m;
    const double4 cn = D.cN[c];
    // an inactive contact (dist > 0) contributes a zero manifold with num_points 0
    d3 na{0, 0, 0}, tca{0, 0, 0}, tcb{0, 0, 0}, rcan{0, 0, 0}, roth{0, 0, 0}, bv{0, 0, 0};
    double e = 0.0, f = 0.0;
    int pts = 0;
    if (!(cn.w > 0.0)) {
      const int2 cab = D.cAB[c];
      const double4 cra = D.cRA[c], crb = D.cRB[c];
      const bool flip = (cab.y >= 0 && cab.y < cab.x);
      const double sn = flip ? 1.0 : -1.0;
      na = d3{sn * cn.x, sn * cn.y, sn * cn.z};  // (exact: a sign change)
      const d3 shift{cn.x * cn.w * 0.5, cn.y * cn.w * 0.5, cn.z * cn.w * 0.5};
      const d3 rAmid = sub3(xyz(cra), shift);
      const d3 rBmid = add3(xyz(crb), shift);
      // (component selects: a select of whole records went through scratch)
      rcan = d3{flip ? rBmid.x : rAmid.x, flip ? rBmid.y : rAmid.y, flip ? rBmid.z : rAmid.z};
      roth = d3{flip ? rAmid.x : rBmid.x, flip ? rAmid.y : rBmid.y, flip ? rAmid.z : rBmid.z};
      tca = cross3(rcan, na);
      tcb = cross3(roth, neg3(na));
      e = cra.w;
      f = crb.w;
      bv = xyz(D.cBV[c]);
      pts = 1;
    }
    ns = add3(ns, na);
    ta = add3(ta, tca);
    tb = add3(tb, tcb);
    ra = add3(ra, rcan);
    rb = add3(rb, roth);
    wv = add3(wv, bv);
    es += e;
    fs += f;
    np += pts;
  }
  D.mNs[m] = tod4(ns, es);
  D.mTA[m] = tod4(ta, fs);
  D.mTB[m] = tod4(tb, 0.0);
  D.mRA[m] = tod4(ra, 0.0);
  D.mRB[m] = tod4(rb, 0.0);
  D.mWV[m] = tod4(wv, 0.0);
  D.mNp[m] = np;
  const int realA = D.realIdx[ab.x];
  const int realB = ab.y >= 0 ? D.realIdx[ab.y] : -1;
  D.mReal[m] = make_int2(realA, realB);
}

// ============================================================================================
// Jones-Plassmann colouring — src/solver_velocity.hpp:277-374, src/solver_position.hpp:155-241
// ============================================================================================
// uncoloured manifolds (-1) are appended to the first work list; inactive / duplicate ones get -2
__device__ __forceinline__ unsigned long long realKey(int realA, int realB) {
  const unsigned a = (unsigned)realA;
  const unsigned b = realB >= 0 ? (unsigned)realB : 0xFFFFFFFFu;
  const unsigned hi = a < b ? a : b, lo = a < b ? b : a;
  return ((unsigned long long)hi << 32) | lo;
}
// a manifold's ledger key (mKey) and support code (mCode: 3 wall, 1 / 2 body B below / above
// A along gravity, 0 otherwise), src/contact_preprocessing.hpp:70-78,213-253; returns the code
__device__ __forceinline__ unsigned char manifoldKeyCode(const Dev& D, int m, d3 gHat) {
  unsigned char code = 0;
  if (D.mNp[m] <= 0) {
    D.mKey[m] = ~0ull;
  } else {
    const int2 ab = D.mAB[m];
    const int2 r = D.mReal[m];
    D.mKey[m] = realKey(r.x, r.y);
    if (ab.y < 0) {
      code = 3;
    } else {
      const d3 dx = sub3(xyz(D.posPred[ab.x]), xyz(D.posPred[ab.y]));
      const double up = -(dx.x * gHat.x + dx.y * gHat.y + dx.z * gHat.z);
      const double thr = 0.3 * sqrt(dot3(dx, dx));
      code = (up > thr) ? 1 : ((up < -thr) ? 2 : 0);
    }
  }
  D.mCode[m] = code;
  return code;
}
// (also rowOf[m] = -1: rows are placed for the coloured manifolds only; and each manifold's
// ledger key and support code with the support list, as k_mkeys, in the same pass)
__global__ __launch_bounds__(TPB) void k_jp_init_manifolds(const Dev* __restrict__ Dp,
                                                            int numManifolds, int* work,
                                                            int* count, int* rowOf, d3 gHat,
                                                            int* supList, int* supCount) {
  const Dev& D = *Dp;
  __shared__ int shWave[WAVES], shBase, shBase2;
  const int2 ch = chunkOf(numManifolds);
  // pass 1: colour state, row slot, key and support code; count the list entries
  int nWork = 0, nSup = 0;
  for (int m = ch.x + (int)threadIdx.x; m < ch.y; m += blockDim.x) {
    const int2 r = D.mReal[m];
    int c = -1;
    if (D.mNp[m] <= 0) c = -2;
    else if (D.mAB[m].y >= 0 && r.x > r.y) c = -2;
    D.mColor[m] = c;
    rowOf[m] = -1;
    const unsigned char code = manifoldKeyCode(D, m, gHat);
    nWork += c == -1 ? 1 : 0;
    nSup += code != 0 ? 1 : 0;
  }
  const int totWork = blockSumI(nWork, shWave);
  const int totSup = blockSumI(nSup, shWave);
  if (threadIdx.x == 0) {
    shBase = totWork ? atomicAdd(count, totWork) : 0;
    shBase2 = totSup ? atomicAdd(supCount, totSup) : 0;
  }
  __syncthreads();
  // pass 2: the entries in chunk order (block-uniform trip count: every thread ranks)
  int runWork = shBase, runSup = shBase2;
  for (int m0 = ch.x; m0 < ch.y; m0 += blockDim.x) {
    const int m = m0 + (int)threadIdx.x;
    const bool in = m < ch.y;
    const bool w = in && D.mColor[m] == -1;
    const bool sp = in && D.mCode[m] != 0;
    int tw, ts;
    const int rw = blockRank(w, shWave, tw);
    const int rs = blockRank(sp, shWave, ts);
    if (w) work[runWork + rw] = m;
    if (sp) supList[runSup + rs] = m;
    runWork += tw;
    runSup += ts;
  }
}
// One JP round over the compacted work list.  winner[] holds, per body, 1 + the largest priority
// among its uncoloured edges; a committing edge clears the slots of its two bodies, a losing edge
// keeps its key there (it re-contends with the same key), so no per-round reset is needed.
// Both round kernels walk their list in grid-stride chunks, so any grid covers it (the host sizes
// grids from the previous colouring's list sizes).
__global__ void k_jp_contend(const int2* ends, const int* work, const int* count,
                             unsigned long long* winner) {
  const int n = *count;
  for (int t = gtid(); t < n; t += gridDim.x * blockDim.x) {
    const int e = work[t];
    const int2 ab = ends[e];
    const unsigned long long key = colorKey1(e);
    atomicMax(&winner[ab.x], key);
    if (ab.y >= 0) atomicMax(&winner[ab.y], key);
  }
}
__global__ void k_jp_commit(const int2* ends, int* colour, const int* work, const int* count,
                            unsigned long long* winner, unsigned long long* mask, int* next,
                            int* nextCount) {
  __shared__ int shCnt, shBase;
  const int n = *count;
  // block-uniform trip count: every thread reaches each blockAppend
  for (int base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
    const int t = base + threadIdx.x;
    int e = -1;
    bool lose = false;
    if (t < n) {
      e = work[t];
      const int2 ab = ends[e];
      const unsigned long long key = colorKey1(e);
      const unsigned long long wa = winner[ab.x];
      const unsigned long long wb = ab.y >= 0 ? winner[ab.y] : key;
      if (wa != key || wb != key) {
        lose = true;
      } else {
        unsigned long long forbidden = mask[ab.x];
        if (ab.y >= 0) forbidden |= mask[ab.y];
        int c = 0;
        while (c < 62 && (forbidden & (1ull << c))) ++c;
        colour[e] = c;
        mask[ab.x] |= 1ull << c;
        winner[ab.x] = 0;
        if (ab.y >= 0) {
          mask[ab.y] |= 1ull << c;
          winner[ab.y] = 0;
        }
      }
    }
    blockAppend(lose, e, next, nextCount, &shCnt, &shBase);
  }
}
// The small rounds of a colouring inside one workgroup, with the edges and their bodies' winner /
// mask slots in LDS.  The bodies are hashed into JPS_HASH slots (linear probing) and loaded from
// the global winner / mask arrays; each round is a contend phase (LDS atomic max) and a commit
// phase separated by workgroup barriers, so a round costs a few microseconds instead of two
// launches.  Round j reads the list of round j0 from `in` (then LDS), writes its remaining count
// to cnt[j + 1] and stops at j1 or once the list is empty.  If the list still holds edges at j1,
// it goes to the global list of round j1 (`in` if j1 - j0 is even, else `out`) and the bodies'
// winner / mask slots go back to the global arrays, so launch-per-round batches can continue.
// More than JPS_EDGES edges or JPS_HASH * 3 / 4 bodies: nothing is changed and cnt[j0 + 1] = -1.
constexpr int JPS_THREADS = 1024;
constexpr int JPS_EDGES = 4096;
constexpr int JPS_HASH = 4096;
__device__ __forceinline__ int jpsSlot(int* hKey, int* used, int body, bool insert) {
  unsigned h = ((unsigned)body * 2654435761u) & (JPS_HASH - 1);
  for (int probe = 0; probe < JPS_HASH; ++probe) {
    const int k = hKey[h];
    if (k == body) return (int)h;
    if (k == -1) {
      if (!insert) return -1;
      if (atomicAdd(used, 1) >= JPS_HASH * 3 / 4) return -1;
      const int old = atomicCAS(&hKey[h], -1, body);
      if (old == -1 || old == body) return (int)h;
      atomicSub(used, 1);
      continue;  // lost the slot to another body: look at it again
    }
    h = (h + 1) & (JPS_HASH - 1);
  }
  return -1;
}
// loadGlobal = 0: the global winner / mask arrays are stale (a colouring starting here), the
// bodies start from zero
// place (a colouring that starts in this launch): when it completes, the launch also sorts its
// edges into colour classes (class starts in colStart[0..64], perm / rowOf as k_color_place),
// instead of the class kernels
__global__ __launch_bounds__(JPS_THREADS) void k_jp_small(const int2* ends, int* colour, int* in,
                                                          int* out, int* cnt, int j0, int j1,
                                                          unsigned long long* winnerG,
                                                          unsigned long long* maskG,
                                                          int loadGlobal, int place, int* perm,
                                                          int* rowOf, int* colStart,
                                                          const int* src) {
  __shared__ int hKey[JPS_HASH];
  __shared__ unsigned long long hWin[JPS_HASH], hMask[JPS_HASH];
  __shared__ int eId[JPS_EDGES];
  __shared__ unsigned eAB[JPS_EDGES];  // slot of a | slot of b << 16 (0xFFFF: boundary)
  __shared__ unsigned short lst[2][JPS_EDGES];
  __shared__ unsigned char eCol[JPS_EDGES];  // colour given to entering edge q
  __shared__ int cCnt[64], cBase[65];
  __shared__ int sNext, sFail, sUsed;
  const int T = blockDim.x, tid = threadIdx.x;
  // src (a colouring that starts here): the entering count, and the round counters this launch
  // owns start from it (k_jp_counts is not launched then)
  const int n0 = src ? *src : cnt[j0];
  if (src) {
    if (tid == 0) cnt[j0] = n0;
    for (int q = j0 + 1 + tid; q <= j1; q += T) cnt[q] = 0;
    __syncthreads();
  }
  if (n0 > JPS_EDGES) {
    if (tid == 0) cnt[j0 + 1] = -1;
    return;
  }
  for (int h = tid; h < JPS_HASH; h += T) hKey[h] = -1;
  if (tid == 0) {
    sFail = 0;
    sUsed = 0;
  }
  __syncthreads();
  for (int q = tid; q < n0; q += T) {
    const int e = in[q];
    const int2 ab = ends[e];
    const int sa = jpsSlot(hKey, &sUsed, ab.x, true);
    const int sb = ab.y >= 0 ? jpsSlot(hKey, &sUsed, ab.y, true) : 0xFFFF;
    if (sa < 0 || sb < 0) sFail = 1;
    eId[q] = e;
    eAB[q] = (unsigned)(sa & 0xFFFF) | ((unsigned)sb << 16);
    lst[0][q] = (unsigned short)q;
  }
  __syncthreads();
  if (sFail) {
    if (tid == 0) cnt[j0 + 1] = -1;
    return;
  }
  for (int h = tid; h < JPS_HASH; h += T) {
    const int b = hKey[h];
    if (b >= 0) {
      hWin[h] = loadGlobal ? winnerG[b] : 0ull;
      hMask[h] = loadGlobal ? maskG[b] : 0ull;
    }
  }
  __syncthreads();
  int cur = 0, n = n0, j = j0;
  for (; j < j1 && n > 0; ++j) {
    for (int q = tid; q < n; q += T) {
      const int k = lst[cur][q];
      const unsigned long long key = colorKey1(eId[k]);
      const unsigned ab = eAB[k];
      atomicMax(&hWin[ab & 0xFFFFu], key);
      if ((ab >> 16) != 0xFFFFu) atomicMax(&hWin[ab >> 16], key);
    }
    if (tid == 0) sNext = 0;
    __syncthreads();
    for (int q = tid; q < n; q += T) {
      const int k = lst[cur][q];
      const unsigned long long key = colorKey1(eId[k]);
      const unsigned ab = eAB[k];
      const int sa = (int)(ab & 0xFFFFu), sb = (int)(ab >> 16);
      const bool hasB = sb != 0xFFFF;
      const unsigned long long wa = hWin[sa];
      const unsigned long long wb = hasB ? hWin[sb] : key;
      if (wa != key || wb != key) {
        lst[cur ^ 1][atomicAdd(&sNext, 1)] = (unsigned short)k;
      } else {
        unsigned long long forbidden = hMask[sa];
        if (hasB) forbidden |= hMask[sb];
        int c = 0;
        while (c < 62 && (forbidden & (1ull << c))) ++c;
        colour[eId[k]] = c;
        eCol[k] = (unsigned char)c;
        hMask[sa] |= 1ull << c;
        hWin[sa] = 0;
        if (hasB) {
          hMask[sb] |= 1ull << c;
          hWin[sb] = 0;
        }
      }
    }
    __syncthreads();
    n = sNext;
    cur ^= 1;
    if (tid == 0) cnt[j + 1] = n;
    __syncthreads();  // sNext is reset by the next round
  }
  if (n > 0) {
    int* dst = ((j - j0) & 1) ? out : in;
    for (int q = tid; q < n; q += T) dst[q] = eId[lst[cur][q]];
    for (int h = tid; h < JPS_HASH; h += T) {
      const int b = hKey[h];
      if (b >= 0) {
        winnerG[b] = hWin[h];
        maskG[b] = hMask[h];
      }
    }
  } else if (place) {
    // every entering edge is coloured: class counts, starts, then each edge's place in its class
    if (tid < 64) cCnt[tid] = 0;
    __syncthreads();
    for (int q = tid; q < n0; q += T) atomicAdd(&cCnt[eCol[q]], 1);
    __syncthreads();
    if (tid < 64) {
      const int v = cCnt[tid];
      int incl = v;
      for (int off = 1; off < 64; off <<= 1) {
        const int o = __shfl_up(incl, off, 64);
        if (tid >= off) incl += o;
      }
      cBase[tid] = incl - v;
      colStart[tid] = incl - v;
      if (tid == 63) colStart[64] = incl;
      cCnt[tid] = 0;
    }
    __syncthreads();
    for (int q = tid; q < n0; q += T) {
      const int c = eCol[q];
      const int k = cBase[c] + atomicAdd(&cCnt[c], 1);
      perm[k] = eId[q];
      rowOf[eId[q]] = k;
    }
  }
}
// round counters of a batch: cnt[0] = the entering list size (from src when given), rest 0
__global__ void k_jp_counts(int* cnt, const int* src, int first, int n) {
  const int i = threadIdx.x;
  if (i < n) cnt[i] = i == 0 ? (src ? *src : first) : 0;
}

// ---- the same rounds over ordered incidence lists ----------------------------------------------
// In the contend / commit rounds an edge is coloured in the first round in which its priority is
// the largest among the uncoloured edges at each of its bodies.  The edges at one body are thus
// coloured at most one per round and in descending priority.  With each body's entering edges
// sorted by descending priority and a cursor at its first uncoloured one, an edge is coloured in
// the round after the one that made it the cursor edge of the last of its distinct bodies (round
// 0 if it starts as the cursor edge of all of them), which is what these kernels run: a round
// colours its list with the same first fit over its bodies' masks, moves those bodies' cursors
// one place on and lists the edges that thereby became cursor edges at all their bodies.  Every
// round colours the same edges with the same colours as a contend / commit round, and each edge
// is read once instead of once per round it waits.  Within a round no two listed edges share a
// body, so a body's mask and cursor have one writer per round.
// Layout: a body's mask, cursor and list end in one 16-byte record; a list entry holds the edge
// and its other body (-1 for an edge of one body); a round's list holds (edge, body, other body),
// so a round reads no edge ends.
struct alignas(16) JpBody {
  unsigned long long mask;
  int cursor, end;
};
constexpr int JPI_THREADS = 1024;  // sort and round kernels: one list append per 1024 threads
__device__ __forceinline__ bool jpTwoBodies(int2 ab) { return ab.y >= 0 && ab.y != ab.x; }
// entry (edge f, other body) has become the cursor entry at one more of f's bodies: true for the
// last of them (an edge of one body arrives once)
__device__ __forceinline__ bool jpArrive(int2 entry, int* arrived) {
  return entry.y < 0 || atomicAdd(&arrived[entry.x], 1) == 1;
}
// block-aggregated append of up to two items per thread (one atomic per block); every thread of
// the block must call it
__device__ __forceinline__ void blockAppend2(bool t0, int4 i0, bool t1, int4 i1, int4* list,
                                             int* count, int* shWave, int* shBase) {
  const unsigned long long m0 = __ballot(t0), m1 = __ballot(t1);
  const int lane = (int)(threadIdx.x & 63), wave = (int)(threadIdx.x >> 6);
  if (lane == 0) shWave[wave] = __popcll(m0) + __popcll(m1);
  __syncthreads();
  if (threadIdx.x == 0) {
    int tot = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
      const int c = shWave[w];
      shWave[w] = tot;
      tot += c;
    }
    *shBase = tot ? atomicAdd(count, tot) : 0;
  }
  __syncthreads();
  const int base = *shBase + shWave[wave];
  const unsigned long long below = (1ull << lane) - 1ull;
  if (t0) list[base + __popcll(m0 & below)] = i0;
  if (t1) list[base + __popcll(m0) + __popcll(m1 & below)] = i1;
  __syncthreads();  // the slots are reused by the next call
}
// incidence counts of the entering edges (deg[] is zero on entry), each entering edge's place in
// the lists of its bodies (slot[t]: at ends.x, at ends.y or -1), arrived[e] = 0
__global__ void k_jpi_count(const int2* ends, const int* work, const int* count, int nHost,
                            int* deg, int2* slot, int* arrived) {
  const int n = count ? *count : nHost;
  const int lane = (int)(threadIdx.x & 63);
  // wave-uniform trip count; a wave's edges are consecutive list entries, so edges listed next
  // to each other with the same first body (the usual case: edges follow the pair order) take
  // their places at that body with one atomic per run
  for (int t0 = gtid() - lane; t0 < n; t0 += gridDim.x * blockDim.x) {
    const int t = t0 + lane;
    const bool act = t < n;  // the active lanes are a prefix of the wave
    int e = 0;
    int2 ab = make_int2(-1, -1);
    if (act) {
      e = work[t];
      ab = ends[e];
    }
    const int prev = __shfl_up(ab.x, 1);
    const bool head = act && (lane == 0 || prev != ab.x);
    const unsigned long long heads = __ballot(head);
    const int nAct = __popcll(__ballot(act));
    // this lane's run: from its head lane hl to the next head (or the last active lane)
    const unsigned long long upTo = heads & ((2ull << lane) - 1ull);
    const int hl = (act && upTo) ? 63 - __clzll(upTo) : 0;
    const unsigned long long after = heads & ~((2ull << hl) - 1ull);
    const int runEnd = after ? __ffsll((long long)after) - 1 : nAct;
    int base = head ? atomicAdd(&deg[ab.x], runEnd - hl) : 0;
    base = __shfl(base, hl);
    if (act) {
      const int sb = jpTwoBodies(ab) ? atomicAdd(&deg[ab.y], 1) : -1;
      slot[t] = make_int2(base + (lane - hl), sb);
      arrived[e] = 0;
    }
  }
}
// every entering edge into the (unordered) lists of its bodies at start[] + its slot (the lists
// are put in priority order afterwards, so the order of the slots does not matter)
__global__ void k_jpi_fill(const int2* ends, const int* work, const int* count, int nHost,
                           const int* start, const int2* slot, int* lists) {
  const int n = count ? *count : nHost;
  for (int t = gtid(); t < n; t += gridDim.x * blockDim.x) {
    const int e = work[t];
    const int2 ab = ends[e];
    const int2 sl = slot[t];
    lists[start[ab.x] + sl.x] = e;
    if (sl.y >= 0) lists[start[ab.y] + sl.y] = e;
  }
}
// the list entry of edge f at body x
__device__ __forceinline__ int2 jpEntry(const int2* ends, int f, int x) {
  const int2 ab = ends[f];
  return make_int2(f, jpTwoBodies(ab) ? (ab.x == x ? ab.y : ab.x) : -1);
}
// n <= R edges of body x sorted by descending priority in registers (a priority's low word is its
// edge + 1); returns the first entry
template <int R>
__device__ __forceinline__ int2 jpiSortRegs(const int2* ends, const int* src, int2* dst, int n,
                                            int x) {
  unsigned long long v[R];
#pragma unroll
  for (int a = 0; a < R; ++a) v[a] = a < n ? colorKey1(src[a]) : 0ull;
  // odd-even transposition sort (fixed indices: stays in registers)
#pragma unroll
  for (int r = 0; r < R; ++r)
#pragma unroll
    for (int a = r & 1; a + 1 < R; a += 2) {
      const unsigned long long hi = v[a] > v[a + 1] ? v[a] : v[a + 1];
      const unsigned long long lo = v[a] > v[a + 1] ? v[a + 1] : v[a];
      v[a] = hi;
      v[a + 1] = lo;
    }
  int2 head = make_int2(0, 0);
#pragma unroll
  for (int a = 0; a < R; ++a)
    if (a < n) {
      const int2 en = jpEntry(ends, (int)(unsigned)(v[a] - 1ull), x);
      dst[a] = en;
      if (a == 0) head = en;
    }
  return head;
}
// One thread per body: sorts a list of up to 16 edges, sets the body's record (mask 0, cursor on
// the first entry) and lists that entry's edge for round 0 if it leads at all its bodies; longer
// lists go to longList for k_jpi_sort_long.
__global__ __launch_bounds__(JPI_THREADS) void k_jpi_sort(const int2* ends, const int* start,
                                                          const int* lists, int2* sorted,
                                                          int numBodies, JpBody* body,
                                                          int* arrived, int4* next,
                                                          int* nextCount, int* longList,
                                                          int* longCount) {
  __shared__ int shWave[JPI_THREADS / 64], shBase, shCnt, shBase2;
  // (no early exit: every thread reaches the block-aggregated appends)
  const int b = gtid();
  const int s = b < numBodies ? start[b] : 0;
  const int n = b < numBodies ? start[b + 1] - s : 0;
  bool lead = false, isLong = false;
  int2 head = make_int2(0, 0);
  if (n > 0) {
    JpBody rec;
    rec.mask = 0ull;
    rec.cursor = s;
    rec.end = s + n;
    body[b] = rec;
    if (n <= 8) {
      head = jpiSortRegs<8>(ends, lists + s, sorted + s, n, b);
    } else if (n <= 16) {
      head = jpiSortRegs<16>(ends, lists + s, sorted + s, n, b);
    } else {
      isLong = true;
    }
    if (!isLong) lead = jpArrive(head, arrived);
  }
  blockAppend2(lead, make_int4(head.x, b, head.y, 0), false, int4{}, next, nextCount, shWave,
               &shBase);
  blockAppend(isLong, b, longList, longCount, &shCnt, &shBase2);
}
// One workgroup per long list (grid-stride over longList): each edge's place is the number of
// the list's edges of higher priority, counted against the list staged through LDS in chunks.
constexpr int JPI_CHUNK = 2048;
constexpr int JPI_LONG_GRID = 512;
__global__ __launch_bounds__(TPB) void k_jpi_sort_long(const int2* ends, const int* start,
                                                       const int* lists, int2* sorted,
                                                       const int* longList, const int* longCount,
                                                       int* arrived, int4* next, int* nextCount) {
  __shared__ unsigned long long sk[JPI_CHUNK];
  const int nl = *longCount;
  for (int i = blockIdx.x; i < nl; i += gridDim.x) {
    const int b = longList[i];
    const int s = start[b], n = start[b + 1] - s;
    // block-uniform trip counts: every thread reaches each barrier
    for (int base = 0; base < n; base += TPB) {
      const int k = base + (int)threadIdx.x;
      const unsigned long long key = k < n ? colorKey1(lists[s + k]) : 0ull;
      int rank = 0;
      for (int c0 = 0; c0 < n; c0 += JPI_CHUNK) {
        const int m = min(JPI_CHUNK, n - c0);
        __syncthreads();
        for (int q = threadIdx.x; q < m; q += TPB) sk[q] = colorKey1(lists[s + c0 + q]);
        __syncthreads();
        if (k < n)
          for (int q = 0; q < m; ++q) rank += sk[q] > key ? 1 : 0;
      }
      if (k < n) {
        const int2 en = jpEntry(ends, (int)(unsigned)(key - 1ull), b);
        sorted[s + rank] = en;
        if (rank == 0 && jpArrive(en, arrived))
          next[atomicAdd(nextCount, 1)] = make_int4(en.x, b, en.y, 0);
      }
    }
  }
}
// One listed edge: its colour (the first fit over its bodies' masks), its bodies' cursors moved
// on, and the entries that thereby reached their last body (take*, n*) for the next list.
__device__ __forceinline__ void jpiColourOne(const int4 w, int* colour, const int2* sorted,
                                             JpBody* body, int* arrived, bool& takeA, int4& nA,
                                             bool& takeB, int4& nB) {
  const JpBody A = body[w.y];
  JpBody B{};
  if (w.z >= 0) B = body[w.z];
  // the first colour free at both bodies; 62 when none of 0..61 is
  const unsigned long long freeBits = ~(A.mask | B.mask);
  const int c = freeBits ? min(__ffsll((long long)freeBits) - 1, 62) : 62;
  colour[w.x] = c;
  const unsigned long long bit = 1ull << c;
  JpBody A1 = A;
  A1.mask |= bit;
  A1.cursor += 1;
  body[w.y] = A1;
  if (A1.cursor < A1.end) {
    const int2 en = sorted[A1.cursor];
    takeA = jpArrive(en, arrived);
    nA = make_int4(en.x, w.y, en.y, 0);
  }
  if (w.z >= 0) {
    JpBody B1 = B;
    B1.mask |= bit;
    B1.cursor += 1;
    body[w.z] = B1;
    if (B1.cursor < B1.end) {
      const int2 en = sorted[B1.cursor];
      takeB = jpArrive(en, arrived);
      nB = make_int4(en.x, w.z, en.y, 0);
    }
  }
}
// One round: colour the listed edges, advance their bodies' cursors, list the next round's edges.
__global__ __launch_bounds__(JPI_THREADS) void k_jpi_round(const int4* work, const int* count,
                                                           int* colour, const int2* sorted,
                                                           JpBody* body, int* arrived,
                                                           int4* next, int* nextCount) {
  __shared__ int shWave[JPI_THREADS / 64], shBase;
  const int n = *count;
  // block-uniform trip count: every thread reaches each block-aggregated append
  for (int base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
    const int t = base + (int)threadIdx.x;
    bool takeA = false, takeB = false;
    int4 nA{}, nB{};
    if (t < n) jpiColourOne(work[t], colour, sorted, body, arrived, takeA, nA, takeB, nB);
    blockAppend2(takeA, nA, takeB, nB, next, nextCount, shWave, &shBase);
  }
}
// Rounds j0 .. j1-1 in one workgroup (short lists): the same round body, rounds separated by
// workgroup barriers, the next list counted in LDS and each round's list size written to cnt[]
// as the launched rounds do.  Round j0 reads listA, the rounds then alternate between the two.
__global__ __launch_bounds__(JPI_THREADS) void k_jpi_tail(int4* listA, int4* listB, int* cnt,
                                                          int j0, int j1, int* colour,
                                                          const int2* sorted, JpBody* body,
                                                          int* arrived) {
  __shared__ int shN, shNext;
  if (threadIdx.x == 0) shN = cnt[j0];
  __syncthreads();
  for (int j = j0; j < j1; ++j) {
    const int n = shN;
    if (n == 0) break;  // (block-uniform) later rounds keep the zero counts they hold
    const int4* work = ((j - j0) & 1) ? listB : listA;
    int4* next = ((j - j0) & 1) ? listA : listB;
    if (threadIdx.x == 0) shNext = 0;
    __syncthreads();
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
      bool takeA = false, takeB = false;
      int4 nA{}, nB{};
      jpiColourOne(work[t], colour, sorted, body, arrived, takeA, nA, takeB, nB);
      if (takeA) next[atomicAdd(&shNext, 1)] = nA;
      if (takeB) next[atomicAdd(&shNext, 1)] = nB;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      cnt[j + 1] = shNext;
      shN = shNext;
    }
    __syncthreads();
  }
}
// Colour classes: pass 1 counts per (block, colour) with LDS atomics; k_color_offsets turns the
// counts into per-block bases; pass 2 places every edge at base + its rank inside the block.  Both
// passes walk the same contiguous chunk per block, so the per-block counts match and a class's
// rows follow the edge order up to the shuffle inside one block's pass over 256 edges.  The order
// inside a class is otherwise free (the class is body-disjoint).
constexpr int CP_BLOCKS = 1024;
// blockCnt is colour-major: blockCnt[c * CP_BLOCKS + b]
__global__ void k_color_count(const int* colour, int numEdges, int* blockCnt) {
  __shared__ int cnt[64];
  if (threadIdx.x < 64) cnt[threadIdx.x] = 0;
  __syncthreads();
  const int chunk = (numEdges + gridDim.x - 1) / gridDim.x;
  const int lo = blockIdx.x * chunk, hi = min(numEdges, lo + chunk);
  for (int e = lo + (int)threadIdx.x; e < hi; e += blockDim.x) {
    const int c = colour[e];
    if (c >= 0) atomicAdd(&cnt[c], 1);
  }
  __syncthreads();
  if (threadIdx.x < 64) blockCnt[threadIdx.x * CP_BLOCKS + blockIdx.x] = cnt[threadIdx.x];
}
// block c: exclusive scan of colour c's per-block counts; its total -> colTot[c]
__global__ void k_color_offsets(int* blockCnt, int nb, int* colTot) {
  __shared__ int sh[TPB];
  int* row = blockCnt + blockIdx.x * CP_BLOCKS;
  constexpr int PER = CP_BLOCKS / TPB;
  int v[PER];
  int sum = 0;
  for (int k = 0; k < PER; ++k) {
    const int b = threadIdx.x * PER + k;
    v[k] = b < nb ? row[b] : 0;
    sum += v[k];
  }
  int tot;
  int ex = blockExclusiveScan(sum, sh, &tot);
  for (int k = 0; k < PER; ++k) {
    const int b = threadIdx.x * PER + k;
    if (b < nb) row[b] = ex;
    ex += v[k];
  }
  if (threadIdx.x == 0) colTot[blockIdx.x] = tot;
}
// one wave: exclusive scan of the 64 colour totals
__global__ void k_color_starts(const int* colTot, int* colStart) {
  const int c = threadIdx.x;
  if (c >= 64) return;
  const int v = colTot[c];
  int incl = v;
  for (int off = 1; off < 64; off <<= 1) {
    const int o = __shfl_up(incl, off, 64);
    if (c >= off) incl += o;
  }
  colStart[c] = incl - v;
  if (c == 63) colStart[64] = incl;
}
// The last colour class (62) may hold rows that share a body; they run one after another, in
// ascending edge order (the order of a serial scan over the edges).  These two launches place
// that class's rows in edge order: contiguous chunks are counted, then each chunk writes its
// members in order at its offset (ballot ranks within a wave, wave totals within the block).
constexpr int CLS_BLOCKS = 256;
__global__ __launch_bounds__(1024) void k_class_chunk_count(const int* colour, int numEdges,
                                                            int cls, int* cnt) {
  __shared__ int total;
  if (threadIdx.x == 0) total = 0;
  __syncthreads();
  const int chunk = (numEdges + gridDim.x - 1) / gridDim.x;
  const int lo = blockIdx.x * chunk, hi = min(numEdges, lo + chunk);
  int c = 0;
  for (int e = lo + (int)threadIdx.x; e < hi; e += blockDim.x) c += colour[e] == cls;
  c = waveSumI(c);
  if ((threadIdx.x & 63) == 0 && c) atomicAdd(&total, c);
  __syncthreads();
  if (threadIdx.x == 0) cnt[blockIdx.x] = total;
}
__global__ __launch_bounds__(1024) void k_class_chunk_place(const int* colour, int numEdges,
                                                            int cls, int rowBegin,
                                                            const int* cnt, int* perm,
                                                            int* rowOf) {
  __shared__ int waveCnt[16];
  __shared__ int base;
  if (threadIdx.x == 0) {
    int b = rowBegin;
    for (int q = 0; q < (int)blockIdx.x; ++q) b += cnt[q];
    base = b;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int chunk = (numEdges + gridDim.x - 1) / gridDim.x;
  const int lo = blockIdx.x * chunk, hi = min(numEdges, lo + chunk);
  for (int e0 = lo; e0 < hi; e0 += 1024) {
    const int e = e0 + (int)threadIdx.x;
    const bool in = e < hi && colour[e] == cls;
    const unsigned long long m = __ballot(in);
    if (lane == 0) waveCnt[wave] = __popcll(m);
    __syncthreads();
    int off = base;
    for (int w = 0; w < wave; ++w) off += waveCnt[w];
    if (in) {
      const int k = off + __popcll(m & ((1ull << lane) - 1ull));
      perm[k] = e;
      rowOf[e] = k;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      int t = 0;
      for (int w = 0; w < 16; ++w) t += waveCnt[w];
      base += t;
    }
    __syncthreads();
  }
}

__global__ void k_color_place(const int* colour, int numEdges, const int* blockBase,
                              const int* colStart, int* perm, int* rowOf) {
  __shared__ int cnt[64];
  __shared__ int base[64];
  if (threadIdx.x < 64) {
    cnt[threadIdx.x] = 0;
    base[threadIdx.x] = colStart[threadIdx.x] + blockBase[threadIdx.x * CP_BLOCKS + blockIdx.x];
  }
  __syncthreads();
  const int chunk = (numEdges + gridDim.x - 1) / gridDim.x;
  const int lo = blockIdx.x * chunk, hi = min(numEdges, lo + chunk);
  for (int e = lo + (int)threadIdx.x; e < hi; e += blockDim.x) {
    const int c = colour[e];
    if (c >= 0) {
      const int k = base[c] + atomicAdd(&cnt[c], 1);
      perm[k] = e;
      rowOf[e] = k;
    } else {
      rowOf[e] = -1;
    }
  }
}

// ============================================================================================
// velocity solve — src/solver_velocity.hpp, src/contact_preprocessing.hpp:70-253
// ============================================================================================
__device__ __forceinline__ int lowerBound(const unsigned long long* keys, int n,
                                          unsigned long long k) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (keys[mid] < k) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}
// first position of `key` in the sorted ledger, or -1.  Keys with the same low real index are
// contiguous (bucket b = key >> 32 spans [bstart[b], bstart[b+1])), so the search stays inside one
// short bucket; the result equals a lower_bound over the whole ledger.
__device__ __forceinline__ int ledgerFind(const unsigned long long* keys, const int* bstart,
                                         int nbuckets, unsigned long long key) {
  const unsigned long long hi = key >> 32;
  if (hi >= (unsigned long long)nbuckets) return -1;
  const int e = bstart[hi + 1];
  for (int p = bstart[hi]; p < e; ++p) {
    const unsigned long long k = keys[p];
    if (k == key) return p;
    if (k > key) return -1;
  }
  return -1;
}

// per manifold: ledger key (markPersistentManifoldsKokkos semantics, duplicates keep their key)
// and the support relation of updateGroundedLevelsKokkos, src/contact_preprocessing.hpp:78-253
// keys, support relation (0 none, 1 B supports A, 2 A supports B, 3 boundary) and the compacted
// list of support manifolds (the only ones the grounded-level sweeps visit)

// velocity rows: gather the manifold (colour-sorted), warm start, persistence, vn0 / vt0
// (gatherWarmLambdaKokkos :113, markPersistentManifoldsKokkos :78, computeVn0Kokkos :390)
// One thread per manifold, in manifold order: the manifold sums and the bodies are read
// coalesced (manifolds follow the pair order) and the row is written at the manifold's row.
__global__ __launch_bounds__(TPB) void k_vrows(const Dev* __restrict__ Dp, Params P, int numManifolds,
                                               int ledgerBuckets, int usePGS) {
  const Dev& D = *Dp;
  const int m = gtid();
  if (m >= numManifolds || D.mColor[m] < 0) return;
  const int k = D.vRowOf[m];
  const int2 r = D.mReal[m];
  const int realA = r.x, realB = r.y;
  const int np = D.mNp[m];
  const double invN = 1.0 / (double)np;
  const double4 ns4 = D.mNs[m], ta4 = D.mTA[m], tb4 = D.mTB[m];
  const d3 Nsum = xyz(ns4), TauA = xyz(ta4), TauB = xyz(tb4);
  double restitution = P.restN;
  {
    const double ra = ns4.w * invN;
    if (ra >= 0.0) restitution = ra;
  }
  double mu = P.mu;
  {
    const double fa = ta4.w * invN;
    if (fa >= 0.0) mu = fa;
  }
  const d3 rAavg = scale3(xyz(D.mRA[m]), invN);
  const d3 rBavg = scale3(xyz(D.mRB[m]), invN);
  const double lenN = sqrt(dot3(Nsum, Nsum));
  const d3 diffCenters = (realB < 0) ? rAavg : sub3(rAavg, rBavg);
  const double alignment = dot3(Nsum, diffCenters);
  const double sgn = (alignment > 0.0) ? 1.0 : -1.0;
  D.vAB[k] = r;
  D.vWV[k] = D.mWV[m];
  D.vN[k] = tod4(Nsum, restitution);
  unsigned char flag = 0;
  if (np == 1) {
    flag |= F_ONE;  // TauA / TauB are recomputed by the rows (rowTaus)
  } else {
    D.vTA[k] = tod4(TauA, 0.0);
    D.vTB[k] = tod4(TauB, invN);
  }
  D.vRA[k] = tod4(rAavg, mu);
  D.vRB[k] = tod4(rBavg, sgn);
  const unsigned char code = D.mCode[m];
  if (code == 1) flag |= F_UP;
  if (code == 2) flag |= F_DOWN;
  double lam = 0.0;
  double4 lamT = make_double4(0, 0, 0, 0);
  double pimp = 0.0;
  if (usePGS) {
    const unsigned long long key = realKey(realA, realB);
    const int lo = ledgerFind(D.lKey, D.bucketStart, ledgerBuckets, key);
    if (lo >= 0) {
      lam = D.lLam[lo];
      const double4 lt = D.lLamT[lo];
      lamT = make_double4(lt.x, lt.y, lt.z, 0.0);
      pimp = lt.w;
      flag |= F_PERSIST;
    }
  }
  D.vLam[k] = lam;
  D.vLamT[k] = make_double4(lamT.x, lamT.y, lamT.z, pimp);  // .w: the warm-start position impulse
  D.vFlag[k] = flag;
  // computeVn0Kokkos, src/solver_velocity.hpp:390-447
  const Body64 ba = D.vw[realA];
  const d3 vA = xyz(ba.v), wA = xyz(ba.w);
  d3 vB{0, 0, 0}, wB{0, 0, 0};
  if (realB >= 0) {
    const Body64 bb = D.vw[realB];
    vB = xyz(bb.v);
    wB = xyz(bb.w);
  } else {
    vB = scale3(xyz(D.mWV[m]), invN);
  }
  const d3 vGrowth = scale3(diffCenters, P.growthRate);
  double vn = dot3(vA, Nsum) + dot3(wA, TauA) + dot3(vB, neg3(Nsum)) + dot3(wB, TauB);
  vn += dot3(vGrowth, Nsum);
  D.vVn0[k] = vn;
  d3 vt{0, 0, 0};
  if (lenN > 1e-9) {
    const d3 nhat = scale3(Nsum, 1.0 / lenN);
    const d3 wBv = (realB >= 0) ? wB : d3{0.0, 0.0, 0.0};
    const d3 vrel = sub3(add3(vA, cross3(wA, rAavg)), add3(vB, cross3(wBv, rBavg)));
    vt = sub3(vrel, scale3(nhat, dot3(vrel, nhat)));
  }
  D.vVt0[k] = tod4(vt, pimp);
}

// Guendelman support levels, double-buffered sweeps (src/contact_preprocessing.hpp:213-253)
__global__ void k_copy_i32(int* dst, const int* src, int n) {
  const int i = gtid();
  if (i < n) dst[i] = src[i];
}
// Grounded sweeps over the support list only.  Levels never decrease across the 8 sweeps, so the
// buffer written by sweep s (it holds sweep s-2's levels, which are <= the current ones) is
// brought up to date for the listed bodies by a max with the levels read, then updated; bodies
// outside the list keep their decayed level in the final buffer.  Same result as a full copy per
// sweep (src/contact_preprocessing.hpp:213-253, double-buffered).
__global__ void k_ground_seed(const Dev* __restrict__ Dp, const int* list, const int* count, const int* src, int* dst) {
  const Dev& D = *Dp;
  const int n = *count;
  for (int q = gtid(); q < n; q += gridDim.x * blockDim.x) {
    const int2 r = D.mReal[list[q]];
    dst[r.x] = src[r.x];
    if (r.y >= 0) dst[r.y] = src[r.y];
  }
}
__global__ void k_ground_sweep_list(const Dev* __restrict__ Dp, const int* list, const int* count, const int* gOld,
                                    int* gNew) {
  const Dev& D = *Dp;
  const int n = *count;
  for (int q = gtid(); q < n; q += gridDim.x * blockDim.x) {
    const int m = list[q];
    const int2 r = D.mReal[m];
    atomicMax(&gNew[r.x], gOld[r.x]);
    if (r.y >= 0) atomicMax(&gNew[r.y], gOld[r.y]);
    const unsigned char code = D.mCode[m];
    if (code == 3) {
      atomicMax(&gNew[r.x], 255);
    } else if (code == 1) {
      const int lvl = gOld[r.y] - 1;
      if (lvl > 0) atomicMax(&gNew[r.x], lvl);
    } else {
      const int lvl = gOld[r.x] - 1;
      if (lvl > 0) atomicMax(&gNew[r.y], lvl);
    }
  }
}
// The seed and the eight sweeps above in one workgroup (a short support list): each sweep is
// separated from the next by a workgroup barrier, so every sweep reads the levels its
// predecessor finished, as the launch-per-sweep form does; the result ends in `ga` (eight
// swaps).  The levels a sweep reads were written by atomics of the previous sweep, so they are
// read with device-scope loads (no stale copies in the vector cache).
constexpr int GROUND_BLOCK_THREADS = 1024;
constexpr int GROUND_BLOCK_MAX = 16384;  // manifolds (the support list's bound)
__device__ __forceinline__ int loadLevel(const int* g, int i) {
  return __hip_atomic_load(g + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__global__ __launch_bounds__(GROUND_BLOCK_THREADS) void k_ground_block(
    const Dev* __restrict__ Dp, const int* list, const int* count, int* ga, int* gb) {
  const Dev& D = *Dp;
  const int n = *count;
  for (int q = threadIdx.x; q < n; q += blockDim.x) {
    const int2 r = D.mReal[list[q]];
    gb[r.x] = ga[r.x];
    if (r.y >= 0) gb[r.y] = ga[r.y];
  }
  __syncthreads();
  int* gOld = ga;
  int* gNew = gb;
  for (int sweep = 0; sweep < 8; ++sweep) {
    for (int q = threadIdx.x; q < n; q += blockDim.x) {
      const int m = list[q];
      const int2 r = D.mReal[m];
      atomicMax(&gNew[r.x], loadLevel(gOld, r.x));
      if (r.y >= 0) atomicMax(&gNew[r.y], loadLevel(gOld, r.y));
      const unsigned char code = D.mCode[m];
      if (code == 3) {
        atomicMax(&gNew[r.x], 255);
      } else if (code == 1) {
        const int lvl = loadLevel(gOld, r.y) - 1;
        if (lvl > 0) atomicMax(&gNew[r.x], lvl);
      } else {
        const int lvl = loadLevel(gOld, r.x) - 1;
        if (lvl > 0) atomicMax(&gNew[r.y], lvl);
      }
    }
    __syncthreads();
    int* t = gOld;
    gOld = gNew;
    gNew = t;
  }
}
__global__ void k_ground_sweep(const Dev* __restrict__ Dp, int numManifolds, const int* gOld, int* gNew) {
  const Dev& D = *Dp;
  const int m = gtid();
  if (m >= numManifolds) return;
  const unsigned char code = D.mCode[m];
  if (code == 0) return;
  const int2 r = D.mReal[m];
  if (code == 3) {
    atomicMax(&gNew[r.x], 255);
  } else if (code == 1) {
    const int lvl = gOld[r.y] - 1;
    if (lvl > 0) atomicMax(&gNew[r.x], lvl);
  } else {
    const int lvl = gOld[r.x] - 1;
    if (lvl > 0) atomicMax(&gNew[r.y], lvl);
  }
}

__device__ __forceinline__ void addVW(Body64& b, d3 dv, d3 dw) {
  b.v.x += dv.x;
  b.v.y += dv.y;
  b.v.z += dv.z;
  b.w.x += dw.x;
  b.w.y += dw.y;
  b.w.z += dw.z;
}

// Where a velocity row finds its bodies' velocities: the global array (GBodies), or a copy of
// the rows' bodies staged in LDS by a one-workgroup sweep loop (LBodies: 6 doubles per body,
// rows hold the two LDS slots).  The arithmetic is the same either way.
struct GBodies {
  Body64* vw;
  __device__ void row(const Dev& D, int k, int2& r, int2& sl) const {
    r = D.vAB[k];
    sl = make_int2(0, 0);
  }
  __device__ Body64 load(int real, int) const { return vw[real]; }
  __device__ void store(int real, int, const Body64& b) const { vw[real] = b; }
  __device__ BodyC cload(const Dev& D, int real) const { return D.bc[real]; }
};
struct PBodies {
  VBody* vb;
  __device__ void row(const Dev& D, int k, int2& r, int2& sl) const {
    r = D.vAB[k];
    sl = make_int2(0, 0);
  }
  __device__ Body64 load(int real, int) const {
    Body64 b;
    b.v = vb[real].v;
    b.w = vb[real].w;
    return b;
  }
  __device__ void store(int real, int, const Body64& b) const {
    vb[real].v = b.v;
    vb[real].w = b.w;
  }
  __device__ BodyC cload(const Dev&, int real) const {
    BodyC c;
    c.q = vb[real].q;
    c.iim = vb[real].iim;
    return c;
  }
};
__global__ void k_vb_pack(const Dev* __restrict__ Dp, VBody* vb, int n) {
  const Dev& D = *Dp;
  const int i = gtid();
  if (i >= n) return;
  // field by field: a local 128-byte-aligned record would be placed in scratch memory
  vb[i].v = D.vw[i].v;
  vb[i].w = D.vw[i].w;
  vb[i].q = D.bc[i].q;
  vb[i].iim = D.bc[i].iim;
}
__global__ void k_vb_unpack(const Dev* __restrict__ Dp, const VBody* vb, int n) {
  const Dev& D = *Dp;
  const int i = gtid();
  if (i >= n) return;
  Body64 b;
  b.v = vb[i].v;
  b.w = vb[i].w;
  D.vw[i] = b;
}
// LDS slot codes of a row's bodies: a body without an LDS slot stays in global memory
constexpr int VL_GLOBAL = 0xFFFE, VL_NONE = 0xFFFF;
struct LBodies {
  double2* bv;             // body s: bv[3s] = (v.x, v.y), bv[3s+1] = (v.z, w.x), bv[3s+2] = (w.y, w.z)
  const unsigned* slots;   // row k - r0: slot of a | slot of b << 16
  const int* bodyOf;       // slot -> real body
  PD_G Body64* vw;         // bodies without a slot
  int r0;
  __device__ void row(const Dev& D, int k, int2& r, int2& sl) const {
    const unsigned p = slots[k - r0];
    sl.x = (int)(p & 0xFFFFu);
    sl.y = (int)(p >> 16);
    if (sl.x == VL_GLOBAL || sl.y == VL_GLOBAL) {
      r = D.vAB[k];
    } else {
      r.x = bodyOf[sl.x];
      r.y = sl.y != VL_NONE ? bodyOf[sl.y] : -1;
    }
  }
  __device__ Body64 load(int real, int s) const {
    if (s == VL_GLOBAL) return vw[real];
    const double2 a = bv[3 * s], b = bv[3 * s + 1], c = bv[3 * s + 2];
    Body64 o;
    o.v = make_double4(a.x, a.y, b.x, 0.0);
    o.w = make_double4(b.y, c.x, c.y, 0.0);
    return o;
  }
  __device__ void store(int real, int s, const Body64& o) const {
    if (s == VL_GLOBAL) {
      vw[real] = o;
      return;
    }
    bv[3 * s] = make_double2(o.v.x, o.v.y);
    bv[3 * s + 1] = make_double2(o.v.z, o.w.x);
    bv[3 * s + 2] = make_double2(o.w.y, o.w.z);
  }
  __device__ BodyC cload(const Dev& D, int real) const { return D.bc[real]; }
};

// A velocity row's constants as k_vrows lays them out: vN = (Nsum, restitution), vRA = (rAavg,
// mu), vRB = (rBavg, sgn), vTA = TauA, vTB = (TauB, 1 / points); |Nsum| is recomputed (the same
// expression k_vrows evaluates).  A one-point row (F_ONE) reads neither vTA nor vTB: its sums are
// one-term sums, TauA = 0 + rA x n and TauB = 0 + rB x (-n) as k_mreduce formed them from the
// contact's own arm and normal, and rA = 0 + arm, Nsum = 0 + n differ from those at most in the
// sign of zero components, which moves no bit of the products' 0 + sums; 1 / points = 1.
__device__ __forceinline__ double rowLenN(d3 Nsum) { return sqrt(dot3(Nsum, Nsum)); }
__device__ __forceinline__ void rowTaus(const Dev& D, int k, unsigned char fl, d3 Nsum, d3 rAavg,
                                        d3 rBavg, d3& TauA, d3& TauB, double& invN) {
  if (fl & F_ONE) {
    const d3 z{0.0, 0.0, 0.0};
    TauA = add3(z, cross3(rAavg, Nsum));
    TauB = add3(z, cross3(rBavg, neg3(Nsum)));
    invN = 1.0;
  } else {
    TauA = xyz(D.vTA[k]);
    const double4 tb4 = D.vTB[k];
    TauB = xyz(tb4);
    invN = tb4.w;
  }
}
// warmStartApplyKokkos, src/solver_velocity.hpp:498-605 (one row)
template <class BS>
__device__ __forceinline__ void warmApplyRowB(const Dev& D, int k, int2 r, int2 sl, const BS& bs) {
  const double p = D.vLam[k];
  const double4 lt4 = D.vLamT[k];
  d3 lt = xyz(lt4);
  const bool hasT = nonzero3(lt);
  if (p == 0.0 && !hasT) return;
  const int realA = r.x, realB = r.y;
  const d3 Nsum = xyz(D.vN[k]);
  const d3 rAavg = xyz(D.vRA[k]);
  const double4 rb4 = D.vRB[k];
  const d3 rBavg = xyz(rb4);
  const double sgn = rb4.w;
  d3 TauA, TauB;
  double invN;
  rowTaus(D, k, D.vFlag[k], Nsum, rAavg, rBavg, TauA, TauB, invN);
  const double lambda = -sgn * p;
  const d3 Jlin = scale3(Nsum, lambda);
  const BodyC ca = bs.cload(D, realA);
  const double invMassA = ca.iim.w;
  Body64 ba = bs.load(realA, sl.x);
  const q4 qA = toq(ca.q);
  const d3 invIA = xyz(ca.iim);
  ba.v.x += Jlin.x * invMassA;
  ba.v.y += Jlin.y * invMassA;
  ba.v.z += Jlin.z * invMassA;
  {
    const d3 dw = angDelta(qA, scale3(TauA, lambda), invIA);
    ba.w.x += dw.x;
    ba.w.y += dw.y;
    ba.w.z += dw.z;
  }
  Body64 bb;
  BodyC cb;
  double invMassB = 0.0;
  if (realB >= 0) {
    cb = bs.cload(D, realB);
    invMassB = cb.iim.w;
    bb = bs.load(realB, sl.y);
    bb.v.x += -Jlin.x * invMassB;
    bb.v.y += -Jlin.y * invMassB;
    bb.v.z += -Jlin.z * invMassB;
    const d3 dw = angDelta(toq(cb.q), scale3(TauB, lambda), xyz(cb.iim));
    bb.w.x += dw.x;
    bb.w.y += dw.y;
    bb.w.z += dw.z;
  }
  if (hasT) {
    const double lenN2 = sqrt(dot3(Nsum, Nsum));
    if (lenN2 > 1e-9) {
      const d3 nhat = scale3(Nsum, 1.0 / lenN2);
      lt = sub3(lt, scale3(nhat, dot3(lt, nhat)));
    }
    D.vLamT[k] = tod4(lt, lt4.w);
    ba.v.x += lt.x * invMassA;
    ba.v.y += lt.y * invMassA;
    ba.v.z += lt.z * invMassA;
    {
      const d3 dw = angDelta(qA, cross3(rAavg, lt), invIA);
      ba.w.x += dw.x;
      ba.w.y += dw.y;
      ba.w.z += dw.z;
    }
    if (realB >= 0) {
      bb.v.x += -lt.x * invMassB;
      bb.v.y += -lt.y * invMassB;
      bb.v.z += -lt.z * invMassB;
      const d3 dw = angDelta(toq(cb.q), cross3(rBavg, scale3(lt, -1.0)), xyz(cb.iim));
      bb.w.x += dw.x;
      bb.w.y += dw.y;
      bb.w.z += dw.z;
    }
  }
  bs.store(realA, sl.x, ba);
  if (realB >= 0) bs.store(realB, sl.y, bb);
}
__device__ __forceinline__ void warmApplyRow(const Dev& D, int k) {
  const GBodies g{D.vw};
  int2 r, sl;
  g.row(D, k, r, sl);
  warmApplyRowB(D, k, r, sl, g);
}

// A velocity row's constants in the PGS / GS updates (body masses, inertia and orientation do
// not change during the velocity solve), evaluated by the row from the bodies it loads.
struct RowPre {
  double wAn, wBn, growth, invLenN;
};
__device__ __forceinline__ RowPre rowPre(const Params& P, int realB, d3 Nsum, d3 TauA, d3 TauB,
                                         d3 rAavg, d3 rBavg, double lenN, double invMassA,
                                         d3 invIA, q4 qA, double invMassB, d3 invIB, q4 qB) {
  RowPre r;
  const double Nsq = dot3(Nsum, Nsum);
  r.wAn = Nsq * invMassA + genInvMass(TauA, invIA, qA);
  r.wBn = Nsq * invMassB + genInvMass(TauB, invIB, qB);
  const d3 diffCenters = (realB < 0) ? rAavg : sub3(rAavg, rBavg);
  r.growth = dot3(scale3(diffCenters, P.growthRate), Nsum);
  r.invLenN = lenN > 0.0 ? 1.0 / lenN : 0.0;
  return r;
}
// PGSManifoldSweep::solveOne, src/solver_velocity.hpp:641-858 (one row)
template <class BS>
__device__ __forceinline__ double pgsRowB(const Dev& D, const Params& P, int k, int2 r, int2 sl,
                          double restVelThreshold, const BS& bs) {
  double amax = 0.0;
  const int realA = r.x, realB = r.y;
  const BodyC ca = bs.cload(D, realA);
  const double invMassA = ca.iim.w;
  const d3 invIA = xyz(ca.iim);
  const q4 qA = toq(ca.q);
  BodyC cb;
  double invMassB = 0.0;
  d3 invIB{0, 0, 0};
  q4 qB{0, 0, 0, 1};
  Body64 ba = bs.load(realA, sl.x);
  Body64 bb;
  d3 vB{0, 0, 0}, wB{0, 0, 0};
  const unsigned char fl = D.vFlag[k];
  const double4 n4 = D.vN[k], ra4 = D.vRA[k], rb4 = D.vRB[k];
  const d3 Nsum = xyz(n4);
  const d3 rAavg = xyz(ra4), rBavg = xyz(rb4);
  d3 TauA, TauB;
  double invN;
  rowTaus(D, k, fl, Nsum, rAavg, rBavg, TauA, TauB, invN);
  if (realB >= 0) {
    cb = bs.cload(D, realB);
    invMassB = cb.iim.w;
    invIB = xyz(cb.iim);
    qB = toq(cb.q);
    bb = bs.load(realB, sl.y);
    vB = xyz(bb.v);
    wB = xyz(bb.w);
  } else {
    vB = scale3(xyz(D.vWV[k]), invN);
  }
  const d3 vA = xyz(ba.v), wA = xyz(ba.w);
  double restitution = n4.w;
  const double lenN = rowLenN(Nsum);
  if (lenN < 1e-9) return 0.0;
  // the row's constants, from the bodies just loaded (cheaper than reading them back)
  const RowPre pre = rowPre(P, realB, Nsum, TauA, TauB, rAavg, rBavg, lenN, invMassA, invIA, qA,
                            invMassB, invIB, qB);
  double vn = dot3(vA, Nsum) + dot3(wA, TauA) + dot3(vB, neg3(Nsum)) + dot3(wB, TauB);
  vn += pre.growth;  // dot3(diffCenters * growthRate, Nsum)
  const double sgn = rb4.w;
  double wA_n = pre.wAn;  // Nsq * invMassA + genInvMass(TauA, invIA, qA)
  double wB_n = pre.wBn;
  bool applyA = true, applyB = true;
  const int sf = (fl & F_SIDE1) ? 1 : ((fl & F_SIDE2) ? 2 : 0);
  if (sf == 1) {
    wB_n = 0.0;
    applyB = false;
    restitution = 0.0;
  } else if (sf == 2) {
    wA_n = 0.0;
    applyA = false;
    restitution = 0.0;
  }
  const double wTotal = wA_n + wB_n;
  if (wTotal <= 0.0) return 0.0;
  const double vn0 = D.vVn0[k];
  const double v0til = sgn * vn0;
  if (fabs(vn0) < restVelThreshold * lenN) restitution = 0.0;
  const double target = (v0til > 0.0) ? -restitution * v0til : 0.0;
  const double vtil = sgn * vn;
  const double dp = (vtil - target) / wTotal;
  const double pOld = D.vLam[k];
  double pNew = pOld + dp;
  if (pNew < 0.0) pNew = 0.0;
  const double dApplied = pNew - pOld;
  const double lamNow = dApplied != 0.0 ? pNew : pOld;  // the row's impulse after the update
  bool dirtyA = false, dirtyB = false;
  if (dApplied != 0.0) {
    D.vLam[k] = pNew;
    amax = fabs(dApplied) * wTotal / lenN;
    const double lambda = -sgn * dApplied;
    const d3 Jlin = scale3(Nsum, lambda);
    if (applyA) {
      ba.v.x += Jlin.x * invMassA;
      ba.v.y += Jlin.y * invMassA;
      ba.v.z += Jlin.z * invMassA;
      const d3 dw = angDelta(qA, scale3(TauA, lambda), invIA);
      ba.w.x += dw.x;
      ba.w.y += dw.y;
      ba.w.z += dw.z;
      dirtyA = true;
    }
    if (realB >= 0 && applyB) {
      bb.v.x += -Jlin.x * invMassB;
      bb.v.y += -Jlin.y * invMassB;
      bb.v.z += -Jlin.z * invMassB;
      const d3 dw = angDelta(qB, scale3(TauB, lambda), invIB);
      bb.w.x += dw.x;
      bb.w.y += dw.y;
      bb.w.z += dw.z;
      dirtyB = true;
    }
  }
  // friction cone, src/solver_velocity.hpp:760-855
  double mu = ra4.w;
  const double4 lto4 = D.vLamT[k];  // .w: the warm-start position impulse
  const d3 ltOld = xyz(lto4);
  const bool haveOld = nonzero3(ltOld);
  if (mu > 0.0 || haveOld) {
    const d3 nhat = scale3(Nsum, pre.invLenN);
    const d3 vA2 = xyz(ba.v), wA2 = xyz(ba.w);
    d3 vB2{0, 0, 0}, wB2{0, 0, 0};
    if (realB >= 0) {
      vB2 = xyz(bb.v);
      wB2 = xyz(bb.w);
    } else {
      vB2 = scale3(xyz(D.vWV[k]), invN);
    }
    const d3 vrel = sub3(add3(vA2, cross3(wA2, rAavg)), add3(vB2, cross3(wB2, rBavg)));
    const d3 vt = sub3(vrel, scale3(nhat, dot3(vrel, nhat)));
    double beta = P.restT;
    if (beta != 0.0 && (sf != 0 || fabs(vn0) < restVelThreshold * lenN)) beta = 0.0;
    d3 vtErr = vt;
    if (beta != 0.0) {
      d3 v0 = xyz(D.vVt0[k]);
      v0 = sub3(v0, scale3(nhat, dot3(v0, nhat)));
      vtErr = add3(vt, scale3(v0, beta));
    }
    const double vtLen = sqrt(dot3(vtErr, vtErr));
    d3 ltNew = ltOld;
    double wT = invMassA + invMassB;
    if (vtLen > 1e-9) {
      const d3 that = scale3(vtErr, 1.0 / vtLen);
      const double wAt = (sf == 2) ? 0.0 : invMassA + genInvMass(cross3(rAavg, that), invIA, qA);
      const double wBt =
          (sf == 1 || realB < 0) ? 0.0 : invMassB + genInvMass(cross3(rBavg, that), invIB, qB);
      wT = wAt + wBt;
      if (wT > 1e-9) ltNew = add3(ltOld, scale3(that, -vtLen / wT));
    }
    ltNew = sub3(ltNew, scale3(nhat, dot3(ltNew, nhat)));
    double nTot = fmax(lamNow, 0.0) * lenN;
    if (fabs(vn0) < restVelThreshold * lenN) nTot += fmax(lto4.w, 0.0);
    const double bound = mu * nTot;
    const double ltLen = sqrt(dot3(ltNew, ltNew));
    if (ltLen > bound) ltNew = (bound > 0.0) ? scale3(ltNew, bound / ltLen) : d3{0.0, 0.0, 0.0};
    const d3 dApp = sub3(ltNew, ltOld);
    if (nonzero3(dApp)) {
      D.vLamT[k] = tod4(ltNew, lto4.w);
      const double dLen = sqrt(dot3(dApp, dApp));
      if (wT > 1e-9 && dLen * wT > amax) amax = dLen * wT;
      if (applyA) {
        ba.v.x += dApp.x * invMassA;
        ba.v.y += dApp.y * invMassA;
        ba.v.z += dApp.z * invMassA;
        const d3 dw = angDelta(qA, cross3(rAavg, dApp), invIA);
        ba.w.x += dw.x;
        ba.w.y += dw.y;
        ba.w.z += dw.z;
        dirtyA = true;
      }
      if (realB >= 0 && applyB) {
        bb.v.x += -dApp.x * invMassB;
        bb.v.y += -dApp.y * invMassB;
        bb.v.z += -dApp.z * invMassB;
        const d3 dw = angDelta(qB, cross3(rBavg, scale3(dApp, -1.0)), invIB);
        bb.w.x += dw.x;
        bb.w.y += dw.y;
        bb.w.z += dw.z;
        dirtyB = true;
      }
    }
  }
  if (dirtyA) bs.store(realA, sl.x, ba);
  if (dirtyB) bs.store(realB, sl.y, bb);
  return amax;
}
__device__ __forceinline__ double pgsRow(const Dev& D, const Params& P, int k, double restVelThreshold) {
  const GBodies g{D.vw};
  int2 r, sl;
  g.row(D, k, r, sl);
  return pgsRowB(D, P, k, r, sl, restVelThreshold, g);
}

// solveVelocityColoredGSKokkos one-shot restitution (g = 0), src/solver_velocity.hpp:989-1133
template <class BS>
__device__ __forceinline__ double gsRowB(const Dev& D, const Params& P, int k, int2 r, int2 sl,
                         double restVelThreshold, const BS& bs) {
  const int realA = r.x, realB = r.y;
  const BodyC ca = bs.cload(D, realA);
  const double invMassA = ca.iim.w;
  const d3 invIA = xyz(ca.iim);
  const q4 qA = toq(ca.q);
  BodyC cb;
  double invMassB = 0.0;
  d3 invIB{0, 0, 0};
  q4 qB{0, 0, 0, 1};
  Body64 ba = bs.load(realA, sl.x);
  Body64 bb;
  d3 vB{0, 0, 0}, wB{0, 0, 0};
  const unsigned char fl = D.vFlag[k];
  const double4 n4 = D.vN[k], ra4 = D.vRA[k], rb4 = D.vRB[k];
  const d3 Nsum = xyz(n4);
  const d3 rAavg = xyz(ra4), rBavg = xyz(rb4);
  d3 TauA, TauB;
  double invN;
  rowTaus(D, k, fl, Nsum, rAavg, rBavg, TauA, TauB, invN);
  if (realB >= 0) {
    cb = bs.cload(D, realB);
    invMassB = cb.iim.w;
    invIB = xyz(cb.iim);
    qB = toq(cb.q);
    bb = bs.load(realB, sl.y);
    vB = xyz(bb.v);
    wB = xyz(bb.w);
  } else {
    vB = scale3(xyz(D.vWV[k]), invN);
  }
  const d3 vA = xyz(ba.v), wA = xyz(ba.w);
  double restitution = n4.w;
  const double lenN = rowLenN(Nsum);
  if (lenN < 1e-9) return 0.0;
  const d3 diffCenters = (realB < 0) ? rAavg : sub3(rAavg, rBavg);
  const RowPre pre = rowPre(P, realB, Nsum, TauA, TauB, rAavg, rBavg, lenN, invMassA, invIA, qA,
                            invMassB, invIB, qB);
  double vn = dot3(vA, Nsum) + dot3(wA, TauA) + dot3(vB, neg3(Nsum)) + dot3(wB, TauB);
  vn += pre.growth;
  const double alignment = dot3(Nsum, diffCenters);
  if (alignment > 0.0) {
    if (vn < 0.0) return 0.0;
  } else {
    if (vn > 0.0) return 0.0;
  }
  const double wA_n = pre.wAn, wB_n = pre.wBn;
  const double wTotal = wA_n + wB_n;
  if (wTotal <= 0.0) return 0.0;
  const double amax = fabs(vn) / lenN;
  if (fabs(vn) < restVelThreshold * lenN) restitution = 0.0;
  const double lambda = (-restitution * vn - vn) / wTotal;
  const d3 Jlin = scale3(Nsum, lambda);
  ba.v.x += Jlin.x * invMassA;
  ba.v.y += Jlin.y * invMassA;
  ba.v.z += Jlin.z * invMassA;
  {
    const d3 dw = angDelta(qA, scale3(TauA, lambda), invIA);
    ba.w.x += dw.x;
    ba.w.y += dw.y;
    ba.w.z += dw.z;
  }
  bs.store(realA, sl.x, ba);
  if (realB >= 0) {
    bb.v.x += -Jlin.x * invMassB;
    bb.v.y += -Jlin.y * invMassB;
    bb.v.z += -Jlin.z * invMassB;
    const d3 dw = angDelta(qB, scale3(TauB, lambda), invIB);
    bb.w.x += dw.x;
    bb.w.y += dw.y;
    bb.w.z += dw.z;
    bs.store(realB, sl.y, bb);
  }
  return amax;
}
__device__ __forceinline__ double gsRow(const Dev& D, const Params& P, int k, double restVelThreshold) {
  const GBodies g{D.vw};
  int2 r, sl;
  g.row(D, k, r, sl);
  return gsRowB(D, P, k, r, sl, restVelThreshold, g);
}
// the same rows on packed body records: RowOf<false> reads vw / bc, RowOf<true> one VBody each
template <bool PK>
struct RowOf {
  __device__ __forceinline__ static void warm(const Dev& D, int k, VBody*) { warmApplyRow(D, k); }
  __device__ __forceinline__ static double pgs(const Dev& D, const Params& P, int k, double thr, VBody*) {
    return pgsRow(D, P, k, thr);
  }
  __device__ __forceinline__ static double gs(const Dev& D, const Params& P, int k, double thr, VBody*) {
    return gsRow(D, P, k, thr);
  }
};
template <>
struct RowOf<true> {
  __device__ __forceinline__ static void warm(const Dev& D, int k, VBody* vb) {
    const PBodies p{vb};
    int2 r, sl;
    p.row(D, k, r, sl);
    warmApplyRowB(D, k, r, sl, p);
  }
  __device__ static double pgs(const Dev& D, const Params& P, int k, double thr, VBody* vb) {
    const PBodies p{vb};
    int2 r, sl;
    p.row(D, k, r, sl);
    return pgsRowB(D, P, k, r, sl, thr, p);
  }
  __device__ static double gs(const Dev& D, const Params& P, int k, double thr, VBody* vb) {
    const PBodies p{vb};
    int2 r, sl;
    p.row(D, k, r, sl);
    return gsRowB(D, P, k, r, sl, thr, p);
  }
};

// computeSideFlagsKokkos, src/solver_velocity.hpp:456-495 (rows are active, non-duplicate)
template <class BS>
__device__ __forceinline__ void sideFlagsRow(const Dev& D, int k, d3 gHat, double riseThr,
                                             double approachThr, const BS& bs) {
  unsigned char fl = D.vFlag[k] & (unsigned char)~(F_SIDE1 | F_SIDE2);
  int2 r, sl;
  bs.row(D, k, r, sl);
  if (r.y >= 0 && (fl & F_PERSIST) && !(fabs(D.vVn0[k]) > approachThr)) {
    if ((fl & F_UP) && -dot3(xyz(bs.load(r.y, sl.y).v), gHat) <= riseThr && D.ground[r.y] > 0)
      fl |= F_SIDE1;
    else if ((fl & F_DOWN) && -dot3(xyz(bs.load(r.x, sl.x).v), gHat) <= riseThr &&
             D.ground[r.x] > 0)
      fl |= F_SIDE2;
  }
  D.vFlag[k] = fl;
}
// `need` (optional) gates the launch like `run`
// vb (optional): the stage's packed body records
__global__ void k_side_flags(const Dev* __restrict__ Dp, int numRows, d3 gHat, double riseThr, double approachThr,
                             const int* run, const int* need, VBody* vb = nullptr) {
  const Dev& D = *Dp;
  const int k = gtid();
  if (k >= numRows || !*run || (need && !*need)) return;
  if (vb)
    sideFlagsRow(D, k, gHat, riseThr, approachThr, PBodies{vb});
  else
    sideFlagsRow(D, k, gHat, riseThr, approachThr, GBodies{D.vw});
}

// mode: 0 = warm start, 1 = PGS sweep, 2 = one-shot GS sweep
// ---- device-side loop control: ctl[0..] ints (see pd_sim.hip), sweeps after a stop are no-ops
enum { CTL_STOP_V = 0, CTL_SWEEPS_V, CTL_STAB_RUN, CTL_STOP_S, CTL_SWEEPS_S, CTL_STOP_P,
       CTL_SWEEPS_P, CTL_NONE, CTL_VFALLBACK, CTL_COUNT };
__global__ void k_sweep_begin(unsigned long long* bits, const int* stop) {
  if (!*stop) *bits = 0ull;
}
// after one sweep: count it and stop once its max is within the tolerance; with `next` set (a
// sweep of the same loop follows) a sweep that does not stop also clears the accumulator for the
// next one, the work of k_sweep_begin
__global__ void k_sweep_check(unsigned long long* bits, int* stop, int* count, double thr,
                              int strict, int next) {
  if (*stop) return;
  *count += 1;
  const double m = bitsd(*bits);
  if (strict ? (m < thr) : (m <= thr)) *stop = 1;
  else if (next) *bits = 0ull;
}
// stabilization gate (src/sim.hpp:279-283): run iff mode != 0 and the residual exceeds vRestS
__global__ void k_stab_gate(const unsigned long long* bits, int* ctl, int mode, double vRestS,
                            const int* need) {
  if (need && !*need) return;
  const int run = (mode != 0 && bitsd(*bits) > vRestS) ? 1 : 0;
  ctl[CTL_STAB_RUN] = run;
  ctl[CTL_STOP_S] = run ? 0 : 1;
}

// PK: bodies in the packed records vb (else vw / bc)
template <int MODE, bool PK = false>
__global__ __launch_bounds__(TPB) void k_vsweep(const Dev* __restrict__ Dp, Params P, int rowBegin, int rowEnd,
                                                double thr, const int* stop, VBody* vb = nullptr) {
  const Dev& D = *Dp;
  if (*stop) return;
  __shared__ unsigned long long sh[WAVES];
  const int k = rowBegin + xcdTid();
  double a = 0.0;
  if (k < rowEnd) {
    if (MODE == 0) RowOf<PK>::warm(D, k, vb);
    else if (MODE == 1) a = RowOf<PK>::pgs(D, P, k, thr, vb);
    else a = RowOf<PK>::gs(D, P, k, thr, vb);
  }
  if (MODE != 0) blockAtomicMaxU64(D.maxApproachBits, dbits(a), sh);
}
// the PGS sweep compiled for at least 4 waves per SIMD (at most 128 VGPRs instead of ~153):
// more rows in flight per CU to hide the body gathers' latency, against register pressure; the
// autotuner picks between the two
template <bool PK = false>
__global__ __launch_bounds__(TPB, 4) void k_vsweep_o4(const Dev* __restrict__ Dp, Params P,
                                                      int rowBegin, int rowEnd, double thr,
                                                      const int* stop, VBody* vb = nullptr) {
  const Dev& D = *Dp;
  if (*stop) return;
  __shared__ unsigned long long sh[WAVES];
  const int k = rowBegin + xcdTid();
  double a = 0.0;
  if (k < rowEnd) a = RowOf<PK>::pgs(D, P, k, thr, vb);
  blockAtomicMaxU64(D.maxApproachBits, dbits(a), sh);
}
// colour 62 may hold manifolds sharing a body (mask saturation): one thread, in row order (the
// class's rows are placed in manifold order)
__global__ __launch_bounds__(64) void k_vsweep_serial(const Dev* __restrict__ Dp, Params P, int rowBegin,
                                                      int rowEnd, int mode, double thr,
                                                      const int* stop, VBody* vb = nullptr) {
  const Dev& D = *Dp;
  if (gtid() != 0 || *stop) return;
  for (int k = rowBegin; k < rowEnd; ++k) {
    if (vb) {
      if (mode == 0) RowOf<true>::warm(D, k, vb);
      else if (mode == 1) atomicMaxD(D.maxApproachBits, RowOf<true>::pgs(D, P, k, thr, vb));
      else atomicMaxD(D.maxApproachBits, RowOf<true>::gs(D, P, k, thr, vb));
    } else {
      if (mode == 0) warmApplyRow(D, k);
      else if (mode == 1) atomicMaxD(D.maxApproachBits, pgsRow(D, P, k, thr));
      else atomicMaxD(D.maxApproachBits, gsRow(D, P, k, thr));
    }
  }
}

// ============================================================================================
// stabilization mode "multilevel" — src/solver_multilevel.hpp, driven by src/sim.hpp:303-355.
// Aggregation levels over the quasi-static rows (one round of random-priority matching per
// level), a colouring of each level's crossing rows, and per loop iteration a fine sweep that
// also tracks the quasi-static residual, then per level restrict / coarse sweeps / prolongate.
// A group's mass and momentum are summed over the aggregation tree (a level-l group = its
// leader's level-(l-1) sum plus its mate's); the coarse colour 62, whose rows may share a
// group, runs in ascending manifold order.
// ============================================================================================
constexpr int ML_MAX_LEVELS = 10;        // kMlMaxLevels, src/solver_multilevel.hpp:44
constexpr unsigned char ML_SKIP = 63;    // kMlSlotSkip, :45
constexpr int ML_GATE_PERSISTENT = 1, ML_GATE_CONE = 2, ML_GATE_SLIP = 4;  // :82-84
struct MlDev {
  unsigned char* elig;         // [rows] coarse-eligible
  unsigned char* slot;         // [rows] this level's colour (ML_SKIP = none)
  int* grp;                    // [numReal] body -> group of the current level
  unsigned long long* winner;  // [numReal] largest key + 1 (0 = none)
  int* mate;                   // [numReal]
  int* lead;                   // [numReal + 1] leader flags
  int* scan;                   // [numReal + 1] their exclusive scan
  int* parent;                 // pooled by level: finer id -> group
  int2* child;                 // pooled by group: (leader, mate or -1)
  double* massG;
  double* invMassG;
  double4* velG;
  double4* velG0;
  double4* mom;                // [numReal + group pool] momentum tree
  unsigned long long* mask;    // [numReal] committed colours per group
  int* col;                    // [manifolds] this level's colour per manifold (-1 none)
};
// mldetail::effMass, src/solver_multilevel.hpp:68-70
__device__ __forceinline__ double mlEffMass(double invMass) { return 1.0 / fmax(invMass, 1e-30); }
// mldetail::eligible, src/solver_multilevel.hpp:88-119 (rows are the coloured manifolds)
__global__ __launch_bounds__(TPB) void k_ml_elig(const Dev* __restrict__ Dp, MlDev ml, int rows, double qsThr,
                                                 int gates, d3 gHat) {
  const Dev& D = *Dp;
  const int k = gtid();
  if (k >= rows) return;
  const int m = D.vPerm[k];
  unsigned char e = D.mNp[m] > 0;
  if (e && (gates & ML_GATE_PERSISTENT) && !(D.vFlag[k] & F_PERSIST)) e = 0;
  if (e && fabs(D.vVn0[k]) > qsThr) e = 0;
  if (e && (gates & ML_GATE_SLIP)) {
    const d3 vt = xyz(D.vVt0[k]);
    if (dot3(vt, vt) > qsThr * qsThr) e = 0;
  }
  if (e && D.vAB[k].y >= 0 && (gates & ML_GATE_CONE)) {
    const int2 ab = D.mAB[m];
    const d3 dx = sub3(xyz(D.posPred[ab.x]), xyz(D.posPred[ab.y]));
    const double up = -(dx.x * gHat.x + dx.y * gHat.y + dx.z * gHat.z);
    e = fabs(up) > 0.3 * sqrt(dot3(dx, dx));
  }
  ml.elig[k] = e;
}
__global__ void k_ml_iota(int* a, int n) {
  const int i = gtid();
  if (i < n) a[i] = i;
}
// ml_match_contend / ml_match_commit, src/solver_multilevel.hpp:169-207
__global__ __launch_bounds__(TPB) void k_ml_match_contend(const Dev* __restrict__ Dp, MlDev ml, int rows) {
  const Dev& D = *Dp;
  const int k = gtid();
  if (k >= rows || !ml.elig[k]) return;
  const int2 r = D.vAB[k];
  if (r.y < 0) return;
  const int gA = ml.grp[r.x], gB = ml.grp[r.y];
  if (gA == gB) return;
  const unsigned long long key = colorKey1(D.vPerm[k]);
  atomicMax(&ml.winner[gA], key);
  atomicMax(&ml.winner[gB], key);
}
__global__ __launch_bounds__(TPB) void k_ml_match_commit(const Dev* __restrict__ Dp, MlDev ml, int rows) {
  const Dev& D = *Dp;
  const int k = gtid();
  if (k >= rows || !ml.elig[k]) return;
  const int2 r = D.vAB[k];
  if (r.y < 0) return;
  const int gA = ml.grp[r.x], gB = ml.grp[r.y];
  if (gA == gB) return;
  const unsigned long long key = colorKey1(D.vPerm[k]);
  if (ml.winner[gA] == key && ml.winner[gB] == key) {
    ml.mate[gA] = gB;
    ml.mate[gB] = gA;
  }
}
// ml_compact, :209-232: leaders are the smaller id of a matched pair and the singletons
__global__ void k_ml_lead(MlDev ml, int ng) {
  const int g = gtid();
  if (g < ng) ml.lead[g] = ml.mate[g] >= g ? 1 : 0;
}
__global__ void k_ml_parent(MlDev ml, int ng, int parentOff, int groupOff) {
  const int g = gtid();
  if (g >= ng) return;
  const int mt = ml.mate[g];
  if (mt >= g) {
    const int id = ml.scan[g];
    ml.parent[parentOff + g] = id;
    ml.child[groupOff + id] = make_int2(g, mt != g ? mt : -1);
  } else {
    ml.parent[parentOff + g] = ml.scan[mt];
  }
}
// ml_grp_compose, :236-242 (and ml_cycle_compose, :376-381)
__global__ void k_ml_compose(MlDev ml, int n, int parentOff) {
  const int i = gtid();
  if (i < n) ml.grp[i] = ml.parent[parentOff + ml.grp[i]];
}
// group masses, :243-256, summed over the tree
__global__ __launch_bounds__(TPB) void k_ml_mass(const Dev* __restrict__ Dp, MlDev ml, int ng, int groupOff,
                                                 int prevOff) {
  const Dev& D = *Dp;
  const int g = gtid();
  if (g >= ng) return;
  const int2 c = ml.child[groupOff + g];
  double m = prevOff < 0 ? mlEffMass(D.bc[c.x].iim.w) : ml.massG[prevOff + c.x];
  if (c.y >= 0) m += prevOff < 0 ? mlEffMass(D.bc[c.y].iim.w) : ml.massG[prevOff + c.y];
  ml.massG[groupOff + g] = m;
  ml.invMassG[groupOff + g] = 1.0 / m;
}
// ml_color_contend / ml_color_commit, :266-312: one round of the crossing-row colouring; the
// losers of the round are counted into *rem
__global__ __launch_bounds__(TPB) void k_ml_col_contend(const Dev* __restrict__ Dp, MlDev ml, int rows) {
  const Dev& D = *Dp;
  const int k = gtid();
  if (k >= rows || !ml.elig[k] || ml.slot[k] != ML_SKIP) return;
  const int2 r = D.vAB[k];
  const int gA = ml.grp[r.x];
  const int gB = r.y >= 0 ? ml.grp[r.y] : -1;
  if (gB == gA) return;
  const unsigned long long key = colorKey1(D.vPerm[k]);
  atomicMax(&ml.winner[gA], key);
  if (gB >= 0) atomicMax(&ml.winner[gB], key);
}
__global__ __launch_bounds__(TPB) void k_ml_col_commit(const Dev* __restrict__ Dp, MlDev ml, int rows,
                                                       int* rem) {
  const Dev& D = *Dp;
  __shared__ int sh[WAVES];
  const int k = gtid();
  int lost = 0;
  if (k < rows && ml.elig[k] && ml.slot[k] == ML_SKIP) {
    const int2 r = D.vAB[k];
    const int gA = ml.grp[r.x];
    const int gB = r.y >= 0 ? ml.grp[r.y] : -1;
    if (gB != gA) {
      const unsigned long long key = colorKey1(D.vPerm[k]);
      if (ml.winner[gA] != key || (gB >= 0 && ml.winner[gB] != key)) {
        lost = 1;
      } else {
        // the only winner at both groups this round: nothing else writes their masks
        unsigned long long forbidden = ml.mask[gA];
        if (gB >= 0) forbidden |= ml.mask[gB];
        int c = 0;
        while (c < ML_SKIP - 1 && (forbidden & (1ull << c))) ++c;
        ml.slot[k] = (unsigned char)c;
        ml.mask[gA] |= 1ull << c;
        if (gB >= 0) ml.mask[gB] |= 1ull << c;
      }
    }
  }
  blockAtomicAddI(rem, lost, sh);
}
// the level's colour per manifold, for the colour-class placement of its rows
__global__ __launch_bounds__(TPB) void k_ml_col_of(const Dev* __restrict__ Dp, MlDev ml, int nm) {
  const Dev& D = *Dp;
  const int m = gtid();
  if (m >= nm) return;
  int c = -1;
  if (D.mColor[m] >= 0) {
    const int k = D.vRowOf[m];
    if (k >= 0 && ml.slot[k] != ML_SKIP) c = ml.slot[k];
  }
  ml.col[m] = c;
}
// restrict, src/solver_multilevel.hpp:382-405: body momenta, then up the tree
__global__ __launch_bounds__(TPB) void k_ml_mom0(const Dev* __restrict__ Dp, MlDev ml, int n) {
  const Dev& D = *Dp;
  const int i = gtid();
  if (i >= n) return;
  const double m = mlEffMass(D.bc[i].iim.w);
  const double4 v = D.vw[i].v;
  ml.mom[i] = make_double4(m * v.x, m * v.y, m * v.z, 0.0);
}
__global__ void k_ml_tree(MlDev ml, int ng, int groupOff, int belowOff, int n) {
  const int g = gtid();
  if (g >= ng) return;
  const int2 c = ml.child[groupOff + g];
  double4 p = ml.mom[belowOff + c.x];
  if (c.y >= 0) {
    const double4 q = ml.mom[belowOff + c.y];
    p.x += q.x;
    p.y += q.y;
    p.z += q.z;
  }
  ml.mom[n + groupOff + g] = p;
}
__global__ void k_ml_restrict(MlDev ml, int ng, int groupOff, int n) {
  const int g = gtid();
  if (g >= ng) return;
  const double invM = 1.0 / ml.massG[groupOff + g];
  const double4 p = ml.mom[n + groupOff + g];
  const double4 v = make_double4(p.x * invM, p.y * invM, p.z * invM, 0.0);
  ml.velG[groupOff + g] = v;
  ml.velG0[groupOff + g] = v;
}
// ml_coarse_pgs body, :405-466: translation only, e = 0, on the shared impulse accumulator
__device__ void mlCoarseRow(const Dev& D, const MlDev& ml, int off, int k,
                            unsigned long long* qsBits) {
  const int2 r = D.vAB[k];
  const int gA = ml.grp[r.x];
  const int gB = r.y >= 0 ? ml.grp[r.y] : -1;
  const d3 Nsum = xyz(D.vN[k]);
  const double lenN = rowLenN(Nsum);
  if (lenN < 1e-9) return;
  const double invN = (D.vFlag[k] & F_ONE) ? 1.0 : D.vTB[k].w;
  const double sgn = D.vRB[k].w;  // the sign of Nsum . (rA - rB), as in the fine row
  const double4 vA4 = ml.velG[off + gA];
  const d3 vA = xyz(vA4);
  const d3 vB = gB >= 0 ? xyz(ml.velG[off + gB]) : scale3(xyz(D.vWV[k]), invN);
  const double vn = dot3(sub3(vA, vB), Nsum);
  const double invMA = ml.invMassG[off + gA];
  const double invMB = gB >= 0 ? ml.invMassG[off + gB] : 0.0;
  const double w = dot3(Nsum, Nsum) * (invMA + invMB);
  if (w <= 0.0) return;
  const double vtil = sgn * vn;
  const double dp = vtil / w;
  const double pOld = D.vLam[k];
  double pNew = pOld + dp;
  if (pNew < 0.0) pNew = 0.0;
  const double d = pNew - pOld;
  if (d == 0.0) return;
  D.vLam[k] = pNew;
  atomicMaxD(qsBits, fabs(d) * w / lenN);
  const double lambda = -sgn * d;
  const d3 J = scale3(Nsum, lambda);
  double4 a = vA4;
  a.x += J.x * invMA;
  a.y += J.y * invMA;
  a.z += J.z * invMA;
  ml.velG[off + gA] = a;
  if (gB >= 0) {
    double4 b = ml.velG[off + gB];
    b.x -= J.x * invMB;
    b.y -= J.y * invMB;
    b.z -= J.z * invMB;
    ml.velG[off + gB] = b;
  }
}
// one coarse colour class (below 62: the rows touch distinct groups)
__global__ __launch_bounds__(TPB) void k_ml_coarse(const Dev* __restrict__ Dp, MlDev ml, const int* perm,
                                                   int b, int e, int off,
                                                   unsigned long long* qsBits) {
  const Dev& D = *Dp;
  const int q = b + gtid();
  if (q >= e) return;
  mlCoarseRow(D, ml, off, D.vRowOf[perm[q]], qsBits);
}
__global__ __launch_bounds__(64) void k_ml_coarse_serial(const Dev* __restrict__ Dp, MlDev ml, const int* perm,
                                                         int b, int e, int off,
                                                         unsigned long long* qsBits) {
  const Dev& D = *Dp;
  if (gtid() != 0) return;
  for (int q = b; q < e; ++q) mlCoarseRow(D, ml, off, D.vRowOf[perm[q]], qsBits);
}
// ml_prolongate, :468-478: every member takes its group's velocity change
__global__ __launch_bounds__(TPB) void k_ml_prolong(const Dev* __restrict__ Dp, MlDev ml, int n, int off) {
  const Dev& D = *Dp;
  const int i = gtid();
  if (i >= n) return;
  const int g = ml.grp[i];
  const double4 a = ml.velG[off + g], a0 = ml.velG0[off + g];
  double4 v = D.vw[i].v;
  v.x += a.x - a0.x;
  v.y += a.y - a0.y;
  v.z += a.z - a0.z;
  D.vw[i].v = v;
}
// the fine smoothing sweep of the multilevel loop: PGS rows that also report the quasi-static
// residual, corrections of rows with |vn0| <= 4 vRest |N| (src/solver_velocity.hpp:721-723,
// 827-831)
__device__ __forceinline__ double mlQs(const Dev& D, int k, double a, double thr) {
  return fabs(D.vVn0[k]) <= 4.0 * thr * rowLenN(xyz(D.vN[k])) ? a : 0.0;
}
__global__ __launch_bounds__(TPB) void k_vsweep_qs(const Dev* __restrict__ Dp, Params P, int rowBegin, int rowEnd,
                                                   double thr, unsigned long long* qsBits) {
  const Dev& D = *Dp;
  __shared__ unsigned long long sh[WAVES];
  const int k = rowBegin + gtid();
  double a = 0.0, q = 0.0;
  if (k < rowEnd) {
    a = pgsRow(D, P, k, thr);
    q = mlQs(D, k, a, thr);
  }
  blockAtomicMaxU64(D.maxApproachBits, dbits(a), sh);
  __syncthreads();
  blockAtomicMaxU64(qsBits, dbits(q), sh);
}
__global__ __launch_bounds__(64) void k_vsweep_qs_serial(const Dev* __restrict__ Dp, Params P, int rowBegin,
                                                         int rowEnd, double thr,
                                                         unsigned long long* qsBits) {
  const Dev& D = *Dp;
  if (gtid() != 0) return;
  for (int k = rowBegin; k < rowEnd; ++k) {
    const double a = pgsRow(D, P, k, thr);
    atomicMaxD(D.maxApproachBits, a);
    atomicMaxD(qsBits, mlQs(D, k, a, thr));
  }
}

// ---- small problems: a whole adaptive sweep loop in one workgroup ---------------------------
// Rows of one colour are body-disjoint, so the colours run in order separated by workgroup
// barriers and each sweep's residual is a block reduction; the stop rule is the loop's own.  The
// result is the same as one launch per colour, without ~(colours x sweeps) launches.
constexpr int BLOCK_SOLVER_THREADS = 1024;
// the velocity rows need ~160 VGPRs: 768 threads (3 waves per SIMD, <= 168 VGPRs) keep them out
// of scratch
constexpr int VBLOCK_THREADS = 768;
constexpr int BLOCK_SOLVER_ROWS = 16384;
struct ClassStarts {
  int s[65];
};
__device__ __forceinline__ unsigned long long blockMaxAll(unsigned long long v,
                                                          unsigned long long* sh) {
  v = waveMaxU64(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  unsigned long long m = 0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) m = sh[w] > m ? sh[w] : m;
  return m;
}
// colour 62 of the manifold colouring in waves (see k_psweep_waves): mode 0 warm start, 1 PGS,
// 2 one-shot GS; PK: bodies in the packed records vb
template <bool PK>
__global__ __launch_bounds__(VBLOCK_THREADS) void k_vsweep_waves(
    const Dev* __restrict__ Dp, Params P, const int* __restrict__ perm,
    const int* __restrict__ waveStart, int numWaves, int mode, double thr, const int* stop,
    VBody* vb) {
  const Dev& D = *Dp;
  __shared__ unsigned long long sh[VBLOCK_THREADS / 64];
  if (*stop) return;
  unsigned long long m = 0;
  for (int w = 0; w < numWaves; ++w) {
    const int b = waveStart[w], e = waveStart[w + 1];
    for (int i = b + (int)threadIdx.x; i < e; i += blockDim.x) {
      const int k = perm[i];
      double a = 0.0;
      if (mode == 0) RowOf<PK>::warm(D, k, vb);
      else if (mode == 1) a = RowOf<PK>::pgs(D, P, k, thr, vb);
      else a = RowOf<PK>::gs(D, P, k, thr, vb);
      const unsigned long long ab = dbits(a);
      m = ab > m ? ab : m;
    }
    __syncthreads();
  }
  const unsigned long long M = blockMaxAll(m, sh);
  if (mode != 0 && threadIdx.x == 0 && M) atomicMax(D.maxApproachBits, M);
}

// The trailing small colour classes [cFrom, numColors) of a per-colour sweep (none of them the
// last class, 62) in one workgroup with a barrier between classes: one launch instead of one
// per small class
template <bool PK>
__global__ __launch_bounds__(VBLOCK_THREADS) void k_vsweep_tail(
    const Dev* __restrict__ Dp, Params P, ClassStarts cs, int cFrom, int numColors, int mode,
    double thr, const int* stop, VBody* vb) {
  const Dev& D = *Dp;
  __shared__ unsigned long long sh[VBLOCK_THREADS / 64];
  if (*stop) return;
  unsigned long long m = 0;
  for (int c = cFrom; c < numColors; ++c) {
    const int b = cs.s[c], e = cs.s[c + 1];
    for (int k = b + (int)threadIdx.x; k < e; k += blockDim.x) {
      double a = 0.0;
      if (mode == 0) RowOf<PK>::warm(D, k, vb);
      else if (mode == 1) a = RowOf<PK>::pgs(D, P, k, thr, vb);
      else a = RowOf<PK>::gs(D, P, k, thr, vb);
      const unsigned long long ab = dbits(a);
      m = ab > m ? ab : m;
    }
    __syncthreads();
  }
  const unsigned long long M = blockMaxAll(m, sh);
  if (mode != 0 && threadIdx.x == 0 && M) atomicMax(D.maxApproachBits, M);
}

// The sweep loop proper (MODE 1 = PGS, 2 = one-shot GS) over rows whose bodies live in `bs`;
// returns after the sweep that meets the stop rule or after `iters` sweeps.
template <int MODE, class BS>
__device__ __forceinline__ void vSweepLoop(const Dev& D, const Params& P, const ClassStarts& cs,
                                           int numColors, int numManifolds, int iters, double thr,
                                           double stopThr, int* stop, int* count, const BS& bs,
                                           unsigned long long* sh) {
  for (int it = 0; it < iters; ++it) {
    unsigned long long m = 0;
    for (int c = 0; c < numColors; ++c) {
      const int b = cs.s[c], e = cs.s[c + 1];
      if (e <= b) continue;
      if (c == 62) {
        if (threadIdx.x == 0)
          for (int k = b; k < e; ++k) {
            int2 r, sl;
            bs.row(D, k, r, sl);
            const unsigned long long a = dbits(MODE == 1 ? pgsRowB(D, P, k, r, sl, thr, bs)
                                                         : gsRowB(D, P, k, r, sl, thr, bs));
            m = a > m ? a : m;
          }
      } else {
        for (int k = b + (int)threadIdx.x; k < e; k += blockDim.x) {
          int2 r, sl;
          bs.row(D, k, r, sl);
          const unsigned long long a = dbits(MODE == 1 ? pgsRowB(D, P, k, r, sl, thr, bs)
                                                       : gsRowB(D, P, k, r, sl, thr, bs));
          m = a > m ? a : m;
        }
      }
      __syncthreads();
    }
    const unsigned long long M = blockMaxAll(m, sh);
    if (threadIdx.x == 0) {
      *D.maxApproachBits = M;
      *count += 1;
    }
    if (bitsd(M) <= stopThr) {
      if (threadIdx.x == 0) *stop = 1;
      return;
    }
  }
}
// MODE 0 = warm start (one pass), 1 = PGS, 2 = one-shot GS.  `run` and `need` (optional) gate
// the whole loop; `stop` / `count` / `bits` are the loop's stop flag, sweep counter and residual.
template <int MODE>
__global__ __launch_bounds__(VBLOCK_THREADS) void k_vsolve_block(
    const Dev* __restrict__ Dp, Params P, ClassStarts cs, int numColors, int numManifolds, int iters, double thr,
    double stopThr, int* stop, int* count, const int* run, const int* need) {
  const Dev& D = *Dp;
  __shared__ unsigned long long sh[VBLOCK_THREADS / 64];
  if ((run && !*run) || (need && !*need) || *stop) return;
  if (MODE == 0) {
    for (int c = 0; c < numColors; ++c) {
      const int b = cs.s[c], e = cs.s[c + 1];
      if (e <= b) continue;
      if (c == 62) {
        if (threadIdx.x == 0)
          for (int k = b; k < e; ++k) warmApplyRow(D, k);
      } else {
        for (int k = b + (int)threadIdx.x; k < e; k += blockDim.x) warmApplyRow(D, k);
      }
      __syncthreads();
    }
    return;
  }
  vSweepLoop<MODE>(D, P, cs, numColors, numManifolds, iters, thr, stopThr, stop, count,
                   GBodies{D.vw}, sh);
}
// The velocity loops with their bodies staged in LDS: a colour pass reads and writes the velocities
// of up to bodyCap (<= VL_BODIES) of the rows' bodies in LDS, the rest in global memory (the
// row constants stay in global memory, read-only in the loop), and the staged velocities go back
// to global memory once.  Staging hashes the rows' distinct bodies into LDS (the hash shares the
// velocity area, which is filled after it); more than VL_ROWS rows or keyCap distinct bodies
// change nothing and set *fallback = 1 (then k_vsolve_block, gated on it, runs the loop);
// otherwise *fallback = 0.
constexpr int VL_ROWS = 8192;
constexpr int VL_BODIES = 2400;
constexpr int VL_HASH = 16384;
static_assert(VL_HASH * (sizeof(int) + sizeof(unsigned short)) <= 3 * VL_BODIES * sizeof(double2),
              "the staging hash fits in the velocity area");
static_assert(VL_BODIES < VL_GLOBAL, "LDS body slots");
constexpr int VL_KEYS = VL_HASH * 3 / 4;  // distinct bodies the hash takes at most
// open-addressing slot of `body` in an LDS hash of VL_HASH keys (-1: full or, on lookup, absent)
__device__ __forceinline__ int vlSlot(int* hKey, int* used, int body, bool insert, int keyCap) {
  unsigned h = ((unsigned)body * 2654435761u) & (VL_HASH - 1);
  for (int probe = 0; probe < VL_HASH; ++probe) {
    const int k = hKey[h];
    if (k == body) return (int)h;
    if (k == -1) {
      if (!insert) return -1;
      if (atomicAdd(used, 1) >= keyCap) return -1;
      const int old = atomicCAS(&hKey[h], -1, body);
      if (old == -1 || old == body) return (int)h;
      atomicSub(used, 1);
      continue;  // lost the slot to another body: look at it again
    }
    h = (h + 1) & (VL_HASH - 1);
  }
  return -1;
}
// The velocity stage one launch runs on the staged bodies: optionally the warm-start pass, the
// sweep loop, and optionally the stabilization gate (src/sim.hpp:279-283) followed, for
// stabilization mode 1, by the side flags and the stabilization loop.
struct VStage {
  int warm;  // 1: warm-start pass first
  int iters;
  double thr, stopThr;
  int *stop, *count;
  int gate;      // 1: stabilization gate after the loop (writes stabRun and stabStop)
  int stabMode;  // the gate's mode; mode 1 also runs the side flags and the loop
  int stabIters;
  double stabThr;
  int *stabRun, *stabStop, *stabCount;
  d3 gHat;
  double riseThr, approachThr;
};
template <int MODE>
__global__ __launch_bounds__(VBLOCK_THREADS) void k_vsolve_lds(
    const Dev* __restrict__ Dp, Params P, ClassStarts cs, int numColors, int numManifolds,
    VStage S, int* fallback, int bodyCap, int keyCap) {
  const Dev& D = *Dp;

  __shared__ unsigned long long sh[VBLOCK_THREADS / 64];
  __shared__ double2 bv[3 * VL_BODIES];
  __shared__ int bodyOf[VL_BODIES];
  __shared__ unsigned slots[VL_ROWS];
  __shared__ int sUsed, sFail, sDense;
  int* hKey = reinterpret_cast<int*>(bv);
  unsigned short* hIdx = reinterpret_cast<unsigned short*>(hKey + VL_HASH);
  if (*S.stop) return;  // (zero at the start of a substep)
  const int T = blockDim.x, tid = threadIdx.x;
  const int r0 = cs.s[0], nRows = cs.s[numColors] - r0;
  if (nRows > VL_ROWS) {
    if (tid == 0) *fallback = 1;
    return;
  }
  for (int h = tid; h < VL_HASH; h += T) hKey[h] = -1;
  if (tid == 0) {
    sUsed = 0;
    sFail = 0;
    sDense = 0;
  }
  __syncthreads();
  for (int q = tid; q < nRows; q += T) {
    const int2 r = D.vAB[r0 + q];
    if (vlSlot(hKey, &sUsed, r.x, true, keyCap) < 0) sFail = 1;
    if (r.y >= 0 && vlSlot(hKey, &sUsed, r.y, true, keyCap) < 0) sFail = 1;
  }
  __syncthreads();
  const bool staged = !sFail;
  if (tid == 0) *fallback = staged ? 0 : 1;
  if (!staged) return;
  // dense LDS slots for the first bodyCap bodies met in hash order, the rest stay global
  for (int h = tid; h < VL_HASH; h += T) {
    const int body = hKey[h];
    if (body < 0) continue;
    const int s = atomicAdd(&sDense, 1);
    if (s < bodyCap) {
      hIdx[h] = (unsigned short)s;
      bodyOf[s] = body;
    } else {
      hIdx[h] = (unsigned short)VL_GLOBAL;
    }
  }
  __syncthreads();
  for (int q = tid; q < nRows; q += T) {
    const int2 r = D.vAB[r0 + q];
    const unsigned sa = hIdx[vlSlot(hKey, &sUsed, r.x, false, keyCap)];
    const unsigned sb = r.y >= 0 ? hIdx[vlSlot(hKey, &sUsed, r.y, false, keyCap)] : VL_NONE;
    slots[q] = sa | (sb << 16);
  }
  __syncthreads();  // the hash is dead: its area now takes the velocities
  const int nb = sDense < bodyCap ? sDense : bodyCap;
  for (int s = tid; s < nb; s += T) {
    const Body64 g = D.vw[bodyOf[s]];
    bv[3 * s] = make_double2(g.v.x, g.v.y);
    bv[3 * s + 1] = make_double2(g.v.z, g.w.x);
    bv[3 * s + 2] = make_double2(g.w.y, g.w.z);
  }
  __syncthreads();
  const LBodies bs{bv, slots, bodyOf, D.vw, r0};
  if (S.warm) {
    for (int c = 0; c < numColors; ++c) {
      const int b = cs.s[c], e = cs.s[c + 1];
      if (e <= b) continue;
      if (c == 62) {
        if (tid == 0)
          for (int k = b; k < e; ++k) {
            int2 r, sl;
            bs.row(D, k, r, sl);
            warmApplyRowB(D, k, r, sl, bs);
          }
      } else {
        for (int k = b + tid; k < e; k += T) {
          int2 r, sl;
          bs.row(D, k, r, sl);
          warmApplyRowB(D, k, r, sl, bs);
        }
      }
      __syncthreads();
    }
  }
  vSweepLoop<MODE>(D, P, cs, numColors, numManifolds, S.iters, S.thr, S.stopThr, S.stop, S.count,
                   bs, sh);
  if (S.gate) {
    __syncthreads();  // the loop's last residual is in *D.maxApproachBits
    const int runS = (S.stabMode != 0 && bitsd(*D.maxApproachBits) > S.stabThr) ? 1 : 0;
    if (tid == 0) {
      *S.stabRun = runS;
      *S.stabStop = runS ? 0 : 1;
    }
    if (runS && S.stabMode == 1) {
      for (int k = r0 + tid; k < r0 + nRows; k += T)
        sideFlagsRow(D, k, S.gHat, S.riseThr, S.approachThr, bs);
      __syncthreads();
      vSweepLoop<1>(D, P, cs, numColors, numManifolds, S.stabIters, S.stabThr, S.stabThr,
                    S.stabStop, S.stabCount, bs, sh);
    }
  }
  __syncthreads();
  for (int s = tid; s < nb; s += T) {
    PD_G Body64* g = D.vw + bodyOf[s];
    const double2 a = bv[3 * s], b = bv[3 * s + 1], c = bv[3 * s + 2];
    g->v.x = a.x;
    g->v.y = a.y;
    g->v.z = b.x;
    g->w.x = b.y;
    g->w.y = c.x;
    g->w.z = c.y;
  }
}
// ---- stabilization mode 4 ("ordered"), src/sim.hpp:356-397 ---------------------------------
constexpr int LEVEL_INF = 1 << 28;
// one relaxation sweep of the height-from-floor levels (src/contact_preprocessing.hpp:255-297);
// repeated until nothing changes, it reaches the exact support distance in any visiting order
__global__ __launch_bounds__(TPB) void k_height_sweep(const Dev* __restrict__ Dp, int numManifolds, d3 gHat,
                                                      int* h, int* changed) {
  const Dev& D = *Dp;
  __shared__ int sh[WAVES];
  const int idx = gtid();
  int ch = 0;
  if (idx < numManifolds && D.mNp[idx] > 0) {
    const int2 ab = D.mAB[idx];
    const int2 r = D.mReal[idx];
    if (ab.y < 0) {
      ch = atomicMin(&h[r.x], 0) > 0;
    } else {
      const d3 dx = sub3(xyz(D.posPred[ab.x]), xyz(D.posPred[ab.y]));
      const double up = -(dx.x * gHat.x + dx.y * gHat.y + dx.z * gHat.z);
      const double thr = 0.3 * sqrt(dot3(dx, dx));
      if (up > thr) {
        const int cand = atomicAdd(&h[r.y], 0);
        if (cand < LEVEL_INF) ch = atomicMin(&h[r.x], cand + 1) > cand + 1;
      } else if (up < -thr) {
        const int cand = atomicAdd(&h[r.x], 0);
        if (cand < LEVEL_INF) ch = atomicMin(&h[r.y], cand + 1) > cand + 1;
      }
    }
  }
  blockAtomicAddI(changed, ch, sh);
}
// key = min level of the two bodies (clamped 1023) * 64 + colour, per velocity row
__global__ void k_level_keys(const Dev* __restrict__ Dp, int numRows, const int* h, int* key, int* cnt) {
  const Dev& D = *Dp;
  const int k = gtid();
  if (k >= numRows) return;
  const int2 r = D.vAB[k];
  const int m = D.vPerm[k];
  int lv = h[r.x];
  if (r.y >= 0) {
    const int hb = h[r.y];
    if (hb < lv) lv = hb;
  }
  if (lv > 1023) lv = 1023;
  const int kk = lv * 64 + D.mColor[m];
  key[k] = kk;
  atomicAdd(&cnt[kk], 1);
}
// rows placed bucket by bucket (order inside a bucket is free: one colour is body-disjoint)
__global__ void k_level_place(int numRows, const int* key, const int* start, int* cursor,
                              int* perm) {
  const int k = gtid();
  if (k >= numRows) return;
  const int kk = key[k];
  perm[start[kk] + atomicAdd(&cursor[kk], 1)] = k;
}
// bucket sweep over perm[b, e) (one colour), PGS
__global__ __launch_bounds__(TPB) void k_vsweep_perm(const Dev* __restrict__ Dp, Params P, const int* perm, int b,
                                                     int e, double thr, int serial) {
  const Dev& D = *Dp;
  __shared__ unsigned long long sh[WAVES];
  double a = 0.0;
  if (serial) {
    if (gtid() == 0)
      for (int q = b; q < e; ++q) {
        const double x = pgsRow(D, P, perm[q], thr);
        a = x > a ? x : a;
      }
  } else {
    const int q = b + gtid();
    if (q < e) a = pgsRow(D, P, perm[q], thr);
  }
  blockAtomicMaxU64(D.maxApproachBits, dbits(a), sh);
}
// the whole ordered loop in one workgroup: buckets ascending then descending per iteration;
// bucket = (begin, end, colour) into perm
__global__ __launch_bounds__(VBLOCK_THREADS) void k_ordered_block(
    const Dev* __restrict__ Dp, Params P, const int* perm, const int4* buckets, int numBuckets, int iters,
    double thr, int* count) {
  const Dev& D = *Dp;
  __shared__ unsigned long long sh[VBLOCK_THREADS / 64];
  for (int it = 0; it < iters; ++it) {
    unsigned long long m = 0;
    for (int pass = 0; pass < 2; ++pass)
      for (int q = 0; q < numBuckets; ++q) {
        const int4 bk = buckets[pass == 0 ? q : numBuckets - 1 - q];
        if (bk.z == 62) {  // colour 62 may share bodies: one lane, row order
          if (threadIdx.x == 0)
            for (int k = bk.x; k < bk.y; ++k) {
              const unsigned long long a = dbits(pgsRow(D, P, perm[k], thr));
              m = a > m ? a : m;
            }
        } else {
          for (int k = bk.x + (int)threadIdx.x; k < bk.y; k += blockDim.x) {
            const unsigned long long a = dbits(pgsRow(D, P, perm[k], thr));
            m = a > m ? a : m;
          }
        }
        __syncthreads();
      }
    const unsigned long long M = blockMaxAll(m, sh);
    if (threadIdx.x == 0) {
      *D.maxApproachBits = M;
      *count += 1;
    }
    if (bitsd(M) <= thr) return;
  }
}

// commitPairKeysLambdaKokkos, src/contact_preprocessing.hpp:160-192: bucket by the low real index
// (key >> 32), then sort each bucket by (key, manifold index) — a stable key sort
// bucketCnt is all zero on entry (k_ledger_scatter clears what this pass counts).  Each live key
// takes its place in its bucket from the count's atomic; the dead keys (~0) share one region,
// their places handed out per block (one atomic per block, then the block's own order).
__global__ __launch_bounds__(TPB) void k_ledger_count(const Dev* __restrict__ Dp, int numManifolds,
                                                      int numReal, int* rank) {
  const Dev& D = *Dp;
  __shared__ int shWave[WAVES], shBase;
  const int2 ch = chunkOf(numManifolds);
  // live entries: rank within their bucket; dead ones (key ~0) counted for the chunk
  int dead = 0;
  for (int m = ch.x + (int)threadIdx.x; m < ch.y; m += blockDim.x) {
    const unsigned long long key = D.mKey[m];
    if (key == ~0ull) ++dead;
    else rank[m] = atomicAdd(&D.bucketCnt[(int)(key >> 32)], 1);
  }
  const int tot = blockSumI(dead, shWave);
  if (threadIdx.x == 0) shBase = tot ? atomicAdd(&D.bucketCnt[numReal], tot) : 0;
  __syncthreads();
  int run = shBase;
  for (int m0 = ch.x; m0 < ch.y; m0 += blockDim.x) {  // (block-uniform trip count)
    const int m = m0 + (int)threadIdx.x;
    const bool isDead = m < ch.y && D.mKey[m] == ~0ull;
    int t;
    const int r = blockRank(isDead, shWave, t);
    if (isDead) rank[m] = run + r;
    run += t;
  }
}
__global__ void k_ledger_scatter(const Dev* __restrict__ Dp, int numManifolds, int numReal,
                                 const int* rank) {
  const Dev& D = *Dp;
  const int m = gtid();
  if (m >= numManifolds) return;
  const unsigned long long key = D.mKey[m];
  const int b = key == ~0ull ? numReal : (int)(key >> 32);
  const int pos = D.bucketStart[b] + rank[m];
  D.lKey[pos] = key;
  D.lIdx[pos] = m;
  D.bucketCnt[b] = 0;  // back to all zero for the next commit
}
__global__ void k_ledger_sort(const Dev* __restrict__ Dp, int numReal) {
  const Dev& D = *Dp;
  const int b = gtid();
  if (b >= numReal) return;
  const int s = D.bucketStart[b], e = D.bucketStart[b + 1];
  for (int a = s + 1; a < e; ++a) {
    const unsigned long long k = D.lKey[a];
    const int ix = D.lIdx[a];
    int c = a - 1;
    while (c >= s && (D.lKey[c] > k || (D.lKey[c] == k && D.lIdx[c] > ix))) {
      D.lKey[c + 1] = D.lKey[c];
      D.lIdx[c + 1] = D.lIdx[c];
      --c;
    }
    D.lKey[c + 1] = k;
    D.lIdx[c + 1] = ix;
  }
}
__global__ void k_ledger_values(const Dev* __restrict__ Dp, int numManifolds) {
  const Dev& D = *Dp;
  const int pos = gtid();
  if (pos >= numManifolds) return;
  const int m = D.lIdx[pos];
  const int k = D.vRowOf[m];
  if (k >= 0) {
    D.lLam[pos] = D.vLam[k];
    const double4 lt = D.vLamT[k];
    D.lLamT[pos] = make_double4(lt.x, lt.y, lt.z, 0.0);
  } else {
    D.lLam[pos] = 0.0;
    D.lLamT[pos] = make_double4(0, 0, 0, 0);
  }
}

// ============================================================================================
// position solve — src/solver_position.hpp:248-331
// ============================================================================================
__global__ void k_jp_init_contacts(int* colour, int n, int* work) {
  const int c = gtid();
  if (c >= n) return;
  colour[c] = -1;
  work[c] = c;
}
// Position rows (colour-sorted).  The lever arms and normal rotated into the predicted frame and
// the effective inverse mass depend only on quatPred / quat / invMass / invI, which the position
// sweeps never change, so they are evaluated once here (same expressions as per sweep in
// src/solver_position.hpp:276-312).  Row: pN = (n, wTotal), pRA = (rA, invMassA),
// pRB = (rB or plane point, invMassB).
// zeroBits (optional): the position loop's residual, cleared before the loop's first sweep
// One thread per contact, in contact order (coalesced reads), written at the contact's row.
__global__ __launch_bounds__(TPB) void k_prows(const Dev* __restrict__ Dp, int numContacts,
                                               unsigned long long* zeroBits = nullptr) {
  const Dev& D = *Dp;
  const int c = gtid();
  if (zeroBits && c == 0) *zeroBits = 0ull;
  if (c >= numContacts || D.cColor[c] < 0) return;
  const int k = D.pRowOf[c];
  const int2 ab = D.cAB[c];
  const int idA = ab.x, idB = ab.y;
  const double4 n4 = D.cN[c], ra4 = D.cRA[c], rb4 = D.cRB[c];
  const BodyC ca = D.bc[idA];
  const q4 qA = toq(D.quatPred[idA]);
  const q4 qAdelta = quatMult(qA, quatInverse(toq(ca.q)));
  const d3 rA = rotate(qAdelta, xyz(ra4));
  d3 rB = xyz(rb4);
  d3 n = xyz(n4);
  double invMassB = 0.0;
  d3 invIB{0, 0, 0};
  if (idB >= 0) {
    const BodyC cb = D.bc[idB];
    const q4 qB = toq(D.quatPred[idB]);
    const q4 qBdelta = quatMult(qB, quatInverse(toq(cb.q)));
    rB = rotate(qBdelta, rB);
    n = rotate(qBdelta, n);
    invMassB = cb.iim.w;
    invIB = xyz(cb.iim);
  } else {
    n = xyz(n4);
  }
  const double invMassA = ca.iim.w;
  const double wTotal = computeW(rA, n, invMassA, xyz(ca.iim)) + computeW(rB, n, invMassB, invIB);
  D.pAB[k] = ab;
  D.pN[k] = tod4(n, wTotal);
  D.pRA[k] = tod4(rA, invMassA);
  D.pRB[k] = tod4(rB, invMassB);
  D.pLam[k] = 0.0;
}
// solvePositionColoredGSKokkos, one contact (src/solver_position.hpp:259-326)
__device__ double posRow(const Dev& D, int k) {
  const int2 ab = D.pAB[k];
  const int idA = ab.x, idB = ab.y;
  const double4 n4 = D.pN[k], ra4 = D.pRA[k], rb4 = D.pRB[k];
  const d3 n = xyz(n4), rA = xyz(ra4), rB = xyz(rb4);
  const double4 pa4 = D.posPred[idA];
  const d3 pA = xyz(pa4);
  double C;
  double4 pb4;
  if (idB < 0) {
    C = dot3(sub3(add3(pA, rA), rB), n);
  } else {
    pb4 = D.posPred[idB];
    C = dot3(sub3(add3(pA, rA), add3(xyz(pb4), rB)), n);
  }
  if (C >= 0.0) return 0.0;
  const double wTotal = n4.w;
  if (wTotal < 1e-6) return 0.0;
  const double dLambda = -C / wTotal;
  const double invMassA = ra4.w, invMassB = rb4.w;
  D.pLam[k] += dLambda;
  D.posPred[idA].x = pa4.x + n.x * dLambda * invMassA;
  D.posPred[idA].y = pa4.y + n.y * dLambda * invMassA;
  D.posPred[idA].z = pa4.z + n.z * dLambda * invMassA;
  if (idB >= 0) {
    D.posPred[idB].x = pb4.x + -n.x * dLambda * invMassB;
    D.posPred[idB].y = pb4.y + -n.y * dLambda * invMassB;
    D.posPred[idB].z = pb4.z + -n.z * dLambda * invMassB;
  }
  return -C;
}
__global__ __launch_bounds__(TPB) void k_psweep(const Dev* __restrict__ Dp, int rowBegin, int rowEnd, const int* stop) {
  const Dev& D = *Dp;
  if (*stop) return;
  __shared__ unsigned long long sh[WAVES];
  const int k = rowBegin + xcdTid();
  const double o = (k < rowEnd) ? posRow(D, k) : 0.0;
  blockAtomicMaxU64(D.maxOverlapBits, dbits(o), sh);
}
// colour 62 in waves: a row's wave is one more than the latest earlier row of the class sharing
// a body, so the rows of a wave are body-disjoint and every row sees the same bodies as in the
// serial pass in row order; the waves run in order, separated by workgroup barriers
__global__ __launch_bounds__(BLOCK_SOLVER_THREADS) void k_psweep_waves(
    const Dev* __restrict__ Dp, const int* __restrict__ perm, const int* __restrict__ waveStart,
    int numWaves, const int* stop) {
  const Dev& D = *Dp;
  __shared__ unsigned long long sh[BLOCK_SOLVER_THREADS / 64];
  if (*stop) return;
  unsigned long long m = 0;
  for (int w = 0; w < numWaves; ++w) {
    const int b = waveStart[w], e = waveStart[w + 1];
    for (int i = b + (int)threadIdx.x; i < e; i += blockDim.x) {
      const unsigned long long a = dbits(posRow(D, perm[i]));
      m = a > m ? a : m;
    }
    __syncthreads();
  }
  const unsigned long long M = blockMaxAll(m, sh);
  if (threadIdx.x == 0 && M) atomicMax(D.maxOverlapBits, M);
}
// one wave of colour 62 with many rows (body-disjoint): the rows perm[b, e) over the grid
__global__ __launch_bounds__(TPB) void k_psweep_perm(const Dev* __restrict__ Dp, const int* __restrict__ perm,
                                                     int b, int e, const int* stop) {
  const Dev& D = *Dp;
  if (*stop) return;
  __shared__ unsigned long long sh[WAVES];
  const int i = b + gtid();
  const double o = (i < e) ? posRow(D, perm[i]) : 0.0;
  blockAtomicMaxU64(D.maxOverlapBits, dbits(o), sh);
}
// colour 62: one thread, in row order (the class's rows are placed in contact order)
__global__ void k_psweep_serial(const Dev* __restrict__ Dp, int rowBegin, int rowEnd, const int* stop) {
  const Dev& D = *Dp;
  if (gtid() != 0 || *stop) return;
  for (int k = rowBegin; k < rowEnd; ++k) atomicMaxD(D.maxOverlapBits, posRow(D, k));
}

// position sweeps (XPBD), strict stop: residual < tol
__global__ __launch_bounds__(BLOCK_SOLVER_THREADS) void k_psolve_block(
    const Dev* __restrict__ Dp, ClassStarts cs, int numColors, int numContacts, int iters, double tol, int* stop,
    int* count) {
  const Dev& D = *Dp;
  __shared__ unsigned long long sh[BLOCK_SOLVER_THREADS / 64];
  if (*stop) return;
  for (int it = 0; it < iters; ++it) {
    unsigned long long m = 0;
    for (int c = 0; c < numColors; ++c) {
      const int b = cs.s[c], e = cs.s[c + 1];
      if (e <= b) continue;
      if (c == 62) {
        if (threadIdx.x == 0)
          for (int k = b; k < e; ++k) {
            const unsigned long long a = dbits(posRow(D, k));
            m = a > m ? a : m;
          }
      } else {
        for (int k = b + (int)threadIdx.x; k < e; k += blockDim.x) {
          const unsigned long long a = dbits(posRow(D, k));
          m = a > m ? a : m;
        }
      }
      __syncthreads();
    }
    const unsigned long long M = blockMaxAll(m, sh);
    if (threadIdx.x == 0) {
      *D.maxOverlapBits = M;
      *count += 1;
    }
    if (bitsd(M) < tol) {
      if (threadIdx.x == 0) *stop = 1;
      return;
    }
  }
}

// The small colours of one position sweep in one workgroup, then the sweep's stop check: the
// colours from cFrom on run in order separated by barriers, the residual is the max of theirs
// and what the launches before them accumulated in *maxOverlapBits; the accumulator is cleared
// for the next sweep unless the loop ends here (stop, or the cap of sweeps reached).
__global__ __launch_bounds__(BLOCK_SOLVER_THREADS) void k_psweep_tail(
    const Dev* __restrict__ Dp, ClassStarts cs, int cFrom, int numColors, int numContacts, double tol, int* stop,
    int* count, int cap) {
  const Dev& D = *Dp;
  __shared__ unsigned long long sh[BLOCK_SOLVER_THREADS / 64];
  if (*stop) return;
  unsigned long long m = 0;
  for (int c = cFrom; c < numColors; ++c) {
    const int b = cs.s[c], e = cs.s[c + 1];
    if (e <= b) continue;
    if (c == 62) {
      if (threadIdx.x == 0)
        for (int k = b; k < e; ++k) {
          const unsigned long long a = dbits(posRow(D, k));
          m = a > m ? a : m;
        }
    } else {
      for (int k = b + (int)threadIdx.x; k < e; k += blockDim.x) {
        const unsigned long long a = dbits(posRow(D, k));
        m = a > m ? a : m;
      }
    }
    __syncthreads();
  }
  unsigned long long M = blockMaxAll(m, sh);
  if (threadIdx.x == 0) {
    const unsigned long long before = *D.maxOverlapBits;
    M = before > M ? before : M;
    const int n = *count + 1;
    *count = n;
    const bool done = bitsd(M) < tol;
    if (done) *stop = 1;
    *D.maxOverlapBits = (done || n >= cap) ? M : 0ull;
  }
}

// commitPosImpulseKokkos, src/contact_preprocessing.hpp:501-538 (sum in contact order; among
// manifolds sharing a key the highest manifold index writes, as a serial loop would)
__global__ void k_commit_pos_impulse(const Dev* __restrict__ Dp, int numManifolds, int numReal, double dt) {
  const Dev& D = *Dp;
  const int m = gtid();
  if (m >= numManifolds) return;
  const int2 rg = D.mRange[m], rg2 = D.mRange2[m];
  double s = 0.0;
  for (int q = 0; q < rg.y + rg2.y; ++q) {
    const int c = q < rg.y ? rg.x + q : rg2.x + (q - rg.y);
    const int row = D.pRowOf[c];  // -1: uncoloured contact, solved by the Jacobi fallback only
    const double l = row >= 0 ? D.pLam[row] : 0.0;
    if (l != 0.0) s += l / dt;
  }
  const unsigned long long key = D.mKey[m];
  if (key == ~0ull) return;
  const int lo = ledgerFind(D.lKey, D.bucketStart, numReal, key);
  if (lo < 0) return;
  int hi = lo + 1;
  while (hi < numManifolds && D.lKey[hi] == key) ++hi;
  if (D.lIdx[hi - 1] == m) D.lLamT[lo].w = s;
}

// ============================================================================================
// colour-saturation fallbacks (colour == -1 after the round cap), src/sim.hpp:245-256,454-462,
// and the Jacobi paths: contributions per element in parallel, summed per body in element order
// (the canonical order of the oracle); the averaged apply is one body per thread.
// ============================================================================================
// applyVelocityDeltasAveragedKokkos, src/solver_velocity.hpp:233-255
__global__ void k_apply_vel_avg(const Dev* __restrict__ Dp, int n) {
  const Dev& D = *Dp;
  const int i = gtid();
  if (i >= n) return;
  const int count = D.dCnt[i];
  if (count <= 0) return;
  const double f = fmin(1.0, 2.0 / (double)count);
  Body64 b = D.vw[i];
  const double4 dv = D.dVel[i], dw = D.dAng[i];
  b.v.x += dv.x * f;
  b.w.x += dw.x * f;
  b.v.y += dv.y * f;
  b.w.y += dw.y * f;
  b.v.z += dv.z * f;
  b.w.z += dw.z * f;
  D.vw[i] = b;
  D.dVel[i] = make_double4(0, 0, 0, 0);
  D.dAng[i] = make_double4(0, 0, 0, 0);
  D.dCnt[i] = 0;
}
// ---- the same three accumulations in parallel, in the same order --------------------------
// The serial loops above add each element's contribution (manifold or contact, ascending) to
// its two bodies' accumulators, and no contribution reads an accumulator.  Here every element's
// contribution is computed by its own thread (the same statements), and each body then sums the
// contributions of its incident elements in ascending element order (A before B for one
// element), which is the serial loop's order of additions into that body: the accumulators end
// bit-identical.  Incidences per body: element * 2 + side (0 = body A, 1 = body B), sorted.
struct JContrib {
  double4 a0, a1, b0, b1;  // A's velocity (or position) and angular parts, then B's
};
// MODE 0: velocity Jacobi over manifolds (bodies real indices); 1: position Jacobi over contacts
// (bodies slots); 2: count-averaged friction over contacts (bodies real indices)
template <int MODE>
__device__ __forceinline__ bool jElementBodies(const Dev& D, int e, int onlyUncoloured, int& a,
                                               int& b) {
  if (MODE == 0) {
    if ((onlyUncoloured && D.mColor[e] != -1) || D.mNp[e] <= 0) return false;
    const int2 r = D.mReal[e];
    if (r.y >= 0 && r.x > r.y) return false;
    a = r.x;
    b = r.y;
    return true;
  }
  if (MODE == 1) {
    if (onlyUncoloured && D.cColor[e] != -1) return false;
    const int2 ab = D.cAB[e];
    a = ab.x;
    b = ab.y;
    return true;
  }
  if (!(D.cLamN[e] > 0.0)) return false;
  const int2 ab = D.cAB[e];
  a = D.realIdx[ab.x];
  b = ab.y >= 0 ? D.realIdx[ab.y] : -1;
  return true;
}
template <int MODE>
__global__ void k_jinc_count(const Dev* __restrict__ Dp, int numElements, int onlyUncoloured,
                             int* cnt) {
  const Dev& D = *Dp;
  const int e = gtid();
  if (e >= numElements) return;
  int a, b;
  if (!jElementBodies<MODE>(D, e, onlyUncoloured, a, b)) return;
  atomicAdd(&cnt[a], 1);
  if (b >= 0) atomicAdd(&cnt[b], 1);
}
template <int MODE>
__global__ void k_jinc_fill(const Dev* __restrict__ Dp, int numElements, int onlyUncoloured,
                            const int* start, int* fill, int* list) {
  const Dev& D = *Dp;
  const int e = gtid();
  if (e >= numElements) return;
  int a, b;
  if (!jElementBodies<MODE>(D, e, onlyUncoloured, a, b)) return;
  list[start[a] + atomicAdd(&fill[a], 1)] = 2 * e;
  if (b >= 0) list[start[b] + atomicAdd(&fill[b], 1)] = 2 * e + 1;
}
// each body's incidences in ascending order (insertion sort of the short segment)
__global__ void k_jinc_sort(int numBodies, const int* start, int* list) {
  const int i = gtid();
  if (i >= numBodies) return;
  const int b = start[i], e = start[i + 1];
  for (int k = b + 1; k < e; ++k) {
    const int v = list[k];
    int j = k - 1;
    while (j >= b && list[j] > v) {
      list[j + 1] = list[j];
      --j;
    }
    list[j + 1] = v;
  }
}
__device__ __forceinline__ double4 d4(double x, double y, double z) {
  return make_double4(x, y, z, 0.0);
}
// solveVelocityKokkos, src/solver_velocity.hpp:38-224: one manifold's contribution
__global__ void k_jvel_contrib(const Dev* __restrict__ Dp, Params P, int numManifolds, double thr,
                               int onlyUncoloured, JContrib* jc, int* jf) {
  const Dev& D = *Dp;
  const int m = gtid();
  if (m >= numManifolds) return;
  jf[m] = 0;
  int realA, realB;
  if (!jElementBodies<0>(D, m, onlyUncoloured, realA, realB)) return;
  const BodyC ca = D.bc[realA];
  const double invMassA = ca.iim.w;
  const d3 invIA = xyz(ca.iim);
  const q4 qA = toq(ca.q);
  double invMassB = 0.0;
  d3 invIB{0, 0, 0};
  q4 qB{0, 0, 0, 1};
  const Body64 ba = D.vw[realA];
  const d3 vA = xyz(ba.v), wA = xyz(ba.w);
  d3 vB{0, 0, 0}, wB{0, 0, 0};
  if (realB >= 0) {
    const BodyC cb = D.bc[realB];
    invMassB = cb.iim.w;
    invIB = xyz(cb.iim);
    qB = toq(cb.q);
    const Body64 bb = D.vw[realB];
    vB = xyz(bb.v);
    wB = xyz(bb.w);
  }
  const double4 ns4 = D.mNs[m];
  const d3 Nsum = xyz(ns4), TauA = xyz(D.mTA[m]), TauB = xyz(D.mTB[m]);
  const double invN = 1.0 / (double)D.mNp[m];
  double restitution = P.restN;
  if (realB < 0) vB = scale3(xyz(D.mWV[m]), invN);
  {
    const double ra = ns4.w * invN;
    if (ra >= 0.0) restitution = ra;
  }
  const d3 rAavg = scale3(xyz(D.mRA[m]), invN);
  const d3 rBavg = scale3(xyz(D.mRB[m]), invN);
  const double lenN = sqrt(dot3(Nsum, Nsum));
  if (lenN < 1e-9) return;
  const d3 diffCenters = (realB < 0) ? rAavg : sub3(rAavg, rBavg);
  const d3 vGrowth = scale3(diffCenters, P.growthRate);
  double vn = dot3(vA, Nsum) + dot3(wA, TauA) + dot3(vB, neg3(Nsum)) + dot3(wB, TauB);
  vn += dot3(vGrowth, Nsum);
  const double alignment = dot3(Nsum, diffCenters);
  if (alignment > 0.0) {
    if (vn < 0.0) return;
  } else {
    if (vn > 0.0) return;
  }
  const double Nsq = dot3(Nsum, Nsum);
  const double wTotal = (Nsq * invMassA + genInvMass(TauA, invIA, qA)) +
                        (Nsq * invMassB + genInvMass(TauB, invIB, qB));
  if (wTotal <= 0.0) return;
  if (fabs(vn) < thr * lenN) restitution = 0.0;
  const double lambda = (-restitution * vn - vn) / wTotal;
  const d3 Jlin = scale3(Nsum, lambda);
  JContrib c;
  c.a0 = d4(Jlin.x * invMassA, Jlin.y * invMassA, Jlin.z * invMassA);
  const d3 dwA = angDelta(qA, scale3(TauA, lambda), invIA);
  c.a1 = d4(dwA.x, dwA.y, dwA.z);
  int f = 1;
  if (realB >= 0) {
    c.b0 = d4(-Jlin.x * invMassB, -Jlin.y * invMassB, -Jlin.z * invMassB);
    const d3 dwB = angDelta(qB, scale3(TauB, lambda), invIB);
    c.b1 = d4(dwB.x, dwB.y, dwB.z);
    f |= 2;
  }
  jc[m] = c;
  jf[m] = f;
}
// solvePositionKokkos, src/solver_position.hpp:34-132: one contact's contribution
__global__ void k_jpos_contrib(const Dev* __restrict__ Dp, int numContacts, int onlyUncoloured,
                               JContrib* jc, int* jf) {
  const Dev& D = *Dp;
  const int c = gtid();
  if (c >= numContacts) return;
  jf[c] = 0;
  int idA, idB;
  if (!jElementBodies<1>(D, c, onlyUncoloured, idA, idB)) return;
  const BodyC ca = D.bc[idA];
  const double invMassA = ca.iim.w;
  const d3 pA = xyz(D.posPred[idA]);
  const q4 qA = toq(D.quatPred[idA]);
  d3 pB{0, 0, 0};
  q4 qB{0, 0, 0, 1};
  BodyC cb;
  double invMassB = 0.0;
  if (idB >= 0) {
    cb = D.bc[idB];
    invMassB = cb.iim.w;
    pB = xyz(D.posPred[idB]);
    qB = toq(D.quatPred[idB]);
  }
  const double4 n4 = D.cN[c], ra4 = D.cRA[c], rb4 = D.cRB[c];
  const q4 qAdelta = quatMult(qA, quatInverse(toq(ca.q)));
  const d3 rA = rotate(qAdelta, xyz(ra4));
  d3 rB = xyz(rb4);
  d3 n = xyz(n4);
  if (idB >= 0) {
    const q4 qBdelta = quatMult(qB, quatInverse(toq(cb.q)));
    rB = rotate(qBdelta, rB);
    n = rotate(qBdelta, n);
  }
  double C;
  if (idB < 0) {
    n = xyz(n4);
    C = dot3(sub3(add3(pA, rA), xyz(rb4)), n);
  } else {
    C = dot3(sub3(add3(pA, rA), add3(pB, rB)), n);
  }
  if (C >= 0.0) return;
  const d3 invIA = xyz(ca.iim);
  const d3 invIB = (idB >= 0) ? xyz(cb.iim) : d3{0, 0, 0};
  const double wTotal = computeW(rA, n, invMassA, invIA) + computeW(rB, n, invMassB, invIB);
  if (wTotal < 1e-6) return;
  const double dLambda = -C / wTotal;
  JContrib q;
  q.a0 = d4(n.x * dLambda * invMassA, n.y * dLambda * invMassA, n.z * dLambda * invMassA);
  int f = 1;
  if (idB >= 0) {
    q.b0 = d4(-n.x * dLambda * invMassB, -n.y * dLambda * invMassB, -n.z * dLambda * invMassB);
    f |= 2;
  }
  jc[c] = q;
  jf[c] = f;
  atomicMaxD(D.maxOverlapBits, -C);
}
// solveContactFrictionKokkos, src/solver_friction.hpp:142-220: one contact's contribution
__global__ void k_jfr_contrib(const Dev* __restrict__ Dp, int nc, double frictionDynamic,
                              JContrib* jc, int* jf) {
  const Dev& D = *Dp;
  const int c = gtid();
  if (c >= nc) return;
  jf[c] = 0;
  int realA, realB;
  if (!jElementBodies<2>(D, c, 0, realA, realB)) return;
  const double lambda_n = D.cLamN[c];
  const int idB = D.cAB[c].y;
  const BodyC ca = D.bc[realA];
  const double invMA = ca.iim.w;
  const d3 invIA = xyz(ca.iim);
  double invMB = 0.0;
  d3 invIB{0.0, 0.0, 0.0};
  const d3 rA = xyz(D.cRA[c]), n = xyz(D.cN[c]);
  const double4 rb4 = D.cRB[c];
  const d3 rB = xyz(rb4);
  const Body64 ba = D.vw[realA];
  const d3 vAc = add3(xyz(ba.v), cross3(xyz(ba.w), rA));
  d3 vBc = xyz(D.cBV[c]);
  if (idB >= 0) {
    const BodyC cb = D.bc[realB];
    invMB = cb.iim.w;
    invIB = xyz(cb.iim);
    const Body64 bb = D.vw[realB];
    vBc = add3(xyz(bb.v), cross3(xyz(bb.w), rB));
  }
  const d3 vrel = sub3(vAc, vBc);
  const double vn = dot3(vrel, n);
  const d3 vt = sub3(vrel, scale3(n, vn));
  const double vt_len = sqrt(dot3(vt, vt));
  if (vt_len < 1e-7) return;
  const d3 t = scale3(vt, 1.0 / vt_len);
  const d3 rnA = cross3(rA, t), rnB = cross3(rB, t);
  const double w_t = invMA + invMB + rnA.x * rnA.x * invIA.x + rnA.y * rnA.y * invIA.y +
                     rnA.z * rnA.z * invIA.z + rnB.x * rnB.x * invIB.x +
                     rnB.y * rnB.y * invIB.y + rnB.z * rnB.z * invIB.z;
  if (w_t < 1e-6) return;
  const double nA = (double)D.frCnt[realA];
  const double nB = (idB >= 0) ? (double)D.frCnt[realB] : 0.0;
  const double inv_n = 1.0 / fmax(fmax(nA, nB), 1.0);
  double lt = -vt_len / w_t;
  const double mu = (rb4.w >= 0.0) ? rb4.w : frictionDynamic;
  const double maxf = mu * lambda_n;
  if (lt < -maxf) lt = -maxf;
  lt *= inv_n;
  JContrib q;
  q.a0 = d4(t.x * lt * invMA, t.y * lt * invMA, t.z * lt * invMA);
  q.a1 = d4(rnA.x * invIA.x * lt, rnA.y * invIA.y * lt, rnA.z * invIA.z * lt);
  int f = 1;
  if (idB >= 0) {
    q.b0 = d4(-t.x * lt * invMB, -t.y * lt * invMB, -t.z * lt * invMB);
    q.b1 = d4(-rnB.x * invIB.x * lt, -rnB.y * invIB.y * lt, -rnB.z * invIB.z * lt);
    f |= 2;
  }
  jc[c] = q;
  jf[c] = f;
}
// each body sums its incident contributions in ascending element order; MODE 0 / 2 add to the
// velocity accumulators (0 also counts), MODE 1 to the position accumulator (and counts)
template <int MODE>
__global__ void k_jgather(const Dev* __restrict__ Dp, int numBodies, const int* start,
                          const int* list, const JContrib* jc, const int* jf) {
  const Dev& D = *Dp;
  const int i = gtid();
  if (i >= numBodies) return;
  const int b = start[i], e = start[i + 1];
  if (b == e) return;
  double4 s0 = MODE == 1 ? D.dPos[i] : D.dVel[i];
  double4 s1 = D.dAng[i];
  int cnt = D.dCnt[i];
  for (int k = b; k < e; ++k) {
    const int code = list[k];
    const int el = code >> 1, side = code & 1;
    if (!(jf[el] & (1 << side))) continue;
    const double4 c0 = side ? jc[el].b0 : jc[el].a0;
    s0.x += c0.x;
    s0.y += c0.y;
    s0.z += c0.z;
    if (MODE != 1) {
      const double4 c1 = side ? jc[el].b1 : jc[el].a1;
      s1.x += c1.x;
      s1.y += c1.y;
      s1.z += c1.z;
    }
    cnt += 1;
  }
  if (MODE == 1) D.dPos[i] = s0;
  else D.dVel[i] = s0;
  if (MODE != 1) D.dAng[i] = s1;
  if (MODE != 2) D.dCnt[i] = cnt;
}

// applyUpdatesKokkos, src/integration.hpp:168-184
__global__ void k_apply_pos_avg(const Dev* __restrict__ Dp, int n) {
  const Dev& D = *Dp;
  const int i = gtid();
  if (i >= n) return;
  const int count = D.dCnt[i];
  if (count <= 0) return;
  const double f = 1.0 / (double)count;
  const double4 dp = D.dPos[i], dv = D.dVel[i];
  D.posPred[i].x = D.posPred[i].x + dp.x * f;
  D.posPred[i].y = D.posPred[i].y + dp.y * f;
  D.posPred[i].z = D.posPred[i].z + dp.z * f;
  D.vw[i].v.x = D.vw[i].v.x + dv.x * f;
  D.vw[i].v.y = D.vw[i].v.y + dv.y * f;
  D.vw[i].v.z = D.vw[i].v.z + dv.z * f;
  D.dPos[i] = make_double4(0, 0, 0, 0);
  D.dVel[i] = make_double4(0, 0, 0, 0);
  D.dCnt[i] = 0;
}

// ============================================================================================
// legacy count-averaged Coulomb friction (zero gravity, or the Jacobi velocity path):
// src/sim.hpp:173-176,224-226,423-429 and src/solver_friction.hpp:32-220
// ============================================================================================
// plane contacts: initial normal impulse from the approach speed (src/sim.hpp:173-176)
__global__ void k_fr_plane_load(const Dev* __restrict__ Dp, int nc) {
  const Dev& D = *Dp;
  const int c = gtid();
  if (c >= nc) return;
  const int2 ab = D.cAB[c];
  double lam = 0.0;
  if (ab.y < 0) {
    const int idA = ab.x;
    const BodyC ca = D.bc[idA];
    const double invMA = ca.iim.w;
    if (invMA > 0.0) {
      const Body64 b = D.vw[idA];
      const d3 n = xyz(D.cN[c]), rA = xyz(D.cRA[c]);
      const d3 vAc = add3(xyz(b.v), cross3(xyz(b.w), rA));
      const double approach = -dot3(sub3(vAc, xyz(D.cBV[c])), n);
      if (approach > 0.0) {
        const double w_n = computeW(rA, n, invMA, xyz(ca.iim));
        lam = (w_n > 1e-6) ? approach / w_n : 0.0;
      }
    }
  }
  D.cLamN[c] = lam;
}
// src/solver_friction.hpp:32-85, once per velocity iteration (each contact its own field)
__global__ void k_fr_accum_normal(const Dev* __restrict__ Dp, int nc, double growthRate) {
  const Dev& D = *Dp;
  const int c = gtid();
  if (c >= nc) return;
  const int2 ab = D.cAB[c];
  const d3 n = xyz(D.cN[c]), rA = xyz(D.cRA[c]);
  if (ab.y < 0) {
    const int idA = ab.x;
    const BodyC ca = D.bc[idA];
    const double invMA = ca.iim.w;
    if (invMA <= 0.0) return;
    const Body64 b = D.vw[idA];
    const d3 vAc = add3(xyz(b.v), cross3(xyz(b.w), rA));
    const double approach = -dot3(sub3(vAc, xyz(D.cBV[c])), n);
    if (approach <= 0.0) return;
    const double w_n = computeW(rA, n, invMA, xyz(ca.iim));
    if (w_n > 1e-6) D.cLamN[c] += approach / w_n;
    return;
  }
  const int realA = D.realIdx[ab.x], realB = D.realIdx[ab.y];
  if (realA > realB) return;
  const d3 rB = xyz(D.cRB[c]);
  const BodyC ca = D.bc[realA], cb = D.bc[realB];
  const Body64 ba = D.vw[realA], bb = D.vw[realB];
  const d3 vAc = add3(xyz(ba.v), cross3(xyz(ba.w), rA));
  const d3 vBc = add3(xyz(bb.v), cross3(xyz(bb.w), rB));
  const d3 vrel = add3(sub3(vAc, vBc), scale3(sub3(rA, rB), growthRate));
  const double approach = -dot3(vrel, n);
  if (approach <= 0.0) return;
  const double w_n = computeW(rA, n, ca.iim.w, xyz(ca.iim)) + computeW(rB, n, cb.iim.w, xyz(cb.iim));
  if (w_n > 1e-6) D.cLamN[c] += approach / w_n;
}
// src/solver_friction.hpp:123-140 (integer counts: atomics are exact)
__global__ void k_fr_count(const Dev* __restrict__ Dp, int nc) {
  const Dev& D = *Dp;
  const int c = gtid();
  if (c >= nc) return;
  if (!(D.cLamN[c] > 0.0)) return;
  const int2 ab = D.cAB[c];
  atomicAdd(&D.frCnt[D.realIdx[ab.x]], 1);
  if (ab.y >= 0) atomicAdd(&D.frCnt[D.realIdx[ab.y]], 1);
}
// src/solver_friction.hpp:142-220 runs as k_jfr_contrib + k_jgather<2> (contact order per
// body), then the impulses are applied unaveraged (src/integration.hpp:97-109)
__global__ void k_fr_apply(const Dev* __restrict__ Dp, int n) {
  const Dev& D = *Dp;
  const int i = gtid();
  if (i >= n) return;
  Body64 b = D.vw[i];
  const double4 dv = D.dVel[i], dw = D.dAng[i];
  b.v.x = b.v.x + dv.x;
  b.v.y = b.v.y + dv.y;
  b.v.z = b.v.z + dv.z;
  b.w.x = b.w.x + dw.x;
  b.w.y = b.w.y + dw.y;
  b.w.z = b.w.z + dw.z;
  D.vw[i] = b;
  D.dVel[i] = make_double4(0, 0, 0, 0);
  D.dAng[i] = make_double4(0, 0, 0, 0);
}

// ============================================================================================
// packed-bed SDF reconstruction, src/output_sdf.hpp:25-197: splat (minimum over particles of the
// exact analytic distance in each particle's band; ordered keys make the minimum order-free),
// then Jacobi fast-iterative Eikonal sweeps, periodic-wrapped
// ============================================================================================
struct SdfGridP {
  int rx, ry, rz, px, py, pz;
  d3 origin, vox;
};
__global__ void k_sdf_splat(const Dev* __restrict__ Dp, Params P, SdfGridP G, unsigned long long* key,
                            unsigned char* fixedCell) {
  const Dev& D = *Dp;
  const int i = gtid();
  if (i >= P.numReal) return;
  const d3 pw = xyz(D.pos[i]);
  const q4 q = toq(D.bc[i].q);
  const double sc = D.scale[i];
  double rbound = 1.0;
  if (P.shape.type == SHAPE_HOLLOW_CYLINDER) {
    const double r = P.shape.p0, h = P.shape.p1;
    rbound = sqrt(r * r + (h * 0.5) * (h * 0.5));
  } else if (P.shape.type == SHAPE_SPHERE || P.shape.type == SHAPE_GRID_SDF) {
    rbound = P.shape.p0;
  }
  rbound *= sc;
  rbound *= 1.2;
  int minx = (int)floor((pw.x - rbound - G.origin.x) / G.vox.x);
  int maxx = (int)ceil((pw.x + rbound - G.origin.x) / G.vox.x);
  int miny = (int)floor((pw.y - rbound - G.origin.y) / G.vox.y);
  int maxy = (int)ceil((pw.y + rbound - G.origin.y) / G.vox.y);
  int minz = (int)floor((pw.z - rbound - G.origin.z) / G.vox.z);
  int maxz = (int)ceil((pw.z + rbound - G.origin.z) / G.vox.z);
  if (!G.px) { minx = minx < 0 ? 0 : minx; maxx = maxx > G.rx - 1 ? G.rx - 1 : maxx; }
  if (!G.py) { miny = miny < 0 ? 0 : miny; maxy = maxy > G.ry - 1 ? G.ry - 1 : maxy; }
  if (!G.pz) { minz = minz < 0 ? 0 : minz; maxz = maxz > G.rz - 1 ? G.rz - 1 : maxz; }
  const bool analytic = P.shape.type == SHAPE_HOLLOW_CYLINDER || P.shape.type == SHAPE_SPHERE ||
                        P.shape.type == SHAPE_GRID_SDF;
  for (int z = minz; z <= maxz; ++z)
    for (int y = miny; y <= maxy; ++y)
      for (int x = minx; x <= maxx; ++x) {
        const d3 vp{G.origin.x + (x + 0.5) * G.vox.x, G.origin.y + (y + 0.5) * G.vox.y,
                    G.origin.z + (z + 0.5) * G.vox.z};
        const d3 plocal = scale3(invRotate(q, sub3(vp, pw)), 1.0 / sc);
        const double distc = analytic ? sdfShape(P.shape, plocal) : 100.0;
        const double dist = distc * sc;
        const int wx = (x % G.rx + G.rx) % G.rx, wy = (y % G.ry + G.ry) % G.ry;
        const int wz = (z % G.rz + G.rz) % G.rz;
        const long vi = (long)wz * G.rx * G.ry + (long)wy * G.rx + wx;
        atomicMin(&key[vi], ordKey(dist));
        if (dist < rbound) fixedCell[vi] = 1;
      }
}
__global__ void k_sdf_from_keys(const unsigned long long* key, double* g, long total) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < total) g[i] = ordVal(key[i]);
}
__global__ void k_sdf_eikonal(SdfGridP G, const double* din, double* dout,
                              const unsigned char* fixedCell, double h) {
  const long total = (long)G.rx * G.ry * G.rz;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= total) return;
  if (fixedCell[idx]) {
    dout[idx] = din[idx];
    return;
  }
  const long sy = G.rx, sz = (long)G.rx * G.ry;
  const int z = (int)(idx / sz);
  const long rem = idx % sz;
  const int y = (int)(rem / sy), x = (int)(rem % sy);
  double xm = 100.0, xp = 100.0, ym = 100.0, yp = 100.0, zm = 100.0, zp = 100.0;
  if (x > 0) xm = din[idx - 1]; else if (G.px) xm = din[idx - 1 + G.rx];
  if (x < G.rx - 1) xp = din[idx + 1]; else if (G.px) xp = din[idx + 1 - G.rx];
  if (y > 0) ym = din[idx - sy]; else if (G.py) ym = din[(long)z * sz + (long)(G.ry - 1) * sy + x];
  if (y < G.ry - 1) yp = din[idx + sy]; else if (G.py) yp = din[(long)z * sz + x];
  if (z > 0) zm = din[idx - sz]; else if (G.pz) zm = din[(long)(G.rz - 1) * sz + (long)y * sy + x];
  if (z < G.rz - 1) zp = din[idx + sz]; else if (G.pz) zp = din[(long)y * sy + x];
  double a = fmin(xm, xp), b = fmin(ym, yp), c = fmin(zm, zp), t;
  if (a > b) { t = a; a = b; b = t; }
  if (b > c) { t = b; b = c; c = t; }
  if (a > b) { t = a; a = b; b = t; }
  double u = 100.0;
  const double u1 = a + h;
  if (u1 <= b) {
    u = u1;
  } else {
    const double d2 = 2.0 * h * h - (a - b) * (a - b);
    const double u2 = (d2 < 0.0) ? 100.0 : (a + b + sqrt(d2)) * 0.5;
    if (u2 <= c) {
      u = u2;
    } else {
      const double s3 = a + b + c, qd = s3 * s3 - 3.0 * (a * a + b * b + c * c - h * h);
      u = (qd < 0.0) ? 100.0 : (s3 + sqrt(qd)) / 3.0;
    }
  }
  dout[idx] = u;
}

}  // namespace pd
