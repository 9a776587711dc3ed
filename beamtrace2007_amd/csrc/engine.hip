// Wavefront beam tracing on gfx950 (MI355X): sub-beam records in HBM queues,
// processed round by round by persistent wave64 kernels.
//
// One round = three launches on the context's stream:
//   admitKernel   one lane per root (a pixel in tiles of 64 for soft shadows,
//                 a tile for primary visibility, a primary hit polygon for
//                 point shadows): per-root setup, direction-sign pre-split,
//                 records appended to the trace queue;
//   traceKernel   one lane per sub-beam, refilled from the queue as soon as its
//                 sub-beam finishes or needs a clip (64-bit ballot + mbcnt
//                 rank, one atomic per wave): kd descent, leaf triangle tests,
//                 restart-trail climb;
//   splitKernel   one lane per clip: the clip runs twice, first counting the
//                 pieces (so the wave reserves exactly the queue slots it
//                 needs with one atomic), then writing them in place - no
//                 per-lane piece buffer.
// The last workgroup of splitKernel to finish schedules the next round
// (admission by reserved budget, as the reference's scheduleKernel,
// cuda/wavefront.cu:327-345), so no single-thread launch is needed.
//
// Results are exact and order-independent: soft-shadow visibility is summed
// as integer multiples of the smallest representable piece area (every
// output area is >= minArea, so each is an exact multiple of ulp(minArea)),
// and the Lambert integral in 2^-60 fixed point. Roots that exceed their
// sub-beam budget are traced again on the GPU with a larger budget (no CPU
// fallback).
//
// The per-sub-beam arithmetic restates src/beam/beam_core.h (BeamCtx) and
// src/beam/wavefront_core.h of the reference; the per-root setups restate
// src/render/soft_core.h and point_core.h; the primary rays restate
// src/ray/ray_core.h. Function comments cite the lines.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <mutex>

#include "engine.h"

// Scan scratch of the allocation-free rasterization (btg_raster_device_noalloc):
// row counts (4 B), span offsets (8 B) and tile sums for kNoAllocMaxPolys
// polygons, 12 MiB + 8 KiB. Every device the library runs on holds its own copy
// until the process ends; host code reaches it only through the address
// hipGetSymbolAddress gives on that device (noAllocScratchOn).
static __device__ __align__(256) unsigned char noAllocScratch[btg::kNoAllocMaxPolys * 12 +
                                                                (btg::kNoAllocMaxPolys / 1024 + 1) * 8];

namespace btg {

namespace {

constexpr int kWave = 64;
constexpr int kXcds = 8;  // MI355X: 8 XCDs of 32 CUs, each with its own L2
constexpr int kStatSlots = 128, kStatWords = 16;  // counter slots (flushCnt), one 128-byte line each
// Threads per workgroup of the wavefront kernels (beams per workgroup), 4
// waves; -DBTG_BLOCK=128 or 512 builds the other sizes (the LDS columns of the
// trace and split kernels scale with it).
#ifndef BTG_BLOCK
#define BTG_BLOCK 256
#endif
constexpr int kBlock = BTG_BLOCK;
constexpr int kRayBlock = 256;  // primary rays: one 16 x 16 pixel tile per workgroup
constexpr int kPrimaryLeafCap = 4;  // leaf cap of primary passes (at most; see prepare)
constexpr int kMbox = 8;            // mailbox entries per lane (a power of two)

// kSoftG: soft shadows in the FP64-guarded mode (double hit point, the
// double build's apex offset, the apex held as a float pair).
enum App { kSoft = 0, kPrimary = 1, kPoint = 2, kSoftG = 3 };
constexpr bool isSoft(int app) { return app == kSoft || app == kSoftG; }
enum RootState { kTraced = 0, kKilled = 1, kNoBeam = 2 };
enum { kNear = 0, kFar = 1, kBoth = 2 };
enum { kKeep = 0, kAllIn = 1, kClip = 2 };
enum { kFinished = 0, kSplit = 1, kContinue = 2 };
enum { kDescend = 0, kTest = 1, kClimb = 2 };
enum { kHeadAdmit = 0, kHeadTrace = 1, kHeadSplit = 2 };
enum { fAnyHit = 1, fFarIsPlane = 2, fCull = 4 };

// A sub-beam record (64 bytes): corners on its root's plane, restart-trail
// continuation (kd node, depth, pending far children), leaf cursor, hit.
// The same information as the reference's WfBeam (wavefront_core.h:45-55).
struct alignas(16) Rec {
  float qx[4], qy[4];
  uint64_t trail;
  uint32_t node;
  int32_t root;
  int32_t hit;
  int32_t cursor;  // next triangle in the leaf; -1: descend from node
  int16_t depth;
  int8_t n;        // corners; 0 = a dropped sliver (only decrements the root's live count)
  int8_t negBits;  // bit a: rays move toward -a
};
static_assert(sizeof(Rec) == 64, "record layout");

// Per-root query (BeamQuery + BeamCtx::setup, beam_core.h:123-132, plus the
// application's shading data).
struct RootQ {
  Plane plane;
  float ext, eps, minArea, farCq;
  int flags;
  int exclude;
  int unitExp;  // soft: exponent of the fixed-point unit of piece areas
  int tri;      // point: receiver triangle
  float lightArea;
  F3 x, ns;     // soft: shading point and normal
  F3 Olo;       // guarded soft: low part of the apex (plane.O + Olo)
  F3 xlo;       // guarded soft: low part of the shading point (x + xlo, the double hit point)
};

// Ints that many waves update with atomics, each in its own 128-byte line:
// atomics on one line are performed one after another at the memory side, so
// counters that share a line (the queue heads and tails of a round) would
// queue behind each other's traffic.
template <int N>
struct LineInts {
  struct alignas(128) Cell {
    int v;
  };
  Cell c[N];
  __host__ __device__ int& operator[](int i) { return c[i].v; }
  __host__ __device__ volatile int& operator[](int i) volatile { return c[i].v; }
};

struct WState {
  int next;
  int admitBegin, admitCount;
  int done;
  int maxQueue;
  int overflow;
  int round;
  unsigned long long traceItems, splitItems;
  LineInts<2> count;  // records in the trace queues (ping-pong by round)
  LineInts<3> head;   // items taken dynamically by admit, trace, split
  alignas(128) int splitCount;
  alignas(128) int shadeCount;  // soft, exact: unoccluded pieces queued this round for the Lambert integral
  alignas(128) int outCount;
  alignas(128) int killed;
  alignas(128) int blocksDone;
  alignas(128) unsigned long long reserved;
  // Finished workgroups of a split launch, counted per group of workgroups
  // (workgroup b in group b % kXcds, one counter per 128-byte line) before
  // blocksDone counts the finished groups: a few hundred workgroups on one
  // counter serialise at the memory side, as the chunk counters would.
  alignas(128) int groupDone[kXcds * 32];
};

// One soft-shadow job of a batch (a single call is a batch of one). Every
// job's scene lives in the context's scene arena: its kd root is node node0,
// its triangles are numbered from triBase. Its pixels are numbered from
// pixBase in the batch's per-pixel arrays, its roots (pixels in tiles of 64)
// from rootBase, its 16 x 16 primary-ray tiles from tileBase.
struct SoftJob {
  int W, H, tilesX;
  uint32_t node0;
  int triBase, pixBase, rootBase, tileBase;
  Box bounds;  // the kd root box of the job's scene
  float offset, halfW, halfH;
  F3 eye, fwd, right, down;
  F3 lc, lU, lV;
};

// An unoccluded soft-shadow piece waiting for its exact Lambert integral:
// corners on its root's plane (lanes past n repeat the last), root, count.
struct ShadeRec {
  float qx[4], qy[4];
  int32_t root;
  int32_t n;
};
static_assert(sizeof(ShadeRec) == 40, "shade record layout");

struct Params {
  DevScene sv;
  uint32_t node0;  // primary / point: kd root and box of the traced scene
  Box rootBox;
  const SoftJob* jobs;  // soft shadows: the batch's jobs, by increasing rootBase
  int numJobs;
  int* jobLeft;  // open passes: roots of each job not finished yet (null otherwise)
  int app;
  int R;
  const int* rootList;
  int budget;
  unsigned long long capacity;
  Rec* tq[2];
  Rec* sq;
  RootQ* roots;
  int* live;
  int* state;
  int* epoch;
  int pass;
  unsigned long long* missFx;
  unsigned long long* gFx;
  unsigned long long* stats;
  unsigned long long* cycles;  // trace-loop phase clocks (BTG_PHASE_CYCLES builds)
  WState* st;
  int W, H;
  Plane camPlane;
  F3 eye, fwd, right, down;
  float halfW, halfH;
  int tilesX;
  F3 lc, lU, lV;
  float offset;
  int exact;
  const int* pixTri;
  const float* pixX;
  const float* pixXlo;  // guarded soft: low parts of the double hit points (null otherwise)
  int* pixRef;          // FP64 soft: the hit's leaf reference per pixel (null otherwise)
  const double* jobCamD;  // FP64 soft: per job eye, fwd, right, down, halfW, halfH in double
  const float* pixN;
  float* vis;
  float* E;
  float* Eu;
  int tile, tilesW, cull;
  int leafCap;  // leaves a sub-beam may visit per trace launch before it yields (0: no cap)
  int mailbox;  // 1: per-lane mailbox of the triangles the lane's sub-beam has passed unchanged
  F3 pointLight;
  const OutPoly* prim;
  OutPoly* out;
  struct ShadeRec* shq;  // soft, exact: pieces waiting for their Lambert integral
  int* outRoot;
  int* outTag;
  int outCap;
};

// ------------------------------------------------------------ wave helpers
__device__ __forceinline__ int lane() { return int(__lane_id()); }
__device__ __forceinline__ int rankBelow(uint64_t mask) {
  return int(__builtin_amdgcn_mbcnt_hi(uint32_t(mask >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(mask), 0u)));
}
// Exclusive prefix of n over the wave, one atomic on `counter` for the
// wave's total; returns this lane's first slot. All 64 lanes must call it.
__device__ __forceinline__ int waveAppend(int* counter, int n) {
  int x = n;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const int y = __shfl_up(x, o, kWave);
    if (lane() >= o) x += y;
  }
  const int total = __shfl(x, kWave - 1, kWave);
  int base = 0;
  if (lane() == kWave - 1 && total > 0) base = atomicAdd(counter, total);
  base = __shfl(base, kWave - 1, kWave);
  return base + x - n;
}

// How a persistent launch deals its items (queue records, clips, roots) to
// waves, 64 at a time: each wave's first chunk statically (wave w takes items
// [64 w, 64 w + 64)), later chunks from the round's atomic counter, numbered
// after the static ones. Thousands of waves incrementing one counter at the
// start of a launch serialise at the memory side (~20 ns each: a short
// round's trace launch took 60 us for that alone), so a round with fewer
// items than the grid has waves x 64 now takes no atomic at all. Workgroups
// are dealt to the XCDs round-robin, so waves are numbered XCD-major:
// consecutive chunks (sibling sub-beams, neighbouring pixels) land on one XCD
// and share its L2.
struct Deal {
  int wave;   // this wave's static chunk
  int waves;  // waves in the grid: items below 64 * waves are dealt statically
};
__device__ __forceinline__ Deal dealOf() {
  const int G = int(gridDim.x), b = int(blockIdx.x);
  const int bx = G % kXcds == 0 ? (b % kXcds) * (G / kXcds) + b / kXcds : b;
  return Deal{bx * (kBlock / kWave) + int(threadIdx.x) / kWave, G * (kBlock / kWave)};
}

// The next chunk of `n` items for the calling wave (all 64 lanes call it):
// the static chunk on the first call, then a chunk of `want` from `counter`.
// Returns the chunk's first item; `exhausted` is set once no item is left.
__device__ __forceinline__ int nextChunk(const Deal& dl, int* counter, int want, int n, bool& first,
                                         bool& exhausted) {
  const int statics = dl.waves * kWave;
  int base;
  if (first) {
    first = false;
    base = dl.wave * kWave;
    if (statics >= n) exhausted = true;
    return base;
  }
  base = 0;
  if (lane() == 0) base = atomicAdd(counter, want);
  base = __shfl(base, 0, kWave) + statics;
  if (base + want >= n) exhausted = true;
  return base;
}

// One slot of `counter` per calling lane, one atomic per wave; callable from
// divergent code (the active lanes share the atomic of their lowest lane).
__device__ __forceinline__ int activeAppend(int* counter) {
  const uint64_t m = __ballot(1);
  const int leader = __builtin_ctzll(m);
  int base = 0;
  if (lane() == leader) base = atomicAdd(counter, __popcll(m));
  base = __shfl(base, leader, kWave);
  return base + rankBelow(m);
}

struct Cnt {
  uint32_t v[sNum];
  double dropped;
  __device__ void zero() {
#pragma unroll
    for (int i = 0; i < sNum; ++i) v[i] = 0;
    dropped = 0;
  }
};

// A wave's counters are added into the stat slot of its workgroup
// (blockIdx mod kStatSlots, one 128-byte line each: the counters in words
// 0..sNum-1, the dropped area as a double in word kStatWords - 1); the host
// sums the slots. With one shared set of counters, the waves of a launch that
// finish together queued their atomics on one line.
__device__ void flushCnt(const Params& p, const Cnt& c) {
  unsigned long long* const slot = p.stats + (blockIdx.x % kStatSlots) * kStatWords;
#pragma unroll
  for (int f = 0; f < sNum; ++f) {
    unsigned long long s = c.v[f];
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) s += __shfl_down(s, o, kWave);
    if (lane() == 0 && s) atomicAdd(slot + f, s);
  }
  double d = c.dropped;
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) d += __shfl_down(d, o, kWave);
  if (lane() == 0 && d != 0) atomicAdd(reinterpret_cast<double*>(slot + kStatWords - 1), d);
}

// ------------------------------------------------------------ beam context
struct Ctx {
  Plane P;
  F3 Olo;  // low part of the apex; a compile-time zero outside kSoftG
  float ext, eps, minArea, farCq;
  int flags, exclude;
  __device__ bool anyHit() const { return flags & fAnyHit; }
  __device__ bool farIsPlane() const { return flags & fFarIsPlane; }
};

__device__ __forceinline__ Ctx ctxOf(const RootQ& q) {
  Ctx c;
  c.P = q.plane;
  c.Olo = f3(0, 0, 0);
  c.ext = q.ext;
  c.eps = q.eps;
  c.minArea = q.minArea;
  c.farCq = q.farCq;
  c.flags = q.flags;
  c.exclude = q.exclude;
  return c;
}

// The flags every root of an application has, as compile-time constants:
// soft and point shadows take any hit and stop at a far plane
// (lightBeamSetup, soft_core.h:38-73; pointShadowSetup, point_core.h:15-39),
// primary visibility keeps the nearest hit with optional back-face culling.
// Kernels are instantiated per application, so each drops the other
// applications' code and registers.
template <int App>
__device__ __forceinline__ Ctx modeCtx(const RootQ& q) {
  Ctx c = ctxOf(q);
  c.flags = App == kPrimary ? (q.flags & fCull) : (fAnyHit | fFarIsPlane);
  if (App == kSoftG) c.Olo = q.Olo;
  return c;
}

// Vectors and coordinates relative to the apex: (p - O) - Olo. Outside the
// guarded mode Olo is the constant +0, and x - (+0) is x for every float x,
// so these are p - O exactly (and compile to it).
__device__ __forceinline__ F3 relApex(const Ctx& c, F3 p) { return (p - c.P.O) - c.Olo; }
__device__ __forceinline__ float relApex(const Ctx& c, float s, int a) { return (s - comp(c.P.O, a)) - comp(c.Olo, a); }

// BeamCtx::setup (beam_core.h:123-132).
__device__ __forceinline__ void setupQuery(RootQ& q, float extent) {
  q.ext = extent > 0 ? extent : 1.0f;
  q.eps = kEpsRel * q.ext;
  q.minArea = q.eps * q.eps;
  q.farCq = dot(q.plane.n, q.plane.d);
}

// CoreBeam (beam_core.h:30-40).
struct Beam {
  C4 qx, qy, aw0, aw1, aw2, farA;
  float farC;
  int n, hit;
  int s0, s1, s2;  // +-1 per axis
  bool hasFar;
};

// Per-axis fields picked by value (a select of loaded values, never of
// addresses, so the beam stays in registers).
__device__ __forceinline__ int sgnOf(const Beam& b, int a) {
  const int s0 = b.s0, s1 = b.s1, s2 = b.s2;
  return a == 0 ? s0 : (a == 1 ? s1 : s2);
}
__device__ __forceinline__ C4 awOf(const Beam& b, int a) {
  C4 r;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const float x0 = b.aw0.v[i], x1 = b.aw1.v[i], x2 = b.aw2.v[i];
    r.v[i] = a == 0 ? x0 : (a == 1 ? x1 : x2);
  }
  return r;
}

// Plane of the current hit cached with the beam (WfWork, wavefront_core.h:58-61).
struct Work {
  Beam b;
  float hitA[3], hitC;
  F3 hitN;
};

// The beam's |w| per axis kept in the lane's LDS column as well (word
// (4 a + i) * kBlock for axis a, corner i): the descent picks an axis per node
// at run time, and reading that axis's four values from LDS costs no VALU
// instruction, where selecting them among the registers of the three axes took
// eight selects per axis.
__device__ __forceinline__ C4 awAt(const float* awCol, int a) {
  const float* q = awCol + a * 4 * kBlock;
  return c4(q[0], q[kBlock], q[2 * kBlock], q[3 * kBlock]);
}
__device__ __forceinline__ void awStore(float* awCol, const Beam& b) {
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    awCol[i * kBlock] = b.aw0.v[i];
    awCol[(4 + i) * kBlock] = b.aw1.v[i];
    awCol[(8 + i) * kBlock] = b.aw2.v[i];
  }
}

// Signed distances from the apex to a box's slabs along each axis, in the
// rays' direction: nIn[k] to the face they enter by, nOut[k] to the face they
// leave by (the Nmin / Nmax of beam_core.h:146-149, 171-173). The descent
// carries them instead of the box: narrowing the box to a child moves one
// face of the split axis onto the split plane, whose distance is the
// descent's own Ns, so only that one value changes per kd step.
struct SlabDist {
  F3 nIn, nOut;
};
__device__ __forceinline__ SlabDist slabDist(const Ctx& c, const Beam& b, const Box& box) {
  SlabDist d;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int sg = sgnOf(b, k);
    const float s = float(sg);
    setComp(d.nIn, k, s * relApex(c, sg > 0 ? comp(box.lo, k) : comp(box.hi, k), k));
    setComp(d.nOut, k, s * relApex(c, sg > 0 ? comp(box.hi, k) : comp(box.lo, k), k));
  }
  return d;
}

// decide (beam_core.h:139-164): near / far / both children of an inner node.
__device__ int decide(const Beam& b, int a, float Ns, const SlabDist& sd, const float* awCol) {
  const C4 awa = awAt(awCol, a);
#pragma unroll
  for (int k = 1; k <= 2; ++k) {
    const int cc = a + k >= 3 ? a + k - 3 : a + k;
    const C4 lhs = scaled(awAt(awCol, cc), Ns);
    if (allBelow(lhs, scaled(awa, comp(sd.nIn, cc)), kKdRel)) return kFar;
    if (allAbove(lhs, scaled(awa, comp(sd.nOut, cc)), kKdRel)) return kNear;
  }
  if (b.hasFar && allAbove(scaled(b.farA, Ns), scaled(awa, b.farC), kKdRel)) return kNear;
  return kBoth;
}

// overlaps (beam_core.h:166-195): false if the beam certainly misses the box
// whose slab distances are sd.
__device__ bool overlaps(const Beam& b, const SlabDist& sd) {
  const float nmin[3] = {sd.nIn.x, sd.nIn.y, sd.nIn.z}, nmax[3] = {sd.nOut.x, sd.nOut.y, sd.nOut.z};
#pragma unroll
  for (int k = 0; k < 3; ++k)
    if (nmax[k] < 0) return false;
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int e = 0; e < 3; ++e) {
      if (k == e) continue;
      if (allBelow(scaled(awOf(b, e), nmax[k]), scaled(awOf(b, k), nmin[e]), kKdRel)) return false;
    }
  if (b.hasFar) {
#pragma unroll
    for (int k = 0; k < 3; ++k)
      if (allAbove(scaled(b.farA, nmin[k]), scaled(awOf(b, k), b.farC), kKdRel)) return false;
  }
  return true;
}
__device__ __forceinline__ bool overlaps(const Ctx& c, const Beam& b, const Box& box) {
  return overlaps(b, slabDist(c, b, box));
}

__device__ __forceinline__ void setAw(Beam& b, const Ctx& c) {
  const Plane& P = c.P;
  C4 w[3] = {affine(b.qx, P.u.x, b.qy, P.v.x, P.d.x), affine(b.qx, P.u.y, b.qy, P.v.y, P.d.y),
             affine(b.qx, P.u.z, b.qy, P.v.z, P.d.z)};
  const int sg[3] = {b.s0, b.s1, b.s2};
  C4 aw[3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const float ws = w[a].v[i] * float(sg[a]);
      aw[a].v[i] = 0.0f > ws ? 0.0f : ws;
    }
  b.aw0 = aw[0];
  b.aw1 = aw[1];
  b.aw2 = aw[2];
}

// beamGeom (beam_core.h:200-221): corners (lanes past n repeat the last) and
// |w| per axis; slivers below minArea are dropped.
// `welded`: p is already the output of cleanPoly, whose last pass over the
// vertices removed nothing, so a further cleanPoly would return it unchanged
// and is skipped.
template <int N>
__device__ __forceinline__ bool beamGeom(const Ctx& c, const Pg<N>& p, int s0, int s1, int s2, Beam& b, float& dropped,
                                         bool welded = false) {
  Pg<N> q = p;
  if (!welded) cleanPoly(q, c.eps);
  const float A = q.n >= 3 ? q.area() : 0.0f;
  if (q.n < 3 || A < c.minArea) {
    dropped = selAbs(p.area());
    return false;
  }
  b.qx = c4(q.x[0], q.x[1], q.x[2], q.n > 3 ? q.x[3] : q.x[2]);
  b.qy = c4(q.y[0], q.y[1], q.y[2], q.n > 3 ? q.y[3] : q.y[2]);
  b.n = q.n;
  b.s0 = s0;
  b.s1 = s1;
  b.s2 = s2;
  setAw(b, c);
  return true;
}

// setStatus (beam_core.h:225-240).
__device__ __forceinline__ void setStatus(const Ctx& c, Beam& b, int hit, const float* A, float C) {
  b.hit = hit;
  if (hit >= 0) {
    const C4 f = affine(b.qx, A[0], b.qy, A[1], A[2]);
#pragma unroll
    for (int i = 0; i < 4; ++i) b.farA.v[i] = __builtin_fabsf(f.v[i]);
    b.farC = selAbs(C);
    b.hasFar = true;
  } else if (c.farIsPlane()) {
    const float c0 = selAbs(c.farCq);
    b.farA = splat(c0);
    b.farC = c0;
    b.hasFar = true;
  } else {
    b.farC = 0;
    b.hasFar = false;
  }
}

// hitPlane (beam_core.h:290-297) from the packed per-triangle normal/vertex.
__device__ __forceinline__ void hitPlane(const Params& p, const Ctx& c, int t, Work& w) {
  const float4 nn = p.sv.triNV[2 * size_t(t)], vv = p.sv.triNV[2 * size_t(t) + 1];
  const F3 nT = f3(nn.x, nn.y, nn.z);
  w.hitN = nT;
  w.hitA[0] = dot(nT, c.P.u);
  w.hitA[1] = dot(nT, c.P.v);
  w.hitA[2] = dot(nT, c.P.d);
  w.hitC = dot(nT, relApex(c, f3(vv.x, vv.y, vv.z)));
}

// wfExpand (wavefront_core.h:92-112). Any-hit applications (soft and point
// shadows) never queue a beam with a hit (occluded pieces end where they are
// found), so their records always carry hit = -1 and the far plane is the
// query's; fixing that at compile time keeps the hit plane out of registers.
template <int App>
__device__ void expand(const Params& p, const Ctx& c, const Rec& r, Work& w) {
  Beam& b = w.b;
  const int n = r.n;
  // corners past n repeat the last (selects, not a runtime index, so the
  // record stays in registers)
  b.qx = c4(r.qx[0], r.qx[1], r.qx[2], n > 3 ? r.qx[3] : r.qx[2]);
  b.qy = c4(r.qy[0], r.qy[1], r.qy[2], n > 3 ? r.qy[3] : r.qy[2]);
  b.n = n;
  b.s0 = (r.negBits & 1) ? -1 : 1;
  b.s1 = (r.negBits & 2) ? -1 : 1;
  b.s2 = (r.negBits & 4) ? -1 : 1;
  setAw(b, c);
  const int hit = App == kPrimary ? r.hit : -1;
  if (hit >= 0) {
    hitPlane(p, c, hit, w);
    setStatus(c, b, hit, w.hitA, w.hitC);
  } else {
    w.hitA[0] = w.hitA[1] = w.hitA[2] = w.hitC = 0;
    w.hitN = f3(0, 0, 0);
    setStatus(c, b, -1, nullptr, 0);
  }
}

// wfStoreGeom (wavefront_core.h:114-121).
__device__ __forceinline__ void storeGeom(const Beam& b, Rec& r) {
  r.n = int8_t(b.n);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    r.qx[i] = b.qx.v[i];
    r.qy[i] = b.qy.v[i];
  }
  r.negBits = int8_t((b.s0 < 0 ? 1 : 0) | (b.s1 < 0 ? 2 : 0) | (b.s2 < 0 ? 4 : 0));
}

// Triangle setup (computeTriInfo + triSetup, beam_core.h:251-286): edge
// planes through the apex in canonical vertex order, plane coefficients.
// The trace computes the pieces lazily - each edge line when its test is
// reached, the plane only when all edges pass - with exactly the operations
// of the full setup, so every value it uses is bit-identical.
struct TriS {
  int tri;
  bool ok;
  float s;
  uint32_t vi[3];
  Ln edge[3];
  float A[3], C;
  F3 r[3], n, m0;
};

// Vertices relative to the apex, first edge normal, orientation and the
// degenerate / back-face rejection.
__device__ __forceinline__ void triBegin(const Ctx& c, const LeafTri& ref, TriS& ti) {
  ti.r[0] = relApex(c, f3(ref.v[0], ref.v[1], ref.v[2]));
  ti.r[1] = relApex(c, f3(ref.v[3], ref.v[4], ref.v[5]));
  ti.r[2] = relApex(c, f3(ref.v[6], ref.v[7], ref.v[8]));
  ti.n = f3(ref.n[0], ref.n[1], ref.n[2]);
  ti.vi[0] = ref.vi[0];
  ti.vi[1] = ref.vi[1];
  ti.vi[2] = ref.vi[2];
  ti.tri = ref.tri;
  ti.ok = false;
  ti.m0 = ref.vi[0] < ref.vi[1] ? cross(ti.r[0], ti.r[1]) : -cross(ti.r[1], ti.r[0]);
  const float D = dot(ti.r[2], ti.m0);
  const float scale = length(ti.m0) * length(ti.r[2]);
  if (!(selAbs(D) > kDegRel * scale)) return;  // apex in the triangle's plane
  if ((c.flags & fCull) && D > 0) return;      // back-facing
  ti.s = D > 0 ? 1.0f : -1.0f;
  ti.ok = true;
}

// Edge line k (k = 0: v0 v1, 1: v1 v2, 2: v2 v0).
template <int k>
__device__ __forceinline__ Ln triEdge(const Ctx& c, const TriS& ti) {
  constexpr int i = k, j = (k + 1) % 3;
  const F3 m = k == 0 ? ti.m0 : (ti.vi[i] < ti.vi[j] ? cross(ti.r[i], ti.r[j]) : -cross(ti.r[j], ti.r[i]));
  const Plane& P = c.P;
  return makeLine(ti.s * dot(m, P.u), ti.s * dot(m, P.v), ti.s * dot(m, P.d), c.ext);
}

__device__ __forceinline__ void triPlane(const Ctx& c, TriS& ti) {
  const Plane& P = c.P;
  ti.A[0] = dot(ti.n, P.u);
  ti.A[1] = dot(ti.n, P.v);
  ti.A[2] = dot(ti.n, P.d);
  ti.C = dot(ti.n, ti.r[0]);
}

// The complete setup (the split kernel needs every line).
__device__ __forceinline__ TriS triSetup(const Ctx& c, const LeafTri& ref) {
  TriS ti;
  triBegin(c, ref, ti);
  if (!ti.ok) return ti;
  ti.edge[0] = triEdge<0>(c, ti);
  ti.edge[1] = triEdge<1>(c, ti);
  ti.edge[2] = triEdge<2>(c, ti);
  triPlane(c, ti);
  return ti;
}

// closerLine (beam_core.h:302-317).
__device__ Ln closerLine(const Ctx& c, const TriS& T, const float* AF, float CF, F3 nF, int fId) {
  const float cosang = dot(T.n, nF);
  const float scaleC = selMax(selAbs(T.C), selAbs(CF));
  const float ca = selAbs(cosang);
  const float diff = T.C - cosang * CF;
  if (ca > 1 - kCopCos && selAbs(diff) <= kCopRel * scaleC) return Ln{0, 0, (fId >= 0 && T.tri < fId) ? 1.0f : -1.0f};
  const float sg = T.C * CF > 0 ? 1.0f : -1.0f;
  return makeLine(sg * (CF * T.A[0] - T.C * AF[0]), sg * (CF * T.A[1] - T.C * AF[1]), sg * (CF * T.A[2] - T.C * AF[2]),
                  c.ext);
}

// separated (beam_core.h:321-342): some beam side plane has all three
// vertices strictly outside. Lane k holds side k -> k+1 (the reference's rot1).
__device__ bool separated(const Ctx& c, const Beam& b, const TriS& ti) {
  const Plane& P = c.P;
  const C4 wx = affine(b.qx, P.u.x, b.qy, P.v.x, P.d.x);
  const C4 wy = affine(b.qx, P.u.y, b.qy, P.v.y, P.d.y);
  const C4 wz = affine(b.qx, P.u.z, b.qy, P.v.z, P.d.z);
  float nx[4], ny[4], nz[4], nl[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int j = (i + 1) & 3;
    nx[i] = wy.v[i] * wz.v[j] - wz.v[i] * wy.v[j];
    ny[i] = wz.v[i] * wx.v[j] - wx.v[i] * wz.v[j];
    nz[i] = wx.v[i] * wy.v[j] - wy.v[i] * wx.v[j];
    nl[i] = __builtin_sqrtf(nx[i] * nx[i] + ny[i] * ny[i] + nz[i] * nz[i]);
  }
  int outAll = b.n == 4 ? 0xF : 0xB;  // a triangle's lane 2 side is degenerate
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const F3 r = ti.r[k];
    const float lr = length(r) * kEpsRel;
    int m = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const float s = nx[i] * r.x + ny[i] * r.y + nz[i] * r.z;
      const float tol = nl[i] * lr;
      if (s < -tol) m |= 1 << i;
    }
    outAll &= m;
  }
  return outAll != 0;
}

// testLine (beam_core.h:356-362): -1 no corner definitely inside; otherwise
// bit 0 set if some corner is definitely outside.
__device__ __forceinline__ int lineTest(const Ctx& c, const Beam& B, const Ln& L) {
  const C4 dv = affine(B.qx, L.a, B.qy, L.b, L.c);
  bool in = false, out = false;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    in = in || dv.v[i] > c.eps;
    out = out || dv.v[i] < -c.eps;
  }
  return in ? (out ? 1 : 0) : -1;
}

// classifyTri (beam_core.h:374-409) with every line the clip needs: the
// three edges in L[0..2], then the closer-than-hit line and the far-plane line
// when they apply (nl of them in all). The slots are fixed so that L stays in
// registers.
__device__ __forceinline__ int classifyTri(const Ctx& c, const Work& w, const TriS& ti, Ln (&L)[5], int& nl) {
  const Beam& B = w.b;
  nl = 0;
  bool allIn = true;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    L[k] = ti.edge[k];
    const int r = lineTest(c, B, ti.edge[k]);
    if (r < 0) return kKeep;
    allIn = allIn && r == 0;
    ++nl;
  }
  const bool hasCloser = !c.anyHit() && B.hit >= 0;
  if (hasCloser) {
    L[3] = closerLine(c, ti, w.hitA, w.hitC, w.hitN, B.hit);
    const int r = lineTest(c, B, L[3]);
    if (r < 0) return kKeep;
    allIn = allIn && r == 0;
    ++nl;
  }
  if (c.farIsPlane()) {
    const float farA[3] = {0, 0, c.farCq};
    const Ln cl = closerLine(c, ti, farA, c.farCq, c.P.n, -1);
    // field-wise selects: an aggregate ?: would store through a chosen slot
    L[3] = Ln{hasCloser ? L[3].a : cl.a, hasCloser ? L[3].b : cl.b, hasCloser ? L[3].c : cl.c};
    L[4] = cl;
    const int r = lineTest(c, B, cl);
    if (r < 0) return kKeep;
    allIn = allIn && r == 0;
    ++nl;
  }
  if (allIn) return kAllIn;
  if (separated(c, B, ti)) return kKeep;
  return kClip;
}

// classifyTri for the trace kernel with the lazy setup: the same decisions as
// classifyTri (beam_core.h:374-409), computing each line only when reached.
__device__ int classifyLazy(const Ctx& c, const Work& w, TriS& ti) {
  const Beam& B = w.b;
  bool allIn = true;
  int r = lineTest(c, B, triEdge<0>(c, ti));
  if (r < 0) return kKeep;
  allIn = r == 0;
  r = lineTest(c, B, triEdge<1>(c, ti));
  if (r < 0) return kKeep;
  allIn = allIn && r == 0;
  r = lineTest(c, B, triEdge<2>(c, ti));
  if (r < 0) return kKeep;
  allIn = allIn && r == 0;
  triPlane(c, ti);
  if (!c.anyHit() && B.hit >= 0) {
    r = lineTest(c, B, closerLine(c, ti, w.hitA, w.hitC, w.hitN, B.hit));
    if (r < 0) return kKeep;
    allIn = allIn && r == 0;
  }
  if (c.farIsPlane()) {
    const float farA[3] = {0, 0, c.farCq};
    r = lineTest(c, B, closerLine(c, ti, farA, c.farCq, c.P.n, -1));
    if (r < 0) return kKeep;
    allIn = allIn && r == 0;
  }
  if (allIn) return kAllIn;
  if (separated(c, B, ti)) return kKeep;
  return kClip;
}

// Line k of L[0..5) at a runtime index (a select chain, not an array index).
__device__ __forceinline__ Ln lineAt(const Ln (&L)[5], int k) {
  Ln r = L[0];
#pragma unroll
  for (int i = 1; i < 5; ++i)
    if (k == i) r = L[i];
  return r;
}

// clipTri (beam_core.h:419-479), pieces delivered to sink.push(list, poly):
// list 0 keeps the beam's status, list 1 takes the triangle. Edges are
// ordered by their positive corner sums first. The reference keeps pentagon
// halves waiting on a stack; its entries always have strictly increasing
// start lines from bottom to top (a piece started at line s only pushes
// halves that start after s), so one register slot per start line holds the
// stack, and popping takes the highest occupied line. One piece-line step per
// iteration, so the sink has a single call site.
template <class Sink>
__device__ __forceinline__ void clipTri(const Ctx& c, const Beam& B, Ln (&L)[5], int nl, Sink& sink, uint32_t& fiveSplits) {
  float score[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const C4 dv = affine(B.qx, L[k].a, B.qy, L[k].b, L[k].c);
    float ps[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) ps[i] = dv.v[i] > 0.0f ? dv.v[i] : 0.0f;
    score[k] = ps[0] + ps[1] + ps[2] + (B.n == 4 ? ps[3] : 0.0f);
  }
  // insertion sort of the three edges by score (stable, as the reference)
  if (score[1] < score[0]) {
    const float t = score[0];
    score[0] = score[1];
    score[1] = t;
    const Ln l = L[0];
    L[0] = L[1];
    L[1] = l;
  }
  if (score[2] < score[1]) {
    const float t = score[1];
    score[1] = score[2];
    score[2] = t;
    const Ln l = L[1];
    L[1] = L[2];
    L[2] = l;
    if (score[1] < score[0]) {
      const float t2 = score[0];
      score[0] = score[1];
      score[1] = t2;
      const Ln l2 = L[0];
      L[0] = L[1];
      L[1] = l2;
    }
  }
  Pg<6> p;
  p.n = B.n;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    p.x[i] = B.qx.v[i];
    p.y[i] = B.qy.v[i];
  }
  p.x[4] = p.x[5] = p.y[4] = p.y[5] = 0;
  // p is welded (the output of cleanPoly) while it is the beam itself (every
  // beam polygon comes out of beamGeom) or a welded split part; a pentagon's
  // quad part and a pending triangle are not
  bool welded = true;
  // pending triangle halves by start line 1..5 (slot 0 unused)
  float px[6][3], py[6][3];
  uint32_t pending = 0;
  int li = 0;
  for (;;) {
    int list = -1;  // the piece handed to the sink in this step
    Pg<6> q;
    bool qWelded = false;
    bool pieceDone = false;
    if (li >= nl) {  // inside every line: the part of the beam the triangle covers
      list = 1;
      q = p;
      qWelded = welded;
      pieceDone = true;
    } else {
      const Ln Lk = lineAt(L, li);
      const int cs = sidesOf(p, Lk, c.eps);
      if (!(cs & 2)) {
        ++li;
      } else if (!(cs & 1)) {
        list = 0;
        q = p;
        qWelded = welded;
        pieceDone = true;
      } else {
        Pg<6> in, out;
        splitPoly(p, Lk, c.eps, in, out);  // pieces of a <= 4-gon have <= 5 corners: never fails
        cleanPoly(out, c.eps);
        if (out.n >= 3) {
          list = 0;
          q = out;
          qWelded = true;
        }
        cleanPoly(in, c.eps);
        if (in.n < 3) {
          pieceDone = true;
        } else if (in.n == 5) {  // quad (0..3) goes on, triangle (0, 3, 4) waits at line li + 1
          ++fiveSplits;
          const int lv = li + 1;
#pragma unroll
          for (int s = 1; s < 6; ++s) {
            if (s == lv) {
              px[s][0] = in.x[0];
              py[s][0] = in.y[0];
              px[s][1] = in.x[3];
              py[s][1] = in.y[3];
              px[s][2] = in.x[4];
              py[s][2] = in.y[4];
            }
          }
          pending |= 1u << lv;
          p = in;
          p.n = 4;
          welded = false;
          li = lv;
        } else {
          p = in;
          welded = true;
          ++li;
        }
      }
    }
    if (list >= 0) sink.push(list, q, qWelded);
    if (pieceDone) {
      if (!pending) break;
      const int lv = 31 - __builtin_clz(pending);
      pending &= ~(1u << lv);
      p.n = 3;
      welded = false;
#pragma unroll
      for (int s = 1; s < 6; ++s) {
        if (s == lv) {
#pragma unroll
          for (int v = 0; v < 3; ++v) {
            p.x[v] = px[s][v];
            p.y[v] = py[s][v];
          }
        }
      }
      li = lv;
    }
  }
}

// Pieces of one clip kept in LDS between the counting and the emitting pass:
// up to kStageSlots per lane, 11 words each (corners 0..4, then
// n | list << 8 | welded << 9;
// pieces have at most 5 corners), word k of slot s of a lane at
// col[(s * kStageWords + k) * kBlock].
#ifndef BTG_STAGE_SLOTS
#define BTG_STAGE_SLOTS 7
#endif
constexpr int kStageSlots = BTG_STAGE_SLOTS, kStageWords = 11;

// Counting pass: records each list produces (pentagons become two) and the
// total area of the hit pieces (summed in piece order, as the reference).
// With a staging column it also keeps the pieces, in push order, so that the
// emitting pass replays them instead of clipping again; a clip with more
// pieces than slots is marked and clipped again.
struct CountSink {
  int recs0 = 0, recs1 = 0;
  float hitArea = 0;
  float* col = nullptr;
  int staged = 0;
  bool overflow = false;
  __device__ __forceinline__ void push(int l, const Pg<6>& q, bool welded) {
    const int k = q.n == 5 ? 2 : 1;
    if (l == 1) {
      recs1 += k;
      hitArea += q.area();
    } else {
      recs0 += k;
    }
    if (!col) return;
    if (staged >= kStageSlots) {
      overflow = true;
      return;
    }
    float* w = col + staged * kStageWords * kBlock;
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      w[v * kBlock] = q.x[v];
      w[(5 + v) * kBlock] = q.y[v];
    }
    w[10 * kBlock] = __int_as_float(q.n | (l << 8) | (welded ? 1 << 9 : 0));
    ++staged;
  }
};

// -------------------------------------------------------------- outputs
__device__ __forceinline__ double pow2(int e) { return __builtin_ldexp(1.0, e); }

// polygonFormFactorG (soft_core.h:18-33) of a 3- or 4-gon, unrolled with
// constant indices so that the corners stay in registers.
__device__ float formFactor(F3 x, F3 n, const F3 (&poly)[4], int count) {
  double sum = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    if (i >= count) break;
    const F3 a = poly[i] - x, b = (i + 1 < count ? poly[(i + 1) & 3] : poly[0]) - x;
    const double la = double(length(a)), lb = double(length(b));
    if (la == 0 || lb == 0) continue;
    double cc = double(dot(a, b)) / (la * lb);
    cc = cc < -1 ? -1 : (cc > 1 ? 1 : cc);
    const double theta = acos(cc);
    const F3 g = cross(a, b);
    const double lg = double(length(g));
    if (lg == 0) continue;
    sum += theta * double(dot(g, n)) / lg;
  }
  return float((sum < 0 ? -sum : sum) * 0.5);
}

#include "trace_fp64.h"

// Signed area of a beam's 3- or 4-gon, as PolyN::area (beam_geom.h:27-33)
// sums it. A triangle's fourth corner repeats its third; the extra term
// x2*y2 - x2*y2 is exactly +0 and leaves the sum unchanged.
__device__ __forceinline__ float area4(const C4& x, const C4& y) {
  float a = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) a += x.v[i] * y.v[(i + 1) & 3] - x.v[(i + 1) & 3] * y.v[i];
  return a * 0.5f;
}

// Appends an output polygon (primary / point shadows); the active lanes of a
// wave share one atomic.
__device__ void appendPoly(const Params& p, int root, const OutPoly& ob) {
  const int k = activeAppend(&p.st->outCount);
  if (k >= p.outCap) return;  // the host traces again with a larger buffer
  p.out[k] = ob;
  p.outRoot[k] = root;
  p.outTag[k] = p.pass;
}

// pointShadowToImage (point_core.h:41-64) of a 3- or 4-gon on the receiver
// plane sp: project to the image plane, make it counter-clockwise, repeat the
// last corner.
__device__ bool pointToImage(const Plane& camPlane, const Plane& sp, const C4& sx, const C4& sy, int n, int T,
                             OutPoly& ip) {
  ip.n = n;
  ip.tri = T;
  ip.area = 0;
  C4 X, Y;
#pragma unroll
  for (int i = 0; i < 3; ++i)
    if (!camPlane.project(sp.point(sx.v[i], sy.v[i]), X.v[i], Y.v[i])) return false;
  if (n > 3) {
    if (!camPlane.project(sp.point(sx.v[3], sy.v[3]), X.v[3], Y.v[3])) return false;
  } else {
    X.v[3] = X.v[2];
    Y.v[3] = Y.v[2];
  }
  const float a = area4(X, Y);
  if (!(a > 0 || a < 0)) return false;
  if (a < 0) {  // reverse the n corners, then repeat the last
    const C4 rx = n > 3 ? c4(X.v[3], X.v[2], X.v[1], X.v[0]) : c4(X.v[2], X.v[1], X.v[0], X.v[0]);
    const C4 ry = n > 3 ? c4(Y.v[3], Y.v[2], Y.v[1], Y.v[0]) : c4(Y.v[2], Y.v[1], Y.v[0], Y.v[0]);
    X = rx;
    Y = ry;
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    ip.x[i] = X.v[i];
    ip.y[i] = Y.v[i];
  }
  return true;
}

// A finished piece (wfOutput, wavefront_core.h:124-132, and the per-app
// handling of cuda/wavefront.cu:193-227).
template <int App>
__device__ __forceinline__ void emitOutput(const Params& p, const Ctx& c, const Beam& b, int root, int tri, Cnt& cnt) {
  const float area = area4(b.qx, b.qy);
  if (tri >= 0) cnt.v[sHits]++;
  cnt.v[sBeams]++;
  if (isSoft(App)) {
    if (tri >= 0) return;  // occluded
    const RootQ& R = p.roots[root];
    if (area > 0) atomicAdd(&p.missFx[root], (unsigned long long)(double(area) * pow2(-R.unitExp)));
    if (p.exact) {  // the Lambert integral is taken by shadeKernel, after this round's trace
      const int k = activeAppend(&p.st->shadeCount);
      if ((unsigned long long)k < p.capacity) {
        ShadeRec sr;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          sr.qx[i] = b.qx.v[i];
          sr.qy[i] = b.qy.v[i];
        }
        sr.root = root;
        sr.n = b.n;
        p.shq[k] = sr;
      } else {
        p.st->overflow = 1;
      }
    }
    return;
  }
  if (tri < 0) return;
  OutPoly ob;
  if (App == kPrimary) {
    // the beam's corners past n already repeat the last (OutBeam's layout)
    ob.n = b.n;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      ob.x[i] = b.qx.v[i];
      ob.y[i] = b.qy.v[i];
    }
    ob.tri = tri;
    ob.area = area;
  } else {
    const RootQ& R = p.roots[root];
    if (!pointToImage(p.camPlane, c.P, b.qx, b.qy, b.n, R.tri, ob)) return;
  }
  appendPoly(p, root, ob);
}

// ---------------------------------------------------------- root budgets
__device__ int jobOfRoot(const Params& p, int k);

// A root has no live sub-beam left: release its budget reservation and, in an
// open pass, count it off its job.
__device__ __forceinline__ void rootDone(const Params& p, int root) {
  atomicAdd(&p.st->reserved, (unsigned long long)(-(long long)p.budget));
  if (p.jobLeft) atomicSub(&p.jobLeft[jobOfRoot(p, root)], 1);
}
__device__ __forceinline__ void liveSub(const Params& p, int root, int n) {
  if (atomicSub(&p.live[root], n) == n) rootDone(p, root);
}
__device__ __forceinline__ void killRoot(const Params& p, int root) {
  if (atomicExch(&p.state[root], int(kKilled)) != int(kKilled)) atomicAdd(&p.st->killed, 1);
}

// ------------------------------------------------------------- the trace
// Kd depths whose ancestors a lane keeps in LDS (deeper ones are reached by
// parent links). The SAH build limits depth to 8 + 1.3 log2(triangles), 36 at
// 2M triangles, unless a deeper limit is asked for (63 at most). A build with
// -DBTG_PATH_LEVELS=4 sends most climbs through the parent-link walk (a test
// build for that path).
#ifndef BTG_PATH_LEVELS
#define BTG_PATH_LEVELS 32
#endif
constexpr int kPathLevels = BTG_PATH_LEVELS;

struct Stepper {
  int phase;
  SlabDist sd;                     // descend: slab distances of r.node's box
  KdNodeD kn;                      // descend: node r.node
  uint32_t leafOffset, leafCount;  // the current leaf
  int validFrom;                   // path levels [validFrom, min(r.depth, kPathLevels)) are valid
};

// A lane's column of the LDS ancestor path: level k at col[k * kBlock] holds
// the far child (the child the rays move into) of r.node's ancestor at depth
// k, which is where a climb to level k goes.
struct Path {
  uint32_t* col;
  __device__ __forceinline__ uint32_t get(int k) const { return col[k * kBlock]; }
  __device__ __forceinline__ void set(int k, uint32_t v) const { col[k * kBlock] = v; }
};

__device__ __forceinline__ int highestBit(uint64_t x) { return 63 - __builtin_clzll(x); }

// Phase timing of the trace loop (build with -DBTG_PHASE_CYCLES): wave clock
// cycles per code region, summed into p.cycles[0..8). PHASE_MARK(k) charges
// the time since the previous mark to region k: 0 refill, 1 descend, 2 climb
// (parent walk), 3 climb (child box test), 4 output emit, 5 triangle tests,
// 6 queue appends, 7 the rest of the traversal loop.
enum { kPhRefill, kPhDescend, kPhWalk, kPhBox, kPhEmit, kPhTest, kPhAppend, kPhRest, kPhases };
#ifdef BTG_PHASE_CYCLES
struct PhaseClock {
  uint64_t t[kPhases];
  uint64_t prev;
};
#define PHASE_ARG , PhaseClock& ph_
#define PHASE_PASS , ph_
#define PHASE_INIT()                                  \
  PhaseClock ph_;                                     \
  for (int k_ = 0; k_ < kPhases; ++k_) ph_.t[k_] = 0; \
  ph_.prev = __builtin_amdgcn_s_memtime()
#define PHASE_MARK(k)                                 \
  do {                                                \
    const uint64_t t_ = __builtin_amdgcn_s_memtime(); \
    ph_.t[k] += t_ - ph_.prev;                        \
    ph_.prev = t_;                                    \
  } while (0)
#define PHASE_FLUSH()                                                           \
  do {                                                                          \
    if (lane() == 0 && p.cycles)                                                \
      for (int k_ = 0; k_ < kPhases; ++k_) atomicAdd(&p.cycles[k_], ph_.t[k_]); \
  } while (0)
#else
#define PHASE_ARG
#define PHASE_PASS
#define PHASE_INIT()
#define PHASE_MARK(k)
#define PHASE_FLUSH()
#endif

// Occupancy of the trace and split kernels: at least N waves per SIMD
// (-DBTG_TRACE_WAVES_PER_EU=N, -DBTG_SPLIT_WAVES_PER_EU=N) caps their
// registers at 512 / N. Two waves per SIMD keep the trace kernel in 224
// registers with nothing spilled; the split kernel spills a little at two
// waves but runs 1.6x faster than with one.
#ifndef BTG_TRACE_WAVES_PER_EU
#define BTG_TRACE_WAVES_PER_EU 2
#endif
#ifndef BTG_SOFT_TRACE_WAVES_PER_EU
#define BTG_SOFT_TRACE_WAVES_PER_EU 3
#endif
#ifndef BTG_SPLIT_WAVES_PER_EU
#define BTG_SPLIT_WAVES_PER_EU 2
#endif
#ifdef BTG_TRACE_WAVES_PER_EU
#define TRACE_OCCUPANCY __attribute__((amdgpu_waves_per_eu(isSoft(App) ? BTG_SOFT_TRACE_WAVES_PER_EU : BTG_TRACE_WAVES_PER_EU)))
#else
#define TRACE_OCCUPANCY
#endif
#ifdef BTG_SPLIT_WAVES_PER_EU
#define SPLIT_OCCUPANCY __attribute__((amdgpu_waves_per_eu(BTG_SPLIT_WAVES_PER_EU)))
#else
#define SPLIT_OCCUPANCY
#endif

// wfAdvance (wavefront_core.h:157-231): descend to a leaf, skip the excluded
// triangle, climb to the next pending far child; returns kContinue with a
// triangle pending at r.cursor, or kFinished after emitting the output.
// The descent records each ancestor's far child in the lane's LDS path, so a
// climb reads the pending child there and fetches its box, its parent and
// the child itself at once, instead of following parent links; the climb
// counter still adds one per level climbed, as the reference's walk does.
// Both children of an inner node are fetched while it is decided.
template <int App>
__device__ int advance(const Params& p, const Ctx& c, Rec& r, Work& w, Stepper& ss, const Path& path,
                       const float* awCol, Cnt& cnt PHASE_ARG) {
  const DevScene& sv = p.sv;
  const Beam& B = w.b;
  for (;;) {
    if (ss.phase == kDescend) {
      KdNodeD kn = ss.kn;
      for (;;) {
        if ((kn.data & 3u) == 3u) break;
        const KdNodeD kl = sv.nodes[r.node + 1], kr = sv.nodes[kn.data >> 2];
        cnt.v[sKd]++;
        const int d = r.depth++;
        const int a = int(kn.data & 3u);
        const float s = kn.split;
        const int sga = sgnOf(B, a);
        const float Ns = float(sga) * relApex(c, s, a);
        const uint32_t left = r.node + 1, right = kn.data >> 2;
        const bool toRight = sga > 0;
        if (d < kPathLevels) path.set(d, toRight ? right : left);
        int dec = kFar;
        if (Ns > 0) dec = decide(B, a, Ns, ss.sd, awCol);
        // kFar: into the child the rays move into; otherwise the apex side
        const bool goRight = (dec == kFar) == toRight;
        if (dec == kBoth) r.trail |= uint64_t(1) << d;
        r.node = goRight ? right : left;
        kn = goRight ? kr : kl;
        // the child's box has the split plane as its entry face on axis a
        // (the far child) or as its exit face (the apex-side child)
        if (dec == kFar)
          setComp(ss.sd.nIn, a, Ns);
        else
          setComp(ss.sd.nOut, a, Ns);
      }
      r.cursor = 0;
      cnt.v[sLeaf]++;
      ss.leafOffset = kn.offset;
      ss.leafCount = kn.data >> 2;
      ss.phase = kTest;
      PHASE_MARK(kPhDescend);
    }
    if (ss.phase == kTest) {
      if (uint32_t(r.cursor) < ss.leafCount) return kContinue;
      ss.phase = kClimb;
    }
    for (;;) {
      if (r.trail == 0) {
        PHASE_MARK(kPhRest);
        emitOutput<App>(p, c, B, r.root, c.anyHit() ? -1 : B.hit, cnt);
        PHASE_MARK(kPhEmit);
        return kFinished;
      }
      const int d = highestBit(r.trail);
      cnt.v[sClimbs] += uint32_t(r.depth - d);
      PHASE_MARK(kPhRest);
      uint32_t cs;
      if (d >= ss.validFrom && d < kPathLevels) {
        cs = path.get(d);
      } else {
        // parent links from the shallowest known ancestor below depth d (the
        // parent of a recorded far child, or r.node) up to depth d
        int depth = r.depth;
        uint32_t node = r.node;
        if (d < ss.validFrom && ss.validFrom < depth && ss.validFrom < kPathLevels) {
          depth = ss.validFrom;
          node = sv.parent[path.get(depth)];
        }
        while (depth > d) {
          node = sv.parent[node];
          --depth;
        }
        const KdNodeD kn = sv.nodes[node];
        cs = sgnOf(B, int(kn.data & 3u)) > 0 ? (kn.data >> 2) : node + 1;
        if (d < kPathLevels) {
          path.set(d, cs);
          ss.validFrom = d;
        }
      }
      PHASE_MARK(kPhWalk);
      r.trail &= (uint64_t(1) << d) - 1;
      const Box cb = sv.nodeBox[cs];
      const uint32_t node = sv.parent[cs];
      const KdNodeD kc = sv.nodes[cs];
      r.node = node;
      r.depth = int16_t(d);
      const SlabDist sd = slabDist(c, B, cb);
      const bool ov = overlaps(B, sd);
      PHASE_MARK(kPhBox);
      if (ov) {
        r.node = cs;
        r.depth = int16_t(d + 1);
        r.cursor = -1;
        ss.phase = kDescend;
        ss.sd = sd;
        ss.kn = kc;
        break;
      }
    }
  }
}

// wfTestOne (wavefront_core.h:236-268). The excluded triangle (the shading
// point's own) is skipped here, uncounted, as the reference skips it before
// testing (wavefront_core.h:198, 276).
template <int App>
__device__ int testOne(const Params& p, const Ctx& c, Rec& r, Work& w, const Stepper& ss, Cnt& cnt) {
  const uint32_t i = uint32_t(r.cursor);
  r.cursor = int32_t(i + 1);
  const LeafTri ref = p.sv.refs[ss.leafOffset + i];
  if (ref.tri == c.exclude) return kContinue;
  TriS ti;
  triBegin(c, ref, ti);
  if (!ti.ok) return kContinue;
  cnt.v[sTests]++;
  const int t = ti.tri;
  if (w.b.hit == t) return kContinue;
  const int res = classifyLazy(c, w, ti);
  if (res == kKeep) return kContinue;
  if (res == kAllIn) {
    if (c.anyHit()) {
      emitOutput<App>(p, c, w.b, r.root, t, cnt);
      return kFinished;
    }
    setStatus(c, w.b, t, ti.A, ti.C);
    w.hitA[0] = ti.A[0];
    w.hitA[1] = ti.A[1];
    w.hitA[2] = ti.A[2];
    w.hitC = ti.C;
    w.hitN = ti.n;
    r.hit = t;
    return kContinue;
  }
  r.cursor = int32_t(i);
  return kSplit;
}

// wfMoreInLeaf (wavefront_core.h:272-279).
__device__ __forceinline__ bool moreInLeaf(const Stepper& ss, const Rec& r) {
  return uint32_t(r.cursor) < ss.leafCount;
}

// Every application tests leaf triangles wave-cooperatively (coopTest); build
// with -DBTG_COOP_TEST=0 for each lane's own sequential loop (testOne).
// Results and counters are the same either way (tested); batch64 runs 1.7%
// faster, a primary pass 18% and a point-shadow pass 13%.
#ifndef BTG_COOP_TEST
#define BTG_COOP_TEST 1
#endif
constexpr bool kCoopTest = BTG_COOP_TEST != 0;

__device__ __forceinline__ F3 shflF3(F3 v, int src) {
  return f3(__shfl(v.x, src, kWave), __shfl(v.y, src, kWave), __shfl(v.z, src, kWave));
}
__device__ __forceinline__ C4 shflC4(const C4& v, int src) {
  C4 r;
#pragma unroll
  for (int i = 0; i < 4; ++i) r.v[i] = __shfl(v.v[i], src, kWave);
  return r;
}

// The leaf tests of the whole wave, dealt 64 at a time. Lanes with `want` own
// a leaf with triangles pending from r.cursor; every (owner, triangle) pair
// becomes one lane's test, run on the owner's beam and query fetched by
// cross-lane reads. For each owner the first triangle in leaf order that is
// not Keep decides, and only the tests up to it are counted, exactly as the
// owner's own sequential loop would. All 64 lanes call it.
// Primary visibility keeps the nearest hit: an AllIn there narrows the beam
// to the triangle's plane (setStatus) and the owner returns kContinue with
// the cursor past it, so the rest of the leaf is tested against the new hit,
// again as the sequential loop does.
// Mailbox (Params::mailbox): each lane keeps, in its column of `mboxLds`
// (kMbox direct-mapped entries, stride kBlock), triangles its current sub-beam
// has been tested against with the result Keep. A Keep leaves the beam
// unchanged, so testing the same triangle again (it is listed in several
// leaves when it straddles a split plane) would give Keep again: such a pair is
// skipped without loading the triangle, counted as a mailbox skip instead of a
// test (as src/beam/beam_core.h:884-912 counts its skips). The column is
// cleared whenever the lane's beam changes: a new record, or (primary) a
// nearer hit. Results are the same with and without it.
template <int App>
__device__ int coopTest(const Params& p, const Ctx& c, Rec& r, Work& w, const Stepper& ss, bool want,
                        Cnt& cnt, int32_t* mboxLds) {
  const int k = want ? int(ss.leafCount) - r.cursor : 0;
  int incl = k;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const int y = __shfl_up(incl, o, kWave);
    if (lane() >= o) incl += y;
  }
  const int total = __shfl(incl, kWave - 1, kWave);
  const int excl = incl - k;
  bool resolved = !(want && k > 0);
  int first = -1, firstTri = -1, counted = 0, skipped = 0;
  bool firstAllIn = false;
  float hA0 = 0, hA1 = 0, hA2 = 0, hC = 0;  // the first AllIn's plane (primary)
  F3 hN = f3(0, 0, 0);
  uint64_t open = __ballot(!resolved);
  for (int base = 0; base < total && open; base += kWave) {
    const int t = base + lane();
    int lo = 0, hi = kWave - 1;
#pragma unroll
    for (int s = 0; s < 6; ++s) {
      const int mid = (lo + hi) >> 1;
      if (__shfl(incl, mid, kWave) > t)
        hi = mid;
      else
        lo = mid + 1 < hi ? mid + 1 : hi;
    }
    const int owner = lo;
    const int oExcl = __shfl(excl, owner, kWave), oCursor = __shfl(r.cursor, owner, kWave);
    const int oIncl = __shfl(incl, owner, kWave);
    int32_t* const ocol = mboxLds + ((int(threadIdx.x) & ~(kWave - 1)) + owner);
    const int oOff = __shfl(int(ss.leafOffset), owner, kWave);
    const bool live = t < total && ((open >> owner) & 1ull);
    Ctx oc;
    oc.P.O = shflF3(c.P.O, owner);
    oc.Olo = App == kSoftG ? shflF3(c.Olo, owner) : f3(0, 0, 0);
    oc.P.u = shflF3(c.P.u, owner);
    oc.P.v = shflF3(c.P.v, owner);
    oc.P.d = shflF3(c.P.d, owner);
    oc.P.n = shflF3(c.P.n, owner);
    oc.P.p0 = f3(0, 0, 0);
    oc.ext = __shfl(c.ext, owner, kWave);
    oc.eps = __shfl(c.eps, owner, kWave);
    oc.minArea = 0;
    oc.farCq = __shfl(c.farCq, owner, kWave);
    oc.flags = App == kPrimary ? __shfl(c.flags, owner, kWave) : c.flags;
    oc.exclude = __shfl(c.exclude, owner, kWave);
    Work ow;
    ow.b.qx = shflC4(w.b.qx, owner);
    ow.b.qy = shflC4(w.b.qy, owner);
    ow.b.n = __shfl(w.b.n, owner, kWave);
    ow.b.hit = -1;
    if (App == kPrimary) {  // the current hit's plane, for the closer-than-hit line
      ow.b.hit = __shfl(w.b.hit, owner, kWave);
#pragma unroll
      for (int k = 0; k < 3; ++k) ow.hitA[k] = __shfl(w.hitA[k], owner, kWave);
      ow.hitC = __shfl(w.hitC, owner, kWave);
      ow.hitN = shflF3(w.hitN, owner);
    }
    int status = 0, tri = -1;  // 0 skipped (uncounted), 1 Keep, 2 AllIn, 3 Clip
    float tA0 = 0, tA1 = 0, tA2 = 0, tC = 0;  // an AllIn triangle's plane (primary)
    F3 tN = f3(0, 0, 0);
    bool boxed = false;
    if (live && p.mailbox) {
      tri = p.sv.refTri[uint32_t(oOff) + uint32_t(oCursor + (t - oExcl))];
      boxed = ocol[(tri & (kMbox - 1)) * kBlock] == tri;
    }
    if (live && !boxed) {
      const uint32_t idx = uint32_t(oCursor + (t - oExcl));
      const LeafTri ref = p.sv.refs[uint32_t(oOff) + idx];
      tri = ref.tri;
      if (ref.tri != oc.exclude) {
        TriS ti;
        triBegin(oc, ref, ti);
        if (ti.ok) {
          if (App == kPrimary && tri == ow.b.hit) {
            status = 1;  // the current hit itself: counted, kept
          } else {
            const int rr = classifyLazy(oc, ow, ti);
            status = rr == kKeep ? 1 : (rr == kAllIn ? 2 : 3);
            if (App == kPrimary && rr == kAllIn) {
              tA0 = ti.A[0];
              tA1 = ti.A[1];
              tA2 = ti.A[2];
              tC = ti.C;
              tN = ti.n;
            }
          }
        }
      }
    }
    const uint64_t hitB = __ballot(status >= 2), cntB = __ballot(status >= 1), allInB = __ballot(status == 2);
    const uint64_t boxB = p.mailbox ? __ballot(boxed) : 0ull;
    if (p.mailbox && status == 1) {  // a Keep before the owner's first other result: the beam passed it unchanged
      const int a = oExcl - base > 0 ? oExcl - base : 0;
      const int b = oIncl - base < kWave ? oIncl - base : kWave;
      const uint64_t below = b == kWave ? ~0ull : ((1ull << b) - 1ull);
      const uint64_t hm = hitB & below & ~((1ull << a) - 1ull);
      if (hm == 0 || lane() < __builtin_ctzll(hm)) ocol[(tri & (kMbox - 1)) * kBlock] = tri;
    }
    int src = lane();
    bool took = false;
    if (!resolved) {
      const int a = excl - base > 0 ? excl - base : 0;
      const int b = incl - base < kWave ? incl - base : kWave;
      if (a < b) {
        const uint64_t below = b == kWave ? ~0ull : ((1ull << b) - 1ull);
        const uint64_t range = below & ~((1ull << a) - 1ull);
        const uint64_t hm = hitB & range;
        uint64_t upto = range;
        if (hm) {
          const int f = __builtin_ctzll(hm);
          upto = range & (f == kWave - 1 ? ~0ull : ((1ull << (f + 1)) - 1ull));
          first = r.cursor + (base + f - excl);
          firstAllIn = ((allInB >> f) & 1ull) != 0;
          src = f;
          took = true;
          resolved = true;
        }
        counted += __popcll(cntB & upto);
        skipped += __popcll(boxB & upto);
      }
      if (!resolved && incl <= base + kWave) resolved = true;  // all of this owner's tests done: Keep
    }
    const int gotTri = __shfl(tri, src, kWave);
    if (took) firstTri = gotTri;
    if (App == kPrimary) {
      const float a0 = __shfl(tA0, src, kWave), a1 = __shfl(tA1, src, kWave), a2 = __shfl(tA2, src, kWave);
      const float cc = __shfl(tC, src, kWave);
      const F3 nn = shflF3(tN, src);
      if (took) {
        hA0 = a0;
        hA1 = a1;
        hA2 = a2;
        hC = cc;
        hN = nn;
      }
    }
    open = __ballot(!resolved);
  }
  if (!want) return kContinue;
  cnt.v[sTests] += uint32_t(counted);
  cnt.v[sMail] += uint32_t(skipped);
  if (first >= 0) {
    if (firstAllIn) {
      if (App == kPrimary) {  // testOne's nearest-hit update
        const float A[3] = {hA0, hA1, hA2};
        setStatus(c, w.b, firstTri, A, hC);
        w.hitA[0] = hA0;
        w.hitA[1] = hA1;
        w.hitA[2] = hA2;
        w.hitC = hC;
        w.hitN = hN;
        r.hit = firstTri;
        r.cursor = first + 1;
        if (p.mailbox) {  // a nearer hit: earlier Keeps were against the farther one
          int32_t* const own = mboxLds + threadIdx.x;
#pragma unroll
          for (int k = 0; k < kMbox; ++k) own[k * kBlock] = -1;
        }
        return kContinue;
      }
      emitOutput<App>(p, c, w.b, r.root, firstTri, cnt);
      return kFinished;
    }
    r.cursor = first;
    return kSplit;
  }
  r.cursor = int(ss.leafCount);
  return kContinue;
}

template <int App>
__global__ void __launch_bounds__(kBlock) TRACE_OCCUPANCY traceKernel(Params p, int cur) {
  Cnt cnt;
  cnt.zero();
  PHASE_INIT();
  const int n = int(min((unsigned long long)p.st->count[cur], p.capacity));
  if (n == 0) return;  // an empty round (a pass draining, or an idle open pass)
  const Rec* q = p.tq[cur];
  const Deal dl = dealOf();
  bool active = false, exhausted = false, first = true;
  int leaves = 0;
  Rec r;
  Work w;
  Ctx c;
  Stepper ss;
  __shared__ uint32_t pathLds[kPathLevels * kBlock];
  __shared__ int32_t mboxLds[kMbox * kBlock];
  __shared__ float awLds[12 * kBlock];
  const Path path{pathLds + threadIdx.x};
  float* const awCol = awLds + threadIdx.x;
  for (;;) {
    const uint64_t idle = __ballot(!active);
    const int nIdle = __popcll(idle);
    if (!exhausted && nIdle > 0) {  // on the first pass every lane is idle: the static chunk of 64
      const int base = nextChunk(dl, &p.st->head[kHeadTrace], nIdle, n, first, exhausted);
      if (!active) {
        const int i = base + rankBelow(idle);
        if (i < n) {
          r = q[i];
          if (r.n == 0 || p.state[r.root] == kKilled) {  // a dropped sliver, or an abandoned root
            liveSub(p, r.root, 1);
          } else {
            c = modeCtx<App>(p.roots[r.root]);
            expand<App>(p, c, r, w);
            awStore(awCol, w.b);
            if (r.cursor < 0) {
              ss.phase = kDescend;
              ss.sd = slabDist(c, w.b, p.sv.nodeBox[r.node]);
              ss.kn = p.sv.nodes[r.node];
            } else {
              const KdNodeD kn = p.sv.nodes[r.node];
              ss.leafOffset = kn.offset;
              ss.leafCount = kn.data >> 2;
              ss.phase = kTest;
            }
            ss.validFrom = r.depth;
            active = true;
            leaves = 0;
            if (p.mailbox) {
#pragma unroll
              for (int k = 0; k < kMbox; ++k) mboxLds[k * kBlock + threadIdx.x] = -1;
            }
          }
        }
      }
    }
    if (!__any(active)) {
      if (exhausted) break;
      continue;
    }
    PHASE_MARK(kPhRefill);
    int res = kContinue;
    if (active) res = advance<App>(p, c, r, w, ss, path, awCol, cnt PHASE_PASS);
    PHASE_MARK(kPhRest);
    if (kCoopTest) {
      const bool want = active && res == kContinue;  // a leaf with triangles pending
      const int cr = coopTest<App>(p, c, r, w, ss, want, cnt, mboxLds);
      if (want) res = cr;
    } else if (active) {
      while (res == kContinue) {
        res = testOne<App>(p, c, r, w, ss, cnt);
        if (res != kContinue || !moreInLeaf(ss, r)) break;
      }
    }
    PHASE_MARK(kPhTest);
    if (active) {
      if (res == kFinished) liveSub(p, r.root, 1);
      if (res != kContinue) active = false;
    }
    // Straggler cap: a sub-beam that has visited leafCap leaves this launch
    // continues in the next round's queue (its record is a full continuation).
    const bool yield = active && p.leafCap > 0 && ++leaves >= p.leafCap;
    const uint64_t yields = __ballot(yield);
    if (yields) {
      int base = 0;
      if (lane() == 0) base = atomicAdd(&p.st->count[cur ^ 1], __popcll(yields));
      base = __shfl(base, 0, kWave);
      if (yield) {
        storeGeom(w.b, r);
        const int slot = base + rankBelow(yields);
        if ((unsigned long long)slot < p.capacity)
          p.tq[cur ^ 1][slot] = r;
        else
          p.st->overflow = 1;
        active = false;
      }
    }
    const uint64_t splits = __ballot(res == kSplit);
    if (splits) {
      int base = 0;
      if (lane() == 0) base = atomicAdd(&p.st->splitCount, __popcll(splits));
      base = __shfl(base, 0, kWave);
      if (res == kSplit) {
        storeGeom(w.b, r);
        p.sq[base + rankBelow(splits)] = r;
      }
    }
    PHASE_MARK(kPhAppend);
  }
  flushCnt(p, cnt);
  PHASE_FLUSH();
}

// ------------------------------------------------------------- the split
// Emitting pass: writes records (or outputs) exactly as wfSplitEmit
// (wavefront_core.h:371-422) into dst[0..c0) for list 0 and dst[c0..) for
// list 1.
template <int App>
struct EmitSink {
  const Params* p;
  const Ctx* c;
  const Beam* parent;
  const Rec* base;
  Rec* dst;
  int k0, k1;  // next record of list 0 / list 1
  int t;
  Cnt* cnt;
  __device__ __forceinline__ void push(int l, const Pg<6>& q0, bool welded) {
    if (q0.n == 5) {
      Pg<6> sub0, sub1;
      splitFive(q0, sub0, sub1);
      cnt->v[sFive]++;
      one(l, sub0, false);
      one(l, sub1, false);
    } else {
      one(l, q0, welded);
    }
  }
  __device__ __forceinline__ void one(int l, const Pg<6>& sub, bool welded) {
    const bool queued = !(l == 1 && c->anyHit());
    const int hit = l == 1 ? t : parent->hit;
    {
      Beam nb;
      float dropped = 0;
      const bool ok = beamGeom(*c, sub, parent->s0, parent->s1, parent->s2, nb, dropped, welded);
      if (!ok) {
        cnt->dropped += double(dropped);
      } else if (!queued) {
        if (isSoft(App)) {  // occluded by t: counted, nothing to accumulate
          cnt->v[sHits]++;
          cnt->v[sBeams]++;
        } else {
          nb.hit = t;
          emitOutput<App>(*p, *c, nb, base->root, t, *cnt);
        }
      }
      if (!queued) return;
      Rec rr = *base;
      if (ok) {
        storeGeom(nb, rr);
        rr.hit = hit;
      } else {
        rr.n = 0;
      }
      if (l == 1)
        dst[k1++] = rr;
      else
        dst[k0++] = rr;
    }
  }
};

__device__ void scheduleNext(const Params& p, int cur);

template <int App>
__global__ void __launch_bounds__(kBlock) SPLIT_OCCUPANCY splitKernel(Params p, int cur) {
  Cnt cnt;
  cnt.zero();
  __shared__ float stageLds[kStageSlots * kStageWords * kBlock];
  float* const stageCol = stageLds + threadIdx.x;
  const int n = int(min((unsigned long long)p.st->splitCount, p.capacity));
  const int nxt = cur ^ 1;
  const Deal dl = dealOf();
  bool first = true, exhausted = false;
  while (n > 0 && !exhausted) {  // an empty round skips to the scheduling below
    const int base = nextChunk(dl, &p.st->head[kHeadSplit], kWave, n, first, exhausted);
    if (base >= n) break;
    const int i = base + lane();
    int slots = 0;
    bool have = false, emit = false, commit = false;
    Rec r;
    Work w;
    Ctx c;
    Ln L[5];
    int nl = 0;
    int recs0 = 0;
    int staged = -1;  // pieces kept in LDS by the counting pass (-1: clip again)
    int tri = -1;
    if (i < n) {
      r = p.sq[i];
      if (p.state[r.root] == kKilled) {
        liveSub(p, r.root, 1);
      } else {
        c = modeCtx<App>(p.roots[r.root]);
        expand<App>(p, c, r, w);
        have = true;
      }
    }
    if (have) {
      const KdNodeD leaf = p.sv.nodes[r.node];
      const LeafTri ref = p.sv.refs[leaf.offset + uint32_t(r.cursor)];
      const TriS ti = triSetup(c, ref);
      tri = ti.tri;
      classifyTri(c, w, ti, L, nl);  // kClip, as found by the trace
      CountSink cs;
      cs.col = stageCol;
      Ln Lc[5];
#pragma unroll
      for (int k = 0; k < 5; ++k) Lc[k] = L[k];
      clipTri(c, w.b, Lc, nl, cs, cnt.v[sFive]);
      commit = cs.hitArea >= c.minArea;
      recs0 = cs.recs0;
      staged = cs.overflow ? -1 : cs.staged;
      slots = commit ? cs.recs0 + (c.anyHit() ? 0 : cs.recs1) : 1;
      const int old = atomicAdd(&p.live[r.root], slots - 1);
      if (old + slots - 1 > p.budget) {
        killRoot(p, r.root);
        liveSub(p, r.root, slots);
        slots = 0;
      } else {
        emit = true;
        if (slots == 0 && old == 1) rootDone(p, r.root);
      }
    }
    const int slot = waveAppend(&p.st->count[nxt], slots);
    if (slots > 0 && (unsigned long long)(slot + slots) > p.capacity) {
      p.st->overflow = 1;
      emit = false;
    }
    if (emit) {
      Rec b = r;
      b.cursor = r.cursor + 1;
      Rec* dst = p.tq[nxt] + slot;
      if (!commit) {
        dst[0] = b;
      } else {
        cnt.v[sSplits]++;
        EmitSink<App> es;
        es.p = &p;
        es.c = &c;
        es.parent = &w.b;
        es.base = &b;
        es.dst = dst;
        es.k0 = 0;
        es.k1 = recs0;
        es.t = tri;
        es.cnt = &cnt;
        if (staged >= 0) {  // replay the counting pass's pieces in their order
          for (int k = 0; k < staged; ++k) {
            const float* rd = stageCol + k * kStageWords * kBlock;
            Pg<6> q;
#pragma unroll
            for (int v = 0; v < 5; ++v) {
              q.x[v] = rd[v * kBlock];
              q.y[v] = rd[(5 + v) * kBlock];
            }
            q.x[5] = q.y[5] = 0;
            const int tag = __float_as_int(rd[10 * kBlock]);
            q.n = tag & 0xff;
            es.push((tag >> 8) & 1, q, ((tag >> 9) & 1) != 0);
          }
        } else {
          uint32_t dummyFive = 0;
          clipTri(c, w.b, L, nl, es, dummyFive);
        }
      }
    }
  }
  flushCnt(p, cnt);
  // The last workgroup to finish schedules the next round: the last of each
  // group of workgroups counts its group done, the last group schedules.
  __syncthreads();
  if (threadIdx.x == 0) {
    const int G = int(gridDim.x), g = int(blockIdx.x) % kXcds;
    const int groups = G < kXcds ? G : kXcds;
    const int members = G / kXcds + (g < G % kXcds ? 1 : 0);
    __threadfence();
    if (atomicAdd(&p.st->groupDone[g * 32], 1) == members - 1) {
      __threadfence();
      if (atomicAdd(&p.st->blocksDone, 1) == groups - 1) {
        __threadfence();
        scheduleNext(p, cur);
      }
    }
  }
}

// scheduleKernel (cuda/wavefront.cu:327-345), run by one thread after a
// round: resets the round's counters and admits roots while their budgets fit.
// Admits the next roots while their budgets fit in the free queue space
// (reserved += n * budget, as the reference's scheduleKernel).
__device__ void admitNext(const Params& p) {
  WState* s = p.st;
  volatile WState* vs = s;
  const unsigned long long reserved = vs->reserved;
  const unsigned long long freeSlots = p.capacity > reserved ? p.capacity - reserved : 0;
  long long nAdmit = (long long)(p.R - vs->next);
  const long long fit = (long long)(freeSlots / (unsigned long long)p.budget);
  if (fit < nAdmit) nAdmit = fit;
  if (nAdmit < 0) nAdmit = 0;
  s->admitBegin = vs->next;
  s->admitCount = int(nAdmit);
  s->next = vs->next + int(nAdmit);
  s->reserved = reserved + (unsigned long long)nAdmit * (unsigned long long)p.budget;
}

__device__ void scheduleNext(const Params& p, int cur) {
  WState* s = p.st;
  const int nxt = cur ^ 1;
  volatile WState* vs = s;
  for (int g = 0; g < kXcds; ++g) s->groupDone[g * 32] = 0;
  if (vs->done) {
    s->blocksDone = 0;
    s->head[0] = s->head[1] = s->head[2] = 0;
    __threadfence();
    return;
  }
  s->traceItems += (unsigned long long)vs->count[cur];
  s->splitItems += (unsigned long long)vs->splitCount;
  s->count[cur] = 0;
  s->splitCount = 0;
  s->shadeCount = 0;
  s->head[0] = s->head[1] = s->head[2] = 0;
  s->blocksDone = 0;
  admitNext(p);
  const int q = vs->count[nxt];
  if (q > vs->maxQueue) s->maxQueue = q;
  s->done = (vs->admitCount == 0 && vs->next >= p.R && q == 0) ? 1 : 0;
  s->round = vs->round + 1;
  __threadfence();
}

// An open pass whose roots all finished is reopened after jobs were
// appended (p.R grew): admission resumes at the first root not yet admitted.
__global__ void reopenKernel(Params p) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  WState* s = p.st;
  if (!s->done) return;
  admitNext(p);
  s->done = s->admitCount == 0 ? 1 : 0;
  __threadfence();
}

__global__ void initKernel(Params p) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    WState* s = p.st;
    s->next = 0;
    s->count[0] = s->count[1] = 0;
    s->splitCount = 0;
    s->shadeCount = 0;
    s->head[0] = s->head[1] = s->head[2] = 0;
    s->done = 0;
    s->killed = 0;
    s->maxQueue = 0;
    s->overflow = 0;
    if (p.pass == 0) s->outCount = 0;  // retry passes append to the first pass's output
    s->blocksDone = 0;
    for (int g = 0; g < kXcds; ++g) s->groupDone[g * 32] = 0;
    s->round = 0;
    s->reserved = 0;
    s->traceItems = s->splitItems = 0;
    long long nAdmit = p.R;
    const long long fit = (long long)(p.capacity / (unsigned long long)p.budget);
    if (fit < nAdmit) nAdmit = fit;
    s->admitBegin = 0;
    s->admitCount = int(nAdmit);
    s->next = int(nAdmit);
    s->reserved = (unsigned long long)nAdmit * (unsigned long long)p.budget;
    s->done = p.R == 0 ? 1 : 0;
  }
}

// ------------------------------------------------------------ the admit
// presplit (beam_core.h:487-546) + wfAdmit (wavefront_core.h:428-471):
// split the root by the lines where a ray-direction component is zero, fan
// each sign region into <= 4-gons, queue the pieces at the kd root.
struct AdmitOut {
  int pushed;
  bool ok;
  uint32_t node0;  // kd root and box of the root's scene
  Box bounds;
};

template <int App>
__device__ bool admitPiece(const Params& p, const Ctx& c, const Pg<8>& qd, const int* sg, int root, int cur,
                           AdmitOut& ao, Cnt& cnt) {
  const uint32_t node0 = ao.node0;
  Beam b;
  float dropped = 0;
  if (!beamGeom(c, qd, sg[0], sg[1], sg[2], b, dropped)) {
    cnt.dropped += double(dropped);
    return true;
  }
  setStatus(c, b, -1, nullptr, 0);
  cnt.v[sRoots]++;
  if (!overlaps(c, b, ao.bounds)) {
    emitOutput<App>(p, c, b, root, -1, cnt);
    return true;
  }
  Rec r;
  storeGeom(b, r);
  r.trail = 0;
  r.node = node0;
  r.root = root;
  r.hit = -1;
  r.cursor = -1;
  r.depth = 0;
  const int slot = atomicAdd(&p.st->count[cur], 1);
  if ((unsigned long long)slot >= p.capacity) {
    p.st->overflow = 1;
    ao.ok = false;
    return false;
  }
  p.tq[cur][slot] = r;
  ao.pushed++;
  return true;
}

// presplit (beam_core.h:487-546): split the root by the three ray-direction
// sign lines (axis 0, 1, 2 in turn), then fan each piece into <= 4-gons. The
// reference walks this depth-3 binary tree with a stack, the inside piece of
// a split before the outside one; here the walk keeps one pending outside
// piece per axis level (a register slot, not a runtime-indexed stack), which
// visits the pieces in the same order.
template <int App>
__device__ void admitPoly(const Params& p, const Ctx& c, const Pg<8>& rootPoly, int root, int cur, AdmitOut& ao,
                          Cnt& cnt) {
  Pg<8> q = rootPoly;
  if (q.n < 3) return;
  if (q.area() < 0) q.reverse();
  const Plane& P = c.P;
  Pg<8> pend0, pend1, pend2;  // outside pieces waiting at axis 0, 1, 2
  uint32_t pending = 0;
  int axis = 0;
  int s0 = 1, s1 = 1, s2 = 1;
  for (;;) {
    bool leafDone = false;
    if (q.n < 3) {
      leafDone = true;
    } else if (axis == 3) {
      // fan into quads (0, k, k+1, k+2) and a last triangle
      for (int k = 1; k + 1 < q.n; k += 2) {
        Pg<8> qd;
        qd.n = 0;
        qd.push(q.x[0], q.y[0]);
        qd.push(q.xAt(k), q.yAt(k));
        qd.push(q.xAt(k + 1), q.yAt(k + 1));
        if (k + 2 < q.n) qd.push(q.xAt(k + 2), q.yAt(k + 2));
        const int sg[3] = {s0, s1, s2};
        if (!admitPiece<App>(p, c, qd, sg, root, cur, ao, cnt)) return;
      }
      leafDone = true;
    } else {
      const Ln L = makeLine(comp(P.u, axis), comp(P.v, axis), comp(P.d, axis), c.ext);
      const int cs = sidesOf(q, L, c.eps);
      if (cs == 3) {
        Pg<8> in, out;
        splitPoly(q, L, c.eps, in, out);
        cleanPoly(in, c.eps);
        cleanPoly(out, c.eps);
        cnt.v[sPresplit]++;
        if (axis == 0) pend0 = out;
        if (axis == 1) pend1 = out;
        if (axis == 2) pend2 = out;
        pending |= 1u << axis;
        q = in;
        if (axis == 0) s0 = 1;
        if (axis == 1) s1 = 1;
        if (axis == 2) s2 = 1;
        ++axis;
      } else if (cs != 0) {
        const int sv = (cs & 1) ? 1 : -1;
        if (axis == 0) s0 = sv;
        if (axis == 1) s1 = sv;
        if (axis == 2) s2 = sv;
        ++axis;
      } else {
        cnt.dropped += double(selAbs(q.area()));
        leafDone = true;
      }
    }
    if (leafDone) {
      if (!pending) return;
      const int lv = 31 - __builtin_clz(pending);  // the deepest pending outside piece
      pending &= ~(1u << lv);
      if (lv == 0) q = pend0;
      if (lv == 1) q = pend1;
      if (lv == 2) q = pend2;
      if (lv == 0) s0 = -1;
      if (lv == 1) s1 = -1;
      if (lv == 2) s2 = -1;
      axis = lv + 1;
    }
  }
}

// lightBeamSetup (soft_core.h:38-73). In the guarded mode the hit point is
// the float pair x + xlo (the double hit point), the apex is lifted off it by
// the double build's offset in double and held as the float pair O + Olo,
// and the horizon line passes through the double hit point.
template <int App>
__device__ bool lightSetup(const SoftJob& p, F3 x, F3 xlo, F3 ns, int tri, RootQ& q, Pg<8>& root) {
  const F3 nL = normalize(cross(p.lU, p.lV));
  if (!(dot(x - p.lc, nL) > 0)) return false;
  float hzc;
  if (App == kSoftG) {
    const double xd = double(x.x) + double(xlo.x), yd = double(x.y) + double(xlo.y), zd = double(x.z) + double(xlo.z);
    const double off = double(p.offset);
    const double ax = xd + double(ns.x) * off, ay = yd + double(ns.y) * off, az = zd + double(ns.z) * off;
    const F3 O = f3(float(ax), float(ay), float(az));
    q.Olo = f3(float(ax - double(O.x)), float(ay - double(O.y)), float(az - double(O.z)));
    q.plane = Plane::make(O, p.lc, p.lU, p.lV, q.Olo);
    hzc = float((double(ns.x) * (double(p.lc.x) - xd) + double(ns.y) * (double(p.lc.y) - yd)) +
                double(ns.z) * (double(p.lc.z) - zd));
  } else {
    const F3 apex = x + ns * p.offset;
    q.plane = Plane::make(apex, p.lc, p.lU, p.lV);
    hzc = dot(ns, p.lc - x);
  }
  q.flags = fAnyHit | fFarIsPlane;
  q.exclude = tri;
  const F3 hu = p.lU * 0.5f, hv = p.lV * 0.5f;
  const F3 corners[4] = {p.lc - hu - hv, p.lc + hu - hv, p.lc + hu + hv, p.lc - hu + hv};
  Pg<8> quad;
  quad.n = 0;
  for (int i = 0; i < 4; ++i) {
    float qx, qy;
    q.plane.coords(corners[i], qx, qy);
    quad.push(qx, qy);
  }
  if (quad.area() < 0) quad.reverse();
  q.lightArea = quad.area();
  const Ln hz{dot(ns, q.plane.u), dot(ns, q.plane.v), hzc};
  int above = 0;
  for (int i = 0; i < quad.n; ++i) above += hz.at(quad.x[i], quad.y[i]) > 0;
  if (above == 0) return false;
  if (above == quad.n) {
    root = quad;
  } else {
    Pg<8> rest;
    splitPoly(quad, hz, 0.0f, root, rest);
  }
  return root.n >= 3 && root.area() > 0;
}

// lightCenterG (soft_core.h:76-84).
__device__ float centerG(const SoftJob& p, F3 x, F3 ns) {
  const F3 nL = normalize(cross(p.lU, p.lV));
  if (!(dot(x - p.lc, nL) > 0)) return 0;
  const F3 w = p.lc - x;
  const float r2 = dot(w, w);
  const F3 wn = w / __builtin_sqrtf(r2);
  const float ci = dot(ns, wn), cl = -dot(nL, wn);
  return (ci > 0 ? ci : 0) * (cl > 0 ? cl : 0) / r2;
}

// pointShadowSetup (point_core.h:15-36).
__device__ bool pointSetup(const Params& p, F3 nT, F3 P0, int T, const float* px, const float* py, int n, RootQ& q,
                           Pg<8>& root) {
  const float sideCam = dot(nT, p.eye - P0), sideL = dot(nT, p.pointLight - P0);
  if (!(sideCam * sideL > 0)) return false;
  const F3 ax = __builtin_fabsf(nT.x) < 0.6f ? f3(1, 0, 0) : (__builtin_fabsf(nT.y) < 0.6f ? f3(0, 1, 0) : f3(0, 0, 1));
  const F3 tu = normalize(cross(nT, ax));
  q.plane = Plane::make(p.pointLight, P0, tu, cross(nT, tu));
  q.flags = fAnyHit | fFarIsPlane;
  q.exclude = T;
  root.n = 0;
  for (int i = 0; i < n; ++i) {
    const F3 d = p.fwd + p.right * px[i] + p.down * py[i];
    const F3 X = p.eye + d * (dot(nT, P0 - p.eye) / dot(nT, d));
    float qx, qy;
    q.plane.coords(X, qx, qy);
    root.push(qx, qy);
  }
  return true;
}

// The job of batch root k (jobs are ordered by rootBase).
__device__ int jobOfRoot(const Params& p, int k) {
  int lo = 0, hi = p.numJobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (p.jobs[mid].rootBase <= k)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// Soft-shadow roots are pixels in tiles of 64 (kRootTileW x kRootTileH): a wave's 64 lanes trace one
// tile. Root k of job J -> its pixel in the batch's per-pixel arrays.
__device__ __forceinline__ bool softPixel(const SoftJob& J, int k, int& pix) {
  const int kl = k - J.rootBase;
  const int t = kl >> 6, l = kl & 63;
  const int x = (t % J.tilesX) * kRootTileW + l % kRootTileW, y = (t / J.tilesX) * kRootTileH + l / kRootTileW;
  pix = J.pixBase + y * J.W + x;
  return x < J.W && y < J.H;
}

template <int App>
__device__ void admitRoot(const Params& p, int k, int cur, Cnt& cnt) {
  p.live[k] = 0;
  p.state[k] = kNoBeam;
  p.epoch[k] = p.pass;
  RootQ R;
  R.flags = 0;
  R.exclude = -1;
  R.unitExp = 0;
  R.tri = -1;
  R.lightArea = 0;
  R.x = R.ns = R.Olo = R.xlo = f3(0, 0, 0);
  Pg<8> poly;
  poly.n = 0;
  uint32_t node0 = p.node0;
  Box bounds = p.rootBox;
  if (isSoft(App)) {
    p.missFx[k] = 0;
    p.gFx[k] = 0;
    const SoftJob& J = p.jobs[jobOfRoot(p, k)];
    node0 = J.node0;
    bounds = J.bounds;
    int pix;
    if (!softPixel(J, k, pix)) {
      rootDone(p, k);
      return;
    }
    const int T = p.pixTri[pix];
    if (T < 0) {
      p.vis[pix] = -1;
      p.E[pix] = 0;
      p.Eu[pix] = 0;
      rootDone(p, k);
      return;
    }
    const F3 x = f3(p.pixX[3 * pix], p.pixX[3 * pix + 1], p.pixX[3 * pix + 2]);
    const F3 ns = f3(p.pixN[3 * pix], p.pixN[3 * pix + 1], p.pixN[3 * pix + 2]);
    const F3 xlo = App == kSoftG ? f3(p.pixXlo[3 * pix], p.pixXlo[3 * pix + 1], p.pixXlo[3 * pix + 2]) : f3(0, 0, 0);
    p.Eu[pix] = length(cross(J.lU, J.lV)) * centerG(J, x, ns);
    if (!lightSetup<App>(J, x, xlo, ns, T, R, poly)) {
      p.vis[pix] = 0;
      p.E[pix] = 0;
      rootDone(p, k);
      return;
    }
    R.x = x;
    R.xlo = xlo;
    R.ns = ns;
  } else if (App == kPrimary) {
    const int tx = k % p.tilesW, ty = k / p.tilesW;
    const int x0 = tx * p.tile, y0 = ty * p.tile;
    const int x1 = min(p.W, x0 + p.tile), y1 = min(p.H, y0 + p.tile);
    const float qx0 = (2 * float(x0) / float(p.W) - 1) * p.halfW, qx1 = (2 * float(x1) / float(p.W) - 1) * p.halfW;
    const float qy0 = (2 * float(y0) / float(p.H) - 1) * p.halfH, qy1 = (2 * float(y1) / float(p.H) - 1) * p.halfH;
    poly.push(qx0, qy0);
    poly.push(qx1, qy0);
    poly.push(qx1, qy1);
    poly.push(qx0, qy1);
    R.plane = p.camPlane;
    R.flags = p.cull ? fCull : 0;
  } else {
    const OutPoly pb = p.prim[k];
    const int T = pb.tri;
    R.tri = T;
    const float4 nn = p.sv.triNV[2 * size_t(T)], vv = p.sv.triNV[2 * size_t(T) + 1];
    if (!pointSetup(p, f3(nn.x, nn.y, nn.z), f3(vv.x, vv.y, vv.z), T, pb.x, pb.y, pb.n, R, poly)) {
      appendPoly(p, k, pb);  // the receiver faces away from the light: all in shadow
      rootDone(p, k);
      return;
    }
  }
  setupQuery(R, poly.extent());
  if (isSoft(App)) {
    // Unit of the exact area sum: ulp(minArea); every output area is a
    // multiple of it (areas >= minArea).
    int e = 0;
    (void)__builtin_frexpf(R.minArea, &e);
    R.unitExp = e - 24 < -149 ? -149 : e - 24;
  }
  p.roots[k] = R;
  p.state[k] = kTraced;
  const Ctx c = modeCtx<App>(R);
  AdmitOut ao{0, true, node0, bounds};
  admitPoly<App>(p, c, poly, k, cur, ao, cnt);
  p.live[k] = ao.pushed;
  if (!ao.ok || ao.pushed > p.budget) killRoot(p, k);
  if (ao.pushed == 0) rootDone(p, k);
}

template <int App>
__global__ void __launch_bounds__(kBlock) admitKernel(Params p, int cur) {
  Cnt cnt;
  cnt.zero();
  const int begin = p.st->admitBegin, n = p.st->admitCount;
  if (n == 0) return;  // most rounds admit nothing: no queue atomics, no counter flush
  const Deal dl = dealOf();
  bool first = true, exhausted = false;
  while (!exhausted) {
    const int base = nextChunk(dl, &p.st->head[kHeadAdmit], kWave, n, first, exhausted);
    if (base >= n) break;
    const int i = base + lane();
    if (i < n) {
      const int slot = begin + i;
      admitRoot<App>(p, p.rootList ? p.rootList[slot] : slot, cur, cnt);
    }
  }
  flushCnt(p, cnt);
}

// The exact Lambert integral of the unoccluded pieces queued this round
// (polygonFormFactorG, soft_core.h:18-33), summed per root in 2^-60 fixed
// point. Kept out of the trace kernel so that its double-precision code does
// not raise the trace kernel's register count.
__global__ void __launch_bounds__(kBlock) shadeKernel(Params p) {
  const int n = int(min((unsigned long long)p.st->shadeCount, p.capacity));
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const ShadeRec sr = p.shq[i];
    const RootQ& R = p.roots[sr.root];
    double G;
    if (p.pixXlo) {  // guarded: in double throughout
      const Plane& pl = R.plane;
      D3 P[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const double qx = sr.qx[k], qy = sr.qy[k];
        P[k] = D3{(double(pl.p0.x) + double(pl.u.x) * qx) + double(pl.v.x) * qy,
                  (double(pl.p0.y) + double(pl.u.y) * qx) + double(pl.v.y) * qy,
                  (double(pl.p0.z) + double(pl.u.z) * qx) + double(pl.v.z) * qy};
      }
      const D3 xd{double(R.x.x) + double(R.xlo.x), double(R.x.y) + double(R.xlo.y), double(R.x.z) + double(R.xlo.z)};
      G = formFactorD(xd, D3{R.ns.x, R.ns.y, R.ns.z}, P, sr.n);
    } else {
      F3 P[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) P[k] = R.plane.point(sr.qx[k], sr.qy[k]);
      G = double(formFactor(R.x, R.ns, P, sr.n));
    }
    atomicAdd(&p.gFx[sr.root], (unsigned long long)rint(G * pow2(60)));
  }
}

// finalizeSoftKernel (cuda/wavefront.cu:539-561) with the exact sums:
// vis = float(missArea) / lightArea as the CPU computes it
// (src/render/render.cpp:291).
// Writes root k's pixel; returns its state.
__device__ __forceinline__ int finalizeRoot(const Params& p, int k) {
  int pix;
  if (!softPixel(p.jobs[jobOfRoot(p, k)], k, pix)) return kNoBeam;
  const int s = p.state[k];
  if (s != kTraced) return s;  // no beam (written at admission) or killed (traced again)
  const RootQ& R = p.roots[k];
  const double miss = double(p.missFx[k]) * pow2(R.unitExp);
  const float V0 = float(miss) / R.lightArea;
  const float m = V0 < 1.0f ? V0 : 1.0f;  // std::max(0, std::min(1, V)) of render.cpp:302
  const float V = 0.0f < m ? m : 0.0f;
  p.vis[pix] = V;
  p.E[pix] = p.exact ? float(double(p.gFx[k]) * pow2(-60)) : p.Eu[pix] * V;
  return s;
}

__global__ void finalizeSoftKernel(Params p) {
  const int i = int(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= p.R) return;
  finalizeRoot(p, p.rootList ? p.rootList[i] : i);
}

// finalizeSoftKernel over the roots [begin, begin + count) of one job of an
// open pass, plus the number of them that were abandoned (over budget).
__global__ void finalizeRangeKernel(Params p, int begin, int count, int* killed) {
  const int i = int(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= count) return;
  if (finalizeRoot(p, begin + i) == kKilled) atomicAdd(killed, 1);
}

__global__ void clearInt(int* x) {
  if (threadIdx.x == 0) *x = 0;
}

// Killed roots, compacted for a retry pass.
__global__ void collectKilled(Params p, int* list, int* count) {
  const int i = int(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= p.R) return;
  const int k = p.rootList ? p.rootList[i] : i;
  if (p.state[k] == kKilled) list[atomicAdd(count, 1)] = k;
}

// ------------------------------------------------------------ primary rays
// rayHitTri (src/ray/ray_core.h:22-37), Moller-Trumbore on a leaf reference.
__device__ __forceinline__ float rayTri(const LeafTri& t, F3 o, F3 d) {
  const F3 v0 = f3(t.v[0], t.v[1], t.v[2]);
  const F3 e1 = f3(t.v[3], t.v[4], t.v[5]) - v0, e2 = f3(t.v[6], t.v[7], t.v[8]) - v0;
  const F3 pv = cross(d, e2);
  const float det = dot(e1, pv);
  if (det == 0) return __builtin_inff();
  const float inv = 1 / det;
  const F3 tv = o - v0;
  const float u = dot(tv, pv) * inv;
  if (u < 0 || u > 1) return __builtin_inff();
  const F3 qv = cross(tv, e1);
  const float v = dot(d, qv) * inv;
  if (v < 0 || u + v > 1) return __builtin_inff();
  return dot(e2, qv) * inv;
}

// rayTraverse<false> (ray_core.h:42-132): nearest hit in (0, inf), coplanar
// ties to the lower id within 4 ulp.
__device__ bool rayNearest(const DevScene& sv, uint32_t node0, const Box& bounds, F3 o, F3 d, int& hitTri, float& hitT,
                           uint32_t& hitRef, Cnt& cnt) {
  cnt.v[sRays]++;
  float t0 = 0, t1 = __builtin_inff();
  float inv[3];
  const float oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
  const float blo[3] = {bounds.lo.x, bounds.lo.y, bounds.lo.z};
  const float bhi[3] = {bounds.hi.x, bounds.hi.y, bounds.hi.z};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    inv[a] = 1 / dd[a];
    if (dd[a] == 0) {
      if (oo[a] < blo[a] || oo[a] > bhi[a]) return false;
      continue;
    }
    float ta = (blo[a] - oo[a]) * inv[a];
    float tb = (bhi[a] - oo[a]) * inv[a];
    if (ta > tb) {
      const float t = ta;
      ta = tb;
      tb = t;
    }
    t0 = selMax(t0, ta);
    t1 = selMin(t1, tb);
  }
  if (t0 > t1) return false;
  uint32_t stNode[64];
  float stT0[64], stT1[64];
  int sp = 0;
  uint32_t node = node0;
  float best = __builtin_inff();
  int bestTri = -1;
  uint32_t bestRef = 0;
  for (;;) {
    const KdNodeD n = sv.nodes[node];
    if ((n.data & 3u) != 3u) {
      cnt.v[sKd]++;
      const int a = int(n.data & 3u);
      const float oa = a == 0 ? o.x : (a == 1 ? o.y : o.z);
      const float da = a == 0 ? d.x : (a == 1 ? d.y : d.z);
      const float ia = a == 0 ? inv[0] : (a == 1 ? inv[1] : inv[2]);
      const bool belowFirst = oa < n.split || (oa == n.split && da <= 0);
      const uint32_t first = belowFirst ? node + 1 : (n.data >> 2);
      const uint32_t second = belowFirst ? (n.data >> 2) : node + 1;
      if (da == 0) {
        node = first;
        continue;
      }
      const float tp = (n.split - oa) * ia;
      if (tp > t1 || tp <= 0) {
        node = first;
      } else if (tp < t0) {
        node = second;
      } else {
        if (sp < 64) {
          stNode[sp] = second;
          stT0[sp] = tp;
          stT1[sp] = t1;
          ++sp;
        }
        node = first;
        t1 = tp;
      }
      continue;
    }
    cnt.v[sLeaf]++;
    const uint32_t count = n.data >> 2;
    for (uint32_t i = 0; i < count; ++i) {
      const LeafTri ref = sv.refs[n.offset + i];
      cnt.v[sTests]++;
      const float th = rayTri(ref, o, d);
      if (!(th > 0 && th < __builtin_inff())) continue;
      const float tol = bestTri < 0 ? 0.0f : best * (4 * 1.1920928955078125e-7f);
      if (bestTri < 0 || th < best - tol || (th <= best + tol && ref.tri < bestTri)) {
        best = th;
        bestTri = ref.tri;
        bestRef = n.offset + i;
      }
    }
    if (bestTri >= 0 && best <= t1) break;
    if (sp == 0) break;
    --sp;
    node = stNode[sp];
    t0 = stT0[sp];
    t1 = stT1[sp];
  }
  if (bestTri < 0) return false;
  cnt.v[sHits]++;
  hitTri = bestTri;
  hitT = best;
  hitRef = bestRef;
  return true;
}

// The hit point in double on the plane of the triangle's float corners, as
// the float pair X + Xlo: o + d t with n = (v1 - v0) x (v2 - v0) and
// t = n.(v0 - o) / n.d, every step in double (the guarded modes' receiver
// point; the float pipeline's o + d t can sit ~1e-7 of the scene off that
// plane, far more than the double build's apex offset). o, d: the ray's
// origin and direction in double (the float ray's, or the FP64 mode's ray).
__device__ void hitPointD(const LeafTri& t, double ox, double oy, double oz, double dx, double dy, double dz, F3& X,
                          F3& Xlo) {
  const double v0x = t.v[0], v0y = t.v[1], v0z = t.v[2];
  const double ax = double(t.v[3]) - v0x, ay = double(t.v[4]) - v0y, az = double(t.v[5]) - v0z;
  const double bx = double(t.v[6]) - v0x, by = double(t.v[7]) - v0y, bz = double(t.v[8]) - v0z;
  const double nx = ay * bz - az * by, ny = az * bx - ax * bz, nz = ax * by - ay * bx;
  const double den = (nx * dx + ny * dy) + nz * dz;
  const double num = (nx * (v0x - ox) + ny * (v0y - oy)) + nz * (v0z - oz);
  const double th = den != 0 ? num / den : 0.0;
  const double px = ox + dx * th, py = oy + dy * th, pz = oz + dz * th;
  X = f3(float(px), float(py), float(pz));
  Xlo = f3(float(px - double(X.x)), float(py - double(X.y)), float(pz - double(X.z)));
}

// The job of a batch pixel: the last job whose first pixel is <= k.
__device__ int jobOfPixel(const Params& p, int k) {
  int lo = 0, hi = p.numJobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (p.jobs[mid].pixBase <= k)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// The FP64 query of a pixel with a primary hit: its double hit point (the
// float pair), the hit's leaf reference and the normal facing the eye.
__device__ void fp64RootOf(const Params& p, const SoftJob& J, int pix, fp64::PixelRoot& R) {
  using fp64::V;
  const V x = fp64::vmk(double(p.pixX[3 * pix]) + double(p.pixXlo[3 * pix]),
                        double(p.pixX[3 * pix + 1]) + double(p.pixXlo[3 * pix + 1]),
                        double(p.pixX[3 * pix + 2]) + double(p.pixXlo[3 * pix + 2]));
  const LeafTri ref = p.sv.refs[uint32_t(p.pixRef[pix])];
  const V nf = fp64::vmk(p.pixN[3 * pix], p.pixN[3 * pix + 1], p.pixN[3 * pix + 2]);
  fp64::pixelRoot(ref, x, nf, fp64::vf(J.lc), fp64::vf(J.lU), fp64::vf(J.lV), double(J.offset), R);
}

// Waves per SIMD the FP64 kernels are built for (build knob BTG_FP64_WAVES;
// 2 and 4 measured slower than 1: the register spills cost more than the
// extra waves hide).
#ifndef BTG_FP64_WAVES
#define BTG_FP64_WAVES 1
#endif
constexpr int kFp64Waves = BTG_FP64_WAVES;

// Triangle tests and kd steps a level-0 or level-1 cell may take before it is
// split into finer cells (build knob BTG_FP64_WORK_CAP): bounds the time one
// lane spends on one task, which sets the pass's tail.
#ifndef BTG_FP64_WORK_CAP
#define BTG_FP64_WORK_CAP 2048
#endif
constexpr uint32_t kFp64WorkCap = BTG_FP64_WORK_CAP;
// Beams and work items of a level-1 cell's pools (build knob BTG_FP64_POOL1).
#ifndef BTG_FP64_POOL1
#define BTG_FP64_POOL1 128
#endif
constexpr int kFp64Pool1 = BTG_FP64_POOL1;

__global__ void zeroWords(unsigned long long* w, size_t n) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) w[i] = 0;
}

// One level of FP64 soft shadows. Tasks: level 0, the pixels 0..numPix-1;
// level 1, 16 per entry of `in` (a pixel: the cells of its 4 x 4 grid);
// level 2, 16 per entry (pixel << 4 | a level-1 cell: the 16 x 16 grid's
// cells inside it). Each lane takes tasks from `next` until none is left,
// tracing with pools of Pn beams and Kn work items in its region of `mem`
// (Tracer::kWords words per thread, interleaved across the block's 64
// threads). A finished cell adds its results to the pixel's fixed-point sums;
// a cell over the work cap or the pools goes to `out` for the next level (on
// level 2, which has no cap, it counts in *failed, as does a full `out`).
template <int Pn, int Kn, int Level>
__global__ void __launch_bounds__(64, kFp64Waves)
    fp64CellKernel(Params p, int numPix, const long long* in, const int* inCount, int inCap, int* next, uint64_t* mem,
                   long long* out, int* outCount, int outCap, unsigned long long* visSum, unsigned long long* gSum,
                   int* failed) {
  using TR = fp64::Tracer<Pn, Kn>;
  TR tr;
  tr.cnt.zero();
  tr.m = fp64::WordMem{mem + size_t(blockIdx.x) * size_t(TR::kWords) * 64, 64, int(threadIdx.x)};
  const int tasks = Level == 0 ? numPix : 16 * min(*inCount, inCap);
  for (;;) {
    const int t = atomicAdd(next, 1);
    if (t >= tasks) break;
    int pix = t, cx = 0, cy = 0;
    if (Level > 0) {
      const long long e = in[t >> 4];
      const int s = t & 15;
      if (Level == 1) {
        pix = int(e);
        cx = s & 3;
        cy = s >> 2;
      } else {
        pix = int(e >> 4);
        const int c1 = int(e & 15);
        cx = (c1 & 3) * 4 + (s & 3);
        cy = (c1 >> 2) * 4 + (s >> 2);
      }
    }
    const int T = p.pixTri[pix];
    if (T < 0) continue;
    const SoftJob& J = p.jobs[jobOfPixel(p, pix)];
    fp64::PixelRoot R;
    fp64RootOf(p, J, pix, R);
    if (R.root.n < 3) continue;
    double miss, G;
    if (fp64::traceCell(tr, p.sv, R, T, p.exact != 0, Level, cx, cy, Level == 2 ? 0xffffffffu : kFp64WorkCap,
                        J.node0, J.bounds, miss, G)) {
      atomicAdd(&visSum[pix], fp64::visFx(miss / R.lightArea));
      if (p.exact) atomicAdd(&gSum[pix], fp64::gFx(G));
    } else if (Level < 2) {
      const int k = atomicAdd(outCount, 1);
      if (k < outCap)
        out[k] = Level == 0 ? (long long)pix : (((long long)pix << 4) | (long long)(cy * 4 + cx));
      else
        atomicAdd(failed, 1);
    } else {
      atomicAdd(failed, 1);
    }
  }
  flushCnt(p, tr.cnt);
}

// The FP64 images from the pixels' sums: vis = visible fraction (-1 without
// a primary hit), E = the exact Lambert integral (or A * G0 * vis), Eu = A * G0.
__global__ void __launch_bounds__(256) fp64FinishKernel(Params p, int numPix, const unsigned long long* visSum,
                                                        const unsigned long long* gSum) {
  const int pix = int(blockIdx.x) * 256 + int(threadIdx.x);
  if (pix >= numPix) return;
  if (p.pixTri[pix] < 0) {
    p.vis[pix] = -1;
    p.E[pix] = 0;
    p.Eu[pix] = 0;
    return;
  }
  fp64::PixelRoot R;
  fp64RootOf(p, p.jobs[jobOfPixel(p, pix)], pix, R);
  p.Eu[pix] = float(R.A * R.G0);
  if (R.root.n < 3) {
    p.vis[pix] = 0;
    p.E[pix] = 0;
    return;
  }
  const double vis = fp64::visOfFx(visSum[pix]);
  p.vis[pix] = float(vis);
  p.E[pix] = p.exact ? float(fp64::gOfFx(gSum[pix])) : float(R.A * R.G0 * vis);
}

// The soft-shadow primary pass (cuda/gpu_render.cu:222-244): one ray through
// each pixel centre; X = o + d t, N = the triangle normal facing the eye.
__global__ void __launch_bounds__(kRayBlock) primaryRayKernel(Params p) {
  Cnt cnt;
  cnt.zero();
  // block -> (job, 16 x 16 tile): the job with the largest tileBase <= block
  const int b = int(blockIdx.x);
  int lo = 0, hi = p.numJobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (p.jobs[mid].tileBase <= b)
      lo = mid;
    else
      hi = mid - 1;
  }
  const SoftJob& J = p.jobs[lo];
  const int t = b - J.tileBase, tx = (J.W + 15) / 16;
  const int x = (t % tx) * 16 + int(threadIdx.x & 15), y = (t / tx) * 16 + int(threadIdx.x >> 4);
  if (x < J.W && y < J.H) {
    const size_t k = size_t(J.pixBase) + size_t(y) * size_t(J.W) + size_t(x);
    const float qx = (2 * (float(x) + 0.5f) / float(J.W) - 1) * J.halfW;
    const float qy = (2 * (float(y) + 0.5f) / float(J.H) - 1) * J.halfH;
    const F3 d = J.fwd + J.right * qx + J.down * qy;
    int th_tri = -1;
    float th = 0;
    uint32_t th_ref = 0;
    int* triOut = const_cast<int*>(p.pixTri);
    float* X = const_cast<float*>(p.pixX);
    float* N = const_cast<float*>(p.pixN);
    triOut[k] = -1;
    if (p.pixRef) p.pixRef[k] = -1;
    if (rayNearest(p.sv, J.node0, J.bounds, J.eye, d, th_tri, th, th_ref, cnt)) {
      triOut[k] = th_tri;
      if (p.pixRef) p.pixRef[k] = int(th_ref);
      F3 hp = J.eye + d * th;
      const float4 nn = p.sv.triNV[2 * size_t(th_tri)];
      F3 nrm = f3(nn.x, nn.y, nn.z);
      bool away = dot(nrm, d) > 0;
      if (p.pixXlo) {
        // FP64 mode (pixRef set): the ray the double build shoots from its
        // double camera, the pixel centre and direction in double (the float
        // ray moves the receiver point by ~1e-7 of the distance, which shows
        // as up to ~2e-5 of visible fraction at sharp penumbrae)
        double ox = J.eye.x, oy = J.eye.y, oz = J.eye.z, dx = d.x, dy = d.y, dz = d.z;
        if (p.pixRef) {
          const double* c = p.jobCamD + size_t(lo) * 14;
          const double qxd = (2 * (double(x) + 0.5) / double(J.W) - 1) * c[12];
          const double qyd = (2 * (double(y) + 0.5) / double(J.H) - 1) * c[13];
          ox = c[0];
          oy = c[1];
          oz = c[2];
          dx = (c[3] + c[6] * qxd) + c[9] * qyd;
          dy = (c[4] + c[7] * qxd) + c[10] * qyd;
          dz = (c[5] + c[8] * qxd) + c[11] * qyd;
          away = (double(nn.x) * dx + double(nn.y) * dy) + double(nn.z) * dz > 0;
        }
        F3 xlo;
        hitPointD(p.sv.refs[th_ref], ox, oy, oz, dx, dy, dz, hp, xlo);
        float* XL = const_cast<float*>(p.pixXlo);
        XL[3 * k] = xlo.x;
        XL[3 * k + 1] = xlo.y;
        XL[3 * k + 2] = xlo.z;
      }
      if (away) nrm = -nrm;
      X[3 * k] = hp.x;
      X[3 * k + 1] = hp.y;
      X[3 * k + 2] = hp.z;
      N[3 * k] = nrm.x;
      N[3 * k + 1] = nrm.y;
      N[3 * k + 2] = nrm.z;
    }
  }
  flushCnt(p, cnt);
}

}  // namespace

// ======================================================================
// Host side
// ======================================================================

#define BTG_TRY(x)                                                                        \
  do {                                                                                    \
    hipError_t e_ = (x);                                                                  \
    if (e_ != hipSuccess) {                                                               \
      err = std::string(#x) + ": " + hipGetErrorString(e_);                               \
      return BTG_ERR_DEVICE_CODE;                                                         \
    }                                                                                     \
  } while (0)

constexpr int BTG_ERR_DEVICE_CODE = -2;
constexpr int BTG_ERR_MEMORY_CODE = -3;
constexpr int BTG_ERR_LIMIT_CODE = -4;
constexpr int BTG_ERR_ARG_CODE = -1;

struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
};

// A resident scene in device format, kept on the host so that the arena can
// be rebuilt when scenes are added or removed.
struct SceneRec {
  bool live = false;
  std::vector<float4> triNV;
  std::vector<KdNodeD> nodes;
  std::vector<Box> boxes;
  std::vector<uint32_t> parent;
  std::vector<LeafTri> refs;
  Box bounds;
  int numTris = 0;
  uint32_t numVerts = 0;
  float offset = 0, offsetGuard = 0;
  uint32_t nodeBase = 0, refBase = 0, triBase = 0, vertBase = 0;  // placement in the arena
};

class Engine {
 public:
  DevScene sv;  // the arena: every resident scene, concatenated
  std::vector<SceneRec> scenes;
  Options opt;
  hipStream_t stream = nullptr;
  hipEvent_t ev[4] = {};
  int cus = 0;
  // scene storage
  void* sceneMem[6] = {};
  // engine buffers (grown on demand)
  DevBuf tq0, tq1, sq, roots, live, state, epoch, missFx, gFx, stats, st, out, outRoot, outTag, list, listCount;
  DevBuf pixTri, pixX, pixXlo, pixN, vis, E, Eu, prim, retry, cycles, shq, jobs;
  // FP64 soft shadows: hit leaf reference per pixel; counters and cell lists;
  // the tracers' arrays; the jobs' double cameras; per-pixel fixed-point sums
  DevBuf pixRef, fpList, fpMem, camD, fpSum;
  DevBuf rasterImg, rasterRows;  // device rasterization: image; row counts, span offsets and tile sums
  void* noAllocScratch = nullptr;  // this device's static rasterization scratch (null: unavailable)
  WState* hostState = nullptr;  // pinned
  uint64_t phaseCycles[8] = {};
  // per-launch timing (Options::profile): event pairs and the kernel kind
  std::vector<hipEvent_t> evPool;
  std::vector<int> evKind;
  size_t evUsed = 0;
  int beginLaunch(int kind) {
    if (!opt.profile) return -1;
    while (evPool.size() < 2 * (evUsed + 1)) {
      hipEvent_t a;
      if (hipEventCreate(&a) != hipSuccess) return -1;
      evPool.push_back(a);
    }
    (void)hipEventRecord(evPool[2 * evUsed], stream);
    evKind.resize(evUsed + 1);
    evKind[evUsed] = kind;
    return int(evUsed++);
  }
  void endLaunch(int slot) {
    if (slot >= 0) (void)hipEventRecord(evPool[2 * size_t(slot) + 1], stream);
  }
  // Sums the recorded launch durations into info (after the stream synced).
  void collectLaunches(RunInfo& info) {
    for (size_t i = 0; i < evUsed; ++i) {
      float ms = 0;
      if (hipEventElapsedTime(&ms, evPool[2 * i], evPool[2 * i + 1]) == hipSuccess) {
        info.kernelMs[evKind[i]] += ms;
        info.launches[evKind[i]]++;
      }
    }
    evUsed = 0;
  }

  int ensure(DevBuf& b, size_t bytes, std::string& err) {
    if (b.bytes >= bytes && b.ptr) return 0;
    if (b.ptr) (void)hipFree(b.ptr);
    b.ptr = nullptr;
    b.bytes = 0;
    const size_t want = bytes ? bytes : 16;
    if (hipMalloc(&b.ptr, want) != hipSuccess) {
      (void)hipGetLastError();
      err = "device allocation of " + std::to_string(want) + " bytes failed";
      return BTG_ERR_MEMORY_CODE;
    }
    b.bytes = want;
    return 0;
  }
  ~Engine() {
    DevBuf* all[] = {&tq0, &tq1, &sq, &roots, &live, &state, &epoch, &missFx, &gFx, &stats, &st, &out,
                     &outRoot, &outTag, &list, &listCount, &rasterImg, &rasterRows, &pixTri, &pixX, &pixXlo, &pixN, &vis, &E, &Eu, &prim, &retry, &cycles, &shq, &jobs, &pixRef, &fpList, &fpMem, &camD, &fpSum};
    for (DevBuf* b : all)
      if (b->ptr) (void)hipFree(b->ptr);
    for (void* m : sceneMem)
      if (m) (void)hipFree(m);
    if (hostState) (void)hipHostFree(hostState);
    for (hipEvent_t e : ev)
      if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : evPool) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

// Persistent grids: as many workgroups as can be resident at once.
template <class K>
int residentGrid(const Engine& e, K kernel) {
  int perCU = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, kernel, kBlock, 0) != hipSuccess || perCU < 1) perCU = 1;
  return e.cus * perCU;
}

// Prepares engine buffers for R roots and the common parameters.
int prepare(Engine& e, int app, int R, int budget, Params& p, std::string& err) {
  // queue records: the configured capacity, but no more than the roots can hold
  const size_t cap = size_t(std::min<long long>(e.opt.capacity, std::max<long long>(
                                                     int64_t(1) << 16, (long long)R * (long long)budget)));
  int rc;
  if ((rc = e.ensure(e.tq0, cap * sizeof(Rec), err)) || (rc = e.ensure(e.tq1, cap * sizeof(Rec), err)) ||
      (rc = e.ensure(e.sq, cap * sizeof(Rec), err)) || (rc = e.ensure(e.roots, size_t(R) * sizeof(RootQ), err)) ||
      (rc = e.ensure(e.live, size_t(R) * 4, err)) || (rc = e.ensure(e.state, size_t(R) * 4, err)) ||
      (rc = e.ensure(e.epoch, size_t(R) * 4, err)) || (rc = e.ensure(e.missFx, size_t(R) * 8, err)) ||
      (rc = e.ensure(e.gFx, size_t(R) * 8, err)) || (rc = e.ensure(e.stats, sizeof(uint64_t) * kStatSlots * kStatWords, err)) ||
      (rc = e.ensure(e.st, sizeof(WState), err)) ||
      (rc = e.ensure(e.cycles, 8 * sizeof(uint64_t), err)) ||
      (rc = e.ensure(e.list, size_t(R) * 4, err)) || (rc = e.ensure(e.listCount, 4, err)))
    return rc;
  p = Params();
  p.sv = e.sv;
  p.node0 = 0;  // primary and point shadows trace scene 0
  p.rootBox = e.scenes[0].bounds;
  p.app = app;
  p.R = R;
  p.rootList = nullptr;
  p.budget = budget;
  p.capacity = (unsigned long long)cap;
  p.tq[0] = static_cast<Rec*>(e.tq0.ptr);
  p.tq[1] = static_cast<Rec*>(e.tq1.ptr);
  p.sq = static_cast<Rec*>(e.sq.ptr);
  p.roots = static_cast<RootQ*>(e.roots.ptr);
  p.live = static_cast<int*>(e.live.ptr);
  p.state = static_cast<int*>(e.state.ptr);
  p.epoch = static_cast<int*>(e.epoch.ptr);
  p.pass = 0;
  p.missFx = static_cast<unsigned long long*>(e.missFx.ptr);
  p.gFx = static_cast<unsigned long long*>(e.gFx.ptr);
  p.stats = static_cast<unsigned long long*>(e.stats.ptr);
  p.cycles = static_cast<unsigned long long*>(e.cycles.ptr);
  p.st = static_cast<WState*>(e.st.ptr);
  // Primary passes are chains of short rounds whose trace launches last as
  // long as their slowest lane: they yield after at most kPrimaryLeafCap
  // leaves (measured on the conference hall: 45 ms against 50 ms at 8).
  p.leafCap = app == kPrimary && e.opt.leafCap > 0 ? std::min(e.opt.leafCap, kPrimaryLeafCap) : e.opt.leafCap;
  p.mailbox = e.opt.mailbox;
  return 0;
}

void setCamera(const HostCam& cam, Params& p) {
  p.W = cam.W;
  p.H = cam.H;
  p.eye = cam.eye;
  p.fwd = cam.fwd;
  p.right = cam.right;
  p.down = cam.down;
  p.halfW = cam.halfW;
  p.halfH = cam.halfH;
  // Camera::plane (render_types.h:28); host code rounds like the device
  // (IEEE single, no contraction).
  p.camPlane = Plane::make(cam.eye, cam.eye + cam.fwd, cam.right, cam.down);
}


// The instantiated graphs of one group of rounds, one per parity of the
// group's first round (the queues ping-pong by round parity), freed on exit.
struct RoundGraphs {
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  ~RoundGraphs() {
    for (hipGraphExec_t x : exec)
      if (x) (void)hipGraphExecDestroy(x);
  }
};

// Runs the rounds of one pass until the state reports done: groups of
// roundsPerCheck rounds (admit, trace, shade, split each), then one state
// read-back. With Options::graphs each group is captured once into a HIP
// graph and replayed; the kernels read all changing state from device memory,
// so the replays are the same launches. Per-launch timing (Options::profile)
// brackets every launch with HIP events, which needs direct launches: events
// recorded by graph nodes give no elapsed times.
template <int App>
int runPassApp(Engine& e, Params& p, RunInfo& info, std::string& err) {
  const int gA = residentGrid(e, admitKernel<App>), gT = residentGrid(e, traceKernel<App>),
            gS = residentGrid(e, splitKernel<App>), gH = residentGrid(e, shadeKernel);
  const bool shade = isSoft(App) && p.exact;
  const int k = std::max(1, e.opt.roundsPerCheck);
  auto group = [&](int first) {
    for (int i = 0; i < k; ++i) {
      const int cur = (first + i) & 1;
      int ls = e.beginLaunch(0);
      admitKernel<App><<<gA, kBlock, 0, e.stream>>>(p, cur);
      e.endLaunch(ls);
      ls = e.beginLaunch(1);
      traceKernel<App><<<gT, kBlock, 0, e.stream>>>(p, cur);
      e.endLaunch(ls);
      if (shade) {
        ls = e.beginLaunch(4);
        shadeKernel<<<gH, kBlock, 0, e.stream>>>(p);
        e.endLaunch(ls);
      }
      ls = e.beginLaunch(2);
      splitKernel<App><<<gS, kBlock, 0, e.stream>>>(p, cur);
      e.endLaunch(ls);
    }
  };
  initKernel<<<1, 64, 0, e.stream>>>(p);
  BTG_TRY(hipGetLastError());
  RoundGraphs gr;
  const bool graphs = e.opt.graphs && !e.opt.profile;
  int rounds = 0;
  for (;;) {
    const int par = rounds & 1;
    if (graphs) {
      if (!gr.exec[par]) {
        BTG_TRY(hipStreamBeginCapture(e.stream, hipStreamCaptureModeThreadLocal));
        group(par);
        hipGraph_t g = nullptr;
        const hipError_t ce = hipStreamEndCapture(e.stream, &g);
        const hipError_t ie = ce == hipSuccess ? hipGraphInstantiate(&gr.exec[par], g, nullptr, nullptr, 0) : ce;
        if (g) (void)hipGraphDestroy(g);
        BTG_TRY(ie);
      }
      BTG_TRY(hipGraphLaunch(gr.exec[par], e.stream));
    } else {
      group(rounds);
    }
    rounds += k;
    BTG_TRY(hipGetLastError());
    BTG_TRY(hipMemcpyAsync(e.hostState, p.st, sizeof(WState), hipMemcpyDeviceToHost, e.stream));
    BTG_TRY(hipStreamSynchronize(e.stream));
    if (e.hostState->overflow) {
      err = "sub-beam queue overflow (capacity " + std::to_string(e.opt.capacity) + ")";
      return BTG_ERR_LIMIT_CODE;
    }
    if (e.hostState->done) break;
    if (rounds > 200000) {
      err = "wavefront engine did not finish";
      return BTG_ERR_LIMIT_CODE;
    }
  }
  info.rounds += e.hostState->round;
  info.maxQueue = std::max(info.maxQueue, e.hostState->maxQueue);
  return 0;
}

int runPass(Engine& e, Params& p, RunInfo& info, std::string& err) {
  if (p.app == kSoft) return runPassApp<kSoft>(e, p, info, err);
  if (p.app == kSoftG) return runPassApp<kSoftG>(e, p, info, err);
  if (p.app == kPrimary) return runPassApp<kPrimary>(e, p, info, err);
  return runPassApp<kPoint>(e, p, info, err);
}

int readStats(Engine& e, uint64_t* dst, double* dropped, std::string& err) {
  static_assert(sNum < kStatWords, "a stat slot holds the counters and the dropped area");
  std::vector<uint64_t> s(size_t(kStatSlots) * kStatWords);
  BTG_TRY(hipMemcpyAsync(e.phaseCycles, e.cycles.ptr, sizeof(e.phaseCycles), hipMemcpyDeviceToHost, e.stream));
  BTG_TRY(hipMemcpyAsync(s.data(), e.stats.ptr, s.size() * sizeof(uint64_t), hipMemcpyDeviceToHost, e.stream));
  BTG_TRY(hipStreamSynchronize(e.stream));
  double d = 0;
  for (int k = 0; k < kStatSlots; ++k) {
    const uint64_t* slot = s.data() + size_t(k) * kStatWords;
    for (int i = 0; i < sNum; ++i) dst[i] += slot[i];
    double x;
    std::memcpy(&x, slot + kStatWords - 1, sizeof(x));
    d += x;
  }
  if (dropped) *dropped += d;
  return 0;
}

int zeroStats(Engine& e, std::string& err) {
  zeroWords<<<kStatSlots * kStatWords / 256, 256, 0, e.stream>>>(static_cast<unsigned long long*>(e.stats.ptr),
                                                               size_t(kStatSlots) * kStatWords);
  zeroWords<<<1, 64, 0, e.stream>>>(static_cast<unsigned long long*>(e.cycles.ptr), 8);
  BTG_TRY(hipGetLastError());
  return 0;
}

// The main pass plus GPU retry passes for roots over budget. `finalize`
// runs after each pass over that pass's roots.
template <class Finalize>
int runAll(Engine& e, Params& p, int baseBudget, RunInfo& info, Finalize finalize, std::string& err) {
  int rc = runPass(e, p, info, err);
  if (rc) return rc;
  if ((rc = finalize(p))) return rc;
  info.killed = e.hostState->killed;
  int budget = baseBudget;
  while (e.hostState->killed > 0) {
    // Collect the killed roots and trace them again with a larger budget.
    clearInt<<<1, 64, 0, e.stream>>>(static_cast<int*>(e.listCount.ptr));
    collectKilled<<<(p.R + 255) / 256, 256, 0, e.stream>>>(p, static_cast<int*>(e.list.ptr),
                                                             static_cast<int*>(e.listCount.ptr));
    BTG_TRY(hipGetLastError());
    int nk = 0;
    BTG_TRY(hipMemcpyAsync(&nk, e.listCount.ptr, 4, hipMemcpyDeviceToHost, e.stream));
    BTG_TRY(hipStreamSynchronize(e.stream));
    if (nk == 0) break;
    if ((unsigned long long)budget >= p.capacity) {
      err = "a root needs more sub-beams than the queue capacity (" + std::to_string(p.capacity) + ")";
      return BTG_ERR_LIMIT_CODE;
    }
    budget = int(std::min<unsigned long long>((unsigned long long)budget * 32, p.capacity));
    // The retry list must not alias the list collectKilled writes next: copy it.
    int rc2;
    Params q = p;
    q.R = nk;
    q.budget = budget;
    q.pass = p.pass + 1;
    if ((rc2 = e.ensure(e.retry, size_t(nk) * 4, err))) return rc2;
    BTG_TRY(hipMemcpyAsync(e.retry.ptr, e.list.ptr, size_t(nk) * 4, hipMemcpyDeviceToDevice, e.stream));
    q.rootList = static_cast<const int*>(e.retry.ptr);
    info.retries++;
    if ((rc = runPass(e, q, info, err))) return rc;
    if ((rc = finalize(q))) return rc;
    p.pass = q.pass;
    p.rootList = q.rootList;
    p.R = q.R;
  }
  return 0;
}

}  // namespace

namespace {

template <class T>
int upload(Engine& e, int slot, const T* src, size_t n, const T*& dst, std::string& err) {
  void* d = nullptr;
  if (hipMalloc(&d, sizeof(T) * (n ? n : 1)) != hipSuccess) {
    (void)hipGetLastError();
    err = "device allocation of the scene failed";
    return BTG_ERR_MEMORY_CODE;
  }
  e.sceneMem[slot] = d;
  // On the context's own (non-blocking) stream, never the legacy stream: on
  // this runtime a legacy-stream operation fails while any stream of the
  // process is capturing a graph (another context's, on another thread).
  if (n) {
    BTG_TRY(hipMemcpyAsync(d, src, sizeof(T) * n, hipMemcpyHostToDevice, e.stream));
    BTG_TRY(hipStreamSynchronize(e.stream));
  }
  dst = static_cast<const T*>(d);
  return 0;
}

}  // namespace

// GpuRenderer::GpuRenderer (cuda/gpu_render.cu:489-528): the scene, kd-tree
// and packed leaf triangles go to device memory once.
namespace {

// Device format of a scene: per-triangle plane data, kd nodes, boxes, parent
// links, and the leaf triangles in kd reference order (buildTriRefs,
// beam_types.h:100-112).
SceneRec makeSceneRec(const btg_scene_desc& d, float offset, float offsetGuard) {
  SceneRec sr;
  sr.live = true;
  sr.triNV.resize(size_t(d.numTris) * 2);
  for (int64_t t = 0; t < d.numTris; ++t) {
    const uint32_t v0 = d.tri3[3 * t];
    sr.triNV[size_t(2 * t)] = make_float4(d.triN[3 * t], d.triN[3 * t + 1], d.triN[3 * t + 2], 0.0f);
    sr.triNV[size_t(2 * t + 1)] = make_float4(d.pos[3 * v0], d.pos[3 * v0 + 1], d.pos[3 * v0 + 2], 0.0f);
  }
  sr.refs.resize(static_cast<size_t>(d.numRefs));
  for (int64_t i = 0; i < d.numRefs; ++i) {
    const uint32_t t = d.triIndices[i];
    LeafTri& r = sr.refs[size_t(i)];
    for (int k = 0; k < 3; ++k) {
      r.vi[k] = d.tri3[3 * size_t(t) + k];
      for (int a = 0; a < 3; ++a) r.v[3 * k + a] = d.pos[3 * size_t(r.vi[k]) + a];
      r.n[k] = d.triN[3 * size_t(t) + k];
    }
    r.tri = int32_t(t);
  }
  const KdNodeD* nodes = static_cast<const KdNodeD*>(d.nodes);
  sr.nodes.assign(nodes, nodes + d.numNodes);
  sr.boxes.resize(static_cast<size_t>(d.numNodes));
  for (int64_t i = 0; i < d.numNodes; ++i) {
    const float* b = d.nodeBox + 6 * i;
    sr.boxes[size_t(i)] = Box{f3(b[0], b[1], b[2]), f3(b[3], b[4], b[5])};
  }
  sr.parent.assign(d.parent, d.parent + d.numNodes);
  sr.bounds = Box{f3(d.kdBounds[0], d.kdBounds[1], d.kdBounds[2]), f3(d.kdBounds[3], d.kdBounds[4], d.kdBounds[5])};
  sr.numTris = int(d.numTris);
  sr.numVerts = uint32_t(d.numVerts);
  sr.offset = offset;
  sr.offsetGuard = offsetGuard;
  return sr;
}

// Concatenates the live scenes into the device arena. Node, leaf-reference,
// triangle and vertex numbers of each scene are shifted by its base; the
// order of a triangle's vertex numbers (which orients shared edges) is kept.
int rebuildArena(Engine& e, std::string& err) {
  uint32_t nodes = 0, refs = 0, tris = 0, verts = 0;
  for (SceneRec& sr : e.scenes) {
    if (!sr.live) continue;
    sr.nodeBase = nodes;
    sr.refBase = refs;
    sr.triBase = tris;
    sr.vertBase = verts;
    const uint64_t n2 = uint64_t(nodes) + sr.nodes.size(), r2 = uint64_t(refs) + sr.refs.size();
    const uint64_t t2 = uint64_t(tris) + uint64_t(sr.numTris), v2 = uint64_t(verts) + sr.numVerts;
    if (n2 >= (1ull << 30) || r2 >= (1ull << 31) || t2 >= (1ull << 31) || v2 >= (1ull << 32)) {
      err = "resident scenes exceed the arena's 30-bit node / 31-bit triangle numbering";
      return BTG_ERR_LIMIT_CODE;
    }
    nodes = uint32_t(n2);
    refs = uint32_t(r2);
    tris = uint32_t(t2);
    verts = uint32_t(v2);
  }
  std::vector<float4> nv;
  std::vector<KdNodeD> kn;
  std::vector<Box> bx;
  std::vector<uint32_t> par;
  std::vector<LeafTri> lt;
  std::vector<int32_t> rt;
  nv.reserve(size_t(tris) * 2);
  kn.reserve(nodes);
  bx.reserve(nodes);
  par.reserve(nodes);
  lt.reserve(refs);
  rt.reserve(refs);
  for (const SceneRec& sr : e.scenes) {
    if (!sr.live) continue;
    nv.insert(nv.end(), sr.triNV.begin(), sr.triNV.end());
    for (size_t i = 0; i < sr.nodes.size(); ++i) {
      KdNodeD n = sr.nodes[i];
      if ((n.data & 3u) != 3u)
        n.data = (n.data & 3u) | (((n.data >> 2) + sr.nodeBase) << 2);
      else
        n.offset += sr.refBase;
      kn.push_back(n);
      par.push_back(i == 0 ? sr.nodeBase : sr.parent[i] + sr.nodeBase);
    }
    bx.insert(bx.end(), sr.boxes.begin(), sr.boxes.end());
    for (LeafTri r : sr.refs) {
      r.tri += int32_t(sr.triBase);
      for (int k = 0; k < 3; ++k) r.vi[k] += sr.vertBase;
      lt.push_back(r);
      rt.push_back(r.tri);
    }
  }
  for (void*& m : e.sceneMem)
    if (m) {
      (void)hipFree(m);
      m = nullptr;
    }
  DevScene& sv = e.sv;
  int rc = upload(e, 0, nv.data(), nv.size(), sv.triNV, err);
  if (!rc) rc = upload(e, 1, kn.data(), kn.size(), sv.nodes, err);
  if (!rc) rc = upload(e, 2, bx.data(), bx.size(), sv.nodeBox, err);
  if (!rc) rc = upload(e, 3, par.data(), par.size(), sv.parent, err);
  if (!rc) rc = upload(e, 4, lt.data(), lt.size(), sv.refs, err);
  if (!rc) rc = upload(e, 5, rt.data(), rt.size(), sv.refTri, err);
  if (rc) return rc;
  const SceneRec& s0 = e.scenes[0];
  sv.bounds = s0.bounds;
  sv.numTris = s0.numTris;
  return 0;
}

}  // namespace

namespace {

// This device's address of noAllocScratch: looked up once per device, with
// that device current, and kept for the process (null if the lookup failed).
void* noAllocScratchOn(int dev) {
  static std::mutex m;
  static std::vector<void*> addr;
  static std::vector<char> looked;
  std::lock_guard<std::mutex> g(m);
  if (dev < 0) return nullptr;
  if (size_t(dev) >= addr.size()) {
    addr.resize(size_t(dev) + 1, nullptr);
    looked.resize(size_t(dev) + 1, 0);
  }
  if (!looked[size_t(dev)]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(noAllocScratch)) != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
    }
    addr[size_t(dev)] = p;
    looked[size_t(dev)] = 1;
  }
  return addr[size_t(dev)];
}

}  // namespace

Engine* engineCreate(const btg_scene_desc& d, float offset, float offsetGuard, std::string& err) {
  Engine* e = new Engine();
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&e->cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
    err = "no HIP device";
    delete e;
    return nullptr;
  }
  e->noAllocScratch = noAllocScratchOn(dev);
  // Queue capacity by device size: the largest power of two whose three
  // sub-beam queues and shade queue (3 x 64 + 40 B per record) fit in 11% of
  // device memory, 2^20..2^28 (2^27 on a 288 GB MI355X: 31 GB), so that a
  // few contexts fit side by side. More roots in flight per round means fewer
  // rounds: the 64-job batch runs 24% faster at 2^27 than at 2^25. Buffers
  // are allocated as passes need them (no more than roots x budget records).
  size_t freeB = 0, totalB = 0;
  if (hipMemGetInfo(&freeB, &totalB) == hipSuccess && totalB > 0) {
    const double fit = double(totalB) * 0.11 / (3 * 64 + 40);
    int64_t cap = int64_t(1) << 20;
    while (cap < (int64_t(1) << 28) && double(cap * 2) <= fit) cap *= 2;
    e->opt.capacity = cap;
  }
  (void)hipGetLastError();
  if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void**>(&e->hostState), sizeof(WState), hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    err = "stream / pinned allocation failed";
    delete e;
    return nullptr;
  }
  for (hipEvent_t& ev : e->ev)
    if (hipEventCreate(&ev) != hipSuccess) {
      err = "event creation failed";
      delete e;
      return nullptr;
    }
  e->scenes.push_back(makeSceneRec(d, offset, offsetGuard));
  if (rebuildArena(*e, err)) {
    delete e;
    return nullptr;
  }
  return e;
}

int engineAddScene(Engine* e, const btg_scene_desc& d, float offset, float offsetGuard, std::string& err) {
  int id = -1;
  for (size_t i = 1; i < e->scenes.size(); ++i)
    if (!e->scenes[i].live) {
      id = int(i);
      break;
    }
  if (id < 0) {
    id = int(e->scenes.size());
    e->scenes.emplace_back();
  }
  e->scenes[size_t(id)] = makeSceneRec(d, offset, offsetGuard);
  BTG_TRY(hipStreamSynchronize(e->stream));
  if (int rc = rebuildArena(*e, err)) {
    e->scenes[size_t(id)] = SceneRec();
    return rc;
  }
  return id;
}

int engineRemoveScene(Engine* e, int scene, std::string& err) {
  if (scene <= 0 || size_t(scene) >= e->scenes.size() || !e->scenes[size_t(scene)].live) {
    err = "no resident scene " + std::to_string(scene) + " (scene 0 stays for the context's lifetime)";
    return BTG_ERR_ARG_CODE;
  }
  BTG_TRY(hipStreamSynchronize(e->stream));
  e->scenes[size_t(scene)] = SceneRec();
  return rebuildArena(*e, err);
}

int64_t engineSceneBytes(const Engine* e, int scene) {
  if (scene < 0 || size_t(scene) >= e->scenes.size() || !e->scenes[size_t(scene)].live) return -1;
  const SceneRec& sr = e->scenes[size_t(scene)];
  return int64_t(sr.triNV.size() * sizeof(float4) + sr.nodes.size() * (sizeof(KdNodeD) + sizeof(Box) + 4) +
                 sr.refs.size() * (sizeof(LeafTri) + sizeof(int32_t)));
}

// roots: RootQ, live, state, epoch, missFx, gFx (as prepare allocates them)
int64_t rootStateBytes() { return int64_t(sizeof(RootQ) + 3 * sizeof(int) + 2 * sizeof(unsigned long long)); }
// pixels: pixTri, pixX, pixN, vis, E, Eu, and pixXlo when guarded
int64_t softPixelBytes(bool guard) { return 4 + 12 + 12 + 4 + 4 + 4 + (guard ? 12 : 0); }

void engineDestroy(Engine* e) { delete e; }

Options& engineOptions(Engine* e) { return e->opt; }

namespace {

// FP64 soft shadows of every pixel of a batch (after its primary pass), in
// three levels of cells (fp64CellKernel): the whole root with pools of 64
// beams and the work cap; the 4 x 4 cells of the roots that overflowed, with
// 128 beams and the cap; the 16 x 16 cells inside those that overflowed
// again, with 384 beams and no cap. The levels run back to back on the stream
// (each reads its task count from the device); one synchronisation at the end
// reads the cells that overflowed their last level.
int softFp64(Engine& e, Params& p, size_t np, std::string& err) {
  using T0 = fp64::Tracer<64, 64>;
  using T1 = fp64::Tracer<kFp64Pool1, kFp64Pool1>;
  using T2 = fp64::Tracer<384, 256>;
  constexpr size_t kBlock0 = size_t(T0::kWords) * 64 * 8, kBlock1 = size_t(T1::kWords) * 64 * 8,
                   kBlock2 = size_t(T2::kWords) * 64 * 8;
  const size_t cus = size_t(std::max(e.cus, 1)), maxBlocks = cus * 4 * kFp64Waves;
  const size_t blocks0 = std::min((np + 63) / 64, maxBlocks);
  const size_t blocks1 = std::min((16 * np + 63) / 64, maxBlocks);
  const size_t blocks2 = std::min((256 * np + 63) / 64, cus);
  const size_t cap2 = std::min<size_t>(16 * np, size_t(1) << 24);  // level-2 entries
  int rc;
  if ((rc = e.ensure(e.fpMem, std::max({blocks0 * kBlock0, blocks1 * kBlock1, blocks2 * kBlock2}), err)) ||
      (rc = e.ensure(e.fpList, 64 + (np + cap2) * sizeof(long long), err)) ||
      (rc = e.ensure(e.fpSum, 2 * np * sizeof(unsigned long long), err)))
    return rc;
  uint64_t* mem = static_cast<uint64_t*>(e.fpMem.ptr);
  int* ctr = static_cast<int*>(e.fpList.ptr);  // next[3], outCount[2], failed
  long long* list1 = reinterpret_cast<long long*>(static_cast<char*>(e.fpList.ptr) + 64);
  long long* list2 = list1 + np;
  unsigned long long* visSum = static_cast<unsigned long long*>(e.fpSum.ptr);
  unsigned long long* gSum = visSum + np;
  zeroWords<<<1, 64, 0, e.stream>>>(reinterpret_cast<unsigned long long*>(ctr), 8);
  zeroWords<<<unsigned(std::min<size_t>((2 * np + 255) / 256, 4096)), 256, 0, e.stream>>>(visSum, 2 * np);
  const int n = int(np);
  int lt = e.beginLaunch(1);
  fp64CellKernel<64, 64, 0><<<unsigned(blocks0), 64, 0, e.stream>>>(p, n, nullptr, nullptr, 0, ctr, mem, list1,
                                                                    ctr + 3, n, visSum, gSum, ctr + 5);
  e.endLaunch(lt);
  lt = e.beginLaunch(1);
  fp64CellKernel<kFp64Pool1, kFp64Pool1, 1><<<unsigned(blocks1), 64, 0, e.stream>>>(p, n, list1, ctr + 3, n, ctr + 1, mem, list2,
                                                                      ctr + 4, int(cap2), visSum, gSum, ctr + 5);
  e.endLaunch(lt);
  lt = e.beginLaunch(1);
  fp64CellKernel<384, 256, 2><<<unsigned(blocks2), 64, 0, e.stream>>>(p, n, list2, ctr + 4, int(cap2), ctr + 2, mem,
                                                                      nullptr, nullptr, 0, visSum, gSum, ctr + 5);
  e.endLaunch(lt);
  fp64FinishKernel<<<unsigned((np + 255) / 256), 256, 0, e.stream>>>(p, n, visSum, gSum);
  BTG_TRY(hipGetLastError());
  int failed = 0;
  BTG_TRY(hipMemcpyAsync(&failed, ctr + 5, 4, hipMemcpyDeviceToHost, e.stream));
  BTG_TRY(hipStreamSynchronize(e.stream));
  if (failed > 0) {
    err = std::to_string(failed) + " cell(s) of 1/256 of a pixel's light need more than 384 live beams in FP64 mode";
    return BTG_ERR_LIMIT_CODE;
  }
  return 0;
}

}  // namespace

int engineSoftBatch(Engine* e, const SoftJobH* jobs, int n, bool exact, RunInfo& info, std::string& err) {
  if (n <= 0) {
    err = "a batch needs at least one job";
    return BTG_ERR_ARG_CODE;
  }
  std::vector<SoftJob> jt(static_cast<size_t>(n));
  long long R = 0, npix = 0, tiles = 0;
  for (int j = 0; j < n; ++j) {
    const SoftJobH& h = jobs[j];
    if (h.scene < 0 || size_t(h.scene) >= e->scenes.size() || !e->scenes[size_t(h.scene)].live) {
      err = "job " + std::to_string(j) + ": no resident scene " + std::to_string(h.scene);
      return BTG_ERR_ARG_CODE;
    }
    if (h.cam.W <= 0 || h.cam.H <= 0) {
      err = "job " + std::to_string(j) + ": image size must be positive";
      return BTG_ERR_ARG_CODE;
    }
    const SceneRec& sr = e->scenes[size_t(h.scene)];
    SoftJob& J = jt[size_t(j)];
    J.W = h.cam.W;
    J.H = h.cam.H;
    J.tilesX = int(rootTilesX(h.cam.W));
    J.node0 = sr.nodeBase;
    J.triBase = int(sr.triBase);
    J.bounds = sr.bounds;
    J.offset = e->opt.precision ? sr.offsetGuard : sr.offset;
    J.halfW = h.cam.halfW;
    J.halfH = h.cam.halfH;
    J.eye = h.cam.eye;
    J.fwd = h.cam.fwd;
    J.right = h.cam.right;
    J.down = h.cam.down;
    J.lc = h.light.c;
    J.lU = h.light.U;
    J.lV = h.light.V;
    J.pixBase = int(npix);
    J.rootBase = int(R);
    J.tileBase = int(tiles);
    npix += (long long)h.cam.W * h.cam.H;
    R += (long long)softRootCount(h.cam.W, h.cam.H);
    tiles += (long long)((h.cam.W + 15) / 16) * ((h.cam.H + 15) / 16);
    if (R > (1LL << 30) || npix > (1LL << 30)) {
      err = "batch too large (more than 2^30 pixels)";
      return BTG_ERR_ARG_CODE;
    }
  }
  Params p;
  const bool guard = e->opt.precision != 0;
  const bool fp64Mode = e->opt.precision == 2;  // no roots: each pixel is traced on its own
  int rc = prepare(*e, guard ? kSoftG : kSoft, fp64Mode ? 0 : int(R), e->opt.budget, p, err);
  if (rc) return rc;
  const size_t np = size_t(npix);
  if ((rc = e->ensure(e->pixTri, np * 4, err)) || (rc = e->ensure(e->pixX, np * 12, err)) ||
      (rc = e->ensure(e->pixN, np * 12, err)) || (rc = e->ensure(e->vis, np * 4, err)) ||
      (rc = e->ensure(e->E, np * 4, err)) || (rc = e->ensure(e->Eu, np * 4, err)) ||
      (rc = e->ensure(e->jobs, jt.size() * sizeof(SoftJob), err)) ||
      (guard && (rc = e->ensure(e->pixXlo, np * 12, err))))
    return rc;
  if (guard) p.pixXlo = static_cast<const float*>(e->pixXlo.ptr);
  // FP64 mode: each job's camera in double (the caller's, or its floats widened)
  std::vector<double> camD;
  if (fp64Mode) {
    camD.resize(size_t(n) * 14);
    for (int j = 0; j < n; ++j) {
      double* c = camD.data() + size_t(j) * 14;
      const HostCam& hc = jobs[j].cam;
      if (jobs[j].camD) {
        std::memcpy(c, jobs[j].camD, 14 * sizeof(double));
      } else {
        const F3 v[4] = {hc.eye, hc.fwd, hc.right, hc.down};
        for (int k = 0; k < 4; ++k) {
          c[3 * k] = v[k].x;
          c[3 * k + 1] = v[k].y;
          c[3 * k + 2] = v[k].z;
        }
        c[12] = hc.halfW;
        c[13] = hc.halfH;
      }
    }
    if ((rc = e->ensure(e->pixRef, np * 4, err)) ||
        (rc = e->ensure(e->camD, camD.size() * sizeof(double), err)))
      return rc;
    BTG_TRY(hipMemcpyAsync(e->camD.ptr, camD.data(), camD.size() * sizeof(double), hipMemcpyHostToDevice, e->stream));
    p.pixRef = static_cast<int*>(e->pixRef.ptr);
    p.jobCamD = static_cast<const double*>(e->camD.ptr);
  }
  BTG_TRY(hipMemcpyAsync(e->jobs.ptr, jt.data(), jt.size() * sizeof(SoftJob), hipMemcpyHostToDevice, e->stream));
  p.jobs = static_cast<const SoftJob*>(e->jobs.ptr);
  p.numJobs = n;
  p.exact = exact ? 1 : 0;
  if (exact) {
    if ((rc = e->ensure(e->shq, size_t(p.capacity) * sizeof(ShadeRec), err))) return rc;
    p.shq = static_cast<ShadeRec*>(e->shq.ptr);
  }
  p.pixTri = static_cast<const int*>(e->pixTri.ptr);
  p.pixX = static_cast<const float*>(e->pixX.ptr);
  p.pixN = static_cast<const float*>(e->pixN.ptr);
  p.vis = static_cast<float*>(e->vis.ptr);
  p.E = static_cast<float*>(e->E.ptr);
  p.Eu = static_cast<float*>(e->Eu.ptr);
  if ((rc = zeroStats(*e, err))) return rc;
  BTG_TRY(hipEventRecord(e->ev[0], e->stream));
  const int lp = e->beginLaunch(3);
  primaryRayKernel<<<unsigned(tiles), kRayBlock, 0, e->stream>>>(p);
  e->endLaunch(lp);
  BTG_TRY(hipGetLastError());
  BTG_TRY(hipEventRecord(e->ev[1], e->stream));
  // the primary pass's counters are reported separately
  if ((rc = readStats(*e, info.rayStats, nullptr, err))) return rc;
  if ((rc = zeroStats(*e, err))) return rc;
  BTG_TRY(hipEventRecord(e->ev[2], e->stream));
  auto fin = [&](Params& q) -> int {
    finalizeSoftKernel<<<(q.R + 255) / 256, 256, 0, e->stream>>>(q);
    BTG_TRY(hipGetLastError());
    return 0;
  };
  if (fp64Mode) {
    if ((rc = softFp64(*e, p, np, err))) return rc;
  } else if ((rc = runAll(*e, p, e->opt.budget, info, fin, err))) {
    return rc;
  }
  BTG_TRY(hipEventRecord(e->ev[3], e->stream));
  BTG_TRY(hipEventSynchronize(e->ev[3]));
  float ms0 = 0, ms1 = 0;
  BTG_TRY(hipEventElapsedTime(&ms0, e->ev[0], e->ev[1]));
  BTG_TRY(hipEventElapsedTime(&ms1, e->ev[2], e->ev[3]));
  info.primaryMs += ms0;
  info.traceMs += ms1;
  if ((rc = readStats(*e, info.stats, &info.dropped, err))) return rc;
  for (int k = 0; k < 8; ++k) info.phaseCycles[k] = e->phaseCycles[k];
  e->collectLaunches(info);
  for (int j = 0; j < n; ++j) {
    const SoftJobH& h = jobs[j];
    const SoftJob& J = jt[size_t(j)];
    const size_t off = size_t(J.pixBase), cnt = size_t(J.W) * size_t(J.H);
    if (h.tri) BTG_TRY(hipMemcpyAsync(h.tri, static_cast<int*>(e->pixTri.ptr) + off, cnt * 4, hipMemcpyDeviceToHost, e->stream));
    if (h.vis) BTG_TRY(hipMemcpyAsync(h.vis, static_cast<float*>(e->vis.ptr) + off, cnt * 4, hipMemcpyDeviceToHost, e->stream));
    if (h.E) BTG_TRY(hipMemcpyAsync(h.E, static_cast<float*>(e->E.ptr) + off, cnt * 4, hipMemcpyDeviceToHost, e->stream));
    if (h.Eu) BTG_TRY(hipMemcpyAsync(h.Eu, static_cast<float*>(e->Eu.ptr) + off, cnt * 4, hipMemcpyDeviceToHost, e->stream));
  }
  BTG_TRY(hipStreamSynchronize(e->stream));
  // triangle ids back to each job's scene numbering
  for (int j = 0; j < n; ++j) {
    const SoftJobH& h = jobs[j];
    const int base = jt[size_t(j)].triBase;
    if (h.tri && base)
      for (size_t i = 0, c = size_t(h.cam.W) * size_t(h.cam.H); i < c; ++i)
        if (h.tri[i] >= 0) h.tri[i] -= base;
  }
  return 0;
}

int engineSoft(Engine* e, const HostCam& cam, const HostLight& L, bool exact, int32_t* tri, float* vis, float* E,
               float* Eu, RunInfo& info, std::string& err) {
  SoftJobH j{0, cam, L, tri, vis, E, Eu};
  return engineSoftBatch(e, &j, 1, exact, info, err);
}

// ------------------------------------------------------------ open passes
Engine* engineShare(Engine* base, std::string& err) {
  Engine* e = new Engine();
  e->cus = base->cus;
  e->opt = base->opt;
  if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void**>(&e->hostState), sizeof(WState), hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    err = "stream / pinned allocation failed";
    delete e;
    return nullptr;
  }
  for (hipEvent_t& ev : e->ev)
    if (hipEventCreate(&ev) != hipSuccess) {
      err = "event creation failed";
      delete e;
      return nullptr;
    }
  // the arena stays owned by `base` (sceneMem stays empty here)
  e->sv = base->sv;
  e->scenes.resize(base->scenes.size());
  for (size_t i = 0; i < base->scenes.size(); ++i) {
    const SceneRec& b = base->scenes[i];
    SceneRec& r = e->scenes[i];
    r.live = b.live;
    r.bounds = b.bounds;
    r.numTris = b.numTris;
    r.numVerts = b.numVerts;
    r.offset = b.offset;
    r.offsetGuard = b.offsetGuard;
    r.nodeBase = b.nodeBase;
    r.refBase = b.refBase;
    r.triBase = b.triBase;
    r.vertBase = b.vertBase;
  }
  return e;
}

struct OpenPass {
  Engine* e = nullptr;  // its own stream and buffers; the arena is shared
  Params p;
  int64_t rootCap = 0, pixCap = 0;
  int jobCap = 0;
  int64_t pixUsed = 0;
  int rounds = 0;  // rounds launched since the last reset (the queue parity)
  std::vector<SoftJob> jobs;
  std::vector<SoftJobH> hostJobs;
  std::vector<int> roots;
  std::vector<char> finished;
  int* jobLeftHost = nullptr;  // pinned
  int* killedHost = nullptr;   // pinned, one per job
  DevBuf jobLeft, killed;
  hipEvent_t stepDone = nullptr;
  bool exact = true;
  bool guard = false;  // FP64-guarded soft shadows (the options' precision at creation)
};

void openDestroy(OpenPass* op) {
  if (!op) return;
  if (op->e) (void)hipStreamSynchronize(op->e->stream);
  if (op->jobLeftHost) (void)hipHostFree(op->jobLeftHost);
  if (op->killedHost) (void)hipHostFree(op->killedHost);
  if (op->jobLeft.ptr) (void)hipFree(op->jobLeft.ptr);
  if (op->killed.ptr) (void)hipFree(op->killed.ptr);
  if (op->stepDone) (void)hipEventDestroy(op->stepDone);
  delete op->e;
  delete op;
}

namespace {
int openInit(OpenPass& op, std::string& err) {
  Engine& e = *op.e;
  const int budget = op.p.budget;
  const unsigned long long cap = op.p.capacity;
  Params& p = op.p;
  p.R = 0;
  p.numJobs = 0;
  p.pass = 0;
  p.rootList = nullptr;
  (void)budget;
  (void)cap;
  op.pixUsed = 0;
  op.rounds = 0;
  op.jobs.clear();
  op.hostJobs.clear();
  op.roots.clear();
  op.finished.clear();
  initKernel<<<1, 64, 0, e.stream>>>(p);
  BTG_TRY(hipGetLastError());
  return 0;
}
}  // namespace

OpenPass* openCreate(Engine* base, int64_t rootCap, int64_t pixCap, int jobCap, bool exact, int64_t queueCap,
                     std::string& err) {
  if (rootCap <= 0 || rootCap > (1LL << 30) || pixCap <= 0 || pixCap > (1LL << 30) || jobCap <= 0) {
    err = "open pass capacities out of range";
    return nullptr;
  }
  OpenPass* op = new OpenPass();
  op->e = engineShare(base, err);
  if (!op->e) {
    delete op;
    return nullptr;
  }
  Engine& e = *op->e;
  if (queueCap > 0) e.opt.capacity = std::min<int64_t>(e.opt.capacity, queueCap);
  op->rootCap = rootCap;
  op->pixCap = pixCap;
  op->jobCap = jobCap;
  op->exact = exact;
  op->guard = e.opt.precision != 0;
  Params& p = op->p;
  int rc = prepare(e, op->guard ? kSoftG : kSoft, int(rootCap), e.opt.budget, p, err);
  const size_t np = size_t(pixCap);
  if (!rc && ((rc = e.ensure(e.pixTri, np * 4, err)) || (rc = e.ensure(e.pixX, np * 12, err)) ||
              (op->guard && (rc = e.ensure(e.pixXlo, np * 12, err))) ||
              (rc = e.ensure(e.pixN, np * 12, err)) || (rc = e.ensure(e.vis, np * 4, err)) ||
              (rc = e.ensure(e.E, np * 4, err)) || (rc = e.ensure(e.Eu, np * 4, err)) ||
              (rc = e.ensure(e.jobs, size_t(jobCap) * sizeof(SoftJob), err)) ||
              (rc = e.ensure(op->jobLeft, size_t(jobCap) * 4, err)) ||
              (rc = e.ensure(op->killed, size_t(jobCap) * 4, err))))
    ;
  if (!rc && exact) rc = e.ensure(e.shq, size_t(p.capacity) * sizeof(ShadeRec), err);
  if (!rc && (hipHostMalloc(reinterpret_cast<void**>(&op->jobLeftHost), size_t(jobCap) * 4, hipHostMallocDefault) !=
                  hipSuccess ||
              hipHostMalloc(reinterpret_cast<void**>(&op->killedHost), size_t(jobCap) * 4, hipHostMallocDefault) !=
                  hipSuccess ||
              hipEventCreate(&op->stepDone) != hipSuccess)) {
    (void)hipGetLastError();
    err = "pinned allocation failed";
    rc = BTG_ERR_MEMORY_CODE;
  }
  if (rc) {
    openDestroy(op);
    return nullptr;
  }
  p.jobs = static_cast<const SoftJob*>(e.jobs.ptr);
  p.jobLeft = static_cast<int*>(op->jobLeft.ptr);
  p.exact = exact ? 1 : 0;
  if (exact) p.shq = static_cast<ShadeRec*>(e.shq.ptr);
  p.pixTri = static_cast<const int*>(e.pixTri.ptr);
  p.pixX = static_cast<const float*>(e.pixX.ptr);
  if (op->guard) p.pixXlo = static_cast<const float*>(e.pixXlo.ptr);
  p.pixN = static_cast<const float*>(e.pixN.ptr);
  p.vis = static_cast<float*>(e.vis.ptr);
  p.E = static_cast<float*>(e.E.ptr);
  p.Eu = static_cast<float*>(e.Eu.ptr);
  op->jobs.reserve(size_t(jobCap));
  if (zeroStats(e, err) || openInit(*op, err)) {
    openDestroy(op);
    return nullptr;
  }
  return op;
}

int openJobs(const OpenPass* op) { return int(op->jobs.size()); }

bool openIdle(const OpenPass* op) {
  for (char f : op->finished)
    if (!f) return false;
  return true;
}

int openReset(OpenPass* op, std::string& err) {
  if (!openIdle(op)) {
    err = "open pass still has jobs in flight";
    return BTG_ERR_ARG_CODE;
  }
  return openInit(*op, err);
}

int openTryAdd(OpenPass* op, const SoftJobH& h, int* slot, std::string& err) {
  Engine& e = *op->e;
  Params& p = op->p;
  if (h.scene < 0 || size_t(h.scene) >= e.scenes.size() || !e.scenes[size_t(h.scene)].live) {
    err = "no resident scene " + std::to_string(h.scene);
    return BTG_ERR_ARG_CODE;
  }
  if (h.cam.W <= 0 || h.cam.H <= 0) {
    err = "image size must be positive";
    return BTG_ERR_ARG_CODE;
  }
  const int tilesX = int(rootTilesX(h.cam.W));
  const int64_t roots = softRootCount(h.cam.W, h.cam.H), pix = int64_t(h.cam.W) * h.cam.H;
  if (roots > op->rootCap || pix > op->pixCap) {
    err = "request larger than a whole open pass";
    return BTG_ERR_LIMIT_CODE;
  }
  if (p.R + roots > op->rootCap || op->pixUsed + pix > op->pixCap || int(op->jobs.size()) >= op->jobCap) return 0;
  const SceneRec& sr = e.scenes[size_t(h.scene)];
  SoftJob J;
  J.W = h.cam.W;
  J.H = h.cam.H;
  J.tilesX = tilesX;
  J.node0 = sr.nodeBase;
  J.triBase = int(sr.triBase);
  J.bounds = sr.bounds;
  J.offset = op->guard ? sr.offsetGuard : sr.offset;
  J.halfW = h.cam.halfW;
  J.halfH = h.cam.halfH;
  J.eye = h.cam.eye;
  J.fwd = h.cam.fwd;
  J.right = h.cam.right;
  J.down = h.cam.down;
  J.lc = h.light.c;
  J.lU = h.light.U;
  J.lV = h.light.V;
  J.pixBase = int(op->pixUsed);
  J.rootBase = p.R;
  J.tileBase = 0;  // the job's primary rays are launched on their own
  const int j = int(op->jobs.size());
  op->jobs.push_back(J);  // capacity reserved: the element stays put until the copy has read it
  op->jobLeftHost[j] = int(roots);
  // The job joins the pass (host lists, root and pixel ranges) only once its
  // uploads and primary rays are queued: a failure here leaves the pass as it
  // was, so it can still drain and end.
  hipError_t he = hipMemcpyAsync(static_cast<SoftJob*>(e.jobs.ptr) + j, &op->jobs[size_t(j)], sizeof(SoftJob),
                                 hipMemcpyHostToDevice, e.stream);
  if (he == hipSuccess)
    he = hipMemcpyAsync(static_cast<int*>(op->jobLeft.ptr) + j, &op->jobLeftHost[j], 4, hipMemcpyHostToDevice,
                        e.stream);
  if (he == hipSuccess) {
    Params q = p;
    q.jobs = static_cast<const SoftJob*>(e.jobs.ptr) + j;
    q.numJobs = 1;
    const int tiles = ((h.cam.W + 15) / 16) * ((h.cam.H + 15) / 16);
    primaryRayKernel<<<unsigned(tiles), kRayBlock, 0, e.stream>>>(q);
    he = hipGetLastError();
  }
  if (he != hipSuccess) {
    op->jobs.pop_back();
    op->jobLeftHost[j] = 0;
    err = std::string("adding a job to an open pass: ") + hipGetErrorString(he);
    return BTG_ERR_DEVICE_CODE;
  }
  op->hostJobs.push_back(h);
  op->roots.push_back(int(roots));
  op->finished.push_back(0);
  p.R += int(roots);
  p.numJobs = j + 1;
  op->pixUsed += pix;
  reopenKernel<<<1, 64, 0, e.stream>>>(p);
  BTG_TRY(hipGetLastError());
  *slot = j;
  return 1;
}

namespace {
template <int App>
void openRounds(OpenPass* op, int rounds) {
  Engine& e = *op->e;
  Params& p = op->p;
  const int gA = residentGrid(e, admitKernel<App>), gT = residentGrid(e, traceKernel<App>),
            gS = residentGrid(e, splitKernel<App>), gH = residentGrid(e, shadeKernel);
  for (int i = 0; i < rounds; ++i, ++op->rounds) {
    const int cur = op->rounds & 1;
    admitKernel<App><<<gA, kBlock, 0, e.stream>>>(p, cur);
    traceKernel<App><<<gT, kBlock, 0, e.stream>>>(p, cur);
    if (p.exact) shadeKernel<<<gH, kBlock, 0, e.stream>>>(p);
    splitKernel<App><<<gS, kBlock, 0, e.stream>>>(p, cur);
  }
}
}  // namespace

int openStep(OpenPass* op, int rounds, std::string& err) {
  Engine& e = *op->e;
  Params& p = op->p;
  if (op->guard)
    openRounds<kSoftG>(op, rounds);
  else
    openRounds<kSoft>(op, rounds);
  BTG_TRY(hipGetLastError());
  BTG_TRY(hipMemcpyAsync(e.hostState, p.st, sizeof(WState), hipMemcpyDeviceToHost, e.stream));
  if (p.numJobs > 0)
    BTG_TRY(hipMemcpyAsync(op->jobLeftHost, op->jobLeft.ptr, size_t(p.numJobs) * 4, hipMemcpyDeviceToHost, e.stream));
  BTG_TRY(hipEventRecord(op->stepDone, e.stream));
  return 0;
}

int openCollect(OpenPass* op, std::vector<int>& completed, std::vector<int>& killed, std::string& err) {
  Engine& e = *op->e;
  Params& p = op->p;
  completed.clear();
  killed.clear();
  BTG_TRY(hipEventSynchronize(op->stepDone));
  if (e.hostState->overflow) {
    err = "sub-beam queue overflow in an open pass";
    return BTG_ERR_LIMIT_CODE;
  }
  for (int j = 0; j < int(op->jobs.size()); ++j)
    if (!op->finished[size_t(j)] && op->jobLeftHost[j] == 0) completed.push_back(j);
  if (completed.empty()) return 0;
  BTG_TRY(hipMemsetAsync(op->killed.ptr, 0, op->jobs.size() * 4, e.stream));
  for (int j : completed) {
    const SoftJob& J = op->jobs[size_t(j)];
    const int cnt = op->roots[size_t(j)];
    finalizeRangeKernel<<<(cnt + 255) / 256, 256, 0, e.stream>>>(p, J.rootBase, cnt,
                                                               static_cast<int*>(op->killed.ptr) + j);
    const SoftJobH& h = op->hostJobs[size_t(j)];
    const size_t off = size_t(J.pixBase), n = size_t(J.W) * size_t(J.H);
    if (h.tri) BTG_TRY(hipMemcpyAsync(h.tri, static_cast<int*>(e.pixTri.ptr) + off, n * 4, hipMemcpyDeviceToHost, e.stream));
    if (h.vis) BTG_TRY(hipMemcpyAsync(h.vis, static_cast<float*>(e.vis.ptr) + off, n * 4, hipMemcpyDeviceToHost, e.stream));
    if (h.E) BTG_TRY(hipMemcpyAsync(h.E, static_cast<float*>(e.E.ptr) + off, n * 4, hipMemcpyDeviceToHost, e.stream));
    if (h.Eu) BTG_TRY(hipMemcpyAsync(h.Eu, static_cast<float*>(e.Eu.ptr) + off, n * 4, hipMemcpyDeviceToHost, e.stream));
  }
  BTG_TRY(hipGetLastError());
  BTG_TRY(hipMemcpyAsync(op->killedHost, op->killed.ptr, op->jobs.size() * 4, hipMemcpyDeviceToHost, e.stream));
  BTG_TRY(hipStreamSynchronize(e.stream));
  for (int j : completed) {
    op->finished[size_t(j)] = 1;
    killed.push_back(op->killedHost[j]);
    const SoftJobH& h = op->hostJobs[size_t(j)];
    const int base = op->jobs[size_t(j)].triBase;
    if (h.tri && base)
      for (size_t i = 0, c = size_t(h.cam.W) * size_t(h.cam.H); i < c; ++i)
        if (h.tri[i] >= 0) h.tri[i] -= base;
  }
  return 0;
}

namespace {

// Downloads the polygons whose tag equals their root's last pass.
int downloadPolys(Engine& e, int R, std::vector<OutPoly>& polys, std::string& err) {
  const int n = std::min(e.hostState->outCount, int(e.out.bytes / sizeof(OutPoly)));
  std::vector<OutPoly> out(static_cast<size_t>(std::max(n, 0)));
  std::vector<int> root(out.size()), tag(out.size()), epoch(static_cast<size_t>(R));
  if (n > 0) {
    BTG_TRY(hipMemcpyAsync(out.data(), e.out.ptr, sizeof(OutPoly) * size_t(n), hipMemcpyDeviceToHost, e.stream));
    BTG_TRY(hipMemcpyAsync(root.data(), e.outRoot.ptr, sizeof(int) * size_t(n), hipMemcpyDeviceToHost, e.stream));
    BTG_TRY(hipMemcpyAsync(tag.data(), e.outTag.ptr, sizeof(int) * size_t(n), hipMemcpyDeviceToHost, e.stream));
  }
  if (R > 0) BTG_TRY(hipMemcpyAsync(epoch.data(), e.epoch.ptr, sizeof(int) * size_t(R), hipMemcpyDeviceToHost, e.stream));
  BTG_TRY(hipStreamSynchronize(e.stream));
  polys.clear();
  polys.reserve(out.size());
  for (int i = 0; i < n; ++i)
    if (tag[size_t(i)] == epoch[size_t(root[size_t(i)])]) polys.push_back(out[size_t(i)]);
  return 0;
}

// Primary and point shadows share the polygon-output driver; the output
// buffer grows and the call repeats if it was too small.
// Output polygon capacity for a pass: 2^20 at least (the visible fragments
// of a scene, not the image size, set the count: 230K for the conference hall
// at any resolution), and never less than an earlier pass grew the buffer to,
// so that a pass is not traced twice for want of room.
size_t outCapFor(const Engine& e, size_t want) {
  return std::max({size_t(1) << 20, want, e.out.bytes / sizeof(OutPoly)});
}

// With download = false the polygons stay on the device (e.out, filtered by
// outTag == epoch[outRoot]) for a device-side consumer.
template <class Setup>
int runPolys(Engine& e, int app, int R, int budget, size_t outCap, const HostCam& cam, Setup setup,
             std::vector<OutPoly>& polys, RunInfo& info, std::string& err, bool download = true) {
  for (;;) {
    Params p;
    int rc = prepare(e, app, R, budget, p, err);
    if (rc) return rc;
    if ((rc = e.ensure(e.out, outCap * sizeof(OutPoly), err)) || (rc = e.ensure(e.outRoot, outCap * 4, err)) ||
        (rc = e.ensure(e.outTag, outCap * 4, err)))
      return rc;
    setCamera(cam, p);
    p.out = static_cast<OutPoly*>(e.out.ptr);
    p.outRoot = static_cast<int*>(e.outRoot.ptr);
    p.outTag = static_cast<int*>(e.outTag.ptr);
    p.outCap = int(std::min<size_t>(outCap, size_t(1) << 30));
    setup(p);
    RunInfo ri;
    if ((rc = zeroStats(e, err))) return rc;
    BTG_TRY(hipEventRecord(e.ev[0], e.stream));
    auto fin = [&](Params&) -> int { return 0; };
    if ((rc = runAll(e, p, budget, ri, fin, err))) return rc;
    BTG_TRY(hipEventRecord(e.ev[1], e.stream));
    BTG_TRY(hipEventSynchronize(e.ev[1]));
    if (e.hostState->outCount > p.outCap) {
      outCap = size_t(e.hostState->outCount) + size_t(e.hostState->outCount) / 4;
      e.evUsed = 0;
      continue;
    }
    float ms = 0;
    BTG_TRY(hipEventElapsedTime(&ms, e.ev[0], e.ev[1]));
    ri.traceMs = ms;
    if ((rc = readStats(e, ri.stats, &ri.dropped, err))) return rc;
    e.collectLaunches(ri);
    if (download && (rc = downloadPolys(e, R, polys, err))) return rc;
    info = ri;
    return 0;
  }
}

}  // namespace

int enginePrimary(Engine* e, const HostCam& cam, bool cull, int tile, std::vector<OutPoly>& out, RunInfo& info,
                  std::string& err) {
  if (tile <= 0 || cam.W <= 0 || cam.H <= 0) {
    err = "tile and image size must be positive";
    return BTG_ERR_ARG_CODE;
  }
  const int tilesW = (cam.W + tile - 1) / tile, tilesH = (cam.H + tile - 1) / tile;
  const size_t outCap = outCapFor(*e, size_t(cam.W) * size_t(cam.H) * 2);
  auto setup = [&](Params& p) {
    p.tile = tile;
    p.tilesW = tilesW;
    p.cull = cull ? 1 : 0;
  };
  return runPolys(*e, kPrimary, tilesW * tilesH, e->opt.primaryBudget, outCap, cam, setup, out, info, err);
}

int enginePoint(Engine* e, const HostCam& cam, const OutPoly* prim, int64_t nPrim, F3 light,
                std::vector<OutPoly>& out, RunInfo& info, std::string& err) {
  out.clear();
  if (nPrim <= 0) return 0;
  if (nPrim > (1LL << 30)) {
    err = "too many primary polygons";
    return BTG_ERR_ARG_CODE;
  }
  for (int64_t i = 0; i < nPrim; ++i)
    if (prim[i].tri < 0 || prim[i].tri >= e->sv.numTris || prim[i].n < 3 || prim[i].n > 4) {
      err = "primary polygon " + std::to_string(i) + " has no valid hit triangle or corner count";
      return BTG_ERR_ARG_CODE;
    }
  int rc = e->ensure(e->prim, size_t(nPrim) * sizeof(OutPoly), err);
  if (rc) return rc;
  BTG_TRY(hipMemcpyAsync(e->prim.ptr, prim, size_t(nPrim) * sizeof(OutPoly), hipMemcpyHostToDevice, e->stream));
  const size_t outCap = outCapFor(*e, size_t(nPrim) * 4);
  auto setup = [&](Params& p) {
    p.pointLight = light;
    p.prim = static_cast<const OutPoly*>(e->prim.ptr);
  };
  return runPolys(*e, kPoint, int(nPrim), e->opt.budget, outCap, cam, setup, out, info, err);
}

// ------------------------------------------------------------ rasterization
namespace {

// x86's double -> int32 conversion (cvttsd2si), which the host rasterizer's
// int(std::ceil(...)) compiles to: truncation, and INT_MIN for NaN and for
// values outside the int range (the device's own conversion clamps instead).
__device__ __forceinline__ int toIntHost(double v) {
  return (v > -2147483649.0 && v < 2147483648.0) ? int(v) : int(0x80000000u);
}

// The edges of a polygon as half-planes A x + B y + C >= 0 in pixel
// coordinates, and its row range: rasterize (src/render/render.cpp:49-95)
// with one sample at each pixel centre. Corners go to pixel coordinates in
// float, as the reference computes them, everything after in double.
struct RasterEdges {
  double A[4], B[4], C[4];
  int n, y0, y1;
};

__device__ __forceinline__ void rasterEdges(const OutPoly& pg, float halfW, float halfH, int W, int H, RasterEdges& r) {
  double px[4], py[4];
  double ylo = 1e300, yhi = -1e300;
  r.n = pg.n;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const float x = (pg.x[i] / halfW + 1) * 0.5f * float(W);
    const float y = (pg.y[i] / halfH + 1) * 0.5f * float(H);
    px[i] = double(x);
    py[i] = double(y);
    if (i < pg.n) {
      ylo = py[i] < ylo ? py[i] : ylo;  // std::min(ylo, py)
      yhi = yhi < py[i] ? py[i] : yhi;  // std::max(yhi, py)
    }
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    // j = i + 1 wrapping at n; corners past n are never used
    const double pxj = i + 1 == pg.n ? px[0] : (i == 3 ? px[0] : px[(i + 1) & 3]);
    const double pyj = i + 1 == pg.n ? py[0] : (i == 3 ? py[0] : py[(i + 1) & 3]);
    r.A[i] = -(pyj - py[i]);
    r.B[i] = pxj - px[i];
    r.C[i] = -(r.A[i] * px[i] + r.B[i] * py[i]);
  }
  const int y0 = toIntHost(ceil(ylo - 0.5)), y1 = toIntHost(floor(yhi - 0.5));
  r.y0 = y0 > 0 ? y0 : 0;
  r.y1 = y1 < H - 1 ? y1 : H - 1;
}

// Per-polygon row counts are scanned into 64-bit span offsets in three
// launches over tiles of kScanTile polygons (counts and tile sums, one block
// scanning the tile sums, per-tile offsets). A polygon covers at most 32768
// rows and a list at most 2^30 polygons, so 64-bit offsets cannot wrap.
constexpr int kScanThreads = 256, kScanItems = 4, kScanTile = kScanThreads * kScanItems;

// Exclusive scan of one value per thread across the block (wave64 shuffles,
// then the waves' totals through LDS); *total receives the block's sum.
// Every thread of the block must call it.
__device__ __forceinline__ uint64_t blockExclusiveScan(uint64_t v, uint64_t* lds, uint64_t* total) {
  const int lane = int(threadIdx.x) & 63, wave = int(threadIdx.x) >> 6, waves = int(blockDim.x) >> 6;
  uint64_t x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint64_t y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  if (lane == 63) lds[wave] = x;
  __syncthreads();
  uint64_t before = 0, all = 0;
  for (int w = 0; w < waves; ++w) {
    const uint64_t s = lds[w];
    before += w < wave ? s : 0;
    all += s;
  }
  __syncthreads();
  *total = all;
  return before + x - v;
}

// Rows a polygon covers (0 if it is filtered out): valid output of its pass
// (tag == epoch of its root, when the epoch filter is given), a hit polygon
// unless this is the shadow pass, 3 or 4 corners. Also each tile's row sum.
__global__ void __launch_bounds__(kScanThreads) rasterRowsKernel(const OutPoly* polys, const int* root, const int* tag,
                                                                 const int* epoch, int n, float halfW, float halfH,
                                                                 int W, int H, int shadow, uint32_t* rows,
                                                                 uint64_t* tileSum) {
  __shared__ uint64_t lds[kScanThreads / 64];
  const int64_t base = int64_t(blockIdx.x) * kScanTile + int64_t(threadIdx.x) * kScanItems;
  uint64_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const int64_t i = base + k;
    if (i >= n) break;
    const OutPoly pg = polys[i];
    bool ok = pg.n >= 3 && pg.n <= 4 && (shadow || pg.tri >= 0);
    if (ok && epoch) ok = tag[i] == epoch[root[i]];
    uint32_t c = 0;
    if (ok) {
      RasterEdges r;
      rasterEdges(pg, halfW, halfH, W, H, r);
      if (r.y1 >= r.y0) c = uint32_t(r.y1 - r.y0 + 1);
    }
    rows[i] = c;
    s += c;
  }
  uint64_t total;
  blockExclusiveScan(s, lds, &total);
  if (threadIdx.x == 0) tileSum[blockIdx.x] = total;
}

// Exclusive scan of the tile sums in place by one block; tileSum[tiles]
// receives the number of spans.
__global__ void __launch_bounds__(1024) rasterTileScanKernel(uint64_t* tileSum, int tiles) {
  __shared__ uint64_t lds[1024 / 64];
  uint64_t carry = 0;
  for (int c = 0; c < tiles; c += int(blockDim.x)) {
    const int i = c + int(threadIdx.x);
    const uint64_t v = i < tiles ? tileSum[i] : 0;
    uint64_t total;
    const uint64_t ex = blockExclusiveScan(v, lds, &total);
    if (i < tiles) tileSum[i] = carry + ex;
    carry += total;
  }
  if (threadIdx.x == 0) tileSum[tiles] = carry;
}

// First span of each polygon: the tile's offset plus the scan within the tile.
__global__ void __launch_bounds__(kScanThreads) rasterOffsKernel(const uint32_t* rows, const uint64_t* tileSum, int n,
                                                                 uint64_t* offs) {
  __shared__ uint64_t lds[kScanThreads / 64];
  const int64_t base = int64_t(blockIdx.x) * kScanTile + int64_t(threadIdx.x) * kScanItems;
  uint32_t r[kScanItems];
  uint64_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    r[k] = base + k < n ? rows[base + k] : 0u;
    s += r[k];
  }
  uint64_t total;
  uint64_t ex = blockExclusiveScan(s, lds, &total) + tileSum[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    if (base + k < n) offs[base + k] = ex;
    ex += r[k];
  }
}

// One (polygon, row) span per thread: the polygon is the last one whose first
// span offset is <= the item (binary search over the exclusive scan).
__global__ void rasterSpansKernel(const OutPoly* polys, const uint64_t* offs, int n, uint32_t total, float halfW,
                                  float halfH, int W, int H, int shadow, int32_t* ids, uint8_t* mask) {
  const uint32_t j = uint32_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= total) return;
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (offs[mid] <= uint64_t(j))
      lo = mid;
    else
      hi = mid - 1;
  }
  const OutPoly pg = polys[lo];
  RasterEdges r;
  rasterEdges(pg, halfW, halfH, W, H, r);
  const uint64_t k = uint64_t(j) - offs[lo];
  if (r.y1 < r.y0 || k > uint64_t(r.y1 - r.y0)) return;  // not with a consistent scan: never leave the rows
  const int y = r.y0 + int(k);
  const double sy = y + 0.5;
  double xl = -1e300, xr = 1e300;
  bool empty = false;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    if (i >= r.n || empty) break;
    const double rr = r.B[i] * sy + r.C[i];
    if (r.A[i] > 0) {
      const double v = -rr / r.A[i];
      xl = xl < v ? v : xl;  // std::max(xl, v)
    } else if (r.A[i] < 0) {
      const double v = -rr / r.A[i];
      xr = v < xr ? v : xr;  // std::min(xr, v)
    } else if (rr < 0) {
      empty = true;
    }
  }
  if (empty || xl > xr) return;
  int x0 = toIntHost(ceil(xl - 0.5)), x1 = toIntHost(floor(xr - 0.5));
  x0 = x0 > 0 ? x0 : 0;
  x1 = x1 < W - 1 ? x1 : W - 1;
  const size_t row = size_t(y) * size_t(W);
  for (int x = x0; x <= x1; ++x) {
    if (shadow)
      mask[row + size_t(x)] = 1;
    else
      atomicMax(&ids[row + size_t(x)], pg.tri);  // polygons sort by tri first: the canonical last writer
  }
}

__global__ void fillInt(int32_t* p, size_t n, int32_t v) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) p[i] = v;
}

__global__ void fillByte(uint8_t* p, size_t n) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) p[i] = 0;
}

// Valid polygons of the last pass (tag == epoch[root]), compacted.
__global__ void compactPolys(const OutPoly* polys, const int* root, const int* tag, const int* epoch, int n,
                             OutPoly* dst, int* count) {
  const int i = int(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n || tag[i] != epoch[root[i]]) return;
  dst[atomicAdd(count, 1)] = polys[i];
}

// Device bytes of the scan scratch for n polygons: row counts (4 B each),
// span offsets (8 B each) and tile sums (8 B per tile of kScanTile, one more).
size_t rasterScratchBytes(int64_t n) {
  const int64_t m = n > 0 ? n : 1;
  const int64_t tiles = (m + kScanTile - 1) / kScanTile;
  return size_t(((m * 4 + 7) & ~int64_t(7)) + m * 8 + (tiles + 1) * 8);
}

// Rasterizes n device polygons (epoch filter optional) into the device image
// `img` (ids or mask, W x H) on e.stream, with `scratch` (rasterScratchBytes(n)
// bytes of device memory) for the scan; allocates nothing. Returns after the
// span launch is queued (the span count is read back to size it).
int rasterInto(Engine& e, const HostCam& cam, const OutPoly* polys, const int* root, const int* tag, const int* epoch,
               int n, bool shadow, void* img, void* scratch, std::string& err) {
  const size_t np = size_t(cam.W) * size_t(cam.H);
  const int fillBlocks = int(std::min<size_t>((np + 255) / 256, 4096));
  if (shadow)
    fillByte<<<fillBlocks, 256, 0, e.stream>>>(static_cast<uint8_t*>(img), np);
  else
    fillInt<<<fillBlocks, 256, 0, e.stream>>>(static_cast<int32_t*>(img), np, -1);
  BTG_TRY(hipGetLastError());
  if (n <= 0) return 0;
  const int tiles = (n + kScanTile - 1) / kScanTile;
  uint32_t* rows = static_cast<uint32_t*>(scratch);
  uint64_t* offs = reinterpret_cast<uint64_t*>(static_cast<char*>(scratch) + ((int64_t(n) * 4 + 7) & ~int64_t(7)));
  uint64_t* tileSum = offs + n;
  rasterRowsKernel<<<tiles, kScanThreads, 0, e.stream>>>(polys, root, tag, epoch, n, cam.halfW, cam.halfH, cam.W,
                                                         cam.H, shadow ? 1 : 0, rows, tileSum);
  BTG_TRY(hipGetLastError());
  rasterTileScanKernel<<<1, 1024, 0, e.stream>>>(tileSum, tiles);
  BTG_TRY(hipGetLastError());
  rasterOffsKernel<<<tiles, kScanThreads, 0, e.stream>>>(rows, tileSum, n, offs);
  BTG_TRY(hipGetLastError());
  uint64_t total = 0;
  BTG_TRY(hipMemcpyAsync(&total, tileSum + tiles, 8, hipMemcpyDeviceToHost, e.stream));
  BTG_TRY(hipStreamSynchronize(e.stream));
  if (total > 0xffffffffull) {
    err = "too many polygon rows to rasterize (" + std::to_string(total) + " spans; at most 2^32 - 1)";
    return BTG_ERR_LIMIT_CODE;
  }
  if (total > 0) {
    rasterSpansKernel<<<unsigned((total + 255) / 256), 256, 0, e.stream>>>(
        polys, offs, n, uint32_t(total), cam.halfW, cam.halfH, cam.W, cam.H, shadow ? 1 : 0,
        static_cast<int32_t*>(img), static_cast<uint8_t*>(img));
    BTG_TRY(hipGetLastError());
  }
  return 0;
}

// Rasterizes n device polygons (epoch filter optional) into the context's
// device image and copies it to the host.
int rasterDevice(Engine& e, const HostCam& cam, const OutPoly* polys, const int* root, const int* tag,
                 const int* epoch, int n, bool shadow, void* hostOut, std::string& err) {
  const size_t np = size_t(cam.W) * size_t(cam.H);
  int rc;
  if ((rc = e.ensure(e.rasterImg, np * (shadow ? 1 : 4), err)) ||
      (rc = e.ensure(e.rasterRows, rasterScratchBytes(n), err)))
    return rc;
  if ((rc = rasterInto(e, cam, polys, root, tag, epoch, n, shadow, e.rasterImg.ptr, e.rasterRows.ptr, err))) return rc;
  BTG_TRY(hipMemcpyAsync(hostOut, e.rasterImg.ptr, np * (shadow ? 1 : 4), hipMemcpyDeviceToHost, e.stream));
  BTG_TRY(hipStreamSynchronize(e.stream));
  return 0;
}

}  // namespace

int engineRaster(Engine* e, const HostCam& cam, const OutPoly* polys, int64_t n, bool shadow, int32_t* ids,
                 uint8_t* mask, std::string& err) {
  if (cam.W <= 0 || cam.H <= 0 || n < 0 || n > (1LL << 30) || (shadow ? !mask : !ids)) {
    err = "invalid raster arguments";
    return BTG_ERR_ARG_CODE;
  }
  int rc = e->ensure(e->prim, size_t(std::max<int64_t>(n, 1)) * sizeof(OutPoly), err);
  if (rc) return rc;
  if (n > 0)
    BTG_TRY(hipMemcpyAsync(e->prim.ptr, polys, size_t(n) * sizeof(OutPoly), hipMemcpyHostToDevice, e->stream));
  return rasterDevice(*e, cam, static_cast<const OutPoly*>(e->prim.ptr), nullptr, nullptr, nullptr, int(n), shadow,
                      shadow ? static_cast<void*>(mask) : static_cast<void*>(ids), err);
}

int engineRasterNoAlloc(Engine* e, const HostCam& cam, const OutPoly* polysDev, int64_t n, bool shadow,
                        int32_t* idsDev, uint8_t* maskDev, std::string& err) {
  if (cam.W <= 0 || cam.H <= 0 || n < 0 || (n > 0 && !polysDev) || (shadow ? !maskDev : !idsDev)) {
    err = "invalid raster arguments";
    return BTG_ERR_ARG_CODE;
  }
  if (n > kNoAllocMaxPolys || rasterScratchBytes(n) > sizeof(noAllocScratch)) {
    err = "allocation-free rasterization takes at most " + std::to_string(kNoAllocMaxPolys) + " polygons (got " +
          std::to_string(n) + ")";
    return BTG_ERR_LIMIT_CODE;
  }
  if (!e->noAllocScratch) {
    err = "the static rasterization scratch has no address on this device";
    return BTG_ERR_DEVICE_CODE;
  }
  const int rc = rasterInto(*e, cam, polysDev, nullptr, nullptr, nullptr, int(n), shadow,
                            shadow ? static_cast<void*>(maskDev) : static_cast<void*>(idsDev), e->noAllocScratch, err);
  if (rc) return rc;
  BTG_TRY(hipStreamSynchronize(e->stream));
  return 0;
}

int engineImages(Engine* e, const HostCam& cam, bool cull, int tile, const F3* light, int32_t* ids, uint8_t* mask,
                 RunInfo& prim, RunInfo& point, double& rasterMs, std::string& err) {
  if (tile <= 0 || cam.W <= 0 || cam.H <= 0 || !ids || (light && !mask)) {
    err = "invalid image arguments";
    return BTG_ERR_ARG_CODE;
  }
  const int tilesW = (cam.W + tile - 1) / tile, tilesH = (cam.H + tile - 1) / tile, R = tilesW * tilesH;
  const size_t outCap = outCapFor(*e, size_t(cam.W) * size_t(cam.H) * 2);
  auto setup = [&](Params& p) {
    p.tile = tile;
    p.tilesW = tilesW;
    p.cull = cull ? 1 : 0;
  };
  std::vector<OutPoly> none;
  int rc = runPolys(*e, kPrimary, R, e->opt.primaryBudget, outCap, cam, setup, none, prim, err, false);
  if (rc) return rc;
  const int nOut = std::min(e->hostState->outCount, int(e->out.bytes / sizeof(OutPoly)));
  auto t0 = std::chrono::steady_clock::now();
  if ((rc = rasterDevice(*e, cam, static_cast<const OutPoly*>(e->out.ptr), static_cast<const int*>(e->outRoot.ptr),
                         static_cast<const int*>(e->outTag.ptr), static_cast<const int*>(e->epoch.ptr), nOut, false,
                         ids, err)))
    return rc;
  rasterMs = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (!light) return 0;
  // the primary hits become the point pass's roots without leaving the device
  if ((rc = e->ensure(e->prim, size_t(std::max(nOut, 1)) * sizeof(OutPoly), err)) ||
      (rc = e->ensure(e->listCount, 4, err)))
    return rc;
  clearInt<<<1, 64, 0, e->stream>>>(static_cast<int*>(e->listCount.ptr));
  if (nOut > 0)
    compactPolys<<<(nOut + 255) / 256, 256, 0, e->stream>>>(
        static_cast<const OutPoly*>(e->out.ptr), static_cast<const int*>(e->outRoot.ptr),
        static_cast<const int*>(e->outTag.ptr), static_cast<const int*>(e->epoch.ptr), nOut,
        static_cast<OutPoly*>(e->prim.ptr), static_cast<int*>(e->listCount.ptr));
  BTG_TRY(hipGetLastError());
  int nPrim = 0;
  BTG_TRY(hipMemcpyAsync(&nPrim, e->listCount.ptr, 4, hipMemcpyDeviceToHost, e->stream));
  BTG_TRY(hipStreamSynchronize(e->stream));
  const size_t np = size_t(cam.W) * size_t(cam.H);
  if (nPrim == 0) {
    std::memset(mask, 0, np);
    return 0;
  }
  const size_t pcap = outCapFor(*e, size_t(nPrim) * 4);
  auto psetup = [&](Params& p) {
    p.pointLight = *light;
    p.prim = static_cast<const OutPoly*>(e->prim.ptr);
  };
  if ((rc = runPolys(*e, kPoint, nPrim, e->opt.budget, pcap, cam, psetup, none, point, err, false))) return rc;
  const int nShadow = std::min(e->hostState->outCount, int(e->out.bytes / sizeof(OutPoly)));
  t0 = std::chrono::steady_clock::now();
  rc = rasterDevice(*e, cam, static_cast<const OutPoly*>(e->out.ptr), static_cast<const int*>(e->outRoot.ptr),
                    static_cast<const int*>(e->outTag.ptr), static_cast<const int*>(e->epoch.ptr), nShadow, true, mask,
                    err);
  rasterMs += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return rc;
}

}  // namespace btg
