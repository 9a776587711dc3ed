// Host-side interface of the wavefront engine (engine.hip), used by the C ABI
// (api.hip). Device layouts of the scene live here too.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "beam_dev.h"
#include "btg.h"

namespace btg {

// kd node as uploaded (src/accel/kdtree.h:13-21): 12 bytes.
struct KdNodeD {
  float split;
  uint32_t data;    // low 2 bits axis (3 = leaf), high bits right child / count
  uint32_t offset;  // leaf: first leaf reference
};
static_assert(sizeof(KdNodeD) == 12, "kd node layout");

// A triangle as referenced from a leaf, packed for one 64-byte load: three
// vertices, unit normal, global vertex indices (orient shared edges
// canonically) and the triangle id (TriRef, src/beam/beam_types.h:78-83).
struct alignas(16) LeafTri {
  float v[9];
  float n[3];
  uint32_t vi[3];
  int32_t tri;
};
static_assert(sizeof(LeafTri) == 64, "leaf triangle layout");

// Scene + kd-tree in device memory.
struct DevScene {
  const float4* triNV = nullptr;   // per triangle: (normal, 0), (vertex 0, 0)
  const KdNodeD* nodes = nullptr;
  const Box* nodeBox = nullptr;
  const uint32_t* parent = nullptr;
  const LeafTri* refs = nullptr;   // in leaf order (kd triIndices order)
  const int32_t* refTri = nullptr; // the triangle id of each leaf reference (the mailbox's 4-byte read)
  Box bounds;                      // kd root box
  int numTris = 0;
};

// Soft-shadow roots are the pixels of tiles of 64 (kRootTileW x kRootTileH;
// build knob BTG_ROOT_TILE_W = 4, 8, 16 or 32), one wave of the admit kernel
// per tile; partial tiles at the image's right and bottom edges keep their
// 64 root slots (the ones outside the image have no beam).
#ifndef BTG_ROOT_TILE_W
#define BTG_ROOT_TILE_W 8
#endif
constexpr int kRootTileW = BTG_ROOT_TILE_W, kRootTileH = 64 / kRootTileW;
static_assert(kRootTileW * kRootTileH == 64, "a root tile is one wave of 64 pixels");
inline int64_t rootTilesX(int64_t W) { return (W + kRootTileW - 1) / kRootTileW; }
inline int64_t softRootCount(int64_t W, int64_t H) {
  return rootTilesX(W) * ((H + kRootTileH - 1) / kRootTileH) * 64;
}

// Per-call run statistics (counters of src/util/stats.h:10-42 that the
// engine produces).
enum StatId { sKd, sLeaf, sTests, sHits, sSplits, sBeams, sRoots, sPresplit, sFive, sClimbs, sRays, sMail, sNum };

struct RunInfo {
  double primaryMs = 0, traceMs = 0;
  int rounds = 0, maxQueue = 0, killed = 0, retries = 0;
  uint64_t stats[sNum] = {};
  uint64_t rayStats[sNum] = {};
  double dropped = 0;
  double kernelMs[5] = {};
  int launches[5] = {};
  uint64_t phaseCycles[8] = {};  // trace loop clocks per region (BTG_PHASE_CYCLES builds)
};

struct Options {
  int budget = 1024;
  int primaryBudget = 16384;
  int64_t capacity = int64_t(1) << 25;  // set by device memory at creation (engineCreate)
  int roundsPerCheck = 8;
  int leafCap = 8;
  int precision = 0;  // soft shadows: 0 float build (bit-exact), 1 FP64-guarded
  int mailbox = 0;  // 1: skip exact repeats of a triangle test by one sub-beam (results unchanged)
  bool profile = false;
  bool graphs = true;  // replay each group of rounds from a captured HIP graph
};

struct HostCam {
  F3 eye, fwd, right, down;
  float halfW, halfH;
  int W, H;
};

struct HostLight {
  F3 c, U, V;
};

struct OutPoly {
  float x[4], y[4];
  int32_t n, tri;
  float area;
};
static_assert(sizeof(OutPoly) == 44, "polygon layout");

class Engine;

// rasterize (src/render/render.cpp:49-95) on the device, one sample at each
// pixel centre: shadow = false writes ids (the tri of the canonical last
// polygon covering each pixel centre, -1 for none; tri < 0 polygons skipped),
// shadow = true writes mask (1 where any polygon covers the centre).
int engineRaster(Engine* e, const HostCam& cam, const OutPoly* polys, int64_t n, bool shadow, int32_t* ids,
                 uint8_t* mask, std::string& err);
// btg_raster of device polygons into a device image with the per-device
// static scratch (no allocation); n <= kNoAllocMaxPolys.
int engineRasterNoAlloc(Engine* e, const HostCam& cam, const OutPoly* polysDev, int64_t n, bool shadow,
                        int32_t* idsDev, uint8_t* maskDev, std::string& err);
constexpr int64_t kNoAllocMaxPolys = int64_t(1) << 20;
// Primary visibility and (light != null) point shadows of one frame with both
// rasterizations on the device: the primary polygons feed the ID image and
// the point pass without a host round trip. rasterMs: wall time of the two
// rasterizations including their downloads.
int engineImages(Engine* e, const HostCam& cam, bool cull, int tile, const F3* light, int32_t* ids,
                 uint8_t* mask, RunInfo& prim, RunInfo& point, double& rasterMs, std::string& err);

// Creates the engine's device state for a scene (copies the arrays); the
// scene becomes resident scene 0. `offset` is the scene's soft-shadow apex
// offset (scene extent x the split epsilon), `offsetGuard` the double
// build's (FP64-guarded mode: with the hit point in double and the apex as a
// float pair).
Engine* engineCreate(const btg_scene_desc& scene, float offset, float offsetGuard, std::string& err);
void engineDestroy(Engine* e);
Options& engineOptions(Engine* e);

// Adds a resident scene (copies the arrays); returns its id >= 1, or a
// negative error code. Removing frees the id for reuse.
int engineAddScene(Engine* e, const btg_scene_desc& scene, float offset, float offsetGuard, std::string& err);
int engineRemoveScene(Engine* e, int scene, std::string& err);
// Device bytes a resident scene occupies (-1: no such scene).
int64_t engineSceneBytes(const Engine* e, int scene);
// Device bytes of one root's state (its query, live count, state, epoch and
// exact sums) and of one soft-shadow pixel's images and hit point (guard: the
// FP64-guarded mode's extra low word of the hit point).
int64_t rootStateBytes();
int64_t softPixelBytes(bool guard);

// A soft-shadow job: resident scene, camera, light, host outputs (any may be
// null: the image stays in device memory), W*H each.
struct SoftJobH {
  int scene;
  HostCam cam;
  HostLight light;
  int32_t* tri;
  float* vis;
  float* E;
  float* Eu;
  const double* camD = nullptr;  // FP64 mode: eye, fwd, right, down, halfW, halfH in double (null: cam widened)
};
// Many jobs in one persistent wavefront pass over all their roots.
int engineSoftBatch(Engine* e, const SoftJobH* jobs, int n, bool exact, RunInfo& info, std::string& err);
// One job on scene 0.
int engineSoft(Engine* e, const HostCam& cam, const HostLight& L, bool exact, int32_t* tri, float* vis, float* E,
               float* Eu, RunInfo& info, std::string& err);
// An engine with its own stream and buffers over another engine's resident
// scenes (which must outlive it and not change while it runs).
Engine* engineShare(Engine* base, std::string& err);

// Open soft-shadow passes (continuous batching): a persistent wavefront pass
// to which jobs are appended while it runs. Capacities: roots, pixels and
// jobs the pass holds until it is reset (after all its jobs finished).
struct OpenPass;
// queueCap > 0 caps the pass's sub-beam queue capacity below the base's.
OpenPass* openCreate(Engine* base, int64_t rootCap, int64_t pixCap, int jobCap, bool exact, int64_t queueCap,
                     std::string& err);
void openDestroy(OpenPass* op);
// Appends a job (launches its primary rays); 1 = added as *slot, 0 = no room
// left in this pass, negative = error.
int openTryAdd(OpenPass* op, const SoftJobH& job, int* slot, std::string& err);
// Launches `rounds` rounds and an asynchronous read-back of the job counters.
int openStep(OpenPass* op, int rounds, std::string& err);
// Waits for the last step; finalizes and downloads the jobs that completed
// since (their slots, and per slot the roots that went over budget).
int openCollect(OpenPass* op, std::vector<int>& completed, std::vector<int>& killed, std::string& err);
int openJobs(const OpenPass* op);
bool openIdle(const OpenPass* op);
int openReset(OpenPass* op, std::string& err);

// Primary visibility, one root per tile: hit polygons.
int enginePrimary(Engine* e, const HostCam& cam, bool cull, int tile, std::vector<OutPoly>& out, RunInfo& info,
                  std::string& err);
// Point shadows for primary hit polygons.
int enginePoint(Engine* e, const HostCam& cam, const OutPoly* prim, int64_t nPrim, F3 light,
                std::vector<OutPoly>& out, RunInfo& info, std::string& err);

}  // namespace btg
