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
  Box bounds;                      // kd root box
  int numTris = 0;
};

// Per-call run statistics (counters of src/util/stats.h:10-42 that the
// engine produces).
enum StatId { sKd, sLeaf, sTests, sHits, sSplits, sBeams, sRoots, sPresplit, sFive, sClimbs, sRays, sNum };

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
  int64_t capacity = int64_t(1) << 25;
  int roundsPerCheck = 8;
  int leafCap = 8;
  bool profile = false;
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

// Creates the engine's device state for a scene (copies the arrays); the
// scene becomes resident scene 0. `offset` is the scene's soft-shadow apex
// offset (scene extent x the split epsilon).
Engine* engineCreate(const btg_scene_desc& scene, float offset, std::string& err);
void engineDestroy(Engine* e);
Options& engineOptions(Engine* e);

// Adds a resident scene (copies the arrays); returns its id >= 1, or a
// negative error code. Removing frees the id for reuse.
int engineAddScene(Engine* e, const btg_scene_desc& scene, float offset, std::string& err);
int engineRemoveScene(Engine* e, int scene, std::string& err);
// Device bytes a resident scene occupies (-1: no such scene).
int64_t engineSceneBytes(const Engine* e, int scene);

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
};
// Many jobs in one persistent wavefront pass over all their roots.
int engineSoftBatch(Engine* e, const SoftJobH* jobs, int n, bool exact, RunInfo& info, std::string& err);
// One job on scene 0.
int engineSoft(Engine* e, const HostCam& cam, const HostLight& L, bool exact, int32_t* tri, float* vis, float* E,
               float* Eu, RunInfo& info, std::string& err);
// Primary visibility, one root per tile: hit polygons.
int enginePrimary(Engine* e, const HostCam& cam, bool cull, int tile, std::vector<OutPoly>& out, RunInfo& info,
                  std::string& err);
// Point shadows for primary hit polygons.
int enginePoint(Engine* e, const HostCam& cam, const OutPoly* prim, int64_t nPrim, F3 light,
                std::vector<OutPoly>& out, RunInfo& info, std::string& err);

}  // namespace btg
