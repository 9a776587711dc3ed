// The C ABI of include/btg.h on top of the wavefront engine (engine.hip).
#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <exception>
#include <mutex>
#include <new>
#include <string>
#include <system_error>
#include <thread>
#include <vector>
#include <unordered_map>
#include <vector>

#include "../host/scene_file.h"
#include "engine.h"

struct btg_ctx {
  btg::Engine* engine = nullptr;
  std::string err;
  float offset = 0;
};

struct btg_polys {
  std::vector<btg::OutPoly> v;
};

namespace {

thread_local std::string g_err;

int failWith(btg_ctx* c, int code, const std::string& m) {
  if (c)
    c->err = m;
  else
    g_err = m;
  return code;
}

btg::HostCam camOf(const btg_camera& c) {
  btg::HostCam h;
  h.eye = btg::f3(c.eye[0], c.eye[1], c.eye[2]);
  h.fwd = btg::f3(c.fwd[0], c.fwd[1], c.fwd[2]);
  h.right = btg::f3(c.right[0], c.right[1], c.right[2]);
  h.down = btg::f3(c.down[0], c.down[1], c.down[2]);
  h.halfW = c.halfW;
  h.halfH = c.halfH;
  h.W = c.W;
  h.H = c.H;
  return h;
}

void fillInfo(const btg::RunInfo& r, double wall, btg_run_info* info) {
  if (!info) return;
  std::memset(info, 0, sizeof(*info));
  info->primarySeconds = r.primaryMs * 1e-3;
  info->traceSeconds = r.traceMs * 1e-3;
  info->totalSeconds = wall;
  info->rounds = r.rounds;
  info->maxQueue = r.maxQueue;
  info->killedRoots = r.killed;
  info->retryPasses = r.retries;
  auto put = [](const uint64_t* s, btg_trace_stats& t) {
    t.kdSteps = s[btg::sKd];
    t.leafVisits = s[btg::sLeaf];
    t.triTests = s[btg::sTests];
    t.hits = s[btg::sHits];
    t.splits = s[btg::sSplits];
    t.beams = s[btg::sBeams];
    t.rootBeams = s[btg::sRoots];
    t.presplitBeams = s[btg::sPresplit];
    t.fiveSplits = s[btg::sFive];
    t.climbs = s[btg::sClimbs];
    t.rays = s[btg::sRays];
    t.mailboxSkips = s[btg::sMail];
  };
  put(r.stats, info->stats);
  info->stats.droppedArea = r.dropped;
  put(r.rayStats, info->rayStats);
  for (int k = 0; k < 5; ++k) {
    info->kernelSeconds[k] = r.kernelMs[k] * 1e-3;
    info->kernelLaunches[k] = r.launches[k];
  }
  for (int k = 0; k < 8; ++k) info->phaseCycles[k] = r.phaseCycles[k];
}

double now() {
  timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return double(ts.tv_sec) + 1e-9 * double(ts.tv_nsec);
}

// Image sides up to 2^15 keep pixel, tile and polygon counts inside int.
constexpr int kMaxSide = 1 << 15;
bool validCam(const btg_camera* c) {
  return c && c->W > 0 && c->H > 0 && c->W <= kMaxSide && c->H <= kMaxSide && c->halfW > 0 && c->halfH > 0;
}

}  // namespace

extern "C" {

int btg_device_query(btg_device_info* info) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    (void)hipGetLastError();
    return 0;
  }
  if (info) {
    hipDeviceProp_t p;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&p, dev) != hipSuccess) return 0;
    std::memset(info, 0, sizeof(*info));
    std::strncpy(info->name, p.name, sizeof(info->name) - 1);
    std::strncpy(info->arch, p.gcnArchName, sizeof(info->arch) - 1);
    info->computeUnits = p.multiProcessorCount;
    info->wavefrontSize = p.warpSize;
    info->memGB = double(p.totalGlobalMem) / double(1ull << 30);
  }
  return 1;
}

const char* btg_last_error(const btg_ctx* ctx) { return ctx ? ctx->err.c_str() : g_err.c_str(); }

namespace {

// Validates a scene description on the host: the kernels trust its indices.
bool validScene(const btg_scene_desc* d, std::string& why) {
  if (d->numTris <= 0 || d->numNodes <= 0 || !d->pos || !d->tri3 || !d->triN || !d->nodes || !d->parent ||
      !d->nodeBox || (d->numRefs > 0 && !d->triIndices)) {
    why = "scene description is incomplete (no triangles or no kd-tree)";
    return false;
  }
  if (d->numTris >= (1LL << 31) || d->numNodes >= (1LL << 30) || d->numRefs >= (1LL << 32)) {
    why = "scene too large for 32-bit indices";
    return false;
  }
  for (int64_t t = 0; t < 3 * d->numTris; ++t)
    if (d->tri3[t] >= uint64_t(d->numVerts)) {
      why = "vertex index out of range";
      return false;
    }
  for (int64_t i = 0; i < d->numRefs; ++i)
    if (d->triIndices[i] >= uint64_t(d->numTris)) {
      why = "leaf triangle out of range";
      return false;
    }
  const btg::KdNodeD* nodes = static_cast<const btg::KdNodeD*>(d->nodes);
  for (int64_t i = 0; i < d->numNodes; ++i) {
    const uint32_t data = nodes[i].data;
    if ((data & 3u) == 3u) {
      if (uint64_t(nodes[i].offset) + (data >> 2) > uint64_t(d->numRefs)) {
        why = "kd leaf references out of range";
        return false;
      }
    } else if ((data >> 2) <= uint64_t(i) + 1 || (data >> 2) >= uint64_t(d->numNodes)) {
      why = "kd child index out of range";
      return false;
    }
    if (i > 0 && d->parent[i] >= uint64_t(d->numNodes)) {
      why = "kd parent out of range";
      return false;
    }
  }
  // Children always follow their parent (left = i + 1, right > i + 1), so one
  // forward pass gives every reachable node its depth. The restart trail holds
  // one bit per inner level (64 bits): inner nodes deeper than 63 are refused.
  // Each child's parent link must name the node that references it (the climb
  // walks those links), which also rules out a node with two parents.
  std::vector<int32_t> depth(size_t(d->numNodes), -1);
  depth[0] = 0;
  for (int64_t i = 0; i < d->numNodes; ++i) {
    const uint32_t data = nodes[i].data;
    if ((data & 3u) == 3u || depth[size_t(i)] < 0) continue;
    if (depth[size_t(i)] > 63) {
      why = "kd-tree deeper than 63 inner levels";
      return false;
    }
    const uint32_t kids[2] = {uint32_t(i + 1), data >> 2};
    for (uint32_t c : kids) {
      if (d->parent[c] != uint32_t(i) || depth[c] >= 0) {
        why = "kd parent links do not match the children";
        return false;
      }
      depth[c] = depth[size_t(i)] + 1;
    }
  }
  return true;
}

// Apex offset of the soft-shadow beams: |scene diagonal| * kEpsRel
// (cuda/gpu_render.cu:574).
float sceneOffset(const btg_scene_desc* d) {
  const btg::F3 diag = btg::f3(d->sceneBounds[3] - d->sceneBounds[0], d->sceneBounds[4] - d->sceneBounds[1],
                               d->sceneBounds[5] - d->sceneBounds[2]);
  return btg::length(diag) * btg::kEpsRel;
}

// The double build's apex offset (kEpsRel = 1e-9, src/core/real.h:31-34), used
// by FP64-guarded soft shadows.
float sceneOffsetGuard(const btg_scene_desc* d) {
  const btg::F3 diag = btg::f3(d->sceneBounds[3] - d->sceneBounds[0], d->sceneBounds[4] - d->sceneBounds[1],
                               d->sceneBounds[5] - d->sceneBounds[2]);
  return btg::length(diag) * btg::kEpsRelDouble;
}

btg::HostLight lightOf(const btg_area_light& l) {
  btg::HostLight L;
  L.c = btg::f3(l.c[0], l.c[1], l.c[2]);
  L.U = btg::f3(l.U[0], l.U[1], l.U[2]);
  L.V = btg::f3(l.V[0], l.V[1], l.V[2]);
  return L;
}

}  // namespace

int btg_create(btg_ctx** out, const btg_scene_desc* d) {
  if (!out || !d) return failWith(nullptr, BTG_ERR_ARG, "null argument");
  *out = nullptr;
  std::string why;
  if (!validScene(d, why)) return failWith(nullptr, BTG_ERR_ARG, why);
  if (!btg_device_query(nullptr)) return failWith(nullptr, BTG_ERR_DEVICE, "no HIP device available");
  try {
    btg_ctx* c = new btg_ctx();
    std::string err;
    c->offset = sceneOffset(d);
    c->engine = btg::engineCreate(*d, c->offset, sceneOffsetGuard(d), err);
    if (!c->engine) {
      delete c;
      return failWith(nullptr, BTG_ERR_DEVICE, err);
    }
    *out = c;
    return BTG_OK;
  } catch (const std::bad_alloc&) {
    return failWith(nullptr, BTG_ERR_MEMORY, "host allocation failed");
  }
}

int btg_scene_add(btg_ctx* ctx, const btg_scene_desc* d, int* sceneId) {
  if (!ctx || !d || !sceneId) return failWith(ctx, BTG_ERR_ARG, "null argument");
  std::string why;
  if (!validScene(d, why)) return failWith(ctx, BTG_ERR_ARG, why);
  try {
    const int id = btg::engineAddScene(ctx->engine, *d, sceneOffset(d), sceneOffsetGuard(d), ctx->err);
    if (id < 0) return id;
    *sceneId = id;
    return BTG_OK;
  } catch (const std::bad_alloc&) {
    return failWith(ctx, BTG_ERR_MEMORY, "host allocation failed");
  }
}

namespace {
// A scene file's arrays as a scene description (the description borrows them).
int readSceneFile(const char* path, btsf::Data& d, btg_scene_desc& s, std::string& err) {
  if (!path) {
    err = "null path";
    return BTG_ERR_ARG;
  }
  if (!btsf::read(path, d, err)) return BTG_ERR_ARG;
  if (d.h.numNodes == 0) {
    err = std::string("scene file ") + path + " holds no kd-tree";
    return BTG_ERR_ARG;
  }
  s = btg_scene_desc{};
  s.pos = d.pos.data();
  s.numVerts = d.h.numVerts;
  s.tri3 = d.tri3.data();
  s.triN = d.triN.data();
  s.numTris = d.h.numTris;
  s.nodes = d.nodes.data();
  s.numNodes = d.h.numNodes;
  s.triIndices = d.triIndices.data();
  s.numRefs = d.h.numRefs;
  s.parent = d.parent.data();
  s.nodeBox = d.nodeBox.data();
  std::memcpy(s.kdBounds, d.h.kdBounds, sizeof s.kdBounds);
  std::memcpy(s.sceneBounds, d.h.sceneBounds, sizeof s.sceneBounds);
  return BTG_OK;
}
}  // namespace

int btg_create_file(btg_ctx** out, const char* path) {
  if (!out) return failWith(nullptr, BTG_ERR_ARG, "null argument");
  *out = nullptr;
  try {
    btsf::Data d;
    btg_scene_desc s;
    std::string err;
    if (int rc = readSceneFile(path, d, s, err)) return failWith(nullptr, rc, err);
    return btg_create(out, &s);
  } catch (const std::bad_alloc&) {
    return failWith(nullptr, BTG_ERR_MEMORY, "host allocation failed");
  }
}

int btg_scene_add_file(btg_ctx* ctx, const char* path, int* sceneId) {
  if (!ctx || !sceneId) return failWith(ctx, BTG_ERR_ARG, "null argument");
  try {
    btsf::Data d;
    btg_scene_desc s;
    if (int rc = readSceneFile(path, d, s, ctx->err)) return rc;
    return btg_scene_add(ctx, &s, sceneId);
  } catch (const std::bad_alloc&) {
    return failWith(ctx, BTG_ERR_MEMORY, "host allocation failed");
  }
}

int btg_scene_remove(btg_ctx* ctx, int sceneId) {
  if (!ctx) return failWith(nullptr, BTG_ERR_ARG, "null context");
  return btg::engineRemoveScene(ctx->engine, sceneId, ctx->err);
}

int btg_footprint(const btg_ctx* ctx, const btg_request* req, int64_t* bytes) {
  if (!ctx || !req || !bytes || !validCam(&req->cam)) return BTG_ERR_ARG;
  const int64_t W = req->cam.W, H = req->cam.H;
  if (req->kind == BTG_REQ_SOFT) {
    // roots: 64 per tile of pixels, partial tiles included (the soft passes' layout)
    const int64_t roots = btg::softRootCount(W, H);
    *bytes = W * H * btg::softPixelBytes(btg::engineOptions(ctx->engine).precision != 0) +
             roots * btg::rootStateBytes();
    return BTG_OK;
  }
  if (req->kind == BTG_REQ_POINT) {
    // the primary pass's roots (16 x 16 tiles) and polygon buffer (polygon, root and tag)
    const int64_t roots = ((W + 15) / 16) * ((H + 15) / 16);
    const int64_t polys = std::max<int64_t>(int64_t(1) << 20, 2 * W * H);
    *bytes = roots * btg::rootStateBytes() + polys * int64_t(sizeof(btg::OutPoly) + 8);
    return BTG_OK;
  }
  return BTG_ERR_ARG;
}

int64_t btg_scene_bytes(const btg_ctx* ctx, int sceneId) {
  return ctx ? btg::engineSceneBytes(ctx->engine, sceneId) : -1;
}

void btg_destroy(btg_ctx* ctx) {
  if (!ctx) return;
  btg::engineDestroy(ctx->engine);
  delete ctx;
}

int btg_get_options(const btg_ctx* ctx, btg_options* o) {
  if (!ctx || !o) return BTG_ERR_ARG;
  const btg::Options& x = btg::engineOptions(ctx->engine);
  o->budget = x.budget;
  o->primaryBudget = x.primaryBudget;
  o->capacity = x.capacity;
  o->roundsPerCheck = x.roundsPerCheck;
  o->profileKernels = x.profile ? 1 : 0;
  o->leafCap = x.leafCap;
  o->precision = x.precision;
  o->useGraphs = x.graphs ? 1 : 0;
  o->mailbox = x.mailbox;
  return BTG_OK;
}

int btg_set_options(btg_ctx* ctx, const btg_options* o) {
  if (!ctx || !o) return BTG_ERR_ARG;
  if (o->budget < 1 || o->primaryBudget < 1 || o->capacity < 1024 || o->capacity > (int64_t(1) << 31) ||
      o->roundsPerCheck < 1 || o->leafCap < 0 || o->precision < 0 || o->precision > 2 || o->useGraphs < 0 ||
      o->useGraphs > 1 || o->mailbox < 0 || o->mailbox > 1)
    return failWith(ctx, BTG_ERR_ARG, "invalid options");
  btg::Options& x = btg::engineOptions(ctx->engine);
  x.budget = o->budget;
  x.primaryBudget = o->primaryBudget;
  x.capacity = o->capacity;
  x.roundsPerCheck = o->roundsPerCheck;
  x.profile = o->profileKernels != 0;
  x.leafCap = o->leafCap;
  x.precision = o->precision;
  x.graphs = o->useGraphs != 0;
  x.mailbox = o->mailbox;
  return BTG_OK;
}

int btg_soft(btg_ctx* ctx, const btg_camera* cam, const btg_area_light* light, int exact, int32_t* tri, float* vis,
             float* E, float* Eu, btg_run_info* info) {
  if (!ctx) return failWith(nullptr, BTG_ERR_ARG, "null context");
  if (!validCam(cam) || !light) return failWith(ctx, BTG_ERR_ARG, "invalid camera or light");
  const double t0 = now();
  btg::RunInfo ri;
  try {
    const int rc = btg::engineSoft(ctx->engine, camOf(*cam), lightOf(*light), exact != 0, tri, vis, E, Eu, ri, ctx->err);
    if (rc) return rc;
  } catch (const std::bad_alloc&) {
    return failWith(ctx, BTG_ERR_MEMORY, "host allocation failed");
  }
  fillInfo(ri, now() - t0, info);
  return BTG_OK;
}

int btg_soft_batch(btg_ctx* ctx, const btg_soft_job* jobs, int numJobs, int exact, btg_run_info* info) {
  if (!ctx) return failWith(nullptr, BTG_ERR_ARG, "null context");
  if (!jobs || numJobs <= 0) return failWith(ctx, BTG_ERR_ARG, "a batch needs at least one job");
  const double t0 = now();
  btg::RunInfo ri;
  try {
    std::vector<btg::SoftJobH> h(static_cast<size_t>(numJobs));
    for (int j = 0; j < numJobs; ++j) {
      if (!validCam(&jobs[j].cam)) return failWith(ctx, BTG_ERR_ARG, "job " + std::to_string(j) + ": invalid camera");
      h[size_t(j)] = btg::SoftJobH{jobs[j].scene, camOf(jobs[j].cam), lightOf(jobs[j].light), jobs[j].tri,
                                   jobs[j].vis, jobs[j].E, jobs[j].Eu,
                                   jobs[j].camD ? jobs[j].camD->eye : nullptr};
    }
    const int rc = btg::engineSoftBatch(ctx->engine, h.data(), numJobs, exact != 0, ri, ctx->err);
    if (rc) return rc;
  } catch (const std::bad_alloc&) {
    return failWith(ctx, BTG_ERR_MEMORY, "host allocation failed");
  }
  fillInfo(ri, now() - t0, info);
  return BTG_OK;
}

int btg_autotune(btg_ctx* ctx, const btg_soft_job* jobs, int numJobs, int exact, double maxSeconds,
                 btg_options* best, btg_autotune_trial* trials, int maxTrials, int* numTrials) {
  if (!ctx || !best) return failWith(ctx, BTG_ERR_ARG, "null argument");
  if (!jobs || numJobs <= 0) return failWith(ctx, BTG_ERR_ARG, "the tuning sample needs at least one job");
  btg::Options& opt = btg::engineOptions(ctx->engine);
  const btg::Options start = opt;
  // Coordinate descent over the schedule knobs (leaf cap, rounds between
  // host checks, mailboxing), one knob at a time, each
  // value timed on the sample (GPU time of the beam kernels and primary
  // rays); a value replaces the incumbent only when it is more than 2%
  // faster, since a scaled-down sample cannot resolve closer calls, and the
  // best value of a knob is kept before the next knob is tuned. No leaf cap
  // (0) is not tried: uncapped stragglers made every measured batch slower
  // (3x on the 64-job batch). The per-root budget is not tuned: it decides
  // which roots are abandoned and traced again, and how many roots the queue
  // capacity admits at once, so a scaled-down sample (fewer roots, retries
  // weighing more) misjudges it: 2048 won on the quarter-side sample of the
  // 64-job batch and lost by 7% on the batch itself.
  struct Knob {
    int btg::Options::*field;
    std::vector<int> values;
  };
  const Knob knobs[] = {{&btg::Options::leafCap, {4, 16, 2, 32, 8}},
                        {&btg::Options::roundsPerCheck, {4, 8, 16, 32}},
                        {&btg::Options::mailbox, {1, 0}}};
  std::vector<btg::SoftJobH> h(static_cast<size_t>(numJobs));
  for (int j = 0; j < numJobs; ++j) {
    if (!validCam(&jobs[j].cam)) return failWith(ctx, BTG_ERR_ARG, "job " + std::to_string(j) + ": invalid camera");
    h[size_t(j)] = btg::SoftJobH{jobs[j].scene, camOf(jobs[j].cam), lightOf(jobs[j].light), nullptr, nullptr,
                                 nullptr, nullptr};
  }
  btg::Options cur = start;
  cur.profile = false;
  auto timeOf = [&](const btg::Options& o, double& sec) -> int {
    opt = o;
    btg::RunInfo ri;
    const int rc = btg::engineSoftBatch(ctx->engine, h.data(), numJobs, exact != 0, ri, ctx->err);
    sec = (ri.primaryMs + ri.traceMs) * 1e-3;
    return rc;
  };
  const double t0 = now();
  int n = 0;
  double bestSec = 0;
  int rc = timeOf(cur, bestSec);  // also warms the code and buffers up
  if (!rc) rc = timeOf(cur, bestSec);
  auto record = [&](const btg::Options& o, double sec) {
    if (trials && n < maxTrials) trials[n] = btg_autotune_trial{o.leafCap, o.budget, o.roundsPerCheck, o.mailbox, sec};
    ++n;
  };
  if (!rc) record(cur, bestSec);
  for (const Knob& k : knobs) {
    if (rc) break;
    const int keep = cur.*k.field;
    int bestVal = keep;
    for (int v : k.values) {
      if (v == keep) continue;
      if (now() - t0 > maxSeconds) break;
      btg::Options o = cur;
      o.*k.field = v;
      double sec = 0;
      if ((rc = timeOf(o, sec))) break;
      record(o, sec);
      if (sec < bestSec * 0.98) {
        bestSec = sec;
        bestVal = v;
      }
    }
    cur.*k.field = bestVal;
  }
  opt = start;  // the caller applies the result with btg_set_options
  if (rc) return rc;
  if (numTrials) *numTrials = n;
  best->budget = cur.budget;
  best->primaryBudget = start.primaryBudget;
  best->capacity = start.capacity;
  best->roundsPerCheck = cur.roundsPerCheck;
  best->profileKernels = start.profile ? 1 : 0;
  best->leafCap = cur.leafCap;
  best->precision = start.precision;
  best->useGraphs = start.graphs ? 1 : 0;
  best->mailbox = cur.mailbox;
  return BTG_OK;
}

int64_t btg_polys_count(const btg_polys* p) { return p ? int64_t(p->v.size()) : 0; }
const btg_poly* btg_polys_data(const btg_polys* p) {
  static_assert(sizeof(btg_poly) == sizeof(btg::OutPoly), "polygon layout");
  return p && !p->v.empty() ? reinterpret_cast<const btg_poly*>(p->v.data()) : nullptr;
}
void btg_polys_free(btg_polys* p) { delete p; }

}  // extern "C"

namespace {

bool polyLess(const btg::OutPoly& a, const btg::OutPoly& b) {
  if (a.tri != b.tri) return a.tri < b.tri;
  if (a.n != b.n) return a.n < b.n;
  for (int k = 0; k < a.n; ++k) {
    if (a.x[k] != b.x[k]) return a.x[k] < b.x[k];
    if (a.y[k] != b.y[k]) return a.y[k] < b.y[k];
  }
  return a.area < b.area;
}

// The canonical order of a polygon list, sorted on up to 8 threads: parts
// sorted on their own threads, then merged pairwise level by level. The order
// is total on distinct polygons and equal polygons are identical bytes, so
// the list equals a single std::sort's.
void sortCanonical(std::vector<btg::OutPoly>& v) {
  const size_t n = v.size();
  const unsigned hw = std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
  size_t parts = 1;
  while (parts * 2 <= hw) parts *= 2;
  if (n < 65536 || parts < 2) {
    std::sort(v.begin(), v.end(), polyLess);
    return;
  }
  std::vector<size_t> b(parts + 1);
  for (size_t i = 0; i <= parts; ++i) b[i] = n * i / parts;
  auto run = [&](auto fn, size_t count) {
    std::vector<std::thread> th;
    for (size_t i = 1; i < count; ++i) {
      try {
        th.emplace_back(fn, i);
      } catch (const std::system_error&) {
        fn(i);  // no thread to spare: this part on the calling thread
      }
    }
    fn(size_t(0));
    for (auto& t : th) t.join();
  };
  run([&](size_t i) { std::sort(v.begin() + long(b[i]), v.begin() + long(b[i + 1]), polyLess); }, parts);
  for (size_t w = 1; w < parts; w *= 2)
    run([&](size_t k) {
      const size_t i = k * 2 * w;
      std::inplace_merge(v.begin() + long(b[i]), v.begin() + long(b[i + w]), v.begin() + long(b[std::min(i + 2 * w, parts)]),
                         polyLess);
    }, parts / (2 * w));
}

}  // namespace

extern "C" {

int btg_primary(btg_ctx* ctx, const btg_camera* cam, int cull, int tile, int canonical, btg_polys** out,
                btg_run_info* info) {
  if (!ctx) return failWith(nullptr, BTG_ERR_ARG, "null context");
  if (!validCam(cam) || !out || tile <= 0) return failWith(ctx, BTG_ERR_ARG, "invalid camera, tile or output");
  *out = nullptr;
  const double t0 = now();
  try {
    btg_polys* P = new btg_polys();
    btg::RunInfo ri;
    const int rc = btg::enginePrimary(ctx->engine, camOf(*cam), cull != 0, tile, P->v, ri, ctx->err);
    if (rc) {
      delete P;
      return rc;
    }
    if (canonical) sortCanonical(P->v);
    *out = P;
    fillInfo(ri, now() - t0, info);
    if (info) info->stats.hits = uint64_t(P->v.size());
    return BTG_OK;
  } catch (const std::bad_alloc&) {
    return failWith(ctx, BTG_ERR_MEMORY, "host allocation failed");
  }
}

int btg_raster(btg_ctx* ctx, const btg_camera* cam, const btg_poly* polys, int64_t count, int shadowPass,
               int32_t* ids, uint8_t* mask) {
  if (!ctx) return failWith(nullptr, BTG_ERR_ARG, "null context");
  if (!validCam(cam) || count < 0 || (count > 0 && !polys) || (shadowPass ? !mask : !ids))
    return failWith(ctx, BTG_ERR_ARG, "invalid camera, polygons or output");
  try {
    return btg::engineRaster(ctx->engine, camOf(*cam), reinterpret_cast<const btg::OutPoly*>(polys), count,
                             shadowPass != 0, ids, mask, ctx->err);
  } catch (const std::bad_alloc&) {
    return failWith(ctx, BTG_ERR_MEMORY, "host allocation failed");
  }
}

int btg_raster_device_noalloc(btg_ctx* ctx, const btg_camera* cam, const btg_poly* polysDev, int64_t count,
                              int shadowPass, int32_t* idsDev, uint8_t* maskDev) {
  if (!ctx) return failWith(nullptr, BTG_ERR_ARG, "null context");
  if (!validCam(cam) || count < 0 || (count > 0 && !polysDev) || (shadowPass ? !maskDev : !idsDev))
    return failWith(ctx, BTG_ERR_ARG, "invalid camera, polygons or output");
  static_assert(BTG_RASTER_NOALLOC_MAX_POLYS == btg::kNoAllocMaxPolys, "allocation-free limit");
  return btg::engineRasterNoAlloc(ctx->engine, camOf(*cam), reinterpret_cast<const btg::OutPoly*>(polysDev), count,
                                  shadowPass != 0, idsDev, maskDev, ctx->err);
}

int btg_images(btg_ctx* ctx, const btg_camera* cam, int cull, int tile, const float* light3, int32_t* ids,
               uint8_t* mask, btg_run_info* primInfo, btg_run_info* pointInfo, double* rasterSeconds) {
  if (!ctx) return failWith(nullptr, BTG_ERR_ARG, "null context");
  if (!validCam(cam) || tile <= 0 || !ids || (light3 && !mask))
    return failWith(ctx, BTG_ERR_ARG, "invalid camera, tile or output");
  const double t0 = now();
  try {
    btg::RunInfo a, b;
    double rasterMs = 0;
    const btg::F3 L = light3 ? btg::f3(light3[0], light3[1], light3[2]) : btg::f3(0, 0, 0);
    const int rc = btg::engineImages(ctx->engine, camOf(*cam), cull != 0, tile, light3 ? &L : nullptr, ids, mask, a, b,
                                     rasterMs, ctx->err);
    if (rc) return rc;
    fillInfo(a, now() - t0, primInfo);
    fillInfo(b, 0, pointInfo);
    if (rasterSeconds) *rasterSeconds = rasterMs * 1e-3;
    return BTG_OK;
  } catch (const std::bad_alloc&) {
    return failWith(ctx, BTG_ERR_MEMORY, "host allocation failed");
  }
}

int btg_point(btg_ctx* ctx, const btg_camera* cam, const btg_poly* prim, int64_t numPrim, const float* light3,
              int canonical, btg_polys** out, btg_run_info* info) {
  if (!ctx) return failWith(nullptr, BTG_ERR_ARG, "null context");
  if (!validCam(cam) || !out || !light3 || numPrim < 0 || (numPrim > 0 && !prim))
    return failWith(ctx, BTG_ERR_ARG, "invalid camera, polygons, light or output");
  *out = nullptr;
  const double t0 = now();
  try {
    btg_polys* P = new btg_polys();
    btg::RunInfo ri;
    const int rc = btg::enginePoint(ctx->engine, camOf(*cam), reinterpret_cast<const btg::OutPoly*>(prim), numPrim,
                                    btg::f3(light3[0], light3[1], light3[2]), P->v, ri, ctx->err);
    if (rc) {
      delete P;
      return rc;
    }
    if (canonical) sortCanonical(P->v);
    *out = P;
    fillInfo(ri, now() - t0, info);
    return BTG_OK;
  } catch (const std::bad_alloc&) {
    return failWith(ctx, BTG_ERR_MEMORY, "host allocation failed");
  }
}

}  // extern "C"

// ------------------------------------------------------------------ server
namespace {

struct SrvReq {
  uint64_t id = 0;
  btg_request req{};
  btg_camera_d camD{};  // the request's double camera, copied at submit (FP64 servers; req.camD points here)
  double tSubmit = 0, tAdmit = 0, tDone = 0;
  int status = BTG_OK;
  bool done = false;
  btg_polys* hits = nullptr;
  btg_polys* polys = nullptr;
};

}  // namespace

struct btg_server {
  btg_ctx* ctx = nullptr;
  int rounds = 8;
  int64_t rootCap = 0, pixCap = 0, queueCap = 0;
  int jobsPerPass = 0;
  btg::OpenPass* pass[2] = {nullptr, nullptr};
  std::unordered_map<int, SrvReq*> slots[2];  // pass slot -> request
  int active = 0;                             // the pass that takes new requests
  btg::Engine* pointEngine = nullptr;
  btg::Engine* fallback = nullptr;            // re-runs a request whose roots went over budget
  std::mutex m;
  std::condition_variable cvSoft, cvPoint, cvDone;
  std::deque<SrvReq*> softQ, pointQ;
  std::unordered_map<uint64_t, SrvReq*> reqs;
  uint64_t nextId = 1;
  bool stop = false;
  std::string softDown;  // set when the soft-shadow loop has ended on an error
  std::string err;
  std::thread softThread, pointThread;
};

namespace {

// Marks a request finished (the server's lock is held by the caller).
void finishLocked(btg_server* s, SrvReq* r, int status, const std::string& why) {
  r->status = status;
  r->done = true;
  r->tDone = now();
  if (status != BTG_OK) s->err = why;
  s->cvDone.notify_all();
}

btg::SoftJobH jobOf(const btg_request& q) {
  return btg::SoftJobH{q.scene, camOf(q.cam), lightOf(q.light), q.tri, q.vis, q.E, q.Eu,
                       q.camD ? q.camD->eye : nullptr};
}

// Fails every request of pass g and starts the pass afresh.
void failPass(btg_server* s, int g, int code, const std::string& why) {
  std::string e2;
  {
    std::lock_guard<std::mutex> lk(s->m);
    for (auto& kv : s->slots[g]) finishLocked(s, kv.second, code, why);
    s->slots[g].clear();
  }
  btg::openDestroy(s->pass[g]);
  s->pass[g] = nullptr;
}

void softLoop(btg_server* s) {
  std::vector<int> completed, killed;
  std::string err;
  for (;;) {
    bool stepped[2] = {false, false};
    {
      std::unique_lock<std::mutex> lk(s->m);
      // admission: requests join the accepting pass in arrival order while
      // their roots and pixels fit; a full pass hands over to the other one
      // once that one is idle
      while (!s->softQ.empty()) {
        SrvReq* r = s->softQ.front();
        btg::OpenPass* acc = s->pass[s->active];
        if (!acc) break;
        if (btg::openIdle(acc) && btg::openJobs(acc) > 0) {
          s->slots[s->active].clear();
          if (btg::openReset(acc, err)) {  // rebuilt below, before it takes requests
            btg::openDestroy(acc);
            s->pass[s->active] = nullptr;
            break;
          }
        }
        int slot = -1;
        const int rc = btg::openTryAdd(acc, jobOf(r->req), &slot, err);
        if (rc == 1) {
          s->slots[s->active][slot] = r;
          r->tAdmit = now();
          s->softQ.pop_front();
          continue;
        }
        if (rc < 0) {
          finishLocked(s, r, rc, err);
          s->softQ.pop_front();
          continue;
        }
        const int o = 1 - s->active;
        if (s->pass[o] && btg::openIdle(s->pass[o])) {
          s->slots[o].clear();
          if (btg::openJobs(s->pass[o]) > 0 && btg::openReset(s->pass[o], err)) {
            btg::openDestroy(s->pass[o]);
            s->pass[o] = nullptr;
            break;
          }
          s->active = o;
          continue;
        }
        break;  // both passes hold work: wait until one drains
      }
      const bool busy0 = s->pass[0] && !btg::openIdle(s->pass[0]);
      const bool busy1 = s->pass[1] && !btg::openIdle(s->pass[1]);
      if (!busy0 && !busy1 && s->pass[0] && s->pass[1]) {  // a missing pass is rebuilt first
        if (s->stop && s->softQ.empty()) return;
        s->cvSoft.wait(lk);
        continue;
      }
    }
    for (int g = 0; g < 2; ++g)
      if (s->pass[g] && !btg::openIdle(s->pass[g])) {
        if (int rc = btg::openStep(s->pass[g], s->rounds, err)) {
          failPass(s, g, rc, err);
          continue;
        }
        stepped[g] = true;
      }
    for (int g = 0; g < 2; ++g) {
      if (!stepped[g]) continue;
      if (int rc = btg::openCollect(s->pass[g], completed, killed, err)) {
        failPass(s, g, rc, err);
        continue;
      }
      for (size_t i = 0; i < completed.size(); ++i) {
        SrvReq* r = nullptr;
        {
          std::lock_guard<std::mutex> lk(s->m);
          const auto it = s->slots[g].find(completed[i]);
          if (it != s->slots[g].end()) {
            r = it->second;
            s->slots[g].erase(it);
          }
        }
        if (!r) continue;  // a slot the server never filled: nothing waits on it
        int rc = BTG_OK;
        if (killed[i] > 0) {  // over budget: trace the request again alone, with GPU retries
          btg::SoftJobH h = jobOf(r->req);
          btg::RunInfo ri;
          rc = btg::engineSoftBatch(s->fallback, &h, 1, true, ri, err);
        }
        std::lock_guard<std::mutex> lk(s->m);
        finishLocked(s, r, rc, err);
      }
    }
    // a failed pass is rebuilt before it takes requests again
    for (int g = 0; g < 2; ++g)
      if (!s->pass[g]) {
        btg::OpenPass* op =
            btg::openCreate(s->ctx->engine, s->rootCap, s->pixCap, s->jobsPerPass, true, s->queueCap, err);
        std::lock_guard<std::mutex> lk(s->m);
        s->pass[g] = op;
        if (!op) {  // no pass left to run on: fail what waits and refuse what comes
          for (SrvReq* r : s->softQ) finishLocked(s, r, BTG_ERR_DEVICE, err);
          s->softQ.clear();
          for (auto& sl : s->slots) {
            for (auto& kv : sl) finishLocked(s, kv.second, BTG_ERR_DEVICE, err);
            sl.clear();
          }
          s->softDown = "soft-shadow passes could not be rebuilt: " + err;
          return;
        }
      }
  }
}

// Soft requests of an FP64 server (precision 2): each batch takes every
// request pending when the previous one ended, up to kFp64BatchJobs requests
// and the server's pixel budget, and runs them through one engineSoftBatch
// call (FP64 cells of all their pixels, dealt dynamically to the device).
constexpr int kFp64BatchJobs = 64;

void fp64Loop(btg_server* s) {
  std::string err;
  std::vector<SrvReq*> batch;
  std::vector<btg::SoftJobH> jobs;
  for (;;) {
    batch.clear();
    jobs.clear();
    {
      std::unique_lock<std::mutex> lk(s->m);
      s->cvSoft.wait(lk, [&] { return s->stop || !s->softQ.empty(); });
      if (s->softQ.empty()) return;
      int64_t pix = 0;
      while (!s->softQ.empty() && int(batch.size()) < kFp64BatchJobs) {
        SrvReq* r = s->softQ.front();
        const int64_t rp = int64_t(r->req.cam.W) * r->req.cam.H;
        if (btg::engineSceneBytes(s->fallback, r->req.scene) < 0 || rp > s->pixCap) {  // fails alone
          s->softQ.pop_front();
          finishLocked(s, r, rp > s->pixCap ? BTG_ERR_LIMIT : BTG_ERR_ARG,
                       rp > s->pixCap ? "request larger than the server's pixel budget"
                                      : "no resident scene " + std::to_string(r->req.scene));
          continue;
        }
        if (!batch.empty() && pix + rp > s->pixCap) break;
        s->softQ.pop_front();
        r->tAdmit = now();
        pix += rp;
        batch.push_back(r);
      }
    }
    std::vector<int> rcs(batch.size(), BTG_OK);
    std::vector<std::string> errs(batch.size());
    try {
      for (SrvReq* r : batch) jobs.push_back(jobOf(r->req));
      btg::RunInfo ri;
      const int rc = btg::engineSoftBatch(s->fallback, jobs.data(), int(jobs.size()), true, ri, err);
      if (rc != BTG_OK && batch.size() > 1) {
        // a batch that fails is run again request by request, so that a
        // request that cannot be traced (e.g. a cell over the pools) fails alone
        for (size_t i = 0; i < batch.size(); ++i) {
          btg::RunInfo r1;
          rcs[i] = btg::engineSoftBatch(s->fallback, &jobs[i], 1, true, r1, errs[i]);
        }
      } else {
        for (size_t i = 0; i < batch.size(); ++i) {
          rcs[i] = rc;
          errs[i] = err;
        }
      }
    } catch (const std::bad_alloc&) {
      for (size_t i = 0; i < batch.size(); ++i) {
        rcs[i] = BTG_ERR_MEMORY;
        errs[i] = "host allocation failed";
      }
    }
    std::lock_guard<std::mutex> lk(s->m);
    for (size_t i = 0; i < batch.size(); ++i) finishLocked(s, batch[i], rcs[i], errs[i]);
  }
}

void pointLoop(btg_server* s) {
  for (;;) {
    SrvReq* r;
    {
      std::unique_lock<std::mutex> lk(s->m);
      s->cvPoint.wait(lk, [&] { return s->stop || !s->pointQ.empty(); });
      if (s->pointQ.empty()) return;
      r = s->pointQ.front();
      s->pointQ.pop_front();
      r->tAdmit = now();
    }
    std::string err;
    btg_polys* hits = nullptr;
    btg_polys* polys = nullptr;
    int rc = BTG_OK;
    try {
      hits = new btg_polys();
      polys = new btg_polys();
      btg::RunInfo ri, ri2;
      const btg::HostCam cam = camOf(r->req.cam);
      rc = btg::enginePrimary(s->pointEngine, cam, false, 16, hits->v, ri, err);
      if (!rc) {
        // the point pass's output set does not depend on the order of its
        // roots, so it runs on the device's order while the hits are sorted
        std::vector<btg::OutPoly> roots = hits->v;
        std::thread sorter([&] { sortCanonical(hits->v); });
        try {
          rc = btg::enginePoint(s->pointEngine, cam, roots.data(), int64_t(roots.size()),
                                btg::f3(r->req.pointLight[0], r->req.pointLight[1], r->req.pointLight[2]),
                                polys->v, ri2, err);
        } catch (...) {
          sorter.join();
          throw;
        }
        sorter.join();
      }
      if (!rc) sortCanonical(polys->v);
    } catch (const std::bad_alloc&) {
      rc = BTG_ERR_MEMORY;
      err = "host allocation failed";
    }
    std::lock_guard<std::mutex> lk(s->m);
    if (rc) {
      delete hits;
      delete polys;
    } else {
      r->hits = hits;
      r->polys = polys;
    }
    finishLocked(s, r, rc, err);
  }
}

}  // namespace

extern "C" {

int btg_server_create(btg_ctx* ctx, const btg_server_config* cfg, btg_server** out) {
  if (!ctx || !out) return failWith(ctx, BTG_ERR_ARG, "null argument");
  *out = nullptr;
  btg_server_config c{};
  if (cfg) c = *cfg;
  // default: 128 GiB, or 45% of the device where that is less. The passes'
  // queues (a quarter of it, up to the context's capacity) set how many roots
  // run at once: on the conference-hall stream 32 / 64 / 128 GiB gave p50
  // 58 / 48 / 43 ms and p99 99 / 86 / 85 ms
  int64_t hbm = c.hbmBytes;
  if (hbm <= 0) {
    hbm = int64_t(128) << 30;
    size_t freeB = 0, totalB = 0;
    if (hipMemGetInfo(&freeB, &totalB) == hipSuccess && totalB > 0)
      hbm = std::min<int64_t>(hbm, int64_t(double(totalB) * 0.45));
    (void)hipGetLastError();
  }
  const int jobs = c.jobsPerPass > 0 ? c.jobsPerPass : 512;
  const int rounds = c.roundsPerStep > 0 ? c.roundsPerStep : 8;
  const btg::Options& opt = btg::engineOptions(ctx->engine);
  const bool fp64 = opt.precision == 2;  // FP64 batches instead of wavefront passes (fp64Loop)
  // per pass: three sub-beam queues (64 B records) and the shade queue (40 B)
  // in at most a quarter of hbmBytes (the largest power-of-two capacity that
  // fits, and no more than the context's), then per pixel its images and per
  // root its state (softPixelBytes, rootStateBytes: a full 8 x 8 tile has a
  // root per pixel)
  int64_t queueCap = int64_t(1) << 16;
  while (queueCap * 2 <= opt.capacity && queueCap * 2 * (3 * 64 + 40) <= hbm / 4) queueCap *= 2;
  const int64_t queues = queueCap * (3 * 64 + 40);
  const int64_t perPass = hbm / 2 - queues;
  if (perPass < (int64_t(1) << 26))
    return failWith(ctx, BTG_ERR_ARG, "hbmBytes too small for two passes at the configured queue capacity");
  const int64_t pixCap = std::min<int64_t>(
      int64_t(1) << 28, perPass / (btg::softPixelBytes(opt.precision != 0) + btg::rootStateBytes()));
  const int64_t rootCap = std::min<int64_t>(int64_t(1) << 30, pixCap + int64_t(jobs) * 16448);
  try {
    btg_server* s = new btg_server();
    s->ctx = ctx;
    s->rounds = rounds;
    s->rootCap = rootCap;
    s->pixCap = pixCap;
    s->queueCap = queueCap;
    s->jobsPerPass = jobs;
    std::string err;
    for (int g = 0; g < 2 && err.empty() && !fp64; ++g) {
      s->pass[g] = btg::openCreate(ctx->engine, rootCap, pixCap, jobs, true, queueCap, err);
      if (!s->pass[g] && err.empty()) err = "open pass creation failed";
    }
    if (err.empty()) s->pointEngine = btg::engineShare(ctx->engine, err);
    if (err.empty()) s->fallback = btg::engineShare(ctx->engine, err);
    if (!err.empty()) {
      for (auto* op : s->pass) btg::openDestroy(op);
      btg::engineDestroy(s->pointEngine);
      btg::engineDestroy(s->fallback);
      delete s;
      return failWith(ctx, BTG_ERR_MEMORY, err);
    }
    btg::engineOptions(s->pointEngine) = opt;
    btg::engineOptions(s->fallback) = opt;
    // The server's passes run beside the caller's threads: no graph capture
    // (while a stream captures, other threads' legacy-stream calls fail on
    // this runtime), rounds are launched directly.
    btg::engineOptions(s->pointEngine).graphs = false;
    btg::engineOptions(s->fallback).graphs = false;
    btg::engineOptions(s->pointEngine).capacity = std::min<int64_t>(opt.capacity, queueCap);
    btg::engineOptions(s->fallback).capacity = std::min<int64_t>(opt.capacity, queueCap);
    s->softThread = fp64 ? std::thread(fp64Loop, s) : std::thread(softLoop, s);
    s->pointThread = std::thread(pointLoop, s);
    *out = s;
    return BTG_OK;
  } catch (const std::exception& ex) {
    return failWith(ctx, BTG_ERR_MEMORY, std::string("server creation failed: ") + ex.what());
  }
}

int btg_server_submit(btg_server* s, const btg_request* req, uint64_t* ticket) {
  if (!s || !req || !ticket) return BTG_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->m);
  if (s->stop) {
    s->err = "server is shutting down";
    return BTG_ERR_ARG;
  }
  if ((req->kind != BTG_REQ_SOFT && req->kind != BTG_REQ_POINT) || !validCam(&req->cam) ||
      (req->kind == BTG_REQ_SOFT && req->exact != 1)) {
    s->err = "invalid request (kind, camera, or a soft request without exact = 1)";
    return BTG_ERR_ARG;
  }
  if (req->kind == BTG_REQ_SOFT && !s->softDown.empty()) {
    s->err = s->softDown;
    return BTG_ERR_DEVICE;
  }
  try {
    SrvReq* r = new SrvReq();
    r->id = s->nextId++;
    r->req = *req;
    if (req->kind == BTG_REQ_SOFT && req->camD) {  // the caller's camera is read now, not when the batch runs
      r->camD = *req->camD;
      r->req.camD = &r->camD;
    } else {
      r->req.camD = nullptr;
    }
    r->tSubmit = now();
    s->reqs[r->id] = r;
    if (req->kind == BTG_REQ_SOFT) {
      s->softQ.push_back(r);
      s->cvSoft.notify_one();
    } else {
      s->pointQ.push_back(r);
      s->cvPoint.notify_one();
    }
    *ticket = r->id;
    return BTG_OK;
  } catch (const std::bad_alloc&) {
    s->err = "host allocation failed";
    return BTG_ERR_MEMORY;
  }
}

int btg_server_wait(btg_server* s, uint64_t ticket, double timeoutSeconds, btg_request_status* st) {
  if (!s || !st) return BTG_ERR_ARG;
  std::unique_lock<std::mutex> lk(s->m);
  if (s->reqs.find(ticket) == s->reqs.end()) {
    s->err = "unknown ticket";
    return BTG_ERR_ARG;
  }
  // The request is looked up again on every wake-up: another waiter on the
  // same ticket may have released it meanwhile.
  auto ready = [&] {
    const auto i = s->reqs.find(ticket);
    return i == s->reqs.end() || i->second->done;
  };
  if (timeoutSeconds < 0)
    s->cvDone.wait(lk, ready);
  else
    s->cvDone.wait_for(lk, std::chrono::duration<double>(timeoutSeconds), ready);
  std::memset(st, 0, sizeof(*st));
  const auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) {
    s->err = "ticket released by another waiter";
    return BTG_ERR_ARG;
  }
  SrvReq* r = it->second;
  if (!r->done) return BTG_OK;
  st->done = 1;
  st->status = r->status;
  st->queueSeconds = (r->tAdmit > 0 ? r->tAdmit : r->tDone) - r->tSubmit;
  st->latencySeconds = r->tDone - r->tSubmit;
  st->hits = r->hits;
  st->polys = r->polys;
  s->reqs.erase(it);
  delete r;
  return BTG_OK;
}

const char* btg_server_error(btg_server* s) {
  if (!s) return "";
  // the worker threads may replace s->err at any time: hand out a copy taken
  // under the lock, valid until this thread's next call
  thread_local std::string copy;
  std::lock_guard<std::mutex> lk(s->m);
  copy = s->err;
  return copy.c_str();
}

void btg_server_destroy(btg_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->m);
    s->stop = true;
    s->cvSoft.notify_all();
    s->cvPoint.notify_all();
  }
  if (s->softThread.joinable()) s->softThread.join();
  if (s->pointThread.joinable()) s->pointThread.join();
  for (auto* op : s->pass) btg::openDestroy(op);
  btg::engineDestroy(s->pointEngine);
  btg::engineDestroy(s->fallback);
  for (auto& kv : s->reqs) {
    delete kv.second->hits;
    delete kv.second->polys;
    delete kv.second;
  }
  delete s;
}

}  // extern "C"
