// Drop-in implementation of the reference's GPU renderer interface
// (cuda/gpu_render.h of joelsalzman/beamtrace2007) on the MI355X library's C
// ABI (include/btg.h). It replaces cuda/gpu_render.cu + cuda/wavefront.cu in a
// build of the reference: compile this file with the reference's headers on
// the include path and link libbtg.so; apps/bt_render.cpp built with
// -DBT_WITH_CUDA then renders with `--device cuda` on the MI355X.
//
// Behaviour kept from the reference (cuda/gpu_render.cu):
//   * the constructor uploads scene, kd-tree and packed leaf triangles once;
//   * softShadows fills res.tri / vis / E / Eunocc and the stats, with
//     primarySeconds = GPU primary rays and traceSeconds = shadow beams;
//   * primary replaces hitBeams, pointShadows appends to shadowPolys;
//   * a device or library error prints a message and aborts
//     (cuda/gpu_common.cuh:15-22).
// Differences: roots over budget are traced again on the GPU (no CPU
// fallback), so lastFallbackSeconds is 0 and lastOverflowPixels counts the
// roots that took a second GPU pass; polygon lists come back in a canonical
// order; useMailbox turns on the library's mailbox (btg_options.mailbox) for
// all three applications (soft shadows: with SoftOptions::mailbox), whose
// results equal those without it.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "btg.h"
#include "cuda/gpu_render.h"

namespace bt {

static_assert(sizeof(Real) == 4, "the MI355X library implements the float build");
static_assert(sizeof(OutBeam) == sizeof(btg_poly), "OutBeam layout (src/beam/beam_types.h:68-73)");
static_assert(sizeof(KdNode) == 12, "KdNode layout (src/accel/kdtree.h:13-21)");
static_assert(sizeof(AABB) == 24, "AABB layout (src/core/vec.h:52-58)");

struct GpuRenderer::Impl {
  btg_ctx* ctx = nullptr;
};

namespace {

[[noreturn]] void die(const char* what, const btg_ctx* ctx) {
  fprintf(stderr, "MI355X beam tracer: %s failed: %s\n", what, btg_last_error(ctx));
  std::abort();
}

btg_camera toCam(const Camera& c) {
  btg_camera b;
  const Vec3* v[4] = {&c.eye, &c.fwd, &c.right, &c.down};
  float* d[4] = {b.eye, b.fwd, b.right, b.down};
  for (int i = 0; i < 4; ++i)
    for (int k = 0; k < 3; ++k) d[i][k] = (*v[i])[k];
  b.halfW = c.halfW;
  b.halfH = c.halfH;
  b.W = c.W;
  b.H = c.H;
  return b;
}

TraceStats toStats(const btg_trace_stats& s) {
  TraceStats t;
  t.kdSteps = s.kdSteps;
  t.leafVisits = s.leafVisits;
  t.triTests = s.triTests;
  t.hits = s.hits;
  t.splits = s.splits;
  t.beams = s.beams;
  t.rootBeams = s.rootBeams;
  t.presplitBeams = s.presplitBeams;
  t.mailboxSkips = s.mailboxSkips;
  t.fiveSplits = s.fiveSplits;
  t.visibleTris = s.visibleTris;
  t.rays = s.rays;
  t.climbs = s.climbs;
  t.droppedArea = s.droppedArea;
  return t;
}

void applyKnobs(btg_ctx* ctx, int budget, int primaryBudget, size_t capacity, bool mailbox) {
  btg_options o;
  if (btg_get_options(ctx, &o) != BTG_OK) die("btg_get_options", ctx);
  o.budget = budget;
  o.primaryBudget = primaryBudget;
  o.capacity = int64_t(capacity);
  o.mailbox = mailbox ? 1 : 0;
  if (btg_set_options(ctx, &o) != BTG_OK) die("btg_set_options", ctx);
}

}  // namespace

bool gpuAvailable(GpuInfo* info) {
  btg_device_info d;
  if (!btg_device_query(&d)) return false;
  if (info) {
    info->name = std::string(d.name) + " (" + d.arch + ")";
    info->smCount = d.computeUnits;
    info->memGB = d.memGB;
  }
  return true;
}

GpuRenderer::GpuRenderer(const Scene& scene, const KdTree& tree) : impl_(new Impl), scene_(scene), tree_(tree) {
  btg_scene_desc d;
  std::memset(&d, 0, sizeof d);
  d.pos = reinterpret_cast<const float*>(scene.pos.data());
  d.numVerts = int64_t(scene.pos.size());
  d.tri3 = reinterpret_cast<const uint32_t*>(scene.tris.data());
  d.triN = reinterpret_cast<const float*>(scene.triN.data());
  d.numTris = scene.numTris();
  d.nodes = tree.nodes.data();
  d.numNodes = int64_t(tree.nodes.size());
  d.triIndices = tree.triIndices.data();
  d.numRefs = int64_t(tree.triIndices.size());
  d.parent = tree.parent.data();
  d.nodeBox = reinterpret_cast<const float*>(tree.nodeBox.data());
  for (int k = 0; k < 3; ++k) {
    d.kdBounds[k] = tree.bounds.lo[k];
    d.kdBounds[3 + k] = tree.bounds.hi[k];
    d.sceneBounds[k] = scene.bounds.lo[k];
    d.sceneBounds[3 + k] = scene.bounds.hi[k];
  }
  if (btg_create(&impl_->ctx, &d) != BTG_OK) die("btg_create", nullptr);
  btg_options o;
  btg_get_options(impl_->ctx, &o);
  // the library's defaults (queues sized for MI355X's HBM) replace the
  // reference's 8M-record / 128-beam setting unless the caller changes them
  wfBudget = o.budget;
  wfPrimaryBudget = o.primaryBudget;
  wfCapacity = size_t(o.capacity);
  if (const char* e = getenv("BT_WF_BUDGET")) wfBudget = std::max(1, atoi(e));
  if (const char* e = getenv("BT_WF_CAPACITY")) wfCapacity = size_t(std::max(1024L, atol(e)));
}

GpuRenderer::~GpuRenderer() {
  btg_destroy(impl_->ctx);
  delete impl_;
}

void GpuRenderer::softShadows(const Camera& cam, const AreaLight& light, const SoftOptions& opt, SoftResult& res) {
  // Post Office mailboxing when both the renderer and the call ask for it, as
  // the reference's GPU soft shadows do (cuda/gpu_render.cu:612)
  applyKnobs(impl_->ctx, wfBudget, wfPrimaryBudget, wfCapacity, useMailbox && opt.mailbox);
  const size_t npix = size_t(cam.W) * size_t(cam.H);
  res.W = cam.W;
  res.H = cam.H;
  res.stats = RenderStats();
  res.tri.resize(npix);
  res.vis.resize(npix);
  res.E.resize(npix);
  res.Eunocc.resize(npix);
  btg_camera c = toCam(cam);
  btg_area_light L;
  for (int k = 0; k < 3; ++k) {
    L.c[k] = light.c[k];
    L.U[k] = light.U[k];
    L.V[k] = light.V[k];
  }
  btg_run_info info;
  if (btg_soft(impl_->ctx, &c, &L, opt.exact ? 1 : 0, res.tri.data(), res.vis.data(), res.E.data(),
               res.Eunocc.data(), &info) != BTG_OK)
    die("btg_soft", impl_->ctx);
  res.stats.trace = toStats(info.stats);
  res.stats.primarySeconds = info.primarySeconds;
  res.stats.traceSeconds = info.traceSeconds;
  lastRounds = info.rounds;
  lastMaxQueue = info.maxQueue;
  lastOverflowPixels = info.killedRoots;
  lastFallbackSeconds = 0;
  lastKernelSeconds = info.traceSeconds;
}

void GpuRenderer::primary(const Camera& cam, bool cull, std::vector<OutBeam>& hitBeams, RenderStats& rs, int tile) {
  applyKnobs(impl_->ctx, wfBudget, wfPrimaryBudget, wfCapacity, useMailbox);
  const auto t0 = std::chrono::steady_clock::now();
  btg_camera c = toCam(cam);
  btg_polys* out = nullptr;
  btg_run_info info;
  if (btg_primary(impl_->ctx, &c, cull ? 1 : 0, tile, 1, &out, &info) != BTG_OK) die("btg_primary", impl_->ctx);
  const int64_t n = btg_polys_count(out);
  hitBeams.resize(size_t(n));
  if (n) std::memcpy(hitBeams.data(), btg_polys_data(out), size_t(n) * sizeof(OutBeam));
  btg_polys_free(out);
  rs.trace = toStats(info.stats);
  rs.trace.hits = hitBeams.size();
  rs.traceSeconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  lastRounds = info.rounds;
  lastMaxQueue = info.maxQueue;
  lastOverflowPixels = info.killedRoots;
  lastFallbackSeconds = 0;
  lastKernelSeconds = info.traceSeconds;
}

void GpuRenderer::pointShadows(const Camera& cam, const std::vector<OutBeam>& primHits, const Vec3& light,
                               std::vector<OutBeam>& shadowPolys, RenderStats& rs) {
  if (primHits.empty()) return;
  applyKnobs(impl_->ctx, wfBudget, wfPrimaryBudget, wfCapacity, useMailbox);
  const auto t0 = std::chrono::steady_clock::now();
  btg_camera c = toCam(cam);
  const float l3[3] = {light.x, light.y, light.z};
  btg_polys* out = nullptr;
  btg_run_info info;
  if (btg_point(impl_->ctx, &c, reinterpret_cast<const btg_poly*>(primHits.data()), int64_t(primHits.size()), l3, 1,
                &out, &info) != BTG_OK)
    die("btg_point", impl_->ctx);
  const int64_t n = btg_polys_count(out);
  const size_t at = shadowPolys.size();
  shadowPolys.resize(at + size_t(n));
  if (n) std::memcpy(shadowPolys.data() + at, btg_polys_data(out), size_t(n) * sizeof(OutBeam));
  btg_polys_free(out);
  rs.trace.add(toStats(info.stats));
  rs.traceSeconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  lastRounds = info.rounds;
  lastMaxQueue = info.maxQueue;
  lastOverflowPixels = info.killedRoots;
  lastFallbackSeconds = 0;
  lastKernelSeconds = info.traceSeconds;
}

}  // namespace bt
