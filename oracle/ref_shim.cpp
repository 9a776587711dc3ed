// C-ABI wrapper around the reference CPU beam tracer (test infrastructure).
//
// This file is compiled together with the reference's own sources, where they
// lie (src/**/*.cpp of joelsalzman/beamtrace2007), into oracle/_ref/ by
// oracle/Makefile: once with Real = float (the reference float build) and once
// with -DBT_REAL_DOUBLE (the double build, bt_render_d). Only tests, smoke()
// and bench.py's cpu_baseline leg load the result. Nothing of the reference is
// copied here; this file only calls the reference's public functions:
//   Scene::finalize (src/scene/scene.cpp:10), KdTree::build
//   (src/accel/kdtree.cpp:237), Camera::make (src/render/render.cpp:14),
//   softShadows (render.cpp:339), beamPrimary (render.cpp:150),
//   beamPointShadows (render.cpp:209), rasterize (render.cpp:49),
//   beamLightVisibility (render.cpp:279), RayTracer (src/ray/ray_tracer.h),
//   WavefrontTracer (src/beam/wavefront.cpp:50).
// All floating-point values cross this boundary as double (exact for float).
#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "beam/wavefront.h"
#include "render/point_core.h"
#include "render/render.h"
#include "util/misc.h"

using namespace bt;

namespace {

struct Polys {
  std::vector<OutBeam> v;
};

Camera makeCam(const double* c, int W, int H) {
  View v;
  v.eye = Vec3(Real(c[0]), Real(c[1]), Real(c[2]));
  v.target = Vec3(Real(c[3]), Real(c[4]), Real(c[5]));
  v.fovY = Real(c[9]);
  Vec3 up{Real(c[6]), Real(c[7]), Real(c[8])};
  return Camera::make(v, up, W, H);
}

AreaLight makeLight(const double* l) {
  AreaLight L;
  L.c = Vec3(Real(l[0]), Real(l[1]), Real(l[2]));
  L.U = Vec3(Real(l[3]), Real(l[4]), Real(l[5]));
  L.V = Vec3(Real(l[6]), Real(l[7]), Real(l[8]));
  return L;
}

void putStats(const TraceStats& s, double* out) {
  if (!out) return;
  out[0] = double(s.kdSteps);
  out[1] = double(s.leafVisits);
  out[2] = double(s.triTests);
  out[3] = double(s.hits);
  out[4] = double(s.splits);
  out[5] = double(s.beams);
  out[6] = double(s.rootBeams);
  out[7] = double(s.presplitBeams);
  out[8] = double(s.mailboxSkips);
  out[9] = double(s.fiveSplits);
  out[10] = double(s.visibleTris);
  out[11] = double(s.rays);
  out[12] = double(s.climbs);
  out[13] = s.droppedArea;
}

OutBeam polyFrom(const double* x, const double* y, int n, int tri, double area) {
  OutBeam ob;
  ob.n = n;
  for (int i = 0; i < 4; ++i) {
    ob.x[i] = Real(x[i]);
    ob.y[i] = Real(y[i]);
  }
  ob.tri = tri;
  ob.area = Real(area);
  return ob;
}

}  // namespace

extern "C" {

int ref_real_bytes() { return int(sizeof(Real)); }

void* ref_scene_new(const float* pos, int64_t nv, const uint32_t* tris, int64_t nt) {
  Scene* s = new Scene();
  s->pos.resize(size_t(nv));
  for (int64_t i = 0; i < nv; ++i) s->pos[size_t(i)] = Vec3(Real(pos[3 * i]), Real(pos[3 * i + 1]), Real(pos[3 * i + 2]));
  for (int64_t t = 0; t < nt; ++t) s->addTri(tris[3 * t], tris[3 * t + 1], tris[3 * t + 2], 0);
  s->finalize();
  return s;
}

void ref_scene_free(void* h) { delete static_cast<Scene*>(h); }

int64_t ref_scene_ntris(void* h) { return static_cast<Scene*>(h)->numTris(); }

void ref_scene_arrays(void* h, uint32_t* tris, double* triN, double* triArea, double* bounds6) {
  const Scene& s = *static_cast<Scene*>(h);
  for (int t = 0; t < s.numTris(); ++t) {
    for (int k = 0; k < 3; ++k) {
      tris[3 * t + k] = s.tris[size_t(t)][size_t(k)];
      triN[3 * t + k] = double(s.triN[size_t(t)][k]);
    }
    triArea[t] = double(s.triArea[size_t(t)]);
  }
  for (int k = 0; k < 3; ++k) {
    bounds6[k] = double(s.bounds.lo[k]);
    bounds6[3 + k] = double(s.bounds.hi[k]);
  }
}

void* ref_kd_new(void* sh, double costIntersect, int singleLeaf) {
  KdTree* t = new KdTree();
  KdBuildParams kp;
  kp.costIntersect = Real(costIntersect);
  kp.singleLeaf = singleLeaf != 0;
  t->build(*static_cast<Scene*>(sh), kp);
  return t;
}

void ref_kd_free(void* h) { delete static_cast<KdTree*>(h); }

void ref_kd_counts(void* h, int64_t* nNodes, int64_t* nIdx) {
  const KdTree& t = *static_cast<KdTree*>(h);
  *nNodes = int64_t(t.nodes.size());
  *nIdx = int64_t(t.triIndices.size());
}

void ref_kd_arrays(void* h, double* split, uint32_t* data, uint32_t* offset, uint32_t* triIndices, uint32_t* parent,
                   double* nodeBox6, double* bounds6) {
  const KdTree& t = *static_cast<KdTree*>(h);
  for (size_t i = 0; i < t.nodes.size(); ++i) {
    split[i] = double(t.nodes[i].split);
    data[i] = t.nodes[i].data;
    offset[i] = t.nodes[i].offset;
    parent[i] = t.parent[i];
    for (int k = 0; k < 3; ++k) {
      nodeBox6[6 * i + k] = double(t.nodeBox[i].lo[k]);
      nodeBox6[6 * i + 3 + k] = double(t.nodeBox[i].hi[k]);
    }
  }
  std::memcpy(triIndices, t.triIndices.data(), t.triIndices.size() * sizeof(uint32_t));
  for (int k = 0; k < 3; ++k) {
    bounds6[k] = double(t.bounds.lo[k]);
    bounds6[3 + k] = double(t.bounds.hi[k]);
  }
}

// Camera::make (render.cpp:14): eye, fwd, right, down, halfW, halfH.
void ref_camera(const double* cam, int W, int H, double* out14) {
  Camera c = makeCam(cam, W, H);
  const Vec3* v[4] = {&c.eye, &c.fwd, &c.right, &c.down};
  for (int i = 0; i < 4; ++i)
    for (int k = 0; k < 3; ++k) out14[3 * i + k] = double((*v[i])[k]);
  out14[12] = double(c.halfW);
  out14[13] = double(c.halfH);
}

// softShadows (render.cpp:339) over the whole image.
void ref_soft(void* sh, void* kh, const double* cam, int W, int H, const double* light, int exact, int mailbox,
              int trail, int threads, int32_t* tri, float* vis, float* E, float* Eu, double* stats, double* seconds) {
  const Scene& s = *static_cast<Scene*>(sh);
  const KdTree& t = *static_cast<KdTree*>(kh);
  Camera c = makeCam(cam, W, H);
  SoftOptions so;
  so.useBeams = true;
  so.exact = exact != 0;
  so.mailbox = mailbox != 0;
  so.trail = trail != 0;
  so.threads = std::max(1, threads);
  SoftResult res;
  softShadows(s, t, c, makeLight(light), so, res);
  const size_t n = size_t(W) * size_t(H);
  std::memcpy(tri, res.tri.data(), n * sizeof(int32_t));
  std::memcpy(vis, res.vis.data(), n * sizeof(float));
  std::memcpy(E, res.E.data(), n * sizeof(float));
  std::memcpy(Eu, res.Eunocc.data(), n * sizeof(float));
  putStats(res.stats.trace, stats);
  if (seconds) {
    seconds[0] = res.stats.primarySeconds;
    seconds[1] = res.stats.traceSeconds;
  }
}

// The per-pixel work of softShadows (render.cpp:351-421) restricted to image
// rows y0, y0 + stride, ... < y1: the bounded CPU-baseline sample (a strided
// sample covers every region of the frame). Same calls, same threading.
void ref_soft_rows(void* sh, void* kh, const double* cam, int W, int H, const double* light, int exact, int mailbox,
                   int trail, int threads, int y0, int y1, int stride, float* vis, double* stats, double* seconds,
                   int32_t* triOut, float* Gout) {
  const Scene& s = *static_cast<Scene*>(sh);
  const KdTree& tree = *static_cast<KdTree*>(kh);
  Camera c = makeCam(cam, W, H);
  AreaLight L = makeLight(light);
  const Real offset = length(s.bounds.diag()) * kEpsRel;
  RayTracer rt(s, tree);
  stride = std::max(1, stride);
  const int rows = (y1 - y0 + stride - 1) / stride;
  std::vector<int> triv(size_t(W) * size_t(rows), -1);
  std::vector<Vec3> X(triv.size()), N(triv.size());
  threads = std::max(1, threads);
  Timer tp;
  {
    std::atomic<int> next(0);
    std::vector<TraceStats> st{size_t(threads), TraceStats()};
    auto worker = [&](int id) {
      for (int r = next++; r < rows; r = next++)
        for (int x = 0; x < W; ++x) {
          const int y = y0 + r * stride;
          size_t k = size_t(r) * size_t(W) + size_t(x);
          Real qx, qy;
          c.pixelToQ(Real(x) + Real(0.5), Real(y) + Real(0.5), qx, qy);
          Ray ray;
          ray.o = c.eye;
          ray.d = c.dir(qx, qy);
          RayHit h;
          if (!rt.intersect(ray, h, st[size_t(id)])) continue;
          triv[k] = h.tri;
          X[k] = ray.o + ray.d * h.t;
          Vec3 n = s.triN[size_t(h.tri)];
          if (dot(n, ray.d) > 0) n = -n;
          N[k] = n;
        }
    };
    std::vector<std::thread> th;
    for (int i = 1; i < threads; ++i) th.emplace_back(worker, i);
    worker(0);
    for (auto& x : th) x.join();
  }
  const double primaryS = tp.seconds();
  if (triOut) std::copy(triv.begin(), triv.end(), triOut);
  Timer ts;
  std::atomic<int> next(0);
  std::vector<TraceStats> st{size_t(threads), TraceStats()};
  auto worker = [&](int id) {
    BeamTracer btr(s, tree);
    btr.useMailbox = mailbox != 0;
    btr.useTrail = trail != 0;
    btr.keepOutput = false;
    for (int r = next++; r < rows; r = next++)
      for (int x = 0; x < W; ++x) {
        size_t k = size_t(r) * size_t(W) + size_t(x);
        vis[k] = -1.f;
        if (Gout) Gout[k] = 0.f;
        if (triv[k] < 0) continue;
        Real Gx = 0;
        vis[k] = float(beamLightVisibility(btr, X[k], N[k], triv[k], L, offset, exact ? &Gx : nullptr));
        if (Gout) Gout[k] = float(Gx);  // the exact irradiance E (render.cpp:298-302)
      }
    st[size_t(id)].add(btr.stats);
  };
  std::vector<std::thread> th;
  for (int i = 1; i < threads; ++i) th.emplace_back(worker, i);
  worker(0);
  for (auto& x : th) x.join();
  TraceStats tot;
  for (auto& x : st) tot.add(x);
  putStats(tot, stats);
  if (seconds) {
    seconds[0] = primaryS;
    seconds[1] = ts.seconds();
  }
}

// Exact visibility of the light from one point (render.cpp:279).
double ref_light_vis(void* sh, void* kh, const double* x, const double* ns, int tri, const double* light,
                     double offset, int mailbox, int trail, double* G, double* stats) {
  BeamTracer btr(*static_cast<Scene*>(sh), *static_cast<KdTree*>(kh));
  btr.useMailbox = mailbox != 0;
  btr.useTrail = trail != 0;
  Real g = 0;
  Real V = beamLightVisibility(btr, Vec3(Real(x[0]), Real(x[1]), Real(x[2])), Vec3(Real(ns[0]), Real(ns[1]), Real(ns[2])),
                               tri, makeLight(light), Real(offset), G ? &g : nullptr);
  if (G) *G = double(g);
  putStats(btr.stats, stats);
  return double(V);
}

// Primary visibility. tile == 0: one root for the whole image (beamPrimary,
// render.cpp:150). tile > 0: one root per tile x tile pixels, traced one
// after another like the GPU engine's roots (tile polygon as in
// cuda/gpu_render.cu:730-742). Returns a polygon list handle (hit beams only
// when tile > 0, all output beams when tile == 0).
void* ref_primary(void* sh, void* kh, const double* cam, int W, int H, int cull, int mailbox, int trail, int tile,
                  double* stats) {
  const Scene& s = *static_cast<Scene*>(sh);
  const KdTree& t = *static_cast<KdTree*>(kh);
  Camera c = makeCam(cam, W, H);
  BeamTracer btr(s, t);
  btr.useMailbox = mailbox != 0;
  btr.useTrail = trail != 0;
  Polys* P = new Polys();
  if (tile <= 0) {
    BeamOutput out;
    RenderStats rs;
    beamPrimary(btr, c, cull != 0, out, rs);
    P->v = out.beams;
    putStats(rs.trace, stats);
    return P;
  }
  BeamQuery q;
  q.plane = c.plane();
  q.mode = BeamMode::Nearest;
  q.cullBackfaces = cull != 0;
  const int tilesX = (W + tile - 1) / tile, tilesY = (H + tile - 1) / tile;
  BeamOutput out;
  for (int i = 0; i < tilesX * tilesY; ++i) {
    int tx = i % tilesX, ty = i / tilesX;
    int x0 = tx * tile, y0 = ty * tile, x1 = std::min(W, x0 + tile), y1 = std::min(H, y0 + tile);
    Real a, b, cc, d;
    c.pixelToQ(Real(x0), Real(y0), a, b);
    c.pixelToQ(Real(x1), Real(y1), cc, d);
    Poly2 p;
    p.push(a, b);
    p.push(cc, b);
    p.push(cc, d);
    p.push(a, d);
    btr.trace(q, &p, 1, out);
    for (const OutBeam& ob : out.beams)
      if (ob.tri >= 0) P->v.push_back(ob);
  }
  putStats(btr.stats, stats);
  return P;
}

// Point-light shadows of the given primary hit polygons (render.cpp:209).
void* ref_point(void* sh, void* kh, const double* cam, int W, int H, const double* px, const double* py,
                const int32_t* pn, const int32_t* ptri, const double* parea, int64_t nPrim, const double* light,
                int mailbox, int trail, double* stats) {
  const Scene& s = *static_cast<Scene*>(sh);
  const KdTree& t = *static_cast<KdTree*>(kh);
  Camera c = makeCam(cam, W, H);
  BeamTracer btr(s, t);
  btr.useMailbox = mailbox != 0;
  btr.useTrail = trail != 0;
  BeamOutput prim;
  for (int64_t i = 0; i < nPrim; ++i) prim.beams.push_back(polyFrom(px + 4 * i, py + 4 * i, pn[i], ptri[i], parea[i]));
  Polys* P = new Polys();
  RenderStats rs;
  beamPointShadows(btr, s, c, prim, Vec3(Real(light[0]), Real(light[1]), Real(light[2])), P->v, rs);
  putStats(rs.trace, stats);
  return P;
}

int64_t ref_polys_count(void* h) { return int64_t(static_cast<Polys*>(h)->v.size()); }

void ref_polys_get(void* h, double* x, double* y, int32_t* n, int32_t* tri, double* area) {
  const Polys& P = *static_cast<Polys*>(h);
  for (size_t i = 0; i < P.v.size(); ++i) {
    const OutBeam& ob = P.v[i];
    for (int k = 0; k < 4; ++k) {
      x[4 * i + k] = double(ob.x[k]);
      y[4 * i + k] = double(ob.y[k]);
    }
    n[i] = ob.n;
    tri[i] = ob.tri;
    area[i] = double(ob.area);
  }
}

void ref_polys_free(void* h) { delete static_cast<Polys*>(h); }

// rasterize (render.cpp:49) with one sample per pixel at the centre.
void ref_raster(const double* cam, int W, int H, const double* px, const double* py, const int32_t* pn,
                const int32_t* ptri, int64_t n, int shadowPass, int32_t* triOut, uint8_t* flagOut) {
  Camera c = makeCam(cam, W, H);
  std::vector<OutBeam> v;
  for (int64_t i = 0; i < n; ++i) v.push_back(polyFrom(px + 4 * i, py + 4 * i, pn[i], ptri[i], 0.0));
  SampleBuffer sb;
  sb.init(W, H, 1);
  if (!shadowPass && triOut) std::memcpy(sb.tri.data(), triOut, sb.tri.size() * sizeof(int32_t));
  if (shadowPass && flagOut) std::memcpy(sb.flag.data(), flagOut, sb.flag.size());
  rasterize(v, c, sb, shadowPass != 0);
  if (triOut) std::memcpy(triOut, sb.tri.data(), sb.tri.size() * sizeof(int32_t));
  if (flagOut) std::memcpy(flagOut, sb.flag.data(), sb.flag.size());
}

// Primary rays through the pixel centres (the soft-shadow primary pass,
// render.cpp:351-381): tri, X and N (facing the eye) per pixel.
void ref_primary_rays(void* sh, void* kh, const double* cam, int W, int H, int32_t* tri, double* X, double* N,
                      double* stats) {
  const Scene& s = *static_cast<Scene*>(sh);
  RayTracer rt(s, *static_cast<KdTree*>(kh));
  Camera c = makeCam(cam, W, H);
  TraceStats st;
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x) {
      size_t k = size_t(y) * size_t(W) + size_t(x);
      Real qx, qy;
      c.pixelToQ(Real(x) + Real(0.5), Real(y) + Real(0.5), qx, qy);
      Ray r;
      r.o = c.eye;
      r.d = c.dir(qx, qy);
      RayHit h;
      tri[k] = -1;
      if (!rt.intersect(r, h, st)) continue;
      tri[k] = h.tri;
      Vec3 p = r.o + r.d * h.t;
      Vec3 n = s.triN[size_t(h.tri)];
      if (dot(n, r.d) > 0) n = -n;
      for (int a = 0; a < 3; ++a) {
        X[3 * k + a] = double(p[a]);
        N[3 * k + a] = double(n[a]);
      }
    }
  putStats(st, stats);
}

}  // extern "C"
