// Host-side input stage of the beam tracer: triangle scenes, the SAH kd-tree
// the beams traverse, pinhole cameras and the polygon rasterizer.
//
// These produce exactly the arrays the reference's GPU path consumes
// (SceneView, src/beam/beam_types.h:86-97): positions, vertex indices, unit
// normals, kd nodes with area-sorted leaves, parent links and node boxes.
// The kd build is bit-identical to the reference's KdTree::build
// (src/accel/kdtree.cpp:237-281) for the same triangles, so the device beams
// take the same splits as the reference CPU build.
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "vecmath.h"

namespace bth {

// Scene::finalize of the reference (src/scene/scene.cpp:10-39): degenerate
// triangles are dropped; normals and areas are computed in double.
struct Scene {
  std::vector<V3> pos;
  std::vector<std::array<uint32_t, 3>> tris;
  std::vector<V3> triN;
  std::vector<float> triArea;
  Box3 bounds;

  uint32_t vert(V3 p) {
    pos.push_back(p);
    return uint32_t(pos.size() - 1);
  }
  void tri(uint32_t a, uint32_t b, uint32_t c) { tris.push_back({a, b, c}); }
  int numTris() const { return int(tris.size()); }
  V3 v(int t, int k) const { return pos[tris[size_t(t)][size_t(k)]]; }
  void finalize();
};

// Geometry helpers for the generators.
void quad(Scene& s, V3 a, V3 b, V3 c, V3 d);
// Grid of nu x nv quads spanning a + s*U + t*V, s,t in [0,1]; cells whose
// centre is inside one of the holes (rectangles in (s,t), 4 floats each) are
// left out. Vertices are shared inside the grid.
void gridQuad(Scene& s, V3 a, V3 U, V3 V, int nu, int nv, const std::vector<std::array<float, 4>>& holes = {});
void box(Scene& s, V3 lo, V3 hi);
void cylinder(Scene& s, V3 base, V3 axis, float r, int seg, bool caps);

struct KdNode {
  float split = 0;
  uint32_t data = 3;    // low 2 bits: axis, or 3 for a leaf; high bits: right child or count
  uint32_t offset = 0;  // leaf: first index into triIndices
  bool leaf() const { return (data & 3u) == 3u; }
  int axis() const { return int(data & 3u); }
  uint32_t hiBits() const { return data >> 2; }
};
static_assert(sizeof(KdNode) == 12, "kd node layout");

struct KdParams {
  float costTraverse = 1.0f;
  float costIntersect = 0.4f;  // beams (the reference's bt_render, apps/bt_render.cpp:170-174)
  float emptyBonus = 0.8f;
  int maxDepth = -1;           // -1: 8 + 1.3 log2(N), at most 63
  int binThreshold = 2048;
  int bins = 64;
  bool singleLeaf = false;
};

struct KdTree {
  std::vector<KdNode> nodes;
  std::vector<uint32_t> triIndices;
  std::vector<uint32_t> parent;
  std::vector<Box3> nodeBox;
  Box3 bounds;
  int depth = 0;
  void build(const Scene& s, const KdParams& p = KdParams());
};

struct View {
  V3 eye, target, up{0, 1, 0};
  float fovY = 60;
};

// Pinhole camera (render_types.h:15-39, render.cpp:14-25).
struct Camera {
  V3 eye, fwd, right, down;
  float halfW = 1, halfH = 1;
  int W = 1, H = 1;
  static Camera make(const View& v, int W, int H);
  void pixelToQ(float px, float py, float& qx, float& qy) const {
    qx = (2 * px / float(W) - 1) * halfW;
    qy = (2 * py / float(H) - 1) * halfH;
  }
};

// The plane on which a beam tree's ray directions live (beam_types.h:14-52).
struct Plane {
  V3 O, p0, u, v, n, d;
  static Plane make(V3 O, V3 p0, V3 u_, V3 v_);
};

// Polygon in image-plane coordinates (OutBeam, beam_types.h:68-73).
struct Poly {
  float x[4], y[4];
  int32_t n;
  int32_t tri;
  float area;
};

// Canonical order of polygons (tie-break for rasterization, like beamLess in
// the reference's tests/test_wavefront.cpp:13-21).
bool polyLess(const Poly& a, const Poly& b);

// One sample per pixel at the centre (render.cpp:49-95, aa = 1): writes tri ids
// (shadowPass = false, polygons with tri < 0 skipped) or sets flags.
void rasterize(const std::vector<Poly>& polys, const Camera& cam, int32_t* triOut, uint8_t* flagOut, bool shadowPass);

// Named synthetic scenes (seeded). Returns false for an unknown name.
struct SceneSpec {
  View view;
  bool hasPoint = false;
  V3 pointLight;
  bool hasArea = false;
  V3 lightC, lightU, lightV;
  bool cull = false;
};
bool generate(const std::string& name, uint64_t seed, Scene& s, SceneSpec& spec, std::string& err);

}  // namespace bth
