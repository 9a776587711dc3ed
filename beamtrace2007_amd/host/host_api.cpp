// C ABI of the host input stage (scenes, kd-trees, cameras, rasterizer):
// what callers without a C++ scene library (Python, tests, the benchmark) use
// to produce the arrays btg_create() uploads. See include/bth.h.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "../../include/bth.h"
#include "scene.h"
#include "scene_file.h"

using namespace bth;

struct bth_scene {
  Scene s;
  SceneSpec spec;
  KdTree kd;
  bool hasKd = false;
};

namespace {
thread_local std::string g_err;

int fail(const std::string& m) {
  g_err = m;
  return -1;
}

View viewFrom(const float* v) {
  View w;
  w.eye = V3(v[0], v[1], v[2]);
  w.target = V3(v[3], v[4], v[5]);
  w.up = V3(v[6], v[7], v[8]);
  w.fovY = v[9];
  return w;
}

Camera camFrom(const float* c14, int W, int H) {
  Camera c;
  c.eye = V3(c14[0], c14[1], c14[2]);
  c.fwd = V3(c14[3], c14[4], c14[5]);
  c.right = V3(c14[6], c14[7], c14[8]);
  c.down = V3(c14[9], c14[10], c14[11]);
  c.halfW = c14[12];
  c.halfH = c14[13];
  c.W = W;
  c.H = H;
  return c;
}
}  // namespace

extern "C" {

const char* bth_last_error(void) { return g_err.c_str(); }

int bth_scene_generate(const char* name, uint64_t seed, bth_scene** out) {
  try {
    bth_scene* h = new bth_scene();
    std::string err;
    if (!generate(name ? name : "", seed, h->s, h->spec, err)) {
      delete h;
      return fail(err);
    }
    *out = h;
    return 0;
  } catch (const std::exception& e) {
    return fail(e.what());
  }
}

int bth_scene_from_arrays(const float* pos, int64_t nv, const uint32_t* tris, int64_t nt, bth_scene** out) {
  try {
    bth_scene* h = new bth_scene();
    h->s.pos.resize(size_t(nv));
    for (int64_t i = 0; i < nv; ++i) h->s.pos[size_t(i)] = V3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
    for (int64_t t = 0; t < nt; ++t) {
      for (int k = 0; k < 3; ++k)
        if (tris[3 * t + k] >= uint64_t(nv)) {
          delete h;
          return fail("triangle vertex index out of range");
        }
      h->s.tri(tris[3 * t], tris[3 * t + 1], tris[3 * t + 2]);
    }
    h->s.finalize();
    *out = h;
    return 0;
  } catch (const std::exception& e) {
    return fail(e.what());
  }
}

void bth_scene_free(bth_scene* h) { delete h; }

int64_t bth_scene_num_tris(const bth_scene* h) { return h->s.numTris(); }
int64_t bth_scene_num_verts(const bth_scene* h) { return int64_t(h->s.pos.size()); }

void bth_scene_arrays(const bth_scene* h, const float** pos, const uint32_t** tri3, const float** triN,
                      const float** triArea, float* bounds6) {
  *pos = reinterpret_cast<const float*>(h->s.pos.data());
  *tri3 = reinterpret_cast<const uint32_t*>(h->s.tris.data());
  *triN = reinterpret_cast<const float*>(h->s.triN.data());
  *triArea = h->s.triArea.data();
  for (int k = 0; k < 3; ++k) {
    bounds6[k] = h->s.bounds.lo[k];
    bounds6[3 + k] = h->s.bounds.hi[k];
  }
}

int bth_scene_spec(const bth_scene* h, float* view10, float* point3, float* light9) {
  const SceneSpec& p = h->spec;
  const float v[10] = {p.view.eye.x,    p.view.eye.y,    p.view.eye.z, p.view.target.x, p.view.target.y,
                       p.view.target.z, p.view.up.x,     p.view.up.y,  p.view.up.z,     p.view.fovY};
  std::memcpy(view10, v, sizeof v);
  const float pl[3] = {p.pointLight.x, p.pointLight.y, p.pointLight.z};
  std::memcpy(point3, pl, sizeof pl);
  const float al[9] = {p.lightC.x, p.lightC.y, p.lightC.z, p.lightU.x, p.lightU.y,
                       p.lightU.z, p.lightV.x, p.lightV.y, p.lightV.z};
  std::memcpy(light9, al, sizeof al);
  return (p.hasPoint ? 1 : 0) | (p.hasArea ? 2 : 0) | (p.cull ? 4 : 0);
}

int bth_kd_build(bth_scene* h, float costIntersect, int singleLeaf) {
  try {
    KdParams kp;
    kp.costIntersect = costIntersect;
    kp.singleLeaf = singleLeaf != 0;
    h->kd.build(h->s, kp);
    h->hasKd = true;
    return 0;
  } catch (const std::exception& e) {
    return fail(e.what());
  }
}

int bth_kd_sizes(const bth_scene* h, int64_t* numNodes, int64_t* numRefs, int* depth) {
  if (!h->hasKd) return fail("kd-tree not built");
  *numNodes = int64_t(h->kd.nodes.size());
  *numRefs = int64_t(h->kd.triIndices.size());
  if (depth) *depth = h->kd.depth;
  return 0;
}

int bth_kd_arrays(const bth_scene* h, const void** nodes, const uint32_t** triIndices, const uint32_t** parent,
                  const float** nodeBox, float* bounds6) {
  if (!h->hasKd) return fail("kd-tree not built");
  *nodes = h->kd.nodes.data();
  *triIndices = h->kd.triIndices.data();
  *parent = h->kd.parent.data();
  *nodeBox = reinterpret_cast<const float*>(h->kd.nodeBox.data());
  for (int k = 0; k < 3; ++k) {
    bounds6[k] = h->kd.bounds.lo[k];
    bounds6[3 + k] = h->kd.bounds.hi[k];
  }
  return 0;
}

int bth_scene_save(const bth_scene* h, const char* path) {
  if (!h || !path) return fail("null argument");
  static_assert(sizeof(V3) == 12 && sizeof(KdNode) == 12 && sizeof(Box3) == 24, "array layouts");
  btsf::View v;
  v.h.specFlags = uint32_t(h->spec.hasPoint ? 1 : 0) | (h->spec.hasArea ? 2 : 0) | (h->spec.cull ? 4 : 0);
  v.h.numVerts = int64_t(h->s.pos.size());
  v.h.numTris = int64_t(h->s.tris.size());
  v.h.numNodes = h->hasKd ? int64_t(h->kd.nodes.size()) : 0;
  v.h.numRefs = h->hasKd ? int64_t(h->kd.triIndices.size()) : 0;
  v.h.kdDepth = h->hasKd ? h->kd.depth : 0;
  for (int k = 0; k < 3; ++k) {
    v.h.sceneBounds[k] = h->s.bounds.lo[k];
    v.h.sceneBounds[3 + k] = h->s.bounds.hi[k];
    v.h.kdBounds[k] = h->hasKd ? h->kd.bounds.lo[k] : 0.0f;
    v.h.kdBounds[3 + k] = h->hasKd ? h->kd.bounds.hi[k] : 0.0f;
  }
  bth_scene_spec(h, v.h.view, v.h.point, v.h.light);
  v.pos = reinterpret_cast<const float*>(h->s.pos.data());
  v.tri3 = reinterpret_cast<const uint32_t*>(h->s.tris.data());
  v.triN = reinterpret_cast<const float*>(h->s.triN.data());
  v.triArea = h->s.triArea.data();
  if (h->hasKd) {
    v.nodes = h->kd.nodes.data();
    v.triIndices = h->kd.triIndices.data();
    v.parent = h->kd.parent.data();
    v.nodeBox = reinterpret_cast<const float*>(h->kd.nodeBox.data());
  }
  std::string err;
  return btsf::write(path, v, err) ? 0 : fail(err);
}

int bth_scene_load(const char* path, bth_scene** out) {
  if (!path || !out) return fail("null argument");
  try {
    btsf::Data d;
    std::string err;
    if (!btsf::read(path, d, err)) return fail(err);
    const btsf::Header& hd = d.h;
    for (uint32_t i : d.tri3)
      if (i >= uint64_t(hd.numVerts)) return fail("scene file: vertex index out of range");
    for (uint32_t i : d.triIndices)
      if (i >= uint64_t(hd.numTris)) return fail("scene file: leaf triangle out of range");
    bth_scene* h = new bth_scene();
    h->s.pos.resize(size_t(hd.numVerts));
    std::memcpy(h->s.pos.data(), d.pos.data(), d.pos.size() * 4);
    h->s.tris.resize(size_t(hd.numTris));
    std::memcpy(h->s.tris.data(), d.tri3.data(), d.tri3.size() * 4);
    h->s.triN.resize(size_t(hd.numTris));
    std::memcpy(h->s.triN.data(), d.triN.data(), d.triN.size() * 4);
    h->s.triArea = std::move(d.triArea);
    for (int k = 0; k < 3; ++k) {
      h->s.bounds.lo[k] = hd.sceneBounds[k];
      h->s.bounds.hi[k] = hd.sceneBounds[3 + k];
    }
    SceneSpec& sp = h->spec;
    sp.view = viewFrom(hd.view);
    sp.hasPoint = (hd.specFlags & 1) != 0;
    sp.hasArea = (hd.specFlags & 2) != 0;
    sp.cull = (hd.specFlags & 4) != 0;
    sp.pointLight = V3(hd.point[0], hd.point[1], hd.point[2]);
    sp.lightC = V3(hd.light[0], hd.light[1], hd.light[2]);
    sp.lightU = V3(hd.light[3], hd.light[4], hd.light[5]);
    sp.lightV = V3(hd.light[6], hd.light[7], hd.light[8]);
    if (hd.numNodes > 0) {
      KdTree& kd = h->kd;
      kd.nodes.resize(size_t(hd.numNodes));
      std::memcpy(kd.nodes.data(), d.nodes.data(), d.nodes.size());
      kd.triIndices = std::move(d.triIndices);
      kd.parent = std::move(d.parent);
      kd.nodeBox.resize(size_t(hd.numNodes));
      std::memcpy(kd.nodeBox.data(), d.nodeBox.data(), d.nodeBox.size() * 4);
      for (int k = 0; k < 3; ++k) {
        kd.bounds.lo[k] = hd.kdBounds[k];
        kd.bounds.hi[k] = hd.kdBounds[3 + k];
      }
      kd.depth = hd.kdDepth;
      h->hasKd = true;
    }
    *out = h;
    return 0;
  } catch (const std::exception& e) {
    return fail(e.what());
  }
}

void bth_camera(const float* view10, int W, int H, float* cam14) {
  const Camera c = Camera::make(viewFrom(view10), W, H);
  const float v[14] = {c.eye.x,   c.eye.y,   c.eye.z,  c.fwd.x,  c.fwd.y,  c.fwd.z,  c.right.x,
                       c.right.y, c.right.z, c.down.x, c.down.y, c.down.z, c.halfW, c.halfH};
  std::memcpy(cam14, v, sizeof v);
}

// render.cpp:14-25 with Real = double: the view's floats widened, then every
// step in double (vec.h's dot, cross, normalize; tan of fovY in double).
void bth_camera_d(const float* view10, int W, int H, double* cam14) {
  struct D {
    double x, y, z;
  };
  const auto cross = [](D a, D b) { return D{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; };
  const auto unit = [](D a) {
    const double l = std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
    return l > 0 ? D{a.x / l, a.y / l, a.z / l} : a;
  };
  const D eye{view10[0], view10[1], view10[2]}, target{view10[3], view10[4], view10[5]};
  const D up{view10[6], view10[7], view10[8]};
  const D fwd = unit(D{target.x - eye.x, target.y - eye.y, target.z - eye.z});
  const D right = unit(cross(fwd, up));
  const D down = cross(fwd, right);
  const double halfH = std::tan(double(view10[9]) * M_PI / 360.0);
  const double halfW = halfH * double(W) / double(H);
  const double v[14] = {eye.x, eye.y, eye.z, fwd.x, fwd.y, fwd.z, right.x, right.y, right.z,
                        down.x, down.y, down.z, halfW, halfH};
  std::memcpy(cam14, v, sizeof v);
}

int bth_rasterize(const float* cam14, int W, int H, const float* x4, const float* y4, const int32_t* n,
                  const int32_t* tri, int64_t count, int shadowPass, int canonical, int32_t* triOut, uint8_t* flagOut) {
  try {
    std::vector<Poly> v(static_cast<size_t>(count));
    for (int64_t i = 0; i < count; ++i) {
      Poly& p = v[size_t(i)];
      for (int k = 0; k < 4; ++k) {
        p.x[k] = x4[4 * i + k];
        p.y[k] = y4[4 * i + k];
      }
      p.n = n[i];
      p.tri = tri[i];
      p.area = 0;
      if (p.n < 0 || p.n > 4) return fail("polygon with more than 4 vertices");
    }
    if (canonical) std::sort(v.begin(), v.end(), polyLess);
    rasterize(v, camFrom(cam14, W, H), triOut, flagOut, shadowPass != 0);
    return 0;
  } catch (const std::exception& e) {
    return fail(e.what());
  }
}

}  // extern "C"
