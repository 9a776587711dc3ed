// Scene finalisation, geometry helpers, camera, rasterizer.
#include <algorithm>
#include <cmath>

#include "scene.h"

namespace bth {

// src/scene/scene.cpp:10-39. The normal and the degeneracy test are computed
// in double so that both reference builds drop the same triangles.
void Scene::finalize() {
  std::vector<std::array<uint32_t, 3>> kept;
  kept.reserve(tris.size());
  triN.clear();
  triArea.clear();
  bounds = Box3();
  for (const auto& t : tris) {
    const V3 &a = pos[t[0]], &b = pos[t[1]], &c = pos[t[2]];
    const double ux = double(b.x) - a.x, uy = double(b.y) - a.y, uz = double(b.z) - a.z;
    const double wx = double(c.x) - a.x, wy = double(c.y) - a.y, wz = double(c.z) - a.z;
    const double nx = uy * wz - uz * wy, ny = uz * wx - ux * wz, nz = ux * wy - uy * wx;
    const double l = std::sqrt(nx * nx + ny * ny + nz * nz);
    const double scale = std::sqrt(ux * ux + uy * uy + uz * uz) * std::sqrt(wx * wx + wy * wy + wz * wz);
    const bool repeated = t[0] == t[1] || t[1] == t[2] || t[0] == t[2];
    if (!(l > 1e-7 * scale) || repeated) continue;
    kept.push_back(t);
    triN.push_back(V3(float(nx / l), float(ny / l), float(nz / l)));
    triArea.push_back(float(0.5 * l));
    bounds.grow(a);
    bounds.grow(b);
    bounds.grow(c);
  }
  tris.swap(kept);
}

void quad(Scene& s, V3 a, V3 b, V3 c, V3 d) {
  uint32_t i0 = s.vert(a), i1 = s.vert(b), i2 = s.vert(c), i3 = s.vert(d);
  s.tri(i0, i1, i2);
  s.tri(i0, i2, i3);
}

void gridQuad(Scene& s, V3 a, V3 U, V3 V, int nu, int nv, const std::vector<std::array<float, 4>>& holes) {
  const uint32_t base = uint32_t(s.pos.size());
  for (int j = 0; j <= nv; ++j)
    for (int i = 0; i <= nu; ++i) s.vert(a + U * (float(i) / float(nu)) + V * (float(j) / float(nv)));
  auto at = [&](int i, int j) { return base + uint32_t(j * (nu + 1) + i); };
  for (int j = 0; j < nv; ++j)
    for (int i = 0; i < nu; ++i) {
      const float cs = (float(i) + 0.5f) / float(nu), ct = (float(j) + 0.5f) / float(nv);
      bool cut = false;
      for (const auto& h : holes) cut = cut || (cs > h[0] && cs < h[2] && ct > h[1] && ct < h[3]);
      if (cut) continue;
      s.tri(at(i, j), at(i + 1, j), at(i + 1, j + 1));
      s.tri(at(i, j), at(i + 1, j + 1), at(i, j + 1));
    }
}

void box(Scene& s, V3 lo, V3 hi) {
  uint32_t v[8];
  for (int i = 0; i < 8; ++i) v[i] = s.vert(V3((i & 1) ? hi.x : lo.x, (i & 2) ? hi.y : lo.y, (i & 4) ? hi.z : lo.z));
  // outward-facing faces, two triangles each
  static const int f[6][4] = {{0, 2, 6, 4}, {1, 5, 7, 3}, {0, 4, 5, 1}, {2, 3, 7, 6}, {0, 1, 3, 2}, {4, 6, 7, 5}};
  for (const auto& q : f) {
    s.tri(v[q[0]], v[q[1]], v[q[2]]);
    s.tri(v[q[0]], v[q[2]], v[q[3]]);
  }
}

void cylinder(Scene& s, V3 base, V3 axis, float r, int seg, bool caps) {
  V3 an = unit3(axis);
  V3 e1 = orthoUnit(an), e2 = cross3(an, e1);
  const uint32_t b0 = uint32_t(s.pos.size());
  for (int i = 0; i < seg; ++i) {
    const double ang = 2 * M_PI * i / seg;
    V3 d = e1 * float(std::cos(ang) * r) + e2 * float(std::sin(ang) * r);
    s.vert(base + d);
    s.vert(base + d + axis);
  }
  for (int i = 0; i < seg; ++i) {
    const uint32_t j = uint32_t((i + 1) % seg);
    const uint32_t lo0 = b0 + 2 * uint32_t(i), hi0 = lo0 + 1, lo1 = b0 + 2 * j, hi1 = lo1 + 1;
    s.tri(lo0, lo1, hi1);
    s.tri(lo0, hi1, hi0);
  }
  if (caps) {
    const uint32_t cb = s.vert(base), ct = s.vert(base + axis);
    for (int i = 0; i < seg; ++i) {
      const uint32_t j = uint32_t((i + 1) % seg);
      s.tri(cb, b0 + 2 * j, b0 + 2 * uint32_t(i));
      s.tri(ct, b0 + 2 * uint32_t(i) + 1, b0 + 2 * j + 1);
    }
  }
}

// render.cpp:14-25: tan in double, everything else in float.
Camera Camera::make(const View& v, int W_, int H_) {
  Camera c;
  c.eye = v.eye;
  c.fwd = unit3(v.target - v.eye);
  c.right = unit3(cross3(c.fwd, v.up));
  c.down = cross3(c.fwd, c.right);
  c.W = W_;
  c.H = H_;
  c.halfH = float(std::tan(double(v.fovY) * M_PI / 360.0));
  c.halfW = c.halfH * float(W_) / float(H_);
  return c;
}

// beam_types.h:40-51.
Plane Plane::make(V3 O, V3 p0, V3 u_, V3 v_) {
  Plane P;
  P.O = O;
  P.p0 = p0;
  V3 n = unit3(cross3(u_, v_));
  if (dot3(n, p0 - O) < 0) n = -n;
  P.u = unit3(u_);
  P.v = cross3(n, P.u);
  P.n = n;
  P.d = p0 - O;
  return P;
}

bool polyLess(const Poly& a, const Poly& b) {
  if (a.tri != b.tri) return a.tri < b.tri;
  if (a.n != b.n) return a.n < b.n;
  for (int k = 0; k < a.n; ++k) {
    if (a.x[k] != b.x[k]) return a.x[k] < b.x[k];
    if (a.y[k] != b.y[k]) return a.y[k] < b.y[k];
  }
  return a.area < b.area;
}

// render.cpp:49-95 with one sample at the pixel centre: each polygon's edges
// become half-planes A x + B y + C >= 0 in pixel coordinates (double); rows
// are spans between the tightest left and right bounds.
void rasterize(const std::vector<Poly>& polys, const Camera& cam, int32_t* triOut, uint8_t* flagOut, bool shadowPass) {
  const double ox = 0.5, oy = 0.5;
  for (const Poly& pg : polys) {
    if (!shadowPass && pg.tri < 0) continue;
    double px[4], py[4], A[4], B[4], C[4];
    double ylo = 1e300, yhi = -1e300;
    for (int i = 0; i < pg.n; ++i) {
      const float x = (pg.x[i] / cam.halfW + 1) * 0.5f * float(cam.W);
      const float y = (pg.y[i] / cam.halfH + 1) * 0.5f * float(cam.H);
      px[i] = double(x);
      py[i] = double(y);
      ylo = std::min(ylo, py[i]);
      yhi = std::max(yhi, py[i]);
    }
    for (int i = 0; i < pg.n; ++i) {
      const int j = i + 1 == pg.n ? 0 : i + 1;
      A[i] = -(py[j] - py[i]);
      B[i] = px[j] - px[i];
      C[i] = -(A[i] * px[i] + B[i] * py[i]);
    }
    const int y0 = std::max(0, int(std::ceil(ylo - oy))), y1 = std::min(cam.H - 1, int(std::floor(yhi - oy)));
    for (int y = y0; y <= y1; ++y) {
      const double sy = y + oy;
      double xl = -1e300, xr = 1e300;
      bool empty = false;
      for (int i = 0; i < pg.n && !empty; ++i) {
        const double r = B[i] * sy + C[i];
        if (A[i] > 0)
          xl = std::max(xl, -r / A[i]);
        else if (A[i] < 0)
          xr = std::min(xr, -r / A[i]);
        else if (r < 0)
          empty = true;
      }
      if (empty || xl > xr) continue;
      const int x0 = std::max(0, int(std::ceil(xl - ox))), x1 = std::min(cam.W - 1, int(std::floor(xr - ox)));
      for (int x = x0; x <= x1; ++x) {
        const size_t k = size_t(y) * size_t(cam.W) + size_t(x);
        if (shadowPass)
          flagOut[k] = 1;
        else
          triOut[k] = pg.tri;
      }
    }
  }
}

}  // namespace bth
