// Seeded synthetic scenes for the benchmark configurations and the tests.
//
//   room      (~900 tris)  10 x 8 x 3 m room, grid-tessellated shell, 12 boxes;
//                          camera at the centre, 60 deg; 8x8-pixel tile roots
//   building  (~2.1M tris) 7 floors of 12 x 8 rooms, tessellated walls with
//                          door and window holes, desks and chairs; point light
//   plant     (~5.4K tris) curved stem with ~290 bent 3x3-grid leaves (2-6 cm)
//                          over a ground quad; 0.5 x 0.5 m area light 1 m above
//   analytic               floor + square occluder under a square light
//                          (closed-form soft shadows)
//   random_tris            interpenetrating random triangles (tests)
//   sphere                 one UV sphere (crack tests)
//
// The generators are this project's own; only their triangle counts follow
// the benchmark descriptions. Random numbers come from PCG32 (O'Neill 2014,
// pcg32_srandom / pcg32_random as published at pcg-random.org).
#include <cmath>
#include <map>

#include "scene.h"

namespace bth {

namespace {

struct Pcg32 {
  uint64_t state = 0, inc = 0;
  Pcg32(uint64_t seed, uint64_t seq) {
    inc = (seq << 1u) | 1u;
    next();
    state += seed;
    next();
  }
  uint32_t next() {
    const uint64_t old = state;
    state = old * 6364136223846793005ULL + inc;
    const uint32_t xs = uint32_t(((old >> 18u) ^ old) >> 27u);
    const uint32_t rot = uint32_t(old >> 59u);
    return (xs >> rot) | (xs << ((32u - rot) & 31u));
  }
  double u01() { return double(next()) * (1.0 / 4294967296.0); }
  double range(double a, double b) { return a + (b - a) * u01(); }
};

V3 lerp3(V3 a, V3 b, float t) { return a + (b - a) * t; }

// ---- room: the primary-visibility configuration ---------------------------
void buildRoom(Scene& s, SceneSpec& spec, uint64_t seed) {
  const float X = 10, Y = 3, Z = 8;
  // floor (normal +y), ceiling (-y), walls facing inward; 12 x 10 / 12 x 3 / 10 x 3 cells
  gridQuad(s, V3(0, 0, 0), V3(0, 0, Z), V3(X, 0, 0), 10, 12);
  gridQuad(s, V3(0, Y, 0), V3(X, 0, 0), V3(0, 0, Z), 12, 10);
  gridQuad(s, V3(0, 0, 0), V3(X, 0, 0), V3(0, Y, 0), 12, 3);
  gridQuad(s, V3(0, 0, Z), V3(0, Y, 0), V3(X, 0, 0), 3, 12);
  gridQuad(s, V3(0, 0, 0), V3(0, Y, 0), V3(0, 0, Z), 3, 10);
  gridQuad(s, V3(X, 0, 0), V3(0, 0, Z), V3(0, Y, 0), 10, 3);
  Pcg32 r(seed, 1);
  int placed = 0;
  while (placed < 12) {
    const float w = float(r.range(0.4, 1.8)), d = float(r.range(0.4, 1.4)), h = float(r.range(0.3, 1.3));
    const float x = float(r.range(0.1, X - 0.1 - w)), z = float(r.range(0.1, Z - 0.1 - d));
    // keep the camera (room centre) clear
    const float cx = x + w / 2 - 5.0f, cz = z + d / 2 - 4.0f;
    if (cx * cx + cz * cz < 2.0f) continue;
    box(s, V3(x, 0, z), V3(x + w, h, z + d));
    ++placed;
  }
  spec.view.eye = V3(5.0f, 1.5f, 4.0f);
  spec.view.target = V3(9.0f, 0.9f, 6.5f);
  spec.view.fovY = 60;
  spec.hasPoint = true;
  spec.pointLight = V3(3.0f, 2.7f, 2.0f);
  spec.hasArea = true;
  spec.lightC = V3(5.0f, 2.95f, 4.0f);
  spec.lightU = V3(1, 0, 0);
  spec.lightV = V3(0, 0, 1);
}

// ---- plant: the exact soft-shadow configuration ---------------------------
void buildPlant(Scene& s, SceneSpec& spec, uint64_t seed) {
  Pcg32 r(seed, 3);
  // ground 2.4 x 2.4 m
  quad(s, V3(-1.2f, 0, -1.2f), V3(-1.2f, 0, 1.2f), V3(1.2f, 0, 1.2f), V3(1.2f, 0, -1.2f));
  // curved stem: a bent tube of 10 rings x 6 sides
  const int rings = 10, sides = 6;
  std::vector<V3> axis;
  for (int i = 0; i <= rings; ++i) {
    const float t = float(i) / float(rings);
    axis.push_back(V3(0.08f * std::sin(2.2f * t), 0.55f * t, 0.05f * (1 - std::cos(2.6f * t))));
  }
  const uint32_t sb = uint32_t(s.pos.size());
  for (int i = 0; i <= rings; ++i) {
    const float rad = 0.012f * (1.0f - 0.5f * float(i) / float(rings));
    for (int k = 0; k < sides; ++k) {
      const double a = 2 * M_PI * k / sides;
      s.vert(axis[size_t(i)] + V3(float(std::cos(a)) * rad, 0, float(std::sin(a)) * rad));
    }
  }
  for (int i = 0; i < rings; ++i)
    for (int k = 0; k < sides; ++k) {
      const uint32_t a = sb + uint32_t(i * sides + k), b = sb + uint32_t(i * sides + (k + 1) % sides);
      s.tri(a, b, b + sides);
      s.tri(a, b + sides, a + sides);
    }
  // leaves: 3 x 3 quad grid each, bent along its length, random orientation
  const int leaves = 292;
  for (int l = 0; l < leaves; ++l) {
    const float t = float(r.range(0.3, 1.0));
    const int ai = std::min(rings - 1, int(t * rings));
    const V3 at = lerp3(axis[size_t(ai)], axis[size_t(ai) + 1], t * rings - float(ai));
    const V3 off(float(r.range(-0.22, 0.22)), float(r.range(-0.08, 0.12)), float(r.range(-0.22, 0.22)));
    const V3 c = at + off * t;
    const V3 dirL = unit3(V3(float(r.range(-1, 1)), float(r.range(-0.4, 0.9)), float(r.range(-1, 1))));
    const V3 side = orthoUnit(dirL);
    const V3 nrm = cross3(dirL, side);
    const float len = float(r.range(0.02, 0.06)), wid = len * float(r.range(0.35, 0.6));
    const float bend = float(r.range(0.1, 0.35)) * len;
    const uint32_t b0 = uint32_t(s.pos.size());
    for (int j = 0; j <= 3; ++j)
      for (int i = 0; i <= 3; ++i) {
        const float u = float(i) / 3.0f - 0.5f, v = float(j) / 3.0f;
        // bent: the leaf curls down along its length and cups across it
        const float h = -bend * v * v + 0.25f * bend * (4 * u * u - 1);
        s.vert(c + dirL * (v * len) + side * (u * wid) + nrm * h);
      }
    for (int j = 0; j < 3; ++j)
      for (int i = 0; i < 3; ++i) {
        const uint32_t a = b0 + uint32_t(j * 4 + i);
        s.tri(a, a + 1, a + 5);
        s.tri(a, a + 5, a + 4);
      }
  }
  spec.view.eye = V3(0.75f, 0.62f, 0.95f);
  spec.view.target = V3(0.0f, 0.22f, 0.0f);
  spec.view.fovY = 45;
  spec.hasPoint = true;
  spec.pointLight = V3(0.3f, 1.5f, 0.2f);
  spec.hasArea = true;
  spec.lightC = V3(0.1f, 1.55f, 0.05f);  // 1 m above the top of the plant
  spec.lightU = V3(0.5f, 0, 0);
  spec.lightV = V3(0, 0, 0.5f);
}

// ---- building: the point-shadow configuration -----------------------------
void buildBuilding(Scene& s, SceneSpec& spec, uint64_t seed, int floors, int nx, int nz, float cellsPerM) {
  Pcg32 r(seed, 5);
  const float Wr = 4.0f, Hr = 3.0f, slab = 0.2f;
  const int cu = int(Wr * cellsPerM), cv = int(Hr * cellsPerM);
  // door: s in [0.38, 0.62], t < 0.7 ; window: s in [0.2, 0.8], t in [0.35, 0.8]
  const std::vector<std::array<float, 4>> door{{0.38f, -1.0f, 0.62f, 0.7f}};
  const std::vector<std::array<float, 4>> window{{0.2f, 0.35f, 0.8f, 0.8f}};
  for (int f = 0; f < floors; ++f) {
    const float y0 = float(f) * (Hr + slab);
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < nz; ++j) {
        const V3 o(float(i) * Wr, y0, float(j) * Wr);
        gridQuad(s, o, V3(0, 0, Wr), V3(Wr, 0, 0), cu / 2, cu / 2);                       // floor
        gridQuad(s, o + V3(0, Hr, 0), V3(Wr, 0, 0), V3(0, 0, Wr), cu / 4, cu / 4);        // ceiling
        gridQuad(s, o, V3(0, 0, Wr), V3(0, Hr, 0), cu, cv, i == 0 ? window : door);       // west wall
        gridQuad(s, o, V3(Wr, 0, 0), V3(0, Hr, 0), cu, cv, j == 0 ? window : door);       // south wall
        if (i == nx - 1) gridQuad(s, o + V3(Wr, 0, 0), V3(0, 0, Wr), V3(0, Hr, 0), cu, cv, window);
        if (j == nz - 1) gridQuad(s, o + V3(0, 0, Wr), V3(Wr, 0, 0), V3(0, Hr, 0), cu, cv, window);
        // desk (top + 4 legs) and chair (seat, back, 4 legs), jittered
        const float dx = float(r.range(0.6, 1.2)), dz = float(r.range(2.2, 2.8));
        box(s, o + V3(dx, 0.72f, dz), o + V3(dx + 1.4f, 0.76f, dz + 0.7f));
        for (int k = 0; k < 4; ++k) {
          const float lx = dx + ((k & 1) ? 1.34f : 0.0f), lz = dz + ((k & 2) ? 0.64f : 0.0f);
          box(s, o + V3(lx, 0, lz), o + V3(lx + 0.06f, 0.72f, lz + 0.06f));
        }
        const float cx = dx + float(r.range(0.3, 0.8)), cz = dz - 0.6f;
        box(s, o + V3(cx, 0.42f, cz), o + V3(cx + 0.45f, 0.46f, cz + 0.45f));
        box(s, o + V3(cx, 0.46f, cz - 0.04f), o + V3(cx + 0.45f, 0.95f, cz));
        for (int k = 0; k < 4; ++k) {
          const float lx = cx + ((k & 1) ? 0.41f : 0.0f), lz = cz + ((k & 2) ? 0.41f : 0.0f);
          box(s, o + V3(lx, 0, lz), o + V3(lx + 0.04f, 0.42f, lz + 0.04f));
        }
      }
  }
  // camera on the 2nd floor looking along a row of rooms through the doors
  spec.view.eye = V3(2.0f, Hr + slab + 1.6f, 1.8f);
  spec.view.target = V3(2.0f + 4 * Wr, Hr + slab + 1.2f, 2.2f + 2 * Wr);
  spec.view.fovY = 70;
  spec.hasPoint = true;
  spec.pointLight = V3(2.6f, Hr + slab + 2.6f, 2.4f);
  spec.hasArea = true;
  spec.lightC = V3(2.0f, Hr + slab + 2.9f, 2.0f);
  spec.lightU = V3(0.8f, 0, 0);
  spec.lightV = V3(0, 0, 0.8f);
}

// ---- test scenes ----------------------------------------------------------
void buildAnalytic(Scene& s, SceneSpec& spec, float h, float occ) {
  const float F = 10, a = occ / 2;
  quad(s, V3(-F, 0, -F), V3(-F, 0, F), V3(F, 0, F), V3(F, 0, -F));
  quad(s, V3(-a, h, -a), V3(-a, h, a), V3(a, h, a), V3(a, h, -a));
  spec.view.eye = V3(0, 5, 5);
  spec.view.target = V3(0, 0, 0);
  spec.view.fovY = 60;
  spec.hasPoint = true;
  spec.pointLight = V3(0.3f, 3, 0.2f);
  spec.hasArea = true;
  spec.lightC = V3(0, 3, 0);
  spec.lightU = V3(1, 0, 0);
  spec.lightV = V3(0, 0, 1);
}

void buildRandomTris(Scene& s, SceneSpec& spec, uint64_t seed, int n) {
  Pcg32 r(seed, 7);
  for (int i = 0; i < n; ++i) {
    const V3 c(float(r.u01()), float(r.u01()), float(r.u01()));
    uint32_t idx[3];
    for (int k = 0; k < 3; ++k)
      idx[k] = s.vert(c + V3(float(r.range(-0.3, 0.3)), float(r.range(-0.3, 0.3)), float(r.range(-0.3, 0.3))));
    s.tri(idx[0], idx[1], idx[2]);
  }
  spec.view.eye = V3(1.8f, 1.6f, 2.2f);
  spec.view.target = V3(0.5f, 0.5f, 0.5f);
  spec.view.fovY = 60;
  spec.hasPoint = true;
  spec.pointLight = V3(0.4f, 2, 0.6f);
  spec.hasArea = true;
  spec.lightC = V3(0.5f, 2.5f, 0.5f);
  spec.lightU = V3(0.5f, 0, 0);
  spec.lightV = V3(0, 0, 0.5f);
}

void buildSphere(Scene& s, SceneSpec& spec, int stacks) {
  const int slices = 2 * stacks;
  const uint32_t top = s.vert(V3(0, 1, 0)), bot = s.vert(V3(0, -1, 0));
  std::vector<uint32_t> ring;
  for (int i = 1; i < stacks; ++i)
    for (int j = 0; j < slices; ++j) {
      const double th = M_PI * i / stacks, ph = 2 * M_PI * j / slices;
      ring.push_back(s.vert(V3(float(std::sin(th) * std::cos(ph)), float(std::cos(th)), float(std::sin(th) * std::sin(ph)))));
    }
  auto at = [&](int i, int j) { return ring[size_t((i - 1) * slices + (j % slices))]; };
  for (int j = 0; j < slices; ++j) {
    s.tri(top, at(1, j + 1), at(1, j));
    s.tri(bot, at(stacks - 1, j), at(stacks - 1, j + 1));
  }
  for (int i = 1; i < stacks - 1; ++i)
    for (int j = 0; j < slices; ++j) {
      s.tri(at(i, j), at(i, j + 1), at(i + 1, j + 1));
      s.tri(at(i, j), at(i + 1, j + 1), at(i + 1, j));
    }
  spec.view.eye = V3(0, 0.6f, 3.2f);
  spec.view.target = V3(0, 0, 0);
  spec.view.fovY = 50;
  spec.hasPoint = true;
  spec.pointLight = V3(2, 3, 2);
  spec.hasArea = true;
  spec.lightC = V3(0, 3, 0);
  spec.lightU = V3(1, 0, 0);
  spec.lightV = V3(0, 0, 1);
}

}  // namespace

bool generate(const std::string& name, uint64_t seed, Scene& s, SceneSpec& spec, std::string& err) {
  s = Scene();
  spec = SceneSpec();
  if (name == "room")
    buildRoom(s, spec, seed);
  else if (name == "plant")
    buildPlant(s, spec, seed);
  else if (name == "building")
    buildBuilding(s, spec, seed, 7, 12, 8, 7.5f);
  else if (name == "building_small")
    buildBuilding(s, spec, seed, 2, 3, 2, 8);
  else if (name == "analytic")
    buildAnalytic(s, spec, 1.0f, 1.0f);
  else if (name == "random_tris")
    buildRandomTris(s, spec, seed, 150);
  else if (name == "sphere")
    buildSphere(s, spec, 24);
  else {
    err = "unknown scene '" + name + "'";
    return false;
  }
  s.finalize();
  if (s.numTris() == 0) {
    err = "scene has no triangles";
    return false;
  }
  return true;
}

}  // namespace bth
