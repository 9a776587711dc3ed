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
//   blob      (~345K tris) noisy level-7 icosphere with limbs on a ground quad
//   atrium    (~66K tris)  two-storey arcades around a courtyard, cloth strips
//   conference (~125K tris) 20 x 12 m hall, 8 tables, 120 chairs
//
// The generators are this project's own; only their triangle counts follow
// the benchmark descriptions. Random numbers come from PCG32 (O'Neill 2014,
// pcg32_srandom / pcg32_random as published at pcg-random.org).
#include <algorithm>
#include <array>
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


// ---- blob: batched-config scene A (~345K tris) ------------------------------
// An icosphere subdivided 7 times (327,680 triangles) pushed in and out by
// seeded multi-octave radial noise (15% amplitude), with tapered limbs
// extruded from it, standing on a ground quad.
struct SphereNoise {
  struct Wave {
    V3 d;
    double f, ph, a;
  };
  std::vector<Wave> w;
  SphereNoise(Pcg32& r, int octaves, int perOctave, double amp) {
    double a = 1.0, norm = 0;
    for (int o = 0; o < octaves; ++o, a *= 0.5) norm += a * perOctave;
    a = 1.0;
    for (int o = 0; o < octaves; ++o, a *= 0.5)
      for (int k = 0; k < perOctave; ++k) {
        const V3 d = unit3(V3(float(r.range(-1, 1)), float(r.range(-1, 1)), float(r.range(-1, 1))));
        w.push_back({d, 2.0 * double(1 << o) * r.range(0.8, 1.25), r.range(0, 2 * M_PI), amp * a / norm});
      }
  }
  double at(V3 p) const {
    double v = 0;
    for (const Wave& x : w) v += x.a * std::sin(x.f * double(dot3(p, x.d)) + x.ph);
    return v;
  }
};

void buildBlob(Scene& s, SceneSpec& spec, uint64_t seed) {
  Pcg32 r(seed, 11);
  const float Rb = 0.8f;
  const V3 centre(0, 1.0f, 0);
  // icosahedron
  const float t = float((1.0 + std::sqrt(5.0)) / 2.0);
  std::vector<V3> v = {V3(-1, t, 0), V3(1, t, 0), V3(-1, -t, 0), V3(1, -t, 0), V3(0, -1, t), V3(0, 1, t),
                       V3(0, -1, -t), V3(0, 1, -t), V3(t, 0, -1), V3(t, 0, 1), V3(-t, 0, -1), V3(-t, 0, 1)};
  for (V3& x : v) x = unit3(x);
  std::vector<std::array<uint32_t, 3>> f = {{0, 11, 5}, {0, 5, 1},  {0, 1, 7},   {0, 7, 10}, {0, 10, 11},
                                             {1, 5, 9},  {5, 11, 4}, {11, 10, 2}, {10, 7, 6}, {7, 1, 8},
                                             {3, 9, 4},  {3, 4, 2},  {3, 2, 6},   {3, 6, 8},  {3, 8, 9},
                                             {4, 9, 5},  {2, 4, 11}, {6, 2, 10},  {8, 6, 7},  {9, 8, 1}};
  for (int level = 0; level < 7; ++level) {
    std::map<std::pair<uint32_t, uint32_t>, uint32_t> mid;
    auto midpoint = [&](uint32_t a, uint32_t b) {
      const auto key = std::make_pair(std::min(a, b), std::max(a, b));
      auto it = mid.find(key);
      if (it != mid.end()) return it->second;
      v.push_back(unit3((v[a] + v[b]) * 0.5f));
      const uint32_t id = uint32_t(v.size() - 1);
      mid.emplace(key, id);
      return id;
    };
    std::vector<std::array<uint32_t, 3>> g;
    g.reserve(f.size() * 4);
    for (const auto& tr : f) {
      const uint32_t a = midpoint(tr[0], tr[1]), b = midpoint(tr[1], tr[2]), c = midpoint(tr[2], tr[0]);
      g.push_back({tr[0], a, c});
      g.push_back({tr[1], b, a});
      g.push_back({tr[2], c, b});
      g.push_back({a, b, c});
    }
    f.swap(g);
  }
  const SphereNoise noise(r, 4, 5, 0.15);
  const uint32_t base = uint32_t(s.pos.size());
  for (const V3& d : v) s.vert(centre + d * (Rb * float(1.0 + noise.at(d))));
  for (const auto& tr : f) s.tri(base + tr[0], base + tr[1], base + tr[2]);
  // limbs: tapered tubes (24 sides x 60 rings) leaving the surface, capped
  const int limbs = 6, sides = 24, rings = 60;
  for (int l = 0; l < limbs; ++l) {
    V3 d = unit3(V3(float(r.range(-1, 1)), float(r.range(-0.3, 1)), float(r.range(-1, 1))));
    const V3 bend = unit3(V3(float(r.range(-1, 1)), float(r.range(-1, 1)), float(r.range(-1, 1))));
    const float len = float(r.range(0.5, 0.9)), rad0 = float(r.range(0.07, 0.12));
    const V3 start = centre + d * (Rb * float(1.0 + noise.at(d)) * 0.9f);
    const uint32_t lb = uint32_t(s.pos.size());
    V3 p = start, dir = d;
    for (int i = 0; i <= rings; ++i) {
      const float u = float(i) / float(rings);
      const V3 e1 = orthoUnit(dir), e2 = cross3(dir, e1);
      const float rad = rad0 * (1.0f - 0.7f * u);
      for (int k = 0; k < sides; ++k) {
        const double a = 2 * M_PI * k / sides;
        s.vert(p + e1 * float(std::cos(a) * rad) + e2 * float(std::sin(a) * rad));
      }
      dir = unit3(dir + bend * (1.2f / float(rings)));
      p = p + dir * (len / float(rings));
    }
    for (int i = 0; i < rings; ++i)
      for (int k = 0; k < sides; ++k) {
        const uint32_t a = lb + uint32_t(i * sides + k), b = lb + uint32_t(i * sides + (k + 1) % sides);
        s.tri(a, b, b + sides);
        s.tri(a, b + sides, a + sides);
      }
    const uint32_t tip = s.vert(p), last = lb + uint32_t(rings * sides);
    for (int k = 0; k < sides; ++k) s.tri(last + uint32_t(k), last + uint32_t((k + 1) % sides), tip);
  }
  quad(s, V3(-4, 0, -4), V3(-4, 0, 4), V3(4, 0, 4), V3(4, 0, -4));
  spec.view.eye = V3(2.6f, 2.0f, 3.0f);
  spec.view.target = V3(0.0f, 0.8f, 0.0f);
  spec.view.fovY = 50;
  spec.hasPoint = true;
  spec.pointLight = V3(1.5f, 4.0f, 1.0f);
  spec.hasArea = true;
  spec.lightC = V3(0.8f, 4.0f, 0.6f);  // 0.25 m square: light scales 0.25-4 give 6 cm to 1 m
  spec.lightU = V3(0.25f, 0, 0);
  spec.lightV = V3(0, 0, 0.25f);
}

// ---- atrium: batched-config scene B (~66K tris) -----------------------------
// Two stories of arcades (12 arches on each long side of a 26 x 10 m open
// courtyard) on 24-segment cylindrical columns, gallery floors, back walls,
// and cloth strips hanging from the upper gallery.
void arch(Scene& s, V3 a, V3 b, float rise, float depth, float thick, int seg) {
  // a semicircular arch band from a to b (horizontal), extruded along z by depth
  const V3 c = (a + b) * 0.5f, half = (b - a) * 0.5f;
  const float r0 = len3(half), r1 = r0 + thick;
  const V3 ex = half / r0, ey(0, 1, 0), ez(0, 0, depth);
  const uint32_t b0 = uint32_t(s.pos.size());
  for (int i = 0; i <= seg; ++i) {
    const double th = M_PI * i / seg;
    const V3 dir = ex * float(std::cos(th)) + ey * float(std::sin(th) * (rise / r0));
    s.vert(c + dir * r0);
    s.vert(c + dir * r1);
    s.vert(c + dir * r0 + ez);
    s.vert(c + dir * r1 + ez);
  }
  for (int i = 0; i < seg; ++i) {
    const uint32_t p = b0 + 4 * uint32_t(i), q = p + 4;
    s.tri(p, q, q + 2);  // intrados
    s.tri(p, q + 2, p + 2);
    s.tri(p + 1, p + 3, q + 3);  // extrados
    s.tri(p + 1, q + 3, q + 1);
    s.tri(p, p + 1, q + 1);  // front face
    s.tri(p, q + 1, q);
    s.tri(p + 2, q + 2, q + 3);  // back face
    s.tri(p + 2, q + 3, p + 3);
  }
}

void buildAtrium(Scene& s, SceneSpec& spec, uint64_t seed) {
  Pcg32 r(seed, 13);
  const float L = 26, Wc = 10, gal = 3.0f, storey = 4.5f, colR = 0.25f;
  const int bays = 12;
  const float bay = L / float(bays);
  // ground: courtyard + galleries
  gridQuad(s, V3(0, 0, -gal), V3(0, 0, Wc + 2 * gal), V3(L, 0, 0), 16, 52);
  for (int st = 0; st < 2; ++st) {
    const float y0 = float(st) * storey;
    for (int side = 0; side < 2; ++side) {
      const float z = side == 0 ? 0.0f : Wc;
      for (int i = 0; i <= bays; ++i) {
        // column: 24 segments, 10 rings
        const V3 bse(float(i) * bay, y0, z);
        const int rings = 10;
        const float h = storey - 1.6f;
        for (int k = 0; k < rings; ++k)
          cylinder(s, bse + V3(0, h * float(k) / float(rings), 0), V3(0, h / float(rings), 0), colR, 24, false);
      }
      for (int i = 0; i < bays; ++i) {
        const float x0 = float(i) * bay + colR, x1 = float(i + 1) * bay - colR;
        const float zz = side == 0 ? -0.3f : Wc - 0.3f;
        arch(s, V3(x0, y0 + storey - 1.6f, zz), V3(x1, y0 + storey - 1.6f, zz), 1.2f, 0.6f, 0.25f, 24);
      }
      // gallery ceiling / upper floor and back wall
      const float zb = side == 0 ? -gal : Wc + gal;
      gridQuad(s, V3(0, y0 + storey, side == 0 ? -gal : Wc), V3(L, 0, 0), V3(0, 0, gal), 26, 6);
      gridQuad(s, V3(0, y0, zb), V3(0, storey, 0), V3(L, 0, 0), 9, 26);
    }
  }
  // cloth strips hanging from the upper gallery edges, gently waving
  const int strips = 84, cu = 4, cv = 40;
  for (int k = 0; k < strips; ++k) {
    const int side = k % 2;
    const float x = float(r.range(0.5, L - 1.2)), w = float(r.range(0.4, 0.9)), hgt = float(r.range(2.5, 4.0));
    const float z = side == 0 ? 0.35f : Wc - 0.35f, amp = float(r.range(0.05, 0.25)), ph = float(r.range(0, 6.28));
    const uint32_t b0 = uint32_t(s.pos.size());
    for (int j = 0; j <= cv; ++j)
      for (int i = 0; i <= cu; ++i) {
        const float u = float(i) / float(cu), v = float(j) / float(cv);
        const float dz = amp * std::sin(3.0f * v * 3.14159f + ph) * (side == 0 ? 1.0f : -1.0f);
        s.vert(V3(x + u * w, 2 * storey - 0.1f - v * hgt, z + dz));
      }
    for (int j = 0; j < cv; ++j)
      for (int i = 0; i < cu; ++i) {
        const uint32_t a = b0 + uint32_t(j * (cu + 1) + i);
        s.tri(a, a + 1, a + uint32_t(cu + 2));
        s.tri(a, a + uint32_t(cu + 2), a + uint32_t(cu + 1));
      }
  }
  spec.view.eye = V3(-2.0f, 2.2f, Wc / 2 + 1.0f);
  spec.view.target = V3(L / 2, 2.8f, Wc / 2);
  spec.view.fovY = 60;
  spec.hasPoint = true;
  spec.pointLight = V3(L / 2, 7.5f, Wc / 2);
  spec.hasArea = true;
  spec.lightC = V3(L / 2, 14.0f, Wc / 2 - 2.0f);
  spec.lightU = V3(3.0f, 0, 0);
  spec.lightV = V3(0, 0, 3.0f);
}

// ---- conference: the served-stream scene (~125K tris) -----------------------
// A 20 x 12 m hall with 8 long tables and 120 instanced chairs (908
// triangles each) at seeded jittered positions.
void chair(Scene& s, V3 o, float yaw) {
  const float c = std::cos(yaw), sn = std::sin(yaw);
  auto X = [&](V3 p) { return o + V3(c * p.x + sn * p.z, p.y, -sn * p.x + c * p.z); };
  auto slab = [&](V3 lo, V3 hi, int nu, int nv) {
    // a box whose top and front faces are finely tessellated (cushions)
    const V3 d = hi - lo;
    const uint32_t before = uint32_t(s.pos.size());
    gridQuad(s, lo + V3(0, d.y, 0), V3(0, 0, d.z), V3(d.x, 0, 0), nv, nu);  // top (+y)
    gridQuad(s, lo, V3(d.x, 0, 0), V3(0, 0, d.z), 1, 1);                  // bottom
    gridQuad(s, lo, V3(0, d.y, 0), V3(d.x, 0, 0), 2, nu);                 // front (-z)
    gridQuad(s, lo + V3(0, 0, d.z), V3(d.x, 0, 0), V3(0, d.y, 0), nu, 2); // back
    gridQuad(s, lo, V3(0, 0, d.z), V3(0, d.y, 0), nv, 2);                 // left
    gridQuad(s, lo + V3(d.x, 0, 0), V3(0, d.y, 0), V3(0, 0, d.z), 2, nv); // right
    for (size_t i = before; i < s.pos.size(); ++i) s.pos[i] = X(s.pos[i]);
  };
  slab(V3(-0.23f, 0.42f, -0.22f), V3(0.23f, 0.48f, 0.22f), 10, 10);  // seat
  slab(V3(-0.23f, 0.48f, 0.18f), V3(0.23f, 0.95f, 0.23f), 12, 2);    // back rest
  for (int k = 0; k < 4; ++k) {
    const V3 leg((k & 1) ? 0.19f : -0.19f, 0, (k & 2) ? 0.18f : -0.18f);
    const uint32_t before = uint32_t(s.pos.size());
    for (int ring = 0; ring < 4; ++ring)
      cylinder(s, leg + V3(0, 0.105f * float(ring), 0), V3(0, 0.105f, 0), 0.018f, 12, false);
    for (size_t i = before; i < s.pos.size(); ++i) s.pos[i] = X(s.pos[i]);
  }
}

void buildConference(Scene& s, SceneSpec& spec, uint64_t seed) {
  Pcg32 r(seed, 17);
  const float X = 20, Y = 4, Z = 12;
  gridQuad(s, V3(0, 0, 0), V3(0, 0, Z), V3(X, 0, 0), 48, 60);
  gridQuad(s, V3(0, Y, 0), V3(X, 0, 0), V3(0, 0, Z), 60, 48);
  gridQuad(s, V3(0, 0, 0), V3(X, 0, 0), V3(0, Y, 0), 40, 8);
  gridQuad(s, V3(0, 0, Z), V3(0, Y, 0), V3(X, 0, 0), 8, 40);
  gridQuad(s, V3(0, 0, 0), V3(0, Y, 0), V3(0, 0, Z), 8, 24);
  gridQuad(s, V3(X, 0, 0), V3(0, 0, Z), V3(0, Y, 0), 24, 8);
  // 8 long tables in two columns of four rows, 15 chairs each
  for (int t = 0; t < 8; ++t) {
    const float tx = t < 4 ? 2.0f : 11.0f, tz = 1.6f + 2.6f * float(t % 4), tl = 7.0f, tw = 0.9f;
    gridQuad(s, V3(tx, 0.75f, tz), V3(0, 0, tw), V3(tl, 0, 0), 4, 28);  // top
    box(s, V3(tx, 0.70f, tz), V3(tx + tl, 0.75f, tz + tw));
    for (int k = 0; k < 4; ++k) {
      const float lx = tx + ((k & 1) ? tl - 0.1f : 0.05f), lz = tz + ((k & 2) ? tw - 0.1f : 0.05f);
      box(s, V3(lx, 0, lz), V3(lx + 0.05f, 0.70f, lz + 0.05f));
    }
    for (int c = 0; c < 15; ++c) {
      const bool front = c < 8;
      const float cx = tx + 0.5f + float(front ? c : c - 8) * (front ? 0.85f : 0.95f) + float(r.range(-0.12, 0.12));
      const float cz = front ? tz - 0.45f + float(r.range(-0.1, 0.1)) : tz + tw + 0.45f + float(r.range(-0.1, 0.1));
      const float yaw = (front ? 3.14159f : 0.0f) + float(r.range(-0.3, 0.3));
      chair(s, V3(cx, 0, cz), yaw);
    }
  }
  spec.view.eye = V3(0.8f, 1.7f, 0.8f);
  spec.view.target = V3(10.0f, 0.7f, 6.5f);
  spec.view.fovY = 65;
  spec.hasPoint = true;
  spec.pointLight = V3(10.0f, 3.6f, 6.0f);
  spec.hasArea = true;
  spec.lightC = V3(10.0f, 3.95f, 6.0f);
  spec.lightU = V3(2.0f, 0, 0);
  spec.lightV = V3(0, 0, 1.0f);
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
  else if (name == "blob")
    buildBlob(s, spec, seed);
  else if (name == "atrium")
    buildAtrium(s, spec, seed);
  else if (name == "conference")
    buildConference(s, spec, seed);
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
