// Test oracle: scalar CPU restatement of the reference beam tracer (see
// bt_oracle.h). TEST INFRASTRUCTURE ONLY — never linked into the product.
//
// Built with -ffp-contract=off like the reference (CMakeLists.txt:17-21).
// Each function cites the reference code it restates. Four-lane corner values
// are plain float[4] ("Lanes"); the reference's SSE R4 performs the same IEEE
// single operations per lane (src/core/simd4.h:37-68), and its min/max are
// the selects "a < b ? a : b" / "a > b ? a : b" (_mm_min_ps/_mm_max_ps).
#include "bt_oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace {

// ---------------------------------------------------------------- numerics
// Tolerances of the float build (src/core/real.h:36-38, beam_core.h:26-27).
constexpr float kEps = 2e-5f;
constexpr float kKd = 1e-5f;
constexpr float kDeg = 2e-6f;
constexpr float kCopCos = 2e-6f;
constexpr float kCopRel = 2e-5f;
constexpr float kInf = INFINITY;

struct V3 {
  float x, y, z;
  float at(int i) const { return i == 0 ? x : (i == 1 ? y : z); }
  float& at(int i) { return i == 0 ? x : (i == 1 ? y : z); }
};
V3 mk(float x, float y, float z) { return V3{x, y, z}; }
V3 add(V3 a, V3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
V3 sub(V3 a, V3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
V3 neg(V3 a) { return mk(-a.x, -a.y, -a.z); }
V3 scl(V3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
V3 dvs(V3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 crs(V3 a, V3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
float len(V3 a) { return std::sqrt(dot(a, a)); }
V3 nrmz(V3 a) {
  float l = len(a);
  return l > 0 ? dvs(a, l) : a;
}
float absSel(float v) { return v < 0 ? -v : v; }  // the reference's "x < 0 ? -x : x" (keeps -0)
float maxSel(float a, float b) { return a > b ? a : b; }
float minSel(float a, float b) { return a < b ? a : b; }

struct Box {
  V3 lo, hi;
};

// Four corner lanes.
struct Lanes {
  float v[4];
};
Lanes splat(float s) { return Lanes{{s, s, s, s}}; }
// a * x + b * y + c per lane, evaluated ((a*x) + (b*y)) + c.
Lanes affine(const Lanes& x, float a, const Lanes& y, float b, float c) {
  Lanes r;
  for (int i = 0; i < 4; ++i) r.v[i] = x.v[i] * a + y.v[i] * b + c;
  return r;
}
Lanes mul(const Lanes& a, float s) {
  Lanes r;
  for (int i = 0; i < 4; ++i) r.v[i] = a.v[i] * s;
  return r;
}
// Tolerance (|l| + |r|) * k per lane.
Lanes tolOf(const Lanes& l, const Lanes& r, float k) {
  Lanes t;
  for (int i = 0; i < 4; ++i) t.v[i] = (std::fabs(l.v[i]) + std::fabs(r.v[i])) * k;
  return t;
}
bool allLess(const Lanes& a, const Lanes& b, const Lanes& t) {  // lt(a, b - t) == all
  for (int i = 0; i < 4; ++i)
    if (!(a.v[i] < b.v[i] - t.v[i])) return false;
  return true;
}
bool allGreater(const Lanes& a, const Lanes& b, const Lanes& t) {  // gt(a, b + t) == all
  for (int i = 0; i < 4; ++i)
    if (!(a.v[i] > b.v[i] + t.v[i])) return false;
  return true;
}

// -------------------------------------------------------------- polygons
// PolyN<N> (src/beam/beam_geom.h:16-52), capacity 8.
struct Poly {
  float x[8], y[8];
  int n = 0;
  void push(float px, float py) {
    if (n < 8) {
      x[n] = px;
      y[n] = py;
      ++n;
    }
  }
  float area() const {
    float a = 0;
    for (int i = 0; i < n; ++i) {
      int j = i + 1 == n ? 0 : i + 1;
      a += x[i] * y[j] - x[j] * y[i];
    }
    return a * 0.5f;
  }
  void reverse() {
    for (int i = 0, j = n - 1; i < j; ++i, --j) {
      std::swap(x[i], x[j]);
      std::swap(y[i], y[j]);
    }
  }
  float extent() const {
    float lx = kInf, hx = -kInf, ly = kInf, hy = -kInf;
    for (int i = 0; i < n; ++i) {
      lx = x[i] < lx ? x[i] : lx;
      hx = x[i] > hx ? x[i] : hx;
      ly = y[i] < ly ? y[i] : ly;
      hy = y[i] > hy ? y[i] : hy;
    }
    float ex = hx - lx, ey = hy - ly;
    return ex > ey ? ex : ey;
  }
};

struct Line {
  float a = 0, b = 0, c = 0;
  float at(float x, float y) const { return a * x + b * y + c; }
};

// makeLine (beam_geom.h:65-76).
Line lineOf(float a, float b, float c, float ext) {
  Line L;
  float nrm = std::sqrt(a * a + b * b);
  if (!(nrm > 0) || nrm * ext * 1e9f < std::fabs(c)) {
    L.c = c > 0 ? 1.0f : (c < 0 ? -1.0f : 0.0f);
    return L;
  }
  L.a = a / nrm;
  L.b = b / nrm;
  L.c = c / nrm;
  return L;
}

// splitPoly (beam_geom.h:82-114); cap = the template capacity N.
bool cutPoly(const Poly& p, const Line& L, float eps, Poly& in, Poly& out, int cap) {
  in.n = out.n = 0;
  float d[8];
  int cls[8];
  int nin = 0, nout = 0;
  for (int i = 0; i < p.n; ++i) {
    d[i] = L.at(p.x[i], p.y[i]);
    cls[i] = d[i] > eps ? 1 : (d[i] < -eps ? -1 : 0);
  }
  for (int i = 0; i < p.n; ++i) {
    int j = i + 1 == p.n ? 0 : i + 1;
    nin += (cls[i] >= 0) + (cls[i] * cls[j] < 0);
    nout += (cls[i] <= 0) + (cls[i] * cls[j] < 0);
  }
  if (nin > cap || nout > cap) return false;
  for (int i = 0; i < p.n; ++i) {
    int j = i + 1 == p.n ? 0 : i + 1;
    if (cls[i] >= 0) in.push(p.x[i], p.y[i]);
    if (cls[i] <= 0) out.push(p.x[i], p.y[i]);
    if (cls[i] * cls[j] < 0) {
      bool iLow = p.x[i] < p.x[j] || (p.x[i] == p.x[j] && p.y[i] < p.y[j]);
      int a = iLow ? i : j, b = iLow ? j : i;
      float t = d[a] / (d[a] - d[b]);
      float X = p.x[a] + t * (p.x[b] - p.x[a]);
      float Y = p.y[a] + t * (p.y[b] - p.y[a]);
      in.push(X, Y);
      out.push(X, Y);
    }
  }
  return true;
}

// cleanPoly (beam_geom.h:121-150).
void weld(Poly& p, float epsSplit) {
  const float e = epsSplit * 1e-3f;
  bool changed = true;
  while (changed && p.n >= 3) {
    changed = false;
    for (int i = 0; i < p.n && p.n >= 3; ++i) {
      int h = i == 0 ? p.n - 1 : i - 1;
      int j = i + 1 == p.n ? 0 : i + 1;
      float ex = p.x[j] - p.x[h], ey = p.y[j] - p.y[h];
      float dx = p.x[i] - p.x[h], dy = p.y[i] - p.y[h];
      float l = std::sqrt(ex * ex + ey * ey);
      float d2 = dx * dx + dy * dy;
      float cr = std::fabs(ex * dy - ey * dx);
      if (d2 <= e * e || cr <= e * l) {
        for (int k = i; k + 1 < p.n; ++k) {
          p.x[k] = p.x[k + 1];
          p.y[k] = p.y[k + 1];
        }
        --p.n;
        changed = true;
        --i;
      }
    }
  }
  if (p.n < 3) p.n = 0;
}

// classifyPoly (beam_core.h:68-78).
int sides(const Poly& p, const Line& L, float eps) {
  int c = 0;
  for (int i = 0; i < p.n; ++i) {
    float d = L.at(p.x[i], p.y[i]);
    if (d > eps)
      c |= 1;
    else if (d < -eps)
      c |= 2;
  }
  return c;
}

// splitFive (beam_core.h:80-89).
void fiveToQuadTri(const Poly& p, Poly& a, Poly& b) {
  a.n = b.n = 0;
  for (int i = 0; i < 4; ++i) a.push(p.x[i], p.y[i]);
  b.push(p.x[0], p.y[0]);
  b.push(p.x[3], p.y[3]);
  b.push(p.x[4], p.y[4]);
}

// ------------------------------------------------------------- scene data
struct KdNode {
  float split;
  uint32_t data, offset;
  bool leaf() const { return (data & 3u) == 3u; }
  int axis() const { return int(data & 3u); }
  uint32_t hi() const { return data >> 2; }
};

struct Plane {  // BeamPlane (beam_types.h:14-52)
  V3 O, p0, u, v, n, d;
  V3 Olo{0, 0, 0};  // guarded soft mode: the apex is the float pair O + Olo
  // Apex-relative vector / coordinate: (p - O) - Olo, which is p - O when Olo
  // is zero (x - (+0) = x).
  V3 rel(V3 p) const { return sub(sub(p, O), Olo); }
  float rel(float s, int a) const { return (s - O.at(a)) - Olo.at(a); }
  V3 point(float qx, float qy) const { return add(add(p0, scl(u, qx)), scl(v, qy)); }
  V3 dir(float qx, float qy) const { return add(add(d, scl(u, qx)), scl(v, qy)); }
  void coords(V3 p, float& qx, float& qy) const {
    V3 r = sub(p, p0);
    qx = dot(r, u);
    qy = dot(r, v);
  }
  bool project(V3 p, float& qx, float& qy) const {
    V3 r = sub(p, O);
    float den = dot(n, r);
    if (!(den > 0)) return false;
    V3 x = add(O, scl(r, dot(n, d) / den));
    coords(x, qx, qy);
    return true;
  }
  static Plane make(V3 O, V3 p0, V3 u_, V3 v_, V3 Olo = V3{0, 0, 0}) {
    Plane P;
    P.O = O;
    P.Olo = Olo;
    P.p0 = p0;
    const V3 d = P.rel(p0);
    V3 n = nrmz(crs(u_, v_));
    if (dot(n, d) < 0) n = neg(n);
    P.u = nrmz(u_);
    P.v = crs(n, P.u);
    P.n = n;
    P.d = d;
    return P;
  }
};

struct Query {
  Plane plane;
  bool anyHit = false;
  bool farIsPlane = false;
  int excludeTri = -1;
  bool cull = false;
};

struct Out {
  float x[4], y[4];
  int n, tri;
  float area;
};

struct Stats {
  uint64_t kd = 0, leaf = 0, tests = 0, hits = 0, splits = 0, beams = 0, roots = 0, presplit = 0, five = 0, rays = 0,
           climbs = 0;
  double dropped = 0;
  void add(const Stats& o) {
    kd += o.kd;
    leaf += o.leaf;
    tests += o.tests;
    hits += o.hits;
    splits += o.splits;
    beams += o.beams;
    roots += o.roots;
    presplit += o.presplit;
    five += o.five;
    rays += o.rays;
    climbs += o.climbs;
    dropped += o.dropped;
  }
  void put(double* o) const {
    if (!o) return;
    const double v[ORC_NSTATS] = {double(kd),       double(leaf), double(tests), double(hits), double(splits),
                                  double(beams),    double(roots), double(presplit), 0.0,        double(five),
                                  0.0,              double(rays), double(climbs), dropped};
    std::memcpy(o, v, sizeof v);
  }
};

}  // namespace

struct orc_scene {
  std::vector<V3> pos, triN;
  std::vector<uint32_t> tri3, idx, parent;
  std::vector<KdNode> nodes;
  std::vector<Box> nodeBox;
  Box bounds;       // kd root box
  Box sceneBounds;  // Scene::bounds (soft-shadow apex offset)
};

struct orc_polys {
  std::vector<Out> v;
};

namespace {

// ------------------------------------------------------------ beam tracer
struct Beam {  // CoreBeam (beam_core.h:30-40)
  Lanes qx, qy, aw[3], farA;
  float farC;
  int n, hit;
  int sgn[3];
  bool hasFar;
};

struct TriSetup {  // CoreTriInfo (beam_core.h:42-51)
  int tri;
  bool ok;
  Line edge[3];
  float A[3], C;
  V3 r[3], n;
};

struct Work {
  int beam;
  uint32_t node;
  int depth;
  uint64_t trail;
  Box box;
};

int topBit(uint64_t x) { return 63 - __builtin_clzll(x); }

// One query at a time, depth-first (BeamCore<S>::trace with the restart
// trail, beam_core.h:569-661; no mailbox, so processLeaf never skips).
class Tracer {
 public:
  explicit Tracer(const orc_scene& s) : S(s) {}

  Stats st;
  std::vector<Out> out;
  bool keep = true;
  double hitArea = 0, missArea = 0;

  void trace(const Query& q, const Poly& rootIn) {
    out.clear();
    hitArea = missArea = 0;
    pool.clear();
    freeL.clear();
    work.clear();
    cache.clear();
    float ext = maxSel(0.0f, rootIn.extent());
    setup(q, ext);
    Poly p = rootIn;
    if (p.n >= 3) {
      if (p.area() < 0) p.reverse();
      presplit(p);
    }
    while (!work.empty()) {
      Work w = work.back();
      work.pop_back();
      uint32_t node = w.node;
      Box box = w.box;
      int depth = w.depth;
      uint64_t trail = w.trail;
      for (;;) {
        const KdNode& kn = S.nodes[node];
        if (kn.leaf()) break;
        st.kd++;
        const int d = depth++;
        const int a = kn.axis();
        const float s = kn.split;
        const Beam& B = pool[size_t(w.beam)];
        const float Ns = float(B.sgn[a]) * Q.plane.rel(s, a);
        const uint32_t left = node + 1, right = kn.hi();
        Box lb = box, rb = box;
        lb.hi.at(a) = s;
        rb.lo.at(a) = s;
        const bool toRight = B.sgn[a] > 0;
        const uint32_t cs = toRight ? right : left, co = toRight ? left : right;
        const Box& bs = toRight ? rb : lb;
        const Box& bo = toRight ? lb : rb;
        if (Ns <= 0) {
          node = cs;
          box = bs;
          continue;
        }
        int dec = decide(B, a, Ns, box);
        if (dec == kNear) {
          node = co;
          box = bo;
        } else if (dec == kFar) {
          node = cs;
          box = bs;
        } else {
          trail |= uint64_t(1) << d;
          node = co;
          box = bo;
        }
      }
      const KdNode& leaf = S.nodes[node];
      if (leaf.hi() == 0) {
        st.leaf++;
        advance(w.beam, node, depth, trail);
        continue;
      }
      processLeaf(w.beam, leaf);
      const std::vector<int> list = cur;
      for (int b : list) advance(b, node, depth, trail);
    }
  }

 private:
  enum { kNear = 0, kFar = 1, kBoth = 2 };
  const orc_scene& S;
  Query Q;
  float eps = 0, minArea = 0, ext = 1, farCq = 0;
  float farA[3] = {0, 0, 0};
  std::vector<Beam> pool;
  std::vector<int> freeL, cur, nxt;
  std::vector<Work> work;
  std::vector<std::pair<int, TriSetup>> cache;  // tiny memo for hit planes

  // BeamCtx::setup (beam_core.h:123-132).
  void setup(const Query& q, float extent) {
    Q = q;
    ext = extent > 0 ? extent : 1.0f;
    eps = kEps * ext;
    minArea = eps * eps;
    farCq = dot(Q.plane.n, Q.plane.d);
    farA[0] = 0;
    farA[1] = 0;
    farA[2] = farCq;
  }

  int alloc() {
    if (!freeL.empty()) {
      int i = freeL.back();
      freeL.pop_back();
      return i;
    }
    pool.emplace_back();
    return int(pool.size()) - 1;
  }

  // decide (beam_core.h:139-164).
  int decide(const Beam& b, int a, float Ns, const Box& bx) const {
    const Plane& P = Q.plane;
    const Lanes& awa = b.aw[a];
    for (int k = 1; k <= 2; ++k) {
      const int c = a + k >= 3 ? a + k - 3 : a + k;
      const float sg = float(b.sgn[c]);
      const float entry = b.sgn[c] > 0 ? bx.lo.at(c) : bx.hi.at(c);
      const float exit = b.sgn[c] > 0 ? bx.hi.at(c) : bx.lo.at(c);
      const float nmin = sg * P.rel(entry, c), nmax = sg * P.rel(exit, c);
      const Lanes lhs = mul(b.aw[c], Ns);
      Lanes rhs = mul(awa, nmin);
      if (allLess(lhs, rhs, tolOf(lhs, rhs, kKd))) return kFar;
      rhs = mul(awa, nmax);
      if (allGreater(lhs, rhs, tolOf(lhs, rhs, kKd))) return kNear;
    }
    if (b.hasFar) {
      const Lanes lhs = mul(b.farA, Ns), rhs = mul(awa, b.farC);
      if (allGreater(lhs, rhs, tolOf(lhs, rhs, kKd))) return kNear;
    }
    return kBoth;
  }

  // overlaps (beam_core.h:166-195).
  bool overlaps(const Beam& b, const Box& bx) const {
    const Plane& P = Q.plane;
    float nmin[3], nmax[3];
    for (int c = 0; c < 3; ++c) {
      const float sg = float(b.sgn[c]);
      nmin[c] = sg * P.rel(b.sgn[c] > 0 ? bx.lo.at(c) : bx.hi.at(c), c);
      nmax[c] = sg * P.rel(b.sgn[c] > 0 ? bx.hi.at(c) : bx.lo.at(c), c);
      if (nmax[c] < 0) return false;
    }
    for (int c = 0; c < 3; ++c)
      for (int e = 0; e < 3; ++e) {
        if (c == e) continue;
        const Lanes lhs = mul(b.aw[e], nmax[c]), rhs = mul(b.aw[c], nmin[e]);
        if (allLess(lhs, rhs, tolOf(lhs, rhs, kKd))) return false;
      }
    if (b.hasFar)
      for (int c = 0; c < 3; ++c) {
        const Lanes lhs = mul(b.farA, nmin[c]), rhs = mul(b.aw[c], b.farC);
        if (allGreater(lhs, rhs, tolOf(lhs, rhs, kKd))) return false;
      }
    return true;
  }

  // beamGeom (beam_core.h:200-221).
  bool geom(const Poly& p, const int sgn[3], Beam& b, float& dropped) const {
    Poly c = p;
    weld(c, eps);
    float A = c.n >= 3 ? c.area() : 0.0f;
    if (c.n < 3 || A < minArea) {
      dropped = absSel(p.area());
      return false;
    }
    const int l3 = c.n > 3 ? 3 : 2;
    b.qx = Lanes{{c.x[0], c.x[1], c.x[2], c.x[l3]}};
    b.qy = Lanes{{c.y[0], c.y[1], c.y[2], c.y[l3]}};
    b.n = c.n;
    const Plane& P = Q.plane;
    for (int a = 0; a < 3; ++a) {
      b.sgn[a] = sgn[a];
      const Lanes w = affine(b.qx, P.u.at(a), b.qy, P.v.at(a), P.d.at(a));
      for (int i = 0; i < 4; ++i) {
        const float ws = w.v[i] * float(sgn[a]);
        b.aw[a].v[i] = 0.0f > ws ? 0.0f : ws;
      }
    }
    return true;
  }

  // setStatus (beam_core.h:225-240).
  void status(Beam& b, int hit, const float* A, float C) const {
    b.hit = hit;
    if (hit >= 0) {
      const Lanes f = affine(b.qx, A[0], b.qy, A[1], A[2]);
      for (int i = 0; i < 4; ++i) b.farA.v[i] = std::fabs(f.v[i]);
      b.farC = absSel(C);
      b.hasFar = true;
    } else if (Q.farIsPlane) {
      const float c0 = absSel(farCq);
      b.farA = splat(c0);
      b.farC = c0;
      b.hasFar = true;
    } else {
      b.farC = 0;
      b.hasFar = false;
    }
  }

  // computeTriInfo + triSetup (beam_core.h:242-286).
  TriSetup setupTri(int t) const {
    TriSetup ti;
    const uint32_t* T = &S.tri3[3 * size_t(t)];
    for (int k = 0; k < 3; ++k) ti.r[k] = Q.plane.rel(S.pos[T[k]]);
    ti.n = S.triN[size_t(t)];
    ti.tri = t;
    ti.ok = false;
    V3 m[3];
    static const int E[3][2] = {{0, 1}, {1, 2}, {2, 0}};
    for (int k = 0; k < 3; ++k) {
      const int i = E[k][0], j = E[k][1];
      m[k] = T[i] < T[j] ? crs(ti.r[i], ti.r[j]) : neg(crs(ti.r[j], ti.r[i]));
    }
    const float D = dot(ti.r[2], m[0]);
    const float scale = len(m[0]) * len(ti.r[2]);
    if (!(absSel(D) > kDeg * scale)) return ti;
    if (Q.cull && D > 0) return ti;
    const float s = D > 0 ? 1.0f : -1.0f;
    const Plane& P = Q.plane;
    for (int k = 0; k < 3; ++k) ti.edge[k] = lineOf(s * dot(m[k], P.u), s * dot(m[k], P.v), s * dot(m[k], P.d), ext);
    ti.A[0] = dot(ti.n, P.u);
    ti.A[1] = dot(ti.n, P.v);
    ti.A[2] = dot(ti.n, P.d);
    ti.C = dot(ti.n, ti.r[0]);
    ti.ok = true;
    return ti;
  }

  // closerLine (beam_core.h:302-317).
  Line closer(const TriSetup& T, const float* AF, float CF, V3 nF, int fId) const {
    const float cosang = dot(T.n, nF);
    const float aC = absSel(T.C), aF = absSel(CF);
    const float scaleC = maxSel(aC, aF);
    const float ca = absSel(cosang);
    const float diff = T.C - cosang * CF;
    if (ca > 1 - kCopCos && absSel(diff) <= kCopRel * scaleC) {
      Line L;
      L.c = (fId >= 0 && T.tri < fId) ? 1.0f : -1.0f;
      return L;
    }
    const float sg = T.C * CF > 0 ? 1.0f : -1.0f;
    return lineOf(sg * (CF * T.A[0] - T.C * AF[0]), sg * (CF * T.A[1] - T.C * AF[1]),
                  sg * (CF * T.A[2] - T.C * AF[2]), ext);
  }

  // separated (beam_core.h:321-342).
  bool separated(const Beam& b, const TriSetup& ti) const {
    const Plane& P = Q.plane;
    const Lanes wx = affine(b.qx, P.u.x, b.qy, P.v.x, P.d.x);
    const Lanes wy = affine(b.qx, P.u.y, b.qy, P.v.y, P.d.y);
    const Lanes wz = affine(b.qx, P.u.z, b.qy, P.v.z, P.d.z);
    Lanes nx, ny, nz, nl;
    for (int i = 0; i < 4; ++i) {
      const int j = (i + 1) & 3;
      nx.v[i] = wy.v[i] * wz.v[j] - wz.v[i] * wy.v[j];
      ny.v[i] = wz.v[i] * wx.v[j] - wx.v[i] * wz.v[j];
      nz.v[i] = wx.v[i] * wy.v[j] - wy.v[i] * wx.v[j];
      nl.v[i] = std::sqrt(nx.v[i] * nx.v[i] + ny.v[i] * ny.v[i] + nz.v[i] * nz.v[i]);
    }
    int outAll = b.n == 4 ? 0xF : 0xB;
    for (int k = 0; k < 3 && outAll; ++k) {
      const V3 r = ti.r[k];
      const float lr = len(r) * kEps;
      int m = 0;
      for (int i = 0; i < 4; ++i) {
        const float s = nx.v[i] * r.x + ny.v[i] * r.y + nz.v[i] * r.z;
        const float tol = nl.v[i] * lr;
        if (s < -tol) m |= 1 << i;
      }
      outAll &= m;
    }
    return outAll != 0;
  }

  // testLine (beam_core.h:356-362): false if no corner is definitely inside.
  bool testLine(const Beam& B, const Line& line, Line* L, int& nl, bool& allIn) const {
    const Lanes dv = affine(B.qx, line.a, B.qy, line.b, line.c);
    bool anyIn = false, anyOut = false;
    for (int i = 0; i < 4; ++i) {
      anyIn = anyIn || dv.v[i] > eps;
      anyOut = anyOut || dv.v[i] < -eps;
    }
    if (!anyIn) return false;
    if (anyOut) allIn = false;
    L[nl++] = line;
    return true;
  }

  enum { kKeep = 0, kAllIn = 1, kClip = 2 };

  const TriSetup& hitTri(int t) {
    for (auto& e : cache)
      if (e.first == t) return e.second;
    if (cache.size() > 64) cache.clear();
    cache.emplace_back(t, setupTri(t));
    return cache.back().second;
  }

  // classifyTri (beam_core.h:374-409).
  int classify(const Beam& B, const TriSetup& ti, Line* L, int& nl) {
    nl = 0;
    bool allIn = true;
    for (int k = 0; k < 3; ++k)
      if (!testLine(B, ti.edge[k], L, nl, allIn)) return kKeep;
    if (!Q.anyHit && B.hit >= 0) {
      const TriSetup& F = hitTri(B.hit);
      if (!testLine(B, closer(ti, F.A, F.C, S.triN[size_t(B.hit)], B.hit), L, nl, allIn)) return kKeep;
    }
    if (Q.farIsPlane)
      if (!testLine(B, closer(ti, farA, farCq, Q.plane.n, -1), L, nl, allIn)) return kKeep;
    if (allIn) return kAllIn;
    if (separated(B, ti)) return kKeep;
    return kClip;
  }

  // clipTri (beam_core.h:419-479) into two piece lists.
  bool clip(const Beam& B, Line* L, int nl, std::vector<Poly>* lists) {
    float score[3];
    for (int k = 0; k < 3; ++k) {
      const Lanes dv = affine(B.qx, L[k].a, B.qy, L[k].b, L[k].c);
      float p[4];
      for (int i = 0; i < 4; ++i) p[i] = dv.v[i] > 0.0f ? dv.v[i] : 0.0f;
      score[k] = p[0] + p[1] + p[2] + (B.n == 4 ? p[3] : 0.0f);
    }
    for (int i = 1; i < 3; ++i)
      for (int j = i; j > 0 && score[j] < score[j - 1]; --j) {
        std::swap(score[j], score[j - 1]);
        std::swap(L[j], L[j - 1]);
      }
    Poly stack[8];
    int stackLine[8];
    int sp = 1;
    stack[0].n = B.n;
    for (int i = 0; i < B.n; ++i) {
      stack[0].x[i] = B.qx.v[i];
      stack[0].y[i] = B.qy.v[i];
    }
    stackLine[0] = 0;
    while (sp > 0) {
      --sp;
      Poly p = stack[sp];
      int li = stackLine[sp];
      bool inside = true;
      for (; li < nl; ++li) {
        const int c = sides(p, L[li], eps);
        if (!(c & 2)) continue;
        if (!(c & 1)) {
          lists[0].push_back(p);
          inside = false;
          break;
        }
        Poly in, outp;
        if (!cutPoly(p, L[li], eps, in, outp, 8)) return false;
        weld(outp, eps);
        if (outp.n >= 3) lists[0].push_back(outp);
        weld(in, eps);
        if (in.n < 3) {
          inside = false;
          break;
        }
        if (in.n == 5) {
          Poly a, b;
          fiveToQuadTri(in, a, b);
          st.five++;
          if (sp >= 8) return false;
          stack[sp] = b;
          stackLine[sp] = li + 1;
          ++sp;
          p = a;
        } else {
          p = in;
        }
      }
      if (inside) lists[1].push_back(p);
    }
    return true;
  }

  // output (beam_core.h:783-793), with the host storage's OutBeam packing
  // (beam_tracer.cpp:56-75).
  void output(const Beam& b, int tri) {
    Poly p;
    p.n = b.n;
    for (int i = 0; i < b.n; ++i) {
      p.x[i] = b.qx.v[i];
      p.y[i] = b.qy.v[i];
    }
    const float area = p.area();
    if (tri >= 0) st.hits++;
    st.beams++;
    if (tri >= 0)
      hitArea += double(area);
    else
      missArea += double(area);
    if (keep) {
      Out o;
      o.n = p.n;
      for (int i = 0; i < 4; ++i) {
        const int k = i < p.n ? i : p.n - 1;
        o.x[i] = p.x[k];
        o.y[i] = p.y[k];
      }
      o.tri = tri;
      o.area = area;
      out.push_back(o);
    }
  }

  void finalize(int bi) {
    const Beam b = pool[size_t(bi)];
    output(b, Q.anyHit ? -1 : b.hit);
    freeL.push_back(bi);
  }

  // makeBeam (beam_core.h:673-688).
  bool makeBeam(const Poly& p, const int sgn[3], int hit, Beam& b) {
    float dropped = 0;
    if (!geom(p, sgn, b, dropped)) {
      st.dropped += double(dropped);
      return false;
    }
    if (hit >= 0) {
      const TriSetup& ti = hitTri(hit);
      status(b, hit, ti.A, ti.C);
    } else {
      status(b, -1, nullptr, 0);
    }
    return true;
  }

  // emit (beam_core.h:801-826).
  void emit(const Poly& p0, const Beam& parent, int newHit, bool hitNow, int tri) {
    Poly pieces[2];
    int np = 1;
    if (p0.n == 5) {
      fiveToQuadTri(p0, pieces[0], pieces[1]);
      st.five++;
      np = 2;
    } else {
      pieces[0] = p0;
    }
    for (int i = 0; i < np; ++i) {
      Beam nb;
      if (!makeBeam(pieces[i], parent.sgn, hitNow ? newHit : parent.hit, nb)) continue;
      if (hitNow && Q.anyHit) {
        output(nb, tri);
        continue;
      }
      const int idx = alloc();
      pool[size_t(idx)] = nb;
      nxt.push_back(idx);
    }
  }

  // intersect (beam_core.h:832-882).
  void intersect(int bi, int t, const TriSetup& ti) {
    st.tests++;
    if (pool[size_t(bi)].hit == t) {
      nxt.push_back(bi);
      return;
    }
    Line L[5];
    int nl = 0;
    const Beam B = pool[size_t(bi)];
    const int r = classify(B, ti, L, nl);
    if (r == kKeep) {
      nxt.push_back(bi);
      return;
    }
    if (r == kAllIn) {
      if (Q.anyHit) {
        output(B, t);
        freeL.push_back(bi);
        return;
      }
      status(pool[size_t(bi)], t, ti.A, ti.C);
      nxt.push_back(bi);
      return;
    }
    std::vector<Poly> lists[2];
    clip(B, L, nl, lists);
    float newArea = 0;
    for (const Poly& p : lists[1]) newArea += p.area();
    if (!(newArea >= minArea)) {
      nxt.push_back(bi);
      return;
    }
    st.splits++;
    freeL.push_back(bi);
    for (const Poly& p : lists[0]) emit(p, B, -1, false, t);
    for (const Poly& p : lists[1]) emit(p, B, t, true, t);
  }

  // processLeaf without mailboxes (beam_core.h:884-912).
  void processLeaf(int bi, const KdNode& leaf) {
    st.leaf++;
    cur.assign(1, bi);
    const uint32_t cnt = leaf.hi();
    for (uint32_t i = 0; i < cnt; ++i) {
      const int t = int(S.idx[leaf.offset + i]);
      if (t == Q.excludeTri) continue;
      const TriSetup ti = setupTri(t);
      if (!ti.ok) continue;
      nxt.clear();
      const std::vector<int> now = cur;
      for (int b : now) intersect(b, t, ti);
      std::swap(cur, nxt);
      if (cur.empty()) break;
    }
  }

  // advance with the restart trail (beam_core.h:924-969).
  void advance(int bi, uint32_t node, int depth, uint64_t trail) {
    const Beam b = pool[size_t(bi)];
    while (trail != 0) {
      const int d = topBit(trail);
      for (; depth > d; --depth) {
        node = S.parent[node];
        st.climbs++;
      }
      trail &= (uint64_t(1) << d) - 1;
      const KdNode& kn = S.nodes[node];
      const uint32_t cs = b.sgn[kn.axis()] > 0 ? kn.hi() : node + 1;
      const Box& bx = S.nodeBox[cs];
      if (overlaps(b, bx)) {
        work.push_back(Work{bi, cs, d + 1, trail, bx});
        return;
      }
    }
    finalize(bi);
  }

  // presplit (beam_core.h:487-546) + addRootPiece (beam_core.h:977-1003).
  void presplit(const Poly& root) {
    struct Item {
      Poly p;
      int axis;
      int sgn[3];
    };
    std::vector<Item> stack;
    stack.push_back(Item{root, 0, {1, 1, 1}});
    const Plane& P = Q.plane;
    while (!stack.empty()) {
      Item it = stack.back();
      stack.pop_back();
      const Poly& p = it.p;
      if (p.n < 3) continue;
      if (it.axis == 3) {
        for (int k = 1; k + 1 < p.n; k += 2) {
          Poly qd;
          qd.push(p.x[0], p.y[0]);
          qd.push(p.x[k], p.y[k]);
          qd.push(p.x[k + 1], p.y[k + 1]);
          if (k + 2 < p.n) qd.push(p.x[k + 2], p.y[k + 2]);
          rootPiece(qd, it.sgn);
        }
        continue;
      }
      const int axis = it.axis;
      const Line L = lineOf(P.u.at(axis), P.v.at(axis), P.d.at(axis), ext);
      const int c = sides(p, L, eps);
      if (c == 3) {
        Poly in, outp;
        cutPoly(p, L, eps, in, outp, 8);
        weld(in, eps);
        weld(outp, eps);
        Item a = it, b = it;
        a.p = in;
        a.axis = axis + 1;
        a.sgn[axis] = 1;
        b.p = outp;
        b.axis = axis + 1;
        b.sgn[axis] = -1;
        stack.push_back(b);
        stack.push_back(a);
        st.presplit++;
      } else if (c != 0) {
        it.axis = axis + 1;
        it.sgn[axis] = (c & 1) ? 1 : -1;
        stack.push_back(it);
      } else {
        st.dropped += double(absSel(p.area()));
      }
    }
  }

  void rootPiece(const Poly& p, const int sgn[3]) {
    Beam b;
    if (!makeBeam(p, sgn, -1, b)) return;
    const int idx = alloc();
    pool[size_t(idx)] = b;
    st.roots++;
    if (overlaps(b, S.bounds))
      work.push_back(Work{idx, 0, 0, 0, S.bounds});
    else
      finalize(idx);
  }
};

// ---------------------------------------------------------------- rays
// rayHitTri (src/ray/ray_core.h:22-37).
float rayTri(const orc_scene& S, V3 o, V3 d, int tri) {
  const uint32_t* T = &S.tri3[3 * size_t(tri)];
  const V3 v0 = S.pos[T[0]];
  const V3 e1 = sub(S.pos[T[1]], v0), e2 = sub(S.pos[T[2]], v0);
  const V3 pv = crs(d, e2);
  const float det = dot(e1, pv);
  if (det == 0) return kInf;
  const float inv = 1 / det;
  const V3 tv = sub(o, v0);
  const float u = dot(tv, pv) * inv;
  if (u < 0 || u > 1) return kInf;
  const V3 qv = crs(tv, e1);
  const float v = dot(d, qv) * inv;
  if (v < 0 || u + v > 1) return kInf;
  return dot(e2, qv) * inv;
}

// rayTraverse<false> (ray_core.h:42-132), tmin 0, tmax inf, no culling.
bool rayNearest(const orc_scene& S, V3 o, V3 d, int& hitTri, float& hitT, Stats& st) {
  st.rays++;
  float t0 = 0, t1 = kInf;
  float inv[3];
  for (int a = 0; a < 3; ++a) {
    inv[a] = 1 / d.at(a);
    if (d.at(a) == 0) {
      if (o.at(a) < S.bounds.lo.at(a) || o.at(a) > S.bounds.hi.at(a)) return false;
      continue;
    }
    float ta = (S.bounds.lo.at(a) - o.at(a)) * inv[a];
    float tb = (S.bounds.hi.at(a) - o.at(a)) * inv[a];
    if (ta > tb) std::swap(ta, tb);
    t0 = maxSel(t0, ta);
    t1 = minSel(t1, tb);
  }
  if (t0 > t1) return false;
  struct Entry {
    uint32_t node;
    float t0, t1;
  };
  Entry stack[64];
  int sp = 0;
  uint32_t node = 0;
  float best = kInf;
  int bestTri = -1;
  for (;;) {
    const KdNode& n = S.nodes[node];
    if (!n.leaf()) {
      st.kd++;
      const int a = n.axis();
      const float oa = o.at(a), da = d.at(a);
      const bool belowFirst = oa < n.split || (oa == n.split && da <= 0);
      const uint32_t first = belowFirst ? node + 1 : n.hi();
      const uint32_t second = belowFirst ? n.hi() : node + 1;
      if (da == 0) {
        node = first;
        continue;
      }
      const float tp = (n.split - oa) * inv[a];
      if (tp > t1 || tp <= 0) {
        node = first;
      } else if (tp < t0) {
        node = second;
      } else {
        if (sp < 64) stack[sp++] = Entry{second, tp, t1};
        node = first;
        t1 = tp;
      }
      continue;
    }
    st.leaf++;
    for (uint32_t i = 0; i < n.hi(); ++i) {
      const int t = int(S.idx[n.offset + i]);
      st.tests++;
      const float th = rayTri(S, o, d, t);
      if (!(th > 0 && th < kInf)) continue;
      const float tol = bestTri < 0 ? 0.0f : best * (4 * 1.1920928955078125e-7f);
      if (bestTri < 0 || th < best - tol || (th <= best + tol && t < bestTri)) {
        best = th;
        bestTri = t;
      }
    }
    if (bestTri >= 0 && best <= t1) break;
    if (sp == 0) break;
    --sp;
    node = stack[sp].node;
    t0 = stack[sp].t0;
    t1 = stack[sp].t1;
  }
  if (bestTri < 0) return false;
  st.hits++;
  hitTri = bestTri;
  hitT = best;
  return true;
}

// ------------------------------------------------------------ applications
struct Cam {
  V3 eye, fwd, right, down;
  float halfW, halfH;
  int W, H;
  Plane plane() const { return Plane::make(eye, add(eye, fwd), right, down); }
  V3 dir(float qx, float qy) const { return add(add(fwd, scl(right, qx)), scl(down, qy)); }
  void pixelToQ(float px, float py, float& qx, float& qy) const {
    qx = (2 * px / float(W) - 1) * halfW;
    qy = (2 * py / float(H) - 1) * halfH;
  }
};

Cam camOf(const float* c, int W, int H) {
  return Cam{mk(c[0], c[1], c[2]), mk(c[3], c[4], c[5]), mk(c[6], c[7], c[8]), mk(c[9], c[10], c[11]), c[12], c[13], W, H};
}

struct Light {
  V3 c, U, V;
  V3 n() const { return nrmz(crs(U, V)); }
  float area() const { return len(crs(U, V)); }
};

Light lightOf(const float* l) { return Light{mk(l[0], l[1], l[2]), mk(l[3], l[4], l[5]), mk(l[6], l[7], l[8])}; }

// polygonFormFactorG (src/render/soft_core.h:18-33).
float formFactor(V3 x, V3 n, const V3* poly, int count) {
  double sum = 0;
  for (int i = 0; i < count; ++i) {
    const V3 a = sub(poly[i], x), b = sub(poly[i + 1 == count ? 0 : i + 1], x);
    const double la = double(len(a)), lb = double(len(b));
    if (la == 0 || lb == 0) continue;
    double c = double(dot(a, b)) / (la * lb);
    c = c < -1 ? -1 : (c > 1 ? 1 : c);
    const double theta = std::acos(c);
    const V3 g = crs(a, b);
    const double lg = double(len(g));
    if (lg == 0) continue;
    sum += theta * double(dot(g, n)) / lg;
  }
  return float((sum < 0 ? -sum : sum) * 0.5);
}

// polygonFormFactorG (soft_core.h:18-33) in double throughout: the guarded
// mode's Lambert integral (x: the double shading point; corners formed in
// double on the light plane from the float piece corners).
double formFactorD(const double* x, V3 n, const double (*P)[3], int count) {
  double sum = 0;
  for (int i = 0; i < count; ++i) {
    const double* pa = P[i];
    const double* pb = P[i + 1 == count ? 0 : i + 1];
    const double a[3] = {pa[0] - x[0], pa[1] - x[1], pa[2] - x[2]}, b[3] = {pb[0] - x[0], pb[1] - x[1], pb[2] - x[2]};
    const double la = std::sqrt((a[0] * a[0] + a[1] * a[1]) + a[2] * a[2]);
    const double lb = std::sqrt((b[0] * b[0] + b[1] * b[1]) + b[2] * b[2]);
    if (la == 0 || lb == 0) continue;
    double c = ((a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]) / (la * lb);
    c = c < -1 ? -1 : (c > 1 ? 1 : c);
    const double theta = std::acos(c);
    const double g[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    const double lg = std::sqrt((g[0] * g[0] + g[1] * g[1]) + g[2] * g[2]);
    if (lg == 0) continue;
    sum += theta * ((g[0] * double(n.x) + g[1] * double(n.y)) + g[2] * double(n.z)) / lg;
  }
  return (sum < 0 ? -sum : sum) * 0.5;
}

// lightBeamSetup (soft_core.h:38-73).
// guard: the engine's FP64-guarded mode (not in the reference): x + xlo is the
// hit point in double, the apex x + ns * offset is formed in double and kept as
// the float pair O + Olo, the horizon line passes through the double point.
bool lightSetup(const Light& L, V3 x, V3 xlo, bool guard, V3 ns, int tri, float offset, Query& q, Poly& root,
                float& lightArea) {
  const V3 nL = L.n();
  if (!(dot(sub(x, L.c), nL) > 0)) return false;
  float hzc;
  if (guard) {
    const double xd = double(x.x) + double(xlo.x), yd = double(x.y) + double(xlo.y), zd = double(x.z) + double(xlo.z);
    const double off = double(offset);
    const double ax = xd + double(ns.x) * off, ay = yd + double(ns.y) * off, az = zd + double(ns.z) * off;
    const V3 O = mk(float(ax), float(ay), float(az));
    const V3 Olo = mk(float(ax - double(O.x)), float(ay - double(O.y)), float(az - double(O.z)));
    q.plane = Plane::make(O, L.c, L.U, L.V, Olo);
    hzc = float((double(ns.x) * (double(L.c.x) - xd) + double(ns.y) * (double(L.c.y) - yd)) +
                double(ns.z) * (double(L.c.z) - zd));
  } else {
    const V3 apex = add(x, scl(ns, offset));
    q.plane = Plane::make(apex, L.c, L.U, L.V);
    hzc = dot(ns, sub(L.c, x));
  }
  q.anyHit = true;
  q.farIsPlane = true;
  q.excludeTri = tri;
  q.cull = false;
  const V3 hu = scl(L.U, 0.5f), hv = scl(L.V, 0.5f);
  const V3 corners[4] = {sub(sub(L.c, hu), hv), sub(add(L.c, hu), hv), add(add(L.c, hu), hv), add(sub(L.c, hu), hv)};
  Poly quad;
  for (int i = 0; i < 4; ++i) {
    float qx, qy;
    q.plane.coords(corners[i], qx, qy);
    quad.push(qx, qy);
  }
  if (quad.area() < 0) quad.reverse();
  lightArea = quad.area();
  Line hz;
  hz.a = dot(ns, q.plane.u);
  hz.b = dot(ns, q.plane.v);
  hz.c = hzc;
  int above = 0;
  for (int i = 0; i < quad.n; ++i) above += hz.at(quad.x[i], quad.y[i]) > 0;
  if (above == 0) return false;
  if (above == quad.n) {
    root = quad;
  } else {
    Poly rest;
    cutPoly(quad, hz, 0, root, rest, 8);
  }
  return root.n >= 3 && root.area() > 0;
}

// lightCenterG (soft_core.h:76-84).
float centerG(const Light& L, V3 x, V3 ns) {
  const V3 nL = L.n();
  if (!(dot(sub(x, L.c), nL) > 0)) return 0;
  const V3 w = sub(L.c, x);
  const float r2 = dot(w, w);
  const V3 wn = dvs(w, std::sqrt(r2));
  const float ci = dot(ns, wn), cl = -dot(nL, wn);
  return (ci > 0 ? ci : 0) * (cl > 0 ? cl : 0) / r2;
}

// beamLightVisibility (render.cpp:279-303).
float lightVis(Tracer& tr, const Light& L, V3 x, V3 xlo, bool guard, V3 ns, int tri, float offset, float* G) {
  if (G) *G = 0;
  Query q;
  Poly root;
  float lightArea;
  if (!lightSetup(L, x, xlo, guard, ns, tri, offset, q, root, lightArea)) return 0;
  tr.keep = G != nullptr;
  tr.trace(q, root);
  const float V = float(tr.missArea) / lightArea;
  if (G && guard) {
    const double xd[3] = {double(x.x) + double(xlo.x), double(x.y) + double(xlo.y), double(x.z) + double(xlo.z)};
    const Plane& pl = q.plane;
    double g = 0;
    for (const Out& ob : tr.out) {
      if (ob.tri >= 0) continue;
      double P[4][3];
      for (int i = 0; i < ob.n; ++i) {
        const double qx = ob.x[i], qy = ob.y[i];
        P[i][0] = (double(pl.p0.x) + double(pl.u.x) * qx) + double(pl.v.x) * qy;
        P[i][1] = (double(pl.p0.y) + double(pl.u.y) * qx) + double(pl.v.y) * qy;
        P[i][2] = (double(pl.p0.z) + double(pl.u.z) * qx) + double(pl.v.z) * qy;
      }
      g += formFactorD(xd, ns, P, ob.n);
    }
    *G = float(g);
  } else if (G) {
    float g = 0;
    for (const Out& ob : tr.out) {
      if (ob.tri >= 0) continue;
      V3 P[4];
      for (int i = 0; i < ob.n; ++i) P[i] = q.plane.point(ob.x[i], ob.y[i]);
      g += formFactor(x, ns, P, ob.n);
    }
    *G = g;
  }
  return std::max(0.0f, std::min(1.0f, V));
}

// The guarded mode's receiver point: the ray's hit on the plane of triangle
// t's float corners, in double, as the float pair X + Xlo (n = (v1 - v0) x
// (v2 - v0), th = n.(v0 - o) / n.d, X = o + d th).
void hitPointD(const orc_scene& S, int t, V3 o, V3 d, V3& X, V3& Xlo) {
  const uint32_t* T = &S.tri3[3 * size_t(t)];
  const V3 v0 = S.pos[T[0]], v1 = S.pos[T[1]], v2 = S.pos[T[2]];
  const double v0x = v0.x, v0y = v0.y, v0z = v0.z;
  const double ax = double(v1.x) - v0x, ay = double(v1.y) - v0y, az = double(v1.z) - v0z;
  const double bx = double(v2.x) - v0x, by = double(v2.y) - v0y, bz = double(v2.z) - v0z;
  const double nx = ay * bz - az * by, ny = az * bx - ax * bz, nz = ax * by - ay * bx;
  const double den = (nx * double(d.x) + ny * double(d.y)) + nz * double(d.z);
  const double num = (nx * (v0x - double(o.x)) + ny * (v0y - double(o.y))) + nz * (v0z - double(o.z));
  const double th = den != 0 ? num / den : 0.0;
  const double px = double(o.x) + double(d.x) * th, py = double(o.y) + double(d.y) * th,
               pz = double(o.z) + double(d.z) * th;
  X = mk(float(px), float(py), float(pz));
  Xlo = mk(float(px - double(X.x)), float(py - double(X.y)), float(pz - double(X.z)));
}

float softOffset(const orc_scene& S, bool guard) {
  const V3 d = sub(S.sceneBounds.hi, S.sceneBounds.lo);
  return len(d) * (guard ? 1e-9f : kEps);  // the double build's kEpsRel when guarded (real.h:31)
}

// pointShadowSetup (point_core.h:15-36).
bool pointSetup(const Cam& c, V3 nT, V3 P0, int T, V3 light, const float* px, const float* py, int n, Query& q,
                Poly& root) {
  const float sideCam = dot(nT, sub(c.eye, P0)), sideL = dot(nT, sub(light, P0));
  if (!(sideCam * sideL > 0)) return false;
  const V3 a = std::fabs(nT.x) < 0.6f ? mk(1, 0, 0) : (std::fabs(nT.y) < 0.6f ? mk(0, 1, 0) : mk(0, 0, 1));
  const V3 tu = nrmz(crs(nT, a));
  q.plane = Plane::make(light, P0, tu, crs(nT, tu));
  q.anyHit = true;
  q.farIsPlane = true;
  q.excludeTri = T;
  q.cull = false;
  root.n = 0;
  for (int i = 0; i < n; ++i) {
    const V3 d = add(add(c.fwd, scl(c.right, px[i])), scl(c.down, py[i]));
    const V3 X = add(c.eye, scl(d, dot(nT, sub(P0, c.eye)) / dot(nT, d)));
    float qx, qy;
    q.plane.coords(X, qx, qy);
    root.push(qx, qy);
  }
  return true;
}

// pointShadowToImage (point_core.h:41-64).
bool pointToImage(const Plane& camPlane, const Plane& sp, const float* sx, const float* sy, int n, int T, Out& ip) {
  ip.n = n;
  ip.tri = T;
  ip.area = 0;
  for (int i = 0; i < n; ++i)
    if (!camPlane.project(sp.point(sx[i], sy[i]), ip.x[i], ip.y[i])) return false;
  Poly p;
  for (int i = 0; i < n; ++i) p.push(ip.x[i], ip.y[i]);
  const float a = p.area();
  if (!(a > 0 || a < 0)) return false;
  if (a < 0) {
    p.reverse();
    for (int i = 0; i < n; ++i) {
      ip.x[i] = p.x[i];
      ip.y[i] = p.y[i];
    }
  }
  for (int i = n; i < 4; ++i) {
    ip.x[i] = ip.x[n - 1];
    ip.y[i] = ip.y[n - 1];
  }
  return true;
}

}  // namespace

extern "C" {

orc_scene* orc_scene_new(const float* pos, int64_t nv, const uint32_t* tri3, const float* triN, int64_t nt,
                         const void* nodes12, int64_t nn, const uint32_t* triIndices, int64_t nr,
                         const uint32_t* parent, const float* nodeBox6, const float* kdBounds6,
                         const float* sceneBounds6) {
  orc_scene* s = new orc_scene();
  s->pos.resize(size_t(nv));
  for (int64_t i = 0; i < nv; ++i) s->pos[size_t(i)] = mk(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
  s->tri3.assign(tri3, tri3 + 3 * nt);
  s->triN.resize(size_t(nt));
  for (int64_t i = 0; i < nt; ++i) s->triN[size_t(i)] = mk(triN[3 * i], triN[3 * i + 1], triN[3 * i + 2]);
  s->nodes.resize(size_t(nn));
  std::memcpy(s->nodes.data(), nodes12, size_t(nn) * 12);
  s->idx.assign(triIndices, triIndices + nr);
  s->parent.assign(parent, parent + nn);
  s->nodeBox.resize(size_t(nn));
  for (int64_t i = 0; i < nn; ++i) {
    const float* b = nodeBox6 + 6 * i;
    s->nodeBox[size_t(i)] = Box{mk(b[0], b[1], b[2]), mk(b[3], b[4], b[5])};
  }
  s->bounds = Box{mk(kdBounds6[0], kdBounds6[1], kdBounds6[2]), mk(kdBounds6[3], kdBounds6[4], kdBounds6[5])};
  s->sceneBounds = Box{mk(sceneBounds6[0], sceneBounds6[1], sceneBounds6[2]),
                       mk(sceneBounds6[3], sceneBounds6[4], sceneBounds6[5])};
  return s;
}

void orc_scene_free(orc_scene* s) { delete s; }

void orc_primary_rays(const orc_scene* s, const float* cam14, int W, int H, int32_t* tri, float* X, float* N,
                      double* stats) {
  const Cam c = camOf(cam14, W, H);
  Stats st;
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x) {
      const size_t k = size_t(y) * size_t(W) + size_t(x);
      float qx, qy;
      c.pixelToQ(float(x) + 0.5f, float(y) + 0.5f, qx, qy);
      const V3 d = c.dir(qx, qy);
      int t;
      float th;
      tri[k] = -1;
      if (!rayNearest(*s, c.eye, d, t, th, st)) continue;
      tri[k] = t;
      const V3 p = add(c.eye, scl(d, th));
      V3 n = s->triN[size_t(t)];
      if (dot(n, d) > 0) n = neg(n);
      X[3 * k] = p.x, X[3 * k + 1] = p.y, X[3 * k + 2] = p.z;
      N[3 * k] = n.x, N[3 * k + 1] = n.y, N[3 * k + 2] = n.z;
    }
  st.put(stats);
}

void orc_soft(const orc_scene* s, const float* cam14, int W, int H, const float* light9, int exact, int threads,
              int y0, int y1, int32_t* tri, float* vis, float* E, float* Eu, double* stats, int guard) {
  const Cam c = camOf(cam14, W, H);
  const Light L = lightOf(light9);
  const float offset = softOffset(*s, guard != 0);
  const float A = L.area();
  if (y1 <= 0) {
    y0 = 0;
    y1 = H;
  }
  threads = std::max(1, threads);
  std::atomic<int> next(y0);
  std::vector<Stats> st(static_cast<size_t>(threads));
  auto worker = [&](int id) {
    Tracer tr(*s);
    Stats rays;  // the primary pass's counters are not reported (render.cpp:382-384)
    for (int y = next++; y < y1; y = next++)
      for (int x = 0; x < W; ++x) {
        const size_t k = size_t(y) * size_t(W) + size_t(x);
        tri[k] = -1;
        vis[k] = -1.f;
        E[k] = 0.f;
        Eu[k] = 0.f;
        float qx, qy;
        c.pixelToQ(float(x) + 0.5f, float(y) + 0.5f, qx, qy);
        const V3 d = c.dir(qx, qy);
        int t;
        float th;
        if (!rayNearest(*s, c.eye, d, t, th, rays)) continue;
        tri[k] = t;
        V3 p = add(c.eye, scl(d, th)), plo = mk(0, 0, 0);
        if (guard) hitPointD(*s, t, c.eye, d, p, plo);
        V3 n = s->triN[size_t(t)];
        if (dot(n, d) > 0) n = neg(n);
        const float G0 = centerG(L, p, n);
        float Gx = 0;
        const float V = lightVis(tr, L, p, plo, guard != 0, n, t, offset, exact ? &Gx : nullptr);
        vis[k] = V;
        Eu[k] = A * G0;
        E[k] = exact ? Gx : A * G0 * V;
      }
    st[size_t(id)].add(tr.st);
  };
  std::vector<std::thread> th;
  for (int i = 1; i < threads; ++i) th.emplace_back(worker, i);
  worker(0);
  for (auto& t : th) t.join();
  Stats tot;
  for (auto& x : st) tot.add(x);
  tot.put(stats);
}

float orc_light_vis(const orc_scene* s, const float* x3, const float* ns3, int tri, const float* light9, float offset,
                    float* G, double* stats) {
  Tracer tr(*s);
  const float V = lightVis(tr, lightOf(light9), mk(x3[0], x3[1], x3[2]), mk(0, 0, 0), false, mk(ns3[0], ns3[1], ns3[2]),
                           tri, offset, G);
  tr.st.put(stats);
  return V;
}

orc_polys* orc_primary(const orc_scene* s, const float* cam14, int W, int H, int cull, int tile, double* stats) {
  return orc_primary_tiles(s, cam14, W, H, cull, tile, 0, tile > 0 ? (H + tile - 1) / tile : 0, stats);
}

// Tiled primary visibility restricted to the tile rows [ty0, ty1) (a strip of
// a large frame; every tile root is traced alone, so a strip's polygons equal
// the whole frame's polygons of those tiles).
orc_polys* orc_primary_tiles(const orc_scene* s, const float* cam14, int W, int H, int cull, int tile, int ty0,
                             int ty1, double* stats) {
  const Cam c = camOf(cam14, W, H);
  Tracer tr(*s);
  Query q;
  q.plane = c.plane();
  q.anyHit = false;
  q.cull = cull != 0;
  orc_polys* P = new orc_polys();
  if (tile <= 0) {
    Poly root;
    root.push(-c.halfW, -c.halfH);
    root.push(c.halfW, -c.halfH);
    root.push(c.halfW, c.halfH);
    root.push(-c.halfW, c.halfH);
    tr.trace(q, root);
    P->v = tr.out;
  } else {
    const int tilesX = (W + tile - 1) / tile, tilesY = (H + tile - 1) / tile;
    ty0 = std::max(0, ty0);
    ty1 = std::min(tilesY, ty1);
    for (int i = ty0 * tilesX; i < ty1 * tilesX; ++i) {
      const int tx = i % tilesX, ty = i / tilesX;
      const int x0 = tx * tile, y0 = ty * tile, x1 = std::min(W, x0 + tile), y1 = std::min(H, y0 + tile);
      float a, b, cc, d;
      c.pixelToQ(float(x0), float(y0), a, b);
      c.pixelToQ(float(x1), float(y1), cc, d);
      Poly root;
      root.push(a, b);
      root.push(cc, b);
      root.push(cc, d);
      root.push(a, d);
      tr.trace(q, root);
      for (const Out& o : tr.out)
        if (o.tri >= 0) P->v.push_back(o);
    }
  }
  tr.st.put(stats);
  return P;
}

orc_polys* orc_point(const orc_scene* s, const float* cam14, int W, int H, const float* px4, const float* py4,
                     const int32_t* pn, const int32_t* ptri, const float* parea, int64_t nPrim, const float* light3,
                     double* stats) {
  const Cam c = camOf(cam14, W, H);
  const Plane camPlane = c.plane();
  const V3 light = mk(light3[0], light3[1], light3[2]);
  Tracer tr(*s);
  orc_polys* P = new orc_polys();
  for (int64_t i = 0; i < nPrim; ++i) {
    const int T = ptri[i];
    if (T < 0) continue;
    Out pb;
    for (int k = 0; k < 4; ++k) {
      pb.x[k] = px4[4 * i + k];
      pb.y[k] = py4[4 * i + k];
    }
    pb.n = pn[i];
    pb.tri = T;
    pb.area = parea[i];
    Query q;
    Poly root;
    const V3 P0 = s->pos[s->tri3[3 * size_t(T)]];
    if (!pointSetup(c, s->triN[size_t(T)], P0, T, light, pb.x, pb.y, pb.n, q, root)) {
      P->v.push_back(pb);
      continue;
    }
    tr.trace(q, root);
    for (const Out& sb : tr.out) {
      if (sb.tri < 0) continue;
      Out ip;
      if (pointToImage(camPlane, q.plane, sb.x, sb.y, sb.n, T, ip)) P->v.push_back(ip);
    }
  }
  tr.st.put(stats);
  return P;
}

int64_t orc_polys_count(const orc_polys* p) { return int64_t(p->v.size()); }

void orc_polys_get(const orc_polys* p, float* x4, float* y4, int32_t* n, int32_t* tri, float* area) {
  for (size_t i = 0; i < p->v.size(); ++i) {
    const Out& o = p->v[i];
    for (int k = 0; k < 4; ++k) {
      x4[4 * i + k] = o.x[k];
      y4[4 * i + k] = o.y[k];
    }
    n[i] = o.n;
    tri[i] = o.tri;
    area[i] = o.area;
  }
}

void orc_polys_free(orc_polys* p) { delete p; }

}  // extern "C"
