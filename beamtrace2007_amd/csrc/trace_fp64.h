// FP64 soft shadows (btg_options.precision = 2), included by engine.hip.
//
// The reference's double build (Real = double, src/core/real.h:20-35) on the
// device: one pixel per thread, the depth-first BeamCore with the restart trail
// and without mailboxes (src/beam/beam_core.h:569-1003), its geometry
// (src/beam/beam_geom.h) and soft-shadow setup (src/render/soft_core.h) in
// double with the double build's tolerances. The scene data are the float
// scene's (vertices, kd-tree: exact in double); the receiver point is the
// guarded primary pass's double hit point and the normal is the triangle's,
// computed in double as Scene::finalize does (src/scene/scene.cpp:18-33).
// The tree is the float build's (the double build makes its own): kd
// decisions are conservative, so the visible fraction is the same up to the
// slivers the other split order drops and the double rounding of the piece
// areas, below 1e-10 of the light.
//
// Every beam of a pixel lives in the thread's region of a device buffer
// (WordMem; bounded pools: a pixel that needs more is traced again with
// larger pools; results do not depend on the pool sizes). For an any-hit query with the light plane as the
// far plane, a beam carries only its corners and ray-direction signs: the far
// plane is the same for all of them and |w| is recomputed from the corners.
// FP64_HD marks every function: device code in the library; a host build of
// the same code (a test harness) defines it as __host__ __device__.
#ifndef FP64_HD
#define FP64_HD __device__
#endif

// polygonFormFactorG as the double build computes it (Real = double): the
// FP64 modes' Lambert integral, every step in double from the double shading
// point and the piece's corners formed in double on the light plane.
struct D3 {
  double x, y, z;
};
FP64_HD inline D3 d3sub(D3 a, D3 b) { return D3{a.x - b.x, a.y - b.y, a.z - b.z}; }
FP64_HD inline double d3dot(D3 a, D3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
FP64_HD inline D3 d3cross(D3 a, D3 b) {
  return D3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
FP64_HD inline double formFactorD(D3 x, D3 n, const D3 (&poly)[4], int count) {
  double sum = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    if (i >= count) break;
    const D3 a = d3sub(poly[i], x), b = d3sub(i + 1 < count ? poly[(i + 1) & 3] : poly[0], x);
    const double la = sqrt(d3dot(a, a)), lb = sqrt(d3dot(b, b));
    if (la == 0 || lb == 0) continue;
    double cc = d3dot(a, b) / (la * lb);
    cc = cc < -1 ? -1 : (cc > 1 ? 1 : cc);
    const double theta = acos(cc);
    const D3 g = d3cross(a, b);
    const double lg = sqrt(d3dot(g, g));
    if (lg == 0) continue;
    sum += theta * d3dot(g, n) / lg;
  }
  return (sum < 0 ? -sum : sum) * 0.5;
}

namespace fp64 {

constexpr double kEps = 1e-9;      // fuzzy split epsilon, relative to the root's extent
constexpr double kKd = 1e-12;      // kd decision margin
constexpr double kDeg = 1e-10;     // edge-on triangles
constexpr double kCopCos = 1e-9;   // coplanar: cosine
constexpr double kCopRel = 1e-8;   // coplanar: offset

struct V {
  double x, y, z;
};
FP64_HD inline V vmk(double x, double y, double z) { return V{x, y, z}; }
FP64_HD inline V vf(F3 a) { return V{double(a.x), double(a.y), double(a.z)}; }
FP64_HD inline V operator+(V a, V b) { return V{a.x + b.x, a.y + b.y, a.z + b.z}; }
FP64_HD inline V operator-(V a, V b) { return V{a.x - b.x, a.y - b.y, a.z - b.z}; }
FP64_HD inline V operator*(V a, double s) { return V{a.x * s, a.y * s, a.z * s}; }
FP64_HD inline V neg(V a) { return V{-a.x, -a.y, -a.z}; }
FP64_HD inline double dot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
FP64_HD inline V crs(V a, V b) {
  return V{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
FP64_HD inline double len(V a) { return sqrt(dot(a, a)); }
FP64_HD inline V nrmz(V a) {
  const double l = len(a);
  return l > 0 ? V{a.x / l, a.y / l, a.z / l} : a;
}
FP64_HD inline double at(V a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
FP64_HD inline double absS(double v) { return v < 0 ? -v : v; }

// Four corner lanes (R4, src/core/simd4.h).
struct L4 {
  double v[4];
};
FP64_HD inline L4 affine(const L4& x, double a, const L4& y, double b, double c) {
  L4 r;
#pragma unroll
  for (int i = 0; i < 4; ++i) r.v[i] = x.v[i] * a + y.v[i] * b + c;
  return r;
}
FP64_HD inline L4 mul(const L4& a, double s) {
  L4 r;
#pragma unroll
  for (int i = 0; i < 4; ++i) r.v[i] = a.v[i] * s;
  return r;
}
// lhs < rhs - (|lhs| + |rhs|) k on every lane; the same with > and +.
FP64_HD inline bool allLess(const L4& a, const L4& b, double k) {
#pragma unroll
  for (int i = 0; i < 4; ++i)
    if (!(a.v[i] < b.v[i] - (fabs(a.v[i]) + fabs(b.v[i])) * k)) return false;
  return true;
}
FP64_HD inline bool allGreater(const L4& a, const L4& b, double k) {
#pragma unroll
  for (int i = 0; i < 4; ++i)
    if (!(a.v[i] > b.v[i] + (fabs(a.v[i]) + fabs(b.v[i])) * k)) return false;
  return true;
}

// PolyN<8> (beam_geom.h:16-52).
struct Poly {
  double x[8], y[8];
  int n;
  FP64_HD void push(double px, double py) {
    if (n < 8) {
      x[n] = px;
      y[n] = py;
      ++n;
    }
  }
  FP64_HD double area() const {
    double a = 0;
    for (int i = 0; i < n; ++i) {
      const int j = i + 1 == n ? 0 : i + 1;
      a += x[i] * y[j] - x[j] * y[i];
    }
    return a * 0.5;
  }
  FP64_HD void reverse() {
    for (int i = 0, j = n - 1; i < j; ++i, --j) {
      const double tx = x[i], ty = y[i];
      x[i] = x[j];
      y[i] = y[j];
      x[j] = tx;
      y[j] = ty;
    }
  }
  FP64_HD double extent() const {
    double lx = INFINITY, hx = -INFINITY, ly = INFINITY, hy = -INFINITY;
    for (int i = 0; i < n; ++i) {
      lx = x[i] < lx ? x[i] : lx;
      hx = x[i] > hx ? x[i] : hx;
      ly = y[i] < ly ? y[i] : ly;
      hy = y[i] > hy ? y[i] : hy;
    }
    const double ex = hx - lx, ey = hy - ly;
    return ex > ey ? ex : ey;
  }
};

struct Line {
  double a, b, c;
  FP64_HD double at(double px, double py) const { return a * px + b * py + c; }
};

// makeLine (beam_geom.h:65-76).
FP64_HD inline Line makeLine(double a, double b, double c, double ext) {
  Line L{0, 0, 0};
  const double nrm = sqrt(a * a + b * b);
  if (!(nrm > 0) || nrm * ext * 1e9 < fabs(c)) {
    L.c = c > 0 ? 1.0 : (c < 0 ? -1.0 : 0.0);
    return L;
  }
  L.a = a / nrm;
  L.b = b / nrm;
  L.c = c / nrm;
  return L;
}

// splitPoly (beam_geom.h:82-114), capacity 8.
FP64_HD inline bool splitPoly(const Poly& p, const Line& L, double eps, Poly& in, Poly& out) {
  in.n = out.n = 0;
  double d[8];
  int cls[8];
  int nin = 0, nout = 0;
  for (int i = 0; i < p.n; ++i) {
    d[i] = L.at(p.x[i], p.y[i]);
    cls[i] = d[i] > eps ? 1 : (d[i] < -eps ? -1 : 0);
  }
  for (int i = 0; i < p.n; ++i) {
    const int j = i + 1 == p.n ? 0 : i + 1;
    nin += (cls[i] >= 0) + (cls[i] * cls[j] < 0);
    nout += (cls[i] <= 0) + (cls[i] * cls[j] < 0);
  }
  if (nin > 8 || nout > 8) return false;
  for (int i = 0; i < p.n; ++i) {
    const int j = i + 1 == p.n ? 0 : i + 1;
    if (cls[i] >= 0) in.push(p.x[i], p.y[i]);
    if (cls[i] <= 0) out.push(p.x[i], p.y[i]);
    if (cls[i] * cls[j] < 0) {
      const bool iLow = p.x[i] < p.x[j] || (p.x[i] == p.x[j] && p.y[i] < p.y[j]);
      const int a = iLow ? i : j, b = iLow ? j : i;
      const double t = d[a] / (d[a] - d[b]);
      const double X = p.x[a] + t * (p.x[b] - p.x[a]);
      const double Y = p.y[a] + t * (p.y[b] - p.y[a]);
      in.push(X, Y);
      out.push(X, Y);
    }
  }
  return true;
}

// cleanPoly (beam_geom.h:121-150).
FP64_HD inline void cleanPoly(Poly& p, double epsSplit) {
  const double e = epsSplit * 1e-3;
  bool changed = true;
  while (changed && p.n >= 3) {
    changed = false;
    for (int i = 0; i < p.n && p.n >= 3; ++i) {
      const int h = i == 0 ? p.n - 1 : i - 1;
      const int j = i + 1 == p.n ? 0 : i + 1;
      const double ex = p.x[j] - p.x[h], ey = p.y[j] - p.y[h];
      const double dx = p.x[i] - p.x[h], dy = p.y[i] - p.y[h];
      const double l = sqrt(ex * ex + ey * ey);
      const double d2 = dx * dx + dy * dy;
      const double cr = fabs(ex * dy - ey * dx);
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

// classifyPoly (beam_core.h:68-78): bit 0 some vertex inside, bit 1 outside.
FP64_HD inline int sides(const Poly& p, const Line& L, double eps) {
  int c = 0;
  for (int i = 0; i < p.n; ++i) {
    const double d = L.at(p.x[i], p.y[i]);
    if (d > eps)
      c |= 1;
    else if (d < -eps)
      c |= 2;
  }
  return c;
}

// splitFive (beam_core.h:80-89): pentagon -> quad (0..3) + triangle (0, 3, 4).
FP64_HD inline void splitFive(const Poly& p, Poly& a, Poly& b) {
  a.n = b.n = 0;
  for (int i = 0; i < 4; ++i) a.push(p.x[i], p.y[i]);
  b.push(p.x[0], p.y[0]);
  b.push(p.x[3], p.y[3]);
  b.push(p.x[4], p.y[4]);
}

// BeamPlane (beam_types.h:14-52).
struct Plane {
  V O, p0, u, v, n, d;
  FP64_HD V rel(V p) const { return p - O; }
  FP64_HD double rel(double s, int a) const { return s - at(O, a); }
  FP64_HD V point(double qx, double qy) const { return (p0 + u * qx) + v * qy; }
  FP64_HD void coords(V p, double& qx, double& qy) const {
    const V r = p - p0;
    qx = dot(r, u);
    qy = dot(r, v);
  }
  FP64_HD static Plane make(V O, V p0, V U, V Vv) {
    Plane P;
    P.O = O;
    P.p0 = p0;
    const V d = p0 - O;
    V n = nrmz(crs(U, Vv));
    if (dot(n, d) < 0) n = neg(n);
    P.u = nrmz(U);
    P.v = crs(n, P.u);
    P.n = n;
    P.d = d;
    return P;
  }
};

// The unit normal of a triangle, in double as Scene::finalize computes it
// (src/scene/scene.cpp:18-33).
FP64_HD inline V unitNormal(V a, V b, V c) {
  const V e1 = b - a, e2 = c - a;
  const V nn = crs(e1, e2);
  const double l = sqrt(nn.x * nn.x + nn.y * nn.y + nn.z * nn.z);
  return l > 0 ? V{nn.x / l, nn.y / l, nn.z / l} : nn;
}

// A sub-beam: corners on the query plane (lanes past n repeat the last) and
// the ray-direction sign per axis.
struct Beam {
  L4 qx, qy;
  int n;
  int sgn[3];
};

// The working form of a beam (CoreBeam, beam_core.h:30-40): |w| per axis.
struct Aw {
  L4 aw[3];
};

// CoreTriInfo (beam_core.h:42-51).
struct Tri {
  bool ok;
  Line edge[3];
  double A[3], C;
  V r[3], n;
};

struct Work {
  int beam;
  uint32_t node;
  int depth;
  uint64_t trail;
  Box box;
};

// A piece of the presplit's stack.
struct PreItem {
  Poly p;
  int axis;
  int sgn[3];
};

// The tracer's arrays live in device memory that the library owns, not in
// private memory: the runtime reserves a queue's private memory for every
// wave the device can hold, which at tens of KB per thread exhausts it once
// a few contexts have run. Records are interleaved by 8-byte word across the
// threads that share a region (a wave: 64; a host build: 1): word w of thread
// `lane` is base[w * threads + lane], so a wave's same-word accesses are one
// contiguous 512-byte run, the layout the hardware gives private memory.
struct WordMem {
  uint64_t* base;
  int threads, lane;
  template <class T>
  FP64_HD T load(uint32_t w) const {
    static_assert(sizeof(T) % 8 == 0, "records are whole words");
    uint64_t u[sizeof(T) / 8];
#pragma unroll
    for (int i = 0; i < int(sizeof(T) / 8); ++i) u[i] = base[size_t(w + uint32_t(i)) * size_t(threads) + size_t(lane)];
    T v;
    __builtin_memcpy(&v, u, sizeof(T));
    return v;
  }
  template <class T>
  FP64_HD void store(uint32_t w, const T& v) const {
    static_assert(sizeof(T) % 8 == 0, "records are whole words");
    uint64_t u[sizeof(T) / 8];
    __builtin_memcpy(u, &v, sizeof(T));
#pragma unroll
    for (int i = 0; i < int(sizeof(T) / 8); ++i) base[size_t(w + uint32_t(i)) * size_t(threads) + size_t(lane)] = u[i];
  }
  FP64_HD int loadInt(uint32_t w) const {
    return int32_t(uint32_t(base[size_t(w) * size_t(threads) + size_t(lane)]));
  }
  FP64_HD void storeInt(uint32_t w, int v) const {
    base[size_t(w) * size_t(threads) + size_t(lane)] = uint64_t(uint32_t(v));
  }
};

// One pixel's query, depth-first. P: beams per pixel, K: pending work items,
// S: pieces per clip list.
template <int P, int K>
struct Tracer {
  static constexpr int S = 16;
  // the thread's arrays: word offsets in its WordMem region
  static constexpr uint32_t wBeam = sizeof(Beam) / 8, wWork = sizeof(Work) / 8, wPoly = sizeof(Poly) / 8,
                            wItem = sizeof(PreItem) / 8;
  static constexpr uint32_t oPool = 0;                      // P beams
  static constexpr uint32_t oFree = oPool + P * wBeam;      // P free beam slots
  static constexpr uint32_t oCurA = oFree + P;              // P + P: the beams left against the current
  static constexpr uint32_t oCurB = oCurA + P;              //   triangle of a leaf and those for the next
  static constexpr uint32_t oWork = oCurB + P;              // K pending work items
  static constexpr uint32_t oList = oWork + K * wWork;      // 2 x S clip pieces (outside, inside)
  static constexpr uint32_t oClip = oList + 2 * S * wPoly;  // clip's stack: 8 polygons, 8 line indices
  static constexpr uint32_t oPre = oClip + 8 * wPoly + 8;   // presplit's stack: 8 items
  static constexpr uint32_t kWords = oPre + 8 * wItem;      // words per thread
  static_assert(sizeof(Beam) % 8 == 0 && sizeof(Work) % 8 == 0 && sizeof(Poly) % 8 == 0 && sizeof(PreItem) % 8 == 0,
                "records are whole words");
  WordMem m;
  const DevScene* sv;
  Plane Q;
  int exclude;
  double eps, minArea, ext, farCq;
  V x, ns;     // shading point and normal (exact Lambert integral of the miss pieces)
  bool exact;  // integrate the miss pieces
  double missArea, G;
  bool overflow;        // a pool, the work stack or the work cap ran out: the trace is void
  uint32_t work;        // triangle tests and kd steps of this trace
  uint32_t workCap = 0xffffffffu;
  Cnt cnt;  // counters of every trace this tracer ran (registers: the owner reads them at the end)
  int nFree, nPool;
  uint32_t oCur, oNxt;  // oCurA / oCurB, swapped after each triangle of a leaf
  int nCur, nNxt;
  int nWork;
  int nList[2];

  FP64_HD Beam beam(int b) const { return m.load<Beam>(oPool + uint32_t(b) * wBeam); }
  FP64_HD void putBeam(int b, const Beam& v) const { m.store(oPool + uint32_t(b) * wBeam, v); }
  FP64_HD Poly piece(int l, int k) const { return m.load<Poly>(oList + uint32_t(l * S + k) * wPoly); }

  FP64_HD void setup(double extent) {
    ext = extent > 0 ? extent : 1.0;
    eps = kEps * ext;
    minArea = eps * eps;
    farCq = dot(Q.n, Q.d);
    missArea = G = 0;
    overflow = false;
    work = 0;
    nFree = nPool = nCur = nNxt = nWork = 0;
    oCur = oCurA;
    oNxt = oCurB;
  }

  FP64_HD int alloc() {
    if (nFree > 0) return m.loadInt(oFree + uint32_t(--nFree));
    if (nPool >= P) {
      overflow = true;
      return -1;
    }
    return nPool++;
  }
  FP64_HD void release(int b) {
    if (nFree < P) m.storeInt(oFree + uint32_t(nFree++), b);
  }
  FP64_HD void pushWork(const Work& w) {
    if (nWork >= K) {
      overflow = true;
      return;
    }
    m.store(oWork + uint32_t(nWork++) * wWork, w);
  }
  FP64_HD void pushNext(int b) {
    if (nNxt >= P) {
      overflow = true;
      return;
    }
    m.storeInt(oNxt + uint32_t(nNxt++), b);
  }

  FP64_HD Aw awOf(const Beam& b) const {
    Aw r;
    for (int a = 0; a < 3; ++a) {
      const L4 w = affine(b.qx, at(Q.u, a), b.qy, at(Q.v, a), at(Q.d, a));
      for (int i = 0; i < 4; ++i) {
        const double ws = w.v[i] * double(b.sgn[a]);
        r.aw[a].v[i] = 0.0 > ws ? 0.0 : ws;
      }
    }
    return r;
  }
  // The light plane as every beam's far plane (setStatus, beam_core.h:225-240).
  FP64_HD double farC() const { return absS(farCq); }

  // decide (beam_core.h:139-164).
  FP64_HD int decide(const Beam& b, const Aw& w, int a, double Ns, const Box& bx) const {
    const L4& awa = w.aw[a];
    for (int k = 1; k <= 2; ++k) {
      const int c = a + k >= 3 ? a + k - 3 : a + k;
      const double sg = double(b.sgn[c]);
      const double entry = b.sgn[c] > 0 ? double(comp(bx.lo, c)) : double(comp(bx.hi, c));
      const double exit = b.sgn[c] > 0 ? double(comp(bx.hi, c)) : double(comp(bx.lo, c));
      const double nmin = sg * Q.rel(entry, c), nmax = sg * Q.rel(exit, c);
      const L4 lhs = mul(w.aw[c], Ns);
      if (allLess(lhs, mul(awa, nmin), kKd)) return 1;     // far
      if (allGreater(lhs, mul(awa, nmax), kKd)) return 0;  // near
    }
    const double fc = farC();
    if (allGreater(mul(L4{{fc, fc, fc, fc}}, Ns), mul(awa, fc), kKd)) return 0;
    return 2;  // both
  }

  // overlaps (beam_core.h:166-195).
  FP64_HD bool overlaps(const Beam& b, const Aw& w, const Box& bx) const {
    double nmin[3], nmax[3];
    for (int c = 0; c < 3; ++c) {
      const double sg = double(b.sgn[c]);
      nmin[c] = sg * Q.rel(b.sgn[c] > 0 ? double(comp(bx.lo, c)) : double(comp(bx.hi, c)), c);
      nmax[c] = sg * Q.rel(b.sgn[c] > 0 ? double(comp(bx.hi, c)) : double(comp(bx.lo, c)), c);
      if (nmax[c] < 0) return false;
    }
    for (int c = 0; c < 3; ++c)
      for (int e = 0; e < 3; ++e) {
        if (c == e) continue;
        if (allLess(mul(w.aw[e], nmax[c]), mul(w.aw[c], nmin[e]), kKd)) return false;
      }
    const double fc = farC();
    for (int c = 0; c < 3; ++c)
      if (allGreater(mul(L4{{fc, fc, fc, fc}}, nmin[c]), mul(w.aw[c], fc), kKd)) return false;
    return true;
  }

  // beamGeom (beam_core.h:200-221): false (and the area dropped) for a sliver.
  FP64_HD bool geom(const Poly& p, const int sgn[3], Beam& b) {
    Poly c = p;
    cleanPoly(c, eps);
    const double A = c.n >= 3 ? c.area() : 0.0;
    if (c.n < 3 || A < minArea) return false;
    const int l3 = c.n > 3 ? 3 : 2;
    b.qx = L4{{c.x[0], c.x[1], c.x[2], c.x[l3]}};
    b.qy = L4{{c.y[0], c.y[1], c.y[2], c.y[l3]}};
    b.n = c.n;
    for (int a = 0; a < 3; ++a) b.sgn[a] = sgn[a];
    return true;
  }

  // computeTriInfo + triSetup (beam_core.h:242-286) of a leaf triangle.
  FP64_HD Tri setupTri(const LeafTri& t) const {
    Tri ti;
    ti.ok = false;
    V p[3];
    for (int k = 0; k < 3; ++k) {
      p[k] = vmk(double(t.v[3 * k]), double(t.v[3 * k + 1]), double(t.v[3 * k + 2]));
      ti.r[k] = Q.rel(p[k]);
    }
    ti.n = unitNormal(p[0], p[1], p[2]);
    V m[3];
    const int E[3][2] = {{0, 1}, {1, 2}, {2, 0}};
    for (int k = 0; k < 3; ++k) {
      const int i = E[k][0], j = E[k][1];
      m[k] = t.vi[i] < t.vi[j] ? crs(ti.r[i], ti.r[j]) : neg(crs(ti.r[j], ti.r[i]));
    }
    const double D = dot(ti.r[2], m[0]);
    const double scale = len(m[0]) * len(ti.r[2]);
    if (!(absS(D) > kDeg * scale)) return ti;
    const double s = D > 0 ? 1.0 : -1.0;
    for (int k = 0; k < 3; ++k) ti.edge[k] = makeLine(s * dot(m[k], Q.u), s * dot(m[k], Q.v), s * dot(m[k], Q.d), ext);
    ti.A[0] = dot(ti.n, Q.u);
    ti.A[1] = dot(ti.n, Q.v);
    ti.A[2] = dot(ti.n, Q.d);
    ti.C = dot(ti.n, ti.r[0]);
    ti.ok = true;
    return ti;
  }

  // closerLine (beam_core.h:302-317) against the light plane.
  FP64_HD Line farLine(const Tri& T) const {
    const double AF[3] = {0, 0, farCq}, CF = farCq;
    const double cosang = dot(T.n, Q.n);
    const double aC = absS(T.C), aF = absS(CF);
    const double scaleC = aC > aF ? aC : aF;
    const double ca = absS(cosang);
    const double diff = T.C - cosang * CF;
    if (ca > 1 - kCopCos && absS(diff) <= kCopRel * scaleC) return Line{0, 0, -1.0};
    const double sg = T.C * CF > 0 ? 1.0 : -1.0;
    return makeLine(sg * (CF * T.A[0] - T.C * AF[0]), sg * (CF * T.A[1] - T.C * AF[1]),
                    sg * (CF * T.A[2] - T.C * AF[2]), ext);
  }

  // separated (beam_core.h:321-342).
  FP64_HD bool separated(const Beam& b, const Tri& ti) const {
    const L4 wx = affine(b.qx, Q.u.x, b.qy, Q.v.x, Q.d.x);
    const L4 wy = affine(b.qx, Q.u.y, b.qy, Q.v.y, Q.d.y);
    const L4 wz = affine(b.qx, Q.u.z, b.qy, Q.v.z, Q.d.z);
    L4 nx, ny, nz, nl;
    for (int i = 0; i < 4; ++i) {
      const int j = (i + 1) & 3;
      nx.v[i] = wy.v[i] * wz.v[j] - wz.v[i] * wy.v[j];
      ny.v[i] = wz.v[i] * wx.v[j] - wx.v[i] * wz.v[j];
      nz.v[i] = wx.v[i] * wy.v[j] - wy.v[i] * wx.v[j];
      nl.v[i] = sqrt(nx.v[i] * nx.v[i] + ny.v[i] * ny.v[i] + nz.v[i] * nz.v[i]);
    }
    int outAll = b.n == 4 ? 0xF : 0xB;
    for (int k = 0; k < 3 && outAll; ++k) {
      const V r = ti.r[k];
      const double lr = len(r) * kEps;
      int m = 0;
      for (int i = 0; i < 4; ++i) {
        const double s = nx.v[i] * r.x + ny.v[i] * r.y + nz.v[i] * r.z;
        if (s < -(nl.v[i] * lr)) m |= 1 << i;
      }
      outAll &= m;
    }
    return outAll != 0;
  }

  // testLine (beam_core.h:356-362).
  FP64_HD bool testLine(const Beam& B, const Line& line, Line* L, int& nl, bool& allIn) const {
    const L4 dv = affine(B.qx, line.a, B.qy, line.b, line.c);
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

  // classifyTri (beam_core.h:374-409) of an any-hit query: 0 Keep, 1 AllIn, 2 Clip.
  FP64_HD int classify(const Beam& B, const Tri& ti, Line* L, int& nl) const {
    nl = 0;
    bool allIn = true;
    for (int k = 0; k < 3; ++k)
      if (!testLine(B, ti.edge[k], L, nl, allIn)) return 0;
    if (!testLine(B, farLine(ti), L, nl, allIn)) return 0;
    if (allIn) return 1;
    if (separated(B, ti)) return 0;
    return 2;
  }

  FP64_HD void listPush(int l, const Poly& p) {
    if (nList[l] >= S) {
      overflow = true;
      return;
    }
    m.store(oList + uint32_t(l * S + nList[l]++) * wPoly, p);
  }

  // clipTri (beam_core.h:419-479) into the piece lists 0 (outside) and 1 (inside).
  FP64_HD bool clip(const Beam& B, Line* L, int nl) {
    nList[0] = nList[1] = 0;
    double score[3];
    for (int k = 0; k < 3; ++k) {
      const L4 dv = affine(B.qx, L[k].a, B.qy, L[k].b, L[k].c);
      double q[4];
      for (int i = 0; i < 4; ++i) q[i] = dv.v[i] > 0.0 ? dv.v[i] : 0.0;
      score[k] = q[0] + q[1] + q[2] + (B.n == 4 ? q[3] : 0.0);
    }
    for (int i = 1; i < 3; ++i)
      for (int j = i; j > 0 && score[j] < score[j - 1]; --j) {
        const double ts = score[j];
        score[j] = score[j - 1];
        score[j - 1] = ts;
        const Line tl = L[j];
        L[j] = L[j - 1];
        L[j - 1] = tl;
      }
    // the pieces still to clip: polygons at oClip, their next line at oClip + 8 wPoly
    int sp = 1;
    {
      Poly p0;
      p0.n = B.n;
      for (int i = 0; i < B.n; ++i) {
        p0.x[i] = B.qx.v[i];
        p0.y[i] = B.qy.v[i];
      }
      m.store(oClip, p0);
      m.storeInt(oClip + 8 * wPoly, 0);
    }
    while (sp > 0) {
      --sp;
      Poly p = m.load<Poly>(oClip + uint32_t(sp) * wPoly);
      int li = m.loadInt(oClip + 8 * wPoly + uint32_t(sp));
      bool inside = true;
      for (; li < nl; ++li) {
        const int c = sides(p, L[li], eps);
        if (!(c & 2)) continue;
        if (!(c & 1)) {
          listPush(0, p);
          inside = false;
          break;
        }
        Poly in, outp;
        if (!splitPoly(p, L[li], eps, in, outp)) return false;
        cleanPoly(outp, eps);
        if (outp.n >= 3) listPush(0, outp);
        cleanPoly(in, eps);
        if (in.n < 3) {
          inside = false;
          break;
        }
        if (in.n == 5) {
          Poly a, b;
          splitFive(in, a, b);
          cnt.v[sFive]++;
          if (sp >= 8) return false;
          m.store(oClip + uint32_t(sp) * wPoly, b);
          m.storeInt(oClip + 8 * wPoly + uint32_t(sp), li + 1);
          ++sp;
          p = a;
        } else {
          p = in;
        }
      }
      if (inside) listPush(1, p);
    }
    return true;
  }

  // output (beam_core.h:783-793): a finished piece; the miss pieces add their
  // area and their exact Lambert integral (polygonFormFactorG, soft_core.h:18-33).
  FP64_HD void output(const Beam& b, int tri) {
    Poly p;
    p.n = b.n;
    for (int i = 0; i < b.n; ++i) {
      p.x[i] = b.qx.v[i];
      p.y[i] = b.qy.v[i];
    }
    const double area = p.area();
    if (tri >= 0) cnt.v[sHits]++;
    cnt.v[sBeams]++;
    if (tri >= 0) return;
    missArea += area;
    if (!exact) return;
    D3 P4[4];
    for (int i = 0; i < 4; ++i) {
      const int k = i < p.n ? i : p.n - 1;
      const V q = Q.point(p.x[k], p.y[k]);
      P4[i] = D3{q.x, q.y, q.z};
    }
    G += formFactorD(D3{x.x, x.y, x.z}, D3{ns.x, ns.y, ns.z}, P4, p.n);
  }

  FP64_HD void finalize(int bi) {
    output(beam(bi), -1);
    release(bi);
  }

  // emit (beam_core.h:801-826) of an any-hit query.
  FP64_HD void emit(const Poly& p0, const Beam& parent, bool hitNow, int tri) {
    Poly pieces[2];
    int np = 1;
    if (p0.n == 5) {
      splitFive(p0, pieces[0], pieces[1]);
      cnt.v[sFive]++;
      np = 2;
    } else {
      pieces[0] = p0;
    }
    for (int i = 0; i < np; ++i) {
      Beam nb;
      if (!geom(pieces[i], parent.sgn, nb)) continue;
      if (hitNow) {
        output(nb, tri);
        continue;
      }
      const int idx = alloc();
      if (idx < 0) return;
      putBeam(idx, nb);
      pushNext(idx);
    }
  }

  // intersect (beam_core.h:832-882).
  FP64_HD void intersect(int bi, int t, const Tri& ti) {
    if (++work > workCap) {
      overflow = true;
      return;
    }
    cnt.v[sTests]++;
    Line L[5];
    int nl = 0;
    const Beam B = beam(bi);
    const int r = classify(B, ti, L, nl);
    if (r == 0) {
      pushNext(bi);
      return;
    }
    if (r == 1) {
      output(B, t);
      release(bi);
      return;
    }
    if (!clip(B, L, nl)) {
      overflow = true;
      return;
    }
    double newArea = 0;
    for (int k = 0; k < nList[1]; ++k) newArea += piece(1, k).area();
    if (!(newArea >= minArea)) {
      pushNext(bi);
      return;
    }
    cnt.v[sSplits]++;
    release(bi);
    // lists are reused by the next clip: copy out before emitting
    const int n0 = nList[0], n1 = nList[1];
    for (int k = 0; k < n0; ++k) emit(piece(0, k), B, false, t);
    for (int k = 0; k < n1; ++k) emit(piece(1, k), B, true, t);
  }

  // advance with the restart trail (beam_core.h:924-969).
  FP64_HD void advance(int bi, uint32_t node, int depth, uint64_t trail) {
    const Beam b = beam(bi);
    const Aw w = awOf(b);
    while (trail != 0) {
      const int d = 63 - __builtin_clzll(trail);
      for (; depth > d; --depth) {
        node = sv->parent[node];
        cnt.v[sClimbs]++;
      }
      trail &= (uint64_t(1) << d) - 1;
      const KdNodeD kn = sv->nodes[node];
      const uint32_t cs = b.sgn[kn.data & 3u] > 0 ? (kn.data >> 2) : node + 1;
      const Box bx = sv->nodeBox[cs];
      if (overlaps(b, w, bx)) {
        pushWork(Work{bi, cs, d + 1, trail, bx});
        return;
      }
    }
    finalize(bi);
  }

  FP64_HD void rootPiece(const Poly& p, const int sgn[3], uint32_t node0, const Box& rootBox) {
    Beam b;
    if (!geom(p, sgn, b)) return;
    const int idx = alloc();
    if (idx < 0) return;
    putBeam(idx, b);
    cnt.v[sRoots]++;
    if (overlaps(b, awOf(b), rootBox))
      pushWork(Work{idx, node0, 0, 0, rootBox});
    else
      finalize(idx);
  }

  // presplit (beam_core.h:487-546) by the ray-direction sign lines per axis.
  FP64_HD void presplit(const Poly& root, uint32_t node0, const Box& rootBox) {
    int sp = 0;
    m.store(oPre, PreItem{root, 0, {1, 1, 1}});
    ++sp;
    while (sp > 0 && !overflow) {
      PreItem it = m.load<PreItem>(oPre + uint32_t(--sp) * wItem);
      const Poly& p = it.p;
      if (p.n < 3) continue;
      if (it.axis == 3) {
        for (int k = 1; k + 1 < p.n; k += 2) {
          Poly qd;
          qd.n = 0;
          qd.push(p.x[0], p.y[0]);
          qd.push(p.x[k], p.y[k]);
          qd.push(p.x[k + 1], p.y[k + 1]);
          if (k + 2 < p.n) qd.push(p.x[k + 2], p.y[k + 2]);
          rootPiece(qd, it.sgn, node0, rootBox);
        }
        continue;
      }
      const int axis = it.axis;
      const Line L = makeLine(at(Q.u, axis), at(Q.v, axis), at(Q.d, axis), ext);
      const int c = sides(p, L, eps);
      if (c == 3) {
        Poly in, outp;
        splitPoly(p, L, eps, in, outp);
        cleanPoly(in, eps);
        cleanPoly(outp, eps);
        if (sp + 2 > 8) {
          overflow = true;
          return;
        }
        PreItem a = it, b = it;
        a.p = in;
        a.axis = axis + 1;
        a.sgn[axis] = 1;
        b.p = outp;
        b.axis = axis + 1;
        b.sgn[axis] = -1;
        m.store(oPre + uint32_t(sp++) * wItem, b);
        m.store(oPre + uint32_t(sp++) * wItem, a);
        cnt.v[sPresplit]++;
      } else if (c != 0) {
        it.axis = axis + 1;
        it.sgn[axis] = (c & 1) ? 1 : -1;
        if (sp >= 8) {
          overflow = true;
          return;
        }
        m.store(oPre + uint32_t(sp++) * wItem, it);
      }
    }
  }

  // BeamCore::trace with the restart trail (beam_core.h:569-661) of the root
  // polygon rootIn; `extent` sets the tolerances (the extent of the whole
  // light's root polygon, also when rootIn is a cell of it). The depth-first
  // order of the reference, written as one loop of small steps (take a work
  // item; one kd step; one beam against one triangle of the leaf
  // (processLeaf without mailboxes, beam_core.h:884-912); advance one of the
  // leaf's surviving beams) so that the lanes of a wave in the same kind of
  // step run together, instead of each waiting for the slowest lane's whole
  // descent or leaf.
  FP64_HD void trace(const Poly& rootIn, uint32_t node0, const Box& rootBox, double extent) {
    setup(0.0 > extent ? 0.0 : extent);
    Poly p = rootIn;
    if (p.n >= 3) {
      if (p.area() < 0) p.reverse();
      presplit(p, node0, rootBox);
    }
    enum { kTake, kDescend, kTest, kAdvance };
    int state = kTake;
    int bi = 0, depth = 0;  // the work item: its beam, node, box, depth and restart trail
    uint32_t node = 0;
    Box box = rootBox;
    uint64_t trail = 0;
    Beam B;
    Aw w;
    uint32_t leafOff = 0, leafN = 0, li = 0;  // the leaf's references and the next one to take
    bool haveTri = false;                     // ti / triId: the triangle the survivors meet now
    Tri ti;
    int triId = 0, k = 0;                     // the next survivor to test or advance
    while (!overflow) {
      if (state == kTake) {
        if (nWork == 0) break;
        const Work wk = m.load<Work>(oWork + uint32_t(--nWork) * wWork);
        bi = wk.beam;
        node = wk.node;
        box = wk.box;
        depth = wk.depth;
        trail = wk.trail;
        B = beam(bi);
        w = awOf(B);
        state = kDescend;
      }
      if (state == kDescend) {
        const KdNodeD kn = sv->nodes[node];
        if ((kn.data & 3u) == 3u) {
          cnt.v[sLeaf]++;
          if ((kn.data >> 2) == 0) {
            advance(bi, node, depth, trail);
            state = kTake;
          } else {
            nCur = 1;
            m.storeInt(oCur, bi);
            leafOff = kn.offset;
            leafN = kn.data >> 2;
            li = 0;
            haveTri = false;
            state = kTest;
          }
        } else {
          if (++work > workCap) {
            overflow = true;
            break;
          }
          cnt.v[sKd]++;
          const int d = depth++;
          const int a = int(kn.data & 3u);
          const float s = kn.split;
          const double Ns = double(B.sgn[a]) * Q.rel(double(s), a);
          const uint32_t left = node + 1, right = kn.data >> 2;
          Box lb = box, rb = box;
          setComp(lb.hi, a, s);
          setComp(rb.lo, a, s);
          const bool toRight = B.sgn[a] > 0;
          const uint32_t cs = toRight ? right : left, co = toRight ? left : right;
          if (Ns <= 0) {
            node = cs;
            box = toRight ? rb : lb;
          } else {
            const int dec = decide(B, w, a, Ns, box);
            if (dec == 0) {
              node = co;
              box = toRight ? lb : rb;
            } else if (dec == 1) {
              node = cs;
              box = toRight ? rb : lb;
            } else {
              trail |= uint64_t(1) << d;
              node = co;
              box = toRight ? lb : rb;
            }
          }
        }
      }
      if (state == kTest) {
        while (!haveTri && li < leafN) {  // the leaf's next triangle the survivors can meet
          const LeafTri ref = sv->refs[leafOff + li];
          if (ref.tri != exclude) {
            ti = setupTri(ref);
            if (ti.ok) {
              haveTri = true;
              triId = ref.tri;
              nNxt = 0;
              k = 0;
              break;
            }
          }
          ++li;
        }
        if (!haveTri) {
          k = 0;
          state = kAdvance;
        } else {
          intersect(m.loadInt(oCur + uint32_t(k)), triId, ti);
          if (++k == nCur) {
            const uint32_t t = oCur;
            oCur = oNxt;
            oNxt = t;
            nCur = nNxt;
            haveTri = false;
            ++li;
            if (nCur == 0) li = leafN;
          }
        }
      }
      if (state == kAdvance) {
        if (k < nCur) advance(m.loadInt(oCur + uint32_t(k++)), node, depth, trail);
        if (k >= nCur) state = kTake;
      }
    }
  }
};

// One receiver point's query: beamLightVisibility (render.cpp:279-303) with
// lightBeamSetup and lightCenterG (soft_core.h:38-84) in double, up to the
// trace. x: the hit point; ref: the hit triangle's leaf reference (its normal
// is taken in double from the corners); nf: the triangle normal facing the
// eye; the light (c, U, V); the apex offset along the normal.
struct PixelRoot {
  V x, ns;
  Plane Q;
  Poly root;         // the light polygon on Q, above the receiver's horizon (n = 0: nothing to trace)
  double lightArea;  // of the whole light polygon
  double ext;        // the root's extent (the trace's tolerances)
  double A, G0;      // the light's area and the centre estimate of G
  bool front;        // the receiver is in front of the light
};

FP64_HD inline void pixelRoot(const LeafTri& ref, V x, V nf, V c, V U, V Vv, double offset, PixelRoot& R) {
  const V nd = unitNormal(vmk(ref.v[0], ref.v[1], ref.v[2]), vmk(ref.v[3], ref.v[4], ref.v[5]),
                          vmk(ref.v[6], ref.v[7], ref.v[8]));
  R.x = x;
  R.ns = dot(nd, nf) < 0 ? neg(nd) : nd;
  const V uv = crs(U, Vv);
  R.A = len(uv);
  const V nL = nrmz(uv);
  R.front = dot(x - c, nL) > 0;
  R.G0 = 0;
  R.root.n = 0;
  R.lightArea = 0;
  R.ext = 0;
  if (!R.front) return;
  {
    const V w = c - x;
    const double r2 = dot(w, w), rl = sqrt(r2);
    const V wn = vmk(w.x / rl, w.y / rl, w.z / rl);
    const double ci = dot(R.ns, wn), cl = -dot(nL, wn);
    R.G0 = (ci > 0 ? ci : 0) * (cl > 0 ? cl : 0) / r2;
  }
  R.Q = Plane::make(x + R.ns * offset, c, U, Vv);
  const double hzc = dot(R.ns, c - x);
  const V hu = U * 0.5, hv = Vv * 0.5;
  const V corners[4] = {(c - hu) - hv, (c + hu) - hv, (c + hu) + hv, (c - hu) + hv};
  Poly quad;
  quad.n = 0;
  for (int i = 0; i < 4; ++i) {
    double qx, qy;
    R.Q.coords(corners[i], qx, qy);
    quad.push(qx, qy);
  }
  if (quad.area() < 0) quad.reverse();
  R.lightArea = quad.area();
  const Line hz{dot(R.ns, R.Q.u), dot(R.ns, R.Q.v), hzc};
  int above = 0;
  for (int i = 0; i < quad.n; ++i) above += hz.at(quad.x[i], quad.y[i]) > 0;
  if (above == 0) return;
  Poly root = quad;
  if (above != quad.n) {
    Poly rest;
    splitPoly(quad, hz, 0, root, rest);
  }
  if (!(root.n >= 3 && root.area() > 0)) return;
  R.root = root;
  const double e0 = root.extent();
  R.ext = e0;
}

// The cells a traced root splits into when it needs too much work: level 0
// is the whole root; level 1 the cells of a 4 x 4 grid over its bounding box
// on the query plane, level 2 of a 16 x 16 grid (each level-1 cell split
// 4 x 4; the two grids share their lines exactly). Every cell is traced with
// the whole root's tolerances and the cells' results add up to the root's.
FP64_HD inline int cellGrid(int level) { return level == 0 ? 1 : (level == 1 ? 4 : 16); }

FP64_HD inline Poly cellOf(const Poly& root, int level, int cx, int cy) {
  if (level == 0) return root;
  double x0 = INFINITY, x1 = -INFINITY, y0 = INFINITY, y1 = -INFINITY;
  for (int i = 0; i < root.n; ++i) {
    x0 = root.x[i] < x0 ? root.x[i] : x0;
    x1 = root.x[i] > x1 ? root.x[i] : x1;
    y0 = root.y[i] < y0 ? root.y[i] : y0;
    y1 = root.y[i] > y1 ? root.y[i] : y1;
  }
  const int g = cellGrid(level);
  const double xa = x0 + (x1 - x0) * (double(cx) / g), xb = cx + 1 == g ? x1 : x0 + (x1 - x0) * (double(cx + 1) / g);
  const double ya = y0 + (y1 - y0) * (double(cy) / g), yb = cy + 1 == g ? y1 : y0 + (y1 - y0) * (double(cy + 1) / g);
  const Line L[4] = {{1, 0, -xa}, {-1, 0, xb}, {0, 1, -ya}, {0, -1, yb}};
  Poly p = root;
  for (int k = 0; k < 4 && p.n >= 3; ++k) {
    Poly in, out;
    splitPoly(p, L[k], 0, in, out);
    p = in;
  }
  if (p.n < 3) p.n = 0;
  return p;
}

// Traces one cell of a pixel's root (cellOf) with at most `workCap` triangle
// tests and kd steps. False when the tracer overflowed (pools, work stack or
// work cap): the cell must be traced again, split or with larger pools.
// Otherwise its missed area and (exact) Lambert integral are in miss and G.
template <class TR>
FP64_HD bool traceCell(TR& tr, const DevScene& sv, const PixelRoot& R, int exclude, bool exact, int level, int cx,
                       int cy, uint32_t workCap, uint32_t node0, const Box& bounds, double& miss,
                       double& G) {
  miss = G = 0;
  const Poly cell = cellOf(R.root, level, cx, cy);
  if (cell.n < 3) return true;
  tr.Q = R.Q;
  tr.sv = &sv;
  tr.exclude = exclude;
  tr.x = R.x;
  tr.ns = R.ns;
  tr.exact = exact;
  tr.workCap = workCap;
  tr.trace(cell, node0, bounds, R.ext);
  if (tr.overflow) return false;
  miss = tr.missArea;
  G = tr.G;
  return true;
}

// The per-pixel sums of the cells' results, in fixed point so that their
// order does not matter: the visible fraction in units of 2^-62 (a cell adds
// its missed area / the light's), G in units of 2^-60 (G <= pi).
FP64_HD inline unsigned long long visFx(double missOverLight) {
  return (unsigned long long)((missOverLight > 0 ? missOverLight : 0.0) * 0x1p62);
}
FP64_HD inline unsigned long long gFx(double G) { return (unsigned long long)((G > 0 ? G : 0.0) * 0x1p60); }
FP64_HD inline double visOfFx(unsigned long long v) {
  const double r = double(v) * 0x1p-62;
  return r > 1 ? 1.0 : r;
}
FP64_HD inline double gOfFx(unsigned long long v) { return double(v) * 0x1p-60; }

}  // namespace fp64
