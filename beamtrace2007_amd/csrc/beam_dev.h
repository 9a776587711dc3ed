// Device-side beam geometry for gfx950: the arithmetic of the reference's
// float build, written for wave64 kernels.
//
// Bit-exactness contract: every float operation here is the one the
// reference's float build performs (src/core, src/beam/beam_geom.h,
// src/beam/beam_core.h), in the same order: the library is compiled with
// -ffp-contract=off (no fused multiply-add), correctly rounded division and
// square root, and IEEE denormals. min/max/abs that the reference writes as
// selects ("a > b ? a : b", "x < 0 ? -x : x") are written as the same
// selects here, never as fminf/fmaxf, so -0 and NaN behave identically.
//
// The four corner rays of a beam are four float lanes ("C4", the reference's
// R4 of src/core/simd4.h); a compare over the corners yields a 4-bit mask.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define BTG_HD __host__ __device__ __forceinline__

namespace btg {

constexpr float kEpsRel = 2e-5f;   // fuzzy split epsilon / beam extent (src/core/real.h:36)
constexpr float kEpsRelDouble = 1e-9f;  // the double build's (real.h:31): FP64-guarded mode
constexpr float kKdRel = 1e-5f;    // kd decision margin (real.h:37)
constexpr float kDegRel = 2e-6f;   // edge-on detection (real.h:38)
constexpr float kCopCos = 2e-6f;   // coplanar tie (src/beam/beam_core.h:26-27)
constexpr float kCopRel = 2e-5f;

// ------------------------------------------------------------------ vectors
struct F3 {
  float x, y, z;
};
BTG_HD F3 f3(float x, float y, float z) { return F3{x, y, z}; }
BTG_HD F3 operator+(F3 a, F3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
BTG_HD F3 operator-(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
BTG_HD F3 operator-(F3 a) { return f3(-a.x, -a.y, -a.z); }
BTG_HD F3 operator*(F3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
BTG_HD F3 operator/(F3 a, float s) { return f3(a.x / s, a.y / s, a.z / s); }
BTG_HD float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
BTG_HD F3 cross(F3 a, F3 b) {
  return f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
BTG_HD float length(F3 a) { return __builtin_sqrtf(dot(a, a)); }
BTG_HD F3 normalize(F3 a) {
  float l = length(a);
  return l > 0 ? a / l : a;
}
// c ? a : b on values. (With two lvalue operands C++'s ?: is an lvalue: the
// compiler then selects between their addresses and loads once, which keeps
// arrays in private memory.)
template <class T>
BTG_HD T pick(bool c, T a, T b) {
  return c ? T(a) : T(b);
}
BTG_HD float comp(F3 a, int i) { return pick(i == 0, a.x, pick(i == 1, a.y, a.z)); }
BTG_HD void setComp(F3& a, int i, float v) {
  if (i == 0)
    a.x = v;
  else if (i == 1)
    a.y = v;
  else
    a.z = v;
}
BTG_HD float selAbs(float v) { return v < 0 ? -v : v; }
BTG_HD float selMax(float a, float b) { return a > b ? a : b; }
BTG_HD float selMin(float a, float b) { return a < b ? a : b; }

struct Box {
  F3 lo, hi;
};

// ------------------------------------------------------------ corner lanes
struct C4 {
  float v[4];
};
BTG_HD C4 c4(float a, float b, float c, float d) { return C4{{a, b, c, d}}; }
BTG_HD C4 splat(float s) { return C4{{s, s, s, s}}; }
// ((x * a) + (y * b)) + c per lane: the reference's "qx * a + qy * b + c".
BTG_HD C4 affine(const C4& x, float a, const C4& y, float b, float c) {
  C4 r;
#pragma unroll
  for (int i = 0; i < 4; ++i) r.v[i] = x.v[i] * a + y.v[i] * b + c;
  return r;
}
BTG_HD C4 scaled(const C4& a, float s) {
  C4 r;
#pragma unroll
  for (int i = 0; i < 4; ++i) r.v[i] = a.v[i] * s;
  return r;
}
// All four lanes: lhs < rhs - (|lhs| + |rhs|) k.
BTG_HD bool allBelow(const C4& l, const C4& r, float k) {
  bool all = true;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const float t = (__builtin_fabsf(l.v[i]) + __builtin_fabsf(r.v[i])) * k;
    all = all && (l.v[i] < r.v[i] - t);
  }
  return all;
}
// All four lanes: lhs > rhs + (|lhs| + |rhs|) k.
BTG_HD bool allAbove(const C4& l, const C4& r, float k) {
  bool all = true;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const float t = (__builtin_fabsf(l.v[i]) + __builtin_fabsf(r.v[i])) * k;
    all = all && (l.v[i] > r.v[i] + t);
  }
  return all;
}

// ---------------------------------------------------------------- polygons
// Convex polygon with room for N vertices (PolyN, src/beam/beam_geom.h:16-52).
//
// Every access below uses a compile-time slot index: loops run over all N
// slots with a guard on n, a push writes its slot through selects, and a
// vertex at a runtime index is read through a select chain. Functions that
// choose between two slots work on a local copy (choosing between two loads
// through a reference would be folded into one load through a chosen
// address). A polygon (and any struct holding one) therefore stays in
// registers instead of private (scratch) memory, which on a persistent kernel
// with thousands of lanes per XCD would spill through L2 to HBM. The float
// operations and their order are exactly those of the straightforward loops.
template <int N>
struct Pg {
  float x[N], y[N];
  int n;
  Pg() = default;
  // element-wise copies (a struct memcpy between polygons defeats the
  // promotion of their storage to registers)
  BTG_HD Pg(const Pg& o) { assign(o); }
  BTG_HD Pg& operator=(const Pg& o) {
    assign(o);
    return *this;
  }
  BTG_HD void assign(const Pg& o) {
#pragma unroll
    for (int k = 0; k < N; ++k) {
      x[k] = o.x[k];
      y[k] = o.y[k];
    }
    n = o.n;
  }
  BTG_HD void clear() { n = 0; }
  BTG_HD float xAt(int i) const {
    float v = x[0];
#pragma unroll
    for (int k = 1; k < N; ++k) v = pick(i == k, x[k], v);
    return v;
  }
  BTG_HD float yAt(int i) const {
    float v = y[0];
#pragma unroll
    for (int k = 1; k < N; ++k) v = pick(i == k, y[k], v);
    return v;
  }
  BTG_HD void push(float px, float py) {
#pragma unroll
    for (int k = 0; k < N; ++k) {
      const bool at = k == n;
      x[k] = pick(at, px, x[k]);
      y[k] = pick(at, py, y[k]);
    }
    n = n < N ? n + 1 : n;
  }
  // Removes vertex i, moving the later ones down.
  BTG_HD void removeAt(int i) {
    Pg<N> q = *this;
#pragma unroll
    for (int k = 0; k + 1 < N; ++k) {
      const bool mv = k >= i;
      q.x[k] = pick(mv, q.x[k + 1], q.x[k]);
      q.y[k] = pick(mv, q.y[k + 1], q.y[k]);
    }
    --q.n;
    *this = q;
  }
  // Signed area, summed vertex by vertex (positive = counter-clockwise).
  BTG_HD float area() const {
    const Pg<N> q = *this;
    float a = 0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      if (i < q.n) {
        const bool wrap = !(i + 1 < q.n);
        const float xj = pick(wrap, q.x[0], q.x[(i + 1) % N]), yj = pick(wrap, q.y[0], q.y[(i + 1) % N]);
        a += q.x[i] * yj - xj * q.y[i];
      }
    }
    return a * 0.5f;
  }
  BTG_HD void reverse() {
    const Pg<N> q = *this;
    Pg<N> r = q;
#pragma unroll
    for (int k = 0; k < N; ++k) {
      if (k < q.n) {
        r.x[k] = q.xAt(q.n - 1 - k);
        r.y[k] = q.yAt(q.n - 1 - k);
      }
    }
    *this = r;
  }
  BTG_HD float extent() const {
    float lx = __builtin_inff(), hx = -__builtin_inff(), ly = __builtin_inff(), hy = -__builtin_inff();
#pragma unroll
    for (int i = 0; i < N; ++i) {
      if (i < n) {
        lx = pick(x[i] < lx, x[i], lx);
        hx = pick(x[i] > hx, x[i], hx);
        ly = pick(y[i] < ly, y[i], ly);
        hy = pick(y[i] > hy, y[i], hy);
      }
    }
    const float ex = hx - lx, ey = hy - ly;
    return ex > ey ? ex : ey;
  }
  template <int M>
  BTG_HD void copyFrom(const Pg<M>& o) {
    n = o.n < N ? o.n : N;
#pragma unroll
    for (int i = 0; i < (M < N ? M : N); ++i) {
      x[i] = o.x[i];
      y[i] = o.y[i];
    }
  }
};

struct Ln {
  float a, b, c;
  Ln() = default;
  BTG_HD Ln(float a_, float b_, float c_) : a(a_), b(b_), c(c_) {}
  BTG_HD Ln(const Ln& o) : a(o.a), b(o.b), c(o.c) {}
  BTG_HD Ln& operator=(const Ln& o) {
    a = o.a;
    b = o.b;
    c = o.c;
    return *this;
  }
  BTG_HD float at(float px, float py) const { return a * px + b * py + c; }
};

// makeLine (beam_geom.h:65-76): unit normal; a line ~1e9 extents away (or
// exactly parallel) becomes the constant sign(c).
BTG_HD Ln makeLine(float a, float b, float c, float ext) {
  const float nrm = __builtin_sqrtf(a * a + b * b);
  if (!(nrm > 0) || nrm * ext * 1e9f < __builtin_fabsf(c)) return Ln{0, 0, c > 0 ? 1.0f : (c < 0 ? -1.0f : 0.0f)};
  return Ln{a / nrm, b / nrm, c / nrm};
}

// splitPoly (beam_geom.h:82-114): vertices within eps of the line go to both
// pieces; crossings are computed from the lexicographically smaller endpoint.
template <int N>
BTG_HD bool splitPoly(const Pg<N>& p0, const Ln& L, float eps, Pg<N>& in, Pg<N>& out) {
  const Pg<N> p = p0;
  in.clear();
  out.clear();
  float d[N];
  int cls[N];
  int nin = 0, nout = 0;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    d[i] = 0;
    cls[i] = 0;
    if (i < p.n) {
      d[i] = L.at(p.x[i], p.y[i]);
      cls[i] = d[i] > eps ? 1 : (d[i] < -eps ? -1 : 0);
    }
  }
#pragma unroll
  for (int i = 0; i < N; ++i) {
    if (i < p.n) {
      const int cj = pick(i + 1 < p.n, cls[(i + 1) % N], cls[0]);
      nin += (cls[i] >= 0) + (cls[i] * cj < 0);
      nout += (cls[i] <= 0) + (cls[i] * cj < 0);
    }
  }
  if (nin > N || nout > N) return false;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    if (i < p.n) {
      const bool wrap = !(i + 1 < p.n);
      const int j = (i + 1) % N;
      const float xi = p.x[i], yi = p.y[i], di = d[i];
      const float xj = pick(wrap, p.x[0], p.x[j]), yj = pick(wrap, p.y[0], p.y[j]), dj = pick(wrap, d[0], d[j]);
      const int cj = pick(wrap, cls[0], cls[j]);
      if (cls[i] >= 0) in.push(xi, yi);
      if (cls[i] <= 0) out.push(xi, yi);
      if (cls[i] * cj < 0) {
        const bool iLow = xi < xj || (xi == xj && yi < yj);
        const float xa = pick(iLow, xi, xj), ya = pick(iLow, yi, yj), da = pick(iLow, di, dj);
        const float xb = pick(iLow, xj, xi), yb = pick(iLow, yj, yi), db = pick(iLow, dj, di);
        const float t = da / (da - db);
        const float X = xa + t * (xb - xa);
        const float Y = ya + t * (yb - ya);
        in.push(X, Y);
        out.push(X, Y);
      }
    }
  }
  return true;
}

// cleanPoly (beam_geom.h:121-150): welds duplicate or collinear vertices at
// 1e-3 of the split epsilon; fewer than 3 vertices -> n = 0. The reference
// repeats passes over the vertices while a pass removed one, re-testing the
// vertex that moves into a removed slot. The first pass runs with constant
// indices; only a polygon that loses a vertex continues on the general loop
// (from the first removed slot, with exactly the reference's sequence).
BTG_HD bool weldTest(float xh, float yh, float xi, float yi, float xj, float yj, float e) {
  const float ex = xj - xh, ey = yj - yh;
  const float dx = xi - xh, dy = yi - yh;
  const float l = __builtin_sqrtf(ex * ex + ey * ey);
  const float d2 = dx * dx + dy * dy;
  const float cr = __builtin_fabsf(ex * dy - ey * dx);
  return d2 <= e * e || cr <= e * l;
}

template <int N>
BTG_HD void cleanPoly(Pg<N>& p0, float epsSplit) {
  const float e = epsSplit * 1e-3f;
  if (p0.n < 3) {
    p0.n = 0;
    return;
  }
  Pg<N> p = p0;
  int first = N;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    if (i < p.n && first == N) {
      const float xh = i == 0 ? p.xAt(p.n - 1) : float(p.x[(i + N - 1) % N]);
      const float yh = i == 0 ? p.yAt(p.n - 1) : float(p.y[(i + N - 1) % N]);
      const bool wrap = !(i + 1 < p.n);
      const float xj = pick(wrap, p.x[0], p.x[(i + 1) % N]), yj = pick(wrap, p.y[0], p.y[(i + 1) % N]);
      if (weldTest(xh, yh, p.x[i], p.y[i], xj, yj, e)) first = i;
    }
  }
  if (first == N) return;  // a pass without removals ends the reference's loop
  p.removeAt(first);
  bool changed = true;
  int i = first;
  while (p.n >= 3) {
    if (i >= p.n) {
      if (!changed) break;
      changed = false;
      i = 0;
      continue;
    }
    const int h = i == 0 ? p.n - 1 : i - 1;
    const int j = i + 1 == p.n ? 0 : i + 1;
    if (weldTest(p.xAt(h), p.yAt(h), p.xAt(i), p.yAt(i), p.xAt(j), p.yAt(j), e)) {
      p.removeAt(i);
      changed = true;
    } else {
      ++i;
    }
  }
  if (p.n < 3) p.n = 0;
  p0 = p;
}

// classifyPoly (beam_core.h:68-78): bit 0 some vertex strictly inside, bit 1
// some vertex strictly outside.
template <int N>
BTG_HD int sidesOf(const Pg<N>& p, const Ln& L, float eps) {
  int c = 0;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    if (i < p.n) {
      const float d = L.at(p.x[i], p.y[i]);
      if (d > eps)
        c |= 1;
      else if (d < -eps)
        c |= 2;
    }
  }
  return c;
}

// splitFive (beam_core.h:80-89): pentagon -> quad (0..3) + triangle (0, 3, 4).
template <int N>
BTG_HD void splitFive(const Pg<N>& p, Pg<N>& a, Pg<N>& b) {
  a = p;
  b = p;
  a.n = 4;
  b.n = 3;
  b.x[1] = p.x[3];
  b.y[1] = p.y[3];
  b.x[2] = p.x[4];
  b.y[2] = p.y[4];
}

// --------------------------------------------------------------- planes
// BeamPlane (src/beam/beam_types.h:14-52).
struct Plane {
  F3 O, p0, u, v, n, d;
  BTG_HD F3 point(float qx, float qy) const { return p0 + u * qx + v * qy; }
  BTG_HD void coords(F3 p, float& qx, float& qy) const {
    const F3 r = p - p0;
    qx = dot(r, u);
    qy = dot(r, v);
  }
  BTG_HD bool project(F3 p, float& qx, float& qy) const {
    const F3 r = p - O;
    const float den = dot(n, r);
    if (!(den > 0)) return false;
    coords(O + r * (dot(n, d) / den), qx, qy);
    return true;
  }
  // Olo: the low part of an apex held as the float pair O + Olo (the
  // FP64-guarded soft mode); zero otherwise, where (p0 - O) - 0 is p0 - O.
  BTG_HD static Plane make(F3 O, F3 p0, F3 u_, F3 v_, F3 Olo = F3{0, 0, 0}) {
    Plane P;
    P.O = O;
    P.p0 = p0;
    const F3 d = (p0 - O) - Olo;
    F3 n = normalize(cross(u_, v_));
    if (dot(n, d) < 0) n = -n;
    P.u = normalize(u_);
    P.v = cross(n, P.u);
    P.n = n;
    P.d = d;
    return P;
  }
};

}  // namespace btg
