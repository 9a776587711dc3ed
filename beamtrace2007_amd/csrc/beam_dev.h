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
BTG_HD float comp(F3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
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
template <int N>
struct Pg {
  float x[N], y[N];
  int n;
  BTG_HD void clear() { n = 0; }
  BTG_HD void push(float px, float py) {
    if (n < N) {
      x[n] = px;
      y[n] = py;
      ++n;
    }
  }
  // Signed area, summed vertex by vertex (positive = counter-clockwise).
  BTG_HD float area() const {
    float a = 0;
    for (int i = 0; i < n; ++i) {
      const int j = i + 1 == n ? 0 : i + 1;
      a += x[i] * y[j] - x[j] * y[i];
    }
    return a * 0.5f;
  }
  BTG_HD void reverse() {
    for (int i = 0, j = n - 1; i < j; ++i, --j) {
      float t = x[i];
      x[i] = x[j];
      x[j] = t;
      t = y[i];
      y[i] = y[j];
      y[j] = t;
    }
  }
  BTG_HD float extent() const {
    float lx = __builtin_inff(), hx = -__builtin_inff(), ly = __builtin_inff(), hy = -__builtin_inff();
    for (int i = 0; i < n; ++i) {
      lx = x[i] < lx ? x[i] : lx;
      hx = x[i] > hx ? x[i] : hx;
      ly = y[i] < ly ? y[i] : ly;
      hy = y[i] > hy ? y[i] : hy;
    }
    const float ex = hx - lx, ey = hy - ly;
    return ex > ey ? ex : ey;
  }
  template <int M>
  BTG_HD void copyFrom(const Pg<M>& o) {
    n = o.n < N ? o.n : N;
    for (int i = 0; i < n; ++i) {
      x[i] = o.x[i];
      y[i] = o.y[i];
    }
  }
};

struct Ln {
  float a, b, c;
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
BTG_HD bool splitPoly(const Pg<N>& p, const Ln& L, float eps, Pg<N>& in, Pg<N>& out) {
  in.clear();
  out.clear();
  float d[N];
  int cls[N];
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
  if (nin > N || nout > N) return false;
  for (int i = 0; i < p.n; ++i) {
    const int j = i + 1 == p.n ? 0 : i + 1;
    if (cls[i] >= 0) in.push(p.x[i], p.y[i]);
    if (cls[i] <= 0) out.push(p.x[i], p.y[i]);
    if (cls[i] * cls[j] < 0) {
      const bool iLow = p.x[i] < p.x[j] || (p.x[i] == p.x[j] && p.y[i] < p.y[j]);
      const int a = iLow ? i : j, b = iLow ? j : i;
      const float t = d[a] / (d[a] - d[b]);
      const float X = p.x[a] + t * (p.x[b] - p.x[a]);
      const float Y = p.y[a] + t * (p.y[b] - p.y[a]);
      in.push(X, Y);
      out.push(X, Y);
    }
  }
  return true;
}

// cleanPoly (beam_geom.h:121-150): welds duplicate or collinear vertices at
// 1e-3 of the split epsilon; fewer than 3 vertices -> n = 0.
template <int N>
BTG_HD void cleanPoly(Pg<N>& p, float epsSplit) {
  const float e = epsSplit * 1e-3f;
  bool changed = true;
  while (changed && p.n >= 3) {
    changed = false;
    for (int i = 0; i < p.n && p.n >= 3; ++i) {
      const int h = i == 0 ? p.n - 1 : i - 1;
      const int j = i + 1 == p.n ? 0 : i + 1;
      const float ex = p.x[j] - p.x[h], ey = p.y[j] - p.y[h];
      const float dx = p.x[i] - p.x[h], dy = p.y[i] - p.y[h];
      const float l = __builtin_sqrtf(ex * ex + ey * ey);
      const float d2 = dx * dx + dy * dy;
      const float cr = __builtin_fabsf(ex * dy - ey * dx);
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

// classifyPoly (beam_core.h:68-78): bit 0 some vertex strictly inside, bit 1
// some vertex strictly outside.
template <int N>
BTG_HD int sidesOf(const Pg<N>& p, const Ln& L, float eps) {
  int c = 0;
  for (int i = 0; i < p.n; ++i) {
    const float d = L.at(p.x[i], p.y[i]);
    if (d > eps)
      c |= 1;
    else if (d < -eps)
      c |= 2;
  }
  return c;
}

// splitFive (beam_core.h:80-89): pentagon -> quad (0..3) + triangle (0, 3, 4).
template <int N>
BTG_HD void splitFive(const Pg<N>& p, Pg<N>& a, Pg<N>& b) {
  a.clear();
  b.clear();
  for (int i = 0; i < 4; ++i) a.push(p.x[i], p.y[i]);
  b.push(p.x[0], p.y[0]);
  b.push(p.x[3], p.y[3]);
  b.push(p.x[4], p.y[4]);
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
  BTG_HD static Plane make(F3 O, F3 p0, F3 u_, F3 v_) {
    Plane P;
    P.O = O;
    P.p0 = p0;
    F3 n = normalize(cross(u_, v_));
    if (dot(n, p0 - O) < 0) n = -n;
    P.u = normalize(u_);
    P.v = cross(n, P.u);
    P.n = n;
    P.d = p0 - O;
    return P;
  }
};

}  // namespace btg
