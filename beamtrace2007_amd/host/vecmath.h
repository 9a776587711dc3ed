// Float 3-vectors and boxes for the host side of the beam tracer.
//
// Every operation rounds exactly as the reference's Vec3 / AABB
// (src/core/vec.h:8-65) do in its float build: component-wise IEEE single
// operations, evaluated left to right, no fused multiply-add (the library is
// built with -ffp-contract=off, as the reference is, CMakeLists.txt:17-21).
// min/max are the "a < b ? a : b" selects of the reference, not fminf/fmaxf,
// so the sign of zero and NaN propagation match as well.
#pragma once

#include <cmath>
#include <cstdint>
#include <limits>

namespace bth {

constexpr float kInfF = std::numeric_limits<float>::infinity();

struct V3 {
  float x = 0, y = 0, z = 0;
  V3() = default;
  V3(float a, float b, float c) : x(a), y(b), z(c) {}
  float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
  float& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};

inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
inline V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator/(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline float dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross3(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float len3(V3 a) { return std::sqrt(dot3(a, a)); }
inline V3 unit3(V3 a) {
  float l = len3(a);
  return l > 0 ? a / l : a;
}
inline float selMin(float a, float b) { return a < b ? a : b; }
inline float selMax(float a, float b) { return a > b ? a : b; }
inline V3 vmin3(V3 a, V3 b) { return {selMin(a.x, b.x), selMin(a.y, b.y), selMin(a.z, b.z)}; }
inline V3 vmax3(V3 a, V3 b) { return {selMax(a.x, b.x), selMax(a.y, b.y), selMax(a.z, b.z)}; }

// A unit vector orthogonal to unit n (the axis picked as in vec.h:41-44).
inline V3 orthoUnit(V3 n) {
  V3 a = std::fabs(n.x) < 0.6f ? V3(1, 0, 0) : (std::fabs(n.y) < 0.6f ? V3(0, 1, 0) : V3(0, 0, 1));
  return unit3(cross3(n, a));
}

struct Box3 {
  V3 lo{kInfF, kInfF, kInfF};
  V3 hi{-kInfF, -kInfF, -kInfF};
  void grow(V3 p) {
    lo = vmin3(lo, p);
    hi = vmax3(hi, p);
  }
  V3 diag() const { return hi - lo; }
  // 2 (dx dy + dy dz + dz dx), summed in that order.
  float area() const {
    V3 d = diag();
    return 2 * (d.x * d.y + d.y * d.z + d.z * d.x);
  }
};

}  // namespace bth
