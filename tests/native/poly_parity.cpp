// Polygon primitives of the GPU library (beamtrace2007_amd/csrc/beam_dev.h,
// compiled here for the host) against the oracle's restatement of
// src/beam/beam_geom.h and beam_core.h (oracle/bt_oracle.cpp, itself checked
// bit for bit against the reference build): area, sidesOf, splitPoly,
// cleanPoly, splitFive, reverse and extent must agree bit for bit.
//
// Inputs are random convex polygons made adversarial on purpose: vertices
// snapped into the fuzzy band of the cutting line, duplicated within the weld
// distance, or placed on the segment between their neighbours, so that every
// branch of the weld and split logic runs.
//
// Build: hipcc -std=c++17 -O2 -ffp-contract=off (host code only; no device
// is used). Prints "ok <cases>" and exits 0, or prints the first mismatch.
#include <cmath>
#include <cstdio>
#include <cstring>

#include "../../beamtrace2007_amd/csrc/beam_dev.h"
#include "../../oracle/bt_oracle.cpp"

namespace {

struct Rng {
  uint64_t s;
  uint32_t next() {
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    return uint32_t(s >> 11);
  }
  float uni() { return float(next() & 0xFFFFFF) / float(0x1000000); }  // [0, 1)
  int below(int n) { return int(next() % uint32_t(n)); }
};

template <int N>
btg::Pg<N> toPg(const Poly& p) {
  btg::Pg<N> q;
  for (int i = 0; i < N; ++i) q.x[i] = q.y[i] = std::nanf("");
  q.n = p.n;
  for (int i = 0; i < p.n; ++i) {
    q.x[i] = p.x[i];
    q.y[i] = p.y[i];
  }
  return q;
}

bool sameBits(float a, float b) { return std::memcmp(&a, &b, sizeof a) == 0; }

template <int N>
bool same(const btg::Pg<N>& q, const Poly& p, const char* what, long caseNo) {
  bool ok = q.n == p.n;
  for (int i = 0; ok && i < p.n; ++i) ok = sameBits(q.x[i], p.x[i]) && sameBits(q.y[i], p.y[i]);
  if (!ok) {
    std::printf("mismatch in %s, case %ld: n %d vs %d\n", what, caseNo, q.n, p.n);
    for (int i = 0; i < N || i < p.n; ++i)
      std::printf("  %d: gpu (%.9g, %.9g)  oracle (%.9g, %.9g)\n", i, i < N ? q.x[i] : 0.f, i < N ? q.y[i] : 0.f,
                  i < p.n ? p.x[i] : 0.f, i < p.n ? p.y[i] : 0.f);
  }
  return ok;
}

// A random convex polygon with n vertices (n in 3..maxN), then perturbed.
Poly randomPoly(Rng& r, int n, const Line& L, float eps) {
  Poly p;
  const float cx = r.uni() * 2 - 1, cy = r.uni() * 2 - 1, rad = 0.01f + r.uni();
  float ang[8];
  for (int i = 0; i < n; ++i) ang[i] = r.uni() * 6.2831853f;
  for (int i = 1; i < n; ++i)
    for (int j = i; j > 0 && ang[j] < ang[j - 1]; --j) std::swap(ang[j], ang[j - 1]);
  for (int i = 0; i < n; ++i) p.push(cx + rad * std::cos(ang[i]), cy + rad * std::sin(ang[i]));
  const float e = eps * 1e-3f;
  for (int i = 0; i < n; ++i) {
    const int mode = r.below(10);
    const int h = (i + n - 1) % n;
    if (mode == 0) {  // into the fuzzy band of L
      const float d = L.at(p.x[i], p.y[i]);
      const float target = (r.uni() * 2 - 1) * eps * 1.5f;
      const float nn = L.a * L.a + L.b * L.b;
      if (nn > 0) {
        p.x[i] -= (d - target) * L.a / nn;
        p.y[i] -= (d - target) * L.b / nn;
      }
    } else if (mode == 1) {  // a near-duplicate of the previous vertex
      p.x[i] = p.x[h] + (r.uni() * 2 - 1) * e;
      p.y[i] = p.y[h] + (r.uni() * 2 - 1) * e;
    } else if (mode == 2) {  // on the segment between its neighbours
      const int j = (i + 1) % n;
      const float t = r.uni();
      p.x[i] = p.x[h] + t * (p.x[j] - p.x[h]);
      p.y[i] = p.y[h] + t * (p.y[j] - p.y[h]);
    } else if (mode == 3) {  // an exact duplicate
      p.x[i] = p.x[h];
      p.y[i] = p.y[h];
    }
  }
  return p;
}

}  // namespace

int main(int argc, char** argv) {
  const long cases = argc > 1 ? std::atol(argv[1]) : 200000;
  Rng r{0x9E3779B97F4A7C15ull};
  for (long k = 0; k < cases; ++k) {
    const float eps = std::ldexp(1.0f, -10 - r.below(12)) * (1 + r.uni());
    const Line L = lineOf(r.uni() * 2 - 1, r.uni() * 2 - 1, (r.uni() * 2 - 1) * 0.5f, 1.0f);
    const btg::Ln Lg{L.a, L.b, L.c};
    const int n = 3 + r.below(4);  // 3..6 (the split pieces' capacity)
    Poly p = randomPoly(r, n, L, eps);
    btg::Pg<6> g = toPg<6>(p);

    if (!sameBits(g.area(), p.area())) {
      std::printf("mismatch in area, case %ld\n", k);
      return 1;
    }
    if (!sameBits(g.extent(), p.extent())) {
      std::printf("mismatch in extent, case %ld\n", k);
      return 1;
    }
    if (btg::sidesOf(g, Lg, eps) != sides(p, L, eps)) {
      std::printf("mismatch in sidesOf, case %ld\n", k);
      return 1;
    }
    {
      Poly in, out;
      btg::Pg<6> gi, go;
      const bool ro = cutPoly(p, L, eps, in, out, 6);
      const bool rg = btg::splitPoly(g, Lg, eps, gi, go);
      if (ro != rg) {
        std::printf("mismatch in splitPoly result, case %ld\n", k);
        return 1;
      }
      if (ro && (!same(gi, in, "splitPoly in", k) || !same(go, out, "splitPoly out", k))) return 1;
      if (ro) {
        weld(in, eps);
        weld(out, eps);
        btg::cleanPoly(gi, eps);
        btg::cleanPoly(go, eps);
        if (!same(gi, in, "cleanPoly(in)", k) || !same(go, out, "cleanPoly(out)", k)) return 1;
        Poly in2 = in, out2 = out;
        weld(in2, eps);
        weld(out2, eps);
        if (!same(gi, in2, "cleanPoly(in) twice", k) || !same(go, out2, "cleanPoly(out) twice", k)) return 1;
      }
    }
    {
      Poly w = p;
      btg::Pg<6> gw = g;
      weld(w, eps);
      btg::cleanPoly(gw, eps);
      if (!same(gw, w, "cleanPoly", k)) return 1;
      // welding is idempotent (the split kernel skips re-welding welded
      // pieces): a second pass of the oracle's weld changes nothing
      Poly w2 = w;
      weld(w2, eps);
      if (!same(gw, w2, "cleanPoly twice", k)) return 1;
    }
    {
      Poly rv = p;
      btg::Pg<6> gr = g;
      rv.reverse();
      gr.reverse();
      if (!same(gr, rv, "reverse", k)) return 1;
    }
    if (n >= 5) {
      Poly five = p;
      five.n = 5;
      btg::Pg<6> g5 = toPg<6>(five);
      Poly a, b;
      btg::Pg<6> ga, gb;
      fiveToQuadTri(five, a, b);
      btg::splitFive(g5, ga, gb);
      if (!same(ga, a, "splitFive quad", k) || !same(gb, b, "splitFive tri", k)) return 1;
    }
    {
      // capacity 8 (the admit kernel's pieces) with a push past capacity
      Poly p8 = randomPoly(r, 3 + r.below(6), L, eps);
      btg::Pg<8> g8 = toPg<8>(p8);
      p8.push(0.25f, 0.5f);
      g8.push(0.25f, 0.5f);
      if (!same(g8, p8, "push (capacity 8)", k)) return 1;
      weld(p8, eps);
      btg::cleanPoly(g8, eps);
      if (!same(g8, p8, "cleanPoly (capacity 8)", k)) return 1;
    }
  }
  std::printf("ok %ld\n", cases);
  return 0;
}
