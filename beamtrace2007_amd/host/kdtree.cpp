// SAH kd-tree build, bit-identical to the reference's KdTree::build
// (src/accel/kdtree.cpp:135-281) for the same scene and parameters:
//   * node order: depth-first, left subtree before right (node i's left child
//     is i + 1, its right child index is stored in `data`);
//   * split search per axis 0, 1, 2: 64 bins above 2048 references, else an
//     exact sweep over sorted (position, end < planar < start) events; the
//     first strictly cheaper candidate wins (kdtree.cpp:148-228);
//   * SAH arithmetic in float in the reference's operation order
//     (kdtree.cpp:148-156); straddling triangles get bounds clipped in double
//     and rounded outward (kdtree.cpp:21-77);
//   * leaves list triangles by area, largest first, ties by id (135-146);
//   * node boxes and parent links by an explicit descent (261-279).
#include <algorithm>
#include <cmath>
#include <limits>

#include "scene.h"

namespace bth {

Box3 clipTriBounds(const Scene& s, uint32_t t, const Box3& bx, const Box3& whole) {
  double P[2][12][3];
  int n = 3, cur = 0;
  for (int k = 0; k < 3; ++k) {
    const V3 p = s.v(int(t), k);
    P[0][k][0] = p.x;
    P[0][k][1] = p.y;
    P[0][k][2] = p.z;
  }
  // Sutherland-Hodgman against lo then hi of each axis, keeping side >= 0.
  for (int axis = 0; axis < 3 && n > 0; ++axis) {
    for (int side = 0; side < 2 && n > 0; ++side) {
      const double plane = side == 0 ? double(bx.lo[axis]) : double(bx.hi[axis]);
      const double sg = side == 0 ? 1.0 : -1.0;
      int m = 0;
      const int nxt = 1 - cur;
      for (int i = 0; i < n; ++i) {
        const double* a = P[cur][i];
        const double* b = P[cur][(i + 1) % n];
        const double da = sg * (a[axis] - plane), db = sg * (b[axis] - plane);
        if (da >= 0) {
          P[nxt][m][0] = a[0];
          P[nxt][m][1] = a[1];
          P[nxt][m][2] = a[2];
          ++m;
        }
        if ((da >= 0) != (db >= 0)) {
          const double f = da / (da - db);
          for (int c = 0; c < 3; ++c) P[nxt][m][c] = a[c] + f * (b[c] - a[c]);
          P[nxt][m][axis] = plane;
          ++m;
        }
      }
      n = m;
      cur = nxt;
    }
  }
  if (n == 0) return whole;
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int i = 0; i < n; ++i)
    for (int c = 0; c < 3; ++c) {
      lo[c] = std::min(lo[c], P[cur][i][c]);
      hi[c] = std::max(hi[c], P[cur][i][c]);
    }
  Box3 r;
  const double ulp = double(std::numeric_limits<float>::epsilon());
  for (int c = 0; c < 3; ++c) {
    const double pad = 4 * ulp * (std::fabs(lo[c]) + std::fabs(hi[c]) + double(bx.hi[c] - bx.lo[c]));
    r.lo[c] = float(lo[c] - pad);
    r.hi[c] = float(hi[c] + pad);
  }
  r.lo = vmax3(r.lo, vmax3(bx.lo, whole.lo));
  r.hi = vmin3(r.hi, vmin3(bx.hi, whole.hi));
  for (int a = 0; a < 3; ++a)
    if (r.lo[a] > r.hi[a]) r.lo[a] = r.hi[a];
  return r;
}

namespace {

struct Ref {
  uint32_t tri;
  Box3 b;
};

struct Event {
  float pos;
  int kind;  // 0 end, 1 planar, 2 start
};

class SahBuilder {
 public:
  SahBuilder(const Scene& s, const KdParams& p, KdTree& t, int maxDepth) : s_(s), p_(p), t_(t), maxDepth_(maxDepth) {}

  void node(std::vector<Ref>& refs, const Box3& bx, int depth) {
    const size_t self = t_.nodes.size();
    t_.nodes.push_back(KdNode());
    t_.depth = std::max(t_.depth, depth);
    const size_t N = refs.size();
    int axis = -1;
    float split = 0, cost = kInfF;
    if (!p_.singleLeaf && N > 1 && depth < maxDepth_) search(refs, bx, axis, split, cost);
    if (axis < 0 || !(cost < p_.costIntersect * float(N))) {
      leaf(self, refs);
      return;
    }
    Box3 lb = bx, rb = bx;
    lb.hi[axis] = split;
    rb.lo[axis] = split;
    std::vector<Ref> L, R;
    L.reserve(N / 2 + 1);
    R.reserve(N / 2 + 1);
    for (const Ref& r : refs) {
      const float lo = r.b.lo[axis], hi = r.b.hi[axis];
      const bool planar = lo == hi && lo == split;
      bool toL = planar || lo < split, toR = planar || hi > split;
      if (!toL && !toR) toL = toR = true;
      if (toL && toR && !planar) {
        L.push_back({r.tri, clipTriBounds(s_, r.tri, lb, r.b)});
        R.push_back({r.tri, clipTriBounds(s_, r.tri, rb, r.b)});
      } else {
        if (toL) L.push_back(r);
        if (toR) R.push_back(r);
      }
    }
    std::vector<Ref>().swap(refs);
    node(L, lb, depth + 1);
    std::vector<Ref>().swap(L);
    const size_t right = t_.nodes.size();
    node(R, rb, depth + 1);
    t_.nodes[self].split = split;
    t_.nodes[self].data = (uint32_t(right) << 2) | uint32_t(axis);
  }

 private:
  void leaf(size_t self, std::vector<Ref>& refs) {
    std::sort(refs.begin(), refs.end(), [&](const Ref& a, const Ref& b) {
      const float aa = s_.triArea[a.tri], ab = s_.triArea[b.tri];
      return aa != ab ? aa > ab : a.tri < b.tri;
    });
    t_.nodes[self].offset = uint32_t(t_.triIndices.size());
    t_.nodes[self].data = (uint32_t(refs.size()) << 2) | 3u;
    for (const Ref& r : refs) t_.triIndices.push_back(r.tri);
  }

  // ct + ci * (A_l / A * n_l + A_r / A * n_r), times the empty bonus.
  float sah(const Box3& bx, int axis, float split, size_t nl, size_t nr) const {
    Box3 l = bx, r = bx;
    l.hi[axis] = split;
    r.lo[axis] = split;
    const float inv = 1 / bx.area();
    float c = p_.costTraverse + p_.costIntersect * (l.area() * inv * float(nl) + r.area() * inv * float(nr));
    if (nl == 0 || nr == 0) c *= p_.emptyBonus;
    return c;
  }

  void search(const std::vector<Ref>& refs, const Box3& bx, int& bestAxis, float& bestSplit, float& bestCost) {
    const size_t N = refs.size();
    if (!(bx.area() > 0)) return;
    for (int axis = 0; axis < 3; ++axis) {
      const float lo = bx.lo[axis], hi = bx.hi[axis];
      if (!(hi > lo)) continue;
      if (N > size_t(p_.binThreshold)) {
        const int B = p_.bins;
        std::vector<size_t> first(size_t(B), 0), last(size_t(B), 0);
        const float scale = float(B) / (hi - lo);
        for (const Ref& r : refs) {
          const int bs = std::min(std::max(int((r.b.lo[axis] - lo) * scale), 0), B - 1);
          const int be = std::min(std::max(int((r.b.hi[axis] - lo) * scale), 0), B - 1);
          first[size_t(bs)]++;
          last[size_t(be)]++;
        }
        size_t nl = 0, nr = N;
        for (int k = 1; k < B; ++k) {
          nl += first[size_t(k - 1)];
          nr -= last[size_t(k - 1)];
          const float split = lo + (hi - lo) * float(k) / float(B);
          const float c = sah(bx, axis, split, nl, nr);
          if (c < bestCost) {
            bestCost = c;
            bestAxis = axis;
            bestSplit = split;
          }
        }
        continue;
      }
      ev_.clear();
      ev_.reserve(2 * N);
      for (const Ref& r : refs) {
        const float a = r.b.lo[axis], b = r.b.hi[axis];
        if (a == b) {
          ev_.push_back({a, 1});
        } else {
          ev_.push_back({a, 2});
          ev_.push_back({b, 0});
        }
      }
      std::sort(ev_.begin(), ev_.end(),
                [](const Event& x, const Event& y) { return x.pos != y.pos ? x.pos < y.pos : x.kind < y.kind; });
      size_t nl = 0, nr = N;
      for (size_t i = 0; i < ev_.size();) {
        const float p = ev_[i].pos;
        size_t ne = 0, np = 0, ns = 0;
        while (i < ev_.size() && ev_[i].pos == p && ev_[i].kind == 0) ++ne, ++i;
        while (i < ev_.size() && ev_[i].pos == p && ev_[i].kind == 1) ++np, ++i;
        while (i < ev_.size() && ev_[i].pos == p && ev_[i].kind == 2) ++ns, ++i;
        nr -= ne + np;
        if (p > lo && p < hi) {
          const float c = sah(bx, axis, p, nl + np, nr + np);
          if (c < bestCost) {
            bestCost = c;
            bestAxis = axis;
            bestSplit = p;
          }
        }
        nl += ns + np;
      }
    }
  }

  const Scene& s_;
  const KdParams& p_;
  KdTree& t_;
  int maxDepth_;
  std::vector<Event> ev_;
};

}  // namespace

void KdTree::build(const Scene& s, const KdParams& p) {
  nodes.clear();
  triIndices.clear();
  depth = 0;
  bounds = s.bounds;
  const V3 pad = bounds.diag() * 1e-4f + V3(1e-6f, 1e-6f, 1e-6f);
  bounds.lo = bounds.lo - pad;
  bounds.hi = bounds.hi + pad;
  std::vector<Ref> refs(size_t(s.numTris()));
  for (int t = 0; t < s.numTris(); ++t) {
    Box3 b;
    for (int k = 0; k < 3; ++k) b.grow(s.v(t, k));
    refs[size_t(t)] = {uint32_t(t), b};
  }
  const int N = std::max(1, s.numTris());
  int md = p.maxDepth >= 0 ? p.maxDepth : int(8 + 1.3 * std::log2(double(N)));
  md = std::min(md, 63);  // inner nodes at depth <= 62: a 64-bit restart trail suffices
  SahBuilder(s, p, *this, md).node(refs, bounds, 0);
  parent.assign(nodes.size(), ~0u);
  nodeBox.assign(nodes.size(), Box3());
  nodeBox[0] = bounds;
  std::vector<uint32_t> todo(1, 0u);
  while (!todo.empty()) {
    const uint32_t i = todo.back();
    todo.pop_back();
    const KdNode& nd = nodes[i];
    if (nd.leaf()) continue;
    const int a = nd.axis();
    const uint32_t l = i + 1, r = nd.hiBits();
    nodeBox[l] = nodeBox[i];
    nodeBox[l].hi[a] = nd.split;
    nodeBox[r] = nodeBox[i];
    nodeBox[r].lo[a] = nd.split;
    parent[l] = parent[r] = i;
    todo.push_back(l);
    todo.push_back(r);
  }
}

}  // namespace bth
