// SAH kd-tree for the host input stage.
//
// The tree must come out bit for bit as the reference's KdTree::build
// (src/accel/kdtree.cpp:135-281) makes it from the same triangles, because the
// beams' kd steps, leaf visits and climbs are part of the parity contract.
// What that fixes, and all this file shares with it:
//   * the float arithmetic of the root padding, the SAH cost and the binned
//     split positions, and the double arithmetic of the clipped bounds of
//     straddling triangles (Sutherland-Hodgman against the child cell, then
//     rounded outward by 4 float ulps), each in the reference's operation order;
//   * the decisions: candidate planes per axis 0, 1, 2 in increasing position,
//     the first strictly cheaper one wins; 64 bins above 2048 references,
//     otherwise every distinct start / end / flat position; a node becomes a
//     leaf unless its best split is cheaper than intersecting everything;
//     which side a reference goes to; leaves sorted by area, largest first,
//     ties by triangle id;
//   * the node numbering: depth first, left subtree first (left child = i + 1).
// The program is this project's own: an explicit task stack instead of
// recursion, references as structure-of-arrays bounds, an exact sweep over
// three separately sorted position lists merged on the fly instead of one
// sorted event list, node boxes and parent links recorded as nodes are made.
//
// Signed zeros: the reference groups event positions with float ==, so -0
// and +0 fall into one group whose stored split takes the sign of whichever
// its unstable sort puts first. Here a group's value is the smallest head of
// the three lists. The two agree whenever no coordinate is -0, which holds for
// every scene the generators make (their trees are pinned by digests).
#include <algorithm>
#include <array>
#include <cmath>
#include <limits>

#include "scene.h"

namespace bth {
namespace {

// References (triangle ids with the bounds they occupy inside a cell) as
// structure of arrays.
struct RefSet {
  std::vector<uint32_t> tri;
  std::array<std::vector<float>, 3> lo, hi;

  size_t size() const { return tri.size(); }
  void reserve(size_t n) {
    tri.reserve(n);
    for (int a = 0; a < 3; ++a) {
      lo[size_t(a)].reserve(n);
      hi[size_t(a)].reserve(n);
    }
  }
  void add(uint32_t t, const Box3& b) {
    tri.push_back(t);
    for (int a = 0; a < 3; ++a) {
      lo[size_t(a)].push_back(b.lo[a]);
      hi[size_t(a)].push_back(b.hi[a]);
    }
  }
  Box3 box(size_t i) const {
    Box3 b;
    for (int a = 0; a < 3; ++a) {
      b.lo[a] = lo[size_t(a)][i];
      b.hi[a] = hi[size_t(a)][i];
    }
    return b;
  }
  void release() {
    RefSet none;
    std::swap(*this, none);
  }
};

// A triangle cut down to an axis-aligned cell in double precision: each of the
// six faces (low then high, axes x, y, z) removes the part outside it; a
// crossing edge gets the point a + f (b - a), f = da / (da - db), with its
// coordinate on the face set to the face exactly.
class CellClip {
 public:
  CellClip(const Scene& s, uint32_t t) {
    for (int k = 0; k < 3; ++k) {
      const V3 p = s.v(int(t), k);
      pts_[0][size_t(k)] = {double(p.x), double(p.y), double(p.z)};
    }
    n_ = 3;
  }

  void clipTo(const Box3& cell) {
    for (int a = 0; a < 3 && n_ > 0; ++a) {
      face(a, double(cell.lo[a]), 1.0);
      if (n_ > 0) face(a, double(cell.hi[a]), -1.0);
    }
  }

  // Bounds of what is left, rounded outward by 4 float ulps of the
  // coordinates' and the cell's size, never beyond the cell or `whole`.
  Box3 bounds(const Box3& cell, const Box3& whole) const {
    if (n_ == 0) return whole;
    const auto& P = pts_[size_t(cur_)];
    std::array<double, 3> lo{1e300, 1e300, 1e300}, hi{-1e300, -1e300, -1e300};
    for (int i = 0; i < n_; ++i)
      for (size_t c = 0; c < 3; ++c) {
        lo[c] = std::min(lo[c], P[size_t(i)][c]);
        hi[c] = std::max(hi[c], P[size_t(i)][c]);
      }
    const double ulp = double(std::numeric_limits<float>::epsilon());
    Box3 r;
    for (int c = 0; c < 3; ++c) {
      const double pad =
          4 * ulp * (std::fabs(lo[size_t(c)]) + std::fabs(hi[size_t(c)]) + double(cell.hi[c] - cell.lo[c]));
      r.lo[c] = float(lo[size_t(c)] - pad);
      r.hi[c] = float(hi[size_t(c)] + pad);
    }
    r.lo = vmax3(r.lo, vmax3(cell.lo, whole.lo));
    r.hi = vmin3(r.hi, vmin3(cell.hi, whole.hi));
    for (int a = 0; a < 3; ++a)
      if (r.lo[a] > r.hi[a]) r.lo[a] = r.hi[a];
    return r;
  }

 private:
  using Pt = std::array<double, 3>;
  // A triangle cut by six planes keeps at most 9 corners; room is left for
  // the extra corners rounding can add to a nearly degenerate piece.
  static constexpr int kCap = 24;

  // Keeps the part where side * (x[axis] - plane) >= 0.
  void face(int axis, double plane, double side) {
    const auto& P = pts_[size_t(cur_)];
    auto& Q = pts_[size_t(1 - cur_)];
    int m = 0;
    for (int i = 0; i < n_ && m + 2 <= kCap; ++i) {
      const Pt& a = P[size_t(i)];
      const Pt& b = P[size_t(i + 1 == n_ ? 0 : i + 1)];
      const double da = side * (a[size_t(axis)] - plane), db = side * (b[size_t(axis)] - plane);
      const bool ina = da >= 0;
      if (ina) Q[size_t(m++)] = a;
      if (ina != (db >= 0)) {
        const double f = da / (da - db);
        Pt x;
        for (size_t c = 0; c < 3; ++c) x[c] = a[c] + f * (b[c] - a[c]);
        x[size_t(axis)] = plane;
        Q[size_t(m++)] = x;
      }
    }
    n_ = m;
    cur_ = 1 - cur_;
  }

  std::array<std::array<Pt, kCap>, 2> pts_;
  int n_ = 0, cur_ = 0;
};

// The best split found so far over all axes of a cell.
struct Best {
  int axis = -1;
  float pos = 0;
  float cost = kInfF;
};

// Work item: the references of one node-to-be with its cell. `rightOf` is the
// inner node whose right child this becomes (-1: the root or a left child,
// whose index needs no patching).
struct Task {
  RefSet refs;
  Box3 cell;
  int depth;
  int64_t rightOf;
  uint32_t parent;
};

class KdBuild {
 public:
  KdBuild(const Scene& s, const KdParams& p, KdTree& t, int maxDepth) : s_(s), p_(p), t_(t), maxDepth_(maxDepth) {}

  void run(RefSet all, const Box3& rootCell) {
    std::vector<Task> todo;
    todo.push_back(Task{std::move(all), rootCell, 0, -1, ~0u});
    while (!todo.empty()) {
      Task task = std::move(todo.back());
      todo.pop_back();
      const uint32_t self = uint32_t(t_.nodes.size());
      t_.nodes.emplace_back();
      t_.nodeBox.push_back(task.cell);
      t_.parent.push_back(task.parent);
      if (task.rightOf >= 0) t_.nodes[size_t(task.rightOf)].data |= self << 2;
      t_.depth = std::max(t_.depth, task.depth);
      const size_t N = task.refs.size();
      Best best;
      if (!p_.singleLeaf && N > 1 && task.depth < maxDepth_) findSplit(task.refs, task.cell, best);
      if (best.axis < 0 || !(best.cost < p_.costIntersect * float(N))) {
        makeLeaf(self, task.refs);
        continue;
      }
      Task left{RefSet(), task.cell, task.depth + 1, -1, self};
      Task right{RefSet(), task.cell, task.depth + 1, int64_t(self), self};
      left.cell.hi[best.axis] = best.pos;
      right.cell.lo[best.axis] = best.pos;
      distribute(task.refs, best, left, right);
      task.refs.release();
      KdNode& nd = t_.nodes[self];
      nd.split = best.pos;
      nd.data = uint32_t(best.axis);  // the right child's index is or-ed in when it is made
      todo.push_back(std::move(right));
      todo.push_back(std::move(left));  // popped first: the left subtree takes the next indices
    }
  }

 private:
  void makeLeaf(uint32_t self, const RefSet& refs) {
    std::vector<uint32_t> ids(refs.tri);
    const std::vector<float>& area = s_.triArea;
    std::sort(ids.begin(), ids.end(), [&area](uint32_t a, uint32_t b) {
      const float fa = area[a], fb = area[b];
      if (fa != fb) return fa > fb;
      return a < b;
    });
    KdNode& nd = t_.nodes[self];
    nd.offset = uint32_t(t_.triIndices.size());
    nd.data = (uint32_t(ids.size()) << 2) | 3u;
    t_.triIndices.insert(t_.triIndices.end(), ids.begin(), ids.end());
  }

  // Each reference goes to the side(s) its bounds reach; one lying flat in the
  // plane goes to both; one that straddles it is clipped to each child cell.
  void distribute(const RefSet& refs, const Best& best, Task& left, Task& right) {
    const size_t N = refs.size();
    left.refs.reserve(N / 2 + 1);
    right.refs.reserve(N / 2 + 1);
    const std::vector<float>& lo = refs.lo[size_t(best.axis)];
    const std::vector<float>& hi = refs.hi[size_t(best.axis)];
    const float s = best.pos;
    for (size_t i = 0; i < N; ++i) {
      const bool flat = lo[i] == hi[i] && lo[i] == s;
      bool l = flat || lo[i] < s, r = flat || hi[i] > s;
      if (!l && !r) l = r = true;
      const Box3 b = refs.box(i);
      if (l && r && !flat) {
        CellClip cl(s_, refs.tri[i]), cr(s_, refs.tri[i]);
        cl.clipTo(left.cell);
        cr.clipTo(right.cell);
        left.refs.add(refs.tri[i], cl.bounds(left.cell, b));
        right.refs.add(refs.tri[i], cr.bounds(right.cell, b));
      } else {
        if (l) left.refs.add(refs.tri[i], b);
        if (r) right.refs.add(refs.tri[i], b);
      }
    }
  }

  // Surface area heuristic of a plane at `pos` with nl references on the left
  // and nr on the right, the empty-side bonus applied.
  float cost(const Box3& cell, int axis, float pos, size_t nl, size_t nr) const {
    Box3 l = cell, r = cell;
    l.hi[axis] = pos;
    r.lo[axis] = pos;
    const float inv = 1 / cell.area();
    float c = p_.costTraverse + p_.costIntersect * (l.area() * inv * float(nl) + r.area() * inv * float(nr));
    if (nl == 0 || nr == 0) c *= p_.emptyBonus;
    return c;
  }

  void consider(const Box3& cell, int axis, float pos, size_t nl, size_t nr, Best& best) const {
    const float c = cost(cell, axis, pos, nl, nr);
    if (c < best.cost) {
      best.cost = c;
      best.axis = axis;
      best.pos = pos;
    }
  }

  void findSplit(const RefSet& refs, const Box3& cell, Best& best) {
    if (!(cell.area() > 0)) return;
    for (int axis = 0; axis < 3; ++axis) {
      if (!(cell.hi[axis] > cell.lo[axis])) continue;
      if (refs.size() > size_t(p_.binThreshold))
        binned(refs, cell, axis, best);
      else
        sweep(refs, cell, axis, best);
    }
  }

  // Planes at the bin edges lo + (hi - lo) k / B, k = 1..B-1; references
  // counted by the bins of their low and high bounds.
  void binned(const RefSet& refs, const Box3& cell, int axis, Best& best) {
    const int B = p_.bins;
    const float lo = cell.lo[axis], hi = cell.hi[axis];
    const float scale = float(B) / (hi - lo);
    startsIn_.assign(size_t(B), 0);
    endsIn_.assign(size_t(B), 0);
    auto binOf = [&](float v) { return std::min(std::max(int((v - lo) * scale), 0), B - 1); };
    const std::vector<float>& rl = refs.lo[size_t(axis)];
    const std::vector<float>& rh = refs.hi[size_t(axis)];
    for (size_t i = 0; i < refs.size(); ++i) {
      startsIn_[size_t(binOf(rl[i]))]++;
      endsIn_[size_t(binOf(rh[i]))]++;
    }
    size_t nl = 0, nr = refs.size();
    for (int k = 1; k < B; ++k) {
      nl += startsIn_[size_t(k - 1)];
      nr -= endsIn_[size_t(k - 1)];
      consider(cell, axis, lo + (hi - lo) * float(k) / float(B), nl, nr, best);
    }
  }

  // Every distinct position where a reference starts, ends or lies flat, in
  // increasing order: at a position the references ending or flat there leave
  // the right side, flat ones count on both sides of a plane there, and those
  // starting or flat there join the left side after it.
  void sweep(const RefSet& refs, const Box3& cell, int axis, Best& best) {
    starts_.clear();
    ends_.clear();
    flats_.clear();
    const std::vector<float>& rl = refs.lo[size_t(axis)];
    const std::vector<float>& rh = refs.hi[size_t(axis)];
    for (size_t i = 0; i < refs.size(); ++i) {
      if (rl[i] == rh[i]) {
        flats_.push_back(rl[i]);
      } else {
        starts_.push_back(rl[i]);
        ends_.push_back(rh[i]);
      }
    }
    std::sort(starts_.begin(), starts_.end());
    std::sort(ends_.begin(), ends_.end());
    std::sort(flats_.begin(), flats_.end());
    const float lo = cell.lo[axis], hi = cell.hi[axis];
    size_t is = 0, ie = 0, iff = 0;
    size_t nl = 0, nr = refs.size();
    for (;;) {
      const bool hs = is < starts_.size(), he = ie < ends_.size(), hf = iff < flats_.size();
      if (!hs && !he && !hf) break;
      float p = kInfF;
      if (he) p = ends_[ie];
      if (hf && (!he || flats_[iff] < p)) p = flats_[iff];
      if (hs && (!(he || hf) || starts_[is] < p)) p = starts_[is];
      size_t ne = 0, nf = 0, ns = 0;
      while (ie < ends_.size() && ends_[ie] == p) ++ie, ++ne;
      while (iff < flats_.size() && flats_[iff] == p) ++iff, ++nf;
      while (is < starts_.size() && starts_[is] == p) ++is, ++ns;
      nr -= ne + nf;
      if (p > lo && p < hi) consider(cell, axis, p, nl + nf, nr + nf, best);
      nl += ns + nf;
    }
  }

  const Scene& s_;
  const KdParams& p_;
  KdTree& t_;
  int maxDepth_;
  std::vector<size_t> startsIn_, endsIn_;
  std::vector<float> starts_, ends_, flats_;
};

}  // namespace

void KdTree::build(const Scene& s, const KdParams& p) {
  nodes.clear();
  triIndices.clear();
  parent.clear();
  nodeBox.clear();
  depth = 0;
  // the root cell: the scene's bounds grown by 1e-4 of the diagonal plus 1e-6
  bounds = s.bounds;
  const V3 pad = bounds.diag() * 1e-4f + V3(1e-6f, 1e-6f, 1e-6f);
  bounds.lo = bounds.lo - pad;
  bounds.hi = bounds.hi + pad;
  RefSet all;
  all.reserve(size_t(s.numTris()));
  for (int t = 0; t < s.numTris(); ++t) {
    Box3 b;
    for (int k = 0; k < 3; ++k) b.grow(s.v(t, k));
    all.add(uint32_t(t), b);
  }
  // 8 + 1.3 log2(N) levels unless given, at most 63, so that inner nodes sit
  // at depth 62 or less and the device's 64-bit restart trail holds a path
  const int n = std::max(1, s.numTris());
  const int levels = std::min(p.maxDepth >= 0 ? p.maxDepth : int(8 + 1.3 * std::log2(double(n))), 63);
  KdBuild(s, p, *this, levels).run(std::move(all), bounds);
}

}  // namespace bth
