// The GPU library's FP64 soft-shadow tracer (beamtrace2007_amd/csrc/trace_fp64.h,
// btg_options.precision = 2) compiled for the host, as a shared library that
// tests/test_fp64_host.py calls through ctypes: the source the device runs,
// checked point by point on the CPU against the reference's double build
// (ref_light_vis) and run with host bounds checks.
//
// Build (host code only; no device is used):
//   hipcc -x hip --offload-arch=gfx950 -std=c++17 -O1 -ffp-contract=off -fPIC -shared
#include <cstdint>
#include <cstring>
#include <memory>
#include <vector>

#define FP64_HD __host__ __device__
#include "../../beamtrace2007_amd/csrc/engine.h"

namespace btg {

// The per-thread counters the tracer increments.
struct Cnt {
  uint32_t v[sNum];
};

#include "../../beamtrace2007_amd/csrc/trace_fp64.h"

}  // namespace btg

using namespace btg;

namespace {

// Shades n points with tracers of First's pools, traced again with Second's
// when those overflow (the kernel's two passes).
template <class First, class Second>
int shadeAll(const void* refs, const void* nodes, const float* nodeBox, const uint32_t* parent, const float* bounds6,
             const float* light9, double offset, int exact, const double* x, const int32_t* ref, const float* nf, int n,
             double* vis, double* E, double* Eu, uint64_t* counters) {
  DevScene sv;
  sv.nodes = static_cast<const KdNodeD*>(nodes);
  sv.nodeBox = reinterpret_cast<const Box*>(nodeBox);
  sv.parent = parent;
  sv.refs = static_cast<const LeafTri*>(refs);
  Box root;
  root.lo = f3(bounds6[0], bounds6[1], bounds6[2]);
  root.hi = f3(bounds6[3], bounds6[4], bounds6[5]);
  const fp64::V c = fp64::vmk(light9[0], light9[1], light9[2]), U = fp64::vmk(light9[3], light9[4], light9[5]),
                Vv = fp64::vmk(light9[6], light9[7], light9[8]);
  // each tracer's arrays in a region of exactly its size (one thread)
  std::vector<uint64_t> firstMem(First::kWords), secondMem(Second::kWords);
  auto first = std::make_unique<First>();
  auto second = std::make_unique<Second>();
  first->m = fp64::WordMem{firstMem.data(), 1, 0};
  second->m = fp64::WordMem{secondMem.data(), 1, 0};
  Cnt cnt;
  std::memset(&cnt, 0, sizeof cnt);
  int retried = 0;
  for (int i = 0; i < n; ++i) {
    const fp64::V xi = fp64::vmk(x[3 * i], x[3 * i + 1], x[3 * i + 2]);
    const fp64::V ni = fp64::vmk(nf[3 * i], nf[3 * i + 1], nf[3 * i + 2]);
    const LeafTri& r = sv.refs[ref[i]];
    if (fp64::shadePoint(*first, sv, xi, r, ni, c, U, Vv, offset, exact != 0, 0u, root, cnt, vis[i], E[i], Eu[i]))
      continue;
    ++retried;
    if (!fp64::shadePoint(*second, sv, xi, r, ni, c, U, Vv, offset, exact != 0, 0u, root, cnt, vis[i], E[i], Eu[i]))
      return -1 - i;
  }
  for (int k = 0; k < sNum; ++k) counters[k] = cnt.v[k];
  return retried;
}

}  // namespace

extern "C" {

// FP64 soft shadows of n receiver points, as the library's FP64 kernel shades
// a pixel (pools of 64 beams, then 384). refs: the leaf references in leaf
// order (64-byte records); nodes (12-byte), nodeBox (6 floats), parent per
// node; bounds6 the kd root box. light9: c, U, V. Per point: x (3 doubles),
// ref (index into refs: the hit triangle), nf (3 floats, the triangle normal
// facing the eye). Writes vis, E and Eu; counters[sNum] (summed). Returns the
// number of points traced again with the larger pools, or -1 - i when point i
// overflowed those too.
int fp64_soft(const void* refs, const void* nodes, const float* nodeBox, const uint32_t* parent, const float* bounds6,
              const float* light9, double offset, int exact, const double* x, const int32_t* ref, const float* nf,
              int n, double* vis, double* E, double* Eu, uint64_t* counters) {
  return shadeAll<fp64::Tracer<64, 64>, fp64::Tracer<384, 256>>(refs, nodes, nodeBox, parent, bounds6, light9, offset,
                                                                exact, x, ref, nf, n, vis, E, Eu, counters);
}

// The same with first pools of 4 beams and 4 work items, so that most points
// overflow and take the second pass.
int fp64_soft_tiny(const void* refs, const void* nodes, const float* nodeBox, const uint32_t* parent,
                   const float* bounds6, const float* light9, double offset, int exact, const double* x,
                   const int32_t* ref, const float* nf, int n, double* vis, double* E, double* Eu,
                   uint64_t* counters) {
  return shadeAll<fp64::Tracer<4, 4>, fp64::Tracer<384, 256>>(refs, nodes, nodeBox, parent, bounds6, light9, offset,
                                                              exact, x, ref, nf, n, vis, E, Eu, counters);
}

}  // extern "C"
