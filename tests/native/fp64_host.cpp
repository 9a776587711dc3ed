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

// Shades n points as the library's FP64 pass does (levels of cells, each
// level with its pools; the first two with the work cap), summing each
// point's cells in the same fixed point. levels[0..1] counts the cells that
// went on to levels 1 and 2.
template <class T0, class T1, class T2>
int shadeAll(const void* refs, const void* nodes, const float* nodeBox, const uint32_t* parent, const float* bounds6,
             const float* light9, double offset, int exact, const double* x, const int32_t* ref, const float* nf, int n,
             uint32_t workCap, double* vis, double* E, double* Eu, uint64_t* counters, int64_t* levels) {
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
  std::vector<uint64_t> mem0(T0::kWords), mem1(T1::kWords), mem2(T2::kWords);
  auto t0 = std::make_unique<T0>();
  auto t1 = std::make_unique<T1>();
  auto t2 = std::make_unique<T2>();
  t0->m = fp64::WordMem{mem0.data(), 1, 0};
  t1->m = fp64::WordMem{mem1.data(), 1, 0};
  t2->m = fp64::WordMem{mem2.data(), 1, 0};
  std::memset(&t0->cnt, 0, sizeof(Cnt));
  std::memset(&t1->cnt, 0, sizeof(Cnt));
  std::memset(&t2->cnt, 0, sizeof(Cnt));
  levels[0] = levels[1] = 0;
  for (int i = 0; i < n; ++i) {
    const fp64::V xi = fp64::vmk(x[3 * i], x[3 * i + 1], x[3 * i + 2]);
    const fp64::V ni = fp64::vmk(nf[3 * i], nf[3 * i + 1], nf[3 * i + 2]);
    const LeafTri& r = sv.refs[ref[i]];
    fp64::PixelRoot R;
    fp64::pixelRoot(r, xi, ni, c, U, Vv, offset, R);
    Eu[i] = R.A * R.G0;
    vis[i] = E[i] = 0;
    if (R.root.n < 3) continue;
    unsigned long long vs = 0, gs = 0;
    double miss, G;
    const auto add = [&] {
      vs += fp64::visFx(miss / R.lightArea);
      if (exact) gs += fp64::gFx(G);
    };
    if (fp64::traceCell(*t0, sv, R, r.tri, exact != 0, 0, 0, 0, workCap, 0u, root, miss, G)) {
      add();
    } else {
      ++levels[0];
      for (int c1 = 0; c1 < 16; ++c1) {
        const int cx1 = c1 & 3, cy1 = c1 >> 2;
        if (fp64::traceCell(*t1, sv, R, r.tri, exact != 0, 1, cx1, cy1, workCap, 0u, root, miss, G)) {
          add();
          continue;
        }
        ++levels[1];
        for (int s = 0; s < 16; ++s) {
          if (!fp64::traceCell(*t2, sv, R, r.tri, exact != 0, 2, cx1 * 4 + (s & 3), cy1 * 4 + (s >> 2), 0xffffffffu,
                               0u, root, miss, G))
            return -1 - i;
          add();
        }
      }
    }
    vis[i] = fp64::visOfFx(vs);
    E[i] = exact ? fp64::gOfFx(gs) : R.A * R.G0 * vis[i];
  }
  for (int k = 0; k < sNum; ++k) counters[k] = uint64_t(t0->cnt.v[k]) + t1->cnt.v[k] + t2->cnt.v[k];
  return 0;
}

}  // namespace

extern "C" {

// FP64 soft shadows of n receiver points, as the library's FP64 pass shades a
// pixel (pools of 64, 128 and 384 beams for the three levels of cells).
// refs: the leaf references in leaf order (64-byte records); nodes (12-byte),
// nodeBox (6 floats), parent per node; bounds6 the kd root box. light9: c, U,
// V. Per point: x (3 doubles), ref (index into refs: the hit triangle), nf (3
// floats, the triangle normal facing the eye). workCap: triangle tests and kd
// steps a level-0 or level-1 cell may take (the library's: 65536). Writes vis,
// E and Eu; counters[sNum] (summed, overflowed traces included); levels[2].
// Returns 0, or -1 - i when a level-2 cell of point i overflowed its pools.
int fp64_soft(const void* refs, const void* nodes, const float* nodeBox, const uint32_t* parent, const float* bounds6,
              const float* light9, double offset, int exact, const double* x, const int32_t* ref, const float* nf,
              int n, uint32_t workCap, double* vis, double* E, double* Eu, uint64_t* counters, int64_t* levels) {
  return shadeAll<fp64::Tracer<64, 64>, fp64::Tracer<128, 128>, fp64::Tracer<384, 256>>(
      refs, nodes, nodeBox, parent, bounds6, light9, offset, exact, x, ref, nf, n, workCap, vis, E, Eu, counters,
      levels);
}

// The same with first pools of 4 beams and 4 work items, so that pool
// overflows also send cells to the next level.
int fp64_soft_tiny(const void* refs, const void* nodes, const float* nodeBox, const uint32_t* parent,
                   const float* bounds6, const float* light9, double offset, int exact, const double* x,
                   const int32_t* ref, const float* nf, int n, uint32_t workCap, double* vis, double* E, double* Eu,
                   uint64_t* counters, int64_t* levels) {
  return shadeAll<fp64::Tracer<4, 4>, fp64::Tracer<8, 8>, fp64::Tracer<384, 256>>(
      refs, nodes, nodeBox, parent, bounds6, light9, offset, exact, x, ref, nf, n, workCap, vis, E, Eu, counters,
      levels);
}

}  // extern "C"
