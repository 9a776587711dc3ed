/* Test oracle: a scalar CPU restatement of the reference beam tracer.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product (beamtrace2007_amd/)
 * links, loads or calls this; only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg do, as the checker.
 *
 * It restates the reference's depth-first BeamCore with the restart trail and
 * without mailboxing (src/beam/beam_core.h:549-1009, useTrail = true,
 * useMailbox = false), which the reference's own wavefront engine reproduces
 * piece for piece (tests/test_wavefront.cpp:40-74), plus the per-root setups
 * of the three applications (src/render/soft_core.h, point_core.h) and the
 * ray tracer of the soft-shadow primary pass (src/ray/ray_core.h). Every
 * floating-point operation follows the reference's float build in kind and
 * order (no contraction), so results are bit-identical to it; pinned by
 * tests/test_oracle.py against oracle/_ref and the golden fixtures.
 */
#ifndef BT_ORACLE_H
#define BT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_scene orc_scene;

/* Raw scene + kd arrays (the layout of include/btg.h's btg_scene_desc). */
orc_scene* orc_scene_new(const float* pos, int64_t numVerts, const uint32_t* tri3, const float* triN, int64_t numTris,
                         const void* nodes12, int64_t numNodes, const uint32_t* triIndices, int64_t numRefs,
                         const uint32_t* parent, const float* nodeBox6, const float* kdBounds6,
                         const float* sceneBounds6);
void orc_scene_free(orc_scene* s);

/* Counters in the order kdSteps, leafVisits, triTests, hits, splits, beams,
 * rootBeams, presplitBeams, mailboxSkips, fiveSplits, visibleTris, rays,
 * climbs, droppedArea (TraceStats, src/util/stats.h:10-42). */
#define ORC_NSTATS 14

/* Exact soft shadows for a W x H image (softShadows, src/render/render.cpp:339
 * with the trail and no mailbox). cam14 = eye, fwd, right, down, halfW,
 * halfH; light9 = centre, U, V. Rows [y0, y1) only (y1 <= 0: all rows);
 * outputs are full-image arrays. threads >= 1. guard = 1: the FP64-guarded
 * variant of the GPU library (double hit point, the double build's apex
 * offset, the apex as a float pair). */
void orc_soft(const orc_scene* s, const float* cam14, int W, int H, const float* light9, int exact, int threads,
              int y0, int y1, int32_t* tri, float* vis, float* E, float* Eu, double* stats, int guard);

/* Primary visibility, one root per tile x tile pixels (tile <= 0: the whole
 * image as one root). Returns a polygon list: hit beams only for tiles,
 * every output beam for the whole image. */
typedef struct orc_polys orc_polys;
orc_polys* orc_primary(const orc_scene* s, const float* cam14, int W, int H, int cull, int tile, double* stats);
/* The same for the tiles of tile rows [ty0, ty1) only (a strip of a frame). */
orc_polys* orc_primary_tiles(const orc_scene* s, const float* cam14, int W, int H, int cull, int tile, int ty0,
                             int ty1, double* stats);
/* Point-light shadows of the primary hit polygons (beamPointShadows,
 * src/render/render.cpp:209-247). */
orc_polys* orc_point(const orc_scene* s, const float* cam14, int W, int H, const float* px4, const float* py4,
                     const int32_t* pn, const int32_t* ptri, const float* parea, int64_t nPrim, const float* light3,
                     double* stats);
int64_t orc_polys_count(const orc_polys* p);
void orc_polys_get(const orc_polys* p, float* x4, float* y4, int32_t* n, int32_t* tri, float* area);
void orc_polys_free(orc_polys* p);

/* Primary rays through the pixel centres: tri, X, N (render.cpp:351-381). */
void orc_primary_rays(const orc_scene* s, const float* cam14, int W, int H, int32_t* tri, float* X, float* N,
                      double* stats);

/* Visibility of the light from one point (beamLightVisibility,
 * src/render/render.cpp:279-303). Returns V; *G = exact integral. */
float orc_light_vis(const orc_scene* s, const float* x3, const float* ns3, int tri, const float* light9, float offset,
                    float* G, double* stats);

#ifdef __cplusplus
}
#endif

#endif
