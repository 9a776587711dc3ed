/* Host input stage of the MI355X beam tracer: synthetic scenes, the SAH
 * kd-tree, cameras and the polygon rasterizer, as a C ABI.
 *
 * These are the caller-side data formats on either side of the GPU path
 * (include/btg.h). In the reference they are C++ classes the GPU renderer
 * receives: Scene (src/scene/scene.h:14-39), KdTree (src/accel/kdtree.h:34-51),
 * Camera (src/render/render_types.h:15-39) and rasterize
 * (src/render/render.h:32). The kd build here is bit-identical to
 * KdTree::build (src/accel/kdtree.cpp:237-281).
 *
 * All functions returning int return 0 on success and -1 on error, with the
 * message in bth_last_error() (per thread). Arrays returned by pointer are
 * owned by the scene handle and stay valid until bth_scene_free().
 */
#ifndef BTH_H
#define BTH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bth_scene bth_scene;

const char* bth_last_error(void);

/* Named, seeded synthetic scene ("room", "building", "plant", "analytic",
 * "random_tris", "sphere", "building_small"), finalized like
 * Scene::finalize (src/scene/scene.cpp:10-39). */
int bth_scene_generate(const char* name, uint64_t seed, bth_scene** out);
/* Scene from raw vertices (3 floats each) and triangles (3 indices each). */
int bth_scene_from_arrays(const float* pos, int64_t numVerts, const uint32_t* tris, int64_t numTris, bth_scene** out);
void bth_scene_free(bth_scene* scene);
int64_t bth_scene_num_tris(const bth_scene* scene);
int64_t bth_scene_num_verts(const bth_scene* scene);
/* pos: 3 floats per vertex; tri3: 3 indices per triangle; triN: unit normal
 * per triangle; triArea; bounds6 = lo xyz, hi xyz. */
void bth_scene_arrays(const bth_scene* scene, const float** pos, const uint32_t** tri3, const float** triN,
                      const float** triArea, float* bounds6);
/* The scene's camera view (eye, target, up, fovY in degrees), point light and
 * area light (centre, U, V). Returns flags: 1 point light, 2 area light,
 * 4 cull back faces. */
int bth_scene_spec(const bth_scene* scene, float* view10, float* point3, float* light9);

/* SAH kd-tree (KdBuildParams, src/accel/kdtree.h:23-32; beams use
 * costIntersect 0.4, apps/bt_render.cpp:170-174). */
int bth_kd_build(bth_scene* scene, float costIntersect, int singleLeaf);
int bth_kd_sizes(const bth_scene* scene, int64_t* numNodes, int64_t* numRefs, int* depth);
/* nodes: 12 bytes each {float split; uint32 data; uint32 offset}
 * (src/accel/kdtree.h:13-21); nodeBox: 6 floats per node; bounds6: root box. */
int bth_kd_arrays(const bth_scene* scene, const void** nodes, const uint32_t** triIndices, const uint32_t** parent,
                  const float** nodeBox, float* bounds6);

/* Resident scene format (.btgs; layout in beamtrace2007_amd/host/scene_file.h):
 * the scene, its view and lights, and (when built) its kd-tree, with a payload
 * hash. No counterpart in the reference, which rebuilds the tree on every
 * start (src/accel/kdtree.cpp:237-281). bth_scene_load verifies the hash and
 * the indices; the loaded scene's kd-tree needs no build. btg_scene_add_file
 * (include/btg.h) uploads such a file straight into a GPU context. */
int bth_scene_save(const bth_scene* scene, const char* path);
int bth_scene_load(const char* path, bth_scene** out);

/* Camera::make (src/render/render.cpp:14-25). view10 as above; cam14 = eye,
 * fwd, right, down (3 floats each), halfW, halfH. */
void bth_camera(const float* view10, int W, int H, float* cam14);

/* The same camera formed in double, as the reference's double build makes it
 * (Real = double) from the same view: eye, fwd, right, down, halfW, halfH in
 * 14 doubles. FP64 soft shadows (btg_options.precision = 2) shoot their
 * primary rays from it when a job carries it (btg_soft_job.camD). */
void bth_camera_d(const float* view10, int W, int H, double* cam14);

/* rasterize (src/render/render.cpp:49-95), one sample at each pixel centre.
 * Polygons: x4/y4 4 floats each, n corners, tri id. canonical != 0 sorts the
 * polygons into a canonical order first (deterministic ties on shared edges).
 * shadowPass = 0 writes tri ids into triOut; 1 sets flagOut = 1. */
int bth_rasterize(const float* cam14, int W, int H, const float* x4, const float* y4, const int32_t* n,
                  const int32_t* tri, int64_t count, int shadowPass, int canonical, int32_t* triOut, uint8_t* flagOut);

#ifdef __cplusplus
}
#endif

#endif
