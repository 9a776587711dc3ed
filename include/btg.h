/* MI355X beam tracer: the GPU path as a C ABI (libbtg.so, gfx950 HIP).
 *
 * Drop-in boundary for the reference's GPU renderer, cuda/gpu_render.h of
 * joelsalzman/beamtrace2007. Each entry point replaces one member of that
 * class; a C++ GpuRenderer with the reference's exact signatures is layered
 * on top in beamtrace2007_amd/shim/gpu_render_hip.cpp.
 *
 * Conventions: plain C types only; every function returning int returns 0 on
 * success and a negative code on error (no abort inside the library), with
 * the message in btg_last_error(ctx) (ctx may be NULL for errors of
 * btg_create / btg_device_query). One context per device; calls on one
 * context are serialised by the caller (as in the reference, gpu_render.h:26).
 * All host arrays are owned by the caller and only read or written during the
 * call.
 */
#ifndef BTG_H
#define BTG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  BTG_OK = 0,
  BTG_ERR_ARG = -1,     /* invalid argument */
  BTG_ERR_DEVICE = -2,  /* HIP runtime error, or no device */
  BTG_ERR_MEMORY = -3,  /* device allocation failed */
  BTG_ERR_LIMIT = -4    /* a root needs more sub-beams than the queues hold */
};

typedef struct btg_ctx btg_ctx;

/* gpuAvailable / GpuInfo (cuda/gpu_render.h:17-24). Returns 1 if a device
 * is usable (and fills *info if given), 0 otherwise. */
typedef struct {
  char name[256];
  char arch[64];
  int computeUnits;
  int wavefrontSize;
  double memGB;
} btg_device_info;
int btg_device_query(btg_device_info* info);

/* The scene and kd-tree the tracer reads (SceneView, src/beam/beam_types.h:86-97,
 * uploaded as in GpuRenderer::GpuRenderer, cuda/gpu_render.cu:489-512).
 * nodes: 12 bytes each {float split; uint32 data; uint32 offset}
 * (src/accel/kdtree.h:13-21); nodeBox: 6 floats (lo, hi) per node; parent:
 * parent index per node (root: ~0u). kdBounds: the kd root box; sceneBounds:
 * the scene's bounds (sets the soft-shadow apex offset, gpu_render.cu:574).
 * The library copies everything to device memory; the caller keeps ownership. */
typedef struct {
  const float* pos;
  int64_t numVerts;
  const uint32_t* tri3;
  const float* triN;
  int64_t numTris;
  const void* nodes;
  int64_t numNodes;
  const uint32_t* triIndices;
  int64_t numRefs;
  const uint32_t* parent;
  const float* nodeBox;
  float kdBounds[6];
  float sceneBounds[6];
} btg_scene_desc;

/* GpuRenderer::GpuRenderer / ~GpuRenderer (cuda/gpu_render.h:28-29). */
int btg_create(btg_ctx** out, const btg_scene_desc* scene);
void btg_destroy(btg_ctx* ctx);
const char* btg_last_error(const btg_ctx* ctx);

/* Engine knobs (cuda/gpu_render.h:48-58): budget = live sub-beams per soft /
 * point root (wfBudget, 128), primaryBudget = per primary tile (4096),
 * capacity = sub-beam records per queue (wfCapacity, 1 << 23),
 * roundsPerCheck = rounds launched between completion checks.
 * Roots over budget are not sent to the CPU: they are traced again on the GPU
 * with a larger budget. */
typedef struct {
  int budget;
  int primaryBudget;
  int64_t capacity;
  int roundsPerCheck;
  int profileKernels; /* 1: time every launch with HIP events (btg_run_info.kernelSeconds) */
  int leafCap;        /* leaves a sub-beam visits per trace launch before it yields to the next round (0: no cap);
                         primary-visibility passes use at most 4 (latency-bound chains of short rounds) */
  int precision;      /* soft shadows: 0 = the reference's float build, bit for bit (default); 1 = FP64-guarded:
                         the receiver point is the primary ray's hit computed in double, the beam apex sits at
                         the double build's offset (1e-9 x scene extent) from it, formed in double and carried as
                         a float pair in every apex-relative difference; the Lambert integral runs in double throughout.
                         Results approach bt_render_d's (mean visibility error 100-1000x below the float build's
                         on the test scenes; bench.py reports it). Same throughput as 0.
                         2 = FP64: the double build's algorithm (bt_render_d: depth-first beam trace, its
                         tolerances, its geometry in double) per pixel, from the primary hit in double; with the
                         job's double camera (btg_soft_job.camD) every visible fraction is bt_render_d's to
                         ~1e-10. Tens of times slower than 0; a server on such a context serves its soft
                         requests in FP64 batches (see btg_server_create). */
  int useGraphs;      /* 1 (default): each group of roundsPerCheck rounds is captured once per pass into a HIP graph
                         and replayed (no counterpart in the reference, which launches every kernel from the host,
                         cuda/wavefront.cu:641-676); 0: plain launches. Results are the same either way.
                         profileKernels = 1 launches directly (it times every launch with HIP events).
                         While a pass captures (microseconds per group), operations that other threads of
                         the process issue on the legacy (null) stream fail on this runtime; a program
                         that uses the legacy stream from other threads meanwhile sets 0. The library
                         itself never uses the legacy stream, and a server's passes never capture. */
  int mailbox;        /* 1: Post Office mailboxing on the GPU (GpuRenderer::useMailbox, cuda/gpu_render.h:48-50;
                         src/beam/beam_core.h:772-781, 884-912): a sub-beam skips a triangle it has already been
                         tested against with the result Keep while unchanged (a triangle straddling a split plane
                         is listed in several leaves); counted in mailboxSkips instead of triTests. Results are the
                         same as with 0 (default). */
} btg_options;
int btg_get_options(const btg_ctx* ctx, btg_options* opt);
int btg_set_options(btg_ctx* ctx, const btg_options* opt);

/* Camera (src/render/render_types.h:15-39): image plane at distance 1 along
 * fwd, q in [-halfW, halfW] x [-halfH, halfH]. W and H: 1..32768 pixels. */
typedef struct {
  float eye[3], fwd[3], right[3], down[3];
  float halfW, halfH;
  int W, H;
} btg_camera;

/* The camera in double, as the reference's double build forms it
 * (src/render/render_types.h:15-39 with Real = double; include/bth.h's
 * bth_camera_d makes it from a view). Only FP64 soft shadows
 * (btg_options.precision = 2) read it: their primary rays then are the
 * double build's, which the float camera's rounding would move by ~1e-8 of
 * the distance (up to ~4e-5 of visible fraction at sharp penumbrae). */
typedef struct {
  double eye[3], fwd[3], right[3], down[3];
  double halfW, halfH;
} btg_camera_d;

/* AreaLight (src/render/soft_core.h:10-14): centre and full edge vectors. */
typedef struct {
  float c[3], U[3], V[3];
} btg_area_light;

/* TraceStats (src/util/stats.h:10-42). */
typedef struct {
  uint64_t kdSteps, leafVisits, triTests, hits, splits, beams, rootBeams, presplitBeams, mailboxSkips, fiveSplits,
      visibleTris, rays, climbs;
  double droppedArea;
} btg_trace_stats;

/* Per-call diagnostics (GpuRenderer::last*, cuda/gpu_render.h:60-64). Times
 * come from HIP events on the library's stream. */
typedef struct {
  double primarySeconds;  /* soft shadows: the primary-ray pass */
  double traceSeconds;    /* beam kernels (all rounds, all passes) */
  double totalSeconds;    /* host wall time of the call */
  int rounds;
  int maxQueue;
  int killedRoots;        /* roots that went over budget in the first pass */
  int retryPasses;        /* extra GPU passes that traced them exactly */
  btg_trace_stats stats;  /* beam counters (engine-independent) */
  btg_trace_stats rayStats;
  /* with profileKernels: summed launch durations and launch counts per kernel
   * (0 admit, 1 trace, 2 split, 3 primary rays, 4 other) */
  double kernelSeconds[5];
  int kernelLaunches[5];
  /* builds with -DBTG_PHASE_CYCLES: trace-loop clocks spent refilling,
   * descending, climbing (parent walk, child box test), emitting, testing
   * triangles, appending and in the rest of the traversal (soft shadows) */
  uint64_t phaseCycles[8];
} btg_run_info;

/* GpuRenderer::softShadows (cuda/gpu_render.h:36): exact soft shadows of a
 * W x H image. Outputs (W*H each, row-major, caller-allocated; any may be
 * NULL): tri = primary hit (-1 background), vis = visible fraction of the
 * light (-1 background, 0 light not visible), E = exact integral of
 * cos cos / r^2 over the visible light if exact != 0, else area * G0 * vis,
 * Eu = area * G0. */
int btg_soft(btg_ctx* ctx, const btg_camera* cam, const btg_area_light* light, int exact, int32_t* tri, float* vis,
             float* E, float* Eu, btg_run_info* info);

/* Resident scenes (no counterpart in the reference, whose GpuRenderer holds
 * one scene; the batched and served forms of the soft-shadow path need
 * several resident at once). The scene given to btg_create is scene 0 for
 * the context's lifetime; btg_scene_add copies another one to the device and
 * returns its id (>= 1); btg_scene_remove frees it (the id may be reused).
 * btg_scene_bytes: device bytes the scene occupies (-1: no such scene). */
int btg_scene_add(btg_ctx* ctx, const btg_scene_desc* scene, int* sceneId);
/* The same from a resident scene file (.btgs, written by bth_scene_save,
 * include/bth.h): the file's hash and indices are verified, its kd-tree is
 * uploaded as stored (no build). No counterpart in the reference. */
int btg_create_file(btg_ctx** out, const char* path);
int btg_scene_add_file(btg_ctx* ctx, const char* path, int* sceneId);
int btg_scene_remove(btg_ctx* ctx, int sceneId);
int64_t btg_scene_bytes(const btg_ctx* ctx, int sceneId);

/* One job of a soft-shadow batch: resident scene, camera, area light, and
 * outputs as btg_soft (W*H each, any may be NULL). */
typedef struct {
  int scene;
  btg_camera cam;
  btg_area_light light;
  int32_t* tri;
  float* vis;
  float* E;
  float* Eu;
  const btg_camera_d* camD; /* NULL, or the same camera in double: with precision = 2 the primary rays are its
                               (W and H stay cam's); other precisions ignore it */
} btg_soft_job;

/* GpuRenderer::softShadows (cuda/gpu_render.h:36) for many jobs at once:
 * every job's roots share one persistent wavefront pass (one primary-ray
 * launch, one set of rounds), so small and uneven jobs keep the device full.
 * Each job's outputs equal those of btg_soft on the same scene, camera and
 * light. info holds the batch's totals. */
int btg_soft_batch(btg_ctx* ctx, const btg_soft_job* jobs, int numJobs, int exact, btg_run_info* info);

/* Schedule autotuner (no counterpart in the reference, whose knobs are fixed,
 * cuda/gpu_render.h:56-58 and cuda/gpu_common.cuh:54-58): times the sample
 * batch `jobs` (typically the real batch at reduced resolution) under the
 * current options, then tunes one knob at a time - leaf cap, rounds between
 * host checks - keeping a value when it is more than 2% faster than the
 * incumbent, within about maxSeconds. The per-root budget and the queue
 * capacity are the caller's (they set how many roots run at once, which a
 * scaled-down sample cannot measure). Writes the best options to `best` (the
 * context's options are left unchanged: apply them with btg_set_options)
 * and, if `trials` is given, up to maxTrials timed settings. Results do not
 * depend on these knobs, only speed does. */
typedef struct {
  int leafCap, budget, roundsPerCheck, mailbox;
  double seconds;
} btg_autotune_trial;
int btg_autotune(btg_ctx* ctx, const btg_soft_job* jobs, int numJobs, int exact, double maxSeconds,
                 btg_options* best, btg_autotune_trial* trials, int maxTrials, int* numTrials);

/* A polygon in image-plane coordinates (OutBeam, src/beam/beam_types.h:68-73):
 * corners past n repeat the last one. 44 bytes. */
typedef struct {
  float x[4], y[4];
  int32_t n;
  int32_t tri;
  float area;
} btg_poly;

typedef struct btg_polys btg_polys;
int64_t btg_polys_count(const btg_polys* p);
const btg_poly* btg_polys_data(const btg_polys* p);
void btg_polys_free(btg_polys* p);

/* GpuRenderer::primary (cuda/gpu_render.h:41): hit beams (tri >= 0) of one
 * root per tile x tile pixels. canonical != 0 sorts the list into a
 * canonical order (tri, n, corners, area); otherwise the order is the
 * device's. *out receives a new list (free with btg_polys_free). */
int btg_primary(btg_ctx* ctx, const btg_camera* cam, int cull, int tile, int canonical, btg_polys** out,
                btg_run_info* info);

/* GpuRenderer::pointShadows (cuda/gpu_render.h:45-46): shadowed polygons
 * (image plane) for the given primary hit polygons and point light; a
 * receiver facing away from the light yields its whole primary polygon. */
int btg_point(btg_ctx* ctx, const btg_camera* cam, const btg_poly* prim, int64_t numPrim, const float* light3,
              int canonical, btg_polys** out, btg_run_info* info);

/* rasterize (src/render/render.cpp:49-95) on the device, one sample at each
 * pixel centre, bit for bit with the reference's CPU post-pass run over the
 * polygons in canonical order: shadowPass = 0 fills ids[W*H] with the tri of
 * the last covering polygon in canonical order (the largest tri, as the order
 * sorts by tri first; -1 where none; polygons with tri < 0 are skipped);
 * shadowPass = 1 fills mask[W*H] with 1 where any polygon covers the pixel
 * centre, else 0. The polygons are uploaded, the image downloaded. */
int btg_raster(btg_ctx* ctx, const btg_camera* cam, const btg_poly* polys, int64_t count, int shadowPass,
               int32_t* ids, uint8_t* mask);

/* btg_raster over device memory with no allocation. No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * polysDev: `count` btg_poly in device memory; idsDev (shadowPass = 0, W*H
 * int32) or maskDev (shadowPass = 1, W*H bytes) in device memory, written
 * with exactly btg_raster's result. The scan scratch is a static device array
 * sized for BTG_RASTER_NOALLOC_MAX_POLYS polygons (12 MiB), one per device,
 * held until the process ends; count above that maximum is BTG_ERR_LIMIT. All
 * calls on one device share that array: calls that overlap on one device race
 * (serialise them). Synchronous on the context's stream; device buffers the
 * caller wrote on other streams must be complete before the call. */
#define BTG_RASTER_NOALLOC_MAX_POLYS (1 << 20)
int btg_raster_device_noalloc(btg_ctx* ctx, const btg_camera* cam, const btg_poly* polysDev, int64_t count,
                              int shadowPass, int32_t* idsDev, uint8_t* maskDev);

/* One frame of primary visibility and (light3 != NULL) point shadows with
 * both rasterizations on the device (bt_render's primary / point output,
 * apps/bt_render.cpp:228-247, without the CPU post-pass): ids[W*H] as
 * btg_raster of btg_primary's polygons, mask[W*H] as btg_raster of
 * btg_point's output for those polygons. The polygons never leave the
 * device. primInfo / pointInfo: the two passes (either may be NULL);
 * rasterSeconds: the two rasterizations with their image downloads. */
int btg_images(btg_ctx* ctx, const btg_camera* cam, int cull, int tile, const float* light3, int32_t* ids,
               uint8_t* mask, btg_run_info* primInfo, btg_run_info* pointInfo, double* rasterSeconds);

/* ---------------------------------------------------------------- serving
 * A request server over one context (no counterpart in the reference, which
 * renders one job per call): scenes stay resident, requests are admitted by
 * the device memory their work needs, and soft-shadow requests join a
 * persistent wavefront pass that is already running (continuous batching:
 * a request that arrives while others are on the device is added to the
 * running pass as soon as its roots and pixels fit, and all in-flight
 * requests' sub-beams share every round). Two such passes alternate, so a
 * pass's space is reclaimed once its requests are done. Point-shadow
 * requests run primary visibility and point shadows on their own stream,
 * overlapping the soft-shadow work. Results equal those of the one-call
 * entry points on the same inputs.
 *
 * On a context set to FP64 soft shadows (btg_options.precision = 2) there is
 * no wavefront pass to join (each pixel is traced on its own): the server
 * serves soft requests in FP64 batches instead, every request pending when
 * the previous batch ends going into the next one (at most 64 requests and
 * the pixels that hbmBytes allows), so requests that arrive together share
 * the device. Each result equals btg_soft_batch's on the same job (the
 * per-pixel sums are fixed point, so the batch a request shares does not
 * change it); with the request's double camera (btg_request.camD) it is the
 * double build's to the rounding of the float outputs.
 *
 * While a server exists, the context must not be used by other calls
 * (its resident scenes must not change). btg_server_submit is thread-safe;
 * each ticket is waited on once. */
typedef struct btg_server btg_server;

typedef struct {
  int64_t hbmBytes;    /* device memory for work in flight, split between two passes (0: 128 GiB, or 45% of the
                          device where that is less); a quarter of it holds the passes' queues, which set how
                          many roots run at once */
  int roundsPerStep;   /* wavefront rounds between host checks (0: 8) */
  int jobsPerPass;     /* request slots of one pass (0: 512) */
} btg_server_config;

enum { BTG_REQ_SOFT = 0, BTG_REQ_POINT = 1 };

typedef struct {
  int kind;                /* BTG_REQ_SOFT or BTG_REQ_POINT */
  int scene;               /* soft: resident scene id; point shadows use scene 0 */
  btg_camera cam;
  btg_area_light light;    /* soft */
  float pointLight[3];     /* point */
  int exact;               /* soft: must be 1 (the server's passes take the exact integral) */
  int32_t* tri;            /* soft outputs as btg_soft (W*H, any may be NULL); written */
  float* vis;              /* before the request reports done; they must stay valid */
  float* E;                /* until then */
  float* Eu;
  const btg_camera_d* camD; /* soft, FP64 servers only: NULL, or the same camera in double (as btg_soft_job.camD;
                               read before btg_server_submit returns) */
} btg_request;

typedef struct {
  int done;                /* 1 once finished (successfully or not) */
  int status;              /* BTG_OK or an error code (message: btg_server_error) */
  double queueSeconds;     /* submit -> its work admitted to the device */
  double latencySeconds;   /* submit -> results in host memory */
  btg_polys* hits;         /* point: primary hit polygons, canonical order; the caller frees them */
  btg_polys* polys;        /* point: shadowed polygons, canonical order; the caller frees them */
} btg_request_status;

/* HBM footprint of a request: the device bytes it holds while in flight,
 * resident scenes excluded. Soft shadows: its images and hit points per pixel
 * plus the state of its roots (64 per 8 x 8 pixel tile), the quantities a
 * server's admission counts against its passes (hbmBytes / 2 each, less the
 * queues). Point shadows: the primary pass's tile roots and polygon buffer; the
 * point pass then adds one root per primary hit polygon. Plays the part of the
 * footprint query of a batching/serving front end (no counterpart in
 * cuda/gpu_render.h, which serves one call at a time). */
int btg_footprint(const btg_ctx* ctx, const btg_request* req, int64_t* bytes);

int btg_server_create(btg_ctx* ctx, const btg_server_config* cfg, btg_server** out);
int btg_server_submit(btg_server* s, const btg_request* req, uint64_t* ticket);
/* Waits up to timeoutSeconds (< 0: no limit; 0: a poll) for a ticket;
 * status->done = 0 on timeout. A finished ticket is released by this call,
 * once: another waiter on the same ticket then returns BTG_ERR_ARG. */
int btg_server_wait(btg_server* s, uint64_t ticket, double timeoutSeconds, btg_request_status* status);
/* The server's last error message; the string stays valid until this
 * thread's next btg_server_error call. */
const char* btg_server_error(btg_server* s);
/* Finishes the requests in flight, then frees the server. */
void btg_server_destroy(btg_server* s);

#ifdef __cplusplus
}
#endif

#endif
