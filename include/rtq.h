/*
 * rtq: ray queries on one AMD Instinct MI355X (gfx950), C ABI.
 *
 * The entry points replace the three surfaces through which the reference renderer feeds and
 * runs its traversal kernels:
 *   upload-BVH / upload-triangles  BVHData::build -> IRenderer::CreateBuffer + UploadBuffer
 *                                  (Renderer/BVHData.cpp:28-48, rtDevice/public/IRenderer.h:30-33)
 *                                  and BVHGeometry::flush (Renderer/BVHGeometry.cpp:69-78)
 *   dispatch-rays                  RayTracingPass::tracePass -> IRenderContext::Dispatch
 *                                  (Renderer/raytracingPass.cpp:89-110, rtDevice/public/IRenderContext.h:70)
 *   read-hits                      the reference keeps hits in a GPU-only buffer
 *                                  (Renderer/raytracingPass.cpp:72-74); reading them back is new API.
 *
 * Conventions: every call returns rtq_status and never aborts; handles are opaque; the caller owns
 * host memory, the library owns the device copies it makes; ray/hit/mask buffers passed to
 * dispatch calls are device pointers owned by the caller; `stream` is a hipStream_t (NULL is the
 * null stream) and dispatch calls are asynchronous on it.
 */
#ifndef RTQ_H
#define RTQ_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum rtq_status {
  RTQ_OK = 0,
  RTQ_ERR_INVALID_ARGUMENT = 1,
  RTQ_ERR_HIP = 2,
  RTQ_ERR_OUT_OF_MEMORY = 3,
  RTQ_ERR_LIMIT = 4,    /* input beyond a documented limit (tree depth, counts, capacities) */
  RTQ_ERR_BAD_BLOB = 5, /* blob and ranges are inconsistent */
  RTQ_ERR_NO_DEVICE = 6,
  RTQ_ERR_BUDGET = 7,     /* serving: a request can never fit the HBM budget */
  RTQ_ERR_QUEUE_FULL = 8, /* serving: admission queue full, retry later */
  RTQ_ERR_CANCELLED = 9,
  RTQ_ERR_TIMEOUT = 10,
  RTQ_ERR_INTERNAL = 11,
  RTQ_ERR_CACHE_MISS = 12 /* scene cache: no file, or one built from other inputs */
} rtq_status;

/* One sub-range of the flattened BVH blob: byte offset and byte size. */
typedef struct rtq_range {
  uint32_t offset;
  uint32_t size;
} rtq_range;

/* The seven ranges BVHBuilder::fillGpuBuffer returns, in its argument order
 * (Renderer/BVHBuilder.h:101; bound to shader slots by BVHData::fillBvhBindings, BVHData.cpp:17-26). */
typedef struct rtq_ranges {
  rtq_range triangles, primitives, materials, lights, nodes, leaf_data, blas;
} rtq_ranges;

/* Ray record, 32 B: the reference Ray {from, dir} (Shaders/core/primitive_cpp.glsl:28-35) plus the
 * traversal's t limit, which the reference passes separately (TMAX for primary rays,
 * baseRaytracingPass.glsl:18; the light distance for shadow rays, bvhLighting.glsl:122). */
typedef struct rtq_ray {
  float org[3];
  float tmax;
  float dir[3];
  uint32_t pad;
} rtq_ray;

/* Closest-hit record, 16 B. tri_id is the builder insertion index of the accepted triangle
 * (RTQ_NO_HIT on a miss); t is the value the reference stores, t_hit - 1e-5f
 * (bvhCollision.glsl:210), or the ray's tmax on a miss; u, v are the Moller-Trumbore
 * barycentrics of the accepted test (collision.glsl:79,84), 0 on a miss. */
typedef struct rtq_hit {
  uint32_t tri_id;
  float t;
  float u;
  float v;
} rtq_hit;

#define RTQ_NO_HIT 0xFFFFFFFFu

/* Frustum of a pinhole camera as the reference's PassData carries it (Shaders/struct_cpp.glsl:10-28,
 * filled by RayTracingPass::fillPassResources, raytracingPass.cpp:39-63). */
typedef struct rtq_frustum {
  float camera_pos[3];
  float corner00[3];
  float corner10[3];
  float corner01[3];
} rtq_frustum;

typedef struct rtq_scene_info {
  uint32_t num_nodes;       /* reference node count (inner + leaves) */
  uint32_t num_inner;       /* inner nodes in the device layout */
  uint32_t num_leaf_slots;  /* inline triangle slots over all leaves */
  uint32_t num_triangles;
  uint32_t num_lights;
  uint32_t max_depth;
  uint64_t device_bytes;    /* HBM held by the scene */
} rtq_scene_info;

typedef struct rtq_device rtq_device;
typedef struct rtq_scene rtq_scene;

const char* rtq_status_string(rtq_status s);
/* Detail of the last failing call on this thread (empty string if none). */
const char* rtq_last_error(void);

rtq_status rtq_device_create(int hip_ordinal, rtq_device** out);
rtq_status rtq_device_destroy(rtq_device* dev);

/* Dispatch tuning; results are identical in every mode. RTQ_OPTION_PERSISTENT:
 *   0 = one lane per ray over a grid of ceil(n/128) blocks (default);
 *   1 = a fixed grid of waves whose lanes refill with the next unclaimed rays of the whole dispatch
 *       after each leaf (best for incoherent rays such as diffuse bounces);
 *   2 = a fixed grid in which each wave owns a contiguous range of rays and refills its lanes from
 *       it (lanes keep neighbouring, coherent rays).
 * RTQ_OPTION_BLOCKS_PER_CU: persistent blocks of 128 lanes per compute unit (1..16), or 0 (default) =
 *   automatic: 7, and 12 for two-level scenes (as many as the compute unit's LDS holds). */
#define RTQ_OPTION_PERSISTENT 1u
#define RTQ_OPTION_BLOCKS_PER_CU 2u
/* RTQ_OPTION_SORT_RAYS: 1 = rtq_dispatch_rays first sorts the rays by a Morton code of the origin over
 * the scene box, then direction octant (rocPRIM radix sort) and traces them in that order, so that a
 * wave's lanes hold rays that start close together and head the same way; results still land at
 * each ray's own index and are identical. The sort workspace belongs to the device and grows on
 * demand; sorting dispatches on different streams take it in turn (each waits on the device for the
 * previous one's traversal to finish), so they are safe from any streams but do not overlap. Default 0. */
#define RTQ_OPTION_SORT_RAYS 3u
/* RTQ_OPTION_LEAF_WAIT: a wave interleaves descent toward the next leaf with the tests of the leaves
 * its lanes have reached; a round of descent ends once this many lanes per 64 of the wave's live lanes
 * are ready for leaf tests (1..64; 64 = every lane reaches its leaf first). 0 (default) = 40 for one
 * closest-hit ray per lane, 16 for one any-hit ray per lane (the value the shadow, record and probe
 * passes always use for their any-hit rays), 40 for refilling dispatches (24 for those of two-level
 * scenes). Results are identical for every value. */
#define RTQ_OPTION_LEAF_WAIT 4u
/* Options apply to dispatches issued after the call; set them while no other thread dispatches on the
 * device (they are plain fields, read when a dispatch is enqueued). */
rtq_status rtq_device_set_option(rtq_device* dev, uint32_t option, uint32_t value);
rtq_status rtq_device_get_option(const rtq_device* dev, uint32_t option, uint32_t* value);

/* Plain device-memory helpers for callers without their own allocator. */
rtq_status rtq_device_alloc(rtq_device* dev, size_t bytes, void** d_ptr);
rtq_status rtq_device_free(rtq_device* dev, void* d_ptr);
rtq_status rtq_copy_to_device(rtq_device* dev, void* d_dst, const void* h_src, size_t bytes, void* stream);
/* Page-locked host memory (hipHostMalloc): request rays and results in it move by DMA at full link
 * rate instead of through the runtime's staging copies of pageable memory. */
rtq_status rtq_host_alloc(rtq_device* dev, size_t bytes, void** h_ptr);
rtq_status rtq_host_free(rtq_device* dev, void* h_ptr);

/* Streams and events of the library's HIP runtime, for callers that need ordering and timing
 * without a second runtime in the process. Streams have the default priority. */
rtq_status rtq_stream_create(rtq_device* dev, void** stream);
rtq_status rtq_stream_destroy(rtq_device* dev, void* stream);
rtq_status rtq_stream_synchronize(rtq_device* dev, void* stream);
rtq_status rtq_event_create(rtq_device* dev, void** event);
rtq_status rtq_event_destroy(rtq_device* dev, void* event);
rtq_status rtq_event_record(void* event, void* stream);
/* Milliseconds between two recorded events; waits for `end`. */
rtq_status rtq_event_elapsed_ms(void* start, void* end, float* ms);

/* Device-to-device copy with the library's own streaming kernel (16-byte loads and stores); also
 * the reference point for HBM counter calibration. `bytes` must be a multiple of 16. */
rtq_status rtq_device_copy(rtq_device* dev, void* d_dst, const void* d_src, size_t bytes, void* stream);

/* upload-BVH + upload-triangles. `blob`/`ranges` are exactly what BVHBuilder::fillGpuBuffer wrote
 * (BVHBuilder.cpp:1319-1534); `positions_xyz` holds the BVHGeometry vertex positions (12 B each).
 * The blob is converted once into the device layout and stays resident in HBM until
 * rtq_scene_destroy. Fails with RTQ_ERR_BAD_BLOB if ranges/offsets point outside the blob or a
 * triangle references a vertex >= num_vertices, and RTQ_ERR_LIMIT if the tree is deeper than 64.
 * A blob whose leaves list BLAS instances (the reference's top-level builder with addBlas,
 * BVHBuilder.cpp:288-301,1521-1533) uploads as a two-level scene: its depth is the top level's + 1 +
 * the deepest BLAS's, and only any-hit rtq_dispatch_rays and culling accept it. */
rtq_status rtq_scene_upload(rtq_device* dev, const void* blob, size_t bytes, const rtq_ranges* ranges,
                            const float* positions_xyz, uint32_t num_vertices, rtq_scene** out);
rtq_status rtq_scene_destroy(rtq_scene* scene);
rtq_status rtq_scene_get_info(const rtq_scene* scene, rtq_scene_info* info);

#define RTQ_TRACE_CLOSEST 1u /* traverseBvh: d_out is rtq_hit[num_rays] */
#define RTQ_TRACE_ANY 2u     /* traverseBvhFast: d_out is uint32_t[num_rays], 1 = occluded */

/* dispatch-rays: traverseBvh (closest hit, Shaders/bvh/bvhTraversal_inline.glsl:3-78, via rayTrace,
 * baseRaytracingPass.glsl:16-27) or traverseBvhFast (any hit, bvhTraversal_inline.glsl:80-151).
 * d_leaf_nid (closest hit only, may be NULL) receives the reference node id of the leaf that held
 * the accepted triangle (ClosestHit_storeNid, bvhBindings.glsl:127-134), 0xFFFFFFFF on a miss.
 * On a two-level scene RTQ_TRACE_ANY runs traverseTlasFast with tlas_collide_fast
 * (bvhTraversal_inline.glsl:80-151, bvhCollision.glsl:321-349: a BLAS is entered when its header box
 * test is > 0, the top level's own triangles are not tested); RTQ_TRACE_CLOSEST returns
 * RTQ_ERR_INVALID_ARGUMENT, since the reference's closest-hit top-level walk calls an undefined HitBox. */
rtq_status rtq_dispatch_rays(rtq_scene* scene, const rtq_ray* d_rays, uint32_t num_rays, uint32_t flags,
                             void* d_out, uint32_t* d_leaf_nid, void* stream);

/* Same traversal, also writing per-ray visit counters d_counters[3*i + {0,1,2}] = inner nodes
 * whose two child boxes were tested (numTraversal, bvhTraversal_inline.glsl:12,19), leaves
 * entered, triangle tests. Used to measure algorithmic bytes; slower than rtq_dispatch_rays. */
rtq_status rtq_dispatch_rays_counted(rtq_scene* scene, const rtq_ray* d_rays, uint32_t num_rays, uint32_t flags,
                                     void* d_out, uint32_t* d_leaf_nid, uint32_t* d_counters, void* stream);

#define RTQ_CAMERA_REFERENCE_ROWS 1u /* store at the reference's row 16*ceil(H/16) - gy; row 0 untouched */

/* dispatch-rays with rays generated in the kernel, as RayTracingPass::tracePass does: one primary
 * ray per pixel from the frustum (cameraPass.comp:46-56), t limit TMAX = 100, closest hit.
 * d_hits / d_leaf_nid hold width*height records, pixel (x, y) at y*width + x, y = 0 at corner00. */
rtq_status rtq_trace_camera(rtq_scene* scene, const rtq_frustum* frustum, uint32_t width, uint32_t height,
                            uint32_t flags, rtq_hit* d_hits, uint32_t* d_leaf_nid, void* stream);

/* Light-probe field placement (LightProbFieldHeader, Shaders/lightprob/lightprob.glsl:16-21, as the
 * camera pass fills it from the scene box and field size, cameraPass.comp:62-66). */
typedef struct rtq_lpf {
  float aabb_min[3];
  float aabb_max[3];
  uint32_t resolution[3];
} rtq_lpf;

/* The camera pass's full TRACING_STEP output (cameraPass.comp:69-75), 8 uint32 per pixel at
 * 2*(y*width + x): [vertexOffset, index01, index2_matId, floatBits(t)] of the closest hit and
 * [shadowMask, probMask, 0, 0] (computeShadowMask with `sun_dir`/`shadow_flags` as in
 * rtq_shadow_masks; computeProbMask when lpf is not NULL, else 0). A miss has t = 100, zero triangle
 * words and zero masks. `flags` as for rtq_trace_camera. */
rtq_status rtq_trace_camera_record(rtq_scene* scene, const rtq_frustum* frustum, uint32_t width, uint32_t height,
                                   uint32_t flags, const float sun_dir[3], uint32_t shadow_flags, const rtq_lpf* lpf,
                                   uint32_t* d_records, void* stream);

/* The light-probe field's TRACING_STEP (genLightProbField.comp:44-69, dispatched by
 * LightProbFieldPass::traceLightProbField, lightProbFieldPass.cpp:81-100): for each of the
 * resolution.x*y*z probes (numbered in 4x4x4 tiles, getLightProbCoord) and each of the 64 directions
 * `dirs` (xyz + pad, the pass's GenLightProbFieldConstants.rays), a closest-hit ray from the probe's
 * position, recorded as rtq_trace_camera_record does (probe mask 0: the pass runs without
 * USE_LPF_MASK) at 8 uint32 per sample, sample = probe * 64 + direction. resolution.x and .y must be
 * at least 4 (the tiling divides by resolution / 4). */
#define RTQ_PROBE_RAYS 64u
rtq_status rtq_trace_probe_field(rtq_scene* scene, const rtq_lpf* lpf, const float dirs[4 * 64],
                                 const float sun_dir[3], uint32_t shadow_flags, uint32_t* d_records, void* stream);

/* read-hits: synchronising copy of `bytes` from device memory written by a dispatch on `stream`. */
rtq_status rtq_read_hits(rtq_device* dev, const void* d_src, void* h_dst, size_t bytes, void* stream);

#define RTQ_SHADOW_SUN 1u /* trace bit 0 toward -sun_dir with t limit TMAX */

/* computeShadowMask (Shaders/bvh/bvhLighting.glsl:143-163) for closest hits with t < 100:
 * pos = org + dir * t; bit 0 = any hit toward -sun_dir (if RTQ_SHADOW_SUN); for each light index L
 * in the hit leaf's light list, bit L+1 = (|light - pos| < radius) ? any hit toward the light with
 * t limit |light - pos| : 1 (evalShadow, :110-131). Lights with L+1 >= 32 have no bit (the shader's
 * shift is undefined there) and are not traced. Misses get mask 0. */
rtq_status rtq_shadow_masks(rtq_scene* scene, const rtq_ray* d_rays, const rtq_hit* d_hits,
                            const uint32_t* d_leaf_nid, uint32_t num_rays, const float sun_dir[3], uint32_t flags,
                            uint32_t* d_masks, void* stream);

/* Per-tile frustum light culling: the light part of cullWithFrustumTile (Shaders/cullFrustum.glsl:
 * 63-100,136-153) over a ceil(width/16) x ceil(height/16) tile grid, tile (x, y) = id y*tiles_x + x.
 * Output is CSR: d_tile_offsets[num_tiles+1], d_tile_lights[...] ascending light indices per tile
 * (the reference's 16-entry atomic-order list is the first 16 of it); d_tile_offsets[num_tiles] is
 * the total, and lists beyond `capacity` entries in all are truncated, never overrun. Each light is
 * tested once per tile. With h_total NULL the call is asynchronous on `stream` (its scratch is
 * stream-ordered device memory). With h_total set it synchronises `stream`, stores the total there
 * and returns RTQ_ERR_LIMIT if the total exceeds `capacity`. */
rtq_status rtq_cull_lights(rtq_scene* scene, const rtq_frustum* frustum, uint32_t width, uint32_t height,
                           uint32_t* d_tile_offsets, uint32_t* d_tile_lights, uint32_t capacity, uint32_t* h_total,
                           void* stream);

/* rtq_cull_lights that allocates nothing. No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). Same kernels and results as rtq_cull_lights with h_total NULL, asynchronous on
 * `stream`; the total is d_tile_offsets[num_tiles]. Its scratch is a fixed 16 MiB device array that
 * every GPU the library runs on holds until the process ends, so a call needs
 * num_tiles * (2 * ceil(num_lights / 64) + 1) <= 4 Mi words (3840x2160 with 4096 lights fits) and
 * returns RTQ_ERR_LIMIT beyond it, launching nothing. Every call on one GPU uses that GPU's one
 * array: calls that overlap on one GPU (other streams, other threads) race and must be ordered by
 * the caller. */
rtq_status rtq_cull_lights_noalloc(rtq_scene* scene, const rtq_frustum* frustum, uint32_t width, uint32_t height,
                                   uint32_t* d_tile_offsets, uint32_t* d_tile_lights, uint32_t capacity,
                                   void* stream);

/* ---- batched form: many scenes and ray sets in one persistent launch ----
 * No counterpart in the reference (it dispatches one pass at a time, raytracingPass.cpp:89-110).
 * Items reference resident scenes by index; each item's rays, outputs and t limits are exactly as in
 * rtq_dispatch_rays, and results are identical to dispatching each item alone. */
typedef struct rtq_batch rtq_batch;

typedef struct rtq_batch_item {
  uint32_t scene;    /* index into the scene array given to rtq_batch_create */
  uint32_t flags;    /* RTQ_TRACE_CLOSEST or RTQ_TRACE_ANY */
  uint32_t num_rays;
  uint32_t reserved; /* 0 */
  const rtq_ray* d_rays;
  void* d_out;            /* rtq_hit[num_rays] or uint32_t[num_rays] */
  uint32_t* d_leaf_nid;   /* closest hit only, may be NULL */
} rtq_batch_item;

/* chunk_rays value selecting the lane-refilling kernel: lanes take single rays from any item as soon
 * as their previous ray is done (best for incoherent rays; chunks keep coherent rays together). */
#define RTQ_BATCH_REFILL 1u
typedef struct rtq_batch_params {
  uint32_t chunk_rays;    /* rays a wave takes per grab: 64, 128, 256 or 512, or RTQ_BATCH_REFILL */
  uint32_t blocks_per_cu; /* persistent 128-lane blocks per compute unit, 1..16 */
  uint32_t stack_depth;   /* per-lane LDS stack entries: 24, 32 or 64, >= the deepest scene */
  uint32_t use_graph;     /* 1: replay a captured hipGraph (needs a non-NULL stream) */
  uint32_t leaf_wait;     /* lanes ready for leaf tests that end a round of descent (RTQ_OPTION_LEAF_WAIT);
                             0 = default: 40 */
} rtq_batch_params;

rtq_status rtq_batch_create(rtq_device* dev, rtq_scene* const* scenes, uint32_t num_scenes,
                            const rtq_batch_item* items, uint32_t num_items, rtq_batch** out);
rtq_status rtq_batch_destroy(rtq_batch* batch);
rtq_status rtq_batch_set_params(rtq_batch* batch, const rtq_batch_params* params);
rtq_status rtq_batch_get_params(const rtq_batch* batch, rtq_batch_params* params);
/* One persistent launch over every item (asynchronous on `stream`). Launches of one batch share its
 * work counter and item table: issue them on one stream, or order them with events. The scenes must
 * outlive the batch (it keeps their device views). */
rtq_status rtq_batch_launch(rtq_batch* batch, void* stream);
/* The same items as one plain rtq_dispatch_rays launch each (for comparison). */
rtq_status rtq_batch_launch_plain(rtq_batch* batch, void* stream);
/* Times every valid (chunk_rays incl. RTQ_BATCH_REFILL, blocks_per_cu, stack_depth) on this batch, `reps`
 * launches each, then leaf_wait in 8..64 for the fastest of them; keeps the fastest in the batch and
 * returns it with its mean time. Synchronises `stream`. */
rtq_status rtq_batch_autotune(rtq_batch* batch, uint32_t reps, rtq_batch_params* best, float* best_ms,
                              void* stream);

/* ---- served form: resident scenes, budgeted admission, continuous batching ----
 * No counterpart in the reference. Scenes stay resident; requests (host rays in, host results
 * out) are admitted while their device buffers fit the HBM budget, batched with whatever else is
 * admitted, and traced by the persistent batch kernel; up to max_inflight batches are on the device
 * at once, so the rays of one batch cross the link while another is traced. Traversal arithmetic is the same as rtq_dispatch_rays, so
 * each result is identical whether the request runs alone or in a mix. rtq_server_submit and
 * rtq_request_* are thread-safe. */
typedef struct rtq_server rtq_server;
typedef struct rtq_request rtq_request;

typedef struct rtq_server_config {
  uint64_t hbm_budget_bytes; /* device bytes for in-flight request buffers (52 B per ray) */
  uint32_t max_queue;        /* waiting requests beyond this make submit return RTQ_ERR_QUEUE_FULL */
  uint32_t max_batch_rays;   /* 0: no cap besides the budget */
  uint32_t plain;            /* 1: one launch per request, no batching (comparison mode) */
  uint32_t max_inflight;     /* batches (or plain launches) on the device at once, 1..8; 0: 3 */
  rtq_batch_params params;   /* kernel settings; chunk_rays 0 = defaults (lane refill, 8 blocks/CU) */
} rtq_server_config;

typedef struct rtq_request_stats {
  double queued_ms;   /* submit -> admitted */
  double latency_ms;  /* submit -> results in host memory (device requests: in device memory) */
  uint32_t batch_id;  /* launch that carried the request */
  uint32_t batch_requests;
} rtq_request_stats;

rtq_status rtq_server_create(rtq_device* dev, rtq_scene* const* scenes, uint32_t num_scenes,
                             const rtq_server_config* cfg, rtq_server** out);
/* Waits for all admitted work, fails queued requests with RTQ_ERR_CANCELLED, frees the server.
 * Request handles not yet released stay valid (all finished) and are released as before; no other
 * call on the server or its requests may run concurrently with this one. */
rtq_status rtq_server_destroy(rtq_server* srv);
/* host_rays, host_out (rtq_hit or uint32_t per ray) and host_leaf_nid (closest hit, may be NULL)
 * must stay valid until the request completes. RTQ_ERR_BUDGET: the request alone exceeds the budget. */
rtq_status rtq_server_submit(rtq_server* srv, uint32_t scene, uint32_t flags, const rtq_ray* host_rays,
                             uint32_t num_rays, void* host_out, uint32_t* host_leaf_nid, rtq_request** req);
/* The same for a request whose rays and results live in device memory on the server's device (rays
 * made and results consumed by other GPU passes, as the reference's passes hand their buffers to
 * each other: cameraPass.comp writes the record buffer the next pass reads, raytracingPass.cpp:112-137).
 * Traced in place: no arena bytes (the budget does not apply), no copies over the host link; the
 * request completes when its results are in d_out / d_leaf_nid. d_rays and closest-hit d_out must be
 * 16-byte aligned, any-hit d_out and d_leaf_nid 4-byte aligned; all must stay valid until completion. */
rtq_status rtq_server_submit_device(rtq_server* srv, uint32_t scene, uint32_t flags, const rtq_ray* d_rays,
                                    uint32_t num_rays, void* d_out, uint32_t* d_leaf_nid, rtq_request** req);
/* Blocks up to timeout_ms (0xFFFFFFFF = forever); returns the request's final status, or
 * RTQ_ERR_TIMEOUT if it is still running. */
rtq_status rtq_request_wait(rtq_request* req, uint32_t timeout_ms);
/* Cancels a request that has not been admitted yet (RTQ_ERR_INVALID_ARGUMENT otherwise). */
rtq_status rtq_request_cancel(rtq_request* req);
rtq_status rtq_request_get_stats(const rtq_request* req, rtq_request_stats* stats);
/* Frees a completed request handle. */
rtq_status rtq_request_release(rtq_request* req);

/* Diagnostic: compares the traversal's reciprocal-based slab division with IEEE division on
 * `count` random operand pairs from the range where the kernels use it; *mismatches receives the
 * number of bitwise differences (0 expected). No counterpart in the reference. Synchronous. */
rtq_status rtq_selftest_division(rtq_device* dev, uint64_t seed, uint64_t count, uint64_t* mismatches);

#ifdef __cplusplus
}
#endif

#endif /* RTQ_H */
