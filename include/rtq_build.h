/*
 * rtq host-side builder and synthetic workloads, C ABI (pure host code, no GPU needed).
 *
 * The builder restates the reference's CPU BVH builder and flattener so that the blob handed to
 * rtq_scene_upload is the one the reference emits:
 *   rtq_builder_add_vertices       BVHGeometry::addTriangleList (Renderer/BVHGeometry.cpp:46-62)
 *   rtq_builder_add_triangle_list  BVHBuilder::addTriangleList (Renderer/BVHBuilder.cpp:265-280)
 *   rtq_builder_add_sphere_light   BVHBuilder::addSphereLight (BVHBuilder.cpp:282-285)
 *   rtq_builder_set_parameters     BVHBuilder::setParameters (BVHBuilder.cpp:378-381)
 *   rtq_builder_add_blas           BVHBuilder::addBlas (BVHBuilder.cpp:288-301)
 *   rtq_builder_set_blas_parameters the parameters of BVHBuilder::buildBlas (BVHBuilder.cpp:383-399)
 *   rtq_builder_build              BVHBuilder::build (BVHBuilder.cpp:408-461), after buildBlas when
 *                                  BLASes were added (BVHData::build, BVHData.cpp:28-40)
 *   rtq_builder_gpu_size           BVHBuilder::getBvhGpuSize (BVHBuilder.cpp:1185-1222)
 *   rtq_builder_fill_gpu_buffer    BVHBuilder::fillGpuBuffer (BVHBuilder.cpp:1319-1534)
 */
#ifndef RTQ_BUILD_H
#define RTQ_BUILD_H

#include <stddef.h>
#include <stdint.h>

#include "rtq.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rtq_builder rtq_builder;

/* Detail of the last failing builder call on this thread (empty string if none). */
const char* rtq_build_last_error(void);

/* BVHBuildParameters (Renderer/BVHBuilder.h:53-61). rtq_builder_create sets the values the
 * reference application builds with (main.cpp:197-201): 25, 4, 4, 0.25, 0.5. */
typedef struct rtq_build_params {
  uint32_t max_depth;
  uint32_t min_obj_per_node;
  uint32_t min_obj_gain;
  float expand_node_volume_threshold;
  float expand_node_dimension_factor;
} rtq_build_params;

typedef struct rtq_build_stats {
  uint32_t num_nodes;
  uint32_t num_leaves;
  uint32_t max_leaf_triangles;
  uint32_t max_depth;
  uint64_t leaf_triangle_slots;
  uint32_t num_duplicated;
  uint32_t num_triangles;
  double build_seconds;
} rtq_build_stats;

rtq_status rtq_builder_create(rtq_builder** out);
rtq_status rtq_builder_destroy(rtq_builder* b);
rtq_status rtq_builder_add_vertices(rtq_builder* b, const float* xyz, uint32_t count, uint32_t* vertex_offset);
/* One Lambertian material of `color` per call; indices relative to vertex_offset must be < 65536
 * and vertex_offset + index < vertex count (RTQ_ERR_INVALID_ARGUMENT otherwise; the reference
 * truncates to 16 bits silently). At most 65535 calls (16-bit material id). */
rtq_status rtq_builder_add_triangle_list(rtq_builder* b, uint32_t vertex_offset, uint32_t num_triangles,
                                         const uint32_t* indices, const float color[3]);
rtq_status rtq_builder_add_sphere_light(rtq_builder* b, const float pos[3], float radius, const float color[3],
                                        float sphere_radius);
rtq_status rtq_builder_set_parameters(rtq_builder* b, const rtq_build_params* p);
/* Two-level scenes. A bottom-level BVH of one triangle list over this builder's vertices (indices as
 * for rtq_builder_add_triangle_list), added as the next BLAS instance (*blas_index); a builder with
 * BLASes builds as the reference's top-level builder: every BLAS first with the BLAS parameters
 * (default: the application's, as rtq_builder_create sets them), then the top level over its own
 * triangles and the BLAS instances' boxes with rtq_builder_set_parameters' values (the application
 * uses 25, 6, 6, 1, 0.225 there, main.cpp:203-206), hoisting into a node the instances both of its
 * children hold (BVHBuilder.cpp:905-940). The blob carries one BlasHeader per instance and the nodes
 * of every BLAS after the top level's. */
rtq_status rtq_builder_add_blas(rtq_builder* b, uint32_t vertex_offset, uint32_t num_triangles,
                                const uint32_t* indices, const float color[3], uint32_t* blas_index);
rtq_status rtq_builder_set_blas_parameters(rtq_builder* b, const rtq_build_params* p);
/* threads <= 1: single thread. Node numbering never depends on threads. */
rtq_status rtq_builder_build(rtq_builder* b, int threads);
rtq_status rtq_builder_gpu_size(const rtq_builder* b, size_t* bytes);
rtq_status rtq_builder_fill_gpu_buffer(const rtq_builder* b, void* blob, size_t bytes, rtq_ranges* ranges);
rtq_status rtq_builder_positions(const rtq_builder* b, const float** xyz, uint32_t* count);
rtq_status rtq_builder_get_stats(const rtq_builder* b, rtq_build_stats* stats);
rtq_status rtq_builder_scene_box(const rtq_builder* b, float mn[3], float mx[3]);

/* ---- synthetic workloads (seeded, deterministic) ---- */
#define RTQ_SCENE_BOX 1     /* closed box: 5 walls of 32x32 quads + 2 boxes of 16x16-quad faces */
#define RTQ_SCENE_SPHERES 2 /* `param` UV spheres of 256 triangles, centres U[-100,100]^3, r ~ LogN(0,0.7) */
#define RTQ_SCENE_SOUP 3    /* `param` triangles from a 64-component Gaussian mixture, edges LogU[0.01,1] */

rtq_status rtq_scene_generate(rtq_builder* b, uint32_t kind, uint64_t seed, uint32_t param);
/* `count` point lights, pos U[lo,hi]^3, radius LogU[rmin,rmax], colour 1, sphere radius 0.1. */
rtq_status rtq_lights_generate(rtq_builder* b, uint64_t seed, uint32_t count, float lo, float hi, float rmin,
                               float rmax);

/* Pinhole camera with the reference projection (fov 70*3.14/180 in fillPassResources,
 * raytracingPass.cpp:43; perspective near 0.1, far TMAX, depth range [0,1], frustum corners at
 * NDC depth 0.5, :53-63). */
rtq_status rtq_camera_frustum(const float pos[3], const float dir[3], const float up[3], float fov_y, float aspect,
                              rtq_frustum* out);
/* Primary rays through pixel (x, y), row y = 0 at corner00 (cameraPass.comp:50-54): target =
 * c00 + stepW*(x + jx) + stepH*(y + jy), dir = normalize(target - cam). jitter U[0,1)^2 per
 * pixel from `seed` when seed != 0, else 0. Ray i = y*width + x. */
rtq_status rtq_frustum_rays(const rtq_frustum* f, uint32_t width, uint32_t height, uint64_t seed, float tmax,
                            rtq_ray* out);
/* Diffuse bounce rays from the builder's triangles: a uniform random triangle, a uniform point on
 * it, origin offset 0.001 along the face normal (rayStorageHelpers.glsl:112), cosine-weighted
 * direction about that normal, t limit tmax. */
rtq_status rtq_surface_rays(const rtq_builder* b, uint64_t seed, uint32_t count, float tmax, rtq_ray* out);
/* Next diffuse bounce from closest hits: for ray i with hits[i].tri_id != RTQ_NO_HIT, origin =
 * org + dir*t + 0.001*n (n faces the incoming ray), cosine direction about n; misses produce a
 * ray with tmax 0 (traced as an immediate miss). */
rtq_status rtq_bounce_rays(const rtq_builder* b, uint64_t seed, const rtq_ray* rays, const rtq_hit* hits,
                           uint32_t count, float tmax, rtq_ray* out);
/* ---- on-disk scene cache (no reference counterpart) ----
 * A built scene's blob, ranges and vertex positions in one file, keyed by a hash of the builder's
 * inputs (positions, triangle lists, materials, lights, build parameters). */
typedef struct rtq_cached_scene rtq_cached_scene;
rtq_status rtq_builder_input_hash(const rtq_builder* b, uint64_t* hash);
/* The builder must be built. Written through a temporary file renamed into place. */
rtq_status rtq_cache_save(const rtq_builder* b, const char* path);
/* expected_hash 0 accepts any key. RTQ_ERR_CACHE_MISS: no readable file, or a valid file for other
 * inputs; RTQ_ERR_BAD_BLOB: not a cache file of this version, truncated or corrupt (checksum). */
rtq_status rtq_cache_load(const char* path, uint64_t expected_hash, rtq_cached_scene** out);
/* Pointers stay valid until rtq_cached_scene_free; they can go straight to rtq_scene_upload. */
rtq_status rtq_cached_scene_view(const rtq_cached_scene* c, const void** blob, size_t* bytes, rtq_ranges* ranges,
                                 const float** positions, uint32_t* num_vertices, uint64_t* hash);
rtq_status rtq_cached_scene_free(rtq_cached_scene* c);

/* `count` (a square, 64 for the probe pass) stratified unit directions on the sphere, 4 floats each
 * (xyz, 0), as LightProbFieldPass::sampleRays lays them out (lightProbFieldPass.cpp:151-180); the
 * jitter is seeded here. */
rtq_status rtq_probe_directions(uint64_t seed, uint32_t count, float* out);

#ifdef __cplusplus
}
#endif

#endif /* RTQ_BUILD_H */
