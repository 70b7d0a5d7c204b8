// C ABI of the host builder and synthetic workloads (include/rtq_build.h).
#include <cstring>
#include <exception>
#include <new>
#include <string>

#include "../../../include/rtq_build.h"
#include "blob_cache.h"
#include "bvh_builder.h"
#include "scenes.h"

struct rtq_builder {
  rtq::BvhBuilder impl;
  bool built = false;
};

struct rtq_cached_scene {
  rtq::CachedScene impl;
};

namespace {
thread_local std::string g_buildError;

rtq_status fail(rtq_status s, const char* msg) {
  g_buildError = msg;
  return s;
}

template <class F>
rtq_status guarded(F&& f) {
  try {
    return f();
  } catch (const std::bad_alloc&) {
    return fail(RTQ_ERR_OUT_OF_MEMORY, "host allocation failed");
  } catch (const std::exception& e) {
    return fail(RTQ_ERR_INTERNAL, e.what());
  }
}
}  // namespace

namespace {
rtq::BuildParams toParams(const rtq_build_params* p) {
  rtq::BuildParams bp;
  bp.maxDepth = p->max_depth;
  bp.minObjPerNode = p->min_obj_per_node;
  bp.minObjGain = p->min_obj_gain;
  bp.expandNodeVolumeThreshold = p->expand_node_volume_threshold;
  bp.expandNodeDimensionFactor = p->expand_node_dimension_factor;
  return bp;
}

// Index checks shared by triangle lists and BLASes.
rtq_status checkList(const rtq_builder* b, uint32_t vertex_offset, uint32_t num_triangles, const uint32_t* indices) {
  const size_t nv = b->impl.positions().size();
  for (size_t i = 0; i < size_t(num_triangles) * 3; ++i) {
    if (indices[i] > 0xFFFFu) return fail(RTQ_ERR_INVALID_ARGUMENT, "triangle index does not fit 16 bits");
    if (size_t(vertex_offset) + indices[i] >= nv) return fail(RTQ_ERR_INVALID_ARGUMENT, "index beyond vertex count");
  }
  return RTQ_OK;
}
}  // namespace

extern "C" {

const char* rtq_build_last_error(void) { return g_buildError.c_str(); }

rtq_status rtq_builder_create(rtq_builder** out) {
  if (!out) return fail(RTQ_ERR_INVALID_ARGUMENT, "out is NULL");
  return guarded([&] {
    *out = new rtq_builder();
    return RTQ_OK;
  });
}

rtq_status rtq_builder_destroy(rtq_builder* b) {
  delete b;
  return RTQ_OK;
}

rtq_status rtq_builder_add_vertices(rtq_builder* b, const float* xyz, uint32_t count, uint32_t* vertex_offset) {
  if (!b || (!xyz && count)) return fail(RTQ_ERR_INVALID_ARGUMENT, "NULL builder or positions");
  return guarded([&] {
    uint32_t off = b->impl.addVertices(count, xyz);
    if (vertex_offset) *vertex_offset = off;
    b->built = false;
    return RTQ_OK;
  });
}

rtq_status rtq_builder_add_triangle_list(rtq_builder* b, uint32_t vertex_offset, uint32_t num_triangles,
                                         const uint32_t* indices, const float color[3]) {
  if (!b || (!indices && num_triangles)) return fail(RTQ_ERR_INVALID_ARGUMENT, "NULL builder or indices");
  rtq_status st = checkList(b, vertex_offset, num_triangles, indices);
  if (st != RTQ_OK) return st;
  // material ids are 16 bits (g_MaxMaterialCount, BVHBuilder.cpp:29); material 0 is the default
  if (b->impl.numMaterials() >= 0x10000u) return fail(RTQ_ERR_LIMIT, "more than 65535 triangle lists");
  return guarded([&] {
    float c[3] = {0.7f, 0.7f, 0.7f};
    if (color) std::memcpy(c, color, sizeof(c));
    b->impl.addTriangleList(vertex_offset, num_triangles, indices, rtq::BvhBuilder::lambertian(c[0], c[1], c[2]));
    b->built = false;
    return RTQ_OK;
  });
}

rtq_status rtq_builder_add_sphere_light(rtq_builder* b, const float pos[3], float radius, const float color[3],
                                        float sphere_radius) {
  if (!b || !pos) return fail(RTQ_ERR_INVALID_ARGUMENT, "NULL builder or position");
  return guarded([&] {
    rtq::SphereLightDesc l{};
    l.pos = {pos[0], pos[1], pos[2]};
    l.radius = radius;
    l.color = color ? rtq::F3{color[0], color[1], color[2]} : rtq::F3{1, 1, 1};
    l.sphereRadius = sphere_radius;
    b->impl.addSphereLight(l);
    b->built = false;
    return RTQ_OK;
  });
}


rtq_status rtq_builder_set_parameters(rtq_builder* b, const rtq_build_params* p) {
  if (!b || !p) return fail(RTQ_ERR_INVALID_ARGUMENT, "NULL builder or params");
  b->impl.setParameters(toParams(p));
  b->built = false;
  return RTQ_OK;
}

rtq_status rtq_builder_set_blas_parameters(rtq_builder* b, const rtq_build_params* p) {
  if (!b || !p) return fail(RTQ_ERR_INVALID_ARGUMENT, "NULL builder or params");
  b->impl.setBlasParameters(toParams(p));
  b->built = false;
  return RTQ_OK;
}

rtq_status rtq_builder_add_blas(rtq_builder* b, uint32_t vertex_offset, uint32_t num_triangles,
                                const uint32_t* indices, const float color[3], uint32_t* blas_index) {
  if (!b || (!indices && num_triangles)) return fail(RTQ_ERR_INVALID_ARGUMENT, "NULL builder or indices");
  if (num_triangles == 0) return fail(RTQ_ERR_INVALID_ARGUMENT, "a BLAS needs triangles");
  rtq_status st = checkList(b, vertex_offset, num_triangles, indices);
  if (st != RTQ_OK) return st;
  // each BLAS brings two materials (its default and the list's); ids are 16 bits
  if (b->impl.numMaterials() + 2 > 0x10000u) return fail(RTQ_ERR_LIMIT, "more than 65536 materials");
  if (b->impl.numBlas() >= (1u << 24)) return fail(RTQ_ERR_LIMIT, "too many BLAS instances");
  return guarded([&] {
    float c[3] = {0.7f, 0.7f, 0.7f};
    if (color) std::memcpy(c, color, sizeof(c));
    const uint32_t k = b->impl.addBlas(vertex_offset, num_triangles, indices, rtq::BvhBuilder::lambertian(c[0], c[1], c[2]));
    if (blas_index) *blas_index = k;
    b->built = false;
    return RTQ_OK;
  });
}

rtq_status rtq_builder_build(rtq_builder* b, int threads) {
  if (!b) return fail(RTQ_ERR_INVALID_ARGUMENT, "NULL builder");
  return guarded([&] {
    b->impl.build(threads);
    b->built = true;
    return RTQ_OK;
  });
}

rtq_status rtq_builder_gpu_size(const rtq_builder* b, size_t* bytes) {
  if (!b || !bytes) return fail(RTQ_ERR_INVALID_ARGUMENT, "NULL argument");
  if (!b->built) return fail(RTQ_ERR_INVALID_ARGUMENT, "builder has not been built");
  *bytes = b->impl.gpuSize();
  return RTQ_OK;
}

rtq_status rtq_builder_fill_gpu_buffer(const rtq_builder* b, void* blob, size_t bytes, rtq_ranges* ranges) {
  if (!b || !blob || !ranges) return fail(RTQ_ERR_INVALID_ARGUMENT, "NULL argument");
  if (!b->built) return fail(RTQ_ERR_INVALID_ARGUMENT, "builder has not been built");
  if (bytes < b->impl.gpuSize()) return fail(RTQ_ERR_INVALID_ARGUMENT, "blob smaller than rtq_builder_gpu_size");
  return guarded([&] {
    rtq::BlobRanges r{};
    b->impl.fillGpuBuffer(blob, &r);
    static_assert(sizeof(rtq::BlobRanges) == sizeof(rtq_ranges), "range layout");
    std::memcpy(ranges, &r, sizeof(r));
    return RTQ_OK;
  });
}

rtq_status rtq_builder_positions(const rtq_builder* b, const float** xyz, uint32_t* count) {
  if (!b || !xyz || !count) return fail(RTQ_ERR_INVALID_ARGUMENT, "NULL argument");
  *xyz = reinterpret_cast<const float*>(b->impl.positions().data());
  *count = uint32_t(b->impl.positions().size());
  return RTQ_OK;
}

rtq_status rtq_builder_get_stats(const rtq_builder* b, rtq_build_stats* s) {
  if (!b || !s) return fail(RTQ_ERR_INVALID_ARGUMENT, "NULL argument");
  const rtq::BuildStats& st = b->impl.stats();
  s->num_nodes = st.numNodes;
  s->num_leaves = st.numLeaves;
  s->max_leaf_triangles = st.maxLeafTriangles;
  s->max_depth = st.maxDepth;
  s->leaf_triangle_slots = st.leafTriangleSlots;
  s->num_duplicated = st.numDuplicated;
  s->num_triangles = b->impl.numTriangles();
  s->build_seconds = st.buildSeconds;
  return RTQ_OK;
}

rtq_status rtq_builder_scene_box(const rtq_builder* b, float mn[3], float mx[3]) {
  if (!b || !mn || !mx) return fail(RTQ_ERR_INVALID_ARGUMENT, "NULL argument");
  rtq::Box3 box = b->impl.sceneBox();
  mn[0] = box.mn.x; mn[1] = box.mn.y; mn[2] = box.mn.z;
  mx[0] = box.mx.x; mx[1] = box.mx.y; mx[2] = box.mx.z;
  return RTQ_OK;
}

rtq_status rtq_scene_generate(rtq_builder* b, uint32_t kind, uint64_t seed, uint32_t param) {
  if (!b || kind < RTQ_SCENE_BOX || kind > RTQ_SCENE_SOUP) return fail(RTQ_ERR_INVALID_ARGUMENT, "bad scene kind");
  return guarded([&] {
    rtq::generateScene(b->impl, kind, seed, param);
    b->built = false;
    return RTQ_OK;
  });
}

rtq_status rtq_lights_generate(rtq_builder* b, uint64_t seed, uint32_t count, float lo, float hi, float rmin,
                               float rmax) {
  if (!b || !(rmin > 0) || !(rmax >= rmin)) return fail(RTQ_ERR_INVALID_ARGUMENT, "bad light parameters");
  return guarded([&] {
    rtq::generateLights(b->impl, seed, count, lo, hi, rmin, rmax);
    b->built = false;
    return RTQ_OK;
  });
}

rtq_status rtq_camera_frustum(const float pos[3], const float dir[3], const float up[3], float fov_y, float aspect,
                              rtq_frustum* out) {
  if (!pos || !dir || !up || !out) return fail(RTQ_ERR_INVALID_ARGUMENT, "NULL argument");
  rtq::Frustum f{};
  rtq::cameraFrustum(pos, dir, up, fov_y, aspect, f);
  static_assert(sizeof(rtq::Frustum) == sizeof(rtq_frustum), "frustum layout");
  std::memcpy(out, &f, sizeof(f));
  return RTQ_OK;
}

rtq_status rtq_frustum_rays(const rtq_frustum* f, uint32_t width, uint32_t height, uint64_t seed, float tmax,
                            rtq_ray* out) {
  if (!f || !out || !width || !height) return fail(RTQ_ERR_INVALID_ARGUMENT, "bad frustum ray arguments");
  rtq::Frustum ff;
  std::memcpy(&ff, f, sizeof(ff));
  static_assert(sizeof(rtq::RayRec) == sizeof(rtq_ray), "ray layout");
  rtq::frustumRays(ff, width, height, seed, tmax, reinterpret_cast<rtq::RayRec*>(out));
  return RTQ_OK;
}

rtq_status rtq_surface_rays(const rtq_builder* b, uint64_t seed, uint32_t count, float tmax, rtq_ray* out) {
  if (!b || (!out && count) || b->impl.numTriangles() == 0) return fail(RTQ_ERR_INVALID_ARGUMENT, "bad arguments");
  return guarded([&] {
    rtq::surfaceRays(b->impl, seed, count, tmax, reinterpret_cast<rtq::RayRec*>(out));
    return RTQ_OK;
  });
}

rtq_status rtq_bounce_rays(const rtq_builder* b, uint64_t seed, const rtq_ray* rays, const rtq_hit* hits,
                           uint32_t count, float tmax, rtq_ray* out) {
  if (!b || ((!rays || !hits || !out) && count)) return fail(RTQ_ERR_INVALID_ARGUMENT, "bad arguments");
  static_assert(sizeof(rtq::HitRec) == sizeof(rtq_hit), "hit layout");
  return guarded([&] {
    rtq::bounceRays(b->impl, seed, reinterpret_cast<const rtq::RayRec*>(rays),
                    reinterpret_cast<const rtq::HitRec*>(hits), count, tmax, reinterpret_cast<rtq::RayRec*>(out));
    return RTQ_OK;
  });
}

rtq_status rtq_builder_input_hash(const rtq_builder* b, uint64_t* hash) {
  if (!b || !hash) return fail(RTQ_ERR_INVALID_ARGUMENT, "NULL argument");
  *hash = rtq::inputHash(b->impl);
  return RTQ_OK;
}

rtq_status rtq_cache_save(const rtq_builder* b, const char* path) {
  if (!b || !path) return fail(RTQ_ERR_INVALID_ARGUMENT, "NULL argument");
  if (!b->built) return fail(RTQ_ERR_INVALID_ARGUMENT, "builder not built");
  return guarded([&] {
    rtq::saveCache(path, b->impl);
    return RTQ_OK;
  });
}

rtq_status rtq_cache_load(const char* path, uint64_t expected_hash, rtq_cached_scene** out) {
  if (!path || !out) return fail(RTQ_ERR_INVALID_ARGUMENT, "NULL argument");
  return guarded([&] {
    rtq::CachedScene c;
    switch (rtq::loadCache(path, expected_hash, c)) {
      case 0: break;
      case 1: return fail(RTQ_ERR_CACHE_MISS, "no readable cache file");
      case 3: return fail(RTQ_ERR_CACHE_MISS, "cache file built from other inputs");
      default: return fail(RTQ_ERR_BAD_BLOB, "not a scene cache of this version, truncated or corrupt");
    }
    *out = new rtq_cached_scene{std::move(c)};
    return RTQ_OK;
  });
}

rtq_status rtq_cached_scene_view(const rtq_cached_scene* c, const void** blob, size_t* bytes, rtq_ranges* ranges,
                                 const float** positions, uint32_t* num_vertices, uint64_t* hash) {
  if (!c) return fail(RTQ_ERR_INVALID_ARGUMENT, "NULL cached scene");
  if (blob) *blob = c->impl.blob.data();
  if (bytes) *bytes = c->impl.blob.size();
  if (ranges) {
    static_assert(sizeof(rtq::BlobRanges) == sizeof(rtq_ranges), "range layout");
    std::memcpy(ranges, &c->impl.ranges, sizeof(rtq_ranges));
  }
  if (positions) *positions = c->impl.positions.empty() ? nullptr : &c->impl.positions[0].x;
  if (num_vertices) *num_vertices = uint32_t(c->impl.positions.size());
  if (hash) *hash = c->impl.key;
  return RTQ_OK;
}

rtq_status rtq_cached_scene_free(rtq_cached_scene* c) {
  delete c;
  return RTQ_OK;
}

rtq_status rtq_probe_directions(uint64_t seed, uint32_t count, float* out) {
  if (!out || count == 0) return fail(RTQ_ERR_INVALID_ARGUMENT, "bad arguments");
  uint32_t side = 1;
  while (side * side < count) ++side;
  if (side * side != count) return fail(RTQ_ERR_INVALID_ARGUMENT, "count must be a square");
  rtq::probeDirections(seed, count, out);
  return RTQ_OK;
}

}  // extern "C"
