// C++ host interface over the rtq C ABI (include/rtq.h, include/rtq_build.h), named after the
// reference classes it stands in for so that renderer code reads the same:
//   BVHBuildParameters  Renderer/BVHBuilder.h:53-61
//   BVHBuilder          Renderer/BVHBuilder.h:63-101 (with BVHGeometry's vertex store, BVHGeometry.h:6-32)
//   BVHData             Renderer/BVHData.h:13-34 (build, then upload-BVH)
//   RayTracingPass      Renderer/raytracingPass.h:35 (tracePass = dispatch-rays) plus read-hits
// Errors: the reference has none in release builds (TIM_ASSERT is empty); here every failing call
// throws rtq::Error carrying the rtq_status.
#pragma once
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "rtq.h"
#include "rtq_build.h"

namespace rtq {

class Error : public std::runtime_error {
 public:
  Error(rtq_status s, const std::string& what) : std::runtime_error(what), status(s) {}
  rtq_status status;
};

inline void check(rtq_status s, const char* what) {
  if (s != RTQ_OK) throw Error(s, std::string(what) + ": " + rtq_status_string(s) + " " + rtq_last_error());
}
inline void checkBuild(rtq_status s, const char* what) {
  if (s != RTQ_OK) throw Error(s, std::string(what) + ": " + rtq_build_last_error());
}

struct vec3 {
  float x, y, z;
};

struct BVHBuildParameters {
  uint32_t maxDepth = 25;
  uint32_t minObjPerNode = 4;
  uint32_t minObjGain = 4;
  float expandNodeVolumeThreshold = 0.25f;
  float expandNodeDimensionFactor = 0.5f;
};

class Device {
 public:
  explicit Device(int ordinal = 0) { check(rtq_device_create(ordinal, &h_), "rtq_device_create"); }
  ~Device() { rtq_device_destroy(h_); }
  Device(const Device&) = delete;
  Device& operator=(const Device&) = delete;
  rtq_device* handle() const { return h_; }
  void* alloc(size_t bytes) {
    void* p = nullptr;
    check(rtq_device_alloc(h_, bytes, &p), "rtq_device_alloc");
    return p;
  }
  void free(void* p) { rtq_device_free(h_, p); }

 private:
  rtq_device* h_ = nullptr;
};

class BVHBuilder {
 public:
  BVHBuilder() { checkBuild(rtq_builder_create(&h_), "rtq_builder_create"); }
  ~BVHBuilder() { rtq_builder_destroy(h_); }
  BVHBuilder(const BVHBuilder&) = delete;
  BVHBuilder& operator=(const BVHBuilder&) = delete;

  // BVHGeometry::addTriangleList: appends vertex positions, returns the vertex offset.
  uint32_t addTriangleList(uint32_t numVertex, const vec3* positions) {
    uint32_t off = 0;
    checkBuild(rtq_builder_add_vertices(h_, &positions[0].x, numVertex, &off), "rtq_builder_add_vertices");
    return off;
  }
  // BVHBuilder::addTriangleList: one Lambertian material per call.
  void addTriangleList(uint32_t vertexOffset, uint32_t numTriangle, const uint32_t* indexData,
                       vec3 color = {0.7f, 0.7f, 0.7f}) {
    const float c[3] = {color.x, color.y, color.z};
    checkBuild(rtq_builder_add_triangle_list(h_, vertexOffset, numTriangle, indexData, c),
               "rtq_builder_add_triangle_list");
  }
  void addSphereLight(vec3 pos, float radius, vec3 color, float sphereRadius) {
    const float p[3] = {pos.x, pos.y, pos.z}, c[3] = {color.x, color.y, color.z};
    checkBuild(rtq_builder_add_sphere_light(h_, p, radius, c, sphereRadius), "rtq_builder_add_sphere_light");
  }
  void setParameters(const BVHBuildParameters& p) {
    rtq_build_params q{p.maxDepth, p.minObjPerNode, p.minObjGain, p.expandNodeVolumeThreshold,
                       p.expandNodeDimensionFactor};
    checkBuild(rtq_builder_set_parameters(h_, &q), "rtq_builder_set_parameters");
  }
  void build(bool useMultipleThreads) {
    checkBuild(rtq_builder_build(h_, useMultipleThreads ? 16 : 1), "rtq_builder_build");
  }
  uint32_t getBvhGpuSize() const {
    size_t n = 0;
    checkBuild(rtq_builder_gpu_size(h_, &n), "rtq_builder_gpu_size");
    return uint32_t(n);
  }
  void fillGpuBuffer(void* data, rtq_ranges& ranges) const {
    checkBuild(rtq_builder_fill_gpu_buffer(h_, data, getBvhGpuSize(), &ranges), "rtq_builder_fill_gpu_buffer");
  }
  std::pair<const float*, uint32_t> positions() const {
    const float* p = nullptr;
    uint32_t n = 0;
    checkBuild(rtq_builder_positions(h_, &p, &n), "rtq_builder_positions");
    return {p, n};
  }
  rtq_builder* handle() const { return h_; }

 private:
  rtq_builder* h_ = nullptr;
};

class BVHData {
 public:
  explicit BVHData(Device& dev) : dev_(dev) {}
  ~BVHData() { rtq_scene_destroy(scene_); }
  BVHData(const BVHData&) = delete;
  BVHData& operator=(const BVHData&) = delete;

  // BVHData::build (BVHData.cpp:28-48): build, size, fill, upload.
  void build(BVHBuilder& builder, const BVHBuildParameters& params) {
    builder.setParameters(params);
    builder.build(true);
    std::vector<uint8_t> blob(builder.getBvhGpuSize());
    rtq_ranges r{};
    builder.fillGpuBuffer(blob.data(), r);
    auto pos = builder.positions();
    upload(blob.data(), blob.size(), r, pos.first, pos.second);
  }
  // upload-BVH for a blob produced elsewhere (e.g. by the reference's own builder).
  void upload(const void* blob, size_t bytes, const rtq_ranges& ranges, const float* positions, uint32_t numVertices) {
    rtq_scene_destroy(scene_);
    scene_ = nullptr;
    check(rtq_scene_upload(dev_.handle(), blob, bytes, &ranges, positions, numVertices, &scene_), "rtq_scene_upload");
  }
  rtq_scene* scene() const { return scene_; }
  rtq_scene_info info() const {
    rtq_scene_info i{};
    check(rtq_scene_get_info(scene_, &i), "rtq_scene_get_info");
    return i;
  }

 private:
  Device& dev_;
  rtq_scene* scene_ = nullptr;
};

class RayTracingPass {
 public:
  RayTracingPass(Device& dev, uint32_t width, uint32_t height) : dev_(dev), w_(width), h_(height) {
    hits_ = static_cast<rtq_hit*>(dev_.alloc(sizeof(rtq_hit) * size_t(w_) * h_));
    nids_ = static_cast<uint32_t*>(dev_.alloc(sizeof(uint32_t) * size_t(w_) * h_));
  }
  ~RayTracingPass() {
    dev_.free(hits_);
    dev_.free(nids_);
    if (records_) dev_.free(records_);
  }
  RayTracingPass(const RayTracingPass&) = delete;
  RayTracingPass& operator=(const RayTracingPass&) = delete;

  // tracePass (raytracingPass.cpp:89-110): one primary ray per pixel, closest hit.
  void tracePass(const BVHData& bvh, const rtq_frustum& frustum, void* stream = nullptr, uint32_t flags = 0) {
    check(rtq_trace_camera(bvh.scene(), &frustum, w_, h_, flags, hits_, nids_, stream), "rtq_trace_camera");
  }
  // read-hits (no reference counterpart).
  std::vector<rtq_hit> readHits(void* stream = nullptr) const {
    std::vector<rtq_hit> out(size_t(w_) * h_);
    check(rtq_read_hits(dev_.handle(), hits_, out.data(), out.size() * sizeof(rtq_hit), stream), "rtq_read_hits");
    return out;
  }
  std::vector<uint32_t> readLeafNids(void* stream = nullptr) const {
    std::vector<uint32_t> out(size_t(w_) * h_);
    check(rtq_read_hits(dev_.handle(), nids_, out.data(), out.size() * 4, stream), "rtq_read_hits");
    return out;
  }
  const rtq_hit* deviceHits() const { return hits_; }
  const uint32_t* deviceLeafNids() const { return nids_; }

  // The camera pass with TRACING_STEP (cameraPass.comp:69-75): m_tracingResult's 2 x uvec4 per pixel
  // (hit triangle words + floatBits(t), shadow mask + probe mask). lpf = NULL leaves the probe mask 0.
  void traceRecord(const BVHData& bvh, const rtq_frustum& frustum, const float sunDir[3], uint32_t shadowFlags,
                   const rtq_lpf* lpf, void* stream = nullptr, uint32_t flags = 0) {
    if (!records_) records_ = static_cast<uint32_t*>(dev_.alloc(32 * size_t(w_) * h_));
    check(rtq_trace_camera_record(bvh.scene(), &frustum, w_, h_, flags, sunDir, shadowFlags, lpf, records_, stream),
          "rtq_trace_camera_record");
  }
  std::vector<uint32_t> readRecords(void* stream = nullptr) const {
    std::vector<uint32_t> out(8 * size_t(w_) * h_);
    if (records_) check(rtq_read_hits(dev_.handle(), records_, out.data(), out.size() * 4, stream), "rtq_read_hits");
    return out;
  }
  const uint32_t* deviceRecords() const { return records_; }

 private:
  Device& dev_;
  uint32_t w_, h_;
  rtq_hit* hits_ = nullptr;
  uint32_t* nids_ = nullptr;
  uint32_t* records_ = nullptr;
};

// LightProbFieldPass::traceLightProbField (lightProbFieldPass.cpp:81-100): the probe field's tracing
// step, 64 rays per probe, records laid out as the pass's tracing-result buffer.
class LightProbFieldPass {
 public:
  LightProbFieldPass(Device& dev, const rtq_lpf& lpf) : dev_(dev), lpf_(lpf) {
    samples_ = size_t(lpf.resolution[0]) * lpf.resolution[1] * lpf.resolution[2] * RTQ_PROBE_RAYS;
    records_ = static_cast<uint32_t*>(dev_.alloc(32 * samples_));
  }
  ~LightProbFieldPass() { dev_.free(records_); }
  LightProbFieldPass(const LightProbFieldPass&) = delete;
  LightProbFieldPass& operator=(const LightProbFieldPass&) = delete;

  // dirs: 64 x (x, y, z, 0), e.g. from rtq_probe_directions
  void traceLightProbField(const BVHData& bvh, const float dirs[4 * 64], const float sunDir[3],
                           uint32_t shadowFlags = RTQ_SHADOW_SUN, void* stream = nullptr) {
    check(rtq_trace_probe_field(bvh.scene(), &lpf_, dirs, sunDir, shadowFlags, records_, stream),
          "rtq_trace_probe_field");
  }
  std::vector<uint32_t> readRecords(void* stream = nullptr) const {
    std::vector<uint32_t> out(8 * samples_);
    check(rtq_read_hits(dev_.handle(), records_, out.data(), out.size() * 4, stream), "rtq_read_hits");
    return out;
  }
  const uint32_t* deviceRecords() const { return records_; }

 private:
  Device& dev_;
  rtq_lpf lpf_;
  size_t samples_ = 0;
  uint32_t* records_ = nullptr;
};

}  // namespace rtq
