// C ABI of the gfx950 ray-query library (include/rtq.h): device handles, blob upload into the
// device layout, dispatch of the traversal kernels, read-back, shadow masks and light culling.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <type_traits>
#include <unordered_map>
#include <vector>

#include <mutex>

#include <rocprim/device/device_radix_sort.hpp>

#include "../../../include/rtq.h"
#include "scene_layout.h"
#include "stack_class.h"
#include "kernels.h"
#include "layout.h"

struct rtq_device {
  int ordinal = 0;
  int numCU = 0;
  uint32_t persistent = 0;
  uint32_t blocksPerCU = 0;  // 0 = automatic (dispatch picks 7, or 12 for two-level scenes)
  uint32_t sortRays = 0;
  uint32_t leafWait = 0;  // 0: the kernels' defaults (rtq::kGridWait, kGridAnyWait, kRefillWait, ...)
  // ray-sort workspace (keys, values and rocPRIM's temporary storage), grown on demand and shared by
  // every sorting dispatch on this device; `sortDone` marks the end of the last dispatch that read it
  std::mutex sortMu;
  void* sortBuf = nullptr;
  size_t sortBytes = 0;
  hipEvent_t sortDone = nullptr;
  bool sortUsed = false;
  // this GPU's copy of the no-allocation scratch array (see `scratch` below), or nullptr with the
  // error its lookup returned
  uint32_t* scratch = nullptr;
  hipError_t scratchError = hipSuccess;
};

// Scratch of rtq_cull_lights_noalloc: a fixed device array, so that the call allocates nothing.
// Every GPU the library runs on holds its own copy, at its own address, until the process ends. In
// host code `scratch` names a host-side stand-in, not device memory: host code only ever uses the
// address hipGetSymbolAddress returns for the current device (rtq_device::scratch).
constexpr uint64_t kNoAllocScratchWords = uint64_t(4) << 20;  // 16 MiB
static __device__ __align__(256) uint32_t scratch[kNoAllocScratchWords];

// Work counters for persistent dispatches: a small ring so that dispatches in flight on different
// streams do not share a counter.
constexpr uint32_t kCounterRing = 64;

struct rtq_scene {
  rtq_device* dev = nullptr;
  float boxMin[3] = {0, 0, 0}, boxMax[3] = {0, 0, 0};  // bounds of the vertex positions (ray sort keys)
  rtq::DevScene view{};
  rtq_scene_info info{};
  std::vector<void*> allocations;
  uint32_t* counters = nullptr;  // kCounterRing groups of rtq::kXcds words, in `allocations`
  std::atomic<uint32_t> nextCounter{0};
};

namespace {

thread_local std::string g_error;

rtq_status fail(rtq_status s, const std::string& msg) {
  g_error = msg;
  return s;
}

rtq_status hipFail(hipError_t e, const char* what) {
  return fail(e == hipErrorOutOfMemory ? RTQ_ERR_OUT_OF_MEMORY : RTQ_ERR_HIP,
              std::string(what) + ": " + hipGetErrorString(e));
}

#define RTQ_HIP(call)                              \
  do {                                             \
    hipError_t e_ = (call);                        \
    if (e_ != hipSuccess) return hipFail(e_, #call); \
  } while (0)

template <class T>
rtq_status uploadArray(rtq_scene* s, const std::vector<T>& host, const T** dptr) {
  size_t bytes = std::max<size_t>(host.size() * sizeof(T), 16);
  void* p = nullptr;
  RTQ_HIP(hipMalloc(&p, bytes));
  s->allocations.push_back(p);
  if (!host.empty()) RTQ_HIP(hipMemcpy(p, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice));
  s->info.device_bytes += bytes;
  *dptr = static_cast<const T*>(p);
  return RTQ_OK;
}

void freeScene(rtq_scene* s) {
  if (!s) return;
  if (s->dev) (void)hipSetDevice(s->dev->ordinal);
  for (void* p : s->allocations) (void)hipFree(p);
  delete s;
}

// Instantiates f with the LDS stack class of a tree of this depth (stack_class.h); false, with
// nothing launched, for a tree deeper than any class holds.
template <class F>
bool withStack(uint32_t depth, F&& f) {
  switch (rtq::gridStackClass(depth)) {
    case 22: f(std::integral_constant<int, 22>{}); return true;
    case 24: f(std::integral_constant<int, 24>{}); return true;
    case 32: f(std::integral_constant<int, 32>{}); return true;
    case 64: f(std::integral_constant<int, 64>{}); return true;
  }
  return false;
}

// Device address of `scratch` on GPU `ordinal` (which must be current), looked up once per GPU and
// kept for the process.
hipError_t scratchAddress(int ordinal, uint32_t** out) {
  static std::mutex mu;
  static std::unordered_map<int, std::pair<uint32_t*, hipError_t>> known;
  std::lock_guard<std::mutex> g(mu);
  auto it = known.find(ordinal);
  if (it == known.end()) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    it = known.emplace(ordinal, std::make_pair(e == hipSuccess ? static_cast<uint32_t*>(p) : nullptr, e)).first;
  }
  *out = it->second.first;
  return it->second.second;
}

// Two-level scenes are traced by any-hit dispatches only: the reference's closest-hit walk of a
// top-level tree (traverseTlas / tlas_collide, bvhTraversal_inline.glsl:3-78, bvhCollision.glsl:
// 287-319) calls HitBox, whose only definition is commented out (collision.glsl:175-195).
rtq_status twoLevelOnlyAny() {
  return fail(RTQ_ERR_INVALID_ARGUMENT,
              "two-level scene: only rtq_dispatch_rays with RTQ_TRACE_ANY (the reference has no closest-hit TLAS walk)");
}

#define RTQ_WITH_STACK(depth, ...)                                                       \
  do {                                                                                   \
    if (!withStack((depth), __VA_ARGS__)) return fail(RTQ_ERR_LIMIT, "tree deeper than 64"); \
  } while (0)

}  // namespace

extern "C" {

const char* rtq_status_string(rtq_status s) {
  switch (s) {
    case RTQ_OK: return "RTQ_OK";
    case RTQ_ERR_INVALID_ARGUMENT: return "RTQ_ERR_INVALID_ARGUMENT";
    case RTQ_ERR_HIP: return "RTQ_ERR_HIP";
    case RTQ_ERR_OUT_OF_MEMORY: return "RTQ_ERR_OUT_OF_MEMORY";
    case RTQ_ERR_LIMIT: return "RTQ_ERR_LIMIT";
    case RTQ_ERR_BAD_BLOB: return "RTQ_ERR_BAD_BLOB";
    case RTQ_ERR_NO_DEVICE: return "RTQ_ERR_NO_DEVICE";
    case RTQ_ERR_BUDGET: return "RTQ_ERR_BUDGET";
    case RTQ_ERR_QUEUE_FULL: return "RTQ_ERR_QUEUE_FULL";
    case RTQ_ERR_CANCELLED: return "RTQ_ERR_CANCELLED";
    case RTQ_ERR_TIMEOUT: return "RTQ_ERR_TIMEOUT";
    case RTQ_ERR_INTERNAL: return "RTQ_ERR_INTERNAL";
    case RTQ_ERR_CACHE_MISS: return "RTQ_ERR_CACHE_MISS";
  }
  return "RTQ_ERR_UNKNOWN";
}

const char* rtq_last_error(void) { return g_error.c_str(); }

rtq_status rtq_device_create(int hip_ordinal, rtq_device** out) {
  if (!out) return fail(RTQ_ERR_INVALID_ARGUMENT, "out is NULL");
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0) return fail(RTQ_ERR_NO_DEVICE, "no HIP device visible");
  if (hip_ordinal < 0 || hip_ordinal >= count) return fail(RTQ_ERR_INVALID_ARGUMENT, "device ordinal out of range");
  RTQ_HIP(hipSetDevice(hip_ordinal));
  hipDeviceProp_t prop;
  RTQ_HIP(hipGetDeviceProperties(&prop, hip_ordinal));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(RTQ_ERR_NO_DEVICE, std::string("device is ") + prop.gcnArchName + ", library is built for gfx950");
  rtq_device* d = new rtq_device();
  d->ordinal = hip_ordinal;
  d->numCU = prop.multiProcessorCount;
  d->scratchError = scratchAddress(hip_ordinal, &d->scratch);
  *out = d;
  return RTQ_OK;
}

rtq_status rtq_device_destroy(rtq_device* dev) {
  if (dev && (dev->sortBuf || dev->sortDone)) {
    (void)hipSetDevice(dev->ordinal);
    (void)hipDeviceSynchronize();
    if (dev->sortBuf) (void)hipFree(dev->sortBuf);
    if (dev->sortDone) (void)hipEventDestroy(dev->sortDone);
  }
  delete dev;
  return RTQ_OK;
}

rtq_status rtq_device_set_option(rtq_device* dev, uint32_t option, uint32_t value) {
  if (!dev) return fail(RTQ_ERR_INVALID_ARGUMENT, "NULL device");
  switch (option) {
    case RTQ_OPTION_PERSISTENT:
      if (value > 2) return fail(RTQ_ERR_INVALID_ARGUMENT, "persistent is 0, 1 or 2");
      dev->persistent = value;
      return RTQ_OK;
    case RTQ_OPTION_BLOCKS_PER_CU:
      if (value > 16) return fail(RTQ_ERR_INVALID_ARGUMENT, "blocks per CU is 0 (automatic) or 1..16");
      dev->blocksPerCU = value;
      return RTQ_OK;
    case RTQ_OPTION_SORT_RAYS:
      if (value > 1) return fail(RTQ_ERR_INVALID_ARGUMENT, "sort is 0 or 1");
      dev->sortRays = value;
      return RTQ_OK;
    case RTQ_OPTION_LEAF_WAIT:
      if (value > 64) return fail(RTQ_ERR_INVALID_ARGUMENT, "leaf wait is 0 (default) or 1..64");
      dev->leafWait = value;
      return RTQ_OK;
  }
  return fail(RTQ_ERR_INVALID_ARGUMENT, "unknown option");
}

rtq_status rtq_device_get_option(const rtq_device* dev, uint32_t option, uint32_t* value) {
  if (!dev || !value) return fail(RTQ_ERR_INVALID_ARGUMENT, "NULL argument");
  switch (option) {
    case RTQ_OPTION_PERSISTENT: *value = dev->persistent; return RTQ_OK;
    case RTQ_OPTION_BLOCKS_PER_CU: *value = dev->blocksPerCU; return RTQ_OK;
    case RTQ_OPTION_SORT_RAYS: *value = dev->sortRays; return RTQ_OK;
    case RTQ_OPTION_LEAF_WAIT: *value = dev->leafWait; return RTQ_OK;
  }
  return fail(RTQ_ERR_INVALID_ARGUMENT, "unknown option");
}

rtq_status rtq_device_alloc(rtq_device* dev, size_t bytes, void** d_ptr) {
  if (!dev || !d_ptr) return fail(RTQ_ERR_INVALID_ARGUMENT, "NULL argument");
  RTQ_HIP(hipSetDevice(dev->ordinal));
  RTQ_HIP(hipMalloc(d_ptr, std::max<size_t>(bytes, 1)));
  return RTQ_OK;
}

rtq_status rtq_device_free(rtq_device* dev, void* d_ptr) {
  if (!dev) return fail(RTQ_ERR_INVALID_ARGUMENT, "NULL device");
  RTQ_HIP(hipSetDevice(dev->ordinal));
  if (d_ptr) RTQ_HIP(hipFree(d_ptr));
  return RTQ_OK;
}

rtq_status rtq_host_alloc(rtq_device* dev, size_t bytes, void** h_ptr) {
  if (!dev || !h_ptr) return fail(RTQ_ERR_INVALID_ARGUMENT, "NULL argument");
  RTQ_HIP(hipSetDevice(dev->ordinal));
  RTQ_HIP(hipHostMalloc(h_ptr, std::max<size_t>(bytes, 1), hipHostMallocDefault));
  return RTQ_OK;
}

rtq_status rtq_host_free(rtq_device* dev, void* h_ptr) {
  if (!dev) return fail(RTQ_ERR_INVALID_ARGUMENT, "NULL device");
  RTQ_HIP(hipSetDevice(dev->ordinal));
  if (h_ptr) RTQ_HIP(hipHostFree(h_ptr));
  return RTQ_OK;
}

rtq_status rtq_copy_to_device(rtq_device* dev, void* d_dst, const void* h_src, size_t bytes, void* stream) {
  if (!dev || (!d_dst && bytes) || (!h_src && bytes)) return fail(RTQ_ERR_INVALID_ARGUMENT, "NULL argument");
  RTQ_HIP(hipSetDevice(dev->ordinal));
  RTQ_HIP(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, static_cast<hipStream_t>(stream)));
  RTQ_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
  return RTQ_OK;
}

rtq_status rtq_device_copy(rtq_device* dev, void* d_dst, const void* d_src, size_t bytes, void* stream) {
  if (!dev || ((!d_dst || !d_src) && bytes)) return fail(RTQ_ERR_INVALID_ARGUMENT, "NULL argument");
  if (bytes % 16 || reinterpret_cast<uintptr_t>(d_dst) % 16 || reinterpret_cast<uintptr_t>(d_src) % 16)
    return fail(RTQ_ERR_INVALID_ARGUMENT, "copy needs 16-byte aligned pointers and size");
  if (bytes == 0) return RTQ_OK;
  RTQ_HIP(hipSetDevice(dev->ordinal));
  hipLaunchKernelGGL(rtq::kCopy16, dim3(8192), dim3(256), 0, static_cast<hipStream_t>(stream),
                     static_cast<uint4*>(d_dst), static_cast<const uint4*>(d_src), bytes / 16);
  RTQ_HIP(hipGetLastError());
  return RTQ_OK;
}

rtq_status rtq_stream_create(rtq_device* dev, void** stream) {
  if (!dev || !stream) return fail(RTQ_ERR_INVALID_ARGUMENT, "NULL argument");
  RTQ_HIP(hipSetDevice(dev->ordinal));
  hipStream_t s = nullptr;
  RTQ_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = s;
  return RTQ_OK;
}

rtq_status rtq_stream_destroy(rtq_device* dev, void* stream) {
  if (!dev) return fail(RTQ_ERR_INVALID_ARGUMENT, "NULL device");
  RTQ_HIP(hipSetDevice(dev->ordinal));
  if (stream) RTQ_HIP(hipStreamDestroy(static_cast<hipStream_t>(stream)));
  return RTQ_OK;
}

rtq_status rtq_stream_synchronize(rtq_device* dev, void* stream) {
  if (!dev) return fail(RTQ_ERR_INVALID_ARGUMENT, "NULL device");
  RTQ_HIP(hipSetDevice(dev->ordinal));
  RTQ_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
  return RTQ_OK;
}

rtq_status rtq_event_create(rtq_device* dev, void** event) {
  if (!dev || !event) return fail(RTQ_ERR_INVALID_ARGUMENT, "NULL argument");
  RTQ_HIP(hipSetDevice(dev->ordinal));
  hipEvent_t e = nullptr;
  RTQ_HIP(hipEventCreate(&e));
  *event = e;
  return RTQ_OK;
}

rtq_status rtq_event_destroy(rtq_device* dev, void* event) {
  if (!dev) return fail(RTQ_ERR_INVALID_ARGUMENT, "NULL device");
  if (event) RTQ_HIP(hipEventDestroy(static_cast<hipEvent_t>(event)));
  return RTQ_OK;
}

rtq_status rtq_event_record(void* event, void* stream) {
  if (!event) return fail(RTQ_ERR_INVALID_ARGUMENT, "NULL event");
  RTQ_HIP(hipEventRecord(static_cast<hipEvent_t>(event), static_cast<hipStream_t>(stream)));
  return RTQ_OK;
}

rtq_status rtq_event_elapsed_ms(void* start, void* end, float* ms) {
  if (!start || !end || !ms) return fail(RTQ_ERR_INVALID_ARGUMENT, "NULL argument");
  RTQ_HIP(hipEventSynchronize(static_cast<hipEvent_t>(end)));
  RTQ_HIP(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(end)));
  return RTQ_OK;
}

rtq_status rtq_read_hits(rtq_device* dev, const void* d_src, void* h_dst, size_t bytes, void* stream) {
  if (!dev || (!d_src && bytes) || (!h_dst && bytes)) return fail(RTQ_ERR_INVALID_ARGUMENT, "NULL argument");
  RTQ_HIP(hipSetDevice(dev->ordinal));
  RTQ_HIP(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, static_cast<hipStream_t>(stream)));
  RTQ_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
  return RTQ_OK;
}

// upload-BVH: decode the blob and build the device layout on the host (scene_layout.h), then copy it
// to HBM where it stays resident.
rtq_status rtq_scene_upload(rtq_device* dev, const void* blobPtr, size_t bytes, const rtq_ranges* ranges,
                            const float* positions, uint32_t numVertices, rtq_scene** out) {
  if (!dev || !blobPtr || !ranges || !out || (!positions && numVertices))
    return fail(RTQ_ERR_INVALID_ARGUMENT, "NULL argument");
  rtq::HostScene h;
  const char* why = "";
  rtq_status st;
  try {
    st = rtq::buildHostScene(static_cast<const uint8_t*>(blobPtr), bytes, ranges, positions, numVertices, h, &why);
  } catch (const std::bad_alloc&) {
    return fail(RTQ_ERR_OUT_OF_MEMORY, "host allocation failed");
  }
  if (st != RTQ_OK) return fail(st, why);

  rtq_scene* s = new rtq_scene();
  s->dev = dev;
  auto bail = [&](rtq_status e) {
    freeScene(s);
    return e;
  };
  hipError_t e = hipSetDevice(dev->ordinal);
  if (e != hipSuccess) return bail(hipFail(e, "hipSetDevice"));
  if ((st = uploadArray(s, h.nodes, &s->view.nodes)) != RTQ_OK) return bail(st);
  if ((st = uploadArray(s, h.tris, &s->view.tris)) != RTQ_OK) return bail(st);
  if ((st = uploadArray(s, h.leafOff, &s->view.leafOff)) != RTQ_OK) return bail(st);
  if ((st = uploadArray(s, h.leafData, &s->view.leafData)) != RTQ_OK) return bail(st);
  if ((st = uploadArray(s, h.lights, &s->view.lights)) != RTQ_OK) return bail(st);
  if ((st = uploadArray(s, h.lightKinds, &s->view.lightKind)) != RTQ_OK) return bail(st);
  if ((st = uploadArray(s, h.triangles, &s->view.triangles)) != RTQ_OK) return bail(st);
  if (h.twoLevel && (st = uploadArray(s, h.insts, &s->view.insts)) != RTQ_OK) return bail(st);
  {
    const uint32_t* dc = nullptr;
    if ((st = uploadArray(s, std::vector<uint32_t>(kCounterRing * rtq::kXcds, 0u), &dc)) != RTQ_OK) return bail(st);
    s->counters = const_cast<uint32_t*>(dc);
  }
  for (int k = 0; k < 3; ++k) {
    s->boxMin[k] = numVertices ? positions[k] : 0.0f;
    s->boxMax[k] = s->boxMin[k];
  }
  for (uint32_t v = 0; v < numVertices; ++v)
    for (int k = 0; k < 3; ++k) {
      const float x = positions[3 * size_t(v) + k];
      if (x < s->boxMin[k]) s->boxMin[k] = x;
      if (x > s->boxMax[k]) s->boxMax[k] = x;
    }
  s->view.root = h.root;
  s->view.numNodes = h.numNodes;
  s->view.leafWords = h.leafWords;
  s->view.numLights = h.numLights;
  s->view.fastBoxes = h.fastBoxes ? 1u : 0u;
  s->view.orderedBoxes = h.orderedBoxes ? 1u : 0u;
  s->view.topSlots = h.topSlots;
  s->view.top = nullptr;
  s->view.twoLevel = h.twoLevel ? 1u : 0u;
  if (!h.twoLevel) s->view.insts = nullptr;
  s->info.num_nodes = h.numNodes;
  s->info.num_inner = h.numInner;
  s->info.num_leaf_slots = uint32_t(h.tris.size());
  s->info.num_triangles = h.numTris;
  s->info.num_lights = h.numLights;
  s->info.max_depth = h.maxDepth;
  *out = s;
  return RTQ_OK;
}

rtq_status rtq_scene_destroy(rtq_scene* scene) {
  freeScene(scene);
  return RTQ_OK;
}

rtq_status rtq_scene_get_info(const rtq_scene* scene, rtq_scene_info* info) {
  if (!scene || !info) return fail(RTQ_ERR_INVALID_ARGUMENT, "NULL argument");
  *info = scene->info;
  return RTQ_OK;
}

// Sort keys and rocPRIM radix sort on `st`; *perm receives the device permutation (lane order -> ray).
// The lock keeps the device's workspace for this dispatch until its kernels are enqueued, and `st`
// first waits for the event the previous sorting dispatch recorded after its traversal kernel (see
// sortRelease), so sorting dispatches on different streams take the workspace one after another.
static rtq_status sortRays(rtq_scene* s, const float4* rays, uint32_t n, hipStream_t st,
                           std::unique_lock<std::mutex>& lock, const uint32_t** perm) {
  rtq_device* dev = s->dev;
  lock = std::unique_lock<std::mutex>(dev->sortMu);
  if (!dev->sortDone) RTQ_HIP(hipEventCreateWithFlags(&dev->sortDone, hipEventDisableTiming));
  if (dev->sortUsed) RTQ_HIP(hipStreamWaitEvent(st, dev->sortDone, 0));
  size_t tempBytes = 0;
  RTQ_HIP(rocprim::radix_sort_pairs(nullptr, tempBytes, static_cast<const uint32_t*>(nullptr),
                                    static_cast<uint32_t*>(nullptr), static_cast<const uint32_t*>(nullptr),
                                    static_cast<uint32_t*>(nullptr), n, 0u, 30u, st));
  const size_t arr = (size_t(n) * 4 + 255) & ~size_t(255);
  const size_t need = 4 * arr + tempBytes;
  if (need > dev->sortBytes) {
    if (dev->sortBuf) {
      RTQ_HIP(hipEventSynchronize(dev->sortDone));
      RTQ_HIP(hipFree(dev->sortBuf));
      dev->sortBuf = nullptr;
      dev->sortBytes = 0;
    }
    RTQ_HIP(hipMalloc(&dev->sortBuf, need));
    dev->sortBytes = need;
  }
  uint8_t* base = static_cast<uint8_t*>(dev->sortBuf);
  uint32_t* keysIn = reinterpret_cast<uint32_t*>(base);
  uint32_t* keysOut = reinterpret_cast<uint32_t*>(base + arr);
  uint32_t* valsIn = reinterpret_cast<uint32_t*>(base + 2 * arr);
  uint32_t* valsOut = reinterpret_cast<uint32_t*>(base + 3 * arr);
  void* temp = base + 4 * arr;
  float sc[3];
  for (int k = 0; k < 3; ++k) {
    const float ext = s->boxMax[k] - s->boxMin[k];
    sc[k] = ext > 0.0f ? 512.0f / ext : 0.0f;
  }
  hipLaunchKernelGGL(rtq::kRayKeys, dim3(uint32_t((uint64_t(n) + 255) / 256)), dim3(256), 0, st, rays, n, s->boxMin[0], s->boxMin[1],
                     s->boxMin[2], sc[0], sc[1], sc[2], keysIn, valsIn);
  RTQ_HIP(hipGetLastError());
  RTQ_HIP(rocprim::radix_sort_pairs(temp, tempBytes, keysIn, keysOut, valsIn, valsOut, n, 0u, 30u, st));
  // marks the sort's end in case the traversal launch fails; sortRelease moves the mark past it
  RTQ_HIP(hipEventRecord(dev->sortDone, st));
  dev->sortUsed = true;
  *perm = valsOut;
  return RTQ_OK;
}

// After the traversal kernel that reads the permutation is enqueued on `st`: the next sorting
// dispatch waits for this point before it rewrites the workspace. Called with the lock held.
static rtq_status sortRelease(rtq_device* dev, hipStream_t st) {
  RTQ_HIP(hipEventRecord(dev->sortDone, st));
  dev->sortUsed = true;
  return RTQ_OK;
}

static rtq_status dispatch(rtq_scene* s, const rtq_ray* rays, uint32_t n, uint32_t flags, void* out, uint32_t* nid,
                           uint32_t* counters, void* stream) {
  if (!s) return fail(RTQ_ERR_INVALID_ARGUMENT, "NULL scene");
  if (n == 0) return RTQ_OK;
  if (!rays || !out) return fail(RTQ_ERR_INVALID_ARGUMENT, "NULL ray or output buffer");
  if (flags != RTQ_TRACE_CLOSEST && flags != RTQ_TRACE_ANY) return fail(RTQ_ERR_INVALID_ARGUMENT, "bad flags");
  RTQ_HIP(hipSetDevice(s->dev->ordinal));
  hipStream_t st = static_cast<hipStream_t>(stream);
  const dim3 grid(uint32_t((uint64_t(n) + rtq::kTraceBlock - 1) / rtq::kTraceBlock)), block(rtq::kTraceBlock);
  const float4* r = reinterpret_cast<const float4*>(rays);
  // two-level scenes take the any-hit kernels below, whose leaf phase enters BLAS instances
  if (s->view.twoLevel && (flags != RTQ_TRACE_ANY || counters)) return twoLevelOnlyAny();
  const uint32_t gridWait = s->dev->leafWait          ? s->dev->leafWait
                            : flags == RTQ_TRACE_ANY ? rtq::kGridAnyWait
                                                     : rtq::kGridWait;
  const uint32_t refillWait = s->dev->leafWait      ? s->dev->leafWait
                              : s->view.twoLevel    ? rtq::kTwoLevelRefillWait
                                                    : rtq::kRefillWait;
  // persistent blocks per CU: the option, or automatically 7 for one-level scenes (soup bounces with
  // 64-ray chunks: 727 / 753 / 752 / 752 Mrays/s at 6 / 7 / 8 / 10 blocks, and 4 % fewer bytes beyond the
  // L2 per ray at 7 than at 8, the span of rays in flight being smaller; with 256-ray chunks, whose wider
  // span crowded the L2, 602 / 600 / 561 at 7 / 8 / 10) and 12 for two-level scenes, whose any-hit walks
  // keep a quarter of their lanes busy and gain from every wave that fits (LDS holds 10 blocks with the
  // 32-entry stack: 949 / 972 / 978 Mrays/s at 8 / 10 / 12)
  const uint32_t blocksPerCU = s->dev->blocksPerCU ? s->dev->blocksPerCU : (s->view.twoLevel ? 12u : 7u);
  const uint32_t* perm = nullptr;
  std::unique_lock<std::mutex> sortLock;
  if (!counters && s->dev->sortRays) {
    rtq_status ss = sortRays(s, r, n, st, sortLock, &perm);
    if (ss != RTQ_OK) return ss;
  }
  if (!counters && s->dev->persistent == 2) {
    const uint64_t lanes = uint64_t(s->dev->numCU) * blocksPerCU * rtq::kTraceBlock;
    const dim3 pgrid(uint32_t(std::min<uint64_t>(lanes, uint64_t(n) + rtq::kTraceBlock - 1) / rtq::kTraceBlock));
    RTQ_WITH_STACK(s->info.max_depth, [&](auto S) {
      constexpr int K = decltype(S)::value;
      if (flags == RTQ_TRACE_CLOSEST)
        hipLaunchKernelGGL((rtq::kTraceWaveRange<K, false>), pgrid, block, 0, st, s->view, r, n, out, nid, perm,
                           refillWait);
      else
        hipLaunchKernelGGL((rtq::kTraceWaveRange<K, true>), pgrid, block, 0, st, s->view, r, n, out, nullptr, perm,
                           refillWait);
    });
    RTQ_HIP(hipGetLastError());
    return perm ? sortRelease(s->dev, st) : RTQ_OK;
  }
  if (!counters && s->dev->persistent) {
    uint32_t* counter = s->counters + rtq::kXcds * (s->nextCounter.fetch_add(1) % kCounterRing);
    hipLaunchKernelGGL(rtq::kResetCounter, dim3(1), dim3(64), 0, st, counter);
    const uint64_t lanes = uint64_t(s->dev->numCU) * blocksPerCU * rtq::kTraceBlock;
    const dim3 pgrid(uint32_t(std::min<uint64_t>(lanes, uint64_t(n) + rtq::kTraceBlock - 1) / rtq::kTraceBlock));
    RTQ_WITH_STACK(s->info.max_depth, [&](auto S) {
      constexpr int K = decltype(S)::value;
      if (flags == RTQ_TRACE_CLOSEST)
        hipLaunchKernelGGL((rtq::kTracePersistent<K, false>), pgrid, block, 0, st, s->view, r, n, out, nid, counter,
                           perm, refillWait);
      else
        hipLaunchKernelGGL((rtq::kTracePersistent<K, true>), pgrid, block, 0, st, s->view, r, n, out, nullptr,
                           counter, perm, refillWait);
    });
    RTQ_HIP(hipGetLastError());
    return perm ? sortRelease(s->dev, st) : RTQ_OK;
  }
  RTQ_WITH_STACK(s->info.max_depth, [&](auto S) {
    constexpr int K = decltype(S)::value;
    if (flags == RTQ_TRACE_CLOSEST) {
      float4* h = static_cast<float4*>(out);
      if (counters)
        hipLaunchKernelGGL((rtq::kTraceClosest<K, true>), grid, block, 0, st, s->view, r, n, h, nid, counters, nullptr,
                           gridWait);
      else
        hipLaunchKernelGGL((rtq::kTraceClosest<K, false>), grid, block, 0, st, s->view, r, n, h, nid, nullptr, perm,
                           gridWait);
    } else {
      uint32_t* o = static_cast<uint32_t*>(out);
      if (counters)
        hipLaunchKernelGGL((rtq::kTraceAny<K, true>), grid, block, 0, st, s->view, r, n, o, counters, nullptr, gridWait);
      else
        hipLaunchKernelGGL((rtq::kTraceAny<K, false>), grid, block, 0, st, s->view, r, n, o, nullptr, perm, gridWait);
    }
  });
  RTQ_HIP(hipGetLastError());
  return perm ? sortRelease(s->dev, st) : RTQ_OK;
}

rtq_status rtq_dispatch_rays(rtq_scene* scene, const rtq_ray* d_rays, uint32_t num_rays, uint32_t flags, void* d_out,
                             uint32_t* d_leaf_nid, void* stream) {
  return dispatch(scene, d_rays, num_rays, flags, d_out, flags == RTQ_TRACE_CLOSEST ? d_leaf_nid : nullptr, nullptr,
                  stream);
}

rtq_status rtq_dispatch_rays_counted(rtq_scene* scene, const rtq_ray* d_rays, uint32_t num_rays, uint32_t flags,
                                     void* d_out, uint32_t* d_leaf_nid, uint32_t* d_counters, void* stream) {
  if (!d_counters && num_rays) return fail(RTQ_ERR_INVALID_ARGUMENT, "NULL counters");
  return dispatch(scene, d_rays, num_rays, flags, d_out, flags == RTQ_TRACE_CLOSEST ? d_leaf_nid : nullptr,
                  d_counters, stream);
}

rtq_status rtq_trace_camera(rtq_scene* s, const rtq_frustum* f, uint32_t width, uint32_t height, uint32_t flags,
                            rtq_hit* d_hits, uint32_t* d_leaf_nid, void* stream) {
  if (!s || !f) return fail(RTQ_ERR_INVALID_ARGUMENT, "NULL scene or frustum");
  if (s->view.twoLevel) return twoLevelOnlyAny();
  if (width == 0 || height == 0) return RTQ_OK;
  if (!d_hits) return fail(RTQ_ERR_INVALID_ARGUMENT, "NULL hit buffer");
  if (flags & ~RTQ_CAMERA_REFERENCE_ROWS) return fail(RTQ_ERR_INVALID_ARGUMENT, "bad flags");
  if (width > (1u << 16) || height > (1u << 16)) return fail(RTQ_ERR_LIMIT, "frame larger than 65536 pixels a side");
  RTQ_HIP(hipSetDevice(s->dev->ordinal));
  rtq::Frustum4 fr;
  std::memcpy(&fr, f, sizeof(fr));
  const uint32_t blocks = ((width + 15u) / 16u) * ((height + 7u) / 8u);
  hipStream_t st = static_cast<hipStream_t>(stream);
  float4* h = reinterpret_cast<float4*>(d_hits);
  const uint32_t ref = flags & RTQ_CAMERA_REFERENCE_ROWS;
  RTQ_WITH_STACK(s->info.max_depth, [&](auto S) {
    constexpr int K = decltype(S)::value;
    hipLaunchKernelGGL((rtq::kTraceCamera<K>), dim3(blocks), dim3(rtq::kTraceBlock), 0, st, s->view, fr, width,
                       height, ref, h, d_leaf_nid);
  });
  RTQ_HIP(hipGetLastError());
  return RTQ_OK;
}

rtq_status rtq_trace_camera_record(rtq_scene* s, const rtq_frustum* f, uint32_t width, uint32_t height,
                                   uint32_t flags, const float sun_dir[3], uint32_t shadowFlags, const rtq_lpf* lpf,
                                   uint32_t* d_records, void* stream) {
  if (!s || !f) return fail(RTQ_ERR_INVALID_ARGUMENT, "NULL scene or frustum");
  if (s->view.twoLevel) return twoLevelOnlyAny();
  if (width == 0 || height == 0) return RTQ_OK;
  if (!d_records) return fail(RTQ_ERR_INVALID_ARGUMENT, "NULL record buffer");
  if (flags & ~RTQ_CAMERA_REFERENCE_ROWS) return fail(RTQ_ERR_INVALID_ARGUMENT, "bad flags");
  if ((shadowFlags & RTQ_SHADOW_SUN) && !sun_dir) return fail(RTQ_ERR_INVALID_ARGUMENT, "sun flag without direction");
  if (width > (1u << 16) || height > (1u << 16)) return fail(RTQ_ERR_LIMIT, "frame larger than 65536 pixels a side");
  rtq::LpfHeader h{};
  if (lpf) {
    std::memcpy(&h, lpf, sizeof(h));
    for (int k = 0; k < 3; ++k)
      if (h.res[k] < 2) return fail(RTQ_ERR_INVALID_ARGUMENT, "probe field resolution below 2");
  }
  float sd[3] = {0, 0, 0};
  if (sun_dir) std::memcpy(sd, sun_dir, sizeof(sd));
  RTQ_HIP(hipSetDevice(s->dev->ordinal));
  rtq::Frustum4 fr;
  std::memcpy(&fr, f, sizeof(fr));
  const uint32_t blocks = ((width + 15u) / 16u) * ((height + 7u) / 8u);
  hipStream_t st = static_cast<hipStream_t>(stream);
  uint4* rec = reinterpret_cast<uint4*>(d_records);
  const uint32_t ref = flags & RTQ_CAMERA_REFERENCE_ROWS;
  RTQ_WITH_STACK(s->info.max_depth, [&](auto S) {
    constexpr int K = decltype(S)::value;
    hipLaunchKernelGGL((rtq::kTraceRecord<K>), dim3(blocks), dim3(rtq::kTraceBlock), 0, st, s->view, fr, width, height,
                       ref, sd[0], sd[1], sd[2], shadowFlags, h, lpf ? 1u : 0u, rec);
  });
  RTQ_HIP(hipGetLastError());
  return RTQ_OK;
}

rtq_status rtq_trace_probe_field(rtq_scene* s, const rtq_lpf* lpf, const float dirs[4 * 64], const float sun_dir[3],
                                 uint32_t shadowFlags, uint32_t* d_records, void* stream) {
  if (!s || !lpf || !dirs) return fail(RTQ_ERR_INVALID_ARGUMENT, "NULL scene, probe field or directions");
  if (s->view.twoLevel) return twoLevelOnlyAny();
  if ((shadowFlags & RTQ_SHADOW_SUN) && !sun_dir) return fail(RTQ_ERR_INVALID_ARGUMENT, "sun flag without direction");
  rtq::LpfHeader h{};
  std::memcpy(&h, lpf, sizeof(h));
  if (h.res[0] < 4 || h.res[1] < 4 || h.res[2] == 0)
    return fail(RTQ_ERR_INVALID_ARGUMENT, "probe field needs resolution.x, .y >= 4 and .z >= 1");
  const uint64_t probes = uint64_t(h.res[0]) * h.res[1] * h.res[2];
  if (probes > (1ull << 24)) return fail(RTQ_ERR_LIMIT, "more than 2^24 probes");
  if (!d_records) return fail(RTQ_ERR_INVALID_ARGUMENT, "NULL record buffer");
  float sd[3] = {0, 0, 0};
  if (sun_dir) std::memcpy(sd, sun_dir, sizeof(sd));
  rtq::ProbeDirs pd;
  std::memcpy(&pd, dirs, sizeof(pd));
  RTQ_HIP(hipSetDevice(s->dev->ordinal));
  // one wave per (64-probe batch, direction), two waves per block
  const uint64_t waves = ((probes + 63u) / 64u) * rtq::kProbeRays;
  const uint32_t blocks = uint32_t((waves + 1u) / 2u);
  hipStream_t st = static_cast<hipStream_t>(stream);
  RTQ_WITH_STACK(s->info.max_depth, [&](auto S) {
    constexpr int K = decltype(S)::value;
    hipLaunchKernelGGL((rtq::kTraceProbes<K>), dim3(blocks), dim3(rtq::kTraceBlock), 0, st, s->view, h, pd,
                       uint32_t(probes), sd[0], sd[1], sd[2], shadowFlags, reinterpret_cast<uint4*>(d_records));
  });
  RTQ_HIP(hipGetLastError());
  return RTQ_OK;
}

rtq_status rtq_shadow_masks(rtq_scene* s, const rtq_ray* d_rays, const rtq_hit* d_hits, const uint32_t* d_leaf_nid,
                            uint32_t n, const float sun_dir[3], uint32_t flags, uint32_t* d_masks, void* stream) {
  if (!s) return fail(RTQ_ERR_INVALID_ARGUMENT, "NULL scene");
  if (s->view.twoLevel) return twoLevelOnlyAny();
  if (n == 0) return RTQ_OK;
  if (!d_rays || !d_hits || !d_leaf_nid || !d_masks) return fail(RTQ_ERR_INVALID_ARGUMENT, "NULL buffer");
  if ((flags & RTQ_SHADOW_SUN) && !sun_dir) return fail(RTQ_ERR_INVALID_ARGUMENT, "sun flag without direction");
  float sd[3] = {0, 0, 0};
  if (sun_dir) std::memcpy(sd, sun_dir, sizeof(sd));
  RTQ_HIP(hipSetDevice(s->dev->ordinal));
  hipStream_t st = static_cast<hipStream_t>(stream);
  const dim3 grid(uint32_t((uint64_t(n) + rtq::kTraceBlock - 1) / rtq::kTraceBlock)), block(rtq::kTraceBlock);
  const float4* r = reinterpret_cast<const float4*>(d_rays);
  const float4* h = reinterpret_cast<const float4*>(d_hits);
  RTQ_WITH_STACK(s->info.max_depth, [&](auto S) {
    constexpr int K = decltype(S)::value;
    hipLaunchKernelGGL((rtq::kShadowMasks<K>), grid, block, 0, st, s->view, r, h, d_leaf_nid, n, sd[0], sd[1], sd[2],
                       flags, d_masks);
  });
  RTQ_HIP(hipGetLastError());
  return RTQ_OK;
}

rtq_status rtq_selftest_division(rtq_device* dev, uint64_t seed, uint64_t count, uint64_t* mismatches) {
  if (!dev || !mismatches) return fail(RTQ_ERR_INVALID_ARGUMENT, "NULL argument");
  RTQ_HIP(hipSetDevice(dev->ordinal));
  const uint32_t blocks = 4096, threads = 256;
  const uint64_t per = (count + uint64_t(blocks) * threads - 1) / (uint64_t(blocks) * threads);
  if (per > 0xFFFFFFFFull) return fail(RTQ_ERR_LIMIT, "count too large");
  unsigned long long* d = nullptr;
  RTQ_HIP(hipMalloc(&d, sizeof(unsigned long long)));
  hipError_t e = hipMemset(d, 0, sizeof(unsigned long long));
  if (e == hipSuccess) {
    hipLaunchKernelGGL(rtq::kDivSelfTest, dim3(blocks), dim3(threads), 0, nullptr, seed, uint32_t(per), d);
    e = hipGetLastError();
  }
  unsigned long long h = 0;
  if (e == hipSuccess) e = hipMemcpy(&h, d, sizeof(h), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess) return hipFail(e, "division self-test");
  *mismatches = h;
  return RTQ_OK;
}

namespace {

// Checks a culling call and fills the launch geometry; RTQ_OK or the error to return.
rtq_status cullSetup(rtq_scene* s, const rtq_frustum* f, uint32_t width, uint32_t height, const uint32_t* d_tile_offsets,
                     const uint32_t* d_tile_lights, uint32_t capacity, rtq::Frustum4& fr, uint32_t& ngx, uint32_t& ngy,
                     uint32_t& groups) {
  if (!s || !f || !d_tile_offsets || (!d_tile_lights && capacity)) return fail(RTQ_ERR_INVALID_ARGUMENT, "NULL argument");
  if (width == 0 || height == 0) return fail(RTQ_ERR_INVALID_ARGUMENT, "empty frame");
  ngx = (width + 15u) / 16u;
  ngy = (height + 15u) / 16u;
  if (uint64_t(ngx) * ngy > (1u << 24)) return fail(RTQ_ERR_LIMIT, "more than 2^24 tiles");
  groups = (s->view.numLights + 63u) / 64u;
  std::memcpy(&fr, f, sizeof(fr));
  return RTQ_OK;
}

// kCullMask -> kScan -> kCullEmit on `st` with their scratch at `work` (cullScratchWords words).
hipError_t cullLaunch(rtq_scene* s, const rtq::Frustum4& fr, uint32_t ngx, uint32_t ngy, uint32_t groups,
                      uint32_t* work, uint32_t* d_tile_offsets, uint32_t* d_tile_lights, uint32_t capacity,
                      hipStream_t st) {
  const uint32_t ntiles = ngx * ngy;
  uint64_t* masks = reinterpret_cast<uint64_t*>(work);
  uint32_t* counts = work + 2ull * ntiles * groups;
  const dim3 grid((ntiles + rtq::kCullBlock / 64 - 1) / (rtq::kCullBlock / 64)), block(rtq::kCullBlock);
  const float4* lights = reinterpret_cast<const float4*>(s->view.lights);
  hipLaunchKernelGGL(rtq::kCullMask, grid, block, 0, st, fr, lights, s->view.numLights, ngx, ngy, groups, masks, counts);
  hipLaunchKernelGGL(rtq::kScan, dim3(1), dim3(1024), 0, st, counts, ntiles, d_tile_offsets);
  hipLaunchKernelGGL(rtq::kCullEmit, grid, block, 0, st, masks, groups, ntiles, d_tile_offsets, d_tile_lights, capacity);
  return hipGetLastError();
}

}  // namespace

rtq_status rtq_cull_lights(rtq_scene* s, const rtq_frustum* f, uint32_t width, uint32_t height,
                           uint32_t* d_tile_offsets, uint32_t* d_tile_lights, uint32_t capacity, uint32_t* h_total,
                           void* stream) {
  rtq::Frustum4 fr;
  uint32_t ngx = 0, ngy = 0, groups = 0;
  rtq_status cs = cullSetup(s, f, width, height, d_tile_offsets, d_tile_lights, capacity, fr, ngx, ngy, groups);
  if (cs != RTQ_OK) return cs;
  const uint32_t ntiles = ngx * ngy;
  RTQ_HIP(hipSetDevice(s->dev->ordinal));
  hipStream_t st = static_cast<hipStream_t>(stream);
  // stream-ordered scratch: the call stays asynchronous
  uint32_t* work = nullptr;
  RTQ_HIP(hipMallocAsync(reinterpret_cast<void**>(&work),
                         sizeof(uint32_t) * std::max<uint64_t>(1, rtq::cullScratchWords(ntiles, groups)), st));
  hipError_t le = cullLaunch(s, fr, ngx, ngy, groups, work, d_tile_offsets, d_tile_lights, capacity, st);
  hipError_t fe = hipFreeAsync(work, st);
  if (le != hipSuccess) {
    (void)hipStreamSynchronize(st);
    return hipFail(le, "culling kernels");
  }
  if (fe != hipSuccess) return hipFail(fe, "hipFreeAsync");
  if (!h_total) return RTQ_OK;
  uint32_t total = 0;
  RTQ_HIP(hipMemcpyAsync(&total, d_tile_offsets + ntiles, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  RTQ_HIP(hipStreamSynchronize(st));
  *h_total = total;
  if (total > capacity) return fail(RTQ_ERR_LIMIT, "tile light lists exceed capacity");
  return RTQ_OK;
}

// rtq_cull_lights with the device's fixed scratch array instead of allocated scratch: no allocation,
// no host synchronisation. Calls that overlap on one GPU share the array and race.
rtq_status rtq_cull_lights_noalloc(rtq_scene* s, const rtq_frustum* f, uint32_t width, uint32_t height,
                                   uint32_t* d_tile_offsets, uint32_t* d_tile_lights, uint32_t capacity,
                                   void* stream) {
  rtq::Frustum4 fr;
  uint32_t ngx = 0, ngy = 0, groups = 0;
  rtq_status cs = cullSetup(s, f, width, height, d_tile_offsets, d_tile_lights, capacity, fr, ngx, ngy, groups);
  if (cs != RTQ_OK) return cs;
  if (rtq::cullScratchWords(ngx * ngy, groups) > kNoAllocScratchWords)
    return fail(RTQ_ERR_LIMIT, "tiles x (2 x light groups of 64 + 1) exceeds the 4 Mi-word no-allocation scratch");
  if (!s->dev->scratch) return hipFail(s->dev->scratchError, "hipGetSymbolAddress(scratch)");
  RTQ_HIP(hipSetDevice(s->dev->ordinal));
  hipError_t le = cullLaunch(s, fr, ngx, ngy, groups, s->dev->scratch, d_tile_offsets, d_tile_lights, capacity,
                             static_cast<hipStream_t>(stream));
  if (le != hipSuccess) return hipFail(le, "culling kernels");
  return RTQ_OK;
}

}  // extern "C"

#include "batch.inc"
#include "serve.inc"
