// gfx950 kernels: closest-hit and any-hit traversal, shadow masks, tile light culling.
//
// Arithmetic follows the reference shaders operation by operation (build with -ffp-contract=off,
// IEEE division and square root, no denormal flushing):
//   GLSL min(x,y) = y < x ? y : x, max(x,y) = x < y ? y : x (the specification's definitions,
//   which also fix the NaN cases); dot = (x*x' + y*y') + z*z'; cross as in collision.glsl.
// Traversal visits nodes in the reference order (bvhTraversal_inline.glsl:3-78): both child boxes
// are tested against the current t, the nearer (d1 < d0 ? child1 : child0) is entered first and the
// other is pushed on a per-lane stack and later entered without re-testing its box, which is what
// the reference's bitstack + sibling backtracking does.
#pragma once
#include <hip/hip_runtime.h>

#include "layout.h"

namespace rtq {

constexpr int kTraceBlock = 128;
constexpr float kTMaxF = 100.0f;
constexpr float kOffset = 10e-6f;   // OFFSET_RAY_COLLISION (system.glsl:45)
constexpr float kEps = 10e-8f;      // EPSILON (collision.glsl:62)

__device__ __forceinline__ float gmin(float x, float y) { return (y < x) ? y : x; }
__device__ __forceinline__ float gmax(float x, float y) { return (x < y) ? y : x; }

// Blocks are handed to the 8 XCDs round-robin; renumber them so that each XCD walks a contiguous
// range of rays (neighbouring rays share nodes, and each XCD has its own L2).
__device__ __forceinline__ uint32_t xcdBlock() {
  const uint32_t nb = gridDim.x, b = blockIdx.x;
  const uint32_t xcd = b & 7u, q = nb >> 3, r = nb & 7u;
  return xcd * q + (xcd < r ? xcd : r) + (b >> 3);
}

struct Ray3 {
  float ox, oy, oz, dx, dy, dz;
};

// CollideBox (collision.glsl:23-40), returnTmin = true.
__device__ __forceinline__ float boxHit(const Ray3& r, float mnx, float mny, float mnz, float mxx, float mxy,
                                        float mxz, float tmax) {
  const float t0x = (mnx - r.ox) / r.dx, t0y = (mny - r.oy) / r.dy, t0z = (mnz - r.oz) / r.dz;
  const float t1x = (mxx - r.ox) / r.dx, t1y = (mxy - r.oy) / r.dy, t1z = (mxz - r.oz) / r.dz;
  const float sx = gmin(t0x, t1x), sy = gmin(t0y, t1y), sz = gmin(t0z, t1z);
  const float bx = gmax(t0x, t1x), by = gmax(t0y, t1y), bz = gmax(t0z, t1z);
  const float tmin = gmax(0.0f, gmax(sx, gmax(sy, sz)));
  const float tm = gmin(tmax, gmin(bx, gmin(by, bz)));
  return (tmin < tm) ? tmin : -1.0f;
}

// CollideTriangle (collision.glsl:58-95) with edges precomputed; u, v written when accepted.
__device__ __forceinline__ float triHit(const Ray3& r, const float4 t0, const float4 t1, const float4 t2, float tmax,
                                        float& uo, float& vo) {
  const float p0x = t0.x, p0y = t0.y, p0z = t0.z;
  const float e1x = t0.w, e1y = t1.x, e1z = t1.y;
  const float e2x = t1.z, e2y = t1.w, e2z = t2.x;
  const float hx = r.dy * e2z - r.dz * e2y;
  const float hy = r.dz * e2x - r.dx * e2z;
  const float hz = r.dx * e2y - r.dy * e2x;
  const float a = (e1x * hx + e1y * hy) + e1z * hz;
  if (fabsf(a) < kEps) return -1.0f;
  const float f = 1.0f / a;
  const float sx = r.ox - p0x, sy = r.oy - p0y, sz = r.oz - p0z;
  const float u = f * ((sx * hx + sy * hy) + sz * hz);
  if (gmin(gmax(u, 0.0f), 1.0f) != u) return -1.0f;
  const float qx = sy * e1z - sz * e1y;
  const float qy = sz * e1x - sx * e1z;
  const float qz = sx * e1y - sy * e1x;
  const float v = f * ((r.dx * qx + r.dy * qy) + r.dz * qz);
  if (v < 0.0f || u + v > 1.0f) return -1.0f;
  const float t = f * ((e2x * qx + e2y * qy) + e2z * qz);
  if (t > kEps && t < tmax) {
    uo = u;
    vo = v;
    return t;
  }
  return -1.0f;
}

struct Counters {
  uint32_t inner = 0, leaves = 0, tests = 0;
};

struct Closest {
  float t;
  uint32_t tri, nid;
  float u, v;
};

// Walk the tree for one ray. ANY: stop at the first accepted triangle (returns true).
// `stk` points at this lane's column of the block's LDS stack (stride kTraceBlock).
template <bool ANY, bool COUNT, int STACK>
__device__ __forceinline__ bool traverse(const DevScene& s, uint32_t* stk, const Ray3& r, float tmax, Closest& hit,
                                         Counters& cnt) {
  uint32_t node = s.root;
  int sp = 0;
  if (!ANY) hit.t = tmax;
  for (;;) {
    while (node < kRefDone) {
      const float4* np = reinterpret_cast<const float4*>(s.nodes + node);
      const float4 a = np[0], b = np[1], c = np[2];
      const uint4 d = reinterpret_cast<const uint4*>(np)[3];
      if (COUNT) cnt.inner++;
      const float lim = ANY ? tmax : hit.t;
      const float d0 = boxHit(r, a.x, a.y, a.z, a.w, b.x, b.y, lim);
      const float d1 = boxHit(r, b.z, b.w, c.x, c.y, c.z, c.w, lim);
      const bool h0 = d0 >= 0.0f, h1 = d1 >= 0.0f;
      if (h0 && h1) {
        const bool n1 = d1 < d0;
        if (sp < STACK) stk[sp * kTraceBlock] = n1 ? d.x : d.y;
        ++sp;
        node = n1 ? d.y : d.x;
      } else if (h0 || h1) {
        node = h0 ? d.x : d.y;
      } else {
        node = sp > 0 ? stk[(--sp) * kTraceBlock] : kRefDone;
      }
    }
    if (node == kRefDone) break;
    if (COUNT) cnt.leaves++;
    if (node != kRefEmpty) {
      uint32_t slot = node & ~kRefLeaf;
      for (;;) {
        const float4* tp = reinterpret_cast<const float4*>(s.tris + slot);
        const float4 t0 = tp[0], t1 = tp[1], t2 = tp[2];
        if (COUNT) cnt.tests++;
        float u = 0.0f, v = 0.0f;
        const float tt = triHit(r, t0, t1, t2, ANY ? tmax : hit.t, u, v);
        if (tt > 0.0f) {
          if (ANY) return true;
          hit.t = tt - kOffset;
          hit.tri = __float_as_uint(t2.y);
          hit.nid = __float_as_uint(t2.z);
          hit.u = u;
          hit.v = v;
        }
        if (__float_as_uint(t2.w) != 0u) break;
        ++slot;
      }
    }
    node = sp > 0 ? stk[(--sp) * kTraceBlock] : kRefDone;
  }
  return false;
}

__device__ __forceinline__ Ray3 loadRay(const float4* rays, uint32_t i, float& tmax) {
  const float4 a = rays[2 * i], b = rays[2 * i + 1];
  tmax = a.w;
  return Ray3{a.x, a.y, a.z, b.x, b.y, b.z};
}

template <int STACK, bool COUNT>
__global__ __launch_bounds__(kTraceBlock) void kTraceClosest(DevScene s, const float4* __restrict__ rays, uint32_t n,
                                                             float4* __restrict__ hits, uint32_t* __restrict__ leafNid,
                                                             uint32_t* __restrict__ counters) {
  __shared__ uint32_t lds[STACK * kTraceBlock];
  const uint32_t i = xcdBlock() * kTraceBlock + threadIdx.x;
  if (i >= n) return;
  float tmax;
  const Ray3 r = loadRay(rays, i, tmax);
  Closest h{tmax, 0xFFFFFFFFu, 0xFFFFFFFFu, 0.0f, 0.0f};
  Counters c;
  traverse<false, COUNT, STACK>(s, lds + threadIdx.x, r, tmax, h, c);
  hits[i] = make_float4(__uint_as_float(h.tri), h.t, h.u, h.v);
  if (leafNid) leafNid[i] = h.nid;
  if (COUNT) {
    counters[3 * i] = c.inner;
    counters[3 * i + 1] = c.leaves;
    counters[3 * i + 2] = c.tests;
  }
}

template <int STACK, bool COUNT>
__global__ __launch_bounds__(kTraceBlock) void kTraceAny(DevScene s, const float4* __restrict__ rays, uint32_t n,
                                                         uint32_t* __restrict__ occluded, uint32_t* __restrict__ counters) {
  __shared__ uint32_t lds[STACK * kTraceBlock];
  const uint32_t i = xcdBlock() * kTraceBlock + threadIdx.x;
  if (i >= n) return;
  float tmax;
  const Ray3 r = loadRay(rays, i, tmax);
  Closest h{};
  Counters c;
  const bool occ = traverse<true, COUNT, STACK>(s, lds + threadIdx.x, r, tmax, h, c);
  occluded[i] = occ ? 1u : 0u;
  if (COUNT) {
    counters[3 * i] = c.inner;
    counters[3 * i + 1] = c.leaves;
    counters[3 * i + 2] = c.tests;
  }
}

// computeShadowMask (bvhLighting.glsl:143-163) for one closest hit per lane.
template <int STACK>
__global__ __launch_bounds__(kTraceBlock) void kShadowMasks(DevScene s, const float4* __restrict__ rays,
                                                            const float4* __restrict__ hits,
                                                            const uint32_t* __restrict__ leafNid, uint32_t n,
                                                            float sunX, float sunY, float sunZ, uint32_t flags,
                                                            uint32_t* __restrict__ masks) {
  __shared__ uint32_t lds[STACK * kTraceBlock];
  const uint32_t i = xcdBlock() * kTraceBlock + threadIdx.x;
  if (i >= n) return;
  float tmax;
  const Ray3 r = loadRay(rays, i, tmax);
  const float t = hits[i].y;
  const uint32_t nid = leafNid[i];
  uint32_t mask = 0;
  uint32_t* stk = lds + threadIdx.x;
  Closest dummy{};
  Counters c;
  if (t < kTMaxF && nid != 0xFFFFFFFFu) {
    // hitPos = from + dir * t (baseRaytracingPass.glsl:91)
    const float px = r.ox + r.dx * t, py = r.oy + r.dy * t, pz = r.oz + r.dz * t;
    if (flags & 1u) {
      const Ray3 sr{px, py, pz, -sunX, -sunY, -sunZ};
      if (traverse<true, false, STACK>(s, stk, sr, kTMaxF, dummy, c)) mask |= 1u;
    }
    const uint32_t node = nid & 0x3FFFFFFFu;
    const uint32_t off = node < s.numNodes ? s.leafOff[node] : 0xFFFFFFFFu;
    if (off != 0xFFFFFFFFu && off < s.leafWords) {
      const uint32_t packed = s.leafData[off];
      // unpackLightCountAndOffset (bvhLighting.glsl:133-141), NodeTriangleStride 3
      uint32_t lo = 3u * (packed & 0xFFFu) + ((packed >> 12) & 0xFFu) + ((packed >> 20) & 0xFFu);
      const uint32_t count = (packed >> 28) & 0xFu;
      lo += 1u + off;
      for (uint32_t k = 0; k < count; ++k) {
        if (lo + k >= s.leafWords) break;
        const uint32_t li = s.leafData[lo + k];
        if (li + 1u >= 32u || li >= s.numLights || s.lightKind[li] != 1u) continue;
        const float4 L = reinterpret_cast<const float4*>(s.lights)[li];
        const float dx = L.x - px, dy = L.y - py, dz = L.z - pz;
        const float len = sqrtf((dx * dx + dy * dy) + dz * dz);
        bool occ = true;
        if (len < L.w) {
          const Ray3 lr{px, py, pz, dx / len, dy / len, dz / len};
          occ = traverse<true, false, STACK>(s, stk, lr, len, dummy, c);
        }
        if (occ) mask |= 1u << (li + 1u);
      }
    }
  }
  masks[i] = mask;
}

// ---- tile light culling ---------------------------------------------------------------------
struct Frustum4 {
  float cam[3], c00[3], c10[3], c01[3];
};

// computePlan (cullFrustum.glsl:9-13): n = normalize(cross(p1 - p2, p3 - p2)), w = -dot(n, p2).
__device__ __forceinline__ void computePlane(const float* p1, const float* p2, const float* p3, float* out) {
  const float ax = p1[0] - p2[0], ay = p1[1] - p2[1], az = p1[2] - p2[2];
  const float bx = p3[0] - p2[0], by = p3[1] - p2[1], bz = p3[2] - p2[2];
  const float cx = ay * bz - az * by, cy = az * bx - ax * bz, cz = ax * by - ay * bx;
  const float len = sqrtf((cx * cx + cy * cy) + cz * cz);
  const float nx = cx / len, ny = cy / len, nz = cz / len;
  out[0] = nx;
  out[1] = ny;
  out[2] = nz;
  out[3] = -((nx * p2[0] + ny * p2[1]) + nz * p2[2]);
}

// Tile corners and the 4 side planes of tile (tx, ty) (cullFrustum.glsl:63-86).
__device__ __forceinline__ void tilePlanes(const Frustum4& f, uint32_t ngx, uint32_t ngy, uint32_t tx, uint32_t ty,
                                           float pl[16]) {
  float sx[3], sy[3], bl[3], br[3], tl[3], tr[3];
  const float fx = float(tx), fy = float(ty), fx1 = float(tx + 1u), fy1 = float(ty + 1u);
  for (int k = 0; k < 3; ++k) {
    sx[k] = (f.c10[k] - f.c00[k]) / float(ngx);
    sy[k] = (f.c01[k] - f.c00[k]) / float(ngy);
  }
  for (int k = 0; k < 3; ++k) {
    bl[k] = (f.c00[k] + sx[k] * fx) + sy[k] * fy;
    br[k] = (f.c00[k] + sx[k] * fx1) + sy[k] * fy;
    tl[k] = (f.c00[k] + sx[k] * fx) + sy[k] * fy1;
    tr[k] = (f.c00[k] + sx[k] * fx1) + sy[k] * fy1;
  }
  for (int k = 0; k < 3; ++k) {
    const float dx = sx[k] * 0.25f, dy = sy[k] * 0.25f;
    bl[k] = (bl[k] - dx) - dy;
    br[k] = (br[k] + dx) - dy;
    tl[k] = (tl[k] - dx) + dy;
    tr[k] = (tr[k] + dx) + dy;
  }
  computePlane(f.cam, tl, bl, pl + 0);
  computePlane(f.cam, br, tr, pl + 4);
  computePlane(f.cam, tr, tl, pl + 8);
  computePlane(f.cam, bl, br, pl + 12);
}

// sphereFrustum4Collision (collision.glsl:219-227).
__device__ __forceinline__ bool sphereInFrustum(const float pl[16], float4 L) {
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    if ((pl[4 * p] * L.x + pl[4 * p + 1] * L.y) + pl[4 * p + 2] * L.z + pl[4 * p + 3] < -L.w) return false;
  }
  return true;
}

constexpr int kCullBlock = 256;

// Pass 1: lights per tile.
__global__ __launch_bounds__(kCullBlock) void kCullCount(Frustum4 f, const float4* __restrict__ lights,
                                                         uint32_t numLights, uint32_t ngx, uint32_t ngy,
                                                         uint32_t* __restrict__ counts) {
  const uint32_t tile = blockIdx.x;
  const uint32_t tx = tile % ngx, ty = tile / ngx;
  float pl[16];
  tilePlanes(f, ngx, ngy, tx, ty, pl);
  uint32_t mine = 0;
  for (uint32_t i = threadIdx.x; i < numLights; i += kCullBlock) mine += sphereInFrustum(pl, lights[i]) ? 1u : 0u;
  __shared__ uint32_t red[kCullBlock / 64];
  // wave reduction then block reduction
  for (int o = 32; o > 0; o >>= 1) mine += __shfl_down(mine, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mine;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t s = 0;
    for (int w = 0; w < kCullBlock / 64; ++w) s += red[w];
    counts[tile] = s;
  }
}

// Exclusive scan of counts[0..n) into offsets[0..n], one block.
__global__ __launch_bounds__(1024) void kScan(const uint32_t* __restrict__ counts, uint32_t n,
                                              uint32_t* __restrict__ offsets) {
  __shared__ uint32_t part[1024];
  const uint32_t per = (n + 1023u) / 1024u;
  const uint32_t b = threadIdx.x * per, e = min(n, b + per);
  uint32_t s = 0;
  for (uint32_t i = b; i < e; ++i) s += counts[i];
  part[threadIdx.x] = s;
  __syncthreads();
  for (uint32_t o = 1; o < 1024; o <<= 1) {
    uint32_t v = threadIdx.x >= o ? part[threadIdx.x - o] : 0u;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  uint32_t run = threadIdx.x ? part[threadIdx.x - 1] : 0u;
  for (uint32_t i = b; i < e; ++i) {
    offsets[i] = run;
    run += counts[i];
  }
  if (threadIdx.x == 1023) offsets[n] = part[1023];
}

// Pass 2: ascending light indices per tile at offsets[tile].
__global__ __launch_bounds__(kCullBlock) void kCullFill(Frustum4 f, const float4* __restrict__ lights,
                                                        uint32_t numLights, uint32_t ngx, uint32_t ngy,
                                                        const uint32_t* __restrict__ offsets,
                                                        uint32_t* __restrict__ out, uint32_t capacity) {
  const uint32_t tile = blockIdx.x;
  const uint32_t tx = tile % ngx, ty = tile / ngx;
  float pl[16];
  tilePlanes(f, ngx, ngy, tx, ty, pl);
  __shared__ uint32_t waveCount[kCullBlock / 64];
  uint32_t base = offsets[tile];
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (uint32_t c = 0; c < numLights; c += kCullBlock) {
    const uint32_t i = c + threadIdx.x;
    const bool in = i < numLights && sphereInFrustum(pl, lights[i]);
    const uint64_t m = __ballot(in);
    if (lane == 0) waveCount[wave] = __popcll(m);
    __syncthreads();
    uint32_t before = 0, total = 0;
    for (uint32_t w = 0; w < kCullBlock / 64; ++w) {
      if (w < wave) before += waveCount[w];
      total += waveCount[w];
    }
    if (in) {
      const uint32_t pos = base + before + uint32_t(__popcll(m & ((1ull << lane) - 1ull)));
      if (pos < capacity) out[pos] = i;
    }
    base += total;
    __syncthreads();
  }
}

}  // namespace rtq
