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
#include "stack_class.h"

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

// ---- slab division without the per-quotient division sequence -----------------------------
// CollideBox divides, so bit parity needs the IEEE fp32 quotient RN(x / d) of every slab, x = bound -
// origin (rounded in fp32 as the shader does). An fp32 division compiles to about ten instructions
// with a quarter-rate reciprocal; here it is one fp64 multiply by the ray's fp64 reciprocal and two
// conversions: q = RN32(RN64(double(x) * RN64(1 / double(d)))).
// Why that is exact: RN64(1/d) and the rounded product each carry a relative error below 2^-53, so
// the fp64 value is within 2^-52 (relative) of x / d. A quotient of two 24-bit significands is never a
// midpoint between adjacent fp32 values, and its distance to one is at least 1 / (D (2K+1)) >= 2^-49
// relative (D the divisor's significand, 2K+1 the midpoint's, both integers below 2^24 and 2^25), so the
// fp64 value and x / d round to the same fp32 number. The fp64 reciprocal is computed once per ray.
// Domain, which keeps every quotient a normal fp32 number: |d| in [2^-30, 2^30] and every non-zero
// x in [2^-60, 2^60]; the latter holds when every box bound and origin coordinate is 0 or has
// magnitude in [2^-37, 2^59] (differences of such values are 0 or >= 2^-60). x = +-0 gives the IEEE
// signed zero. Lanes outside the domain take the IEEE division.
__device__ __forceinline__ bool fastDirComponent(float d) {
  const float a = fabsf(d);
  return a >= 0x1p-30f && a <= 0x1p30f;
}
__device__ __forceinline__ bool fastCoord(float v) {
  const float a = fabsf(v);
  return a == 0.0f || (a >= 0x1p-37f && a <= 0x1p59f);
}
__device__ __forceinline__ double recip64(float d) { return 1.0 / double(d); }
__device__ __forceinline__ float quotient(float x, double r) { return float(double(x) * r); }

struct RayRcp {
  double rx, ry, rz;
};

// CollideBox with quotient(); no NaN can arise (finite x, non-zero d, no overflow), so min/max of
// the slab values use the hardware min/max, which agree with the GLSL selects on non-NaN inputs up
// to the sign of a zero result, which the final max(0, .) and the < comparison ignore.
__device__ __forceinline__ float boxHitFast(const Ray3& r, const RayRcp& q, float mnx, float mny, float mnz,
                                            float mxx, float mxy, float mxz, float tmax) {
  const float t0x = quotient(mnx - r.ox, q.rx), t0y = quotient(mny - r.oy, q.ry), t0z = quotient(mnz - r.oz, q.rz);
  const float t1x = quotient(mxx - r.ox, q.rx), t1y = quotient(mxy - r.oy, q.ry), t1z = quotient(mxz - r.oz, q.rz);
  const float tmin = fmaxf(0.0f, fmaxf(fminf(t0x, t1x), fmaxf(fminf(t0y, t1y), fminf(t0z, t1z))));
  const float tm = fminf(tmax, fminf(fmaxf(t0x, t1x), fminf(fmaxf(t0y, t1y), fmaxf(t0z, t1z))));
  return (tmin < tm) ? tmin : -1.0f;
}

// boxHitFast for a ray whose direction octant is OCT (bit k set: component k negative) and a box with
// min <= max on every axis. quotient() is monotone in x (x -> RN32(RN64(x * r)) with one fixed r), so
// on an axis where d > 0 the min-plane value is the smaller of the two and on one where d < 0 the
// larger: the per-axis min and max of boxHitFast are known without comparing, and the entry and exit
// distances are the same values (up to the sign of a zero, which the comparisons ignore).
template <int OCT>
__device__ __forceinline__ float boxHitOrd(const Ray3& r, const RayRcp& q, float mnx, float mny, float mnz,
                                           float mxx, float mxy, float mxz, float tmax) {
  constexpr bool nx = (OCT & 1) != 0, ny = (OCT & 2) != 0, nz = (OCT & 4) != 0;
  const float enx = quotient((nx ? mxx : mnx) - r.ox, q.rx), fax = quotient((nx ? mnx : mxx) - r.ox, q.rx);
  const float eny = quotient((ny ? mxy : mny) - r.oy, q.ry), fay = quotient((ny ? mny : mxy) - r.oy, q.ry);
  const float enz = quotient((nz ? mxz : mnz) - r.oz, q.rz), faz = quotient((nz ? mnz : mxz) - r.oz, q.rz);
  const float tmin = fmaxf(0.0f, fmaxf(enx, fmaxf(eny, enz)));
  const float tm = fminf(tmax, fminf(fax, fminf(fay, faz)));
  return (tmin < tm) ? tmin : -1.0f;
}

// CollideTriangle (collision.glsl:58-95) with edges precomputed. Every quantity is computed with the
// shader's operations in its order, and the early returns are folded into one final condition, so the
// wave runs one straight-line block (its loads issue together) instead of three divergent branches; an
// accepted result is the same bits either way. The conditions, as the shader tests them:
//   |a| < EPSILON rejects (NaN a is not rejected here, but then u is NaN and rejected below);
//   clamp(u, 0, 1) != u rejects, with GLSL min/max this is exactly !(u >= 0 && u <= 1) (NaN fails both,
//   -0 passes both); v < 0 or u + v > 1 rejects; t > EPSILON && t < tmax accepts.
// u and v are written in every case; callers read them only when the result is > 0.
__device__ __forceinline__ float triHit(const Ray3& r, const float4 t0, const float4 t1, const float4 t2, float tmax,
                                        float& uo, float& vo) {
  const float p0x = t0.x, p0y = t0.y, p0z = t0.z;
  const float e1x = t0.w, e1y = t1.x, e1z = t1.y;
  const float e2x = t1.z, e2y = t1.w, e2z = t2.x;
  const float hx = r.dy * e2z - r.dz * e2y;
  const float hy = r.dz * e2x - r.dx * e2z;
  const float hz = r.dx * e2y - r.dy * e2x;
  const float a = (e1x * hx + e1y * hy) + e1z * hz;
  const float f = 1.0f / a;
  const float sx = r.ox - p0x, sy = r.oy - p0y, sz = r.oz - p0z;
  const float u = f * ((sx * hx + sy * hy) + sz * hz);
  const float qx = sy * e1z - sz * e1y;
  const float qy = sz * e1x - sx * e1z;
  const float qz = sx * e1y - sy * e1x;
  const float v = f * ((r.dx * qx + r.dy * qy) + r.dz * qz);
  const float t = f * ((e2x * qx + e2y * qy) + e2z * qz);
  const bool accept = !(fabsf(a) < kEps) && u >= 0.0f && u <= 1.0f && !(v < 0.0f) && !(u + v > 1.0f) && t > kEps &&
                      t < tmax;
  uo = u;
  vo = v;
  return accept ? t : -1.0f;
}

struct Counters {
  uint32_t inner = 0, leaves = 0, tests = 0;
};

struct Closest {
  float t;
  uint32_t tri, nid;
  float u, v;
};

// ---- leaf tests ------------------------------------------------------------------------------------
// The triangles of one leaf (bvh_collide / bvh_collide_fast, bvhCollision.glsl:177-284), in slot order,
// each against the current t. The leaf ref carries the first slot and min(n, 16) - 1 for a leaf of n
// triangles, so the loads of up to GROUP slots known to be in the leaf are issued together and the
// tests then run in order: each slot's loads no longer wait for the previous slot's test (vector loads
// return in issue order, so a one-slot loop waits once per slot). The last-slot flag ends the leaf
// as before; past 16 slots the count is unknown and the walk goes on one group after another.
// Returns true when `any` and a triangle is accepted; closest hits update `hit`.
// Group sizes, measured (soup-3M bounces, spheres-1M primary rays): the lane-refilling kernels run at
// 4 waves per SIMD with registers to spare and gain most from 4 (bounce 1: 518 -> 563 Mrays/s; 6 and 8
// gave the same, 16 less); the kernels tracing one ray per lane to the end keep 7 waves per SIMD only
// below 73 registers, and 2 is the most they take without losing speed.
constexpr int kGridLeafGroup = 2;    // kernels tracing one ray per lane to the end
constexpr int kRefillLeafGroup = 4;  // lane-refilling kernels

template <int GROUP, bool COUNT>
__device__ __forceinline__ bool leafTests(const DevTri* __restrict__ tris, uint32_t ref, const Ray3& r, bool any,
                                          float tmax, Closest& hit, uint32_t& tests) {
  uint32_t slot = ref & kLeafSlotMask;
  uint32_t rest = ((ref >> kLeafCountShift) & 15u) + 1u;  // slots known to remain, this one included
  for (;;) {
    const uint32_t g = rest < uint32_t(GROUP) ? rest : uint32_t(GROUP);
    float4 t[GROUP][3];
#pragma unroll
    for (int j = 0; j < GROUP; ++j) {
      if (j == 0 || uint32_t(j) < g) {
        const float4* tp = reinterpret_cast<const float4*>(tris + slot + j);
        t[j][0] = tp[0];
        t[j][1] = tp[1];
        t[j][2] = tp[2];
      }
    }
#pragma unroll
    for (int j = 0; j < GROUP; ++j) {
      if (j > 0 && uint32_t(j) >= g) break;
      if (COUNT) ++tests;
      float u = 0.0f, v = 0.0f;
      const float tt = triHit(r, t[j][0], t[j][1], t[j][2], any ? tmax : hit.t, u, v);
      if (tt > 0.0f) {
        if (any) return true;
        hit.t = tt - kOffset;
        hit.tri = __float_as_uint(t[j][2].y);
        hit.nid = __float_as_uint(t[j][2].z);
        hit.u = u;
        hit.v = v;
      }
      if (__float_as_uint(t[j][2].w) != 0u) return false;
    }
    slot += g;
    rest -= g;
    if (rest == 0u) rest = 1u;  // a leaf of more than 16 slots: one more at least
  }
}

// ---- descent rounds ----------------------------------------------------------------------------
// A wave alternates between descent (each lane steps down inner nodes toward its next leaf) and a
// leaf phase (lanes at a leaf test its triangles). If every descent ran until all lanes had reached a
// leaf, the lanes that got there first would idle while the deepest one descends, and on incoherent
// rays most instructions would run with a few lanes on. So a round of descent ends as soon as the
// lanes ready for the leaf phase (at a leaf or finished, plus `extra` lanes that are ready anyway, such
// as free lanes about to take a new ray) reach `wait` per 64 of the wave's live and extra lanes: a wave
// whose other lanes have ended (hit lanes only in a shadow pass, the last rays of a refilling launch)
// keeps the same fraction instead of running every lane to its leaf. Lanes still at an inner node keep
// their node and stack and go on in the next round. Each ray's own sequence of operations, and so every
// result, is the same for any `wait`; only the interleaving of the lanes changes. wait = 64 gives the
// plain order (every lane reaches its leaf before any leaf is tested). The threshold scaled to the live
// lanes instead of a fixed count, measured: soup bounces 754 -> 781 Mrays/s, two-level any hit 1048 ->
// 1138, probe field 0.63 -> 0.57 ms, 1080p primary and shadow passes unchanged.
// default when each lane traces one closest-hit ray (1080p sphere frame, tiled primary rays: 3951 /
// 3997 / 4006 / 3958 / 3815 / 3665 Mrays/s at 24 / 32 / 40 / 48 / 56 / 64, the means of three runs)
constexpr uint32_t kGridWait = 40;
// default when each lane traces one any-hit ray (shadow rays, probe segments, any-hit grid dispatches),
// which ends at its first hit (1080p sphere frame: shadow pass 0.345 / 0.336 / 0.361 ms and probe field
// 0.573 / 0.575 / 0.630 ms at 24 / 16 / 8)
constexpr uint32_t kGridAnyWait = 16;
// default when lanes refill with new rays (soup bounces 748 / 780 / 790 / 773 Mrays/s at 24 / 32 / 40 /
// 48 with the threshold scaled to the live lanes)
constexpr uint32_t kRefillWait = 40;
// default of the lane-refilling batch kernel, whose served launches mix closest- and any-hit rays of
// many scenes (served capacity on device requests 2086 / 2134 / 2098 Mrays/s at 32 / 40 / 48; a
// single-scene any-hit refill walk loses at 40: two-level 917 -> 866)
constexpr uint32_t kBatchRefillWait = 40;
// default of a lane-refilling any-hit walk of a two-level scene, whose rays end at very different
// times (1024 BLAS instances, 4 Mi rays: 1083 / 1141 / 1141 Mrays/s at 16 / 24 / 32)
constexpr uint32_t kTwoLevelRefillWait = 24;

// Top-of-tree staging: a kernel that traces one scene copies the first s.topSlots node records (the
// top levels, laid out breadth first) into LDS once per block and reads them there. Every thread of
// the block must call this before any of them returns.
template <uint32_t TOP = kTopSlots>
__device__ __forceinline__ DevScene stageTop(const DevScene& s) {
  DevScene v = s;
  if constexpr (TOP > 0) {
    __shared__ DevNode top[TOP];
    const uint32_t words = s.topSlots * 4u;
    for (uint32_t k = threadIdx.x; k < words; k += blockDim.x)
      reinterpret_cast<float4*>(top)[k] = reinterpret_cast<const float4*>(s.nodes)[k];
    __syncthreads();
    v.top = top;
  } else {
    v.topSlots = 0;
    v.top = nullptr;
  }
  return v;
}

// One descent step at inner node `node` whose record is (a, b, c, d): both child boxes are tested
// against lim, then the walk goes to the nearer of two hit children (deferring the other on the
// stack), to the only hit child, or to the stack top. FASTONLY: the reciprocal slab form for every
// calling lane (decided per wave); otherwise `fast` chooses per lane. OCT >= 0 (with FASTONLY): every
// lane's direction lies in octant OCT and the scene's boxes are ordered, so boxHitOrd replaces
// boxHitFast.
template <int STACK, bool FASTONLY, int OCT>
__device__ __forceinline__ void nodeStep(const float4 a, const float4 b, const float4 c, const uint4 d, uint32_t* stk,
                                         const Ray3& r, const RayRcp& rq, bool fast, float lim, uint32_t& node,
                                         int& sp) {
  float d0, d1;
  if constexpr (OCT >= 0) {
    d0 = boxHitOrd<OCT>(r, rq, a.x, a.y, a.z, a.w, b.x, b.y, lim);
    d1 = boxHitOrd<OCT>(r, rq, b.z, b.w, c.x, c.y, c.z, c.w, lim);
  } else if (FASTONLY || fast) {
    d0 = boxHitFast(r, rq, a.x, a.y, a.z, a.w, b.x, b.y, lim);
    d1 = boxHitFast(r, rq, b.z, b.w, c.x, c.y, c.z, c.w, lim);
  } else {
    d0 = boxHit(r, a.x, a.y, a.z, a.w, b.x, b.y, lim);
    d1 = boxHit(r, b.z, b.w, c.x, c.y, c.z, c.w, lim);
  }
  const bool h0 = d0 >= 0.0f, h1 = d1 >= 0.0f;
  if (h0 && h1) {
    const bool n1 = d1 < d0;
    if (sp < STACK) {
      stk[sp * kTraceBlock] = n1 ? d.x : d.y;
      ++sp;
      node = n1 ? d.y : d.x;
    } else {
      // a deeper tree than the stack holds: the lane stops here instead of writing past its
      // stack column (unreachable, every launch sizes STACK to the tree depth, stack_class.h)
      node = kRefDone;
      sp = 0;
    }
  } else if (h0 || h1) {
    node = h0 ? d.x : d.y;
  } else {
    node = sp > 0 ? stk[(--sp) * kTraceBlock] : kRefDone;
  }
}

// Ready lanes that end a round: `wait` per 64 of the wave's `lanes` (live plus extra), at least one.
__device__ __forceinline__ uint32_t roundNeed(uint32_t wait, uint32_t lanes) {
  return max(1u, (wait * lanes + 63u) / 64u);
}

// One round of descent for the calling lanes (see nodeStep for FASTONLY and OCT).
template <int STACK, bool FASTONLY, bool COUNT, int OCT = -1>
__device__ __forceinline__ void descendRound(const DevNode* __restrict__ nodes, const DevNode* top, uint32_t topSlots,
                                             uint32_t* stk, const Ray3& r, const RayRcp& rq, bool fast, float lim,
                                             uint32_t& node, int& sp, uint32_t wait, uint32_t extra,
                                             uint32_t& inner) {
  const uint32_t live = uint32_t(__popcll(__ballot(1)));
  const uint32_t need = roundNeed(wait, live + extra);
  for (;;) {
    const bool down = node < kRefDone;
    const uint64_t m = __ballot(down);
    if (m == 0 || live - uint32_t(__popcll(m)) + extra >= need) break;
    if (down) {
      const DevNode* src = nodes;
      if constexpr (kTopSlots > 0) {
        if (node < topSlots) src = top;  // the block's LDS copy of the top levels
      }
      const float4* np = reinterpret_cast<const float4*>(src + node);
      const float4 a = np[0], b = np[1], c = np[2];
      const uint4 d = reinterpret_cast<const uint4*>(np)[3];
      if (COUNT) ++inner;
      nodeStep<STACK, FASTONLY, OCT>(a, b, c, d, stk, r, rq, fast, lim, node, sp);
    }
  }
}

// Closest-hit walks of the lane-refilling kernels hold the record of the lane's current inner node in
// registers across steps, rounds and leaf phases: its loads are issued as soon as the node is known
// (after a step, when a ray starts, or when a leaf is reached: the stack top is then the node that
// follows the leaf, since a closest-hit walk always goes on), so the record of the node popped after a
// leaf arrives while the leaf is tested (soup bounces: +2 %).
struct NodeRec {
  float4 a, b, c;
  uint4 d;
};

__device__ __forceinline__ void loadRec(const DevNode* __restrict__ nodes, uint32_t node, NodeRec& q) {
  const float4* np = reinterpret_cast<const float4*>(nodes + node);
  q.a = np[0];
  q.b = np[1];
  q.c = np[2];
  q.d = reinterpret_cast<const uint4*>(np)[3];
}

template <int STACK, bool FASTONLY>
__device__ __forceinline__ void descendRoundHeld(const DevNode* __restrict__ nodes, NodeRec& q, uint32_t* stk,
                                                 const Ray3& r, const RayRcp& rq, bool fast, float lim,
                                                 uint32_t& node, int& sp, uint32_t wait, uint32_t extra) {
  const uint32_t live = uint32_t(__popcll(__ballot(1)));
  const uint32_t need = roundNeed(wait, live + extra);
  for (;;) {
    const bool down = node < kRefDone;
    const uint64_t m = __ballot(down);
    if (m == 0 || live - uint32_t(__popcll(m)) + extra >= need) break;
    if (down) {
      nodeStep<STACK, FASTONLY, -1>(q.a, q.b, q.c, q.d, stk, r, rq, fast, lim, node, sp);
      if (node < kRefDone) loadRec(nodes, node, q);
    }
  }
}

// ---- two-level any hit (traverseTlasFast, bvhTraversal_inline.glsl:80-151 with TLAS_COLLIDE) -------
// The top-level tree and every BLAS are inner-node records of one array, walked by the same descent.
// At a top-level leaf, tlas_collide_fast (bvhCollision.glsl:321-349) takes each listed instance whose
// BlasHeader box CollideBox reports strictly positive (> 0: a ray starting inside the box gets 0 and
// skips that BLAS) and walks the BLAS from its root; the leaf's own triangles are never tested. Here an
// instance ref (kRefLeaf | kRefInst | k, the leaf's instances from k on) is a step of the descent like
// an inner node: the step tests instance k's header box; if it accepts, the walk leaves a resume ref for
// the instances after k on its stack (unless k is the leaf's last) and goes on with that BLAS's root,
// whose walk pops back to the resume ref when it is done; if not, it goes on with instance k + 1, or
// the stack top after the last one. Lanes at instances thus step with the lanes at nodes instead of
// scanning a leaf's instance list in a loop of their own. An any-hit answer (is some triangle reachable
// this way accepted?) does not depend on the visit order. The slab division takes the reciprocal form
// when the lane's ray and the scene qualify (the header boxes are part of that check), with the same
// results as the IEEE division.
__device__ __forceinline__ bool isInstRef(uint32_t node) {
  return node > kRefDone && node != kRefEmpty && (node & kRefInst) != 0;
}

template <int STACK>
__device__ __forceinline__ void instanceStep(const DevInst* __restrict__ insts, uint32_t* stk, const Ray3& r,
                                             const RayRcp& rq, bool fast, float tmax, uint32_t& node, int& sp) {
  const uint32_t k = node & kRefIndex;
  const float4* ip = reinterpret_cast<const float4*>(insts + k);
  const float4 lo = ip[0], hi = ip[1];
  const float t = fast ? boxHitFast(r, rq, lo.x, lo.y, lo.z, hi.x, hi.y, hi.z, tmax)
                       : boxHit(r, lo.x, lo.y, lo.z, hi.x, hi.y, hi.z, tmax);
  const bool last = __float_as_uint(hi.w) != 0u;
  if (t > 0.0f) {
    if (!last) {
      if (sp >= STACK) {  // unreachable: STACK covers both levels and the resume entry
        node = kRefDone;
        sp = 0;
        return;
      }
      stk[sp * kTraceBlock] = kRefLeaf | kRefInst | (k + 1u);
      ++sp;
    }
    node = __float_as_uint(lo.w);
  } else if (!last) {
    node = kRefLeaf | kRefInst | (k + 1u);
  } else {
    node = sp > 0 ? stk[(--sp) * kTraceBlock] : kRefDone;
  }
}

// A round of descent over a two-level scene: lanes at inner nodes take a node step, lanes at instance
// refs an instance step; a round ends as descendRound's does, with lanes at triangle leaves ready.
template <int STACK, bool FASTONLY, int OCT = -1>
__device__ __forceinline__ void descendRoundTwoLevel(const DevScene& s, uint32_t* stk, const Ray3& r,
                                                     const RayRcp& rq, bool fast, float tmax, uint32_t& node, int& sp,
                                                     uint32_t wait, uint32_t extra) {
  const uint32_t live = uint32_t(__popcll(__ballot(1)));
  const uint32_t need = roundNeed(wait, live + extra);
  for (;;) {
    const bool inst = isInstRef(node);
    const bool down = node < kRefDone || inst;
    const uint64_t m = __ballot(down);
    if (m == 0 || live - uint32_t(__popcll(m)) + extra >= need) break;
    if (inst) {
      instanceStep<STACK>(s.insts, stk, r, rq, FASTONLY || fast, tmax, node, sp);
    } else if (down) {
      const float4* np = reinterpret_cast<const float4*>(s.nodes + node);
      const float4 a = np[0], b = np[1], c = np[2];
      const uint4 d = reinterpret_cast<const uint4*>(np)[3];
      nodeStep<STACK, FASTONLY, OCT>(a, b, c, d, stk, r, rq, fast, tmax, node, sp);
    }
  }
}

// Walk the tree for one ray. ANY: stop at the first accepted triangle (returns true).
// `stk` points at this lane's column of the block's LDS stack (stride kTraceBlock).
template <bool ANY, bool COUNT, int STACK, bool FAST, int OCT = -1>
__device__ __forceinline__ bool traverse(const DevScene& s, uint32_t* stk, const Ray3& r, float tmax, Closest& hit,
                                         Counters& cnt, uint32_t wait) {
  RayRcp rq{};
  if (FAST) rq = RayRcp{recip64(r.dx), recip64(r.dy), recip64(r.dz)};
  uint32_t node = s.root;
  int sp = 0;
  if (!ANY) hit.t = tmax;
  for (;;) {
    if (ANY && !COUNT && s.twoLevel)
      descendRoundTwoLevel<STACK, FAST, OCT>(s, stk, r, rq, FAST, tmax, node, sp, wait, 0u);
    else
      descendRound<STACK, FAST, COUNT, OCT>(s.nodes, s.top, s.topSlots, stk, r, rq, FAST, ANY ? tmax : hit.t, node,
                                            sp, wait, 0u, cnt.inner);
    if (node == kRefDone) break;
    if (node < kRefDone || (ANY && isInstRef(node))) continue;  // still descending: next round
    if (COUNT) cnt.leaves++;
    if (node != kRefEmpty && leafTests<kGridLeafGroup, COUNT>(s.tris, node, r, ANY, tmax, hit, cnt.tests)) return true;
    node = sp > 0 ? stk[(--sp) * kTraceBlock] : kRefDone;
    if (node == kRefDone) break;
  }
  return false;
}

// Whether this ray may use boxHitFast (see quotient()); the scene flag covers the box bounds.
__device__ __forceinline__ bool fastRay(uint32_t fastBoxes, const Ray3& r, float tmax) {
  return fastBoxes && fastDirComponent(r.dx) && fastDirComponent(r.dy) && fastDirComponent(r.dz) &&
         fastCoord(r.ox) && fastCoord(r.oy) && fastCoord(r.oz) && tmax == tmax;
}
__device__ __forceinline__ bool fastLane(const DevScene& s, const Ray3& r, float tmax) {
  return fastRay(s.fastBoxes, r, tmax);
}

// One traversal routine per wave: the fast slab division when every active lane qualifies, and its
// octant form when, in addition, the scene's boxes are ordered and all the wave's rays head into one
// octant (camera tiles, sun rays).
template <bool ANY, bool COUNT, int STACK>
__device__ __forceinline__ bool traverseAuto(const DevScene& s, uint32_t* stk, const Ray3& r, float tmax,
                                             Closest& hit, Counters& cnt,
                                             uint32_t wait = ANY ? kGridAnyWait : kGridWait) {
  if (!__all(fastLane(s, r, tmax) ? 1 : 0)) return traverse<ANY, COUNT, STACK, false>(s, stk, r, tmax, hit, cnt, wait);
  const uint32_t oct = (r.dx < 0.0f ? 1u : 0u) | (r.dy < 0.0f ? 2u : 0u) | (r.dz < 0.0f ? 4u : 0u);
  const uint32_t oct0 = __builtin_amdgcn_readfirstlane(oct);
  if (s.orderedBoxes && __all(oct == oct0 ? 1 : 0)) {
    switch (oct0) {
      case 0: return traverse<ANY, COUNT, STACK, true, 0>(s, stk, r, tmax, hit, cnt, wait);
      case 1: return traverse<ANY, COUNT, STACK, true, 1>(s, stk, r, tmax, hit, cnt, wait);
      case 2: return traverse<ANY, COUNT, STACK, true, 2>(s, stk, r, tmax, hit, cnt, wait);
      case 3: return traverse<ANY, COUNT, STACK, true, 3>(s, stk, r, tmax, hit, cnt, wait);
      case 4: return traverse<ANY, COUNT, STACK, true, 4>(s, stk, r, tmax, hit, cnt, wait);
      case 5: return traverse<ANY, COUNT, STACK, true, 5>(s, stk, r, tmax, hit, cnt, wait);
      case 6: return traverse<ANY, COUNT, STACK, true, 6>(s, stk, r, tmax, hit, cnt, wait);
      default: return traverse<ANY, COUNT, STACK, true, 7>(s, stk, r, tmax, hit, cnt, wait);
    }
  }
  return traverse<ANY, COUNT, STACK, true>(s, stk, r, tmax, hit, cnt, wait);
}

__device__ __forceinline__ Ray3 loadRay(const float4* rays, uint32_t i, float& tmax) {
  const float4 a = rays[2 * size_t(i)], b = rays[2 * size_t(i) + 1];
  tmax = a.w;
  return Ray3{a.x, a.y, a.z, b.x, b.y, b.z};
}

// perm (optional): lane order -> ray index, so that neighbouring lanes trace rays a sort has put
// together; results still land at the ray's own index.
template <int STACK, bool COUNT>
__global__ __launch_bounds__(kTraceBlock) void kTraceClosest(DevScene s0, const float4* __restrict__ rays, uint32_t n,
                                                             float4* __restrict__ hits, uint32_t* __restrict__ leafNid,
                                                             uint32_t* __restrict__ counters,
                                                             const uint32_t* __restrict__ perm, uint32_t wait) {
  __shared__ uint32_t lds[STACK * kTraceBlock];
  const DevScene s = stageTop(s0);
  const uint32_t lanePos = xcdBlock() * kTraceBlock + threadIdx.x;
  if (lanePos >= n) return;
  const uint32_t i = perm ? perm[lanePos] : lanePos;
  float tmax;
  const Ray3 r = loadRay(rays, i, tmax);
  Closest h{tmax, 0xFFFFFFFFu, 0xFFFFFFFFu, 0.0f, 0.0f};
  Counters c;
  traverseAuto<false, COUNT, STACK>(s, lds + threadIdx.x, r, tmax, h, c, wait);
  hits[i] = make_float4(__uint_as_float(h.tri), h.t, h.u, h.v);
  if (leafNid) leafNid[i] = h.nid;
  if (COUNT) {
    counters[3 * size_t(i)] = c.inner;
    counters[3 * size_t(i) + 1] = c.leaves;
    counters[3 * size_t(i) + 2] = c.tests;
  }
}

template <int STACK, bool COUNT>
__global__ __launch_bounds__(kTraceBlock) void kTraceAny(DevScene s0, const float4* __restrict__ rays, uint32_t n,
                                                         uint32_t* __restrict__ occluded, uint32_t* __restrict__ counters,
                                                         const uint32_t* __restrict__ perm, uint32_t wait) {
  __shared__ uint32_t lds[STACK * kTraceBlock];
  const DevScene s = stageTop(s0);
  const uint32_t lanePos = xcdBlock() * kTraceBlock + threadIdx.x;
  if (lanePos >= n) return;
  const uint32_t i = perm ? perm[lanePos] : lanePos;
  float tmax;
  const Ray3 r = loadRay(rays, i, tmax);
  Closest h{};
  Counters c;
  const bool occ = traverseAuto<true, COUNT, STACK>(s, lds + threadIdx.x, r, tmax, h, c, wait);
  occluded[i] = occ ? 1u : 0u;
  if (COUNT) {
    counters[3 * size_t(i)] = c.inner;
    counters[3 * size_t(i) + 1] = c.leaves;
    counters[3 * size_t(i) + 2] = c.tests;
  }
}

// ---- persistent traversal with lane refill ------------------------------------------------
// A fixed grid of waves; each lane traces one ray at a time. After every round of descent and leaf
// tests the lanes whose ray has finished take new rays, so lanes no longer idle while the slowest ray
// of their wave finishes. Each ray still follows the reference visit order exactly; only the
// assignment of rays to lanes changes. Two sources of rays:
//   RANGE = false: chunks of kRefillChunk consecutive rays taken from the per-XCD chunk queues
//     (takeChunk) when the wave's current chunk runs out (one atomic per chunk, so waves balance
//     dynamically);
//   RANGE = true: one fixed contiguous range per wave (no atomics).
// Either way a wave's lanes hold nearby rays of the buffer. Chunk size, measured (soup-3M bounce 1 /
// bounce 2 and two-level any hit, Mrays/s): 16: 608 / 849, 806; 32: 622 / 915, 979; 64: 625 / 947,
// 1048; 128: 607 / 945, 1040; 256: 572 / 907, 966; 512: 495 / 810, 875. An XCD's waves take
// consecutive chunks, so its rays in flight span about (waves x chunk) neighbouring rays of the buffer:
// on sorted bounce rays 64-ray chunks cut the bytes read beyond the L2 per ray from 8.68 to 6.85 KB
// against 256. On unsorted rays (the two-level workload: bytes unchanged) they balance the waves' last
// chunks at a finer grain. Below 64 a chunk no longer fills a wave's idle lanes in one round.
constexpr uint32_t kRefillChunk = 64;
// chunk of the lane-refilling batch kernel (kBatchRefill), whose launches hold many items of many
// scenes: served capacity on device-buffer requests 2075 / 2122 / 2133-2139 Mrays/s at 64 / 128 / 256
constexpr uint32_t kBatchRefillChunk = 256;

// Chunk queues, one per XCD. Blocks are handed to the 8 XCDs round-robin (block b runs on XCD b % 8), and
// each XCD has its own L2. A wave first takes chunks from its own XCD's share of the rays, a contiguous
// eighth, and once that share is used up moves on to the next XCD's share, and so on; so while there is
// work an XCD traces the rays of its own share: in a batch, the items (scenes) of that share; after a ray
// sort, one region of space. `counters` holds kXcds words, zero at launch; `q` and `tried` are the wave's
// (uniform) position in this walk; [next, end) receives up to `grab` consecutive units (rays, or whole
// chunks of a batch). Returns false once every share is used up.
constexpr uint32_t kXcds = 8;

__device__ __forceinline__ bool takeChunk(uint32_t* __restrict__ counters, uint32_t n, uint32_t grab, uint32_t& q,
                                          uint32_t& tried, uint32_t& next, uint32_t& end) {
  const uint32_t lane = threadIdx.x & 63u;
  while (tried < kXcds) {
    const uint32_t lo = uint32_t(uint64_t(n) * q / kXcds), hi = uint32_t(uint64_t(n) * (q + 1u) / kXcds);
    uint32_t got = 0;
    if (lane == 0) got = atomicAdd(counters + q, grab);
    got = __builtin_amdgcn_readfirstlane(__shfl(got, 0, 64));
    if (uint64_t(lo) + got < hi) {
      next = lo + got;
      end = uint32_t(min(uint64_t(hi), uint64_t(next) + grab));
      return true;
    }
    q = (q + 1u) % kXcds;
    ++tried;
  }
  return false;
}

template <int STACK, bool ANY, bool RANGE>
__device__ __forceinline__ void refillTrace(const DevScene& s, uint32_t* stk, const float4* __restrict__ rays,
                                            uint32_t n, void* __restrict__ out, uint32_t* __restrict__ leafNid,
                                            uint32_t* __restrict__ counter, const uint32_t* __restrict__ perm,
                                            uint32_t wait) {
  const uint32_t lane = threadIdx.x & 63u;
  bool active = false, fast = false, exhausted = false;
  uint32_t rid = 0, node = kRefDone;
  uint32_t next = 0, end = 0;  // the wave's current source of rays [next, end)
  uint32_t q = blockIdx.x % kXcds, tried = 0;
  if (RANGE) {
    const uint32_t waves = gridDim.x * (kTraceBlock / 64u);
    const uint32_t wave = xcdBlock() * (kTraceBlock / 64u) + (threadIdx.x >> 6);
    next = uint32_t(uint64_t(n) * wave / waves);
    end = uint32_t(uint64_t(n) * (wave + 1u) / waves);
    exhausted = next >= end;
  }
  int sp = 0;
  float tmax = 0.0f;
  Ray3 r{};
  RayRcp rq{};
  Closest hit{};
  NodeRec rec{};
  for (;;) {
    const uint64_t idle = __ballot(!active);
    if (idle && !exhausted) {
      if (!RANGE && next >= end && !takeChunk(counter, n, kRefillChunk, q, tried, next, end)) exhausted = true;
      if (next < end) {
        const uint32_t want = uint32_t(__popcll(idle));
        const uint32_t base = next;
        next = (end - next > want) ? next + want : end;
        if (RANGE && next >= end) exhausted = true;
        if (!active) {
          const uint32_t k = base + uint32_t(__popcll(idle & ((1ull << lane) - 1ull)));
          if (k < next) {
            rid = perm ? perm[k] : k;
            r = loadRay(rays, rid, tmax);
            fast = fastLane(s, r, tmax);
            if (fast) rq = RayRcp{recip64(r.dx), recip64(r.dy), recip64(r.dz)};
            hit = Closest{tmax, 0xFFFFFFFFu, 0xFFFFFFFFu, 0.0f, 0.0f};
            node = s.root;
            sp = 0;
            if (!ANY && node < kRefDone) loadRec(s.nodes, node, rec);
            active = true;
          }
        }
      }
    }
    if (!__any(active)) break;
    // free lanes count as ready while rays remain: they take new rays after this round
    const uint32_t extra = exhausted ? 0u : uint32_t(__popcll(__ballot(!active)));
    if (!active) continue;
    bool accepted = false;
    const float lim = ANY ? tmax : hit.t;
    const bool allFast = __all(fast ? 1 : 0);
    if constexpr (ANY) {
      // any hit: a leaf may end the walk, so nothing is loaded ahead (measured: holding records is
      // 1 % slower here)
      uint32_t unused = 0;
      if (s.twoLevel) {
        if (allFast)
          descendRoundTwoLevel<STACK, true>(s, stk, r, rq, fast, lim, node, sp, wait, extra);
        else
          descendRoundTwoLevel<STACK, false>(s, stk, r, rq, fast, lim, node, sp, wait, extra);
      } else if (allFast) {
        descendRound<STACK, true, false>(s.nodes, s.top, s.topSlots, stk, r, rq, fast, lim, node, sp, wait, extra, unused);
      } else {
        descendRound<STACK, false, false>(s.nodes, s.top, s.topSlots, stk, r, rq, fast, lim, node, sp, wait, extra, unused);
      }
    } else {
      if (allFast)
        descendRoundHeld<STACK, true>(s.nodes, rec, stk, r, rq, fast, lim, node, sp, wait, extra);
      else
        descendRoundHeld<STACK, false>(s.nodes, rec, stk, r, rq, fast, lim, node, sp, wait, extra);
    }
    if (ANY && isInstRef(node)) {
      // still at a top-level leaf's instances when the round ended: the next round goes on
    } else if (node > kRefDone) {
      // closest hit: the node after this leaf is the stack top (the walk always goes on), and its record
      // is loaded while the leaf is tested
      const uint32_t top = sp > 0 ? stk[(sp - 1) * kTraceBlock] : kRefDone;
      if (!ANY && top < kRefDone) loadRec(s.nodes, top, rec);
      uint32_t unusedTests = 0;
      if (node != kRefEmpty) accepted = leafTests<kRefillLeafGroup, false>(s.tris, node, r, ANY, tmax, hit, unusedTests);
      if (!accepted && sp > 0) {
        node = top;
        --sp;
      } else {
        node = kRefDone;
      }
    }
    if (node == kRefDone) {
      if (ANY) {
        static_cast<uint32_t*>(out)[rid] = accepted ? 1u : 0u;
      } else {
        static_cast<float4*>(out)[rid] = make_float4(__uint_as_float(hit.tri), hit.t, hit.u, hit.v);
        if (leafNid) leafNid[rid] = hit.nid;
      }
      active = false;
    }
  }
}

template <int STACK, bool ANY>
__global__ __launch_bounds__(kTraceBlock) void kTracePersistent(DevScene s0, const float4* __restrict__ rays,
                                                                uint32_t n, void* __restrict__ out,
                                                                uint32_t* __restrict__ leafNid,
                                                                uint32_t* __restrict__ counter,
                                                                const uint32_t* __restrict__ perm, uint32_t wait) {
  __shared__ uint32_t lds[STACK * kTraceBlock];
  const DevScene s = stageTop(s0);
  refillTrace<STACK, ANY, false>(s, lds + threadIdx.x, rays, n, out, leafNid, counter, perm, wait);
}

template <int STACK, bool ANY>
__global__ __launch_bounds__(kTraceBlock) void kTraceWaveRange(DevScene s0, const float4* __restrict__ rays,
                                                               uint32_t n, void* __restrict__ out,
                                                               uint32_t* __restrict__ leafNid,
                                                               const uint32_t* __restrict__ perm, uint32_t wait) {
  __shared__ uint32_t lds[STACK * kTraceBlock];
  const DevScene s = stageTop(s0);
  refillTrace<STACK, ANY, true>(s, lds + threadIdx.x, rays, n, out, leafNid, nullptr, perm, wait);
}

// ---- ray sort keys -------------------------------------------------------------------------------
// Key of ray i: a 27-bit Morton code of the origin quantised to 512 cells per axis over the scene box
// (out-of-box and NaN coordinates clamp to the border cells) in bits 3-29, the direction octant in bits
// 0-2. Sorting by it puts rays that start close together on neighbouring lanes, those of one cell
// grouped by direction. (The octant above the Morton code traced the soup's bounces 7 % slower.)
__device__ __forceinline__ uint32_t spread3(uint32_t v) {  // 9 bits -> every third of 27 bits
  v &= 0x1FFu;
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}
__global__ __launch_bounds__(256) void kRayKeys(const float4* __restrict__ rays, uint32_t n, float mnx, float mny,
                                                float mnz, float sx, float sy, float sz, uint32_t* __restrict__ keys,
                                                uint32_t* __restrict__ vals) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n) return;
  const float4 o = rays[2 * size_t(i)], d = rays[2 * size_t(i) + 1];
  const float q[3] = {(o.x - mnx) * sx, (o.y - mny) * sy, (o.z - mnz) * sz};
  uint32_t c[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) c[k] = q[k] >= 0.0f ? (q[k] < 511.0f ? uint32_t(q[k]) : 511u) : 0u;
  const uint32_t oct = (d.x < 0.0f ? 1u : 0u) | (d.y < 0.0f ? 2u : 0u) | (d.z < 0.0f ? 4u : 0u);
  keys[i] = (((spread3(c[0]) << 2) | (spread3(c[1]) << 1) | spread3(c[2])) << 3) | oct;
  vals[i] = i;
}

// ---- camera rays generated in the kernel --------------------------------------------------
struct Frustum4 {
  float cam[3], c00[3], c10[3], c01[3];
};

// cameraPass.comp:50-54: stepW = (c10 - c00) * invW, stepH = (c01 - c00) * invH,
// target = c00 + stepW * gx + stepH * gy, dir = normalize(target - cam), with
// normalize(v) = v / sqrt(dot(v, v)) and invW = 1.f / W as fillPassResources computes it
// (raytracingPass.cpp:47).
__device__ __forceinline__ Ray3 cameraRay(const Frustum4& f, uint32_t W, uint32_t H, uint32_t gx, uint32_t gy) {
  const float invW = 1.0f / float(W), invH = 1.0f / float(H);
  const float fx = float(gx), fy = float(gy);
  float d[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float sw = (f.c10[k] - f.c00[k]) * invW;
    const float sh = (f.c01[k] - f.c00[k]) * invH;
    const float target = (f.c00[k] + sw * fx) + sh * fy;
    d[k] = target - f.cam[k];
  }
  const float len = sqrtf((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
  return Ray3{f.cam[0], f.cam[1], f.cam[2], d[0] / len, d[1] / len, d[2] / len};
}

// One block covers a 16x8 pixel region, each wave an 8x8 tile: neighbouring lanes trace
// neighbouring pixels. referenceRows = 1 reproduces cameraPass.comp:46's storage row
// 16*ceil(H/16) - gy (the generating row differs from the stored one and row 0 is never written).
template <int STACK>
__global__ __launch_bounds__(kTraceBlock) void kTraceCamera(DevScene s0, Frustum4 f, uint32_t W, uint32_t H,
                                                            uint32_t referenceRows, float4* __restrict__ hits,
                                                            uint32_t* __restrict__ leafNid) {
  __shared__ uint32_t lds[STACK * kTraceBlock];
  const DevScene s = stageTop(s0);
  const uint32_t tilesX = (W + 15u) / 16u;
  const uint32_t b = xcdBlock();
  const uint32_t bx = b % tilesX, by = b / tilesX;
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint32_t x = bx * 16u + wave * 8u + (lane & 7u), y = by * 8u + (lane >> 3);
  if (x >= W || y >= H) return;
  uint32_t gy = y;
  if (referenceRows) {
    if (y == 0) return;
    gy = 16u * ((H + 15u) / 16u) - y;
  }
  const Ray3 r = cameraRay(f, W, H, x, gy);
  Closest h{kTMaxF, 0xFFFFFFFFu, 0xFFFFFFFFu, 0.0f, 0.0f};
  Counters c;
  traverseAuto<false, false, STACK>(s, lds + threadIdx.x, r, kTMaxF, h, c);
  const size_t i = size_t(y) * W + x;
  hits[i] = make_float4(__uint_as_float(h.tri), h.t, h.u, h.v);
  if (leafNid) leafNid[i] = h.nid;
}

// computeShadowMask (bvhLighting.glsl:143-163) at hit position p of the hit leaf `nid`.
template <int STACK>
__device__ __forceinline__ uint32_t shadowMask(const DevScene& s, uint32_t* stk, float px, float py, float pz,
                                               uint32_t nid, float sunX, float sunY, float sunZ, uint32_t flags) {
  uint32_t mask = 0;
  Closest dummy{};
  Counters c;
  if (flags & 1u) {
    const Ray3 sr{px, py, pz, -sunX, -sunY, -sunZ};
    if (traverseAuto<true, false, STACK>(s, stk, sr, kTMaxF, dummy, c)) mask |= 1u;
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
        occ = traverseAuto<true, false, STACK>(s, stk, lr, len, dummy, c);
      }
      if (occ) mask |= 1u << (li + 1u);
    }
  }
  return mask;
}

// computeProbMask (bvhLighting.glsl:165-184) with getLightProbCoordUVW / getLightProbPosition
// (lightprob/lightprobHelpers.glsl:63-77): 8 unnormalised any-hit segments of t limit 1 toward the
// corners of the probe cell around p.
template <int STACK>
__device__ __forceinline__ uint32_t probMask(const DevScene& s, uint32_t* stk, float px, float py, float pz,
                                             const LpfHeader& lpf) {
  const float p[3] = {px, py, pz};
  float step[3], coordf[3];
  uint32_t coord[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    step[k] = (lpf.mx[k] - lpf.mn[k]) / float(lpf.res[k]);
    float uvw = ((p[k] - lpf.mn[k]) / step[k]) - 0.5f;
    uvw = gmax(uvw, 0.0f);
    uvw = gmin(uvw, float(lpf.res[k]) - 1.0f);
    coordf[k] = uvw;
    // int() of NaN is undefined in GLSL; fixed to 0 (the position is finite for every hit t < TMAX)
    const uint32_t c = uint32_t(coordf[k] == coordf[k] ? int(coordf[k]) : 0);
    const uint32_t lim = lpf.res[k] - 2u;
    coord[k] = c < lim ? c : lim;
  }
  uint32_t mask = 0;
  Closest dummy{};
  Counters cnt;
  for (uint32_t i = 0; i < 8; ++i) {
    const uint32_t o[3] = {i & 1u, (i >> 1) & 1u, (i >> 2) & 1u};
    float d[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const float probe = lpf.mn[k] + step[k] * (float(coord[k] + o[k]) + 0.5f);
      d[k] = probe - p[k];
    }
    const Ray3 pr{px, py, pz, d[0], d[1], d[2]};
    if (traverseAuto<true, false, STACK>(s, stk, pr, 1.0f, dummy, cnt)) mask |= 1u << i;
  }
  return mask;
}

// One TRACING_STEP record: the closest hit's ClosestHit_getHitData (bvhBindings.glsl:85-92) =
// [vertexOffset, index01, index2_matId, floatBits(t)] and computeAdditionalHitData
// (baseRaytracingPass.glsl:85-97) = [shadowMask, probMask, 0, 0] when t < TMAX. A miss keeps t = TMAX
// and zero triangle words (the shader leaves whatever its shared memory held) and masks 0.
template <int STACK>
__device__ __forceinline__ void traceRecord(const DevScene& s, uint32_t* stk, const Ray3& r, float sunX, float sunY,
                                            float sunZ, uint32_t shadowFlags, const LpfHeader& lpf,
                                            uint32_t useLpf, uint4& r0, uint4& r1) {
  Closest h{kTMaxF, 0xFFFFFFFFu, 0xFFFFFFFFu, 0.0f, 0.0f};
  Counters c;
  traverseAuto<false, false, STACK>(s, stk, r, kTMaxF, h, c);
  r0 = make_uint4(0u, 0u, 0u, __float_as_uint(h.t));
  r1 = make_uint4(0u, 0u, 0u, 0u);
  if (h.tri != 0xFFFFFFFFu) {
    r0.x = s.triangles[3 * size_t(h.tri)];
    r0.y = s.triangles[3 * size_t(h.tri) + 1];
    r0.z = s.triangles[3 * size_t(h.tri) + 2];
  }
  if (h.t < kTMaxF && h.nid != 0xFFFFFFFFu) {
    // hitPos = from + dir * t (baseRaytracingPass.glsl:91)
    const float px = r.ox + r.dx * h.t, py = r.oy + r.dy * h.t, pz = r.oz + r.dz * h.t;
    r1.x = shadowMask<STACK>(s, stk, px, py, pz, h.nid, sunX, sunY, sunZ, shadowFlags);
    if (useLpf) r1.y = probMask<STACK>(s, stk, px, py, pz, lpf);
  }
}

// The camera pass's TRACING_STEP output (cameraPass.comp:69-75): per pixel two uvec4,
// [vertexOffset, index01, index2_matId, floatBits(t)] and [shadowMask, probMask, 0, 0]. A miss keeps
// t = TMAX and zero triangle words (the shader leaves whatever its shared memory held) and masks 0.
template <int STACK>
__global__ __launch_bounds__(kTraceBlock) void kTraceRecord(DevScene s0, Frustum4 f, uint32_t W, uint32_t H,
                                                            uint32_t referenceRows, float sunX, float sunY,
                                                            float sunZ, uint32_t shadowFlags, LpfHeader lpf,
                                                            uint32_t useLpf, uint4* __restrict__ records) {
  __shared__ uint32_t lds[STACK * kTraceBlock];
  const DevScene s = stageTop(s0);
  const uint32_t tilesX = (W + 15u) / 16u;
  const uint32_t b = xcdBlock();
  const uint32_t bx = b % tilesX, by = b / tilesX;
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint32_t x = bx * 16u + wave * 8u + (lane & 7u), y = by * 8u + (lane >> 3);
  if (x >= W || y >= H) return;
  uint32_t gy = y;
  if (referenceRows) {
    if (y == 0) return;
    gy = 16u * ((H + 15u) / 16u) - y;
  }
  uint4 r0, r1;
  traceRecord<STACK>(s, lds + threadIdx.x, cameraRay(f, W, H, x, gy), sunX, sunY, sunZ, shadowFlags, lpf, useLpf,
                     r0, r1);
  const size_t i = size_t(y) * W + x;
  records[2 * i] = r0;
  records[2 * i + 1] = r1;
}

// ---- light-probe field tracing step (genLightProbField.comp:44-69 with TRACING_STEP) ----------
constexpr uint32_t kProbeRays = 64;  // NUM_RAYS_PER_PROB (lightprob/lightprob.glsl)
struct ProbeDirs {
  float4 d[kProbeRays];
};

// getLightProbCoord (lightprob/lightprobHelpers.glsl:35-52, LP_TILE_RES 4): probes are numbered in
// 4x4x4 tiles, tiles in x, y, z order. getLightProbPosition (:62-68): min + step * (coord + 0.5).
__device__ __forceinline__ void probePosition(const LpfHeader& lpf, uint32_t index, float& px, float& py,
                                              float& pz) {
  const uint32_t tx = lpf.res[0] / 4u, ty = lpf.res[1] / 4u;
  uint32_t tile = index / 64u;
  const uint32_t inTile = index % 64u;
  const uint32_t cz = tile / (tx * ty);
  tile = tile % (tx * ty);
  const uint32_t cy = tile / tx, cx = tile % tx;
  const uint32_t c[3] = {cx * 4u + inTile % 4u, cy * 4u + (inTile / 4u) % 4u, cz * 4u + (inTile / 16u) % 4u};
  float p[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float step = (lpf.mx[k] - lpf.mn[k]) / float(lpf.res[k]);
    p[k] = lpf.mn[k] + step * (float(c[k]) + 0.5f);
  }
  px = p[0];
  py = p[1];
  pz = p[2];
}

// One wave traces 64 consecutive probes along one direction, as a workgroup of the reference does;
// the record of ray `rayId` of probe `p` lands at sample p * 64 + rayId. The probe mask is not
// computed in this pass (the reference dispatches it without USE_LPF_MASK).
template <int STACK>
__global__ __launch_bounds__(kTraceBlock) void kTraceProbes(DevScene s0, LpfHeader lpf, ProbeDirs dirs,
                                                            uint32_t numProbes, float sunX, float sunY, float sunZ,
                                                            uint32_t shadowFlags, uint4* __restrict__ records) {
  __shared__ uint32_t lds[STACK * kTraceBlock];
  const DevScene s = stageTop(s0);
  const uint32_t wave = xcdBlock() * (kTraceBlock / 64u) + (threadIdx.x >> 6);
  const uint32_t rayId = wave % kProbeRays;
  const uint32_t probe = (wave / kProbeRays) * 64u + (threadIdx.x & 63u);
  if (probe >= numProbes) return;
  float px, py, pz;
  probePosition(lpf, probe, px, py, pz);
  const float4 d = dirs.d[rayId];
  uint4 r0, r1;
  traceRecord<STACK>(s, lds + threadIdx.x, Ray3{px, py, pz, d.x, d.y, d.z}, sunX, sunY, sunZ, shadowFlags, lpf, 0u,
                     r0, r1);
  const size_t i = size_t(probe) * kProbeRays + rayId;
  records[2 * i] = r0;
  records[2 * i + 1] = r1;
}

// computeShadowMask (bvhLighting.glsl:143-163) for one closest hit per lane.
template <int STACK>
__global__ __launch_bounds__(kTraceBlock) void kShadowMasks(DevScene s0, const float4* __restrict__ rays,
                                                            const float4* __restrict__ hits,
                                                            const uint32_t* __restrict__ leafNid, uint32_t n,
                                                            float sunX, float sunY, float sunZ, uint32_t flags,
                                                            uint32_t* __restrict__ masks) {
  __shared__ uint32_t lds[STACK * kTraceBlock];
  const DevScene s = stageTop(s0);
  const uint32_t i = xcdBlock() * kTraceBlock + threadIdx.x;
  if (i >= n) return;
  float tmax;
  const Ray3 r = loadRay(rays, i, tmax);
  const float t = hits[i].y;
  const uint32_t nid = leafNid[i];
  uint32_t mask = 0;
  if (t < kTMaxF && nid != 0xFFFFFFFFu) {
    // hitPos = from + dir * t (baseRaytracingPass.glsl:91)
    const float px = r.ox + r.dx * t, py = r.oy + r.dy * t, pz = r.oz + r.dz * t;
    mask = shadowMask<STACK>(s, lds + threadIdx.x, px, py, pz, nid, sunX, sunY, sunZ, flags);
  }
  masks[i] = mask;
}

// ---- batched form: many scenes and ray sets in one persistent launch --------------------------
// Work is split into chunks of `chunkRays` consecutive rays of one item; items are sorted by
// chunkBegin. Each wave takes the next chunk from the per-XCD chunk queues (so an XCD works through
// the items of its own share first) until all chunks are gone, so every wave reaches the exit test and
// the grid drains.
struct BatchItemDev {
  uint32_t scene, flags, numRays, chunkBegin;
  const float4* rays;
  void* out;
  uint32_t* nid;
  uint32_t rayBegin;  // rays of all earlier items
  // the scene fields a lane needs to start a ray (setItemScene), so that the lane-refilling kernel
  // starts one from the item record alone instead of a further dependent load of the scene record
  uint32_t root;
  const DevNode* nodes;
  const DevTri* tris;
  uint32_t fastBoxes;
  uint32_t pad[3];
};
static_assert(sizeof(BatchItemDev) == 80, "BatchItemDev is 80 B");

inline void setItemScene(BatchItemDev& d, const DevScene& v) {
  d.root = v.root;
  d.nodes = v.nodes;
  d.tris = v.tris;
  d.fastBoxes = v.fastBoxes;
}

// Zero the kXcds chunk-queue words of a refilling launch (the chunked batch kernel uses the first).
__global__ void kResetCounter(uint32_t* counter) {
  if (threadIdx.x < kXcds && blockIdx.x == 0) counter[threadIdx.x] = 0u;
}

template <int STACK>
__global__ __launch_bounds__(kTraceBlock) void kBatchPersistent(const DevScene* __restrict__ scenes,
                                                                const BatchItemDev* __restrict__ items,
                                                                uint32_t numItems, uint32_t totalChunks,
                                                                uint32_t chunkRays, uint32_t* __restrict__ counter,
                                                                uint32_t wait) {
  __shared__ uint32_t lds[STACK * kTraceBlock];
  const uint32_t lane = threadIdx.x & 63u;
  uint32_t* stk = lds + threadIdx.x;
  uint32_t q = blockIdx.x % kXcds, tried = 0;
  for (;;) {
    uint32_t chunk = 0, chunkEnd = 0;  // one chunk at a time from the per-XCD queues
    if (!takeChunk(counter, totalChunks, 1u, q, tried, chunk, chunkEnd)) break;
    uint32_t lo = 0, hi = numItems;
    while (hi - lo > 1u) {
      const uint32_t mid = (lo + hi) >> 1;
      if (items[mid].chunkBegin <= chunk) lo = mid;
      else hi = mid;
    }
    const BatchItemDev it = items[lo];
    DevScene s = scenes[it.scene];
    s.topSlots = 0;  // the blocks of this kernel trace many scenes: nothing is staged in LDS
    s.top = nullptr;
    const uint32_t base = (chunk - it.chunkBegin) * chunkRays;
    for (uint32_t g = 0; g < chunkRays; g += 64u) {
      const uint32_t i = base + g + lane;
      if (i < it.numRays) {
        float tmax;
        const Ray3 r = loadRay(it.rays, i, tmax);
        Counters c;
        if (it.flags == 1u) {
          Closest h{tmax, 0xFFFFFFFFu, 0xFFFFFFFFu, 0.0f, 0.0f};
          traverseAuto<false, false, STACK>(s, stk, r, tmax, h, c, wait);
          reinterpret_cast<float4*>(it.out)[i] = make_float4(__uint_as_float(h.tri), h.t, h.u, h.v);
          if (it.nid) it.nid[i] = h.nid;
        } else {
          Closest h{};
          const bool occ = traverseAuto<true, false, STACK>(s, stk, r, tmax, h, c, wait);
          reinterpret_cast<uint32_t*>(it.out)[i] = occ ? 1u : 0u;
        }
      }
    }
  }
}

// Lane-refilling form of the batch: lanes take single rays from the concatenation of all items (the
// wave draws chunks of kBatchRefillChunk rays from the per-XCD chunk queues, so an XCD works through the
// items of its own share first; a binary search over rayBegin maps a ray to its item), so a lane that
// finishes early picks up the next ray, whatever its item, scene or mode, instead of idling until its
// wave's slowest ray is done. The scene pointers and the closest/any mode are per lane. Each ray still
// follows the reference visit order exactly; every wave reaches the exit test once every share is used
// up.
template <int STACK>
__global__ __launch_bounds__(kTraceBlock) void kBatchRefill(const BatchItemDev* __restrict__ items,
                                                            uint32_t numItems, uint32_t totalRays,
                                                            uint32_t* __restrict__ counter, uint32_t wait) {
  __shared__ uint32_t lds[STACK * kTraceBlock];
  uint32_t* stk = lds + threadIdx.x;
  const uint32_t lane = threadIdx.x & 63u;
  bool active = false, fast = false, exhausted = false, anyRay = false;
  uint32_t rid = 0, item = 0, node = kRefDone;
  uint32_t next = 0, end = 0;  // the wave's current chunk of rays [next, end)
  uint32_t q = blockIdx.x % kXcds, tried = 0;
  int sp = 0;
  float tmax = 0.0f;
  const DevNode* nodes = nullptr;
  const DevTri* tris = nullptr;
  Ray3 r{};
  RayRcp rq{};
  Closest hit{};
  NodeRec rec{};
  for (;;) {
    const uint64_t idle = __ballot(!active);
    if (idle && !exhausted && next >= end && !takeChunk(counter, totalRays, kBatchRefillChunk, q, tried, next, end))
      exhausted = true;
    if (idle && next < end) {
      const uint32_t want = uint32_t(__popcll(idle));
      const uint32_t base = next;
      next = (end - next > want) ? next + want : end;
      // The item holding `base` (the last one whose rayBegin is <= base), searched once by the wave: the
      // rays it hands out are consecutive, so each lane steps forward from there past the few item
      // boundaries they may cross, instead of a per-lane search of dependent loads before every ray.
      uint32_t first = 0, hi = numItems;
      while (hi - first > 1u) {
        const uint32_t mid = (first + hi) >> 1;
        if (items[mid].rayBegin <= base) first = mid;
        else hi = mid;
      }
      first = __builtin_amdgcn_readfirstlane(first);
      if (!active) {
        const uint32_t k = base + uint32_t(__popcll(idle & ((1ull << lane) - 1ull)));
        if (k < next) {
          uint32_t lo = first;
          while (lo + 1u < numItems && items[lo + 1u].rayBegin <= k) ++lo;
          item = lo;
          const BatchItemDev& it = items[lo];
          rid = k - it.rayBegin;
          nodes = it.nodes;
          tris = it.tris;
          r = loadRay(it.rays, rid, tmax);
          fast = fastRay(it.fastBoxes, r, tmax);
          if (fast) rq = RayRcp{recip64(r.dx), recip64(r.dy), recip64(r.dz)};
          anyRay = it.flags != 1u;
          hit = Closest{tmax, 0xFFFFFFFFu, 0xFFFFFFFFu, 0.0f, 0.0f};  // any hit keeps t = tmax throughout
          node = it.root;
          sp = 0;
          if (node < kRefDone) loadRec(nodes, node, rec);
          active = true;
        }
      }
    }
    if (!__any(active)) break;
    const uint32_t extra = exhausted ? 0u : uint32_t(__popcll(__ballot(!active)));
    if (!active) continue;
    bool accepted = false;
    // lanes of one block trace different scenes here, so nothing is staged in LDS
    if (__all(fast ? 1 : 0))
      descendRoundHeld<STACK, true>(nodes, rec, stk, r, rq, fast, hit.t, node, sp, wait, extra);
    else
      descendRoundHeld<STACK, false>(nodes, rec, stk, r, rq, fast, hit.t, node, sp, wait, extra);
    if (node > kRefDone) {
      // a closest-hit walk goes on with the stack top: its record is loaded while the leaf is tested
      const uint32_t top = sp > 0 ? stk[(sp - 1) * kTraceBlock] : kRefDone;
      if (!anyRay && top < kRefDone) loadRec(nodes, top, rec);
      uint32_t unusedTests = 0;
      // any hit keeps hit.t = tmax, so the closest-hit bound serves both modes
      if (node != kRefEmpty)
        accepted = leafTests<kRefillLeafGroup, false>(tris, node, r, anyRay, hit.t, hit, unusedTests);
      if (!accepted && sp > 0) {
        node = top;
        --sp;
        if (anyRay && node < kRefDone) loadRec(nodes, node, rec);
      } else {
        node = kRefDone;
      }
    }
    if (node == kRefDone) {
      const BatchItemDev& it = items[item];
      if (anyRay) {
        static_cast<uint32_t*>(it.out)[rid] = accepted ? 1u : 0u;
      } else {
        static_cast<float4*>(it.out)[rid] = make_float4(__uint_as_float(hit.tri), hit.t, hit.u, hit.v);
        if (it.nid) it.nid[rid] = hit.nid;
      }
      active = false;
    }
  }
}

// ---- streaming copy ----------------------------------------------------------------------------
__global__ __launch_bounds__(256) void kCopy16(uint4* __restrict__ dst, const uint4* __restrict__ src, size_t n) {
  const size_t stride = size_t(gridDim.x) * 256u;
  for (size_t i = size_t(blockIdx.x) * 256u + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

// ---- self-test of quotient() against IEEE division --------------------------------------------
__device__ __forceinline__ uint32_t mix32(uint64_t& s) {
  s = s * 6364136223846793005ull + 1442695040888963407ull;
  uint64_t z = s;
  z = (z ^ (z >> 33)) * 0xff51afd7ed558ccdull;
  z = (z ^ (z >> 33)) * 0xc4ceb9fe1a85ec53ull;
  return uint32_t(z >> 32);
}

// Random (x, d) in the range where quotient() claims equality with x / d: |d| in [2^-30, 2^30),
// |x| in [2^-60, 2^60), random signs and mantissas, a quarter of the mantissas forced to the
// extremes (all ones / all zeros) where rounding ties are nearest. Counts bitwise differences.
__global__ __launch_bounds__(256) void kDivSelfTest(uint64_t seed, uint32_t perThread,
                                                    unsigned long long* __restrict__ mismatches) {
  const uint64_t gid = uint64_t(blockIdx.x) * 256u + threadIdx.x;
  uint64_t st = seed ^ (gid * 0x9E3779B97F4A7C15ull);
  unsigned long long bad = 0;
  for (uint32_t k = 0; k < perThread; ++k) {
    const uint32_t a = mix32(st), b = mix32(st), c = mix32(st);
    uint32_t md = a & 0x7FFFFFu, mx = b & 0x7FFFFFu;
    if ((c & 3u) == 0) md = (c & 4u) ? 0x7FFFFFu : 0u;
    if ((c & 24u) == 0) mx = (c & 32u) ? 0x7FFFFFu : 0u;
    const uint32_t ed = 127u - 30u + ((c >> 8) % 60u);
    const uint32_t ex = 127u - 60u + ((c >> 16) % 120u);
    const float d = __uint_as_float(((a >> 31) << 31) | (ed << 23) | md);
    const float x = __uint_as_float(((b >> 31) << 31) | (ex << 23) | mx);
    const float ieee = x / d;
    const float fast = quotient(x, recip64(d));
    bad += (__float_as_uint(ieee) != __float_as_uint(fast)) ? 1ull : 0ull;
  }
  atomicAdd(mismatches, bad);
}

// ---- tile light culling ---------------------------------------------------------------------
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

// sphereFrustum4Collision (collision.glsl:219-227): culled if dot(n, c) + w < -r on any plane. All four
// tests are evaluated (no early exit, so the wave runs one straight block); the result is the same.
__device__ __forceinline__ bool sphereInFrustum(const float pl[16], float4 L) {
  bool in = true;
#pragma unroll
  for (int p = 0; p < 4; ++p) in &= !((pl[4 * p] * L.x + pl[4 * p + 1] * L.y) + pl[4 * p + 2] * L.z + pl[4 * p + 3] < -L.w);
  return in;
}

constexpr int kCullBlock = 256;

// Scratch of one culling call: a 64-bit light mask per (tile, group of 64 lights), tile-major, then
// the per-tile light counts. Every word is written by kCullMask before kScan and kCullEmit read it,
// so the scratch needs no clearing between calls.
__host__ __device__ constexpr uint64_t cullScratchWords(uint32_t ntiles, uint32_t groups) {
  return 2ull * ntiles * groups + ntiles;
}

// Pass 1: one wave per tile tests every light once against the tile's four planes. Lane l of the
// wave keeps the ballot of group g where g % 64 == l, and the 64 words of a round of groups are
// stored together (one coalesced store); lane 0 writes the tile's light count.
__global__ __launch_bounds__(kCullBlock) void kCullMask(Frustum4 f, const float4* __restrict__ lights,
                                                        uint32_t numLights, uint32_t ngx, uint32_t ngy,
                                                        uint32_t groups, uint64_t* __restrict__ masks,
                                                        uint32_t* __restrict__ counts) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t tile = blockIdx.x * (kCullBlock / 64u) + (threadIdx.x >> 6);
  if (tile >= ngx * ngy) return;
  float pl[16];
  tilePlanes(f, ngx, ngy, tile % ngx, tile / ngx, pl);
  uint32_t count = 0;
  uint64_t* row = masks + size_t(tile) * groups;
  for (uint32_t g0 = 0; g0 < groups; g0 += 64u) {
    uint64_t mine = 0;
    const uint32_t gEnd = min(groups, g0 + 64u);
    for (uint32_t g = g0; g < gEnd; ++g) {
      const uint32_t i = g * 64u + lane;
      const bool in = i < numLights && sphereInFrustum(pl, lights[i]);
      const uint64_t m = __ballot(in);
      count += uint32_t(__popcll(m));
      if (lane == g - g0) mine = m;
    }
    if (g0 + lane < gEnd) row[g0 + lane] = mine;
  }
  if (lane == 0) counts[tile] = count;
}

// Pass 2: exclusive scan of counts[0..n) into offsets[0..n], one block; offsets[n] is the total.
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

// Pass 3: one wave per tile writes the ascending light indices of its mask at offsets[tile], the
// lists truncated at `capacity` entries in all. Lane l expands mask word l of each round of 64 words,
// at the position a wave prefix sum of the words' population counts gives it.
__global__ __launch_bounds__(kCullBlock) void kCullEmit(const uint64_t* __restrict__ masks, uint32_t groups,
                                                        uint32_t ntiles, const uint32_t* __restrict__ offsets,
                                                        uint32_t* __restrict__ out, uint32_t capacity) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t tile = blockIdx.x * (kCullBlock / 64u) + (threadIdx.x >> 6);
  if (tile >= ntiles) return;
  const uint64_t* row = masks + size_t(tile) * groups;
  uint32_t base = offsets[tile];
  for (uint32_t g0 = 0; g0 < groups; g0 += 64u) {
    const uint32_t g = g0 + lane;
    uint64_t m = g < groups ? row[g] : 0ull;
    const uint32_t c = uint32_t(__popcll(m));
    uint32_t incl = c;  // inclusive prefix sum over the lanes
    for (uint32_t o = 1; o < 64u; o <<= 1) {
      const uint32_t v = __shfl_up(incl, o, 64);
      if (lane >= o) incl += v;
    }
    uint32_t pos = base + incl - c;
    while (m) {
      const uint32_t b = uint32_t(__builtin_ctzll(m));
      m &= m - 1ull;
      if (pos < capacity) out[pos] = g * 64u + b;
      ++pos;
    }
    base += __shfl(incl, 63, 64);
  }
}

}  // namespace rtq
