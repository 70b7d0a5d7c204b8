// Device-resident scene layout shared by the upload code and the kernels.
//
// Inner nodes: one 64 B record per inner node of the reference tree, laid out depth first with the
// two children of a node side by side in one 128 B pair (the root alone in pair 0; the half of a pair
// whose child is a leaf stays empty):
//   a = (b0.min.x, b0.min.y, b0.min.z, b0.max.x)
//   b = (b0.max.y, b0.max.z, b1.min.x, b1.min.y)
//   c = (b1.min.z, b1.max.x, b1.max.y, b1.max.z)
//   d = (ref child0, ref child1, reference node id, 0)
// A child ref is an inner index (< kRefDone), a leaf ref kRefLeaf | (min(n, 16) - 1) << 26 | first
// triangle slot for a leaf of n triangles, or kRefEmpty for a leaf holding no triangle. Leaves need no
// record: visiting one reads its triangle slots, and the count lets a walk issue the loads of several
// slots at once (the last-slot flag still ends the leaf).
//
// Triangle slots: one 48 B record per inline triangle of every leaf, leaves in the same depth-first
// order, triangles in blob (sorted leaf) order:
//   t0 = (p0.x, p0.y, p0.z, e1.x), t1 = (e1.y, e1.z, e2.x, e2.y),
//   t2 = (e2.z, triangle id bits, leaf reference node id bits, 1 if last slot of its leaf else 0)
// with e1 = p1 - p0, e2 = p2 - p0 computed once in fp32 on the host: the same IEEE subtractions
// CollideTriangle performs (collision.glsl:63-64), so the test's results are unchanged.
#pragma once
#include <cstdint>

namespace rtq {

// Inner-node records at the start of the layout that hold the top levels of the tree (breadth first)
// and that single-scene kernels copy into LDS once per block. 0 disables the staging.
#ifndef RTQ_TOP_SLOTS
#define RTQ_TOP_SLOTS 0
#endif
constexpr uint32_t kTopSlots = RTQ_TOP_SLOTS;

constexpr uint32_t kRefLeaf = 0x80000000u;
constexpr uint32_t kRefEmpty = 0xFFFFFFFFu;
constexpr uint32_t kRefDone = 0x7FFFFFFFu;  // traversal finished (never a valid inner index)
// Two-level scenes: a top-level leaf's ref is kRefLeaf | kRefInst | index of the next instance to
// consider (its first one when the walk reaches the leaf); every other leaf ref then has a slot below
// 2^30, so bit 30 tells the two kinds apart.
constexpr uint32_t kRefInst = 0x40000000u;
constexpr uint32_t kRefIndex = 0x3FFFFFFFu;
// Triangle leaf refs: bits 26-29 hold min(triangles in the leaf, 16) - 1, bits 0-25 the first slot
// (a scene holds fewer than 2^26 slots).
constexpr uint32_t kLeafCountShift = 26;
constexpr uint32_t kLeafSlotMask = (1u << kLeafCountShift) - 1u;
constexpr uint32_t kMaxSlots = 1u << kLeafCountShift;

struct alignas(16) DevNode {
  float a[4], b[4], c[4];
  uint32_t d[4];
};
static_assert(sizeof(DevNode) == 64, "DevNode is 64 B");

struct alignas(16) DevTri {
  float t0[4], t1[4], t2[4];
};
static_assert(sizeof(DevTri) == 48, "DevTri is 48 B");

// One BLAS instance of a top-level leaf, 32 B: the BlasHeader box, the device ref of the BLAS's
// root and whether it is the leaf's last instance. A two-level scene's top-level leaf refs are
// kRefLeaf | kRefInst | index of the leaf's first instance (kRefEmpty for a leaf without instances);
// the instances of one leaf are consecutive.
struct alignas(16) DevInst {
  float mn[3];
  uint32_t root;
  float mx[3];
  uint32_t last;
};
static_assert(sizeof(DevInst) == 32, "DevInst is 32 B");

// Sphere light for shadow rays and culling: (x, y, z, attenuation radius); kind in w of `kinds`.
struct DevScene {
  const DevNode* nodes;
  const DevTri* tris;
  const uint32_t* leafOff;   // nid.w per reference node (leaf data word offset or 0xFFFFFFFF)
  const uint32_t* leafData;  // the reference leaf u32 stream (light lists are read from it)
  const float* lights;       // 4 floats per light: pos.xyz, radius
  const uint32_t* lightKind; // PackedLight.iparam per light
  uint32_t root;             // root ref
  uint32_t numNodes;         // reference node count
  uint32_t leafWords;
  uint32_t numLights;
  uint32_t fastBoxes;        // 1 if every inner-node box bound qualifies for the fast slab division
  uint32_t orderedBoxes;     // 1 if every box has min <= max on each axis (no empty or NaN bounds)
  const uint32_t* triangles; // the blob's Triangle records (3 words each), indexed by triangle id
  uint32_t topSlots;         // leading node records a kernel may read from its LDS copy (<= kTopSlots)
  const DevNode* top;        // that LDS copy, set by the kernel; nullptr outside a staging kernel
  const DevInst* insts;      // two-level scenes: the top-level leaves' BLAS instances (else nullptr)
  uint32_t twoLevel;         // 1: `root` is a top-level tree whose leaves hold BLAS instances
};

// Light-probe field placement as cameraPass.comp builds LightProbFieldHeader (cameraPass.comp:62-66).
struct LpfHeader {
  float mn[3], mx[3];
  uint32_t res[3];
};

}  // namespace rtq
