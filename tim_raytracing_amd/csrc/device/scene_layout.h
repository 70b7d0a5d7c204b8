// Host side of upload-BVH: decode the fillGpuBuffer blob (BVHBuilder.cpp:1319-1534) as the shaders
// bind it (BVHData::fillBvhBindings, BVHData.cpp:17-26) and build the device layout of layout.h in
// host memory. Plain C++ (no HIP), so the decode and the layout can be checked on a CPU.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstring>
#include <deque>
#include <unordered_map>
#include <vector>

#include "../../../include/rtq.h"
#include "layout.h"

namespace rtq {

struct HostScene {
  std::vector<DevNode> nodes;        // inner-node records, depth-first sibling pairs
  std::vector<DevTri> tris;          // triangle slots, leaves in the same depth-first order
  std::vector<uint32_t> leafOff;     // nid.w per reference node
  std::vector<uint32_t> leafData;    // the reference leaf word stream
  std::vector<float> lights;         // pos.xyz, radius per light
  std::vector<uint32_t> lightKinds;  // PackedLight.iparam per light
  std::vector<uint32_t> triangles;   // the blob's Triangle records, 3 words each
  uint32_t root = kRefEmpty;         // root ref
  uint32_t topSlots = 0;             // leading node records holding the top levels, breadth first
  uint32_t numNodes = 0, numInner = 0, leafWords = 0, numLights = 0, numTris = 0, maxDepth = 0;
  bool fastBoxes = true;             // every box bound qualifies for the reciprocal slab form
  bool orderedBoxes = true;          // every box has min <= max on each axis
  bool twoLevel = false;             // top-level tree over BLAS instances (insts)
  std::vector<DevInst> insts;        // the top-level leaves' instances, leaf by leaf
};

namespace detail {
struct TriKey {
  uint32_t a, b, c;
  bool operator==(const TriKey& o) const { return a == o.a && b == o.b && c == o.c; }
};
struct TriKeyHash {
  size_t operator()(const TriKey& k) const {
    uint64_t h = 1469598103934665603ull;
    h = (h ^ k.a) * 1099511628211ull;
    h = (h ^ k.b) * 1099511628211ull;
    h = (h ^ k.c) * 1099511628211ull;
    return size_t(h ^ (h >> 32));
  }
};

using TriIndexMap = std::unordered_map<TriKey, uint32_t, TriKeyHash>;

// The triangle slots of blob leaf `id` appended to h.tris (layout.h); *ref receives the leaf ref
// (first slot and count), or stays kRefEmpty for a leaf without triangles.
inline rtq_status appendLeafSlots(const uint32_t* nodes, const uint32_t* leaf, uint32_t leafWords,
                                  const float* positions, uint32_t numVertices, const TriIndexMap& triIndex,
                                  uint32_t id, HostScene& h, uint32_t& ref, const char** err) {
  auto fail = [&](rtq_status s, const char* m) {
    *err = m;
    return s;
  };
  const uint32_t off = nodes[16 * id + 3];
  if (off == 0xFFFFFFFFu) return RTQ_OK;
  if (off >= leafWords) return fail(RTQ_ERR_BAD_BLOB, "leaf data offset beyond the leaf range");
  const uint32_t count = leaf[off] & 0xFFFu;
  if (count == 0) return RTQ_OK;
  if (size_t(off) + 1 + 3 * size_t(count) > leafWords) return fail(RTQ_ERR_BAD_BLOB, "leaf triangles overrun");
  if (h.tris.size() + count > kMaxSlots) return fail(RTQ_ERR_LIMIT, "more than 2^26 triangle slots in a scene");
  ref = kRefLeaf | ((std::min<uint32_t>(count, 16u) - 1u) << kLeafCountShift) | uint32_t(h.tris.size());
  for (uint32_t k = 0; k < count; ++k) {
    const uint32_t* t = leaf + off + 1 + 3 * k;
    const uint32_t vi[3] = {t[0] + (t[1] & 0xFFFFu), t[0] + (t[1] >> 16), t[0] + (t[2] & 0xFFFFu)};
    float p[3][3];
    for (int q = 0; q < 3; ++q) {
      if (vi[q] >= numVertices) return fail(RTQ_ERR_BAD_BLOB, "triangle vertex beyond position count");
      p[q][0] = positions[3 * size_t(vi[q])];
      p[q][1] = positions[3 * size_t(vi[q]) + 1];
      p[q][2] = positions[3 * size_t(vi[q]) + 2];
    }
    auto found = triIndex.find(TriKey{t[0], t[1], t[2]});
    const uint32_t triId = found == triIndex.end() ? RTQ_NO_HIT : found->second;
    // e1 = p1 - p0, e2 = p2 - p0: the IEEE subtractions CollideTriangle performs (collision.glsl:63-64)
    const float e1[3] = {p[1][0] - p[0][0], p[1][1] - p[0][1], p[1][2] - p[0][2]};
    const float e2[3] = {p[2][0] - p[0][0], p[2][1] - p[0][1], p[2][2] - p[0][2]};
    DevTri d{};
    d.t0[0] = p[0][0]; d.t0[1] = p[0][1]; d.t0[2] = p[0][2]; d.t0[3] = e1[0];
    d.t1[0] = e1[1]; d.t1[1] = e1[2]; d.t1[2] = e2[0]; d.t1[3] = e2[1];
    const uint32_t w[4] = {0, triId, id, k + 1 == count ? 1u : 0u};
    std::memcpy(&d.t2, w, sizeof(w));
    d.t2[0] = e2[2];
    h.tris.push_back(d);
  }
  return RTQ_OK;
}

// Two-level blob (a top-level tree whose leaves list BLAS instances, BVHBuilder.cpp:1319-1534 with
// BLASes): every tree gets one record per inner node in depth-first order, the top level's first. A
// top-level leaf's ref points at its instances (header box, root ref of the BLAS, last flag), in the
// order its leaf data lists them; the traversal that reads it (traverseTlasFast,
// bvhTraversal_inline.glsl:80-151 with tlas_collide_fast, bvhCollision.glsl:321-349) tests only those
// instances: a top-level leaf's own triangles and the instances hoisted into inner nodes are not
// traced, and are not laid out. A BLAS whose root has no children is laid out as a leaf (the shader
// would descend into it as an inner node from rootIndex, which has no leaf bit). maxDepth is the
// top level's depth plus the deepest BLAS's: a walk's stack holds both at a top-level leaf.
// The reciprocal slab form needs every box bound to be 0 or of magnitude in [2^-37, 2^59].
inline bool fastBound(float x) {
  const float ax = std::fabs(x);
  return ax == 0.0f || (ax >= 0x1p-37f && ax <= 0x1p59f);
}
// fastBoxes and orderedBoxes over the inner-node records: box 0 = min (a0, a1, a2), max (a3, b0, b1);
// box 1 = min (b2, b3, c0), max (c1, c2, c3).
inline void boxFlags(HostScene& h) {
  h.fastBoxes = true;
  h.orderedBoxes = true;
  for (const DevNode& d : h.nodes) {
    for (int k = 0; k < 4; ++k)
      for (float x : {d.a[k], d.b[k], d.c[k]})
        if (!fastBound(x)) h.fastBoxes = false;
    if (!(d.a[0] <= d.a[3] && d.a[1] <= d.b[0] && d.a[2] <= d.b[1] && d.b[2] <= d.c[1] && d.b[3] <= d.c[2] &&
          d.c[0] <= d.c[3]))
      h.orderedBoxes = false;
  }
}

inline rtq_status buildTwoLevel(const uint8_t* blob, size_t bytes, const rtq_ranges* ranges, const float* positions,
                                uint32_t numVertices, HostScene& out, const char** err) {
  auto fail = [&](rtq_status s, const char* m) {
    *err = m;
    return s;
  };
  if (size_t(ranges->blas.offset) + ranges->blas.size > bytes || ranges->blas.offset % 4 != 0)
    return fail(RTQ_ERR_BAD_BLOB, "the BLAS header range points outside the blob");
  HostScene h;
  h.twoLevel = true;
  const uint32_t numNodes = ranges->nodes.size / 64u;
  const uint32_t leafWords = ranges->leaf_data.size / 4u;
  const uint32_t numTris = ranges->triangles.size / 12u;
  const uint32_t numLights = ranges->lights.size / 56u;
  const uint32_t numHeaders = ranges->blas.size / 32u;
  const uint32_t* nodes = reinterpret_cast<const uint32_t*>(blob + ranges->nodes.offset);
  const uint32_t* leaf = reinterpret_cast<const uint32_t*>(blob + ranges->leaf_data.offset);
  const uint32_t* tris = reinterpret_cast<const uint32_t*>(blob + ranges->triangles.offset);
  const uint32_t* lightsRaw = reinterpret_cast<const uint32_t*>(blob + ranges->lights.offset);
  const uint8_t* headers = blob + ranges->blas.offset;
  TriIndexMap triIndex;
  triIndex.reserve(size_t(numTris) * 2);
  for (uint32_t i = 0; i < numTris; ++i) triIndex.emplace(TriKey{tris[3 * i], tris[3 * i + 1], tris[3 * i + 2]}, i);

  auto isLeafNode = [&](uint32_t id) { return (nodes[16 * id + 2] & 0x3FFFFFFFu) == 0x3FFFFFFFu; };
  std::vector<uint8_t> seen(numNodes, 0);
  // Lays out the tree under blob node rootId; leafRef(id, ref&) gives a leaf child's ref. Records are
  // indexed, never referenced, across leafRef calls, which may append records of other trees.
  auto layoutTree = [&](uint32_t rootId, auto&& leafRef, uint32_t& rootRef, uint32_t& depth) -> rtq_status {
    depth = 0;
    if (rootId >= numNodes) return fail(RTQ_ERR_BAD_BLOB, "root id beyond node count");
    if (isLeafNode(rootId)) {
      if (seen[rootId]) return fail(RTQ_ERR_BAD_BLOB, "node reached twice");
      seen[rootId] = 1;
      rootRef = kRefEmpty;
      return leafRef(rootId, rootRef);
    }
    std::vector<std::pair<uint32_t, uint32_t>> inner;  // (blob id, record)
    std::unordered_map<uint32_t, uint32_t> record;
    std::vector<std::pair<uint32_t, uint32_t>> st{{rootId, 0u}};
    while (!st.empty()) {
      const auto [id, d] = st.back();
      st.pop_back();
      if (id >= numNodes) return fail(RTQ_ERR_BAD_BLOB, "child id beyond node count");
      if (seen[id]) return fail(RTQ_ERR_BAD_BLOB, "node reached twice");
      seen[id] = 1;
      if (d > 64) return fail(RTQ_ERR_LIMIT, "tree deeper than 64 levels");
      if (isLeafNode(id)) {
        depth = std::max(depth, d);
        continue;
      }
      const uint32_t l = nodes[16 * id + 2] & 0x3FFFFFFFu;
      if (l + 1u >= numNodes) return fail(RTQ_ERR_BAD_BLOB, "child id beyond node count");
      record[id] = uint32_t(h.nodes.size());
      inner.push_back({id, uint32_t(h.nodes.size())});
      h.nodes.push_back(DevNode{});
      st.push_back({l + 1u, d + 1u});
      st.push_back({l, d + 1u});
    }
    for (const auto& [id, rec] : inner) {
      const uint32_t l = nodes[16 * id + 2] & 0x3FFFFFFFu;
      uint32_t ref[2];
      for (uint32_t k = 0; k < 2; ++k) {
        const uint32_t c = l + k;
        if (isLeafNode(c)) {
          ref[k] = kRefEmpty;
          rtq_status e = leafRef(c, ref[k]);
          if (e != RTQ_OK) return e;
        } else {
          ref[k] = record.at(c);
        }
      }
      const float* f = reinterpret_cast<const float*>(nodes + 16 * id);
      const float* n0xy = f + 4;
      const float* n1xy = f + 8;
      const float* nz = f + 12;
      DevNode& dn = h.nodes[rec];
      dn.a[0] = n0xy[0]; dn.a[1] = n0xy[1]; dn.a[2] = nz[0]; dn.a[3] = n0xy[2];
      dn.b[0] = n0xy[3]; dn.b[1] = nz[1]; dn.b[2] = n1xy[0]; dn.b[3] = n1xy[1];
      dn.c[0] = nz[2]; dn.c[1] = n1xy[2]; dn.c[2] = n1xy[3]; dn.c[3] = nz[3];
      dn.d[0] = ref[0];
      dn.d[1] = ref[1];
      dn.d[2] = id;
      dn.d[3] = 0;
    }
    rootRef = record.at(rootId);
    return RTQ_OK;
  };

  std::unordered_map<uint32_t, uint32_t> blasRootRef;  // blob root id -> device ref
  uint32_t blasDepth = 0;
  auto blasLeaf = [&](uint32_t id, uint32_t& ref) -> rtq_status {
    return appendLeafSlots(nodes, leaf, leafWords, positions, numVertices, triIndex, id, h, ref, err);
  };
  auto topLeaf = [&](uint32_t id, uint32_t& ref) -> rtq_status {
    const uint32_t off = nodes[16 * id + 3];
    if (off == 0xFFFFFFFFu) return RTQ_OK;
    if (off >= leafWords) return fail(RTQ_ERR_BAD_BLOB, "leaf data offset beyond the leaf range");
    const uint32_t ntri = leaf[off] & 0xFFFu, nb = (leaf[off] >> 12) & 0xFFu;
    if (nb == 0) return RTQ_OK;
    if (size_t(off) + 1 + 3 * size_t(ntri) + nb > leafWords) return fail(RTQ_ERR_BAD_BLOB, "leaf data overrun");
    std::vector<DevInst> list(nb);
    for (uint32_t k = 0; k < nb; ++k) {
      const uint32_t b = leaf[off + 1 + 3 * ntri + k];
      if (b >= numHeaders) return fail(RTQ_ERR_BAD_BLOB, "BLAS index beyond the header range");
      float hb[8];
      std::memcpy(hb, headers + 32 * size_t(b), 32);
      uint32_t root;
      std::memcpy(&root, headers + 32 * size_t(b) + 28, 4);
      auto it = blasRootRef.find(root);
      if (it == blasRootRef.end()) {
        uint32_t rref = kRefEmpty, d = 0;
        rtq_status e = layoutTree(root, blasLeaf, rref, d);
        if (e != RTQ_OK) return e;
        blasDepth = std::max(blasDepth, d);
        it = blasRootRef.emplace(root, rref).first;
      }
      DevInst& in = list[k];
      in.mn[0] = hb[0]; in.mn[1] = hb[1]; in.mn[2] = hb[2];
      in.mx[0] = hb[4]; in.mx[1] = hb[5]; in.mx[2] = hb[6];
      in.root = it->second;
      in.last = k + 1 == nb ? 1u : 0u;
    }
    if (h.insts.size() + list.size() > kRefIndex) return fail(RTQ_ERR_LIMIT, "more than 2^30 instance entries");
    ref = kRefLeaf | kRefInst | uint32_t(h.insts.size());
    h.insts.insert(h.insts.end(), list.begin(), list.end());
    return RTQ_OK;
  };
  uint32_t topDepth = 0;
  rtq_status st = layoutTree(0u, topLeaf, h.root, topDepth);
  if (st != RTQ_OK) return st;
  // a lane's stack holds the top-level walk's deferred children, one resume entry for the rest of the
  // current leaf's instances, and the BLAS walk's deferred children
  h.maxDepth = topDepth + 1 + blasDepth;
  if (h.maxDepth > 64) return fail(RTQ_ERR_LIMIT, "top-level plus BLAS depth above 63");
  h.numInner = uint32_t(h.nodes.size());
  boxFlags(h);
  for (const DevInst& in : h.insts)
    for (float x : {in.mn[0], in.mn[1], in.mn[2], in.mx[0], in.mx[1], in.mx[2]})
      if (!fastBound(x)) h.fastBoxes = false;
  h.leafOff.resize(numNodes);
  for (uint32_t id = 0; id < numNodes; ++id) h.leafOff[id] = nodes[16 * id + 3];
  h.leafData.assign(leaf, leaf + leafWords);
  h.lights.resize(size_t(numLights) * 4);
  h.lightKinds.resize(numLights);
  for (uint32_t i = 0; i < numLights; ++i) {
    const uint32_t* L = lightsRaw + 14 * size_t(i);
    h.lightKinds[i] = L[0];
    std::memcpy(&h.lights[4 * size_t(i)], L + 1, 16);
  }
  h.triangles.assign(tris, tris + 3 * size_t(numTris));
  h.numNodes = numNodes;
  h.leafWords = leafWords;
  h.numLights = numLights;
  h.numTris = numTris;
  out = std::move(h);
  return RTQ_OK;
}
}  // namespace detail

// Returns RTQ_OK, or RTQ_ERR_BAD_BLOB / RTQ_ERR_LIMIT with *err naming the problem.
inline rtq_status buildHostScene(const uint8_t* blob, size_t bytes, const rtq_ranges* ranges, const float* positions,
                                 uint32_t numVertices, HostScene& out, const char** err) {
  auto fail = [&](rtq_status s, const char* m) {
    *err = m;
    return s;
  };
  auto inBlob = [&](const rtq_range& r) { return size_t(r.offset) + r.size <= bytes && r.offset % 4 == 0; };
  if (!inBlob(ranges->triangles) || !inBlob(ranges->lights) || !inBlob(ranges->nodes) || !inBlob(ranges->leaf_data))
    return fail(RTQ_ERR_BAD_BLOB, "a range points outside the blob");
  HostScene h;
  const uint32_t numNodes = ranges->nodes.size / 64u;
  const uint32_t leafWords = ranges->leaf_data.size / 4u;
  const uint32_t numTris = ranges->triangles.size / 12u;
  const uint32_t numLights = ranges->lights.size / 56u;
  if (numNodes == 0) return fail(RTQ_ERR_BAD_BLOB, "no nodes");
  const uint32_t* nodes = reinterpret_cast<const uint32_t*>(blob + ranges->nodes.offset);
  const uint32_t* leaf = reinterpret_cast<const uint32_t*>(blob + ranges->leaf_data.offset);
  const uint32_t* tris = reinterpret_cast<const uint32_t*>(blob + ranges->triangles.offset);
  const uint32_t* lightsRaw = reinterpret_cast<const uint32_t*>(blob + ranges->lights.offset);

  // a leaf listing BLAS instances makes this a two-level blob
  for (uint32_t id = 0; id < numNodes; ++id) {
    const uint32_t off = nodes[16 * id + 3];
    if (off != 0xFFFFFFFFu && off < leafWords && ((leaf[off] >> 12) & 0xFFu) != 0)
      return detail::buildTwoLevel(blob, bytes, ranges, positions, numVertices, out, err);
  }

  // triangle triple -> first triangle index (the hit record's tri_id)
  std::unordered_map<detail::TriKey, uint32_t, detail::TriKeyHash> triIndex;
  triIndex.reserve(size_t(numTris) * 2);
  for (uint32_t i = 0; i < numTris; ++i)
    triIndex.emplace(detail::TriKey{tris[3 * i], tris[3 * i + 1], tris[3 * i + 2]}, i);

  // Walk the tree from the root through child pointers (bvhGetter.glsl:17-25).
  std::vector<uint8_t> isLeaf(numNodes, 0), seen(numNodes, 0);
  const bool rootIsLeaf = numNodes == 1;
  {
    struct Item { uint32_t id, depth; bool leaf; };
    std::vector<Item> st{{0u, 0u, rootIsLeaf}};
    while (!st.empty()) {
      const Item it = st.back();
      st.pop_back();
      if (it.id >= numNodes) return fail(RTQ_ERR_BAD_BLOB, "child id beyond node count");
      if (seen[it.id]) return fail(RTQ_ERR_BAD_BLOB, "node reached twice");
      seen[it.id] = 1;
      if (it.depth > 64) return fail(RTQ_ERR_LIMIT, "tree deeper than 64 levels");
      if (it.leaf) {
        isLeaf[it.id] = 1;
        h.maxDepth = std::max(h.maxDepth, it.depth);
        continue;
      }
      const uint32_t packed = nodes[16 * it.id + 2];
      const uint32_t l = packed & 0x3FFFFFFFu;
      if (l + 1u >= numNodes) return fail(RTQ_ERR_BAD_BLOB, "child id beyond node count");
      st.push_back({l + 1u, it.depth + 1u, (packed & 0x40000000u) != 0});
      st.push_back({l, it.depth + 1u, (packed & 0x80000000u) != 0});
    }
  }

  // Device order: the two children of an inner node side by side in one 128-byte pair (a pop then
  // usually finds its node in the line its sibling already brought into the cache); the first kTopSlots
  // records hold the top levels breadth first (kernels copy them into LDS), the rest follow depth first,
  // and the triangle slots of the leaves in the order their parents are laid out. Leaf children need no
  // record; their half of a pair stays empty. Traversal order is fixed by the child refs alone, so the
  // layout changes no result.
  std::vector<uint32_t> innerIndex(numNodes, 0xFFFFFFFFu);
  std::vector<uint32_t> leafSlot(numNodes, kRefEmpty);
  std::vector<uint32_t> slotOwner;  // device inner slot -> reference node id (0xFFFFFFFF: empty)
  auto addLeaf = [&](uint32_t id) -> rtq_status {
    return detail::appendLeafSlots(nodes, leaf, leafWords, positions, numVertices, triIndex, id, h, leafSlot[id], err);
  };
  rtq_status st = RTQ_OK;
  if (rootIsLeaf) {
    if ((st = addLeaf(0)) != RTQ_OK) return st;
    h.root = leafSlot[0];
  } else {
    innerIndex[0] = 0;
    slotOwner = {0u, 0xFFFFFFFFu};  // the root alone in the first pair
    std::deque<uint32_t> wide{0u};  // breadth first while the top records fill
    std::vector<uint32_t> todo;     // depth first afterwards
    while (!wide.empty() || !todo.empty()) {
      const bool fromWide = !wide.empty();
      const uint32_t id = fromWide ? wide.front() : todo.back();
      if (fromWide) wide.pop_front();
      else todo.pop_back();
      const uint32_t l = nodes[16 * id + 2] & 0x3FFFFFFFu;
      const uint32_t c[2] = {l, l + 1u};
      if (!isLeaf[c[0]] || !isLeaf[c[1]]) {
        const uint32_t base = uint32_t(slotOwner.size());
        slotOwner.push_back(0xFFFFFFFFu);
        slotOwner.push_back(0xFFFFFFFFu);
        for (uint32_t k = 0; k < 2; ++k)
          if (!isLeaf[c[k]]) {
            innerIndex[c[k]] = base + k;
            slotOwner[base + k] = c[k];
          }
      }
      for (int k = 0; k < 2; ++k)
        if (isLeaf[c[k]] && (st = addLeaf(c[k])) != RTQ_OK) return st;
      const bool wideChildren = fromWide && slotOwner.size() < kTopSlots;
      for (int k = 0; k < 2; ++k)
        if (!isLeaf[c[k]] && wideChildren) wide.push_back(c[k]);
      for (int k = 1; k >= 0; --k)  // walk child 0's subtree first
        if (!isLeaf[c[k]] && !wideChildren) todo.push_back(c[k]);
    }
    h.root = 0;
    h.topSlots = uint32_t(std::min<size_t>(kTopSlots, slotOwner.size()));
  }
  auto ref = [&](uint32_t id) { return isLeaf[id] ? leafSlot[id] : innerIndex[id]; };
  h.nodes.assign(slotOwner.size(), DevNode{});
  for (size_t k = 0; k < slotOwner.size(); ++k) {
    const uint32_t id = slotOwner[k];
    if (id == 0xFFFFFFFFu) continue;
    ++h.numInner;
    const uint32_t* n = nodes + 16 * id;
    const float* f = reinterpret_cast<const float*>(n);
    const float* n0xy = f + 4;
    const float* n1xy = f + 8;
    const float* nz = f + 12;
    DevNode& d = h.nodes[k];
    d.a[0] = n0xy[0]; d.a[1] = n0xy[1]; d.a[2] = nz[0]; d.a[3] = n0xy[2];
    d.b[0] = n0xy[3]; d.b[1] = nz[1]; d.b[2] = n1xy[0]; d.b[3] = n1xy[1];
    d.c[0] = nz[2]; d.c[1] = n1xy[2]; d.c[2] = n1xy[3]; d.c[3] = nz[3];
    const uint32_t l = n[2] & 0x3FFFFFFFu;
    d.d[0] = ref(l);
    d.d[1] = ref(l + 1);
    d.d[2] = id;
    d.d[3] = 0;
  }
  detail::boxFlags(h);
  h.leafOff.resize(numNodes);
  for (uint32_t id = 0; id < numNodes; ++id) h.leafOff[id] = nodes[16 * id + 3];
  h.leafData.assign(leaf, leaf + leafWords);
  h.lights.resize(size_t(numLights) * 4);
  h.lightKinds.resize(numLights);
  for (uint32_t i = 0; i < numLights; ++i) {
    const uint32_t* L = lightsRaw + 14 * size_t(i);
    h.lightKinds[i] = L[0];
    std::memcpy(&h.lights[4 * size_t(i)], L + 1, 16);
  }
  h.triangles.assign(tris, tris + 3 * size_t(numTris));
  h.numNodes = numNodes;
  h.leafWords = leafWords;
  h.numLights = numLights;
  h.numTris = numTris;
  out = std::move(h);
  return RTQ_OK;
}

}  // namespace rtq
