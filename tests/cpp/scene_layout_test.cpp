// CPU test of upload-BVH's host half (tim_raytracing_amd/csrc/device/scene_layout.h): the device
// layout built from builder blobs keeps every reference node, box, child link, leaf and triangle, lays
// sibling inner nodes out as aligned pairs in depth-first order, and rejects broken blobs. Prints "ok".
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <set>
#include <vector>

#include "../../include/rtq_build.h"
#include "../../tim_raytracing_amd/csrc/device/scene_layout.h"
#include "../../tim_raytracing_amd/csrc/device/stack_class.h"

namespace {

int g_fail = 0;
#define CHECK(c)                                                     \
  do {                                                               \
    if (!(c)) {                                                      \
      std::fprintf(stderr, "%s:%d: CHECK(%s)\n", __FILE__, __LINE__, #c); \
      ++g_fail;                                                      \
      return;                                                        \
    }                                                                \
  } while (0)

struct Built {
  std::vector<uint8_t> blob;
  rtq_ranges ranges{};
  std::vector<float> pos;
};

Built build(uint32_t kind, uint64_t seed, uint32_t param, const std::vector<float>* tri = nullptr) {
  rtq_builder* b = nullptr;
  rtq_builder_create(&b);
  if (tri) {
    uint32_t off = 0;
    rtq_builder_add_vertices(b, tri->data(), uint32_t(tri->size() / 3), &off);
    std::vector<uint32_t> idx(tri->size() / 3);
    for (uint32_t i = 0; i < idx.size(); ++i) idx[i] = i;
    rtq_builder_add_triangle_list(b, off, uint32_t(idx.size() / 3), idx.data(), nullptr);
  } else {
    rtq_scene_generate(b, kind, seed, param);
  }
  rtq_builder_build(b, 2);
  Built out;
  size_t bytes = 0;
  rtq_builder_gpu_size(b, &bytes);
  out.blob.resize(bytes);
  rtq_builder_fill_gpu_buffer(b, out.blob.data(), bytes, &out.ranges);
  const float* xyz = nullptr;
  uint32_t n = 0;
  rtq_builder_positions(b, &xyz, &n);
  out.pos.assign(xyz, xyz + 3 * size_t(n));
  rtq_builder_destroy(b);
  return out;
}

void checkScene(const Built& s) {
  rtq::HostScene h;
  const char* why = "";
  CHECK(rtq::buildHostScene(s.blob.data(), s.blob.size(), &s.ranges, s.pos.data(), uint32_t(s.pos.size() / 3), h,
                            &why) == RTQ_OK);
  const uint32_t* nodes = reinterpret_cast<const uint32_t*>(s.blob.data() + s.ranges.nodes.offset);
  const uint32_t* leaf = reinterpret_cast<const uint32_t*>(s.blob.data() + s.ranges.leaf_data.offset);
  const uint32_t* tris = reinterpret_cast<const uint32_t*>(s.blob.data() + s.ranges.triangles.offset);
  const uint32_t numNodes = s.ranges.nodes.size / 64;
  CHECK(h.numNodes == numNodes && h.nodes.size() % 2 == 0);
  std::set<uint32_t> refNodes, usedSlots, innerSlots;
  std::vector<uint32_t> pairOrder;  // pair bases in depth-first (child 0 first) order
  std::vector<uint32_t> slotDepth(h.nodes.size(), 0xFFFFFFFFu);
  uint32_t inner = 0, leaves = 0;
  bool ordered = true;  // every reachable inner node's two boxes have min <= max per axis
  // walk device refs and reference links side by side
  struct Item { uint32_t ref, id, depth; };
  std::vector<Item> st;
  if (numNodes == 1) {
    st.push_back({h.root, 0, 0});
  } else {
    CHECK(h.root == 0);
    st.push_back({0u, 0u, 0u});
  }
  while (!st.empty()) {
    const Item it = st.back();
    st.pop_back();
    CHECK(refNodes.insert(it.id).second);
    if (it.ref == rtq::kRefEmpty || (it.ref & rtq::kRefLeaf)) {
      ++leaves;
      const uint32_t off = nodes[16 * it.id + 3];
      const uint32_t count = off == 0xFFFFFFFFu ? 0u : (leaf[off] & 0xFFFu);
      if (count == 0) {
        CHECK(it.ref == rtq::kRefEmpty);
        continue;
      }
      CHECK(it.ref != rtq::kRefEmpty);
      const uint32_t first = it.ref & rtq::kLeafSlotMask;
      CHECK(((it.ref >> rtq::kLeafCountShift) & 15u) == std::min(count, 16u) - 1u);  // leaf size in the ref
      CHECK((it.ref & rtq::kRefInst) == 0);
      for (uint32_t k = 0; k < count; ++k) {
        CHECK(first + k < h.tris.size());
        CHECK(usedSlots.insert(first + k).second);
        const rtq::DevTri& d = h.tris[first + k];
        uint32_t w[4];
        std::memcpy(w, d.t2, 16);
        CHECK(w[2] == it.id && w[3] == (k + 1 == count ? 1u : 0u));
        const uint32_t* t = leaf + off + 1 + 3 * k;
        CHECK(w[1] < s.ranges.triangles.size / 12);
        CHECK(std::memcmp(tris + 3 * w[1], t, 12) == 0);
        const uint32_t v0 = t[0] + (t[1] & 0xFFFFu);
        CHECK(std::memcmp(d.t0, &s.pos[3 * size_t(v0)], 12) == 0);
      }
      continue;
    }
    ++inner;
    CHECK(it.ref < h.nodes.size());
    CHECK(innerSlots.insert(it.ref).second);
    slotDepth[it.ref] = it.depth;
    const rtq::DevNode& d = h.nodes[it.ref];
    CHECK(d.d[2] == it.id);
    const float* f = reinterpret_cast<const float*>(nodes + 16 * it.id);
    const float want[12] = {f[4], f[5], f[12], f[6], f[7], f[13], f[8], f[9], f[14], f[10], f[11], f[15]};
    const float got[12] = {d.a[0], d.a[1], d.a[2], d.a[3], d.b[0], d.b[1], d.b[2], d.b[3], d.c[0], d.c[1], d.c[2], d.c[3]};
    CHECK(std::memcmp(want, got, sizeof(want)) == 0);
    for (int k = 0; k < 3; ++k) ordered = ordered && want[k] <= want[3 + k] && want[6 + k] <= want[9 + k];
    const uint32_t packed = nodes[16 * it.id + 2];
    const uint32_t l = packed & 0x3FFFFFFFu;
    const bool leafChild[2] = {(packed & 0x80000000u) != 0, (packed & 0x40000000u) != 0};
    uint32_t base = 0xFFFFFFFFu;
    for (int k = 0; k < 2; ++k) {
      if (leafChild[k]) {
        CHECK(d.d[k] == rtq::kRefEmpty || (d.d[k] & rtq::kRefLeaf));
        continue;
      }
      CHECK(d.d[k] < h.nodes.size() && (d.d[k] & 1u) == uint32_t(k));  // child k in half k of its pair
      const uint32_t b = d.d[k] - uint32_t(k);
      CHECK(base == 0xFFFFFFFFu || base == b);                         // both inner children share the pair
      base = b;
    }
    if (base != 0xFFFFFFFFu) pairOrder.push_back(base);
    for (int k = 1; k >= 0; --k) st.push_back({d.d[k], l + uint32_t(k), it.depth + 1});
  }
  CHECK(inner == h.numInner);
  CHECK(h.orderedBoxes == ordered);
  CHECK(usedSlots.size() == h.tris.size());
  CHECK(h.nodes.size() == (numNodes == 1 ? 0u : 2 + 2 * pairOrder.size()));
  // the first topSlots records hold the top levels breadth first: depths never decrease along them and
  // no record after them is shallower than any of them
  CHECK(h.topSlots == std::min<size_t>(rtq::kTopSlots, h.nodes.size()));
  uint32_t topMax = 0;
  for (uint32_t k = 0; k < h.topSlots; ++k) {
    if (slotDepth[k] == 0xFFFFFFFFu) continue;
    CHECK(slotDepth[k] >= topMax);
    topMax = slotDepth[k];
  }
  for (size_t k = h.topSlots; k < h.nodes.size(); ++k)
    if (slotDepth[k] != 0xFFFFFFFFu) CHECK(slotDepth[k] >= topMax);
  if (rtq::kTopSlots == 0) {
    // all depth first: pairs appear in walk order right after the root's pair
    for (size_t i = 0; i < pairOrder.size(); ++i) CHECK(pairOrder[i] == 2 + 2 * i);
  } else {
    // pairs are dense whatever the order
    std::set<uint32_t> bases(pairOrder.begin(), pairOrder.end());
    CHECK(bases.size() == pairOrder.size());
    for (size_t i = 0; i < pairOrder.size(); ++i) CHECK(bases.count(uint32_t(2 + 2 * i)) == 1);
  }
  // every slot that holds no reachable inner node is an all-zero record
  static const rtq::DevNode zero{};
  for (size_t k = 0; k < h.nodes.size(); ++k)
    if (!innerSlots.count(uint32_t(k))) CHECK(std::memcmp(&h.nodes[k], &zero, sizeof(zero)) == 0);
}

// The kernels' stack discipline in its worst case, where both child boxes of every inner node pass:
// the far child is pushed, the near one entered, and a finished leaf pops. The deepest stack such a
// walk reaches never exceeds the tree depth, so the stack class chosen for that depth always holds it.
void checkStackBound(const Built& s) {
  rtq::HostScene h;
  const char* why = "";
  CHECK(rtq::buildHostScene(s.blob.data(), s.blob.size(), &s.ranges, s.pos.data(), uint32_t(s.pos.size() / 3), h,
                            &why) == RTQ_OK);
  std::vector<uint32_t> stk;
  size_t deepest = 0;
  uint32_t node = h.root;
  for (;;) {
    while (node < rtq::kRefDone) {
      const rtq::DevNode& d = h.nodes[node];
      stk.push_back(d.d[1]);  // far child deferred
      deepest = std::max(deepest, stk.size());
      node = d.d[0];
    }
    if (stk.empty()) break;
    node = stk.back();
    stk.pop_back();
  }
  CHECK(deepest <= h.maxDepth);
  CHECK(rtq::gridStackClass(h.maxDepth) != 0 && rtq::stackHolds(rtq::gridStackClass(h.maxDepth), uint32_t(deepest)));
  CHECK(rtq::batchStackClass(h.maxDepth) != 0 && rtq::stackHolds(rtq::batchStackClass(h.maxDepth), uint32_t(deepest)));
}

// The depth -> stack mapping: every depth the upload accepts gets a class that holds it, the classes
// are the ones the kernels are built for, and deeper trees get none.
void checkStackClasses() {
  for (uint32_t d = 0; d <= rtq::kMaxTreeDepth; ++d) {
    const uint32_t g = rtq::gridStackClass(d), b = rtq::batchStackClass(d);
    CHECK(g == 22 || g == 24 || g == 32 || g == 64);
    CHECK(b == 24 || b == 32 || b == 64);
    CHECK(rtq::stackHolds(g, d) && rtq::stackHolds(b, d));
    if (d > 0) CHECK(g >= rtq::gridStackClass(d - 1) && b >= rtq::batchStackClass(d - 1));
  }
  CHECK(rtq::gridStackClass(21) == 22 && rtq::gridStackClass(23) == 24 && rtq::gridStackClass(25) == 32);
  CHECK(rtq::gridStackClass(rtq::kMaxTreeDepth + 1) == 0 && rtq::batchStackClass(rtq::kMaxTreeDepth + 1) == 0);
  CHECK(!rtq::stackHolds(24, 25));
}

void checkRejects(Built s) {
  rtq::HostScene h;
  const char* why = "";
  rtq_ranges bad = s.ranges;
  bad.nodes.offset = uint32_t(s.blob.size());  // outside the blob
  CHECK(rtq::buildHostScene(s.blob.data(), s.blob.size(), &bad, s.pos.data(), uint32_t(s.pos.size() / 3), h, &why) ==
        RTQ_ERR_BAD_BLOB);
  // a child id beyond the node count
  uint32_t* nodes = reinterpret_cast<uint32_t*>(s.blob.data() + s.ranges.nodes.offset);
  const uint32_t keep = nodes[2];
  nodes[2] = (keep & 0xC0000000u) | 0x3FFFFFF0u;
  CHECK(rtq::buildHostScene(s.blob.data(), s.blob.size(), &s.ranges, s.pos.data(), uint32_t(s.pos.size() / 3), h,
                            &why) == RTQ_ERR_BAD_BLOB);
  nodes[2] = keep;
  // too few positions for the triangles
  CHECK(rtq::buildHostScene(s.blob.data(), s.blob.size(), &s.ranges, s.pos.data(), 2, h, &why) == RTQ_ERR_BAD_BLOB);
  CHECK(rtq::buildHostScene(s.blob.data(), s.blob.size(), &s.ranges, s.pos.data(), uint32_t(s.pos.size() / 3), h,
                            &why) == RTQ_OK);
}

// The ordered-box flag (which lets the kernels skip the per-axis min/max of the slab values): set for
// builder output, cleared by one inverted box or one NaN bound.
void checkOrderedFlag(Built s) {
  rtq::HostScene h;
  const char* why = "";
  const uint32_t n3 = uint32_t(s.pos.size() / 3);
  CHECK(rtq::buildHostScene(s.blob.data(), s.blob.size(), &s.ranges, s.pos.data(), n3, h, &why) == RTQ_OK);
  CHECK(h.orderedBoxes && h.fastBoxes);
  float* f = reinterpret_cast<float*>(s.blob.data() + s.ranges.nodes.offset);  // node 0 (the root) is inner
  std::swap(f[14], f[15]);  // box 1 z: min above max
  CHECK(f[14] != f[15]);
  CHECK(rtq::buildHostScene(s.blob.data(), s.blob.size(), &s.ranges, s.pos.data(), n3, h, &why) == RTQ_OK);
  CHECK(!h.orderedBoxes && h.fastBoxes);
  std::swap(f[14], f[15]);
  const float keep = f[4];
  f[4] = __builtin_nanf("");  // box 0 min x
  CHECK(rtq::buildHostScene(s.blob.data(), s.blob.size(), &s.ranges, s.pos.data(), n3, h, &why) == RTQ_OK);
  CHECK(!h.orderedBoxes && !h.fastBoxes);
  f[4] = keep;
  CHECK(rtq::buildHostScene(s.blob.data(), s.blob.size(), &s.ranges, s.pos.data(), n3, h, &why) == RTQ_OK);
  CHECK(h.orderedBoxes && h.fastBoxes);
}

}  // namespace

// A two-level scene: nblas clusters of small triangles as BLAS instances plus a few top-level triangles.
Built buildTwoLevelScene(uint32_t nblas, uint32_t trisPer, uint32_t seed) {
  rtq_builder* b = nullptr;
  rtq_builder_create(&b);
  const rtq_build_params top{25, 6, 6, 1.0f, 0.225f};
  rtq_builder_set_parameters(b, &top);
  uint32_t state = seed * 2654435761u + 1u;
  auto rnd = [&] {  // uniform in [-1, 1)
    state = state * 1664525u + 1013904223u;
    return float(state >> 8) / float(1u << 23) - 1.0f;
  };
  std::vector<uint32_t> idx(3 * trisPer);
  for (uint32_t i = 0; i < idx.size(); ++i) idx[i] = i;
  for (uint32_t k = 0; k < nblas; ++k) {
    const float c[3] = {30 * rnd(), 30 * rnd(), 30 * rnd()};
    std::vector<float> v;
    for (uint32_t t = 0; t < trisPer; ++t) {
      const float tc[3] = {c[0] + 3 * rnd(), c[1] + 3 * rnd(), c[2] + 3 * rnd()};
      for (int q = 0; q < 3; ++q)
        for (int a = 0; a < 3; ++a) v.push_back(tc[a] + 0.5f * rnd());
    }
    uint32_t off = 0, bi = 0;
    rtq_builder_add_vertices(b, v.data(), uint32_t(v.size() / 3), &off);
    rtq_builder_add_blas(b, off, trisPer, idx.data(), nullptr, &bi);
  }
  const std::vector<float> own = {-40, -40, -45, 40, -40, -44, 0, 40, -46, -40, 40, -45, 40, 40, -44, 0, -40, -46};
  uint32_t off = 0;
  rtq_builder_add_vertices(b, own.data(), 6, &off);
  rtq_builder_add_triangle_list(b, off, 2, idx.data(), nullptr);
  rtq_builder_build(b, 2);
  Built out;
  size_t bytes = 0;
  rtq_builder_gpu_size(b, &bytes);
  out.blob.resize(bytes);
  rtq_builder_fill_gpu_buffer(b, out.blob.data(), bytes, &out.ranges);
  const float* xyz = nullptr;
  uint32_t n = 0;
  rtq_builder_positions(b, &xyz, &n);
  out.pos.assign(xyz, xyz + 3 * size_t(n));
  rtq_builder_destroy(b);
  return out;
}

// Two-level decode: every top-level inner node and every BLAS inner node is one record with the blob's
// boxes; a top-level leaf's ref names its instance run (header boxes, the BLAS root's ref, last flag);
// BLAS leaves name their triangle slots; each BLAS is laid out once however many leaves list it.
void checkTwoLevel(const Built& s, uint32_t nblas) {
  rtq::HostScene h;
  const char* why = "";
  CHECK(rtq::buildHostScene(s.blob.data(), s.blob.size(), &s.ranges, s.pos.data(), uint32_t(s.pos.size() / 3), h,
                            &why) == RTQ_OK);
  CHECK(h.twoLevel);
  const uint32_t* nodes = reinterpret_cast<const uint32_t*>(s.blob.data() + s.ranges.nodes.offset);
  const uint32_t* leaf = reinterpret_cast<const uint32_t*>(s.blob.data() + s.ranges.leaf_data.offset);
  const uint8_t* hdr = s.blob.data() + s.ranges.blas.offset;
  CHECK(s.ranges.blas.size == 32 * nblas);
  std::vector<int> blasRef(nblas, -1);
  std::set<uint32_t> recs, slots;
  uint32_t listed = 0;
  // walk(ref, id, blas): blas < 0 for the top level
  struct Item { uint32_t ref, id; int blas; };
  std::vector<Item> st{{h.root, 0u, -1}};
  std::set<std::pair<uint32_t, uint32_t>> blasWalked;
  while (!st.empty()) {
    const Item it = st.back();
    st.pop_back();
    const bool isLeaf = (nodes[16 * it.id + 2] & 0x3FFFFFFFu) == 0x3FFFFFFFu;
    if (!isLeaf) {
      CHECK(it.ref < h.nodes.size() && recs.insert(it.ref).second);
      const rtq::DevNode& d = h.nodes[it.ref];
      CHECK(d.d[2] == it.id);
      const float* f = reinterpret_cast<const float*>(nodes + 16 * it.id);
      const float want[12] = {f[4], f[5], f[12], f[6], f[7], f[13], f[8], f[9], f[14], f[10], f[11], f[15]};
      const float got[12] = {d.a[0], d.a[1], d.a[2], d.a[3], d.b[0], d.b[1], d.b[2], d.b[3], d.c[0], d.c[1], d.c[2], d.c[3]};
      CHECK(std::memcmp(want, got, sizeof(want)) == 0);
      const uint32_t l = nodes[16 * it.id + 2] & 0x3FFFFFFFu;
      st.push_back({d.d[1], l + 1, it.blas});
      st.push_back({d.d[0], l, it.blas});
      continue;
    }
    const uint32_t off = nodes[16 * it.id + 3];
    const uint32_t ntri = off == 0xFFFFFFFFu ? 0u : leaf[off] & 0xFFFu;
    const uint32_t nb = off == 0xFFFFFFFFu ? 0u : (leaf[off] >> 12) & 0xFFu;
    if (it.blas < 0) {
      if (nb == 0) {
        CHECK(it.ref == rtq::kRefEmpty);  // the top level's own triangles are not laid out
        continue;
      }
      CHECK((it.ref & rtq::kRefLeaf) && (it.ref & rtq::kRefInst));
      const uint32_t first = it.ref & rtq::kRefIndex;
      for (uint32_t k = 0; k < nb; ++k) {
        CHECK(first + k < h.insts.size() && slots.insert(first + k).second);
        const rtq::DevInst& in = h.insts[first + k];
        const uint32_t b = leaf[off + 1 + 3 * ntri + k];
        CHECK(b < nblas);
        float hb[8];
        uint32_t root;
        std::memcpy(hb, hdr + 32 * b, 32);
        std::memcpy(&root, hdr + 32 * b + 28, 4);
        const float want[6] = {hb[0], hb[1], hb[2], hb[4], hb[5], hb[6]};
        const float got[6] = {in.mn[0], in.mn[1], in.mn[2], in.mx[0], in.mx[1], in.mx[2]};
        CHECK(std::memcmp(want, got, sizeof(want)) == 0);
        CHECK(in.last == (k + 1 == nb ? 1u : 0u));
        CHECK(blasRef[b] < 0 || uint32_t(blasRef[b]) == in.root);
        blasRef[b] = int(in.root);
        ++listed;
        if (blasWalked.insert({b, root}).second) st.push_back({in.root, root, int(b)});
      }
      continue;
    }
    CHECK(nb == 0);
    if (ntri == 0) {
      CHECK(it.ref == rtq::kRefEmpty);
      continue;
    }
    const uint32_t first = it.ref & rtq::kLeafSlotMask;
    CHECK((it.ref & rtq::kRefLeaf) != 0 && (it.ref & rtq::kRefInst) == 0);
    CHECK(((it.ref >> rtq::kLeafCountShift) & 15u) == std::min(ntri, 16u) - 1u);
    for (uint32_t k = 0; k < ntri; ++k) {
      CHECK(first + k < h.tris.size());
      uint32_t w[4];
      std::memcpy(w, h.tris[first + k].t2, 16);
      CHECK(w[2] == it.id && w[3] == (k + 1 == ntri ? 1u : 0u));
      const uint32_t* t = leaf + off + 1 + 3 * k;
      const uint32_t v0 = t[0] + (t[1] & 0xFFFFu);
      CHECK(std::memcmp(h.tris[first + k].t0, &s.pos[3 * size_t(v0)], 12) == 0);
    }
  }
  CHECK(listed == h.insts.size() && slots.size() == h.insts.size());
  CHECK(recs.size() == h.nodes.size());
  CHECK(h.maxDepth <= rtq::kMaxTreeDepth);
  // the reciprocal slab form is allowed exactly when every node and header bound qualifies
  bool fast = true, ordered = true;
  for (const rtq::DevNode& d : h.nodes) {
    for (int k = 0; k < 4; ++k)
      for (float x : {d.a[k], d.b[k], d.c[k]}) fast = fast && rtq::detail::fastBound(x);
    ordered = ordered && d.a[0] <= d.a[3] && d.a[1] <= d.b[0] && d.a[2] <= d.b[1] && d.b[2] <= d.c[1] &&
              d.b[3] <= d.c[2] && d.c[0] <= d.c[3];
  }
  for (const rtq::DevInst& in : h.insts)
    for (float x : {in.mn[0], in.mn[1], in.mn[2], in.mx[0], in.mx[1], in.mx[2]}) fast = fast && rtq::detail::fastBound(x);
  CHECK(h.fastBoxes == fast && h.orderedBoxes == ordered);
}

// The stack bound of the two-level walk: the deepest top-level leaf's deferred children, one resume
// entry, then the deepest BLAS's, never exceed maxDepth.
void checkTwoLevelStack(const Built& s) {
  rtq::HostScene h;
  const char* why = "";
  CHECK(rtq::buildHostScene(s.blob.data(), s.blob.size(), &s.ranges, s.pos.data(), uint32_t(s.pos.size() / 3), h,
                            &why) == RTQ_OK);
  // worst case: every box passes, so the walk defers one child per inner level it descends
  struct Item { uint32_t ref, depth; };
  uint32_t deepest = 0;
  std::vector<Item> st{{h.root, 0}};
  while (!st.empty()) {
    const Item it = st.back();
    st.pop_back();
    if (it.ref == rtq::kRefEmpty) continue;
    if (it.ref & rtq::kRefLeaf) {
      if (it.ref & rtq::kRefInst) {
        for (uint32_t k = it.ref & rtq::kRefIndex;; ++k) {
          const uint32_t extra = h.insts[k].last ? 0u : 1u;
          st.push_back({h.insts[k].root, it.depth + extra});
          if (h.insts[k].last) break;
        }
      }
      deepest = std::max(deepest, it.depth);
      continue;
    }
    const rtq::DevNode& d = h.nodes[it.ref];
    st.push_back({d.d[0], it.depth + 1});
    st.push_back({d.d[1], it.depth + 1});
  }
  CHECK(deepest <= h.maxDepth);
}

int main() {
  checkScene(build(RTQ_SCENE_BOX, 1, 0));
  checkScene(build(RTQ_SCENE_SPHERES, 2, 64));
  checkScene(build(RTQ_SCENE_SOUP, 3, 3000));
  const std::vector<float> one = {0, 0, 0, 1, 0, 0, 0, 1, 0};
  checkScene(build(0, 0, 0, &one));  // a single leaf as the root
  const std::vector<float> two = {0, 0, 0, 1, 0, 0, 0, 1, 0, 5, 5, 5, 6, 5, 5, 5, 6, 5};
  checkScene(build(0, 0, 0, &two));
  checkStackClasses();
  checkStackBound(build(RTQ_SCENE_BOX, 1, 0));
  checkStackBound(build(RTQ_SCENE_SPHERES, 2, 64));
  checkStackBound(build(RTQ_SCENE_SOUP, 3, 3000));
  checkRejects(build(RTQ_SCENE_SPHERES, 4, 8));
  checkOrderedFlag(build(RTQ_SCENE_SPHERES, 5, 16));
  checkOrderedFlag(build(RTQ_SCENE_SOUP, 6, 2000));
  checkTwoLevel(buildTwoLevelScene(20, 200, 1), 20);
  checkTwoLevel(buildTwoLevelScene(3, 50, 2), 3);   // a top level that is one leaf
  checkTwoLevel(buildTwoLevelScene(40, 3, 3), 40);  // BLASes that are one leaf each
  checkTwoLevelStack(buildTwoLevelScene(20, 200, 1));
  checkTwoLevelStack(buildTwoLevelScene(3, 50, 2));
  if (g_fail) return 1;
  std::puts("ok");
  return 0;
}
