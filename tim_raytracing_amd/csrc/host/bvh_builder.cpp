// BVH builder restatement. Every float operation below is written in the reference's evaluation
// order (compile with -ffp-contract=off, no fast-math) so that node boxes, split decisions and the
// emitted blob are the ones Renderer/BVHBuilder.cpp produces for the same input.
#include "bvh_builder.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <iterator>
#include <limits>
#include <mutex>
#include <stdexcept>
#include <thread>
#include <unordered_set>

#include "tribox.h"

namespace rtq {
namespace {

// Planted mistakes for the builder's pins, compiled only into test builds (-DRTQ_BUILDER_PLANT=k); a
// library built without the macro has none. 1 = axis ties won by the earlier axis, 2 = the gain reject
// with || instead of &&, 3 = a split candidate kept on an equal score, 4 = the bisection moving away from
// balance, 5 = a straddling item grown into the side whose volume grows more, 6 = the duplicate threshold
// taken against the larger child volume, 7 = the light/node test inclusive (<=) instead of strict.
#ifndef RTQ_BUILDER_PLANT
#define RTQ_BUILDER_PLANT 0
#endif
constexpr int kPlant = RTQ_BUILDER_PLANT;

// linalg's component min/max (timCore/linalg.h:152-153): a<b ? a : b and a<b ? b : a.
inline float lmin(float a, float b) { return a < b ? a : b; }
inline float lmax(float a, float b) { return a < b ? b : a; }
inline F3 lmin3(F3 a, F3 b) { return {lmin(a.x, b.x), lmin(a.y, b.y), lmin(a.z, b.z)}; }
inline F3 lmax3(F3 a, F3 b) { return {lmax(a.x, b.x), lmax(a.y, b.y), lmax(a.z, b.z)}; }
inline F3 sub(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline F3 add(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline F3 mulS(F3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline float comp(F3 a, int axis) { return axis == 0 ? a.x : (axis == 1 ? a.y : a.z); }
// linalg::sum folds from a zero seed (linalg.h:397): ((0 + x) + y) + z.
inline float lsum(F3 a) { return ((0.0f + a.x) + a.y) + a.z; }
inline float ldot(F3 a, F3 b) { return ((0.0f + a.x * b.x) + a.y * b.y) + a.z * b.z; }

// Only the moving axis survives (the _movingAxis lambdas, BVHBuilder.cpp:789-793).
inline F3 keepAxis(F3 v, int axis) {
  return {axis == 0 ? v.x : 0.0f, axis == 1 ? v.y : 0.0f, axis == 2 ? v.z : 0.0f};
}
// The moving axis is zeroed (the _fixedAxis lambdas).
inline F3 dropAxis(F3 v, int axis) {
  return {axis == 0 ? 0.0f : v.x, axis == 1 ? 0.0f : v.y, axis == 2 ? 0.0f : v.z};
}

// mergeBox / intersectionBox (BVHBuilder.cpp:177-191); argument order matters for signed zeros.
inline Box3 mergeBox(const Box3& a, const Box3& b) { return {lmin3(a.mn, b.mn), lmax3(a.mx, b.mx)}; }
inline Box3 intersectBox(const Box3& a, const Box3& b) { return {lmax3(a.mn, b.mn), lmin3(a.mx, b.mx)}; }

// getBoxVolume (BVHBuilder.cpp:81-85).
inline float boxVolume(const Box3& b) {
  F3 d = sub(b.mx, b.mn);
  return d.x * d.y * d.z;
}

// checkBox (BVHBuilder.cpp:87-90).
inline bool checkBox(const Box3& b) { return b.mn.x <= b.mx.x && b.mn.y <= b.mx.y && b.mn.z <= b.mx.z; }

// adjustAABB (BVHBuilder.cpp:401-406).
inline Box3 adjustBox(Box3 b) {
  const float e = float(1e-6);
  b.mx = {b.mx.x + e, b.mx.y + e, b.mx.z + e};
  b.mn = {b.mn.x - e, b.mn.y - e, b.mn.z - e};
  return b;
}

// pointBoxCollision (BVHBuilder.cpp:58-63): inclusive containment.
inline bool pointInBox(F3 p, const Box3& b) {
  return !(p.x < b.mn.x || p.y < b.mn.y || p.z < b.mn.z || p.x > b.mx.x || p.y > b.mx.y || p.z > b.mx.z);
}

// boxBoxCollision != Disjoint (BVHBuilder.cpp:41-47): boxes that only touch are disjoint.
inline bool boxesOverlap(const Box3& a, const Box3& b) {
  return !(a.mn.x >= b.mx.x || a.mn.y >= b.mx.y || a.mn.z >= b.mx.z || b.mn.x >= a.mx.x || b.mn.y >= a.mx.y ||
           b.mn.z >= a.mx.z);
}

// The TLAS's hoisting of items common to both children (mergeCommonItems, BVHBuilder.cpp:907-935):
// both lists are sorted, their intersection is removed from each and returned.
std::vector<u32> mergeCommon(std::vector<u32>& left, std::vector<u32>& right) {
  std::sort(left.begin(), left.end());
  std::sort(right.begin(), right.end());
  std::vector<u32> common;
  std::set_intersection(left.begin(), left.end(), right.begin(), right.end(), std::back_inserter(common));
  std::vector<u32> newLeft, newRight;
  std::set_difference(left.begin(), left.end(), common.begin(), common.end(), std::back_inserter(newLeft));
  std::set_difference(right.begin(), right.end(), common.begin(), common.end(), std::back_inserter(newRight));
  left = std::move(newLeft);
  right = std::move(newRight);
  return common;
}

// sphereBoxCollision (BVHBuilder.cpp:33-39).
inline bool sphereTouchesBox(F3 c, float r, const Box3& b) {
  F3 closest = lmin3(lmax3(c, b.mn), b.mx);
  F3 d = sub(closest, c);
  return kPlant == 7 ? ldot(d, d) <= r * r : ldot(d, d) < r * r;
}

// computeMaxDistanceBetweenBoxDimensions (BVHBuilder.cpp:944-950): maxelem folds from x.
inline float maxDimDistance(const Box3& a, const Box3& b) {
  F3 d1 = sub(a.mx, a.mn), d2 = sub(b.mx, b.mn);
  F3 d = sub(d1, d2);
  F3 ad = {std::fabs(d.x), std::fabs(d.y), std::fabs(d.z)};
  return lmax(lmax(lmax(ad.x, ad.x), ad.y), ad.z);
}

// std::min / std::max over an initializer list keep the first extreme element.
inline float firstMin3(float a, float b, float c) {
  float r = a;
  if (b < r) r = b;
  if (c < r) r = c;
  return r;
}
inline float firstMax3(float a, float b, float c) {
  float r = a;
  if (r < b) r = b;
  if (r < c) r = c;
  return r;
}

inline u32 alignUp(u32 x, u32 a) { return x + (x % a == 0 ? 0 : a - x % a); }

constexpr u32 kBufferAlignment = 32;  // m_bufferAlignment (BVHBuilder.h:161)

// Leaf sort key of removeDuplicates (BVHBuilder.cpp:543-562): vertex offset, material, then the
// ascending vertex triple packed into a u64 (triangleVerticesUniqueID, :536-541).
struct LeafKey {
  u32 vo, mat;
  u64 verts;
  bool operator<(const LeafKey& o) const {
    if (vo != o.vo) return vo < o.vo;
    if (mat != o.mat) return mat < o.mat;
    return verts < o.verts;
  }
  bool operator==(const LeafKey& o) const { return vo == o.vo && mat == o.mat && verts == o.verts; }
};

inline LeafKey leafKey(const TriangleRec& t) {
  uint16_t ids[3] = {uint16_t(t.index01 & 0xFFFFu), uint16_t(t.index01 >> 16), uint16_t(t.index2_matId & 0xFFFFu)};
  std::sort(ids, ids + 3);
  return {t.vertexOffset, t.index2_matId >> 16, u64(ids[0]) + (u64(ids[1]) << 16) + (u64(ids[2]) << 32)};
}

}  // namespace

struct BvhBuilder::TreeNode {
  Box3 extent{};
  std::unique_ptr<TreeNode> left, right;
  std::vector<u32> tris, lights, blas;
};

// SplitData (BVHBuilder.h:188-206), triangle lists only.
struct BvhBuilder::Split {
  u32 numL = 0, numR = 0, numBoth = 0, uniqueL = 0, uniqueR = 0;
  Box3 leftBox{}, rightBox{};
  std::vector<u32> itemsL, itemsR;
};

struct BvhBuilder::Ctx {
  int countThreads = 1;
};

BvhBuilder::BvhBuilder() {
  // The constructor registers material 0, a white Lambertian (BVHBuilder.h:73).
  materials_.push_back(lambertian(1, 1, 1));
}
BvhBuilder::~BvhBuilder() = default;

MaterialRec BvhBuilder::lambertian(float r, float g, float b) {
  MaterialRec m{};
  m.type_ids[0] = kMaterialLambert;
  m.type_ids[1] = m.type_ids[2] = m.type_ids[3] = 0xFFFFFFFFu;
  m.color[0] = r;
  m.color[1] = g;
  m.color[2] = b;
  m.color[3] = 0;
  return m;
}

u32 BvhBuilder::addVertices(u32 count, const float* xyz) {
  u32 off = u32(positions_.size());
  for (u32 i = 0; i < count; ++i) positions_.push_back({xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]});
  return off;
}

void BvhBuilder::addTriangleList(u32 vertexOffset, u32 numTriangles, const u32* idx, const MaterialRec& material) {
  u32 matId = u32(materials_.size());
  materials_.push_back(material);
  for (u32 i = 0; i < numTriangles; ++i) {
    // BVHGeometry::TriangleData stores 16-bit indices (BVHGeometry.h:12-16).
    u32 i0 = idx[3 * i] & 0xFFFFu, i1 = idx[3 * i + 1] & 0xFFFFu, i2 = idx[3 * i + 2] & 0xFFFFu;
    triangles_.push_back({vertexOffset, i0 + (i1 << 16), i2 + (matId << 16)});
  }
}

void BvhBuilder::addSphereLight(const SphereLightDesc& l) { lights_.push_back(l); }

u32 BvhBuilder::addBlas(u32 vertexOffset, u32 numTriangles, const u32* idx, const MaterialRec& material) {
  auto b = std::make_unique<BvhBuilder>();
  b->sharedPositions_ = &positions_;
  b->addTriangleList(vertexOffset, numTriangles, idx, material);
  blasMaterialOffset_.push_back(u32(materials_.size()));
  materials_.insert(materials_.end(), b->materials_.begin(), b->materials_.end());
  instances_.push_back({u32(blas_.size()), {}});
  blas_.push_back(std::move(b));
  return u32(instances_.size() - 1);
}

// ---------------------------------------------------------------------------------------------
// Split classification: fillSplitData<FillItems> (BVHBuilder.cpp:952-1094).
template <bool Fill>
bool BvhBuilder::classify(const Ctx&, Split& out, const Box3& parent, const Box3& leftBox, const Box3& rightBox,
                          const std::vector<u32>& items, long bound) const {
  out.leftBox = leftBox;
  out.rightBox = rightBox;
  const float leftVolume = boxVolume(leftBox);
  const float rightVolume = boxVolume(rightBox);
  bool leftInit = false, rightInit = false;
  Box3 curLeft = leftBox, curRight = rightBox;
  const float thr = params_.expandNodeVolumeThreshold;
  const float meanSize = meanTriangleSize_;
  const long n = long(items.size());
  long k = 0;

  for (u32 id : items) {
    if (!Fill && bound >= 0 && (k & 7) == 0) {
      // each remaining item adds at most one to uniqueL or uniqueR
      const long lo = long(out.uniqueL < out.uniqueR ? out.uniqueL : out.uniqueR);
      if (lo + (n - k) <= bound) return false;
    }
    ++k;
    const Box3& aabb = aabb_[id];
    const bool tri = isTriangle(id);
    const bool inL = tri ? triangleBoxOverlap(v0_[id], v1_[id], v2_[id], leftBox) : boxesOverlap(aabb, leftBox);
    const bool inR = tri ? triangleBoxOverlap(v0_[id], v1_[id], v2_[id], rightBox) : boxesOverlap(aabb, rightBox);

    auto toRight = [&]() {
      out.numR++;
      out.uniqueR++;
      if constexpr (Fill) {
        out.itemsR.push_back(id);
        out.rightBox = intersectBox(rightInit ? mergeBox(aabb, out.rightBox) : aabb, curRight);
        rightInit = true;
      }
    };
    auto toLeft = [&]() {
      out.numL++;
      out.uniqueL++;
      if constexpr (Fill) {
        out.itemsL.push_back(id);
        out.leftBox = intersectBox(leftInit ? mergeBox(aabb, out.leftBox) : aabb, curLeft);
        leftInit = true;
      }
    };
    auto toBoth = [&]() {
      out.numBoth++;
      out.numL++;
      out.numR++;
      if constexpr (Fill) {
        out.itemsL.push_back(id);
        out.itemsR.push_back(id);
        out.leftBox = intersectBox(leftInit ? mergeBox(aabb, out.leftBox) : aabb, curLeft);
        leftInit = true;
        out.rightBox = intersectBox(rightInit ? mergeBox(aabb, out.rightBox) : aabb, curRight);
        rightInit = true;
      }
    };

    if (!inL && inR) {
      toRight();
    } else if (inL && !inR) {
      toLeft();
    } else if (!inL && !inR) {
      // dropped: the reference's "floating point precision" case (:1001-1005)
    } else {
      const float smallest = kPlant == 6 ? (rightVolume < leftVolume ? leftVolume : rightVolume)
                                         : (rightVolume < leftVolume) ? rightVolume : leftVolume;  // std::min
      Box3 newLeft = intersectBox(mergeBox(leftBox, aabb), parent);
      Box3 newRight = intersectBox(mergeBox(rightBox, aabb), parent);
      float diffLeft = boxVolume(newLeft) - leftVolume;
      float diffRight = boxVolume(newRight) - rightVolume;
      if (kPlant == 5 ? diffRight < diffLeft : diffLeft < diffRight) {
        float grow = maxDimDistance(newLeft, leftBox);
        if ((diffLeft / smallest) >= thr || grow > meanSize) {
          toBoth();
        } else {
          out.numL++;
          out.uniqueL++;
          if constexpr (Fill) {
            out.itemsL.push_back(id);
            out.leftBox = intersectBox(leftInit ? mergeBox(out.leftBox, aabb) : aabb, parent);
            curLeft = intersectBox(mergeBox(aabb, curLeft), parent);
            leftInit = true;
          }
        }
      } else {
        float grow = maxDimDistance(newRight, rightBox);
        if ((diffRight / smallest) >= thr || grow > meanSize) {
          toBoth();
        } else {
          out.numR++;
          out.uniqueR++;
          if constexpr (Fill) {
            out.itemsR.push_back(id);
            out.rightBox = intersectBox(rightInit ? mergeBox(out.rightBox, aabb) : aabb, parent);
            curRight = intersectBox(mergeBox(aabb, curRight), parent);
            rightInit = true;
          }
        }
      }
    }
  }
  return true;
}

// Count-only classification. The counts are sums of per-item decisions that do not depend on
// each other, so large item lists are split across threads with identical results.
void BvhBuilder::countSplit(const Ctx& ctx, Split& out, const Box3& parent, const Box3& leftBox,
                            const Box3& rightBox, const std::vector<u32>& items) const {
  const size_t n = items.size();
  int T = ctx.countThreads;
  if (T <= 1 || n < 200000) {
    classify<false>(ctx, out, parent, leftBox, rightBox, items);
    return;
  }
  T = int(std::min<size_t>(size_t(T), n / 50000));
  std::vector<Split> part(T);
  std::vector<std::vector<u32>> chunks(T);
  std::vector<std::thread> th;
  for (int t = 0; t < T; ++t) {
    size_t b = n * t / T, e = n * (t + 1) / T;
    chunks[t].assign(items.begin() + b, items.begin() + e);
  }
  for (int t = 1; t < T; ++t)
    th.emplace_back([&, t]() { classify<false>(ctx, part[t], parent, leftBox, rightBox, chunks[t]); });
  classify<false>(ctx, part[0], parent, leftBox, rightBox, chunks[0]);
  for (auto& x : th) x.join();
  out.leftBox = leftBox;
  out.rightBox = rightBox;
  for (auto& p : part) {
    out.numL += p.numL;
    out.numR += p.numR;
    out.numBoth += p.numBoth;
    out.uniqueL += p.uniqueL;
    out.uniqueR += p.uniqueR;
  }
}

// computeSplitScore (BVHBuilder.cpp:503-507) is min(uniqueLeft, uniqueRight) converted to float; the
// searches below compare that minimum as an integer, which orders splits identically while the counts
// stay below 2^24 (every integer up to 2^24 is a float), so nodes of more than 16 M items could break a
// tie the reference's float comparison would see differently.

// searchBestSplit (BVHBuilder.cpp:1096-1183).
bool BvhBuilder::searchSplit(const Ctx& ctx, const TreeNode* node, const std::vector<u32>& items, int axis,
                             Split& out, long need) const {
  const u32 n = u32(items.size());
  const Box3& ext = node->extent;
  const F3 boxDim = sub(ext.mx, ext.mn);

  if (n > 256) {
    float lo = 0, hi = 1;
    for (u32 i = 0; i < 32; ++i) {
      float newCut = (lo + hi) * 0.5f;
      F3 step = keepAxis(mulS(boxDim, newCut), axis);
      F3 fixedStep = dropAxis(boxDim, axis);
      Box3 leftBox = {ext.mn, add(add(ext.mn, fixedStep), step)};
      Box3 rightBox = {add(ext.mn, step), ext.mx};
      out = Split{};
      countSplit(ctx, out, ext, leftBox, rightBox, items);
      u32 absDiff = out.numL < out.numR ? out.numR - out.numL : out.numL - out.numR;
      u32 tol = n / (16 * 1024);
      if (tol < 1) tol = 1;
      if (absDiff <= tol) break;
      if ((out.numL > out.numR) != (kPlant == 4))
        hi = newCut;
      else
        lo = newCut;
    }
    return true;
  } else {
    // The reference keeps the first candidate whose score (min of the unique counts) strictly
    // beats all earlier ones. A candidate is skipped as soon as its count so far plus the items
    // left cannot beat the current best or reach `need` (below `need` the axis loses the pick
    // anyway); the kept candidate, and hence the tree, is the same.
    const F3 fixedStep = dropAxis(boxDim, axis);
    long best = -1;
    bool anyValid = false, recorded = false;
    auto processStep = [&](F3 step) {
      Box3 leftBox = {ext.mn, add(add(ext.mn, fixedStep), step)};
      Box3 rightBox = {add(ext.mn, step), ext.mx};
      if (!(checkBox(leftBox) && checkBox(rightBox))) return;
      anyValid = true;
      const long kept = (kPlant == 3 && best >= 0) ? best - 1 : best;  // plant 3: equal scores replace
      const long bound = kept > need - 1 ? kept : need - 1;
      Split s;
      if (!classify<false>(ctx, s, ext, leftBox, rightBox, items, bound)) return;
      const long cur = long(s.uniqueL < s.uniqueR ? s.uniqueL : s.uniqueR);
      if (cur > bound) {
        best = cur;
        out = std::move(s);
        recorded = true;
      }
    };
    processStep(keepAxis(mulS(boxDim, 0.5f), axis));
    const float delta = float(1e-5);
    // the reference walks the BLAS instances' bounds first, then the triangles' (:1160-1181)
    for (int pass = 0; pass < 2; ++pass)
      for (u32 id : items) {
        if (isTriangle(id) != (pass == 1)) continue;
        const Box3& b = aabb_[id];
        F3 lo3 = {b.mn.x - delta, b.mn.y - delta, b.mn.z - delta};
        F3 hi3 = {b.mx.x + delta, b.mx.y + delta, b.mx.z + delta};
        processStep(keepAxis(sub(lo3, ext.mn), axis));
        processStep(keepAxis(sub(hi3, ext.mn), axis));
      }
    // no valid candidate: `out` stays default-constructed, as in the reference
    return recorded || !anyValid;
  }
}

// fillLeafData (BVHBuilder.cpp:734-760); items == nullptr gives the light-only fill used for inner
// nodes (:846).
void BvhBuilder::fillLeaf(TreeNode* node, const std::vector<u32>* items) {
  if (items) {
    const size_t cap = size_t(1) << kTriangleBits;
    for (u32 id : *items) {
      if (!isTriangle(id)) continue;
      if (node->tris.size() >= cap) break;
      node->tris.push_back(id);
    }
    std::vector<std::pair<LeafKey, u32>> kv;
    kv.reserve(node->tris.size());
    for (u32 id : node->tris) kv.push_back({leafKey(triangles_[id]), id});
    std::sort(kv.begin(), kv.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    auto last = std::unique(kv.begin(), kv.end(), [](const auto& a, const auto& b) { return a.first == b.first; });
    kv.erase(last, kv.end());
    node->tris.clear();
    for (auto& p : kv) node->tris.push_back(p.second);
    const size_t blasCap = size_t(1) << kBlasBits;
    for (u32 id : *items)
      if (!isTriangle(id) && node->blas.size() < blasCap) node->blas.push_back(id - u32(triangles_.size()));
  }
  const size_t lightCap = size_t(1) << kLightBits;
  for (u32 i = 0; i < lights_.size() && node->lights.size() < lightCap; ++i) {
    if (sphereTouchesBox(lights_[i].pos, lights_[i].radius, node->extent)) node->lights.push_back(i);
  }
}

// addObjectsRec (BVHBuilder.cpp:762-942), triangle lists only.
void BvhBuilder::buildRec(Ctx& ctx, u32 depth, u32 numUnique, std::vector<u32>& items, TreeNode* node,
                          bool parallel) {
  node->tris.clear();
  node->lights.clear();
  node->blas.clear();
  const u32 n = u32(items.size());

  if (numUnique <= params_.minObjPerNode || depth >= params_.maxDepth) {
    fillLeaf(node, &items);
    return;
  }

  Split best;
  if (n < 1024) {
    // getBestSplit(getBestSplit(0, 1), 2) with getBestSplit(a, b) = score(a) > score(b) ? a : b
    // (BVHBuilder.cpp:795-802): a later axis wins ties, so it only matters if it reaches the
    // current winner's score, which is what `need` passes down.
    Split s[3];
    auto score = [&](int a) { return long(s[a].uniqueL < s[a].uniqueR ? s[a].uniqueL : s[a].uniqueR); };
    searchSplit(ctx, node, items, 0, s[0], 0);
    int cur = 0;
    for (int a = 1; a < 3; ++a)
      if (searchSplit(ctx, node, items, a, s[a], score(cur)) &&
          (kPlant == 1 ? score(a) > score(cur) : !(score(cur) > score(a))))
        cur = a;
    best = std::move(s[cur]);
  } else {
    searchSplit(ctx, node, items, int(depth % 3), best, 0);
  }

  if (depth > 0) {
    const bool noGainL = n - best.numL < params_.minObjGain, noGainR = n - best.numR < params_.minObjGain;
    if ((kPlant == 2 ? (noGainL || noGainR) : (noGainL && noGainR)) || best.uniqueL == 0 || best.uniqueR == 0) {
      fillLeaf(node, &items);
      return;
    }
  }

  fillLeaf(node, nullptr);

  Split sd;
  classify<true>(ctx, sd, node->extent, best.leftBox, best.rightBox, items);
  node->left = std::make_unique<TreeNode>();
  node->right = std::make_unique<TreeNode>();
  node->left->extent = adjustBox(sd.leftBox);
  node->right->extent = adjustBox(sd.rightBox);
  std::vector<u32>().swap(items);

  TreeNode* L = node->left.get();
  TreeNode* R = node->right.get();
  if (parallel && n > 10000) {
    std::thread th([&]() { buildRec(ctx, depth + 1, sd.uniqueL, sd.itemsL, L, parallel); });
    buildRec(ctx, depth + 1, sd.uniqueR, sd.itemsR, R, parallel);
    th.join();
  } else {
    buildRec(ctx, depth + 1, sd.uniqueL, sd.itemsL, L, false);
    buildRec(ctx, depth + 1, sd.uniqueR, sd.itemsR, R, false);
  }
  if (isTlas()) node->blas = mergeCommon(L->blas, R->blas);  // BVHBuilder.cpp:905-940
}

// Node numbering of a single-thread build: the root is 0 and every split allocates its two
// children as an adjacent pair before descending left, then right (BVHBuilder.cpp:851-868).
void BvhBuilder::flatten(std::unique_ptr<TreeNode> root) {
  nodes_.clear();
  nodes_.emplace_back();
  struct Frame { TreeNode* t; int id; u32 depth; };
  std::vector<Frame> stack{{root.get(), 0, 0}};
  stats_ = BuildStats{};
  std::unordered_set<u32> seen;
  while (!stack.empty()) {
    Frame f = stack.back();
    stack.pop_back();
    FlatNode& fn = nodes_[f.id];
    fn.extent = f.t->extent;
    fn.tris = std::move(f.t->tris);
    fn.lights = std::move(f.t->lights);
    fn.blas = std::move(f.t->blas);
    if (f.t->left) {
      int l = int(nodes_.size());
      nodes_.emplace_back();
      nodes_.emplace_back();
      FlatNode& p = nodes_[f.id];  // re-fetch after growth
      p.left = l;
      p.right = l + 1;
      nodes_[l].parent = f.id;
      nodes_[l + 1].parent = f.id;
      nodes_[l].sibling = l + 1;
      nodes_[l + 1].sibling = l;
      // pairs are numbered depth-first, left subtree completely before the right one
      stack.push_back({f.t->right.get(), l + 1, f.depth + 1});
      stack.push_back({f.t->left.get(), l, f.depth + 1});
    } else {
      stats_.numLeaves++;
      stats_.maxDepth = std::max(stats_.maxDepth, f.depth);
      stats_.maxLeafTriangles = std::max(stats_.maxLeafTriangles, u32(fn.tris.size()));
      stats_.leafTriangleSlots += fn.tris.size();
      for (u32 t : fn.tris) seen.insert(t);
    }
  }
  stats_.numNodes = u32(nodes_.size());
  stats_.numDuplicated = u32(stats_.leafTriangleSlots - seen.size());
}

// BVHBuilder::build (BVHBuilder.cpp:408-461).
void BvhBuilder::build(int threads) {
  auto t0 = std::chrono::steady_clock::now();
  // buildBlas (BVHBuilder.cpp:383-399): every BLAS with the BLAS parameters. The BLASes are
  // independent, so with several threads each thread builds whole BLASes (single-threaded each; the
  // node numbering never depends on the thread count).
  for (auto& b : blas_) b->setParameters(blasParams_);
  if (threads > 1 && blas_.size() > 1) {
    std::atomic<size_t> next{0};
    std::exception_ptr failure;
    std::mutex failMu;
    auto work = [&] {
      for (size_t k; (k = next.fetch_add(1)) < blas_.size();) {
        try {
          blas_[k]->build(1);
        } catch (...) {
          std::lock_guard<std::mutex> g(failMu);
          if (!failure) failure = std::current_exception();
        }
      }
    };
    std::vector<std::thread> pool;
    const size_t n = std::min(blas_.size(), size_t(threads));
    for (size_t t = 1; t < n; ++t) pool.emplace_back(work);
    work();
    for (auto& t : pool) t.join();
    if (failure) std::rethrow_exception(failure);
  } else {
    for (auto& b : blas_) b->build(threads);
  }
  const size_t nT = triangles_.size(), nB = instances_.size();
  v0_.resize(nT);
  v1_.resize(nT);
  v2_.resize(nT);
  aabb_.resize(nT + nB);
  const float big = std::numeric_limits<float>::max();
  Box3 tight = {{big, big, big}, {-big, -big, -big}};
  meanTriangleSize_ = 0;
  for (size_t i = 0; i < nT; ++i) {
    const TriangleRec& t = triangles_[i];
    F3 p0 = pos().at(t.vertexOffset + (t.index01 & 0xFFFFu));
    F3 p1 = pos().at(t.vertexOffset + (t.index01 >> 16));
    F3 p2 = pos().at(t.vertexOffset + (t.index2_matId & 0xFFFFu));
    v0_[i] = p0;
    v1_[i] = p1;
    v2_[i] = p2;
    Box3 b;  // getAABB(Triangle) (:148-158)
    b.mn = {firstMin3(p0.x, p1.x, p2.x), firstMin3(p0.y, p1.y, p2.y), firstMin3(p0.z, p1.z, p2.z)};
    b.mx = {firstMax3(p0.x, p1.x, p2.x), firstMax3(p0.y, p1.y, p2.y), firstMax3(p0.z, p1.z, p2.z)};
    aabb_[i] = b;
    tight.mn = lmin3(tight.mn, b.mn);
    tight.mx = lmax3(tight.mx, b.mx);
    meanTriangleSize_ += lsum(sub(b.mx, b.mn)) / 3;
  }
  for (size_t i = 0; i < nB; ++i) {  // BLAS instances after the triangles (:433-438)
    instances_[i].aabb = blas_[instances_[i].blasId]->sceneBox_;
    aabb_[nT + i] = instances_[i].aabb;
    tight.mn = lmin3(tight.mn, instances_[i].aabb.mn);
    tight.mx = lmax3(tight.mx, instances_[i].aabb.mx);
  }
  sceneBox_ = adjustBox(tight);
  // over the builder's own triangles only: a top level without triangles divides 0 by 0 (:442), and
  // the resulting NaN makes the growth-distance test of the split never fire
  meanTriangleSize_ /= float(nT);
  meanTriangleSize_ *= params_.expandNodeDimensionFactor;

  std::vector<u32> items(nT + nB);
  for (size_t i = 0; i < nT + nB; ++i) items[i] = u32(i);

  auto root = std::make_unique<TreeNode>();
  root->extent = sceneBox_;
  Ctx ctx;
  ctx.countThreads = threads > 1 ? threads : 1;
  buildRec(ctx, 0, u32(nT + nB), items, root.get(), threads > 1);
  flatten(std::move(root));
  if (nodes_.size() >= kNidMask) throw std::runtime_error("node count exceeds NID_MASK");
  stats_.buildSeconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// getBvhGpuSize (BVHBuilder.cpp:1185-1222) for a scene of triangles, sphere lights and BLAS instances:
// the node records of the top level and of every BLAS are each aligned on their own here, while the
// blob stores them as one array, so the size can exceed what is written (the rest stays zero).
size_t BvhBuilder::gpuSize() const {
  u32 tris = u32(triangles_.size());
  for (const auto& b : blas_) tris += b->numTriangles();
  u32 size = alignUp(u32(materials_.size() * sizeof(MaterialRec)), kBufferAlignment);
  size += alignUp(u32(tris * sizeof(TriangleRec)), kBufferAlignment);
  size += alignUp(0, kBufferAlignment);  // primitives
  size += alignUp(u32(lights_.size() * sizeof(PackedLightRec)), kBufferAlignment);
  size += alignUp(u32(instances_.size() * 32u), kBufferAlignment);  // BlasHeader
  auto nodesSize = [](const std::vector<FlatNode>& nodes) {
    u32 sz = alignUp(u32(nodes.size() * sizeof(PackedNodeRec)), kBufferAlignment);
    for (const auto& n : nodes) {
      sz += u32(1 + n.lights.size() + n.blas.size()) * 4u;
      sz += u32(n.tris.size()) * u32(sizeof(TriangleRec));
    }
    return sz;
  };
  size += nodesSize(nodes_);
  for (const auto& b : blas_) size += nodesSize(b->nodes_);
  return size;
}

// fillGpuBuffer (BVHBuilder.cpp:1319-1534). Bytes the reference leaves uninitialised (section
// padding, unused light parameters, BlasHeader.matId) are zero here. In a top-level blob the nodes of
// every BLAS follow the top level's, renumbered by their offset, and their leaves inline the BLAS's
// triangle records with the BLAS's material offset, as they stand in the triangle section (the
// reference copies m_triangles[t] of the top-level builder there, whose index range does not hold
// them: BVHBuilder.cpp:1490-1495).
void BvhBuilder::fillGpuBuffer(void* outPtr, BlobRanges* r) const {
  const size_t total = gpuSize();
  uint8_t* base = static_cast<uint8_t*>(outPtr);
  std::memset(base, 0, total);
  u32 off = 0;
  auto clamp32 = [](u32 s) { return s < kBufferAlignment ? kBufferAlignment : s; };

  const u32 headerBytes = u32(instances_.size() * 32u);
  r->blas = {off, clamp32(headerBytes)};
  const u32 headerOff = off;
  off += alignUp(headerBytes, kBufferAlignment);

  std::vector<u32> blasTriOffset(blas_.size());
  u32 numTris = u32(triangles_.size());
  for (size_t i = 0; i < blas_.size(); ++i) {
    blasTriOffset[i] = numTris;
    numTris += blas_[i]->numTriangles();
  }
  auto blasTriangle = [&](size_t i, u32 t) {  // fillTriangles(out, matIdOffset) (:1536-1546)
    TriangleRec rec = blas_[i]->triangles_[t];
    rec.index2_matId += blasMaterialOffset_[i] << 16;
    return rec;
  };
  const u32 triBytes = u32(numTris * sizeof(TriangleRec));
  r->triangles = {off, clamp32(triBytes)};
  if (!triangles_.empty()) std::memcpy(base + off, triangles_.data(), triangles_.size() * sizeof(TriangleRec));
  for (size_t i = 0; i < blas_.size(); ++i)
    for (u32 t = 0; t < blas_[i]->numTriangles(); ++t) {
      const TriangleRec rec = blasTriangle(i, t);
      std::memcpy(base + off + size_t(blasTriOffset[i] + t) * sizeof(TriangleRec), &rec, sizeof(rec));
    }
  off += alignUp(triBytes, kBufferAlignment);

  r->primitives = {off, clamp32(0)};
  off += alignUp(0, kBufferAlignment);

  u32 matBytes = u32(materials_.size() * sizeof(MaterialRec));
  r->materials = {off, clamp32(matBytes)};
  std::memcpy(base + off, materials_.data(), matBytes);
  off += alignUp(matBytes, kBufferAlignment);

  u32 lightBytes = u32(lights_.size() * sizeof(PackedLightRec));
  r->lights = {off, clamp32(lightBytes)};
  for (size_t i = 0; i < lights_.size(); ++i) {  // packSphereLight (:1246-1258)
    PackedLightRec pl{};
    const SphereLightDesc& l = lights_[i];
    pl.iparam = kLightSphere;
    pl.fparam[0] = l.pos.x;
    pl.fparam[1] = l.pos.y;
    pl.fparam[2] = l.pos.z;
    pl.fparam[3] = l.radius;
    pl.fparam[4] = l.color.x;
    pl.fparam[5] = l.color.y;
    pl.fparam[6] = l.color.z;
    pl.fparam[7] = l.sphereRadius;
    std::memcpy(base + off + i * sizeof(PackedLightRec), &pl, sizeof(pl));
  }
  off += alignUp(lightBytes, kBufferAlignment);

  // every node of the top level, then of each BLAS (merged and renumbered, :1441-1457)
  struct Src { const FlatNode* n; u32 shift; int blas; };
  std::vector<Src> all;
  for (const FlatNode& n : nodes_) all.push_back({&n, 0u, -1});
  std::vector<u32> blasRoot(blas_.size());
  for (size_t i = 0; i < blas_.size(); ++i) {
    blasRoot[i] = u32(all.size());
    for (const FlatNode& n : blas_[i]->nodes_) all.push_back({&n, blasRoot[i], int(i)});
  }
  u32 nodeBytes = u32(all.size() * sizeof(PackedNodeRec));
  r->nodes = {off, clamp32(nodeBytes)};
  uint8_t* nodeOut = base + off;
  u32 leafBase = off + alignUp(nodeBytes, kBufferAlignment);
  u32* leafBegin = reinterpret_cast<u32*>(base + leafBase);
  u32* cur = leafBegin;

  for (size_t ni = 0; ni < all.size(); ++ni) {
    const FlatNode& n = *all[ni].n;
    const std::vector<FlatNode>& tree = all[ni].blas < 0 ? nodes_ : blas_[all[ni].blas]->nodes_;
    const u32 shift = all[ni].shift;
    auto isLeaf = [&](int i) { return i >= 0 && tree[i].left < 0; };
    u32 leafIndex = u32(cur - leafBegin);
    u32 packed = u32(n.tris.size()) + (u32(n.blas.size()) << kTriangleBits) + (0u << (kTriangleBits + kBlasBits)) +
                 (u32(n.lights.size()) << (kTriangleBits + kBlasBits + kPrimitiveBits));
    *cur++ = packed;
    for (u32 t : n.tris) {
      const TriangleRec rec = all[ni].blas < 0 ? triangles_[t] : blasTriangle(size_t(all[ni].blas), t);
      std::memcpy(cur, &rec, sizeof(TriangleRec));
      cur += 3;
    }
    for (u32 b : n.blas) *cur++ = b;
    for (u32 l : n.lights) *cur++ = l;

    // packNodeData (:1283-1317)
    PackedNodeRec rec{};
    u32 left = n.left >= 0 ? u32(n.left) + shift : kNidMask;
    u32 parent = n.parent >= 0 ? u32(n.parent) + shift : kNidMask;
    u32 sibling = n.sibling >= 0 ? u32(n.sibling) + shift : kNidMask;
    if (isLeaf(n.sibling)) sibling |= kNidLeafBit;
    u32 child = left | (isLeaf(n.left) ? kNidLeftLeafBit : 0u) | (isLeaf(n.right) ? kNidRightLeafBit : 0u);
    rec.nid[0] = parent;
    rec.nid[1] = sibling;
    rec.nid[2] = child;
    rec.nid[3] = packed != 0 ? leafIndex : 0xFFFFFFFFu;
    if (n.left >= 0) {
      const Box3& b0 = tree[n.left].extent;
      const Box3& b1 = tree[n.right].extent;
      float n0[4] = {b0.mn.x, b0.mn.y, b0.mx.x, b0.mx.y};
      float n1[4] = {b1.mn.x, b1.mn.y, b1.mx.x, b1.mx.y};
      float nz[4] = {b0.mn.z, b0.mx.z, b1.mn.z, b1.mx.z};
      std::memcpy(rec.n0xy, n0, 16);
      std::memcpy(rec.n1xy, n1, 16);
      std::memcpy(rec.nz, nz, 16);
    }
    std::memcpy(nodeOut + ni * sizeof(PackedNodeRec), &rec, sizeof(rec));
  }
  r->leafData = {leafBase, u32(reinterpret_cast<uint8_t*>(cur) - reinterpret_cast<uint8_t*>(leafBegin))};

  // BlasHeader {minExtent, matId, maxExtent, rootIndex} per instance (:1521-1533)
  for (size_t i = 0; i < instances_.size(); ++i) {
    const Box3& b = instances_[i].aabb;
    const float mn[3] = {b.mn.x, b.mn.y, b.mn.z}, mx[3] = {b.mx.x, b.mx.y, b.mx.z};
    const u32 root = blasRoot[instances_[i].blasId];
    uint8_t* h = base + headerOff + i * 32u;
    std::memcpy(h, mn, 12);
    std::memcpy(h + 16, mx, 12);
    std::memcpy(h + 28, &root, 4);
  }
}

std::vector<u32> BvhBuilder::leafTriangleIds() const {
  std::vector<u32> out;
  for (const auto& n : nodes_)
    for (u32 t : n.tris) out.push_back(t);
  return out;
}

}  // namespace rtq
