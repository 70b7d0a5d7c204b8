// CPU BVH builder: a restatement of the reference's top-down binary builder with clipped,
// duplicating splits (Renderer/BVHBuilder.cpp) and of its flattener (fillGpuBuffer), so that the
// device path consumes exactly the blob the reference emits.
//
// Scope: triangle lists, sphere lights and bottom-level BVH instances (a builder given BLASes is a
// top-level builder, as BVHBuilder(name, geometry, isTlas = true) is). Analytic primitives and
// triangle strips are compiled out of the reference's traced path (bvhCollision.glsl:219-238,
// INLINE_STRIPS 0 at BVHBuilder.cpp:14) and are not built here; their blob sections are emitted empty,
// exactly as the reference emits them for a scene without them.
#pragma once
#include <cstddef>
#include <memory>
#include <vector>

#include "bvh_types.h"

namespace rtq {

// BVHBuildParameters (BVHBuilder.h:53-61). Defaults are the values the reference application
// builds every scene with (main.cpp:197-201), not the struct's own initialisers.
struct BuildParams {
  u32 maxDepth = 25;
  u32 minObjPerNode = 4;
  u32 minObjGain = 4;
  float expandNodeVolumeThreshold = 0.25f;
  float expandNodeDimensionFactor = 0.5f;
};

struct Range { u32 offset, size; };

// The 7 sub-ranges fillGpuBuffer returns, in its argument order (BVHBuilder.h:101).
struct BlobRanges { Range triangles, primitives, materials, lights, nodes, leafData, blas; };

struct BuildStats {
  u32 numNodes = 0, numLeaves = 0, maxLeafTriangles = 0, maxDepth = 0;
  u64 leafTriangleSlots = 0;  // sum of leaf triangle counts (duplicates included)
  u32 numDuplicated = 0;      // leafTriangleSlots - distinct triangles referenced
  double buildSeconds = 0;
};

class BvhBuilder {
 public:
  BvhBuilder();
  ~BvhBuilder();

  // BVHGeometry::addTriangleList (BVHGeometry.cpp:46-62): appends positions and returns the vertex
  // offset of the first one.
  u32 addVertices(u32 count, const float* xyz);
  // BVHBuilder::addTriangleList (BVHBuilder.cpp:265-280): one material per call; indices are
  // relative to vertexOffset and must fit in 16 bits.
  void addTriangleList(u32 vertexOffset, u32 numTriangles, const u32* indices, const MaterialRec& material);
  // BVHBuilder::addSphereLight (BVHBuilder.cpp:282-285).
  void addSphereLight(const SphereLightDesc& light);

  void setParameters(const BuildParams& p) { params_ = p; }
  const BuildParams& parameters() const { return params_; }

  // BVHBuilder::addBlas (BVHBuilder.cpp:288-301) for a bottom-level BVH of one triangle list over this
  // builder's vertices (one BLAS per loaded shape, as Scene::loadBlas makes them, Scene.cpp:229-263):
  // the BLAS's materials (its default material and the list's) are appended to this builder's and the
  // instance is the next BLAS index. Returns that index.
  u32 addBlas(u32 vertexOffset, u32 numTriangles, const u32* indices, const MaterialRec& material);
  // The parameters buildBlas builds every BLAS with (BVHBuilder.cpp:383-399); setParameters sets the
  // top level's, as BVHData::build does (BVHData.cpp:31-35).
  void setBlasParameters(const BuildParams& p) { blasParams_ = p; }
  bool isTlas() const { return !blas_.empty(); }
  u32 numBlas() const { return u32(blas_.size()); }
  const BvhBuilder& blas(u32 i) const { return *blas_[i]; }
  const BuildParams& blasParameters() const { return blasParams_; }

  // BVHBuilder::build (BVHBuilder.cpp:408-461), after buildBlas for a top-level builder. threads > 1
  // builds large subtrees concurrently (the reference forks above 10 000 items, :879-903); node
  // numbering is always the single-thread order, so the blob does not depend on threads.
  void build(int threads);

  // getBvhGpuSize (BVHBuilder.cpp:1185-1222) and fillGpuBuffer (:1319-1534).
  size_t gpuSize() const;
  void fillGpuBuffer(void* out, BlobRanges* ranges) const;

  // Default materials (BVHBuilder.cpp:193-200).
  static MaterialRec lambertian(float r, float g, float b);

  const std::vector<F3>& positions() const { return positions_; }
  const std::vector<TriangleRec>& triangles() const { return triangles_; }
  const std::vector<MaterialRec>& materials() const { return materials_; }
  const std::vector<SphereLightDesc>& lights() const { return lights_; }
  u32 numTriangles() const { return u32(triangles_.size()); }
  u32 numLights() const { return u32(lights_.size()); }
  u32 numMaterials() const { return u32(materials_.size()); }
  u32 numNodes() const { return u32(nodes_.size()); }
  const BuildStats& stats() const { return stats_; }
  Box3 sceneBox() const { return sceneBox_; }

  // Per-leaf-slot triangle insertion index, in blob leaf order (one entry per inline triangle).
  std::vector<u32> leafTriangleIds() const;

  struct FlatNode {
    Box3 extent;
    int parent = -1, sibling = -1, left = -1, right = -1;
    std::vector<u32> tris, lights, blas;  // blas: BLAS instance indices (top-level builders)
  };
  const std::vector<FlatNode>& nodes() const { return nodes_; }

 private:
  struct TreeNode;
  struct Split;
  struct Ctx;
  void buildRec(Ctx& ctx, u32 depth, u32 numUnique, std::vector<u32>& items, TreeNode* node, bool parallel);
  void fillLeaf(TreeNode* node, const std::vector<u32>* items);
  // Returns false when every valid candidate of this axis scores below `need` (the axis then cannot
  // win the cross-axis pick and `out` is meaningless); otherwise `out` is the reference's choice.
  bool searchSplit(const Ctx& ctx, const TreeNode* node, const std::vector<u32>& items, int axis, Split& out,
                   long need) const;
  // Fill = false with bound >= 0: stops (returns false) as soon as min(uniqueL, uniqueR) can no
  // longer exceed `bound` whatever the remaining items do.
  template <bool Fill>
  bool classify(const Ctx& ctx, Split& out, const Box3& parent, const Box3& left, const Box3& right,
                const std::vector<u32>& items, long bound = -1) const;
  void countSplit(const Ctx& ctx, Split& out, const Box3& parent, const Box3& left, const Box3& right,
                  const std::vector<u32>& items) const;
  void flatten(std::unique_ptr<TreeNode> root);

  // Items are triangles [0, numTriangles) and, in a top-level builder, BLAS instances after them.
  bool isTriangle(u32 id) const { return id < u32(triangles_.size()); }
  const std::vector<F3>& pos() const { return sharedPositions_ ? *sharedPositions_ : positions_; }

  struct BlasInstance {
    u32 blasId;
    Box3 aabb;  // the BLAS's adjusted scene box (getAABB)
  };

  BuildParams params_;
  BuildParams blasParams_;
  std::vector<std::unique_ptr<BvhBuilder>> blas_;
  std::vector<u32> blasMaterialOffset_;  // m_blasMaterialIdOffset
  std::vector<BlasInstance> instances_;
  const std::vector<F3>* sharedPositions_ = nullptr;  // a BLAS reads its top-level builder's vertices
  std::vector<F3> positions_;
  std::vector<TriangleRec> triangles_;
  std::vector<MaterialRec> materials_;
  std::vector<SphereLightDesc> lights_;
  std::vector<FlatNode> nodes_;
  Box3 sceneBox_{};
  float meanTriangleSize_ = 0;
  BuildStats stats_;

  // per-triangle caches (positions gathered once; identical values to the reference's gathers)
  std::vector<F3> v0_, v1_, v2_;
  std::vector<Box3> aabb_;
};

}  // namespace rtq
