// Wire formats shared by the BVH builder, the device upload path and the scalar oracle.
//
// Every struct here is byte-for-byte the std430 layout the reference's shaders read from the
// flattened BVH blob (Shaders/core/primitive_cpp.glsl:50-117, Shaders/struct_cpp.glsl:52-55).
#pragma once
#include <cstdint>

namespace rtq {

using u32 = uint32_t;
using u64 = uint64_t;

struct F3 { float x, y, z; };
struct Box3 { F3 mn, mx; };

// Triangle {vertexOffset, index01, index2_matId}: primitive_cpp.glsl:50-55, packed by
// BVHBuilder::addTriangle (BVHBuilder.cpp:247-255).
struct TriangleRec { u32 vertexOffset, index01, index2_matId; };
static_assert(sizeof(TriangleRec) == 12, "Triangle is 12 B");

// Material {uvec4 type_ids; vec4 color; vec4 params}: primitive_cpp.glsl:67-72.
struct MaterialRec { u32 type_ids[4]; float color[4]; float params[4]; };
static_assert(sizeof(MaterialRec) == 48, "Material is 48 B");

// PackedLight {uint iparam; float fparam[13]}: primitive_cpp.glsl:113-117.
struct PackedLightRec { u32 iparam; float fparam[13]; };
static_assert(sizeof(PackedLightRec) == 56, "PackedLight is 56 B");

// PackedBVHNode {uvec4 nid; vec4 n0xy; vec4 n1xy; vec4 nz}: primitive_cpp.glsl:80-86,
// written by BVHBuilder::packNodeData (BVHBuilder.cpp:1283-1317).
struct PackedNodeRec { u32 nid[4]; float n0xy[4]; float n1xy[4]; float nz[4]; };
static_assert(sizeof(PackedNodeRec) == 64, "PackedBVHNode is 64 B");

// SphereLight {pos, radius, color, sphereRadius}: primitive_cpp.glsl:96-102.
struct SphereLightDesc { F3 pos; float radius; F3 color; float sphereRadius; };

// Node-id bits: struct_cpp.glsl:52-55.
constexpr u32 kNidLeafBit = 0x80000000u;
constexpr u32 kNidMask = 0x3FFFFFFFu;
constexpr u32 kNidLeftLeafBit = 0x80000000u;
constexpr u32 kNidRightLeafBit = 0x40000000u;

// Leaf count word bit widths: primitive_cpp.glsl:18-21; packing order BVHBuilder.cpp:1483-1486.
constexpr u32 kTriangleBits = 12, kBlasBits = 8, kPrimitiveBits = 8, kLightBits = 4;

// Scalar constants, all fp32 in the shaders.
constexpr float kTMax = 100.0f;               // struct_cpp.glsl:5
constexpr float kOffsetRayCollision = 10e-6f; // system.glsl:45
constexpr float kTriEpsilon = 10e-8f;         // collision.glsl:62

constexpr u32 kLightSphere = 1;      // primitive_cpp.glsl:8
constexpr u32 kMaterialLambert = 2;  // primitive_cpp.glsl:12

}  // namespace rtq
