// On-disk cache of a built scene: the fillGpuBuffer blob, its seven ranges and the vertex positions,
// keyed by a hash of everything the build depends on (positions, triangle lists, materials, lights,
// build parameters). Large batched or served scene sets load it instead of rebuilding.
//
// File layout (little endian): CacheHeader, positions (12 B each), blob. The header's checksum covers
// positions and blob; a file whose magic, version, sizes or checksum disagree is rejected.
#pragma once
#include <string>
#include <vector>

#include "bvh_builder.h"

namespace rtq {

struct CacheHeader {
  char magic[8];        // "RTQBVH\0\1"
  u32 version;          // kCacheVersion
  u32 headerBytes;      // sizeof(CacheHeader)
  u64 key;              // inputHash() of the builder that produced the blob
  u64 checksum;         // hashBytes over positions then blob
  u64 blobBytes;
  u32 numVertices;
  u32 ranges[14];       // offset, size of the seven ranges in BlobRanges order
  u32 pad;
};
static_assert(sizeof(CacheHeader) == 104, "cache header is 104 B");

constexpr u32 kCacheVersion = 1;

struct CachedScene {
  u64 key = 0;
  std::vector<F3> positions;
  std::vector<unsigned char> blob;
  BlobRanges ranges{};
};

u64 hashBytes(const void* data, size_t bytes, u64 h);
// Hash of the builder's inputs and parameters (not of its output).
u64 inputHash(const BvhBuilder& b);
// Writes to `path` through a temporary file renamed into place. Throws std::runtime_error on I/O errors.
void saveCache(const std::string& path, const BvhBuilder& b);
// 0: loaded; 1: file missing or unreadable; 2: not a cache file of this version or corrupt;
// 3: a valid cache for another key (expectedKey != 0 and different).
int loadCache(const std::string& path, u64 expectedKey, CachedScene& out);

}  // namespace rtq
