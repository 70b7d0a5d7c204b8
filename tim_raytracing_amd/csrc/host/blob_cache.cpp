// On-disk cache of built scenes (see blob_cache.h).
#include "blob_cache.h"

#include <cstdio>
#include <cstring>
#include <stdexcept>

namespace rtq {
namespace {

constexpr char kMagic[8] = {'R', 'T', 'Q', 'B', 'V', 'H', 0, 1};

u64 mix(u64 h, u64 w) {
  h ^= w;
  h *= 0x9E3779B97F4A7C15ull;
  return h ^ (h >> 29);
}

struct File {
  FILE* f = nullptr;
  ~File() {
    if (f) std::fclose(f);
  }
};

}  // namespace

u64 hashBytes(const void* data, size_t bytes, u64 h) {
  const unsigned char* p = static_cast<const unsigned char*>(data);
  size_t i = 0;
  for (; i + 8 <= bytes; i += 8) {
    u64 w;
    std::memcpy(&w, p + i, 8);
    h = mix(h, w);
  }
  u64 tail = 0;
  if (bytes > i) std::memcpy(&tail, p + i, bytes - i);
  return mix(mix(h, tail), bytes);
}

u64 inputHash(const BvhBuilder& b) {
  u64 h = 0x5254514256480001ull;
  const BuildParams& p = b.parameters();
  const u32 pw[3] = {p.maxDepth, p.minObjPerNode, p.minObjGain};
  const float pf[2] = {p.expandNodeVolumeThreshold, p.expandNodeDimensionFactor};
  h = hashBytes(pw, sizeof(pw), h);
  h = hashBytes(pf, sizeof(pf), h);
  h = hashBytes(b.positions().data(), b.positions().size() * sizeof(F3), h);
  h = hashBytes(b.triangles().data(), b.triangles().size() * sizeof(TriangleRec), h);
  h = hashBytes(b.materials().data(), b.materials().size() * sizeof(MaterialRec), h);
  h = hashBytes(b.lights().data(), b.lights().size() * sizeof(SphereLightDesc), h);
  if (b.numBlas()) {  // two-level: the BLAS parameters and every BLAS's triangles
    const BuildParams& q = b.blasParameters();
    const u32 qw[4] = {q.maxDepth, q.minObjPerNode, q.minObjGain, b.numBlas()};
    const float qf[2] = {q.expandNodeVolumeThreshold, q.expandNodeDimensionFactor};
    h = hashBytes(qw, sizeof(qw), h);
    h = hashBytes(qf, sizeof(qf), h);
    for (u32 i = 0; i < b.numBlas(); ++i)
      h = hashBytes(b.blas(i).triangles().data(), b.blas(i).triangles().size() * sizeof(TriangleRec), h);
  }
  return h;
}

void saveCache(const std::string& path, const BvhBuilder& b) {
  CacheHeader hd{};
  std::memcpy(hd.magic, kMagic, 8);
  hd.version = kCacheVersion;
  hd.headerBytes = sizeof(CacheHeader);
  hd.key = inputHash(b);
  std::vector<unsigned char> blob(b.gpuSize());
  BlobRanges r{};
  b.fillGpuBuffer(blob.data(), &r);
  const Range* rr = &r.triangles;
  for (int k = 0; k < 7; ++k) {
    hd.ranges[2 * k] = rr[k].offset;
    hd.ranges[2 * k + 1] = rr[k].size;
  }
  hd.blobBytes = blob.size();
  hd.numVertices = u32(b.positions().size());
  hd.checksum = hashBytes(blob.data(), blob.size(),
                          hashBytes(b.positions().data(), b.positions().size() * sizeof(F3), 0));
  const std::string tmp = path + ".tmp";
  {
    File f;
    f.f = std::fopen(tmp.c_str(), "wb");
    if (!f.f) throw std::runtime_error("cannot create " + tmp);
    bool ok = std::fwrite(&hd, sizeof(hd), 1, f.f) == 1;
    if (ok && hd.numVertices)
      ok = std::fwrite(b.positions().data(), sizeof(F3), hd.numVertices, f.f) == hd.numVertices;
    if (ok && !blob.empty()) ok = std::fwrite(blob.data(), 1, blob.size(), f.f) == blob.size();
    ok = (std::fflush(f.f) == 0) && ok;
    if (!ok) {
      std::remove(tmp.c_str());
      throw std::runtime_error("write failed: " + tmp);
    }
  }
  if (std::rename(tmp.c_str(), path.c_str()) != 0) {
    std::remove(tmp.c_str());
    throw std::runtime_error("cannot rename " + tmp + " to " + path);
  }
}

int loadCache(const std::string& path, u64 expectedKey, CachedScene& out) {
  File f;
  f.f = std::fopen(path.c_str(), "rb");
  if (!f.f) return 1;
  CacheHeader hd{};
  if (std::fread(&hd, sizeof(hd), 1, f.f) != 1) return 2;
  if (std::memcmp(hd.magic, kMagic, 8) != 0 || hd.version != kCacheVersion || hd.headerBytes != sizeof(hd)) return 2;
  if (expectedKey != 0 && hd.key != expectedKey) return 3;
  if (std::fseek(f.f, 0, SEEK_END) != 0) return 2;
  const long end = std::ftell(f.f);
  const u64 want = sizeof(hd) + u64(hd.numVertices) * sizeof(F3) + hd.blobBytes;
  if (end < 0 || u64(end) != want) return 2;
  for (int k = 0; k < 7; ++k)
    if (u64(hd.ranges[2 * k]) + hd.ranges[2 * k + 1] > hd.blobBytes) return 2;
  if (std::fseek(f.f, sizeof(hd), SEEK_SET) != 0) return 2;
  CachedScene c;
  c.key = hd.key;
  c.positions.resize(hd.numVertices);
  c.blob.resize(hd.blobBytes);
  if (hd.numVertices && std::fread(c.positions.data(), sizeof(F3), hd.numVertices, f.f) != hd.numVertices) return 2;
  if (hd.blobBytes && std::fread(c.blob.data(), 1, hd.blobBytes, f.f) != hd.blobBytes) return 2;
  const u64 sum = hashBytes(c.blob.data(), c.blob.size(),
                            hashBytes(c.positions.data(), c.positions.size() * sizeof(F3), 0));
  if (sum != hd.checksum) return 2;
  Range* rr = &c.ranges.triangles;
  for (int k = 0; k < 7; ++k) rr[k] = Range{hd.ranges[2 * k], hd.ranges[2 * k + 1]};
  out = std::move(c);
  return 0;
}

}  // namespace rtq
