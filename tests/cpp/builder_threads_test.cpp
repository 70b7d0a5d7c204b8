// Threaded builds of the host library (linked from its sources, built with -fsanitize=thread by
// tests/test_sanitizers.py): a one-level soup and a two-level scene, each built on 1 and on 8 threads;
// the blobs must be identical and ThreadSanitizer must report nothing. Prints "ok" and exits 0.
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/rtq_build.h"

namespace {

std::vector<uint8_t> blobOf(rtq_builder* b) {
  size_t bytes = 0;
  rtq_ranges r{};
  std::vector<uint8_t> out;
  if (rtq_builder_gpu_size(b, &bytes) != RTQ_OK) return out;
  out.resize(bytes);
  if (rtq_builder_fill_gpu_buffer(b, out.data(), bytes, &r) != RTQ_OK) out.clear();
  return out;
}

std::vector<uint8_t> soup(int threads) {
  rtq_builder* b = nullptr;
  rtq_builder_create(&b);
  rtq_scene_generate(b, RTQ_SCENE_SOUP, 9, 20000);
  rtq_lights_generate(b, 9, 64, -100, 100, 0.5f, 20);
  std::vector<uint8_t> out;
  if (rtq_builder_build(b, threads) == RTQ_OK) out = blobOf(b);
  rtq_builder_destroy(b);
  return out;
}

std::vector<uint8_t> twoLevel(int threads) {
  rtq_builder* b = nullptr;
  rtq_builder_create(&b);
  const rtq_build_params top{25, 6, 6, 1.0f, 0.225f};
  rtq_builder_set_parameters(b, &top);
  uint32_t state = 12345u;
  auto rnd = [&] {
    state = state * 1664525u + 1013904223u;
    return float(state >> 8) / float(1u << 23) - 1.0f;
  };
  std::vector<uint32_t> idx(3 * 400);
  for (uint32_t i = 0; i < idx.size(); ++i) idx[i] = i;
  for (int k = 0; k < 48; ++k) {
    const float c[3] = {30 * rnd(), 30 * rnd(), 30 * rnd()};
    std::vector<float> v;
    for (int t = 0; t < 400; ++t) {  // small triangles scattered around the cluster centre
      const float tc[3] = {c[0] + 3 * rnd(), c[1] + 3 * rnd(), c[2] + 3 * rnd()};
      for (int q = 0; q < 3; ++q)
        for (int a = 0; a < 3; ++a) v.push_back(tc[a] + 0.3f * rnd());
    }
    uint32_t off = 0, bi = 0;
    rtq_builder_add_vertices(b, v.data(), uint32_t(v.size() / 3), &off);
    rtq_builder_add_blas(b, off, 400, idx.data(), nullptr, &bi);
  }
  std::vector<uint8_t> out;
  if (rtq_builder_build(b, threads) == RTQ_OK) out = blobOf(b);
  rtq_builder_destroy(b);
  return out;
}

}  // namespace

int main() {
  const std::vector<uint8_t> s1 = soup(1), s8 = soup(8);
  if (s1.empty() || s1 != s8) {
    std::fprintf(stderr, "soup: threaded build differs (%zu vs %zu bytes)\n", s1.size(), s8.size());
    return 1;
  }
  const std::vector<uint8_t> t1 = twoLevel(1), t8 = twoLevel(8);
  if (t1.empty() || t1 != t8) {
    std::fprintf(stderr, "two-level: threaded build differs (%zu vs %zu bytes)\n", t1.size(), t8.size());
    return 1;
  }
  std::puts("ok");
  return 0;
}
