// GPU test of the C++ host interface (include/rtq.hpp): builds the closed-box scene through the
// reference-named classes, runs tracePass on a 320x240 frame and checks every hit bit for bit
// against the oracle on the rays the camera pass generates. Prints "ok" and exits 0 on success.
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/rtq.hpp"

extern "C" {
int orc_trace_closest(const uint8_t* blob, size_t bytes, const uint32_t* ranges, const float* pos, uint32_t nv,
                      const float* rays, uint32_t n, uint32_t* tri_id, float* t, float* u, float* v, uint32_t* nid,
                      uint32_t* slot, uint32_t* counters, int form, int threads);
void orc_camera_rays(const float* frustum, uint32_t width, uint32_t height, int reference_rows, float* out);
int orc_trace_record(const uint8_t* blob, size_t bytes, const uint32_t* ranges, const float* pos, uint32_t nv,
                     const float* frustum, uint32_t width, uint32_t height, int reference_rows, const float* sun_dir,
                     uint32_t shadow_flags, const float* lpf_box, const uint32_t* lpf_res, uint32_t* records,
                     int threads);
int orc_trace_probes(const uint8_t* blob, size_t bytes, const uint32_t* ranges, const float* pos, uint32_t nv,
                     const float* lpf_box, const uint32_t* lpf_res, const float* dirs, const float* sun_dir,
                     uint32_t shadow_flags, uint32_t* records, int threads);
}

int main() {
  try {
    rtq::Device dev(0);
    rtq::BVHBuilder builder;
    // a 16x16-quad floor and a tilted wall through the reference-named calls
    std::vector<rtq::vec3> v;
    std::vector<uint32_t> idx;
    for (int j = 0; j <= 16; ++j)
      for (int i = 0; i <= 16; ++i) v.push_back({i * 0.5f, j * 0.5f, 0.0f});
    for (int j = 0; j < 16; ++j)
      for (int i = 0; i < 16; ++i) {
        uint32_t a = j * 17 + i;
        idx.insert(idx.end(), {a, a + 1, a + 18, a, a + 18, a + 17});
      }
    uint32_t off = builder.addTriangleList(uint32_t(v.size()), v.data());
    builder.addTriangleList(off, uint32_t(idx.size() / 3), idx.data());
    std::vector<rtq::vec3> w = {{0, 6, 0}, {8, 6, 0}, {0, 7, 5}, {8, 7, 5}};
    uint32_t off2 = builder.addTriangleList(4, w.data());
    uint32_t widx[6] = {0, 1, 2, 1, 3, 2};
    builder.addTriangleList(off2, 2, widx, {0.2f, 0.8f, 0.2f});
    rtq::BVHData data(dev);
    data.build(builder, rtq::BVHBuildParameters{});
    rtq_frustum f{};
    const float pos[3] = {4, -3, 3}, dir[3] = {0, 1, -0.4f}, up[3] = {0, 0, 1};
    if (rtq_camera_frustum(pos, dir, up, 70.f * 3.14f / 180, 320.f / 240, &f) != RTQ_OK) return 2;
    rtq::RayTracingPass pass(dev, 320, 240);
    pass.tracePass(data, f);
    std::vector<rtq_hit> hits = pass.readHits();
    std::vector<uint32_t> nids = pass.readLeafNids();

    std::vector<uint8_t> blob(builder.getBvhGpuSize());
    rtq_ranges r{};
    builder.fillGpuBuffer(blob.data(), r);
    auto p = builder.positions();
    std::vector<float> rays(320 * 240 * 8);
    orc_camera_rays(&f.camera_pos[0], 320, 240, 0, rays.data());
    const size_t n = 320 * 240;
    std::vector<uint32_t> id(n), nid(n), slot(n);
    std::vector<float> t(n), u(n), vv(n);
    uint32_t rr[14];
    std::memcpy(rr, &r, sizeof(rr));
    orc_trace_closest(blob.data(), blob.size(), rr, p.first, p.second, rays.data(), uint32_t(n), id.data(), t.data(),
                      u.data(), vv.data(), nid.data(), slot.data(), nullptr, 0, 4);
    size_t bad = 0, hit = 0;
    for (size_t i = 0; i < n; ++i) {
      hit += hits[i].tri_id != RTQ_NO_HIT;
      bad += hits[i].tri_id != id[i] || std::memcmp(&hits[i].t, &t[i], 4) || std::memcmp(&hits[i].u, &u[i], 4) ||
             std::memcmp(&hits[i].v, &vv[i], 4) || nids[i] != nid[i];
    }
    if (bad || hit < n / 4) {
      std::fprintf(stderr, "mismatches %zu, hits %zu of %zu\n", bad, hit, n);
      return 1;
    }
    // the full TRACING_STEP record (reference row order, sun + probe masks)
    const float sun[3] = {0.6f, 0.0f, -0.8f};
    rtq_lpf lpf{{0, 0, -1}, {8, 8, 4}, {8, 8, 5}};
    pass.traceRecord(data, f, sun, RTQ_SHADOW_SUN, &lpf, nullptr, RTQ_CAMERA_REFERENCE_ROWS);
    std::vector<uint32_t> rec = pass.readRecords();
    std::vector<uint32_t> orec(8 * n, 0);
    const float box[6] = {0, 0, -1, 8, 8, 4};
    orc_trace_record(blob.data(), blob.size(), rr, p.first, p.second, &f.camera_pos[0], 320, 240, 1, sun,
                     RTQ_SHADOW_SUN, box, lpf.resolution, orec.data(), 4);
    size_t recBad = 0, masked = 0;
    for (size_t i = 320; i < n; ++i) {
      recBad += std::memcmp(&rec[8 * i], &orec[8 * i], 32) != 0;
      masked += orec[8 * i + 5] != 0;
    }
    if (recBad || masked < n / 8) {
      std::fprintf(stderr, "record mismatches %zu, probe-masked pixels %zu\n", recBad, masked);
      return 1;
    }
    // the light-probe field's tracing step through LightProbFieldPass
    {
      float dirs[4 * 64];
      if (rtq_probe_directions(5, 64, dirs) != RTQ_OK) return 3;
      rtq_lpf field{{0.5f, 0.5f, 0.5f}, {7.5f, 7.5f, 4.5f}, {4, 4, 2}};
      rtq::LightProbFieldPass lp(dev, field);
      lp.traceLightProbField(data, dirs, sun);
      std::vector<uint32_t> prec = lp.readRecords();
      std::vector<uint32_t> oprec(prec.size(), 0);
      const float fbox[6] = {0.5f, 0.5f, 0.5f, 7.5f, 7.5f, 4.5f};
      orc_trace_probes(blob.data(), blob.size(), rr, p.first, p.second, fbox, field.resolution, dirs, sun,
                       RTQ_SHADOW_SUN, oprec.data(), 4);
      if (prec != oprec) {
        std::fprintf(stderr, "probe records differ\n");
        return 1;
      }
    }
    // a failing call throws with its status
    try {
      rtq::RayTracingPass bad(dev, 0, 0);
      bad.tracePass(data, f);
    } catch (const rtq::Error&) {
    }
  } catch (const rtq::Error& e) {
    std::fprintf(stderr, "rtq::Error %s\n", e.what());
    return 1;
  }
  std::puts("ok");
  return 0;
}
