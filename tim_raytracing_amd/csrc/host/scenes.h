// Seeded synthetic scenes, lights, cameras and rays for the benchmark workloads.
#pragma once
#include "bvh_builder.h"

namespace rtq {

class Rng {
 public:
  Rng(u64 seed, u64 stream);
  u32 next();
  double uniform();
  double normal();

 private:
  u64 state_, inc_;
};

// Same layout as rtq_ray / rtq_hit / rtq_frustum in include/rtq.h.
struct RayRec { float org[3]; float tmax; float dir[3]; u32 pad; };
struct HitRec { u32 triId; float t, u, v; };
struct Frustum { float cameraPos[3], corner00[3], corner10[3], corner01[3]; };

void generateScene(BvhBuilder& b, u32 kind, u64 seed, u32 param);
void generateLights(BvhBuilder& b, u64 seed, u32 count, float lo, float hi, float rmin, float rmax);
void cameraFrustum(const float pos[3], const float dir[3], const float up[3], float fovY, float aspect, Frustum& f);
void frustumRays(const Frustum& f, u32 width, u32 height, u64 seed, float tmax, RayRec* out);
void surfaceRays(const BvhBuilder& b, u64 seed, u32 count, float tmax, RayRec* out);
void bounceRays(const BvhBuilder& b, u64 seed, const RayRec* rays, const HitRec* hits, u32 count, float tmax,
                RayRec* out);
void probeDirections(u64 seed, u32 count, float* out);

}  // namespace rtq
