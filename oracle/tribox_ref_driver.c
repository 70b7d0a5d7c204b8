/* TEST INFRASTRUCTURE ONLY: exports the reference's own triangle/box SAT test
 * (Renderer/TriBoxCollision.hpp:174-324, included from the reference tree at build time) so the
 * builder's restatement in tim_raytracing_amd/csrc/host/tribox.h can be compared bit for bit. */
#include "TriBoxCollision.hpp"

int ref_tri_box_overlap(float* center, float* half, float* verts9) {
  float v[3][3] = {{verts9[0], verts9[1], verts9[2]}, {verts9[3], verts9[4], verts9[5]},
                   {verts9[6], verts9[7], verts9[8]}};
  return triBoxOverlap(center, half, v);
}
