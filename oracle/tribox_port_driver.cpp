// TEST INFRASTRUCTURE ONLY: exports the builder's triangle/box test (tribox.h) with the same C
// signature as oracle/tribox_ref_driver.c so tests can run both on the same inputs.
#include "../tim_raytracing_amd/csrc/host/tribox.h"

extern "C" int port_tri_box_overlap(const float* center, const float* half, const float* v) {
  rtq::F3 p0{v[0], v[1], v[2]}, p1{v[3], v[4], v[5]}, p2{v[6], v[7], v[8]};
  return rtq::satTriangleBox(center, half, p0, p1, p2) ? 1 : 0;
}

extern "C" int port_triangle_box_collision(const float* mn, const float* mx, const float* v) {
  rtq::F3 p0{v[0], v[1], v[2]}, p1{v[3], v[4], v[5]}, p2{v[6], v[7], v[8]};
  rtq::Box3 b{{mn[0], mn[1], mn[2]}, {mx[0], mx[1], mx[2]}};
  return rtq::triangleBoxOverlap(p0, p1, p2, b) ? 1 : 0;
}
