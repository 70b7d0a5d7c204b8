// Triangle/box overlap as the reference builder evaluates it.
//
// triangleVerticesBoxCollision (Renderer/BVHBuilder.cpp:65-79) first accepts when the third
// vertex lies inside the box (inclusive): its condition is `(in(p0) || in(p1)), in(p2)`, and the
// comma operator keeps only in(p2). Otherwise it runs the separating-axis test of Akenine-Moller,
// "Fast 3D Triangle-Box Overlap Testing" (2001), with the box given as centre and half extents
// (TriBoxCollision.hpp:174-324): the nine edge-cross-axis tests first (edge 0, 1, 2; axes x, y, z
// in the order the paper's optimised version uses), then the three box-axis tests, then the
// triangle plane against the box. Each expression below keeps that operation order, so with
// contraction off the float results match the reference bit for bit.
#pragma once
#include <cmath>

#include "bvh_types.h"

namespace rtq {

namespace tribox_detail {

// One edge-cross-axis projection: the triangle projects to [min(pa,pb), max(pa,pb)] and the box
// to [-rad, rad]; `pbFirst` selects the comparison orientation of the paper's Z12 variant.
inline bool separated(float pa, float pb, float rad, bool pbFirst) {
  float mn, mx;
  if (pbFirst) {
    if (pb < pa) { mn = pb; mx = pa; } else { mn = pa; mx = pb; }
  } else {
    if (pa < pb) { mn = pa; mx = pb; } else { mn = pb; mx = pa; }
  }
  return mn > rad || mx < -rad;
}

// planeBoxOverlap (TriBoxCollision.hpp:65-109).
inline bool planeTouchesBox(const float n[3], const float v[3], const float h[3]) {
  float vmin[3], vmax[3];
  for (int q = 0; q < 3; ++q) {
    float x = v[q];
    if (n[q] > 0.0f) {
      vmin[q] = -h[q] - x;
      vmax[q] = h[q] - x;
    } else {
      vmin[q] = h[q] - x;
      vmax[q] = -h[q] - x;
    }
  }
  if (n[0] * vmin[0] + n[1] * vmin[1] + n[2] * vmin[2] > 0.0f) return false;
  if (n[0] * vmax[0] + n[1] * vmax[1] + n[2] * vmax[2] >= 0.0f) return true;
  return false;
}

}  // namespace tribox_detail

// Separating-axis overlap of a triangle and a box given by centre c and half extents h.
inline bool satTriangleBox(const float c[3], const float h[3], const F3& p0, const F3& p1, const F3& p2) {
  using tribox_detail::separated;
  const float a0[3] = {p0.x - c[0], p0.y - c[1], p0.z - c[2]};
  const float a1[3] = {p1.x - c[0], p1.y - c[1], p1.z - c[2]};
  const float a2[3] = {p2.x - c[0], p2.y - c[1], p2.z - c[2]};
  const float e0[3] = {a1[0] - a0[0], a1[1] - a0[1], a1[2] - a0[2]};
  const float e1[3] = {a2[0] - a1[0], a2[1] - a1[1], a2[2] - a1[2]};
  const float e2[3] = {a0[0] - a2[0], a0[1] - a2[1], a0[2] - a2[2]};

  // x-axis family: p = ez*vy - ey*vz ; y family: p = -ez*vx + ex*vz ; z family: p = ey*vx - ex*vy
  auto px = [](const float* e, const float* v) { return e[2] * v[1] - e[1] * v[2]; };
  auto py = [](const float* e, const float* v) { return -e[2] * v[0] + e[0] * v[2]; };
  auto pz = [](const float* e, const float* v) { return e[1] * v[0] - e[0] * v[1]; };

  {  // edge 0: X01, Y02, Z12
    float fx = std::fabs(e0[0]), fy = std::fabs(e0[1]), fz = std::fabs(e0[2]);
    if (separated(px(e0, a0), px(e0, a2), fz * h[1] + fy * h[2], false)) return false;
    if (separated(py(e0, a0), py(e0, a2), fz * h[0] + fx * h[2], false)) return false;
    if (separated(pz(e0, a1), pz(e0, a2), fy * h[0] + fx * h[1], true)) return false;
  }
  {  // edge 1: X01, Y02, Z0
    float fx = std::fabs(e1[0]), fy = std::fabs(e1[1]), fz = std::fabs(e1[2]);
    if (separated(px(e1, a0), px(e1, a2), fz * h[1] + fy * h[2], false)) return false;
    if (separated(py(e1, a0), py(e1, a2), fz * h[0] + fx * h[2], false)) return false;
    if (separated(pz(e1, a0), pz(e1, a1), fy * h[0] + fx * h[1], false)) return false;
  }
  {  // edge 2: X2, Y1, Z12
    float fx = std::fabs(e2[0]), fy = std::fabs(e2[1]), fz = std::fabs(e2[2]);
    if (separated(px(e2, a0), px(e2, a1), fz * h[1] + fy * h[2], false)) return false;
    if (separated(py(e2, a0), py(e2, a1), fz * h[0] + fx * h[2], false)) return false;
    if (separated(pz(e2, a1), pz(e2, a2), fy * h[0] + fx * h[1], true)) return false;
  }
  // box axes: triangle extent per axis (first-found min/max as FINDMINMAX scans)
  for (int q = 0; q < 3; ++q) {
    float mn = a0[q], mx = a0[q];
    if (a1[q] < mn) mn = a1[q];
    if (a1[q] > mx) mx = a1[q];
    if (a2[q] < mn) mn = a2[q];
    if (a2[q] > mx) mx = a2[q];
    if (mn > h[q] || mx < -h[q]) return false;
  }
  // triangle plane: normal = e0 x e1
  const float n[3] = {e0[1] * e1[2] - e0[2] * e1[1], e0[2] * e1[0] - e0[0] * e1[2], e0[0] * e1[1] - e0[1] * e1[0]};
  return tribox_detail::planeTouchesBox(n, a0, h);
}

// triangleVerticesBoxCollision != Disjoint.
inline bool triangleBoxOverlap(const F3& p0, const F3& p1, const F3& p2, const Box3& b) {
  const bool p2Inside =
      !(p2.x < b.mn.x || p2.y < b.mn.y || p2.z < b.mn.z || p2.x > b.mx.x || p2.y > b.mx.y || p2.z > b.mx.z);
  if (p2Inside) return true;
  const float c[3] = {(b.mn.x + b.mx.x) * 0.5f, (b.mn.y + b.mx.y) * 0.5f, (b.mn.z + b.mx.z) * 0.5f};
  const float h[3] = {(b.mx.x - b.mn.x) * 0.5f, (b.mx.y - b.mn.y) * 0.5f, (b.mx.z - b.mn.z) * 0.5f};
  return satTriangleBox(c, h, p0, p1, p2);
}

}  // namespace rtq
