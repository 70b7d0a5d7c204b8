// Seeded synthetic scenes and ray sets for the benchmark workloads. None of this exists in the
// reference (it loads OBJ assets, Renderer/Scene.cpp:185-227); the generators only produce the
// builder input (positions + 16-bit indexed triangle lists + sphere lights) and ray buffers.
#include "scenes.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>

namespace rtq {

// ---- random numbers: SplitMix64 seeding a PCG-XSH-RR 32-bit generator -----------------------
static u64 splitmix64(u64& s) {
  u64 z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

Rng::Rng(u64 seed, u64 stream) {
  u64 s = seed ^ (stream * 0xD1B54A32D192ED03ull);
  state_ = splitmix64(s);
  inc_ = splitmix64(s) | 1ull;
}

u32 Rng::next() {
  u64 old = state_;
  state_ = old * 6364136223846793005ull + inc_;
  u32 xorshifted = u32(((old >> 18) ^ old) >> 27);
  u32 rot = u32(old >> 59);
  return (xorshifted >> rot) | (xorshifted << ((32 - rot) & 31));
}

double Rng::uniform() {  // [0,1) with 53 bits
  u64 hi = next(), lo = next();
  return double(((hi << 32) | lo) >> 11) * (1.0 / 9007199254740992.0);
}

double Rng::normal() {  // Box-Muller
  double u1 = uniform(), u2 = uniform();
  if (u1 < 1e-300) u1 = 1e-300;
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
}

// ---- small double-precision vector helpers for generation -----------------------------------
struct D3 { double x, y, z; };
static D3 d3(double x, double y, double z) { return {x, y, z}; }
static D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
static D3 operator-(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static D3 operator*(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
static double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static D3 cross(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
static D3 normalize(D3 a) { return a * (1.0 / std::sqrt(dot(a, a))); }

// Random rotation from a uniform unit quaternion (Shoemake).
static void randomRotation(Rng& rng, double m[3][3]) {
  double u1 = rng.uniform(), u2 = rng.uniform(), u3 = rng.uniform();
  double a = std::sqrt(1 - u1), b = std::sqrt(u1);
  double qx = a * std::sin(2 * M_PI * u2), qy = a * std::cos(2 * M_PI * u2);
  double qz = b * std::sin(2 * M_PI * u3), qw = b * std::cos(2 * M_PI * u3);
  m[0][0] = 1 - 2 * (qy * qy + qz * qz); m[0][1] = 2 * (qx * qy - qz * qw); m[0][2] = 2 * (qx * qz + qy * qw);
  m[1][0] = 2 * (qx * qy + qz * qw); m[1][1] = 1 - 2 * (qx * qx + qz * qz); m[1][2] = 2 * (qy * qz - qx * qw);
  m[2][0] = 2 * (qx * qz - qy * qw); m[2][1] = 2 * (qy * qz + qx * qw); m[2][2] = 1 - 2 * (qx * qx + qy * qy);
}

// One addTriangleList call: vertices appended to the geometry store, triangles indexed from them.
static void emitList(BvhBuilder& b, const std::vector<D3>& verts, const std::vector<u32>& idx, float gray) {
  if (verts.size() > 65536) throw std::runtime_error("triangle list exceeds 16-bit indices");
  std::vector<float> xyz(verts.size() * 3);
  for (size_t i = 0; i < verts.size(); ++i) {
    xyz[3 * i] = float(verts[i].x);
    xyz[3 * i + 1] = float(verts[i].y);
    xyz[3 * i + 2] = float(verts[i].z);
  }
  u32 off = b.addVertices(u32(verts.size()), xyz.data());
  b.addTriangleList(off, u32(idx.size() / 3), idx.data(), BvhBuilder::lambertian(gray, gray, gray));
}

// A grid of n x n quads spanning origin + s*du + t*dv, s,t in [0,1]; 2 triangles per quad.
static void gridFace(std::vector<D3>& verts, std::vector<u32>& idx, D3 origin, D3 du, D3 dv, int n) {
  u32 base = u32(verts.size());
  for (int j = 0; j <= n; ++j)
    for (int i = 0; i <= n; ++i) verts.push_back(origin + du * (double(i) / n) + dv * (double(j) / n));
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      u32 a = base + u32(j * (n + 1) + i), c = a + 1, d = a + u32(n + 1), e = d + 1;
      idx.insert(idx.end(), {a, c, e, a, e, d});
    }
}

// Closed-box scene: walls of [0,10]^3 open toward -y (floor, ceiling, left, right, back), each a
// 32x32 grid of quads; two axis-aligned inner boxes whose faces are 16x16 grids. 16 384 triangles.
static void sceneBox(BvhBuilder& b) {
  const int W = 32;
  struct Wall { D3 o, du, dv; };
  const Wall walls[5] = {
      {d3(0, 0, 0), d3(10, 0, 0), d3(0, 10, 0)},   // floor z=0
      {d3(0, 0, 10), d3(0, 10, 0), d3(10, 0, 0)},  // ceiling z=10
      {d3(0, 0, 0), d3(0, 0, 10), d3(0, 10, 0)},   // left x=0
      {d3(10, 0, 0), d3(0, 10, 0), d3(0, 0, 10)},  // right x=10
      {d3(0, 10, 0), d3(10, 0, 0), d3(0, 0, 10)},  // back y=10
  };
  for (const Wall& w : walls) {
    std::vector<D3> v;
    std::vector<u32> idx;
    gridFace(v, idx, w.o, w.du, w.dv, W);
    emitList(b, v, idx, 0.7f);
  }
  const D3 boxes[2][2] = {{d3(1.5, 4.0, 0.2), d3(4.0, 6.5, 3.5)}, {d3(5.5, 6.0, 0.2), d3(8.5, 8.0, 6.0)}};
  for (const auto& bx : boxes) {
    D3 mn = bx[0], mx = bx[1], e = mx - mn;
    std::vector<D3> v;
    std::vector<u32> idx;
    const int F = 16;
    gridFace(v, idx, mn, d3(e.x, 0, 0), d3(0, e.y, 0), F);              // bottom
    gridFace(v, idx, d3(mn.x, mn.y, mx.z), d3(0, e.y, 0), d3(e.x, 0, 0), F);  // top
    gridFace(v, idx, mn, d3(0, 0, e.z), d3(e.x, 0, 0), F);              // front y=min
    gridFace(v, idx, d3(mn.x, mx.y, mn.z), d3(e.x, 0, 0), d3(0, 0, e.z), F);  // back y=max
    gridFace(v, idx, mn, d3(0, e.y, 0), d3(0, 0, e.z), F);              // x=min
    gridFace(v, idx, d3(mx.x, mn.y, mn.z), d3(0, 0, e.z), d3(0, e.y, 0), F);  // x=max
    emitList(b, v, idx, 0.5f);
  }
}

// UV sphere with 16 longitudes and 9 latitude bands: 2 poles + 8 rings of 16 = 130 vertices,
// 16 + 7*32 + 16 = 256 triangles.
static void uvSphere(std::vector<D3>& v, std::vector<u32>& idx) {
  const int L = 16, B = 9;
  v.push_back(d3(0, 0, 1));
  for (int k = 1; k < B; ++k) {
    double th = M_PI * k / B;
    for (int j = 0; j < L; ++j) {
      double ph = 2 * M_PI * j / L;
      v.push_back(d3(std::sin(th) * std::cos(ph), std::sin(th) * std::sin(ph), std::cos(th)));
    }
  }
  v.push_back(d3(0, 0, -1));
  auto ring = [&](int k, int j) { return u32(1 + (k - 1) * L + (j % L)); };
  for (int j = 0; j < L; ++j) idx.insert(idx.end(), {0u, ring(1, j), ring(1, j + 1)});
  for (int k = 1; k < B - 1; ++k)
    for (int j = 0; j < L; ++j) {
      u32 a = ring(k, j), c = ring(k, j + 1), d = ring(k + 1, j), e = ring(k + 1, j + 1);
      idx.insert(idx.end(), {a, d, e, a, e, c});
    }
  u32 south = u32(v.size() - 1);
  for (int j = 0; j < L; ++j) idx.insert(idx.end(), {south, ring(B - 1, j + 1), ring(B - 1, j)});
}

static void sceneSpheres(BvhBuilder& b, u64 seed, u32 count) {
  std::vector<D3> unitV;
  std::vector<u32> unitIdx;
  uvSphere(unitV, unitIdx);
  Rng rng(seed, 1);
  for (u32 s = 0; s < count; ++s) {
    D3 c = d3(rng.uniform() * 200 - 100, rng.uniform() * 200 - 100, rng.uniform() * 200 - 100);
    double r = std::exp(0.7 * rng.normal());
    double m[3][3];
    randomRotation(rng, m);
    std::vector<D3> v(unitV.size());
    for (size_t i = 0; i < unitV.size(); ++i) {
      D3 p = unitV[i];
      D3 q = d3(m[0][0] * p.x + m[0][1] * p.y + m[0][2] * p.z, m[1][0] * p.x + m[1][1] * p.y + m[1][2] * p.z,
                m[2][0] * p.x + m[2][1] * p.y + m[2][2] * p.z);
      v[i] = c + q * r;
    }
    emitList(b, v, unitIdx, 0.7f);
  }
}

// Triangle soup: 64 Gaussian clusters (means U[-100,100]^3, sigma LogU[2,20], equal weights),
// triangle centre from a cluster, edge length LogU[0.01,1], random orientation; at most 21 845
// triangles (65 535 vertices) per list.
static void sceneSoup(BvhBuilder& b, u64 seed, u32 count) {
  Rng rng(seed, 3);
  const int K = 64;
  D3 mean[K];
  double sigma[K];
  for (int k = 0; k < K; ++k) {
    mean[k] = d3(rng.uniform() * 200 - 100, rng.uniform() * 200 - 100, rng.uniform() * 200 - 100);
    sigma[k] = std::exp(std::log(2.0) + rng.uniform() * (std::log(20.0) - std::log(2.0)));
  }
  const u32 perList = 21845;
  for (u32 done = 0; done < count;) {
    u32 n = std::min(perList, count - done);
    std::vector<D3> v;
    std::vector<u32> idx;
    v.reserve(3 * n);
    for (u32 t = 0; t < n; ++t) {
      int k = int(rng.next() % K);
      D3 c = mean[k] + d3(rng.normal(), rng.normal(), rng.normal()) * sigma[k];
      double len = std::exp(std::log(0.01) + rng.uniform() * (std::log(1.0) - std::log(0.01)));
      double m[3][3];
      randomRotation(rng, m);
      // equilateral-ish triangle in the local xy plane, circumradius len/sqrt(3)
      double R = len / std::sqrt(3.0);
      for (int q = 0; q < 3; ++q) {
        double a = 2 * M_PI * q / 3 + rng.uniform() * 0.5;
        D3 p = d3(R * std::cos(a), R * std::sin(a), 0);
        D3 w = d3(m[0][0] * p.x + m[0][1] * p.y + m[0][2] * p.z, m[1][0] * p.x + m[1][1] * p.y + m[1][2] * p.z,
                  m[2][0] * p.x + m[2][1] * p.y + m[2][2] * p.z);
        v.push_back(c + w);
      }
      u32 base = 3 * t;
      idx.insert(idx.end(), {base, base + 1, base + 2});
    }
    emitList(b, v, idx, 0.7f);
    done += n;
  }
}

void generateScene(BvhBuilder& b, u32 kind, u64 seed, u32 param) {
  switch (kind) {
    case 1: sceneBox(b); break;
    case 2: sceneSpheres(b, seed, param); break;
    case 3: sceneSoup(b, seed, param); break;
    default: throw std::invalid_argument("unknown scene kind");
  }
}

void generateLights(BvhBuilder& b, u64 seed, u32 count, float lo, float hi, float rmin, float rmax) {
  Rng rng(seed, 7);
  for (u32 i = 0; i < count; ++i) {
    SphereLightDesc l{};
    l.pos = {float(lo + (hi - lo) * rng.uniform()), float(lo + (hi - lo) * rng.uniform()),
             float(lo + (hi - lo) * rng.uniform())};
    l.radius = float(std::exp(std::log(double(rmin)) + rng.uniform() * (std::log(double(rmax)) - std::log(double(rmin)))));
    l.color = {1, 1, 1};
    l.sphereRadius = 0.1f;
    b.addSphereLight(l);
  }
}

// Frustum corners at NDC depth 0.5 of a [0,1]-depth perspective with near 0.1 and far 100: the
// view depth there is 2nf/(n+f). Computed in double and rounded once.
void cameraFrustum(const float pos[3], const float dir[3], const float up[3], float fovY, float aspect, Frustum& f) {
  D3 eye = d3(pos[0], pos[1], pos[2]);
  D3 fwd = normalize(d3(dir[0], dir[1], dir[2]));
  D3 upv = d3(up[0], up[1], up[2]);
  D3 right = normalize(cross(fwd, upv));
  D3 trueUp = cross(right, fwd);
  const double n = 0.1, fa = 100.0;
  const double depth = 2 * n * fa / (n + fa);
  const double ty = depth * std::tan(double(fovY) / 2), tx = ty * aspect;
  D3 centre = eye + fwd * depth;
  D3 c00 = centre - right * tx - trueUp * ty;
  D3 c10 = centre + right * tx - trueUp * ty;
  D3 c01 = centre - right * tx + trueUp * ty;
  auto put = [](float* o, D3 v) { o[0] = float(v.x); o[1] = float(v.y); o[2] = float(v.z); };
  put(f.cameraPos, eye);
  put(f.corner00, c00);
  put(f.corner10, c10);
  put(f.corner01, c01);
}

// cameraPass.comp:50-54 with per-pixel jitter; fp32 like the shader.
void frustumRays(const Frustum& f, u32 width, u32 height, u64 seed, float tmax, RayRec* out) {
  const float invW = 1.0f / float(width), invH = 1.0f / float(height);
  float stepW[3], stepH[3];
  for (int c = 0; c < 3; ++c) {
    stepW[c] = (f.corner10[c] - f.corner00[c]) * invW;
    stepH[c] = (f.corner01[c] - f.corner00[c]) * invH;
  }
  Rng rng(seed, 11);
  for (u32 y = 0; y < height; ++y)
    for (u32 x = 0; x < width; ++x) {
      float jx = 0, jy = 0;
      if (seed != 0) {
        jx = float(rng.uniform());
        jy = float(rng.uniform());
      }
      float fx = float(x) + jx, fy = float(y) + jy;
      float d[3];
      for (int c = 0; c < 3; ++c) {
        float target = f.corner00[c] + stepW[c] * fx + stepH[c] * fy;
        d[c] = target - f.cameraPos[c];
      }
      float len = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      RayRec& r = out[size_t(y) * width + x];
      for (int c = 0; c < 3; ++c) {
        r.org[c] = f.cameraPos[c];
        r.dir[c] = d[c] / len;
      }
      r.tmax = tmax;
      r.pad = 0;
    }
}

static void cosineDirection(Rng& rng, D3 n, D3& out) {
  D3 a = std::fabs(n.x) > 0.9 ? d3(0, 1, 0) : d3(1, 0, 0);
  D3 t = normalize(cross(n, a));
  D3 bt = cross(n, t);
  double u1 = rng.uniform(), u2 = rng.uniform();
  double r = std::sqrt(u1), ph = 2 * M_PI * u2;
  out = normalize(t * (r * std::cos(ph)) + bt * (r * std::sin(ph)) + n * std::sqrt(std::max(0.0, 1 - u1)));
}

static D3 faceNormal(const BvhBuilder& b, u32 tri, D3& p0, D3& p1, D3& p2) {
  const TriangleRec& t = b.triangles()[tri];
  const auto& P = b.positions();
  F3 a = P[t.vertexOffset + (t.index01 & 0xFFFFu)], c = P[t.vertexOffset + (t.index01 >> 16)],
     e = P[t.vertexOffset + (t.index2_matId & 0xFFFFu)];
  p0 = d3(a.x, a.y, a.z);
  p1 = d3(c.x, c.y, c.z);
  p2 = d3(e.x, e.y, e.z);
  D3 n = cross(p1 - p0, p2 - p0);
  double l = std::sqrt(dot(n, n));
  return l > 0 ? n * (1.0 / l) : d3(0, 0, 1);
}

static void putRay(RayRec& r, D3 o, D3 d, float tmax) {
  r.org[0] = float(o.x); r.org[1] = float(o.y); r.org[2] = float(o.z);
  r.dir[0] = float(d.x); r.dir[1] = float(d.y); r.dir[2] = float(d.z);
  r.tmax = tmax;
  r.pad = 0;
}

void surfaceRays(const BvhBuilder& b, u64 seed, u32 count, float tmax, RayRec* out) {
  Rng rng(seed, 13);
  const u32 nT = b.numTriangles();
  for (u32 i = 0; i < count; ++i) {
    u32 tri = u32(rng.uniform() * nT);
    if (tri >= nT) tri = nT - 1;
    D3 p0, p1, p2;
    D3 n = faceNormal(b, tri, p0, p1, p2);
    if (rng.next() & 1) n = n * -1.0;
    double s = rng.uniform(), t = rng.uniform();
    if (s + t > 1) { s = 1 - s; t = 1 - t; }
    D3 o = p0 + (p1 - p0) * s + (p2 - p0) * t + n * 0.001;
    D3 d;
    cosineDirection(rng, n, d);
    putRay(out[i], o, d, tmax);
  }
}

void bounceRays(const BvhBuilder& b, u64 seed, const RayRec* rays, const HitRec* hits, u32 count, float tmax,
                RayRec* out) {
  for (u32 i = 0; i < count; ++i) {
    if (hits[i].triId == 0xFFFFFFFFu || hits[i].triId >= b.numTriangles()) {
      putRay(out[i], d3(rays[i].org[0], rays[i].org[1], rays[i].org[2]),
             d3(rays[i].dir[0], rays[i].dir[1], rays[i].dir[2]), 0.0f);
      continue;
    }
    D3 p0, p1, p2;
    D3 n = faceNormal(b, hits[i].triId, p0, p1, p2);
    D3 din = d3(rays[i].dir[0], rays[i].dir[1], rays[i].dir[2]);
    if (dot(n, din) > 0) n = n * -1.0;
    D3 o = d3(rays[i].org[0], rays[i].org[1], rays[i].org[2]) + din * double(hits[i].t) + n * 0.001;
    Rng local(seed ^ (u64(i) * 0x9E3779B97F4A7C15ull), 19);
    D3 d;
    cosineDirection(local, n, d);
    putRay(out[i], o, d, tmax);
  }
}

// Stratified uniform directions on the sphere as LightProbFieldPass::sampleRays builds the probe
// rays (lightProbFieldPass.cpp:151-180): cell (i, j) of a sqrt(count)^2 grid, a jittered point r
// in it, theta = 2*pi*r.x, phi = acos(1 - 2*r.y), direction (sin phi cos theta, sin phi sin theta,
// cos phi); index i*sqrt(count) + j. The jitter comes from this library's generator.
void probeDirections(u64 seed, u32 count, float* out) {
  u32 side = 1;
  while (side * side < count) ++side;
  Rng rng(seed, 0x70726f6265ull);
  const float kPi = 3.14159265358979f;
  for (u32 i = 0; i < side; ++i)
    for (u32 j = 0; j < side; ++j) {
      const u32 k = i * side + j;
      const float rx = (float(i) + float(rng.uniform())) / float(side);
      const float ry = (float(j) + float(rng.uniform())) / float(side);
      if (k >= count) continue;
      const float theta = 2.0f * kPi * rx, phi = std::acos(1.0f - 2.0f * ry);
      out[4 * k] = std::sin(phi) * std::cos(theta);
      out[4 * k + 1] = std::sin(phi) * std::sin(theta);
      out[4 * k + 2] = std::cos(phi);
      out[4 * k + 3] = 0.0f;
    }
}

}  // namespace rtq
