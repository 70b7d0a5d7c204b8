/*
 * rtq oracle: scalar CPU restatement of the reference's GLSL ray-query path.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker the HIP kernels are compared against; it is
 * loaded only by tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py, never by
 * the product path.
 *
 * It reads the flattened blob exactly as the shaders bind it (BVHData::fillBvhBindings,
 * Renderer/BVHData.cpp:17-26): PackedBVHNode records, the leaf u32 stream with inline triangles,
 * PackedLight records and the vertex position array. Traversal is the stackless bitstack form of
 * the shaders, not the device's stack form, so the two are independent restatements.
 *
 * GLSL leaves min/max of NaN and the order of dot/cross terms to the driver; the fixed
 * conventions here (shared with the device) are:
 *   min(x,y) = y < x ? y : x,  max(x,y) = x < y ? y : x   (the GLSL specification's definitions)
 *   dot(a,b) = (a.x*b.x + a.y*b.y) + a.z*b.z, cross as written in collision.glsl, no FMA.
 * Build with -ffp-contract=off and without fast-math.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef uint32_t u32;

#define NID_LEAF_BIT 0x80000000u   /* struct_cpp.glsl:52 */
#define NID_MASK 0x3FFFFFFFu       /* struct_cpp.glsl:53 */
#define NID_LEFT_LEAF_BIT 0x80000000u
#define NID_RIGHT_LEAF_BIT 0x40000000u
#define TMAX 100.0f                /* struct_cpp.glsl:5 */
#define OFFSET_RAY_COLLISION 10e-6f /* system.glsl:45, an fp32 literal in GLSL */
#define TRI_EPSILON 10e-8f          /* collision.glsl:62 */

#ifdef ORC_PLANT
/* Planted mistakes, used only by the tests that check the pins catch them:
 * 1 = tie-break to the far child (d1 <= d0), 2 = no t shrink on acceptance, 3 = u clamp dropped,
 * 4 = probe cell without the half-cell shift, 5 = probe cell not capped at resolution - 2,
 * 6 = probe position without the half-step offset, 7 = tile corners not widened by a quarter step,
 * 8 = reference storage row off by one (16 * ceil(H / 16) - y - 1), 9 = BLAS instance box accepted
 * at 0 (>= 0 instead of > 0), 10 = a top-level leaf's own triangles tested as well. */
static const int kPlant = ORC_PLANT;
#else
static const int kPlant = 0;
#endif

typedef struct { float x, y, z; } v3;

static inline float gmin(float x, float y) { return (y < x) ? y : x; }
static inline float gmax(float x, float y) { return (x < y) ? y : x; }
static inline v3 vsub(v3 a, v3 b) { v3 r = {a.x - b.x, a.y - b.y, a.z - b.z}; return r; }
static inline float vdot(v3 a, v3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static inline v3 vcross(v3 a, v3 b) {
  v3 r = {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
  return r;
}

/* ---------------------------------------------------------------------------------------------
 * CollideBox (Shaders/core/collision.glsl:23-40) with STORE_RAY_INVDIR 0 (system.glsl:28): IEEE
 * division per slab; returns tmin if tmin < tmax' else -1. */
float orc_collide_box(const float* o, const float* d, const float* mn, const float* mx, float tmax) {
  float t0x = (mn[0] - o[0]) / d[0], t0y = (mn[1] - o[1]) / d[1], t0z = (mn[2] - o[2]) / d[2];
  float t1x = (mx[0] - o[0]) / d[0], t1y = (mx[1] - o[1]) / d[1], t1z = (mx[2] - o[2]) / d[2];
  float sx = gmin(t0x, t1x), sy = gmin(t0y, t1y), sz = gmin(t0z, t1z);
  float bx = gmax(t0x, t1x), by = gmax(t0y, t1y), bz = gmax(t0z, t1z);
  float tmin = gmax(0.0f, gmax(sx, gmax(sy, sz)));
  float tm = gmin(tmax, gmin(bx, gmin(by, bz)));
  return (tmin < tm) ? tmin : -1.0f;
}

/* CollideTriangle (collision.glsl:58-95), IGNORE_BACKFACE off; also returns u, v. */
float orc_collide_triangle(const float* o, const float* d, const float* a, const float* b, const float* c,
                           float tmax, float* uo, float* vo) {
  v3 O = {o[0], o[1], o[2]}, D = {d[0], d[1], d[2]};
  v3 p0 = {a[0], a[1], a[2]}, p1 = {b[0], b[1], b[2]}, p2 = {c[0], c[1], c[2]};
  v3 e1 = vsub(p1, p0), e2 = vsub(p2, p0);
  v3 h = vcross(D, e2);
  float A = vdot(e1, h);
  if (fabsf(A) < TRI_EPSILON) return -1.0f;
  float f = 1.0f / A;
  v3 s = vsub(O, p0);
  float u = f * vdot(s, h);
  if (kPlant != 3) {
    /* clamp(u, 0, 1) != u, with clamp = min(max(u, 0), 1) */
    if (gmin(gmax(u, 0.0f), 1.0f) != u) return -1.0f;
  }
  v3 q = vcross(s, e1);
  float v = f * vdot(D, q);
  if (v < 0.0f || u + v > 1.0f) return -1.0f;
  float t = f * vdot(e2, q);
  if (t > TRI_EPSILON && t < tmax) {
    if (uo) *uo = u;
    if (vo) *vo = v;
    return t;
  }
  return -1.0f;
}

/* Double-precision Moller-Trumbore for error statements (no epsilon rejections beyond the
 * parallel test). Returns 1 and t,u,v if the line hits the triangle's plane inside it. */
int orc_intersect_f64(const float* o, const float* d, const float* a, const float* b, const float* c, double* t,
                      double* u, double* v) {
  double O[3] = {o[0], o[1], o[2]}, D[3] = {d[0], d[1], d[2]};
  double e1[3], e2[3], s[3], h[3], q[3];
  for (int k = 0; k < 3; ++k) {
    e1[k] = (double)b[k] - a[k];
    e2[k] = (double)c[k] - a[k];
    s[k] = O[k] - a[k];
  }
  h[0] = D[1] * e2[2] - D[2] * e2[1];
  h[1] = D[2] * e2[0] - D[0] * e2[2];
  h[2] = D[0] * e2[1] - D[1] * e2[0];
  double A = e1[0] * h[0] + e1[1] * h[1] + e1[2] * h[2];
  if (A == 0.0) return 0;
  double f = 1.0 / A;
  double uu = f * (s[0] * h[0] + s[1] * h[1] + s[2] * h[2]);
  q[0] = s[1] * e1[2] - s[2] * e1[1];
  q[1] = s[2] * e1[0] - s[0] * e1[2];
  q[2] = s[0] * e1[1] - s[1] * e1[0];
  double vv = f * (D[0] * q[0] + D[1] * q[1] + D[2] * q[2]);
  double tt = f * (e2[0] * q[0] + e2[1] * q[1] + e2[2] * q[2]);
  *t = tt;
  *u = uu;
  *v = vv;
  return 1;
}

/* ---------------------------------------------------------------------------------------------
 * Blob view. `ranges` is the 7 (offset, size) pairs of rtq_ranges in fillGpuBuffer order:
 * triangles, primitives, materials, lights, nodes, leaf_data, blas. */
typedef struct {
  const uint8_t* blob;
  const u32* nodes;  /* 16 words per PackedBVHNode */
  u32 num_nodes;
  const u32* leaf;   /* g_BvhLeafData */
  u32 leaf_words;
  const u32* tris;   /* g_BvhTriangleData, 3 words each */
  u32 num_tris;
  const float* lights; /* PackedLight: 14 words each (iparam + 13 floats) */
  u32 num_lights;
  const float* pos;  /* g_positionData */
  u32 num_vertices;
  const uint8_t* blas; /* g_blasHeader: 32 bytes each (minExtent, matId, maxExtent, rootIndex) */
  u32 num_blas;
  /* triple -> first triangle index (open addressing) */
  u32* map_keys;     /* 3 words per bucket */
  u32* map_vals;
  u32 map_cap;
} orc_view;

static u32 hash3(u32 a, u32 b, u32 c) {
  uint64_t h = 1469598103934665603ull;
  h = (h ^ a) * 1099511628211ull;
  h = (h ^ b) * 1099511628211ull;
  h = (h ^ c) * 1099511628211ull;
  return (u32)(h ^ (h >> 32));
}

static int view_init(orc_view* v, const uint8_t* blob, size_t bytes, const u32* r, const float* pos, u32 nv) {
  memset(v, 0, sizeof(*v));
  static const int used[4] = {0, 3, 4, 5}; /* triangles, lights, nodes, leaf data */
  for (int k = 0; k < 4; ++k) {
    int i = used[k];
    if ((size_t)r[2 * i] + r[2 * i + 1] > bytes) return -1;
  }
  v->blob = blob;
  v->nodes = (const u32*)(blob + r[8]);
  v->num_nodes = r[9] / 64u;
  v->leaf = (const u32*)(blob + r[10]);
  v->leaf_words = r[11] / 4u;
  v->tris = (const u32*)(blob + r[0]);
  v->num_tris = r[1] / 12u;
  v->lights = (const float*)(blob + r[6]);
  v->num_lights = r[7] / 56u;
  v->pos = pos;
  v->num_vertices = nv;
  if ((size_t)r[12] + r[13] <= bytes) {
    v->blas = blob + r[12];
    v->num_blas = r[13] / 32u;
  }
  u32 cap = 16;
  while (cap < 2 * v->num_tris + 16) cap <<= 1;
  v->map_cap = cap;
  v->map_keys = (u32*)malloc(sizeof(u32) * 3 * cap);
  v->map_vals = (u32*)malloc(sizeof(u32) * cap);
  if (!v->map_keys || !v->map_vals) return -2;
  for (u32 i = 0; i < cap; ++i) v->map_vals[i] = 0xFFFFFFFFu;
  for (u32 i = 0; i < v->num_tris; ++i) {
    const u32* t = v->tris + 3 * i;
    u32 hh = hash3(t[0], t[1], t[2]) & (cap - 1);
    for (;;) {
      if (v->map_vals[hh] == 0xFFFFFFFFu) {
        v->map_keys[3 * hh] = t[0];
        v->map_keys[3 * hh + 1] = t[1];
        v->map_keys[3 * hh + 2] = t[2];
        v->map_vals[hh] = i;
        break;
      }
      if (v->map_keys[3 * hh] == t[0] && v->map_keys[3 * hh + 1] == t[1] && v->map_keys[3 * hh + 2] == t[2]) break;
      hh = (hh + 1) & (cap - 1);
    }
  }
  return 0;
}

static void view_free(orc_view* v) {
  free(v->map_keys);
  free(v->map_vals);
}

static u32 lookup_tri(const orc_view* v, const u32* t) {
  u32 hh = hash3(t[0], t[1], t[2]) & (v->map_cap - 1);
  for (;;) {
    if (v->map_vals[hh] == 0xFFFFFFFFu) return 0xFFFFFFFFu;
    if (v->map_keys[3 * hh] == t[0] && v->map_keys[3 * hh + 1] == t[1] && v->map_keys[3 * hh + 2] == t[2])
      return v->map_vals[hh];
    hh = (hh + 1) & (v->map_cap - 1);
  }
}

/* bvh_getChildId (bvhGetter.glsl:17-25) */
static void get_children(const orc_view* v, u32 nid, u32* c0, u32* c1) {
  u32 packed = v->nodes[16 * nid + 2];
  u32 l = packed & NID_MASK, r = l + 1;
  *c0 = l | ((packed & NID_LEFT_LEAF_BIT) ? NID_LEAF_BIT : 0u);
  *c1 = r | ((packed & NID_RIGHT_LEAF_BIT) ? NID_LEAF_BIT : 0u);
}

/* bvh_getNodeBoxes (bvhGetter.glsl:27-42) */
static void get_boxes(const orc_view* v, u32 nid, float b0mn[3], float b0mx[3], float b1mn[3], float b1mx[3]) {
  const float* f = (const float*)(v->nodes + 16 * nid);
  const float* n0xy = f + 4;
  const float* n1xy = f + 8;
  const float* nz = f + 12;
  b0mn[0] = n0xy[0]; b0mn[1] = n0xy[1]; b0mx[0] = n0xy[2]; b0mx[1] = n0xy[3];
  b1mn[0] = n1xy[0]; b1mn[1] = n1xy[1]; b1mx[0] = n1xy[2]; b1mx[1] = n1xy[3];
  b0mn[2] = nz[0]; b0mx[2] = nz[1]; b1mn[2] = nz[2]; b1mx[2] = nz[3];
}

/* loadTriangleVertices (bvhBindings.glsl:57-62) via getIndex0..2 (primitive_cpp.glsl:121-134) */
static int load_vertices(const orc_view* v, const u32* t, float p[3][3]) {
  u32 idx[3] = {t[0] + (t[1] & 0xFFFFu), t[0] + ((t[1] & 0xFFFF0000u) >> 16), t[0] + (t[2] & 0xFFFFu)};
  for (int k = 0; k < 3; ++k) {
    if (idx[k] >= v->num_vertices) return -1;
    p[k][0] = v->pos[3 * idx[k]];
    p[k][1] = v->pos[3 * idx[k] + 1];
    p[k][2] = v->pos[3 * idx[k] + 2];
  }
  return 0;
}

typedef struct {
  float t;       /* ClosestHit.t */
  u32 slot;      /* leaf-data word index of the accepted inline triangle */
  u32 nid;       /* ClosestHit_storeNid */
  float u, v;
  u32 inner, leaves, tests;
} orc_hit;

/* bvh_collide (bvhCollision.glsl:177-217): inline triangles of the leaf in blob order; an
 * accepted t replaces hit.t with t - OFFSET_RAY_COLLISION. A leaf whose nid.w is 0xFFFFFFFF has
 * no data word; it is treated as holding nothing. */
static void leaf_closest(const orc_view* v, u32 nid, const float* o, const float* d, orc_hit* h) {
  nid &= NID_MASK;
  u32 off = v->nodes[16 * nid + 3];
  h->leaves++;
  if (off == 0xFFFFFFFFu || off >= v->leaf_words) return;
  u32 n = v->leaf[off] & 0xFFFu;
  for (u32 i = 0; i < n; ++i) {
    u32 w = 1 + off + 3 * i;
    float p[3][3], u = 0, vv = 0;
    if (w + 3 > v->leaf_words || load_vertices(v, v->leaf + w, p)) continue;
    h->tests++;
    float t = orc_collide_triangle(o, d, p[0], p[1], p[2], h->t, &u, &vv);
    if (t > 0) {
      h->t = (kPlant == 2) ? t : t - OFFSET_RAY_COLLISION;
      h->slot = w;
      h->nid = nid;
      h->u = u;
      h->v = vv;
    }
  }
}

/* bvh_collide_fast (bvhCollision.glsl:241-284); the BLAS loop is inert (leaf BLAS count 0). */
static int leaf_any(const orc_view* v, u32 nid, const float* o, const float* d, float tmax, orc_hit* h) {
  nid &= NID_MASK;
  u32 off = v->nodes[16 * nid + 3];
  h->leaves++;
  if (off == 0xFFFFFFFFu || off >= v->leaf_words) return 0;
  u32 n = v->leaf[off] & 0xFFFu;
  for (u32 i = 0; i < n; ++i) {
    u32 w = 1 + off + 3 * i;
    float p[3][3];
    if (w + 3 > v->leaf_words || load_vertices(v, v->leaf + w, p)) continue;
    h->tests++;
    if (orc_collide_triangle(o, d, p[0], p[1], p[2], tmax, 0, 0) > 0) return 1;
  }
  return 0;
}

static inline int near_is_child1(float d0, float d1) {
  return kPlant == 1 ? (d1 <= d0) : (d1 < d0);
}

/* traverseBvh / traverseBvhFast (bvhTraversal_inline.glsl:3-78, 80-151): MBVH2 bitstack with
 * parent/sibling backtracking. any = 0: closest hit into h (returns 0); any = 1: returns 1 on the
 * first accepted triangle with t in (1e-7, tmax). */
static int traverse_bitstack(const orc_view* v, const float* o, const float* d, float tmax, int any, orc_hit* h) {
  u32 root = (v->num_nodes == 1) ? NID_LEAF_BIT : 0u;
  u32 node = root, bitstack = 0, parent = 0, sibling = 0;
  for (;;) {
    while (!(node & NID_LEAF_BIT)) {
      h->inner++;
      u32 c0, c1;
      get_children(v, node, &c0, &c1);
      float b0mn[3], b0mx[3], b1mn[3], b1mx[3];
      get_boxes(v, node, b0mn, b0mx, b1mn, b1mx);
      float lim = any ? tmax : h->t;
      float d0 = orc_collide_box(o, d, b0mn, b0mx, lim);
      float d1 = orc_collide_box(o, d, b1mn, b1mx, lim);
      if (d0 < 0 && d1 < 0) break;
      bitstack <<= 1;
      if (d0 >= 0 && d1 >= 0) {
        node = near_is_child1(d0, d1) ? c1 : c0;
        bitstack |= 1u;
      } else {
        node = (d0 >= 0) ? c0 : c1;
      }
    }
    if (node & NID_LEAF_BIT) {
      if (any) {
        if (leaf_any(v, node, o, d, tmax, h)) return 1;
      } else {
        leaf_closest(v, node, o, d, h);
      }
    }
    parent = v->nodes[16 * (node & NID_MASK) + 0];
    sibling = v->nodes[16 * (node & NID_MASK) + 1];
    while ((bitstack & 1u) == 0) {
      if (bitstack == 0) return 0;
      node = parent;
      parent = v->nodes[16 * (node & NID_MASK) + 0];
      sibling = v->nodes[16 * (node & NID_MASK) + 1];
      bitstack >>= 1;
    }
    node = sibling;
    bitstack ^= 1u;
  }
}

/* The same visit order with an explicit stack of deferred far children (pushed when both child
 * boxes are hit, popped without re-testing): the form the device uses. */
static int traverse_stack(const orc_view* v, const float* o, const float* d, float tmax, int any, orc_hit* h) {
  u32 stack[64];
  int sp = 0;
  u32 node = (v->num_nodes == 1) ? NID_LEAF_BIT : 0u;
  for (;;) {
    int haveNode = 1;
    while (!(node & NID_LEAF_BIT)) {
      h->inner++;
      u32 c0, c1;
      get_children(v, node, &c0, &c1);
      float b0mn[3], b0mx[3], b1mn[3], b1mx[3];
      get_boxes(v, node, b0mn, b0mx, b1mn, b1mx);
      float lim = any ? tmax : h->t;
      float d0 = orc_collide_box(o, d, b0mn, b0mx, lim);
      float d1 = orc_collide_box(o, d, b1mn, b1mx, lim);
      if (d0 < 0 && d1 < 0) { haveNode = 0; break; }
      if (d0 >= 0 && d1 >= 0) {
        int n1 = near_is_child1(d0, d1);
        if (sp >= 64) return -1;
        stack[sp++] = n1 ? c0 : c1;
        node = n1 ? c1 : c0;
      } else {
        node = (d0 >= 0) ? c0 : c1;
      }
    }
    if (haveNode) {
      if (any) {
        if (leaf_any(v, node, o, d, tmax, h)) return 1;
      } else {
        leaf_closest(v, node, o, d, h);
      }
    }
    if (sp == 0) return 0;
    node = stack[--sp];
  }
}

/* ---------------------------------------------------------------------------------------------
 * Two-level any hit: traverseTlasFast (bvhTraversal_inline.glsl:80-151 compiled with TLAS_COLLIDE)
 * over the top-level tree, whose leaves run tlas_collide_fast (bvhCollision.glsl:321-349): for each
 * BLAS index listed after the leaf's triangles, CollideBox against its BlasHeader box with the fixed
 * t limit, and if that is > 0 (strictly: a ray starting inside the box gets 0 and skips the BLAS)
 * traverseBvhFast from the header's rootIndex, whose leaves run bvh_collide_fast. The top-level
 * leaf's own triangles are not tested, nor are the instances hoisted into inner nodes (only leaves
 * call tlas_collide_fast). The top-level root is node 0, with NID_LEAF_BIT when it has no children
 * (the shader's numNodes == 1 case). A BLAS whose root has no children is walked as a leaf; the
 * shader would descend into it from rootIndex, which carries no leaf bit, and read past the node
 * array. */
static int walk_any(const orc_view* v, u32 root, const float* o, const float* d, float tmax, int top);

static int tlas_leaf_any(const orc_view* v, u32 nid, const float* o, const float* d, float tmax) {
  orc_hit h;
  memset(&h, 0, sizeof(h));
  nid &= NID_MASK;
  u32 off = v->nodes[16 * nid + 3];
  if (off == 0xFFFFFFFFu || off >= v->leaf_words) return 0;
  u32 packed = v->leaf[off], ntri = packed & 0xFFFu, nb = (packed >> 12) & 0xFFu;
  if (kPlant == 10 && leaf_any(v, nid, o, d, tmax, &h)) return 1;
  for (u32 i = 0; i < nb; ++i) {
    u32 w = 1 + off + 3 * ntri + i;
    if (w >= v->leaf_words) return 0;
    u32 b = v->leaf[w];
    if (b >= v->num_blas) continue;
    const float* hb = (const float*)(v->blas + 32 * (size_t)b);
    u32 root;
    memcpy(&root, v->blas + 32 * (size_t)b + 28, 4);
    float box = orc_collide_box(o, d, hb, hb + 4, tmax);
    if (kPlant == 9 ? box >= 0.0f : box > 0.0f) {
      if (root >= v->num_nodes) continue;
      u32 start = ((v->nodes[16 * root + 2] & NID_MASK) == NID_MASK) ? (root | NID_LEAF_BIT) : root;
      if (walk_any(v, start, o, d, tmax, 0)) return 1;
    }
  }
  return 0;
}

static int walk_any(const orc_view* v, u32 root, const float* o, const float* d, float tmax, int top) {
  orc_hit h;
  memset(&h, 0, sizeof(h));
  u32 node = root, bitstack = 0, parent = 0, sibling = 0;
  for (;;) {
    while (!(node & NID_LEAF_BIT)) {
      u32 c0, c1;
      get_children(v, node, &c0, &c1);
      float b0mn[3], b0mx[3], b1mn[3], b1mx[3];
      get_boxes(v, node, b0mn, b0mx, b1mn, b1mx);
      float d0 = orc_collide_box(o, d, b0mn, b0mx, tmax);
      float d1 = orc_collide_box(o, d, b1mn, b1mx, tmax);
      if (d0 < 0 && d1 < 0) break;
      bitstack <<= 1;
      if (d0 >= 0 && d1 >= 0) {
        node = near_is_child1(d0, d1) ? c1 : c0;
        bitstack |= 1u;
      } else {
        node = (d0 >= 0) ? c0 : c1;
      }
    }
    if (node & NID_LEAF_BIT) {
      if (top ? tlas_leaf_any(v, node, o, d, tmax) : leaf_any(v, node, o, d, tmax, &h)) return 1;
    }
    parent = v->nodes[16 * (node & NID_MASK) + 0];
    sibling = v->nodes[16 * (node & NID_MASK) + 1];
    while ((bitstack & 1u) == 0) {
      if (bitstack == 0) return 0;
      node = parent;
      parent = v->nodes[16 * (node & NID_MASK) + 0];
      sibling = v->nodes[16 * (node & NID_MASK) + 1];
      bitstack >>= 1;
    }
    node = sibling;
    bitstack ^= 1u;
  }
}

typedef struct {
  const orc_view* v;
  const float* rays;
  u32* occluded;
  u32 begin, end;
} orc_tlas_job;

static void* run_tlas(void* arg) {
  orc_tlas_job* j = (orc_tlas_job*)arg;
  const orc_view* v = j->v;
  u32 root = ((v->nodes[2] & NID_MASK) == NID_MASK) ? NID_LEAF_BIT : 0u;
  for (u32 i = j->begin; i < j->end; ++i) {
    const float* r = j->rays + 8 * (size_t)i;
    j->occluded[i] = walk_any(v, root, r, r + 4, r[3], 1) ? 1u : 0u;
  }
  return 0;
}

/* Two-level any hit for n rays (t limit = the ray's tmax). */
int orc_trace_any_tlas(const uint8_t* blob, size_t bytes, const u32* ranges, const float* pos, u32 nv,
                       const float* rays, u32 n, u32* occluded, int threads) {
  orc_view view;
  /* a blob is two-level when some leaf lists an instance (a one-level blob still carries a 32-byte
   * placeholder BLAS section) */
  int two_level = 0;
  if (view_init(&view, blob, bytes, ranges, pos, nv) == 0 && view.blas && view.num_nodes) {
    for (u32 i = 0; i < view.num_nodes && !two_level; ++i) {
      u32 off = view.nodes[16 * i + 3];
      if ((view.nodes[16 * i + 2] & NID_MASK) == NID_MASK && off < view.leaf_words && ((view.leaf[off] >> 12) & 0xFFu))
        two_level = 1;
    }
  }
  if (!two_level) {
    view_free(&view);
    return -1;
  }
  if (threads < 1) threads = 1;
  if (threads > 256) threads = 256;
  if ((u32)threads > n) threads = n ? (int)n : 1;
  pthread_t th[256];
  orc_tlas_job jobs[256];
  for (int t = 0; t < threads; ++t) {
    orc_tlas_job jj = {&view, rays, occluded, (u32)((uint64_t)n * t / threads), (u32)((uint64_t)n * (t + 1) / threads)};
    jobs[t] = jj;
  }
  for (int t = 1; t < threads; ++t) pthread_create(&th[t], 0, run_tlas, &jobs[t]);
  run_tlas(&jobs[0]);
  for (int t = 1; t < threads; ++t) pthread_join(th[t], 0);
  view_free(&view);
  return 0;
}

/* ---------------------------------------------------------------------------------------------
 * Batch entry points. Rays are rtq_ray records (8 words: org.xyz, tmax, dir.xyz, pad). */
typedef struct {
  const orc_view* v;
  const float* rays;
  u32 begin, end;
  int mode;      /* 0 closest, 1 any */
  int form;      /* 0 bitstack, 1 stack */
  u32* tri_id;
  float* t;
  float* u;
  float* vv;
  u32* nid;
  u32* slot;
  u32* counters;
  u32* occluded;
} orc_job;

static void* run_job(void* arg) {
  orc_job* j = (orc_job*)arg;
  for (u32 i = j->begin; i < j->end; ++i) {
    const float* r = j->rays + 8 * (size_t)i;
    const float* o = r;
    float tmax = r[3];
    const float* d = r + 4;
    orc_hit h;
    memset(&h, 0, sizeof(h));
    h.t = tmax;
    h.slot = 0xFFFFFFFFu;
    h.nid = 0xFFFFFFFFu;
    int res = j->form ? traverse_stack(j->v, o, d, tmax, j->mode, &h)
                      : traverse_bitstack(j->v, o, d, tmax, j->mode, &h);
    if (j->mode == 0) {
      u32 id = 0xFFFFFFFFu;
      if (h.slot != 0xFFFFFFFFu) id = lookup_tri(j->v, j->v->leaf + h.slot);
      if (j->tri_id) j->tri_id[i] = id;
      if (j->t) j->t[i] = h.t;
      if (j->u) j->u[i] = h.slot != 0xFFFFFFFFu ? h.u : 0.0f;
      if (j->vv) j->vv[i] = h.slot != 0xFFFFFFFFu ? h.v : 0.0f;
      if (j->nid) j->nid[i] = h.nid;
      if (j->slot) j->slot[i] = h.slot;
    } else {
      if (j->occluded) j->occluded[i] = res == 1 ? 1u : 0u;
    }
    if (j->counters) {
      j->counters[3 * (size_t)i] = h.inner;
      j->counters[3 * (size_t)i + 1] = h.leaves;
      j->counters[3 * (size_t)i + 2] = h.tests;
    }
  }
  return 0;
}

static int run_parallel(orc_job* proto, u32 n, int threads) {
  if (threads < 1) threads = 1;
  if (threads > 256) threads = 256;
  if ((u32)threads > n) threads = n ? (int)n : 1;
  pthread_t th[256];
  orc_job jobs[256];
  for (int t = 0; t < threads; ++t) {
    jobs[t] = *proto;
    jobs[t].begin = (u32)((uint64_t)n * t / threads);
    jobs[t].end = (u32)((uint64_t)n * (t + 1) / threads);
  }
  for (int t = 1; t < threads; ++t) pthread_create(&th[t], 0, run_job, &jobs[t]);
  run_job(&jobs[0]);
  for (int t = 1; t < threads; ++t) pthread_join(th[t], 0);
  return 0;
}

/* Closest hit for n rays. Outputs may be NULL. form: 0 = bitstack (reference form), 1 = stack. */
int orc_trace_closest(const uint8_t* blob, size_t bytes, const u32* ranges, const float* pos, u32 nv,
                      const float* rays, u32 n, u32* tri_id, float* t, float* u, float* v, u32* nid, u32* slot,
                      u32* counters, int form, int threads) {
  orc_view view;
  if (view_init(&view, blob, bytes, ranges, pos, nv)) { view_free(&view); return -1; }
  orc_job j;
  memset(&j, 0, sizeof(j));
  j.v = &view; j.rays = rays; j.mode = 0; j.form = form;
  j.tri_id = tri_id; j.t = t; j.u = u; j.vv = v; j.nid = nid; j.slot = slot; j.counters = counters;
  run_parallel(&j, n, threads);
  view_free(&view);
  return 0;
}

/* Any hit for n rays (t limit = the ray's tmax). */
int orc_trace_any(const uint8_t* blob, size_t bytes, const u32* ranges, const float* pos, u32 nv, const float* rays,
                  u32 n, u32* occluded, u32* counters, int form, int threads) {
  orc_view view;
  if (view_init(&view, blob, bytes, ranges, pos, nv)) { view_free(&view); return -1; }
  orc_job j;
  memset(&j, 0, sizeof(j));
  j.v = &view; j.rays = rays; j.mode = 1; j.form = form; j.occluded = occluded; j.counters = counters;
  run_parallel(&j, n, threads);
  view_free(&view);
  return 0;
}

/* Brute force over every triangle of the blob's triangle section in index order, fp32
 * CollideTriangle with the ray's tmax and no shrink: the minimum t and the first index reaching
 * it, plus how many triangles hit within `window` of that minimum. */
int orc_brute_closest(const uint8_t* blob, size_t bytes, const u32* ranges, const float* pos, u32 nv,
                      const float* rays, u32 n, float window, u32* tri_id, float* t, u32* near_count) {
  orc_view view;
  if (view_init(&view, blob, bytes, ranges, pos, nv)) { view_free(&view); return -1; }
  for (u32 i = 0; i < n; ++i) {
    const float* r = rays + 8 * (size_t)i;
    float best = r[3];
    u32 bid = 0xFFFFFFFFu;
    for (u32 k = 0; k < view.num_tris; ++k) {
      float p[3][3];
      if (load_vertices(&view, view.tris + 3 * k, p)) continue;
      float tt = orc_collide_triangle(r, r + 4, p[0], p[1], p[2], r[3], 0, 0);
      if (tt > 0 && tt < best) { best = tt; bid = k; }
    }
    u32 cnt = 0;
    if (bid != 0xFFFFFFFFu) {
      for (u32 k = 0; k < view.num_tris; ++k) {
        float p[3][3];
        if (load_vertices(&view, view.tris + 3 * k, p)) continue;
        float tt = orc_collide_triangle(r, r + 4, p[0], p[1], p[2], r[3], 0, 0);
        if (tt > 0 && tt <= best + window) cnt++;
      }
    }
    tri_id[i] = bid;
    t[i] = best;
    near_count[i] = cnt;
  }
  view_free(&view);
  return 0;
}

/* ---------------------------------------------------------------------------------------------
 * computeShadowMask (Shaders/bvh/bvhLighting.glsl:143-163) for closest hits with t < TMAX.
 * evalShadow (:110-131): sphere lights only; rays toward the light are normalised by their length
 * and limited to it. Lights whose bit L+1 would not fit 32 bits are skipped. */
typedef struct {
  const orc_view* v;
  const float* rays;
  const float* hit_t;
  const u32* hit_nid;
  const float* sun;
  u32 flags;
  u32* masks;
  u32 begin, end;
} orc_shadow_job;

/* FP64 intersection of each ray with its recorded closest-hit triangle, for error statements of the
 * fp32 path: for ray i with tri_id[i] != 0xFFFFFFFF, orc_intersect_f64 of that triangle (vertices
 * through the blob's Triangle record, loadTriangleVertices); NaN for misses. Returns -1 on a bad blob,
 * a triangle id beyond the triangle range or a vertex beyond the positions. */
int orc_hits_f64(const uint8_t* blob, size_t bytes, const u32* ranges, const float* pos, u32 nv, const float* rays,
                 u32 n, const u32* tri_id, double* t, double* u, double* v) {
  orc_view view;
  if (view_init(&view, blob, bytes, ranges, pos, nv)) { view_free(&view); return -1; }
  int rc = 0;
  for (u32 i = 0; i < n; ++i) {
    t[i] = u[i] = v[i] = NAN;
    if (tri_id[i] == 0xFFFFFFFFu) continue;
    float p[3][3];
    if (tri_id[i] >= view.num_tris || load_vertices(&view, view.tris + 3 * (size_t)tri_id[i], p)) { rc = -1; break; }
    const float* r = rays + 8 * (size_t)i;
    if (!orc_intersect_f64(r, r + 4, p[0], p[1], p[2], t + i, u + i, v + i)) t[i] = u[i] = v[i] = NAN;
  }
  view_free(&view);
  return rc;
}

static int any_hit(const orc_view* v, const float* o, const float* d, float tmax) {
  orc_hit h;
  memset(&h, 0, sizeof(h));
  return traverse_bitstack(v, o, d, tmax, 1, &h);
}

/* computeShadowMask at hit position pos of the hit leaf nid. */
static u32 shadow_mask_at(const orc_view* v, const float pos[3], u32 hit_nid, const float* sun, u32 flags) {
  u32 mask = 0;
  if (flags & 1u) {
    float sd[3] = {-sun[0], -sun[1], -sun[2]};
    mask |= any_hit(v, pos, sd, TMAX) ? 1u : 0u;
  }
  u32 nid = hit_nid & NID_MASK;
  if (nid >= v->num_nodes) return mask;
  u32 off = v->nodes[16 * nid + 3];
  if (off != 0xFFFFFFFFu && off < v->leaf_words) {
    u32 packed = v->leaf[off];
    /* unpackLightCountAndOffset (:133-141) with NodeTriangleStride 3 */
    u32 lo = 3u * (packed & 0xFFFu) + ((packed >> 12) & 0xFFu) + ((packed >> 20) & 0xFFu);
    u32 count = (packed >> 28) & 0xFu;
    lo += 1u + off;
    for (u32 k = 0; k < count; ++k) {
      if (lo + k >= v->leaf_words) break;
      u32 li = v->leaf[lo + k];
      if (li + 1u >= 32u || li >= v->num_lights) continue;
      const float* L = v->lights + 14 * (size_t)li;
      if (((const u32*)L)[0] != 1u) continue; /* Light_Sphere */
      float lp[3] = {L[1], L[2], L[3]};
      float radius = L[4];
      float dir[3] = {lp[0] - pos[0], lp[1] - pos[1], lp[2] - pos[2]};
      float len = sqrtf((dir[0] * dir[0] + dir[1] * dir[1]) + dir[2] * dir[2]);
      int occ;
      if (len < radius) {
        float nd[3] = {dir[0] / len, dir[1] / len, dir[2] / len};
        occ = any_hit(v, pos, nd, len);
      } else {
        occ = 1;
      }
      if (occ) mask |= 1u << (li + 1u);
    }
  }
  return mask;
}

static void* run_shadow(void* arg) {
  orc_shadow_job* j = (orc_shadow_job*)arg;
  const orc_view* v = j->v;
  for (u32 i = j->begin; i < j->end; ++i) {
    const float* r = j->rays + 8 * (size_t)i;
    float t = j->hit_t[i];
    u32 mask = 0;
    if (t < TMAX && j->hit_nid[i] != 0xFFFFFFFFu) {
      float pos[3] = {r[0] + r[4] * t, r[1] + r[5] * t, r[2] + r[6] * t};
      mask = shadow_mask_at(v, pos, j->hit_nid[i], j->sun, j->flags);
    }
    j->masks[i] = mask;
  }
  return 0;
}

int orc_shadow_masks(const uint8_t* blob, size_t bytes, const u32* ranges, const float* pos, u32 nv,
                     const float* rays, const float* hit_t, const u32* hit_nid, u32 n, const float* sun_dir,
                     u32 flags, u32* masks, int threads) {
  orc_view view;
  if (view_init(&view, blob, bytes, ranges, pos, nv)) { view_free(&view); return -1; }
  if (threads < 1) threads = 1;
  if (threads > 256) threads = 256;
  pthread_t th[256];
  orc_shadow_job jobs[256];
  for (int t = 0; t < threads; ++t) {
    orc_shadow_job jj = {&view, rays, hit_t, hit_nid, sun_dir, flags, masks,
                         (u32)((uint64_t)n * t / threads), (u32)((uint64_t)n * (t + 1) / threads)};
    jobs[t] = jj;
  }
  for (int t = 1; t < threads; ++t) pthread_create(&th[t], 0, run_shadow, &jobs[t]);
  run_shadow(&jobs[0]);
  for (int t = 1; t < threads; ++t) pthread_join(th[t], 0);
  view_free(&view);
  return 0;
}

/* ---------------------------------------------------------------------------------------------
 * Primary rays of the camera pass (Shaders/cameraPass.comp:46-54): stepW = (c10 - c00) * invW,
 * invW = 1.f / W (raytracingPass.cpp:47), target = c00 + stepW * gx + stepH * gy,
 * dir = normalize(target - cam) with normalize(v) = v / sqrt(dot(v, v)); t limit TMAX.
 * reference_rows = 1: the ray for stored row y (>= 1) is generated from gy = 16*ceil(H/16) - y;
 * row 0 gets tmax = -1 (the reference never writes it). Output: rtq_ray records, y*W + x. */
void orc_camera_rays(const float* frustum, u32 width, u32 height, int reference_rows, float* out) {
  const float* cam = frustum;
  const float* c00 = frustum + 3;
  const float* c10 = frustum + 6;
  const float* c01 = frustum + 9;
  const float invW = 1.0f / (float)width, invH = 1.0f / (float)height;
  for (u32 y = 0; y < height; ++y)
    for (u32 x = 0; x < width; ++x) {
      float* r = out + 8 * ((size_t)y * width + x);
      u32 gy = y;
      if (reference_rows) gy = 16u * ((height + 15u) / 16u) - y - (kPlant == 8 ? 1u : 0u);
      float d[3];
      for (int k = 0; k < 3; ++k) {
        float sw = (c10[k] - c00[k]) * invW;
        float sh = (c01[k] - c00[k]) * invH;
        float target = (c00[k] + sw * (float)x) + sh * (float)gy;
        d[k] = target - cam[k];
      }
      float len = sqrtf((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
      r[0] = cam[0];
      r[1] = cam[1];
      r[2] = cam[2];
      r[3] = (reference_rows && y == 0) ? -1.0f : TMAX;
      r[4] = d[0] / len;
      r[5] = d[1] / len;
      r[6] = d[2] / len;
      r[7] = 0.0f;
    }
}

/* ---------------------------------------------------------------------------------------------
 * computeProbMask (Shaders/bvh/bvhLighting.glsl:165-184): the probe cell around pos from
 * getLightProbCoordUVW (lightprob/lightprobHelpers.glsl:70-76) with step = computeLpfStep
 * (:10-14), cell corner = min(int(uvw), resolution - 2), and 8 any-hit segments of t limit 1 toward
 * getLightProbPosition (:62-68) of each corner, the direction left unnormalised (createShadowRay,
 * bvhLighting.glsl:87-98). int() of NaN is undefined in GLSL; it is fixed to 0 here. */
static u32 prob_mask_at(const orc_view* v, const float pos[3], const float* lpf_box, const u32* res) {
  const float* mn = lpf_box;
  const float* mx = lpf_box + 3;
  u32 coord[3];
  for (int k = 0; k < 3; ++k) {
    float step = (mx[k] - mn[k]) / (float)res[k];
    float uvw = ((pos[k] - mn[k]) / step) - (kPlant == 4 ? 0.0f : 0.5f);
    uvw = gmin(gmax(uvw, 0.0f), (float)res[k] - 1.0f);
    int c = (uvw == uvw) ? (int)uvw : 0;
    u32 lim = kPlant == 5 ? res[k] - 1u : res[k] - 2u;
    coord[k] = ((u32)c < lim) ? (u32)c : lim;
  }
  u32 mask = 0;
  for (u32 i = 0; i < 8; ++i) {
    u32 off[3] = {i & 1u, (i >> 1) & 1u, (i >> 2) & 1u};
    float d[3];
    for (int k = 0; k < 3; ++k) {
      float step = (mx[k] - mn[k]) / (float)res[k];
      float probe = mn[k] + step * ((float)(coord[k] + off[k]) + 0.5f);
      d[k] = probe - pos[k];
    }
    if (any_hit(v, pos, d, 1.0f)) mask |= 1u << i;
  }
  return mask;
}

/* The camera pass's TRACING_STEP (Shaders/cameraPass.comp:69-75): per pixel 8 words,
 * ClosestHit_getHitData (bvhBindings.glsl:85-92) = the hit leaf's inline triangle triple and
 * floatBits(t), then computeAdditionalHitData (baseRaytracingPass.glsl:85-97) = shadow and probe
 * masks when t < TMAX. A miss writes zero triangle words (the shader's are stale) and zero masks.
 * lpf_box = (min.xyz, max.xyz) and lpf_res, or NULL for no probe mask. Row 0 is left untouched
 * when reference_rows is set. */
typedef struct {
  const orc_view* v;
  const float* rays;
  u32 width, height;
  int reference_rows;
  const float* sun;
  u32 shadow_flags;
  const float* lpf_box;
  const u32* lpf_res;
  u32* records;
  u32 begin, end;
} orc_record_job;

static void* run_record(void* arg) {
  orc_record_job* j = (orc_record_job*)arg;
  const orc_view* v = j->v;
  for (u32 i = j->begin; i < j->end; ++i) {
    if (j->reference_rows && i < j->width) continue;
    const float* r = j->rays + 8 * (size_t)i;
    orc_hit h;
    memset(&h, 0, sizeof(h));
    h.t = TMAX;
    h.slot = 0xFFFFFFFFu;
    h.nid = 0xFFFFFFFFu;
    traverse_bitstack(v, r, r + 4, TMAX, 0, &h);
    u32* out = j->records + 8 * (size_t)i;
    memset(out, 0, 8 * sizeof(u32));
    if (h.slot != 0xFFFFFFFFu) {
      out[0] = v->leaf[h.slot];
      out[1] = v->leaf[h.slot + 1];
      out[2] = v->leaf[h.slot + 2];
    }
    memcpy(&out[3], &h.t, 4);
    if (h.t < TMAX && h.nid != 0xFFFFFFFFu) {
      float pos[3] = {r[0] + r[4] * h.t, r[1] + r[5] * h.t, r[2] + r[6] * h.t};
      out[4] = shadow_mask_at(v, pos, h.nid, j->sun, j->shadow_flags);
      if (j->lpf_box) out[5] = prob_mask_at(v, pos, j->lpf_box, j->lpf_res);
    }
  }
  return 0;
}

int orc_trace_record(const uint8_t* blob, size_t bytes, const u32* ranges, const float* pos, u32 nv,
                     const float* frustum, u32 width, u32 height, int reference_rows, const float* sun_dir,
                     u32 shadow_flags, const float* lpf_box, const u32* lpf_res, u32* records, int threads) {
  orc_view view;
  if (view_init(&view, blob, bytes, ranges, pos, nv)) { view_free(&view); return -1; }
  size_t n = (size_t)width * height;
  float* rays = (float*)malloc(sizeof(float) * 8 * (n ? n : 1));
  if (!rays) { view_free(&view); return -2; }
  orc_camera_rays(frustum, width, height, reference_rows, rays);
  if (threads < 1) threads = 1;
  if (threads > 256) threads = 256;
  pthread_t th[256];
  orc_record_job jobs[256];
  for (int t = 0; t < threads; ++t) {
    orc_record_job jj = {&view, rays, width, height, reference_rows, sun_dir, shadow_flags, lpf_box, lpf_res,
                         records, (u32)(n * t / threads), (u32)(n * (t + 1) / threads)};
    jobs[t] = jj;
  }
  for (int t = 1; t < threads; ++t) pthread_create(&th[t], 0, run_record, &jobs[t]);
  run_record(&jobs[0]);
  for (int t = 1; t < threads; ++t) pthread_join(th[t], 0);
  free(rays);
  view_free(&view);
  return 0;
}

/* ---------------------------------------------------------------------------------------------
 * The light-probe field's TRACING_STEP (Shaders/genLightProbField.comp:44-69): probe index ->
 * coordinate by getLightProbCoord with LP_TILE_RES 4 (lightprob/lightprobHelpers.glsl:35-52),
 * position by getLightProbPosition (:62-68), one closest-hit ray per (probe, direction) from that
 * position, the record as in orc_trace_record with the probe mask 0 (the pass has no USE_LPF_MASK).
 * Sample s = probe * 64 + direction. */
void orc_probe_position(const float* lpf_box, const u32* res, u32 index, float out[3]) {
  u32 tres[2] = {res[0] / 4u, res[1] / 4u};
  u32 tile = index / 64u, in_tile = index % 64u;
  u32 c[3];
  c[2] = tile / (tres[0] * tres[1]);
  tile = tile % (tres[0] * tres[1]);
  c[1] = tile / tres[0];
  c[0] = tile % tres[0];
  c[0] = c[0] * 4u + in_tile % 4u;
  c[1] = c[1] * 4u + (in_tile / 4u) % 4u;
  c[2] = c[2] * 4u + (in_tile / 16u) % 4u;
  for (int k = 0; k < 3; ++k) {
    float dim = lpf_box[3 + k] - lpf_box[k];
    float step = dim / (float)res[k];
    out[k] = lpf_box[k] + step * ((float)c[k] + (kPlant == 6 ? 0.0f : 0.5f));
  }
}

typedef struct {
  const orc_view* v;
  const float* lpf_box;
  const u32* lpf_res;
  const float* dirs;
  const float* sun;
  u32 shadow_flags;
  u32* records;
  u32 begin, end; /* probe range */
} orc_probe_job;

static void* run_probes(void* arg) {
  orc_probe_job* j = (orc_probe_job*)arg;
  const orc_view* v = j->v;
  for (u32 p = j->begin; p < j->end; ++p) {
    float o[3];
    orc_probe_position(j->lpf_box, j->lpf_res, p, o);
    for (u32 k = 0; k < 64; ++k) {
      const float* d = j->dirs + 4 * k;
      orc_hit h;
      memset(&h, 0, sizeof(h));
      h.t = TMAX;
      h.slot = 0xFFFFFFFFu;
      h.nid = 0xFFFFFFFFu;
      traverse_bitstack(v, o, d, TMAX, 0, &h);
      u32* out = j->records + 8 * ((size_t)p * 64 + k);
      memset(out, 0, 8 * sizeof(u32));
      if (h.slot != 0xFFFFFFFFu) {
        out[0] = v->leaf[h.slot];
        out[1] = v->leaf[h.slot + 1];
        out[2] = v->leaf[h.slot + 2];
      }
      memcpy(&out[3], &h.t, 4);
      if (h.t < TMAX && h.nid != 0xFFFFFFFFu) {
        float pos[3] = {o[0] + d[0] * h.t, o[1] + d[1] * h.t, o[2] + d[2] * h.t};
        out[4] = shadow_mask_at(v, pos, h.nid, j->sun, j->shadow_flags);
      }
    }
  }
  return 0;
}

int orc_trace_probes(const uint8_t* blob, size_t bytes, const u32* ranges, const float* pos, u32 nv,
                     const float* lpf_box, const u32* lpf_res, const float* dirs, const float* sun_dir,
                     u32 shadow_flags, u32* records, int threads) {
  if (lpf_res[0] < 4 || lpf_res[1] < 4) return -3;
  orc_view view;
  if (view_init(&view, blob, bytes, ranges, pos, nv)) { view_free(&view); return -1; }
  u32 n = lpf_res[0] * lpf_res[1] * lpf_res[2];
  if (threads < 1) threads = 1;
  if (threads > 256) threads = 256;
  pthread_t th[256];
  orc_probe_job jobs[256];
  for (int t = 0; t < threads; ++t) {
    orc_probe_job jj = {&view, lpf_box, lpf_res, dirs, sun_dir, shadow_flags, records,
                        (u32)((uint64_t)n * t / threads), (u32)((uint64_t)n * (t + 1) / threads)};
    jobs[t] = jj;
  }
  for (int t = 1; t < threads; ++t) pthread_create(&th[t], 0, run_probes, &jobs[t]);
  run_probes(&jobs[0]);
  for (int t = 1; t < threads; ++t) pthread_join(th[t], 0);
  view_free(&view);
  return 0;
}

/* ---------------------------------------------------------------------------------------------
 * Tile frustum light culling: the light part of cullWithFrustumTile (Shaders/cullFrustum.glsl:
 * 63-100, 136-153) with computePlan (:9-13) and sphereFrustum4Collision (collision.glsl:219-227).
 * normalize(v) = v / sqrt(dot(v, v)). Output: full ascending list per tile (CSR). */
static void plane(const float* p1, const float* p2, const float* p3, float out[4]) {
  v3 a = {p1[0] - p2[0], p1[1] - p2[1], p1[2] - p2[2]};
  v3 b = {p3[0] - p2[0], p3[1] - p2[1], p3[2] - p2[2]};
  v3 c = vcross(a, b);
  float len = sqrtf(vdot(c, c));
  v3 n = {c.x / len, c.y / len, c.z / len};
  v3 P2 = {p2[0], p2[1], p2[2]};
  out[0] = n.x;
  out[1] = n.y;
  out[2] = n.z;
  out[3] = -vdot(n, P2);
}

void orc_tile_planes(const float* frustum, u32 width, u32 height, u32 tx, u32 ty, float planes[16]) {
  const float* cam = frustum;
  const float* c00 = frustum + 3;
  const float* c10 = frustum + 6;
  const float* c01 = frustum + 9;
  u32 ngx = (width + 15u) / 16u, ngy = (height + 15u) / 16u;
  float sx[3], sy[3], bl[3], br[3], tl[3], tr[3];
  float fx = (float)tx, fy = (float)ty, fx1 = (float)(tx + 1u), fy1 = (float)(ty + 1u);
  for (int k = 0; k < 3; ++k) {
    sx[k] = (c10[k] - c00[k]) / (float)ngx;
    sy[k] = (c01[k] - c00[k]) / (float)ngy;
  }
  for (int k = 0; k < 3; ++k) {
    bl[k] = (c00[k] + sx[k] * fx) + sy[k] * fy;
    br[k] = (c00[k] + sx[k] * fx1) + sy[k] * fy;
    tl[k] = (c00[k] + sx[k] * fx) + sy[k] * fy1;
    tr[k] = (c00[k] + sx[k] * fx1) + sy[k] * fy1;
  }
  for (int k = 0; k < 3; ++k) {
    float dx = sx[k] * (kPlant == 7 ? 0.0f : 0.25f), dy = sy[k] * (kPlant == 7 ? 0.0f : 0.25f);
    bl[k] = (bl[k] - dx) - dy;
    br[k] = (br[k] + dx) - dy;
    tl[k] = (tl[k] - dx) + dy;
    tr[k] = (tr[k] + dx) + dy;
  }
  plane(cam, tl, bl, planes + 0);
  plane(cam, br, tr, planes + 4);
  plane(cam, tr, tl, planes + 8);
  plane(cam, bl, br, planes + 12);
}

/* lights: PackedLight records (14 words). Returns the total list length; lists beyond capacity
 * are not written. */
uint64_t orc_cull_lights(const float* lights, u32 num_lights, const float* frustum, u32 width, u32 height,
                         u32* offsets, u32* lists, uint64_t capacity) {
  u32 ngx = (width + 15u) / 16u, ngy = (height + 15u) / 16u;
  uint64_t total = 0;
  for (u32 ty = 0; ty < ngy; ++ty)
    for (u32 tx = 0; tx < ngx; ++tx) {
      float pl[16];
      orc_tile_planes(frustum, width, height, tx, ty, pl);
      offsets[ty * ngx + tx] = (u32)total;
      for (u32 i = 0; i < num_lights; ++i) {
        const float* L = lights + 14 * (size_t)i;
        v3 c = {L[1], L[2], L[3]};
        float r = L[4];
        int inside = 1;
        for (int p = 0; p < 4; ++p) {
          v3 n = {pl[4 * p], pl[4 * p + 1], pl[4 * p + 2]};
          if (vdot(n, c) + pl[4 * p + 3] < -r) { inside = 0; break; }
        }
        if (inside) {
          if (total < capacity) lists[total] = i;
          total++;
        }
      }
    }
  offsets[ngx * ngy] = (u32)total;
  return total;
}
