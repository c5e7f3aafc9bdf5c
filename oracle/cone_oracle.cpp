// TEST INFRASTRUCTURE ONLY (see oracle_math.h): a CPU restatement of the
// reference's cone-visibility path, used as the checker by tests/, smoke()
// and bench.py's cpu_baseline leg. Never linked by the product.
//
// Every function follows the reference's GPU semantics, the WGSL in
// crates/continuum-gpu/src/mask_cones.wgsl (cited as "WGSL :line"), which
// differ in documented ways from the CPU model in crates/continuum/src/cone3.rs
// (restated separately in cone3_oracle.cpp). One query walks alone (no
// packets), exactly in the order the WGSL walks.
#include <algorithm>
#include <array>
#include <atomic>
#include <cstring>
#include <thread>
#include <vector>

#include "oracle_math.h"

namespace orc {

constexpr uint32_t NONE = 0xffffffffu;
constexpr uint32_t GRID = 8;
constexpr uint32_t BINS = 64;
constexpr uint32_t STACK = 48;  // WGSL :111
constexpr float NEAR = 1e-4f;   // WGSL :211

// WGSL constant expressions are evaluated in abstract (f64) precision and
// rounded to f32 once.
constexpr float ONE_MINUS_1EM5 = (float)(1.0 - 1e-5);

struct Node {
  float lo[4];
  float hi[4];
  uint32_t info[4];
};

// The scene as the reference's `flatten`/`mesh_hierarchy` lay it out
// (crates/continuum-gpu/src/hierarchy.rs:10-66, lib.rs:3361-3440).
struct Scene {
  const float* tris;  // 3 vec4 per triangle
  const uint32_t* ranges;  // per object: first, count, blas root, xform
  uint32_t n_objects;
  const Node* blas;
  const uint32_t* blas_items;
  const float* blas_cov;  // 27 per blas node
  const Node* nodes;      // object hierarchy
  const uint32_t* items;
  const float* xforms;  // 16 per transform
};

struct Options {
  uint32_t hierarchy;       // 0 every triangle, 1 object BVH, 2 + triangle BVHs
  float prefilter;          // bins (0: exact)
  float wide_extent;        // WGSL :1047 WIDE_EXTENT
  uint32_t light_cone_cap;  // WGSL :1051
  uint32_t wide_cone_cap;   // WGSL :1055
  uint32_t profile_stage;   // WGSL :21-26
  uint32_t object_cov;      // light.objects.x: first object node's coverage in blas_cov (NONE: none)
};

struct Light {
  float corner[4];
  float edge_u[4];
  float edge_v[4];
  float radiance[4];
};

struct Cone {
  V3 apex, right, up, forward;
  float window[4];  // u0, v0, u1, v1
  V3 far_normal;
  float far_offset;
  V3 horizon;
  float bound[4];
  float bound_sin;
};

// The per-invocation private state the WGSL keeps in module-scope `var<private>`s.
struct Query {
  float prefilter_bins = 0.0f;
  float tri_weight = 1.0f;
  // WGSL :583-586: a transmissive object's factor (0: none) and whether it
  // is a volume (light passes two surfaces).
  float tri_transmit = 0.0f;
  bool tri_volume = false;
  uint32_t cone_work = 0;
  uint32_t cone_cap = NONE;
  bool wide_clipped = false;
  bool wide_window = false;
  bool stack_overflow = false;
  // WGSL :587-591: a triangle the cone does not see (object, triangle), and
  // that triangle while the walk is in its object.
  uint32_t skip_obj = NONE, skip_tri = NONE, cover_skip = NONE;
  // WGSL :653 `independent_occluders`: sky cones combine meshes as
  // independent occluders, 1 - (1 - a)(1 - b), and walk children in a fixed
  // order (:1195-1201); light cones take the maximum, nearest child first.
  bool independent = false;
  uint32_t stack[STACK];
  const Scene* s;
  const Options* o;
};

static inline V3 xyz(const float* p) { return V3{p[0], p[1], p[2]}; }

// ---- transforms (WGSL :58-81)
static V3 rotate_to_object(const Scene& s, uint32_t x, V3 d) {
  const float* r = s.xforms + 16 * x;
  return xyz(r) * d.x + xyz(r + 4) * d.y + xyz(r + 8) * d.z;
}
static V3 translation(const Scene& s, uint32_t x) {
  const float* r = s.xforms + 16 * x;
  return V3{r[3], r[7], r[11]};
}
static V3 point_to_object(const Scene& s, uint32_t m, V3 p) {
  uint32_t x = s.ranges[4 * m + 3];
  if (x == NONE) return p;
  return rotate_to_object(s, x, p - translation(s, x)) * s.xforms[16 * x + 13];
}

// ---- cone geometry (WGSL :159-255)
static V3 direction(const Cone& c, float s, float t) {
  return c.forward + c.right * (c.window[0] + s * (c.window[2] - c.window[0])) +
         c.up * (c.window[1] + t * (c.window[3] - c.window[1]));
}

static V3 to_local(const Cone& c, V3 p) {
  V3 d = p - c.apex;
  return V3{dot(c.right, d), dot(c.up, d), dot(c.forward, d)};
}

// WGSL :159-168
static bool in_bound(const Cone& c, V3 p, float r) {
  if (c.bound[3] < -1.0f) return true;
  V3 d = p - c.apex;
  float len = length(d);
  if (len <= r) return true;
  float s = r / len;
  float limit = c.bound[3] * std::sqrt(1.0f - s * s) - c.bound_sin * s;
  return dot(xyz(c.bound), d) >= limit * len;
}

// WGSL :174-186
static Cone bound_cone(Cone c, V3 axis) {
  float cos_half = 1.0f;
  for (uint32_t k = 0; k < 4; k++) {
    V3 d = normalize(direction(c, (float)(k & 1u), (float)(k >> 1u)));
    cos_half = fmin_(cos_half, dot(axis, d));
  }
  float cos_b = clampf(cos_half - 1e-5f, -1.0f, 1.0f);
  c.bound[0] = axis.x;
  c.bound[1] = axis.y;
  c.bound[2] = axis.z;
  c.bound[3] = cos_b;
  c.bound_sin = std::sqrt(fmax_(1.0f - cos_b * cos_b, 0.0f));
  return c;
}

// WGSL :193-201
static float lambert_edge_unit(V3 n, V3 an, V3 bn) {
  V3 g = cross(an, bn);
  float len = length(g);
  if (len > 1e-12f) return atan2_c(len, dot(an, bn)) * dot(n, g / len);
  return 0.0f;
}
static float lambert_edge(V3 n, V3 a, V3 b) { return lambert_edge_unit(n, normalize(a), normalize(b)); }

// WGSL :213-232
static bool may_touch(const Cone& c, V3 v0, V3 v1, V3 v2) {
  float f0 = c.far_offset - dot(c.far_normal, v0);
  float f1 = c.far_offset - dot(c.far_normal, v1);
  float f2 = c.far_offset - dot(c.far_normal, v2);
  if (f0 < 0.0f && f1 < 0.0f && f2 < 0.0f) return false;
  if (dot(c.horizon, v0 - c.apex) < 0.0f && dot(c.horizon, v1 - c.apex) < 0.0f &&
      dot(c.horizon, v2 - c.apex) < 0.0f)
    return false;
  V3 p0 = to_local(c, v0), p1 = to_local(c, v1), p2 = to_local(c, v2);
  if (p0.z < NEAR && p1.z < NEAR && p2.z < NEAR) return false;
  const float* w = c.window;
  if (p0.x - w[0] * p0.z < 0.0f && p1.x - w[0] * p1.z < 0.0f && p2.x - w[0] * p2.z < 0.0f) return false;
  if (w[2] * p0.z - p0.x < 0.0f && w[2] * p1.z - p1.x < 0.0f && w[2] * p2.z - p2.x < 0.0f) return false;
  if (p0.y - w[1] * p0.z < 0.0f && p1.y - w[1] * p1.z < 0.0f && p2.y - w[1] * p2.z < 0.0f) return false;
  if (w[3] * p0.z - p0.y < 0.0f && w[3] * p1.z - p1.y < 0.0f && w[3] * p2.z - p2.y < 0.0f) return false;
  return true;
}

static V3 box_corner(V3 lo, V3 hi, uint32_t k) {
  return V3{(k & 1u) ? hi.x : lo.x, (k & 2u) ? hi.y : lo.y, (k & 4u) ? hi.z : lo.z};
}

// WGSL :234-255
static bool may_touch_box(const Cone& c, V3 lo, V3 hi) {
  bool far_out = true, near_out = true, out_u0 = true, out_u1 = true, out_v0 = true, out_v1 = true,
       below = true;
  for (uint32_t k = 0; k < 8; k++) {
    V3 corner = box_corner(lo, hi, k);
    if (c.far_offset - dot(c.far_normal, corner) >= 0.0f) far_out = false;
    if (dot(c.horizon, corner - c.apex) >= 0.0f) below = false;
    V3 p = to_local(c, corner);
    if (p.z >= NEAR) near_out = false;
    if (p.x - c.window[0] * p.z >= 0.0f) out_u0 = false;
    if (c.window[2] * p.z - p.x >= 0.0f) out_u1 = false;
    if (p.y - c.window[1] * p.z >= 0.0f) out_v0 = false;
    if (c.window[3] * p.z - p.y >= 0.0f) out_v1 = false;
  }
  if (far_out || near_out || out_u0 || out_u1 || out_v0 || out_v1 || below) return false;
  return in_bound(c, (lo + hi) * 0.5f, 0.5f * length(hi - lo));
}

// ---- polygons and the accumulation rasterizer (WGSL :257-451)
struct Poly3 {
  V3 v[10];
  uint32_t n;
};
struct Poly2 {
  V2 v[10];
  uint32_t n;
};

// A clip keeps at most 10 vertices (the reference's MAX_POLY, WGSL :11); in
// exact arithmetic these clips never produce more.
static Poly3 clip_plane3(const Poly3& p, V3 normal, float offset) {
  Poly3 out;
  out.n = 0;
  for (uint32_t i = 0; i < p.n; i++) {
    V3 a = p.v[i];
    V3 b = p.v[(i + 1) % p.n];
    float da = offset - dot(normal, a);
    float db = offset - dot(normal, b);
    if (da >= 0.0f && out.n < 10) out.v[out.n++] = a;
    if (da >= 0.0f && db < 0.0f && out.n < 10) out.v[out.n++] = a + (b - a) * (da / (da - db));
    if (da < 0.0f && db >= 0.0f && out.n < 10) out.v[out.n++] = b + (a - b) * (db / (db - da));
  }
  return out;
}

static Poly2 clip_line2(const Poly2& p, V2 normal, float offset) {
  Poly2 out;
  out.n = 0;
  for (uint32_t i = 0; i < p.n; i++) {
    V2 a = p.v[i];
    V2 b = p.v[(i + 1) % p.n];
    float da = offset - dot2(normal, a);
    float db = offset - dot2(normal, b);
    if (da >= 0.0f && out.n < 10) out.v[out.n++] = a;
    if (da >= 0.0f && db < 0.0f && out.n < 10) out.v[out.n++] = a + (b - a) * (da / (da - db));
    if (da < 0.0f && db >= 0.0f && out.n < 10) out.v[out.n++] = b + (a - b) * (db / (db - da));
  }
  return out;
}

struct Acc {
  float a[68];
  uint32_t span_lo, span_hi;
};

// WGSL :297-348
static void draw_edge(Acc& acc, V2 q0, V2 q1, float weight) {
  if (std::fabs(q0.y - q1.y) <= 1.1920929e-7f) return;
  float dir = 1.0f;
  V2 p0 = q0, p1 = q1;
  if (q0.y > q1.y) {
    dir = -1.0f;
    p0 = q1;
    p1 = q0;
  }
  float dxdy = (p1.x - p0.x) / (p1.y - p0.y);
  float x = p0.x;
  uint32_t y_end = std::min((uint32_t)std::ceil(p1.y), GRID);
  for (uint32_t y = (uint32_t)p0.y; y < y_end; y++) {
    uint32_t row = y * GRID;
    acc.span_lo = std::min(acc.span_lo, row);
    acc.span_hi = std::max(acc.span_hi, row + GRID + 1u);
    float dy = fmin_((float)(y + 1u), p1.y) - fmax_((float)y, p0.y);
    float x_next = clampf(x + dxdy * dy, 0.0f, (float)GRID);
    float d = dy * dir * weight;
    float x0 = fmin_(x, x_next);
    float x1 = fmax_(x, x_next);
    float x0_floor = std::floor(x0);
    uint32_t x0i = (uint32_t)x0_floor;
    float x1_ceil = std::ceil(x1);
    uint32_t x1i = (uint32_t)x1_ceil;
    if (x1i <= x0i + 1u) {
      float xmf = 0.5f * (x + x_next) - x0_floor;
      acc.a[row + x0i] += d - d * xmf;
      acc.a[row + x0i + 1u] += d * xmf;
    } else {
      float s = 1.0f / (x1 - x0);
      float x0f = x0 - x0_floor;
      float a0 = 0.5f * s * (1.0f - x0f) * (1.0f - x0f);
      float x1f = x1 - x1_ceil + 1.0f;
      float am = 0.5f * s * x1f * x1f;
      acc.a[row + x0i] += d * a0;
      if (x1i == x0i + 2u) {
        acc.a[row + x0i + 1u] += d * (1.0f - a0 - am);
      } else {
        float a1 = s * (1.5f - x0f);
        acc.a[row + x0i + 1u] += d * (a1 - a0);
        for (uint32_t xi = x0i + 2u; xi < x1i - 1u; xi++) acc.a[row + xi] += d * s;
        float a2 = a1 + (float)(x1i - x0i - 3u) * s;
        acc.a[row + x1i - 1u] += d * (1.0f - a2 - am);
      }
      acc.a[row + x1i] += d * am;
    }
    x = x_next;
  }
}

static V2 clamp2(V2 v, float g) { return V2{clampf(v.x, 0.0f, g), clampf(v.y, 0.0f, g)}; }

static void draw_poly(Acc& acc, const Poly2& q, float weight) {
  const float g = (float)GRID;
  for (uint32_t i = 0; i < q.n; i++) {
    V2 a = clamp2(q.v[i], g);
    V2 b = clamp2(q.v[(i + 1) % q.n], g);
    draw_edge(acc, a, b, weight);
  }
}

// WGSL :350-451
static bool rasterize(const Cone& c, Query& qs, Acc& acc, V3 v0, V3 v1, V3 v2) {
  Poly3 p;
  p.v[0] = v0;
  p.v[1] = v1;
  p.v[2] = v2;
  p.n = 3;
  float f0 = c.far_offset - dot(c.far_normal, v0);
  float f1 = c.far_offset - dot(c.far_normal, v1);
  float f2 = c.far_offset - dot(c.far_normal, v2);
  if (f0 < 0.0f || f1 < 0.0f || f2 < 0.0f) p = clip_plane3(p, c.far_normal, c.far_offset);
  bool all_near_ok = true, all_in_window = true;
  const float* w = c.window;
  for (uint32_t i = 0; i < p.n; i++) {
    p.v[i] = to_local(c, p.v[i]);
    V3 v = p.v[i];
    if (v.z < NEAR) all_near_ok = false;
    if (v.x < w[0] * v.z || v.x > w[2] * v.z || v.y < w[1] * v.z || v.y > w[3] * v.z) all_in_window = false;
  }
  if (!all_near_ok) p = clip_plane3(p, V3{0.0f, 0.0f, -1.0f}, -NEAR);
  if (p.n < 3) return false;
  float su = (float)GRID / (w[2] - w[0]);
  float sv = (float)GRID / (w[3] - w[1]);
  const float g = (float)GRID;
  Poly2 q;
  if (all_near_ok && all_in_window) {
    q.n = p.n;
    for (uint32_t i = 0; i < p.n; i++)
      q.v[i] = V2{(p.v[i].x / p.v[i].z - w[0]) * su, (p.v[i].y / p.v[i].z - w[1]) * sv};
  } else {
    float d = dot(to_local(c, v0), cross(to_local(c, v1), to_local(c, v2)));
    if (d == 0.0f) return false;
    float side = d > 0.0f ? 1.0f : -1.0f;
    q.n = 4;
    q.v[0] = V2{0.0f, 0.0f};
    q.v[2] = V2{g, g};
    if (side > 0.0f) {
      q.v[1] = V2{g, 0.0f};
      q.v[3] = V2{0.0f, g};
    } else {
      q.v[1] = V2{0.0f, g};
      q.v[3] = V2{g, 0.0f};
    }
    for (uint32_t i = 0; i < p.n; i++) {
      V3 e = cross(p.v[i], p.v[(i + 1) % p.n]);
      float len = length(e);
      if (len == 0.0f) continue;
      V3 k = e * (side / len);
      q = clip_line2(q, V2{-k.x / su, -k.y / sv}, k.x * w[0] + k.y * w[1] + k.z);
    }
  }
  bool inside = true;
  for (uint32_t i = 0; i < q.n; i++)
    if (q.v[i].x < 0.0f || q.v[i].y < 0.0f || q.v[i].x > g || q.v[i].y > g) inside = false;
  if (!inside) {
    q = clip_line2(q, V2{-1.0f, 0.0f}, 0.0f);
    q = clip_line2(q, V2{1.0f, 0.0f}, g);
    q = clip_line2(q, V2{0.0f, -1.0f}, 0.0f);
    q = clip_line2(q, V2{0.0f, 1.0f}, g);
  }
  if (q.n < 3) return false;
  draw_poly(acc, q, qs.tri_weight);
  return true;
}

// ---- prefiltered splats (WGSL :455-574)
static float cov_sample(const Scene& s, uint32_t n, int i, int j, int k) {
  return s.blas_cov[n * 27u + (uint32_t)(9 * (i + 1) + 3 * (j + 1) + (k + 1))];
}

// WGSL :462-486
static float node_front_area(const Scene& s, uint32_t n, V3 v) {
  V3 a = vabs(v);
  uint32_t axis = 0;
  if (a.y > a.x && a.y >= a.z)
    axis = 1;
  else if (a.z > a.x && a.z > a.y)
    axis = 2;
  float mainv, sv, tv;
  if (axis == 0) {
    mainv = v.x;
    sv = v.y / a.x;
    tv = v.z / a.x;
  } else if (axis == 1) {
    mainv = v.y;
    sv = v.x / a.y;
    tv = v.z / a.y;
  } else {
    mainv = v.z;
    sv = v.x / a.z;
    tv = v.y / a.z;
  }
  int m = mainv >= 0.0f ? 1 : -1;
  int s0 = sv < 0.0f ? -1 : 0;
  int t0 = tv < 0.0f ? -1 : 0;
  float fs = sv - (float)s0;
  float ft = tv - (float)t0;
  float f[4];
  for (int c = 0; c < 4; c++) {
    int gs = s0 + (c & 1);
    int gt = t0 + (c >> 1);
    if (axis == 0)
      f[c] = cov_sample(s, n, m, gs, gt);
    else if (axis == 1)
      f[c] = cov_sample(s, n, gs, m, gt);
    else
      f[c] = cov_sample(s, n, gs, gt, m);
  }
  return mixf(mixf(f[0], f[1], fs), mixf(f[2], f[3], fs), ft);
}

// WGSL :490-512
static Poly2 hull8(const V2 pts[8]) {
  uint32_t start = 0;
  for (uint32_t k = 1; k < 8; k++)
    if (pts[k].x < pts[start].x || (pts[k].x == pts[start].x && pts[k].y < pts[start].y)) start = k;
  Poly2 out;
  out.n = 0;
  uint32_t cur = start;
  for (;;) {
    out.v[out.n++] = pts[cur];
    uint32_t next = (cur + 1) % 8;
    for (uint32_t k = 0; k < 8; k++) {
      V2 e = pts[next] - pts[cur];
      V2 f = pts[k] - pts[cur];
      float cr = e.x * f.y - e.y * f.x;
      if (cr > 0.0f || (cr == 0.0f && dot2(f, f) > dot2(e, e))) next = k;
    }
    cur = next;
    if (cur == start || out.n >= 8) break;
  }
  return out;
}

static Poly2 clip_to_grid(Poly2 q) {
  const float g = (float)GRID;
  q = clip_line2(q, V2{-1.0f, 0.0f}, 0.0f);
  q = clip_line2(q, V2{1.0f, 0.0f}, g);
  q = clip_line2(q, V2{0.0f, -1.0f}, 0.0f);
  q = clip_line2(q, V2{0.0f, 1.0f}, g);
  return q;
}

// WGSL :519-574: 0 culled, 1 splatted with nothing drawn, 2 splatted and drawn, 3 descend.
static uint32_t node_cone_test(const Cone& c, Query& qs, Acc& acc, uint32_t n, V3 lo, V3 hi) {
  const float g = (float)GRID;
  V2 pts[8];
  V2 mn{3.4e38f, 3.4e38f}, mx{-3.4e38f, -3.4e38f};
  bool far_out = true, near_out = true, out_u0 = true, out_u1 = true, out_v0 = true, out_v1 = true,
       below = true;
  bool inside = qs.prefilter_bins > 0.0f;
  for (uint32_t k = 0; k < 8; k++) {
    V3 corner = box_corner(lo, hi, k);
    if (dot(c.horizon, corner - c.apex) >= 0.0f) below = false;
    bool in_far = c.far_offset - dot(c.far_normal, corner) >= 0.0f;
    if (in_far)
      far_out = false;
    else
      inside = false;
    V3 p = to_local(c, corner);
    if (p.z >= NEAR)
      near_out = false;
    else
      inside = false;
    if (p.x - c.window[0] * p.z >= 0.0f) out_u0 = false;
    if (c.window[2] * p.z - p.x >= 0.0f) out_u1 = false;
    if (p.y - c.window[1] * p.z >= 0.0f) out_v0 = false;
    if (c.window[3] * p.z - p.y >= 0.0f) out_v1 = false;
    if (inside) {
      V2 q{(p.x / p.z - c.window[0]) / (c.window[2] - c.window[0]) * g,
           (p.y / p.z - c.window[1]) / (c.window[3] - c.window[1]) * g};
      pts[k] = q;
      mn = V2{fmin_(mn.x, q.x), fmin_(mn.y, q.y)};
      mx = V2{fmax_(mx.x, q.x), fmax_(mx.y, q.y)};
    }
  }
  if (far_out || near_out || out_u0 || out_u1 || out_v0 || out_v1 || below) return 0;
  if (!in_bound(c, (lo + hi) * 0.5f, 0.5f * length(hi - lo))) return 0;
  if (!inside || fmax_(mx.x - mn.x, mx.y - mn.y) > qs.prefilter_bins) return 3;
  if (mx.x <= 0.0f || mx.y <= 0.0f || mn.x >= g || mn.y >= g) return 1;
  V3 e = hi - lo;
  V3 v = normalize(c.apex - (lo + hi) * 0.5f);
  float silhouette = std::fabs(v.x) * e.y * e.z + std::fabs(v.y) * e.x * e.z + std::fabs(v.z) * e.x * e.y;
  if (silhouette <= 1e-30f) return 1;
  float alpha = node_front_area(*qs.s, n, v) / silhouette;
  if (alpha <= 0.0f) return 1;
  Poly2 q = clip_to_grid(hull8(pts));
  if (q.n < 3) return 1;
  draw_poly(acc, q, alpha);
  return 2;
}

// x^5 for Schlick's reflectance, by products: ((x x)(x x)) x. (The WGSL's
// `pow(x, 5.0)` is left to the driver; the kernels use these products too.)
static inline float pow5(float x) {
  const float x2 = x * x;
  return (x2 * x2) * x;
}

// WGSL :594-616
static bool cover_triangle(const Cone& c, Query& qs, Acc& acc, uint32_t t) {
  if (qs.o->profile_stage == 4 || t == qs.cover_skip) return false;
  const float* tr = qs.s->tris + 12u * t;
  V3 v0 = xyz(tr), v1 = xyz(tr + 4), v2 = xyz(tr + 8);
  V3 n = cross(v1 - v0, v2 - v0);
  if (dot(n, c.apex - v0) <= 0.0f) return false;
  if (dot(n, direction(c, 0.0f, 0.0f)) >= 0.0f && dot(n, direction(c, 1.0f, 0.0f)) >= 0.0f &&
      dot(n, direction(c, 0.0f, 1.0f)) >= 0.0f && dot(n, direction(c, 1.0f, 1.0f)) >= 0.0f)
    return false;
  if (!may_touch(c, v0, v1, v2)) return false;
  if (qs.tri_transmit > 0.0f) {
    // WGSL :608-614: thin glass passes what Fresnel's reflectance (Schlick's,
    // F0 0.04, at the cone's centre direction) leaves, times the factor.
    const float cos_t = std::fabs(dot(normalize(n), normalize(direction(c, 0.5f, 0.5f))));
    const float f = 0.04f + 0.96f * pow5(1.0f - cos_t);
    float passes = qs.tri_transmit * (1.0f - f);
    if (qs.tri_volume) passes *= 1.0f - f;
    qs.tri_weight = 1.0f - passes;
  }
  return rasterize(c, qs, acc, v0, v1, v2);
}

// WGSL :623-638
static Cone cone_to_object(const Scene& s, const Cone& c, uint32_t m) {
  uint32_t x = s.ranges[4 * m + 3];
  if (x == NONE) return c;
  Cone o;
  o.apex = point_to_object(s, m, c.apex);
  o.right = rotate_to_object(s, x, c.right);
  o.up = rotate_to_object(s, x, c.up);
  o.forward = rotate_to_object(s, x, c.forward);
  for (int k = 0; k < 4; k++) o.window[k] = c.window[k];
  o.far_normal = rotate_to_object(s, x, c.far_normal);
  o.far_offset = (c.far_offset - dot(c.far_normal, translation(s, x))) * s.xforms[16 * x + 13];
  o.horizon = rotate_to_object(s, x, c.horizon);
  V3 b = rotate_to_object(s, x, xyz(c.bound));
  o.bound[0] = b.x;
  o.bound[1] = b.y;
  o.bound[2] = b.z;
  o.bound[3] = c.bound[3];
  o.bound_sin = c.bound_sin;
  return o;
}

// Combines one mesh's accumulation buffer into `cov` (WGSL :809-828).
static void combine_mesh(const Acc& acc, float* cov, uint32_t& full, bool snap, bool independent) {
  float running = 0.0f;
  uint32_t end = std::min(acc.span_hi, BINS);
  for (uint32_t b = acc.span_lo; b < end; b++) {
    running += acc.a[b];
    float covered = std::fabs(running);
    if (snap) {
      if (covered >= ONE_MINUS_1EM5) covered = 1.0f;
    } else {
      covered = fmin_(covered, 1.0f);
    }
    if (covered > 1e-6f) {
      float before = cov[b];
      if (independent)
        cov[b] = 1.0f - (1.0f - before) * (1.0f - fmin_(covered, 1.0f));
      else
        cov[b] = fmax_(before, fmin_(covered, 1.0f));
      if (before < 1.0f && cov[b] >= 1.0f) full++;
    }
  }
}

// WGSL :719-829 (one invocation alone: `on` is its own packet)
static void add_mesh_coverage(const Cone& world_cone, Query& qs, uint32_t m, float* cov, uint32_t& full,
                              uint32_t base) {
  if (qs.o->profile_stage == 3) return;
  const Scene& s = *qs.s;
  Cone c = cone_to_object(s, world_cone, m);
  const uint32_t* range = s.ranges + 4 * m;
  qs.cover_skip = m == qs.skip_obj ? qs.skip_tri : NONE;
  qs.tri_weight = 1.0f;
  qs.tri_transmit = 0.0f;
  qs.tri_volume = false;
  if (range[3] != NONE) {
    // WGSL :733-742: blended (clear > 0), transmissive (clear < 0), or a
    // transmissive volume (clear <= -2).
    float clear = s.xforms[16 * range[3] + 14];
    qs.tri_weight = 1.0f - fmax_(clear, 0.0f);
    qs.tri_transmit = fmax_(-clear, 0.0f);
    if (clear <= -2.0f) {
      qs.tri_transmit = -2.0f - clear;
      qs.tri_volume = true;
    }
  }
  Acc acc;
  for (int b = 0; b < 68; b++) acc.a[b] = 0.0f;
  acc.span_lo = 68;
  acc.span_hi = 0;
  bool touched = false;
  if (qs.o->hierarchy == 2) {
    if (range[2] != NONE) {
      uint32_t sp = base + 1;
      if (base >= STACK) {
        qs.stack_overflow = true;
        return;
      }
      qs.stack[base] = range[2];
      while (sp > base) {
        sp--;
        uint32_t n = qs.stack[sp];
        bool mine = qs.cone_work < qs.cone_cap;
        if (mine) qs.cone_work++;
        if (!mine) continue;
        const Node& node = s.blas[n];
        uint32_t test = node_cone_test(c, qs, acc, n, xyz(node.lo), xyz(node.hi));
        if (test == 2) touched = true;
        if (test != 3) continue;
        if (node.info[1] > 0) {
          qs.cone_work += node.info[1];
          for (uint32_t k = node.info[0]; k < node.info[0] + node.info[1]; k++)
            if (cover_triangle(c, qs, acc, s.blas_items[k])) touched = true;
        } else {
          if (sp + 2 > STACK) {
            qs.stack_overflow = true;
            return;
          }
          qs.stack[sp] = node.info[0];
          qs.stack[sp + 1] = node.info[0] + 1;
          sp += 2;
        }
      }
    }
  } else {
    for (uint32_t t = range[0]; t < range[0] + range[1]; t++)
      if (cover_triangle(c, qs, acc, t)) touched = true;
  }
  if (touched) combine_mesh(acc, cov, full, true, qs.independent);
}

// WGSL :1096-1150 (object-level splat; light cones combine by the maximum)
static bool splat_object_node(const Cone& c, Query& qs, uint32_t n, V3 lo, V3 hi, float* cov, uint32_t& full);

// WGSL :1160-1227
static void scene_coverage(const Cone& c, Query& qs, float* cov, uint32_t& full) {
  const Scene& s = *qs.s;
  qs.cone_work = 0;
  if (qs.o->hierarchy >= 1) {
    uint32_t sp = 1;
    qs.stack[0] = 0;
    while (sp > 0) {
      sp--;
      uint32_t n = qs.stack[sp];
      bool mine = full < BINS && qs.cone_work < qs.cone_cap;
      if (mine) qs.cone_work++;
      if (!mine) continue;
      const Node& node = s.nodes[n];
      bool descend = false;
      if (may_touch_box(c, xyz(node.lo), xyz(node.hi))) {
        if (node.info[2] == 0 && splat_object_node(c, qs, n, xyz(node.lo), xyz(node.hi), cov, full)) {
        } else {
          descend = true;
        }
      }
      if (!descend) continue;
      if (node.info[1] > 0) {
        for (uint32_t k = node.info[0]; k < node.info[0] + node.info[1]; k++)
          add_mesh_coverage(c, qs, s.items[k], cov, full, sp);
      } else {
        uint32_t nearc = node.info[0], farc = node.info[0] + 1;
        if (!qs.independent) {
          const Node& a = s.nodes[nearc];
          const Node& b = s.nodes[farc];
          bool use_bound = c.bound[3] >= -1.0f && (c.bound[0] != 0.0f || c.bound[1] != 0.0f || c.bound[2] != 0.0f);
          V3 axis = use_bound ? xyz(c.bound) : c.forward;
          float da = dot(axis, (xyz(a.lo) + xyz(a.hi)) * 0.5f - c.apex);
          float db = dot(axis, (xyz(b.lo) + xyz(b.hi)) * 0.5f - c.apex);
          if (db < da) {
            nearc = node.info[0] + 1;
            farc = node.info[0];
          }
        }
        if (sp + 2 > STACK) {
          qs.stack_overflow = true;
          return;
        }
        qs.stack[sp] = farc;
        qs.stack[sp + 1] = nearc;
        sp += 2;
      }
    }
  } else {
    for (uint32_t m = 0; m < s.n_objects; m++)
      if (full < BINS) add_mesh_coverage(c, qs, m, cov, full, 0);
  }
}

// Cephes expf (S. L. Moshier): x = n ln2 + r with ln2 split in two parts,
// then a degree-6 polynomial for e^r.
static float exp_c(float x) {
  if (x > 88.72283905206835f) return INFINITY;
  if (x < -103.278929903431851103f) return 0.0f;
  float z = std::floor(1.44269504088896341f * x + 0.5f);
  x = x - z * 0.693359375f;
  x = x - z * -2.12194440e-4f;
  int n = (int)z;
  float xx = x * x;
  float p = ((((((1.9875691500e-4f * x + 1.3981999507e-3f) * x + 8.3334519073e-3f) * x + 4.1665795894e-2f) * x +
               1.6666665459e-1f) *
                  x +
              5.0000001201e-1f) *
                 xx +
             x) +
            1.0f;
  return std::ldexp(p, n);
}

static bool splat_object_node(const Cone& c, Query& qs, uint32_t n, V3 lo, V3 hi, float* cov, uint32_t& full) {
  if (qs.o->object_cov == NONE || qs.prefilter_bins <= 0.0f) return false;
  const float g = (float)GRID;
  V2 pts[8];
  V2 mn{3.4e38f, 3.4e38f}, mx{-3.4e38f, -3.4e38f};
  for (uint32_t k = 0; k < 8; k++) {
    V3 corner = box_corner(lo, hi, k);
    if (c.far_offset - dot(c.far_normal, corner) < 0.0f) return false;
    V3 p = to_local(c, corner);
    if (p.z < NEAR) return false;
    V2 q{(p.x / p.z - c.window[0]) / (c.window[2] - c.window[0]) * g,
         (p.y / p.z - c.window[1]) / (c.window[3] - c.window[1]) * g};
    pts[k] = q;
    mn = V2{fmin_(mn.x, q.x), fmin_(mn.y, q.y)};
    mx = V2{fmax_(mx.x, q.x), fmax_(mx.y, q.y)};
  }
  if (fmax_(mx.x - mn.x, mx.y - mn.y) > qs.prefilter_bins) return false;
  if (mx.x <= 0.0f || mx.y <= 0.0f || mn.x >= g || mn.y >= g) return true;
  V3 e = hi - lo;
  V3 v = normalize(c.apex - (lo + hi) * 0.5f);
  float silhouette = std::fabs(v.x) * e.y * e.z + std::fabs(v.y) * e.x * e.z + std::fabs(v.z) * e.x * e.y;
  if (silhouette <= 1e-30f) return true;
  float alpha = 1.0f - exp_c(-node_front_area(*qs.s, qs.o->object_cov + n, v) / silhouette);
  if (alpha <= 0.0f) return true;
  Poly2 q = clip_to_grid(hull8(pts));
  if (q.n < 3) return true;
  Acc acc;
  for (int b = 0; b < 68; b++) acc.a[b] = 0.0f;
  acc.span_lo = 68;
  acc.span_hi = 0;
  draw_poly(acc, q, alpha);
  combine_mesh(acc, cov, full, false, qs.independent);
  return true;
}

// ---- weights (WGSL :1232-1272, :1331-1410)
static float bin_weight(const Cone& c, V3 normal, uint32_t b) {
  float col = (float)(b % GRID), row = (float)(b / GRID), g = (float)GRID;
  V3 d0 = direction(c, col / g, row / g);
  V3 d1 = direction(c, (col + 1.0f) / g, row / g);
  V3 d2 = direction(c, (col + 1.0f) / g, (row + 1.0f) / g);
  V3 d3 = direction(c, col / g, (row + 1.0f) / g);
  return std::fabs(0.5f * (lambert_edge(normal, d0, d1) + lambert_edge(normal, d1, d2) +
                           lambert_edge(normal, d2, d3) + lambert_edge(normal, d3, d0)));
}

static float clipped_quad_weight(V3 normal, const V3 d[4], bool allow_full) {
  V3 poly[8];
  uint32_t n = 0;
  for (uint32_t k = 0; k < 4; k++) {
    V3 a = d[k], e = d[(k + 1) % 4];
    float sa = dot(normal, a), se = dot(normal, e);
    if (sa >= 0.0f) poly[n++] = a;
    if ((sa >= 0.0f) != (se >= 0.0f)) poly[n++] = a + (e - a) * (sa / (sa - se));
  }
  if (!allow_full && n < 3) return 0.0f;
  float sum = 0.0f;
  for (uint32_t k = 0; k < n; k++) sum += lambert_edge(normal, poly[k], poly[(k + 1) % n]);
  return std::fabs(0.5f * sum);
}

static float clipped_bin_weight(const Cone& c, V3 normal, uint32_t b) {
  float col = (float)(b % GRID), row = (float)(b / GRID), g = (float)GRID;
  V3 d[4];
  d[0] = direction(c, col / g, row / g);
  d[1] = direction(c, (col + 1.0f) / g, row / g);
  d[2] = direction(c, (col + 1.0f) / g, (row + 1.0f) / g);
  d[3] = direction(c, col / g, (row + 1.0f) / g);
  uint32_t above = 0;
  for (uint32_t k = 0; k < 4; k++)
    if (dot(normal, d[k]) > 0.0f) above++;
  if (above == 4) return bin_weight(c, normal, b);
  if (above == 0) return 0.0f;
  return clipped_quad_weight(normal, d, true);
}

static float clipped_window_weight(const Cone& c, V3 normal) {
  V3 d[4];
  d[0] = direction(c, 0.0f, 0.0f);
  d[1] = direction(c, 1.0f, 0.0f);
  d[2] = direction(c, 1.0f, 1.0f);
  d[3] = direction(c, 0.0f, 1.0f);
  return clipped_quad_weight(normal, d, false);
}

static float cone_visible(const Cone& c, V3 normal, const float* cov) {
  V3 d00 = direction(c, 0.0f, 0.0f), d10 = direction(c, 1.0f, 0.0f);
  V3 d11 = direction(c, 1.0f, 1.0f), d01 = direction(c, 0.0f, 1.0f);
  if (dot(normal, d00) > 0.0f && dot(normal, d10) > 0.0f && dot(normal, d11) > 0.0f && dot(normal, d01) > 0.0f) {
    float blocked = 0.0f;
    for (uint32_t b = 0; b < BINS; b++) {
      float covered = cov[b];
      if (covered > 0.0f) blocked += covered * bin_weight(c, normal, b);
    }
    float total = std::fabs(0.5f * (lambert_edge(normal, d00, d10) + lambert_edge(normal, d10, d11) +
                                    lambert_edge(normal, d11, d01) + lambert_edge(normal, d01, d00)));
    return fmax_(total - blocked, 0.0f);
  }
  float blocked = 0.0f;
  for (uint32_t b = 0; b < BINS; b++) {
    float covered = cov[b];
    if (covered > 0.0f) blocked += covered * clipped_bin_weight(c, normal, b);
  }
  return fmax_(clipped_window_weight(c, normal) - blocked, 0.0f);
}

static float window_open(const Cone& c, V3 normal) {
  V3 d00 = direction(c, 0.0f, 0.0f), d10 = direction(c, 1.0f, 0.0f);
  V3 d11 = direction(c, 1.0f, 1.0f), d01 = direction(c, 0.0f, 1.0f);
  if (dot(normal, d00) > 0.0f && dot(normal, d10) > 0.0f && dot(normal, d11) > 0.0f && dot(normal, d01) > 0.0f)
    return std::fabs(0.5f * (lambert_edge(normal, d00, d10) + lambert_edge(normal, d10, d11) +
                             lambert_edge(normal, d11, d01) + lambert_edge(normal, d01, d00)));
  return clipped_window_weight(c, normal);
}

// ---- lights (WGSL :850-1042, :2125-2190)
static bool is_sphere(const Light& r) { return r.corner[3] > 1.5f; }

static V3 light_center(const Light& r) {
  if (is_sphere(r)) return xyz(r.corner);
  return xyz(r.corner) + (xyz(r.edge_u) + xyz(r.edge_v)) * 0.5f;
}

static float spot_falloff(const Light& r, V3 w) {
  if (!is_sphere(r) || r.edge_v[3] < 0.5f) return 1.0f;
  float a = clampf(dot(xyz(r.edge_v), w) * r.edge_u[1] + r.edge_u[2], 0.0f, 1.0f);
  return a * a;
}

static bool faces_light(const Light& r, V3 point, V3 normal) {
  V3 to_light = light_center(r) - point;
  if (is_sphere(r)) return dot(normal, to_light) > -r.edge_u[0];
  return dot(normal, to_light) > 0.0f;
}

// Share of each bin of a sphere light's window inside the inscribed disk:
// the exact area of the circle of radius 4 (bin units) about the window's
// centre within each bin, rounded to 6 decimals (WGSL :888-897).
static const float DISK[64] = {
    0.000000f, 0.160355f, 0.694958f, 0.957934f, 0.957934f, 0.694958f, 0.160355f, 0.000000f,
    0.160355f, 0.939877f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 0.939877f, 0.160355f,
    0.694958f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 0.694958f,
    0.957934f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 0.957934f,
    0.957934f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 0.957934f,
    0.694958f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 0.694958f,
    0.160355f, 0.939877f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 0.939877f, 0.160355f,
    0.000000f, 0.160355f, 0.694958f, 0.957934f, 0.957934f, 0.694958f, 0.160355f, 0.000000f,
};

struct LightCone {
  Cone c;
  bool ok;
};

static Cone zero_cone() {
  Cone c;
  std::memset(&c, 0, sizeof(c));
  return c;
}

// WGSL :902-931
static LightCone light_cone_sphere(Query& qs, V3 apex, const Light& r) {
  LightCone out;
  out.c = zero_cone();
  out.ok = false;
  qs.prefilter_bins = qs.o->prefilter;
  V3 to_center = xyz(r.corner) - apex;
  float d = length(to_center);
  float radius = r.edge_u[0];
  if (d <= radius * 1.0001f) return out;
  V3 forward = to_center / d;
  float s = radius / d;
  float t = s / std::sqrt(fmax_(1.0f - s * s, 1e-12f));
  qs.wide_clipped = 2.0f * t > qs.o->wide_extent;
  qs.wide_window = qs.wide_clipped;
  t = fmin_(t, 0.5f * qs.o->wide_extent);
  V3 helper = std::fabs(forward.x) > 0.9f ? V3{0.0f, 1.0f, 0.0f} : V3{1.0f, 0.0f, 0.0f};
  V3 right = normalize(cross(helper, forward));
  out.c.apex = apex;
  out.c.right = right;
  out.c.up = cross(forward, right);
  out.c.forward = forward;
  out.c.window[0] = -t;
  out.c.window[1] = -t;
  out.c.window[2] = t;
  out.c.window[3] = t;
  out.c.far_normal = forward;
  out.c.far_offset = dot(forward, apex) + (d - radius * s) * ONE_MINUS_1EM5;
  out.c = bound_cone(out.c, forward);
  out.ok = true;
  return out;
}

// WGSL :937-955
static float sphere_unshadowed(V3 point, V3 normal, const Light& r) {
  V3 to_center = xyz(r.corner) - point;
  float d2 = dot(to_center, to_center);
  float radius = r.edge_u[0];
  if (d2 <= radius * radius) return 0.0f;
  float sin2 = radius * radius / d2;
  float cos_t = clampf(dot(normal, to_center) * (1.0f / std::sqrt(d2)), -1.0f, 1.0f);
  if (cos_t * cos_t > sin2) return 3.14159265f * sin2 * fmax_(cos_t, 0.0f);
  float sin_t = std::sqrt(1.0f - cos_t * cos_t);
  float x = std::sqrt(1.0f / sin2 - 1.0f);
  float y = clampf(-x * cos_t / sin_t, -1.0f, 1.0f);
  float q = sin_t * std::sqrt(1.0f - y * y);
  return fmax_((cos_t * acos_c(y) - x * q) * sin2 + atan_c(q / x), 0.0f);
}

// WGSL :960-972
static float sphere_open_share(const Query& qs, const Cone& c, V3 normal, const float* cov) {
  float total = 0.0f, open = 0.0f;
  const float g = (float)GRID;
  for (uint32_t b = 0; b < BINS; b++) {
    float m = qs.wide_clipped ? 1.0f : DISK[b];
    if (m <= 0.0f) continue;
    V3 d = direction(c, ((float)(b % GRID) + 0.5f) / g, ((float)(b / GRID) + 0.5f) / g);
    float l2 = dot(d, d);
    float w = m * fmax_(dot(normal, d), 0.0f) / (l2 * l2);
    total += w;
    open += w * (1.0f - cov[b]);
  }
  if (total <= 0.0f) return 1.0f;
  return open / total;
}

// WGSL :974-1042
static LightCone light_cone_rect(Query& qs, V3 apex, const Light& r) {
  if (is_sphere(r)) return light_cone_sphere(qs, apex, r);
  LightCone out;
  out.c = zero_cone();
  out.ok = false;
  qs.prefilter_bins = qs.o->prefilter;
  V3 corner0 = xyz(r.corner), edge_u = xyz(r.edge_u), edge_v = xyz(r.edge_v);
  V3 forward = normalize(cross(edge_u, edge_v));
  if (r.corner[3] > 0.5f && dot(forward, apex - corner0) <= 0.0f) return out;
  if (dot(forward, corner0 - apex) < 0.0f) forward = -forward;
  float dist = dot(forward, corner0 - apex);
  if (dist <= 1e-6f) return out;
  V3 right = normalize(edge_u);
  V3 up = cross(forward, right);
  V3 rel = corner0 - apex;
  V2 lo{3.4e38f, 3.4e38f}, hi{-3.4e38f, -3.4e38f};
  for (uint32_t k = 0; k < 4; k++) {
    V3 corner = rel + edge_u * (float)(k & 1u) + edge_v * (float)((k >> 1u) & 1u);
    V2 uv{dot(right, corner) / dist, dot(up, corner) / dist};
    lo = V2{fmin_(lo.x, uv.x), fmin_(lo.y, uv.y)};
    hi = V2{fmax_(hi.x, uv.x), fmax_(hi.y, uv.y)};
  }
  V2 extent = hi - lo;
  V2 foot{fmin_(fmax_(0.0f, lo.x), hi.x), fmin_(fmax_(0.0f, lo.y), hi.y)};
  V3 mid = normalize(V3{0.5f * (lo.x + hi.x), 0.5f * (lo.y + hi.y), 1.0f});
  float cos_wide = 1.0f;
  for (uint32_t k = 0; k < 4; k++) {
    V3 q{(k & 1u) ? hi.x : lo.x, (k & 2u) ? hi.y : lo.y, 1.0f};
    cos_wide = fmin_(cos_wide, dot(mid, normalize(q)));
  }
  float size = qs.o->wide_extent;
  qs.wide_window = extent.x > size || extent.y > size;
  qs.wide_clipped = qs.wide_window && cos_wide < 0.8660254f;
  if (qs.wide_clipped) {
    bool wx = extent.x > size, wy = extent.y > size;
    V2 keep{clampf(foot.x - 0.5f * size, lo.x, hi.x - size), clampf(foot.y - 0.5f * size, lo.y, hi.y - size)};
    V2 nlo{wx ? keep.x : lo.x, wy ? keep.y : lo.y};
    V2 nhi{wx ? keep.x + size : hi.x, wy ? keep.y + size : hi.y};
    lo = nlo;
    hi = nhi;
  }
  out.c.apex = apex;
  out.c.right = right;
  out.c.up = up;
  out.c.forward = forward;
  out.c.window[0] = lo.x;
  out.c.window[1] = lo.y;
  out.c.window[2] = hi.x;
  out.c.window[3] = hi.y;
  out.c.far_normal = forward;
  out.c.far_offset = dot(forward, apex) + dist * ONE_MINUS_1EM5;
  V3 axis = normalize(rel + (edge_u + edge_v) * 0.5f);
  if (qs.wide_clipped) axis = normalize(direction(out.c, 0.5f, 0.5f));
  out.c = bound_cone(out.c, axis);
  out.ok = true;
  return out;
}

// WGSL :2148-2190
static float rect_open(V3 point, V3 normal, const Light& r) {
  V3 u = xyz(r.edge_u), v = xyz(r.edge_v);
  V3 c0 = xyz(r.corner) - point;
  float facing = dot(normalize(cross(u, v)), c0);
  if ((r.corner[3] > 0.5f && facing >= 0.0f) || std::fabs(facing) <= 1e-6f) return 0.0f;
  V3 d1 = c0 + u;
  V3 d2 = c0 + u + v;
  V3 d3 = c0 + v;
  if (dot(normal, c0) >= 0.0f && dot(normal, d1) >= 0.0f && dot(normal, d2) >= 0.0f && dot(normal, d3) >= 0.0f) {
    V3 a0 = normalize(c0), a1 = normalize(d1), a2 = normalize(d2), a3 = normalize(d3);
    float sum = 0.0f;
    sum += lambert_edge_unit(normal, a0, a1);
    sum += lambert_edge_unit(normal, a1, a2);
    sum += lambert_edge_unit(normal, a2, a3);
    sum += lambert_edge_unit(normal, a3, a0);
    return std::fabs(0.5f * sum);
  }
  V3 d[4] = {c0, c0 + u, c0 + u + v, c0 + v};
  return clipped_quad_weight(normal, d, false);
}

// WGSL :1070-1076
static float wide_visible(const Query& qs, const Cone& c, V3 apex, V3 normal, const Light& r, float visible) {
  if (!qs.wide_clipped) return visible;
  float kept = window_open(c, normal);
  return (kept > 1e-9f ? visible / kept : 1.0f) * rect_open(apex, normal, r);
}

struct Result {
  float visible;
  uint64_t full, touched;
  bool capped;
  bool overflow;
};

// WGSL :1286-1321 and :2125-2128 (hit_irradiance_rect).
static Result hit_irradiance_rect(const Scene& s, const Options& o, V3 point, V3 normal, const Light& r,
                                  float* cov_out, uint32_t skip_obj = NONE, uint32_t skip_tri = NONE) {
  Query qs;
  qs.s = &s;
  qs.o = &o;
  qs.skip_obj = skip_obj;
  qs.skip_tri = skip_tri;
  Result res{0.0f, 0, 0, false, false};
  float spot = spot_falloff(r, normalize(point - light_center(r)));
  bool on = spot > 0.0f && faces_light(r, point, normal);
  float cov[64];
  for (int b = 0; b < 64; b++) cov[b] = 0.0f;
  float value = 0.0f;
  if (o.profile_stage != 1) {
    LightCone lc = light_cone_rect(qs, point, r);
    bool traced = on && lc.ok;
    uint32_t full = 0;
    qs.cone_cap = qs.wide_window ? o.wide_cone_cap : o.light_cone_cap;
    qs.cone_work = 0;
    if (o.profile_stage != 2 && traced) scene_coverage(lc.c, qs, cov, full);
    if (qs.cone_work >= qs.cone_cap && qs.wide_window && !qs.wide_clipped) {
      for (int b = 0; b < 64; b++) cov[b] = 1.0f;
      full = BINS;
    }
    res.capped = qs.cone_work >= qs.cone_cap;
    if (traced) {
      if (is_sphere(r))
        value = sphere_unshadowed(point, normal, r) * sphere_open_share(qs, lc.c, normal, cov);
      else
        value = wide_visible(qs, lc.c, point, normal, r, cone_visible(lc.c, normal, cov));
    }
  }
  res.visible = spot * value;
  for (int b = 0; b < 64; b++) {
    if (cov[b] >= 1.0f) res.full |= (uint64_t)1 << b;
    if (cov[b] > 0.0f) res.touched |= (uint64_t)1 << b;
  }
  res.overflow = qs.stack_overflow;
  if (cov_out)
    for (int b = 0; b < 64; b++) cov_out[b] = cov[b];
  return res;
}

// ---- Sky light: a uniform sky through six cube-face mask cones (WGSL
// :1323-1510), faces on the world axes (`sky_frame` NONE), no skipped object.

// WGSL :1326: far enough that every scene lies inside a sky cone.
static const float SKY_FAR = 1e7f;

struct SkyFace {
  Cone c;
  bool above;
};

// WGSL :1444-1478: cube face `f` around `apex`, its horizon `normal` when
// `horizon_cull` (params.w bit 2 clear), and whether any of its directions
// lies above that horizon (and `on`).
static SkyFace sky_face(bool on, uint32_t f, V3 apex, V3 normal, bool horizon_cull) {
  const uint32_t axis = f / 2u;
  const float sign = (f & 1u) ? -1.0f : 1.0f;
  float fw[3] = {0.0f, 0.0f, 0.0f}, rt[3] = {0.0f, 0.0f, 0.0f};
  fw[axis] = sign;
  rt[(axis + 1u) % 3u] = 1.0f;
  const V3 forward{fw[0], fw[1], fw[2]}, right{rt[0], rt[1], rt[2]};
  SkyFace face;
  face.c = zero_cone();
  face.c.apex = apex;
  face.c.right = right;
  face.c.up = cross(forward, right);
  face.c.forward = forward;
  face.c.window[0] = -1.0f;
  face.c.window[1] = -1.0f;
  face.c.window[2] = 1.0f;
  face.c.window[3] = 1.0f;
  face.c.far_normal = forward;
  face.c.far_offset = dot(forward, apex) + SKY_FAR;
  if (horizon_cull) face.c.horizon = normal;
  face.c = bound_cone(face.c, forward);
  bool above = false;
  for (uint32_t k = 0; k < 4; k++)
    if (dot(normal, direction(face.c, (float)(k & 1u), (float)(k >> 1u))) > 0.0f) above = true;
  face.above = above && on;
  return face;
}

struct SkyOptions {
  bool max_combine = false;   // params.w bit 1 (`sky_max`, a diagnostic)
  bool horizon_cull = true;   // params.w bit 2 clear (`sky_horizon_cull`)
  float prefilter_scale = 1.0f;  // light.objects.y (`sky_prefilter`)
};

// WGSL :1482-1488 (`sky_face_coverage`) and :1492-1501 (`sky_face_light`):
// the face's irradiance per unit sky radiance; its coverage into `cov`.
static float sky_face_irradiance(const Scene& s, const Options& o, const SkyOptions& so, bool on, uint32_t f,
                                 V3 apex, V3 normal, float* cov, uint64_t* full_mask) {
  if (o.profile_stage == 1) return 0.0f;
  Query qs;
  qs.s = &s;
  qs.o = &o;
  const float scale = so.prefilter_scale > 0.0f ? so.prefilter_scale : 1.0f;
  qs.prefilter_bins = o.prefilter * scale;
  const SkyFace face = sky_face(on, f, apex, normal, so.horizon_cull);
  for (uint32_t b = 0; b < BINS; b++) cov[b] = 0.0f;
  uint32_t full = 0;
  qs.independent = !so.max_combine;
  // A lane with `above` false follows the packet without acting: no coverage.
  if (o.profile_stage != 2 && face.above) scene_coverage(face.c, qs, cov, full);
  uint64_t fm = 0;
  for (uint32_t b = 0; b < BINS; b++)
    if (cov[b] >= 1.0f) fm |= 1ull << b;
  if (full_mask) *full_mask = fm;
  if (!face.above) return 0.0f;
  float blocked = 0.0f;
  for (uint32_t b = 0; b < BINS; b++)
    if (cov[b] > 0.0f) blocked += cov[b] * clipped_bin_weight(face.c, normal, b);
  return fmax_(clipped_window_weight(face.c, normal) - blocked, 0.0f);
}

// ---- The lighting cache: lightmap blocks of 8 x 8 texels re-baked when
// meshes move (lightmap.rs:256-333; WGSL :2408-2520; live.rs:515-720, 815-900).

constexpr uint32_t BLOCK_PLANES = 16;  // lightmap.rs:83

// glam's Vec3::normalize (v * (1 / length)), for host-side planes.
static V3 glam_normalize(V3 a) { return a * (1.0f / length(a)); }

// Lightmap::block_planes (lightmap.rs:256-333) for one block: the hull of
// the block's box and the light's corners, pushed out by 1e-3; or one plane
// that excludes everything when no texel can receive the light.
static void block_planes_one(const float* texels, const float* box, const Light& light, float* out) {
  const float MARGIN = 1e-3f;
  const V3 corner = xyz(light.corner), eu = xyz(light.edge_u), ev = xyz(light.edge_v);
  const bool one_sided = light.corner[3] > 0.5f;
  const V3 center = corner + (eu + ev) * 0.5f;
  const V3 emits = glam_normalize(cross(eu, ev));
  const V3 light_corners[4] = {corner, corner + eu, corner + ev, corner + eu + ev};
  std::vector<std::array<float, 4>> planes;
  bool lit = false;
  for (int t = 0; t < 64 && !lit; t++) {
    V3 p = xyz(texels + 8 * t), n = xyz(texels + 8 * t + 4);
    bool nonzero = n.x != 0.0f || n.y != 0.0f || n.z != 0.0f;
    lit = nonzero && dot(n, center - p) > -1e-3f && (!one_sided || dot(emits, p - corner) > -1e-3f);
  }
  if (!lit) {
    planes.push_back({0.0f, 0.0f, 0.0f, -1.0f});
  } else {
    const V3 lo = xyz(box), hi = xyz(box + 3);
    std::vector<V3> pts;
    for (int k = 0; k < 8; k++) {
      V3 p{(k & 1) ? hi.x : lo.x, (k & 2) ? hi.y : lo.y, (k & 4) ? hi.z : lo.z};
      bool seen = false;
      for (const V3& q : pts) seen = seen || (q.x == p.x && q.y == p.y && q.z == p.z);
      if (!seen) pts.push_back(p);
    }
    for (const V3& c : light_corners) pts.push_back(c);
    for (size_t i = 0; i < pts.size(); i++)
      for (size_t j = i + 1; j < pts.size(); j++)
        for (size_t k = j + 1; k < pts.size(); k++) {
          V3 n = cross(pts[j] - pts[i], pts[k] - pts[i]);
          if (dot(n, n) < 1e-14f) continue;
          n = glam_normalize(n);
          float d = dot(n, pts[i]);
          float below = 0.0f, above = 0.0f;
          for (const V3& p : pts) {
            below = std::fmin(below, dot(n, p) - d);
            above = std::fmax(above, dot(n, p) - d);
          }
          if (above <= 1e-6f) {
          } else if (below >= -1e-6f) {
            n = -n;
            d = -d;
          } else {
            continue;
          }
          bool dup = false;
          for (const auto& q : planes)
            dup = dup || (dot(V3{q[0], q[1], q[2]}, n) > 1.0f - 1e-6f && std::fabs(q[3] - MARGIN - d) < 1e-6f);
          if (dup) continue;
          planes.push_back({n.x, n.y, n.z, d + MARGIN});
        }
  }
  planes.resize(BLOCK_PLANES, {0.0f, 0.0f, 0.0f, 0.0f});
  for (uint32_t p = 0; p < BLOCK_PLANES; p++)
    for (int c = 0; c < 4; c++) out[4 * p + c] = planes[p][c];
}

// WGSL :2430-2452 mark_blocks for one block: true unless a plane separates
// every swept box from it.
static bool block_marked(const float* planes, const float* boxes, uint32_t n_boxes) {
  for (uint32_t k = 0; k < n_boxes; k++) {
    const V3 lo = xyz(boxes + 8 * k), hi = xyz(boxes + 8 * k + 4);
    const V3 center = (lo + hi) * 0.5f, extent = (hi - lo) * 0.5f;
    bool separated = false;
    for (uint32_t p = 0; p < BLOCK_PLANES && !separated; p++) {
      const float* q = planes + 4 * p;
      const V3 n = xyz(q);
      separated = dot(n, center) - dot(vabs(n), extent) > q[3];
    }
    if (!separated) return true;
  }
  return false;
}

// WGSL :2468-2487 bake_static for texel `t` of a block with a slot: the
// coverage of the meshes in the scene's (static) object hierarchy, uncapped.
static void static_coverage(const Scene& s, const Options& o, const float* texel, const Light& light, float* cov) {
  Query qs;
  qs.s = &s;
  qs.o = &o;
  for (int b = 0; b < 64; b++) cov[b] = 0.0f;
  const V3 point = xyz(texel), normal = xyz(texel + 4);
  const V3 to_light = xyz(light.corner) + (xyz(light.edge_u) + xyz(light.edge_v)) * 0.5f - point;
  if (!(dot(normal, to_light) > 0.0f)) return;
  LightCone lc = light_cone_rect(qs, point, light);
  if (!lc.ok) return;
  uint32_t full = 0;
  scene_coverage(lc.c, qs, cov, full);
}

// WGSL :2491-2510 cached_irradiance: the static coverage plus every moving
// mesh (in order) whose padded bounds the cone may touch, then the window's
// visible weight.
static float cached_irradiance(const Scene& s, const Options& o, const float* texel, const Light& light,
                               const float* static_cov, const uint32_t* moving, const float* moving_boxes,
                               uint32_t n_moving) {
  Query qs;
  qs.s = &s;
  qs.o = &o;
  const V3 apex = xyz(texel), normal = xyz(texel + 4);
  const V3 to_light = xyz(light.corner) + (xyz(light.edge_u) + xyz(light.edge_v)) * 0.5f - apex;
  if (dot(normal, to_light) <= 0.0f) return 0.0f;
  LightCone lc = light_cone_rect(qs, apex, light);
  if (!lc.ok) return 0.0f;
  float cov[64];
  uint32_t full = 0;
  for (int b = 0; b < 64; b++) {
    cov[b] = static_cov[b];
    if (cov[b] >= 1.0f) full++;
  }
  for (uint32_t k = 0; k < n_moving; k++) {
    if (full == BINS) break;
    if (!may_touch_box(lc.c, xyz(moving_boxes + 8 * k), xyz(moving_boxes + 8 * k + 4))) continue;
    add_mesh_coverage(lc.c, qs, moving[k], cov, full, 0);
  }
  return wide_visible(qs, lc.c, apex, normal, light, cone_visible(lc.c, normal, cov));
}

}  // namespace orc

using namespace orc;

static Scene make_scene(const float* tris, const uint32_t* ranges, uint32_t n_objects, const float* blas,
                        const uint32_t* blas_items, const float* blas_cov, const float* nodes, const uint32_t* items,
                        const float* xforms) {
  Scene s;
  s.tris = tris;
  s.ranges = ranges;
  s.n_objects = n_objects;
  s.blas = reinterpret_cast<const Node*>(blas);
  s.blas_items = blas_items;
  s.blas_cov = blas_cov;
  s.nodes = reinterpret_cast<const Node*>(nodes);
  s.items = items;
  s.xforms = xforms;
  return s;
}

static Options make_options(const uint32_t* opts) {
  Options o;
  o.hierarchy = opts[0];
  std::memcpy(&o.prefilter, &opts[1], 4);
  std::memcpy(&o.wide_extent, &opts[2], 4);
  o.light_cone_cap = opts[3];
  o.wide_cone_cap = opts[4];
  o.profile_stage = opts[5];
  o.object_cov = opts[6];
  return o;
}

// fn(i) for i in [0, n), over `threads` threads in chunks of 256.
template <class F>
static void parallel_for(uint64_t n, int threads, F fn) {
  std::atomic<uint64_t> next{0};
  auto work = [&]() {
    for (;;) {
      uint64_t i0 = next.fetch_add(256);
      if (i0 >= n) break;
      for (uint64_t i = i0; i < std::min(n, i0 + 256); i++) fn(i);
    }
  };
  if (threads <= 1) {
    work();
    return;
  }
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; t++) pool.emplace_back(work);
  for (auto& t : pool) t.join();
}

extern "C" {

// One query per (receiver i, light l), stored at l * n + i. Receivers are
// (point vec4, normal vec4), the reference's ReceiverGpu; one whose normal.w
// is negative traces nothing (a pixel that sees the sky or a light). Lights
// are the reference's LightGpu (64 B).
// opts: [hierarchy, prefilter(f32 bits), wide_extent(f32 bits),
// light_cone_cap, wide_cone_cap, profile_stage, object_cov]. Returns the
// number of queries whose stack would overflow the reference's 48 entries.
int oracle_irradiance(const float* tris, const uint32_t* ranges, uint32_t n_objects, const float* blas,
                      const uint32_t* blas_items, const float* blas_cov, const float* nodes, const uint32_t* items,
                      const float* xforms, const float* receivers, uint64_t n, const float* lights, uint32_t n_lights,
                      const uint32_t* opts, float* out_visible, uint64_t* out_full, uint64_t* out_touched,
                      float* out_cov, uint8_t* out_capped, int threads) {
  Scene s = make_scene(tris, ranges, n_objects, blas, blas_items, blas_cov, nodes, items, xforms);
  Options o;
  o.hierarchy = opts[0];
  std::memcpy(&o.prefilter, &opts[1], 4);
  std::memcpy(&o.wide_extent, &opts[2], 4);
  o.light_cone_cap = opts[3];
  o.wide_cone_cap = opts[4];
  o.profile_stage = opts[5];
  o.object_cov = opts[6];
  const uint64_t total = n * (uint64_t)n_lights;
  std::atomic<uint64_t> next{0};
  std::atomic<int> overflows{0};
  auto work = [&]() {
    for (;;) {
      uint64_t q0 = next.fetch_add(256);
      if (q0 >= total) break;
      uint64_t q1 = std::min(total, q0 + 256);
      for (uint64_t q = q0; q < q1; q++) {
        uint64_t i = q % n;
        uint32_t l = (uint32_t)(q / n);
        const float* rc = receivers + 8 * i;
        const Light& light = reinterpret_cast<const Light*>(lights)[l];
        Result r{0.0f, 0, 0, false, false};
        if (rc[7] >= 0.0f) {
          r = hit_irradiance_rect(s, o, xyz(rc), xyz(rc + 4), light, out_cov ? out_cov + 64 * q : nullptr);
        } else if (out_cov) {
          for (int b = 0; b < 64; b++) out_cov[64 * q + b] = 0.0f;
        }
        out_visible[q] = r.visible;
        out_full[q] = r.full;
        out_touched[q] = r.touched;
        if (out_capped) out_capped[q] = r.capped ? 1 : 0;
        if (r.overflow) overflows++;
      }
    }
  };
  if (threads <= 1) {
    work();
  } else {
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; t++) pool.emplace_back(work);
    for (auto& t : pool) t.join();
  }
  return overflows.load();
}

// Single building blocks, for the tests that pin them.
float oracle_atan2(float y, float x) { return atan2_c(y, x); }
// WGSL :1416-1426 (`sky_irradiance_if`) per receiver (ReceiverGpu; a
// negative normal.w traces nothing): out[i] the sum over the six faces in
// order, out_faces[f * n + i] each face's share, out_full[f * n + i] the
// face's full bins (coverage 1). sky_opts: [max_combine, horizon_cull,
// prefilter_scale bits].
void oracle_sky(const float* tris, const uint32_t* ranges, uint32_t n_objects, const float* blas,
                const uint32_t* blas_items, const float* blas_cov, const float* nodes, const uint32_t* items,
                const float* xforms, const float* receivers, uint64_t n, const uint32_t* opts,
                const uint32_t* sky_opts, float* out, float* out_faces, uint64_t* out_full, int threads) {
  Scene s = make_scene(tris, ranges, n_objects, blas, blas_items, blas_cov, nodes, items, xforms);
  Options o = make_options(opts);
  SkyOptions so;
  so.max_combine = sky_opts[0] != 0;
  so.horizon_cull = sky_opts[1] != 0;
  std::memcpy(&so.prefilter_scale, &sky_opts[2], 4);
  parallel_for(n, threads, [&](uint64_t i) {
    const float* rc = receivers + 8 * i;
    const bool on = rc[7] >= 0.0f;
    float total = 0.0f;
    for (uint32_t f = 0; f < 6; f++) {
      float cov[64];
      uint64_t fm = 0;
      const float e = sky_face_irradiance(s, o, so, on, f, xyz(rc), xyz(rc + 4), cov, &fm);
      total += e;
      if (out_faces) out_faces[f * n + i] = e;
      if (out_full) out_full[f * n + i] = fm;
    }
    out[i] = total;
  });
}

// One sky face's coverage for one point (tests): cov[64].
float oracle_sky_face(const float* tris, const uint32_t* ranges, uint32_t n_objects, const float* blas,
                      const uint32_t* blas_items, const float* blas_cov, const float* nodes, const uint32_t* items,
                      const float* xforms, const float* point, const float* normal, uint32_t face,
                      const uint32_t* opts, const uint32_t* sky_opts, float* cov) {
  Scene s = make_scene(tris, ranges, n_objects, blas, blas_items, blas_cov, nodes, items, xforms);
  Options o = make_options(opts);
  SkyOptions so;
  so.max_combine = sky_opts[0] != 0;
  so.horizon_cull = sky_opts[1] != 0;
  std::memcpy(&so.prefilter_scale, &sky_opts[2], 4);
  return sky_face_irradiance(s, o, so, true, face, xyz(point), xyz(normal), cov, nullptr);
}

float oracle_acos(float y) { return acos_c(y); }
float oracle_exp(float x) { return exp_c(x); }

// The accumulation rasterizer on one polygon (points in bin units), for
// closed-form checks: writes 68 accumulator cells and the coverage of the
// 64 bins after the running sum (no snap).
void oracle_draw_polygon(const float* pts, uint32_t n, float weight, float* acc_out, float* cov_out) {
  Acc acc;
  for (int b = 0; b < 68; b++) acc.a[b] = 0.0f;
  acc.span_lo = 68;
  acc.span_hi = 0;
  Poly2 q;
  q.n = n;
  for (uint32_t i = 0; i < n; i++) q.v[i] = V2{pts[2 * i], pts[2 * i + 1]};
  draw_poly(acc, q, weight);
  for (int b = 0; b < 68; b++) acc_out[b] = acc.a[b];
  float running = 0.0f;
  for (int b = 0; b < 64; b++) {
    running += acc.a[b];
    cov_out[b] = running;
  }
}

// The light cone's frame and window for one apex (for closed-form checks):
// out = apex3 right3 up3 forward3 window4 far_normal3 far_offset bound4 bound_sin, ok.
int oracle_light_cone(const float* apex, const float* light, const uint32_t* opts, float* out) {
  Options o{};
  o.hierarchy = opts[0];
  std::memcpy(&o.prefilter, &opts[1], 4);
  std::memcpy(&o.wide_extent, &opts[2], 4);
  Query qs;
  qs.o = &o;
  LightCone lc = light_cone_rect(qs, xyz(apex), *reinterpret_cast<const Light*>(light));
  const Cone& c = lc.c;
  float v[] = {c.apex.x,      c.apex.y,      c.apex.z,      c.right.x,       c.right.y,       c.right.z,
               c.up.x,        c.up.y,        c.up.z,        c.forward.x,     c.forward.y,     c.forward.z,
               c.window[0],   c.window[1],   c.window[2],   c.window[3],     c.far_normal.x,  c.far_normal.y,
               c.far_normal.z, c.far_offset, c.bound[0],    c.bound[1],      c.bound[2],      c.bound[3],
               c.bound_sin};
  std::memcpy(out, v, sizeof(v));
  return lc.ok ? (qs.wide_clipped ? 3 : (qs.wide_window ? 2 : 1)) : 0;
}

// Bin weights of a light cone for a receiver normal (64 values) and the
// window's total weight; returns 0 if no cone.
int oracle_bin_weights(const float* apex, const float* normal, const float* light, float wide_extent,
                       float* out64, float* out_total) {
  Options o{};
  o.wide_extent = wide_extent;
  Query qs;
  qs.o = &o;
  LightCone lc = light_cone_rect(qs, xyz(apex), *reinterpret_cast<const Light*>(light));
  if (!lc.ok) return 0;
  for (uint32_t b = 0; b < 64; b++) out64[b] = clipped_bin_weight(lc.c, xyz(normal), b);
  *out_total = window_open(lc.c, xyz(normal));
  return 1;
}

// One query (WGSL :2125 hit_irradiance_rect with `on` = true) with the
// triangle `skip_tri` of object `skip_obj` unseen (NONE: none), for the view
// cache's restatement: visible, then the full and touched masks.
void oracle_query_skip(const float* tris, const uint32_t* ranges, uint32_t n_objects, const float* blas,
                       const uint32_t* blas_items, const float* blas_cov, const float* nodes, const uint32_t* items,
                       const float* xforms, const uint32_t* opts, const float* point, const float* normal,
                       const float* light, uint32_t skip_obj, uint32_t skip_tri, float* visible, uint64_t* masks) {
  Scene s = make_scene(tris, ranges, n_objects, blas, blas_items, blas_cov, nodes, items, xforms);
  Options o = make_options(opts);
  Result r = hit_irradiance_rect(s, o, xyz(point), xyz(normal), *reinterpret_cast<const Light*>(light), nullptr,
                                 skip_obj, skip_tri);
  *visible = r.visible;
  masks[0] = r.full;
  masks[1] = r.touched;
}

// WGSL :2135-2141 `unshadowed_rect`.
float oracle_unshadowed_rect(const float* point, const float* normal, const float* light) {
  const Light& r = *reinterpret_cast<const Light*>(light);
  V3 p = xyz(point), n = xyz(normal);
  if (!faces_light(r, p, n)) return 0.0f;
  if (is_sphere(r)) return spot_falloff(r, normalize(p - xyz(r.corner))) * sphere_unshadowed(p, n, r);
  return rect_open(p, n, r);
}

// WGSL :880 `faces_light`, :866 `light_center`.
int oracle_faces_light(const float* light, const float* point, const float* normal) {
  return faces_light(*reinterpret_cast<const Light*>(light), xyz(point), xyz(normal)) ? 1 : 0;
}

void oracle_light_center(const float* light, float* out) {
  V3 c = light_center(*reinterpret_cast<const Light*>(light));
  out[0] = c.x, out[1] = c.y, out[2] = c.z;
}

float oracle_rect_open(const float* point, const float* normal, const float* light) {
  return rect_open(xyz(point), xyz(normal), *reinterpret_cast<const Light*>(light));
}

float oracle_sphere_unshadowed(const float* point, const float* normal, const float* light) {
  return sphere_unshadowed(xyz(point), xyz(normal), *reinterpret_cast<const Light*>(light));
}

// Lightmap::block_planes: texels are 64 receivers (point vec4, normal vec4)
// per block, `boxes` 6 floats (lo, hi) per block, the light a LightGpu rect;
// 16 planes (n, w) per block out.
void oracle_block_planes(const float* texels, const float* boxes, uint32_t n_blocks, const float* light,
                         float* planes) {
  const Light& l = *reinterpret_cast<const Light*>(light);
  for (uint32_t b = 0; b < n_blocks; b++)
    block_planes_one(texels + 512 * (size_t)b, boxes + 6 * (size_t)b, l, planes + 64 * (size_t)b);
}

// mark_blocks: marked[b] = 1 when no plane of block b separates some swept
// box (8 floats each: lo vec4, hi vec4) from it.
void oracle_mark_blocks(const float* planes, uint32_t n_blocks, const float* boxes, uint32_t n_boxes,
                        uint8_t* marked) {
  for (uint32_t b = 0; b < n_blocks; b++) marked[b] = block_marked(planes + 64 * (size_t)b, boxes, n_boxes) ? 1 : 0;
}

// bake_static over every block with a slot (block_slot[b] != NONE): the
// static scene's coverage per texel, 64 bins each, at
// cov[(slot * 64 + texel) * 64 + bin] (the reference's layout). The scene's
// object hierarchy (nodes/items) holds the static meshes only. opts as
// oracle_irradiance's.
void oracle_bake_static(const float* tris, const uint32_t* ranges, uint32_t n_objects, const float* blas,
                        const uint32_t* blas_items, const float* blas_cov, const float* nodes, const uint32_t* items,
                        const float* xforms, const float* texels, const uint32_t* block_slot, uint32_t n_blocks,
                        const float* light, const uint32_t* opts, float* cov, int threads) {
  Scene s = make_scene(tris, ranges, n_objects, blas, blas_items, blas_cov, nodes, items, xforms);
  Options o = make_options(opts);
  const Light& l = *reinterpret_cast<const Light*>(light);
  parallel_for(n_blocks * 64ull, threads, [&](uint64_t i) {
    uint32_t slot = block_slot[i / 64];
    if (slot == NONE) return;
    static_coverage(s, o, texels + 8 * i, l, cov + 64 * (64 * (size_t)slot + i % 64));
  });
}

// rebake over `blocks`: every texel of each block from its cached static
// coverage plus the moving meshes (`moving` mesh indices with padded bounds,
// 8 floats each), into irradiance[block * 64 + texel].
void oracle_rebake(const float* tris, const uint32_t* ranges, uint32_t n_objects, const float* blas,
                   const uint32_t* blas_items, const float* blas_cov, const float* xforms, const float* texels,
                   const uint32_t* block_slot, const float* static_cov, const uint32_t* blocks, uint32_t n_list,
                   const float* light, const uint32_t* moving, const float* moving_boxes, uint32_t n_moving,
                   const uint32_t* opts, float* irradiance, int threads) {
  Scene s = make_scene(tris, ranges, n_objects, blas, blas_items, blas_cov, nullptr, nullptr, xforms);
  Options o = make_options(opts);
  const Light& l = *reinterpret_cast<const Light*>(light);
  parallel_for(n_list * 64ull, threads, [&](uint64_t j) {
    uint32_t block = blocks[j / 64], t = (uint32_t)(j % 64);
    uint32_t slot = block_slot[block];
    uint64_t i = 64ull * block + t;
    irradiance[i] = slot == NONE ? 0.0f
                                 : cached_irradiance(s, o, texels + 8 * i, l, static_cov + 64 * (64 * (size_t)slot + t),
                                                     moving, moving_boxes, n_moving);
  });
}

}  // extern "C"

// ---- Camera rays: shading points for the plain per-view launch
// (WGSL :1524-1760 `camera_dir`, `ray_triangle`, `hit_triangle`,
// `ray_box_entry`, `push_children`, `nearest_in_mesh`, `primary_hit`;
// :2062-2115 `surface_offset`, `rect_distance`, `seen_light`). Camera rays
// cull back faces (`camera_hit` :2222).
namespace orc {

struct Hit {
  float distance;
  uint32_t mesh, tri;
  V3 normal;
};

constexpr float EDGE_TOLERANCE = 1e-5f;

static V3 rotate_to_world(const Scene& s, uint32_t x, V3 d) {
  const float* r = s.xforms + 16 * x;
  return V3{dot(xyz(r), d), dot(xyz(r + 4), d), dot(xyz(r + 8), d)};
}

static float ray_triangle(V3 o, V3 d, V3 v0, V3 v1, V3 v2) {
  V3 e1 = v1 - v0, e2 = v2 - v0;
  V3 p = cross(d, e2);
  float det = dot(e1, p);
  if (std::fabs(det) < 1e-12f) return -1.0f;
  float inv = 1.0f / det;
  V3 s = o - v0;
  float u = dot(s, p) * inv;
  if (u < -EDGE_TOLERANCE || u > 1.0f + EDGE_TOLERANCE) return -1.0f;
  V3 q = cross(s, e1);
  float v = dot(d, q) * inv;
  if (v < -EDGE_TOLERANCE || u + v > 1.0f + EDGE_TOLERANCE) return -1.0f;
  return dot(e2, q) * inv;
}

static Hit hit_triangle(const Scene& s, V3 eye, V3 dir, uint32_t m, uint32_t t, Hit best) {
  const float* tr = s.tris + 12u * t;
  V3 v0 = xyz(tr), v1 = xyz(tr + 4), v2 = xyz(tr + 8);
  V3 n = cross(v1 - v0, v2 - v0);
  if (dot(n, dir) > 0.0f) return best;  // cull_back
  float d = ray_triangle(eye, dir, v0, v1, v2);
  if (d > 0.0f && m == best.mesh && std::fabs(d - best.distance) <= 2e-6f * d) {
    bool facing = dot(n, dir) < 0.0f;
    if (facing != (dot(best.normal, dir) < 0.0f)) {
      if (facing) return Hit{d, m, t, normalize(n)};
      return best;
    }
  }
  if (d > 0.0f &&
      (d < best.distance || (d == best.distance && (m < best.mesh || (m == best.mesh && t < best.tri)))))
    return Hit{d, m, t, normalize(cross(v1 - v0, v2 - v0))};
  return best;
}

static float ray_box_entry(V3 eye, V3 inv, V3 lo, V3 hi, float limit) {
  V3 t0{(lo.x - eye.x) * inv.x, (lo.y - eye.y) * inv.y, (lo.z - eye.z) * inv.z};
  V3 t1{(hi.x - eye.x) * inv.x, (hi.y - eye.y) * inv.y, (hi.z - eye.z) * inv.z};
  float nearv = fmax_(fmax_(fmin_(t0.x, t1.x), fmin_(t0.y, t1.y)), fmin_(t0.z, t1.z));
  float farv = fmin_(fmin_(fmax_(t0.x, t1.x), fmax_(t0.y, t1.y)), fmax_(t0.z, t1.z));
  if (farv < fmax_(nearv, 0.0f) || nearv > limit) return -1.0f;
  return fmax_(nearv, 0.0f);
}

struct RayWalk {
  uint32_t stack[STACK];
  bool overflow = false;
  bool nearest_first = true;
};

static void push_children(RayWalk& w, V3 eye, V3 inv, uint32_t left, const Node& l, const Node& r, float limit,
                          uint32_t& sp) {
  if (sp + 2 > STACK) {
    w.overflow = true;
    return;
  }
  if (!w.nearest_first) {
    w.stack[sp] = left;
    w.stack[sp + 1] = left + 1;
    sp += 2;
    return;
  }
  float tl = ray_box_entry(eye, inv, xyz(l.lo), xyz(l.hi), limit);
  float tr = ray_box_entry(eye, inv, xyz(r.lo), xyz(r.hi), limit);
  if (tl >= 0.0f && tr >= 0.0f) {
    if (tl <= tr) {
      w.stack[sp] = left + 1;
      w.stack[sp + 1] = left;
    } else {
      w.stack[sp] = left;
      w.stack[sp + 1] = left + 1;
    }
    sp += 2;
  } else if (tl >= 0.0f) {
    w.stack[sp++] = left;
  } else if (tr >= 0.0f) {
    w.stack[sp++] = left + 1;
  }
}

static V3 recip3(V3 d) { return V3{1.0f / d.x, 1.0f / d.y, 1.0f / d.z}; }

static Hit nearest_in_mesh(const Scene& s, uint32_t hierarchy, RayWalk& w, V3 world_eye, V3 world_dir, uint32_t m,
                           Hit best, uint32_t base) {
  Hit hit = best;
  const uint32_t* range = s.ranges + 4 * m;
  V3 eye = world_eye, dir = world_dir;
  if (range[3] != NONE) {
    eye = point_to_object(s, m, world_eye);
    dir = rotate_to_object(s, range[3], world_dir) * s.xforms[16 * range[3] + 13];
  }
  if (hierarchy == 2) {
    if (range[2] == NONE) return hit;
    V3 inv = recip3(dir);
    if (base >= STACK) {
      w.overflow = true;
      return hit;
    }
    uint32_t sp = base + 1;
    w.stack[base] = range[2];
    while (sp > base && !w.overflow) {
      sp--;
      const Node& node = s.blas[w.stack[sp]];
      if (ray_box_entry(eye, inv, xyz(node.lo), xyz(node.hi), hit.distance) < 0.0f) continue;
      if (node.info[1] > 0) {
        for (uint32_t k = node.info[0]; k < node.info[0] + node.info[1]; k++)
          hit = hit_triangle(s, eye, dir, m, s.blas_items[k], hit);
      } else {
        push_children(w, eye, inv, node.info[0], s.blas[node.info[0]], s.blas[node.info[0] + 1], hit.distance, sp);
      }
    }
  } else {
    for (uint32_t t = range[0]; t < range[0] + range[1]; t++) hit = hit_triangle(s, eye, dir, m, t, hit);
  }
  if (range[3] != NONE && hit.mesh == m && best.mesh != m) hit.normal = normalize(rotate_to_world(s, range[3], hit.normal));
  return hit;
}

static Hit primary_hit(const Scene& s, uint32_t hierarchy, RayWalk& w, V3 eye, V3 dir) {
  Hit hit{3.4e38f, NONE, NONE, V3{0.0f, 0.0f, 0.0f}};
  if (hierarchy >= 1) {
    V3 inv = recip3(dir);
    uint32_t sp = 1;
    w.stack[0] = 0;
    while (sp > 0 && !w.overflow) {
      sp--;
      const Node& node = s.nodes[w.stack[sp]];
      if (ray_box_entry(eye, inv, xyz(node.lo), xyz(node.hi), hit.distance) < 0.0f) continue;
      if (node.info[1] > 0) {
        for (uint32_t k = node.info[0]; k < node.info[0] + node.info[1]; k++)
          hit = nearest_in_mesh(s, hierarchy, w, eye, dir, s.items[k], hit, sp);
      } else {
        push_children(w, eye, inv, node.info[0], s.nodes[node.info[0]], s.nodes[node.info[0] + 1], hit.distance, sp);
      }
    }
  } else {
    for (uint32_t m = 0; m < s.n_objects; m++) hit = nearest_in_mesh(s, hierarchy, w, eye, dir, m, hit, 0);
  }
  if (dot(hit.normal, dir) > 0.0f) hit.normal = -hit.normal;
  return hit;
}

static float rect_distance(const Light& l, V3 eye, V3 dir) {
  if (is_sphere(l)) {
    V3 oc = eye - xyz(l.corner);
    float b = dot(oc, dir);
    float c = dot(oc, oc) - l.edge_u[0] * l.edge_u[0];
    float disc = b * b - c;
    if (c <= 0.0f || disc < 0.0f) return 3.4e38f;
    float t = -b - std::sqrt(disc);
    if (t <= 0.0f) return 3.4e38f;
    return t;
  }
  if (l.corner[3] < 0.5f) return 3.4e38f;
  V3 u = xyz(l.edge_u), v = xyz(l.edge_v);
  V3 n = cross(u, v);
  float facing = dot(dir, n);
  if (facing >= 0.0f) return 3.4e38f;
  float t = dot(xyz(l.corner) - eye, n) / facing;
  if (t <= 0.0f) return 3.4e38f;
  V3 q = eye + dir * t - xyz(l.corner);
  float ss = dot(q, u) / dot(u, u);
  float rr = dot(q, v) / dot(v, v);
  if (ss < 0.0f || ss > 1.0f || rr < 0.0f || rr > 1.0f) return 3.4e38f;
  return t;
}

// WGSL :2047-2058 `shading_normal` (with `barycentric` :1780-1793,
// `point_to_object` :74-78, `normal_to_world` :81-85): triangle t's corner
// normals (9 floats per triangle in `normals`) at world point p, the face's
// where they sum to nothing, in world space on the side `toward` points to.
static V3 shading_normal(const Scene& s, const float* normals, uint32_t m, uint32_t t, V3 p_world, V3 toward) {
  const uint32_t x = s.ranges[4 * m + 3];
  V3 p = point_to_object(s, m, p_world);
  const float* tr = s.tris + 12u * t;
  V3 v0 = xyz(tr), e1 = xyz(tr + 4) - v0, e2 = xyz(tr + 8) - v0;
  V3 d = p - v0;
  float d11 = dot(e1, e1), d12 = dot(e1, e2), d22 = dot(e2, e2);
  float denom = d11 * d22 - d12 * d12;
  float bs = (d22 * dot(d, e1) - d12 * dot(d, e2)) / denom;
  float bu = (d11 * dot(d, e2) - d12 * dot(d, e1)) / denom;
  float bt = 1.0f - bs - bu;
  V3 face = cross(e1, e2);
  const float* nr = normals + 9u * t;
  V3 sn = xyz(nr) * bt + xyz(nr + 3) * bs + xyz(nr + 6) * bu;
  V3 pick = dot(sn, sn) > 1e-20f ? sn : face;
  V3 n = x == NONE ? pick : rotate_to_world(s, x, pick);
  V3 geometric = x == NONE ? face : rotate_to_world(s, x, face);
  n = normalize(n);
  if (dot(geometric, toward) < 0.0f) n = -n;
  return n;
}

// camera_dir (WGSL :1769-1774). cam: eye4 forward4 right4 up4 params4.
static V3 camera_dir(const float* cam, uint32_t x, uint32_t y) {
  const float* p = cam + 16;
  float aspect = p[0] / p[1];
  float sx = (((float)x + 0.5f + cam[11]) / p[0] * 2.0f - 1.0f) * p[2] * aspect;
  float sy = (1.0f - ((float)y + 0.5f + cam[15]) / p[1] * 2.0f) * p[2];
  return normalize(xyz(cam + 4) + xyz(cam + 8) * sx + xyz(cam + 12) * sy);
}

}  // namespace orc

extern "C" {

// Per pixel of a w x h view: the camera ray's nearest front-facing hit and
// the shading point a cone starts from. Writes hits[4*p] = (mesh, tri,
// distance bits, seen light index + 1) and receivers[8*p] = (point, 0,
// normal, 0 where a surface is shaded, -1 where the pixel sees the sky or a
// light and no cone is traced). The shading normal is the hit
// triangle's face normal toward the camera (flat shading), or with
// `normals` (nullable; 9 floats per triangle) the interpolated one
// (`shading_normal`, as `per_pixel_hit_color` takes it, WGSL :2280). Returns
// the number of rays whose stack would overflow 48 entries.
int oracle_primary(const float* tris, const uint32_t* ranges, uint32_t n_objects, const float* blas,
                   const uint32_t* blas_items, const float* nodes, const uint32_t* items, const float* xforms,
                   uint32_t hierarchy, uint32_t nearest_first, const float* cam, uint32_t w, uint32_t h,
                   const float* lights, uint32_t n_lights, uint32_t* hits, float* receivers, int threads,
                   const float* normals) {
  Scene s = make_scene(tris, ranges, n_objects, blas, blas_items, nullptr, nodes, items, xforms);
  const Light* ls = reinterpret_cast<const Light*>(lights);
  std::atomic<uint64_t> next{0};
  std::atomic<int> overflows{0};
  const uint64_t total = (uint64_t)w * h;
  auto work = [&]() {
    for (;;) {
      uint64_t p0 = next.fetch_add(64);
      if (p0 >= total) break;
      uint64_t p1 = std::min(total, p0 + 64);
      for (uint64_t px = p0; px < p1; px++) {
        uint32_t x = (uint32_t)(px % w), y = (uint32_t)(px / w);
        V3 dir = camera_dir(cam, x, y);
        V3 eye = xyz(cam);
        RayWalk rw;
        rw.nearest_first = nearest_first != 0;
        Hit hit = primary_hit(s, hierarchy, rw, eye, dir);
        if (rw.overflow) overflows++;
        uint32_t seen = 0;
        float nearest = hit.distance;
        for (uint32_t l = 0; l < n_lights; l++) {
          float t = rect_distance(ls[l], eye, dir);
          if (t < nearest) {
            nearest = t;
            seen = l + 1;
          }
        }
        uint32_t dbits;
        std::memcpy(&dbits, &hit.distance, 4);
        hits[4 * px] = hit.mesh;
        hits[4 * px + 1] = hit.tri;
        hits[4 * px + 2] = dbits;
        hits[4 * px + 3] = seen;
        float* r = receivers + 8 * px;
        bool lit = seen == 0 && hit.mesh != NONE;
        V3 point{0.0f, 0.0f, 0.0f}, normal{0.0f, 0.0f, 0.0f};
        if (lit) {
          V3 surface = eye + dir * hit.distance;
          point = surface + hit.normal * fmax_(1e-4f, 2e-6f * hit.distance);
          normal = normals ? shading_normal(s, normals, hit.mesh, hit.tri, surface, -dir) : hit.normal;
        }
        r[0] = point.x, r[1] = point.y, r[2] = point.z, r[3] = 0.0f;
        r[4] = normal.x, r[5] = normal.y, r[6] = normal.z, r[7] = lit ? 0.0f : -1.0f;
      }
    }
  };
  if (threads <= 1) {
    work();
  } else {
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; t++) pool.emplace_back(work);
    for (auto& t : pool) t.join();
  }
  return overflows.load();
}

}  // extern "C"

// ---- Moving lights in the view cache: the guard cone and the lattice of a
// moving rect light (WGSL :4457-4530 `guard_cone`, :4532 `below_horizon`,
// :4689-4905 `rect_corner`, `lattice_frame`, `lattice_kappa`,
// `lattice_cone`, `lattice_areas`, `lattice_share`, `lattice_ratio`). Walks
// of these cones take no node cap: the WGSL sets `cone_cap` back to
// unlimited after each light cone (:1305).

namespace orc {

constexpr float GUARD = 3.0f;           // WGSL :4482
constexpr float LATTICE_MARGIN = 0.03f; // WGSL :4704

static V3 rect_corner(const Light& r, uint32_t k) {
  return xyz(r.corner) + xyz(r.edge_u) * (float)(k & 1u) + xyz(r.edge_v) * (float)((k >> 1u) & 1u);
}

struct GuardCone {
  Cone c;
  float alpha;  // the light's angular radius around its centre, seen from the apex
};

// WGSL :4499-4529
static GuardCone guard_cone(V3 apex, const Light& r) {
  GuardCone out;
  out.c = zero_cone();
  V3 to = light_center(r) - apex;
  float d = fmax_(length(to), 1e-6f);
  V3 forward = to / d;
  float alpha = 0.0f;
  if (is_sphere(r)) {
    alpha = asin_c(fmin_(r.edge_u[0] / d, 1.0f));
  } else {
    for (uint32_t k = 0; k < 4; k++)
      alpha = fmax_(alpha, acos_c(clampf(dot(forward, normalize(rect_corner(r, k) - apex)), -1.0f, 1.0f)));
  }
  float wide = fmin_(GUARD * alpha, 1.4f);
  float t = tan_c(wide);
  V3 helper = std::fabs(forward.x) > 0.9f ? V3{0.0f, 1.0f, 0.0f} : V3{1.0f, 0.0f, 0.0f};
  V3 right = normalize(cross(helper, forward));
  out.c.apex = apex;
  out.c.right = right;
  out.c.up = cross(forward, right);
  out.c.forward = forward;
  out.c.window[0] = -t;
  out.c.window[1] = -t;
  out.c.window[2] = t;
  out.c.window[3] = t;
  out.c.far_normal = forward;
  out.c.far_offset = dot(forward, apex) + SKY_FAR;
  out.c = bound_cone(out.c, forward);
  out.alpha = alpha;
  return out;
}

// WGSL :4532-4539
static float below_horizon(V3 apex, V3 n, const Light& r) {
  V3 to = light_center(r) - apex;
  float d = fmax_(length(to), 1e-6f);
  float below = -dot(n, to);
  if (is_sphere(r)) below -= r.edge_u[0];
  return asin_c(clampf(below / d, -1.0f, 1.0f));
}

struct LatticeFrame {
  V3 forward, right, up;
  V2 beta;  // a bin's size in the window's plane, per axis
  float alpha;
  bool ok;
};

// WGSL :4727-4760
static LatticeFrame lattice_frame(V3 apex, const Light& then) {
  LatticeFrame f;
  std::memset(&f, 0, sizeof(f));
  f.ok = false;
  V3 to = light_center(then) - apex;
  float d = length(to);
  if (d <= 1e-6f) return f;
  V3 forward = to / d;
  V3 eu = xyz(then.edge_u), ev = xyz(then.edge_v);
  V3 right = eu - forward * dot(forward, eu);
  if (length(right) <= 1e-6f * length(eu)) right = ev - forward * dot(forward, ev);
  if (length(right) <= 0.0f) return f;
  right = normalize(right);
  V3 up = cross(forward, right);
  V2 lo{3.4e38f, 3.4e38f}, hi{-3.4e38f, -3.4e38f};
  float alpha = 0.0f;
  for (uint32_t k = 0; k < 4; k++) {
    V3 p = rect_corner(then, k) - apex;
    float z = dot(forward, p);
    if (z <= 1e-6f * d) return f;
    V2 uv{dot(right, p) / z, dot(up, p) / z};
    lo = V2{fmin_(lo.x, uv.x), fmin_(lo.y, uv.y)};
    hi = V2{fmax_(hi.x, uv.x), fmax_(hi.y, uv.y)};
    alpha = fmax_(alpha, atan2_c(length(cross(forward, p)), z));
  }
  f.forward = forward;
  f.right = right;
  f.up = up;
  const float share = (1.0f + LATTICE_MARGIN) / 7.0f;
  f.beta = V2{fmax_(hi.x - lo.x, 1e-12f) * share, fmax_(hi.y - lo.y, 1e-12f) * share};
  f.alpha = alpha * 1.0001f;
  f.ok = true;
  return f;
}

// WGSL :4766-4773; (1 + t^2)^1.5 as q * sqrt(q).
static float lattice_kappa(const LatticeFrame& f, V3 n) {
  if (!f.ok) return 3.0e38f;
  float reach = fmin_(GUARD * f.alpha, 1.4f);
  float theta = atan2_c(length(cross(n, f.forward)), dot(n, f.forward));
  if (theta + reach >= 1.5706f) return 3.0e38f;
  float t = tan_c(reach);
  float q = 1.0f + t * t;
  return cos_c(fmax_(theta - reach, 0.0f)) / cos_c(theta + reach) * (q * std::sqrt(q)) * 1.0001f;
}

struct Lattice {
  Cone c;
  Poly2 q;  // the light's footprint in bins, corners in order around it
  bool ok;
};

// WGSL :4786-4826
static Lattice lattice_cone(V3 apex, const Light& now, const Light& then) {
  Lattice out;
  std::memset(&out, 0, sizeof(out));
  out.ok = false;
  LatticeFrame f = lattice_frame(apex, then);
  if (!f.ok) return out;
  V3 n = normalize(cross(xyz(now.edge_u), xyz(now.edge_v)));
  if (now.corner[3] > 0.5f && dot(n, apex - xyz(now.corner)) <= 0.0f) return out;
  V2 pts[4];
  V2 lo{3.4e38f, 3.4e38f}, hi{-3.4e38f, -3.4e38f};
  for (uint32_t k = 0; k < 4; k++) {
    V3 p = rect_corner(now, k) - apex;
    float z = dot(f.forward, p);
    if (z <= 1e-6f) return out;
    pts[k] = V2{dot(f.right, p) / z, dot(f.up, p) / z};
    lo = V2{fmin_(lo.x, pts[k].x), fmin_(lo.y, pts[k].y)};
    hi = V2{fmax_(hi.x, pts[k].x), fmax_(hi.y, pts[k].y)};
  }
  V2 origin{std::floor(lo.x / f.beta.x) * f.beta.x, std::floor(lo.y / f.beta.y) * f.beta.y};
  V2 end{origin.x + 8.0f * f.beta.x, origin.y + 8.0f * f.beta.y};
  if (hi.x > end.x || hi.y > end.y) return out;
  V3 far_n = n;
  if (dot(far_n, xyz(now.corner) - apex) < 0.0f) far_n = -far_n;
  float dist = dot(far_n, xyz(now.corner) - apex);
  if (dist <= 1e-6f) return out;
  out.c.apex = apex;
  out.c.right = f.right;
  out.c.up = f.up;
  out.c.forward = f.forward;
  out.c.window[0] = origin.x;
  out.c.window[1] = origin.y;
  out.c.window[2] = end.x;
  out.c.window[3] = end.y;
  out.c.far_normal = far_n;
  out.c.far_offset = dot(far_n, apex) + dist * ONE_MINUS_1EM5;
  out.c = bound_cone(out.c, normalize(light_center(now) - apex));
  const uint32_t order[4] = {0, 1, 3, 2};
  out.q.n = 4;
  for (uint32_t i = 0; i < 4; i++)
    out.q.v[i] = V2{(pts[order[i]].x - origin.x) / f.beta.x, (pts[order[i]].y - origin.y) / f.beta.y};
  out.ok = true;
  return out;
}

// WGSL :4830-4848: the footprint's area in each bin (in bins).
static void lattice_areas(const Poly2& q, float m[64]) {
  Acc acc;
  for (uint32_t b = 0; b < 68; b++) acc.a[b] = 0.0f;
  acc.span_lo = 68;
  acc.span_hi = 0;
  const float g = (float)GRID;
  for (uint32_t i = 0; i < q.n; i++) draw_edge(acc, clamp2(q.v[i], g), clamp2(q.v[(i + 1) % q.n], g), 1.0f);
  float running = 0.0f;
  for (uint32_t b = 0; b < BINS; b++) {
    running += acc.a[b];
    m[b] = fmin_(std::fabs(running), 1.0f);
  }
}

// WGSL :4853-4869
static float lattice_share(const Cone& c, V3 normal, const float m[64], const float cov[64]) {
  const float g = (float)GRID;
  float total = 0.0f, open = 0.0f;
  for (uint32_t b = 0; b < BINS; b++) {
    float area = m[b];
    if (area <= 0.0f) continue;
    V3 d = direction(c, ((float)(b % GRID) + 0.5f) / g, ((float)(b / GRID) + 0.5f) / g);
    float l2 = dot(d, d);
    float w = area * fmax_(dot(normal, d), 0.0f) / (l2 * l2);
    total += w;
    open += w * (1.0f - cov[b]);
  }
  if (total <= 0.0f) return 1.0f;
  return open / total;
}

}  // namespace orc

extern "C" {

// The guard cone's walk from `apex` toward `light` (WGSL :4576-4583): 1 if
// any bin is touched; its angular radius to `alpha`.
int oracle_guard_trace(const float* tris, const uint32_t* ranges, uint32_t n_objects, const float* blas,
                       const uint32_t* blas_items, const float* blas_cov, const float* nodes, const uint32_t* items,
                       const float* xforms, const uint32_t* opts, const float* apex, const float* light,
                       int trace, float* alpha) {
  Scene s = make_scene(tris, ranges, n_objects, blas, blas_items, blas_cov, nodes, items, xforms);
  Options o = make_options(opts);
  GuardCone g = guard_cone(xyz(apex), *reinterpret_cast<const Light*>(light));
  *alpha = g.alpha;
  if (!trace) return 0;
  Query qs;
  qs.s = &s;
  qs.o = &o;
  qs.prefilter_bins = o.prefilter;
  qs.cone_cap = NONE;
  float cov[64];
  for (int b = 0; b < 64; b++) cov[b] = 0.0f;
  uint32_t full = 0;
  scene_coverage(g.c, qs, cov, full);
  for (int b = 0; b < 64; b++)
    if (cov[b] > 0.0f) return 1;
  return 0;
}

float oracle_asin(float x) { return asin_c(x); }
float oracle_cos(float x) { return cos_c(x); }
float oracle_tan(float x) { return tan_c(x); }

float oracle_below_horizon(const float* apex, const float* normal, const float* light) {
  return below_horizon(xyz(apex), xyz(normal), *reinterpret_cast<const Light*>(light));
}

// Lattice frame of pose `then` seen from `apex`: out = forward3 right3 up3
// beta2 alpha ok (13 floats); returns kappa for `normal`.
float oracle_lattice_frame(const float* apex, const float* then, const float* normal, float* out) {
  LatticeFrame f = lattice_frame(xyz(apex), *reinterpret_cast<const Light*>(then));
  if (out) {
    const float v[13] = {f.forward.x, f.forward.y, f.forward.z, f.right.x, f.right.y, f.right.z, f.up.x, f.up.y,
                         f.up.z, f.beta.x, f.beta.y, f.alpha, f.ok ? 1.0f : 0.0f};
    std::memcpy(out, v, sizeof(v));
  }
  return lattice_kappa(f, xyz(normal));
}

// WGSL :4875-4899 `lattice_ratio`: the ratio (-1 where the lattice cannot
// hold the footprint or `on` is false) and flags[0] = last_empty,
// flags[1] = last_full. With `areas`, the footprint's 64 bin areas.
float oracle_lattice_ratio(const float* tris, const uint32_t* ranges, uint32_t n_objects, const float* blas,
                           const uint32_t* blas_items, const float* blas_cov, const float* nodes,
                           const uint32_t* items, const float* xforms, const uint32_t* opts, int on,
                           const float* apex, const float* normal, const float* now, const float* then,
                           int* flags, float* areas) {
  Scene s = make_scene(tris, ranges, n_objects, blas, blas_items, blas_cov, nodes, items, xforms);
  Options o = make_options(opts);
  Lattice lc = lattice_cone(xyz(apex), *reinterpret_cast<const Light*>(now), *reinterpret_cast<const Light*>(then));
  bool traced = on && lc.ok;
  float m[64], cov[64];
  lattice_areas(lc.q, m);
  if (areas) std::memcpy(areas, m, sizeof(m));
  uint32_t full = 0;
  for (int b = 0; b < 64; b++) {
    cov[b] = traced && m[b] <= 0.0f ? 1.0f : 0.0f;
    if (cov[b] >= 1.0f) full++;
  }
  if (traced) {
    Query qs;
    qs.s = &s;
    qs.o = &o;
    qs.prefilter_bins = o.prefilter;
    qs.cone_cap = NONE;
    scene_coverage(lc.c, qs, cov, full);
  }
  bool last_empty = true, last_full = true;
  for (int b = 0; b < 64; b++) {
    if (cov[b] > 0.0f && m[b] > 0.0f) last_empty = false;
    if (cov[b] < 1.0f) last_full = false;
  }
  flags[0] = last_empty ? 1 : 0;
  flags[1] = last_full ? 1 : 0;
  if (!traced) return -1.0f;
  return lattice_share(lc.c, xyz(normal), m, cov);
}

}  // extern "C"
