// TEST INFRASTRUCTURE ONLY (see oracle_math.h): the reference's CPU model of
// a mask cone, crates/continuum/src/{cone3.rs,coverage.rs,mask.rs}, restated
// so the reference's own known-answer tests and its GPU-vs-CPU check
// (crates/continuum-gltf/tests/gpu.rs:294-336, 1e-4 of the light's weight)
// can be run against this build. glam's conventions: normalize(v) =
// v * (1 / length(v)); dot and cross by components. Rust's f32::atan2 is
// libm's atan2f.
#include <cmath>
#include <cstdint>
#include <algorithm>
#include <cstring>
#include <vector>

#include "oracle_math.h"

namespace c3 {
using orc::V2;
using orc::V3;
using orc::cross;
using orc::dot;

constexpr int GRID = 8;
constexpr int BINS = 64;

static inline V3 gnorm(V3 v) {
  float len = std::sqrt(dot(v, v));
  return v * (1.0f / len);
}

// mask.rs:24-37
static uint64_t mask_range(uint32_t lo, uint32_t hi) {
  if (lo > hi || lo >= 64) return 0;
  if (hi > 63) hi = 63;
  uint32_t width = hi - lo + 1;
  if (width == 64) return ~(uint64_t)0;
  return (((uint64_t)1 << width) - 1) << lo;
}

// mask.rs:42-52 (an interval over 64 bins of a 1D cone)
static uint64_t mask_from_interval(float t0, float t1) {
  float a = t0 <= t1 ? t0 : t1, b = t0 <= t1 ? t1 : t0;
  const float n = 64.0f;
  float lo = std::fmax(std::ceil(a * n - 0.5f), 0.0f);
  float hi = std::fmin(std::floor(b * n - 0.5f), n - 1.0f);
  if (hi < lo) return 0;
  return mask_range((uint32_t)lo, (uint32_t)hi);
}

// coverage.rs:22-128
struct Coverage {
  uint64_t full = 0;
  float partial[BINS] = {0};
  void cover(int bin, float fraction) {
    if (fraction >= 1.0f)
      full |= mask_range(bin, bin);
    else if (fraction > partial[bin])
      partial[bin] = fraction;
  }
  void add_interval(float t0, float t1) {
    const float n = (float)BINS;
    float a = t0 <= t1 ? t0 : t1, b = t0 <= t1 ? t1 : t0;
    float x0 = std::fmin(std::fmax(a * n, 0.0f), n), x1 = std::fmin(std::fmax(b * n, 0.0f), n);
    if (x1 <= x0) return;
    int first = (int)std::floor(x0);
    if (first > BINS - 1) first = BINS - 1;
    int last = (int)std::ceil(x1) - 1;
    if (last < 0) last = 0;
    if (last > BINS - 1) last = BINS - 1;
    if (first == last) {
      cover(first, x1 - x0);
      return;
    }
    cover(first, (float)(first + 1) - x0);
    cover(last, x1 - (float)last);
    if (last > first + 1) full |= mask_range(first + 1, last - 1);
  }
  float visible_weighted(const float* w) const {
    float v = 0.0f;
    for (int b = 0; b < BINS; b++)
      if (!((full >> b) & 1)) v += (1.0f - partial[b]) * w[b];
    return v;
  }
  float visible_fraction() const {
    float v = 0.0f;
    for (int b = 0; b < BINS; b++)
      if (!((full >> b) & 1)) v += 1.0f - partial[b];
    return v / (float)BINS;
  }
  float open_length(float t0, float t1) const {
    const float n = (float)BINS;
    float a = t0 <= t1 ? t0 : t1, b = t0 <= t1 ? t1 : t0;
    float x0 = std::fmin(std::fmax(a * n, 0.0f), n), x1 = std::fmin(std::fmax(b * n, 0.0f), n);
    if (x1 <= x0) return 0.0f;
    int first = (int)std::floor(x0);
    if (first > BINS - 1) first = BINS - 1;
    int last = (int)std::ceil(x1) - 1;
    if (last < 0) last = 0;
    if (last > BINS - 1) last = BINS - 1;
    float open = 0.0f;
    for (int bin = first; bin <= last; bin++)
      if (!((full >> bin) & 1)) open += std::fmax(std::fmin(x1, (float)(bin + 1)) - std::fmax(x0, (float)bin), 0.0f);
    return open / n;
  }
};

// cone3.rs:70-82
static float projected_solid_angle(V3 n, const V3* verts, int count) {
  float sum = 0.0f;
  for (int i = 0; i < count; i++) {
    V3 a = gnorm(verts[i]), b = gnorm(verts[(i + 1) % count]);
    V3 g = cross(a, b);
    float len = std::sqrt(dot(g, g));
    if (len > 1e-12f) sum += std::atan2(len, dot(a, b)) * dot(n, g * (1.0f / len));
  }
  return std::fabs(0.5f * sum);
}

struct Cone {
  V3 apex, right, up, forward;
  float window[4];
  float near_;
  V3 far_normal;
  float far_offset;
  float weights[BINS];
  V3 direction(float s, float t) const {
    return forward + right * (window[0] + s * (window[2] - window[0])) + up * (window[1] + t * (window[3] - window[1]));
  }
  V3 to_local(V3 p) const {
    V3 d = p - apex;
    return V3{dot(right, d), dot(up, d), dot(forward, d)};
  }
  float far_sd(V3 p) const { return far_offset - dot(far_normal, p); }
};

// cone3.rs:197-248
static bool toward_rect_light(V3 apex, V3 corner, V3 eu, V3 ev, V3 rn, Cone& c) {
  V3 forward = gnorm(cross(eu, ev));
  if (dot(forward, corner - apex) < 0.0f) forward = -forward;
  float dist = dot(forward, corner - apex);
  if (dist <= 1e-6f) return false;
  V3 right = gnorm(eu);
  V3 up = cross(forward, right);
  V3 rel = corner - apex;
  V3 corners[4] = {rel, rel + eu, rel + ev, rel + eu + ev};
  float u0 = 3.40282347e38f, v0 = 3.40282347e38f, u1 = -3.40282347e38f, v1 = -3.40282347e38f;
  for (V3 k : corners) {
    float u = dot(right, k) / dist, v = dot(up, k) / dist;
    u0 = std::fmin(u0, u), v0 = std::fmin(v0, v), u1 = std::fmax(u1, u), v1 = std::fmax(v1, v);
  }
  c.apex = apex, c.right = right, c.up = up, c.forward = forward;
  c.window[0] = u0, c.window[1] = v0, c.window[2] = u1, c.window[3] = v1;
  c.near_ = 1e-4f;
  c.far_normal = forward;
  c.far_offset = dot(forward, apex) + dist * (float)(1.0 - 1e-5);
  for (int bin = 0; bin < BINS; bin++) {
    float col = (float)(bin % GRID), row = (float)(bin / GRID), g = (float)GRID;
    V3 quad[4] = {c.direction(col / g, row / g), c.direction((col + 1.0f) / g, row / g),
                  c.direction((col + 1.0f) / g, (row + 1.0f) / g), c.direction(col / g, (row + 1.0f) / g)};
    c.weights[bin] = projected_solid_angle(rn, quad, 4);
  }
  return true;
}

// cone3.rs:266-283
static bool may_touch(const Cone& c, const V3* t) {
  if (c.far_sd(t[0]) < 0.0f && c.far_sd(t[1]) < 0.0f && c.far_sd(t[2]) < 0.0f) return false;
  V3 l[3] = {c.to_local(t[0]), c.to_local(t[1]), c.to_local(t[2])};
  if (l[0].z < c.near_ && l[1].z < c.near_ && l[2].z < c.near_) return false;
  const float* w = c.window;
  auto outside = [&](int which) {
    for (int i = 0; i < 3; i++) {
      V3 p = l[i];
      float f = which == 0 ? p.x - w[0] * p.z : which == 1 ? w[2] * p.z - p.x : which == 2 ? p.y - w[1] * p.z : w[3] * p.z - p.y;
      if (!(f < 0.0f)) return false;
    }
    return true;
  };
  return !(outside(0) || outside(1) || outside(2) || outside(3));
}

template <class P>
struct Poly {
  P v[10];
  int n;
};

static V2 add2(V2 a, V2 b) { return V2{a.x + b.x, a.y + b.y}; }
static V2 sub2(V2 a, V2 b) { return V2{a.x - b.x, a.y - b.y}; }
static V2 mul2(V2 a, float s) { return V2{a.x * s, a.y * s}; }

// cone3.rs:110-129 (Sutherland-Hodgman, crossings from `a`)
template <class P, class D, class Add, class Sub, class Mul>
static Poly<P> clip(const Poly<P>& p, D d, Add add, Sub sub, Mul mul) {
  Poly<P> out;
  out.n = 0;
  for (int i = 0; i < p.n; i++) {
    P a = p.v[i], b = p.v[(i + 1) % p.n];
    float da = d(a), db = d(b);
    if (da >= 0.0f) out.v[out.n++] = a;
    if ((da >= 0.0f) != (db >= 0.0f)) out.v[out.n++] = add(a, mul(sub(b, a), da / (da - db)));
  }
  return out;
}

// cone3.rs:140-187
static void draw_edge(float* buf, V2 p0, V2 p1) {
  if (std::fabs(p0.y - p1.y) <= 1.1920929e-7f) return;
  float dir = 1.0f;
  if (!(p0.y < p1.y)) {
    dir = -1.0f;
    V2 t = p0;
    p0 = p1;
    p1 = t;
  }
  float dxdy = (p1.x - p0.x) / (p1.y - p0.y);
  float x = p0.x;
  int y_end = (int)std::ceil(p1.y);
  if (y_end > GRID) y_end = GRID;
  for (int y = (int)p0.y; y < y_end; y++) {
    int row = y * GRID;
    float dy = std::fmin((float)(y + 1), p1.y) - std::fmax((float)y, p0.y);
    float x_next = std::fmin(std::fmax(x + dxdy * dy, 0.0f), (float)GRID);
    float d = dy * dir;
    float x0 = x < x_next ? x : x_next, x1 = x < x_next ? x_next : x;
    float x0_floor = std::floor(x0);
    int x0i = (int)x0_floor;
    float x1_ceil = std::ceil(x1);
    int x1i = (int)x1_ceil;
    if (x1i <= x0i + 1) {
      float xmf = 0.5f * (x + x_next) - x0_floor;
      buf[row + x0i] += d - d * xmf;
      buf[row + x0i + 1] += d * xmf;
    } else {
      float s = 1.0f / (x1 - x0);
      float x0f = x0 - x0_floor;
      float a0 = 0.5f * s * (1.0f - x0f) * (1.0f - x0f);
      float x1f = x1 - x1_ceil + 1.0f;
      float am = 0.5f * s * x1f * x1f;
      buf[row + x0i] += d * a0;
      if (x1i == x0i + 2) {
        buf[row + x0i + 1] += d * (1.0f - a0 - am);
      } else {
        float a1 = s * (1.5f - x0f);
        buf[row + x0i + 1] += d * (a1 - a0);
        for (int xi = x0i + 2; xi < x1i - 1; xi++) buf[row + xi] += d * s;
        float a2 = a1 + (float)(x1i - x0i - 3) * s;
        buf[row + x1i - 1] += d * (1.0f - a2 - am);
      }
      buf[row + x1i] += d * am;
    }
    x = x_next;
  }
}

// cone3.rs:360-394
static bool rasterize(const Cone& c, float* buf, const V3* tri) {
  Poly<V3> p;
  p.n = 3;
  for (int i = 0; i < 3; i++) p.v[i] = tri[i];
  auto add3 = [](V3 a, V3 b) { return a + b; };
  auto sub3 = [](V3 a, V3 b) { return a - b; };
  auto mul3 = [](V3 a, float s) { return a * s; };
  p = clip(p, [&](V3 q) { return c.far_sd(q); }, add3, sub3, mul3);
  for (int i = 0; i < p.n; i++) p.v[i] = c.to_local(p.v[i]);
  p = clip(p, [&](V3 q) { return q.z - c.near_; }, add3, sub3, mul3);
  if (p.n < 3) return false;
  const float* w = c.window;
  float su = (float)GRID / (w[2] - w[0]), sv = (float)GRID / (w[3] - w[1]);
  Poly<V2> q;
  q.n = p.n;
  for (int i = 0; i < p.n; i++) q.v[i] = V2{(p.v[i].x / p.v[i].z - w[0]) * su, (p.v[i].y / p.v[i].z - w[1]) * sv};
  const float g = (float)GRID;
  q = clip(q, [](V2 a) { return a.x; }, add2, sub2, mul2);
  q = clip(q, [&](V2 a) { return g - a.x; }, add2, sub2, mul2);
  q = clip(q, [](V2 a) { return a.y; }, add2, sub2, mul2);
  q = clip(q, [&](V2 a) { return g - a.y; }, add2, sub2, mul2);
  if (q.n < 3) return false;
  auto at = [&](int i) {
    V2 v = q.v[i % q.n];
    return V2{std::fmin(std::fmax(v.x, 0.0f), g), std::fmin(std::fmax(v.y, 0.0f), g)};
  };
  for (int i = 0; i < q.n; i++) draw_edge(buf, at(i), at(i + 1));
  return true;
}

// cone3.rs:317-339
static void add_mesh(const Cone& c, Coverage& m, const float* tris, int n, bool cull) {
  float buf[BINS + 4] = {0};
  bool touched = false;
  for (int t = 0; t < n; t++) {
    V3 tri[3] = {V3{tris[9 * t], tris[9 * t + 1], tris[9 * t + 2]}, V3{tris[9 * t + 3], tris[9 * t + 4], tris[9 * t + 5]},
                 V3{tris[9 * t + 6], tris[9 * t + 7], tris[9 * t + 8]}};
    V3 normal = cross(tri[1] - tri[0], tri[2] - tri[0]);
    if (dot(normal, c.apex - tri[0]) <= 0.0f) continue;
    if (cull && !may_touch(c, tri)) continue;  // gpu.rs:326 culls with may_touch first
    touched |= rasterize(c, buf, tri);
  }
  if (touched) {
    float running = 0.0f;
    for (int bin = 0; bin < BINS; bin++) {
      running += buf[bin];
      float covered = std::fabs(running);
      if (covered > 1e-6f) m.cover(bin, std::fmin(covered, 1.0f));
    }
  }
}

// cone3.rs:419-437 (Moller-Trumbore, no edge tolerance)
static bool ray_triangle(V3 o, V3 d, const float* t9, float& t_out) {
  V3 v0{t9[0], t9[1], t9[2]}, v1{t9[3], t9[4], t9[5]}, v2{t9[6], t9[7], t9[8]};
  V3 e1 = v1 - v0, e2 = v2 - v0;
  V3 p = cross(d, e2);
  float det = dot(e1, p);
  if (std::fabs(det) < 1e-12f) return false;
  float inv = 1.0f / det;
  V3 s = o - v0;
  float u = dot(s, p) * inv;
  if (!(u >= 0.0f && u <= 1.0f)) return false;
  V3 q = cross(s, e1);
  float v = dot(d, q) * inv;
  if (v < 0.0f || u + v > 1.0f) return false;
  float t = dot(e2, q) * inv;
  t_out = t;
  return t > 0.0f;
}


// cone3.rs:286-307: whether an axis-aligned box may meet the cone's slab and
// window (conservative; the hierarchy walk's box test).
static bool may_touch_box(const Cone& c, V3 lo, V3 hi) {
  V3 corner[8];
  for (int i = 0; i < 8; i++)
    corner[i] = V3{(i & 1) ? hi.x : lo.x, (i & 2) ? hi.y : lo.y, (i & 4) ? hi.z : lo.z};
  bool all_far = true;
  for (int i = 0; i < 8; i++)
    if (!(c.far_sd(corner[i]) < 0.0f)) all_far = false;
  if (all_far) return false;
  V3 l[8];
  bool all_near = true;
  for (int i = 0; i < 8; i++) {
    l[i] = c.to_local(corner[i]);
    if (!(l[i].z < c.near_)) all_near = false;
  }
  if (all_near) return false;
  const float* w = c.window;
  for (int which = 0; which < 4; which++) {
    bool out = true;
    for (int i = 0; i < 8 && out; i++) {
      V3 p = l[i];
      float f = which == 0 ? p.x - w[0] * p.z : which == 1 ? w[2] * p.z - p.x : which == 2 ? p.y - w[1] * p.z
                                                                                        : w[3] * p.z - p.y;
      if (!(f < 0.0f)) out = false;
    }
    if (out) return false;
  }
  return true;
}

// bvh.rs:25-165: the CPU model's hierarchy over every triangle of every
// mesh: median splits on the longest axis of the triangles' centroids
// ((a + b + c) / 3), keyed by the sum of the three vertices' coordinates on
// that axis, at most 4 triangles a leaf. Equal keys fall in index order (the
// reference leaves their order to select_nth_unstable_by).
struct CpuBvh {
  struct Node {
    V3 lo, hi;
    uint32_t start, count;
  };
  std::vector<Node> nodes;
  std::vector<uint32_t> tri;  // triangle (global index) at each leaf position
  std::vector<uint32_t> mesh_of;
  const float* tris9 = nullptr;

  V3 vert(uint32_t t, int k) const {
    const float* p = tris9 + 9 * (size_t)t + 3 * k;
    return V3{p[0], p[1], p[2]};
  }

  void split(uint32_t node, uint32_t start, uint32_t end) {
    V3 lo{3.40282347e38f, 3.40282347e38f, 3.40282347e38f}, hi{-3.40282347e38f, -3.40282347e38f, -3.40282347e38f};
    V3 clo = lo, chi = hi;
    for (uint32_t i = start; i < end; i++) {
      V3 a = vert(tri[i], 0), b = vert(tri[i], 1), cc = vert(tri[i], 2);
      for (V3 p : {a, b, cc}) {
        lo = V3{std::fmin(lo.x, p.x), std::fmin(lo.y, p.y), std::fmin(lo.z, p.z)};
        hi = V3{std::fmax(hi.x, p.x), std::fmax(hi.y, p.y), std::fmax(hi.z, p.z)};
      }
      V3 sum = a + b + cc;
      V3 cen{sum.x / 3.0f, sum.y / 3.0f, sum.z / 3.0f};
      clo = V3{std::fmin(clo.x, cen.x), std::fmin(clo.y, cen.y), std::fmin(clo.z, cen.z)};
      chi = V3{std::fmax(chi.x, cen.x), std::fmax(chi.y, cen.y), std::fmax(chi.z, cen.z)};
    }
    nodes[node].lo = lo;
    nodes[node].hi = hi;
    if (end - start <= 4) {
      nodes[node].start = start;
      nodes[node].count = end - start;
      return;
    }
    V3 e = chi - clo;
    int axis = (e.x >= e.y && e.x >= e.z) ? 0 : (e.y >= e.z ? 1 : 2);
    auto key = [&](uint32_t t) {
      auto co = [&](V3 v) { return axis == 0 ? v.x : axis == 1 ? v.y : v.z; };
      return co(vert(t, 0)) + co(vert(t, 1)) + co(vert(t, 2));
    };
    uint32_t mid = (start + end) / 2;
    std::stable_sort(tri.begin() + start, tri.begin() + end, [&](uint32_t a, uint32_t b) {
      float ka = key(a), kb = key(b);
      return ka < kb;
    });
    uint32_t left = (uint32_t)nodes.size();
    nodes.push_back(Node{});
    nodes.push_back(Node{});
    nodes[node].start = left;
    nodes[node].count = 0;
    split(left, start, mid);
    split(left + 1, mid, end);
  }

  void build(const float* t9, const uint32_t* counts, uint32_t n_meshes) {
    tris9 = t9;
    uint32_t total = 0;
    for (uint32_t m = 0; m < n_meshes; m++) {
      for (uint32_t k = 0; k < counts[m]; k++) mesh_of.push_back(m);
      total += counts[m];
    }
    tri.resize(total);
    for (uint32_t i = 0; i < total; i++) tri[i] = i;
    nodes.assign(1, Node{});
    split(0, 0, total);
  }

  // bvh.rs:116-165 `collect`: every triangle whose node boxes pass and which
  // passes the triangle test, in the walk's order (children pushed left then
  // right, so the right child is taken first); `visited` counts nodes.
  template <class NodeTest, class TriTest>
  std::vector<uint32_t> collect(NodeTest node_test, TriTest tri_test, uint64_t& visited) const {
    std::vector<uint32_t> found;
    std::vector<uint32_t> stack{0};
    visited = 0;
    while (!stack.empty()) {
      uint32_t i = stack.back();
      stack.pop_back();
      visited++;
      const Node& n = nodes[i];
      if (!node_test(n.lo, n.hi)) continue;
      if (n.count > 0) {
        for (uint32_t k = n.start; k < n.start + n.count; k++)
          if (tri_test(tri[k])) found.push_back(tri[k]);
      } else {
        stack.push_back(n.start);
        stack.push_back(n.start + 1);
      }
    }
    return found;
  }
};

// experiments/room3d/src/lib.rs:121-171 `TileFrustum`: the prismatoid
// between a floor tile [x0, y0, x1, y1] at z = 0 and a horizontal,
// axis-aligned light at height h, as six planes n . p <= o.
struct TileFrustum {
  V3 n[6];
  float o[6];
  TileFrustum(const float tile[4], V3 corner, V3 eu, V3 ev) {
    V3 a = corner, b = corner + eu + ev;
    float lx0 = std::fmin(a.x, b.x), ly0 = std::fmin(a.y, b.y), lx1 = std::fmax(a.x, b.x),
          ly1 = std::fmax(a.y, b.y), h = a.z;
    auto side = [&](int k, float t, float l, V3 axis, float sign) {
      V3 nn = (axis - V3{0.0f, 0.0f, 1.0f} * ((l - t) / h)) * sign;
      n[k] = nn;
      o[k] = sign * t;
    };
    side(0, tile[0], lx0, V3{1, 0, 0}, -1.0f);
    side(1, tile[2], lx1, V3{1, 0, 0}, 1.0f);
    side(2, tile[1], ly0, V3{0, 1, 0}, -1.0f);
    side(3, tile[3], ly1, V3{0, 1, 0}, 1.0f);
    n[4] = V3{-0.0f, -0.0f, -1.0f};
    o[4] = 0.0f;
    n[5] = V3{0.0f, 0.0f, 1.0f};
    o[5] = h;
  }
  bool outside_one_plane(const V3* p, int count) const {
    for (int k = 0; k < 6; k++) {
      bool all = true;
      for (int i = 0; i < count && all; i++)
        if (!(dot(n[k], p[i]) > o[k])) all = false;
      if (all) return true;
    }
    return false;
  }
  bool may_touch_box(V3 lo, V3 hi) const {
    V3 corner[8];
    for (int i = 0; i < 8; i++)
      corner[i] = V3{(i & 1) ? hi.x : lo.x, (i & 2) ? hi.y : lo.y, (i & 4) ? hi.z : lo.z};
    return !outside_one_plane(corner, 8);
  }
  bool may_touch(const V3* t) const { return !outside_one_plane(t, 3); }
};

}  // namespace c3

using namespace c3;

extern "C" {

uint64_t oracle_mask_range(uint32_t lo, uint32_t hi) { return mask_range(lo, hi); }
uint64_t oracle_mask_from_interval(float t0, float t1) { return mask_from_interval(t0, t1); }

// CoverageMask over a list of intervals (pairs): returns visible_fraction.
float oracle_coverage_visible(const float* intervals, uint32_t n) {
  Coverage m;
  for (uint32_t i = 0; i < n; i++) m.add_interval(intervals[2 * i], intervals[2 * i + 1]);
  return m.visible_fraction();
}

float oracle_coverage_open_length(const float* intervals, uint32_t n, float t0, float t1) {
  Coverage m;
  for (uint32_t i = 0; i < n; i++) m.add_interval(intervals[2 * i], intervals[2 * i + 1]);
  return m.open_length(t0, t1);
}

// The CPU model's visible irradiance for receivers (point3, normal3 per
// receiver, 6 floats) against meshes (9 floats per triangle, counts per mesh):
// out[i] = visible, out_total[i] = the cone's total weight (0 if no cone).
void oracle_cone3_visible(const float* recv, uint32_t n, const float* light12, const float* tris,
                          const uint32_t* counts, uint32_t n_meshes, int cull, float* out, float* out_total) {
  V3 corner{light12[0], light12[1], light12[2]}, eu{light12[4], light12[5], light12[6]},
      ev{light12[8], light12[9], light12[10]};
  for (uint32_t i = 0; i < n; i++) {
    V3 p{recv[6 * i], recv[6 * i + 1], recv[6 * i + 2]}, nn{recv[6 * i + 3], recv[6 * i + 4], recv[6 * i + 5]};
    Cone c;
    if (!toward_rect_light(p, corner, eu, ev, nn, c)) {
      out[i] = 0.0f;
      out_total[i] = 0.0f;
      continue;
    }
    Coverage m;
    size_t off = 0;
    for (uint32_t k = 0; k < n_meshes; k++) {
      add_mesh(c, m, tris + 9 * off, (int)counts[k], cull != 0);
      off += counts[k];
    }
    out[i] = m.visible_weighted(c.weights);
    float total = 0.0f;
    for (int b = 0; b < BINS; b++) total += c.weights[b];
    out_total[i] = total;
  }
}

// Brute force (cone3.rs:515-537): stratified area sampling of a rectangular
// light (corner + two edges) with `k`x`k` rays, cosine at both ends; a ray is
// blocked by any triangle hit before the light.
double oracle_sampled_visible(const float* p3, const float* light12, const float* tris, uint64_t n_tris,
                              uint32_t k) {
  V3 p{p3[0], p3[1], p3[2]};
  V3 corner{light12[0], light12[1], light12[2]}, eu{light12[4], light12[5], light12[6]},
      ev{light12[8], light12[9], light12[10]};
  V3 ln = cross(eu, ev);
  double area = std::sqrt((double)dot(ln, ln));
  V3 lnu = ln * (1.0f / std::sqrt(dot(ln, ln)));
  double sum = 0.0;
  for (uint32_t i = 0; i < k; i++)
    for (uint32_t j = 0; j < k; j++) {
      float a = ((float)i + 0.5f) / (float)k, b = ((float)j + 0.5f) / (float)k;
      V3 q = corner + eu * a + ev * b;
      V3 d = q - p;
      float r2 = dot(d, d);
      bool blocked = false;
      for (uint64_t t = 0; t < n_tris && !blocked; t++) {
        float th;
        if (ray_triangle(p, d, tris + 9 * t, th) && th < 1.0f) blocked = true;
      }
      if (!blocked) {
        double r = std::sqrt((double)r2);
        double cos_r = std::fabs((double)dot(d, lnu)) / r;
        double cos_p = (double)d.z / r;  // receivers face +z in these tests
        sum += cos_p * cos_r / (double)r2 * (area / ((double)k * k));
      }
    }
  return sum;
}


// bvh.rs:116 `Bvh::candidates` for the cone toward a rect light (corner,
// edge_u, edge_v as light12's first three vec4s) from `apex` with receiver
// normal `normal`: per mesh, the number of candidate triangles
// (out_per_mesh, n_meshes entries); out_tris (nullable, capacity `cap`)
// the candidates grouped by mesh in the walk's order (`group_by_mesh`: a
// stable sort by mesh); returns the number of candidates, -1 when there is
// no cone. *visited: nodes the walk visited.
int64_t oracle_bvh_candidates(const float* tris9, const uint32_t* counts, uint32_t n_meshes, const float* apex3,
                              const float* light12, const float* normal3, uint32_t* out_per_mesh, uint32_t* out_tris,
                              uint64_t cap, uint64_t* visited) {
  CpuBvh bvh;
  bvh.build(tris9, counts, n_meshes);
  V3 apex{apex3[0], apex3[1], apex3[2]}, nn{normal3[0], normal3[1], normal3[2]};
  V3 corner{light12[0], light12[1], light12[2]}, eu{light12[4], light12[5], light12[6]},
      ev{light12[8], light12[9], light12[10]};
  Cone c;
  if (!toward_rect_light(apex, corner, eu, ev, nn, c)) return -1;
  uint64_t v = 0;
  std::vector<uint32_t> found = bvh.collect([&](V3 lo, V3 hi) { return may_touch_box(c, lo, hi); },
                                            [&](uint32_t t) {
                                              V3 tv[3] = {bvh.vert(t, 0), bvh.vert(t, 1), bvh.vert(t, 2)};
                                              return may_touch(c, tv);
                                            },
                                            v);
  std::stable_sort(found.begin(), found.end(),
                   [&](uint32_t a, uint32_t b) { return bvh.mesh_of[a] < bvh.mesh_of[b]; });
  for (uint32_t m = 0; m < n_meshes; m++) out_per_mesh[m] = 0;
  for (uint64_t i = 0; i < found.size(); i++) {
    out_per_mesh[bvh.mesh_of[found[i]]]++;
    if (out_tris && i < cap) out_tris[i] = found[i];
  }
  if (visited) *visited = v;
  return (int64_t)found.size();
}

// cone3.rs:267 `may_touch` of every triangle (9 floats each) for the same cone.
int oracle_cone3_may_touch(const float* tris9, uint64_t n, const float* apex3, const float* light12,
                           const float* normal3, uint8_t* out) {
  V3 apex{apex3[0], apex3[1], apex3[2]}, nn{normal3[0], normal3[1], normal3[2]};
  V3 corner{light12[0], light12[1], light12[2]}, eu{light12[4], light12[5], light12[6]},
      ev{light12[8], light12[9], light12[10]};
  Cone c;
  if (!toward_rect_light(apex, corner, eu, ev, nn, c)) return -1;
  for (uint64_t t = 0; t < n; t++) {
    const float* p = tris9 + 9 * t;
    V3 tv[3] = {V3{p[0], p[1], p[2]}, V3{p[3], p[4], p[5]}, V3{p[6], p[7], p[8]}};
    out[t] = may_touch(c, tv) ? 1 : 0;
  }
  return 0;
}

// TileFrustum's planes (6 x (n.xyz, o)) for a floor tile and a horizontal
// light, and may_touch_box for boxes (6 floats each) / may_touch for
// triangles (9 floats each).
void oracle_tile_frustum(const float* tile4, const float* light12, float* planes24, const float* boxes6,
                         uint64_t n_boxes, uint8_t* box_out, const float* tris9, uint64_t n_tris, uint8_t* tri_out) {
  V3 corner{light12[0], light12[1], light12[2]}, eu{light12[4], light12[5], light12[6]},
      ev{light12[8], light12[9], light12[10]};
  TileFrustum f(tile4, corner, eu, ev);
  for (int k = 0; k < 6 && planes24; k++) {
    planes24[4 * k] = f.n[k].x, planes24[4 * k + 1] = f.n[k].y, planes24[4 * k + 2] = f.n[k].z;
    planes24[4 * k + 3] = f.o[k];
  }
  for (uint64_t i = 0; i < n_boxes; i++) {
    const float* b = boxes6 + 6 * i;
    box_out[i] = f.may_touch_box(V3{b[0], b[1], b[2]}, V3{b[3], b[4], b[5]}) ? 1 : 0;
  }
  for (uint64_t t = 0; t < n_tris; t++) {
    const float* p = tris9 + 9 * t;
    V3 tv[3] = {V3{p[0], p[1], p[2]}, V3{p[3], p[4], p[5]}, V3{p[6], p[7], p[8]}};
    tri_out[t] = f.may_touch(tv) ? 1 : 0;
  }
}

}  // extern "C"
