// The cone query's device side, shared by the query kernels and the lighting
// cache's kernels: lights, weights, the packet walk over the object and
// triangle hierarchies, and one whole query.
//
// Cone-visibility queries on gfx950: one lane per (shading point, light)
// query, 64 queries per wave walking the scene's hierarchies as one packet.
//
// Follows the reference's GPU path, crates/continuum-gpu/src/mask_cones.wgsl:
// `hit_irradiance_rect` :2125 -> `cone_irradiance_rect` :1286 ->
// `scene_coverage_if` :1160 (object hierarchy) -> `add_mesh_coverage_if` :719
// (triangle hierarchies) -> `node_cone_test` :519 / `cover_triangle` :594 ->
// `rasterize` :350 -> `draw_edge` :297, then `cone_visible` :1244.
//
// Packets (WGSL :674-711 with PACKETS): the wave keeps one traversal stack.
// Entry i lives in lane i of three VGPRs (node index, and the 64-bit mask of
// lanes that need it); a pop is a v_readlane into SGPRs, so every node and
// triangle the wave visits is fetched once with scalar loads. Each lane acts
// only on the nodes its own walk would visit, in the order it would visit
// them, so a lane's result equals its solo walk. The one packet-wide choice
// is the object hierarchy's nearest-first child order, taken from the first
// lane as the reference's `packet_first` does; it cannot change coverage
// (meshes combine by the maximum), only which nodes a capped walk reaches.
//
// Per-query state: the 64 bins' coverage in VGPRs (every access unrolled to
// constant indices), the 68-cell accumulation buffer in LDS, interleaved by
// lane.
#pragma once
#include "cnq_kernels.h"

namespace cnq {

// ---- lights (WGSL :843-1042)
struct LightCone {
  Cone c;
  bool ok;
  bool wide_clipped;
  bool wide_window;
};

__device__ __forceinline__ bool is_sphere(const LightRect& r) { return r.corner.w > 1.5f; }

__device__ __forceinline__ f3 light_center(const LightRect& r) {
  if (is_sphere(r)) return f3{r.corner.x, r.corner.y, r.corner.z};
  return f3{r.corner.x, r.corner.y, r.corner.z} +
         (f3{r.edge_u.x, r.edge_u.y, r.edge_u.z} + f3{r.edge_v.x, r.edge_v.y, r.edge_v.z}) * 0.5f;
}

__device__ __forceinline__ float spot_falloff(const LightRect& r, f3 w) {
  if (!is_sphere(r) || r.edge_v.w < 0.5f) return 1.0f;
  float a = clampf(dot3(f3{r.edge_v.x, r.edge_v.y, r.edge_v.z}, w) * r.edge_u.y + r.edge_u.z, 0.0f, 1.0f);
  return a * a;
}

__device__ __forceinline__ bool faces_light(const LightRect& r, f3 point, f3 normal) {
  f3 to_light = light_center(r) - point;
  if (is_sphere(r)) return dot3(normal, to_light) > -r.edge_u.x;
  return dot3(normal, to_light) > 0.0f;
}

// WGSL :902-931
__device__ __forceinline__ LightCone light_cone_sphere(f3 apex, const LightRect& r, float wide_extent) {
  LightCone out;
  out.ok = false;
  out.wide_clipped = false;
  out.wide_window = false;
  out.c = Cone{};
  f3 center{r.corner.x, r.corner.y, r.corner.z};
  f3 to_center = center - apex;
  float d = len3(to_center);
  float radius = r.edge_u.x;
  if (d <= radius * 1.0001f) return out;
  f3 forward = to_center / d;
  float s = radius / d;
  float t = s / sqrtf(mx(1.0f - s * s, 1e-12f));
  out.wide_clipped = 2.0f * t > wide_extent;
  out.wide_window = out.wide_clipped;
  t = mn(t, 0.5f * wide_extent);
  f3 helper = fabsf(forward.x) > 0.9f ? f3{0.0f, 1.0f, 0.0f} : f3{1.0f, 0.0f, 0.0f};
  f3 right = norm3(cross3(helper, forward));
  out.c.apex = apex;
  out.c.right = right;
  out.c.up = cross3(forward, right);
  out.c.forward = forward;
  out.c.w0 = -t;
  out.c.w1 = -t;
  out.c.w2 = t;
  out.c.w3 = t;
  out.c.far_normal = forward;
  out.c.far_offset = dot3(forward, apex) + (d - radius * s) * ONE_MINUS_1EM5;
  bound_cone(out.c, forward);
  out.ok = true;
  return out;
}

// WGSL :974-1042
__device__ __forceinline__ LightCone light_cone_rect(f3 apex, const LightRect& r, float wide_extent) {
  if (is_sphere(r)) return light_cone_sphere(apex, r, wide_extent);
  LightCone out;
  out.ok = false;
  out.wide_clipped = false;
  out.wide_window = false;
  out.c = Cone{};
  f3 corner0{r.corner.x, r.corner.y, r.corner.z};
  f3 edge_u{r.edge_u.x, r.edge_u.y, r.edge_u.z};
  f3 edge_v{r.edge_v.x, r.edge_v.y, r.edge_v.z};
  f3 forward = norm3(cross3(edge_u, edge_v));
  if (r.corner.w > 0.5f && dot3(forward, apex - corner0) <= 0.0f) return out;
  if (dot3(forward, corner0 - apex) < 0.0f) forward = -forward;
  float dist = dot3(forward, corner0 - apex);
  if (dist <= 1e-6f) return out;
  f3 right = norm3(edge_u);
  f3 up = cross3(forward, right);
  f3 rel = corner0 - apex;
  f2 lo{3.4e38f, 3.4e38f}, hi{-3.4e38f, -3.4e38f};
#pragma unroll
  for (uint32_t k = 0; k < 4; k++) {
    f3 corner = rel + edge_u * (float)(k & 1u) + edge_v * (float)((k >> 1u) & 1u);
    f2 uv{dot3(right, corner) / dist, dot3(up, corner) / dist};
    lo = f2{mn(lo.x, uv.x), mn(lo.y, uv.y)};
    hi = f2{mx(hi.x, uv.x), mx(hi.y, uv.y)};
  }
  f2 extent = hi - lo;
  f2 foot{mn(mx(0.0f, lo.x), hi.x), mn(mx(0.0f, lo.y), hi.y)};
  f3 mid = norm3(f3{0.5f * (lo.x + hi.x), 0.5f * (lo.y + hi.y), 1.0f});
  float cos_wide = 1.0f;
#pragma unroll
  for (uint32_t k = 0; k < 4; k++) {
    f3 q{(k & 1u) ? hi.x : lo.x, (k & 2u) ? hi.y : lo.y, 1.0f};
    cos_wide = mn(cos_wide, dot3(mid, norm3(q)));
  }
  float size = wide_extent;
  out.wide_window = extent.x > size || extent.y > size;
  out.wide_clipped = out.wide_window && cos_wide < 0.8660254f;
  if (out.wide_clipped) {
    bool wx = extent.x > size, wy = extent.y > size;
    f2 keep{clampf(foot.x - 0.5f * size, lo.x, hi.x - size), clampf(foot.y - 0.5f * size, lo.y, hi.y - size)};
    f2 nlo{wx ? keep.x : lo.x, wy ? keep.y : lo.y};
    f2 nhi{wx ? keep.x + size : hi.x, wy ? keep.y + size : hi.y};
    lo = nlo;
    hi = nhi;
  }
  out.c.apex = apex;
  out.c.right = right;
  out.c.up = up;
  out.c.forward = forward;
  out.c.w0 = lo.x;
  out.c.w1 = lo.y;
  out.c.w2 = hi.x;
  out.c.w3 = hi.y;
  out.c.far_normal = forward;
  out.c.far_offset = dot3(forward, apex) + dist * ONE_MINUS_1EM5;
  f3 axis = norm3(rel + (edge_u + edge_v) * 0.5f);
  if (out.wide_clipped) axis = norm3(direction(out.c, 0.5f, 0.5f));
  bound_cone(out.c, axis);
  out.ok = true;
  return out;
}

// ---- weights (WGSL :1232-1272, :1331-1410, :2148-2190)
__device__ __forceinline__ float bin_weight(const Cone& c, f3 normal, uint32_t b) {
  float col = (float)(b % GRID), row = (float)(b / GRID), g = (float)GRID;
  f3 d0 = direction(c, col / g, row / g);
  f3 d1 = direction(c, (col + 1.0f) / g, row / g);
  f3 d2 = direction(c, (col + 1.0f) / g, (row + 1.0f) / g);
  f3 d3 = direction(c, col / g, (row + 1.0f) / g);
  return fabsf(0.5f * (lambert_edge(normal, d0, d1) + lambert_edge(normal, d1, d2) + lambert_edge(normal, d2, d3) +
                       lambert_edge(normal, d3, d0)));
}

// A quad of directions clipped to the horizon of `normal`, then Lambert's
// edge integrals; its polygon (up to 8 vertices) at P3A.
__device__ __forceinline__ float clipped_quad_weight(Lane& L, f3 normal, f3 d0, f3 d1, f3 d2, f3 d3,
                                                     bool allow_small) {
  uint32_t n = 0;
  f3 d[4] = {d0, d1, d2, d3};
#pragma unroll
  for (uint32_t k = 0; k < 4; k++) {
    const f3 a = d[k], e = d[(k + 1) & 3u];
    const float sa = dot3(normal, a), se = dot3(normal, e);
    if (sa >= 0.0f) L.put3(P3A, n++, a);
    if ((sa >= 0.0f) != (se >= 0.0f)) L.put3(P3A, n++, a + (e - a) * (sa / (sa - se)));
  }
  if (!allow_small && n < 3) return 0.0f;
  float sum = 0.0f;
  if (n > 0) {
    const f3 first = L.p3(P3A, 0);
    f3 a = first;
    for (uint32_t k = 0; k < n; k++) {
      const f3 b = k + 1 < n ? L.p3(P3A, k + 1) : first;
      sum += lambert_edge(normal, a, b);
      a = b;
    }
  }
  return fabsf(0.5f * sum);
}

__device__ __forceinline__ float clipped_bin_weight(Lane& L, const Cone& c, f3 normal, uint32_t b) {
  float col = (float)(b % GRID), row = (float)(b / GRID), g = (float)GRID;
  f3 d0 = direction(c, col / g, row / g);
  f3 d1 = direction(c, (col + 1.0f) / g, row / g);
  f3 d2 = direction(c, (col + 1.0f) / g, (row + 1.0f) / g);
  f3 d3 = direction(c, col / g, (row + 1.0f) / g);
  uint32_t above = (dot3(normal, d0) > 0.0f) + (dot3(normal, d1) > 0.0f) + (dot3(normal, d2) > 0.0f) +
                   (dot3(normal, d3) > 0.0f);
  if (above == 4) return bin_weight(c, normal, b);
  if (above == 0) return 0.0f;
  return clipped_quad_weight(L, normal, d0, d1, d2, d3, true);
}

// The window's weight: one quad when wholly above the horizon, else clipped
// (`cone_visible` :1249-1262, `window_open` :1078, `clipped_window_weight` :1375).
__device__ __forceinline__ float window_weight4(Lane& L, const Cone& c, f3 normal, bool& all_above) {
  f3 d00 = direction(c, 0.0f, 0.0f), d10 = direction(c, 1.0f, 0.0f);
  f3 d11 = direction(c, 1.0f, 1.0f), d01 = direction(c, 0.0f, 1.0f);
  all_above = dot3(normal, d00) > 0.0f && dot3(normal, d10) > 0.0f && dot3(normal, d11) > 0.0f &&
              dot3(normal, d01) > 0.0f;
  if (all_above)
    return fabsf(0.5f * (lambert_edge(normal, d00, d10) + lambert_edge(normal, d10, d11) +
                         lambert_edge(normal, d11, d01) + lambert_edge(normal, d01, d00)));
  return clipped_quad_weight(L, normal, d00, d10, d11, d01, false);
}

// `cone_visible` (WGSL :1244-1272) over the bins set in `touched` (a bin
// with zero coverage adds nothing), coverage read from the accumulator cells.
__device__ __forceinline__ float cone_visible(Lane& L, const Cone& c, f3 normal, uint64_t touched) {
  bool all_above;
  const float total = window_weight4(L, c, normal, all_above);
  float blocked = 0.0f;
  while (touched) {
    const uint32_t b = (uint32_t)__builtin_ctzll(touched);
    touched &= touched - 1;
    const float covered = L.get(b);
    blocked += covered * (all_above ? bin_weight(c, normal, b) : clipped_bin_weight(L, c, normal, b));
  }
  return mx(total - blocked, 0.0f);
}

__device__ __forceinline__ float rect_open(Lane& L, f3 point, f3 normal, const LightRect& r) {
  f3 u{r.edge_u.x, r.edge_u.y, r.edge_u.z}, v{r.edge_v.x, r.edge_v.y, r.edge_v.z};
  f3 c0 = f3{r.corner.x, r.corner.y, r.corner.z} - point;
  float facing = dot3(norm3(cross3(u, v)), c0);
  if ((r.corner.w > 0.5f && facing >= 0.0f) || fabsf(facing) <= 1e-6f) return 0.0f;
  f3 d1 = c0 + u;
  f3 d2 = c0 + u + v;
  f3 d3 = c0 + v;
  if (dot3(normal, c0) >= 0.0f && dot3(normal, d1) >= 0.0f && dot3(normal, d2) >= 0.0f && dot3(normal, d3) >= 0.0f) {
    f3 a0 = norm3(c0), a1 = norm3(d1), a2 = norm3(d2), a3 = norm3(d3);
    float sum = 0.0f;
    sum += lambert_edge_unit(normal, a0, a1);
    sum += lambert_edge_unit(normal, a1, a2);
    sum += lambert_edge_unit(normal, a2, a3);
    sum += lambert_edge_unit(normal, a3, a0);
    return fabsf(0.5f * sum);
  }
  return clipped_quad_weight(L, normal, c0, c0 + u, c0 + u + v, c0 + v, false);
}

// WGSL :937-955
__device__ __forceinline__ float sphere_unshadowed(f3 point, f3 normal, const LightRect& r) {
  f3 to_center = f3{r.corner.x, r.corner.y, r.corner.z} - point;
  float d2 = dot3(to_center, to_center);
  float radius = r.edge_u.x;
  if (d2 <= radius * radius) return 0.0f;
  float sin2 = radius * radius / d2;
  float cos_t = clampf(dot3(normal, to_center) * (1.0f / sqrtf(d2)), -1.0f, 1.0f);
  if (cos_t * cos_t > sin2) return 3.14159265f * sin2 * mx(cos_t, 0.0f);
  float sin_t = sqrtf(1.0f - cos_t * cos_t);
  float x = sqrtf(1.0f / sin2 - 1.0f);
  float y = clampf(-x * cos_t / sin_t, -1.0f, 1.0f);
  float q = sin_t * sqrtf(1.0f - y * y);
  return mx((cos_t * acos_c(y) - x * q) * sin2 + atan_c(q / x), 0.0f);
}

// Share of each bin of a sphere light's window inside the inscribed disk:
// the exact area of the circle of radius 4 (bin units) about the window's
// centre within each bin, rounded to 6 decimals (WGSL :888-897).
__constant__ float DISK[64] = {
    0.000000f, 0.160355f, 0.694958f, 0.957934f, 0.957934f, 0.694958f, 0.160355f, 0.000000f,
    0.160355f, 0.939877f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 0.939877f, 0.160355f,
    0.694958f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 0.694958f,
    0.957934f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 0.957934f,
    0.957934f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 0.957934f,
    0.694958f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 0.694958f,
    0.160355f, 0.939877f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 0.939877f, 0.160355f,
    0.000000f, 0.160355f, 0.694958f, 0.957934f, 0.957934f, 0.694958f, 0.160355f, 0.000000f,
};

// WGSL :960-972 (coverage read from the accumulator cells)
__device__ __forceinline__ float sphere_open_share(Lane& L, const Cone& c, f3 normal, bool wide_clipped) {
  float total = 0.0f, open = 0.0f;
  const float g = (float)GRID;
  for (uint32_t b = 0; b < BINS; b++) {
    float m = wide_clipped ? 1.0f : DISK[b];
    if (m <= 0.0f) continue;
    f3 d = direction(c, ((float)(b % GRID) + 0.5f) / g, ((float)(b / GRID) + 0.5f) / g);
    float l2 = dot3(d, d);
    float w = m * mx(dot3(normal, d), 0.0f) / (l2 * l2);
    total += w;
    open += w * (1.0f - L.get(b));
  }
  if (total <= 0.0f) return 1.0f;
  return open / total;
}

// ---- the walk
__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

__device__ __forceinline__ uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

__device__ __forceinline__ uint32_t uni(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }

// The wave's traversal stack: entry i in lane i of three VGPRs.
struct Stack {
  uint32_t node, lo, hi;
  __device__ __forceinline__ void put(uint32_t sp, uint32_t n, uint64_t m) {
    if (lane_id() == sp) {
      node = n;
      lo = (uint32_t)m;
      hi = (uint32_t)(m >> 32);
    }
  }
  __device__ __forceinline__ uint32_t node_at(uint32_t sp) const {
    return (uint32_t)__builtin_amdgcn_readlane((int)node, (int)sp);
  }
  __device__ __forceinline__ uint64_t mask_at(uint32_t sp) const {
    return (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)lo, (int)sp) |
           ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)hi, (int)sp) << 32);
  }
};

constexpr uint32_t STACK_LIMIT = 64;

// The scene and launch parameters, from the kernel's own (noalias) arguments.
struct Ctx {
  const NodeC* __restrict__ blas;
  const NodeC* __restrict__ nodes;
  const float* __restrict__ tri_leaf;
  const float* __restrict__ tris;
  const uint32_t* __restrict__ ranges;
  const uint32_t* __restrict__ items;
  const uint32_t* __restrict__ blas_items;
  const float* __restrict__ blas_cov;
  const float* __restrict__ xforms;
  uint32_t n_objects, hierarchy, profile_stage, object_cov;
};

// Per-lane walk state.
struct Walk {
  // The 64 coverage values, in the lane's private memory: a walk reads and
  // writes them only when a mesh's accumulator is folded in (`combine`) and
  // at the end, so keeping them out of registers leaves the triangle and
  // clipping code room without spills (the caller provides the array).
  float* cov;
  uint32_t full;
  uint32_t cone_work;
  uint32_t cone_cap;
  float prefilter_bins;
  // Sky walks (SKY) only: combine by the maximum with nearest-first order
  // (params.w bit 1, a diagnostic) instead of as independent occluders.
  bool sky_max;
  bool overflow;   // deeper than the reference's 48 entries
  bool truncated;  // deeper than this stack's 64 (the walk stopped)
  // A walk that only asks whether anything meets the cone (a guard cone)
  // stops once a mesh's coverage lands in a bin: the answer is settled.
  bool stop_on_touch = false;
  uint32_t work[6];  // object nodes tested, blas nodes tested, triangles tested, triangles drawn,
                     // meshes entered, blas nodes splatted
};

// A triangle the cone does not see (object, triangle in `tris`): a cache
// texel's own, when it looks for its light below its horizon (WGSL
// :587-591, :728). Read only by walks instantiated with SKIP.
struct Skip {
  uint32_t obj, tri;
};

// WGSL :623-638 (the cone in an object's space, for transform `xf`)
__device__ __forceinline__ Cone cone_to_object(const Cone& c, const float* __restrict__ xf) {
  f3 r0 = ld3(xf), r1 = ld3(xf + 4), r2 = ld3(xf + 8);
  f3 t{xf[3], xf[7], xf[11]};
  float inv_scale = xf[13];
  auto rot = [&](f3 d) { return r0 * d.x + r1 * d.y + r2 * d.z; };
  Cone o;
  o.apex = rot(c.apex - t) * inv_scale;
  o.right = rot(c.right);
  o.up = rot(c.up);
  o.forward = rot(c.forward);
  o.w0 = c.w0;
  o.w1 = c.w1;
  o.w2 = c.w2;
  o.w3 = c.w3;
  o.far_normal = rot(c.far_normal);
  o.far_offset = (c.far_offset - dot3(c.far_normal, t)) * inv_scale;
  o.baxis = rot(c.baxis);
  o.bcos = c.bcos;
  o.bsin = c.bsin;
  o.horizon = rot(c.horizon);
  return o;
}

__device__ __forceinline__ float cov_sample(const float* __restrict__ cov27, int i, int j, int k) {
  return cov27[9 * (i + 1) + 3 * (j + 1) + (k + 1)];
}

// WGSL :462-486
__device__ __forceinline__ float node_front_area(const float* __restrict__ cov27, f3 v) {
  f3 a{fabsf(v.x), fabsf(v.y), fabsf(v.z)};
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
#pragma unroll
  for (int c = 0; c < 4; c++) {
    int gs = s0 + (c & 1);
    int gt = t0 + (c >> 1);
    if (axis == 0)
      f[c] = cov_sample(cov27, m, gs, gt);
    else if (axis == 1)
      f[c] = cov_sample(cov27, gs, m, gt);
    else
      f[c] = cov_sample(cov27, gs, gt, m);
  }
  return mixf(mixf(f[0], f[1], fs), mixf(f[2], f[3], fs), ft);
}

// Point `k` (per lane) of the 8 in registers.
__device__ __forceinline__ f2 pick8(const f2 (&p)[8], uint32_t k) {
  f2 r = p[0];
#pragma unroll
  for (uint32_t j = 1; j < 8; j++) r = k == j ? p[j] : r;
  return r;
}

// WGSL :490-512: clockwise hull of the 8 points (2D, in registers), written
// at Q2A; returns its vertex count.
__device__ __forceinline__ uint32_t hull8(Lane& L, const f2 (&p)[8]) {
  uint32_t start = 0;
  f2 ps = p[0];
#pragma unroll
  for (uint32_t k = 1; k < 8; k++) {
    if (p[k].x < ps.x || (p[k].x == ps.x && p[k].y < ps.y)) {
      start = k;
      ps = p[k];
    }
  }
  uint32_t n = 0;
  uint32_t cur = start;
  f2 pc = ps;  // p[cur]
  for (;;) {
    L.put2(Q2A, n++, pc);
    uint32_t next = (cur + 1) & 7u;
    f2 pn = pick8(p, next);
#pragma unroll
    for (uint32_t k = 0; k < 8; k++) {
      const f2 e = pn - pc;
      const f2 f = p[k] - pc;
      const float cr = e.x * f.y - e.y * f.x;
      if (cr > 0.0f || (cr == 0.0f && dot2(f, f) > dot2(e, e))) {
        next = k;
        pn = p[k];
      }
    }
    cur = next;
    pc = pn;
    if (cur == start || n >= 8) break;
  }
  return n;
}

// The splat of a node's 8 projected corners with weight `alpha`: their hull
// clipped to the grid, drawn into the accumulator. When every corner lies in
// the grid, so does the hull, and the four clips would return it unchanged
// (each keeps every vertex and adds none), so they are skipped.
__device__ __forceinline__ bool draw_hull(Lane& L, const f2 (&p)[8], float alpha) {
  bool inside = true;
#pragma unroll
  for (uint32_t k = 0; k < 8; k++) inside = inside && in_grid(p[k]);
  uint32_t n = hull8(L, p);
  uint32_t at = Q2A;
  if (!inside) n = clip_to_grid(L, at, n);
  if (n < 3) return false;
  draw_poly(L, at, n, alpha);
  return true;
}

// WGSL :519-574: 0 culled, 1 splatted with nothing drawn, 2 splatted and drawn, 3 descend.
// `before_draw` runs (for this lane) just before a splat draws into the
// accumulator, so whatever the lane still has queued goes in first.
template <bool SKY = false, typename F>
__device__ __forceinline__ uint32_t node_cone_test(const Cone& c, Lane& L, float prefilter_bins, f3 lo, f3 hi,
                                                   const float* __restrict__ cov27, F&& before_draw) {
  const float g = (float)GRID;
  f2 mnv{3.4e38f, 3.4e38f}, mxv{-3.4e38f, -3.4e38f};
  bool far_out = true, near_out = true, out_u0 = true, out_u1 = true, out_v0 = true, out_v1 = true;
  bool below = SKY;
  bool inside = prefilter_bins > 0.0f;
  // The bounding-cone test first (see may_touch_box).
  const int quick = in_bound_quick(c, lo, hi);
  if (quick == 0) return 0;
#pragma unroll
  for (uint32_t k = 0; k < 8; k++) {
    f3 corner = box_corner(lo, hi, k);
    if (SKY && dot3(c.horizon, corner - c.apex) >= 0.0f) below = false;
    bool in_far = c.far_offset - dot3(c.far_normal, corner) >= 0.0f;
    if (in_far)
      far_out = false;
    else
      inside = false;
    f3 p = to_local(c, corner);
    if (p.z >= NEAR)
      near_out = false;
    else
      inside = false;
    if (p.x - c.w0 * p.z >= 0.0f) out_u0 = false;
    if (c.w2 * p.z - p.x >= 0.0f) out_u1 = false;
    if (p.y - c.w1 * p.z >= 0.0f) out_v0 = false;
    if (c.w3 * p.z - p.y >= 0.0f) out_v1 = false;
  }
  if (far_out || near_out || out_u0 || out_u1 || out_v0 || out_v1 || below) return 0;
  if (quick == 2 && !in_bound(c, (lo + hi) * 0.5f, 0.5f * len3(hi - lo))) return 0;
  if (!inside) return 3;
  // Most nodes that get here project far wider than the prefilter's bins.
  // A bound from approximate reciprocals settles those without the exact
  // divisions, and only where it proves the exact extent exceeds the
  // threshold: with a = px * rcp(pz) (within 3e-7 of px / pz relative) and
  // G = GRID * rcp(w2 - w0), every exact projection lies within
  // 4.4e-7 * G * (U + |w0|) of G * (a - w0), U bounding |a|; so an
  // approximate extent above the threshold by the margins below is one the
  // exact arithmetic also finds above it, and the result is the same.
  {
    const float du = c.w2 - c.w0, dv = c.w3 - c.w1;
    if (du > 0.0f && dv > 0.0f) {
      float ulo = 3.4e38f, uhi = -3.4e38f, vlo = 3.4e38f, vhi = -3.4e38f, umag = 0.0f, vmag = 0.0f;
#pragma unroll
      for (uint32_t k = 0; k < 8; k++) {
        const f3 p = to_local(c, box_corner(lo, hi, k));
        const float r = __builtin_amdgcn_rcpf(p.z);
        const float u = p.x * r, v = p.y * r;
        ulo = mn(ulo, u), uhi = mx(uhi, u), vlo = mn(vlo, v), vhi = mx(vhi, v);
        umag = mx(umag, fabsf(u)), vmag = mx(vmag, fabsf(v));
      }
      const float gu = g * __builtin_amdgcn_rcpf(du), gv = g * __builtin_amdgcn_rcpf(dv);
      const float eu = (uhi - ulo) * gu * (1.0f - 1e-6f) - 2e-6f * gu * (umag * 1.000001f + fabsf(c.w0));
      const float ev = (vhi - vlo) * gv * (1.0f - 1e-6f) - 2e-6f * gv * (vmag * 1.000001f + fabsf(c.w1));
      if (eu > prefilter_bins * 1.000001f || ev > prefilter_bins * 1.000001f) return 3;
    }
  }
  // Every corner between the near and far planes: their projections (the
  // same arithmetic the WGSL does in its corner loop, done only for nodes
  // that survive the culls). Most such nodes are too large to splat, so the
  // extent is found first and the points are stored only for a splat.
#pragma unroll
  for (uint32_t k = 0; k < 8; k++) {
    const f3 p = to_local(c, box_corner(lo, hi, k));
    const f2 q{(p.x / p.z - c.w0) / (c.w2 - c.w0) * g, (p.y / p.z - c.w1) / (c.w3 - c.w1) * g};
    mnv = f2{mn(mnv.x, q.x), mn(mnv.y, q.y)};
    mxv = f2{mx(mxv.x, q.x), mx(mxv.y, q.y)};
  }
  if (mx(mxv.x - mnv.x, mxv.y - mnv.y) > prefilter_bins) return 3;
  if (mxv.x <= 0.0f || mxv.y <= 0.0f || mnv.x >= g || mnv.y >= g) return 1;
  f2 pts[8];
#pragma unroll
  for (uint32_t k = 0; k < 8; k++) {
    const f3 p = to_local(c, box_corner(lo, hi, k));
    pts[k] = f2{(p.x / p.z - c.w0) / (c.w2 - c.w0) * g, (p.y / p.z - c.w1) / (c.w3 - c.w1) * g};
  }
  f3 e = hi - lo;
  f3 v = norm3(c.apex - (lo + hi) * 0.5f);
  float silhouette = fabsf(v.x) * e.y * e.z + fabsf(v.y) * e.x * e.z + fabsf(v.z) * e.x * e.y;
  if (silhouette <= 1e-30f) return 1;
  float alpha = node_front_area(cov27, v) / silhouette;
  if (alpha <= 0.0f) return 1;
  before_draw();
  return draw_hull(L, pts, alpha) ? 2 : 1;
}

// How a mesh's triangles cover (WGSL :578-586, set per mesh :730-742): the
// blend weight of an alpha-blended object, or a transmissive object's
// factor (0: opaque) and whether it is a volume (two surfaces).
struct Cover {
  float weight;
  float transmit;
  bool volume;
};

__device__ __forceinline__ Cover cover_of(float clear) {
  Cover k{1.0f - mx(clear, 0.0f), mx(-clear, 0.0f), false};
  if (clear <= -2.0f) {
    k.transmit = -2.0f - clear;
    k.volume = true;
  }
  return k;
}

// x^5 as ((x x)(x x)) x, for Schlick's reflectance.
__device__ __forceinline__ float pow5(float x) {
  const float x2 = x * x;
  return (x2 * x2) * x;
}

// WGSL :594-607: whether the cone draws the triangle (facing it, not wholly
// behind the window, touching the cone).
template <bool SKY = false>
__device__ __forceinline__ bool triangle_drawn(const Cone& c, f3 v0, f3 v1, f3 v2, uint32_t stage,
                                               uint32_t& drawn_count) {
  if (stage == 4) return false;
  f3 n = cross3(v1 - v0, v2 - v0);
  if (dot3(n, c.apex - v0) <= 0.0f) return false;
  if (dot3(n, direction(c, 0.0f, 0.0f)) >= 0.0f && dot3(n, direction(c, 1.0f, 0.0f)) >= 0.0f &&
      dot3(n, direction(c, 0.0f, 1.0f)) >= 0.0f && dot3(n, direction(c, 1.0f, 1.0f)) >= 0.0f)
    return false;
  if (!may_touch<SKY>(c, v0, v1, v2)) return false;
  drawn_count++;
  return true;
}

// WGSL :608-614: the weight a drawn triangle covers with. Thin glass passes
// what Fresnel's reflectance (Schlick's, F0 0.04, at the cone's centre
// direction) leaves, times the factor.
__device__ __forceinline__ float triangle_weight(const Cone& c, f3 v0, f3 v1, f3 v2, const Cover& k) {
  float weight = k.weight;
  if (k.transmit > 0.0f) {
    const f3 n = cross3(v1 - v0, v2 - v0);
    const float cos_t = fabsf(dot3(norm3(n), norm3(direction(c, 0.5f, 0.5f))));
    const float f = 0.04f + 0.96f * pow5(1.0f - cos_t);
    float passes = k.transmit * (1.0f - f);
    if (k.volume) passes *= 1.0f - f;
    weight = 1.0f - passes;
  }
  return weight;
}

// WGSL :594-616
template <bool SKY = false>
__device__ __forceinline__ bool cover_triangle(const Cone& c, Lane& L, f3 v0, f3 v1, f3 v2, const Cover& k,
                                               uint32_t stage, uint32_t& drawn_count) {
  if (!triangle_drawn<SKY>(c, v0, v1, v2, stage, drawn_count)) return false;
  return rasterize(c, L, v0, v1, v2, triangle_weight(c, v0, v1, v2, k));
}

// Adds one mesh's accumulation buffer to the coverage (WGSL :809-828): by
// the maximum, or with SKY (unless `sky_max`) as independent occluders,
// 1 - (1 - before)(1 - covered) (WGSL :820-821, :1141-1142).
template <bool SKY = false>
__device__ __forceinline__ void combine(Walk& w, Lane& L, bool snap) {
  float running = 0.0f;
  bool landed = false;
  const uint32_t end = L.span_hi < BINS ? L.span_hi : BINS;
  for (uint32_t b = L.span_lo; b < end; b++) {
    {
      running += L.get(b);
      float covered = fabsf(running);
      if (snap) {
        if (covered >= ONE_MINUS_1EM5) covered = 1.0f;
      } else {
        covered = mn(covered, 1.0f);
      }
      if (covered > 1e-6f) {
        const float before = w.cov[b];
        if (SKY && !w.sky_max)
          w.cov[b] = 1.0f - (1.0f - before) * (1.0f - mn(covered, 1.0f));
        else
          w.cov[b] = mx(before, mn(covered, 1.0f));
        if (before < 1.0f && w.cov[b] >= 1.0f) w.full++;
        landed = true;
      }
    }
  }
  if (w.stop_on_touch && landed) w.full = BINS;
}

__device__ __forceinline__ void clear_acc(Lane& L) {
#pragma unroll
  for (uint32_t b = 0; b < ACC_CELLS; b++) L.set(b, 0.0f);
  L.span_lo = ACC_CELLS;
  L.span_hi = 0;
}

// A lane's triangles waiting to be drawn into its accumulator (walks with
// QUEUE): up to DRAW_QUEUE leaves (the leaf's first triangle and a bit per
// triangle), first in, first out. A mesh's walk tests each leaf it opens
// with the whole packet and queues what each lane draws; a round draws one
// triangle per lane, its oldest, whenever a lane's queue is full and at the
// mesh's end, so rounds serve lanes that draw in different leaves at once.
// A lane draws its triangles in its walk's order, and a splat waits for the
// lane's queued triangles, so its accumulator sums the same terms in the
// same order. (The leaves must hold at most 32 triangles.)
constexpr uint32_t DRAW_QUEUE = 4;
struct DrawQueue {
  uint32_t base[DRAW_QUEUE], bits[DRAW_QUEUE], n;
  __device__ __forceinline__ void push(uint32_t k0, uint32_t drawn) {
#pragma unroll
    for (uint32_t i = 0; i < DRAW_QUEUE; i++) {
      base[i] = n == i ? k0 : base[i];
      bits[i] = n == i ? drawn : bits[i];
    }
    n++;
  }
  // The oldest queued triangle, taken off the queue (n > 0).
  __device__ __forceinline__ uint32_t pop() {
    const uint32_t k = base[0] + (uint32_t)__builtin_ctz(bits[0]);
    bits[0] &= bits[0] - 1u;
    if (bits[0] == 0) {
#pragma unroll
      for (uint32_t i = 0; i + 1 < DRAW_QUEUE; i++) {
        base[i] = base[i + 1];
        bits[i] = bits[i + 1];
      }
      n--;
    }
    return k;
  }
};

// One round: each lane with a queued triangle draws its oldest.
__device__ __forceinline__ void draw_round(const Ctx& x, const Cone& c, Lane& L, DrawQueue& q, const Cover& weight,
                                           bool& touched) {
  if (q.n) {
    const uint32_t k = q.pop();
    const float* t = x.tri_leaf + 9 * (size_t)k;
    const f3 v0{t[0], t[1], t[2]}, v1{t[3], t[4], t[5]}, v2{t[6], t[7], t[8]};
    if (rasterize(c, L, v0, v1, v2, triangle_weight(c, v0, v1, v2, weight))) touched = true;
  }
}

// WGSL :719-829 for mesh `m` (uniform), for the lanes with `on`.
template <bool COUNT, bool SKIP = false, bool SKY = false, bool QUEUE = false>
__device__ __forceinline__ void add_mesh_coverage(const Ctx& x, const Cone& wc, uint32_t m, bool on, Walk& w,
                                                  Stack& st, uint32_t base, Lane& L, Skip skip = Skip{NONE, NONE}) {
  if (x.profile_stage == 3) return;
  if (!ballot(on)) return;
  if (COUNT && on) w.work[4]++;
  const uint32_t first = x.ranges[4 * m], count = x.ranges[4 * m + 1], root = x.ranges[4 * m + 2],
                 xform = x.ranges[4 * m + 3];
  Cone c = wc;
  Cover weight{1.0f, 0.0f, false};
  if (xform != NONE) {
    c = cone_to_object(wc, x.xforms + 16 * (size_t)xform);
    weight = cover_of(x.xforms[16 * (size_t)xform + 14]);
  }
  const uint32_t cover_skip = (SKIP && m == skip.obj) ? skip.tri : NONE;
  clear_acc(L);
  bool touched = false;
  if (x.hierarchy == 2) {
    DrawQueue q;
    q.n = 0;
#pragma unroll
    for (uint32_t i = 0; i < DRAW_QUEUE; i++) q.base[i] = 0, q.bits[i] = 0;
    if (root != NONE) {
      if (base + 1 > STACK_LIMIT) {
        w.truncated = true;
        return;
      }
      st.put(base, root, ballot(on));
      uint32_t sp = base + 1;
      for (;;) {
        if (QUEUE) {
          // Draw rounds: while a lane's queue is full (so each has room for
          // the one leaf a step adds), and at the walk's end until every
          // queue is empty (one place, so the drawing code is emitted once).
          const bool end = sp <= base;
          while (ballot(end ? q.n != 0 : q.n == DRAW_QUEUE)) draw_round(x, c, L, q, weight, touched);
          if (end) break;
        } else if (sp <= base) {
          break;
        }
        sp--;
        const uint32_t n = st.node_at(sp);
        const uint64_t need = st.mask_at(sp);
        bool mine = ((need >> lane_id()) & 1u) && w.cone_work < w.cone_cap;
        if (mine) w.cone_work++;
        if (!ballot(mine)) continue;
        const NodeC node = x.blas[n];
        bool descend = false;
        if (mine) {
          if (COUNT) w.work[1]++;
          const uint32_t test = node_cone_test<SKY>(
              c, L, w.prefilter_bins, f3{node.lo[0], node.lo[1], node.lo[2]}, f3{node.hi[0], node.hi[1], node.hi[2]},
              x.blas_cov + 27 * (size_t)n, [&]() {
                if (QUEUE)
                  while (q.n) draw_round(x, c, L, q, weight, touched);
              });
          if (test == 2) touched = true;
          if (COUNT && (test == 1 || test == 2)) w.work[5]++;
          descend = test == 3;
        }
        const uint64_t dm = ballot(descend);
        if (!dm) continue;
        if (node.b > 0) {
          // The leaf's triangles are tested one at a time (uniform loads);
          // then each lane draws the ones it passed, in order, a triangle a
          // round, so the wave takes as many rounds as its busiest lane
          // draws triangles rather than one for each triangle any lane
          // draws. A lane's accumulator sums its triangles' edges in the
          // same order either way.
          // (Leaves of more than 32 triangles, which only a caller's own
          // hierarchy can hold, go in runs of 32.)
          const uint32_t k1 = node.a + node.b;
          if (descend) w.cone_work += node.b;
          for (uint32_t k0 = node.a; k0 < k1; k0 += 32u) {
            const uint32_t kn = k1 - k0 < 32u ? k1 : k0 + 32u;
            uint32_t drawn = 0;  // bit k - k0: this lane draws triangle k
            for (uint32_t k = k0; k < kn; k++) {
              const float* t = x.tri_leaf + 9 * (size_t)k;
              const f3 v0{t[0], t[1], t[2]}, v1{t[3], t[4], t[5]}, v2{t[6], t[7], t[8]};
              if (descend && (!SKIP || x.blas_items[k] != cover_skip)) {
                if (COUNT) w.work[2]++;
                if (triangle_drawn<SKY>(c, v0, v1, v2, x.profile_stage, w.work[3])) drawn |= 1u << (k - k0);
              }
            }
            if (QUEUE) {
              if (drawn) q.push(k0, drawn);
              continue;
            }
            while (ballot(drawn != 0)) {
              if (drawn) {
                const uint32_t k = k0 + (uint32_t)__builtin_ctz(drawn);
                drawn &= drawn - 1u;
                const float* t = x.tri_leaf + 9 * (size_t)k;
                const f3 v0{t[0], t[1], t[2]}, v1{t[3], t[4], t[5]}, v2{t[6], t[7], t[8]};
                if (rasterize(c, L, v0, v1, v2, triangle_weight(c, v0, v1, v2, weight))) touched = true;
              }
            }
          }
        } else {
          if (sp + 2 > REF_STACK) w.overflow = w.overflow || descend;
          if (sp + 2 > STACK_LIMIT) {
            if (descend) w.truncated = true;
            continue;
          }
          st.put(sp, node.a, dm);
          st.put(sp + 1, node.a + 1, dm);
          sp += 2;
        }
      }
    }
  } else if (on) {
    for (uint32_t t = first; t < first + count; t++) {
      const float* tr = x.tris + 12 * (size_t)t;
      if (SKIP && t == cover_skip) continue;
      if (COUNT) w.work[2]++;
      if (cover_triangle<SKY>(c, L, ld3(tr), ld3(tr + 4), ld3(tr + 8), weight, x.profile_stage, w.work[3]))
        touched = true;
    }
  }
  if (touched) combine<SKY>(w, L, true);
}

// WGSL :1096-1150 (object-level splat): true when the node was splatted
// (coverage combined if anything was drawn); false to descend.
template <bool SKY = false>
__device__ __forceinline__ bool splat_object_node(const Ctx& x, const Cone& c, Walk& w, uint32_t n, f3 lo, f3 hi,
                                                  Lane& L) {
  if (x.object_cov == NONE || w.prefilter_bins <= 0.0f) return false;
  const float g = (float)GRID;
  f2 mnv{3.4e38f, 3.4e38f}, mxv{-3.4e38f, -3.4e38f};
  for (uint32_t k = 0; k < 8; k++) {
    f3 corner = box_corner(lo, hi, k);
    if (c.far_offset - dot3(c.far_normal, corner) < 0.0f) return false;
    f3 p = to_local(c, corner);
    if (p.z < NEAR) return false;
    f2 q{(p.x / p.z - c.w0) / (c.w2 - c.w0) * g, (p.y / p.z - c.w1) / (c.w3 - c.w1) * g};
    mnv = f2{mn(mnv.x, q.x), mn(mnv.y, q.y)};
    mxv = f2{mx(mxv.x, q.x), mx(mxv.y, q.y)};
  }
  if (mx(mxv.x - mnv.x, mxv.y - mnv.y) > w.prefilter_bins) return false;
  if (mxv.x <= 0.0f || mxv.y <= 0.0f || mnv.x >= g || mnv.y >= g) return true;
  // A splat: the hull's points (as computed above).
  f2 pts[8];
#pragma unroll
  for (uint32_t k = 0; k < 8; k++) {
    const f3 p = to_local(c, box_corner(lo, hi, k));
    pts[k] = f2{(p.x / p.z - c.w0) / (c.w2 - c.w0) * g, (p.y / p.z - c.w1) / (c.w3 - c.w1) * g};
  }
  f3 e = hi - lo;
  f3 v = norm3(c.apex - (lo + hi) * 0.5f);
  float silhouette = fabsf(v.x) * e.y * e.z + fabsf(v.y) * e.x * e.z + fabsf(v.z) * e.x * e.y;
  if (silhouette <= 1e-30f) return true;
  float alpha = 1.0f - exp_c(-node_front_area(x.blas_cov + 27 * (size_t)(x.object_cov + n), v) / silhouette);
  if (alpha <= 0.0f) return true;
  clear_acc(L);
  if (draw_hull(L, pts, alpha)) combine<SKY>(w, L, false);
  return true;
}

// WGSL :1160-1227 for the lanes with `traced`. Sky walks (SKY, unless
// `sky_max`) keep the children's fixed order, as their products would round
// differently in another.
template <bool COUNT, bool SKIP = false, bool SKY = false, bool QUEUE = false>
__device__ __forceinline__ void scene_coverage(const Ctx& x, const Cone& c, bool traced, Walk& w, Lane& L,
                                               Skip skip = Skip{NONE, NONE}) {
  w.cone_work = 0;
  Stack st{0u, 0u, 0u};
  if (x.hierarchy >= 1) {
    st.put(0, 0, ballot(traced));
    uint32_t sp = 1;
    while (sp > 0) {
      sp--;
      const uint32_t n = st.node_at(sp);
      const uint64_t need = st.mask_at(sp);
      bool mine = ((need >> lane_id()) & 1u) && w.full < BINS && w.cone_work < w.cone_cap;
      if (mine) w.cone_work++;
      if (!ballot(mine)) continue;
      const NodeC node = x.nodes[n];
      const f3 lo{node.lo[0], node.lo[1], node.lo[2]}, hi{node.hi[0], node.hi[1], node.hi[2]};
      bool descend = false;
      if (mine) {
        if (COUNT) w.work[0]++;
        if (may_touch_box<SKY>(c, lo, hi)) {
          const bool splatted = (node.b & NODE_NO_SPLAT) == 0 && splat_object_node<SKY>(x, c, w, n, lo, hi, L);
          descend = !splatted;
        }
      }
      const uint64_t dm = ballot(descend);
      if (!dm) continue;
      const uint32_t count = node.b & ~NODE_NO_SPLAT;
      if (count > 0) {
        for (uint32_t k = node.a; k < node.a + count; k++)
          add_mesh_coverage<COUNT, SKIP, SKY, QUEUE>(x, c, x.items[k], descend, w, st, sp, L, skip);
      } else {
        uint32_t nearc = node.a, farc = node.a + 1;
        if (!SKY || uni(w.sky_max ? 1u : 0u)) {
          const NodeC na = x.nodes[nearc];
          const NodeC nb = x.nodes[farc];
          const bool use_bound =
              c.bcos >= -1.0f && (c.baxis.x != 0.0f || c.baxis.y != 0.0f || c.baxis.z != 0.0f);
          // (componentwise selects keep the cone in registers)
          const f3 axis{use_bound ? c.baxis.x : c.forward.x, use_bound ? c.baxis.y : c.forward.y,
                        use_bound ? c.baxis.z : c.forward.z};
          const float da =
              dot3(axis, (f3{na.lo[0], na.lo[1], na.lo[2]} + f3{na.hi[0], na.hi[1], na.hi[2]}) * 0.5f - c.apex);
          const float db =
              dot3(axis, (f3{nb.lo[0], nb.lo[1], nb.lo[2]} + f3{nb.hi[0], nb.hi[1], nb.hi[2]}) * 0.5f - c.apex);
          // packet_first: the first lane's order for the whole packet.
          if (uni(db < da ? 1u : 0u)) {
            nearc = node.a + 1;
            farc = node.a;
          }
        }
        if (sp + 2 > REF_STACK) w.overflow = w.overflow || descend;
        if (sp + 2 > STACK_LIMIT) {
          if (descend) w.truncated = true;
          continue;
        }
        st.put(sp, farc, dm);
        st.put(sp + 1, nearc, dm);
        sp += 2;
      }
    }
  } else {
    for (uint32_t m = 0; m < x.n_objects; m++)
      add_mesh_coverage<COUNT, SKIP, SKY, QUEUE>(x, c, m, traced && w.full < BINS, w, st, 0, L, skip);
  }
}

// One query, `valid` or not (every lane of the wave takes part in the walk):
// WGSL :2125-2128 `hit_irradiance_rect` around `cone_irradiance_rect` :1286.
struct QueryOut {
  float visible;
  uint64_t full, touched;
  bool capped, overflow, truncated;
};

// `active`: the lane traces (a valid query at a point that is not a sky or
// light pixel); `skip`: the triangle this cone does not see (with SKIP).
template <bool COUNT, bool SKIP = false, bool QUEUE = false>
__device__ __forceinline__ QueryOut evaluate_at(const Ctx& x, const QueryParams& p, Lane& L, bool valid, f3 point,
                                                f3 normal, bool active, const LightRect& r,
                                                float* __restrict__ cov_out, uint64_t* __restrict__ work_out,
                                                Skip skip = Skip{NONE, NONE}) {
  Walk w;
  float cov[BINS];
  w.cov = cov;
#pragma unroll
  for (uint32_t b = 0; b < BINS; b++) w.cov[b] = 0.0f;
  w.full = 0;
  w.cone_work = 0;
  w.overflow = false;
  w.truncated = false;
  w.prefilter_bins = p.prefilter;
#pragma unroll
  for (int k = 0; k < 6; k++) w.work[k] = 0;

  const float spot = spot_falloff(r, norm3(point - light_center(r)));
  const bool on = active && spot > 0.0f && faces_light(r, point, normal);
  float value = 0.0f;
  bool capped = false;
  LightCone lc;
  lc.ok = false;
  lc.wide_clipped = false;
  lc.wide_window = false;
  if (p.profile_stage != 1) {
    lc = light_cone_rect(point, r, p.wide_extent);
    const bool traced = on && lc.ok;
    w.cone_cap = lc.wide_window ? p.wide_cone_cap : p.light_cone_cap;
    if (p.profile_stage != 2) scene_coverage<COUNT, SKIP, false, QUEUE>(x, lc.c, traced, w, L, skip);
    capped = traced && w.cone_work >= w.cone_cap;
    if (w.cone_work >= w.cone_cap && lc.wide_window && !lc.wide_clipped) {
#pragma unroll
      for (uint32_t b = 0; b < BINS; b++) w.cov[b] = 1.0f;
      w.full = BINS;
    }
    if (!traced) {
#pragma unroll
      for (uint32_t b = 0; b < BINS; b++) w.cov[b] = 0.0f;
    }
  }
  QueryOut o;
  o.full = 0;
  o.touched = 0;
#pragma unroll
  for (uint32_t b = 0; b < BINS; b++) {
    o.full |= (uint64_t)(w.cov[b] >= 1.0f) << b;
    o.touched |= (uint64_t)(w.cov[b] > 0.0f) << b;
    L.set(b, w.cov[b]);  // the coverage, for the weights' loop (the accumulator is free now)
  }
  if (p.profile_stage != 1 && on && lc.ok) {
    if (is_sphere(r)) {
      value = sphere_unshadowed(point, normal, r) * sphere_open_share(L, lc.c, normal, lc.wide_clipped);
    } else {
      value = cone_visible(L, lc.c, normal, o.touched);
      if (lc.wide_clipped) {
        bool all_above;
        const float kept = window_weight4(L, lc.c, normal, all_above);
        value = (kept > 1e-9f ? value / kept : 1.0f) * rect_open(L, point, normal, r);
      }
    }
  }
  o.visible = spot * value;
  o.capped = capped;
  o.overflow = w.overflow;
  o.truncated = w.truncated;
  if (valid && cov_out) {
#pragma unroll
    for (uint32_t b = 0; b < BINS; b++) cov_out[b] = w.cov[b];
  }
  if (COUNT && work_out) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (int k = 0; k < 6; k++) {
      uint32_t v = valid ? w.work[k] : 0u;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) v += (uint32_t)__shfl_xor((int)v, off, 64);
      if (lane == 0) atomicAdd(reinterpret_cast<unsigned long long*>(&work_out[k]), (unsigned long long)v);
    }
  }
  return o;
}

// One query at receiver `rc` (the reference's ReceiverGpu: point, normal;
// a negative normal.w traces nothing).
template <bool COUNT>
__device__ __forceinline__ QueryOut evaluate_query(const Ctx& x, const QueryParams& p, Lane& L, bool valid,
                                                   const float* __restrict__ rc, const LightRect& r,
                                                   float* __restrict__ cov_out, uint64_t* __restrict__ work_out) {
  return evaluate_at<COUNT>(x, p, L, valid, ld3(rc), ld3(rc + 4), valid && rc[7] >= 0.0f, r, cov_out, work_out);
}

// ---- Sky light: a uniform sky through six cube-face mask cones (WGSL
// :1323-1510), faces on the world axes, no skipped object.

// WGSL :1326: far enough that every scene lies inside a sky cone.
constexpr float SKY_FAR = 1e7f;

// WGSL :1444-1478: cube face `f` (axis f / 2, negative when f is odd) around
// `apex`; `above`: any of its directions lies above the horizon of `normal`.
__device__ __forceinline__ Cone sky_face(uint32_t f, f3 apex, f3 normal, bool horizon_cull, bool& above) {
  const uint32_t axis = f >> 1;
  const float sign = (f & 1u) ? -1.0f : 1.0f;
  const f3 forward{axis == 0 ? sign : 0.0f, axis == 1 ? sign : 0.0f, axis == 2 ? sign : 0.0f};
  const uint32_t ra = axis == 2 ? 0u : axis + 1u;
  const f3 right{ra == 0 ? 1.0f : 0.0f, ra == 1 ? 1.0f : 0.0f, ra == 2 ? 1.0f : 0.0f};
  Cone c = Cone{};
  c.apex = apex;
  c.right = right;
  c.up = cross3(forward, right);
  c.forward = forward;
  c.w0 = -1.0f;
  c.w1 = -1.0f;
  c.w2 = 1.0f;
  c.w3 = 1.0f;
  c.far_normal = forward;
  c.far_offset = dot3(forward, apex) + SKY_FAR;
  c.horizon = horizon_cull ? normal : f3{0.0f, 0.0f, 0.0f};
  bound_cone(c, forward);
  above = false;
#pragma unroll
  for (uint32_t k = 0; k < 4; k++)
    if (dot3(normal, direction(c, (float)(k & 1u), (float)(k >> 1u))) > 0.0f) above = true;
  return c;
}

struct SkyOut {
  float value;    // irradiance per unit sky radiance through the face
  uint64_t full;  // bins with coverage 1
};

// WGSL :1505-1511 (`sky_face_irradiance`): face `f` at `point` for the lanes
// with `active` (every lane of the wave takes part in the walk).
template <bool COUNT>
__device__ __forceinline__ SkyOut evaluate_sky_face(const Ctx& x, const QueryParams& p, const SkyParams& sp,
                                                    Lane& L, bool valid, f3 point, f3 normal, bool active,
                                                    uint32_t f, uint64_t* __restrict__ work_out) {
  Walk w;
  float cov[BINS];
  w.cov = cov;
#pragma unroll
  for (uint32_t b = 0; b < BINS; b++) w.cov[b] = 0.0f;
  w.full = 0;
  w.cone_work = 0;
  w.cone_cap = 0xffffffffu;  // sky walks have no cap (WGSL :1059)
  w.overflow = false;
  w.truncated = false;
  w.sky_max = sp.max_combine != 0;
  w.prefilter_bins = p.prefilter * (sp.prefilter_scale > 0.0f ? sp.prefilter_scale : 1.0f);
#pragma unroll
  for (int k = 0; k < 6; k++) w.work[k] = 0;
  SkyOut o{0.0f, 0ull};
  if (p.profile_stage == 1) return o;
  bool above;
  const Cone c = sky_face(f, point, normal, sp.horizon_cull != 0, above);
  above = above && active;
  if (p.profile_stage != 2) scene_coverage<COUNT, false, true>(x, c, above, w, L);
#pragma unroll
  for (uint32_t b = 0; b < BINS; b++) o.full |= (uint64_t)(w.cov[b] >= 1.0f) << b;
  // WGSL :1492-1501 (`sky_face_light`)
  if (above) {
    float blocked = 0.0f;
    for (uint32_t b = 0; b < BINS; b++) {
      const float covered = w.cov[b];
      if (covered > 0.0f) blocked += covered * clipped_bin_weight(L, c, normal, b);
    }
    bool all_above;
    o.value = mx(window_weight4(L, c, normal, all_above) - blocked, 0.0f);
  }
  if (COUNT && work_out) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (int k = 0; k < 6; k++) {
      uint32_t v = valid ? w.work[k] : 0u;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) v += (uint32_t)__shfl_xor((int)v, off, 64);
      if (lane == 0) atomicAdd(reinterpret_cast<unsigned long long*>(&work_out[k]), (unsigned long long)v);
    }
  }
  return o;
}

}  // namespace cnq
