// Camera rays for the plain per-view launch: per pixel, the nearest
// front-facing hit and the shading point its cones start from.
//
// Follows crates/continuum-gpu/src/mask_cones.wgsl: `camera_dir` :1769,
// `primary_hit` :1732, `nearest_in_mesh` :1647, `push_children` :1617,
// `ray_box_entry` :1605, `hit_triangle` :1575, `ray_triangle` :1539 (camera
// rays cull back faces, `camera_hit` :2222), `seen_light` :2098,
// `rect_distance` :2071, `surface_offset` :2062. Each lane walks alone (ray
// order decides coincident-face ties), its stack in LDS interleaved by lane.
#include "cnq_kernels.h"

namespace cnq {

struct Hit {
  float distance;
  uint32_t mesh, tri;
  f3 normal;
};

constexpr float EDGE_TOLERANCE = 1e-5f;
// The ray stack's depth: 64 entries a lane by default (16 KB of LDS a
// 64-lane block), or 32 (8 KB: more waves in flight); rays deeper than the
// depth stop and are counted as truncated.
constexpr uint32_t RSTACK = 64;

__device__ __forceinline__ float ray_triangle(f3 o, f3 d, f3 v0, f3 v1, f3 v2) {
  f3 e1 = v1 - v0, e2 = v2 - v0;
  f3 p = cross3(d, e2);
  float det = dot3(e1, p);
  if (fabsf(det) < 1e-12f) return -1.0f;
  float inv = 1.0f / det;
  f3 s = o - v0;
  float u = dot3(s, p) * inv;
  if (u < -EDGE_TOLERANCE || u > 1.0f + EDGE_TOLERANCE) return -1.0f;
  f3 q = cross3(s, e1);
  float v = dot3(d, q) * inv;
  if (v < -EDGE_TOLERANCE || u + v > 1.0f + EDGE_TOLERANCE) return -1.0f;
  return dot3(e2, q) * inv;
}

__device__ __forceinline__ Hit hit_triangle(f3 eye, f3 dir, uint32_t m, uint32_t t, f3 v0, f3 v1, f3 v2, Hit best) {
  f3 n = cross3(v1 - v0, v2 - v0);
  if (dot3(n, dir) > 0.0f) return best;
  float d = ray_triangle(eye, dir, v0, v1, v2);
  if (d > 0.0f && m == best.mesh && fabsf(d - best.distance) <= 2e-6f * d) {
    bool facing = dot3(n, dir) < 0.0f;
    if (facing != (dot3(best.normal, dir) < 0.0f)) {
      if (facing) return Hit{d, m, t, norm3(n)};
      return best;
    }
  }
  if (d > 0.0f &&
      (d < best.distance || (d == best.distance && (m < best.mesh || (m == best.mesh && t < best.tri)))))
    return Hit{d, m, t, norm3(cross3(v1 - v0, v2 - v0))};
  return best;
}

__device__ __forceinline__ float ray_box_entry(f3 eye, f3 inv, const NodeC& nd, float limit) {
  float t0x = (nd.lo[0] - eye.x) * inv.x, t0y = (nd.lo[1] - eye.y) * inv.y, t0z = (nd.lo[2] - eye.z) * inv.z;
  float t1x = (nd.hi[0] - eye.x) * inv.x, t1y = (nd.hi[1] - eye.y) * inv.y, t1z = (nd.hi[2] - eye.z) * inv.z;
  float nearv = mx(mx(mn(t0x, t1x), mn(t0y, t1y)), mn(t0z, t1z));
  float farv = mn(mn(mx(t0x, t1x), mx(t0y, t1y)), mx(t0z, t1z));
  if (farv < mx(nearv, 0.0f) || nearv > limit) return -1.0f;
  return mx(nearv, 0.0f);
}

struct RayStack {
  uint32_t* base;  // this lane's entry 0 in LDS; entry i at base[i * 64]
  bool overflow;
  bool truncated;
  bool nearest_first;
  uint32_t depth;  // entries
  __device__ __forceinline__ void set(uint32_t i, uint32_t v) { base[i * 64u] = v; }
  __device__ __forceinline__ uint32_t get(uint32_t i) const { return base[i * 64u]; }
};

__device__ __forceinline__ void push_children(RayStack& st, f3 eye, f3 inv, uint32_t left, const NodeC& l,
                                              const NodeC& r, float limit, uint32_t& sp) {
  if (sp + 2 > REF_STACK) st.overflow = true;
  if (sp + 2 > st.depth) {
    st.truncated = true;
    return;
  }
  if (!st.nearest_first) {
    st.set(sp, left);
    st.set(sp + 1, left + 1);
    sp += 2;
    return;
  }
  float tl = ray_box_entry(eye, inv, l, limit);
  float tr = ray_box_entry(eye, inv, r, limit);
  if (tl >= 0.0f && tr >= 0.0f) {
    if (tl <= tr) {
      st.set(sp, left + 1);
      st.set(sp + 1, left);
    } else {
      st.set(sp, left);
      st.set(sp + 1, left + 1);
    }
    sp += 2;
  } else if (tl >= 0.0f) {
    st.set(sp, left);
    sp += 1;
  } else if (tr >= 0.0f) {
    st.set(sp, left + 1);
    sp += 1;
  }
}

__device__ __forceinline__ f3 recip3(f3 d) { return f3{1.0f / d.x, 1.0f / d.y, 1.0f / d.z}; }

__device__ __forceinline__ Hit nearest_in_mesh(const PrimaryArgs& a, RayStack& st, f3 world_eye, f3 world_dir, uint32_t m, Hit best,
                               uint32_t base) {
  const SceneArgs& s = a.s;
  Hit hit = best;
  const uint32_t first = s.ranges[4 * m], count = s.ranges[4 * m + 1], root = s.ranges[4 * m + 2],
                 xform = s.ranges[4 * m + 3];
  f3 eye = world_eye, dir = world_dir;
  const float* xf = s.xforms + 16 * (size_t)(xform == NONE ? 0 : xform);
  f3 r0 = ld3(xf), r1 = ld3(xf + 4), r2 = ld3(xf + 8);
  if (xform != NONE) {
    f3 t{xf[3], xf[7], xf[11]};
    eye = (r0 * (world_eye - t).x + r1 * (world_eye - t).y + r2 * (world_eye - t).z) * xf[13];
    dir = (r0 * world_dir.x + r1 * world_dir.y + r2 * world_dir.z) * xf[13];
  }
  if (a.hierarchy == 2) {
    if (root == NONE) return hit;
    f3 inv = recip3(dir);
    if (base >= st.depth) {
      st.truncated = true;
      return hit;
    }
    uint32_t sp = base + 1;
    st.set(base, root);
    while (sp > base && !st.truncated) {
      sp--;
      const NodeC node = s.blas[st.get(sp)];
      if (ray_box_entry(eye, inv, node, hit.distance) < 0.0f) continue;
      if (node.b > 0) {
        for (uint32_t k = node.a; k < node.a + node.b; k++) {
          const float* t = s.tri_leaf + 9 * (size_t)k;
          hit = hit_triangle(eye, dir, m, s.blas_items[k], f3{t[0], t[1], t[2]}, f3{t[3], t[4], t[5]},
                             f3{t[6], t[7], t[8]}, hit);
        }
      } else {
        push_children(st, eye, inv, node.a, s.blas[node.a], s.blas[node.a + 1], hit.distance, sp);
      }
    }
  } else {
    for (uint32_t t = first; t < first + count; t++) {
      const float* tr = s.tris + 12 * (size_t)t;
      hit = hit_triangle(eye, dir, m, t, ld3(tr), ld3(tr + 4), ld3(tr + 8), hit);
    }
  }
  if (xform != NONE && hit.mesh == m && best.mesh != m)
    hit.normal = norm3(f3{dot3(r0, hit.normal), dot3(r1, hit.normal), dot3(r2, hit.normal)});
  return hit;
}

__device__ __forceinline__ float rect_distance(const LightRect& l, f3 eye, f3 dir) {
  if (l.corner.w > 1.5f) {
    f3 oc = eye - f3{l.corner.x, l.corner.y, l.corner.z};
    float b = dot3(oc, dir);
    float c = dot3(oc, oc) - l.edge_u.x * l.edge_u.x;
    float disc = b * b - c;
    if (c <= 0.0f || disc < 0.0f) return 3.4e38f;
    float t = -b - sqrtf(disc);
    if (t <= 0.0f) return 3.4e38f;
    return t;
  }
  if (l.corner.w < 0.5f) return 3.4e38f;
  f3 u{l.edge_u.x, l.edge_u.y, l.edge_u.z}, v{l.edge_v.x, l.edge_v.y, l.edge_v.z};
  f3 corner{l.corner.x, l.corner.y, l.corner.z};
  f3 n = cross3(u, v);
  float facing = dot3(dir, n);
  if (facing >= 0.0f) return 3.4e38f;
  float t = dot3(corner - eye, n) / facing;
  if (t <= 0.0f) return 3.4e38f;
  f3 q = eye + dir * t - corner;
  float ss = dot3(q, u) / dot3(u, u);
  float rr = dot3(q, v) / dot3(v, v);
  if (ss < 0.0f || ss > 1.0f || rr < 0.0f || rr > 1.0f) return 3.4e38f;
  return t;
}

// WGSL :2047-2058 `shading_normal` (with `barycentric` :1780-1793 and
// `normal_to_world` :81-85): the corners' normals of triangle `t` of object
// `m` interpolated at world point `p` (the face's where that is degenerate),
// in world space, on the side `toward` points to.
__device__ __forceinline__ f3 shading_normal(const SceneArgs& s, const float* __restrict__ normals, uint32_t m,
                                             uint32_t t, f3 p_world, f3 toward) {
  const uint32_t xform = s.ranges[4 * m + 3];
  const float* xf = s.xforms + 16 * (size_t)(xform == NONE ? 0 : xform);
  const f3 r0 = ld3(xf), r1 = ld3(xf + 4), r2 = ld3(xf + 8);
  f3 p = p_world;
  if (xform != NONE) {
    const f3 d = p_world - f3{xf[3], xf[7], xf[11]};
    p = (r0 * d.x + r1 * d.y + r2 * d.z) * xf[13];
  }
  const float* tr = s.tris + 12 * (size_t)t;
  const f3 v0 = ld3(tr), e1 = ld3(tr + 4) - v0, e2 = ld3(tr + 8) - v0;
  const f3 d = p - v0;
  const float d11 = dot3(e1, e1), d12 = dot3(e1, e2), d22 = dot3(e2, e2);
  const float denom = d11 * d22 - d12 * d12;
  const float bs = (d22 * dot3(d, e1) - d12 * dot3(d, e2)) / denom;
  const float bu = (d11 * dot3(d, e2) - d12 * dot3(d, e1)) / denom;
  const float bt = 1.0f - bs - bu;
  const f3 face = cross3(e1, e2);
  const float* nr = normals + 9 * (size_t)t;
  const f3 sn = ld3(nr) * bt + ld3(nr + 3) * bs + ld3(nr + 6) * bu;
  const f3 pick = dot3(sn, sn) > 1e-20f ? sn : face;
  f3 n = pick, geometric = face;
  if (xform != NONE) {
    n = f3{dot3(r0, pick), dot3(r1, pick), dot3(r2, pick)};
    geometric = f3{dot3(r0, face), dot3(r1, face), dot3(r2, face)};
  }
  n = norm3(n);
  if (dot3(geometric, toward) < 0.0f) n = -n;
  return n;
}

template <uint32_t DEPTH>
__global__ __launch_bounds__(64) void primary_kernel(PrimaryArgs a) {
  __shared__ uint32_t s_stack[DEPTH * 64];
  const uint32_t lane = threadIdx.x & 63u;
  // 8x8 pixel tiles: one wave per tile, so a wave's rays are neighbours.
  const uint32_t tile = blockIdx.x;
  uint32_t vi = 0;
  if (a.n_views > 1) {
    uint32_t lo = 0, hi = a.n_views;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (a.views[mid].ptile_off <= tile)
        lo = mid;
      else
        hi = mid;
    }
    vi = lo;
  }
  const ViewDesc v = a.n_views > 1 ? a.views[vi] : a.v0;
  const uint32_t local = tile - v.ptile_off;
  const uint32_t tiles_x = (v.w + 7) / 8;
  const uint32_t tx = local % tiles_x, ty = local / tiles_x;
  const uint32_t x = tx * 8 + (lane & 7u), y = ty * 8 + (lane >> 3);
  if (x >= v.w || y >= v.h) return;
  const uint64_t px = v.recv_off + (uint64_t)y * v.w + x;
  const float* cam = a.cameras + 20 * (size_t)vi;
  const float* p = cam + 16;
  float aspect = p[0] / p[1];
  float sx = (((float)x + 0.5f + cam[11]) / p[0] * 2.0f - 1.0f) * p[2] * aspect;
  float sy = (1.0f - ((float)y + 0.5f + cam[15]) / p[1] * 2.0f) * p[2];
  const f3 dir = norm3(ld3(cam + 4) + ld3(cam + 8) * sx + ld3(cam + 12) * sy);
  const f3 eye = ld3(cam);
  RayStack st{s_stack + lane, false, false, a.nearest_first != 0, DEPTH};
  Hit hit{3.4e38f, NONE, NONE, f3{0.0f, 0.0f, 0.0f}};
  const SceneArgs& s = a.s;
  if (a.hierarchy >= 1) {
    f3 inv = recip3(dir);
    uint32_t sp = 1;
    st.set(0, 0);
    while (sp > 0 && !st.truncated) {
      sp--;
      const NodeC node = s.nodes[st.get(sp)];
      if (ray_box_entry(eye, inv, node, hit.distance) < 0.0f) continue;
      const uint32_t count = node.b & ~NODE_NO_SPLAT;
      if (count > 0) {
        for (uint32_t k = node.a; k < node.a + count; k++) hit = nearest_in_mesh(a, st, eye, dir, s.items[k], hit, sp);
      } else {
        push_children(st, eye, inv, node.a, s.nodes[node.a], s.nodes[node.a + 1], hit.distance, sp);
      }
    }
  } else {
    for (uint32_t m = 0; m < s.n_objects; m++) hit = nearest_in_mesh(a, st, eye, dir, m, hit, 0);
  }
  if (dot3(hit.normal, dir) > 0.0f) hit.normal = -hit.normal;
  uint32_t seen = 0;
  float nearest = hit.distance;
  const LightRect* ls = reinterpret_cast<const LightRect*>(a.lights) + v.light_off;
  for (uint32_t l = 0; l < v.n_lights; l++) {
    float t = rect_distance(ls[l], eye, dir);
    if (t < nearest) {
      nearest = t;
      seen = l + 1;
    }
  }
  uint32_t* h = a.hits + 4 * px;
  h[0] = hit.mesh;
  h[1] = hit.tri;
  h[2] = __float_as_uint(hit.distance);
  h[3] = seen;
  const bool lit = seen == 0 && hit.mesh != NONE;
  f3 point{0.0f, 0.0f, 0.0f}, normal{0.0f, 0.0f, 0.0f};
  if (lit) {
    f3 surface = eye + dir * hit.distance;
    point = surface + hit.normal * mx(1e-4f, 2e-6f * hit.distance);
    normal = a.normals ? shading_normal(s, a.normals, hit.mesh, hit.tri, surface, -dir) : hit.normal;
  }
  float* r = a.receivers + 8 * px;
  r[0] = point.x, r[1] = point.y, r[2] = point.z, r[3] = 0.0f;
  r[4] = normal.x, r[5] = normal.y, r[6] = normal.z, r[7] = lit ? 0.0f : -1.0f;
  if (a.counters && (st.overflow || st.truncated)) atomicAdd(&a.counters[0], 1u);
}

hipError_t launch_primary(const PrimaryArgs& a, hipStream_t stream, uint32_t depth) {
  if (a.n_tiles == 0) return hipSuccess;
  if (depth == 32)
    hipLaunchKernelGGL(primary_kernel<32>, dim3(a.n_tiles), dim3(64), 0, stream, a);
  else
    hipLaunchKernelGGL(primary_kernel<RSTACK>, dim3(a.n_tiles), dim3(64), 0, stream, a);
  return hipGetLastError();
}

}  // namespace cnq
