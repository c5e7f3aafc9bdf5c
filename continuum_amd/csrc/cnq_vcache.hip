// The view cache's page passes on gfx950: pages of 9 x 9 texels on geometry
// charts, keyed (object, chart | mip << 24, page x | page y << 16) and found
// through an open-addressing page table; each texel holds, per light, the
// irradiance (or its shadowed share of the unshadowed light) through one
// mask cone from the texel's place on the surface.
//
// Follows crates/continuum-gpu/src/mask_cones.wgsl: `vc_hash`/`vc_find`
// :3123-3150, `vc_grid` :3181, `uv_barycentric` :3250, `vc_surface(_at)`
// :3274-3326, `page_texel` :3338, `vc_update` :3394, `vc_bake` :3432-3554
// (static lights), `vc_fill` :3693-3836, and the movers' stale passes
// `world_box` :3867, `light_box` :3887, `hull_meets` :3903, `boxes_meet`
// :3919, `may_be_shadowed` :3967, `mark_invalid` :3988, `append_page`
// :3998, `append_jobs` :4018, `fail_entry` :4029, `vc_stale` :4233-4379,
// `vc_stale_texels` :4385-4432, and the moving lights' error ledger :4457-5164
// (guard cones, `ledger_texel`, `ledger_reset`, `light_shift`, the rect
// lights' lattice, `vc_ledger`); host side crates/continuum-gpu/src/vcache.rs
// (planning :3940-4050, eviction :3425-3440, `set_moved` :2426-2492, moving
// lights :1399-1430 and :2255-2380).
//
// Organisation on this GPU: bake jobs are consecutive (page, texel, light)
// triples, so a wave's 64 lanes are 64 neighbouring texels of one page and
// light (at most two): one packet through the cone walk of cnq_walk.h. The
// fill runs one wave per page (a lane per texel); the stale passes one lane
// per slot and per candidate texel, appending with atomics as the WGSL does.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <unordered_map>

#include "cnq_internal.h"
#include "cnq_walk.h"
#include "vcache_host.h"

namespace cnq {

constexpr uint32_t VC_STRIDE = vc::STRIDE;
constexpr uint32_t VC_TEXELS = vc::PAGE_TEXELS;

// mover_stats words (WGSL): 0 pages appended, 1 invalid marks, 2 visible
// pages tested, 3 appends that did not fit, 4 candidates, 5 jobs appended,
// 6 pages made stale for a re-bake, 7 (here) the bake's job counter; the
// moving lights' ledger: 8 pages re-baked, 9 texels re-baked, 10 pages set
// aside, 11 pages checked, 12 pages re-baking mixed texels, 13 pages held
// whole, 14 jobs asked, 21 pages held on their lattice, 22 lattices that
// may have grown, 23 unbounded texels of held pages.
constexpr uint32_t VC_STATS = 24;
// Frames of light poses kept for the ledger (vcache.rs:161 POSE_RING).
constexpr uint32_t VC_POSE_RING = 1024;
constexpr uint32_t LEDGER_WORDS_HOST = 48;  // words of a ledger record (vcache.rs:162)
// The request set's entries (vcache.rs:85 REQUEST_SET) and the most
// requests one shading pass lists.
constexpr uint32_t VC_REQUEST_SET = 1u << 18;
constexpr uint32_t VC_REQUEST_CAP = 1u << 16;
constexpr uint32_t VC_EVICT_PER_FRAME = 2048;  // vcache.rs:41

struct VcArgs {
  const vc::Chart* __restrict__ charts;
  const uint32_t* __restrict__ cells;
  const uint32_t* __restrict__ cell_tris;
  const float* __restrict__ tri_uv;    // 6 per triangle
  const float* __restrict__ normals;   // 9 per triangle
  const float* __restrict__ tris;      // 12 per triangle
  const uint32_t* __restrict__ ranges;
  const float* __restrict__ xforms;
  const float* __restrict__ lights;    // 16 per light
  const float* __restrict__ moved;     // 4 per vec4
  uint32_t* page_table;
  uint32_t* slot_info;                 // 12 per slot
  uint32_t* slot_stale;
  uint32_t* slot_used;
  float* pool;
  uint32_t* bake_list;                 // 4 per entry
  float* bake_bounds;                  // 8 per entry
  uint32_t* bake_jobs;                 // 2 per job
  uint32_t* candidates;                // 2 per candidate
  uint32_t* invalid_list;
  uint32_t* stats;
  uint32_t n_slots, n_lights, channels, n_movers;
  uint32_t plan_pages, plan_jobs, append_pages, append_jobs;
  uint32_t candidate_cap, invalid_cap, table_mask;
  uint32_t frame, visible_frames, light_ratio;
  uint32_t bake_group;                 // most jobs a bake wave takes at a time (64)
  uint32_t bake_resident;              // waves the GPU holds at once
  float inv_base;
  // Moving lights (WGSL :4457-5164): per light its index among the moving
  // ones (NONE: static); ledger records (slot x moving x 48 words); each
  // record's certified bound; the light poses of the last VC_POSE_RING
  // frames (frame % ring x lights x 16 floats).
  const uint32_t* __restrict__ moving_index;
  uint32_t* ledger;
  float* ledger_bound;
  const float* __restrict__ poses;
  uint32_t n_moving;
  uint32_t ledger_limit;   // jobs the ledger may append; top bit: a moving light moved
  uint32_t ledger_record;  // 0: bakes leave the records alone (a check's bake)
  float epsilon;
};

namespace {

// WGSL :3123-3129 `vc_hash` (vcache.rs:904-913)
__device__ __forceinline__ uint32_t key_hash(uint32_t k0, uint32_t k1, uint32_t k2) {
  uint32_t h = (k0 * 0x9E3779B1u) ^ (k1 * 0x85EBCA77u) ^ (k2 * 0xC2B2AE3Du);
  h ^= h >> 15;
  h *= 0x2C1B3C6Du;
  h ^= h >> 12;
  return h;
}

struct Box3 {
  f3 lo, hi;
};

__device__ __forceinline__ float surface_offset(float distance) { return mx(1e-4f, 2e-6f * distance); }

__device__ __forceinline__ f3 row3(const float* xf, int r) { return f3{xf[4 * r], xf[4 * r + 1], xf[4 * r + 2]}; }

__device__ __forceinline__ f3 rotate_to_world(const float* xf, f3 d) {
  return f3{dot3(row3(xf, 0), d), dot3(row3(xf, 1), d), dot3(row3(xf, 2), d)};
}

__device__ __forceinline__ f3 translation(const float* xf) { return f3{xf[3], xf[7], xf[11]}; }

__device__ __forceinline__ f3 vabs3(f3 a) { return f3{fabsf(a.x), fabsf(a.y), fabsf(a.z)}; }
__device__ __forceinline__ f3 vmax3(f3 a, f3 b) { return f3{mx(a.x, b.x), mx(a.y, b.y), mx(a.z, b.z)}; }
__device__ __forceinline__ float max_el(f3 a) { return mx(a.x, mx(a.y, a.z)); }

__device__ __forceinline__ float object_scale(const VcArgs& a, uint32_t m) {
  const uint32_t x = a.ranges[4 * m + 3];
  return x == NONE ? 1.0f : a.xforms[16 * (size_t)x + 12];
}

// WGSL :3181-3184
__device__ __forceinline__ void vc_grid(const VcArgs& a, uint32_t c, uint32_t m, uint32_t mip, uint32_t n[2]) {
  const vc::Chart ch = a.charts[c];
  const float s = object_scale(a, m);
  const float inv = ldexpf(a.inv_base, -(int)mip);
  const float nx = ceilf((ch.size[0] * s) * inv), ny = ceilf((ch.size[1] * s) * inv);
  const uint32_t ux = nx >= 4294967296.0f ? 0xffffffffu : (uint32_t)nx;
  const uint32_t uy = ny >= 4294967296.0f ? 0xffffffffu : (uint32_t)ny;
  n[0] = ux > 1u ? ux : 1u;
  n[1] = uy > 1u ? uy : 1u;
}

// WGSL :3250-3260
__device__ __forceinline__ void uv_barycentric(const VcArgs& a, uint32_t t, f2 uv, float b[3]) {
  const float* q = a.tri_uv + 6 * (size_t)t;
  const f2 o{q[0], q[1]};
  const f2 e1 = f2{q[2], q[3]} - o, e2 = f2{q[4], q[5]} - o, d = uv - o;
  const float det = e1.x * e2.y - e1.y * e2.x;
  if (fabsf(det) < 1e-30f) {
    b[0] = b[1] = b[2] = -1.0f;
    return;
  }
  const float s = (d.x * e2.y - d.y * e2.x) / det;
  const float u = (e1.x * d.y - e1.y * d.x) / det;
  b[0] = 1.0f - s - u;
  b[1] = s;
  b[2] = u;
}

struct TexelSurface {
  f3 apex, normal;
  bool ok;
  uint32_t tri;
};

// WGSL :3283-3326 (objects' lighting levels are always present here)
__device__ __forceinline__ TexelSurface vc_surface_at(const VcArgs& a, uint32_t m, uint32_t c, f2 uv) {
  const vc::Chart ch = a.charts[c];
  const uint32_t ux = (uint32_t)(uv.x * (float)ch.grid[0]), uy = (uint32_t)(uv.y * (float)ch.grid[1]);
  const uint32_t cx = ux < ch.grid[0] - 1u ? ux : ch.grid[0] - 1u;
  const uint32_t cy = uy < ch.grid[1] - 1u ? uy : ch.grid[1] - 1u;
  const uint32_t cell = ch.first_cell + cy * ch.grid[0] + cx;
  const uint32_t first = a.cells[2 * cell], count = a.cells[2 * cell + 1];
  uint32_t best_t = NONE;
  float best_b[3] = {0.0f, 0.0f, 0.0f};
  float best_score = -3.4e38f;
  for (uint32_t q = first; q < first + count; q++) {
    const uint32_t t = a.cell_tris[q];
    float b[3];
    uv_barycentric(a, t, uv, b);
    const float score = mn(b[0], mn(b[1], b[2]));
    if (score > best_score) {
      best_score = score;
      best_t = t;
      best_b[0] = b[0], best_b[1] = b[1], best_b[2] = b[2];
    }
  }
  float b0[3];
  for (int k = 0; k < 3; k++) b0[k] = clampf(best_b[k], 0.0f, 1.0f);
  const float sum = mx(b0[0] + b0[1] + b0[2], 1e-20f);
  const float bx = b0[0] / sum, by = b0[1] / sum, bz = b0[2] / sum;
  const uint32_t t = best_t != NONE ? best_t : 0u;
  const float* tr = a.tris + 12 * (size_t)t;
  const f3 v0 = ld3(tr), v1 = ld3(tr + 4), v2 = ld3(tr + 8);
  const float* nr = a.normals + 9 * (size_t)t;
  const f3 p = v0 * bx + v1 * by + v2 * bz;
  const f3 interpolated = ld3(nr) * bx + ld3(nr + 3) * by + ld3(nr + 6) * bz;
  const f3 face = cross3(v1 - v0, v2 - v0);
  const f3 nrm = dot3(interpolated, interpolated) > 1e-20f ? interpolated : face;
  f3 pw = p;
  f3 nw = norm3(nrm);
  f3 geometric = face;
  const uint32_t x = a.ranges[4 * m + 3];
  if (x != NONE) {
    const float* xf = a.xforms + 16 * (size_t)x;
    pw = rotate_to_world(xf, p) * xf[12] + translation(xf);
    nw = norm3(rotate_to_world(xf, nrm));
    geometric = rotate_to_world(xf, face);
  }
  TexelSurface s;
  s.apex = pw + norm3(geometric) * surface_offset(len3(pw));
  s.normal = nw;
  s.ok = best_t != NONE;
  s.tri = t;
  return s;
}

struct PageTexel {
  uint32_t e[4];
  uint32_t m, c, n[2], texel[2];
  bool inside;
};

// WGSL :3338-3350
__device__ __forceinline__ PageTexel page_texel(const VcArgs& a, uint32_t page_index, uint32_t k) {
  PageTexel t;
  const uint32_t* e = a.bake_list + 4 * (size_t)page_index;
  t.e[0] = e[0], t.e[1] = e[1], t.e[2] = e[2], t.e[3] = e[3];
  t.m = t.e[0];
  t.c = t.e[1] & vc::KEY_CHART_MASK;
  const uint32_t mip = t.e[1] >> vc::KEY_MIP_SHIFT;
  vc_grid(a, t.c, t.m, mip, t.n);
  t.texel[0] = (t.e[2] & 0xffffu) * vc::PAGE + k % VC_STRIDE;
  t.texel[1] = (t.e[2] >> 16) * vc::PAGE + k / VC_STRIDE;
  t.inside = t.texel[0] < t.n[0] && t.texel[1] < t.n[1];
  return t;
}

// ---- the movers' tests

// WGSL :3867-3882
__device__ __forceinline__ Box3 world_box(const VcArgs& a, uint32_t m, f3 lo, f3 hi) {
  Box3 b{lo, hi};
  const uint32_t x = a.ranges[4 * m + 3];
  if (x != NONE) {
    const float* xf = a.xforms + 16 * (size_t)x;
    const float s = xf[12];
    const f3 c = (lo + hi) * 0.5f;
    const f3 e = (hi - lo) * 0.5f;
    const f3 wc = rotate_to_world(xf, c) * s + translation(xf);
    const f3 we = f3{dot3(vabs3(row3(xf, 0)), e), dot3(vabs3(row3(xf, 1)), e), dot3(vabs3(row3(xf, 2)), e)} * s;
    b = Box3{wc - we, wc + we};
  }
  const f3 far = vmax3(vmax3(vabs3(b.lo), vabs3(b.hi)), f3{1.0f, 1.0f, 1.0f});
  const float pad = 2.0f * surface_offset(len3(far)) + 1e-5f * max_el(far);
  return Box3{b.lo - f3{pad, pad, pad}, b.hi + f3{pad, pad, pad}};
}

__device__ __forceinline__ f3 light_center_of(const float* r) {
  if (r[3] > 1.5f) return ld3(r);
  return ld3(r) + (ld3(r + 4) + ld3(r + 8)) * 0.5f;
}

// WGSL :3887-3898
__device__ __forceinline__ Box3 light_box(const float* r) {
  const f3 c = light_center_of(r);
  float radius;
  if (r[3] > 1.5f) radius = 1.5f * r[4];
  else radius = 0.5f * (len3(ld3(r + 4)) + len3(ld3(r + 8)));
  const f3 far = vmax3(vabs3(c), f3{1.0f, 1.0f, 1.0f});
  radius = radius * 1.001f + 1e-5f * max_el(far) + 1e-4f;
  return Box3{c - f3{radius, radius, radius}, c + f3{radius, radius, radius}};
}

__device__ __forceinline__ float comp(f3 v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }

// WGSL :3903-3917
__device__ __forceinline__ bool hull_meets(const Box3& p, const Box3& l, const Box3& m) {
  float s0 = 0.0f, s1 = 1.0f;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const float b0 = comp(l.lo, a) - comp(p.lo, a);
    const float c0 = comp(m.hi, a) - comp(p.lo, a);
    if (b0 > 0.0f) s1 = mn(s1, c0 / b0);
    else if (b0 < 0.0f) s0 = mx(s0, c0 / b0);
    else if (c0 < 0.0f) return false;
    const float b1 = comp(l.hi, a) - comp(p.hi, a);
    const float c1 = comp(m.lo, a) - comp(p.hi, a);
    if (b1 > 0.0f) s0 = mx(s0, c1 / b1);
    else if (b1 < 0.0f) s1 = mn(s1, c1 / b1);
    else if (c1 > 0.0f) return false;
  }
  return s0 <= s1 + 1e-6f;
}

__device__ __forceinline__ bool boxes_meet(const Box3& a, const Box3& b) {
  return a.lo.x <= b.hi.x && a.lo.y <= b.hi.y && a.lo.z <= b.hi.z && b.lo.x <= a.hi.x && b.lo.y <= a.hi.y &&
         b.lo.z <= a.hi.z;
}

__device__ __forceinline__ Box3 moved_box(const VcArgs& a, uint32_t i) {
  return Box3{ld3(a.moved + 4 * (size_t)i), ld3(a.moved + 4 * (size_t)(i + 1))};
}

__device__ __forceinline__ uint32_t moved_w(const VcArgs& a, uint32_t i) {
  return __float_as_uint(a.moved[4 * (size_t)i + 3]);
}

// WGSL :3967-3982
__device__ __forceinline__ bool may_be_shadowed(const VcArgs& a, const Box3& p, uint32_t l, uint32_t s) {
  const uint32_t lights_n = a.n_lights;
  if (!boxes_meet(p, moved_box(a, 2 * l))) return false;
  const Box3 lb = light_box(a.lights + 16 * (size_t)l);
  const uint32_t first = 2 * lights_n + 2, stride = 4 + 4 * lights_n;
  for (uint32_t k = 0; k < a.n_movers; k++) {
    const uint32_t e = first + k * stride;
    const uint32_t slice = moved_w(a, e + 3);
    if (slice != 0u && s % (slice >> 16) != (slice & 0xffffu)) continue;
    const uint32_t r = e + 4 + 4 * l;
    if (boxes_meet(p, moved_box(a, r)) && hull_meets(p, lb, moved_box(a, e))) return true;
    if (boxes_meet(p, moved_box(a, r + 2)) && hull_meets(p, lb, moved_box(a, e + 2))) return true;
  }
  return false;
}

// WGSL :3988-3996
__device__ __forceinline__ void mark_invalid(const VcArgs& a, uint32_t s) {
  const uint32_t old = atomicExch(&a.slot_stale[s], 1u);
  if (old == 0u || old == 2u || old == 3u) {
    const uint32_t q = atomicAdd(&a.stats[1], 1u);
    if (q < a.invalid_cap) a.invalid_list[q] = s;
  }
}

// WGSL :3998-4012
__device__ __forceinline__ uint32_t append_page(const VcArgs& a, uint32_t s, const uint32_t key[3], f3 lo, f3 hi) {
  const uint32_t p = atomicAdd(&a.stats[0], 1u);
  if (p >= a.append_pages) {
    atomicAdd(&a.stats[3], 1u);
    return NONE;
  }
  const uint32_t entry = a.plan_pages + p;
  uint32_t* e = a.bake_list + 4 * (size_t)entry;
  e[0] = key[0], e[1] = key[1], e[2] = key[2], e[3] = s;
  float* b = a.bake_bounds + 8 * (size_t)entry;
  b[0] = lo.x, b[1] = lo.y, b[2] = lo.z, b[3] = 0.0f;
  b[4] = hi.x, b[5] = hi.y, b[6] = hi.z, b[7] = 0.0f;
  return entry;
}

// WGSL :4018-4027: one add; past the capacity, what of the range lies
// within it is written as empty jobs.
__device__ __forceinline__ uint32_t append_jobs(const VcArgs& a, uint32_t count) {
  const uint32_t old = atomicAdd(&a.stats[5], count);
  if (old + count <= a.append_jobs) return a.plan_jobs + old;
  const uint32_t end = old + count < a.append_jobs ? old + count : a.append_jobs;
  for (uint32_t k = old; k < end; k++) {
    a.bake_jobs[2 * (size_t)(a.plan_jobs + k)] = NONE;
    a.bake_jobs[2 * (size_t)(a.plan_jobs + k) + 1] = 0u;
  }
  return NONE;
}

// WGSL :4029-4035
__device__ __forceinline__ void fail_entry(const VcArgs& a, uint32_t entry, uint32_t s) {
  a.bake_bounds[8 * (size_t)entry + 3] = 1.0f;
  atomicExch(&a.slot_stale[s], 1u);
  const uint32_t q = atomicAdd(&a.stats[1], 1u);
  if (q < a.invalid_cap) a.invalid_list[q] = s;
}

// ---- moving lights through the error ledger (WGSL :4457-5164)

constexpr uint32_t LEDGER_WORDS = 48;      // WGSL :4481, vcache.rs:162
constexpr uint32_t MAX_OFFSET = 1023;      // WGSL :4483
constexpr float GUARD = 3.0f;              // WGSL :4482
constexpr float LATTICE_MARGIN = 0.03f;    // WGSL :4704
constexpr float LATTICE_KAPPA_MAX = 8.0f;  // WGSL :4708

__device__ __forceinline__ uint32_t ledger_base(const VcArgs& a, uint32_t slot, uint32_t j) {
  return (slot * a.n_moving + j) * LEDGER_WORDS;
}

__device__ __forceinline__ const float* pose_at(const VcArgs& a, uint32_t frame, uint32_t l) {
  return a.poses + ((size_t)(frame % VC_POSE_RING) * a.n_lights + l) * 16;
}

__device__ __forceinline__ f3 rect_corner_of(const float* r, uint32_t k) {
  return ld3(r) + ld3(r + 4) * (float)(k & 1u) + ld3(r + 8) * (float)((k >> 1u) & 1u);
}

__device__ __forceinline__ f3 vmin3(f3 a, f3 b) { return f3{mn(a.x, b.x), mn(a.y, b.y), mn(a.z, b.z)}; }

// WGSL :4666-4687: the largest angle a point of light pose `a` moves to
// reach pose `b`, seen from anywhere in box `page` (4: unbounded).
__device__ __forceinline__ float light_shift(const float* la_, const float* lb_, const Box3& page) {
  float moved = 0.0f;
  if (la_[3] > 1.5f) {
    moved = len3(ld3(lb_) - ld3(la_)) + fabsf(lb_[4] - la_[4]);
  } else {
#pragma unroll
    for (uint32_t k = 0; k < 4; k++) moved = mx(moved, len3(rect_corner_of(lb_, k) - rect_corner_of(la_, k)));
  }
  if (moved <= 0.0f) return 0.0f;
  const Box3 la = light_box(la_), lb = light_box(lb_);
  const Box3 span{vmin3(la.lo, lb.lo), vmax3(la.hi, lb.hi)};
  const f3 gap = vmax3(f3{0.0f, 0.0f, 0.0f}, vmax3(page.lo - span.hi, span.lo - page.hi));
  const float d = len3(gap);
  if (moved >= 2.0f * d) return 4.0f;
  return 2.0f * asin_c(moved / (2.0f * d)) * 1.0001f;
}

struct LatticeReach {
  float near, far, cos_min, half_diagonal;
};

// WGSL :4907-4942
__device__ __forceinline__ LatticeReach lattice_reach(const float* a, const float* b, const Box3& page) {
  LatticeReach out;
  const Box3 la = light_box(a), lb = light_box(b);
  const Box3 span{vmin3(la.lo, lb.lo), vmax3(la.hi, lb.hi)};
  const f3 gap = vmax3(f3{0.0f, 0.0f, 0.0f}, vmax3(page.lo - span.hi, span.lo - page.hi));
  out.near = len3(gap);
  const f3 au = ld3(a + 4), av = ld3(a + 8), bu = ld3(b + 4), bv = ld3(b + 8);
  out.half_diagonal = 0.5f * mx(len3(au + av), len3(bu + bv));
  out.half_diagonal = mx(out.half_diagonal, 0.5f * mx(len3(au - av), len3(bu - bv)));
  const f3 pc = (page.lo + page.hi) * 0.5f;
  const float ph = 0.5f * len3(page.hi - page.lo);
  const f3 ca = light_center_of(a) - pc, cb = light_center_of(b) - pc;
  out.far = mx(len3(ca), len3(cb)) + ph + out.half_diagonal;
  const f3 na = norm3(cross3(au, av)), nb = norm3(cross3(bu, bv));
  const float cos_c = mn(fabsf(dot3(na, norm3(ca))), fabsf(dot3(nb, norm3(cb))));
  out.cos_min = out.near > 0.0f ? cos_c - asin_c(mn((ph + out.half_diagonal) / out.near, 1.0f)) : -1.0f;
  return out;
}

// WGSL :4949-4971
__device__ __forceinline__ float lattice_bound(const float* a, const float* b, const Box3& page, float kappa) {
  const float delta = light_shift(a, b, page);
  if (delta <= 0.0f) return 0.0f;
  if (delta >= 1.0f || kappa >= 1.0e30f) return 1.0f;
  const LatticeReach r = lattice_reach(a, b, page);
  if (r.cos_min <= 0.0f || r.near <= 2.0f * r.half_diagonal) return 1.0f;
  const f3 au = ld3(a + 4), av = ld3(a + 8), bu = ld3(b + 4), bv = ld3(b + 8);
  const float area = mn(len3(cross3(au, av)), len3(cross3(bu, bv)));
  const float omega = area * r.cos_min / (r.far * r.far);
  const float lu = mx(len3(au), len3(bu)), lv = mx(len3(av), len3(bv));
  const float perimeter =
      4.0f * (asin_c(mn(lu / (2.0f * r.near), 1.0f)) + asin_c(mn(lv / (2.0f * r.near), 1.0f)));
  const float rho = mn(GUARD * asin_c(mn(r.half_diagonal / r.near, 1.0f)) * 1.0001f, 1.4f);
  const float t = tan_c(rho);
  const float sec2 = 1.0f + t * t;
  const float moved = sec2 * delta;
  const float sw = sec2 * perimeter * moved + 3.14159265f * moved * moved;
  return mn(1.5f * kappa * sw / omega * 1.0001f, 1.0f);
}

// WGSL :4981-5000
__device__ __forceinline__ bool lattice_grew(const float* a, const float* b, const Box3& page) {
  const LatticeReach r = lattice_reach(a, b, page);
  if (r.cos_min <= 0.0f || r.near <= 2.0f * r.half_diagonal) return true;
  const f3 ca = light_center_of(a), cb = light_center_of(b);
  const float t = len3(cb - ca);
  float rest = 0.0f;
#pragma unroll
  for (uint32_t k = 0; k < 4; k++) rest = mx(rest, len3((rect_corner_of(b, k) - cb) - (rect_corner_of(a, k) - ca)));
  const float diameter = 2.0f * asin_c(mn(r.half_diagonal / r.near, 1.0f));
  const float rho = mn(GUARD * 0.5f * diameter * 1.0001f, 1.4f);
  const float tr = tan_c(rho);
  const float sec2 = 1.0f + tr * tr;
  const float delta = light_shift(a, b, page);
  const float spread = sec2 * (3.0f * (t / r.near) * (2.0f * r.half_diagonal / r.near) + rest / r.near) +
                       3.0f * tr * sec2 * sec2 * diameter * delta;
  const f3 au = ld3(a + 4), av = ld3(a + 8), bu = ld3(b + 4), bv = ld3(b + 8);
  const float area = mn(len3(cross3(au, av)), len3(cross3(bu, bv)));
  const float width = area * r.cos_min / (r.far * r.far) / (sec2 * diameter);
  return 2.0f * spread * 1.0001f > LATTICE_MARGIN * width;
}

// WGSL :4643-4652
__device__ __forceinline__ void ledger_reset(const VcArgs& a, uint32_t slot, uint32_t j) {
  uint32_t* r = a.ledger + ledger_base(a, slot, j);
  r[0] = a.frame;
  r[1] = 0x7f7fffffu;
  r[2] = 0u;
  r[3] = a.frame;
  r[4] = 0u;
  for (uint32_t w = 5; w < LEDGER_WORDS; w++) r[w] = 0u;
  a.ledger_bound[(size_t)slot * a.n_moving + j] = 0.0f;
}

__device__ __forceinline__ float hashed_share(uint32_t slot, uint32_t frame, uint32_t salt) {
  return 0.5f + 0.5f * (float)(key_hash(slot, frame, salt) & 0xffffu) / 65535.0f;
}

// The guard cone (WGSL :4499-4529) around the direction to light `r`'s
// centre from `apex`, and the light's angular radius there.
struct GuardCone {
  Cone c;
  float alpha;
};

__device__ __forceinline__ GuardCone guard_cone(f3 apex, const LightRect& r) {
  GuardCone out;
  out.c = Cone{};
  const f3 to = light_center(r) - apex;
  const float d = mx(len3(to), 1e-6f);
  const f3 forward = to / d;
  float alpha = 0.0f;
  if (is_sphere(r)) {
    alpha = asin_c(mn(r.edge_u.x / d, 1.0f));
  } else {
    const float* rf = reinterpret_cast<const float*>(&r);
#pragma unroll
    for (uint32_t k = 0; k < 4; k++)
      alpha = mx(alpha, acos_c(clampf(dot3(forward, norm3(rect_corner_of(rf, k) - apex)), -1.0f, 1.0f)));
  }
  const float wide = mn(GUARD * alpha, 1.4f);
  const float t = tan_c(wide);
  const f3 helper = fabsf(forward.x) > 0.9f ? f3{0.0f, 1.0f, 0.0f} : f3{1.0f, 0.0f, 0.0f};
  const f3 right = norm3(cross3(helper, forward));
  out.c.apex = apex;
  out.c.right = right;
  out.c.up = cross3(forward, right);
  out.c.forward = forward;
  out.c.w0 = -t;
  out.c.w1 = -t;
  out.c.w2 = t;
  out.c.w3 = t;
  out.c.far_normal = forward;
  out.c.far_offset = dot3(forward, apex) + SKY_FAR;
  bound_cone(out.c, forward);
  out.alpha = alpha;
  return out;
}

// WGSL :4532-4539
__device__ __forceinline__ float below_horizon(f3 apex, f3 n, const LightRect& r) {
  const f3 to = light_center(r) - apex;
  const float d = mx(len3(to), 1e-6f);
  float below = -dot3(n, to);
  if (is_sphere(r)) below -= r.edge_u.x;
  return asin_c(clampf(below / d, -1.0f, 1.0f));
}

struct LatticeFrame {
  f3 forward, right, up;
  float beta_x, beta_y, alpha;
  bool ok;
};

// WGSL :4727-4760
__device__ __forceinline__ LatticeFrame lattice_frame(f3 apex, const float* then) {
  LatticeFrame f;
  f.ok = false;
  f.forward = f.right = f.up = f3{0.0f, 0.0f, 0.0f};
  f.beta_x = f.beta_y = f.alpha = 0.0f;
  const f3 to = light_center_of(then) - apex;
  const float d = len3(to);
  if (d <= 1e-6f) return f;
  const f3 forward = to / d;
  const f3 eu = ld3(then + 4), ev = ld3(then + 8);
  f3 right = eu - forward * dot3(forward, eu);
  if (len3(right) <= 1e-6f * len3(eu)) right = ev - forward * dot3(forward, ev);
  if (len3(right) <= 0.0f) return f;
  right = norm3(right);
  const f3 up = cross3(forward, right);
  float lox = 3.4e38f, loy = 3.4e38f, hix = -3.4e38f, hiy = -3.4e38f, alpha = 0.0f;
  for (uint32_t k = 0; k < 4; k++) {
    const f3 p = rect_corner_of(then, k) - apex;
    const float z = dot3(forward, p);
    if (z <= 1e-6f * d) return f;
    const float ux = dot3(right, p) / z, uy = dot3(up, p) / z;
    lox = mn(lox, ux), loy = mn(loy, uy), hix = mx(hix, ux), hiy = mx(hiy, uy);
    alpha = mx(alpha, atan2_c(len3(cross3(forward, p)), z));
  }
  f.forward = forward;
  f.right = right;
  f.up = up;
  const float share = (1.0f + LATTICE_MARGIN) / 7.0f;
  f.beta_x = mx(hix - lox, 1e-12f) * share;
  f.beta_y = mx(hiy - loy, 1e-12f) * share;
  f.alpha = alpha * 1.0001f;
  f.ok = true;
  return f;
}

// WGSL :4766-4773, (1 + t^2)^1.5 as q sqrt(q)
__device__ __forceinline__ float lattice_kappa(const LatticeFrame& f, f3 n) {
  if (!f.ok) return 3.0e38f;
  const float reach = mn(GUARD * f.alpha, 1.4f);
  const float theta = atan2_c(len3(cross3(n, f.forward)), dot3(n, f.forward));
  if (theta + reach >= 1.5706f) return 3.0e38f;
  const float t = tan_c(reach);
  const float q = 1.0f + t * t;
  return cos_c(mx(theta - reach, 0.0f)) / cos_c(theta + reach) * (q * sqrtf(q)) * 1.0001f;
}

// The lattice cone (WGSL :4786-4826) of rect light `now` on the lattice of
// pose `then`, and the footprint's corners in bins.
struct LatticeCone {
  Cone c;
  f2 q[4];
  bool ok;
};

__device__ __forceinline__ LatticeCone lattice_cone(f3 apex, const float* now, const float* then) {
  LatticeCone out;
  out.ok = false;
  out.c = Cone{};
#pragma unroll
  for (int i = 0; i < 4; i++) out.q[i] = f2{0.0f, 0.0f};
  const LatticeFrame f = lattice_frame(apex, then);
  if (!f.ok) return out;
  const f3 n = norm3(cross3(ld3(now + 4), ld3(now + 8)));
  if (now[3] > 0.5f && dot3(n, apex - ld3(now)) <= 0.0f) return out;
  f2 pts[4];
  float lox = 3.4e38f, loy = 3.4e38f, hix = -3.4e38f, hiy = -3.4e38f;
  for (uint32_t k = 0; k < 4; k++) {
    const f3 p = rect_corner_of(now, k) - apex;
    const float z = dot3(f.forward, p);
    if (z <= 1e-6f) return out;
    pts[k] = f2{dot3(f.right, p) / z, dot3(f.up, p) / z};
    lox = mn(lox, pts[k].x), loy = mn(loy, pts[k].y), hix = mx(hix, pts[k].x), hiy = mx(hiy, pts[k].y);
  }
  const float ox = floorf(lox / f.beta_x) * f.beta_x, oy = floorf(loy / f.beta_y) * f.beta_y;
  const float ex = ox + 8.0f * f.beta_x, ey = oy + 8.0f * f.beta_y;
  if (hix > ex || hiy > ey) return out;
  f3 far_n = n;
  if (dot3(far_n, ld3(now) - apex) < 0.0f) far_n = f3{-far_n.x, -far_n.y, -far_n.z};
  const float dist = dot3(far_n, ld3(now) - apex);
  if (dist <= 1e-6f) return out;
  out.c.apex = apex;
  out.c.right = f.right;
  out.c.up = f.up;
  out.c.forward = f.forward;
  out.c.w0 = ox;
  out.c.w1 = oy;
  out.c.w2 = ex;
  out.c.w3 = ey;
  out.c.far_normal = far_n;
  out.c.far_offset = dot3(far_n, apex) + dist * ONE_MINUS_1EM5;
  bound_cone(out.c, norm3(light_center_of(now) - apex));
  const uint32_t order[4] = {0, 1, 3, 2};
#pragma unroll
  for (int i = 0; i < 4; i++)
    out.q[i] = f2{(pts[order[i]].x - ox) / f.beta_x, (pts[order[i]].y - oy) / f.beta_y};
  out.ok = true;
  return out;
}

// The footprint drawn into the lane's accumulator (WGSL :4830-4848 before
// its prefix sum): bin b's area is min(|sum of cells 0..b|, 1).
__device__ __forceinline__ void lattice_draw(Lane& L, const f2 q[4]) {
  clear_acc(L);
  const float g = (float)GRID;
#pragma unroll
  for (uint32_t i = 0; i < 4; i++) {
    const f2 a{clampf(q[i].x, 0.0f, g), clampf(q[i].y, 0.0f, g)};
    const f2 e{clampf(q[(i + 1) % 4].x, 0.0f, g), clampf(q[(i + 1) % 4].y, 0.0f, g)};
    draw_edge(L, a, e, 1.0f);
  }
}

struct LatticeOut {
  float ratio;  // -1 where the lattice cannot hold the footprint, or not traced
  bool empty, full;
};

// WGSL :4875-4899 `lattice_ratio` with :4853-4869 `lattice_share`, for the
// lanes with `on` (every lane takes part in the walk).
template <bool COUNT, bool QUEUE = false>
__device__ __forceinline__ LatticeOut lattice_ratio(const Ctx& x, const QueryParams& p, Lane& L, bool on, f3 apex,
                                                    f3 normal, const float* now, const float* then, float* cov,
                                                    uint64_t* work) {
  const LatticeCone lc = lattice_cone(apex, now, then);
  const bool traced = on && lc.ok;
  Walk w;
  w.cov = cov;
  w.full = 0;
  lattice_draw(L, lc.q);
  float running = 0.0f;
  for (uint32_t b = 0; b < BINS; b++) {
    running += L.get(b);
    const float m = mn(fabsf(running), 1.0f);
    w.cov[b] = traced && m <= 0.0f ? 1.0f : 0.0f;
    if (w.cov[b] >= 1.0f) w.full++;
  }
  w.cone_work = 0;
  w.cone_cap = NONE;
  w.prefilter_bins = p.prefilter;
  w.sky_max = false;
  w.overflow = false;
  w.truncated = false;
#pragma unroll
  for (int k = 0; k < 6; k++) w.work[k] = 0;
  if (uni(ballot(traced) != 0ull ? 1u : 0u)) scene_coverage<COUNT, false, false, QUEUE>(x, lc.c, traced, w, L);
  if (COUNT && work) {
#pragma unroll
    for (int k = 0; k < 6; k++) {
      uint32_t v = w.work[k];
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) v += (uint32_t)__shfl_xor((int)v, off, 64);
      if (lane_id() == 0) atomicAdd(reinterpret_cast<unsigned long long*>(&work[k]), (unsigned long long)v);
    }
  }
  // The footprint again (the walk used the accumulator), then the share.
  lattice_draw(L, lc.q);
  LatticeOut o;
  o.empty = true;
  o.full = true;
  running = 0.0f;
  const float g = (float)GRID;
  float total = 0.0f, open = 0.0f;
  for (uint32_t b = 0; b < BINS; b++) {
    running += L.get(b);
    const float m = mn(fabsf(running), 1.0f);
    const float cb = w.cov[b];
    if (cb > 0.0f && m > 0.0f) o.empty = false;
    if (cb < 1.0f) o.full = false;
    if (m > 0.0f) {
      const f3 d = direction(lc.c, ((float)(b % GRID) + 0.5f) / g, ((float)(b / GRID) + 0.5f) / g);
      const float l2 = dot3(d, d);
      const float wt = m * mx(dot3(normal, d), 0.0f) / (l2 * l2);
      total += wt;
      open += wt * (1.0f - cb);
    }
  }
  o.ratio = !traced ? -1.0f : (total <= 0.0f ? 1.0f : open / total);
  return o;
}

// The guard cone's walk (WGSL :4576-4583) for the lanes with `trace`:
// whether any bin is touched. The reference walks the whole guard cone and
// then looks for a touched bin; a lane here stops at the first mesh or
// splat that lands coverage in a bin, which settles the same answer.
template <bool COUNT, bool QUEUE = false>
__device__ __forceinline__ bool guard_touched(const Ctx& x, const QueryParams& p, Lane& L, const Cone& c,
                                              bool trace, float* cov, uint64_t* work) {
  Walk w;
  w.cov = cov;
#pragma unroll
  for (uint32_t b = 0; b < BINS; b++) w.cov[b] = 0.0f;
  w.full = 0;
  w.cone_work = 0;
  w.cone_cap = NONE;
  w.prefilter_bins = p.prefilter;
  w.sky_max = false;
  w.overflow = false;
  w.truncated = false;
  w.stop_on_touch = true;
#pragma unroll
  for (int k = 0; k < 6; k++) w.work[k] = 0;
  if (uni(ballot(trace) != 0ull ? 1u : 0u)) scene_coverage<COUNT, false, false, QUEUE>(x, c, trace, w, L);
  if (COUNT && work) {
#pragma unroll
    for (int k = 0; k < 6; k++) {
      uint32_t v = w.work[k];
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) v += (uint32_t)__shfl_xor((int)v, off, 64);
      if (lane_id() == 0) atomicAdd(reinterpret_cast<unsigned long long*>(&work[k]), (unsigned long long)v);
    }
  }
  bool any = false;
  for (uint32_t b = 0; b < BINS; b++) any = any || w.cov[b] > 0.0f;
  return trace && any;
}


}  // namespace

// WGSL :3394-3398, and this build's use marks and cleared slots: updates
// are [n_table x (entry, key, slot)] [n_used slots] [n_cleared slots].
__global__ __launch_bounds__(256) void vc_update_kernel(VcArgs a, const uint32_t* __restrict__ upd, uint32_t n_table,
                                                        uint32_t n_used, uint32_t n_cleared) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i < n_table) {
    const uint32_t* u = upd + 5 * (size_t)i;
    uint32_t* e = a.page_table + 4 * (size_t)u[0];
    e[0] = u[1], e[1] = u[2], e[2] = u[3], e[3] = u[4];
  } else if (i < n_table + n_used) {
    a.slot_used[upd[5 * (size_t)n_table + (i - n_table)]] = a.frame;
  } else if (i < n_table + n_used + n_cleared) {
    const uint32_t s = upd[5 * (size_t)n_table + n_used + (i - n_table - n_used)];
    for (int k = 0; k < 12; k++) a.slot_info[12 * (size_t)s + k] = 0u;
  }
}

// WGSL :4233-4379, rigid movers (no deforming catch-up, no sky): one lane
// per slot.
__global__ __launch_bounds__(256) void vc_stale_kernel(VcArgs a) {
  const uint32_t s = blockIdx.x * 256u + threadIdx.x;
  if (s >= a.n_slots) return;
  const uint32_t* info = a.slot_info + 12 * (size_t)s;
  if (info[3] == 0u) return;
  const uint32_t mark = __atomic_load_n(&a.slot_stale[s], __ATOMIC_RELAXED);
  if (mark == 1u) return;
  const uint32_t m = info[0];
  const uint32_t key[3] = {info[0], info[1], info[2]};
  const bool visible = a.slot_used[s] + a.visible_frames >= a.frame;
  if (visible) atomicAdd(&a.stats[2], 1u);
  const f3 lo{__uint_as_float(info[4]), __uint_as_float(info[5]), __uint_as_float(info[6])};
  const f3 hi{__uint_as_float(info[8]), __uint_as_float(info[9]), __uint_as_float(info[10])};
  const Box3 page = world_box(a, m, lo, hi);
  const uint32_t lights_n = a.n_lights;
  const uint32_t first = 2 * lights_n + 2, stride = 4 + 4 * lights_n;
  // A page on a moved object is stale whole.
  bool own = false;
  uint32_t own_mask = 0;
  if (boxes_meet(page, moved_box(a, 2 * lights_n))) {
    for (uint32_t k = 0; k < a.n_movers; k++) {
      if (moved_w(a, first + k * stride) == m) {
        own = true;
        own_mask = moved_w(a, first + k * stride + 2);
      }
    }
  }
  if (own && own_mask != 0u) {
    if (!visible) {
      mark_invalid(a, s);
      return;
    }
    const uint32_t entry = append_page(a, s, key, lo, hi);
    if (entry == NONE) {
      mark_invalid(a, s);
      return;
    }
    const uint32_t all = lights_n >= 32 ? 0xffffffffu : (1u << lights_n) - 1u;
    const uint32_t updated = (uint32_t)__builtin_popcount(own_mask & all);
    const uint32_t j = append_jobs(a, VC_TEXELS * updated);
    if (j == NONE) {
      atomicAdd(&a.stats[3], 1u);
      fail_entry(a, entry, s);
      return;
    }
    atomicAdd(&a.stats[6], 1u);
    uint32_t at = j;
    for (uint32_t l = 0; l < lights_n; l++) {
      if ((own_mask & (1u << l)) == 0u) continue;
      for (uint32_t k = 0; k < VC_TEXELS; k++) {
        a.bake_jobs[2 * (size_t)(at + k)] = entry;
        a.bake_jobs[2 * (size_t)(at + k) + 1] = k | (l << 8);
      }
      at += VC_TEXELS;
    }
    return;
  }
  bool candidate = false;
  for (uint32_t l = 0; l < lights_n && !candidate; l++) candidate = may_be_shadowed(a, page, l, s);
  if (!candidate) return;
  if (mark != 0u) {
    mark_invalid(a, s);
    return;
  }
  const uint32_t c = atomicAdd(&a.stats[4], 1u);
  if (c >= a.candidate_cap) {
    mark_invalid(a, s);
    return;
  }
  uint32_t entry = NONE;
  if (visible) {
    entry = append_page(a, s, key, lo, hi);
    if (entry == NONE) {
      a.candidates[2 * (size_t)c] = NONE;
      a.candidates[2 * (size_t)c + 1] = NONE;
      // Only rigid movers reach here, so the reference's `lag_only` is false.
      mark_invalid(a, s);
      return;
    }
  }
  a.candidates[2 * (size_t)c] = s;
  a.candidates[2 * (size_t)c + 1] = entry;
}

// WGSL :4385-4432: one lane per candidate page's texel.
__global__ __launch_bounds__(256) void vc_stale_texels_kernel(VcArgs a) {
  const uint32_t j = blockIdx.x * 256u + threadIdx.x;
  const uint32_t c = j / VC_TEXELS, k = j % VC_TEXELS;
  const uint32_t listed = __atomic_load_n(&a.stats[4], __ATOMIC_RELAXED);
  if (c >= (listed < a.candidate_cap ? listed : a.candidate_cap)) return;
  const uint32_t s = a.candidates[2 * (size_t)c], cand_entry = a.candidates[2 * (size_t)c + 1];
  if (s == NONE) return;
  const uint32_t* info = a.slot_info + 12 * (size_t)s;
  const uint32_t m = info[0];
  const uint32_t chart = info[1] & vc::KEY_CHART_MASK, mip = info[1] >> vc::KEY_MIP_SHIFT;
  uint32_t n[2];
  vc_grid(a, chart, m, mip, n);
  const uint32_t tx = (info[2] & 0xffffu) * vc::PAGE + k % VC_STRIDE;
  const uint32_t ty = (info[2] >> 16) * vc::PAGE + k / VC_STRIDE;
  if (tx >= n[0] || ty >= n[1]) return;
  const TexelSurface surface =
      vc_surface_at(a, m, chart, f2{((float)tx + 0.5f) / (float)n[0], ((float)ty + 0.5f) / (float)n[1]});
  if (!surface.ok) return;
  const f3 far = vmax3(vabs3(surface.apex), f3{1.0f, 1.0f, 1.0f});
  const float pad = 1e-5f * max_el(far);
  const Box3 apex{surface.apex - f3{pad, pad, pad}, surface.apex + f3{pad, pad, pad}};
  for (uint32_t l = 0; l < a.n_lights; l++) {
    if (!may_be_shadowed(a, apex, l, s)) continue;
    if (cand_entry == NONE) {
      mark_invalid(a, s);
      return;
    }
    const uint32_t at = append_jobs(a, 1u);
    if (at == NONE) {
      atomicAdd(&a.stats[3], 1u);
      fail_entry(a, cand_entry, s);
      return;
    }
    a.bake_jobs[2 * (size_t)at] = cand_entry;
    a.bake_jobs[2 * (size_t)at + 1] = k | (l << 8);
    if (atomicExch(&a.slot_stale[s], 1u) == 0u) atomicAdd(&a.stats[6], 1u);
  }
}

// WGSL :4655-4662 `vc_ledger_reset`: one lane per planned page, fresh
// records for every moving light (version: this frame).
__global__ __launch_bounds__(64) void vc_ledger_reset_kernel(VcArgs a) {
  const uint32_t p = blockIdx.x * 64u + threadIdx.x;
  if (p >= a.plan_pages) return;
  const uint32_t slot = a.bake_list[4 * (size_t)p + 3];
  for (uint32_t j = 0; j < a.n_moving; j++) ledger_reset(a, slot, j);
}

// WGSL :5007-5164 `vc_ledger`: one lane per slot. A page read in the frame
// before re-bakes its texels whose bound exceeds ε (the whole page past its
// slack or its pose ring); a page not read is set aside (kept, shown as
// missing) until a pixel wants it.
__global__ __launch_bounds__(256) void vc_ledger_kernel(VcArgs a) {
  const uint32_t s = blockIdx.x * 256u + threadIdx.x;
  if (s >= a.n_slots) return;
  const uint32_t* info = a.slot_info + 12 * (size_t)s;
  const uint32_t set_aside = __atomic_load_n(&a.slot_stale[s], __ATOMIC_RELAXED);
  if (info[3] == 0u || set_aside == 1u) return;
  const bool visible = a.slot_used[s] + 1u >= a.frame;
  const bool lit = (a.ledger_limit >> 31) == 1u;
  if (set_aside == 2u && !visible) return;
  if (set_aside == 0u && !lit) return;
  if (visible) atomicAdd(&a.stats[11], 1u);
  const f3 lo{__uint_as_float(info[4]), __uint_as_float(info[5]), __uint_as_float(info[6])};
  const f3 hi{__uint_as_float(info[8]), __uint_as_float(info[9]), __uint_as_float(info[10])};
  const Box3 page = world_box(a, info[0], lo, hi);
  const uint32_t key[3] = {info[0], info[1], info[2]};
  const uint32_t limit = a.ledger_limit & 0x7fffffffu;
  uint32_t entry = NONE;
  bool counted = false;
  for (uint32_t l = 0; l < a.n_lights; l++) {
    const uint32_t j = a.moving_index[l];
    if (j == NONE) continue;
    uint32_t* rec = a.ledger + ledger_base(a, s, j);
    const uint32_t version = rec[0];
    if (version == 0u) continue;
    const float* now = a.lights + 16 * (size_t)l;
    const float slack = __uint_as_float(rec[1]);
    const uint32_t page_flags = rec[4];
    const bool mixed_page = (page_flags & 1u) != 0u;
    const uint32_t age = a.frame - version;
    bool whole = age + 1u >= (VC_POSE_RING < MAX_OFFSET ? VC_POSE_RING : MAX_OFFSET);
    const bool lattice = !(now[3] > 1.5f);
    if (lattice && !whole && (page_flags & 8u) != 0u) {
      whole = true;
      if (visible) atomicAdd(&a.stats[22], 1u);
    }
    uint32_t stale[3] = {0u, 0u, 0u};
    uint32_t count = 0;
    float bound = 0.0f;
    bool held = false;
    if (!whole) {
      const float d0 = light_shift(pose_at(a, version, l), now, page);
      if (d0 > slack * hashed_share(s, version, 0x5bd1e995u)) {
        whole = true;
      } else if (!mixed_page) {
        if (visible) atomicAdd(&a.stats[13], 1u);
        a.ledger_bound[(size_t)s * a.n_moving + j] = 0.0f;
        continue;
      } else {
        if (visible) atomicAdd(&a.stats[12], 1u);
        if (lattice) {
          const uint32_t since = rec[3];
          const float kappa = __uint_as_float(rec[2]);
          if (a.frame - since < VC_POSE_RING) {
            const float b = lattice_bound(pose_at(a, since, l), now, page, kappa);
            held = b <= a.epsilon * hashed_share(s, since, 0x27d4eb2du);
            if (held && lattice_grew(pose_at(a, version, l), now, page)) {
              held = false;
              if (visible) atomicAdd(&a.stats[22], 1u);
            }
            if (held) bound = b;
          }
        }
        const uint32_t want = held ? 7u : 3u;
        for (uint32_t k = 0; k < VC_TEXELS; k++) {
          const uint32_t half = (rec[5 + k / 2] >> (16u * (k % 2u))) & 0xffffu;
          if ((half & want) == want) {
            stale[k / 32] |= 1u << (k % 32u);
            count++;
          }
        }
        if (held && visible) {
          atomicAdd(&a.stats[21], 1u);
          atomicAdd(&a.stats[23], count);
        }
        whole = count == VC_TEXELS && !held;
      }
    }
    if (whole) count = VC_TEXELS;
    a.ledger_bound[(size_t)s * a.n_moving + j] = bound;
    if (count == 0u) continue;
    if (!visible) {
      if (atomicExch(&a.slot_stale[s], 2u) == 0u) atomicAdd(&a.stats[10], 1u);
      return;
    }
    if (entry == NONE) {
      entry = append_page(a, s, key, lo, hi);
      if (entry == NONE) {
        atomicExch(&a.slot_stale[s], 2u);
        atomicAdd(&a.stats[10], 1u);
        return;
      }
    }
    uint32_t first = NONE;
    if (atomicAdd(&a.stats[14], count) + count <= limit) first = append_jobs(a, count);
    if (first == NONE) {
      atomicAdd(&a.stats[10], 1u);
      a.bake_bounds[8 * (size_t)entry + 3] = 1.0f;
      atomicExch(&a.slot_stale[s], 2u);
      return;
    }
    if (whole) {
      ledger_reset(a, s, j);
      for (uint32_t k = 0; k < VC_TEXELS; k++) {
        a.bake_jobs[2 * (size_t)(first + k)] = entry;
        a.bake_jobs[2 * (size_t)(first + k) + 1] = k | (l << 8);
      }
    } else {
      uint32_t at = 0;
      for (uint32_t k = 0; k < VC_TEXELS; k++) {
        if ((stale[k / 32] & (1u << (k % 32u))) == 0u) continue;
        a.bake_jobs[2 * (size_t)(first + at)] = entry;
        a.bake_jobs[2 * (size_t)(first + at) + 1] = k | (l << 8);
        at++;
      }
      if (!held) rec[3] = a.frame;
    }
    atomicAdd(&a.stats[9], count);
    if (!counted) {
      counted = true;
      atomicAdd(&a.stats[8], 1u);
    }
  }
  if (set_aside == 2u && entry == NONE) atomicExch(&a.slot_stale[s], 0u);
}

namespace {

// WGSL :4547-4638 `ledger_texel` for the lanes with `rec_on` (a listed
// texel of a moving light, records kept): its state (2 clear, 3 mixed, 0
// fixed or off), the page's slack (the smallest guard, less the light's
// shift since the version), κ on a lattice, and the page's flags. The guard
// cone's walk runs as a packet for the lanes baked fresh at the version.
template <bool COUNT, bool QUEUE = false>
__device__ __forceinline__ void ledger_texel(const VcArgs& a, const Ctx& x, const QueryParams& p, Lane& L,
                                             bool rec_on, bool on, uint32_t slot, uint32_t j, uint32_t l, uint32_t k,
                                             f3 apex, f3 n, const LightRect& r, float open, const float* then,
                                             bool lattice, bool overflow, float* cov, uint64_t* work) {
  const uint32_t base = rec_on ? ledger_base(a, slot, j) : 0u;
  const uint32_t version = rec_on ? a.ledger[base] : a.frame;
  const uint32_t age = a.frame - version;
  const uint32_t offset = age < MAX_OFFSET ? age : MAX_OFFSET;
  uint32_t state = 0;
  float guard = 3.4e38f;
  float reach = 0.0f, margin = 3.4e38f, bin = 0.0f, kappa = 3.0e38f;
  if (rec_on && lattice && on && open > 0.0f) {
    const LatticeFrame f = lattice_frame(apex, then);
    if (f.ok) {
      reach = mn(GUARD * f.alpha, 1.4f);
      bin = sqrtf(f.beta_x * f.beta_x + f.beta_y * f.beta_y);
      margin = f.alpha + bin;
      kappa = lattice_kappa(f, n);
    }
  }
  const GuardCone g = guard_cone(apex, r);
  const bool fresh = rec_on && on && open > 0.0f && offset == 0u;
  const bool any = guard_touched<COUNT, QUEUE>(x, p, L, g.c, fresh, cov, work);
  if (!rec_on) return;
  if (on && open > 0.0f && offset != 0u) {
    state = 3;
    if (lattice) guard = reach - margin;
  } else if (on && open > 0.0f) {
    state = any ? 3u : 2u;
    if (-below_horizon(apex, n, r) <= mx(2.0f * mn(GUARD * g.alpha, 1.4f), 0.1f)) state = 3;
    guard = mn(GUARD * g.alpha, 1.4f) - g.alpha;
    if (lattice) guard = mn(mn(GUARD * g.alpha, 1.4f), reach) - mx(margin, g.alpha + bin);
  } else if (on) {
    const float below = below_horizon(apex, n, r);
    if (below >= 2.0f * g.alpha)
      guard = below;
    else
      state = 3;
  }
  if ((state != 3u || lattice) && guard < 3.0e38f) {
    float slack = guard;
    if (offset != 0u) slack = guard - light_shift(pose_at(a, version, l), reinterpret_cast<const float*>(&r), Box3{apex, apex});
    atomicMin(&a.ledger[base + 1], __float_as_uint(mx(slack, 0.0f)));
  }
  uint32_t code = 0;
  if (lattice && state == 3u) {
    if (on && open > 0.0f && kappa <= LATTICE_KAPPA_MAX)
      atomicMax(&a.ledger[base + 2], __float_as_uint(kappa));
    else
      code = 4;
  }
  const uint32_t shift = 16u * (k % 2u), word = base + 5 + k / 2;
  atomicAnd(&a.ledger[word], ~(0xffffu << shift));
  atomicOr(&a.ledger[word], ((offset << 6) | code | state) << shift);
  const uint32_t flags = (state == 3u ? 1u : 0u) | (offset != 0u ? 2u : 0u) | (overflow ? 8u : 0u);
  if (flags != 0u) atomicOr(&a.ledger[base + 4], flags);
}

}  // namespace

// WGSL :3432-3554: one wave per 64 consecutive jobs, a grid-stride loop
// over the planned and appended jobs. With COUNT, the walks' work goes to
// `work` (as the cone kernel's) and capped walks to counters[0]. With
// MOVING (a cache with moving lights), a moving light's texel bakes as WGSL
// :3448-3530 and :3551-3553 say: a rect light on its page's lattice, ratios
// within ε of 0 or 1 snapped, and its ledger record kept. With QUEUE (leaves
// of at most 32 triangles), each lane queues the triangles it draws across
// a mesh's leaves and draws them in shared rounds (the walk's draw queue):
// a bake packet's texels, spread over a page, mostly draw different
// triangles.
template <bool COUNT, bool MOVING, bool QUEUE = false>
__global__ __launch_bounds__(64, WALK_WAVES) void vc_bake_kernel(VcArgs a, const NodeC* __restrict__ blas,
                                                     const NodeC* __restrict__ nodes,
                                                     const float* __restrict__ tri_leaf,
                                                     const uint32_t* __restrict__ items,
                                                     const uint32_t* __restrict__ blas_items,
                                                     const float* __restrict__ blas_cov, QueryParams p,
                                                     uint64_t* __restrict__ work, uint32_t* __restrict__ counters) {
  __shared__ float s_acc[ACC_CELLS * 64];
  const uint32_t lane = lane_id();
  float poly[POLY_SLOTS];
  float wcov[MOVING ? BINS : 1];  // the lattice and guard walks' coverage
  Lane L{s_acc + lane, poly, POLY_STRIDE, ACC_CELLS, 0u};
  const Ctx x{blas, nodes, tri_leaf, a.tris, a.ranges, items, blas_items, blas_cov, a.xforms,
              p.n_objects, p.hierarchy, p.profile_stage, p.object_cov};
  const uint32_t appended = a.stats[5] < a.append_jobs ? a.stats[5] : a.append_jobs;
  const uint32_t total = a.plan_jobs + appended;
  // Waves take jobs from a counter (stats[7], zero at the frame's start),
  // consecutive jobs on their first lanes: 64 at a time while many remain,
  // and fewer as the bake nears its end (down to 4 once fewer jobs remain
  // than 4 for each wave the GPU holds). A packet's walk visits the union
  // of its lanes' nodes, and a narrower packet's union is shorter, so the
  // last waves, which the frame waits for, finish sooner. A bake smaller
  // than one round of 64-job packets (a view's few new pages) shrinks its
  // packets as if four times as many jobs remained: waves then take several
  // narrow packets each, and its few long walks no longer set its length;
  // larger bakes keep their wide packets.
  const uint32_t spread = (uint64_t)total < (uint64_t)a.bake_group * a.bake_resident ? 4u : 1u;
  for (;;) {
    uint32_t start = 0, group = 0;
    if (lane == 0) {
      const uint32_t done = __atomic_load_n(&a.stats[7], __ATOMIC_RELAXED);
      const uint32_t left = done < total ? total - done : 0u;
      group = a.bake_group;
      while (group > 4u && (uint64_t)left < (uint64_t)group * a.bake_resident * spread) group >>= 1;
      start = atomicAdd(&a.stats[7], group);
    }
    start = uni(start);
    group = uni(group);
    if (start >= total) break;
    const uint32_t j = start + lane;
    const bool mine = lane < group && j < total;
    const uint32_t* job = a.bake_jobs + 2 * (size_t)(mine ? j : total - 1u);
    const uint32_t jx = job[0], jy = job[1];
    const bool listed = mine && jx != NONE;
    const uint32_t k = jy & 0xffu;
    const uint32_t l = (jy >> 8) < a.n_lights - 1u ? (jy >> 8) : a.n_lights - 1u;
    const PageTexel t = page_texel(a, jx == NONE ? 0u : jx, k);
    const TexelSurface s = vc_surface_at(
        a, t.m, t.c, f2{((float)t.texel[0] + 0.5f) / (float)t.n[0], ((float)t.texel[1] + 0.5f) / (float)t.n[1]});
    const bool on = listed && t.inside && s.ok;
    const LightRect r = reinterpret_cast<const LightRect*>(a.lights)[l];
    // unshadowed_rect (WGSL :2135-2141)
    float open = 0.0f;
    if (faces_light(r, s.apex, s.normal))
      open = is_sphere(r) ? spot_falloff(r, norm3(s.apex - f3{r.corner.x, r.corner.y, r.corner.z})) *
                                sphere_unshadowed(s.apex, s.normal, r)
                          : rect_open(L, s.apex, s.normal, r);
    // A moving light (WGSL :3448-3472): a rect one bakes on its page's
    // lattice, in the frame of its pose at the page's version, unless its
    // bins' weights spread too far; both walks run for every lane.
    uint32_t mj = NONE;
    const float* lf = a.lights + 16 * (size_t)l;
    const float* then = lf;
    bool on_lattice = false, overflow = false, lattice = false;
    LatticeOut lat{-1.0f, true, false};
    if (MOVING) {
      mj = a.moving_index[l];
      lattice = mj != NONE && !is_sphere(r);
      if (lattice && listed) {
        const uint32_t version = a.ledger[ledger_base(a, t.e[3], mj)];
        if (version != 0u && a.frame - version < VC_POSE_RING) then = pose_at(a, version, l);
      }
      on_lattice = lattice;
      if (lattice && on && open > 0.0f)
        on_lattice = lattice_kappa(lattice_frame(s.apex, then), s.normal) <= LATTICE_KAPPA_MAX;
      lat = lattice_ratio<COUNT, QUEUE>(x, p, L, on && on_lattice && open > 0.0f, s.apex, s.normal, lf, then, wcov,
                                 COUNT ? work : nullptr);
      overflow = on && on_lattice && open > 0.0f && lat.ratio < 0.0f;
    }
    const f3 toward = norm3(light_center(r) - s.apex);
    const bool back = on && open <= 0.0f && mj == NONE && faces_light(r, s.apex, toward);
    float seen = open;
    if (back) {
      seen = 0.0f;
      if (faces_light(r, s.apex, toward))
        seen = is_sphere(r) ? spot_falloff(r, norm3(s.apex - f3{r.corner.x, r.corner.y, r.corner.z})) *
                                  sphere_unshadowed(s.apex, toward, r)
                            : rect_open(L, s.apex, toward, r);
    }
    const Skip skip{back ? t.m : NONE, back ? s.tri : NONE};
    const QueryOut o = evaluate_at<COUNT, true, QUEUE>(x, p, L, listed, s.apex, back ? toward : s.normal,
                                                on && (!on_lattice || overflow), r, nullptr, COUNT ? work : nullptr,
                                                skip);
    if (COUNT) {
      const uint64_t capped = ballot(listed && o.capped);
      if (lane == 0 && capped) atomicAdd(&counters[0], (uint32_t)__builtin_popcountll(capped));
    }
    const float value = o.visible;
    const bool last_empty = o.touched == 0ull, last_full = o.full == ~0ull;
    float ratio = !on ? 0.0f : (seen > 0.0f ? clampf(value / seen, 0.0f, 1.0f) : 1.0f);
    if (on && seen > 0.0f && last_empty) ratio = 1.0f;
    if (on && seen > 0.0f && last_full) ratio = 0.0f;
    // A moving light's ratio within ε of 0 or 1 is that (WGSL :3515-3520).
    if (MOVING && mj != NONE) {
      if (ratio < a.epsilon) ratio = 0.0f;
      if (ratio > 1.0f - a.epsilon) ratio = 1.0f;
    }
    float light_value = back ? 0.0f : value;
    if (MOVING && on_lattice && !overflow) {
      ratio = !on ? 0.0f : (open > 0.0f ? lat.ratio : 1.0f);
      if (on && open > 0.0f && lat.empty) ratio = 1.0f;
      if (on && open > 0.0f && lat.full) ratio = 0.0f;
      light_value = ratio * open;
    }
    if (listed && t.inside)
      a.pool[((size_t)t.e[3] * VC_TEXELS + k) * a.channels + l] = a.light_ratio ? ratio : (on ? light_value : 0.0f);
    if (MOVING)
      ledger_texel<COUNT, QUEUE>(a, x, p, L, mj != NONE && listed && t.inside && a.ledger_record != 0u, on, t.e[3], mj, l,
                          k, s.apex, s.normal, r, open, then, lattice, overflow, wcov, COUNT ? work : nullptr);
  }
}

// WGSL :3693-3836 without sky or bounce light: one wave per baked page.
__global__ __launch_bounds__(64) void vc_fill_kernel(VcArgs a) {
  const uint32_t lane = lane_id();
  const uint32_t appended = a.stats[0] < a.append_pages ? a.stats[0] : a.append_pages;
  const uint32_t count = a.plan_pages + appended;
  for (uint32_t p = blockIdx.x; p < count; p += gridDim.x) {
    const uint32_t* e = a.bake_list + 4 * (size_t)p;
    const float* bb = a.bake_bounds + 8 * (size_t)p;
    if (bb[3] != 0.0f) continue;  // a page whose stale texels did not all fit stays invalid
    const uint32_t slot = e[3];
    uint32_t* info = a.slot_info + 12 * (size_t)slot;
    const float kind = bb[7];
    if (lane == 0) {
      const uint32_t prev_w = info[3];
      const bool same = prev_w != 0u && info[0] == e[0] && info[1] == e[1] && info[2] == e[2];
      const bool fresh = p < a.plan_pages && (kind == 0.0f || kind == 2.0f || kind == 5.0f);
      const bool keep = same && !fresh;
      const uint32_t charge = keep ? info[7] : 0u;
      const uint32_t moving = (fresh && a.n_movers > 0u) ? 4u : 0u;
      info[0] = e[0], info[1] = e[1], info[2] = e[2];
      info[3] = 1u | (keep ? (prev_w & 0xffffff06u) : 0u) | moving;
      info[4] = __float_as_uint(bb[0]), info[5] = __float_as_uint(bb[1]), info[6] = __float_as_uint(bb[2]);
      info[7] = charge;
      for (int q = 0; q < 4; q++) info[8 + q] = __float_as_uint(bb[4 + q]);
      if (kind == 5.0f) {
      } else if (kind >= 3.0f) {
        atomicCAS(&a.slot_stale[slot], 3u, 0u);
      } else {
        atomicExch(&a.slot_stale[slot], kind == 2.0f ? 3u : 0u);
      }
      if (p < a.plan_pages) a.slot_used[slot] = a.frame;
    }
    const uint32_t c = e[1] & vc::KEY_CHART_MASK, mip = e[1] >> vc::KEY_MIP_SHIFT;
    uint32_t n[2];
    vc_grid(a, c, e[0], mip, n);
    const uint32_t px = (e[2] & 0xffffu) * vc::PAGE, py = (e[2] >> 16) * vc::PAGE;
    float* o = a.pool + (size_t)slot * VC_TEXELS * a.channels;
    const uint32_t sky = a.channels - 1u;
    for (uint32_t k = lane; k < VC_TEXELS; k += 64u) {
      // No sky here: its channel is zero, in the border as inside.
      o[k * a.channels + sky] = 0.0f;
      const uint32_t tx = px + k % VC_STRIDE, ty = py + k / VC_STRIDE;
      if (tx >= n[0] || ty >= n[1]) {
        const uint32_t ix = (tx < n[0] - 1u ? tx : n[0] - 1u) - px;
        const uint32_t iy = (ty < n[1] - 1u ? ty : n[1] - 1u) - py;
        const uint32_t nearest = iy * VC_STRIDE + ix;
        for (uint32_t ch = 0; ch < sky; ch++) o[k * a.channels + ch] = o[nearest * a.channels + ch];
      }
    }
  }
}

// ---- shading from the cache (WGSL :5907-6130, the direct light's part)

struct ShadeArgs {
  const float* __restrict__ camera;  // 20 floats: eye, forward, right, up, params
  const uint32_t* __restrict__ hits; // 4 per pixel: object, triangle, distance bits, seen light + 1
  const uint32_t* __restrict__ tri_chart;
  float* __restrict__ out;           // per light and pixel: l * w * h + p
  uint32_t* __restrict__ out_page;   // per pixel: the shown page's mip, top bit set when none was found
  uint32_t* request_set;             // fingerprints (zeroed before the pass)
  uint32_t* request_prio;            // per set entry: the most urgent priority asked (all ones before)
  uint32_t* requests;                // 4 per request: key, then its set entry (its priority after the gather)
  uint32_t* counts;                  // [requests, pixels that fell back]
  uint32_t w, h, set_mask, request_cap, tile, refine;
  float pixel_angle, pixels_per_texel;
};

namespace {

// WGSL :3134-3160: a page marked stale (1) or set aside by the ledger (2)
// is missing; one set aside is marked wanted, so the next frame's ledger
// pass re-bakes what it lacks.
__device__ __forceinline__ uint32_t vc_find(const VcArgs& a, uint32_t k0, uint32_t k1, uint32_t k2) {
  uint32_t at = key_hash(k0, k1, k2) & a.table_mask;
  for (uint32_t probe = 0; probe < vc::MAX_PROBE; probe++) {
    const uint32_t* e = a.page_table + 4 * (size_t)at;
    if (e[0] == vc::EMPTY) return NONE;
    if (e[0] == k0 && e[1] == k1 && e[2] == k2) {
      const uint32_t stale = a.slot_stale[e[3]];
      if (stale == 2u) a.slot_used[e[3]] = a.frame;
      return (stale != 0u && stale != 3u) ? NONE : e[3];
    }
    at = (at + 1u) & a.table_mask;
  }
  return NONE;
}

__device__ __forceinline__ f3 point_to_object(const VcArgs& a, uint32_t m, f3 p) {
  const uint32_t x = a.ranges[4 * m + 3];
  if (x == NONE) return p;
  const float* xf = a.xforms + 16 * (size_t)x;
  const f3 d = p - translation(xf);
  // rotate_to_object: the rows' transpose (WGSL :61)
  return (row3(xf, 0) * d.x + row3(xf, 1) * d.y + row3(xf, 2) * d.z) * xf[13];
}

// WGSL :1780-1793
__device__ __forceinline__ void barycentric(const VcArgs& a, uint32_t m, uint32_t t, f3 world_p, float b[3]) {
  const f3 p = point_to_object(a, m, world_p);
  const float* tr = a.tris + 12 * (size_t)t;
  const f3 v0 = ld3(tr);
  const f3 e1 = ld3(tr + 4) - v0, e2 = ld3(tr + 8) - v0;
  const f3 d = p - v0;
  const float d11 = dot3(e1, e1), d12 = dot3(e1, e2), d22 = dot3(e2, e2);
  const float denom = d11 * d22 - d12 * d12;
  const float sv = (d22 * dot3(d, e1) - d12 * dot3(d, e2)) / denom;
  const float u = (d11 * dot3(d, e2) - d12 * dot3(d, e1)) / denom;
  b[0] = 1.0f - sv - u;
  b[1] = sv;
  b[2] = u;
}

__device__ __forceinline__ f3 normal_to_world(const VcArgs& a, uint32_t m, f3 n) {
  const uint32_t x = a.ranges[4 * m + 3];
  if (x == NONE) return n;
  return rotate_to_world(a.xforms + 16 * (size_t)x, n);
}

// WGSL :3187-3196
__device__ __forceinline__ uint32_t vc_top_mip(const VcArgs& a, uint32_t c, uint32_t m) {
  const vc::Chart ch = a.charts[c];
  const float ratio = mx(ch.size[0], ch.size[1]) * object_scale(a, m) * a.inv_base;
  uint32_t top = 0;
  if (ratio > 1.0f) {
    const float l = ceilf(log2f(ratio));
    top = l < 30.0f ? (uint32_t)l : 30u;
  }
  uint32_t n[2];
  if (top > 0) {
    vc_grid(a, c, m, top - 1, n);
    if (n[0] == 1 && n[1] == 1) top -= 1;
  }
  vc_grid(a, c, m, top, n);
  if (n[0] != 1 || n[1] != 1) top += 1;
  return top;
}

struct PageRef {
  uint32_t key[3];
  uint32_t base[2];
  float frac[2];
  bool coarsest;
};

// WGSL :3218-3231
__device__ __forceinline__ PageRef vc_page(const VcArgs& a, uint32_t m, uint32_t c, f2 uv, uint32_t mip) {
  uint32_t n[2];
  vc_grid(a, c, m, mip, n);
  const float tx = clampf(uv.x * (float)n[0] - 0.5f, 0.0f, (float)(n[0] - 1u));
  const float ty = clampf(uv.y * (float)n[1] - 0.5f, 0.0f, (float)(n[1] - 1u));
  const uint32_t bx = (uint32_t)floorf(tx), by = (uint32_t)floorf(ty);
  const uint32_t px = bx / vc::PAGE, py = by / vc::PAGE;
  PageRef r;
  r.key[0] = m;
  r.key[1] = c | (mip << vc::KEY_MIP_SHIFT);
  r.key[2] = px | (py << 16);
  r.base[0] = bx - px * vc::PAGE;
  r.base[1] = by - py * vc::PAGE;
  r.frac[0] = tx - floorf(tx);
  r.frac[1] = ty - floorf(ty);
  r.coarsest = n[0] == 1 && n[1] == 1;
  return r;
}

// WGSL :3242-3247
__device__ __forceinline__ float vc_bilinear(const VcArgs& a, uint32_t slot, const PageRef& r, uint32_t ch) {
  const uint32_t o = slot * VC_TEXELS + r.base[1] * VC_STRIDE + r.base[0];
  const float* p = a.pool;
  const uint32_t C = a.channels;
  const float x0 = mixf(p[(size_t)o * C + ch], p[(size_t)(o + 1) * C + ch], r.frac[0]);
  const float x1 = mixf(p[(size_t)(o + VC_STRIDE) * C + ch], p[(size_t)(o + VC_STRIDE + 1) * C + ch], r.frac[0]);
  return mixf(x0, x1, r.frac[1]);
}

// WGSL :5907-5925 (a strong compare-and-swap: no spurious failures to
// retry). A page several pixels ask for keeps the most urgent (smallest)
// of their priorities, whichever pixel lists it.
__device__ __forceinline__ void vc_request(const ShadeArgs& s, const uint32_t k[3], uint32_t priority) {
  const uint32_t fp = key_hash(k[0], k[1], k[2]) | 1u;
  uint32_t i = key_hash(k[2], k[0], k[1]) & s.set_mask;
  for (uint32_t probe = 0; probe < 32u; probe++) {
    const uint32_t old = atomicCAS(&s.request_set[i], 0u, fp);
    if (old == 0u) {
      atomicMin(&s.request_prio[i], priority);
      const uint32_t n = atomicAdd(&s.counts[0], 1u);
      if (n < s.request_cap) {
        uint32_t* r = s.requests + 4 * (size_t)n;
        r[0] = k[0], r[1] = k[1], r[2] = k[2], r[3] = i;
      }
      return;
    }
    if (old == fp) {
      atomicMin(&s.request_prio[i], priority);
      return;
    }
    i = (i + 1u) & s.set_mask;
  }
}

}  // namespace

// WGSL :5937-6024 and :6078-6130 for the lights' direct part: one lane per
// pixel of the camera pass's hits; per light the cached share times the
// unshadowed light at the pixel (or, without light_ratio, the cached value),
// the unshadowed light where no page is resident. One pixel per request
// tile (a different one each frame) asks for the page one to `refine` mips
// finer than the one it found, or for the coarsest missing one.
__global__ __launch_bounds__(256) void vc_shade_kernel(VcArgs a, ShadeArgs s) {
  const uint32_t p = blockIdx.x * 256u + threadIdx.x;
  const uint32_t n_px = s.w * s.h;
  const bool in_bounds = p < n_px;
  const uint32_t x = in_bounds ? p % s.w : 0u, y = in_bounds ? p / s.w : 0u;
  const uint32_t sample = (a.frame * 7u) % (s.tile * s.tile);
  const bool is_sample = in_bounds && x % s.tile == sample % s.tile && y % s.tile == sample / s.tile;
  uint32_t req_key[3] = {0u, 0u, 0u}, priority = 0u;
  bool fell_back = false;
  if (in_bounds) {
    const float* cam = s.camera;
    const float* pp = cam + 16;
    const float aspect = pp[0] / pp[1];
    const float sx = (((float)x + 0.5f + cam[11]) / pp[0] * 2.0f - 1.0f) * pp[2] * aspect;
    const float sy = (1.0f - ((float)y + 0.5f + cam[15]) / pp[1] * 2.0f) * pp[2];
    const f3 dir = norm3(ld3(cam + 4) + ld3(cam + 8) * sx + ld3(cam + 12) * sy);
    const uint32_t* v = s.hits + 4 * (size_t)p;
    const uint32_t m = v[0], t = v[1];
    if (m != NONE && v[3] == 0u) {
      const float distance = __uint_as_float(v[2]);
      const float* tr = a.tris + 12 * (size_t)t;
      const f3 v0 = ld3(tr);
      const f3 face = cross3(ld3(tr + 4) - v0, ld3(tr + 8) - v0);
      f3 hn = norm3(face);
      if (a.ranges[4 * m + 3] != NONE) hn = norm3(normal_to_world(a, m, hn));
      if (dot3(hn, dir) > 0.0f) hn = -hn;
      const f3 surface = ld3(cam) + dir * distance;
      const uint32_t c = s.tri_chart[t];
      float b[3];
      barycentric(a, m, t, surface, b);
      const float* q = a.tri_uv + 6 * (size_t)t;
      const f2 uv = f2{q[0], q[1]} * b[0] + f2{q[2], q[3]} * b[1] + f2{q[4], q[5]} * b[2];
      // WGSL :3198-3202 `vc_mip`
      const float ratio = distance * s.pixel_angle * s.pixels_per_texel * a.inv_base;
      uint32_t want = 0;
      if (ratio > 1.0f) {
        const float lr = rintf(log2f(ratio));
        want = lr < 30.0f ? (uint32_t)lr : 30u;
      }
      const uint32_t top = vc_top_mip(a, c, m);
      const uint32_t wanted = want < top ? want : top;
      // WGSL :6024-6070 `vc_find_shown`
      uint32_t mip = wanted, slot = NONE;
      uint32_t finer[4][3] = {{0u, 0u, 0u}, {0u, 0u, 0u}, {0u, 0u, 0u}, {0u, 0u, 0u}};
      PageRef r;
      bool found = false;
      for (;;) {
        r = vc_page(a, m, c, uv, mip);
        slot = vc_find(a, r.key[0], r.key[1], r.key[2]);
        if (slot != NONE) {
          a.slot_used[slot] = a.frame;
          found = true;
          if (mip > wanted) {
            const uint32_t step = s.refine;
            const uint32_t* pick = finer[0];
            if (step >= 2u && mip >= wanted + 2u) pick = finer[1];
            if (step >= 3u && mip >= wanted + 3u) pick = finer[2];
            if (step >= 4u && mip >= wanted + 4u) pick = finer[3];
            req_key[0] = pick[0], req_key[1] = pick[1], req_key[2] = pick[2];
            const uint32_t gap = mip - wanted;
            priority = 32u - (gap < 30u ? gap : 30u);
          }
          break;
        }
        for (int k = 0; k < 3; k++) {
          finer[3][k] = finer[2][k];
          finer[2][k] = finer[1][k];
          finer[1][k] = finer[0][k];
          finer[0][k] = r.key[k];
        }
        if (r.coarsest || mip >= 30u) break;
        mip += 1;
      }
      if (!found) {
        req_key[0] = finer[0][0], req_key[1] = finer[0][1], req_key[2] = finer[0][2];
        priority = 1u;
      }
      fell_back = !found;
      const f3 point = surface + hn * surface_offset(distance);
      // WGSL :2046-2059 `shading_normal`, toward the camera
      const float* nr = a.normals + 9 * (size_t)t;
      const f3 sn = ld3(nr) * b[0] + ld3(nr + 3) * b[1] + ld3(nr + 6) * b[2];
      f3 normal = norm3(normal_to_world(a, m, dot3(sn, sn) > 1e-20f ? sn : face));
      if (dot3(normal_to_world(a, m, face), -dir) < 0.0f) normal = -normal;
      for (uint32_t l = 0; l < a.n_lights; l++) {
        const float* lr = a.lights + 16 * (size_t)l;
        float e;
        if (found && !a.light_ratio) {
          e = vc_bilinear(a, slot, r, l);
        } else {
          const float share = found ? vc_bilinear(a, slot, r, l) : 1.0f;
          e = 0.0f;
          if (share > 0.0f) {
            // unshadowed_rect (WGSL :2135)
            float open = 0.0f;
            const LightRect lt = reinterpret_cast<const LightRect*>(a.lights)[l];
            if (faces_light(lt, point, normal)) {
              if (lr[3] > 1.5f) {
                open = spot_falloff(lt, norm3(point - ld3(lr))) * sphere_unshadowed(point, normal, lt);
              } else {
                // rect_open without its clipped case's workspace: the
                // horizon clip runs in the lane's registers
                float poly[POLY_SLOTS];
                float acc_dummy[1];
                Lane L{acc_dummy, poly, POLY_STRIDE, 0u, 0u};
                open = rect_open(L, point, normal, lt);
              }
            }
            e = share * open;
          }
        }
        s.out[(size_t)l * n_px + p] = e;
      }
      s.out_page[p] = found ? mip : (0x80000000u | mip);
    } else {
      for (uint32_t l = 0; l < a.n_lights; l++) s.out[(size_t)l * n_px + p] = 0.0f;
      s.out_page[p] = NONE;
    }
  }
  if (is_sample && priority > 0u) vc_request(s, req_key, priority);
  const uint64_t fb = ballot(fell_back);
  if (lane_id() == 0 && fb) atomicAdd(&s.counts[1], (uint32_t)__builtin_popcountll(fb));
}

// After the shading pass: each listed request takes its entry's priority.
__global__ __launch_bounds__(256) void vc_request_gather_kernel(ShadeArgs s) {
  const uint32_t n = blockIdx.x * 256u + threadIdx.x;
  const uint32_t count = s.counts[0] < s.request_cap ? s.counts[0] : s.request_cap;
  if (n >= count) return;
  uint32_t* r = s.requests + 4 * (size_t)n;
  r[3] = s.request_prio[r[3]];
}

// The border texels of each listed page (entry p of the bake list) for
// channels [ch0, ch1): copied from their nearest inside texel, as
// vc_fill_kernel copies every channel (one wave per page).
__global__ __launch_bounds__(64) void vc_border_kernel(VcArgs a, uint32_t n, uint32_t ch0, uint32_t ch1) {
  const uint32_t lane = lane_id();
  for (uint32_t p = blockIdx.x; p < n; p += gridDim.x) {
    const uint32_t* e = a.bake_list + 4 * (size_t)p;
    const uint32_t c = e[1] & vc::KEY_CHART_MASK, mip = e[1] >> vc::KEY_MIP_SHIFT;
    uint32_t g[2];
    vc_grid(a, c, e[0], mip, g);
    const uint32_t px = (e[2] & 0xffffu) * vc::PAGE, py = (e[2] >> 16) * vc::PAGE;
    float* o = a.pool + (size_t)e[3] * VC_TEXELS * a.channels;
    for (uint32_t k = lane; k < VC_TEXELS; k += 64u) {
      const uint32_t tx = px + k % VC_STRIDE, ty = py + k / VC_STRIDE;
      if (tx >= g[0] || ty >= g[1]) {
        const uint32_t ix = (tx < g[0] - 1u ? tx : g[0] - 1u) - px;
        const uint32_t iy = (ty < g[1] - 1u ? ty : g[1] - 1u) - py;
        const uint32_t nearest = iy * VC_STRIDE + ix;
        for (uint32_t ch = ch0; ch < ch1; ch++) o[k * a.channels + ch] = o[nearest * a.channels + ch];
      }
    }
  }
}

// `verify_cache`'s comparison (instanced.rs:3721-3729): the f32 bit
// patterns of each listed slot's texels in the pool and in a fresh bake,
// the differing ones counted (one wave per slot).
__global__ __launch_bounds__(64) void vc_compare_kernel(const float* __restrict__ pool,
                                                        const float* __restrict__ fresh,
                                                        const uint32_t* __restrict__ slots, uint32_t n,
                                                        uint32_t per_slot, unsigned long long* __restrict__ differ) {
  for (uint32_t p = blockIdx.x; p < n; p += gridDim.x) {
    const size_t base = (size_t)slots[p] * per_slot;
    uint32_t d = 0;
    for (uint32_t i = threadIdx.x; i < per_slot; i += 64u)
      d += __float_as_uint(pool[base + i]) != __float_as_uint(fresh[base + i]) ? 1u : 0u;
    for (int off = 32; off > 0; off >>= 1) d += (uint32_t)__shfl_xor((int)d, off, 64);
    if (threadIdx.x == 0 && d) atomicAdd(differ, (unsigned long long)d);
  }
}

// WGSL :3134-3160 (`vc_find`, a page shown unless marked), then the page's
// texels: one wave per key.
__global__ __launch_bounds__(64) void vc_read_kernel(VcArgs a, const uint32_t* __restrict__ keys, uint32_t n_keys,
                                                     float* __restrict__ out, uint32_t* __restrict__ out_slot) {
  const uint32_t lane = lane_id();
  const uint32_t i = blockIdx.x;
  if (i >= n_keys) return;
  const uint32_t k0 = keys[3 * (size_t)i], k1 = keys[3 * (size_t)i + 1], k2 = keys[3 * (size_t)i + 2];
  const uint32_t kk[3] = {k0, k1, k2};
  uint32_t h = (kk[0] * 0x9E3779B1u) ^ (kk[1] * 0x85EBCA77u) ^ (kk[2] * 0xC2B2AE3Du);
  h ^= h >> 15;
  h *= 0x2C1B3C6Du;
  h ^= h >> 12;
  uint32_t at = h & a.table_mask;
  uint32_t slot = NONE;
  for (uint32_t probe = 0; probe < vc::MAX_PROBE; probe++) {
    const uint32_t* e = a.page_table + 4 * (size_t)at;
    if (e[0] == vc::EMPTY) break;
    if (e[0] == k0 && e[1] == k1 && e[2] == k2) {
      const uint32_t stale = a.slot_stale[e[3]];
      slot = (stale != 0u && stale != 3u) ? NONE : e[3];
      break;
    }
    at = (at + 1u) & a.table_mask;
  }
  if (lane == 0) out_slot[i] = slot;
  const uint32_t row = VC_TEXELS * a.channels;
  float* dst = out + (size_t)i * row;
  for (uint32_t q = lane; q < row; q += 64u) dst[q] = slot == NONE ? 0.0f : a.pool[(size_t)slot * row + q];
}

}  // namespace cnq

// ---------------------------------------------------------------- host side

using namespace cnq::detail;
namespace vc = cnq::vc;
using cnq::NONE;

namespace {

struct Key {
  uint32_t k[3];
  bool operator==(const Key& o) const { return k[0] == o.k[0] && k[1] == o.k[1] && k[2] == o.k[2]; }
};
struct KeyHash {
  size_t operator()(const Key& a) const { return vc::page_hash(a.k) * 0x9E3779B97F4A7C15ull ^ a.k[2]; }
};

}  // namespace

struct cnq_vcache {
  cnq_scene* scene = nullptr;
  cnq_vcache_config cfg;
  cnq_options opts;
  uint32_t n_lights = 0, channels = 0;
  float inv_base = 1.0f;
  std::vector<float> lights;
  std::vector<vc::Box> light_boxes;
  vc::ChartTables tables;
  std::vector<float> scale;  // per object
  vc::PageTable table;
  std::unordered_map<Key, std::pair<uint32_t, uint32_t>, KeyHash> map;  // key -> (slot, entry)
  std::vector<Key> slot_key;
  std::vector<uint32_t> last_used, alloc_frame, free_slots;
  uint32_t slot_extent = 0;
  uint32_t frame = 1;
  // The next frame's plan.
  std::vector<uint32_t> plan_list, plan_jobs, used_marks, cleared;
  std::vector<float> plan_bounds;
  // Objects moved since the last frame.
  std::vector<float> moved;
  uint32_t n_movers = 0;
  std::vector<std::pair<uint32_t, uint32_t>> mover_ids;
  uint64_t last_stats[10] = {0};
  // Moving lights through the ledger (cnq_vcache_set_moving_lights): per
  // light its index among them (NONE: static), ε, whether one moved since
  // the last frame, whether the table changed on the host since its upload,
  // whether objects ever moved (the two do not mix), and the last frame's
  // ledger counters.
  std::vector<uint32_t> moving_index;
  uint32_t n_moving = 0;
  float epsilon = 0.01f;
  bool lights_moved = false, lights_dirty = false, had_movers = false;
  uint64_t last_ledger[10] = {0};
  DevBuf<uint32_t> d_moving_index, d_ledger;
  DevBuf<float> d_ledger_bound, d_poses;
  uint32_t shade_requests = 0;  // requests the last shading pass listed
  DevBuf<float> d_camera;
  std::vector<uint32_t> upd_host;  // staging for flush_updates (alive until the stream syncs)

  DevBuf<vc::Chart> d_charts;
  DevBuf<uint32_t> d_cells, d_cell_tris, d_page_table, d_slot_info, d_slot_stale, d_slot_used, d_bake_list,
      d_bake_jobs, d_candidates, d_invalid, d_stats, d_updates;
  // Shading from the cache: each triangle's chart, the request set and list.
  DevBuf<uint32_t> d_tri_chart, d_request_set, d_request_prio, d_requests, d_shade_counts;
  DevBuf<float> d_tri_uv, d_normals, d_lights, d_pool, d_bake_bounds, d_moved;
  uint32_t list_cap = 0, job_cap = 0;  // bake list / jobs capacity (plan + append)

  void release() {
    d_charts.release(), d_cells.release(), d_cell_tris.release(), d_page_table.release(), d_slot_info.release();
    d_slot_stale.release(), d_slot_used.release(), d_bake_list.release(), d_bake_jobs.release();
    d_candidates.release(), d_invalid.release(), d_stats.release(), d_updates.release(), d_tri_uv.release();
    d_normals.release(), d_lights.release(), d_pool.release(), d_bake_bounds.release(), d_moved.release();
    d_tri_chart.release(), d_request_set.release(), d_request_prio.release(), d_requests.release();
    d_shade_counts.release();
    d_camera.release();
    d_moving_index.release(), d_ledger.release(), d_ledger_bound.release(), d_poses.release();
  }
  uint32_t append_jobs() const { return cfg.append_pages * vc::PAGE_TEXELS * n_lights; }

  cnq::VcArgs args() {
    cnq::VcArgs a;
    a.charts = d_charts.p;
    a.cells = d_cells.p;
    a.cell_tris = d_cell_tris.p;
    a.tri_uv = d_tri_uv.p;
    a.normals = d_normals.p;
    a.tris = scene->tris.p;
    a.ranges = scene->ranges.p;
    a.xforms = scene->xforms.p;
    a.lights = d_lights.p;
    a.moved = d_moved.p;
    a.page_table = d_page_table.p;
    a.slot_info = d_slot_info.p;
    a.slot_stale = d_slot_stale.p;
    a.slot_used = d_slot_used.p;
    a.pool = d_pool.p;
    a.bake_list = d_bake_list.p;
    a.bake_bounds = d_bake_bounds.p;
    a.bake_jobs = d_bake_jobs.p;
    a.candidates = d_candidates.p;
    a.invalid_list = d_invalid.p;
    a.stats = d_stats.p;
    a.n_slots = slot_extent;
    a.n_lights = n_lights;
    a.channels = channels;
    a.n_movers = n_movers;
    a.plan_pages = (uint32_t)(plan_list.size() / 4);
    a.plan_jobs = (uint32_t)(plan_jobs.size() / 2);
    a.append_pages = cfg.append_pages;
    a.append_jobs = append_jobs();
    a.candidate_cap = cfg.candidates;
    a.invalid_cap = cfg.pages;
    a.table_mask = table.len() - 1;
    a.frame = frame;
    a.visible_frames = cfg.visible_frames;
    a.light_ratio = cfg.light_ratio;
    a.bake_group = 64;
    a.bake_resident = 1;
    a.inv_base = inv_base;
    a.moving_index = d_moving_index.p;
    a.ledger = d_ledger.p;
    a.ledger_bound = d_ledger_bound.p;
    a.poses = d_poses.p;
    a.n_moving = n_moving;
    a.ledger_limit = 0;
    a.ledger_record = 1;
    a.epsilon = epsilon;
    return a;
  }

  // Frees slot `slot` (its key leaves the table; its slot info is cleared
  // before the next frame's passes).
  void evict_slot(uint32_t slot) {
    const Key k = slot_key[slot];
    auto it = map.find(k);
    if (it == map.end()) return;
    const uint32_t entry = it->second.second;
    map.erase(it);
    auto& m = map;
    table.remove(entry, [&m](const uint32_t* moved_key, uint32_t i) {
      Key mk{{moved_key[0], moved_key[1], moved_key[2]}};
      m[mk].second = i;
    });
    slot_key[slot] = Key{{vc::EMPTY, vc::EMPTY, vc::EMPTY}};
    free_slots.push_back(slot);
    cleared.push_back(slot);
  }
};

namespace {

// The bake over `a`'s jobs on `st`: the work-counting and moving-light
// instantiations as the scene and the cache ask.
hipError_t launch_bake(cnq_vcache* c, const cnq::VcArgs& a, const cnq::QueryParams& p, uint32_t grid,
                       hipStream_t st) {
  cnq_scene* s = c->scene;
  const bool moving = c->n_moving > 0;
  if (s->count_work) {
    if (moving)
      hipLaunchKernelGGL((cnq::vc_bake_kernel<true, true>), dim3(grid), dim3(64), 0, st, a, s->blas.p, s->nodes.p,
                         s->tri_leaf.p, s->items.p, s->blas_items.p, s->blas_cov.p, p, s->work.p, s->counters.p);
    else
      hipLaunchKernelGGL((cnq::vc_bake_kernel<true, false>), dim3(grid), dim3(64), 0, st, a, s->blas.p, s->nodes.p,
                         s->tri_leaf.p, s->items.p, s->blas_items.p, s->blas_cov.p, p, s->work.p, s->counters.p);
  } else if (s->max_leaf <= 32) {
    if (moving)
      hipLaunchKernelGGL((cnq::vc_bake_kernel<false, true, true>), dim3(grid), dim3(64), 0, st, a, s->blas.p,
                         s->nodes.p, s->tri_leaf.p, s->items.p, s->blas_items.p, s->blas_cov.p, p, nullptr, nullptr);
    else
      hipLaunchKernelGGL((cnq::vc_bake_kernel<false, false, true>), dim3(grid), dim3(64), 0, st, a, s->blas.p,
                         s->nodes.p, s->tri_leaf.p, s->items.p, s->blas_items.p, s->blas_cov.p, p, nullptr, nullptr);
  } else {
    if (moving)
      hipLaunchKernelGGL((cnq::vc_bake_kernel<false, true>), dim3(grid), dim3(64), 0, st, a, s->blas.p, s->nodes.p,
                         s->tri_leaf.p, s->items.p, s->blas_items.p, s->blas_cov.p, p, nullptr, nullptr);
    else
      hipLaunchKernelGGL((cnq::vc_bake_kernel<false, false>), dim3(grid), dim3(64), 0, st, a, s->blas.p, s->nodes.p,
                         s->tri_leaf.p, s->items.p, s->blas_items.p, s->blas_cov.p, p, nullptr, nullptr);
  }
  return hipGetLastError();
}

// Sends the page-table entries changed since the last flush, the use marks
// and the cleared slots (`vc_update`), on `st`; the caller synchronizes.
int flush_updates(cnq_vcache* c, hipStream_t st) {
  const std::vector<uint32_t> dirty = c->table.take_dirty();
  std::vector<uint32_t>& upd = c->upd_host;
  upd.clear();
  for (uint32_t i : dirty) {
    const uint32_t* e = &c->table.entries[4 * (size_t)i];
    upd.insert(upd.end(), {i, e[0], e[1], e[2], e[3]});
  }
  upd.insert(upd.end(), c->used_marks.begin(), c->used_marks.end());
  upd.insert(upd.end(), c->cleared.begin(), c->cleared.end());
  const uint32_t n_table = (uint32_t)dirty.size(), n_used = (uint32_t)c->used_marks.size(),
                 n_cleared = (uint32_t)c->cleared.size();
  c->used_marks.clear();
  c->cleared.clear();
  const uint32_t n_upd = n_table + n_used + n_cleared;
  if (n_upd == 0) return CNQ_OK;
  if (upd.size() > c->d_updates.n) {
    c->d_updates.release();
    HIP_TRY(c->d_updates.alloc(upd.size() + 1024), "table updates");
  }
  HIP_TRY(hipMemcpyAsync(c->d_updates.p, upd.data(), upd.size() * 4, hipMemcpyHostToDevice, st), "table updates");
  cnq::VcArgs a = c->args();
  hipLaunchKernelGGL(cnq::vc_update_kernel, dim3((n_upd + 255) / 256), dim3(256), 0, st, a, c->d_updates.p, n_table,
                     n_used, n_cleared);
  HIP_TRY(hipGetLastError(), "vc_update launch");
  return CNQ_OK;
}

uint32_t next_pow2(uint64_t v) {
  uint64_t p = 1;
  while (p < v) p <<= 1;
  return (uint32_t)p;
}

}  // namespace

extern "C" {

void cnq_vcache_config_default(cnq_vcache_config* c) {
  if (!c) return;
  c->pages = 4096;
  c->append_pages = 256;
  c->candidates = 4096;
  c->base_texel = 0.05f;
  c->light_ratio = 1;
  c->visible_frames = 1;
}

int cnq_vcache_create(cnq_scene* s, const cnq_vcache_surfaces* sf, const float* lights, uint32_t n_lights,
                      const cnq_vcache_config* cfg, const cnq_options* opts, cnq_vcache** out) {
  if (!s || !sf || !lights || !cfg || !out) return fail(CNQ_INVALID_ARG, "null argument");
  *out = nullptr;
  int rc = check_options(opts);
  if (rc != CNQ_OK) return rc;
  if (n_lights == 0 || n_lights > vc::MAX_LIGHTS)
    return fail(CNQ_INVALID_ARG, "the view cache takes 1 to 16 lights (lib.rs:751 MAX_LIGHTS)");
  if (cfg->pages == 0 || cfg->pages > (1u << 24)) return fail(CNQ_INVALID_ARG, "pages must be 1 .. 2^24");
  if (!(cfg->base_texel > 0.0f)) return fail(CNQ_INVALID_ARG, "base_texel must be > 0");
  if (s->n_tris == 0) return fail(CNQ_INVALID_ARG, "the scene has no triangles");
  if (sf->n_charts == 0 || sf->n_charts >= (1u << 24))
    return fail(CNQ_INVALID_ARG, "1 .. 2^24 - 1 charts (page keys hold 24 bits)");
  if (!sf->tri_chart || !sf->tri_uv || !sf->normals || !sf->chart_size || !sf->chart_grid)
    return fail(CNQ_INVALID_ARG, "surface arrays are null");
  rc = check_object_cov(s, opts);
  if (rc != CNQ_OK) return rc;
  HIP_TRY(hipSetDevice(s->dev->ordinal), "hipSetDevice");
  auto c = new cnq_vcache();
  c->scene = s;
  c->cfg = *cfg;
  c->opts = *opts;
  c->n_lights = n_lights;
  c->channels = n_lights + 1;
  c->inv_base = 1.0f / cfg->base_texel;
  c->lights.assign(lights, lights + 16 * (size_t)n_lights);
  for (uint32_t l = 0; l < n_lights; l++) c->light_boxes.push_back(vc::light_box(lights + 16 * (size_t)l));
  // Triangles (object space) for the cells' boxes, from the device copy.
  std::vector<float> tris(12 * (size_t)s->n_tris);
  std::vector<float> xf(16 * (size_t)s->n_xforms);
  hipError_t e = hipMemcpy(tris.data(), s->tris.p, tris.size() * 4, hipMemcpyDeviceToHost);
  if (e == hipSuccess && !xf.empty()) e = hipMemcpy(xf.data(), s->xforms.p, xf.size() * 4, hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    delete c;
    return hip_fail(e, "scene copy");
  }
  if (!vc::chart_tables(tris.data(), s->n_tris, sf->tri_chart, sf->tri_uv, sf->chart_size, sf->chart_grid,
                        sf->n_charts, c->tables)) {
    delete c;
    return fail(CNQ_INVALID_ARG, "a triangle's chart is out of range, or a chart has an empty grid");
  }
  c->scale.resize(s->n_objects);
  for (uint32_t m = 0; m < s->n_objects; m++) {
    const uint32_t x = s->h_ranges[4 * m + 3];
    c->scale[m] = x == NONE ? 1.0f : xf[16 * (size_t)x + 12];
  }
  const uint32_t P = cfg->pages;
  c->table = vc::PageTable(std::max<uint32_t>(next_pow2(4ull * P), 1024u));
  c->slot_key.assign(P, Key{{vc::EMPTY, vc::EMPTY, vc::EMPTY}});
  c->last_used.assign(P, 0);
  c->alloc_frame.assign(P, 0);
  for (uint32_t k = P; k > 0; k--) c->free_slots.push_back(k - 1);
  c->list_cap = 64 + cfg->append_pages;
  c->job_cap = 64 * vc::PAGE_TEXELS * n_lights + c->append_jobs();
  const size_t pool = (size_t)P * vc::PAGE_TEXELS * c->channels;
  const uint32_t stats_zero[cnq::VC_STATS] = {0};
  e = c->d_charts.upload(c->tables.charts.data(), c->tables.charts.size());
  if (e == hipSuccess) e = c->d_cells.upload(c->tables.cells.data(), c->tables.cells.size());
  if (e == hipSuccess) e = c->d_cell_tris.upload(c->tables.cell_tris.data(), c->tables.cell_tris.size());
  if (e == hipSuccess) e = c->d_tri_uv.upload(sf->tri_uv, 6 * (size_t)s->n_tris);
  if (e == hipSuccess) e = c->d_normals.upload(sf->normals, 9 * (size_t)s->n_tris);
  if (e == hipSuccess) e = c->d_tri_chart.upload(sf->tri_chart, s->n_tris);
  if (e == hipSuccess) e = c->d_request_set.alloc(cnq::VC_REQUEST_SET);
  if (e == hipSuccess) e = c->d_request_prio.alloc(cnq::VC_REQUEST_SET);
  if (e == hipSuccess) e = c->d_requests.alloc(4 * (size_t)cnq::VC_REQUEST_CAP);
  if (e == hipSuccess) e = c->d_shade_counts.alloc(2);
  if (e == hipSuccess) e = c->d_lights.upload(lights, 16 * (size_t)n_lights);
  if (e == hipSuccess) e = c->d_page_table.upload(c->table.entries.data(), c->table.entries.size());
  if (e == hipSuccess) e = c->d_slot_info.alloc(12 * (size_t)P);
  if (e == hipSuccess) e = hipMemset(c->d_slot_info.p, 0, c->d_slot_info.bytes());
  if (e == hipSuccess) e = c->d_slot_stale.alloc(P);
  if (e == hipSuccess) e = hipMemset(c->d_slot_stale.p, 0, c->d_slot_stale.bytes());
  if (e == hipSuccess) e = c->d_slot_used.alloc(P);
  if (e == hipSuccess) e = hipMemset(c->d_slot_used.p, 0, c->d_slot_used.bytes());
  if (e == hipSuccess) e = c->d_pool.alloc(pool);
  if (e == hipSuccess) e = hipMemset(c->d_pool.p, 0, c->d_pool.bytes());
  if (e == hipSuccess) e = c->d_bake_list.alloc(4 * (size_t)c->list_cap);
  if (e == hipSuccess) e = hipMemset(c->d_bake_list.p, 0, c->d_bake_list.bytes());
  if (e == hipSuccess) e = c->d_bake_bounds.alloc(8 * (size_t)c->list_cap);
  if (e == hipSuccess) e = hipMemset(c->d_bake_bounds.p, 0, c->d_bake_bounds.bytes());
  if (e == hipSuccess) e = c->d_bake_jobs.alloc(2 * (size_t)c->job_cap);
  if (e == hipSuccess) e = hipMemset(c->d_bake_jobs.p, 0, c->d_bake_jobs.bytes());
  if (e == hipSuccess) e = c->d_candidates.alloc(2 * (size_t)std::max<uint32_t>(cfg->candidates, 1));
  if (e == hipSuccess) e = c->d_invalid.alloc(P);
  if (e == hipSuccess) e = c->d_stats.upload(stats_zero, cnq::VC_STATS);
  if (e == hipSuccess) e = c->d_moved.alloc(4 * (2 * (size_t)n_lights + 2));
  if (e == hipSuccess) e = hipMemset(c->d_moved.p, 0, c->d_moved.bytes());
  if (e != hipSuccess) {
    c->release();
    delete c;
    return hip_fail(e, "view cache allocation");
  }
  *out = c;
  return CNQ_OK;
}

void cnq_vcache_destroy(cnq_vcache* c) {
  if (!c) return;
  (void)hipSetDevice(c->scene->dev->ordinal);
  (void)hipDeviceSynchronize();
  c->release();
  delete c;
}

int cnq_vcache_grid(cnq_vcache* c, uint32_t object, uint32_t chart, uint32_t mip, uint32_t out[2]) {
  if (!c || !out) return fail(CNQ_INVALID_ARG, "null argument");
  if (object >= c->scene->n_objects || chart >= c->tables.charts.size() || mip > 30)
    return fail(CNQ_INVALID_ARG, "object, chart or mip out of range");
  vc::grid(c->tables.charts[chart], c->scale[object], c->inv_base, mip, out);
  return CNQ_OK;
}

int cnq_vcache_request(cnq_vcache* c, const uint32_t* keys, uint32_t n, uint32_t* slots) {
  if (!c || (n && (!keys || !slots))) return fail(CNQ_INVALID_ARG, "null argument");
  for (uint32_t i = 0; i < n; i++) {
    const uint32_t* k = keys + 3 * (size_t)i;
    const uint32_t m = k[0], chart = k[1] & vc::KEY_CHART_MASK, mip = k[1] >> vc::KEY_MIP_SHIFT;
    if (m >= c->scene->n_objects || chart >= c->tables.charts.size() || mip > 30)
      return fail(CNQ_INVALID_ARG, "key " + std::to_string(i) + ": object, chart or mip out of range");
  }
  for (uint32_t i = 0; i < n; i++) {
    const uint32_t* k = keys + 3 * (size_t)i;
    const Key key{{k[0], k[1], k[2]}};
    slots[i] = NONE;
    auto it = c->map.find(key);
    if (it != c->map.end()) {
      const uint32_t slot = it->second.first;
      if (c->last_used[slot] != c->frame) c->used_marks.push_back(slot);
      c->last_used[slot] = c->frame;
      slots[i] = slot;
      continue;
    }
    const uint32_t m = k[0], chart = k[1] & vc::KEY_CHART_MASK, mip = k[1] >> vc::KEY_MIP_SHIFT;
    const uint32_t px = k[2] & 0xffffu, py = k[2] >> 16;
    // vcache.rs:3943-3957: the grid one larger (the shaders' f32 grid may
    // be), the texels inside it and one past them.
    uint32_t g[2];
    vc::grid(c->tables.charts[chart], c->scale[m], c->inv_base, mip, g);
    uint32_t listed[2];
    bool empty = false;
    const uint32_t pp[2] = {px, py};
    for (int a = 0; a < 2; a++) {
      const uint32_t nn = g[a] + 1;
      const uint32_t start = std::min(pp[a] * vc::PAGE, nn);
      const uint32_t inside = std::min(nn - start, vc::STRIDE);
      // The reference's slack of one texel covers a shader grid one larger
      // than the host's; here both are the same f32 products, so a page
      // starting past the grid holds no texel and is refused.
      if (inside == 0 || (uint64_t)pp[a] * vc::PAGE >= g[a]) empty = true;
      listed[a] = std::min(inside + 1, vc::STRIDE);
    }
    if (empty || c->free_slots.empty()) continue;
    const uint32_t slot = c->free_slots.back();
    c->free_slots.pop_back();
    const uint32_t entry = c->table.insert(key.k, slot);
    if (entry == vc::EMPTY) {
      c->free_slots.push_back(slot);
      continue;
    }
    c->slot_extent = std::max(c->slot_extent, slot + 1);
    c->map[key] = {slot, entry};
    c->slot_key[slot] = key;
    c->last_used[slot] = c->frame;
    c->alloc_frame[slot] = c->frame;
    const uint32_t index = (uint32_t)(c->plan_list.size() / 4);
    c->plan_list.insert(c->plan_list.end(), {k[0], k[1], k[2], slot});
    float b[8];
    vc::page_bounds(c->tables, chart, c->scale[m], c->inv_base, mip, px, py, b);
    b[3] = 0.0f;
    b[7] = 0.0f;  // kind 0: lights baked, no sky or bounce samples
    c->plan_bounds.insert(c->plan_bounds.end(), b, b + 8);
    for (uint32_t l = 0; l < c->n_lights; l++)
      for (uint32_t y = 0; y < listed[1]; y++)
        for (uint32_t x = 0; x < listed[0]; x++)
          c->plan_jobs.insert(c->plan_jobs.end(), {index, (y * vc::STRIDE + x) | (l << 8)});
    slots[i] = slot;
  }
  return CNQ_OK;
}

int cnq_vcache_evict(cnq_vcache* c, const uint32_t* keys, uint32_t n) {
  if (!c || (n && !keys)) return fail(CNQ_INVALID_ARG, "null argument");
  for (uint32_t i = 0; i < n; i++) {
    const uint32_t* k = keys + 3 * (size_t)i;
    auto it = c->map.find(Key{{k[0], k[1], k[2]}});
    if (it == c->map.end()) continue;
    const uint32_t slot = it->second.first;
    if (c->alloc_frame[slot] == c->frame) {
      // Planned this frame and not baked yet: take it out of the plan.
      return fail(CNQ_INVALID_ARG, "a page planned this frame is evicted only after its frame");
    }
    c->evict_slot(slot);
  }
  // The device's table follows at once, so reads no longer find the pages.
  HIP_TRY(hipSetDevice(c->scene->dev->ordinal), "hipSetDevice");
  const int rc = flush_updates(c, nullptr);
  if (rc != CNQ_OK) return rc;
  HIP_TRY(hipStreamSynchronize(nullptr), "view cache eviction");
  return CNQ_OK;
}

int cnq_vcache_set_moved(cnq_vcache* c, const cnq_moved* moved, uint32_t n, const float* scene_box) {
  if (!c || (n && !moved)) return fail(CNQ_INVALID_ARG, "null argument");
  if (c->n_moving && n)
    return fail(CNQ_INVALID_ARG, "moving lights and moving objects together are not supported (vcache.rs:1410)");
  if (n) c->had_movers = true;
  std::vector<vc::Moved> list(n);
  for (uint32_t i = 0; i < n; i++) {
    if (moved[i].object >= c->scene->n_objects) return fail(CNQ_INVALID_ARG, "moved object out of range");
    list[i].object = moved[i].object;
    std::memcpy(list[i].old_box.lo, moved[i].old_lo, 12);
    std::memcpy(list[i].old_box.hi, moved[i].old_hi, 12);
    std::memcpy(list[i].new_box.lo, moved[i].new_lo, 12);
    std::memcpy(list[i].new_box.hi, moved[i].new_hi, 12);
  }
  vc::Box sb;
  if (scene_box) {
    std::memcpy(sb.lo, scene_box, 12);
    std::memcpy(sb.hi, scene_box + 3, 12);
  } else {
    // instanced.rs:1395-1406: the object hierarchy's root box with the
    // movers' old boxes, padded.
    cnq::NodeC root;
    HIP_TRY(hipSetDevice(c->scene->dev->ordinal), "hipSetDevice");
    HIP_TRY(hipMemcpy(&root, c->scene->nodes.p, sizeof(root), hipMemcpyDeviceToHost), "root box");
    float lo[3], hi[3];
    for (int a = 0; a < 3; a++) lo[a] = root.lo[a], hi[a] = root.hi[a];
    for (const vc::Moved& m : list)
      for (int a = 0; a < 3; a++) {
        lo[a] = std::min(lo[a], m.old_box.lo[a]);
        hi[a] = std::max(hi[a], m.old_box.hi[a]);
      }
    float e = 1.0f;
    for (int a = 0; a < 3; a++) e = std::max(e, std::max(std::fabs(lo[a]), std::fabs(hi[a])));
    const float pad = 1e-3f * e;
    for (int a = 0; a < 3; a++) sb.lo[a] = lo[a] - pad, sb.hi[a] = hi[a] + pad;
  }
  c->moved = vc::pack_moved(list, c->light_boxes, sb, c->mover_ids);
  c->n_movers = n;
  return CNQ_OK;
}

int cnq_vcache_frame(cnq_vcache* c, void* stream, uint64_t stats[10]) {
  if (!c) return fail(CNQ_INVALID_ARG, "null argument");
  cnq_scene* s = c->scene;
  HIP_TRY(hipSetDevice(s->dev->ordinal), "hipSetDevice");
  hipStream_t st = (hipStream_t)stream;
  const uint32_t plan_pages = (uint32_t)(c->plan_list.size() / 4), plan_jobs = (uint32_t)(c->plan_jobs.size() / 2);
  // Capacities for this frame's plan and appends.
  if (plan_pages + c->cfg.append_pages > c->list_cap) {
    const uint32_t cap = plan_pages + c->cfg.append_pages + 64;
    c->d_bake_list.release(), c->d_bake_bounds.release();
    HIP_TRY(c->d_bake_list.alloc(4 * (size_t)cap), "bake list");
    HIP_TRY(c->d_bake_bounds.alloc(8 * (size_t)cap), "bake bounds");
    HIP_TRY(hipMemset(c->d_bake_list.p, 0, c->d_bake_list.bytes()), "bake list");
    c->list_cap = cap;
  }
  if ((uint64_t)plan_jobs + c->append_jobs() > c->job_cap) {
    const uint64_t cap = (uint64_t)plan_jobs + c->append_jobs() + 64;
    if (cap > 0xffffffffull) return fail(CNQ_CAPACITY, "too many bake jobs in one frame");
    c->d_bake_jobs.release();
    HIP_TRY(c->d_bake_jobs.alloc(2 * cap), "bake jobs");
    HIP_TRY(hipMemset(c->d_bake_jobs.p, 0, c->d_bake_jobs.bytes()), "bake jobs");
    c->job_cap = (uint32_t)cap;
  }
  const size_t moved_floats = std::max<size_t>(c->moved.size(), 4 * (2 * (size_t)c->n_lights + 2));
  if (moved_floats > c->d_moved.n) {
    c->d_moved.release();
    HIP_TRY(c->d_moved.alloc(moved_floats), "moved boxes");
  }
  if (plan_pages) {
    HIP_TRY(hipMemcpyAsync(c->d_bake_list.p, c->plan_list.data(), c->plan_list.size() * 4, hipMemcpyHostToDevice, st),
            "bake list");
    HIP_TRY(hipMemcpyAsync(c->d_bake_bounds.p, c->plan_bounds.data(), c->plan_bounds.size() * 4,
                           hipMemcpyHostToDevice, st), "bake bounds");
  }
  if (plan_jobs)
    HIP_TRY(hipMemcpyAsync(c->d_bake_jobs.p, c->plan_jobs.data(), c->plan_jobs.size() * 4, hipMemcpyHostToDevice, st),
            "bake jobs");
  if (!c->moved.empty())
    HIP_TRY(hipMemcpyAsync(c->d_moved.p, c->moved.data(), c->moved.size() * 4, hipMemcpyHostToDevice, st), "moved");
  HIP_TRY(fill_u32(c->d_stats.p, 0, cnq::VC_STATS, st), "stats clear");
  // Page-table entries changed since the last frame, use marks, cleared slots.
  int rc = flush_updates(c, st);
  if (rc != CNQ_OK) return rc;
  cnq::VcArgs a = c->args();
  if (c->moved.empty()) a.n_movers = 0;
  if (c->n_moving) {
    // The lights as moved since the last frame, this frame's poses
    // (vcache.rs:2261-2266), and fresh ledger records for the planned pages
    // (WGSL :4655-4662).
    if (c->lights_dirty) {
      HIP_TRY(hipMemcpyAsync(c->d_lights.p, c->lights.data(), 64 * (size_t)c->n_lights, hipMemcpyHostToDevice, st),
              "lights");
      c->lights_dirty = false;
    }
    HIP_TRY(hipMemcpyAsync(c->d_poses.p + (size_t)(c->frame % cnq::VC_POSE_RING) * c->n_lights * 16,
                           c->lights.data(), 64 * (size_t)c->n_lights, hipMemcpyHostToDevice, st),
            "light poses");
    if (plan_pages) {
      hipLaunchKernelGGL(cnq::vc_ledger_reset_kernel, dim3((plan_pages + 63) / 64), dim3(64), 0, st, a);
      HIP_TRY(hipGetLastError(), "vc_ledger_reset launch");
    }
  }
  if (a.n_movers > 0 && c->slot_extent > 0) {
    hipLaunchKernelGGL(cnq::vc_stale_kernel, dim3((c->slot_extent + 255) / 256), dim3(256), 0, st, a);
    HIP_TRY(hipGetLastError(), "vc_stale launch");
    const uint64_t threads = (uint64_t)c->cfg.candidates * vc::PAGE_TEXELS;
    if (threads) {
      hipLaunchKernelGGL(cnq::vc_stale_texels_kernel, dim3((uint32_t)((threads + 255) / 256)), dim3(256), 0, st, a);
      HIP_TRY(hipGetLastError(), "vc_stale_texels launch");
    }
  }
  // The ledger runs every frame with moving lights (instanced.rs:176-185):
  // it re-bakes what a move took past ε, and pages set aside once wanted.
  if (c->n_moving && c->slot_extent > 0) {
    a.ledger_limit = std::min<uint32_t>(a.append_jobs, 0x7fffffffu) | (c->lights_moved ? 0x80000000u : 0u);
    hipLaunchKernelGGL(cnq::vc_ledger_kernel, dim3((c->slot_extent + 255) / 256), dim3(256), 0, st, a);
    HIP_TRY(hipGetLastError(), "vc_ledger launch");
  }
  const bool appends = a.n_movers > 0 || c->n_moving > 0;
  const cnq::QueryParams p = query_params(s, &c->opts, 64);
  // As many waves as the GPU holds at once (4 SIMDs a compute unit,
  // WALK_WAVES each), or fewer for a small bake; they take jobs from a
  // counter, fewer at a time near the end (vc_bake_kernel).
  const uint64_t jobs = (uint64_t)plan_jobs + (appends ? a.append_jobs : 0);
  const uint32_t resident = s->dev->compute_units * 4 * (uint32_t)cnq::WALK_WAVES;
  a.bake_group = 64;
  a.bake_resident = resident;
  if (jobs) {
    const uint32_t grid = (uint32_t)std::min<uint64_t>((jobs + 3) / 4, resident);
    HIP_TRY(launch_bake(c, a, p, grid, st), "vc_bake launch");
  }
  const uint32_t page_blocks = std::min<uint32_t>(plan_pages + (appends ? c->cfg.append_pages : 0), 4096);
  if (page_blocks) {
    hipLaunchKernelGGL(cnq::vc_fill_kernel, dim3(page_blocks), dim3(64), 0, st, a);
    HIP_TRY(hipGetLastError(), "vc_fill launch");
  }
  uint32_t dstats[cnq::VC_STATS];
  HIP_TRY(hipMemcpyAsync(dstats, c->d_stats.p, sizeof(dstats), hipMemcpyDeviceToHost, st), "stats");
  HIP_TRY(hipStreamSynchronize(st), "view cache frame");
  // Pages marked invalid leave the cache (vcache.rs:3425-3440); past the
  // list's capacity, every marked slot does (:3444-3452).
  std::vector<uint32_t> invalid;
  uint64_t evicted = 0;
  if (dstats[1] > 0) {
    if (dstats[1] <= c->cfg.pages) {
      invalid.resize(dstats[1]);
      HIP_TRY(hipMemcpy(invalid.data(), c->d_invalid.p, invalid.size() * 4, hipMemcpyDeviceToHost), "invalid list");
    } else {
      std::vector<uint32_t> marks(c->slot_extent);
      HIP_TRY(hipMemcpy(marks.data(), c->d_slot_stale.p, marks.size() * 4, hipMemcpyDeviceToHost), "stale marks");
      for (uint32_t sl = 0; sl < c->slot_extent; sl++)
        if (marks[sl] != 0) invalid.push_back(sl);
    }
  }
  // In slot order (the list's order is the atomics'), so the free list, and
  // the slots the next pages take, do not depend on it.
  std::sort(invalid.begin(), invalid.end());
  for (uint32_t slot : invalid) {
    if (c->slot_key[slot].k[0] == vc::EMPTY || c->alloc_frame[slot] >= c->frame) continue;
    c->evict_slot(slot);
    evicted++;
  }
  // The device's table follows the evictions at once (the reference sends
  // them with the next frame; reads between frames see the same pages).
  rc = flush_updates(c, st);
  if (rc != CNQ_OK) return rc;
  HIP_TRY(hipStreamSynchronize(st), "view cache evictions");
  const uint64_t st_out[10] = {plan_pages,
                               plan_jobs,
                               std::min(dstats[0], c->cfg.append_pages),
                               std::min<uint64_t>(dstats[5], c->append_jobs()),
                               dstats[4],
                               dstats[1],
                               dstats[6],
                               dstats[3],
                               dstats[2],
                               evicted};
  std::memcpy(c->last_stats, st_out, sizeof(st_out));
  if (stats) std::memcpy(stats, st_out, sizeof(st_out));
  const uint32_t ledger_words[10] = {8, 9, 10, 11, 12, 13, 14, 21, 22, 23};
  for (int i = 0; i < 10; i++) c->last_ledger[i] = dstats[ledger_words[i]];
  c->lights_moved = false;
  c->plan_list.clear(), c->plan_bounds.clear(), c->plan_jobs.clear(), c->used_marks.clear();
  c->moved.clear();
  c->n_movers = 0;
  c->frame++;
  return CNQ_OK;
}

int cnq_vcache_read(cnq_vcache* c, const uint32_t* keys, uint32_t n, float* texels, uint32_t* slots) {
  if (!c || (n && (!keys || !texels || !slots))) return fail(CNQ_INVALID_ARG, "null argument");
  if (n == 0) return CNQ_OK;
  HIP_TRY(hipSetDevice(c->scene->dev->ordinal), "hipSetDevice");
  DevBuf<uint32_t> dk, ds;
  DevBuf<float> dt;
  const size_t row = (size_t)vc::PAGE_TEXELS * c->channels;
  hipError_t e = dk.upload(keys, 3 * (size_t)n);
  if (e == hipSuccess) e = ds.alloc(n);
  if (e == hipSuccess) e = dt.alloc(row * n);
  if (e == hipSuccess) {
    cnq::VcArgs a = c->args();
    hipLaunchKernelGGL(cnq::vc_read_kernel, dim3(n), dim3(64), 0, 0, a, dk.p, n, dt.p, ds.p);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(texels, dt.p, row * n * 4, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(slots, ds.p, 4 * (size_t)n, hipMemcpyDeviceToHost);
  dk.release(), ds.release(), dt.release();
  if (e != hipSuccess) return hip_fail(e, "view cache read");
  return CNQ_OK;
}

namespace {

// Bake lists for `pages` (key, slot) of the cache: per page its entry and
// box, and a job per listed texel (inside the grid and one past) for each
// light in [l0, l1), as cnq_vcache_request lists a new page's.
struct BakeLists {
  std::vector<uint32_t> list, jobs, slots;
  std::vector<float> bounds;
};

void list_pages(const cnq_vcache* c, const std::vector<std::pair<Key, uint32_t>>& pages, uint32_t l0, uint32_t l1,
                BakeLists& b) {
  for (const auto& pg : pages) {
    const uint32_t* k = pg.first.k;
    const uint32_t slot = pg.second;
    const uint32_t m = k[0], chart = k[1] & vc::KEY_CHART_MASK, mip = k[1] >> vc::KEY_MIP_SHIFT;
    const uint32_t pp[2] = {k[2] & 0xffffu, k[2] >> 16};
    uint32_t g[2], listed[2];
    vc::grid(c->tables.charts[chart], c->scale[m], c->inv_base, mip, g);
    for (int ax = 0; ax < 2; ax++) {
      const uint32_t nn = g[ax] + 1;
      const uint32_t inside = std::min(nn - std::min(pp[ax] * vc::PAGE, nn), vc::STRIDE);
      listed[ax] = std::min(inside + 1, vc::STRIDE);
    }
    const uint32_t index = (uint32_t)(b.list.size() / 4);
    b.list.insert(b.list.end(), {k[0], k[1], k[2], slot});
    float bb[8];
    vc::page_bounds(c->tables, chart, c->scale[m], c->inv_base, mip, pp[0], pp[1], bb);
    bb[3] = 0.0f;
    bb[7] = 0.0f;
    b.bounds.insert(b.bounds.end(), bb, bb + 8);
    for (uint32_t l = l0; l < l1; l++)
      for (uint32_t y = 0; y < listed[1]; y++)
        for (uint32_t x = 0; x < listed[0]; x++) b.jobs.insert(b.jobs.end(), {index, (y * vc::STRIDE + x) | (l << 8)});
    b.slots.push_back(slot);
  }
}

// The bake of `b` into `pool` on `st` (device copies of the lists kept in
// the DevBufs given), then the border texels of channels [ch0, ch1) when
// ch0 < ch1.
int bake_lists(cnq_vcache* c, const BakeLists& b, float* pool, hipStream_t st, uint32_t ch0, uint32_t ch1,
               DevBuf<float>& dbounds, DevBuf<uint32_t>& dlist, DevBuf<uint32_t>& djobs, DevBuf<uint32_t>& dstats,
               bool record = true) {
  cnq_scene* s = c->scene;
  const uint32_t zeros[cnq::VC_STATS] = {0};
  hipError_t e = dbounds.upload(b.bounds.data(), b.bounds.size());
  if (e == hipSuccess) e = dlist.upload(b.list.data(), b.list.size());
  if (e == hipSuccess) e = djobs.upload(b.jobs.data(), b.jobs.size());
  if (e == hipSuccess) e = dstats.upload(zeros, cnq::VC_STATS);
  if (e != hipSuccess) return hip_fail(e, "bake lists");
  cnq::VcArgs a = c->args();
  a.pool = pool;
  a.bake_list = dlist.p;
  a.bake_bounds = dbounds.p;
  a.bake_jobs = djobs.p;
  a.stats = dstats.p;
  a.n_movers = 0;
  a.plan_pages = (uint32_t)b.slots.size();
  a.plan_jobs = (uint32_t)(b.jobs.size() / 2);
  a.append_pages = 0;
  a.append_jobs = 0;
  a.ledger_record = record ? 1u : 0u;
  const uint32_t resident = s->dev->compute_units * 4 * (uint32_t)cnq::WALK_WAVES;
  a.bake_group = 64;
  a.bake_resident = resident;
  const cnq::QueryParams p = query_params(s, &c->opts, 64);
  if (a.plan_jobs) {
    const uint32_t grid = (uint32_t)std::min<uint64_t>((a.plan_jobs + 3) / 4, resident);
    HIP_TRY(launch_bake(c, a, p, grid, st), "bake launch");
  }
  if (ch0 < ch1 && a.plan_pages) {
    hipLaunchKernelGGL(cnq::vc_border_kernel, dim3(std::min<uint32_t>(a.plan_pages, 4096)), dim3(64), 0, st, a,
                       a.plan_pages, ch0, ch1);
    HIP_TRY(hipGetLastError(), "border launch");
  }
  return CNQ_OK;
}

}  // namespace

int cnq_vcache_set_lights(cnq_vcache* c, const float* lights, uint32_t first, uint32_t n, void* stream,
                          uint64_t stats[2]) {
  if (!c || (n && !lights)) return fail(CNQ_INVALID_ARG, "null argument");
  if ((uint64_t)first + n > c->n_lights) return fail(CNQ_INVALID_ARG, "lights past the cache's light count");
  for (uint32_t l = first; c->n_moving && l < first + n; l++)
    if (c->moving_index[l] != cnq::NONE) return fail(CNQ_INVALID_ARG, "a moving light moves by cnq_vcache_move_light");
  if (stats) stats[0] = stats[1] = 0;
  if (n == 0) return CNQ_OK;
  HIP_TRY(hipSetDevice(c->scene->dev->ordinal), "hipSetDevice");
  hipStream_t st = (hipStream_t)stream;
  HIP_TRY(hipStreamSynchronize(st), "light edit");  // no pass on the stream still reads the old lights
  std::copy(lights, lights + 16 * (size_t)n, c->lights.begin() + 16 * (size_t)first);
  for (uint32_t l = first; l < first + n; l++) c->light_boxes[l] = vc::light_box(&c->lights[16 * (size_t)l]);
  HIP_TRY(hipMemcpy(c->d_lights.p + 16 * (size_t)first, lights, 64 * (size_t)n, hipMemcpyHostToDevice), "lights");
  // Every resident page's channels of these lights, baked afresh.
  std::vector<std::pair<Key, uint32_t>> pages;
  for (const auto& kv : c->map) pages.push_back({kv.first, kv.second.first});
  BakeLists b;
  list_pages(c, pages, first, first + n, b);
  DevBuf<float> dbounds;
  DevBuf<uint32_t> dlist, djobs, dstats;
  int rc = bake_lists(c, b, c->d_pool.p, st, first, first + n, dbounds, dlist, djobs, dstats);
  hipError_t e = hipStreamSynchronize(st);
  dbounds.release(), dlist.release(), djobs.release(), dstats.release();
  if (rc != CNQ_OK) return rc;
  if (e != hipSuccess) return hip_fail(e, "light edit");
  if (stats) stats[0] = b.slots.size(), stats[1] = b.jobs.size() / 2;
  return CNQ_OK;
}

int cnq_vcache_verify(cnq_vcache* c, void* stream, uint64_t out[2]) {
  if (!c || !out) return fail(CNQ_INVALID_ARG, "null argument");
  if (!c->plan_list.empty()) return fail(CNQ_INVALID_ARG, "pages are planned but not baked: run cnq_vcache_frame first");
  cnq_scene* s = c->scene;
  HIP_TRY(hipSetDevice(s->dev->ordinal), "hipSetDevice");
  hipStream_t st = (hipStream_t)stream;
  out[0] = out[1] = 0;
  // The pages shown: resident and not marked stale or invalid.
  std::vector<uint32_t> marks(c->slot_extent);
  HIP_TRY(hipStreamSynchronize(st), "verify");
  if (c->slot_extent)
    HIP_TRY(hipMemcpy(marks.data(), c->d_slot_stale.p, 4 * (size_t)c->slot_extent, hipMemcpyDeviceToHost), "marks");
  std::vector<std::pair<Key, uint32_t>> pages;
  for (const auto& kv : c->map) {
    const uint32_t slot = kv.second.first;
    if (slot < c->slot_extent && marks[slot] == 0) pages.push_back({kv.first, slot});
  }
  if (pages.empty()) return CNQ_OK;
  BakeLists b;
  list_pages(c, pages, 0, c->n_lights, b);
  if (b.jobs.size() / 2 > 0xffffffffull) return fail(CNQ_CAPACITY, "too many texels to verify at once");
  const uint32_t n = (uint32_t)b.slots.size(), per_slot = vc::PAGE_TEXELS * c->channels;
  DevBuf<float> fresh, dbounds;
  DevBuf<uint32_t> dlist, djobs, dslots, dstats;
  DevBuf<unsigned long long> differ;
  auto cleanup = [&]() {
    fresh.release(), dbounds.release(), dlist.release(), djobs.release(), dslots.release(), dstats.release();
    differ.release();
  };
  const unsigned long long zero = 0;
  hipError_t e = fresh.alloc(c->d_pool.n);
  if (e == hipSuccess) e = hipMemcpyAsync(fresh.p, c->d_pool.p, c->d_pool.bytes(), hipMemcpyDeviceToDevice, st);
  if (e == hipSuccess) e = dslots.upload(b.slots.data(), b.slots.size());
  if (e == hipSuccess) e = differ.upload(&zero, 1);
  if (e != hipSuccess) {
    cleanup();
    return hip_fail(e, "verify buffers");
  }
  int rc = bake_lists(c, b, fresh.p, st, 0, 0, dbounds, dlist, djobs, dstats, false);
  if (rc != CNQ_OK) {
    cleanup();
    return rc;
  }
  hipLaunchKernelGGL(cnq::vc_compare_kernel, dim3(std::min<uint32_t>(n, 4096)), dim3(64), 0, st, c->d_pool.p, fresh.p,
                     dslots.p, n, per_slot, differ.p);
  e = hipGetLastError();
  unsigned long long d = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(&d, differ.p, sizeof(d), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  cleanup();
  if (e != hipSuccess) return hip_fail(e, "verify");
  out[0] = d;
  out[1] = (uint64_t)n * per_slot;
  return CNQ_OK;
}

int cnq_vcache_set_moving_lights(cnq_vcache* c, const uint32_t* lights, uint32_t n, float epsilon) {
  if (!c || (n && !lights)) return fail(CNQ_INVALID_ARG, "null argument");
  if (c->frame != 1 || c->n_moving) return fail(CNQ_INVALID_ARG, "moving lights are set once, before the first frame");
  if (c->had_movers || c->n_movers)
    return fail(CNQ_INVALID_ARG, "moving lights and moving objects together are not supported (vcache.rs:1410)");
  if (!(epsilon >= 0.0f && epsilon < 0.5f)) return fail(CNQ_INVALID_ARG, "epsilon in [0, 0.5)");
  if (!c->cfg.light_ratio) return fail(CNQ_INVALID_ARG, "moving lights need light_ratio (cached ratios)");
  if (n == 0) return CNQ_OK;
  std::vector<uint32_t> index(c->n_lights, cnq::NONE);
  for (uint32_t j = 0; j < n; j++) {
    if (lights[j] >= c->n_lights) return fail(CNQ_INVALID_ARG, "moving light past the light table");
    if (index[lights[j]] != cnq::NONE) return fail(CNQ_INVALID_ARG, "a moving light listed twice");
    index[lights[j]] = j;
  }
  HIP_TRY(hipSetDevice(c->scene->dev->ordinal), "hipSetDevice");
  const size_t records = (size_t)c->cfg.pages * n;
  hipError_t e = c->d_moving_index.upload(index.data(), index.size());
  if (e == hipSuccess) e = c->d_ledger.alloc(records * cnq::LEDGER_WORDS_HOST);
  if (e == hipSuccess) e = hipMemset(c->d_ledger.p, 0, c->d_ledger.bytes());
  if (e == hipSuccess) e = c->d_ledger_bound.alloc(records);
  if (e == hipSuccess) e = hipMemset(c->d_ledger_bound.p, 0, c->d_ledger_bound.bytes());
  if (e == hipSuccess) e = c->d_poses.alloc((size_t)cnq::VC_POSE_RING * c->n_lights * 16);
  if (e == hipSuccess) e = hipMemset(c->d_poses.p, 0, c->d_poses.bytes());
  if (e != hipSuccess) {
    c->d_moving_index.release(), c->d_ledger.release(), c->d_ledger_bound.release(), c->d_poses.release();
    return hip_fail(e, "ledger buffers");
  }
  c->moving_index = index;
  c->n_moving = n;
  c->epsilon = epsilon;
  return CNQ_OK;
}

int cnq_vcache_move_light(cnq_vcache* c, uint32_t index, const float* light) {
  if (!c || !light) return fail(CNQ_INVALID_ARG, "null argument");
  if (index >= c->n_lights || !c->n_moving || c->moving_index[index] == cnq::NONE)
    return fail(CNQ_INVALID_ARG, "not a moving light (cnq_vcache_set_moving_lights)");
  if ((light[3] > 1.5f) != (c->lights[16 * (size_t)index + 3] > 1.5f))
    return fail(CNQ_INVALID_ARG, "a moving light keeps its kind (rect or sphere)");
  std::copy(light, light + 16, c->lights.begin() + 16 * (size_t)index);
  c->light_boxes[index] = vc::light_box(&c->lights[16 * (size_t)index]);
  c->lights_moved = true;
  c->lights_dirty = true;
  return CNQ_OK;
}

int cnq_vcache_ledger_stats(cnq_vcache* c, uint64_t out[10]) {
  if (!c || !out) return fail(CNQ_INVALID_ARG, "null argument");
  std::memcpy(out, c->last_ledger, sizeof(c->last_ledger));
  return CNQ_OK;
}

int cnq_vcache_ledger_state(cnq_vcache* c, uint32_t* records, float* bounds) {
  if (!c) return fail(CNQ_INVALID_ARG, "null argument");
  if (!c->n_moving) return fail(CNQ_INVALID_ARG, "no moving lights");
  HIP_TRY(hipSetDevice(c->scene->dev->ordinal), "hipSetDevice");
  if (records) HIP_TRY(hipMemcpy(records, c->d_ledger.p, c->d_ledger.bytes(), hipMemcpyDeviceToHost), "ledger");
  if (bounds)
    HIP_TRY(hipMemcpy(bounds, c->d_ledger_bound.p, c->d_ledger_bound.bytes(), hipMemcpyDeviceToHost), "bounds");
  return CNQ_OK;
}

int cnq_vcache_verify_ledger(cnq_vcache* c, uint32_t light, void* stream, float out[6]) {
  if (!c || !out) return fail(CNQ_INVALID_ARG, "null argument");
  for (int i = 0; i < 6; i++) out[i] = 0.0f;
  if (light >= c->n_lights || !c->n_moving || c->moving_index[light] == cnq::NONE)
    return fail(CNQ_INVALID_ARG, "not a moving light");
  if (!c->plan_list.empty()) return fail(CNQ_INVALID_ARG, "pages are planned but not baked: run cnq_vcache_frame first");
  cnq_scene* s = c->scene;
  HIP_TRY(hipSetDevice(s->dev->ordinal), "hipSetDevice");
  hipStream_t st = (hipStream_t)stream;
  HIP_TRY(hipStreamSynchronize(st), "verify ledger");
  const uint32_t j = c->moving_index[light];
  std::vector<uint32_t> marks(c->slot_extent);
  if (c->slot_extent)
    HIP_TRY(hipMemcpy(marks.data(), c->d_slot_stale.p, 4 * (size_t)c->slot_extent, hipMemcpyDeviceToHost), "marks");
  std::vector<std::pair<Key, uint32_t>> pages;
  for (const auto& kv : c->map) {
    const uint32_t slot = kv.second.first;
    if (slot < c->slot_extent && marks[slot] == 0) pages.push_back({kv.first, slot});
  }
  if (pages.empty()) return CNQ_OK;
  std::sort(pages.begin(), pages.end(),
            [](const std::pair<Key, uint32_t>& x, const std::pair<Key, uint32_t>& y) { return x.second < y.second; });
  BakeLists b;
  list_pages(c, pages, light, light + 1, b);
  DevBuf<float> fresh, dbounds;
  DevBuf<uint32_t> dlist, djobs, dstats;
  auto cleanup = [&]() { fresh.release(), dbounds.release(), dlist.release(), djobs.release(), dstats.release(); };
  hipError_t e = fresh.alloc(c->d_pool.n);
  if (e == hipSuccess) e = hipMemcpyAsync(fresh.p, c->d_pool.p, c->d_pool.bytes(), hipMemcpyDeviceToDevice, st);
  if (e != hipSuccess) {
    cleanup();
    return hip_fail(e, "verify ledger buffers");
  }
  int rc = bake_lists(c, b, fresh.p, st, 0, 0, dbounds, dlist, djobs, dstats, false);
  if (rc != CNQ_OK) {
    cleanup();
    return rc;
  }
  std::vector<float> before(c->d_pool.n), after(c->d_pool.n), bound((size_t)c->cfg.pages * c->n_moving);
  e = hipStreamSynchronize(st);
  if (e == hipSuccess) e = hipMemcpy(before.data(), c->d_pool.p, c->d_pool.bytes(), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(after.data(), fresh.p, c->d_pool.bytes(), hipMemcpyDeviceToHost);
  if (e == hipSuccess)
    e = hipMemcpy(bound.data(), c->d_ledger_bound.p, c->d_ledger_bound.bytes(), hipMemcpyDeviceToHost);
  cleanup();
  if (e != hipSuccess) return hip_fail(e, "verify ledger");
  uint32_t n_pages = 0, violations = 0, over = 0, changed = 0;
  float largest = 0.0f, largest_bound = 0.0f;
  for (const auto& pg : pages) {
    const uint32_t* k = pg.first.k;
    const uint32_t slot = pg.second;
    const uint32_t chart = k[1] & vc::KEY_CHART_MASK, mip = k[1] >> vc::KEY_MIP_SHIFT;
    uint32_t g[2];
    vc::grid(c->tables.charts[chart], c->scale[k[0]], c->inv_base, mip, g);
    const uint32_t px = (k[2] & 0xffffu) * vc::PAGE, py = (k[2] >> 16) * vc::PAGE;
    float change = 0.0f;
    for (uint32_t t = 0; t < vc::PAGE_TEXELS; t++) {
      if (px + t % vc::STRIDE >= g[0] || py + t / vc::STRIDE >= g[1]) continue;  // border texels copy inside ones
      const size_t at = ((size_t)slot * vc::PAGE_TEXELS + t) * c->channels + light;
      change = std::max(change, std::fabs(before[at] - after[at]));
    }
    const float bd = bound[(size_t)slot * c->n_moving + j];
    n_pages++;
    if (change > bd + 1e-5f) violations++;
    if (change > c->epsilon) over++;
    if (change > 0.0f) changed++;
    largest = std::max(largest, change);
    largest_bound = std::max(largest_bound, bd);
  }
  out[0] = (float)n_pages, out[1] = (float)violations, out[2] = (float)over, out[3] = (float)changed;
  out[4] = largest, out[5] = largest_bound;
  return CNQ_OK;
}

int cnq_vcache_state(cnq_vcache* c, uint32_t* slot_info, uint32_t* slot_stale, uint32_t* slot_used, float* pool,
                     uint32_t* table_device, uint32_t* table_host, uint32_t* sizes) {
  if (!c) return fail(CNQ_INVALID_ARG, "null argument");
  HIP_TRY(hipSetDevice(c->scene->dev->ordinal), "hipSetDevice");
  const uint32_t P = c->cfg.pages;
  if (sizes) {
    sizes[0] = P;
    sizes[1] = c->channels;
    sizes[2] = c->table.len();
    sizes[3] = c->slot_extent;
    sizes[4] = c->frame;
  }
  if (slot_info) HIP_TRY(hipMemcpy(slot_info, c->d_slot_info.p, 48 * (size_t)P, hipMemcpyDeviceToHost), "slot info");
  if (slot_stale) HIP_TRY(hipMemcpy(slot_stale, c->d_slot_stale.p, 4 * (size_t)P, hipMemcpyDeviceToHost), "stale");
  if (slot_used) HIP_TRY(hipMemcpy(slot_used, c->d_slot_used.p, 4 * (size_t)P, hipMemcpyDeviceToHost), "used");
  if (pool) HIP_TRY(hipMemcpy(pool, c->d_pool.p, c->d_pool.bytes(), hipMemcpyDeviceToHost), "pool");
  if (table_device)
    HIP_TRY(hipMemcpy(table_device, c->d_page_table.p, c->d_page_table.bytes(), hipMemcpyDeviceToHost), "table");
  if (table_host) std::memcpy(table_host, c->table.entries.data(), c->table.entries.size() * 4);
  return CNQ_OK;
}


int cnq_vcache_shade(cnq_vcache* c, const float* camera, uint32_t width, uint32_t height, const uint32_t* hits,
                     float* out, uint32_t* out_page, const cnq_vcache_shade_params* sp, void* stream,
                     uint32_t counts[2]) {
  if (!c || !camera || !hits || !out || !out_page || !sp) return fail(CNQ_INVALID_ARG, "null argument");
  if (width == 0 || height == 0 || (uint64_t)width * height > 0xffffffffull)
    return fail(CNQ_INVALID_ARG, "bad view size");
  if (sp->request_tile == 0 || sp->refine_levels == 0 || sp->refine_levels > 4)
    return fail(CNQ_INVALID_ARG, "request_tile >= 1, refine_levels 1..4");
  HIP_TRY(hipSetDevice(c->scene->dev->ordinal), "hipSetDevice");
  hipStream_t st = (hipStream_t)stream;
  // The camera goes by value through a small device copy.
  if (c->d_camera.n < 20) HIP_TRY(c->d_camera.alloc(20), "camera");
  HIP_TRY(hipMemcpyAsync(c->d_camera.p, camera, 80, hipMemcpyHostToDevice, st), "camera");
  HIP_TRY(fill_u32(c->d_request_set.p, 0, cnq::VC_REQUEST_SET, st), "request set clear");
  HIP_TRY(fill_u32(c->d_request_prio.p, 0xffffffffu, cnq::VC_REQUEST_SET, st), "request priorities clear");
  HIP_TRY(fill_u32(c->d_shade_counts.p, 0, 2, st), "shade counts clear");
  cnq::VcArgs a = c->args();
  cnq::ShadeArgs s;
  s.camera = c->d_camera.p;
  s.hits = hits;
  s.tri_chart = c->d_tri_chart.p;
  s.out = out;
  s.out_page = out_page;
  s.request_set = c->d_request_set.p;
  s.request_prio = c->d_request_prio.p;
  s.requests = c->d_requests.p;
  s.counts = c->d_shade_counts.p;
  s.w = width;
  s.h = height;
  s.set_mask = cnq::VC_REQUEST_SET - 1;
  s.request_cap = cnq::VC_REQUEST_CAP;
  s.tile = sp->request_tile;
  s.refine = sp->refine_levels;
  s.pixel_angle = sp->pixel_angle;
  s.pixels_per_texel = sp->pixels_per_texel;
  const uint64_t px = (uint64_t)width * height;
  hipLaunchKernelGGL(cnq::vc_shade_kernel, dim3((uint32_t)((px + 255) / 256)), dim3(256), 0, st, a, s);
  HIP_TRY(hipGetLastError(), "vc_shade launch");
  hipLaunchKernelGGL(cnq::vc_request_gather_kernel, dim3(cnq::VC_REQUEST_CAP / 256), dim3(256), 0, st, s);
  HIP_TRY(hipGetLastError(), "request gather launch");
  uint32_t got[2];
  HIP_TRY(hipMemcpyAsync(got, c->d_shade_counts.p, 8, hipMemcpyDeviceToHost, st), "shade counts");
  HIP_TRY(hipStreamSynchronize(st), "vc_shade");
  c->shade_requests = std::min(got[0], cnq::VC_REQUEST_CAP);
  if (counts) counts[0] = got[0], counts[1] = got[1];
  return CNQ_OK;
}

int cnq_vcache_plan(cnq_vcache* c, uint32_t budget, uint64_t stats[4]) {
  if (!c) return fail(CNQ_INVALID_ARG, "null argument");
  HIP_TRY(hipSetDevice(c->scene->dev->ordinal), "hipSetDevice");
  const uint32_t P = c->cfg.pages;
  // The pages pixels used (vcache.rs:3485 `merge_used`).
  std::vector<uint32_t> used(P);
  HIP_TRY(hipMemcpy(used.data(), c->d_slot_used.p, 4 * (size_t)P, hipMemcpyDeviceToHost), "slot use");
  for (uint32_t k = 0; k < P; k++) c->last_used[k] = std::max(c->last_used[k], used[k]);
  std::vector<uint32_t> req(4 * (size_t)c->shade_requests);
  if (!req.empty())
    HIP_TRY(hipMemcpy(req.data(), c->d_requests.p, req.size() * 4, hipMemcpyDeviceToHost), "requests");
  c->shade_requests = 0;
  // vcache.rs:3740-3790: the non-resident pages by (priority, key), at most
  // `budget` of them, in that order.
  struct Order {
    uint32_t priority, k0, k1, k2;
    bool operator<(const Order& o) const {
      if (priority != o.priority) return priority < o.priority;
      if (k0 != o.k0) return k0 < o.k0;
      if (k1 != o.k1) return k1 < o.k1;
      return k2 < o.k2;
    }
  };
  std::vector<Order> order;
  for (size_t i = 0; i < req.size(); i += 4) {
    const Key k{{req[i], req[i + 1], req[i + 2]}};
    if (c->map.count(k)) continue;
    order.push_back(Order{req[i + 3], req[i], req[i + 1], req[i + 2]});
  }
  const uint32_t cap = std::max<uint32_t>(budget, vc::PAGE_TEXELS);
  if (order.size() > cap) {
    std::nth_element(order.begin(), order.begin() + cap, order.end());
    order.resize(cap);
  }
  std::sort(order.begin(), order.end());
  uint64_t spent = 0, planned = 0, evicted = 0;
  bool evicted_now = false;
  for (const Order& o : order) {
    const uint32_t k[3] = {o.k0, o.k1, o.k2};
    const Key key{{k[0], k[1], k[2]}};
    if (c->map.count(key)) continue;  // a fingerprint collision can list a page twice
    const uint32_t m = k[0], chart = k[1] & vc::KEY_CHART_MASK, mip = k[1] >> vc::KEY_MIP_SHIFT;
    if (m >= c->scene->n_objects || chart >= c->tables.charts.size() || mip > 30) continue;
    const uint32_t pp[2] = {k[2] & 0xffffu, k[2] >> 16};
    uint32_t g[2];
    vc::grid(c->tables.charts[chart], c->scale[m], c->inv_base, mip, g);
    uint32_t inside[2], listed[2];
    bool empty = false;
    for (int ax = 0; ax < 2; ax++) {
      const uint32_t nn = g[ax] + 1;
      inside[ax] = std::min(nn - std::min(pp[ax] * vc::PAGE, nn), vc::STRIDE);
      if (inside[ax] == 0 || (uint64_t)pp[ax] * vc::PAGE >= g[ax]) empty = true;
      listed[ax] = std::min(inside[ax] + 1, vc::STRIDE);
    }
    if (empty) continue;
    const uint64_t cost = (uint64_t)inside[0] * inside[1] * c->n_lights;
    if (spent + cost > budget && !c->plan_list.empty()) break;
    if (c->free_slots.empty() && !evicted_now) {
      // vcache.rs:3541-3557 `evict_now(2)`: pages unused for more than two
      // frames, least recently used first, up to a frame's eviction count.
      evicted_now = true;
      std::vector<std::pair<uint32_t, uint32_t>> stale;
      for (uint32_t sl = 0; sl < P; sl++)
        if (c->slot_key[sl].k[0] != vc::EMPTY && c->last_used[sl] + 2 < c->frame) stale.push_back({c->last_used[sl], sl});
      std::sort(stale.begin(), stale.end());
      for (size_t q = 0; q < stale.size() && q < cnq::VC_EVICT_PER_FRAME; q++) {
        c->evict_slot(stale[q].second);
        evicted++;
      }
    }
    if (c->free_slots.empty()) break;
    const uint32_t slot = c->free_slots.back();
    c->free_slots.pop_back();
    const uint32_t entry = c->table.insert(k, slot);
    if (entry == vc::EMPTY) {
      c->free_slots.push_back(slot);
      continue;
    }
    c->slot_extent = std::max(c->slot_extent, slot + 1);
    c->map[key] = {slot, entry};
    c->slot_key[slot] = key;
    c->last_used[slot] = c->frame;
    c->alloc_frame[slot] = c->frame;
    const uint32_t index = (uint32_t)(c->plan_list.size() / 4);
    c->plan_list.insert(c->plan_list.end(), {k[0], k[1], k[2], slot});
    float b[8];
    vc::page_bounds(c->tables, chart, c->scale[m], c->inv_base, mip, pp[0], pp[1], b);
    b[3] = 0.0f;
    b[7] = 0.0f;
    c->plan_bounds.insert(c->plan_bounds.end(), b, b + 8);
    for (uint32_t l = 0; l < c->n_lights; l++)
      for (uint32_t y = 0; y < listed[1]; y++)
        for (uint32_t x = 0; x < listed[0]; x++)
          c->plan_jobs.insert(c->plan_jobs.end(), {index, (y * vc::STRIDE + x) | (l << 8)});
    spent += cost;
    planned++;
  }
  if (stats) {
    stats[0] = req.size() / 4;
    stats[1] = planned;
    stats[2] = spent;
    stats[3] = evicted;
  }
  return CNQ_OK;
}

}  // extern "C"
