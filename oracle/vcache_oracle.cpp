// TEST INFRASTRUCTURE ONLY (see oracle_math.h): a CPU restatement of the
// reference's view-cache page passes, used as the checker by tests/. Never
// linked by the product.
//
// Follows crates/continuum-gpu/src/vcache.rs (page table :904-1093, chart
// tables :688-790, grid :2104, page_bounds :2676, set_moved :2426-2492,
// planning :3940-4050, eviction :3425-3452), hierarchy.rs:418-457
// (shadow_region), instanced.rs:512-522 (host light boxes), and the WGSL
// passes in crates/continuum-gpu/src/mask_cones.wgsl (cited "WGSL :line"):
// vc_find :3134, vc_grid :3181, uv_barycentric :3250, vc_surface_at :3283,
// page_texel :3338, vc_update :3394, vc_bake :3432 (static lights),
// vc_fill :3693, world_box :3867, light_box :3887, hull_meets :3903,
// boxes_meet :3919, may_be_shadowed :3967, mark_invalid :3988, append_page
// :3998, append_jobs :4018, fail_entry :4029, vc_stale :4233,
// vc_stale_texels :4385. The passes run one invocation after another, in
// index order; where the GPU's atomics order appends differently, only the
// order of list entries differs, never a texel's value.
//
// Moving lights go through the error ledger (WGSL :4457-5164: `ledger_texel`
// :4547, `ledger_reset` :4643, `light_shift` :4666, the lattice's bounds
// :4907-5000, `vc_ledger` :5007; the guard and lattice cones are restated
// in cone_oracle.cpp), as in vc_bake :3448-3530 and vcache.rs:2255-2380.
// Left out as the product leaves them out: sky and bounce channels (the sky
// channel is zero) and their ledgers, deforming
// objects' catch-up slices (every move is rigid, so the reference's `lag_only` is
// false wherever it is asked and no page is deferred), area texels, wide
// windows' ray pass, absent lighting levels.
#include <algorithm>
#include <array>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <vector>

#include "oracle_math.h"

extern "C" void oracle_query_skip(const float* tris, const uint32_t* ranges, uint32_t n_objects, const float* blas,
                                  const uint32_t* blas_items, const float* blas_cov, const float* nodes,
                                  const uint32_t* items, const float* xforms, const uint32_t* opts,
                                  const float* point, const float* normal, const float* light, uint32_t skip_obj,
                                  uint32_t skip_tri, float* visible, uint64_t* masks);
extern "C" float oracle_unshadowed_rect(const float* point, const float* normal, const float* light);
extern "C" int oracle_faces_light(const float* light, const float* point, const float* normal);
extern "C" void oracle_light_center(const float* light, float* out);
extern "C" int oracle_guard_trace(const float* tris, const uint32_t* ranges, uint32_t n_objects, const float* blas,
                                  const uint32_t* blas_items, const float* blas_cov, const float* nodes,
                                  const uint32_t* items, const float* xforms, const uint32_t* opts, const float* apex,
                                  const float* light, int trace, float* alpha);
extern "C" float oracle_below_horizon(const float* apex, const float* normal, const float* light);
extern "C" float oracle_lattice_frame(const float* apex, const float* then, const float* normal, float* out);
extern "C" float oracle_lattice_ratio(const float* tris, const uint32_t* ranges, uint32_t n_objects,
                                      const float* blas, const uint32_t* blas_items, const float* blas_cov,
                                      const float* nodes, const uint32_t* items, const float* xforms,
                                      const uint32_t* opts, int on, const float* apex, const float* normal,
                                      const float* now, const float* then, int* flags, float* areas);

namespace vco {

using orc::V3;

constexpr uint32_t NONE = 0xffffffffu;
constexpr uint32_t PAGE = 8, STRIDE = 9, TEXELS = 81;
constexpr uint32_t MAX_PROBE = 64;
constexpr uint32_t MIP_SHIFT = 24, CHART_MASK = 0xffffffu;

// Rust's f32::min / f32::max.
static float rmin(float a, float b) { return std::isnan(a) ? b : (std::isnan(b) ? a : std::min(a, b)); }
static float rmax(float a, float b) { return std::isnan(a) ? b : (std::isnan(b) ? a : std::max(a, b)); }
static uint32_t sat_u32(float f) {
  if (!(f > 0.0f)) return 0;
  if (f >= 4294967296.0f) return NONE;
  return (uint32_t)f;
}
static V3 ld(const float* p) { return V3{p[0], p[1], p[2]}; }
static float comp(V3 v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }

// vcache.rs:904-913
uint32_t hash3(const uint32_t* k) {
  uint32_t h = (k[0] * 0x9E3779B1u) ^ (k[1] * 0x85EBCA77u) ^ (k[2] * 0xC2B2AE3Du);
  h ^= h >> 15;
  h *= 0x2C1B3C6Du;
  h ^= h >> 12;
  return h;
}

// vcache.rs:992-1093
struct Table {
  std::vector<uint32_t> e;  // 4 per entry
  uint32_t len() const { return (uint32_t)(e.size() / 4); }
  void init(uint32_t n) {
    e.assign(4 * (size_t)n, 0);
    for (uint32_t i = 0; i < n; i++) e[4 * i] = NONE;
  }
  uint32_t insert(const uint32_t* k, uint32_t slot) {
    uint32_t mask = len() - 1, i = hash3(k) & mask;
    for (uint32_t p = 0; p < MAX_PROBE; p++) {
      if (e[4 * i] == NONE) {
        e[4 * i] = k[0], e[4 * i + 1] = k[1], e[4 * i + 2] = k[2], e[4 * i + 3] = slot;
        return i;
      }
      i = (i + 1) & mask;
    }
    return NONE;
  }
  // Knuth's Algorithm R; `moved` gets (key, new entry).
  template <class F>
  void remove(uint32_t i, F moved) {
    uint32_t mask = len() - 1;
    e[4 * i] = NONE, e[4 * i + 1] = e[4 * i + 2] = e[4 * i + 3] = 0;
    uint32_t j = i;
    for (;;) {
      j = (j + 1) & mask;
      if (e[4 * j] == NONE) return;
      uint32_t k[3] = {e[4 * j], e[4 * j + 1], e[4 * j + 2]};
      uint32_t home = hash3(k) & mask;
      bool stays = i <= j ? (i < home && home <= j) : (i < home || home <= j);
      if (!stays) {
        for (int q = 0; q < 4; q++) e[4 * i + q] = e[4 * j + q];
        moved(k, i);
        e[4 * j] = NONE, e[4 * j + 1] = e[4 * j + 2] = e[4 * j + 3] = 0;
        i = j;
      }
    }
  }
  uint32_t find(const uint32_t* k) const {
    uint32_t mask = len() - 1, i = hash3(k) & mask;
    for (uint32_t p = 0; p < MAX_PROBE; p++) {
      if (e[4 * i] == NONE) return NONE;
      if (e[4 * i] == k[0] && e[4 * i + 1] == k[1] && e[4 * i + 2] == k[2]) return e[4 * i + 3];
      i = (i + 1) & mask;
    }
    return NONE;
  }
};

struct BoxF {
  float lo[3], hi[3];
};

struct State {
  // scene (caller-owned arrays)
  const float *tris, *blas, *blas_cov, *nodes, *xforms;
  const uint32_t *ranges, *blas_items, *items;
  uint32_t n_objects, n_tris;
  uint32_t opts[8];
  // surfaces
  std::vector<float> tri_uv, normals;
  std::vector<uint32_t> tri_chart;
  std::vector<float> chart_size;
  std::vector<uint32_t> chart_grid, first_cell, cells, cell_tris;
  std::vector<float> cell_bounds;
  uint32_t n_charts;
  // lights, config
  std::vector<float> lights;
  uint32_t n_lights, channels;
  uint32_t pages, append_pages, candidate_cap, light_ratio, visible_frames;
  float inv_base;
  // host manager
  Table table;
  std::map<std::vector<uint32_t>, std::pair<uint32_t, uint32_t>> map;
  std::vector<std::vector<uint32_t>> slot_key;
  std::vector<uint32_t> last_used, alloc_frame, free_slots;
  uint32_t slot_extent = 0, frame = 1;
  std::vector<uint32_t> plan_list, plan_jobs, used_marks, cleared;
  std::vector<float> plan_bounds;
  std::vector<float> moved;
  uint32_t n_movers = 0;
  std::vector<std::pair<uint32_t, uint32_t>> mover_ids;
  // "device" state
  std::vector<uint32_t> slot_info, slot_stale, slot_used;
  std::vector<float> pool;
  std::vector<uint32_t> bake_list, bake_jobs, candidates, invalid_list;
  std::vector<float> bake_bounds;
  // WGSL mover_stats: 0-7 as the stale passes (7 unused here), 8-14 and
  // 21-23 the moving lights' ledger.
  uint32_t stats[24];
  // moving lights (WGSL :4457-5164): per light its index among the moving
  // ones (NONE: static), the ledger records (slot x moving x 48 words), the
  // poses ring (POSE_RING frames x lights x 16 floats), each page's certified
  // bound, ε, and whether a moving light moved since the last frame.
  std::vector<uint32_t> moving_index, ledger;
  std::vector<float> poses, ledger_bound;
  uint32_t n_moving = 0;
  float epsilon = 0.01f;
  bool lights_moved = false;
  uint32_t last_ledger[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
};

static float object_scale(const State& s, uint32_t m) {
  uint32_t x = s.ranges[4 * m + 3];
  return x == NONE ? 1.0f : s.xforms[16 * x + 12];
}

// vcache.rs:2104-2111 (and WGSL :3181, the same products)
static void grid_of(const State& s, uint32_t c, uint32_t m, uint32_t mip, uint32_t n[2]) {
  float inv = std::ldexp(s.inv_base, -(int)mip);
  float sc = object_scale(s, m);
  for (int a = 0; a < 2; a++) {
    uint32_t v = sat_u32(std::ceil((s.chart_size[2 * c + a] * sc) * inv));
    n[a] = v > 1 ? v : 1;
  }
}

// vcache.rs:704-790
static bool build_tables(State& s, const uint32_t* tri_chart) {
  s.first_cell.resize(s.n_charts);
  uint64_t total = 0;
  for (uint32_t c = 0; c < s.n_charts; c++) {
    if (s.chart_grid[2 * c] == 0 || s.chart_grid[2 * c + 1] == 0) return false;
    s.first_cell[c] = (uint32_t)total;
    total += (uint64_t)s.chart_grid[2 * c] * s.chart_grid[2 * c + 1];
  }
  std::vector<std::vector<uint32_t>> lists(total);
  for (uint32_t t = 0; t < s.n_tris; t++) {
    uint32_t c = tri_chart[t];
    if (c >= s.n_charts) return false;
    const float* uv = &s.tri_uv[6 * t];
    uint32_t g[2] = {s.chart_grid[2 * c], s.chart_grid[2 * c + 1]};
    uint32_t r0[2], r1[2];
    for (int a = 0; a < 2; a++) {
      float lo = std::floor(rmin(rmin(uv[a], uv[2 + a]), uv[4 + a]) * (float)g[a] + 1e-3f);
      float hi = std::floor(rmax(rmax(uv[a], uv[2 + a]), uv[4 + a]) * (float)g[a] - 1e-3f);
      r0[a] = std::min(sat_u32(rmax(lo, 0.0f)), g[a] - 1);
      r1[a] = std::min(sat_u32(rmax(rmax(hi, lo), 0.0f)), g[a] - 1);
    }
    for (uint32_t y = r0[1]; y <= r1[1]; y++)
      for (uint32_t x = r0[0]; x <= r1[0]; x++) lists[s.first_cell[c] + y * g[0] + x].push_back(t);
  }
  s.cells.clear();
  s.cell_tris.clear();
  s.cell_bounds.clear();
  for (auto& list : lists) {
    s.cells.push_back((uint32_t)s.cell_tris.size());
    s.cells.push_back((uint32_t)list.size());
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (uint32_t t : list)
      for (int v = 0; v < 3; v++)
        for (int a = 0; a < 3; a++) {
          lo[a] = rmin(lo[a], s.tris[12 * t + 4 * v + a]);
          hi[a] = rmax(hi[a], s.tris[12 * t + 4 * v + a]);
        }
    s.cell_bounds.insert(s.cell_bounds.end(), {lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]});
    s.cell_tris.insert(s.cell_tris.end(), list.begin(), list.end());
  }
  if (s.cell_tris.empty()) s.cell_tris.push_back(0);
  return true;
}

// vcache.rs:2676-2698
static void page_bounds(const State& s, uint32_t m, uint32_t c, uint32_t mip, uint32_t px, uint32_t py, float out[8]) {
  uint32_t n[2];
  grid_of(s, c, m, mip, n);
  uint32_t g[2] = {s.chart_grid[2 * c], s.chart_grid[2 * c + 1]};
  uint32_t pg[2] = {px, py}, a0[2], a1[2];
  for (int a = 0; a < 2; a++) {
    uint32_t x0 = pg[a] * PAGE, x1 = std::min(pg[a] * PAGE + PAGE, n[a] - 1);
    auto cell = [&](uint32_t x) { return sat_u32(((float)x + 0.5f) / (float)n[a] * (float)g[a]); };
    uint32_t c0 = cell(x0);
    a0[a] = c0 == 0 ? 0 : c0 - 1;
    a1[a] = std::min(cell(x1) + 1, g[a] - 1);
  }
  float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
  for (uint32_t y = a0[1]; y <= a1[1]; y++)
    for (uint32_t x = a0[0]; x <= a1[0]; x++) {
      uint64_t cell = (uint64_t)s.first_cell[c] + (uint64_t)y * g[0] + x;
      if (cell >= s.cell_bounds.size() / 6) continue;
      for (int q = 0; q < 3; q++) {
        lo[q] = rmin(lo[q], s.cell_bounds[6 * cell + q]);
        hi[q] = rmax(hi[q], s.cell_bounds[6 * cell + 3 + q]);
      }
    }
  if (lo[0] > hi[0]) lo[0] = lo[1] = lo[2] = hi[0] = hi[1] = hi[2] = 0.0f;
  out[0] = lo[0], out[1] = lo[1], out[2] = lo[2], out[3] = 0.0f;
  out[4] = hi[0], out[5] = hi[1], out[6] = hi[2], out[7] = 0.0f;
}

// hierarchy.rs:418-457
bool shadow_region(const BoxF& m, const BoxF& l, const BoxF& sc, BoxF& out) {
  float dl[3], dh[3];
  for (int a = 0; a < 3; a++) dl[a] = m.lo[a] - l.hi[a], dh[a] = m.hi[a] - l.lo[a];
  float t0 = 0.0f, t1 = INFINITY;
  for (int a = 0; a < 3; a++) {
    float b = dl[a], c = sc.hi[a] - m.lo[a];
    if (b > 0.0f) t1 = rmin(t1, c / b);
    else if (b < 0.0f) t0 = rmax(t0, c / b);
    else if (c < 0.0f) return false;
    b = dh[a], c = sc.lo[a] - m.hi[a];
    if (b > 0.0f) t0 = rmax(t0, c / b);
    else if (b < 0.0f) t1 = rmin(t1, c / b);
    else if (c > 0.0f) return false;
  }
  if (t0 > t1 * (1.0f + 1e-6f) + 1e-6f) return false;
  if (!std::isfinite(t1)) {
    out = sc;
    return true;
  }
  float lo[3], hi[3], big = 0.0f;
  for (int a = 0; a < 3; a++) {
    lo[a] = rmax(rmin(m.lo[a] + dl[a] * t0, m.lo[a] + dl[a] * t1), sc.lo[a]);
    hi[a] = rmin(rmax(m.hi[a] + dh[a] * t0, m.hi[a] + dh[a] * t1), sc.hi[a]);
    big = rmax(big, rmax(std::fabs(lo[a]), std::fabs(hi[a])));
  }
  float pad = 1e-4f * rmax(big, 1.0f);
  for (int a = 0; a < 3; a++) out.lo[a] = lo[a] - pad, out.hi[a] = hi[a] + pad;
  return true;
}

// instanced.rs:512-522
BoxF host_light_box(const float* r) {
  float c[3], rad;
  if (r[3] > 1.5f) {
    for (int a = 0; a < 3; a++) c[a] = r[a];
    rad = 1.5f * r[4];
  } else {
    for (int a = 0; a < 3; a++) c[a] = r[a] + (r[4 + a] + r[8 + a]) * 0.5f;
    float lu = std::sqrt(r[4] * r[4] + r[5] * r[5] + r[6] * r[6]);
    float lv = std::sqrt(r[8] * r[8] + r[9] * r[9] + r[10] * r[10]);
    rad = 0.5f * (lu + lv);
  }
  float big = rmax(rmax(rmax(std::fabs(c[0]), std::fabs(c[1])), std::fabs(c[2])), 1.0f);
  rad = rad * 1.01f + 2e-5f * big + 1e-3f;
  BoxF b;
  for (int a = 0; a < 3; a++) b.lo[a] = c[a] - rad, b.hi[a] = c[a] + rad;
  return b;
}

// vcache.rs:2426-2492 for rigid moves (every light's mask and own bit set).
static void set_moved(State& s, const float* moved, uint32_t n, const BoxF& sc) {
  auto bits = [](uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
  };
  BoxF empty{{FLT_MAX, FLT_MAX, FLT_MAX}, {-FLT_MAX, -FLT_MAX, -FLT_MAX}};
  auto grow = [](BoxF a, const BoxF& b) {
    for (int q = 0; q < 3; q++) a.lo[q] = rmin(a.lo[q], b.lo[q]), a.hi[q] = rmax(a.hi[q], b.hi[q]);
    return a;
  };
  std::vector<BoxF> unions(s.n_lights, empty);
  BoxF all = empty;
  std::vector<float> entries;
  uint32_t every = s.n_lights >= 32 ? NONE : (1u << s.n_lights) - 1u;
  for (uint32_t i = 0; i < n; i++) {
    const float* m = moved + 13 * i;  // object bits, old lo, old hi, new lo, new hi
    uint32_t object;
    std::memcpy(&object, m, 4);
    BoxF old{{m[1], m[2], m[3]}, {m[4], m[5], m[6]}}, nw{{m[7], m[8], m[9]}, {m[10], m[11], m[12]}};
    all = grow(grow(all, old), nw);
    uint32_t id = (uint32_t)s.mover_ids.size();
    bool seen = false;
    for (auto& p : s.mover_ids)
      if (p.first == object) id = p.second, seen = true;
    if (!seen) s.mover_ids.push_back({object, id});
    entries.insert(entries.end(), {old.lo[0], old.lo[1], old.lo[2], bits(object)});
    entries.insert(entries.end(), {old.hi[0], old.hi[1], old.hi[2], bits(id)});
    entries.insert(entries.end(), {nw.lo[0], nw.lo[1], nw.lo[2], bits(every)});
    entries.insert(entries.end(), {nw.hi[0], nw.hi[1], nw.hi[2], 0.0f});
    for (uint32_t l = 0; l < s.n_lights; l++) {
      BoxF lb = host_light_box(&s.lights[16 * l]);
      for (const BoxF* b : {&old, &nw}) {
        BoxF r;
        if (!shadow_region(*b, lb, sc, r)) r = empty;
        if (r.lo[0] <= r.hi[0]) unions[l] = grow(unions[l], r);
        entries.insert(entries.end(), {r.lo[0], r.lo[1], r.lo[2], 0.0f});
        entries.insert(entries.end(), {r.hi[0], r.hi[1], r.hi[2], 0.0f});
      }
    }
  }
  s.moved.clear();
  for (auto& u : unions) {
    s.moved.insert(s.moved.end(), {u.lo[0], u.lo[1], u.lo[2], 0.0f});
    s.moved.insert(s.moved.end(), {u.hi[0], u.hi[1], u.hi[2], 0.0f});
  }
  s.moved.insert(s.moved.end(), {all.lo[0], all.lo[1], all.lo[2], 0.0f});
  s.moved.insert(s.moved.end(), {all.hi[0], all.hi[1], all.hi[2], 0.0f});
  s.moved.insert(s.moved.end(), entries.begin(), entries.end());
  s.n_movers = n;
}

// ---- the WGSL passes

struct Box3 {
  V3 lo, hi;
};

static float surface_offset(float d) { return orc::fmax_(1e-4f, 2e-6f * d); }

static V3 rot_world(const float* xf, V3 d) {
  return V3{orc::dot(ld(xf), d), orc::dot(ld(xf + 4), d), orc::dot(ld(xf + 8), d)};
}

// WGSL :3867-3882
static Box3 world_box(const State& s, uint32_t m, V3 lo, V3 hi) {
  Box3 b{lo, hi};
  uint32_t x = s.ranges[4 * m + 3];
  if (x != NONE) {
    const float* xf = s.xforms + 16 * x;
    float sc = xf[12];
    V3 c = (lo + hi) * 0.5f, e = (hi - lo) * 0.5f;
    V3 wc = rot_world(xf, c) * sc + V3{xf[3], xf[7], xf[11]};
    V3 we = V3{orc::dot(orc::vabs(ld(xf)), e), orc::dot(orc::vabs(ld(xf + 4)), e), orc::dot(orc::vabs(ld(xf + 8)), e)} * sc;
    b = Box3{wc - we, wc + we};
  }
  V3 far = orc::vmax(orc::vmax(orc::vabs(b.lo), orc::vabs(b.hi)), V3{1.0f, 1.0f, 1.0f});
  float pad = 2.0f * surface_offset(orc::length(far)) + 1e-5f * orc::fmax_(far.x, orc::fmax_(far.y, far.z));
  return Box3{b.lo - V3{pad, pad, pad}, b.hi + V3{pad, pad, pad}};
}

// WGSL :3887-3898
static Box3 light_box(const float* r) {
  float cc[3];
  oracle_light_center(r, cc);
  V3 c{cc[0], cc[1], cc[2]};
  float radius = r[3] > 1.5f ? 1.5f * r[4] : 0.5f * (orc::length(ld(r + 4)) + orc::length(ld(r + 8)));
  V3 far = orc::vmax(orc::vabs(c), V3{1.0f, 1.0f, 1.0f});
  radius = radius * 1.001f + 1e-5f * orc::fmax_(far.x, orc::fmax_(far.y, far.z)) + 1e-4f;
  return Box3{c - V3{radius, radius, radius}, c + V3{radius, radius, radius}};
}

// WGSL :3903-3917
bool hull_meets(const Box3& p, const Box3& l, const Box3& m) {
  float s0 = 0.0f, s1 = 1.0f;
  for (int a = 0; a < 3; a++) {
    float b0 = comp(l.lo, a) - comp(p.lo, a), c0 = comp(m.hi, a) - comp(p.lo, a);
    if (b0 > 0.0f) s1 = orc::fmin_(s1, c0 / b0);
    else if (b0 < 0.0f) s0 = orc::fmax_(s0, c0 / b0);
    else if (c0 < 0.0f) return false;
    float b1 = comp(l.hi, a) - comp(p.hi, a), c1 = comp(m.lo, a) - comp(p.hi, a);
    if (b1 > 0.0f) s0 = orc::fmax_(s0, c1 / b1);
    else if (b1 < 0.0f) s1 = orc::fmin_(s1, c1 / b1);
    else if (c1 > 0.0f) return false;
  }
  return s0 <= s1 + 1e-6f;
}

static bool boxes_meet(const Box3& a, const Box3& b) {
  return a.lo.x <= b.hi.x && a.lo.y <= b.hi.y && a.lo.z <= b.hi.z && b.lo.x <= a.hi.x && b.lo.y <= a.hi.y &&
         b.lo.z <= a.hi.z;
}

static Box3 moved_box(const State& s, uint32_t i) { return Box3{ld(&s.moved[4 * i]), ld(&s.moved[4 * (i + 1)])}; }
static uint32_t moved_w(const State& s, uint32_t i) {
  uint32_t u;
  std::memcpy(&u, &s.moved[4 * i + 3], 4);
  return u;
}

// WGSL :3967-3982
static bool may_be_shadowed(const State& s, const Box3& p, uint32_t l, uint32_t slot) {
  if (!boxes_meet(p, moved_box(s, 2 * l))) return false;
  Box3 lb = light_box(&s.lights[16 * l]);
  uint32_t first = 2 * s.n_lights + 2, stride = 4 + 4 * s.n_lights;
  for (uint32_t k = 0; k < s.n_movers; k++) {
    uint32_t e = first + k * stride;
    uint32_t slice = moved_w(s, e + 3);
    if (slice != 0 && slot % (slice >> 16) != (slice & 0xffffu)) continue;
    uint32_t r = e + 4 + 4 * l;
    if (boxes_meet(p, moved_box(s, r)) && hull_meets(p, lb, moved_box(s, e))) return true;
    if (boxes_meet(p, moved_box(s, r + 2)) && hull_meets(p, lb, moved_box(s, e + 2))) return true;
  }
  return false;
}

struct Surface {
  V3 apex, normal;
  bool ok;
  uint32_t tri;
};

// WGSL :3250-3260
static void uv_barycentric(const State& s, uint32_t t, float u0, float v0, float b[3]) {
  const float* q = &s.tri_uv[6 * t];
  float e1x = q[2] - q[0], e1y = q[3] - q[1], e2x = q[4] - q[0], e2y = q[5] - q[1];
  float dx = u0 - q[0], dy = v0 - q[1];
  float det = e1x * e2y - e1y * e2x;
  if (std::fabs(det) < 1e-30f) {
    b[0] = b[1] = b[2] = -1.0f;
    return;
  }
  float sv = (dx * e2y - dy * e2x) / det;
  float uv = (e1x * dy - e1y * dx) / det;
  b[0] = 1.0f - sv - uv, b[1] = sv, b[2] = uv;
}

// WGSL :3283-3326
static Surface surface_at(const State& s, uint32_t m, uint32_t c, float u, float v) {
  uint32_t g[2] = {s.chart_grid[2 * c], s.chart_grid[2 * c + 1]};
  uint32_t cx = std::min(sat_u32(u * (float)g[0]), g[0] - 1), cy = std::min(sat_u32(v * (float)g[1]), g[1] - 1);
  uint32_t cell = s.first_cell[c] + cy * g[0] + cx;
  uint32_t first = s.cells[2 * cell], count = s.cells[2 * cell + 1];
  uint32_t best_t = NONE;
  float best_b[3] = {0.0f, 0.0f, 0.0f}, best_score = -3.4e38f;
  for (uint32_t q = first; q < first + count; q++) {
    uint32_t t = s.cell_tris[q];
    float b[3];
    uv_barycentric(s, t, u, v, b);
    float score = orc::fmin_(b[0], orc::fmin_(b[1], b[2]));
    if (score > best_score) {
      best_score = score, best_t = t;
      best_b[0] = b[0], best_b[1] = b[1], best_b[2] = b[2];
    }
  }
  float b0[3];
  for (int k = 0; k < 3; k++) b0[k] = orc::clampf(best_b[k], 0.0f, 1.0f);
  float sum = orc::fmax_(b0[0] + b0[1] + b0[2], 1e-20f);
  float bx = b0[0] / sum, by = b0[1] / sum, bz = b0[2] / sum;
  uint32_t t = best_t != NONE ? best_t : 0;
  const float* tr = s.tris + 12 * t;
  V3 v0 = ld(tr), v1 = ld(tr + 4), v2 = ld(tr + 8);
  const float* nr = &s.normals[9 * t];
  V3 p = v0 * bx + v1 * by + v2 * bz;
  V3 in = ld(nr) * bx + ld(nr + 3) * by + ld(nr + 6) * bz;
  V3 face = orc::cross(v1 - v0, v2 - v0);
  V3 nrm = orc::dot(in, in) > 1e-20f ? in : face;
  V3 pw = p, nw = orc::normalize(nrm), geometric = face;
  uint32_t x = s.ranges[4 * m + 3];
  if (x != NONE) {
    const float* xf = s.xforms + 16 * x;
    pw = rot_world(xf, p) * xf[12] + V3{xf[3], xf[7], xf[11]};
    nw = orc::normalize(rot_world(xf, nrm));
    geometric = rot_world(xf, face);
  }
  Surface out;
  out.apex = pw + orc::normalize(geometric) * surface_offset(orc::length(pw));
  out.normal = nw;
  out.ok = best_t != NONE;
  out.tri = t;
  return out;
}

static Surface texel_surface(const State& s, uint32_t m, uint32_t c, const uint32_t n[2], uint32_t tx, uint32_t ty) {
  return surface_at(s, m, c, ((float)tx + 0.5f) / (float)n[0], ((float)ty + 0.5f) / (float)n[1]);
}

static void mark_invalid(State& s, uint32_t slot) {
  uint32_t old = s.slot_stale[slot];
  s.slot_stale[slot] = 1;
  if (old == 0 || old == 2 || old == 3) {
    uint32_t q = s.stats[1]++;
    if (q < s.pages) s.invalid_list[q] = slot;
  }
}

static uint32_t append_page(State& s, uint32_t slot, const uint32_t* key, V3 lo, V3 hi) {
  uint32_t p = s.stats[0]++;
  if (p >= s.append_pages) {
    s.stats[3]++;
    return NONE;
  }
  uint32_t entry = (uint32_t)(s.plan_list.size() / 4) + p;
  s.bake_list[4 * entry] = key[0], s.bake_list[4 * entry + 1] = key[1], s.bake_list[4 * entry + 2] = key[2];
  s.bake_list[4 * entry + 3] = slot;
  float* b = &s.bake_bounds[8 * entry];
  b[0] = lo.x, b[1] = lo.y, b[2] = lo.z, b[3] = 0.0f, b[4] = hi.x, b[5] = hi.y, b[6] = hi.z, b[7] = 0.0f;
  return entry;
}

static uint32_t append_jobs_cap(const State& s) { return s.append_pages * TEXELS * s.n_lights; }

static uint32_t append_jobs(State& s, uint32_t count) {
  uint32_t old = s.stats[5];
  s.stats[5] += count;
  uint32_t plan = (uint32_t)(s.plan_jobs.size() / 2), cap = append_jobs_cap(s);
  if (old + count <= cap) return plan + old;
  for (uint32_t k = old; k < std::min(old + count, cap); k++) {
    s.bake_jobs[2 * (plan + k)] = NONE;
    s.bake_jobs[2 * (plan + k) + 1] = 0;
  }
  return NONE;
}

static void fail_entry(State& s, uint32_t entry, uint32_t slot) {
  s.bake_bounds[8 * entry + 3] = 1.0f;
  s.slot_stale[slot] = 1;
  uint32_t q = s.stats[1]++;
  if (q < s.pages) s.invalid_list[q] = slot;
}

static float f_of(uint32_t u) {
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

// ---- moving lights through the error ledger (WGSL :4457-5164)

constexpr uint32_t LEDGER_WORDS = 48;        // WGSL :4481, vcache.rs:162
constexpr uint32_t MAX_OFFSET = 1023;        // WGSL :4483
constexpr uint32_t POSE_RING = 1024;         // vcache.rs:161
constexpr float GUARD = 3.0f;                // WGSL :4482
constexpr float LATTICE_MARGIN = 0.03f;      // WGSL :4704
constexpr float LATTICE_KAPPA_MAX = 8.0f;    // WGSL :4708

static uint32_t u_of(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return u;
}

static bool sphere_light(const float* r) { return r[3] > 1.5f; }

static V3 light_center_v(const float* r) {
  float cc[3];
  oracle_light_center(r, cc);
  return V3{cc[0], cc[1], cc[2]};
}

static V3 rect_corner(const float* r, uint32_t k) {
  return ld(r) + ld(r + 4) * (float)(k & 1u) + ld(r + 8) * (float)((k >> 1u) & 1u);
}

static const float* pose_at(const State& s, uint32_t frame, uint32_t l) {
  return &s.poses[((size_t)(frame % POSE_RING) * s.n_lights + l) * 16];
}

static uint32_t ledger_base(const State& s, uint32_t slot, uint32_t j) {
  return (slot * s.n_moving + j) * LEDGER_WORDS;
}

// WGSL :4666-4687: the largest angle a point of light pose `a` moves to
// reach pose `b`, seen from anywhere in box `page` (4: unbounded).
static float light_shift(const float* a, const float* b, const Box3& page) {
  float moved = 0.0f;
  if (sphere_light(a)) {
    moved = orc::length(ld(b) - ld(a)) + std::fabs(b[4] - a[4]);
  } else {
    for (uint32_t k = 0; k < 4; k++) moved = orc::fmax_(moved, orc::length(rect_corner(b, k) - rect_corner(a, k)));
  }
  if (moved <= 0.0f) return 0.0f;
  Box3 la = light_box(a), lb = light_box(b);
  Box3 span{orc::vmin(la.lo, lb.lo), orc::vmax(la.hi, lb.hi)};
  V3 gap = orc::vmax(V3{0.0f, 0.0f, 0.0f}, orc::vmax(page.lo - span.hi, span.lo - page.hi));
  float d = orc::length(gap);
  if (moved >= 2.0f * d) return 4.0f;
  return 2.0f * orc::asin_c(moved / (2.0f * d)) * 1.0001f;
}

struct LatticeReach {
  float near, far, cos_min, half_diagonal;
};

// WGSL :4907-4942
static LatticeReach lattice_reach(const float* a, const float* b, const Box3& page) {
  LatticeReach out;
  Box3 la = light_box(a), lb = light_box(b);
  Box3 span{orc::vmin(la.lo, lb.lo), orc::vmax(la.hi, lb.hi)};
  V3 gap = orc::vmax(V3{0.0f, 0.0f, 0.0f}, orc::vmax(page.lo - span.hi, span.lo - page.hi));
  out.near = orc::length(gap);
  V3 au = ld(a + 4), av = ld(a + 8), bu = ld(b + 4), bv = ld(b + 8);
  out.half_diagonal = 0.5f * orc::fmax_(orc::length(au + av), orc::length(bu + bv));
  out.half_diagonal = orc::fmax_(out.half_diagonal, 0.5f * orc::fmax_(orc::length(au - av), orc::length(bu - bv)));
  V3 pc = (page.lo + page.hi) * 0.5f;
  float ph = 0.5f * orc::length(page.hi - page.lo);
  V3 ca = light_center_v(a) - pc, cb = light_center_v(b) - pc;
  out.far = orc::fmax_(orc::length(ca), orc::length(cb)) + ph + out.half_diagonal;
  V3 na = orc::normalize(orc::cross(au, av)), nb = orc::normalize(orc::cross(bu, bv));
  float cos_c = orc::fmin_(std::fabs(orc::dot(na, orc::normalize(ca))), std::fabs(orc::dot(nb, orc::normalize(cb))));
  if (out.near > 0.0f)
    out.cos_min = cos_c - orc::asin_c(orc::fmin_((ph + out.half_diagonal) / out.near, 1.0f));
  else
    out.cos_min = -1.0f;
  return out;
}

// WGSL :4949-4971
static float lattice_bound(const float* a, const float* b, const Box3& page, float kappa) {
  float delta = light_shift(a, b, page);
  if (delta <= 0.0f) return 0.0f;
  if (delta >= 1.0f || kappa >= 1.0e30f) return 1.0f;
  LatticeReach r = lattice_reach(a, b, page);
  if (r.cos_min <= 0.0f || r.near <= 2.0f * r.half_diagonal) return 1.0f;
  V3 au = ld(a + 4), av = ld(a + 8), bu = ld(b + 4), bv = ld(b + 8);
  float area = orc::fmin_(orc::length(orc::cross(au, av)), orc::length(orc::cross(bu, bv)));
  float omega = area * r.cos_min / (r.far * r.far);
  float lu = orc::fmax_(orc::length(au), orc::length(bu));
  float lv = orc::fmax_(orc::length(av), orc::length(bv));
  float perimeter = 4.0f * (orc::asin_c(orc::fmin_(lu / (2.0f * r.near), 1.0f)) +
                            orc::asin_c(orc::fmin_(lv / (2.0f * r.near), 1.0f)));
  float rho = orc::fmin_(GUARD * orc::asin_c(orc::fmin_(r.half_diagonal / r.near, 1.0f)) * 1.0001f, 1.4f);
  float t = orc::tan_c(rho);
  float sec2 = 1.0f + t * t;
  float moved = sec2 * delta;
  float sw = sec2 * perimeter * moved + 3.14159265f * moved * moved;
  return orc::fmin_(1.5f * kappa * sw / omega * 1.0001f, 1.0f);
}

// WGSL :4981-5000
static bool lattice_grew(const float* a, const float* b, const Box3& page) {
  LatticeReach r = lattice_reach(a, b, page);
  if (r.cos_min <= 0.0f || r.near <= 2.0f * r.half_diagonal) return true;
  V3 ca = light_center_v(a), cb = light_center_v(b);
  float t = orc::length(cb - ca);
  float rest = 0.0f;
  for (uint32_t k = 0; k < 4; k++)
    rest = orc::fmax_(rest, orc::length((rect_corner(b, k) - cb) - (rect_corner(a, k) - ca)));
  float diameter = 2.0f * orc::asin_c(orc::fmin_(r.half_diagonal / r.near, 1.0f));
  float rho = orc::fmin_(GUARD * 0.5f * diameter * 1.0001f, 1.4f);
  float tr = orc::tan_c(rho);
  float sec2 = 1.0f + tr * tr;
  float delta = light_shift(a, b, page);
  float spread = sec2 * (3.0f * (t / r.near) * (2.0f * r.half_diagonal / r.near) + rest / r.near) +
                 3.0f * tr * sec2 * sec2 * diameter * delta;
  V3 au = ld(a + 4), av = ld(a + 8), bu = ld(b + 4), bv = ld(b + 8);
  float area = orc::fmin_(orc::length(orc::cross(au, av)), orc::length(orc::cross(bu, bv)));
  float width = area * r.cos_min / (r.far * r.far) / (sec2 * diameter);
  return 2.0f * spread * 1.0001f > LATTICE_MARGIN * width;
}

// WGSL :4643-4652
static void ledger_reset(State& s, uint32_t slot, uint32_t j) {
  uint32_t base = ledger_base(s, slot, j);
  s.ledger[base] = s.frame;
  s.ledger[base + 1] = 0x7f7fffffu;
  s.ledger[base + 2] = 0;
  s.ledger[base + 3] = s.frame;
  s.ledger[base + 4] = 0;
  for (uint32_t w = 5; w < LEDGER_WORDS; w++) s.ledger[base + w] = 0;
  s.ledger_bound[(size_t)slot * s.n_moving + j] = 0.0f;
}

static void guard_of(const State& s, V3 apex, const float* r, bool trace, bool& any, float& alpha) {
  float p[3] = {apex.x, apex.y, apex.z};
  any = oracle_guard_trace(s.tris, s.ranges, s.n_objects, s.blas, s.blas_items, s.blas_cov, s.nodes, s.items,
                           s.xforms, s.opts, p, r, trace ? 1 : 0, &alpha) != 0;
}

// WGSL :4547-4638: texel k of slot `slot`'s record for moving light j
// (light index `l`, pose `r` now, `then` at the page's version).
static void ledger_texel(State& s, bool on, uint32_t slot, uint32_t j, uint32_t l, uint32_t k, V3 apex, V3 n,
                         const float* r, float open, const float* then, bool lattice, bool overflow) {
  uint32_t base = ledger_base(s, slot, j);
  uint32_t version = s.ledger[base];
  uint32_t offset = std::min(s.frame - version, MAX_OFFSET);
  uint32_t state = 0;
  float guard = 3.4e38f;
  float reach = 0.0f, margin = 3.4e38f, bin = 0.0f, kappa = 3.0e38f;
  float ap[3] = {apex.x, apex.y, apex.z}, np[3] = {n.x, n.y, n.z};
  if (lattice && on && open > 0.0f) {
    float f[13];
    float kap = oracle_lattice_frame(ap, then, np, f);
    if (f[12] != 0.0f) {
      reach = orc::fmin_(GUARD * f[11], 1.4f);
      bin = std::sqrt(f[9] * f[9] + f[10] * f[10]);
      margin = f[11] + bin;
      kappa = kap;
    }
  }
  if (on && open > 0.0f && offset != 0) {
    state = 3;
    if (lattice) guard = reach - margin;
  } else if (on && open > 0.0f) {
    bool any;
    float alpha;
    guard_of(s, apex, r, true, any, alpha);
    state = any ? 3 : 2;
    if (-oracle_below_horizon(ap, np, r) <= orc::fmax_(2.0f * orc::fmin_(GUARD * alpha, 1.4f), 0.1f)) state = 3;
    guard = orc::fmin_(GUARD * alpha, 1.4f) - alpha;
    if (lattice) guard = orc::fmin_(orc::fmin_(GUARD * alpha, 1.4f), reach) - orc::fmax_(margin, alpha + bin);
  } else if (on) {
    float below = oracle_below_horizon(ap, np, r);
    bool any;
    float alpha;
    guard_of(s, apex, r, false, any, alpha);
    if (below >= 2.0f * alpha)
      guard = below;
    else
      state = 3;
  }
  if ((state != 3 || lattice) && guard < 3.0e38f) {
    float slack = guard;
    if (offset != 0) {
      Box3 point{apex, apex};
      slack = guard - light_shift(pose_at(s, version, l), r, point);
    }
    s.ledger[base + 1] = std::min(s.ledger[base + 1], u_of(orc::fmax_(slack, 0.0f)));
  }
  uint32_t code = 0;
  if (lattice && state == 3) {
    if (on && open > 0.0f && kappa <= LATTICE_KAPPA_MAX)
      s.ledger[base + 2] = std::max(s.ledger[base + 2], u_of(kappa));
    else
      code = 4;
  }
  uint32_t shift = 16 * (k % 2), word = base + 5 + k / 2;
  s.ledger[word] &= ~(0xffffu << shift);
  s.ledger[word] |= ((offset << 6) | code | state) << shift;
  uint32_t flags = (state == 3 ? 1u : 0u) | (offset != 0 ? 2u : 0u) | (overflow ? 8u : 0u);
  if (flags) s.ledger[base + 4] |= flags;
}

static float share_of(uint32_t slot, uint32_t frame, uint32_t salt) {
  uint32_t k[3] = {slot, frame, salt};
  return 0.5f + 0.5f * (float)(hash3(k) & 0xffffu) / 65535.0f;
}

// WGSL :5007-5164: one slot's check of every moving light's record.
static void vc_ledger(State& s, uint32_t slot, bool lit) {
  const uint32_t* info = &s.slot_info[12 * slot];
  uint32_t set_aside = s.slot_stale[slot];
  if (info[3] == 0 || set_aside == 1) return;
  bool visible = s.slot_used[slot] + 1 >= s.frame;
  if (set_aside == 2 && !visible) return;
  if (set_aside == 0 && !lit) return;
  if (visible) s.stats[11]++;
  V3 lo{f_of(info[4]), f_of(info[5]), f_of(info[6])}, hi{f_of(info[8]), f_of(info[9]), f_of(info[10])};
  Box3 page = world_box(s, info[0], lo, hi);
  const uint32_t key[3] = {info[0], info[1], info[2]};
  const uint32_t limit = std::min(append_jobs_cap(s), 0x7fffffffu);
  uint32_t entry = NONE;
  bool counted = false;
  for (uint32_t l = 0; l < s.n_lights; l++) {
    uint32_t j = s.moving_index[l];
    if (j == NONE) continue;
    uint32_t base = ledger_base(s, slot, j);
    uint32_t version = s.ledger[base];
    if (version == 0) continue;
    const float* now = &s.lights[16 * l];
    float slack = f_of(s.ledger[base + 1]);
    uint32_t page_flags = s.ledger[base + 4];
    bool mixed_page = (page_flags & 1u) != 0;
    uint32_t age = s.frame - version;
    bool whole = age + 1 >= std::min(POSE_RING, MAX_OFFSET);
    bool lattice = !sphere_light(now);
    if (lattice && !whole && (page_flags & 8u)) {
      whole = true;
      if (visible) s.stats[22]++;
    }
    uint32_t stale[3] = {0, 0, 0}, count = 0;
    float bound = 0.0f;
    bool held = false;
    if (!whole) {
      float d0 = light_shift(pose_at(s, version, l), now, page);
      if (d0 > slack * share_of(slot, version, 0x5bd1e995u)) {
        whole = true;
      } else if (!mixed_page) {
        if (visible) s.stats[13]++;
        s.ledger_bound[(size_t)slot * s.n_moving + j] = 0.0f;
        continue;
      } else {
        if (visible) s.stats[12]++;
        if (lattice) {
          uint32_t since = s.ledger[base + 3];
          float kappa = f_of(s.ledger[base + 2]);
          if (s.frame - since < POSE_RING) {
            float b = lattice_bound(pose_at(s, since, l), now, page, kappa);
            held = b <= s.epsilon * share_of(slot, since, 0x27d4eb2du);
            if (held && lattice_grew(pose_at(s, version, l), now, page)) {
              held = false;
              if (visible) s.stats[22]++;
            }
            if (held) bound = b;
          }
        }
        uint32_t want = held ? 7u : 3u;
        for (uint32_t k = 0; k < TEXELS; k++) {
          uint32_t half = (s.ledger[base + 5 + k / 2] >> (16 * (k % 2))) & 0xffffu;
          if ((half & want) == want) {
            stale[k / 32] |= 1u << (k % 32);
            count++;
          }
        }
        if (held && visible) {
          s.stats[21]++;
          s.stats[23] += count;
        }
        whole = count == TEXELS && !held;
      }
    }
    if (whole) count = TEXELS;
    s.ledger_bound[(size_t)slot * s.n_moving + j] = bound;
    if (count == 0) continue;
    if (!visible) {
      if (s.slot_stale[slot] == 0) s.stats[10]++;
      s.slot_stale[slot] = 2;
      return;
    }
    if (entry == NONE) {
      entry = append_page(s, slot, key, lo, hi);
      if (entry == NONE) {
        s.slot_stale[slot] = 2;
        s.stats[10]++;
        return;
      }
    }
    uint32_t first = NONE;
    uint32_t before = s.stats[14];
    s.stats[14] += count;
    if (before + count <= limit) first = append_jobs(s, count);
    if (first == NONE) {
      s.stats[10]++;
      s.bake_bounds[8 * entry + 3] = 1.0f;
      s.slot_stale[slot] = 2;
      return;
    }
    if (whole) {
      ledger_reset(s, slot, j);
      for (uint32_t k = 0; k < TEXELS; k++)
        s.bake_jobs[2 * (first + k)] = entry, s.bake_jobs[2 * (first + k) + 1] = k | (l << 8);
    } else {
      uint32_t at = 0;
      for (uint32_t k = 0; k < TEXELS; k++)
        if (stale[k / 32] & (1u << (k % 32))) {
          s.bake_jobs[2 * (first + at)] = entry, s.bake_jobs[2 * (first + at) + 1] = k | (l << 8);
          at++;
        }
      if (!held) s.ledger[base + 3] = s.frame;
    }
    s.stats[9] += count;
    if (!counted) {
      counted = true;
      s.stats[8]++;
    }
  }
  if (set_aside == 2 && entry == NONE) s.slot_stale[slot] = 0;
}

// WGSL :4233-4379 (rigid moves, no sky)
static void vc_stale(State& s, uint32_t slot) {
  const uint32_t* info = &s.slot_info[12 * slot];
  if (info[3] == 0) return;
  uint32_t mark = s.slot_stale[slot];
  if (mark == 1) return;
  uint32_t m = info[0];
  uint32_t key[3] = {info[0], info[1], info[2]};
  bool visible = s.slot_used[slot] + s.visible_frames >= s.frame;
  if (visible) s.stats[2]++;
  V3 lo{f_of(info[4]), f_of(info[5]), f_of(info[6])}, hi{f_of(info[8]), f_of(info[9]), f_of(info[10])};
  Box3 page = world_box(s, m, lo, hi);
  uint32_t L = s.n_lights, first = 2 * L + 2, stride = 4 + 4 * L;
  bool own = false;
  uint32_t own_mask = 0;
  if (boxes_meet(page, moved_box(s, 2 * L)))
    for (uint32_t k = 0; k < s.n_movers; k++)
      if (moved_w(s, first + k * stride) == m) own = true, own_mask = moved_w(s, first + k * stride + 2);
  if (own && own_mask != 0) {
    if (!visible) return mark_invalid(s, slot);
    uint32_t entry = append_page(s, slot, key, lo, hi);
    if (entry == NONE) return mark_invalid(s, slot);
    uint32_t all = L >= 32 ? NONE : (1u << L) - 1u;
    uint32_t updated = (uint32_t)__builtin_popcount(own_mask & all);
    uint32_t j = append_jobs(s, TEXELS * updated);
    if (j == NONE) {
      s.stats[3]++;
      return fail_entry(s, entry, slot);
    }
    s.stats[6]++;
    uint32_t at = j;
    for (uint32_t l = 0; l < L; l++) {
      if (!(own_mask & (1u << l))) continue;
      for (uint32_t k = 0; k < TEXELS; k++) s.bake_jobs[2 * (at + k)] = entry, s.bake_jobs[2 * (at + k) + 1] = k | (l << 8);
      at += TEXELS;
    }
    return;
  }
  bool candidate = false;
  for (uint32_t l = 0; l < L && !candidate; l++) candidate = may_be_shadowed(s, page, l, slot);
  if (!candidate) return;
  if (mark != 0) return mark_invalid(s, slot);
  uint32_t c = s.stats[4]++;
  if (c >= s.candidate_cap) return mark_invalid(s, slot);
  uint32_t entry = NONE;
  if (visible) {
    entry = append_page(s, slot, key, lo, hi);
    if (entry == NONE) {
      s.candidates[2 * c] = NONE, s.candidates[2 * c + 1] = NONE;
      return mark_invalid(s, slot);
    }
  }
  s.candidates[2 * c] = slot, s.candidates[2 * c + 1] = entry;
}

// WGSL :4385-4432
static void vc_stale_texel(State& s, uint32_t j) {
  uint32_t c = j / TEXELS, k = j % TEXELS;
  if (c >= std::min(s.stats[4], s.candidate_cap)) return;
  uint32_t slot = s.candidates[2 * c], entry = s.candidates[2 * c + 1];
  if (slot == NONE) return;
  const uint32_t* info = &s.slot_info[12 * slot];
  uint32_t m = info[0], chart = info[1] & CHART_MASK, mip = info[1] >> MIP_SHIFT;
  uint32_t n[2];
  grid_of(s, chart, m, mip, n);
  uint32_t tx = (info[2] & 0xffffu) * PAGE + k % STRIDE, ty = (info[2] >> 16) * PAGE + k / STRIDE;
  if (tx >= n[0] || ty >= n[1]) return;
  Surface sf = texel_surface(s, m, chart, n, tx, ty);
  if (!sf.ok) return;
  V3 far = orc::vmax(orc::vabs(sf.apex), V3{1.0f, 1.0f, 1.0f});
  float pad = 1e-5f * orc::fmax_(far.x, orc::fmax_(far.y, far.z));
  Box3 apex{sf.apex - V3{pad, pad, pad}, sf.apex + V3{pad, pad, pad}};
  for (uint32_t l = 0; l < s.n_lights; l++) {
    if (!may_be_shadowed(s, apex, l, slot)) continue;
    if (entry == NONE) return mark_invalid(s, slot);
    uint32_t at = append_jobs(s, 1);
    if (at == NONE) {
      s.stats[3]++;
      return fail_entry(s, entry, slot);
    }
    s.bake_jobs[2 * at] = entry, s.bake_jobs[2 * at + 1] = k | (l << 8);
    if (s.slot_stale[slot] == 0) s.stats[6]++;
    s.slot_stale[slot] = 1;
  }
}

// WGSL :3432-3554: one job. A moving light's texel (WGSL :3448-3530) bakes
// on its page's lattice when the light is a rect one, snaps ratios within ε
// of 0 or 1 when not, and records its ledger state (`record`); `pool` is
// the pool written (the cache's, or a check's copy).
static void vc_bake_job(State& s, uint32_t j, uint32_t total, bool record = true, float* pool = nullptr) {
  if (!pool) pool = s.pool.data();
  uint32_t jx = s.bake_jobs[2 * std::min(j, total - 1)], jy = s.bake_jobs[2 * std::min(j, total - 1) + 1];
  bool listed = j < total && jx != NONE;
  if (!listed) return;  // an unlisted invocation writes nothing
  uint32_t k = jy & 0xffu, l = std::min(jy >> 8, s.n_lights - 1);
  const uint32_t* e = &s.bake_list[4 * jx];
  uint32_t m = e[0], c = e[1] & CHART_MASK, mip = e[1] >> MIP_SHIFT;
  uint32_t n[2];
  grid_of(s, c, m, mip, n);
  uint32_t tx = (e[2] & 0xffffu) * PAGE + k % STRIDE, ty = (e[2] >> 16) * PAGE + k / STRIDE;
  bool inside = tx < n[0] && ty < n[1];
  if (!inside) return;
  Surface sf = texel_surface(s, m, c, n, tx, ty);
  bool on = sf.ok;
  const float* light = &s.lights[16 * l];
  float apex[3] = {sf.apex.x, sf.apex.y, sf.apex.z}, nrm[3] = {sf.normal.x, sf.normal.y, sf.normal.z};
  float open = oracle_unshadowed_rect(apex, nrm, light);
  const uint32_t mj = s.n_moving ? s.moving_index[l] : NONE;
  const bool lattice = mj != NONE && !sphere_light(light);
  const float* then = light;
  if (lattice) {
    uint32_t version = s.ledger[ledger_base(s, e[3], mj)];
    if (version != 0 && s.frame - version < POSE_RING) then = pose_at(s, version, l);
  }
  bool on_lattice = lattice;
  if (lattice && on && open > 0.0f) on_lattice = oracle_lattice_frame(apex, then, nrm, nullptr) <= LATTICE_KAPPA_MAX;
  int lat_flags[2] = {1, 0};
  float lat = -1.0f;
  if (lattice)
    lat = oracle_lattice_ratio(s.tris, s.ranges, s.n_objects, s.blas, s.blas_items, s.blas_cov, s.nodes, s.items,
                               s.xforms, s.opts, on && on_lattice && open > 0.0f, apex, nrm, light, then, lat_flags,
                               nullptr);
  const bool overflow = on && on_lattice && open > 0.0f && lat < 0.0f;
  float cc[3];
  oracle_light_center(light, cc);
  V3 toward = orc::normalize(V3{cc[0], cc[1], cc[2]} - sf.apex);
  float tw[3] = {toward.x, toward.y, toward.z};
  bool back = on && open <= 0.0f && mj == NONE && oracle_faces_light(light, apex, tw);
  float seen = back ? oracle_unshadowed_rect(apex, tw, light) : open;
  float value = 0.0f;
  uint64_t masks[2] = {0, 0};
  if (on && (!on_lattice || overflow))
    oracle_query_skip(s.tris, s.ranges, s.n_objects, s.blas, s.blas_items, s.blas_cov, s.nodes, s.items, s.xforms,
                      s.opts, apex, back ? tw : nrm, light, back ? m : NONE, back ? sf.tri : NONE, &value, masks);
  bool last_empty = masks[1] == 0, last_full = masks[0] == ~0ull;
  float ratio = !on ? 0.0f : (seen > 0.0f ? orc::clampf(value / seen, 0.0f, 1.0f) : 1.0f);
  if (on && seen > 0.0f && last_empty) ratio = 1.0f;
  if (on && seen > 0.0f && last_full) ratio = 0.0f;
  if (mj != NONE) {
    if (ratio < s.epsilon) ratio = 0.0f;
    if (ratio > 1.0f - s.epsilon) ratio = 1.0f;
  }
  float light_value = back ? 0.0f : value;
  if (on_lattice && !overflow) {
    ratio = !on ? 0.0f : (open > 0.0f ? lat : 1.0f);
    if (on && open > 0.0f && lat_flags[0]) ratio = 1.0f;
    if (on && open > 0.0f && lat_flags[1]) ratio = 0.0f;
    light_value = ratio * open;
  }
  pool[((size_t)e[3] * TEXELS + k) * s.channels + l] = s.light_ratio ? ratio : (on ? light_value : 0.0f);
  if (mj != NONE && record)
    ledger_texel(s, on, e[3], mj, l, k, sf.apex, sf.normal, light, open, then, lattice, overflow);
}

// WGSL :3693-3836 without sky or bounce light: one page.
static void vc_fill(State& s, uint32_t p) {
  uint32_t plan_pages = (uint32_t)(s.plan_list.size() / 4);
  const uint32_t* e = &s.bake_list[4 * p];
  const float* bb = &s.bake_bounds[8 * p];
  if (bb[3] != 0.0f) return;
  uint32_t slot = e[3];
  uint32_t* info = &s.slot_info[12 * slot];
  float kind = bb[7];
  uint32_t prev_w = info[3];
  bool same = prev_w != 0 && info[0] == e[0] && info[1] == e[1] && info[2] == e[2];
  bool fresh = p < plan_pages && (kind == 0.0f || kind == 2.0f || kind == 5.0f);
  bool keep = same && !fresh;
  uint32_t charge = keep ? info[7] : 0;
  uint32_t moving = fresh && s.n_movers > 0 ? 4 : 0;
  uint32_t b[8];
  std::memcpy(b, bb, 32);
  info[0] = e[0], info[1] = e[1], info[2] = e[2], info[3] = 1u | (keep ? prev_w & 0xffffff06u : 0u) | moving;
  info[4] = b[0], info[5] = b[1], info[6] = b[2], info[7] = charge;
  info[8] = b[4], info[9] = b[5], info[10] = b[6], info[11] = b[7];
  if (kind == 5.0f) {
  } else if (kind >= 3.0f) {
    if (s.slot_stale[slot] == 3) s.slot_stale[slot] = 0;
  } else {
    s.slot_stale[slot] = kind == 2.0f ? 3 : 0;
  }
  uint32_t c = e[1] & CHART_MASK, mip = e[1] >> MIP_SHIFT;
  uint32_t n[2];
  grid_of(s, c, e[0], mip, n);
  uint32_t px = (e[2] & 0xffffu) * PAGE, py = (e[2] >> 16) * PAGE;
  float* o = &s.pool[(size_t)slot * TEXELS * s.channels];
  uint32_t sky = s.channels - 1;
  for (uint32_t k = 0; k < TEXELS; k++) o[k * s.channels + sky] = 0.0f;
  for (uint32_t k = 0; k < TEXELS; k++) {
    uint32_t tx = px + k % STRIDE, ty = py + k / STRIDE;
    if (tx >= n[0] || ty >= n[1]) {
      uint32_t ix = std::min(tx, n[0] - 1) - px, iy = std::min(ty, n[1] - 1) - py;
      uint32_t nearest = iy * STRIDE + ix;
      for (uint32_t ch = 0; ch < s.channels; ch++) o[k * s.channels + ch] = o[nearest * s.channels + ch];
    }
  }
  if (p < plan_pages) s.slot_used[slot] = s.frame;
}

// ---- shading from the cache (WGSL :5907-6130, the direct light's part)

static uint32_t find_shown_slot(const State& s, const uint32_t* k) {
  uint32_t slot = s.table.find(k);
  if (slot != NONE && s.slot_stale[slot] != 0 && s.slot_stale[slot] != 3) slot = NONE;
  return slot;
}

// WGSL :3187-3196
static uint32_t top_mip(const State& s, uint32_t c, uint32_t m) {
  float ratio = orc::fmax_(s.chart_size[2 * c], s.chart_size[2 * c + 1]) * object_scale(s, m) * s.inv_base;
  uint32_t top = 0;
  if (ratio > 1.0f) {
    float l = std::ceil(std::log2(ratio));
    top = l < 30.0f ? (uint32_t)l : 30u;
  }
  uint32_t n[2];
  if (top > 0) {
    grid_of(s, c, m, top - 1, n);
    if (n[0] == 1 && n[1] == 1) top -= 1;
  }
  grid_of(s, c, m, top, n);
  if (n[0] != 1 || n[1] != 1) top += 1;
  return top;
}

struct PageRef {
  uint32_t key[3], base[2];
  float frac[2];
  bool coarsest;
};

// WGSL :3218-3231
static PageRef page_ref(const State& s, uint32_t m, uint32_t c, float u, float v, uint32_t mip) {
  uint32_t n[2];
  grid_of(s, c, m, mip, n);
  float tx = orc::clampf(u * (float)n[0] - 0.5f, 0.0f, (float)(n[0] - 1)), ty = orc::clampf(v * (float)n[1] - 0.5f, 0.0f, (float)(n[1] - 1));
  uint32_t bx = (uint32_t)std::floor(tx), by = (uint32_t)std::floor(ty);
  uint32_t px = bx / PAGE, py = by / PAGE;
  PageRef r;
  r.key[0] = m, r.key[1] = c | (mip << MIP_SHIFT), r.key[2] = px | (py << 16);
  r.base[0] = bx - px * PAGE, r.base[1] = by - py * PAGE;
  r.frac[0] = tx - std::floor(tx), r.frac[1] = ty - std::floor(ty);
  r.coarsest = n[0] == 1 && n[1] == 1;
  return r;
}

// WGSL :3242-3247
static float bilinear(const State& s, uint32_t slot, const PageRef& r, uint32_t ch) {
  size_t o = (size_t)slot * TEXELS + r.base[1] * STRIDE + r.base[0], C = s.channels;
  float a = orc::mixf(s.pool[o * C + ch], s.pool[(o + 1) * C + ch], r.frac[0]);
  float b = orc::mixf(s.pool[(o + STRIDE) * C + ch], s.pool[(o + STRIDE + 1) * C + ch], r.frac[0]);
  return orc::mixf(a, b, r.frac[1]);
}

static V3 point_to_object(const State& s, uint32_t m, V3 p) {
  uint32_t x = s.ranges[4 * m + 3];
  if (x == NONE) return p;
  const float* xf = s.xforms + 16 * x;
  V3 d = p - V3{xf[3], xf[7], xf[11]};
  return (ld(xf) * d.x + ld(xf + 4) * d.y + ld(xf + 8) * d.z) * xf[13];
}

static V3 normal_world(const State& s, uint32_t m, V3 n) {
  uint32_t x = s.ranges[4 * m + 3];
  return x == NONE ? n : rot_world(s.xforms + 16 * x, n);
}

struct Request {
  uint32_t k[3], priority;
};

// WGSL :5907-5925 on the CPU: a set of fingerprints, first come first
// listed; a page several pixels ask for keeps the smallest priority.
struct RequestSet {
  std::vector<uint32_t> set, prio;
  std::vector<Request> list;
  std::vector<uint32_t> entry;  // per listed request, its set entry
  void request(const uint32_t* k, uint32_t priority) {
    uint32_t mask = (uint32_t)set.size() - 1;
    uint32_t fp = hash3(k) | 1u;
    uint32_t k2[3] = {k[2], k[0], k[1]};
    uint32_t i = hash3(k2) & mask;
    for (uint32_t probe = 0; probe < 32; probe++) {
      if (set[i] == 0) {
        set[i] = fp;
        prio[i] = std::min(prio[i], priority);
        list.push_back(Request{{k[0], k[1], k[2]}, 0});
        entry.push_back(i);
        return;
      }
      if (set[i] == fp) {
        prio[i] = std::min(prio[i], priority);
        return;
      }
      i = (i + 1) & mask;
    }
  }
};

}  // namespace vco

namespace vco {

static void evict_slot(State& s, uint32_t slot) {
  auto it = s.map.find(s.slot_key[slot]);
  if (it == s.map.end()) return;
  uint32_t entry = it->second.second;
  s.map.erase(it);
  s.table.remove(entry, [&](const uint32_t* k, uint32_t i) { s.map[std::vector<uint32_t>(k, k + 3)].second = i; });
  s.slot_key[slot].clear();
  s.free_slots.push_back(slot);
  s.cleared.push_back(slot);
}

}  // namespace vco

using namespace vco;

extern "C" {

void* oracle_vc_create(const float* tris, uint32_t n_tris, const uint32_t* ranges, uint32_t n_objects,
                       const float* blas, const uint32_t* blas_items, const float* blas_cov, const float* nodes,
                       const uint32_t* items, const float* xforms, const uint32_t* opts, const uint32_t* tri_chart,
                       const float* tri_uv, const float* normals, uint32_t n_charts, const float* chart_size,
                       const uint32_t* chart_grid, const float* lights, uint32_t n_lights, const uint32_t* cfg,
                       float base_texel) {
  State* s = new State();
  s->tris = tris, s->ranges = ranges, s->n_objects = n_objects, s->blas = blas, s->blas_items = blas_items;
  s->blas_cov = blas_cov, s->nodes = nodes, s->items = items, s->xforms = xforms, s->n_tris = n_tris;
  std::memcpy(s->opts, opts, sizeof(s->opts));
  s->tri_uv.assign(tri_uv, tri_uv + 6 * (size_t)n_tris);
  s->tri_chart.assign(tri_chart, tri_chart + n_tris);
  s->normals.assign(normals, normals + 9 * (size_t)n_tris);
  s->n_charts = n_charts;
  s->chart_size.assign(chart_size, chart_size + 2 * (size_t)n_charts);
  s->chart_grid.assign(chart_grid, chart_grid + 2 * (size_t)n_charts);
  s->lights.assign(lights, lights + 16 * (size_t)n_lights);
  s->n_lights = n_lights, s->channels = n_lights + 1;
  s->pages = cfg[0], s->append_pages = cfg[1], s->candidate_cap = cfg[2], s->light_ratio = cfg[3];
  s->visible_frames = cfg[4];
  s->inv_base = 1.0f / base_texel;
  if (!build_tables(*s, tri_chart)) {
    delete s;
    return nullptr;
  }
  uint32_t len = 1;
  while (len < 4 * s->pages) len <<= 1;
  s->table.init(std::max<uint32_t>(len, 1024));
  s->slot_key.assign(s->pages, {});
  s->last_used.assign(s->pages, 0);
  s->alloc_frame.assign(s->pages, 0);
  for (uint32_t k = s->pages; k > 0; k--) s->free_slots.push_back(k - 1);
  s->slot_info.assign(12 * (size_t)s->pages, 0);
  s->slot_stale.assign(s->pages, 0);
  s->slot_used.assign(s->pages, 0);
  s->pool.assign((size_t)s->pages * TEXELS * s->channels, 0.0f);
  s->invalid_list.assign(s->pages, 0);
  s->candidates.assign(2 * (size_t)std::max<uint32_t>(s->candidate_cap, 1), 0);
  return s;
}

void oracle_vc_destroy(void* h) { delete static_cast<State*>(h); }

// The scene after a move (caller-owned arrays, as at creation).
void oracle_vc_set_scene(void* h, const uint32_t* ranges, const float* blas, const uint32_t* blas_items,
                         const float* blas_cov, const float* nodes, const uint32_t* items, const float* xforms) {
  State& s = *static_cast<State*>(h);
  s.ranges = ranges, s.blas = blas, s.blas_items = blas_items, s.blas_cov = blas_cov, s.nodes = nodes;
  s.items = items, s.xforms = xforms;
}

void oracle_vc_grid(void* h, uint32_t object, uint32_t chart, uint32_t mip, uint32_t* out) {
  grid_of(*static_cast<State*>(h), chart, object, mip, out);
}

// vcache.rs:3940-4050 without a budget: as cnq_vcache_request.
void oracle_vc_request(void* h, const uint32_t* keys, uint32_t n, uint32_t* slots) {
  State& s = *static_cast<State*>(h);
  for (uint32_t i = 0; i < n; i++) {
    const uint32_t* k = keys + 3 * i;
    std::vector<uint32_t> key(k, k + 3);
    slots[i] = NONE;
    auto it = s.map.find(key);
    if (it != s.map.end()) {
      uint32_t slot = it->second.first;
      if (s.last_used[slot] != s.frame) s.used_marks.push_back(slot);
      s.last_used[slot] = s.frame;
      slots[i] = slot;
      continue;
    }
    uint32_t m = k[0], c = k[1] & CHART_MASK, mip = k[1] >> MIP_SHIFT, px = k[2] & 0xffffu, py = k[2] >> 16;
    uint32_t g[2];
    grid_of(s, c, m, mip, g);
    uint32_t pg[2] = {px, py}, listed[2];
    bool empty = false;
    for (int a = 0; a < 2; a++) {
      uint32_t nn = g[a] + 1;
      uint32_t inside = std::min(nn - std::min(pg[a] * PAGE, nn), STRIDE);
      if (inside == 0 || (uint64_t)pg[a] * PAGE >= g[a]) empty = true;
      listed[a] = std::min(inside + 1, STRIDE);
    }
    if (empty || s.free_slots.empty()) continue;
    uint32_t slot = s.free_slots.back();
    s.free_slots.pop_back();
    uint32_t entry = s.table.insert(k, slot);
    if (entry == NONE) {
      s.free_slots.push_back(slot);
      continue;
    }
    s.slot_extent = std::max(s.slot_extent, slot + 1);
    s.map[key] = {slot, entry};
    s.slot_key[slot] = key;
    s.last_used[slot] = s.frame;
    s.alloc_frame[slot] = s.frame;
    uint32_t index = (uint32_t)(s.plan_list.size() / 4);
    s.plan_list.insert(s.plan_list.end(), {k[0], k[1], k[2], slot});
    float b[8];
    page_bounds(s, m, c, mip, px, py, b);
    s.plan_bounds.insert(s.plan_bounds.end(), b, b + 8);
    for (uint32_t l = 0; l < s.n_lights; l++)
      for (uint32_t y = 0; y < listed[1]; y++)
        for (uint32_t x = 0; x < listed[0]; x++) s.plan_jobs.insert(s.plan_jobs.end(), {index, (y * STRIDE + x) | (l << 8)});
    slots[i] = slot;
  }
}

void oracle_vc_evict(void* h, const uint32_t* keys, uint32_t n) {
  State& s = *static_cast<State*>(h);
  for (uint32_t i = 0; i < n; i++) {
    auto it = s.map.find(std::vector<uint32_t>(keys + 3 * i, keys + 3 * i + 3));
    if (it != s.map.end()) evict_slot(s, it->second.first);
  }
  // As the product: the evicted slots' info clears at once.
  for (uint32_t slot : s.cleared) std::fill(&s.slot_info[12 * slot], &s.slot_info[12 * slot + 12], 0u);
  s.cleared.clear();
}

// moved: per object 13 floats (object bits, old lo, old hi, new lo, new hi); scene box 6 floats.
void oracle_vc_set_moved(void* h, const float* moved, uint32_t n, const float* scene_box) {
  State& s = *static_cast<State*>(h);
  BoxF sc{{scene_box[0], scene_box[1], scene_box[2]}, {scene_box[3], scene_box[4], scene_box[5]}};
  set_moved(s, moved, n, sc);
}

void oracle_vc_moved(void* h, float* out, uint32_t* n_floats) {
  State& s = *static_cast<State*>(h);
  if (out) std::memcpy(out, s.moved.data(), s.moved.size() * 4);
  *n_floats = (uint32_t)s.moved.size();
}

// One frame, as cnq_vcache_frame: stats as there.
void oracle_vc_frame(void* h, uint64_t* out_stats) {
  State& s = *static_cast<State*>(h);
  uint32_t plan_pages = (uint32_t)(s.plan_list.size() / 4), plan_jobs = (uint32_t)(s.plan_jobs.size() / 2);
  uint32_t app_jobs = append_jobs_cap(s);
  s.bake_list.assign(4 * (size_t)(plan_pages + s.append_pages + 1), 0);
  s.bake_bounds.assign(8 * (size_t)(plan_pages + s.append_pages + 1), 0.0f);
  s.bake_jobs.assign(2 * ((size_t)plan_jobs + app_jobs + 1), 0);
  std::copy(s.plan_list.begin(), s.plan_list.end(), s.bake_list.begin());
  std::copy(s.plan_bounds.begin(), s.plan_bounds.end(), s.bake_bounds.begin());
  std::copy(s.plan_jobs.begin(), s.plan_jobs.end(), s.bake_jobs.begin());
  std::memset(s.stats, 0, sizeof(s.stats));
  // vc_update (WGSL :3394) and this build's use marks and cleared slots.
  for (uint32_t slot : s.used_marks) s.slot_used[slot] = s.frame;
  for (uint32_t slot : s.cleared) std::fill(&s.slot_info[12 * slot], &s.slot_info[12 * slot + 12], 0u);
  s.cleared.clear();
  if (s.n_moving) {
    // This frame's light poses (vcache.rs:2261-2266), then fresh ledger
    // records for the planned pages (WGSL :4655-4662 `vc_ledger_reset`).
    std::copy(s.lights.begin(), s.lights.end(), s.poses.begin() + (size_t)(s.frame % POSE_RING) * s.n_lights * 16);
    for (uint32_t p = 0; p < plan_pages; p++)
      for (uint32_t j = 0; j < s.n_moving; j++) ledger_reset(s, s.bake_list[4 * p + 3], j);
  }
  bool moving = !s.moved.empty() && s.n_movers > 0;
  if (!moving) s.n_movers = 0;
  if (moving) {
    for (uint32_t slot = 0; slot < s.slot_extent; slot++) vc_stale(s, slot);
    uint64_t threads = (uint64_t)s.candidate_cap * TEXELS;
    for (uint64_t j = 0; j < threads; j++) vc_stale_texel(s, (uint32_t)j);
  }
  // The ledger runs every frame with moving lights (instanced.rs:176-185).
  if (s.n_moving)
    for (uint32_t slot = 0; slot < s.slot_extent; slot++) vc_ledger(s, slot, s.lights_moved);
  uint32_t total = plan_jobs + std::min(s.stats[5], app_jobs);
  for (uint32_t j = 0; j < total; j++) vc_bake_job(s, j, total);
  uint32_t pages = plan_pages + std::min(s.stats[0], s.append_pages);
  for (uint32_t p = 0; p < pages; p++) vc_fill(s, p);
  std::vector<uint32_t> invalid;
  if (s.stats[1] <= s.pages) {
    invalid.assign(s.invalid_list.begin(), s.invalid_list.begin() + s.stats[1]);
  } else {
    for (uint32_t slot = 0; slot < s.slot_extent; slot++)
      if (s.slot_stale[slot] != 0) invalid.push_back(slot);
  }
  std::sort(invalid.begin(), invalid.end());
  uint64_t evicted = 0;
  for (uint32_t slot : invalid) {
    if (s.slot_key[slot].empty() || s.alloc_frame[slot] >= s.frame) continue;
    evict_slot(s, slot);
    evicted++;
  }
  // The product sends evictions to its device at once: their slots' info clears now.
  for (uint32_t slot : s.cleared) std::fill(&s.slot_info[12 * slot], &s.slot_info[12 * slot + 12], 0u);
  s.cleared.clear();
  uint64_t st[10] = {plan_pages, plan_jobs, std::min(s.stats[0], s.append_pages), std::min(s.stats[5], app_jobs),
                     s.stats[4], s.stats[1], s.stats[6], s.stats[3], s.stats[2], evicted};
  if (out_stats) std::memcpy(out_stats, st, sizeof(st));
  const uint32_t ledger_words[10] = {8, 9, 10, 11, 12, 13, 14, 21, 22, 23};
  for (int i = 0; i < 10; i++) s.last_ledger[i] = s.stats[ledger_words[i]];
  s.plan_list.clear(), s.plan_bounds.clear(), s.plan_jobs.clear(), s.used_marks.clear();
  s.moved.clear();
  s.n_movers = 0;
  s.lights_moved = false;
  s.frame++;
}

// WGSL :4666-4687 `light_shift` from pose `a` to pose `b` seen from the
// box (lo3, hi3).
float oracle_vc_light_shift(const float* a, const float* b, const float* box) {
  return light_shift(a, b, Box3{V3{box[0], box[1], box[2]}, V3{box[3], box[4], box[5]}});
}

// Moving lights (vcache.rs:1399-1430): lights[i] of n move through the
// ledger with threshold `epsilon`; before the first frame, with no movers.
int oracle_vc_set_moving(void* h, const uint32_t* lights, uint32_t n, float epsilon) {
  State& s = *static_cast<State*>(h);
  if (s.frame != 1 || n == 0) return -1;
  s.moving_index.assign(s.n_lights, NONE);
  for (uint32_t j = 0; j < n; j++) {
    if (lights[j] >= s.n_lights || s.moving_index[lights[j]] != NONE) return -1;
    s.moving_index[lights[j]] = j;
  }
  s.n_moving = n;
  s.epsilon = epsilon;
  s.ledger.assign((size_t)s.pages * n * LEDGER_WORDS, 0);
  s.ledger_bound.assign((size_t)s.pages * n, 0.0f);
  s.poses.assign((size_t)POSE_RING * s.n_lights * 16, 0.0f);
  return 0;
}

// vcache.rs:2375-2379 `set_light_pose`: light `index` is `light` from the next frame.
int oracle_vc_move_light(void* h, uint32_t index, const float* light) {
  State& s = *static_cast<State*>(h);
  if (index >= s.n_lights || s.n_moving == 0 || s.moving_index[index] == NONE) return -1;
  std::copy(light, light + 16, s.lights.begin() + 16 * (size_t)index);
  s.lights_moved = true;
  return 0;
}

// The last frame's ledger counters: pages re-baked, texels re-baked, pages
// set aside, pages checked, pages re-baking mixed texels, pages held whole,
// jobs asked, pages held on their lattice, lattices that may have grown,
// unbounded texels of held pages (WGSL mover_stats 8-14, 21-23).
void oracle_vc_ledger_stats(void* h, uint64_t* out) {
  State& s = *static_cast<State*>(h);
  for (int i = 0; i < 10; i++) out[i] = s.last_ledger[i];
}

void oracle_vc_ledger_state(void* h, uint32_t* ledger, float* bound) {
  State& s = *static_cast<State*>(h);
  if (ledger) std::memcpy(ledger, s.ledger.data(), s.ledger.size() * 4);
  if (bound) std::memcpy(bound, s.ledger_bound.data(), s.ledger_bound.size() * 4);
}

// instanced.rs:3741-3850 `verify_ledger`: every readable page's texels of
// moving light `light` baked afresh (its ledger untouched) against the
// cache: out = pages, pages past their certified bound (+1e-5), pages past
// ε, pages changed at all, the largest change, the largest bound.
void oracle_vc_verify_ledger(void* h, uint32_t light, float* out) {
  State& s = *static_cast<State*>(h);
  for (int i = 0; i < 6; i++) out[i] = 0.0f;
  if (light >= s.n_lights || s.n_moving == 0 || s.moving_index[light] == NONE) return;
  const uint32_t j = s.moving_index[light];
  std::vector<float> fresh(s.pool);
  std::vector<uint32_t> list, jobs;
  std::vector<uint32_t> slots;
  for (uint32_t slot = 0; slot < s.slot_key.size(); slot++) {
    if (s.slot_key[slot].empty() || s.slot_stale[slot] != 0) continue;
    const uint32_t* k = s.slot_key[slot].data();
    uint32_t index = (uint32_t)(list.size() / 4);
    list.insert(list.end(), {k[0], k[1], k[2], slot});
    for (uint32_t t = 0; t < TEXELS; t++) jobs.insert(jobs.end(), {index, t | (light << 8)});
    slots.push_back(slot);
  }
  std::swap(s.bake_list, list);
  std::swap(s.bake_jobs, jobs);
  const uint32_t total = (uint32_t)(s.bake_jobs.size() / 2);
  for (uint32_t q = 0; q < total; q++) vc_bake_job(s, q, total, false, fresh.data());
  std::swap(s.bake_list, list);
  std::swap(s.bake_jobs, jobs);
  uint32_t pages = 0, violations = 0, over = 0, changed = 0;
  float largest = 0.0f, largest_bound = 0.0f;
  for (uint32_t p = 0; p < slots.size(); p++) {
    const uint32_t slot = slots[p];
    const uint32_t* k = s.slot_key[slot].data();
    uint32_t c = k[1] & CHART_MASK, mip = k[1] >> MIP_SHIFT, g[2];
    grid_of(s, c, k[0], mip, g);
    uint32_t px = (k[2] & 0xffffu) * PAGE, py = (k[2] >> 16) * PAGE;
    float change = 0.0f;
    for (uint32_t t = 0; t < TEXELS; t++) {
      if (px + t % STRIDE >= g[0] || py + t / STRIDE >= g[1]) continue;  // border texels copy inside ones
      size_t at = ((size_t)slot * TEXELS + t) * s.channels + light;
      change = orc::fmax_(change, std::fabs(s.pool[at] - fresh[at]));
    }
    float bound = s.ledger_bound[(size_t)slot * s.n_moving + j];
    pages++;
    if (change > bound + 1e-5f) violations++;
    if (change > s.epsilon) over++;
    if (change > 0.0f) changed++;
    largest = orc::fmax_(largest, change);
    largest_bound = orc::fmax_(largest_bound, bound);
  }
  out[0] = (float)pages, out[1] = (float)violations, out[2] = (float)over, out[3] = (float)changed;
  out[4] = largest, out[5] = largest_bound;
}

// WGSL :3134 vc_find, then the page's texels.
// cnq_vcache_set_lights: lights [first, first + n) replaced, and those
// lights' channels of every resident page baked afresh (inside texels, then
// the border texels from their nearest inside texel, as vc_fill copies).
void oracle_vc_set_lights(void* h, const float* lights, uint32_t first, uint32_t n) {
  State& s = *static_cast<State*>(h);
  std::copy(lights, lights + 16 * (size_t)n, s.lights.begin() + 16 * (size_t)first);
  std::vector<uint32_t> list, jobs;
  for (uint32_t slot = 0; slot < s.slot_key.size(); slot++) {
    if (s.slot_key[slot].empty()) continue;
    const uint32_t* k = s.slot_key[slot].data();
    uint32_t m = k[0], c = k[1] & CHART_MASK, mip = k[1] >> MIP_SHIFT, pp[2] = {k[2] & 0xffffu, k[2] >> 16};
    uint32_t g[2], listed[2];
    grid_of(s, c, m, mip, g);
    for (int ax = 0; ax < 2; ax++) {
      uint32_t nn = g[ax] + 1;
      uint32_t inside = std::min(nn - std::min(pp[ax] * PAGE, nn), STRIDE);
      listed[ax] = std::min(inside + 1, STRIDE);
    }
    uint32_t index = (uint32_t)(list.size() / 4);
    list.insert(list.end(), {k[0], k[1], k[2], slot});
    for (uint32_t l = first; l < first + n; l++)
      for (uint32_t y = 0; y < listed[1]; y++)
        for (uint32_t x = 0; x < listed[0]; x++) jobs.insert(jobs.end(), {index, (y * STRIDE + x) | (l << 8)});
  }
  std::swap(s.bake_list, list);
  std::swap(s.bake_jobs, jobs);
  const uint32_t total = (uint32_t)(s.bake_jobs.size() / 2);
  for (uint32_t j = 0; j < total; j++) vc_bake_job(s, j, total);
  for (size_t p = 0; p < s.bake_list.size() / 4; p++) {
    const uint32_t* e = &s.bake_list[4 * p];
    uint32_t c = e[1] & CHART_MASK, mip = e[1] >> MIP_SHIFT, g[2];
    grid_of(s, c, e[0], mip, g);
    uint32_t px = (e[2] & 0xffffu) * PAGE, py = (e[2] >> 16) * PAGE;
    float* o = &s.pool[(size_t)e[3] * TEXELS * s.channels];
    for (uint32_t k = 0; k < TEXELS; k++) {
      uint32_t tx = px + k % STRIDE, ty = py + k / STRIDE;
      if (tx >= g[0] || ty >= g[1]) {
        uint32_t ix = std::min(tx, g[0] - 1) - px, iy = std::min(ty, g[1] - 1) - py;
        for (uint32_t ch = first; ch < first + n; ch++)
          o[k * s.channels + ch] = o[(iy * STRIDE + ix) * s.channels + ch];
      }
    }
  }
  std::swap(s.bake_list, list);
  std::swap(s.bake_jobs, jobs);
}

void oracle_vc_read(void* h, const uint32_t* keys, uint32_t n, float* texels, uint32_t* slots) {
  State& s = *static_cast<State*>(h);
  size_t row = (size_t)TEXELS * s.channels;
  for (uint32_t i = 0; i < n; i++) {
    uint32_t slot = s.table.find(keys + 3 * i);
    if (slot != NONE && s.slot_stale[slot] != 0 && s.slot_stale[slot] != 3) slot = NONE;
    slots[i] = slot;
    for (size_t q = 0; q < row; q++) texels[i * row + q] = slot == NONE ? 0.0f : s.pool[slot * row + q];
  }
}

void oracle_vc_state(void* h, uint32_t* slot_info, uint32_t* slot_stale, uint32_t* slot_used, float* pool,
                     uint32_t* table, uint32_t* sizes) {
  State& s = *static_cast<State*>(h);
  if (sizes) {
    sizes[0] = s.pages, sizes[1] = s.channels, sizes[2] = s.table.len(), sizes[3] = s.slot_extent, sizes[4] = s.frame;
  }
  if (slot_info) std::memcpy(slot_info, s.slot_info.data(), s.slot_info.size() * 4);
  if (slot_stale) std::memcpy(slot_stale, s.slot_stale.data(), s.slot_stale.size() * 4);
  if (slot_used) std::memcpy(slot_used, s.slot_used.data(), s.slot_used.size() * 4);
  if (pool) std::memcpy(pool, s.pool.data(), s.pool.size() * 4);
  if (table) std::memcpy(table, s.table.e.data(), s.table.e.size() * 4);
}

// Building blocks, for the tests that pin them.
uint32_t oracle_vc_hash(const uint32_t* k) { return hash3(k); }

// The reference's own page-table test (vcache.rs:4308-4349): `steps`
// random inserts and removes with its LCG; returns the number of entries
// deletions moved, or -1 if a live key was not findable or a removed one
// still was.
int64_t oracle_vc_table_check(uint32_t len, uint32_t steps) {
  Table t;
  t.init(len);
  uint64_t rng = 12345;
  auto next = [&]() {
    rng = rng * 6364136223846793005ull + 1442695040888963407ull;
    return (uint32_t)(rng >> 33);
  };
  std::map<std::vector<uint32_t>, uint32_t> live;
  std::vector<std::vector<uint32_t>> order;  // live keys, for picking one
  int64_t moves = 0;
  for (uint32_t step = 0; step < steps; step++) {
    bool remove = live.size() > 400 || (!live.empty() && next() % 3 == 0);
    if (remove) {
      uint32_t pick = next() % (uint32_t)order.size();
      std::vector<uint32_t> k = order[pick];
      order.erase(order.begin() + pick);
      uint32_t entry = live[k];
      live.erase(k);
      t.remove(entry, [&](const uint32_t* mk, uint32_t i) {
        live[std::vector<uint32_t>(mk, mk + 3)] = i;
        moves++;
      });
      if (t.find(k.data()) != NONE) return -1;
    } else {
      std::vector<uint32_t> k = {next() % 64, next() % 4, next() % 64};
      if (live.count(k)) continue;
      uint32_t entry = t.insert(k.data(), step);
      if (entry != NONE) {
        live[k] = entry;
        order.push_back(k);
      }
    }
    if (step % 997 == 0) {
      for (auto& kv : live) {
        const uint32_t* e = &t.e[4 * kv.second];
        if (e[0] != kv.first[0] || e[1] != kv.first[1] || e[2] != kv.first[2]) return -1;
        if (t.find(kv.first.data()) != e[3]) return -1;
      }
    }
  }
  return moves;
}

int oracle_vc_shadow_region(const float* m, const float* l, const float* sc, float* out) {
  BoxF mb{{m[0], m[1], m[2]}, {m[3], m[4], m[5]}}, lb{{l[0], l[1], l[2]}, {l[3], l[4], l[5]}};
  BoxF sb{{sc[0], sc[1], sc[2]}, {sc[3], sc[4], sc[5]}}, r;
  if (!shadow_region(mb, lb, sb, r)) return 0;
  for (int a = 0; a < 3; a++) out[a] = r.lo[a], out[3 + a] = r.hi[a];
  return 1;
}

int oracle_vc_hull_meets(const float* p, const float* l, const float* m) {
  Box3 pb{ld(p), ld(p + 3)}, lb{ld(l), ld(l + 3)}, mb{ld(m), ld(m + 3)};
  return hull_meets(pb, lb, mb) ? 1 : 0;
}

// A texel's place: apex, normal, ok, triangle (8 floats).
void oracle_vc_surface(void* h, uint32_t m, uint32_t c, uint32_t mip, uint32_t tx, uint32_t ty, float* out) {
  State& s = *static_cast<State*>(h);
  uint32_t n[2];
  grid_of(s, c, m, mip, n);
  Surface sf = texel_surface(s, m, c, n, tx, ty);
  out[0] = sf.apex.x, out[1] = sf.apex.y, out[2] = sf.apex.z;
  out[3] = sf.normal.x, out[4] = sf.normal.y, out[5] = sf.normal.z;
  out[6] = sf.ok ? 1.0f : 0.0f;
  std::memcpy(&out[7], &sf.tri, 4);
}

void oracle_vc_world_box(void* h, uint32_t m, const float* lo, const float* hi, float* out) {
  Box3 b = world_box(*static_cast<State*>(h), m, ld(lo), ld(hi));
  out[0] = b.lo.x, out[1] = b.lo.y, out[2] = b.lo.z, out[3] = b.hi.x, out[4] = b.hi.y, out[5] = b.hi.z;
}

void oracle_vc_page_bounds(void* h, uint32_t m, uint32_t c, uint32_t mip, uint32_t px, uint32_t py, float* out) {
  page_bounds(*static_cast<State*>(h), m, c, mip, px, py, out);
}


// Shading from the cache for one view's hits (4 u32 per pixel), as
// cnq_vcache_shade: out (lights x pixels), out_page, requests (4 u32 each,
// in pixel order), counts [requests, pixels without a page].
void oracle_vc_shade(void* h, const float* cam, uint32_t w, uint32_t hgt, const uint32_t* hits, const float* params,
                     float* out, uint32_t* out_page, uint32_t* requests, uint32_t request_cap, uint32_t* counts) {
  State& s = *static_cast<State*>(h);
  float ppt = params[0], angle = params[3];
  uint32_t refine, tile;
  std::memcpy(&refine, &params[1], 4);
  std::memcpy(&tile, &params[2], 4);
  RequestSet rs;
  rs.set.assign(1u << 18, 0u);
  rs.prio.assign(1u << 18, 0xffffffffu);
  uint32_t fell = 0;
  const uint32_t npx = w * hgt;
  const uint32_t sample = (s.frame * 7u) % (tile * tile);
  for (uint32_t p = 0; p < npx; p++) {
    uint32_t x = p % w, y = p / w;
    bool is_sample = x % tile == sample % tile && y % tile == sample / tile;
    uint32_t req[3] = {0, 0, 0}, priority = 0;
    const float* pp = cam + 16;
    float aspect = pp[0] / pp[1];
    float sx = (((float)x + 0.5f + cam[11]) / pp[0] * 2.0f - 1.0f) * pp[2] * aspect;
    float sy = (1.0f - ((float)y + 0.5f + cam[15]) / pp[1] * 2.0f) * pp[2];
    V3 dir = orc::normalize(ld(cam + 4) + ld(cam + 8) * sx + ld(cam + 12) * sy);
    const uint32_t* v = hits + 4 * (size_t)p;
    uint32_t m = v[0], t = v[1];
    if (m == NONE || v[3] != 0) {
      for (uint32_t l = 0; l < s.n_lights; l++) out[(size_t)l * npx + p] = 0.0f;
      out_page[p] = NONE;
      continue;
    }
    float distance;
    std::memcpy(&distance, &v[2], 4);
    const float* tr = s.tris + 12 * t;
    V3 v0 = ld(tr), face = orc::cross(ld(tr + 4) - v0, ld(tr + 8) - v0);
    V3 hn = orc::normalize(face);
    if (s.ranges[4 * m + 3] != NONE) hn = orc::normalize(normal_world(s, m, hn));
    if (orc::dot(hn, dir) > 0.0f) hn = -hn;
    V3 surface = ld(cam) + dir * distance;
    uint32_t c = s.tri_chart[t];
    // WGSL :1780-1793 barycentric, :2339 chart_uv
    V3 po = point_to_object(s, m, surface);
    V3 e1 = ld(tr + 4) - v0, e2 = ld(tr + 8) - v0, d = po - v0;
    float d11 = orc::dot(e1, e1), d12 = orc::dot(e1, e2), d22 = orc::dot(e2, e2);
    float denom = d11 * d22 - d12 * d12;
    float bs = (d22 * orc::dot(d, e1) - d12 * orc::dot(d, e2)) / denom;
    float bu = (d11 * orc::dot(d, e2) - d12 * orc::dot(d, e1)) / denom;
    float b[3] = {1.0f - bs - bu, bs, bu};
    const float* q = &s.tri_uv[6 * t];
    float u = q[0] * b[0] + q[2] * b[1] + q[4] * b[2], vv = q[1] * b[0] + q[3] * b[1] + q[5] * b[2];
    // WGSL :3198-3202 vc_mip
    float ratio = distance * angle * ppt * s.inv_base;
    uint32_t want = 0;
    if (ratio > 1.0f) {
      float lr = std::nearbyint(std::log2(ratio));
      want = lr < 30.0f ? (uint32_t)lr : 30u;
    }
    uint32_t top = top_mip(s, c, m), wanted = std::min(want, top);
    uint32_t mip = wanted, slot = NONE, finer[4][3] = {};
    PageRef r;
    bool found = false;
    for (;;) {
      r = page_ref(s, m, c, u, vv, mip);
      slot = find_shown_slot(s, r.key);
      if (slot != NONE) {
        s.slot_used[slot] = s.frame;
        found = true;
        if (mip > wanted) {
          const uint32_t* pick = finer[0];
          if (refine >= 2 && mip >= wanted + 2) pick = finer[1];
          if (refine >= 3 && mip >= wanted + 3) pick = finer[2];
          if (refine >= 4 && mip >= wanted + 4) pick = finer[3];
          req[0] = pick[0], req[1] = pick[1], req[2] = pick[2];
          priority = 32u - std::min(mip - wanted, 30u);
        }
        break;
      }
      for (int k = 0; k < 3; k++) {
        finer[3][k] = finer[2][k], finer[2][k] = finer[1][k], finer[1][k] = finer[0][k], finer[0][k] = r.key[k];
      }
      if (r.coarsest || mip >= 30) break;
      mip++;
    }
    if (!found) {
      req[0] = finer[0][0], req[1] = finer[0][1], req[2] = finer[0][2];
      priority = 1;
      fell++;
    }
    V3 point = surface + hn * surface_offset(distance);
    const float* nr = &s.normals[9 * t];
    V3 sn = ld(nr) * b[0] + ld(nr + 3) * b[1] + ld(nr + 6) * b[2];
    V3 normal = orc::normalize(normal_world(s, m, orc::dot(sn, sn) > 1e-20f ? sn : face));
    if (orc::dot(normal_world(s, m, face), -dir) < 0.0f) normal = -normal;
    float pt[3] = {point.x, point.y, point.z}, nm[3] = {normal.x, normal.y, normal.z};
    for (uint32_t l = 0; l < s.n_lights; l++) {
      float e;
      if (found && !s.light_ratio) {
        e = bilinear(s, slot, r, l);
      } else {
        float share = found ? bilinear(s, slot, r, l) : 1.0f;
        e = share > 0.0f ? share * oracle_unshadowed_rect(pt, nm, &s.lights[16 * l]) : 0.0f;
      }
      out[(size_t)l * npx + p] = e;
    }
    out_page[p] = found ? mip : (0x80000000u | mip);
    if (is_sample && priority > 0) rs.request(req, priority);
  }
  uint32_t n = (uint32_t)std::min<size_t>(rs.list.size(), request_cap);
  for (uint32_t i = 0; i < n; i++) {
    requests[4 * i] = rs.list[i].k[0], requests[4 * i + 1] = rs.list[i].k[1], requests[4 * i + 2] = rs.list[i].k[2];
    requests[4 * i + 3] = rs.prio[rs.entry[i]];
  }
  counts[0] = (uint32_t)rs.list.size();
  counts[1] = fell;
}

// VirtualCache::plan without sky, bounce or lighting levels (vcache.rs:
// 3704-3810, 3928-4050, `evict_now` :3541 with age 2), as cnq_vcache_plan,
// over `n` requests (4 u32 each). stats: [requests, planned, cones, evicted].
void oracle_vc_plan(void* h, const uint32_t* requests, uint32_t n, uint32_t budget, uint64_t* stats) {
  State& s = *static_cast<State*>(h);
  for (uint32_t k = 0; k < s.pages; k++) s.last_used[k] = std::max(s.last_used[k], s.slot_used[k]);
  std::vector<std::array<uint32_t, 4>> order;
  for (uint32_t i = 0; i < n; i++) {
    const uint32_t* r = requests + 4 * i;
    if (s.map.count(std::vector<uint32_t>(r, r + 3))) continue;
    order.push_back({r[3], r[0], r[1], r[2]});
  }
  uint32_t cap = std::max(budget, TEXELS);
  if (order.size() > cap) {
    std::nth_element(order.begin(), order.begin() + cap, order.end());
    order.resize(cap);
  }
  std::sort(order.begin(), order.end());
  uint64_t spent = 0, planned = 0, evicted = 0;
  bool evicted_now = false;
  for (auto& o : order) {
    uint32_t k[3] = {o[1], o[2], o[3]};
    std::vector<uint32_t> key(k, k + 3);
    if (s.map.count(key)) continue;
    uint32_t m = k[0], c = k[1] & CHART_MASK, mip = k[1] >> MIP_SHIFT, pg[2] = {k[2] & 0xffffu, k[2] >> 16};
    if (m >= s.n_objects || c >= s.n_charts || mip > 30) continue;
    uint32_t g[2], inside[2], listed[2];
    grid_of(s, c, m, mip, g);
    bool empty = false;
    for (int a = 0; a < 2; a++) {
      uint32_t nn = g[a] + 1;
      inside[a] = std::min(nn - std::min(pg[a] * PAGE, nn), STRIDE);
      if (inside[a] == 0 || (uint64_t)pg[a] * PAGE >= g[a]) empty = true;
      listed[a] = std::min(inside[a] + 1, STRIDE);
    }
    if (empty) continue;
    uint64_t cost = (uint64_t)inside[0] * inside[1] * s.n_lights;
    if (spent + cost > budget && !s.plan_list.empty()) break;
    if (s.free_slots.empty() && !evicted_now) {
      evicted_now = true;
      std::vector<std::pair<uint32_t, uint32_t>> stale;
      for (uint32_t sl = 0; sl < s.pages; sl++)
        if (!s.slot_key[sl].empty() && s.last_used[sl] + 2 < s.frame) stale.push_back({s.last_used[sl], sl});
      std::sort(stale.begin(), stale.end());
      for (size_t q = 0; q < stale.size() && q < 2048; q++) {
        evict_slot(s, stale[q].second);
        evicted++;
      }
      for (uint32_t sl : s.cleared) std::fill(&s.slot_info[12 * sl], &s.slot_info[12 * sl + 12], 0u);
      s.cleared.clear();
    }
    if (s.free_slots.empty()) break;
    uint32_t slot = s.free_slots.back();
    s.free_slots.pop_back();
    uint32_t entry = s.table.insert(k, slot);
    if (entry == NONE) {
      s.free_slots.push_back(slot);
      continue;
    }
    s.slot_extent = std::max(s.slot_extent, slot + 1);
    s.map[key] = {slot, entry};
    s.slot_key[slot] = key;
    s.last_used[slot] = s.frame;
    s.alloc_frame[slot] = s.frame;
    uint32_t index = (uint32_t)(s.plan_list.size() / 4);
    s.plan_list.insert(s.plan_list.end(), {k[0], k[1], k[2], slot});
    float b[8];
    page_bounds(s, m, c, mip, pg[0], pg[1], b);
    s.plan_bounds.insert(s.plan_bounds.end(), b, b + 8);
    for (uint32_t l = 0; l < s.n_lights; l++)
      for (uint32_t y = 0; y < listed[1]; y++)
        for (uint32_t x = 0; x < listed[0]; x++) s.plan_jobs.insert(s.plan_jobs.end(), {index, (y * STRIDE + x) | (l << 8)});
    spent += cost;
    planned++;
  }
  stats[0] = n, stats[1] = planned, stats[2] = spent, stats[3] = evicted;
}

}  // extern "C"
