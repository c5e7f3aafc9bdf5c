// Host side of the view cache's page passes (see vcache_host.h). Every
// float operation follows the reference's host code (glam, f32) in order,
// so page boxes and shadow regions are the reference's bit for bit.
#include "vcache_host.h"

#include <cfloat>
#include <cmath>
#include <cstring>

namespace cnq {
namespace vc {

namespace {

// Rust's f32::min / f32::max (a NaN operand yields the other).
inline float rmin(float a, float b) { return std::isnan(a) ? b : (std::isnan(b) ? a : (b < a ? b : a)); }
inline float rmax(float a, float b) { return std::isnan(a) ? b : (std::isnan(b) ? a : (b > a ? b : a)); }

// Rust's saturating `f as u32`.
inline uint32_t to_u32(float f) {
  if (!(f > 0.0f)) return 0;
  if (f >= 4294967296.0f) return 0xffffffffu;
  return (uint32_t)f;
}

inline float len3(const float* v) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

}  // namespace

uint32_t page_hash(const uint32_t k[3]) {
  uint32_t h = (k[0] * 0x9E3779B1u) ^ (k[1] * 0x85EBCA77u) ^ (k[2] * 0xC2B2AE3Du);
  h ^= h >> 15;
  h *= 0x2C1B3C6Du;
  h ^= h >> 12;
  return h;
}

PageTable::PageTable(uint32_t len) : entries(4 * (size_t)len, 0), is_dirty(len, 0) {
  for (uint32_t i = 0; i < len; i++) entries[4 * (size_t)i] = EMPTY;
}

void PageTable::mark(uint32_t i) {
  if (!is_dirty[i]) {
    is_dirty[i] = 1;
    dirty.push_back(i);
  }
}

std::vector<uint32_t> PageTable::take_dirty() {
  for (uint32_t i : dirty) is_dirty[i] = 0;
  std::vector<uint32_t> out;
  out.swap(dirty);
  return out;
}

uint32_t PageTable::insert(const uint32_t k[3], uint32_t slot) {
  const uint32_t mask = len() - 1;
  uint32_t i = page_hash(k) & mask;
  for (uint32_t probe = 0; probe < MAX_PROBE; probe++) {
    uint32_t* e = &entries[4 * (size_t)i];
    if (e[0] == EMPTY) {
      e[0] = k[0], e[1] = k[1], e[2] = k[2], e[3] = slot;
      mark(i);
      return i;
    }
    i = (i + 1) & mask;
  }
  return EMPTY;
}

void PageTable::remove(uint32_t i, const std::function<void(const uint32_t*, uint32_t)>& moved) {
  const uint32_t mask = len() - 1;
  auto clear = [&](uint32_t at) {
    uint32_t* e = &entries[4 * (size_t)at];
    e[0] = EMPTY, e[1] = 0, e[2] = 0, e[3] = 0;
    mark(at);
  };
  clear(i);
  uint32_t j = i;
  for (;;) {
    j = (j + 1) & mask;
    const uint32_t* e = &entries[4 * (size_t)j];
    if (e[0] == EMPTY) return;
    const uint32_t k[3] = {e[0], e[1], e[2]};
    const uint32_t home = page_hash(k) & mask;
    // It stays if its home lies cyclically in (i, j].
    const bool stays = i <= j ? (i < home && home <= j) : (i < home || home <= j);
    if (!stays) {
      std::memcpy(&entries[4 * (size_t)i], e, 16);
      mark(i);
      moved(k, i);
      clear(j);
      i = j;
    }
  }
}

uint32_t PageTable::find(const uint32_t k[3]) const {
  const uint32_t mask = len() - 1;
  uint32_t i = page_hash(k) & mask;
  for (uint32_t probe = 0; probe < MAX_PROBE; probe++) {
    const uint32_t* e = &entries[4 * (size_t)i];
    if (e[0] == EMPTY) return EMPTY;
    if (e[0] == k[0] && e[1] == k[1] && e[2] == k[2]) return e[3];
    i = (i + 1) & mask;
  }
  return EMPTY;
}

bool chart_tables(const float* tris, uint64_t n_tris, const uint32_t* tri_chart, const float* tri_uv,
                  const float* chart_size, const uint32_t* chart_grid, uint32_t n_charts, ChartTables& t) {
  t = ChartTables();
  t.charts.resize(n_charts);
  uint64_t n_cells = 0;
  for (uint32_t c = 0; c < n_charts; c++) {
    Chart& ch = t.charts[c];
    ch.size[0] = chart_size[2 * c];
    ch.size[1] = chart_size[2 * c + 1];
    ch.grid[0] = chart_grid[2 * c];
    ch.grid[1] = chart_grid[2 * c + 1];
    if (ch.grid[0] == 0 || ch.grid[1] == 0) return false;
    ch.first_cell = (uint32_t)n_cells;
    ch.pad = 0;
    n_cells += (uint64_t)ch.grid[0] * ch.grid[1];
  }
  if (n_cells >= 0xffffffffull) return false;
  std::vector<std::vector<uint32_t>> lists(n_cells);
  for (uint64_t k = 0; k < n_tris; k++) {
    const uint32_t chart = tri_chart[k];
    if (chart >= n_charts) return false;
    const Chart& c = t.charts[chart];
    const float* uv = tri_uv + 6 * k;
    const float gx = (float)c.grid[0], gy = (float)c.grid[1];
    const float lox = std::floor(rmin(rmin(uv[0], uv[2]), uv[4]) * gx + 1e-3f);
    const float loy = std::floor(rmin(rmin(uv[1], uv[3]), uv[5]) * gy + 1e-3f);
    const float hix = std::floor(rmax(rmax(uv[0], uv[2]), uv[4]) * gx - 1e-3f);
    const float hiy = std::floor(rmax(rmax(uv[1], uv[3]), uv[5]) * gy - 1e-3f);
    auto clampc = [](float v, uint32_t n) {
      const uint32_t x = to_u32(rmax(v, 0.0f));
      return x < n - 1 ? x : n - 1;
    };
    const uint32_t x0 = clampc(lox, c.grid[0]), x1 = clampc(rmax(hix, lox), c.grid[0]);
    const uint32_t y0 = clampc(loy, c.grid[1]), y1 = clampc(rmax(hiy, loy), c.grid[1]);
    for (uint32_t y = y0; y <= y1; y++)
      for (uint32_t x = x0; x <= x1; x++) lists[c.first_cell + y * c.grid[0] + x].push_back((uint32_t)k);
  }
  t.cells.resize(2 * n_cells);
  t.cell_bounds.resize(6 * n_cells);
  for (uint64_t cell = 0; cell < n_cells; cell++) {
    t.cells[2 * cell] = (uint32_t)t.cell_tris.size();
    t.cells[2 * cell + 1] = (uint32_t)lists[cell].size();
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (uint32_t tri : lists[cell]) {
      for (int v = 0; v < 3; v++) {
        const float* p = tris + 12 * (size_t)tri + 4 * v;
        for (int a = 0; a < 3; a++) {
          lo[a] = rmin(lo[a], p[a]);
          hi[a] = rmax(hi[a], p[a]);
        }
      }
    }
    std::memcpy(&t.cell_bounds[6 * cell], lo, 12);
    std::memcpy(&t.cell_bounds[6 * cell + 3], hi, 12);
    t.cell_tris.insert(t.cell_tris.end(), lists[cell].begin(), lists[cell].end());
  }
  if (t.cells.empty()) t.cells.assign(2, 0);
  if (t.cell_tris.empty()) t.cell_tris.push_back(0);
  return true;
}

void grid(const Chart& c, float scale, float inv_base, uint32_t mip, uint32_t out[2]) {
  uint32_t bits = (127u - mip) << 23;  // 2^-mip exactly
  float two_mip;
  std::memcpy(&two_mip, &bits, 4);
  const float inv = inv_base * two_mip;
  for (int a = 0; a < 2; a++) {
    const float n = std::ceil((c.size[a] * scale) * inv);
    const uint32_t u = to_u32(n);
    out[a] = u > 1 ? u : 1;
  }
}

void page_bounds(const ChartTables& t, uint32_t c, float scale, float inv_base, uint32_t mip, uint32_t px,
                 uint32_t py, float out[8]) {
  const Chart& ch = t.charts[c];
  uint32_t n[2];
  grid(ch, scale, inv_base, mip, n);
  auto range = [](uint32_t p, uint32_t n, uint32_t g, uint32_t& a, uint32_t& b) {
    const uint32_t x0 = p * PAGE;
    const uint32_t x1 = (p * PAGE + PAGE) < (n - 1) ? (p * PAGE + PAGE) : (n - 1);
    auto cell = [&](uint32_t x) { return to_u32(((float)x + 0.5f) / (float)n * (float)g); };
    const uint32_t c0 = cell(x0), c1 = cell(x1) + 1;
    a = c0 > 0 ? c0 - 1 : 0;
    b = c1 < g - 1 ? c1 : g - 1;
  };
  uint32_t x0, x1, y0, y1;
  range(px, n[0], ch.grid[0], x0, x1);
  range(py, n[1], ch.grid[1], y0, y1);
  float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
  const uint64_t n_cells = t.cell_bounds.size() / 6;
  for (uint32_t y = y0; y <= y1; y++) {
    for (uint32_t x = x0; x <= x1; x++) {
      const uint64_t cell = (uint64_t)ch.first_cell + (uint64_t)y * ch.grid[0] + x;
      if (cell >= n_cells) continue;
      const float* b = &t.cell_bounds[6 * cell];
      for (int a = 0; a < 3; a++) {
        lo[a] = rmin(lo[a], b[a]);
        hi[a] = rmax(hi[a], b[3 + a]);
      }
    }
  }
  if (lo[0] > hi[0]) {
    for (int a = 0; a < 3; a++) lo[a] = hi[a] = 0.0f;
  }
  out[0] = lo[0], out[1] = lo[1], out[2] = lo[2], out[3] = 0.0f;
  out[4] = hi[0], out[5] = hi[1], out[6] = hi[2], out[7] = 0.0f;
}

bool shadow_region(const Box& m, const Box& l, const Box& s, Box& out) {
  float d_lo[3], d_hi[3];
  for (int a = 0; a < 3; a++) {
    d_lo[a] = m.lo[a] - l.hi[a];
    d_hi[a] = m.hi[a] - l.lo[a];
  }
  float t0 = 0.0f, t1 = INFINITY;
  for (int a = 0; a < 3; a++) {
    {
      const float b = d_lo[a], c = s.hi[a] - m.lo[a];
      if (b > 0.0f) t1 = rmin(t1, c / b);
      else if (b < 0.0f) t0 = rmax(t0, c / b);
      else if (c < 0.0f) return false;
    }
    {
      const float b = d_hi[a], c = s.lo[a] - m.hi[a];
      if (b > 0.0f) t0 = rmax(t0, c / b);
      else if (b < 0.0f) t1 = rmin(t1, c / b);
      else if (c > 0.0f) return false;
    }
  }
  if (t0 > t1 * (1.0f + 1e-6f) + 1e-6f) return false;
  if (!std::isfinite(t1)) {
    out = s;
    return true;
  }
  float lo[3], hi[3];
  for (int a = 0; a < 3; a++) {
    lo[a] = rmax(rmin(m.lo[a] + d_lo[a] * t0, m.lo[a] + d_lo[a] * t1), s.lo[a]);
    hi[a] = rmin(rmax(m.hi[a] + d_hi[a] * t0, m.hi[a] + d_hi[a] * t1), s.hi[a]);
  }
  float e = 0.0f;
  for (int a = 0; a < 3; a++) e = a == 0 ? rmax(std::fabs(lo[a]), std::fabs(hi[a])) : rmax(e, rmax(std::fabs(lo[a]), std::fabs(hi[a])));
  const float pad = 1e-4f * rmax(e, 1.0f);
  for (int a = 0; a < 3; a++) {
    out.lo[a] = lo[a] - pad;
    out.hi[a] = hi[a] + pad;
  }
  return true;
}

Box light_box(const float* light) {
  float c[3], r;
  if (light[3] > 1.5f) {
    for (int a = 0; a < 3; a++) c[a] = light[a];
    r = 1.5f * light[4];
  } else {
    for (int a = 0; a < 3; a++) c[a] = light[a] + (light[4 + a] + light[8 + a]) * 0.5f;
    r = 0.5f * (len3(light + 4) + len3(light + 8));
  }
  const float e = rmax(rmax(std::fabs(c[0]), rmax(std::fabs(c[1]), std::fabs(c[2]))), 1.0f);
  r = r * 1.01f + 2e-5f * e + 1e-3f;
  Box b;
  for (int a = 0; a < 3; a++) {
    b.lo[a] = c[a] - r;
    b.hi[a] = c[a] + r;
  }
  return b;
}

std::vector<float> pack_moved(const std::vector<Moved>& moved, const std::vector<Box>& light_boxes,
                              const Box& scene, std::vector<std::pair<uint32_t, uint32_t>>& ids) {
  const Box empty{{FLT_MAX, FLT_MAX, FLT_MAX}, {-FLT_MAX, -FLT_MAX, -FLT_MAX}};
  auto grow = [](Box a, const Box& b) {
    for (int k = 0; k < 3; k++) {
      a.lo[k] = rmin(a.lo[k], b.lo[k]);
      a.hi[k] = rmax(a.hi[k], b.hi[k]);
    }
    return a;
  };
  auto push = [](std::vector<float>& v, const float* p, uint32_t wbits) {
    float w;
    std::memcpy(&w, &wbits, 4);
    v.insert(v.end(), {p[0], p[1], p[2], w});
  };
  const uint32_t lights = (uint32_t)light_boxes.size();
  const uint32_t all = lights >= 32 ? 0xffffffffu : (1u << lights) - 1u;
  std::vector<Box> unions(lights, empty);
  Box every = empty;
  std::vector<float> entries;
  for (const Moved& m : moved) {
    every = grow(grow(every, m.old_box), m.new_box);
    uint32_t id = (uint32_t)ids.size();
    bool found = false;
    for (auto& p : ids)
      if (p.first == m.object) id = p.second, found = true;
    if (!found) ids.emplace_back(m.object, id);
    push(entries, m.old_box.lo, m.object);
    push(entries, m.old_box.hi, id);
    push(entries, m.new_box.lo, all);
    push(entries, m.new_box.hi, 0);
    for (uint32_t l = 0; l < lights; l++) {
      for (const Box* b : {&m.old_box, &m.new_box}) {
        Box r;
        if (!shadow_region(*b, light_boxes[l], scene, r)) r = empty;
        if (r.lo[0] <= r.hi[0]) unions[l] = grow(unions[l], r);
        push(entries, r.lo, 0);
        push(entries, r.hi, 0);
      }
    }
  }
  std::vector<float> out;
  for (const Box& u : unions) {
    push(out, u.lo, 0);
    push(out, u.hi, 0);
  }
  push(out, every.lo, 0);
  push(out, every.hi, 0);
  out.insert(out.end(), entries.begin(), entries.end());
  return out;
}

}  // namespace vc
}  // namespace cnq
