// Host side of the view cache's page passes: the page table the shaders
// probe, the charts' cell tables that place texels on surfaces, page boxes,
// and the moved objects' shadow regions the stale pass tests against.
#pragma once
#include <stdint.h>

#include <functional>
#include <vector>

namespace cnq {
namespace vc {

constexpr uint32_t PAGE = 8;                        // vcache.rs:35
constexpr uint32_t STRIDE = PAGE + 1;               // texels along a page's side, with its border
constexpr uint32_t PAGE_TEXELS = STRIDE * STRIDE;   // vcache.rs:37
constexpr uint32_t EMPTY = 0xffffffffu;             // vcache.rs:38
constexpr uint32_t MAX_PROBE = 64;                  // vcache.rs:82, WGSL vc_find
constexpr uint32_t KEY_MIP_SHIFT = 24;              // WGSL :3215
constexpr uint32_t KEY_CHART_MASK = 0xffffffu;
constexpr uint32_t MAX_LIGHTS = 16;                 // lib.rs:751

// The page-table hash the shaders and the host share (vcache.rs:904-913,
// WGSL vc_hash :3123).
uint32_t page_hash(const uint32_t k[3]);

// Open addressing with linear probing over (key, slot) entries, as the
// shaders read it (vcache.rs:992-1093): no tombstones, deletion moves later
// entries of a probe run back (Knuth's Algorithm R).
struct PageTable {
  std::vector<uint32_t> entries;  // 4 per entry: key, slot
  std::vector<uint32_t> dirty;
  std::vector<uint8_t> is_dirty;

  explicit PageTable(uint32_t len = 0);
  uint32_t len() const { return (uint32_t)(entries.size() / 4); }
  // The entry given to `k` (absent), or EMPTY past MAX_PROBE probes.
  uint32_t insert(const uint32_t k[3], uint32_t slot);
  // Empties entry `i`; `moved(key, new entry)` reports each entry moved back.
  void remove(uint32_t i, const std::function<void(const uint32_t*, uint32_t)>& moved);
  // The slot of `k`, probed as the shaders do, or EMPTY.
  uint32_t find(const uint32_t k[3]) const;
  std::vector<uint32_t> take_dirty();

 private:
  void mark(uint32_t i);
};

// Per chart: size (object units, u and v), grid of cells, first cell.
struct Chart {
  float size[2];
  uint32_t grid[2];
  uint32_t first_cell;
  uint32_t pad;
};

// vcache.rs:688-790 `chart_tables`: each triangle listed in every cell its
// chart coordinates' bounds meet (shrunk by a thousandth of a cell), and
// each cell's object-space box.
struct ChartTables {
  std::vector<Chart> charts;
  std::vector<uint32_t> cells;      // 2 per cell: first entry in cell_tris, count
  std::vector<uint32_t> cell_tris;
  std::vector<float> cell_bounds;   // 6 per cell: lo, hi (lo above hi: no triangle)
};

// tris: 12 floats per triangle (vec4 per vertex); tri_chart: global chart of
// each; tri_uv: 6 per triangle. Returns false on an out-of-range chart.
bool chart_tables(const float* tris, uint64_t n_tris, const uint32_t* tri_chart, const float* tri_uv,
                  const float* chart_size, const uint32_t* chart_grid, uint32_t n_charts, ChartTables& out);

// Texels along u and v of chart `c` on an object of scale `scale` at `mip`,
// as the shaders compute them (vcache.rs:2104-2111, WGSL vc_grid :3181).
void grid(const Chart& c, float scale, float inv_base, uint32_t mip, uint32_t out[2]);

// The object-space box of a page's texels (vcache.rs:2676-2698): the boxes
// of the chart cells its texel centres fall in, one cell wider each way.
void page_bounds(const ChartTables& t, uint32_t c, float scale, float inv_base, uint32_t mip, uint32_t px,
                 uint32_t py, float out[8]);

struct Box {
  float lo[3], hi[3];
};

// hierarchy.rs:418-457: the box of the points m + t (m - l), t >= 0, where
// a cone from a point to light box `l` can meet box `m`, clipped to `s`.
// false if none.
bool shadow_region(const Box& m, const Box& l, const Box& s, Box& out);

// instanced.rs:512-522: the box around what a light's cones can reach near
// it (lights are the reference's LightGpu, 16 floats).
Box light_box(const float* light);

struct Moved {
  uint32_t object;
  Box old_box, new_box;
};

// VirtualCache::set_moved (vcache.rs:2426-2492) for rigid moves: per light
// the union of its shadow regions; the moved boxes' union; then per object
// (old lo | object, old hi | mover id, new lo | lights it updates, new hi |
// 0) and per light its old and new shadow regions. `ids` keeps each moved
// object's stable index among movers.
std::vector<float> pack_moved(const std::vector<Moved>& moved, const std::vector<Box>& light_boxes,
                              const Box& scene, std::vector<std::pair<uint32_t, uint32_t>>& ids);

}  // namespace vc
}  // namespace cnq
