// TEST INFRASTRUCTURE ONLY (see oracle_math.h): a CPU restatement of the
// reference's GPU hierarchy builder, crates/continuum-gpu/src/hierarchy.rs
// (`split` :128-179, `triangle_hierarchy` :527-571, `object_hierarchy`
// :198-253, `flatten` :29-66, `refit_triangle_hierarchy` :579-619,
// `padded` :762, `mesh_bounds` :751, `coverage_directions` :736).
//
// The reference splits each node with Rust's `select_nth_unstable_by`, whose
// permutation of the two halves is an implementation detail of the Rust
// toolchain. Here a node's items are ordered by the total order (key, item
// index) and split at the median; each side keeps that order. The tree's
// shape and numbering depend only on the item count, exactly as in the
// reference; which items land in a leaf and in what order follows the
// (key, index) rule.
#include <algorithm>
#include <array>
#include <cstring>
#include <vector>

#include "oracle_math.h"

namespace orc {

constexpr uint32_t BNONE = 0xffffffffu;
constexpr float FMAX = 3.40282347e38f;

struct NodeGpu {
  float lo[4];
  float hi[4];
  uint32_t info[4];
};

struct Box {
  V3 lo, hi;
};

// glam's f32::max/min ignore NaN; none occur here.
static inline float gmax(float a, float b) { return a > b ? a : (b > a ? b : a); }
static inline float gmin(float a, float b) { return a < b ? a : (b < a ? b : a); }
static inline V3 gvmin(V3 a, V3 b) { return V3{gmin(a.x, b.x), gmin(a.y, b.y), gmin(a.z, b.z)}; }
static inline V3 gvmax(V3 a, V3 b) { return V3{gmax(a.x, b.x), gmax(a.y, b.y), gmax(a.z, b.z)}; }

// hierarchy.rs:762-765
static Box padded(Box b) {
  V3 m = gvmax(vabs(b.lo), vabs(b.hi));
  float e = gmax(gmax(m.x, m.y), m.z);
  float pad = 1e-5f * gmax(e, 1.0f);
  return Box{V3{b.lo.x - pad, b.lo.y - pad, b.lo.z - pad}, V3{b.hi.x + pad, b.hi.y + pad, b.hi.z + pad}};
}

// hierarchy.rs:751-756 over one triangle (9 floats) or a run of them.
static Box tri_bounds(const float* t, size_t n_tris) {
  Box b{V3{FMAX, FMAX, FMAX}, V3{-FMAX, -FMAX, -FMAX}};
  for (size_t i = 0; i < 3 * n_tris; i++) {
    V3 p{t[3 * i], t[3 * i + 1], t[3 * i + 2]};
    b.lo = gvmin(b.lo, p);
    b.hi = gvmax(b.hi, p);
  }
  return b;
}

// Rust's f32::total_cmp as an integer key.
static inline int32_t total_key(float f) {
  int32_t i;
  std::memcpy(&i, &f, 4);
  return i ^ (int32_t)(((uint32_t)(i >> 31)) >> 1);
}

struct Builder {
  std::vector<NodeGpu>* nodes;
  const std::vector<Box>* bounds;
  size_t leaf;
  uint32_t node_base, item_base;

  // hierarchy.rs:128-179
  void split(size_t node, uint32_t* items, size_t count, size_t start) {
    const std::vector<Box>& bd = *bounds;
    V3 lo{FMAX, FMAX, FMAX}, hi{-FMAX, -FMAX, -FMAX};
    V3 clo{FMAX, FMAX, FMAX}, chi{-FMAX, -FMAX, -FMAX};
    for (size_t i = 0; i < count; i++) {
      const Box& b = bd[items[i]];
      lo = gvmin(lo, b.lo);
      hi = gvmax(hi, b.hi);
      V3 c = (b.lo + b.hi) * 0.5f;
      clo = gvmin(clo, c);
      chi = gvmax(chi, c);
    }
    NodeGpu& nd = (*nodes)[node];
    nd.lo[0] = lo.x, nd.lo[1] = lo.y, nd.lo[2] = lo.z, nd.lo[3] = 0.0f;
    nd.hi[0] = hi.x, nd.hi[1] = hi.y, nd.hi[2] = hi.z, nd.hi[3] = 0.0f;
    if (count <= leaf) {
      nd.info[0] = item_base + (uint32_t)start;
      nd.info[1] = (uint32_t)count;
      nd.info[2] = nd.info[3] = 0;
      return;
    }
    V3 e = chi - clo;
    int axis = (e.x >= e.y && e.x >= e.z) ? 0 : (e.y >= e.z ? 1 : 2);
    size_t mid = count / 2;
    auto key = [&](uint32_t m) {
      const Box& b = bd[m];
      float a = axis == 0 ? b.lo.x : (axis == 1 ? b.lo.y : b.lo.z);
      float c = axis == 0 ? b.hi.x : (axis == 1 ? b.hi.y : b.hi.z);
      return total_key(a + c);
    };
    std::sort(items, items + count, [&](uint32_t a, uint32_t b) {
      int32_t ka = key(a), kb = key(b);
      return ka != kb ? ka < kb : a < b;
    });
    size_t left = nodes->size();
    nodes->push_back(NodeGpu{});
    nodes->push_back(NodeGpu{});
    NodeGpu& nd2 = (*nodes)[node];
    nd2.info[0] = node_base + (uint32_t)left;
    nd2.info[1] = nd2.info[2] = nd2.info[3] = 0;
    split(left, items, mid, start);
    split(left + 1, items + mid, count - mid, start + mid);
  }
};

static void coverage_directions(V3 dirs[27]) {
  for (int n = 0; n < 27; n++) {
    V3 d{(float)(n / 9) - 1.0f, (float)(n / 3 % 3) - 1.0f, (float)(n % 3) - 1.0f};
    float len = std::sqrt(dot(d, d));
    float rcp = 1.0f / len;
    dirs[n] = (std::isfinite(rcp) && rcp > 0.0f) ? d * rcp : V3{0.0f, 0.0f, 0.0f};
  }
}

static void tri_cov_add(const float* t, const V3 dirs[27], float* out) {
  V3 a{t[0], t[1], t[2]}, b{t[3], t[4], t[5]}, c{t[6], t[7], t[8]};
  V3 an = cross(b - a, c - a) * 0.5f;
  for (int k = 0; k < 27; k++) out[k] += gmax(dot(an, dirs[k]), 0.0f);
}

// hierarchy.rs:527-571. `mesh` holds n triangles as 9 floats each.
static void triangle_hierarchy(const float* mesh, size_t n, uint32_t root, uint32_t first,
                               std::vector<NodeGpu>& nodes, std::vector<uint32_t>& items_out,
                               std::vector<float>& cov) {
  std::vector<Box> bounds(n);
  for (size_t i = 0; i < n; i++) bounds[i] = padded(tri_bounds(mesh + 9 * i, 1));
  std::vector<uint32_t> items(n);
  for (size_t i = 0; i < n; i++) items[i] = (uint32_t)i;
  nodes.assign(1, NodeGpu{});
  Builder b{&nodes, &bounds, 8, root, first};
  b.split(0, items.data(), n, 0);
  V3 dirs[27];
  coverage_directions(dirs);
  cov.assign(nodes.size() * 27, 0.0f);
  for (size_t nn = nodes.size(); nn-- > 0;) {
    uint32_t x = nodes[nn].info[0], count = nodes[nn].info[1];
    float* out = cov.data() + nn * 27;
    if (count > 0) {
      size_t st = x - first;
      for (size_t k = st; k < st + count; k++) tri_cov_add(mesh + 9 * (size_t)items[k], dirs, out);
    } else {
      size_t left = x - root;
      for (int k = 0; k < 27; k++) out[k] = cov[left * 27 + k] + cov[(left + 1) * 27 + k];
    }
  }
  items_out.resize(n);
  for (size_t i = 0; i < n; i++) items_out[i] = first + items[i];
}

// hierarchy.rs:198-253, over the objects `keep` selects (null: all);
// indices stay global.
static void object_hierarchy(const std::vector<Box>& bounds, std::vector<NodeGpu>& nodes,
                             std::vector<uint32_t>& items, const uint8_t* keep = nullptr) {
  items.clear();
  for (size_t i = 0; i < bounds.size(); i++)
    if (!keep || keep[i]) items.push_back((uint32_t)i);
  nodes.assign(1, NodeGpu{});
  if (items.empty()) {
    items.push_back(0);
    nodes[0].info[0] = 0;
    nodes[0].info[1] = 1;
    return;
  }
  auto diagonal = [&](uint32_t m) { return length(bounds[m].hi - bounds[m].lo); };
  std::vector<float> sizes(items.size());
  for (size_t i = 0; i < items.size(); i++) sizes[i] = diagonal(items[i]);
  size_t mid = sizes.size() / 2;
  std::vector<float> sorted = sizes;
  std::sort(sorted.begin(), sorted.end(), [](float a, float b) { return total_key(a) < total_key(b); });
  float median = sorted[mid];
  std::vector<uint32_t> big, small;
  for (uint32_t m : items) (diagonal(m) > 64.0f * median ? big : small).push_back(m);
  if (big.empty() || small.empty()) {
    Builder b{&nodes, &bounds, 2, 0, 0};
    b.split(0, items.data(), items.size(), 0);
    return;
  }
  size_t big_count = big.size();
  items = big;
  items.insert(items.end(), small.begin(), small.end());
  nodes.push_back(NodeGpu{});
  nodes.push_back(NodeGpu{});
  Builder b{&nodes, &bounds, 2, 0, 0};
  b.split(1, items.data(), big_count, 0);
  b.split(2, items.data() + big_count, items.size() - big_count, big_count);
  const NodeGpu &a = nodes[1], &c = nodes[2];
  for (int k = 0; k < 3; k++) {
    nodes[0].lo[k] = gmin(a.lo[k], c.lo[k]);
    nodes[0].hi[k] = gmax(a.hi[k], c.hi[k]);
  }
  nodes[0].lo[3] = nodes[0].hi[3] = 0.0f;
  nodes[0].info[0] = 1;
  nodes[0].info[1] = nodes[0].info[2] = nodes[0].info[3] = 0;
}

// The reference's `Flat` with the object hierarchy.
struct OFlat {
  std::vector<float> tris;
  std::vector<uint32_t> ranges;
  std::vector<NodeGpu> blas;
  std::vector<uint32_t> blas_items;
  std::vector<float> blas_cov;
  std::vector<NodeGpu> nodes;
  std::vector<uint32_t> items;
};

// hierarchy.rs:29-66 `flatten` + :183 `mesh_hierarchy`.
static void flatten(const float* tris, const uint32_t* counts, uint32_t n_meshes, OFlat& f) {
  std::vector<Box> mesh_boxes;
  size_t off = 0;
  for (uint32_t m = 0; m < n_meshes; m++) {
    const float* mesh = tris + 9 * off;
    size_t n = counts[m];
    uint32_t first = (uint32_t)(f.tris.size() / 12);
    uint32_t root = BNONE;
    if (n > 0) {
      root = (uint32_t)f.blas.size();
      std::vector<NodeGpu> nodes;
      std::vector<uint32_t> items;
      std::vector<float> cov;
      triangle_hierarchy(mesh, n, root, first, nodes, items, cov);
      f.blas.insert(f.blas.end(), nodes.begin(), nodes.end());
      f.blas_items.insert(f.blas_items.end(), items.begin(), items.end());
      f.blas_cov.insert(f.blas_cov.end(), cov.begin(), cov.end());
    }
    f.ranges.insert(f.ranges.end(), {first, (uint32_t)n, root, BNONE});
    for (size_t i = 0; i < 3 * n; i++) f.tris.insert(f.tris.end(), {mesh[3 * i], mesh[3 * i + 1], mesh[3 * i + 2], 0.0f});
    mesh_boxes.push_back(padded(tri_bounds(mesh, n)));
    off += n;
  }
  if (f.tris.empty()) f.tris.assign(12, 0.0f);
  if (f.ranges.empty()) f.ranges.insert(f.ranges.end(), {0, 0, BNONE, BNONE});
  if (f.blas.empty()) {
    f.blas.push_back(NodeGpu{});
    f.blas_items.push_back(0);
    f.blas_cov.assign(27, 0.0f);
  }
  object_hierarchy(mesh_boxes, f.nodes, f.items);
}

// hierarchy.rs:262-301 `Placement`: world = scale * R p + t, R by rows.
struct Placed {
  uint32_t geometry;
  V3 row[3];
  float scale;
  V3 t;
  float clear;
  // glam's Mat3 * Vec3: the columns scaled by p's components, summed in order.
  V3 apply(V3 p) const {
    V3 c0{row[0].x, row[1].x, row[2].x}, c1{row[0].y, row[1].y, row[2].y}, c2{row[0].z, row[1].z, row[2].z};
    return c0 * p.x + c1 * p.y + c2 * p.z;
  }
  // rotation.transpose() * d: the rows scaled by d's components.
  V3 apply_t(V3 d) const { return row[0] * d.x + row[1] * d.y + row[2] * d.z; }
};

// hierarchy.rs:624-653 `front_area` (host lerp: a + (b - a) t).
static float host_front_area(const float* cov, V3 v) {
  V3 a = vabs(v);
  int axis = (a.y > a.x && a.y >= a.z) ? 1 : ((a.z > a.x && a.z > a.y) ? 2 : 0);
  float mainv = axis == 0 ? v.x : (axis == 1 ? v.y : v.z);
  float s = axis == 0 ? v.y / a.x : (axis == 1 ? v.x / a.y : v.x / a.z);
  float t = axis == 0 ? v.z / a.x : (axis == 1 ? v.z / a.y : v.y / a.z);
  int m = mainv >= 0.0f ? 1 : -1;
  int s0 = s < 0.0f ? -1 : 0, t0 = t < 0.0f ? -1 : 0;
  float fs = s - (float)s0, ft = t - (float)t0;
  auto sample = [&](int i, int j, int k) { return cov[9 * (i + 1) + 3 * (j + 1) + (k + 1)]; };
  float f[4];
  for (int c = 0; c < 4; c++) {
    int gs = s0 + (c & 1), gt = t0 + (c >> 1);
    f[c] = axis == 0 ? sample(m, gs, gt) : (axis == 1 ? sample(gs, m, gt) : sample(gs, gt, m));
  }
  auto lerp = [](float x, float y, float w) { return x + (y - x) * w; };
  return lerp(lerp(f[0], f[1], fs), lerp(f[2], f[3], fs), ft);
}

// hierarchy.rs:374-384 `union_root`
static NodeGpu union_root(const NodeGpu& a, const NodeGpu& b) {
  NodeGpu r{};
  for (int k = 0; k < 3; k++) {
    r.lo[k] = gmin(a.lo[k], b.lo[k]);
    r.hi[k] = gmax(a.hi[k], b.hi[k]);
  }
  r.info[0] = 1;
  r.info[2] = 1;
  return r;
}

// hierarchy.rs:331-411: the static objects' hierarchy under node 1, the
// movers' subtree under node 2 (its other nodes from first_node).
static void object_hierarchy_with_movers(const std::vector<Box>& bounds, const std::vector<uint32_t>& movable,
                                         std::vector<NodeGpu>& nodes, std::vector<uint32_t>& items,
                                         uint32_t& first_node, uint32_t& first_item) {
  std::vector<uint8_t> keep(bounds.size(), 1);
  for (uint32_t m : movable) keep[m] = 0;
  std::vector<NodeGpu> fixed;
  object_hierarchy(bounds, fixed, items, keep.data());
  first_node = (uint32_t)fixed.size() + 2;
  first_item = (uint32_t)items.size();
  // the movers' subtree over their bounds, in `movable` order
  std::vector<Box> mb;
  for (uint32_t m : movable) mb.push_back(bounds[m]);
  std::vector<NodeGpu> local(1, NodeGpu{});
  std::vector<uint32_t> order(movable.size());
  for (size_t k = 0; k < order.size(); k++) order[k] = (uint32_t)k;
  Builder b{&local, &mb, 2, first_node - 1, first_item};
  b.split(0, order.data(), order.size(), 0);
  for (NodeGpu& n : local) n.info[2] = 1;
  nodes.assign(first_node, NodeGpu{});
  for (size_t i = 0; i < fixed.size(); i++) {
    NodeGpu n = fixed[i];
    if (n.info[1] == 0) n.info[0] = n.info[0] == 0 ? 1 : n.info[0] + 2;
    nodes[i == 0 ? 1 : i + 2] = n;
  }
  nodes[2] = local[0];
  nodes.insert(nodes.end(), local.begin() + 1, local.end());
  nodes[0] = union_root(nodes[1], nodes[2]);
  for (uint32_t k : order) items.push_back(movable[k]);
}

static void flatten_instanced(const float* tris, const uint32_t* counts, uint32_t n_geometries,
                              const std::vector<Placed>& objects, bool object_prefilter, OFlat& f,
                              std::vector<float>& xforms, uint32_t& object_cov,
                              const std::vector<uint32_t>& movable = {}) {
  flatten(tris, counts, n_geometries, f);
  // Blended geometries: coverage times the first object's opacity.
  std::vector<std::pair<uint32_t, uint32_t>> roots;
  for (uint32_t g = 0; g < n_geometries; g++)
    if (f.ranges[4 * g + 2] != BNONE) roots.push_back({f.ranges[4 * g + 2], g});
  std::sort(roots.begin(), roots.end());
  for (size_t k = 0; k < roots.size(); k++) {
    auto it = std::find_if(objects.begin(), objects.end(), [&](const Placed& o) { return o.geometry == roots[k].second; });
    if (it == objects.end() || it->clear == 0.0f) continue;
    uint32_t end = k + 1 < roots.size() ? roots[k + 1].first : (uint32_t)f.blas.size();
    float opacity = it->clear > 0.0f ? 1.0f - it->clear
                    : it->clear <= -2.0f ? 1.0f - (0.96f * (-2.0f - it->clear)) * (0.96f * (-2.0f - it->clear))
                                         : 1.0f - 0.96f * -it->clear;
    for (size_t c = 27 * (size_t)roots[k].first; c < 27 * (size_t)end; c++) f.blas_cov[c] *= opacity;
  }
  std::vector<uint32_t> granges = f.ranges;
  std::vector<Box> local;
  size_t off = 0;
  for (uint32_t g = 0; g < n_geometries; g++) {
    local.push_back(tri_bounds(tris + 9 * off, counts[g]));
    off += counts[g];
  }
  f.ranges.clear();
  std::vector<Box> bounds;
  for (size_t k = 0; k < objects.size(); k++) {
    const Placed& o = objects[k];
    f.ranges.insert(f.ranges.end(), {granges[4 * o.geometry], granges[4 * o.geometry + 1], granges[4 * o.geometry + 2],
                                     (uint32_t)k});
    xforms.insert(xforms.end(), {o.row[0].x, o.row[0].y, o.row[0].z, o.t.x, o.row[1].x, o.row[1].y, o.row[1].z, o.t.y,
                                 o.row[2].x, o.row[2].y, o.row[2].z, o.t.z, o.scale, 1.0f / o.scale, o.clear, 0.0f});
    Box w{V3{FMAX, FMAX, FMAX}, V3{-FMAX, -FMAX, -FMAX}};
    const Box& lb = local[o.geometry];
    for (int c = 0; c < 8; c++) {
      V3 p{(c & 1) ? lb.hi.x : lb.lo.x, (c & 2) ? lb.hi.y : lb.lo.y, (c & 4) ? lb.hi.z : lb.lo.z};
      V3 q = o.apply(p) * o.scale + o.t;
      w.lo = gvmin(w.lo, q);
      w.hi = gvmax(w.hi, q);
    }
    bounds.push_back(padded(w));
  }
  if (f.ranges.empty()) f.ranges.insert(f.ranges.end(), {0, 0, BNONE, BNONE});
  if (xforms.empty()) xforms.assign(16, 0.0f);
  if (movable.empty()) {
    object_hierarchy(bounds, f.nodes, f.items);
  } else {
    uint32_t fn, fi;
    object_hierarchy_with_movers(bounds, movable, f.nodes, f.items, fn, fi);
  }
  object_cov = BNONE;
  if (!object_prefilter) return;
  V3 dirs[27];
  coverage_directions(dirs);
  std::vector<std::array<float, 27>> per(objects.size());
  for (size_t k = 0; k < objects.size(); k++) {
    per[k].fill(0.0f);
    uint32_t root = f.ranges[4 * k + 2];
    if (root == BNONE) continue;
    for (int n = 0; n < 27; n++) {
      if (dirs[n].x == 0.0f && dirs[n].y == 0.0f && dirs[n].z == 0.0f) continue;
      per[k][n] = objects[k].scale * objects[k].scale *
                  host_front_area(f.blas_cov.data() + 27 * (size_t)root, objects[k].apply_t(dirs[n]));
    }
  }
  std::vector<float> cov(27 * f.nodes.size(), 0.0f);
  for (size_t n = f.nodes.size(); n-- > 0;) {
    uint32_t x = f.nodes[n].info[0], count = f.nodes[n].info[1];
    if (count > 0) {
      for (uint32_t i = x; i < x + count; i++)
        for (int c = 0; c < 27; c++) cov[27 * n + c] += per[f.items[i]][c];
    } else {
      for (int c = 0; c < 27; c++) cov[27 * n + c] = cov[27 * (size_t)x + c] + cov[27 * (size_t)(x + 1) + c];
    }
  }
  object_cov = (uint32_t)(f.blas_cov.size() / 27);
  f.blas_cov.insert(f.blas_cov.end(), cov.begin(), cov.end());
}

}  // namespace orc

using namespace orc;

extern "C" {

// flatten (hierarchy.rs:29-66) + mesh_hierarchy (:183-196) for meshes given
// as 9 floats per triangle, `counts[m]` triangles each, in order. Two calls:
// with null outputs it returns the sizes in `sizes` = [n_tris, n_ranges,
// n_blas, n_blas_items, n_nodes, n_items]; then it fills them.
int oracle_flatten(const float* tris, const uint32_t* counts, uint32_t n_meshes, uint64_t* sizes, float* out_tris,
                   uint32_t* out_ranges, float* out_blas, uint32_t* out_blas_items, float* out_blas_cov,
                   float* out_nodes, uint32_t* out_items) {
  OFlat f;
  flatten(tris, counts, n_meshes, f);
  if (!out_tris) {
    sizes[0] = f.tris.size() / 12;
    sizes[1] = f.ranges.size() / 4;
    sizes[2] = f.blas.size();
    sizes[3] = f.blas_items.size();
    sizes[4] = f.nodes.size();
    sizes[5] = f.items.size();
    return 0;
  }
  std::memcpy(out_tris, f.tris.data(), f.tris.size() * 4);
  std::memcpy(out_ranges, f.ranges.data(), f.ranges.size() * 4);
  std::memcpy(out_blas, f.blas.data(), f.blas.size() * sizeof(NodeGpu));
  std::memcpy(out_blas_items, f.blas_items.data(), f.blas_items.size() * 4);
  std::memcpy(out_blas_cov, f.blas_cov.data(), f.blas_cov.size() * 4);
  std::memcpy(out_nodes, f.nodes.data(), f.nodes.size() * sizeof(NodeGpu));
  std::memcpy(out_items, f.items.data(), f.items.size() * 4);
  return 0;
}

// flatten_instanced (hierarchy.rs:464-522) + object_hierarchy over the
// objects' world bounds (lib.rs:1984) + object_coverage (hierarchy.rs:661)
// when object_prefilter. Placements: 15 words each (geometry u32, rotation
// row-major 9 f32, scale, translation 3, clear). Two calls: with null
// outputs, sizes = [n_tris, n_ranges, n_blas, n_blas_items, n_nodes,
// n_items, n_cov_rows, object_cov]; then it fills them (xforms: 16 per object).
int oracle_flatten_instanced(const float* tris, const uint32_t* counts, uint32_t n_geometries, const uint32_t* placements,
                             uint64_t n_objects, const uint32_t* movable, uint32_t n_movable, int object_prefilter,
                             uint64_t* sizes, float* out_tris,
                             uint32_t* out_ranges, float* out_blas, uint32_t* out_blas_items, float* out_blas_cov,
                             float* out_nodes, uint32_t* out_items, float* out_xforms) {
  std::vector<Placed> objs(n_objects);
  for (uint64_t k = 0; k < n_objects; k++) {
    const uint32_t* w = placements + 15 * k;
    objs[k].geometry = w[0];
    float fw[14];
    std::memcpy(fw, w + 1, 56);
    objs[k].row[0] = V3{fw[0], fw[1], fw[2]};
    objs[k].row[1] = V3{fw[3], fw[4], fw[5]};
    objs[k].row[2] = V3{fw[6], fw[7], fw[8]};
    objs[k].scale = fw[9];
    objs[k].t = V3{fw[10], fw[11], fw[12]};
    objs[k].clear = fw[13];
  }
  OFlat f;
  std::vector<float> xforms;
  uint32_t object_cov = BNONE;
  flatten_instanced(tris, counts, n_geometries, objs, object_prefilter != 0, f, xforms, object_cov,
                    std::vector<uint32_t>(movable, movable + n_movable));
  if (!out_tris) {
    sizes[0] = f.tris.size() / 12;
    sizes[1] = f.ranges.size() / 4;
    sizes[2] = f.blas.size();
    sizes[3] = f.blas_items.size();
    sizes[4] = f.nodes.size();
    sizes[5] = f.items.size();
    sizes[6] = f.blas_cov.size() / 27;
    sizes[7] = object_cov;
    return 0;
  }
  std::memcpy(out_tris, f.tris.data(), f.tris.size() * 4);
  std::memcpy(out_ranges, f.ranges.data(), f.ranges.size() * 4);
  std::memcpy(out_blas, f.blas.data(), f.blas.size() * sizeof(NodeGpu));
  std::memcpy(out_blas_items, f.blas_items.data(), f.blas_items.size() * 4);
  std::memcpy(out_blas_cov, f.blas_cov.data(), f.blas_cov.size() * 4);
  std::memcpy(out_nodes, f.nodes.data(), f.nodes.size() * sizeof(NodeGpu));
  std::memcpy(out_items, f.items.data(), f.items.size() * 4);
  std::memcpy(out_xforms, xforms.data(), xforms.size() * 4);
  return 0;
}

// refit_triangle_hierarchy (hierarchy.rs:579-619) over a whole flattened
// scene: for every mesh with a hierarchy, recompute its nodes' boxes and
// coverage from `tris` (3 vec4 per triangle, the new pose) in place.
void oracle_refit(const float* tris, const uint32_t* ranges, uint32_t n_ranges, float* blas_f,
                  const uint32_t* blas_items, float* blas_cov, uint32_t n_blas) {
  NodeGpu* blas = reinterpret_cast<NodeGpu*>(blas_f);
  V3 dirs[27];
  coverage_directions(dirs);
  // Each mesh's nodes run from its root to the next root (or the end).
  std::vector<uint32_t> roots;
  for (uint32_t r = 0; r < n_ranges; r++)
    if (ranges[4 * r + 2] != BNONE) roots.push_back(ranges[4 * r + 2]);
  std::sort(roots.begin(), roots.end());
  roots.erase(std::unique(roots.begin(), roots.end()), roots.end());
  for (size_t ri = 0; ri < roots.size(); ri++) {
    uint32_t root = roots[ri];
    uint32_t end = ri + 1 < roots.size() ? roots[ri + 1] : n_blas;
    for (uint32_t nn = end; nn-- > root;) {
      uint32_t x = blas[nn].info[0], count = blas[nn].info[1];
      V3 lo{FMAX, FMAX, FMAX}, hi{-FMAX, -FMAX, -FMAX};
      float* out = blas_cov + (size_t)nn * 27;
      for (int k = 0; k < 27; k++) out[k] = 0.0f;
      if (count > 0) {
        for (uint32_t k = x; k < x + count; k++) {
          uint32_t g = blas_items[k];
          const float* t = tris + 12 * (size_t)g;
          float t9[9] = {t[0], t[1], t[2], t[4], t[5], t[6], t[8], t[9], t[10]};
          Box b = padded(tri_bounds(t9, 1));
          lo = gvmin(lo, b.lo);
          hi = gvmax(hi, b.hi);
          tri_cov_add(t9, dirs, out);
        }
      } else {
        for (uint32_t c = x; c <= x + 1; c++) {
          lo = gvmin(lo, V3{blas[c].lo[0], blas[c].lo[1], blas[c].lo[2]});
          hi = gvmax(hi, V3{blas[c].hi[0], blas[c].hi[1], blas[c].hi[2]});
        }
        for (int k = 0; k < 27; k++) out[k] = blas_cov[(size_t)x * 27 + k] + blas_cov[(size_t)(x + 1) * 27 + k];
      }
      blas[nn].lo[0] = lo.x, blas[nn].lo[1] = lo.y, blas[nn].lo[2] = lo.z, blas[nn].lo[3] = 0.0f;
      blas[nn].hi[0] = hi.x, blas[nn].hi[1] = hi.y, blas[nn].hi[2] = hi.z, blas[nn].hi[3] = 0.0f;
    }
  }
}

// mesh_hierarchy_of (hierarchy.rs:188-196): the object hierarchy over the
// padded bounds of the meshes `keep` selects (null: all), mesh indices
// global. Meshes as in oracle_flatten. Two calls: with null outputs it
// returns [n_nodes, n_items] in `sizes`; then it fills them.
int oracle_mesh_hierarchy(const float* tris, const uint32_t* counts, uint32_t n_meshes, const uint8_t* keep,
                          uint64_t* sizes, float* out_nodes, uint32_t* out_items) {
  std::vector<Box> boxes;
  size_t off = 0;
  for (uint32_t m = 0; m < n_meshes; m++) {
    boxes.push_back(padded(tri_bounds(tris + 9 * off, counts[m])));
    off += counts[m];
  }
  std::vector<NodeGpu> nodes;
  std::vector<uint32_t> items;
  object_hierarchy(boxes, nodes, items, keep);
  if (!out_nodes) {
    sizes[0] = nodes.size();
    sizes[1] = items.size();
    return 0;
  }
  std::memcpy(out_nodes, nodes.data(), nodes.size() * sizeof(NodeGpu));
  std::memcpy(out_items, items.data(), items.size() * 4);
  return 0;
}

}  // extern "C"
