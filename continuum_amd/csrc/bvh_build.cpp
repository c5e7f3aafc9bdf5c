// Host BVH builder: the reference's `flatten` + `mesh_hierarchy`
// (crates/continuum-gpu/src/hierarchy.rs:29-66, :183-253), producing its
// exact layout (NodeGpu 48 B, items, 27 coverage samples per node, ranges).
//
// `split` (hierarchy.rs:128-179) takes the median on the longest centroid
// axis with Rust's `select_nth_unstable_by`; the permutation that leaves
// within each half is the Rust toolchain's choice. Here both halves are
// ordered by (key, item index), key = lo + hi of the padded box on the split
// axis, and a leaf keeps the order of its parent's split axis. Node
// numbering, leaf sizes and tree shape are the reference's (they depend only
// on the item count).
//
// Three index lists, each sorted once by (key on one axis, index), are
// stably partitioned down the tree, so the build is O(n log n) without a
// sort per node.
#include "bvh_build.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace cnq {
namespace host {

static inline float gmax(float a, float b) { return a > b ? a : (b > a ? b : a); }
static inline float gmin(float a, float b) { return a < b ? a : (b < a ? b : a); }

static inline int32_t total_key(float f) {
  int32_t i;
  std::memcpy(&i, &f, 4);
  return i ^ (int32_t)(((uint32_t)(i >> 31)) >> 1);
}

Box padded(Box b) {
  float e = gmax(gmax(gmax(std::fabs(b.lo[0]), std::fabs(b.hi[0])), gmax(std::fabs(b.lo[1]), std::fabs(b.hi[1]))),
                 gmax(std::fabs(b.lo[2]), std::fabs(b.hi[2])));
  float pad = 1e-5f * gmax(e, 1.0f);
  Box o;
  for (int k = 0; k < 3; k++) {
    o.lo[k] = b.lo[k] - pad;
    o.hi[k] = b.hi[k] + pad;
  }
  return o;
}

Box tri_run_bounds(const float* t, size_t n_tris) {
  Box b;
  for (int k = 0; k < 3; k++) {
    b.lo[k] = FMAX;
    b.hi[k] = -FMAX;
  }
  for (size_t i = 0; i < 3 * n_tris; i++)
    for (int k = 0; k < 3; k++) {
      b.lo[k] = gmin(b.lo[k], t[3 * i + k]);
      b.hi[k] = gmax(b.hi[k], t[3 * i + k]);
    }
  return b;
}

namespace {

struct Splitter {
  std::vector<NodeGpu>& nodes;
  const std::vector<Box>& bounds;
  size_t leaf;
  uint32_t node_base, item_base;
  std::vector<uint32_t> lists[3];
  std::vector<uint8_t> left_flag;
  std::vector<uint32_t> tmp;
  std::vector<uint32_t>& out_items;  // final item order by position

  Splitter(std::vector<NodeGpu>& n, const std::vector<Box>& b, size_t leaf_, uint32_t nb, uint32_t ib,
           const uint32_t* items, size_t count, std::vector<uint32_t>& out)
      : nodes(n), bounds(b), leaf(leaf_), node_base(nb), item_base(ib), out_items(out) {
    size_t maxi = 0;
    for (size_t i = 0; i < count; i++) maxi = std::max(maxi, (size_t)items[i] + 1);
    left_flag.assign(maxi, 0);
    tmp.resize(count);
    for (int a = 0; a < 3; a++) {
      std::vector<uint64_t> keyed(count);
      for (size_t i = 0; i < count; i++) {
        const Box& bx = bounds[items[i]];
        uint32_t k = (uint32_t)total_key(bx.lo[a] + bx.hi[a]) ^ 0x80000000u;  // signed order as unsigned
        keyed[i] = ((uint64_t)k << 32) | items[i];
      }
      std::sort(keyed.begin(), keyed.end());
      lists[a].resize(count);
      for (size_t i = 0; i < count; i++) lists[a][i] = (uint32_t)keyed[i];
    }
    out_items.assign(items, items + count);
  }

  // Segment [start, start + count) of every list holds this node's items.
  // `order` is the list giving the node's own item order (its parent's split
  // axis), or -1 for the root (input order).
  void split(size_t node, size_t start, size_t count, int order) {
    float lo[3] = {FMAX, FMAX, FMAX}, hi[3] = {-FMAX, -FMAX, -FMAX};
    float clo[3] = {FMAX, FMAX, FMAX}, chi[3] = {-FMAX, -FMAX, -FMAX};
    const uint32_t* it = order >= 0 ? lists[order].data() + start : out_items.data() + start;
    for (size_t i = 0; i < count; i++) {
      const Box& b = bounds[it[i]];
      for (int k = 0; k < 3; k++) {
        lo[k] = gmin(lo[k], b.lo[k]);
        hi[k] = gmax(hi[k], b.hi[k]);
        float c = (b.lo[k] + b.hi[k]) * 0.5f;
        clo[k] = gmin(clo[k], c);
        chi[k] = gmax(chi[k], c);
      }
    }
    NodeGpu& nd = nodes[node];
    for (int k = 0; k < 3; k++) {
      nd.lo[k] = lo[k];
      nd.hi[k] = hi[k];
    }
    nd.lo[3] = nd.hi[3] = 0.0f;
    if (count <= leaf) {
      if (order >= 0) std::memcpy(out_items.data() + start, it, count * 4);
      nd.info[0] = item_base + (uint32_t)start;
      nd.info[1] = (uint32_t)count;
      nd.info[2] = nd.info[3] = 0;
      return;
    }
    float e0 = chi[0] - clo[0], e1 = chi[1] - clo[1], e2 = chi[2] - clo[2];
    int axis = (e0 >= e1 && e0 >= e2) ? 0 : (e1 >= e2 ? 1 : 2);
    size_t mid = count / 2;
    const uint32_t* la = lists[axis].data() + start;
    for (size_t i = 0; i < count; i++) left_flag[la[i]] = i < mid ? 1 : 0;
    for (int k = 0; k < 3; k++) {
      if (k == axis) continue;
      uint32_t* lk = lists[k].data() + start;
      size_t nl = 0, nr = mid;
      for (size_t i = 0; i < count; i++) {
        uint32_t m = lk[i];
        if (left_flag[m])
          tmp[nl++] = m;
        else
          tmp[nr++] = m;
      }
      std::memcpy(lk, tmp.data(), count * 4);
    }
    size_t left = nodes.size();
    nodes.push_back(NodeGpu{});
    nodes.push_back(NodeGpu{});
    NodeGpu& nd2 = nodes[node];
    nd2.info[0] = node_base + (uint32_t)left;
    nd2.info[1] = nd2.info[2] = nd2.info[3] = 0;
    split(left, start, mid, axis);
    split(left + 1, start + mid, count - mid, axis);
  }
};

void coverage_directions(float dirs[27][3]) {
  for (int n = 0; n < 27; n++) {
    float d[3] = {(float)(n / 9) - 1.0f, (float)(n / 3 % 3) - 1.0f, (float)(n % 3) - 1.0f};
    float len = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    float rcp = 1.0f / len;
    bool ok = std::isfinite(rcp) && rcp > 0.0f;
    for (int k = 0; k < 3; k++) dirs[n][k] = ok ? d[k] * rcp : 0.0f;
  }
}

void tri_cov_add(const float* t, const float dirs[27][3], float* out) {
  float e1[3] = {t[3] - t[0], t[4] - t[1], t[5] - t[2]};
  float e2[3] = {t[6] - t[0], t[7] - t[1], t[8] - t[2]};
  float an[3] = {(e1[1] * e2[2] - e1[2] * e2[1]) * 0.5f, (e1[2] * e2[0] - e1[0] * e2[2]) * 0.5f,
                 (e1[0] * e2[1] - e1[1] * e2[0]) * 0.5f};
  for (int k = 0; k < 27; k++) out[k] += gmax(an[0] * dirs[k][0] + an[1] * dirs[k][1] + an[2] * dirs[k][2], 0.0f);
}

}  // namespace

void split_tree(std::vector<NodeGpu>& nodes, size_t root_node, const std::vector<Box>& bounds, uint32_t* items,
                size_t count, size_t leaf, uint32_t node_base, uint32_t item_base) {
  std::vector<uint32_t> out;
  Splitter s(nodes, bounds, leaf, node_base, item_base, items, count, out);
  s.split(root_node, 0, count, -1);
  std::memcpy(items, out.data(), count * 4);
}

void triangle_hierarchy(const float* mesh, size_t n, uint32_t root, uint32_t first, std::vector<NodeGpu>& nodes,
                        std::vector<uint32_t>& items_out, std::vector<float>& cov) {
  std::vector<Box> bounds(n);
  for (size_t i = 0; i < n; i++) bounds[i] = padded(tri_run_bounds(mesh + 9 * i, 1));
  std::vector<uint32_t> items(n);
  for (size_t i = 0; i < n; i++) items[i] = (uint32_t)i;
  nodes.assign(1, NodeGpu{});
  split_tree(nodes, 0, bounds, items.data(), n, 8, root, first);
  float dirs[27][3];
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

void object_hierarchy(const std::vector<Box>& bounds, std::vector<NodeGpu>& nodes, std::vector<uint32_t>& items,
                      const uint8_t* keep) {
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
  auto diagonal = [&](uint32_t m) {
    float d[3] = {bounds[m].hi[0] - bounds[m].lo[0], bounds[m].hi[1] - bounds[m].lo[1],
                  bounds[m].hi[2] - bounds[m].lo[2]};
    return std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  };
  std::vector<float> sizes(items.size());
  for (size_t i = 0; i < items.size(); i++) sizes[i] = diagonal(items[i]);
  size_t mid = sizes.size() / 2;
  std::nth_element(sizes.begin(), sizes.begin() + mid, sizes.end(),
                   [](float a, float b) { return total_key(a) < total_key(b); });
  float median = sizes[mid];
  std::vector<uint32_t> big, small;
  for (uint32_t m : items) (diagonal(m) > 64.0f * median ? big : small).push_back(m);
  if (big.empty() || small.empty()) {
    split_tree(nodes, 0, bounds, items.data(), items.size(), 2, 0, 0);
    return;
  }
  size_t big_count = big.size();
  items = big;
  items.insert(items.end(), small.begin(), small.end());
  nodes.push_back(NodeGpu{});
  nodes.push_back(NodeGpu{});
  split_tree(nodes, 1, bounds, items.data(), big_count, 2, 0, 0);
  // The second subtree's items sit after the first's.
  {
    std::vector<uint32_t> out;
    Splitter s(nodes, bounds, 2, 0, (uint32_t)big_count, items.data() + big_count, items.size() - big_count, out);
    s.split(2, 0, items.size() - big_count, -1);
    std::memcpy(items.data() + big_count, out.data(), out.size() * 4);
  }
  const NodeGpu &a = nodes[1], &c = nodes[2];
  for (int k = 0; k < 3; k++) {
    nodes[0].lo[k] = gmin(a.lo[k], c.lo[k]);
    nodes[0].hi[k] = gmax(a.hi[k], c.hi[k]);
  }
  nodes[0].lo[3] = nodes[0].hi[3] = 0.0f;
  nodes[0].info[0] = 1;
  nodes[0].info[1] = nodes[0].info[2] = nodes[0].info[3] = 0;
}

void flatten(const float* tris, const uint32_t* counts, uint32_t n_meshes, Flat& f) {
  f = Flat{};
  size_t total = 0;
  for (uint32_t m = 0; m < n_meshes; m++) total += counts[m];
  f.tris.reserve(12 * total);
  std::vector<Box> mesh_boxes;
  size_t off = 0;
  for (uint32_t m = 0; m < n_meshes; m++) {
    const float* mesh = tris + 9 * off;
    size_t n = counts[m];
    uint32_t first = (uint32_t)(f.tris.size() / 12);
    uint32_t root = NONE_U32;
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
    f.ranges.insert(f.ranges.end(), {first, (uint32_t)n, root, NONE_U32});
    for (size_t i = 0; i < 3 * n; i++) f.tris.insert(f.tris.end(), {mesh[3 * i], mesh[3 * i + 1], mesh[3 * i + 2], 0.0f});
    mesh_boxes.push_back(padded(tri_run_bounds(mesh, n)));
    off += n;
  }
  if (f.tris.empty()) f.tris.assign(12, 0.0f);
  if (f.ranges.empty()) f.ranges.insert(f.ranges.end(), {0, 0, NONE_U32, NONE_U32});
  if (f.blas.empty()) {
    f.blas.push_back(NodeGpu{});
    f.blas_items.push_back(0);
    f.blas_cov.assign(27, 0.0f);
  }
  object_hierarchy(mesh_boxes, f.nodes, f.items);
}

namespace {

// M * p with M row-major, summed as glam's Mat3 * Vec3 (column by column).
void mat_vec(const float* m, const float* p, float* out) {
  for (int i = 0; i < 3; i++) out[i] = m[3 * i] * p[0] + m[3 * i + 1] * p[1] + m[3 * i + 2] * p[2];
}

// M^T * p, summed the same way.
void mat_t_vec(const float* m, const float* p, float* out) {
  for (int j = 0; j < 3; j++) out[j] = m[j] * p[0] + m[3 + j] * p[1] + m[6 + j] * p[2];
}

// hierarchy.rs:620-653 `front_area`: the host's bilinear lookup (a + (b - a) t).
float front_area(const float* cov, const float* v) {
  const float a[3] = {std::fabs(v[0]), std::fabs(v[1]), std::fabs(v[2])};
  int axis = 0;
  if (a[1] > a[0] && a[1] >= a[2])
    axis = 1;
  else if (a[2] > a[0] && a[2] > a[1])
    axis = 2;
  float mainv, s, t;
  if (axis == 0) {
    mainv = v[0], s = v[1] / a[0], t = v[2] / a[0];
  } else if (axis == 1) {
    mainv = v[1], s = v[0] / a[1], t = v[2] / a[1];
  } else {
    mainv = v[2], s = v[0] / a[2], t = v[1] / a[2];
  }
  const int m = mainv >= 0.0f ? 1 : -1;
  const int s0 = s < 0.0f ? -1 : 0, t0 = t < 0.0f ? -1 : 0;
  const float fs = s - (float)s0, ft = t - (float)t0;
  float f[4];
  for (int c = 0; c < 4; c++) {
    const int gs = s0 + (c & 1), gt = t0 + (c >> 1);
    int i, j, k;
    if (axis == 0)
      i = m, j = gs, k = gt;
    else if (axis == 1)
      i = gs, j = m, k = gt;
    else
      i = gs, j = gt, k = m;
    f[c] = cov[9 * (i + 1) + 3 * (j + 1) + (k + 1)];
  }
  auto lerp = [](float x, float y, float w) { return x + (y - x) * w; };
  return lerp(lerp(f[0], f[1], fs), lerp(f[2], f[3], fs), ft);
}

}  // namespace

Box world_bounds(const Placement& o, const Box& lb) {
  Box w;
  for (int a = 0; a < 3; a++) {
    w.lo[a] = FMAX;
    w.hi[a] = -FMAX;
  }
  for (int c = 0; c < 8; c++) {
    const float p[3] = {(c & 1) ? lb.hi[0] : lb.lo[0], (c & 2) ? lb.hi[1] : lb.lo[1], (c & 4) ? lb.hi[2] : lb.lo[2]};
    float q[3];
    mat_vec(o.rotation, p, q);
    for (int a = 0; a < 3; a++) {
      const float x = q[a] * o.scale + o.translation[a];
      w.lo[a] = gmin(w.lo[a], x);
      w.hi[a] = gmax(w.hi[a], x);
    }
  }
  return padded(w);
}

NodeGpu union_root(const NodeGpu& a, const NodeGpu& b) {
  NodeGpu r{};
  for (int k = 0; k < 3; k++) {
    r.lo[k] = gmin(a.lo[k], b.lo[k]);
    r.hi[k] = gmax(a.hi[k], b.hi[k]);
  }
  r.info[0] = 1;
  r.info[2] = 1;  // NO_SPLAT
  return r;
}

void mover_subtree(const MoverTree& tree, const std::vector<Box>& bounds, NodeGpu& root, std::vector<NodeGpu>& region,
                   std::vector<uint32_t>& items) {
  std::vector<NodeGpu> local(1, NodeGpu{});
  std::vector<uint32_t> order(bounds.size());
  for (size_t k = 0; k < order.size(); k++) order[k] = (uint32_t)k;
  split_tree(local, 0, bounds, order.data(), order.size(), 2, tree.first_node - 1, tree.first_item);
  for (NodeGpu& n : local) n.info[2] = 1;  // NO_SPLAT: their boxes change
  items.resize(order.size());
  for (size_t k = 0; k < order.size(); k++) items[k] = tree.objects[order[k]];
  root = local[0];
  region.assign(local.begin() + 1, local.end());
}

void object_hierarchy_with_movers(const std::vector<Box>& bounds, const std::vector<uint32_t>& movable,
                                  std::vector<NodeGpu>& nodes, std::vector<uint32_t>& items, MoverTree& tree) {
  std::vector<uint8_t> keep(bounds.size(), 1);
  for (uint32_t m : movable) keep[m] = 0;
  std::vector<NodeGpu> fixed;
  object_hierarchy(bounds, fixed, items, keep.data());
  auto place = [](uint32_t i) { return i == 0 ? 1u : i + 2; };
  tree = MoverTree{};
  tree.first_node = (uint32_t)fixed.size() + 2;
  tree.first_item = (uint32_t)items.size();
  tree.objects = movable;
  std::vector<Box> mover_bounds;
  for (uint32_t m : movable) mover_bounds.push_back(bounds[m]);
  NodeGpu root;
  std::vector<NodeGpu> region;
  std::vector<uint32_t> mover_items;
  mover_subtree(tree, mover_bounds, root, region, mover_items);
  tree.nodes = (uint32_t)region.size();
  nodes.assign(tree.first_node, NodeGpu{});
  for (size_t i = 0; i < fixed.size(); i++) {
    NodeGpu n = fixed[i];
    if (n.info[1] == 0) n.info[0] = place(n.info[0]);
    nodes[place((uint32_t)i)] = n;
  }
  nodes[2] = root;
  nodes.insert(nodes.end(), region.begin(), region.end());
  nodes[0] = union_root(nodes[1], nodes[2]);
  tree.fixed_root = nodes[1];
  items.insert(items.end(), mover_items.begin(), mover_items.end());
}

void flatten_instanced(const float* tris9, const uint32_t* counts, uint32_t n_geometries, const Placement* objects,
                       size_t n_objects, bool object_prefilter, Flat& f, std::vector<float>& xforms,
                       uint32_t& object_cov, const uint32_t* movable, size_t n_movable, MoverTree* tree) {
  flatten(tris9, counts, n_geometries, f);
  const size_t n_blas = f.blas.size();
  // A blended geometry's coverage scales by its first object's opacity.
  std::vector<std::pair<uint32_t, uint32_t>> roots;  // (root, geometry)
  for (uint32_t g = 0; g < n_geometries; g++)
    if (f.ranges[4 * g + 2] != NONE_U32) roots.push_back({f.ranges[4 * g + 2], g});
  std::sort(roots.begin(), roots.end());
  for (size_t k = 0; k < roots.size(); k++) {
    const uint32_t root = roots[k].first, g = roots[k].second;
    const Placement* o = nullptr;
    for (size_t i = 0; i < n_objects && !o; i++)
      if (objects[i].geometry == g) o = &objects[i];
    if (!o || o->clear == 0.0f) continue;
    const uint32_t end = k + 1 < roots.size() ? roots[k + 1].first : (uint32_t)n_blas;
    float opacity;
    if (o->clear > 0.0f) {
      opacity = 1.0f - o->clear;
    } else if (o->clear <= -2.0f) {
      const float x = 0.96f * (-2.0f - o->clear);
      opacity = 1.0f - x * x;
    } else {
      opacity = 1.0f - 0.96f * -o->clear;
    }
    for (size_t c = 27 * (size_t)root; c < 27 * (size_t)end; c++) f.blas_cov[c] *= opacity;
  }
  // Per object: its geometry's range with a transform, and its world bounds.
  const std::vector<uint32_t> geometry_ranges = f.ranges;
  std::vector<Box> local(n_geometries);
  size_t off = 0;
  for (uint32_t g = 0; g < n_geometries; g++) {
    local[g] = tri_run_bounds(tris9 + 9 * off, counts[g]);
    off += counts[g];
  }
  f.ranges.clear();
  xforms.clear();
  std::vector<Box> bounds(n_objects);
  for (size_t k = 0; k < n_objects; k++) {
    const Placement& o = objects[k];
    const uint32_t* gr = &geometry_ranges[4 * (size_t)o.geometry];
    f.ranges.insert(f.ranges.end(), {gr[0], gr[1], gr[2], (uint32_t)k});
    const float* r = o.rotation;
    xforms.insert(xforms.end(), {r[0], r[1], r[2], o.translation[0], r[3], r[4], r[5], o.translation[1], r[6], r[7],
                                 r[8], o.translation[2], o.scale, 1.0f / o.scale, o.clear, 0.0f});
    bounds[k] = world_bounds(o, local[o.geometry]);
  }
  if (f.ranges.empty()) f.ranges.insert(f.ranges.end(), {0, 0, NONE_U32, NONE_U32});
  if (xforms.empty()) xforms.assign(16, 0.0f);
  if (n_movable && tree) {
    object_hierarchy_with_movers(bounds, std::vector<uint32_t>(movable, movable + n_movable), f.nodes, f.items, *tree);
  } else {
    object_hierarchy(bounds, f.nodes, f.items);
  }
  object_cov = NONE_U32;
  if (!object_prefilter) return;
  // Each object's 27 samples from its geometry's root coverage, seen from
  // each sample direction turned into the object, times scale squared.
  float dirs[27][3];
  coverage_directions(dirs);
  std::vector<float> per_object(27 * n_objects, 0.0f);
  for (size_t k = 0; k < n_objects; k++) {
    const uint32_t root = f.ranges[4 * k + 2];
    if (root == NONE_U32) continue;
    const float* cov = f.blas_cov.data() + 27 * (size_t)root;
    const Placement& o = objects[k];
    for (int n = 0; n < 27; n++) {
      if (dirs[n][0] == 0.0f && dirs[n][1] == 0.0f && dirs[n][2] == 0.0f) continue;
      float v[3];
      mat_t_vec(o.rotation, dirs[n], v);
      per_object[27 * k + n] = o.scale * o.scale * front_area(cov, v);
    }
  }
  // Children follow their parents: a reverse sweep sees them first.
  const size_t nn = f.nodes.size();
  std::vector<float> ocov(27 * nn, 0.0f);
  for (size_t n = nn; n-- > 0;) {
    const uint32_t x = f.nodes[n].info[0], count = f.nodes[n].info[1];
    float* out = ocov.data() + 27 * n;
    if (count > 0) {
      for (uint32_t i = x; i < x + count; i++)
        for (int c = 0; c < 27; c++) out[c] += per_object[27 * (size_t)f.items[i] + c];
    } else {
      for (int c = 0; c < 27; c++) out[c] = ocov[27 * (size_t)x + c] + ocov[27 * (size_t)(x + 1) + c];
    }
  }
  object_cov = (uint32_t)(f.blas_cov.size() / 27);
  f.blas_cov.insert(f.blas_cov.end(), ocov.begin(), ocov.end());
}

}  // namespace host
}  // namespace cnq
