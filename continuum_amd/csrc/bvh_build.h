// Host BVH builder in the reference's flattened layout (see bvh_build.cpp).
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <vector>

namespace cnq {
namespace host {

constexpr uint32_t NONE_U32 = 0xffffffffu;
constexpr float FMAX = 3.40282347e38f;

// The reference's NodeGpu (hierarchy.rs:70-76): 48 bytes.
struct NodeGpu {
  float lo[4];
  float hi[4];
  uint32_t info[4];
};
static_assert(sizeof(NodeGpu) == 48, "NodeGpu is 48 bytes");

struct Box {
  float lo[3], hi[3];
};

// The reference's `Flat` plus the mesh hierarchy (hierarchy.rs:10-26).
struct Flat {
  std::vector<float> tris;  // 12 per triangle (vec4 per vertex)
  std::vector<uint32_t> ranges;
  std::vector<NodeGpu> blas;
  std::vector<uint32_t> blas_items;
  std::vector<float> blas_cov;
  std::vector<NodeGpu> nodes;
  std::vector<uint32_t> items;
};

Box padded(Box b);
Box tri_run_bounds(const float* tris9, size_t n_tris);
void split_tree(std::vector<NodeGpu>& nodes, size_t root_node, const std::vector<Box>& bounds, uint32_t* items,
                size_t count, size_t leaf, uint32_t node_base, uint32_t item_base);
void triangle_hierarchy(const float* mesh9, size_t n, uint32_t root, uint32_t first, std::vector<NodeGpu>& nodes,
                        std::vector<uint32_t>& items, std::vector<float>& cov);
// Over the objects keep[i] != 0 selects (null: every object); indices stay global.
void object_hierarchy(const std::vector<Box>& bounds, std::vector<NodeGpu>& nodes, std::vector<uint32_t>& items,
                      const uint8_t* keep = nullptr);
void flatten(const float* tris9, const uint32_t* counts, uint32_t n_meshes, Flat& out);

// A copy of a geometry placed in the world (hierarchy.rs:262-301
// `Placement`): world = scale * rotation * p + translation, rotation
// row-major and orthonormal; `clear` is one less the opacity of a blended
// object (0: opaque).
struct Placement {
  uint32_t geometry;
  float rotation[9];
  float scale;
  float translation[3];
  float clear;
};

// `flatten_instanced` (hierarchy.rs:464-522) and the object hierarchy over
// the objects' padded world bounds (lib.rs:1984); with `object_prefilter`,
// the object hierarchy's coverage samples (`object_coverage`,
// hierarchy.rs:661-700) appended to blas_cov from row `object_cov` (else
// object_cov = NONE). xforms: 16 floats per object (`Placement::rows`).
// With movable objects (lib.rs:1986-1993), the object hierarchy is
// `object_hierarchy_with_movers` and `tree` describes the movers' subtree.
void flatten_instanced(const float* tris9, const uint32_t* counts, uint32_t n_geometries, const Placement* objects,
                       size_t n_objects, bool object_prefilter, Flat& out, std::vector<float>& xforms,
                       uint32_t& object_cov, const uint32_t* movable = nullptr, size_t n_movable = 0,
                       struct MoverTree* tree = nullptr);

// Where the movable objects' subtree lives (hierarchy.rs:308-325
// `MoverTree`): its root is node 2, its other nodes `nodes` from
// `first_node`, its items from `first_item`; node 1 roots the static objects.
struct MoverTree {
  NodeGpu fixed_root;
  uint32_t first_node = 0, nodes = 0, first_item = 0;
  std::vector<uint32_t> objects;
};

// `object_hierarchy_with_movers` (hierarchy.rs:331-372), `union_root`
// (:374-384) and `mover_subtree` (:388-411).
void object_hierarchy_with_movers(const std::vector<Box>& bounds, const std::vector<uint32_t>& movable,
                                  std::vector<NodeGpu>& nodes, std::vector<uint32_t>& items, MoverTree& tree);
NodeGpu union_root(const NodeGpu& a, const NodeGpu& b);
void mover_subtree(const MoverTree& tree, const std::vector<Box>& mover_bounds, NodeGpu& root,
                   std::vector<NodeGpu>& region, std::vector<uint32_t>& items);
// A placement's padded world bounds of an object-space box (hierarchy.rs:288-301).
Box world_bounds(const Placement& p, const Box& local);

}  // namespace host
}  // namespace cnq
