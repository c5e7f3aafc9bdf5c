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
void flatten_instanced(const float* tris9, const uint32_t* counts, uint32_t n_geometries, const Placement* objects,
                       size_t n_objects, bool object_prefilter, Flat& out, std::vector<float>& xforms,
                       uint32_t& object_cov);

}  // namespace host
}  // namespace cnq
