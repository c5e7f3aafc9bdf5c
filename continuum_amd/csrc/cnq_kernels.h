// Device-resident scene layout and kernel arguments shared by the HIP
// kernels and the host library.
//
// HBM layout (per scene, resident across calls):
//   blas      NodeC[]   32 B per triangle-hierarchy node, the reference's node
//                       numbering (hierarchy.rs `NodeGpu`, 48 B) compacted:
//                       box, then (a, b) = (left child, 0) for an interior
//                       node or (first item, count) for a leaf.
//   tri_leaf  float[9]  36 B per triangle, in `blas_items` order, so a leaf's
//                       triangles are contiguous and fetched without the
//                       item indirection.
//   tris      float[12] 48 B per triangle in the reference's order (vec4 per
//                       vertex), for walks without triangle hierarchies and
//                       for camera hits' triangle indices.
//   nodes     NodeC[]   object hierarchy; b's top bit marks a node without
//                       prefiltered coverage (`NodeGpu::info.z` = NO_SPLAT).
//   items, ranges (4 u32 per object), xforms (16 f32 per transform),
//   blas_cov (27 f32 per node), blas_items: as the reference lays them out.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cnq_device.h"

namespace cnq {

constexpr uint32_t NODE_NO_SPLAT = 0x80000000u;

struct NodeC {
  float lo[3];
  float hi[3];
  uint32_t a;
  uint32_t b;
};
static_assert(sizeof(NodeC) == 32, "NodeC is 32 bytes");

struct LightRect {
  float4 corner, edge_u, edge_v, radiance;
};

struct SceneArgs {
  const float* __restrict__ tris;
  const float* __restrict__ tri_leaf;
  const uint32_t* __restrict__ blas_items;
  const uint32_t* __restrict__ ranges;
  const NodeC* __restrict__ blas;
  const float* __restrict__ blas_cov;
  const NodeC* __restrict__ nodes;
  const uint32_t* __restrict__ items;
  const float* __restrict__ xforms;
  uint32_t n_objects;
};

// The cone kernel's scalar parameters (its arrays are separate, noalias
// kernel arguments).
struct QueryParams {
  uint32_t n_objects, hierarchy, light_cone_cap, wide_cone_cap, profile_stage, object_cov;
  float prefilter, wide_extent;
  uint32_t n_views, n_items, item_begin, tile_w_log2;
};

// One view of a batch: a w x h grid of receivers (a camera view, or a list
// as n x 1) lit by lights [light_off, light_off + n_lights). Its work items
// are (light, tile) pairs from item_off, tiles_x tiles per row and `tiles`
// per light; its query (light l, pixel p) is stored at out_off + l * w * h + p.
// Its camera pass's 8 x 8 tiles are numbered from ptile_off.
struct ViewDesc {
  uint64_t recv_off;
  uint64_t out_off;
  uint32_t w, h;
  uint32_t light_off, n_lights;
  uint32_t tiles_x, tiles;
  uint32_t item_off, ptile_off;
};

struct ItemArgs {
  SceneArgs s;
  const float* __restrict__ receivers;
  const float* __restrict__ lights;
  const ViewDesc* __restrict__ views;  // device copy (used when p.n_views > 1)
  ViewDesc v0;                         // the view when p.n_views == 1
  float* __restrict__ out_visible;
  uint64_t* __restrict__ out_full;
  uint64_t* __restrict__ out_touched;
  float* __restrict__ out_cov;
  uint32_t* __restrict__ counters;
  uint64_t* __restrict__ work;
  uint32_t* __restrict__ work_counter;  // persistent mode: the next item (zeroed before launch)
  // Items [item_begin, n_items): wave w (direct mode) or the w-th grab
  // (persistent mode) takes item item_begin + order[w] when `order` is
  // given (a permutation of the launch's items), and writes the item's time
  // (wall clock, 10 ns units) to item_cost[item - item_begin] when given.
  const uint32_t* __restrict__ order;
  uint32_t* __restrict__ item_cost;
  // Nullable: per query q < n_flags of this launch's outputs, bit 0 the walk
  // stopped at its cap, bit 1 it went deeper than the reference's 48-entry
  // stack, bit 2 this build truncated it.
  uint8_t* __restrict__ query_flags;
  uint64_t n_flags;
  QueryParams p;
};

// The camera pass over one or many views: view v's pixels are the receivers
// (and hits) from views[v].recv_off, its camera is 20 floats at
// cameras + 20 v, and the lights it may see are its own range. One block per
// 8 x 8 tile; views[v].item_off numbers the tiles (tiles_x per row).
struct PrimaryArgs {
  SceneArgs s;
  const float* __restrict__ cameras;  // eye, forward, right, up, params (20 floats per view)
  const float* __restrict__ lights;
  const ViewDesc* __restrict__ views;  // device copy (used when n_views > 1)
  ViewDesc v0;
  uint32_t n_views, n_tiles;
  uint32_t hierarchy, nearest_first;
  uint32_t* __restrict__ hits;   // 4 per pixel
  float* __restrict__ receivers; // 8 per pixel
  uint32_t* __restrict__ counters;  // nullable: [stack overflow]
  // Nullable: per triangle (the scene's order) its corners' shading normals,
  // 9 floats; the receivers then take the interpolated shading normal.
  const float* __restrict__ normals = nullptr;
};

struct RefitArgs {
  const float* __restrict__ tris;        // new pose, 12 floats per triangle, reference order
  const uint32_t* __restrict__ blas_items;
  NodeC* blas;
  float* blas_cov;
  float* tri_leaf;
  const uint32_t* __restrict__ leaves;   // the leaves to refit
  uint32_t n_leaves;
};

// The lighting cache: a lightmap's texels in blocks of 8 x 8 (64
// consecutive receivers), re-baked when meshes move.
//   planes      float4[16][n_blocks]  plane-major, so a wave of consecutive
//                                     blocks reads each plane coalesced
//   block_slot  u32 per block: its slot in static_cov, or NONE
//   slot_block  u32 per slot: the block
//   static_cov  f32[slot][bin][texel]: the static meshes' coverage, bin-major
//               within a block so a wave's 64 texels read each bin coalesced
constexpr uint32_t BLOCK_PLANES = 16;  // lightmap.rs:83
constexpr uint32_t MAX_SWEPT = 16;     // live.rs:220
constexpr uint32_t MAX_MOVING = 16;    // live.rs:175

struct SweptBoxes {
  float4 lo[MAX_SWEPT], hi[MAX_SWEPT];
  uint32_t n;
};

struct MovingMeshes {
  float4 lo[MAX_MOVING], hi[MAX_MOVING];  // padded bounds
  uint32_t mesh[MAX_MOVING];
  uint32_t n;
};

struct CacheArgs {
  SceneArgs s;                     // the whole scene (walks of moving meshes)
  const NodeC* __restrict__ static_nodes;  // the object hierarchy of the static meshes
  const uint32_t* __restrict__ static_items;
  const float* __restrict__ texels;        // 8 floats per texel (point, normal)
  const float4* __restrict__ planes;
  const uint32_t* __restrict__ block_slot;
  const uint32_t* __restrict__ slot_block;
  float* __restrict__ static_cov;
  uint32_t* __restrict__ dirty;            // marked blocks
  uint32_t* __restrict__ n_dirty;          // their count (zeroed before marking)
  float* __restrict__ irradiance;          // per texel
  LightRect light;
  QueryParams p;
  uint32_t n_blocks, n_slots;
};

// The sky walk's options: the reference's params.w bits 1 (`sky_max`: meshes
// combine by the maximum, a diagnostic) and 2 (`sky_horizon_cull` off: walk
// geometry below the receiver's horizon too), and its sky prefilter scale
// (light.objects.y, `sky_prefilter`).
struct SkyParams {
  uint32_t max_combine, horizon_cull;
  float prefilter_scale;
};

// The sky over a w x h grid of receivers (a view, or a list as n x 1):
// work items are (face, tile), face-major, a tile being tile_w x 64 / tile_w
// receivers (64 x 1 for a list); face f of receiver i = y * w + x goes to
// out_faces[f * n + i] (n = w * h; its full bins to out_full, when given),
// then out[i] sums the six.
struct SkyArgs {
  SceneArgs s;
  const float* __restrict__ receivers;
  float* __restrict__ out;
  float* __restrict__ out_faces;
  uint64_t* __restrict__ out_full;
  uint64_t* __restrict__ work;
  // Wave w takes item order[w] when given; item times go to item_cost
  // when given (as the cone kernel's direct mode).
  const uint32_t* __restrict__ order;
  uint32_t* __restrict__ item_cost;
  uint32_t n;
  uint32_t w = 0, h = 1, tile_w_log2 = 6;  // the grid (w 0: a list of n) and its tiles
  QueryParams p;
  SkyParams sky;
};

// `waves`: waves per SIMD in flight (2, the most the walk's registers and
// LDS allow, or 1: the launch reserves LDS so a CU holds four waves).
hipError_t launch_cone_items(const ItemArgs& a, bool count_work, uint32_t persistent_blocks, hipStream_t stream,
                             uint32_t waves = 2);
hipError_t launch_sky(const SkyArgs& a, bool count_work, hipStream_t stream);
// The sky launch's work items (six faces times its tiles).
uint64_t sky_items(const SkyArgs& a);
// The next launch's order from this one's item times: longest first, by
// powers of two, when the times resemble the last launch's (`last`: each
// item's last power of two, 0xff for none), else natural order (one block;
// the order within a power of two is arbitrary).
hipError_t launch_item_order(const uint32_t* cost, uint8_t* last, uint32_t* order, uint32_t n, hipStream_t stream);
// `depth`: the rays' stack, 64 entries (default) or 32.
hipError_t launch_primary(const PrimaryArgs& a, hipStream_t stream, uint32_t depth = 64);
// Leaves first, then the internal nodes one height at a time: `inner` lists
// them by height, level_off[h]..level_off[h + 1] (n_levels + 1 entries, on
// the host and, as level_off_d, on the device) those of height h + 1.
hipError_t launch_refit(const RefitArgs& a, const uint32_t* inner, const uint32_t* level_off,
                        const uint32_t* level_off_d, uint32_t n_levels, hipStream_t stream);
hipError_t launch_bake_static(const CacheArgs& a, hipStream_t stream);
hipError_t launch_mark_blocks(const CacheArgs& a, const SweptBoxes& boxes, hipStream_t stream);
hipError_t launch_rebake(const CacheArgs& a, const MovingMeshes& moving, uint32_t grid, hipStream_t stream);

}  // namespace cnq
