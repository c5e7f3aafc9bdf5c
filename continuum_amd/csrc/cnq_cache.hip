// The lighting cache on gfx950: lightmap texels, stored in blocks of 8 x 8,
// keep direct irradiance from one rectangular light; when meshes move, only
// the blocks whose cones the swept bounds can reach are re-baked, from a
// cache of the static meshes' per-texel coverage plus the moving meshes.
//
// Follows crates/continuum-gpu/src/mask_cones.wgsl :2408-2520
// (`mark_blocks` :2430, `bake_static` :2468, `cached_irradiance` :2491,
// `rebake` :2514) as crates/continuum-gpu/src/live.rs:515-720 and :815-900
// drive them. One wave per block, one lane per texel: a block's 64 texels
// are one packet through the walk in cnq_walk.h.
#include "cnq_walk.h"

namespace cnq {

// WGSL :2430-2452: one lane per block; a block is marked unless, for every
// swept box, one of its planes separates them. Marked blocks are appended
// to `dirty` with one atomic per wave.
__global__ __launch_bounds__(256) void mark_blocks_kernel(const float4* __restrict__ planes, uint32_t n_blocks,
                                                          SweptBoxes sw, uint32_t* __restrict__ dirty,
                                                          uint32_t* __restrict__ n_dirty) {
  const uint32_t b = blockIdx.x * 256u + threadIdx.x;
  bool marked = false;
  if (b < n_blocks) {
    float4 pl[BLOCK_PLANES];
#pragma unroll
    for (uint32_t p = 0; p < BLOCK_PLANES; p++) pl[p] = planes[(size_t)p * n_blocks + b];
    for (uint32_t k = 0; k < sw.n && !marked; k++) {
      const f3 lo{sw.lo[k].x, sw.lo[k].y, sw.lo[k].z}, hi{sw.hi[k].x, sw.hi[k].y, sw.hi[k].z};
      const f3 center = (lo + hi) * 0.5f, extent = (hi - lo) * 0.5f;
      bool separated = false;
#pragma unroll
      for (uint32_t p = 0; p < BLOCK_PLANES; p++) {
        const f3 n{pl[p].x, pl[p].y, pl[p].z};
        const f3 an{fabsf(pl[p].x), fabsf(pl[p].y), fabsf(pl[p].z)};
        separated = separated || dot3(n, center) - dot3(an, extent) > pl[p].w;
      }
      marked = !separated;
    }
  }
  const uint64_t m = ballot(marked);
  if (!m) return;
  const uint32_t lane = lane_id();
  uint32_t base = 0;
  if (lane == 0) base = atomicAdd(n_dirty, (uint32_t)__builtin_popcountll(m));
  base = (uint32_t)__shfl((int)base, 0, 64);
  if (marked) {
    const uint32_t below = (uint32_t)__builtin_popcountll(m & ((1ull << lane) - 1ull));
    dirty[base + below] = b;
  }
}

// Walk state for a texel's cone, every bin empty, no cap.
__device__ __forceinline__ void uncapped_walk(Walk& w, float prefilter) {
#pragma unroll
  for (uint32_t b = 0; b < BINS; b++) w.cov[b] = 0.0f;
  w.full = 0;
  w.cone_work = 0;
  w.cone_cap = NONE;
  w.prefilter_bins = prefilter;
  w.overflow = false;
  w.truncated = false;
#pragma unroll
  for (int k = 0; k < 6; k++) w.work[k] = 0;
}

// The reference's facing test of the cache kernels: the light's middle
// above the texel's horizon.
__device__ __forceinline__ bool faces_rect(const LightRect& r, f3 point, f3 normal) {
  const f3 mid = f3{r.corner.x, r.corner.y, r.corner.z} +
                 (f3{r.edge_u.x, r.edge_u.y, r.edge_u.z} + f3{r.edge_v.x, r.edge_v.y, r.edge_v.z}) * 0.5f;
  return dot3(normal, mid - point) > 0.0f;
}

// WGSL :2468-2487: one wave per block with a slot; the static meshes'
// coverage of each texel's cone, uncapped, into static_cov.
__global__ __launch_bounds__(64, WALK_WAVES) void bake_static_kernel(
    const NodeC* __restrict__ blas, const NodeC* __restrict__ static_nodes, const float* __restrict__ tri_leaf,
    const float* __restrict__ tris, const uint32_t* __restrict__ ranges, const uint32_t* __restrict__ static_items,
    const float* __restrict__ blas_cov, const float* __restrict__ xforms, const float* __restrict__ texels,
    const uint32_t* __restrict__ slot_block, float* __restrict__ static_cov, LightRect light, QueryParams p) {
  __shared__ float s_acc[ACC_CELLS * 64];
  const uint32_t lane = lane_id();
  float poly[POLY_SLOTS];
  Lane L{s_acc + lane, poly, POLY_STRIDE, ACC_CELLS, 0u};
  const Ctx x{blas, static_nodes, tri_leaf, tris, ranges, static_items, nullptr, blas_cov, xforms,
              p.n_objects, 2u, p.profile_stage, NONE};
  const uint32_t slot = blockIdx.x;
  const uint64_t i = 64ull * slot_block[slot] + lane;
  const float* rc = texels + 8 * i;
  const f3 point = ld3(rc), normal = ld3(rc + 4);
  Walk w;
  float cov[BINS];
  w.cov = cov;
  uncapped_walk(w, p.prefilter);
  const bool facing = faces_rect(light, point, normal);
  const LightCone lc = light_cone_rect(point, light, p.wide_extent);
  scene_coverage<false>(x, lc.c, facing && lc.ok, w, L);
  float* out = static_cov + (size_t)slot * BINS * 64 + lane;
#pragma unroll
  for (uint32_t b = 0; b < BINS; b++) out[b * 64] = w.cov[b];
}

// WGSL :2491-2520: one wave per marked block (a grid-stride loop over the
// count mark_blocks left): each texel's cached static coverage, plus every
// moving mesh whose padded bounds its cone may touch, then the window's
// visible weight, into the texel's irradiance.
__global__ __launch_bounds__(64, WALK_WAVES) void rebake_kernel(
    const NodeC* __restrict__ blas, const float* __restrict__ tri_leaf, const float* __restrict__ tris,
    const uint32_t* __restrict__ ranges, const float* __restrict__ blas_cov, const float* __restrict__ xforms,
    const float* __restrict__ texels, const uint32_t* __restrict__ block_slot, const float* __restrict__ static_cov,
    const uint32_t* __restrict__ dirty, const uint32_t* __restrict__ n_dirty, float* __restrict__ irradiance,
    LightRect light, QueryParams p, MovingMeshes mv) {
  __shared__ float s_acc[ACC_CELLS * 64];
  const uint32_t lane = lane_id();
  float poly[POLY_SLOTS];
  Lane L{s_acc + lane, poly, POLY_STRIDE, ACC_CELLS, 0u};
  const Ctx x{blas, nullptr, tri_leaf, tris, ranges, nullptr, nullptr, blas_cov, xforms,
              p.n_objects, 2u, p.profile_stage, NONE};
  const uint32_t count = *n_dirty;
  for (uint32_t j = blockIdx.x; j < count; j += gridDim.x) {
    const uint32_t block = dirty[j];
    const uint32_t slot = block_slot[block];
    const uint64_t i = 64ull * block + lane;
    float value = 0.0f;
    if (slot != NONE) {
      const float* rc = texels + 8 * i;
      const f3 point = ld3(rc), normal = ld3(rc + 4);
      const bool facing = faces_rect(light, point, normal);
      const LightCone lc = light_cone_rect(point, light, p.wide_extent);
      const bool on = facing && lc.ok;
      Walk w;
      float cov[BINS];
      w.cov = cov;
      uncapped_walk(w, p.prefilter);
      const float* cached = static_cov + (size_t)slot * BINS * 64 + lane;
#pragma unroll
      for (uint32_t b = 0; b < BINS; b++) {
        w.cov[b] = on ? cached[b * 64] : 0.0f;
        if (w.cov[b] >= 1.0f) w.full++;
      }
      Stack st{0u, 0u, 0u};
      for (uint32_t k = 0; k < mv.n; k++) {
        const bool need = on && w.full < BINS &&
                          may_touch_box(lc.c, f3{mv.lo[k].x, mv.lo[k].y, mv.lo[k].z}, f3{mv.hi[k].x, mv.hi[k].y, mv.hi[k].z});
        add_mesh_coverage<false>(x, lc.c, mv.mesh[k], need, w, st, 0, L);
      }
      uint64_t touched = 0;
#pragma unroll
      for (uint32_t b = 0; b < BINS; b++) {
        touched |= (uint64_t)(w.cov[b] > 0.0f) << b;
        L.set(b, w.cov[b]);
      }
      if (on) {
        value = cone_visible(L, lc.c, normal, touched);
        if (lc.wide_clipped) {
          bool all_above;
          const float kept = window_weight4(L, lc.c, normal, all_above);
          value = (kept > 1e-9f ? value / kept : 1.0f) * rect_open(L, point, normal, light);
        }
      }
    }
    irradiance[i] = value;
  }
}

hipError_t launch_bake_static(const CacheArgs& a, hipStream_t stream) {
  if (a.n_slots == 0) return hipSuccess;
  const SceneArgs& s = a.s;
  hipLaunchKernelGGL(bake_static_kernel, dim3(a.n_slots), dim3(64), 0, stream, s.blas, a.static_nodes, s.tri_leaf,
                     s.tris, s.ranges, a.static_items, s.blas_cov, s.xforms, a.texels, a.slot_block, a.static_cov,
                     a.light, a.p);
  return hipGetLastError();
}

hipError_t launch_mark_blocks(const CacheArgs& a, const SweptBoxes& boxes, hipStream_t stream) {
  if (a.n_blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(mark_blocks_kernel, dim3((a.n_blocks + 255) / 256), dim3(256), 0, stream, a.planes, a.n_blocks,
                     boxes, a.dirty, a.n_dirty);
  return hipGetLastError();
}

hipError_t launch_rebake(const CacheArgs& a, const MovingMeshes& moving, uint32_t grid, hipStream_t stream) {
  if (a.n_blocks == 0 || grid == 0) return hipSuccess;
  const SceneArgs& s = a.s;
  hipLaunchKernelGGL(rebake_kernel, dim3(grid), dim3(64), 0, stream, s.blas, s.tri_leaf, s.tris, s.ranges,
                     s.blas_cov, s.xforms, a.texels, a.block_slot, a.static_cov, a.dirty, a.n_dirty, a.irradiance,
                     a.light, a.p, moving);
  return hipGetLastError();
}

}  // namespace cnq
