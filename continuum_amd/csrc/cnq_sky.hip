// The sky light over a list of receivers: the reference's `sky_irradiance`
// (crates/continuum-gpu/src/mask_cones.wgsl :1411-1426), six cube-face mask
// cones per receiver through the scene with meshes as independent occluders
// (:1482-1488), each face's light its horizon-clipped window weight less its
// covered bins' weights (:1492-1501), summed over the faces in order.
//
// Organisation on this GPU: one wave per (face, tile of receivers),
// face-major, so a launch has six times as many waves as a light's and a
// wave's cones share a face (one direction range: one packet through the
// walk of cnq_walk.h); a tile is 64 consecutive receivers of a list, or a
// tile_w x 64 / tile_w patch of a view's pixels, whose cones start close
// together. A second pass sums each receiver's faces in the reference's
// order, so the result does not depend on which wave ends first.
#include <hip/hip_runtime.h>

#include "cnq_walk.h"

namespace cnq {

template <bool COUNT>
__global__ __launch_bounds__(64, WALK_WAVES) void sky_items_kernel(
    const NodeC* __restrict__ blas, const NodeC* __restrict__ nodes, const float* __restrict__ tri_leaf,
    const float* __restrict__ tris, const uint32_t* __restrict__ ranges, const uint32_t* __restrict__ items,
    const float* __restrict__ blas_cov, const float* __restrict__ xforms, const float* __restrict__ receivers,
    float* __restrict__ out_faces, uint64_t* __restrict__ out_full, uint64_t* __restrict__ work_out,
    const uint32_t* __restrict__ order, uint32_t* __restrict__ item_cost, uint32_t n, uint32_t w, uint32_t h,
    uint32_t tw_log2, QueryParams p, SkyParams sp) {
  __shared__ float s_acc[ACC_CELLS * 64];
  const uint32_t lane = lane_id();
  float poly[POLY_SLOTS];
  Lane L{s_acc + lane, poly, POLY_STRIDE, ACC_CELLS, 0u};
  const Ctx x{blas, nodes, tri_leaf, tris, ranges, items, nullptr, blas_cov, xforms,
              p.n_objects, p.hierarchy, p.profile_stage, p.object_cov};
  const uint32_t tiles_x = (w + (1u << tw_log2) - 1u) >> tw_log2;
  const uint32_t tiles = tiles_x * ((h + (64u >> tw_log2) - 1u) / (64u >> tw_log2));
  const uint32_t item = order ? order[blockIdx.x] : blockIdx.x;
  const uint64_t started = item_cost ? wall_clock64() : 0;
  const uint32_t face = item / tiles, tile = item - face * tiles;
  const uint32_t tx = tile % tiles_x, ty = tile / tiles_x;
  const uint32_t px = (tx << tw_log2) + (lane & ((1u << tw_log2) - 1u)), py = (ty << (6u - tw_log2)) + (lane >> tw_log2);
  const bool valid = px < w && py < h;
  const uint32_t i = valid ? py * w + px : 0u;
  const float* rc = receivers + 8 * (size_t)(valid ? i : 0u);
  const bool active = valid && rc[7] >= 0.0f;
  const SkyOut o = evaluate_sky_face<COUNT>(x, p, sp, L, valid, ld3(rc), ld3(rc + 4), active, face, work_out);
  if (valid) {
    out_faces[(size_t)face * n + i] = o.value;
    if (out_full) out_full[(size_t)face * n + i] = o.full;
  }
  if (item_cost && lane == 0) {
    const uint64_t t = wall_clock64() - started;
    item_cost[item] = t < 0xffffffffull ? (uint32_t)t : 0xffffffffu;
  }
}

// WGSL :1422-1424: total += each face's light, faces 0..5 in order.
__global__ __launch_bounds__(256) void sky_sum_kernel(const float* __restrict__ faces, float* __restrict__ out,
                                                      uint32_t n) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n) return;
  float total = 0.0f;
#pragma unroll
  for (uint32_t f = 0; f < 6; f++) total += faces[(size_t)f * n + i];
  out[i] = total;
}

uint64_t sky_items(const SkyArgs& a) {
  const uint32_t w = a.w ? a.w : a.n, h = a.w ? a.h : 1u, tw = a.w ? (1u << a.tile_w_log2) : 64u;
  return 6ull * ((w + tw - 1u) / tw) * ((h + 64u / tw - 1u) / (64u / tw));
}

hipError_t launch_sky(const SkyArgs& a, bool count_work, hipStream_t stream) {
  if (a.n == 0) return hipSuccess;
  const SceneArgs& s = a.s;
  const uint32_t w = a.w ? a.w : a.n, h = a.w ? a.h : 1u, twl = a.w ? a.tile_w_log2 : 6u;
  const dim3 grid((uint32_t)sky_items(a)), block(64);
  if (count_work)
    hipLaunchKernelGGL(sky_items_kernel<true>, grid, block, 0, stream, s.blas, s.nodes, s.tri_leaf, s.tris, s.ranges,
                       s.items, s.blas_cov, s.xforms, a.receivers, a.out_faces, a.out_full, a.work, a.order, a.item_cost,
                       a.n, w, h, twl, a.p, a.sky);
  else
    hipLaunchKernelGGL(sky_items_kernel<false>, grid, block, 0, stream, s.blas, s.nodes, s.tri_leaf, s.tris,
                       s.ranges, s.items, s.blas_cov, s.xforms, a.receivers, a.out_faces, a.out_full, nullptr, a.order,
                       a.item_cost, a.n, w, h, twl, a.p, a.sky);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(sky_sum_kernel, dim3((a.n + 255u) / 256u), dim3(256), 0, stream, a.out_faces, a.out, a.n);
  return hipGetLastError();
}

}  // namespace cnq
