// Cone-visibility query kernel on gfx950: work items of 64 queries (a tile
// of one view's shading points against one light), one wave each, in a
// direct grid or a persistent one that takes items from a counter.
// The per-query walk is in cnq_walk.h.
#include "cnq_walk.h"

namespace cnq {


// The view holding work item `item` (uniform): binary search over the
// views' first items.
__device__ __forceinline__ uint32_t find_view(const ViewDesc* __restrict__ views, uint32_t n_views, uint32_t item) {
  uint32_t lo = 0, hi = n_views;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (views[mid].item_off <= item)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// Work items: (view, light, tile of 64 receivers). A view is a w x h grid
// of receivers (a camera view, or a list as n x 1); a tile is tile_w x
// tile_h of them, so a wave's packet covers a compact patch of pixels. The
// query of (view, light l, pixel y * w + x) is stored at
// view.out_off + l * w * h + y * w + x. Direct mode: one item per block;
// persistent mode: a fixed grid takes items from `work_counter` in order.
// FAST: the launch walks both hierarchies (hierarchy 2) with no profiling
// stage, so those parameters are constants here and the other paths are not
// compiled into this variant (a smaller kernel, kinder to the instruction
// cache); other launches take the general variant. ONE: one wave per SIMD,
// the clipper's workspace in LDS (lane-interleaved) instead of private
// memory, and up to 512 registers a lane.
template <bool COUNT, bool PERSISTENT, bool FAST, bool ONE = false>
__global__ __launch_bounds__(64, ONE ? 1 : WALK_WAVES) void cone_items_kernel(
    const NodeC* __restrict__ blas, const NodeC* __restrict__ nodes, const float* __restrict__ tri_leaf,
    const float* __restrict__ tris, const uint32_t* __restrict__ ranges, const uint32_t* __restrict__ items,
    const float* __restrict__ blas_cov, const float* __restrict__ xforms, const float* __restrict__ receivers,
    const float* __restrict__ lights, const ViewDesc* __restrict__ views, float* __restrict__ out_visible,
    uint64_t* __restrict__ out_full, uint64_t* __restrict__ out_touched, float* __restrict__ out_cov,
    uint32_t* __restrict__ counters, uint64_t* __restrict__ work_out, uint32_t* __restrict__ work_counter,
    const uint32_t* __restrict__ order, uint32_t* __restrict__ item_cost, uint8_t* __restrict__ query_flags,
    uint64_t n_flags, QueryParams pa, ViewDesc v0) {
  __shared__ float s_acc[ACC_CELLS * 64];
  __shared__ float s_poly[ONE ? POLY_SLOTS * 64 : 1];
  const uint32_t lane = lane_id();
  float poly[ONE ? 1 : POLY_SLOTS];
  Lane L{s_acc + lane, ONE ? s_poly + lane : poly, ONE ? POLY_STRIDE_LDS : POLY_STRIDE, ACC_CELLS, 0u};
  QueryParams p = pa;
  if (FAST) {
    p.hierarchy = 2;
    p.profile_stage = 0;
  }
  const Ctx x{blas, nodes, tri_leaf, tris, ranges, items, nullptr, blas_cov, xforms,
              p.n_objects, p.hierarchy, p.profile_stage, p.object_cov};
  const uint32_t tw_log2 = p.tile_w_log2;
  for (uint32_t round = 0;; round++) {
    uint32_t item;
    if (PERSISTENT) {
      uint32_t got = 0;
      if (lane == 0) got = atomicAdd(work_counter, 1u);
      got = uni(got);
      if (got >= p.n_items - p.item_begin) break;
      item = p.item_begin + (order ? order[got] : got);
    } else {
      item = p.item_begin + (order ? order[blockIdx.x] : blockIdx.x);
    }
    if (item >= p.n_items) break;
    const uint64_t started = item_cost ? wall_clock64() : 0;
    const ViewDesc v = p.n_views == 1 ? v0 : views[find_view(views, p.n_views, item)];
    const uint32_t local = item - v.item_off;
    const uint32_t light = local / v.tiles;
    const uint32_t tile = local - light * v.tiles;
    const uint32_t tx = tile % v.tiles_x, ty = tile / v.tiles_x;
    const uint32_t xpix = (tx << tw_log2) + (lane & ((1u << tw_log2) - 1u));
    const uint32_t ypix = (ty << (6u - tw_log2)) + (lane >> tw_log2);
    const bool valid = xpix < v.w && ypix < v.h;
    const uint64_t pixel = valid ? (uint64_t)ypix * v.w + xpix : 0;
    const float* rc = receivers + 8 * (v.recv_off + pixel);
    const LightRect r = reinterpret_cast<const LightRect*>(lights)[v.light_off + light];
    const uint64_t q = v.out_off + (uint64_t)light * ((uint64_t)v.w * v.h) + pixel;
    const QueryOut o = evaluate_query<COUNT>(x, p, L, valid, rc, r, out_cov ? out_cov + 64 * q : nullptr, work_out);
    if (valid) {
      out_visible[q] = o.visible;
      out_full[q] = o.full;
      out_touched[q] = o.touched;
      if (query_flags && q < n_flags)
        query_flags[q] = (uint8_t)((o.capped ? 1u : 0u) | (o.overflow ? 2u : 0u) | (o.truncated ? 4u : 0u));
    }
    const uint64_t cap_m = ballot(valid && o.capped), of_m = ballot(valid && o.overflow),
                   tr_m = ballot(valid && o.truncated);
    if (lane == 0 && counters) {
      if (cap_m) atomicAdd(&counters[0], (uint32_t)__builtin_popcountll(cap_m));
      if (of_m) atomicAdd(&counters[1], (uint32_t)__builtin_popcountll(of_m));
      if (tr_m) atomicAdd(&counters[2], (uint32_t)__builtin_popcountll(tr_m));
    }
    if (item_cost && lane == 0) {
      const uint64_t t = wall_clock64() - started;
      item_cost[item - p.item_begin] = t < 0xffffffffull ? (uint32_t)t : 0xffffffffu;
    }
    if (!PERSISTENT) break;
  }
}

#define CNQ_ITEMS_ARGS                                                                                         \
  s.blas, s.nodes, s.tri_leaf, s.tris, s.ranges, s.items, s.blas_cov, s.xforms, a.receivers, a.lights, a.views, \
      a.out_visible, a.out_full, a.out_touched, a.out_cov, a.counters, a.work, a.work_counter, a.order, a.item_cost, \
      a.query_flags, a.n_flags, a.p, a.v0

template <bool FAST, bool ONE>
void launch_items_variant(const ItemArgs& a, bool count_work, uint32_t persistent_blocks, hipStream_t stream) {
  const SceneArgs& s = a.s;
  const dim3 block(64);
  const uint32_t n = a.p.n_items - a.p.item_begin;
  if (persistent_blocks) {
    const dim3 grid(persistent_blocks < n ? persistent_blocks : n);
    if (count_work)
      hipLaunchKernelGGL((cone_items_kernel<true, true, FAST>), grid, block, 0, stream, CNQ_ITEMS_ARGS);
    else
      hipLaunchKernelGGL((cone_items_kernel<false, true, FAST, ONE>), grid, block, 0, stream, CNQ_ITEMS_ARGS);
  } else {
    const dim3 grid(n);
    if (count_work)
      hipLaunchKernelGGL((cone_items_kernel<true, false, FAST>), grid, block, 0, stream, CNQ_ITEMS_ARGS);
    else
      hipLaunchKernelGGL((cone_items_kernel<false, false, FAST, ONE>), grid, block, 0, stream, CNQ_ITEMS_ARGS);
  }
}

// Occupancy: two waves per SIMD (the clipper's workspace in private memory),
// or with `waves` 1 the one-wave variant (workspace in LDS), for launches
// without the work counters that walk both hierarchies unprofiled.
hipError_t launch_cone_items(const ItemArgs& a, bool count_work, uint32_t persistent_blocks, hipStream_t stream,
                             uint32_t waves) {
  if (a.p.n_items <= a.p.item_begin) return hipSuccess;
  const bool fast = a.p.hierarchy == 2 && a.p.profile_stage == 0;
  if (fast && waves == 1 && !count_work)
    launch_items_variant<true, true>(a, count_work, persistent_blocks, stream);
  else if (fast)
    launch_items_variant<true, false>(a, count_work, persistent_blocks, stream);
  else
    launch_items_variant<false, false>(a, count_work, persistent_blocks, stream);
  return hipGetLastError();
}

// One block: items bucketed by the power of two of their time, buckets
// longest first (within a bucket the order is the atomics'), when this
// launch's heaviest items (the top two powers of two) were mostly heavy in
// the last launch too (the same view, or one that moved a little), or when
// there is no last launch; otherwise natural order, since another view's
// order would only scatter
// the waves that run together. Comparing the heavy sets, not every item's
// time, keeps the decision stable where the order itself shifts the times
// a little. `last` keeps each item's bucket (0xff: none yet). A permutation
// of [0, n) whatever the times.
__device__ __forceinline__ uint32_t time_bucket(uint32_t t) { return 32u - (uint32_t)__builtin_clz(t | 1u); }

__global__ __launch_bounds__(1024) void item_order_kernel(const uint32_t* __restrict__ cost,
                                                         uint8_t* __restrict__ last, uint32_t* __restrict__ order,
                                                         uint32_t n) {
  __shared__ uint32_t count[33], next[33], top_new, top_old, unknown, heavy, kept;
  const uint32_t t = threadIdx.x;
  if (t < 33) count[t] = 0;
  if (t == 0) top_new = 0, top_old = 0, unknown = 0, heavy = 0, kept = 0;
  __syncthreads();
  uint32_t my_top_new = 0, my_top_old = 0, my_unknown = 0;
  for (uint32_t i = t; i < n; i += 1024u) {
    const uint32_t b = time_bucket(cost[i]), was = last[i];
    my_top_new = b > my_top_new ? b : my_top_new;
    if (was == 0xffu)
      my_unknown++;
    else
      my_top_old = was > my_top_old ? was : my_top_old;
    atomicAdd(&count[b], 1u);
  }
  atomicMax(&top_new, my_top_new);
  atomicMax(&top_old, my_top_old);
  atomicAdd(&unknown, my_unknown);
  __syncthreads();
  uint32_t my_heavy = 0, my_kept = 0;
  for (uint32_t i = t; i < n; i += 1024u) {
    const uint32_t b = time_bucket(cost[i]), was = last[i];
    if (b + 1u >= top_new) {
      my_heavy++;
      if (was != 0xffu && was + 1u >= top_old) my_kept++;
    }
    last[i] = (uint8_t)b;
  }
  atomicAdd(&heavy, my_heavy);
  atomicAdd(&kept, my_kept);
  __syncthreads();
  // After the first launch of a view (no history) the next is taken to be
  // the same view; after that the heavy sets decide.
  const bool similar = unknown == n || (unknown == 0 && 2u * kept >= heavy);
  if (t == 0) {
    uint32_t at = 0;
    for (int b = 32; b >= 0; b--) {
      next[b] = at;
      at += count[b];
    }
  }
  __syncthreads();
  for (uint32_t i = t; i < n; i += 1024u) {
    if (similar) {
      const uint32_t pos = atomicAdd(&next[time_bucket(cost[i])], 1u);
      order[pos] = i;
    } else {
      order[i] = i;
    }
  }
}

hipError_t launch_item_order(const uint32_t* cost, uint8_t* last, uint32_t* order, uint32_t n, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(item_order_kernel, dim3(1), dim3(1024), 0, stream, cost, last, order, n);
  return hipGetLastError();
}

}  // namespace cnq
